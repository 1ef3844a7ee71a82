// Learned closure kernels (CDNA4 / gfx950): bf16 MFMA convolutions with FP32
// accumulation for the forward pass, the input gradient and the weight
// gradient; reductions, the optimizer, and the fine <-> coarse plumbing.
//
// Convolution as an implicit GEMM: D[co][px] = sum_k W[co][k] * X[k][px],
// k = (tap, ci). One wavefront computes 32 output channels x 32 pixels of one
// row with v_mfma_f32_32x32x16_bf16 (16 input channels of one tap per
// instruction), or 16 channels x 2 x 16 pixels with v_mfma_f32_16x16x32_bf16
// (the 16-channel output layer, 32 input channels per instruction). Operand
// fragments are 16-byte loads: weights [tap][co][ci] give a lane 8
// consecutive ci of one co; NHWC activations give a lane 8 consecutive ci of
// one pixel. The accumulator leaves each lane one pixel and groups of 4
// consecutive channels, stored as 8-byte NHWC vectors.
//
// Weight gradient: dW[tap][co][ci] = sum_px dPre[px][co] * X[px + tap][ci],
// a GEMM whose reduction runs over pixels, on v_mfma_f32_16x16x32_bf16 with
// fragments gathered pixel by pixel; every workgroup sums its pixels into a
// partial dW, and a second pass adds the partials in a fixed order.
#include <cmath>

#include "nemo_closure.h"
#include "nemo_hip_common.h"
#include "nemo_point.h"

namespace nemo {
namespace cl {
namespace {

typedef __attribute__((ext_vector_type(8))) __bf16 bf16x8;
typedef __attribute__((ext_vector_type(4))) __bf16 bf16x4;
typedef __attribute__((ext_vector_type(16))) float f32x16;
typedef __attribute__((ext_vector_type(4))) float f32x4;

constexpr float kInvSqrt2 = 0.70710678118654752f;
constexpr float kInvSqrt2Pi = 0.39894228040143268f;

// GELU, erf form: 0.5 x (1 + erf(x / sqrt 2)), and its derivative
// Phi(x) + x phi(x).
__device__ __forceinline__ float gelu(float x) { return 0.5f * x * (1.0f + erff(x * kInvSqrt2)); }
__device__ __forceinline__ float gelu_grad(float x)
{
    return 0.5f * (1.0f + erff(x * kInvSqrt2)) + x * expf(-0.5f * x * x) * kInvSqrt2Pi;
}

__device__ __forceinline__ size_t pad_index(int y, int x, int nx)
{
    return (size_t)(y + 2) * (size_t)(nx + 4) + (size_t)(x + 2);
}

// Activations are stored in planes of 16 channels: channel c of padded pixel
// p at (c / 16) * plane + p * 16 + c % 16, so that the 8-channel fragments
// of 32 consecutive pixels form one contiguous kilobyte.
__device__ __forceinline__ size_t plane_elems(int nx, int ny)
{
    return (size_t)(ny + 4) * (size_t)(nx + 4) * 16;
}

// Weights live in LDS for the whole launch (rows of cin bf16 padded by 8 so
// that the 16-byte fragment reads of 16 lanes spread over the banks);
// activations come from global memory through the L1, where a 32-pixel row
// segment serves all 25 taps of its tile. Each wavefront computes two tiles
// per iteration, so every weight fragment read from LDS feeds two MFMAs.
template <int CIN, int COUT, int EPI>
__global__ __launch_bounds__(256) void k_conv(ConvArgs a)
{
    static_assert(CIN == 16 || CIN == 32, "input channels");
    static_assert((COUT == 32) || (COUT == 16 && CIN == 32), "output channels");
    constexpr int ROW = CIN + 8;
    __shared__ __attribute__((aligned(16))) __bf16 wsh[TAPS * COUT * ROW];
    for (int q = (int)threadIdx.x; q < TAPS * COUT * (CIN / 8); q += (int)blockDim.x) {
        const int r = q / (CIN / 8), c8 = q - r * (CIN / 8);
        *reinterpret_cast<bf16x8 *>(&wsh[r * ROW + c8 * 8]) =
            *reinterpret_cast<const bf16x8 *>(a.w + (size_t)r * CIN + c8 * 8);
    }
    __syncthreads();

    const int lane = (int)(threadIdx.x & 63);
    const int wave = (int)(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6));
    const int nwaves = (int)(gridDim.x * (blockDim.x >> 6));
    const int nx = a.nx, Wp = nx + 4;
    const int ntx = nx >> 5, items = a.ny * ntx;
    const size_t PL = plane_elems(nx, a.ny);

    // hidden / input-gradient epilogue of one 32-channel x 32-pixel tile
    auto store32 = [&](const f32x16 &acc, int y, int x0) {
        const size_t px = pad_index(y, x0 + (lane & 31), nx) * 16;
#pragma unroll
        for (int g = 0; g < 4; g++) {
            const int co = g * 8 + (lane >> 5) * 4;
            const size_t o = (size_t)(co >> 4) * PL + px + (co & 15);
            float v[4] = {acc[4 * g], acc[4 * g + 1], acc[4 * g + 2], acc[4 * g + 3]};
            bf16x4 r;
            if constexpr (EPI == EPI_HIDDEN) {
                bf16x4 p;
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    v[k] += a.bias[co + k];
                    p[k] = (__bf16)v[k];
                    r[k] = (__bf16)gelu(v[k]);
                }
                if (a.pre) *reinterpret_cast<bf16x4 *>(a.pre + o) = p;
            } else {  // EPI_DX: gradient times gelu'(pre) of the layer below
                const bf16x4 gp = *reinterpret_cast<const bf16x4 *>(a.gpre + o);
#pragma unroll
                for (int k = 0; k < 4; k++) r[k] = (__bf16)(v[k] * gelu_grad((float)gp[k]));
            }
            *reinterpret_cast<bf16x4 *>(a.out + o) = r;
        }
    };
    // output-layer epilogue: used channels 0..2 sit in lanes 0..15
    auto store16 = [&](const f32x4 &lo, const f32x4 &hi, int y, int x0) {
        if ((lane >> 4) != 0) return;
        float4 r0, r1;
        r0.x = lo[0] + a.bias[0];
        r0.y = lo[1] + a.bias[1];
        r0.z = lo[2] + a.bias[2];
        r0.w = 0.0f;
        r1.x = hi[0] + a.bias[0];
        r1.y = hi[1] + a.bias[1];
        r1.z = hi[2] + a.bias[2];
        r1.w = 0.0f;
        const size_t p = (size_t)y * nx + x0 + (lane & 15);
        if (a.gate) {
            const float g0 = a.gate[p], g1 = a.gate[p + 16];
            r0.x *= g0;
            r0.y *= g0;
            r0.z *= g0;
            r1.x *= g1;
            r1.y *= g1;
            r1.z *= g1;
        }
        float4 *dst = reinterpret_cast<float4 *>(a.outf) + p;
        dst[0] = r0;
        dst[16] = r1;
    };

    for (int it = 2 * wave; it < items; it += 2 * nwaves) {
        const bool two = it + 1 < items;  // wavefront-uniform
        const int ya = it / ntx, xa = (it - ya * ntx) * 32;
        const int it2 = two ? it + 1 : it;
        const int yb = it2 / ntx, xb = (it2 - yb * ntx) * 32;
        // The activation fragments of one kernel row (dy) are loaded one row
        // ahead of the MFMAs that use them (two register buffers), so a
        // wavefront keeps a whole row's loads in flight.
        if constexpr (COUT == 32) {
            constexpr int NC = CIN / 16;
            const int pl = lane & 31, kg = (lane >> 5) * 8;
            f32x16 acc0 = {}, acc1 = {};
            const __bf16 *sa = a.in + (size_t)(ya * Wp + xa + pl) * 16 + kg;
            const __bf16 *sb = a.in + (size_t)(yb * Wp + xb + pl) * 16 + kg;
            bf16x8 buf[2][5][NC][2];
            auto load_row = [&](int dy, int slot) {
#pragma unroll
                for (int dx = 0; dx < 5; dx++)
#pragma unroll
                    for (int c = 0; c < NC; c++) {
                        const size_t o = (size_t)c * PL + (size_t)(dy * Wp + dx) * 16;
                        buf[slot][dx][c][0] = *reinterpret_cast<const bf16x8 *>(sa + o);
                        buf[slot][dx][c][1] = *reinterpret_cast<const bf16x8 *>(sb + o);
                    }
            };
            load_row(0, 0);
#pragma unroll
            for (int dy = 0; dy < 5; dy++) {
                if (dy < 4) load_row(dy + 1, (dy + 1) & 1);
                // keep the next row's loads ahead of this row's MFMAs
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int dx = 0; dx < 5; dx++) {
                    const __bf16 *ws = wsh + ((dy * 5 + dx) * 32 + pl) * ROW + kg;
#pragma unroll
                    for (int c = 0; c < NC; c++) {
                        const bf16x8 w = *reinterpret_cast<const bf16x8 *>(ws + c * 16);
                        acc0 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(w, buf[dy & 1][dx][c][0],
                                                                       acc0, 0, 0, 0);
                        acc1 = __builtin_amdgcn_mfma_f32_32x32x16_bf16(w, buf[dy & 1][dx][c][1],
                                                                       acc1, 0, 0, 0);
                    }
                }
            }
            store32(acc0, ya, xa);
            if (two) store32(acc1, yb, xb);
        } else {
            const int pl = lane & 15, kg = (lane >> 4) * 8;
            f32x4 a0 = {}, a1 = {}, b0 = {}, b1 = {};
            // lane k-group kg covers channels kg..kg+7: plane kg / 16
            const size_t kof = (size_t)(kg >> 4) * PL + (kg & 15);
            const __bf16 *sa = a.in + kof + (size_t)(ya * Wp + xa + pl) * 16;
            const __bf16 *sb = a.in + kof + (size_t)(yb * Wp + xb + pl) * 16;
            bf16x8 buf[2][5][4];
            auto load_row = [&](int dy, int slot) {
#pragma unroll
                for (int dx = 0; dx < 5; dx++) {
                    const size_t o = (size_t)(dy * Wp + dx) * 16;
                    buf[slot][dx][0] = *reinterpret_cast<const bf16x8 *>(sa + o);
                    buf[slot][dx][1] = *reinterpret_cast<const bf16x8 *>(sa + o + 16 * 16);
                    buf[slot][dx][2] = *reinterpret_cast<const bf16x8 *>(sb + o);
                    buf[slot][dx][3] = *reinterpret_cast<const bf16x8 *>(sb + o + 16 * 16);
                }
            };
            load_row(0, 0);
#pragma unroll
            for (int dy = 0; dy < 5; dy++) {
                if (dy < 4) load_row(dy + 1, (dy + 1) & 1);
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int dx = 0; dx < 5; dx++) {
                    const bf16x8 w = *reinterpret_cast<const bf16x8 *>(
                        wsh + ((dy * 5 + dx) * 16 + pl) * ROW + kg);
                    const int s = dy & 1;
                    a0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w, buf[s][dx][0], a0, 0, 0, 0);
                    a1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w, buf[s][dx][1], a1, 0, 0, 0);
                    b0 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w, buf[s][dx][2], b0, 0, 0, 0);
                    b1 = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w, buf[s][dx][3], b1, 0, 0, 0);
                }
            }
            store16(a0, a1, ya, xa);
            if (two) store16(b0, b1, yb, xb);
        }
    }
}

__device__ __forceinline__ bf16x8 gather8(const __bf16 *p, size_t stride)
{
    bf16x8 r;
#pragma unroll
    for (int j = 0; j < 8; j++) r[j] = p[j * stride];
    return r;
}

// Partial weight gradient of one workgroup. Wave w handles the taps w,
// w+4, ...; the dPre fragment of a 32-pixel chunk is shared by its taps.
template <int CO, int CI>
__global__ __launch_bounds__(256) void k_dw(const __bf16 *__restrict__ dpre,
                                            const __bf16 *__restrict__ in,
                                            float *__restrict__ partial, int nx, int ny)
{
    constexpr int MB = CO / 16, NB = CI / 16, MAXT = 7;
    const int lane = (int)(threadIdx.x & 63);
    const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int pl = lane & 15, kg = (lane >> 4) * 8;
    const int Wp = nx + 4, ntx = nx >> 5, items = ny * ntx;
    const size_t PL = plane_elems(nx, ny);
    f32x4 acc[MAXT][MB][NB];
#pragma unroll
    for (int t = 0; t < MAXT; t++)
#pragma unroll
        for (int mb = 0; mb < MB; mb++)
#pragma unroll
            for (int nb = 0; nb < NB; nb++) acc[t][mb][nb] = f32x4{0.f, 0.f, 0.f, 0.f};

    for (int it = (int)blockIdx.x; it < items; it += (int)gridDim.x) {
        const int y = it / ntx, x0 = (it - y * ntx) * 32;
        bf16x8 A[MB];
        const __bf16 *dp = dpre + ((size_t)(y + 2) * Wp + x0 + kg + 2) * 16 + pl;
#pragma unroll
        for (int mb = 0; mb < MB; mb++) A[mb] = gather8(dp + mb * PL, 16);
#pragma unroll
        for (int t = 0; t < MAXT; t++) {
            const int tap = wv + 4 * t;
            if (tap < TAPS) {
                const int dy = tap / 5, dx = tap - 5 * (tap / 5);
                const __bf16 *ip = in + ((size_t)(y + dy) * Wp + x0 + kg + dx) * 16 + pl;
#pragma unroll
                for (int nb = 0; nb < NB; nb++) {
                    const bf16x8 B = gather8(ip + nb * PL, 16);
#pragma unroll
                    for (int mb = 0; mb < MB; mb++)
                        acc[t][mb][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                            A[mb], B, acc[t][mb][nb], 0, 0, 0);
                }
            }
        }
    }
    float *out = partial + (size_t)blockIdx.x * TAPS * CO * CI;
#pragma unroll
    for (int t = 0; t < MAXT; t++) {
        const int tap = wv + 4 * t;
        if (tap < TAPS) {
#pragma unroll
            for (int mb = 0; mb < MB; mb++)
#pragma unroll
                for (int nb = 0; nb < NB; nb++)
#pragma unroll
                    for (int r = 0; r < 4; r++) {
                        const int co = mb * 16 + (lane >> 4) * 4 + r;
                        out[((size_t)tap * CO + co) * CI + nb * 16 + pl] = acc[t][mb][nb][r];
                    }
        }
    }
}

template <int CO>
__global__ __launch_bounds__(256) void k_bias_partial(const __bf16 *__restrict__ dpre,
                                                      float *__restrict__ partial, int nx,
                                                      int ny)
{
    constexpr int PL = 256 / CO;  // pixel lanes
    __shared__ float red[256];
    const int c = (int)threadIdx.x % CO, q = (int)threadIdx.x / CO;
    const int npx = nx * ny;
    float s = 0.0f;
    for (int p = (int)blockIdx.x * PL + q; p < npx; p += (int)gridDim.x * PL) {
        const int y = p / nx, x = p - y * nx;
        s += (float)dpre[(size_t)(c >> 4) * plane_elems(nx, ny) + pad_index(y, x, nx) * 16 +
                         (c & 15)];
    }
    red[threadIdx.x] = s;
    __syncthreads();
    if (q == 0) {
        float t = 0.0f;
        for (int k = 0; k < PL; k++) t += red[k * CO + c];
        partial[(size_t)blockIdx.x * CO + c] = t;
    }
}

__global__ void k_reduce(const float *__restrict__ partial, int nblk, int n,
                         float *__restrict__ out)
{
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= n) return;
    float s = 0.0f;
    for (int b = 0; b < nblk; b++) s += partial[(size_t)b * n + i];
    out[i] = s;
}

// Global gradient norm, and the factor that scales it down to max_norm
// (1 when max_norm <= 0 or the norm is below it): norm_scale = {norm, factor}.
__global__ __launch_bounds__(256) void k_grad_norm(const float *__restrict__ g, int n,
                                                   float max_norm, float *__restrict__ norm_scale)
{
    __shared__ float red[256];
    float s = 0.0f;
    for (int i = (int)threadIdx.x; i < n; i += 256) s += g[i] * g[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const float nrm = sqrtf(red[0]);
        norm_scale[0] = nrm;
        // a non-finite gradient skips the update
        norm_scale[1] = !isfinite(nrm) ? 0.0f
                        : (max_norm > 0.0f && nrm > max_norm) ? max_norm / nrm : 1.0f;
    }
}

__global__ void k_optimizer(int kind, float *__restrict__ w, const float *__restrict__ g,
                            float *__restrict__ m, float *__restrict__ v, int n, float lr,
                            float b1, float b2, float eps, float bc1, float bc2,
                            const float *__restrict__ norm_scale)
{
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i >= n) return;
    if (norm_scale[1] == 0.0f) return;  // non-finite gradient: no update
    const float gi = g[i] * norm_scale[1];
    if (kind == 1) {  // Adam
        const float mi = b1 * m[i] + (1.0f - b1) * gi;
        const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
        m[i] = mi;
        v[i] = vi;
        w[i] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
    } else {          // SGD with momentum b1
        const float mi = b1 * m[i] + gi;
        m[i] = mi;
        w[i] -= lr * mi;
    }
}

__global__ void k_weights_bf16(const float *__restrict__ w, int cin, int cout,
                               __bf16 *__restrict__ fwd, __bf16 *__restrict__ flip)
{
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const int n = TAPS * cin * cout;
    if (i >= n) return;
    const int tap = i / (cin * cout), r = i - tap * cin * cout;
    const int co = r / cin, ci = r - co * cin;
    const __bf16 b = (__bf16)w[i];
    fwd[i] = b;
    if (flip) flip[((size_t)(TAPS - 1 - tap) * cin + ci) * cout + co] = b;
}

__global__ void k_pack(CoarseState c, Norm nm, __bf16 *__restrict__ in0,
                       float *__restrict__ gate, int nx, int ny)
{
    const int p = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (p >= nx * ny) return;
    const int y = p / nx, x = p - y * nx;
    const size_t k = (size_t)(y + 1) * c.W + (x + 1);
    bf16x8 lo = {}, hi = {};
    const float u = ((float)c.u[k] - nm.in_shift[0]) / nm.in_scale[0];
    const float v = ((float)c.v[k] - nm.in_shift[1]) / nm.in_scale[1];
    const float s = ((float)c.ssh[k] - nm.in_shift[2]) / nm.in_scale[2];
    lo[0] = (__bf16)u;
    lo[1] = (__bf16)v;
    lo[2] = (__bf16)s;
    lo[3] = (__bf16)(((float)c.ht[k] - nm.in_shift[3]) / nm.in_scale[3]);
    if (gate) {
        const float act = (fabsf(u) + fabsf(v) + fabsf(s)) / nm.gate_eps;
        gate[p] = nm.gate_eps > 0.0f ? 1.0f - expf(-act * act) : 1.0f;
    }
    bf16x8 *dst = reinterpret_cast<bf16x8 *>(in0 + pad_index(y, x, nx) * CIN0);
    dst[0] = lo;
    dst[1] = hi;
}

__global__ void k_target(const double *__restrict__ fssh, const double *__restrict__ fu,
                         const double *__restrict__ fv, CoarseState c, Norm nm,
                         float *__restrict__ target, int nx, int ny)
{
    const int p = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (p >= nx * ny) return;
    const int y = p / nx, x = p - y * nx;
    const size_t k = (size_t)(y + 1) * c.W + (x + 1);
    float4 t;
    t.x = (float)(fu[k] - c.u[k]) / nm.out_scale[0];
    t.y = (float)(fv[k] - c.v[k]) / nm.out_scale[1];
    t.z = (float)(fssh[k] - c.ssh[k]) / nm.out_scale[2];
    t.w = 0.0f;
    reinterpret_cast<float4 *>(target)[p] = t;
}

// A wet T cell whose ssh the step does not clamp to the tide (bc_ssh sets
// every wet cell with an open (-1) 4-neighbour): the closure corrects and
// learns the ssh of these cells only.
__device__ __forceinline__ bool free_ssh_cell(const int *tm, size_t k, int W)
{
    return tm[k] > 0 && tm[k - 1] >= 0 && tm[k + 1] >= 0 && tm[k - W] >= 0 && tm[k + W] >= 0;
}

__global__ __launch_bounds__(256) void k_loss(const float *__restrict__ pred,
                                              const float *__restrict__ target,
                                              __bf16 *__restrict__ dpre4,
                                              float *__restrict__ partial, int nx, int ny,
                                              const int *__restrict__ tmask, int W,
                                              const float *__restrict__ gate)
{
    __shared__ float red[3][256];
    const int npx = nx * ny;
    const float scale = 2.0f / (3.0f * (float)npx);
    float s0 = 0.0f, s1 = 0.0f, s2 = 0.0f;
    for (int p = (int)(blockIdx.x * blockDim.x + threadIdx.x); p < npx;
         p += (int)(gridDim.x * blockDim.x)) {
        const float4 a = reinterpret_cast<const float4 *>(pred)[p];
        const float4 b = reinterpret_cast<const float4 *>(target)[p];
        const int y = p / nx, x = p - y * nx;
        const bool ssh_free =
            !tmask || free_ssh_cell(tmask, (size_t)(y + 1) * W + (x + 1), W);
        const float d0 = a.x - b.x, d1 = a.y - b.y, d2 = ssh_free ? a.z - b.z : 0.0f;
        s0 += d0 * d0;
        s1 += d1 * d1;
        s2 += d2 * d2;
        // the output is gated (pred = gate * raw output): d loss / d raw
        const float gs = gate ? scale * gate[p] : scale;
        bf16x8 lo = {}, hi = {};
        lo[0] = (__bf16)(gs * d0);
        lo[1] = (__bf16)(gs * d1);
        lo[2] = (__bf16)(gs * d2);
        bf16x8 *dst = reinterpret_cast<bf16x8 *>(dpre4 + pad_index(y, x, nx) * COUT4);
        dst[0] = lo;
        dst[1] = hi;
    }
    red[0][threadIdx.x] = s0;
    red[1][threadIdx.x] = s1;
    red[2][threadIdx.x] = s2;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k)
            for (int c = 0; c < 3; c++) red[c][threadIdx.x] += red[c][threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x < 3) partial[blockIdx.x * 3 + threadIdx.x] = red[threadIdx.x][0];
}

__global__ void k_filter(const double *__restrict__ fssh, const double *__restrict__ fu,
                         const double *__restrict__ fv, int Wf, int f,
                         double *__restrict__ cssh, double *__restrict__ cu,
                         double *__restrict__ cv, int Wc, int nx, int ny)
{
    const int p = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (p >= nx * ny) return;
    const int J = p / nx + 1, I = p - (J - 1) * nx + 1;
    const int i0 = f * (I - 1) + 1, j0 = f * (J - 1) + 1;  // first fine cell
    double s = 0.0, su = 0.0, sv = 0.0;
    for (int b = 0; b < f; b++) {
        for (int a = 0; a < f; a++) s += fssh[(size_t)(j0 + b) * Wf + i0 + a];
        su += fu[(size_t)(j0 + b) * Wf + f * I];
        sv += fv[(size_t)(f * J) * Wf + i0 + b];
    }
    const size_t k = (size_t)J * Wc + I;
    cssh[k] = s / (double)(f * f);
    cu[k] = su / (double)f;
    cv[k] = sv / (double)f;
}

__global__ void k_set_fields(int mode, const double *__restrict__ sssh,
                             const double *__restrict__ su, const double *__restrict__ sv,
                             const float *__restrict__ pred, Norm nm,
                             const int *__restrict__ flag, CoarseState c, int nx, int ny)
{
    const int p = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (p >= nx * ny) return;
    if (flag && *flag == 0) return;
    const int J = p / nx + 1, I = p - (J - 1) * nx + 1;
    const size_t k = (size_t)J * c.W + I;
    const bool wet = c.tmask[k] > 0;
    float4 q = {0.f, 0.f, 0.f, 0.f};
    if (mode == 1) q = reinterpret_cast<const float4 *>(pred)[p];
    if (mode == 0 && wet) c.ssh[k] = sssh[k];
    if (mode == 1 && free_ssh_cell(c.tmask, k, c.W)) c.ssh[k] += (double)(q.z * nm.out_scale[2]);
    if (I < nx && wet && c.tmask[k + 1] > 0) {
        if (mode == 0) c.u[k] = su[k];
        else c.u[k] += (double)(q.x * nm.out_scale[0]);
    }
    if (J < ny && wet && c.tmask[k + c.W] > 0) {
        if (mode == 0) c.v[k] = sv[k];
        else c.v[k] += (double)(q.y * nm.out_scale[1]);
    }
}

__global__ void k_face_ssh(const int *__restrict__ flag, CoarseState c, int nx, int ny)
{
    const int p = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (p >= nx * ny) return;
    if (flag && *flag == 0) return;
    const int J = p / nx + 1, I = p - (J - 1) * nx + 1;
    const size_t k = (size_t)J * c.W + I;
    double val;
    if (I < nx && next_ssh_pt(c.ssh[k], c.ssh[k + 1], c.area_t[k], c.area_t[k + 1],
                              c.area_u[k], c.tmask[k], c.tmask[k + 1], &val))
        c.sshu[k] = val;
    if (J < ny && next_ssh_pt(c.ssh[k], c.ssh[k + c.W], c.area_t[k], c.area_t[k + c.W],
                              c.area_v[k], c.tmask[k], c.tmask[k + c.W], &val))
        c.sshv[k] = val;
}

__global__ __launch_bounds__(256) void k_guard_partial(const float *__restrict__ pred,
                                                       Norm nm, CoarseState c, double g,
                                                       double *__restrict__ partial,
                                                       int nx, int ny)
{
    __shared__ double red[3][256];
    double e0 = 0.0, e1 = 0.0, bad = 0.0;
    for (int p = (int)(blockIdx.x * blockDim.x + threadIdx.x); p < nx * ny;
         p += (int)(gridDim.x * blockDim.x)) {
        const int J = p / nx + 1, I = p - (J - 1) * nx + 1;
        const size_t k = (size_t)J * c.W + I;
        const float4 q = reinterpret_cast<const float4 *>(pred)[p];
        const double s = c.ssh[k], u = c.u[k], v = c.v[k], h = c.ht[k];
        const bool wet = c.tmask[k] > 0;
        const double s1 =
            free_ssh_cell(c.tmask, k, c.W) ? s + (double)(q.z * nm.out_scale[2]) : s;
        const double u1 = (I < nx && wet && c.tmask[k + 1] > 0)
                              ? u + (double)(q.x * nm.out_scale[0]) : u;
        const double v1 = (J < ny && wet && c.tmask[k + c.W] > 0)
                              ? v + (double)(q.y * nm.out_scale[1]) : v;
        e0 += g * s * s + h * (u * u + v * v);
        const double t = g * s1 * s1 + h * (u1 * u1 + v1 * v1);
        if (isfinite(t)) e1 += t;
        else bad += 1.0;
    }
    red[0][threadIdx.x] = e0;
    red[1][threadIdx.x] = e1;
    red[2][threadIdx.x] = bad;
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k)
            for (int r = 0; r < 3; r++) red[r][threadIdx.x] += red[r][threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x < 3) partial[blockIdx.x * 3 + threadIdx.x] = red[threadIdx.x][0];
}

__global__ __launch_bounds__(256) void k_guard_decide(const double *__restrict__ partial,
                                                      int nblk, double rel, double abs_tol,
                                                      int *__restrict__ flag,
                                                      int *__restrict__ trips)
{
    __shared__ double red[3][256];
    double a[3] = {0.0, 0.0, 0.0};
    for (int b = (int)threadIdx.x; b < nblk; b += 256)
        for (int r = 0; r < 3; r++) a[r] += partial[b * 3 + r];
    for (int r = 0; r < 3; r++) red[r][threadIdx.x] = a[r];
    __syncthreads();
    for (int k = 128; k > 0; k >>= 1) {
        if ((int)threadIdx.x < k)
            for (int r = 0; r < 3; r++) red[r][threadIdx.x] += red[r][threadIdx.x + k];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const double e0 = red[0][0], e1 = red[1][0], bad = red[2][0];
        const int ok = bad == 0.0 && isfinite(e1) && e1 <= (1.0 + rel) * e0 + abs_tol;
        *flag = ok;
        if (!ok) *trips += 1;
    }
}

inline dim3 grid_for(long n, int block = 256)
{
    return dim3((unsigned)((n + block - 1) / block));
}

int conv_grid()
{
    static int g = [] {
        int dev = 0, cus = 0;
        NEMO_HIP_CHECK(hipGetDevice(&dev));
        NEMO_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        return cus * 2;  // 4 wavefronts per workgroup; 2 workgroups fit a CU's LDS
    }();
    return g;
}

}  // namespace

void launch_conv(hipStream_t s, int cin, int cout, Epilogue epi, const ConvArgs &a)
{
    if (a.nx % 32 != 0 || a.nx <= 0 || a.ny <= 0)
        NEMO_FATAL("closure conv: nx = %d must be a positive multiple of 32", a.nx);
    const dim3 grid((unsigned)conv_grid()), block(256);
    if (cin == 16 && cout == 32 && epi == EPI_HIDDEN)
        k_conv<16, 32, EPI_HIDDEN><<<grid, block, 0, s>>>(a);
    else if (cin == 32 && cout == 32 && epi == EPI_HIDDEN)
        k_conv<32, 32, EPI_HIDDEN><<<grid, block, 0, s>>>(a);
    else if (cin == 32 && cout == 16 && epi == EPI_OUTPUT)
        k_conv<32, 16, EPI_OUTPUT><<<grid, block, 0, s>>>(a);
    else if (cin == 16 && cout == 32 && epi == EPI_DX)
        k_conv<16, 32, EPI_DX><<<grid, block, 0, s>>>(a);
    else if (cin == 32 && cout == 32 && epi == EPI_DX)
        k_conv<32, 32, EPI_DX><<<grid, block, 0, s>>>(a);
    else
        NEMO_FATAL("closure conv: unsupported %d -> %d channels, epilogue %d", cin, cout,
                   (int)epi);
    NEMO_LAUNCH_CHECK();
}

int dw_blocks() { return 256; }

void launch_dw(hipStream_t s, int co, int ci, const __bf16 *dpre, const __bf16 *in,
               float *partial, int nx, int ny)
{
    if (nx % 32 != 0) NEMO_FATAL("closure dW: nx = %d not a multiple of 32", nx);
    const dim3 grid((unsigned)dw_blocks()), block(256);
    if (co == 32 && ci == 32)
        k_dw<32, 32><<<grid, block, 0, s>>>(dpre, in, partial, nx, ny);
    else if (co == 32 && ci == 16)
        k_dw<32, 16><<<grid, block, 0, s>>>(dpre, in, partial, nx, ny);
    else if (co == 16 && ci == 32)
        k_dw<16, 32><<<grid, block, 0, s>>>(dpre, in, partial, nx, ny);
    else
        NEMO_FATAL("closure dW: unsupported %d x %d", co, ci);
    NEMO_LAUNCH_CHECK();
}

int bias_blocks() { return 128; }

void launch_bias_partial(hipStream_t s, int co, const __bf16 *dpre, float *partial,
                         int nx, int ny)
{
    const dim3 grid((unsigned)bias_blocks()), block(256);
    if (co == 32)
        k_bias_partial<32><<<grid, block, 0, s>>>(dpre, partial, nx, ny);
    else if (co == 16)
        k_bias_partial<16><<<grid, block, 0, s>>>(dpre, partial, nx, ny);
    else
        NEMO_FATAL("closure bias gradient: %d channels", co);
    NEMO_LAUNCH_CHECK();
}

void launch_reduce(hipStream_t s, const float *partial, int nblk, int n, float *out)
{
    k_reduce<<<grid_for(n), dim3(256), 0, s>>>(partial, nblk, n, out);
    NEMO_LAUNCH_CHECK();
}

void launch_optimizer(hipStream_t s, int kind, float *w, const float *g, float *m,
                      float *v, int n, float lr, float b1, float b2, float eps,
                      float bc1, float bc2, float max_norm, float *norm_scale)
{
    k_grad_norm<<<dim3(1), dim3(256), 0, s>>>(g, n, max_norm, norm_scale);
    NEMO_LAUNCH_CHECK();
    k_optimizer<<<grid_for(n), dim3(256), 0, s>>>(kind, w, g, m, v, n, lr, b1, b2, eps,
                                                  bc1, bc2, norm_scale);
    NEMO_LAUNCH_CHECK();
}

void launch_weights_bf16(hipStream_t s, const float *w, int cin, int cout, __bf16 *fwd,
                         __bf16 *flip)
{
    k_weights_bf16<<<grid_for((long)TAPS * cin * cout), dim3(256), 0, s>>>(w, cin, cout,
                                                                          fwd, flip);
    NEMO_LAUNCH_CHECK();
}

void launch_pack(hipStream_t s, const CoarseState &c, const Norm &nm, __bf16 *in0,
                 float *gate, int nx, int ny)
{
    k_pack<<<grid_for((long)nx * ny), dim3(256), 0, s>>>(c, nm, in0, gate, nx, ny);
    NEMO_LAUNCH_CHECK();
}

void launch_target(hipStream_t s, const double *fssh, const double *fu, const double *fv,
                   const CoarseState &c, const Norm &nm, float *target, int nx, int ny)
{
    k_target<<<grid_for((long)nx * ny), dim3(256), 0, s>>>(fssh, fu, fv, c, nm, target,
                                                           nx, ny);
    NEMO_LAUNCH_CHECK();
}

int loss_blocks() { return 256; }

void launch_loss(hipStream_t s, const float *pred, const float *target, __bf16 *dpre4,
                 float *partial, int nx, int ny, const int *tmask, int W,
                 const float *gate)
{
    k_loss<<<dim3((unsigned)loss_blocks()), dim3(256), 0, s>>>(pred, target, dpre4,
                                                               partial, nx, ny, tmask, W,
                                                               gate);
    NEMO_LAUNCH_CHECK();
}

void launch_filter(hipStream_t s, const double *fssh, const double *fu, const double *fv,
                   int Wf, int factor, double *cssh, double *cu, double *cv, int Wc,
                   int nx, int ny)
{
    k_filter<<<grid_for((long)nx * ny), dim3(256), 0, s>>>(fssh, fu, fv, Wf, factor,
                                                           cssh, cu, cv, Wc, nx, ny);
    NEMO_LAUNCH_CHECK();
}

void launch_set_fields(hipStream_t s, int mode, const double *sssh, const double *su,
                       const double *sv, const float *pred, const Norm &nm,
                       const int *flag, const CoarseState &c, int nx, int ny)
{
    k_set_fields<<<grid_for((long)nx * ny), dim3(256), 0, s>>>(mode, sssh, su, sv, pred,
                                                               nm, flag, c, nx, ny);
    NEMO_LAUNCH_CHECK();
}

void launch_face_ssh(hipStream_t s, const int *flag, const CoarseState &c, int nx, int ny)
{
    k_face_ssh<<<grid_for((long)nx * ny), dim3(256), 0, s>>>(flag, c, nx, ny);
    NEMO_LAUNCH_CHECK();
}

int guard_blocks() { return 512; }

void launch_guard_partial(hipStream_t s, const float *pred, const Norm &nm,
                          const CoarseState &c, double g, double *partial, int nx, int ny)
{
    k_guard_partial<<<dim3((unsigned)guard_blocks()), dim3(256), 0, s>>>(pred, nm, c, g,
                                                                         partial, nx, ny);
    NEMO_LAUNCH_CHECK();
}

void launch_guard_decide(hipStream_t s, const double *partial, int nblk, double rel,
                         double abs_tol, int *flag, int *trips)
{
    k_guard_decide<<<dim3(1), dim3(256), 0, s>>>(partial, nblk, rel, abs_tol, flag, trips);
    NEMO_LAUNCH_CHECK();
}

}  // namespace cl
}  // namespace nemo
