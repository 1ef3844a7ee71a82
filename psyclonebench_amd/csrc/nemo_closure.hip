// Learned closure kernels (CDNA4 / gfx950): bf16 MFMA convolutions with FP32
// accumulation for the forward pass, the input gradient and the weight
// gradient; reductions, the optimizer, and the fine <-> coarse plumbing.
//
// Convolution as an implicit GEMM on v_mfma_f32_32x32x16_bf16 (32 output
// channels x 32 pixels x 16 input channels): the weights sit in LDS for the
// whole persistent launch, and one activation fragment (32 pixels of one
// kernel row) feeds the five horizontal taps of that row through five
// accumulators that are shifted into place at the end (k_conv_sh). The
// output layer packs (tap column, channel) into the 16 rows of
// v_mfma_f32_16x16x32_bf16 (k_conv_out). Activation fragments are 16-byte
// loads of 8 consecutive channels of one pixel from a 16-channel plane.
//
// Weight gradient: dW[tap][co][ci] = sum_px dPre[px][co] * X[px + tap][ci],
// a GEMM whose reduction runs over pixels, on v_mfma_f32_16x16x32_bf16 with
// its operands staged pixel-fastest in LDS (k_dw); every workgroup sums its
// pixels into a partial dW, and a second pass adds the partials in a fixed
// order.
#include <cmath>
#include <type_traits>

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

// GELU in its tanh form: 0.5 x (1 + tanh(u / 2)) = x s(u), s the logistic
// sigmoid, u = 2 sqrt(2 / pi) (x + 0.044715 x^3). One hardware exp2 and one
// reciprocal, no branches: the epilogues evaluate 32 of these per pixel and
// layer. exp(-u) = exp2(x (A + B x^2)).
constexpr float kLog2e = 1.44269504088896341f;
constexpr float kGeluK = 2.0f * 0.79788456080286536f;  // 2 sqrt(2 / pi)
constexpr float kGeluC = 0.044715f;
constexpr float kGeluA = -kGeluK * kLog2e;
constexpr float kGeluB = kGeluA * kGeluC;

__device__ __forceinline__ float gelu_sigmoid(float x)
{
    const float e = __builtin_amdgcn_exp2f(x * fmaf(x * x, kGeluB, kGeluA));  // exp(-u)
    return __builtin_amdgcn_rcpf(1.0f + e);
}
__device__ __forceinline__ float gelu(float x) { return x * gelu_sigmoid(x); }
// d/dx x s(u) = s + x s (1 - s) du/dx, du/dx = K (1 + 3 C x^2)
__device__ __forceinline__ float gelu_grad(float x)
{
    const float s = gelu_sigmoid(x);
    const float du = kGeluK * fmaf(3.0f * kGeluC, x * x, 1.0f);
    return fmaf(x * s * (1.0f - s), du, s);
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


// ---- shift-accumulate convolutions ------------------------------------------
//
// One activation fragment (32 input pixels of one kernel row, 16 channels)
// feeds the five horizontal taps of that row: accumulator dx collects
// D_dx[co][n] = sum over dy, ci of W[dy][dx][co][ci] * X[y + dy][x0 + n][ci]
// (padded coordinates), and output pixel x0 + j is sum over dx of
// D_dx[co][j + dx], formed after the reduction by four DPP wavefront shifts.
// A tile of 32 input pixels therefore yields 28 outputs, and the activation
// traffic per output is a fifth of one load per tap. Tiles sit at
// x0 = min(28 t, nx - 28); the last one overlaps its neighbour and stores
// only the pixels x >= 28 t.
//
// Work items (row y, tile t) are split into eight contiguous bands, one per
// XCD (workgroups are dispatched to the XCDs round robin), so that the five
// kernel rows an output row reads come from its own XCD's L2.

// lane i <- lane i + 1 over the whole wavefront (DPP wave_shl:1); lane 63 <- 0
__device__ __forceinline__ float wave_shl1(float v)
{
    return __builtin_bit_cast(
        float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), 0x130, 0xf, 0xf, true));
}

// Items [first, first + count) of this workgroup's XCD band; the workgroup is
// number `local` of `nlocal` on that band. gridDim.x is a multiple of 8.
struct Band {
    int first, count, local, nlocal;
};
__device__ __forceinline__ Band xcd_band(int items)
{
    const int xcd = (int)(blockIdx.x & 7);
    const int first = (int)((long)items * xcd / 8);
    const int last = (int)((long)items * (xcd + 1) / 8);
    return Band{first, last - first, (int)(blockIdx.x >> 3), (int)(gridDim.x >> 3)};
}

// 32 output channels. KIND 1: 16 input channels. KIND 2: 32 input channels.
// (The network input layer, 4 channels, is k_conv_in.)
// EPI_HIDDEN: bias, GELU, optional pre-activation; EPI_DX: times gelu'(gpre).
// NT tiles per wavefront pass, WPB wavefronts per workgroup, two workgroups
// per CU (the LDS weights):
// - NT = 2, WPB = 4: two wavefronts per SIMD (at most 256 registers), each
//   weight fragment read from LDS once for two tiles;
// - NT = 1, WPB = 8: four wavefronts per SIMD (at most 128 registers), one
//   LDS weight read per MFMA, twice the wavefronts to hide the activation
//   loads and overlap one wavefront's epilogue with another's MFMAs.
// Every tile sees the same MFMA sequence in both forms: the same bits.
template <int KIND, int EPI, int NT, int WPB>
__global__ __launch_bounds__(64 * WPB) __attribute__((amdgpu_waves_per_eu(4 / NT, 4 / NT)))
void k_conv_sh(ConvArgs a)
{
    static_assert(KIND == 1 || KIND == 2, "16 or 32 input channels");
    static_assert(EPI == EPI_HIDDEN || EPI == EPI_DX, "epilogue");
    static_assert((NT == 2 && WPB == 4) || (NT == 1 && WPB == 8), "launch form");
    constexpr int NC = KIND == 2 ? 2 : 1;
    constexpr int CIN = KIND == 2 ? 32 : 16;
    constexpr int ROW = CIN + 8;  // LDS row of one output channel, padded by 8
    constexpr int NROWS = TAPS * 32;
    __shared__ __attribute__((aligned(16))) __bf16 wsh[NROWS * ROW];
    // the bias sits in LDS too: an LDS read in the epilogue waits on the LDS
    // counter only, where a global load would wait for every store and
    // every prefetched fragment before it
    __shared__ __attribute__((aligned(16))) float bsh[32];
    // Weight fill (every thread of the workgroup): every global load is
    // issued before the first LDS store, so the workgroup waits for one
    // memory latency, not one per loop trip.
    const int tid = (int)threadIdx.x;
    if constexpr (EPI == EPI_HIDDEN) {
        if (tid < 32) bsh[tid] = a.bias[tid];
    }
    {
        constexpr int NTHR = 64 * WPB;
        constexpr int NTOT = TAPS * 32 * (CIN / 8), NQ = (NTOT + NTHR - 1) / NTHR;
        bf16x8 v[NQ];
#pragma unroll
        for (int i = 0; i < NQ; i++) {
            const int q = tid + NTHR * i;
            if (q < NTOT) v[i] = *reinterpret_cast<const bf16x8 *>(a.w + (size_t)q * 8);
        }
#pragma unroll
        for (int i = 0; i < NQ; i++) {
            const int q = tid + NTHR * i;
            const int r = q / (CIN / 8), c8 = q - r * (CIN / 8);
            if (q < NTOT) *reinterpret_cast<bf16x8 *>(&wsh[r * ROW + c8 * 8]) = v[i];
        }
    }
    __syncthreads();

    const int lane = (int)(threadIdx.x & 63), n = lane & 31, h = lane >> 5;
    const int nx = a.nx, Wp = nx + 4, T = (nx + 27) / 28;
    constexpr int PXS = 16;  // elements per padded pixel of a plane
    const size_t PL = plane_elems(nx, a.ny), RS = (size_t)Wp * PXS;  // row stride
    const Band bd = xcd_band(a.ny * T);
    const int wib = bd.local * WPB + (int)(threadIdx.x >> 6), nwb = bd.nlocal * WPB;

    // (the lambdas are forced inline: an out-of-line call would move the
    // accumulators to scratch memory)
    // tile (row y, column tile t): items are y * T + t
    struct Pos {
        int y, t;
    };
    auto next_pos = [&](Pos p, int dy, int dt) __attribute__((always_inline)) {
        p.y += dy;
        p.t += dt;
        if (p.t >= T) {
            p.t -= T;
            p.y++;
        }
        return p;
    };
    auto finish = [&](f32x16(&acc)[5], Pos p) __attribute__((always_inline)) {
        const int y = p.y, t = p.t, x0 = min(28 * t, nx - 28);
        f32x16 s = acc[4];
#pragma unroll
        for (int dx = 3; dx >= 0; dx--)
#pragma unroll
            for (int r = 0; r < 16; r++) s[r] = wave_shl1(s[r]) + acc[dx][r];
        const int x = x0 + n;
        if (n >= 28 || x < 28 * t) return;
        const size_t px = pad_index(y, x, nx) * 16;
#pragma unroll
        for (int g = 0; g < 4; g++) {
            const int co = g * 8 + h * 4;
            const size_t o = (size_t)(co >> 4) * PL + px + (co & 15);
            bf16x4 r;
            if constexpr (EPI == EPI_HIDDEN) {
                const f32x4 bias = *reinterpret_cast<const f32x4 *>(&bsh[co]);
                bf16x4 p;
#pragma unroll
                for (int k = 0; k < 4; k++) {
                    const float v = s[4 * g + k] + bias[k];
                    p[k] = (__bf16)v;
                    r[k] = (__bf16)gelu(v);
                }
                if (a.pre) *reinterpret_cast<bf16x4 *>(a.pre + o) = p;
            } else {
                const bf16x4 gp = *reinterpret_cast<const bf16x4 *>(a.gpre + o);
#pragma unroll
                for (int k = 0; k < 4; k++)
                    r[k] = (__bf16)(s[4 * g + k] * gelu_grad((float)gp[k]));
            }
            *reinterpret_cast<bf16x4 *>(a.out + o) = r;
        }
    };

    // A wavefront takes the groups of NT consecutive tiles at band offsets
    // k = NT wib, NT (wib + nwb), ...; the activation fragments of the next
    // group are in flight while the current group runs its MFMAs and
    // epilogue.
    int k = NT * wib;
    if (k >= bd.count) return;
    const size_t coff = (size_t)h * 8;
    // positions of the group at k, advanced by NT nwb items per group (no
    // division in the loop)
    const int step = NT * nwb, step_y = step / T, step_t = step - step_y * T;
    Pos pos{(bd.first + k) / T, (bd.first + k) - (bd.first + k) / T * T};
    // tile j of the group at p (the group's last tile when the band ends
    // before it)
    auto member = [&](Pos p, int kk, int j) __attribute__((always_inline)) {
        return j > 0 && kk + j < bd.count ? next_pos(p, 0, j) : p;
    };
    auto group_src = [&](Pos p, int kk, const __bf16 *(&src)[NT])
                         __attribute__((always_inline)) {
#pragma unroll
        for (int j = 0; j < NT; j++) {
            const Pos q = member(p, kk, j);
            const int xq = min(28 * q.t, nx - 28);
            src[j] = a.in + ((size_t)q.y * Wp + xq + n) * PXS + coff;
        }
    };
    const __bf16 *src[NT];
    group_src(pos, k, src);
    f32x16 acc[NT][5];
    auto finish_group = [&](Pos p, int kk) __attribute__((always_inline)) {
        finish(acc[0], p);
        if constexpr (NT == 2) {
            if (kk + 1 < bd.count) finish(acc[1], next_pos(p, 0, 1));  // wavefront-uniform
        }
    };

    // Two register slots of one kernel row each: row dy of a group sits in
    // slot (dy + P) & 1, and row 0 of the next group is loaded with row 4
    // of this one into slot (P + 1) & 1, so consecutive groups alternate P.
    bf16x8 buf[2][NC][NT];
    auto load_row = [&](const __bf16 *const (&sp)[NT], int dy, int slot)
                        __attribute__((always_inline)) {
#pragma unroll
        for (int c = 0; c < NC; c++) {
            const size_t o = (size_t)c * PL + (size_t)dy * RS;
#pragma unroll
            for (int j = 0; j < NT; j++)
                buf[slot][c][j] = *reinterpret_cast<const bf16x8 *>(sp[j] + o);
        }
    };
    auto pair = [&](auto parity) __attribute__((always_inline)) {
        constexpr int P = decltype(parity)::value;
        const bool more = k + NT * nwb < bd.count;  // wavefront-uniform
        const __bf16 *nsrc[NT];
#pragma unroll
        for (int j = 0; j < NT; j++) nsrc[j] = src[j];
        const Pos cur = pos;
        if (more) {
            pos = next_pos(pos, step_y, step_t);
            group_src(pos, k + NT * nwb, nsrc);
        }
#pragma unroll
        for (int j = 0; j < NT; j++)
#pragma unroll
            for (int dx = 0; dx < 5; dx++) acc[j][dx] = f32x16{};
#pragma unroll
        for (int dy = 0; dy < 5; dy++) {
            // the last group loads its own row 0 again (unused): every path has
            // the same loads outstanding, so the waits stay exact
            if (dy < 4)
                load_row(src, dy + 1, (dy + 1 + P) & 1);
            else
                load_row(nsrc, 0, (P + 1) & 1);
            __builtin_amdgcn_sched_barrier(0);
            // channel-chunk-outer order: an accumulator comes back every
            // 5 NT MFMAs, past the latency of the previous one
#pragma unroll
            for (int c = 0; c < NC; c++)
#pragma unroll
                for (int dx = 0; dx < 5; dx++) {
                    const bf16x8 w = *reinterpret_cast<const bf16x8 *>(
                        wsh + ((dy * 5 + dx) * 32 + n) * ROW + h * 8 + c * 16);
#pragma unroll
                    for (int j = 0; j < NT; j++)
                        acc[j][dx] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
                            w, buf[(dy + P) & 1][c][j], acc[j][dx], 0, 0, 0);
                }
        }
        finish_group(cur, k);
        k += NT * nwb;
#pragma unroll
        for (int j = 0; j < NT; j++) src[j] = nsrc[j];
    };
    load_row(src, 0, 0);
    while (true) {
        pair(std::integral_constant<int, 0>{});
        if (k >= bd.count) break;
        pair(std::integral_constant<int, 1>{});
        if (k >= bd.count) break;
    }
}

// Network input layer (4 -> 32 channels) as an implicit GEMM with the taps in
// K: 25 taps x 4 channels (100, padded to 112 = 7 chunks of 16), so one
// accumulator holds the whole 5x5 sum of a 32-pixel tile and no horizontal
// shift-sum or tile overlap is needed. Each half h of the wavefront carries
// two taps of a chunk (8 K values): chunk c < 5 holds kernel row c, columns
// 2h and 2h + 1; chunk 5 column 4 of rows 2h and 2h + 1; chunk 6 tap (4, 4)
// in half 0 and zeros. As A, lane (n, h) holds the weights of output
// channel n at its two taps (the 7 fragments stay in registers for the
// whole launch); as B, the 4 input channels of pixel x0 + n at those taps:
// two 8-byte loads whose addresses differ between chunks by a
// wavefront-uniform row offset. The layer's MFMA work is small and its
// epilogue (GELU of 32 channels a pixel and 64 bytes of stores) large, so
// the kernel keeps its registers under 128 for 4 wavefronts per SIMD
// instead of sharing one loaded fragment among five accumulators. Tiles sit
// at x0 = min(32 t, nx - 32); the last one overlaps its neighbour and stores
// only x >= 32 t.
constexpr int IN_CHUNKS = 7;

// Kernel tap (dy * 5 + dx) of chunk c, half h, slot q; -1: a zero K slot.
__device__ __forceinline__ int in_tap(int c, int h, int q)
{
    if (c < 5) return c * 5 + 2 * h + q;
    if (c == 5) return (2 * h + q) * 5 + 4;
    return h == 0 && q == 0 ? 24 : -1;
}

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 4))) void k_conv_in(
    ConvArgs a)
{
    // the bias sits in LDS: an LDS read in the epilogue waits on the LDS
    // counter only, where a global load would wait for every store and
    // every prefetched fragment before it
    __shared__ __attribute__((aligned(16))) float bsh[HID];
    if (threadIdx.x < HID) bsh[threadIdx.x] = a.bias[threadIdx.x];
    __syncthreads();
    const int lane = (int)(threadIdx.x & 63), n = lane & 31, h = lane >> 5;
    const int nx = a.nx, Wp = nx + 4, T = (nx + 31) / 32;
    const size_t PL = plane_elems(nx, a.ny);
    // A fragments: W[tap][co = n][0..3] at this half's two taps of each chunk
    bf16x8 wf[IN_CHUNKS];
#pragma unroll
    for (int c = 0; c < IN_CHUNKS; c++) {
        const bf16x4 z = {};
        bf16x4 w[2];
#pragma unroll
        for (int q = 0; q < 2; q++) {
            const int tap = in_tap(c, h, q);
            w[q] = tap >= 0 ? *reinterpret_cast<const bf16x4 *>(
                                  a.w + ((size_t)tap * HID + n) * CIN0)
                            : z;
        }
        wf[c] = __builtin_shufflevector(w[0], w[1], 0, 1, 2, 3, 4, 5, 6, 7);
    }
    const Band bd = xcd_band(a.ny * T);
    // the wavefront's index is uniform: position arithmetic stays scalar
    const int wib = __builtin_amdgcn_readfirstlane(bd.local * 4 + (int)(threadIdx.x >> 6));
    const int nwb = bd.nlocal * 4;
    const int RS = Wp * IN0C;  // elements per padded input row
    struct Pos {
        int y, t;
    };
    // item (row y, column tile t) = y * T + t, advanced by nwb items a trip
    const int step_y = nwb / T, step_t = nwb - step_y * T;
    auto advance = [&](Pos p) __attribute__((always_inline)) {
        p.y += step_y;
        p.t += step_t;
        if (p.t >= T) {
            p.t -= T;
            p.y++;
        }
        return p;
    };
    auto load = [&](Pos q, bf16x8(&b)[IN_CHUNKS]) __attribute__((always_inline)) {
        const int y = q.y, t = q.t, x0 = min(32 * t, nx - 32);
        // padded pixel (y, x0 + n): the tile's tap (0, 0) for this lane
        const __bf16 *px = a.in + ((size_t)y * Wp + x0 + n) * IN0C;
        const __bf16 *rows = px + 2 * h * IN0C;          // chunks 0..4: columns 2h, 2h + 1
        const __bf16 *col4 = px + 4 * IN0C + 2 * h * RS;  // chunk 5: column 4, rows 2h, 2h + 1
        // two 8-byte loads a chunk: a pixel's 4 channels are 8-byte aligned only
#pragma unroll
        for (int c = 0; c < 5; c++) {
            const bf16x4 u0 = *reinterpret_cast<const bf16x4 *>(rows + c * RS);
            const bf16x4 u1 = *reinterpret_cast<const bf16x4 *>(rows + c * RS + IN0C);
            b[c] = __builtin_shufflevector(u0, u1, 0, 1, 2, 3, 4, 5, 6, 7);
        }
        const bf16x4 r0 = *reinterpret_cast<const bf16x4 *>(col4);
        const bf16x4 r1 = *reinterpret_cast<const bf16x4 *>(col4 + RS);
        b[5] = __builtin_shufflevector(r0, r1, 0, 1, 2, 3, 4, 5, 6, 7);
        // tap (4, 4) in both halves; half 1 zeroes it where the MFMA reads it,
        // so that no VALU work waits for this load here
        const bf16x4 z = {};
        const bf16x4 t44 = *reinterpret_cast<const bf16x4 *>(px + 4 * RS + 4 * IN0C);
        b[6] = __builtin_shufflevector(t44, z, 0, 1, 2, 3, 4, 5, 6, 7);
    };
    int k = wib;
    if (k >= bd.count) return;
    Pos pos{(bd.first + k) / T, (bd.first + k) - (bd.first + k) / T * T};
    // two fragment sets: the next item's loads go to the other set while
    // this item runs its MFMAs and epilogue (trips alternate P, no copies)
    bf16x8 b[2][IN_CHUNKS];
    load(pos, b[0]);
    auto item = [&](auto parity) __attribute__((always_inline)) {
        constexpr int P = decltype(parity)::value;
        const Pos cur = pos;
        // the last trip loads its own item again (unused): every path has the
        // same loads outstanding, so the waits before the MFMAs name exactly
        // this item's fragments
        if (k + nwb < bd.count) pos = advance(pos);  // wavefront-uniform
        load(pos, b[P ^ 1]);
        f32x16 acc = {};
#pragma unroll
        for (int c = 0; c < IN_CHUNKS - 1; c++)
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(wf[c], b[P][c], acc, 0, 0, 0);
        const bf16x8 z8 = {};
        acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(wf[IN_CHUNKS - 1],
                                                      h ? z8 : b[P][IN_CHUNKS - 1], acc, 0, 0,
                                                      0);
        const int y = cur.y, t = cur.t, x0 = min(32 * t, nx - 32);
        const int x = x0 + n;
        if (x >= 32 * t) {
            const size_t px = pad_index(y, x, nx) * 16;
#pragma unroll
            for (int g = 0; g < 4; g++) {
                const int co = g * 8 + h * 4;
                const size_t o = (size_t)(co >> 4) * PL + px + (co & 15);
                const f32x4 bias = *reinterpret_cast<const f32x4 *>(&bsh[co]);
                bf16x4 p, r;
#pragma unroll
                for (int j = 0; j < 4; j++) {
                    const float v = acc[4 * g + j] + bias[j];
                    p[j] = (__bf16)v;
                    r[j] = (__bf16)gelu(v);
                }
                if (a.pre) *reinterpret_cast<bf16x4 *>(a.pre + o) = p;
                *reinterpret_cast<bf16x4 *>(a.out + o) = r;
            }
        }
        k += nwb;
    };
    while (true) {
        item(std::integral_constant<int, 0>{});
        if (k >= bd.count) break;
        item(std::integral_constant<int, 1>{});
        if (k >= bd.count) break;
    }
}

// A wet T cell whose ssh the step does not clamp to the tide (bc_ssh sets
// every wet cell with an open (-1) 4-neighbour).
__device__ __forceinline__ bool free_ssh_cell(const int *tm, size_t k, int W)
{
    return tm[k] > 0 && tm[k - 1] >= 0 && tm[k + 1] >= 0 && tm[k - W] >= 0 && tm[k + W] >= 0;
}

// A cell whose ssh the closure corrects and learns: every wet cell
// (ssh_wet), or the free cells only.
__device__ __forceinline__ bool ssh_cell(const int *tm, size_t k, int W, int ssh_wet)
{
    return ssh_wet ? tm[k] > 0 : free_ssh_cell(tm, k, W);
}

// The coarse state at one interior point that the energy guard reads: ssh,
// u, v, depth and the T mask of the point and its 4 neighbours.
struct GuardPoint {
    double s, u, v, h;
    int tc, te, tw, tn, ts;
};

__device__ __forceinline__ GuardPoint guard_point(const CoarseState &c, size_t k)
{
    return GuardPoint{c.ssh[k],      c.u[k],         c.v[k],
                      c.ht[k],       c.tmask[k],     c.tmask[k + 1],
                      c.tmask[k - 1], c.tmask[k + c.W], c.tmask[k - c.W]};
}

// Energy-guard terms of interior point (I, J) whose network output is q:
// E before and after the correction that k_set_fields applies there, and a
// count of the non-finite or out-of-bound applied components.
__device__ __forceinline__ void guard_terms(const float4 &q, const GuardSums &gs,
                                            const GuardPoint &p, int I, int J, int nx, int ny,
                                            double &e0, double &e1, double &bad)
{
    const Norm &nm = gs.nm;
    const double g = gs.g;
    const bool wet = p.tc > 0;
    // ssh_cell() on the loaded mask values
    const bool cs = nm.ssh_wet ? wet : wet && p.tw >= 0 && p.te >= 0 && p.ts >= 0 && p.tn >= 0;
    const bool cu = I < nx && wet && p.te > 0;
    const bool cv = J < ny && wet && p.tn > 0;
    const double s = p.s, u = p.u, v = p.v, h = p.h;
    const double s1 = cs ? s + (double)(q.z * nm.out_scale[2]) : s;
    const double u1 = cu ? u + (double)(q.x * nm.out_scale[0]) : u;
    const double v1 = cv ? v + (double)(q.y * nm.out_scale[1]) : v;
    if (nm.max_out > 0.0f && ((cs && fabsf(q.z) > nm.max_out) ||
                              (cu && fabsf(q.x) > nm.max_out) ||
                              (cv && fabsf(q.y) > nm.max_out)))
        bad += 1.0;
    e0 += g * s * s + h * (u * u + v * v);
    const double t = g * s1 * s1 + h * (u1 * u1 + v1 * v1);
    if (isfinite(t)) e1 += t;
    else bad += 1.0;
}

// Output layer (32 -> 3 used channels): the 16 rows of v_mfma_f32_16x16x32_bf16
// hold (dx, co) = dx * 3 + co, so one instruction covers a whole kernel row
// of 16 input pixels and 32 channels; the shifted sum over dx goes through a
// per-wavefront LDS tile. 16 input pixels give 12 outputs. One wavefront
// owns one column tile over OUT_RB output rows: it loads each of the
// OUT_RB + 4 input rows once and keeps them in registers for the five kernel
// rows that read them, and it keeps the five weight fragments in registers,
// so no LDS fill and no workgroup barrier. One item per wavefront (the grid
// covers every item: out_grid()). GUARD: the workgroup also sums the energy
// guard's terms over the points it writes (guard_terms, in a fixed order:
// lanes by a shuffle tree, then the 4 wavefronts) into gs.partial[block].
constexpr int OUT_RB = 8;

template <bool GUARD>
__global__ __launch_bounds__(256) void k_conv_out(ConvArgs a, GuardSums gs)
{
    __shared__ float dsh[4][4][16][17];  // [wave][row][dx * 3 + co][pixel]
    __shared__ double red[3][4];
    const int lane = (int)(threadIdx.x & 63), wv = (int)(threadIdx.x >> 6);
    const int n = lane & 15, q4 = lane >> 4, kg = q4 * 8;
    const int nx = a.nx, ny = a.ny, Wp = nx + 4, T = (nx + 11) / 12;
    const int nbands = (ny + OUT_RB - 1) / OUT_RB;
    const Band bd = xcd_band(nbands * T);
    const int k = bd.local * 4 + wv;
    double e0 = 0.0, e1 = 0.0, bad = 0.0;  // GUARD: this lane's terms
    if (k < bd.count) {  // wavefront-uniform
    const int it = bd.first + k, band = it / T, t = it - band * T;
    const int y0 = band * OUT_RB, nr = min(OUT_RB, ny - y0), x0 = min(12 * t, nx - 12);
    const size_t PL = plane_elems(nx, ny), RS = (size_t)Wp * 16;

    // A[m = 3 dx + co][k = ci] of kernel row dy; row 15 is zero
    bf16x8 w[5];
#pragma unroll
    for (int dy = 0; dy < 5; dy++) {
        const int dx = n / 3, co = n - 3 * (n / 3);
        w[dy] = n < 15 ? *reinterpret_cast<const bf16x8 *>(
                             a.w + ((size_t)(dy * 5 + dx) * COUT4 + co) * 32 + kg)
                       : bf16x8{};
    }
    // padded input rows y0 .. y0 + nr + 3: 16 pixels, channels kg .. kg + 7
    const __bf16 *src = a.in + (size_t)(kg >> 4) * PL + (kg & 15) + ((size_t)y0 * Wp + x0 + n) * 16;
    bf16x8 rows[OUT_RB + 4];
#pragma unroll
    for (int r = 0; r < OUT_RB + 4; r++)
        rows[r] = r < nr + 4 ? *reinterpret_cast<const bf16x8 *>(src + r * RS) : bf16x8{};
    // GUARD: the state at this lane's output points (row 4 g4 + q4 of the
    // band), loaded with the input rows so that their latency hides behind
    // the MFMAs; lanes without a point read interior point (1, 1)
    GuardPoint gp[OUT_RB / 4];
    if constexpr (GUARD) {
#pragma unroll
        for (int g4 = 0; g4 < OUT_RB / 4; g4++) {
            const int yy = 4 * g4 + q4, x = x0 + n;
            const bool mine = yy < nr && n < 12 && x >= 12 * t;
            gp[g4] = guard_point(gs.c, mine ? (size_t)(y0 + yy + 1) * gs.c.W + (x + 1)
                                            : (size_t)gs.c.W + 1);
        }
    }

    float(*d)[16][17] = dsh[wv];
    const int nc = min(n, 11);
#pragma unroll
    for (int g4 = 0; g4 < OUT_RB / 4; g4++) {
        f32x4 acc[4] = {};
#pragma unroll
        for (int dy = 0; dy < 5; dy++)
#pragma unroll
            for (int j = 0; j < 4; j++)
                acc[j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(w[dy], rows[4 * g4 + j + dy],
                                                                 acc[j], 0, 0, 0);
#pragma unroll
        for (int j = 0; j < 4; j++)
#pragma unroll
            for (int r = 0; r < 4; r++) d[j][q4 * 4 + r][n] = acc[j][r];
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        // lane: output row 4 g4 + q4 of the band, pixel n of the tile
        float v[3];
#pragma unroll
        for (int co = 0; co < 3; co++) {
            float s = 0.0f;
#pragma unroll
            for (int dx = 0; dx < 5; dx++) s += d[q4][dx * 3 + co][nc + dx];
            v[co] = s + a.bias[co];
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        const int yy = 4 * g4 + q4, x = x0 + n;
        if (yy < nr && n < 12 && x >= 12 * t) {
            const size_t p = (size_t)(y0 + yy) * nx + x;
            const float gt = a.gate ? a.gate[p] : 1.0f;
            float4 r;
            r.x = v[0] * gt;
            r.y = v[1] * gt;
            r.z = v[2] * gt;
            r.w = 0.0f;
            reinterpret_cast<float4 *>(a.outf)[p] = r;
            if constexpr (GUARD)
                guard_terms(r, gs, gp[g4], x + 1, y0 + yy + 1, nx, ny, e0, e1, bad);
        }
    }
    }
    if constexpr (GUARD) {
        for (int off = 32; off > 0; off >>= 1) {
            e0 += __shfl_xor(e0, off);
            e1 += __shfl_xor(e1, off);
            bad += __shfl_xor(bad, off);
        }
        if (lane == 0) {
            red[0][wv] = e0;
            red[1][wv] = e1;
            red[2][wv] = bad;
        }
        __syncthreads();
        if (threadIdx.x < 3) {
            const int r = (int)threadIdx.x;
            gs.partial[blockIdx.x * 3 + r] = (red[r][0] + red[r][1]) + (red[r][2] + red[r][3]);
        }
    }
}

// Workgroups of k_conv_out: 4 wavefronts each, every item covered once.
dim3 out_grid(int nx, int ny)
{
    const long items = (long)((nx + 11) / 12) * ((ny + OUT_RB - 1) / OUT_RB);
    const long per_xcd = (items + 7) / 8;
    return dim3((unsigned)(8 * ((per_xcd + 3) / 4)));
}

// Weight gradient, dW[tap][co][ci] = sum over pixels of dPre[px][co] *
// X[px + tap][ci]: a GEMM whose K runs over pixels, on
// v_mfma_f32_16x16x32_bf16 (16 co x 16 ci x 32 pixels). Its operands want 8
// consecutive pixels of one channel per lane, while the activations store
// channels fastest, so a workgroup stages them in LDS transposed ([channel]
// [pixel]) as it marches down a 32-pixel column strip: per output row one
// dPre row and one new input row, the input rows in a ring of the 5 kernel
// rows. Each input row is staged at 4 pixel shifts so that the fragment of
// every horizontal tap is an aligned read (dx 0..3: 16 bytes from shift dx;
// dx 4: two 8-byte reads from shift 0). Wave w accumulates the taps w, w+4,
// ...; the global loads of the row DW_PF rows ahead are in flight during
// the current row's MFMAs. Every workgroup writes its partial dW; k_reduce adds the partials
// in a fixed order. CI == CIN0 is layer 0: its input is the IN0C-channel
// network input, and staged channels IN0C..15 stay zero.
constexpr int DW_LEN = 40;  // staged pixels per row and shift (36 used)
constexpr int DW_PF = 4;    // rows whose global loads are in flight

template <int CO, int CI>
__global__ __launch_bounds__(256) void k_dw(const __bf16 *__restrict__ dpre,
                                            const __bf16 *__restrict__ in,
                                            float *__restrict__ partial, int nx, int ny,
                                            int rows_per_block)
{
    constexpr int MB = CO / 16, NB = CI / 16, MAXT = 7;
    constexpr bool NET_IN = CI == CIN0;
    constexpr int XQ = NET_IN ? 36 : 36 * (CI / 8);  // 16-byte (8-byte) chunks of an input row
    constexpr int DQ = 32 * (CO / 8);                // 16-byte chunks of a dPre row
    static_assert(XQ <= 256 && DQ <= 256, "one chunk per thread");
    __shared__ __attribute__((aligned(16))) __bf16 xs[5][4][CI][DW_LEN];
    __shared__ __attribute__((aligned(16))) __bf16 dsh[CO][32];
    const int tid = (int)threadIdx.x, lane = tid & 63;
    const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int n16 = lane & 15, kg = (lane >> 4) * 8;
    const int Wp = nx + 4, nstrip = nx >> 5;
    const size_t PL = plane_elems(nx, ny);
    const int strip = (int)blockIdx.x % nstrip, chunk = (int)blockIdx.x / nstrip;
    const int x0 = strip * 32, ya = chunk * rows_per_block;
    const int yb = min(ya + rows_per_block, ny);

    if constexpr (NET_IN) {
        // channels IN0C..15 are never staged: zero them once
        bf16x8 *z = reinterpret_cast<bf16x8 *>(&xs[0][0][0][0]);
        for (int q = tid; q < (int)(sizeof(xs) / 16); q += 256) z[q] = bf16x8{};
        __syncthreads();
    }

    auto load_x = [&](int r, bf16x8 &xr) __attribute__((always_inline)) {  // padded input row r
        if (tid < XQ) {
            if constexpr (NET_IN) {
                const bf16x4 v =
                    *reinterpret_cast<const bf16x4 *>(in + ((size_t)r * Wp + x0 + tid) * IN0C);
                xr = __builtin_shufflevector(v, bf16x4{}, 0, 1, 2, 3, 4, 5, 6, 7);
            } else {
                const int p = tid / (CI / 8), c8 = tid - p * (CI / 8);
                xr = *reinterpret_cast<const bf16x8 *>(
                    in + (size_t)(c8 >> 1) * PL + ((size_t)r * Wp + x0 + p) * 16 + (c8 & 1) * 8);
            }
        }
    };
    auto store_x = [&](int r, const bf16x8 &xr) __attribute__((always_inline)) {
        if (tid < XQ) {
            const int slot = r % 5;
            const int p = NET_IN ? tid : tid / (CI / 8);
            const int cb = NET_IN ? 0 : (tid - p * (CI / 8)) * 8;
            constexpr int NCH = NET_IN ? IN0C : 8;
#pragma unroll
            for (int s = 0; s < 4; s++) {
                const int j = p - s;
                if (j >= 0) {
#pragma unroll
                    for (int k = 0; k < NCH; k++) xs[slot][s][cb + k][j] = xr[k];
                }
            }
        }
    };
    auto load_d = [&](int y, bf16x8 &dr) __attribute__((always_inline)) {  // interior row y
        if (tid < DQ) {
            const int p = tid / (CO / 8), c8 = tid - p * (CO / 8);
            dr = *reinterpret_cast<const bf16x8 *>(
                dpre + (size_t)(c8 >> 1) * PL + pad_index(y, x0 + p, nx) * 16 + (c8 & 1) * 8);
        }
    };
    auto store_d = [&](const bf16x8 &dr) __attribute__((always_inline)) {
        if (tid < DQ) {
            const int p = tid / (CO / 8), cb = (tid - p * (CO / 8)) * 8;
#pragma unroll
            for (int k = 0; k < 8; k++) dsh[cb + k][p] = dr[k];
        }
    };

    f32x4 acc[MAXT][MB][NB];
#pragma unroll
    for (int t = 0; t < MAXT; t++)
#pragma unroll
        for (int mb = 0; mb < MB; mb++)
#pragma unroll
            for (int nb = 0; nb < NB; nb++) acc[t][mb][nb] = f32x4{0.f, 0.f, 0.f, 0.f};

    // the ring holds padded input rows y .. y+4 while output row y is reduced;
    // registers xr[u] / dr[u] hold the rows that arrive DW_PF rows later
    for (int r = ya; r < ya + 4; r++) {
        bf16x8 xr = {};
        load_x(r, xr);
        store_x(r, xr);
    }
    bf16x8 xr[DW_PF], dr[DW_PF];
#pragma unroll
    for (int u = 0; u < DW_PF; u++) {
        xr[u] = bf16x8{};
        dr[u] = bf16x8{};
        if (ya + u < yb) {
            load_x(ya + 4 + u, xr[u]);
            load_d(ya + u, dr[u]);
        }
    }
    for (int y0 = ya; y0 < yb; y0 += DW_PF) {
#pragma unroll
        for (int u = 0; u < DW_PF; u++) {
            const int y = y0 + u;
            if (y < yb) {  // workgroup-uniform
                store_x(y + 4, xr[u]);
                store_d(dr[u]);
                __syncthreads();
                if (y + DW_PF < yb) {
                    load_x(y + 4 + DW_PF, xr[u]);
                    load_d(y + DW_PF, dr[u]);
                }
                bf16x8 A[MB];
#pragma unroll
                for (int mb = 0; mb < MB; mb++)
                    A[mb] = *reinterpret_cast<const bf16x8 *>(&dsh[mb * 16 + n16][kg]);
#pragma unroll
                for (int t = 0; t < MAXT; t++) {
                    const int tap = wv + 4 * t;
                    if (tap < TAPS) {  // wavefront-uniform
                        const int dy = tap / 5, dx = tap - 5 * (tap / 5);
                        const int slot = (y + dy) % 5;
#pragma unroll
                        for (int nb = 0; nb < NB; nb++) {
                            bf16x8 B;
                            if (dx < 4) {
                                B = *reinterpret_cast<const bf16x8 *>(
                                    &xs[slot][dx][nb * 16 + n16][kg]);
                            } else {
                                const bf16x4 lo = *reinterpret_cast<const bf16x4 *>(
                                    &xs[slot][0][nb * 16 + n16][4 + kg]);
                                const bf16x4 hi = *reinterpret_cast<const bf16x4 *>(
                                    &xs[slot][0][nb * 16 + n16][8 + kg]);
                                B = __builtin_shufflevector(lo, hi, 0, 1, 2, 3, 4, 5, 6, 7);
                            }
#pragma unroll
                            for (int mb = 0; mb < MB; mb++)
                                acc[t][mb][nb] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                                    A[mb], B, acc[t][mb][nb], 0, 0, 0);
                        }
                    }
                }
                __syncthreads();
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
                        out[((size_t)tap * CO + co) * CI + nb * 16 + n16] = acc[t][mb][nb][r];
                    }
        }
    }
}

// Partial bias gradient: per channel, the sum over this workgroup's pixels of
// dPre. A thread reads one 16-byte chunk (8 channels of one pixel) per
// iteration, CO / 8 chunks a pixel, so a wavefront's loads cover contiguous
// pixels of each plane; the sums of the threads that share a chunk are added
// in LDS in a fixed order.
template <int CO>
__global__ __launch_bounds__(256) void k_bias_partial(const __bf16 *__restrict__ dpre,
                                                      float *__restrict__ partial, int nx,
                                                      int ny)
{
    constexpr int NCH = CO / 8;     // chunks per pixel
    constexpr int PPI = 256 / NCH;  // pixels per workgroup and iteration
    __shared__ float red[256][9];
    const int ch = (int)threadIdx.x % NCH, q = (int)threadIdx.x / NCH;
    const size_t coff = (size_t)(ch >> 1) * plane_elems(nx, ny) + (size_t)(ch & 1) * 8;
    const int npx = nx * ny;
    float s[8] = {};
#pragma unroll 4
    for (int p = (int)blockIdx.x * PPI + q; p < npx; p += (int)gridDim.x * PPI) {
        const int y = p / nx, x = p - y * nx;
        const bf16x8 v = *reinterpret_cast<const bf16x8 *>(dpre + coff + pad_index(y, x, nx) * 16);
#pragma unroll
        for (int k = 0; k < 8; k++) s[k] += (float)v[k];
    }
#pragma unroll
    for (int k = 0; k < 8; k++) red[threadIdx.x][k] = s[k];
    __syncthreads();
    if ((int)threadIdx.x < CO) {
        const int c = (int)threadIdx.x, chunk = c >> 3, k = c & 7;
        float t = 0.0f;
        for (int qq = 0; qq < PPI; qq++) t += red[qq * NCH + chunk][k];
        partial[(size_t)blockIdx.x * CO + c] = t;
    }
}

// out[i] = sum over the nblk partials: a workgroup owns 64 consecutive
// outputs, its wavefront w sums the w-th quarter of the partials (64-lane
// coalesced rows, 8 loads in flight), and the 4 quarter sums are added in
// order, so the result does not depend on timing.
__global__ __launch_bounds__(256) void k_reduce(const float *__restrict__ partial, int nblk,
                                                int n, float *__restrict__ out)
{
    __shared__ float red[4][64];
    const int lane = (int)threadIdx.x & 63, w = (int)threadIdx.x >> 6;
    const int i = (int)blockIdx.x * 64 + lane;
    const int q = (nblk + 3) / 4, b0 = w * q, b1 = min(b0 + q, nblk);
    float s = 0.0f;
    if (i < n) {
#pragma unroll 8
        for (int b = b0; b < b1; b++) s += partial[(size_t)b * n + i];
    }
    red[w][lane] = s;
    __syncthreads();
    if (w == 0 && i < n) out[i] = ((red[0][lane] + red[1][lane]) + red[2][lane]) + red[3][lane];
}

// Global gradient norm, and the factor that scales it down to max_norm
// (1 when max_norm <= 0 or the norm is below it): norm_scale = {norm, factor}.
// One workgroup of 1024 threads, 8 loads in flight per thread, then a tree
// in LDS: a fixed order, and a few memory latencies for the ~77K values.
__global__ __launch_bounds__(1024) void k_grad_norm(const float *__restrict__ g, int n,
                                                    float max_norm,
                                                    float *__restrict__ norm_scale)
{
    __shared__ float red[1024];
    float s = 0.0f;
#pragma unroll 8
    for (int i = (int)threadIdx.x; i < n; i += 1024) s += g[i] * g[i];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int k = 512; k > 0; k >>= 1) {
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
    bf16x4 lo;
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
    static_assert(IN0C == 4, "one 8-byte pixel");
    *reinterpret_cast<bf16x4 *>(in0 + pad_index(y, x, nx) * IN0C) = lo;
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


__global__ __launch_bounds__(256) void k_loss(const float *__restrict__ pred,
                                              const float *__restrict__ target,
                                              __bf16 *__restrict__ dpre4,
                                              float *__restrict__ partial, int nx, int ny,
                                              const int *__restrict__ tmask, int W,
                                              int ssh_wet,
                                              const float *__restrict__ gate, LossWeights lw)
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
            !tmask || ssh_cell(tmask, (size_t)(y + 1) * W + (x + 1), W, ssh_wet);
        const float d0 = a.x - b.x, d1 = a.y - b.y, d2 = ssh_free ? a.z - b.z : 0.0f;
        s0 += d0 * d0;
        s1 += d1 * d1;
        s2 += d2 * d2;
        // the output is gated (pred = gate * raw output): d loss / d raw
        const float gs = gate ? scale * gate[p] : scale;
        bf16x8 lo = {}, hi = {};
        lo[0] = (__bf16)(gs * lw.w[0] * d0);
        lo[1] = (__bf16)(gs * lw.w[1] * d1);
        lo[2] = (__bf16)(gs * lw.w[2] * d2);
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
    if (mode == 1 && ssh_cell(c.tmask, k, c.W, nm.ssh_wet))
        c.ssh[k] += (double)(q.z * nm.out_scale[2]);
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

// Adds the partial sums in workgroup order (the same sum every call) and
// decides. A separate launch rather than a last-workgroup epilogue: the
// device-scope fence that epilogue needs writes back each XCD's L2.
// One workgroup of DECIDE_THREADS threads; each thread adds the partials
// b = tid, tid + DECIDE_THREADS, ... with DECIDE_UNROLL of them loaded
// together (one memory latency for a few thousand partials), then a tree in
// LDS.
constexpr int DECIDE_THREADS = 1024, DECIDE_UNROLL = 4;

__global__ __launch_bounds__(DECIDE_THREADS) void k_guard_decide(
    const double *__restrict__ partial, int nblk, double rel, double abs_tol,
    int *__restrict__ flag, int *__restrict__ trips)
{
    __shared__ double red[3][DECIDE_THREADS];
    double a[3] = {0.0, 0.0, 0.0};
    for (int b0 = (int)threadIdx.x; b0 < nblk; b0 += DECIDE_UNROLL * DECIDE_THREADS) {
        double v[DECIDE_UNROLL][3];
#pragma unroll
        for (int u = 0; u < DECIDE_UNROLL; u++) {
            const int b = b0 + u * DECIDE_THREADS;
#pragma unroll
            for (int r = 0; r < 3; r++) v[u][r] = b < nblk ? partial[b * 3 + r] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < DECIDE_UNROLL; u++)
#pragma unroll
            for (int r = 0; r < 3; r++) a[r] += v[u][r];
    }
    for (int r = 0; r < 3; r++) red[r][threadIdx.x] = a[r];
    __syncthreads();
    for (int k = DECIDE_THREADS / 2; k > 0; k >>= 1) {
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

int cu_count()
{
    static int n = [] {
        int dev = 0, cus = 0;
        NEMO_HIP_CHECK(hipGetDevice(&dev));
        NEMO_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
        return cus;
    }();
    return n;
}

// Persistent grid of `per_cu` 4-wavefront workgroups per CU, rounded up to a
// multiple of 8 for the per-XCD bands.
dim3 conv_grid(int per_cu) { return dim3((unsigned)((cu_count() * per_cu + 7) / 8 * 8)); }

}  // namespace

namespace {
// one tile per wavefront: 61.4 against 62.3 us per 32 -> 32 layer at 1024^2
// on MI355X
int g_conv_form = 1;
}

void set_conv_form(int form)
{
    if (form != 0 && form != 1)
        NEMO_FATAL("closure conv: form %d (0: two tiles per wavefront, 1: one)", form);
    g_conv_form = form;
}

int conv_form() { return g_conv_form; }

void launch_conv(hipStream_t s, int cin, int cout, Epilogue epi, const ConvArgs &a)
{
    if (a.nx < 32 || a.ny <= 0)
        NEMO_FATAL("closure conv: interior %d x %d (nx must be at least 32)", a.nx, a.ny);
    // k_conv_sh: persistent, 2 workgroups per CU (LDS weights; 4 wavefronts
    // of 256 registers or 8 of 128); k_conv_in: persistent, 4 workgroups
    // per CU (128 registers); k_conv_out: one item per wavefront (out_grid)
    const dim3 grid = conv_grid(2), block(256), block8(512);
    const bool one = g_conv_form == 1;
    if (cin == 16 && cout == 32 && epi == EPI_HIDDEN)  // the network input: 4 used channels
        k_conv_in<<<conv_grid(4), block, 0, s>>>(a);
    else if (cin == 32 && cout == 32 && epi == EPI_HIDDEN && one)
        k_conv_sh<2, EPI_HIDDEN, 1, 8><<<grid, block8, 0, s>>>(a);
    else if (cin == 32 && cout == 32 && epi == EPI_HIDDEN)
        k_conv_sh<2, EPI_HIDDEN, 2, 4><<<grid, block, 0, s>>>(a);
    else if (cin == 32 && cout == 16 && epi == EPI_OUTPUT)
        k_conv_out<false><<<out_grid(a.nx, a.ny), block, 0, s>>>(a, GuardSums{});
    else if (cin == 16 && cout == 32 && epi == EPI_DX && one)
        k_conv_sh<1, EPI_DX, 1, 8><<<grid, block8, 0, s>>>(a);
    else if (cin == 16 && cout == 32 && epi == EPI_DX)
        k_conv_sh<1, EPI_DX, 2, 4><<<grid, block, 0, s>>>(a);
    else if (cin == 32 && cout == 32 && epi == EPI_DX && one)
        k_conv_sh<2, EPI_DX, 1, 8><<<grid, block8, 0, s>>>(a);
    else if (cin == 32 && cout == 32 && epi == EPI_DX)
        k_conv_sh<2, EPI_DX, 2, 4><<<grid, block, 0, s>>>(a);
    else
        NEMO_FATAL("closure conv: unsupported %d -> %d channels, epilogue %d", cin, cout,
                   (int)epi);
    NEMO_LAUNCH_CHECK();
}

// dW workgroups: nx / 32 column strips x row chunks, about 512 in all, two
// per CU (at least 8 rows a chunk, so the 4-row ring fill stays a small
// overhead)
static int dw_rows_per_block(int nx, int ny)
{
    const long nstrip = nx / 32;
    return (int)std::max(8L, (ny * nstrip + 511) / 512);
}

int dw_blocks(int nx, int ny)
{
    const int rpb = dw_rows_per_block(nx, ny);
    return (nx / 32) * ((ny + rpb - 1) / rpb);
}

void launch_dw(hipStream_t s, int co, int ci, const __bf16 *dpre, const __bf16 *in,
               float *partial, int nx, int ny)
{
    if (nx % 32 != 0 || nx <= 0 || ny <= 0)
        NEMO_FATAL("closure dW: interior %d x %d (nx a positive multiple of 32)", nx, ny);
    const int rpb = dw_rows_per_block(nx, ny);
    const dim3 grid((unsigned)dw_blocks(nx, ny)), block(256);
    if (co == 32 && ci == 32)
        k_dw<32, 32><<<grid, block, 0, s>>>(dpre, in, partial, nx, ny, rpb);
    else if (co == 32 && ci == 16)
        k_dw<32, 16><<<grid, block, 0, s>>>(dpre, in, partial, nx, ny, rpb);
    else if (co == 16 && ci == 32)
        k_dw<16, 32><<<grid, block, 0, s>>>(dpre, in, partial, nx, ny, rpb);
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
    k_reduce<<<dim3((unsigned)((n + 63) / 64)), dim3(256), 0, s>>>(partial, nblk, n, out);
    NEMO_LAUNCH_CHECK();
}

void launch_optimizer(hipStream_t s, int kind, float *w, const float *g, float *m,
                      float *v, int n, float lr, float b1, float b2, float eps,
                      float bc1, float bc2, float max_norm, float *norm_scale)
{
    k_grad_norm<<<dim3(1), dim3(1024), 0, s>>>(g, n, max_norm, norm_scale);
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
                 float *partial, int nx, int ny, const int *tmask, int W, int ssh_wet,
                 const float *gate, const LossWeights &lw)
{
    k_loss<<<dim3((unsigned)loss_blocks()), dim3(256), 0, s>>>(pred, target, dpre4,
                                                               partial, nx, ny, tmask, W,
                                                               ssh_wet, gate, lw);
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

// one partial per workgroup of the output layer
int guard_blocks(int nx, int ny) { return (int)out_grid(nx, ny).x; }

void launch_conv_out_guarded(hipStream_t s, const ConvArgs &a, const GuardSums &gs)
{
    if (a.nx < 32 || a.ny <= 0)
        NEMO_FATAL("closure conv: interior %d x %d (nx must be at least 32)", a.nx, a.ny);
    k_conv_out<true><<<out_grid(a.nx, a.ny), dim3(256), 0, s>>>(a, gs);
    NEMO_LAUNCH_CHECK();
}

void launch_guard_decide(hipStream_t s, const double *partial, int nblk, double rel,
                         double abs_tol, int *flag, int *trips)
{
    k_guard_decide<<<dim3(1), dim3(DECIDE_THREADS), 0, s>>>(partial, nblk, rel, abs_tol, flag,
                                                            trips);
    NEMO_LAUNCH_CHECK();
}

}  // namespace cl
}  // namespace nemo
