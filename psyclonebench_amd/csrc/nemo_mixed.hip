// Mixed-precision NEMOLite2D step (CDNA4 / gfx950): FP32 storage and
// arithmetic for velocities, face ssh, depths and metrics; the T-cell ssh is
// a compensated pair (hi, lo) of FP32 words whose update is accumulated in
// FP64, and the tidal elevation is evaluated in FP64 on the host. Bytes per
// interior update: 21 x 4 (fields, metrics) + 1 (the T mask, one byte a
// point) + 5 x 4 (writes) + 4 + 4 (the lo word read and written) = 113,
// against 212 in FP64.
//
// The arithmetic is nemo_point.h's *_val functions instantiated for float,
// except continuity, whose flux divergence is formed in FP32 (the same
// expression tree as continuity_kern.c:110-116) and then added to the
// (hi, lo) pair in FP64. The kernel walks rows like k_fused_wave (61 output
// columns per 64-lane wavefront, neighbours by cross-lane permute).
#include "nemo_hip_common.h"
#include "nemo_mixed.h"
#include "nemo_wave.h"

namespace nemo {
namespace {

template <int R>
__global__ __launch_bounds__(64) void k_mixed_wave(MixedIn f, StepScalars<float> p,
                                                   MixedOut o, int W, Bounds b,
                                                   int nstrip,
                                                   const double *__restrict__ tide_tab,
                                                   int tide_k, double tide_val)
{
    const int logical = wave::xcd_tile((int)blockIdx.x, (int)gridDim.x);
    const int band = logical / nstrip, strip = logical - band * nstrip;
    wave::tile_mixed<R>(f, p, o, W, 0u, b, strip, band,
                        wave::split(tide_tab ? tide_tab[tide_k] : tide_val));
}

// The same with OW owned columns per wavefront: 48 (192-byte strips, every
// strip's stores start on a 64-byte boundary) or 56 (224-byte strips,
// 32-byte boundaries, fewer neighbour lanes).
template <int R, int OW>
__global__ __launch_bounds__(64) void k_mixed_wave_a(MixedIn f, StepScalars<float> p,
                                                     MixedOut o, int W, Bounds b,
                                                     int nstrip,
                                                     const double *__restrict__ tide_tab,
                                                     int tide_k, double tide_val)
{
    const int logical = wave::xcd_tile((int)blockIdx.x, (int)gridDim.x);
    const int band = logical / nstrip, strip = logical - band * nstrip;
    wave::tile_mixed<R, OW>(f, p, o, W, 0u, b, strip, band,
                            wave::split(tide_tab ? tide_tab[tide_k] : tide_val));
}

// The OW-column march with the loads issued PF + 1 rows ahead (tile_mixed's
// PF = 1 or 2; OW = 61, 56 or 48).
template <int R, int OW, int PF = 1>
__global__ __launch_bounds__(64) void k_mixed_wave_p(MixedIn f, StepScalars<float> p,
                                                     MixedOut o, int W, Bounds b,
                                                     int nstrip,
                                                     const double *__restrict__ tide_tab,
                                                     int tide_k, double tide_val)
{
    const int logical = wave::xcd_tile((int)blockIdx.x, (int)gridDim.x);
    const int band = logical / nstrip, strip = logical - band * nstrip;
    wave::tile_mixed<R, OW, PF>(f, p, o, W, 0u, b, strip, band,
                               wave::split(tide_tab ? tide_tab[tide_k] : tide_val));
}

// Flather over the same segment lists as the FP64 path, FP32 values.
__global__ __launch_bounds__(64) void k_flather_f32(FlatherList lu, FlatherList lv,
                                                    float *ua, float *va,
                                                    const float *ua_prev,
                                                    const float *va_prev,
                                                    const float *__restrict__ su,
                                                    const float *__restrict__ sv)
{
    int t = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const FlatherList *L = &lu;
    float *vel = ua;
    const float *vel_prev = ua_prev;
    const float *ssh = su;
    if (t >= lu.nseg) {
        t -= lu.nseg;
        if (t >= lv.nseg) return;
        L = &lv;
        vel = va;
        vel_prev = va_prev;
        ssh = sv;
    }
    const int k1 = L->seg_ptr[t + 1];
    for (int k = L->seg_ptr[t]; k < k1; k++) {
        const int c = L->idx[k], off = L->off[k];
        const float v_nb = off ? (L->prev[k] ? vel_prev[c + off] : vel[c + off]) : 0.0f;
        const float s_nb = off ? ssh[c + off] : 0.0f;
        vel[c] = v_nb + (float)L->coef[k] * (ssh[c] - s_nb);
    }
}

__global__ void k_to_f32(const double *__restrict__ src, float *__restrict__ dst,
                         long n)
{
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k < n) dst[k] = (float)src[k];
}

__global__ void k_split_f64(const double *__restrict__ src, float *__restrict__ hi,
                            float *__restrict__ lo, long n)
{
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const double s = src[k];
    const float h = (float)s;
    hi[k] = h;
    lo[k] = (float)(s - (double)h);
}

__global__ void k_to_f64(const float *__restrict__ src, const float *__restrict__ lo,
                         double *__restrict__ dst, long n)
{
    const long k = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    dst[k] = lo ? (double)src[k] + (double)lo[k] : (double)src[k];
}

inline dim3 blocks_for(long n) { return dim3((unsigned)((n + 255) / 256)); }

}  // namespace

void launch_mixed_step(hipStream_t s, int rows, int cols, int prefetch, const MixedIn &f,
                       const StepScalars<float> &p, const MixedOut &o, int W,
                       Bounds b, const double *tide_tab, int tide_k, double tide_val)
{
    if (cols != wave::OUT && cols != 56 && cols != 48)
        NEMO_FATAL("mixed step: unsupported columns per wavefront %d", cols);
    const int ncol = b.xe - b.xs + 3, nrow = b.ye - b.ys + 3;
    const int nstrip = (ncol + cols - 1) / cols;
    if (prefetch) {
        switch (rows * 4 + prefetch) {
#define NEMO_MIXED_PF_CASE(RR, PF)                                                   \
    case RR * 4 + PF: {                                                              \
        const dim3 grid((unsigned)(nstrip * ((nrow + RR - 1) / RR)));                \
        if (cols == 48)                                                              \
            k_mixed_wave_p<RR, 48, PF><<<grid, dim3(64), 0, s>>>(                    \
                f, p, o, W, b, nstrip, tide_tab, tide_k, tide_val);                  \
        else if (cols == 56)                                                         \
            k_mixed_wave_p<RR, 56, PF><<<grid, dim3(64), 0, s>>>(                    \
                f, p, o, W, b, nstrip, tide_tab, tide_k, tide_val);                  \
        else                                                                         \
            k_mixed_wave_p<RR, wave::OUT, PF><<<grid, dim3(64), 0, s>>>(             \
                f, p, o, W, b, nstrip, tide_tab, tide_k, tide_val);                  \
        break;                                                                       \
    }
            NEMO_MIXED_PF_CASE(4, 1)
            NEMO_MIXED_PF_CASE(8, 1)
            NEMO_MIXED_PF_CASE(4, 2)
            NEMO_MIXED_PF_CASE(8, 2)
#undef NEMO_MIXED_PF_CASE
        default:
            NEMO_FATAL("mixed step: the prefetching kernel takes 4 or 8 rows and depth 1 or 2, "
                       "not %d rows, depth %d", rows, prefetch);
        }
        NEMO_LAUNCH_CHECK();
        return;
    }
    switch (rows) {
#define NEMO_MIXED_CASE(RR)                                                          \
    case RR: {                                                                       \
        const dim3 grid((unsigned)(nstrip * ((nrow + RR - 1) / RR)));                \
        if (cols == 48)                                                              \
            k_mixed_wave_a<RR, 48><<<grid, dim3(64), 0, s>>>(                        \
                f, p, o, W, b, nstrip, tide_tab, tide_k, tide_val);                  \
        else if (cols == 56)                                                         \
            k_mixed_wave_a<RR, 56><<<grid, dim3(64), 0, s>>>(                        \
                f, p, o, W, b, nstrip, tide_tab, tide_k, tide_val);                  \
        else                                                                         \
            k_mixed_wave<RR><<<grid, dim3(64), 0, s>>>(f, p, o, W, b, nstrip,        \
                                                       tide_tab, tide_k, tide_val);   \
        break;                                                                       \
    }
        NEMO_MIXED_CASE(4)
        NEMO_MIXED_CASE(8)
        NEMO_MIXED_CASE(16)
        NEMO_MIXED_CASE(32)
        NEMO_MIXED_CASE(48)
        NEMO_MIXED_CASE(64)
        NEMO_MIXED_CASE(80)
        NEMO_MIXED_CASE(96)
        NEMO_MIXED_CASE(128)
#undef NEMO_MIXED_CASE
    default: NEMO_FATAL("mixed step: unsupported rows per band %d", rows);
    }
    NEMO_LAUNCH_CHECK();
}

void launch_flather_f32(hipStream_t s, const FlatherList &lu, const FlatherList &lv,
                        float *ua, float *va, const float *ua_prev, const float *va_prev,
                        const float *su, const float *sv)
{
    const int n = lu.nseg + lv.nseg;
    if (n == 0) return;
    k_flather_f32<<<dim3((unsigned)((n + 63) / 64)), dim3(64), 0, s>>>(
        lu, lv, ua, va, ua_prev, va_prev, su, sv);
    NEMO_LAUNCH_CHECK();
}

void launch_to_f32(hipStream_t s, const double *src, float *dst, long n)
{
    k_to_f32<<<blocks_for(n), dim3(256), 0, s>>>(src, dst, n);
    NEMO_LAUNCH_CHECK();
}

void launch_split_f64(hipStream_t s, const double *src, float *hi, float *lo, long n)
{
    k_split_f64<<<blocks_for(n), dim3(256), 0, s>>>(src, hi, lo, n);
    NEMO_LAUNCH_CHECK();
}

void launch_to_f64(hipStream_t s, const float *src, const float *lo, double *dst, long n)
{
    k_to_f64<<<blocks_for(n), dim3(256), 0, s>>>(src, lo, dst, n);
    NEMO_LAUNCH_CHECK();
}

}  // namespace nemo
