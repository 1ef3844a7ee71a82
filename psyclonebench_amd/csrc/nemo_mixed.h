// Mixed-precision step kernels (nemo_mixed.hip): FP32 fields, compensated
// (hi, lo) FP32 ssh with FP64 accumulation.
#pragma once

#include "nemo_step.h"

namespace nemo {

struct MixedIn {
    const float *__restrict__ ssh_hi;  // sshn_t, high word
    const float *__restrict__ ssh_lo;  // sshn_t, low word
    const float *__restrict__ sshn_u;
    const float *__restrict__ sshn_v;
    const float *__restrict__ un;
    const float *__restrict__ vn;
    const float *__restrict__ ht;
    const float *__restrict__ hu;
    const float *__restrict__ hv;
    const float *__restrict__ ssha_u;
    const float *__restrict__ ssha_v;
    const float *__restrict__ area_t;
    const float *__restrict__ area_u;
    const float *__restrict__ area_v;
    const float *__restrict__ dx_u;
    const float *__restrict__ dx_v;
    const float *__restrict__ dx_t;
    const float *__restrict__ dy_u;
    const float *__restrict__ dy_v;
    const float *__restrict__ dy_t;
    const float *__restrict__ fcor_u;
    const float *__restrict__ fcor_v;
    const signed char *__restrict__ tmask;  // one byte a point (mask_bytes)
};

struct MixedOut {
    float *ssh_hi, *ssh_lo, *u, *v, *sshu, *sshv;
};

// rows per band: 4, 8, 16, 32, 48, 64, 80, 96 or 128; output columns per
// wavefront: 61 (overlapped strips), 56 (32-byte-aligned stores) or 48
// (64-byte-aligned stores)
// prefetch 1: the kernel that issues its loads two rows ahead (rows 4 or 8)
void launch_mixed_step(hipStream_t s, int rows, int cols, int prefetch, const MixedIn &f,
                       const StepScalars<float> &p, const MixedOut &o, int W,
                       Bounds b, const double *tide_tab, int tide_k, double tide_val);
// ua_prev / va_prev: the previous step's velocities (see FlatherList)
void launch_flather_f32(hipStream_t s, const FlatherList &lu, const FlatherList &lv,
                        float *ua, float *va, const float *ua_prev, const float *va_prev,
                        const float *su, const float *sv);
void launch_to_f32(hipStream_t s, const double *src, float *dst, long n);
void launch_split_f64(hipStream_t s, const double *src, float *hi, float *lo, long n);
void launch_to_f64(hipStream_t s, const float *src, const float *lo, double *dst, long n);

}  // namespace nemo
