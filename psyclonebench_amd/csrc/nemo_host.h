// Host-side helpers shared by the FP64 PSy layer (nemo_psy.cpp) and the
// mixed-precision and ensemble APIs: argument order of the reference C ABI,
// geometry checks, the library stream, device allocation and the Flather
// segment lists.
#pragma once

#include <cmath>
#include <vector>

#include "nemo_hip_common.h"
#include "nemo_step.h"

namespace nemo {

// Argument order of the reference C ABI (psykal_cpp/time_step.cpp:12-37).
enum Arr {
    A_SSHA_T, A_SSHN_T, A_SSHN_U, A_SSHN_V, A_HU, A_HV, A_UN, A_VN, A_UA,
    A_HT, A_SSHA_U, A_VA, A_SSHA_V, A_TMASK, A_AREA_T, A_AREA_U, A_AREA_V,
    A_DX_U, A_DX_V, A_DX_T, A_DY_U, A_DY_V, A_DY_T, A_GPHIU, A_GPHIV,
    A_COUNT
};

struct Geometry {
    int xs, xe, ys, ye, W, total;
};

struct Physics {
    double rdt, cbfr, visc, omega, d2r, g;
};

// The library's one stream (default priority, non-blocking w.r.t. stream 0).
hipStream_t lib_stream();

// bc_ssh elevation of step istep with the tide set by nemo_hip_set_tide
// (host libm, as boundary_conditions_kern.c:187-192).
double tide_at(int istep, double rdt);

// Work queued after every step of psy_steps_with_hook, on the step's stream,
// over the state the step has just written: ssh, u, v, face ssh (the current
// ping-pong buffers), and the constant ht, tmask and cell / face areas.
struct StepHook {
    void (*enqueue)(void *user, hipStream_t s, double *ssh, double *u, double *v,
                    double *sshu, double *sshv, const double *ht, const int *tmask,
                    const double *area_t, const double *area_u, const double *area_v);
    void *user;
    // identifies what enqueue queues (its buffers and the values it passes
    // to kernels); graphs captured with one key are replayed for that key
    unsigned long long key;
};

// nsteps fused steps istep, istep+1, ... of the device context whose ssha_t
// slot is `ssha_t_slot` (made by c_invoke_time_step in the fused form), each
// followed by hook.enqueue, as one hipGraph replay; the caller's slots are
// republished as c_invoke_time_steps does. Exits if there is no such fused
// context.
void psy_steps_with_hook(void **ssha_t_slot, int nsteps, int istep, const StepHook &hook);

template <typename T>
inline T *dev_alloc(size_t n)
{
    void *p = nullptr;
    NEMO_HIP_CHECK(hipMalloc(&p, n * sizeof(T)));
    return static_cast<T *>(p);
}

struct HostFlather {
    std::vector<int> seg_ptr{0}, idx, off, prev;
    std::vector<double> coef;
};

// Faces where bc_flather_u (stride 1) or bc_flather_v (stride W) writes,
// grouped per row (U) or per column (V) in the reference's sweep order.
// boundary_conditions_kern.c:282-291 / :322-331.
inline HostFlather build_flather(const int *tm, const double *h, const Geometry &G,
                          double g, bool along_i)
{
    HostFlather L;
    const int W = G.W;
    const int stride = along_i ? 1 : W;
    // U faces: ji in [xs-1, xe], jj in [ys-1, ye+1];
    // V faces: ji in [xs-1, xe+1], jj in [ys-1, ye]
    const int x0 = G.xs - 1, y0 = G.ys - 1;
    const int x1 = along_i ? G.xe : G.xe + 1;
    const int y1 = along_i ? G.ye + 1 : G.ye;
    const int outer0 = along_i ? y0 : x0, outer1 = along_i ? y1 : x1;
    const int inner0 = along_i ? x0 : y0, inner1 = along_i ? x1 : y1;
    // a face of the sweep's region that the sweep writes (and only it:
    // such a face has one open and one wet cell, so neither momentum nor
    // bc_solid writes it)
    auto is_flather = [&](int ji, int jj) {
        if (ji < x0 || ji > x1 || jj < y0 || jj > y1) return false;
        const long c = (long)jj * W + ji;
        const int mc = tm[c], mp = tm[c + stride];
        return mc + mp > -1 && (mc < 0 || mp < 0);
    };
    for (int o = outer0; o <= outer1; o++) {
        const size_t before = L.idx.size();
        for (int in = inner0; in <= inner1; in++) {
            const int ji = along_i ? in : o, jj = along_i ? o : in;
            const long c = (long)jj * W + ji;
            const int mc = tm[c], mp = tm[c + stride];
            if (mc + mp <= -1) continue;
            int d;
            if (mc < 0)
                d = stride;
            else if (mp < 0)
                d = -stride;
            else
                continue;
            const long q = c + d;
            L.idx.push_back((int)c);
            L.off.push_back(q >= 0 && q < G.total ? d : 0);
            L.prev.push_back(d > 0 && (along_i ? is_flather(ji + 1, jj)
                                               : is_flather(ji, jj + 1)));
            L.coef.push_back(std::sqrt(g / h[c]));
        }
        if (L.idx.size() > before) L.seg_ptr.push_back((int)L.idx.size());
    }
    return L;
}

struct DeviceFlather {
    int *seg_ptr = nullptr, *idx = nullptr, *off = nullptr, *prev = nullptr;
    double *coef = nullptr;
    int nseg = 0;

    void upload(const HostFlather &L, hipStream_t s)
    {
        release();
        nseg = (int)L.seg_ptr.size() - 1;
        if (nseg == 0) return;
        seg_ptr = dev_alloc<int>(L.seg_ptr.size());
        idx = dev_alloc<int>(L.idx.size());
        off = dev_alloc<int>(L.off.size());
        prev = dev_alloc<int>(L.prev.size());
        coef = dev_alloc<double>(L.coef.size());
        NEMO_HIP_CHECK(hipMemcpyAsync(seg_ptr, L.seg_ptr.data(),
                                      L.seg_ptr.size() * sizeof(int),
                                      hipMemcpyHostToDevice, s));
        NEMO_HIP_CHECK(hipMemcpyAsync(idx, L.idx.data(), L.idx.size() * sizeof(int),
                                      hipMemcpyHostToDevice, s));
        NEMO_HIP_CHECK(hipMemcpyAsync(off, L.off.data(), L.off.size() * sizeof(int),
                                      hipMemcpyHostToDevice, s));
        NEMO_HIP_CHECK(hipMemcpyAsync(prev, L.prev.data(), L.prev.size() * sizeof(int),
                                      hipMemcpyHostToDevice, s));
        NEMO_HIP_CHECK(hipMemcpyAsync(coef, L.coef.data(),
                                      L.coef.size() * sizeof(double),
                                      hipMemcpyHostToDevice, s));
        NEMO_HIP_CHECK(hipStreamSynchronize(s));
    }
    void release()
    {
        if (seg_ptr) NEMO_HIP_CHECK(hipFree(seg_ptr));
        if (idx) NEMO_HIP_CHECK(hipFree(idx));
        if (off) NEMO_HIP_CHECK(hipFree(off));
        if (prev) NEMO_HIP_CHECK(hipFree(prev));
        if (coef) NEMO_HIP_CHECK(hipFree(coef));
        seg_ptr = idx = off = prev = nullptr;
        coef = nullptr;
        nseg = 0;
    }
    FlatherList view() const { return FlatherList{seg_ptr, idx, off, prev, coef, nseg}; }
};

inline void check_geometry(const Geometry &G)
{
    if (G.W <= 0 || G.xs < 1 || G.ys < 1 || G.xe < G.xs || G.ye < G.ys)
        NEMO_FATAL("invalid internal bounds %d..%d x %d..%d (width %d)", G.xs,
                   G.xe, G.ys, G.ye, G.W);
    if (G.xe + 2 > G.W)
        NEMO_FATAL("width %d too small for internal xstop %d", G.W, G.xe);
    if ((long)G.W * (G.ye + 2) > (long)G.total)
        NEMO_FATAL("total_size %d too small for width %d and ystop %d", G.total,
                   G.W, G.ye);
}

// The fused and mixed kernels address every array with a 32-bit byte offset
// from a uniform base (ld32 / st32 in nemo_point.h). True when an array of
// G.total elements of elem_bytes each fits that addressing.
inline bool offsets_fit_32(const Geometry &G, size_t elem_bytes)
{
    return (unsigned long long)G.total * elem_bytes <= (1ull << 32);
}

// Coriolis factor (2*omega)*sin(gphi*d2r) per point, host libm, the same
// expression as momentum_u_kern.c:251 / momentum_v_kern.c:230.
inline std::vector<double> coriolis(const double *gphi, int n, double omega, double d2r)
{
    std::vector<double> f(n);
    for (int k = 0; k < n; k++) f[k] = 2. * omega * std::sin(gphi[k] * d2r);
    return f;
}

// The T mask as one signed byte a point, for the FP32 kernels (a quarter of
// the int32 mask's bytes). The step only compares mask values and adds two of
// them, so a byte copy gives the same results for every value in -64..63; the
// reference's masks hold -1, 0 and 1.
inline std::vector<signed char> mask_bytes(const int *tm, size_t n)
{
    std::vector<signed char> b(n);
    for (size_t k = 0; k < n; k++) {
        if (tm[k] < -64 || tm[k] > 63)
            NEMO_FATAL("T mask value %d at element %zu: the FP32 step keeps the mask "
                       "in one byte (-64..63)", tm[k], k);
        b[k] = (signed char)tm[k];
    }
    return b;
}

// The fused form copies nothing; it shares one buffer per after/now pair.
// That reproduces every value the reference computes as long as no later
// sweep reads a now-field outside the reference's copy region
// (time_step.cpp:115-122). The only such read is momentum_u's un(j+1) on
// the last T row, guarded by a positive tmask there.
inline bool fused_allowed(const int *tm, const Geometry &G)
{
    const int jl = G.ye + 1;
    for (int ji = 0; ji < G.W; ji++)
        if (tm[(long)jl * G.W + ji] > 0) return false;
    return true;
}

}  // namespace nemo
