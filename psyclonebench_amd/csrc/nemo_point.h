// Per-point NEMOLite2D physics for the HIP kernels, templated on the storage
// and arithmetic type T (double for the reference-exact path, float for the
// mixed-precision path).
//
// Each function evaluates its expression tree in the same order as the
// reference kernel it cites, so that with FMA contraction off the double
// instantiation rounds exactly like the reference C code. The only
// restructuring is arithmetic-neutral:
//   * the Coriolis factor 2*omega*sin(gphi*d2r) comes in precomputed per
//     point (fcor_u / fcor_v, evaluated on the host with the same libm call
//     and the same left-to-right order as momentum_u_kern.c:251);
//   * 1 + cbfr*rdt comes in as one scalar (same value at every point);
//   * the kernels return the new value instead of storing it, so that the
//     fused step can decide where it goes.
// Index convention: point (ji, jj) at c = jj*W + ji; U(i,j) is the east face
// and V(i,j) the north face of T(i,j).
#pragma once

#include <hip/hip_runtime.h>

namespace nemo {

// Old-state fields of one step (the *n arrays of the reference) and the
// constant inputs. All read-only inside a step.
template <typename T>
struct StepIn {
    const T *__restrict__ sshn_t;
    const T *__restrict__ sshn_u;
    const T *__restrict__ sshn_v;
    const T *__restrict__ un;
    const T *__restrict__ vn;
    const T *__restrict__ ht;
    const T *__restrict__ hu;
    const T *__restrict__ hv;
    const T *__restrict__ ssha_u;
    const T *__restrict__ ssha_v;
    const T *__restrict__ area_t;
    const T *__restrict__ area_u;
    const T *__restrict__ area_v;
    const T *__restrict__ dx_u;
    const T *__restrict__ dx_v;
    const T *__restrict__ dx_t;
    const T *__restrict__ dy_u;
    const T *__restrict__ dy_v;
    const T *__restrict__ dy_t;
    const T *__restrict__ fcor_u;
    const T *__restrict__ fcor_v;
    const int *__restrict__ tmask;
};

template <typename T>
struct StepScalars {
    T rdt;
    T visc;
    T g;
    T fric;  // 1 + cbfr*rdt
};

template <typename T>
__device__ __forceinline__ T upwind_hi(T x)
{
    return T(0.5) + copysign(T(0.5), x);
}
template <typename T>
__device__ __forceinline__ T upwind_lo(T x)
{
    return T(0.5) - copysign(T(0.5), x);
}

// continuity_kern.c:110-116: new ssh from the flux divergence.
template <typename T>
__device__ __forceinline__ T continuity_pt(const StepIn<T> &f, int c, int W,
                                           T rdt)
{
    const int w = c - 1, s = c - W;
    const T flux_e = (f.sshn_u[c] + f.hu[c]) * f.un[c];
    const T flux_w = (f.sshn_u[w] + f.hu[w]) * f.un[w];
    const T flux_n = (f.sshn_v[c] + f.hv[c]) * f.vn[c];
    const T flux_s = (f.sshn_v[s] + f.hv[s]) * f.vn[s];
    return f.sshn_t[c] + (flux_w - flux_e + flux_s - flux_n) * rdt / f.area_t[c];
}

// boundary_conditions_kern.c:189-202: a wet cell with an open (-1)
// 4-neighbour is clamped to the tidal elevation.
__device__ __forceinline__ bool tide_cell(const int *__restrict__ tm, int c,
                                          int W)
{
    return tm[c] > 0 &&
           (tm[c - W] < 0 || tm[c + W] < 0 || tm[c + 1] < 0 || tm[c - 1] < 0);
}

// momentum_u_kern.c:181-261. Returns false (and leaves *out alone) for
// faces the reference skips.
template <typename T>
__device__ __forceinline__ bool momentum_u_pt(const StepIn<T> &f,
                                              const StepScalars<T> &p, int c,
                                              int W, T *out)
{
    const int *__restrict__ tm = f.tmask;
    const int e = c + 1, w = c - 1, s = c - W, n = c + W;
    const int se = s + 1, ne = n + 1;
    const int m_c = tm[c], m_e = tm[e];
    if (m_c + m_e <= 0) return false;
    if (m_c <= 0 || m_e <= 0) return false;

    const T *__restrict__ un = f.un;
    const T *__restrict__ vn = f.vn;
    const T un_c = un[c], un_e = un[e], un_w = un[w];
    const T hu_c = f.hu[c], sshu_c = f.sshn_u[c];
    const T ht_e = f.ht[e], ht_c = f.ht[c];
    const T sshn_e = f.sshn_t[e], sshn_c = f.sshn_t[c];

    const T u_e = T(0.5) * (un_c + un_e) * f.dy_t[e];
    const T dep_e = ht_e + sshn_e;
    const T u_w = T(0.5) * (un_c + un_w) * f.dy_t[c];
    const T dep_w = ht_c + sshn_c;

    const T vmean_s = T(0.5) * (vn[s] + vn[se]);
    const T v_s = T(0.5) * vmean_s * (f.dx_v[s] + f.dx_v[se]);
    const T dep_s = T(0.5) * (f.hv[s] + f.sshn_v[s] + f.hv[se] + f.sshn_v[se]);

    const T vmean_n = T(0.5) * (vn[c] + vn[e]);
    const T v_n = T(0.5) * vmean_n * (f.dx_v[c] + f.dx_v[e]);
    const T dep_n = T(0.5) * (f.hv[c] + f.sshn_v[c] + f.hv[e] + f.sshn_v[e]);

    const bool shut_s = tm[s] <= 0 || tm[se] <= 0;
    const bool shut_n = tm[n] <= 0 || tm[ne] <= 0;

    // first-order upwind advection
    const T up_w = upwind_lo(u_w) * un_c + upwind_hi(u_w) * un_w;
    const T up_e = upwind_hi(u_e) * un_c + upwind_lo(u_e) * un_e;
    T up_s, up_n, shear_s, shear_n;
    if (shut_s) {
        up_s = upwind_lo(v_s) * un_c;
        shear_s = T(0);
    } else {
        const T un_s = un[s];
        up_s = upwind_lo(v_s) * un_c + upwind_hi(v_s) * un_s;
        shear_s = (un_c - un_s) / (f.dy_u[c] + f.dy_u[s]) *
                  (hu_c + sshu_c + f.hu[s] + f.sshn_u[s]);
    }
    if (shut_n) {
        up_n = upwind_hi(v_n) * un_c;
        shear_n = T(0);
    } else {
        const T un_n = un[n];
        up_n = upwind_hi(v_n) * un_c + upwind_lo(v_n) * un_n;
        shear_n = (un_n - un_c) / (f.dy_u[c] + f.dy_u[n]) *
                  (hu_c + sshu_c + f.hu[n] + f.sshn_u[n]);
    }
    const T adv = up_w * u_w * dep_w - up_e * u_e * dep_e + up_s * v_s * dep_s -
                  up_n * v_n * dep_n;

    // viscosity (slip walls zero the cross-face shear above)
    const T strain_e = (un_e - un_c) / f.dx_t[e] * (ht_e + sshn_e);
    const T strain_w = (un_c - un_w) / f.dx_t[c] * (ht_c + sshn_c);
    T vis = (strain_e - strain_w) * f.dy_u[c] +
            (shear_n - shear_s) * f.dx_u[c] * T(0.5);
    vis = p.visc * vis;

    const T cor = T(0.5) * (f.fcor_u[c] * (vmean_s + vmean_n)) * f.area_u[c] *
                  (hu_c + sshu_c);
    const T hpg = -p.g * (hu_c + sshu_c) * f.dy_u[c] * (sshn_e - sshn_c);

    *out = (un_c * (hu_c + sshu_c) + p.rdt * (adv + vis + cor + hpg) / f.area_u[c]) /
           (hu_c + f.ssha_u[c]) / p.fric;
    return true;
}

// momentum_v_kern.c:155-240 (including its T(i)+T(i+1) first test, :155).
template <typename T>
__device__ __forceinline__ bool momentum_v_pt(const StepIn<T> &f,
                                              const StepScalars<T> &p, int c,
                                              int W, T *out)
{
    const int *__restrict__ tm = f.tmask;
    const int e = c + 1, w = c - 1, s = c - W, n = c + W;
    const int nw = n - 1, ne = n + 1;
    const int m_c = tm[c];
    if (m_c + tm[e] <= 0) return false;
    if (m_c <= 0 || tm[n] <= 0) return false;

    const T *__restrict__ un = f.un;
    const T *__restrict__ vn = f.vn;
    const T vn_c = vn[c], vn_n = vn[n], vn_s = vn[s];
    const T hv_c = f.hv[c], sshv_c = f.sshn_v[c];
    const T ht_n = f.ht[n], ht_c = f.ht[c];
    const T sshn_n = f.sshn_t[n], sshn_c = f.sshn_t[c];

    const T v_n = T(0.5) * (vn_c + vn_n) * f.dx_t[n];
    const T dep_n = ht_n + sshn_n;
    const T v_s = T(0.5) * (vn_c + vn_s) * f.dx_t[c];
    const T dep_s = ht_c + sshn_c;

    const T umean_w = T(0.5) * (un[w] + un[nw]);
    const T u_w = T(0.5) * umean_w * (f.dy_u[w] + f.dy_u[nw]);
    const T dep_w = T(0.50) * (f.hu[w] + f.sshn_u[w] + f.hu[nw] + f.sshn_u[nw]);

    const T umean_e = T(0.5) * (un[c] + un[n]);
    const T u_e = T(0.5) * umean_e * (f.dy_u[c] + f.dy_u[n]);
    const T dep_e = T(0.50) * (f.hu[c] + f.sshn_u[c] + f.hu[n] + f.sshn_u[n]);

    const bool shut_w = tm[w] <= 0 || tm[nw] <= 0;
    const bool shut_e = tm[e] <= 0 || tm[ne] <= 0;

    const T up_s = upwind_lo(v_s) * vn_c + upwind_hi(v_s) * vn_s;
    const T up_n = upwind_hi(v_n) * vn_c + upwind_lo(v_n) * vn_n;
    T up_w, up_e, shear_w, shear_e;
    if (shut_w) {
        up_w = upwind_lo(u_w) * vn_c;
        shear_w = T(0);
    } else {
        const T vn_w = vn[w];
        up_w = upwind_lo(u_w) * vn_c + upwind_hi(u_w) * vn_w;
        shear_w = (vn_c - vn_w) / (f.dx_v[c] + f.dx_v[w]) *
                  (hv_c + sshv_c + f.hv[w] + f.sshn_v[w]);
    }
    if (shut_e) {
        up_e = upwind_hi(u_e) * vn_c;
        shear_e = T(0);
    } else {
        const T vn_e = vn[e];
        up_e = upwind_hi(u_e) * vn_c + upwind_lo(u_e) * vn_e;
        shear_e = (vn_e - vn_c) / (f.dx_v[c] + f.dx_v[e]) *
                  (hv_c + sshv_c + f.hv[e] + f.sshn_v[e]);
    }
    const T adv = up_w * u_w * dep_w - up_e * u_e * dep_e + up_s * v_s * dep_s -
                  up_n * v_n * dep_n;

    const T strain_n = (vn_n - vn_c) / f.dy_t[n] * (ht_n + sshn_n);
    const T strain_s = (vn_c - vn_s) / f.dy_t[c] * (ht_c + sshn_c);
    T vis = (strain_n - strain_s) * f.dx_v[c] +
            (shear_e - shear_w) * f.dy_v[c] * T(0.5);
    vis = p.visc * vis;

    const T cor = -T(0.5) * (f.fcor_v[c] * (umean_e + umean_w)) * f.area_v[c] *
                  (hv_c + sshv_c);
    const T hpg = -p.g * (hv_c + sshv_c) * f.dx_v[c] * (sshn_n - sshn_c);

    *out = (vn_c * (hv_c + sshv_c) + p.rdt * (adv + vis + cor + hpg) / f.area_v[c]) /
           (hv_c + f.ssha_v[c]) / p.fric;
    return true;
}

// time_update_kern.c:101-112 (stride 1, U faces) and :143-153 (stride W, V
// faces): face ssh from the new T-cell ssh `ssh`. False = no write.
// m_c, m_p: tmask of the two T cells; once their sum is positive and their
// product is not, exactly one of them is dry.
template <typename T>
__device__ __forceinline__ bool next_ssh_pt(T ssh_c, T ssh_p, T area_c,
                                            T area_p, T area_face, int m_c,
                                            int m_p, T *out)
{
    if (m_c + m_p <= 0) return false;
    if (m_c * m_p > 0) {
        const T vol = area_c * ssh_c + area_p * ssh_p;
        *out = T(0.5) * vol / area_face;
    } else if (m_c <= 0) {
        *out = ssh_p;
    } else {
        *out = ssh_c;
    }
    return true;
}

}  // namespace nemo
