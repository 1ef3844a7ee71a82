// Per-point NEMOLite2D physics for the HIP kernels, templated on the storage
// and arithmetic type T (double for the reference-exact path, float for the
// mixed-precision path).
//
// Each *_val function evaluates its expression tree in the same order as the
// reference kernel it cites, so that with FMA contraction off the double
// instantiation rounds exactly like the reference C code. They take their
// stencil as values (a *Vals struct) so that a kernel may bring the values
// from wherever it keeps them (registers of a row march, global memory);
// the *_pt wrappers gather the stencil from global memory.
// The only restructuring is arithmetic-neutral:
//   * the Coriolis factor 2*omega*sin(gphi*d2r) comes in precomputed per
//     point (fcor_u / fcor_v, evaluated on the host with the same libm call
//     and the same left-to-right order as momentum_u_kern.c:251);
//   * 1 + cbfr*rdt comes in as one scalar (same value at every point);
//   * the functions return the new value instead of storing it.
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

// Load / store at element offset `off` from a kernel-argument base pointer,
// with the byte offset formed in 32 bits: the compiler then addresses with
// the uniform base in SGPRs plus a 32-bit VGPR offset instead of a 64-bit
// add per access. Every allocation here is far below 4 GiB.
template <typename T>
__device__ __forceinline__ T ld32(const T *__restrict__ base, unsigned off)
{
    return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) +
                                        (unsigned)(off * (unsigned)sizeof(T)));
}
template <typename T>
__device__ __forceinline__ void st32(T *__restrict__ base, unsigned off, T v)
{
    *reinterpret_cast<T *>(reinterpret_cast<char *>(base) +
                           (unsigned)(off * (unsigned)sizeof(T))) = v;
}

// Division used by the physics. FP64 keeps the IEEE quotient (the reference
// rounding); FP32 (the mixed-precision form, which has an accuracy target
// rather than a bitwise one) uses the hardware reciprocal, 2 ulp.
template <typename T>
__device__ __forceinline__ T divide(T a, T b)
{
    return a / b;
}
template <>
__device__ __forceinline__ float divide<float>(float a, float b)
{
    return a * __builtin_amdgcn_rcpf(b);
}

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

// ---- continuity ------------------------------------------------------------
// Stencil of T(i,j): its U faces i-1, i and V faces j-1, j.
template <typename T>
struct ContVals {
    T sshn;                 // sshn_t(i,j)
    T su_e, hu_e, un_e;     // U(i,j)
    T su_w, hu_w, un_w;     // U(i-1,j)
    T sv_n, hv_n, vn_n;     // V(i,j)
    T sv_s, hv_s, vn_s;     // V(i,j-1)
    T area;                 // area_t(i,j)
};

// continuity_kern.c:110-116: new ssh from the flux divergence.
template <typename T>
__device__ __forceinline__ T continuity_val(const ContVals<T> &v, T rdt)
{
    const T flux_e = (v.su_e + v.hu_e) * v.un_e;
    const T flux_w = (v.su_w + v.hu_w) * v.un_w;
    const T flux_n = (v.sv_n + v.hv_n) * v.vn_n;
    const T flux_s = (v.sv_s + v.hv_s) * v.vn_s;
    return v.sshn + divide((flux_w - flux_e + flux_s - flux_n) * rdt, v.area);
}

template <typename T>
__device__ __forceinline__ T continuity_pt(const StepIn<T> &f, int c, int W,
                                           T rdt)
{
    const int w = c - 1, s = c - W;
    ContVals<T> v;
    v.sshn = f.sshn_t[c];
    v.su_e = f.sshn_u[c];
    v.hu_e = f.hu[c];
    v.un_e = f.un[c];
    v.su_w = f.sshn_u[w];
    v.hu_w = f.hu[w];
    v.un_w = f.un[w];
    v.sv_n = f.sshn_v[c];
    v.hv_n = f.hv[c];
    v.vn_n = f.vn[c];
    v.sv_s = f.sshn_v[s];
    v.hv_s = f.hv[s];
    v.vn_s = f.vn[s];
    v.area = f.area_t[c];
    return continuity_val(v, rdt);
}

// boundary_conditions_kern.c:189-202: a wet cell with an open (-1)
// 4-neighbour is clamped to the tidal elevation.
__device__ __forceinline__ bool tide_applies(int m_c, int m_s, int m_n, int m_e,
                                         int m_w)
{
    return m_c > 0 && (m_s < 0 || m_n < 0 || m_e < 0 || m_w < 0);
}

__device__ __forceinline__ bool tide_cell(const int *__restrict__ tm, int c,
                                          int W)
{
    return tide_applies(tm[c], tm[c - W], tm[c + W], tm[c + 1], tm[c - 1]);
}

// ---- momentum_u ------------------------------------------------------------
// Stencil of U(i,j). Suffixes: c = (i,j), e = (i+1,j), w = (i-1,j),
// s = (i,j-1), n = (i,j+1), se = (i+1,j-1), ne = (i+1,j+1).
template <typename T>
struct MomUVals {
    int m_c, m_e, m_s, m_se, m_n, m_ne;  // tmask of T points
    T un_c, un_e, un_w, un_s, un_n;      // U points
    T vn_s, vn_se, vn_c, vn_e;           // V points (i,j-1) (i+1,j-1) (i,j) (i+1,j)
    T e1v_s, e1v_se, e1v_c, e1v_e;       // dx_v at the same V points
    T hv_s, sv_s, hv_se, sv_se;          // hv, sshn_v at V(i,j-1), V(i+1,j-1)
    T hv_c, sv_c, hv_e, sv_e;            // ... at V(i,j), V(i+1,j)
    T e2t_c, e2t_e, e1t_c, e1t_e;        // dy_t, dx_t at T(i,j), T(i+1,j)
    T ht_c, ht_e, st_c, st_e;            // ht, sshn_t at T(i,j), T(i+1,j)
    T hu_c, su_c, hu_s, su_s, hu_n, su_n;  // hu, sshn_u at U(i,j), U(i,j-1), U(i,j+1)
    T e2u_c, e2u_s, e2u_n;               // dy_u at the same U points
    T e1u_c;                             // dx_u(i,j)
    T fcor, area, ssha;                  // fcor_u, area_u, ssha_u at U(i,j)
};

// momentum_u_kern.c:181-261. Returns false (and leaves *out alone) for
// faces the reference skips.
template <typename T>
__device__ __forceinline__ bool momentum_u_val(const MomUVals<T> &v,
                                               const StepScalars<T> &p, T *out)
{
    if (v.m_c + v.m_e <= 0) return false;
    if (v.m_c <= 0 || v.m_e <= 0) return false;

    const T u_e = T(0.5) * (v.un_c + v.un_e) * v.e2t_e;
    const T dep_e = v.ht_e + v.st_e;
    const T u_w = T(0.5) * (v.un_c + v.un_w) * v.e2t_c;
    const T dep_w = v.ht_c + v.st_c;

    const T vmean_s = T(0.5) * (v.vn_s + v.vn_se);
    const T v_s = T(0.5) * vmean_s * (v.e1v_s + v.e1v_se);
    const T dep_s = T(0.5) * (v.hv_s + v.sv_s + v.hv_se + v.sv_se);

    const T vmean_n = T(0.5) * (v.vn_c + v.vn_e);
    const T v_n = T(0.5) * vmean_n * (v.e1v_c + v.e1v_e);
    const T dep_n = T(0.5) * (v.hv_c + v.sv_c + v.hv_e + v.sv_e);

    const bool shut_s = v.m_s <= 0 || v.m_se <= 0;
    const bool shut_n = v.m_n <= 0 || v.m_ne <= 0;

    // first-order upwind advection
    const T up_w = upwind_lo(u_w) * v.un_c + upwind_hi(u_w) * v.un_w;
    const T up_e = upwind_hi(u_e) * v.un_c + upwind_lo(u_e) * v.un_e;
    T up_s, up_n, shear_s, shear_n;
    if (shut_s) {
        up_s = upwind_lo(v_s) * v.un_c;
        shear_s = T(0);
    } else {
        up_s = upwind_lo(v_s) * v.un_c + upwind_hi(v_s) * v.un_s;
        shear_s = divide(v.un_c - v.un_s, v.e2u_c + v.e2u_s) *
                  (v.hu_c + v.su_c + v.hu_s + v.su_s);
    }
    if (shut_n) {
        up_n = upwind_hi(v_n) * v.un_c;
        shear_n = T(0);
    } else {
        up_n = upwind_hi(v_n) * v.un_c + upwind_lo(v_n) * v.un_n;
        shear_n = divide(v.un_n - v.un_c, v.e2u_c + v.e2u_n) *
                  (v.hu_c + v.su_c + v.hu_n + v.su_n);
    }
    const T adv = up_w * u_w * dep_w - up_e * u_e * dep_e + up_s * v_s * dep_s -
                  up_n * v_n * dep_n;

    // viscosity (slip walls zero the cross-face shear above)
    const T strain_e = divide(v.un_e - v.un_c, v.e1t_e) * (v.ht_e + v.st_e);
    const T strain_w = divide(v.un_c - v.un_w, v.e1t_c) * (v.ht_c + v.st_c);
    T vis = (strain_e - strain_w) * v.e2u_c +
            (shear_n - shear_s) * v.e1u_c * T(0.5);
    vis = p.visc * vis;

    const T cor = T(0.5) * (v.fcor * (vmean_s + vmean_n)) * v.area *
                  (v.hu_c + v.su_c);
    const T hpg = -p.g * (v.hu_c + v.su_c) * v.e2u_c * (v.st_e - v.st_c);

    *out = divide(divide(v.un_c * (v.hu_c + v.su_c) +
                             divide(p.rdt * (adv + vis + cor + hpg), v.area),
                         v.hu_c + v.ssha),
                  p.fric);
    return true;
}

template <typename T>
__device__ __forceinline__ bool momentum_u_pt(const StepIn<T> &f,
                                              const StepScalars<T> &p, int c,
                                              int W, T *out)
{
    const int *__restrict__ tm = f.tmask;
    const int e = c + 1, w = c - 1, s = c - W, n = c + W;
    const int se = s + 1, ne = n + 1;
    MomUVals<T> v;
    v.m_c = tm[c];
    v.m_e = tm[e];
    if (v.m_c + v.m_e <= 0 || v.m_c <= 0 || v.m_e <= 0) return false;
    v.m_s = tm[s];
    v.m_se = tm[se];
    v.m_n = tm[n];
    v.m_ne = tm[ne];
    v.un_c = f.un[c];
    v.un_e = f.un[e];
    v.un_w = f.un[w];
    v.un_s = f.un[s];
    v.un_n = f.un[n];
    v.vn_s = f.vn[s];
    v.vn_se = f.vn[se];
    v.vn_c = f.vn[c];
    v.vn_e = f.vn[e];
    v.e1v_s = f.dx_v[s];
    v.e1v_se = f.dx_v[se];
    v.e1v_c = f.dx_v[c];
    v.e1v_e = f.dx_v[e];
    v.hv_s = f.hv[s];
    v.sv_s = f.sshn_v[s];
    v.hv_se = f.hv[se];
    v.sv_se = f.sshn_v[se];
    v.hv_c = f.hv[c];
    v.sv_c = f.sshn_v[c];
    v.hv_e = f.hv[e];
    v.sv_e = f.sshn_v[e];
    v.e2t_c = f.dy_t[c];
    v.e2t_e = f.dy_t[e];
    v.e1t_c = f.dx_t[c];
    v.e1t_e = f.dx_t[e];
    v.ht_c = f.ht[c];
    v.ht_e = f.ht[e];
    v.st_c = f.sshn_t[c];
    v.st_e = f.sshn_t[e];
    v.hu_c = f.hu[c];
    v.su_c = f.sshn_u[c];
    v.hu_s = f.hu[s];
    v.su_s = f.sshn_u[s];
    v.hu_n = f.hu[n];
    v.su_n = f.sshn_u[n];
    v.e2u_c = f.dy_u[c];
    v.e2u_s = f.dy_u[s];
    v.e2u_n = f.dy_u[n];
    v.e1u_c = f.dx_u[c];
    v.fcor = f.fcor_u[c];
    v.area = f.area_u[c];
    v.ssha = f.ssha_u[c];
    return momentum_u_val(v, p, out);
}

// ---- momentum_v ------------------------------------------------------------
// Stencil of V(i,j). Suffixes: c = (i,j), n = (i,j+1), s = (i,j-1),
// w = (i-1,j), e = (i+1,j), nw = (i-1,j+1), ne = (i+1,j+1).
template <typename T>
struct MomVVals {
    int m_c, m_e, m_n, m_w, m_nw, m_ne;  // tmask of T points
    T vn_c, vn_n, vn_s, vn_w, vn_e;      // V points
    T un_w, un_nw, un_c, un_n;           // U points (i-1,j) (i-1,j+1) (i,j) (i,j+1)
    T e2u_w, e2u_nw, e2u_c, e2u_n;       // dy_u at the same U points
    T hu_w, su_w, hu_nw, su_nw;          // hu, sshn_u at U(i-1,j), U(i-1,j+1)
    T hu_c, su_c, hu_n, su_n;            // ... at U(i,j), U(i,j+1)
    T e1t_c, e1t_n, e2t_c, e2t_n;        // dx_t, dy_t at T(i,j), T(i,j+1)
    T ht_c, ht_n, st_c, st_n;            // ht, sshn_t at T(i,j), T(i,j+1)
    T e1v_c, e1v_w, e1v_e;               // dx_v at V(i,j), V(i-1,j), V(i+1,j)
    T hv_c, sv_c, hv_w, sv_w, hv_e, sv_e;  // hv, sshn_v at the same V points
    T e2v_c;                             // dy_v(i,j)
    T fcor, area, ssha;                  // fcor_v, area_v, ssha_v at V(i,j)
};

// momentum_v_kern.c:155-240 (including its T(i)+T(i+1) first test, :155).
template <typename T>
__device__ __forceinline__ bool momentum_v_val(const MomVVals<T> &v,
                                               const StepScalars<T> &p, T *out)
{
    if (v.m_c + v.m_e <= 0) return false;
    if (v.m_c <= 0 || v.m_n <= 0) return false;

    const T v_n = T(0.5) * (v.vn_c + v.vn_n) * v.e1t_n;
    const T dep_n = v.ht_n + v.st_n;
    const T v_s = T(0.5) * (v.vn_c + v.vn_s) * v.e1t_c;
    const T dep_s = v.ht_c + v.st_c;

    const T umean_w = T(0.5) * (v.un_w + v.un_nw);
    const T u_w = T(0.5) * umean_w * (v.e2u_w + v.e2u_nw);
    const T dep_w = T(0.50) * (v.hu_w + v.su_w + v.hu_nw + v.su_nw);

    const T umean_e = T(0.5) * (v.un_c + v.un_n);
    const T u_e = T(0.5) * umean_e * (v.e2u_c + v.e2u_n);
    const T dep_e = T(0.50) * (v.hu_c + v.su_c + v.hu_n + v.su_n);

    const bool shut_w = v.m_w <= 0 || v.m_nw <= 0;
    const bool shut_e = v.m_e <= 0 || v.m_ne <= 0;

    const T up_s = upwind_lo(v_s) * v.vn_c + upwind_hi(v_s) * v.vn_s;
    const T up_n = upwind_hi(v_n) * v.vn_c + upwind_lo(v_n) * v.vn_n;
    T up_w, up_e, shear_w, shear_e;
    if (shut_w) {
        up_w = upwind_lo(u_w) * v.vn_c;
        shear_w = T(0);
    } else {
        up_w = upwind_lo(u_w) * v.vn_c + upwind_hi(u_w) * v.vn_w;
        shear_w = divide(v.vn_c - v.vn_w, v.e1v_c + v.e1v_w) *
                  (v.hv_c + v.sv_c + v.hv_w + v.sv_w);
    }
    if (shut_e) {
        up_e = upwind_hi(u_e) * v.vn_c;
        shear_e = T(0);
    } else {
        up_e = upwind_hi(u_e) * v.vn_c + upwind_lo(u_e) * v.vn_e;
        shear_e = divide(v.vn_e - v.vn_c, v.e1v_c + v.e1v_e) *
                  (v.hv_c + v.sv_c + v.hv_e + v.sv_e);
    }
    const T adv = up_w * u_w * dep_w - up_e * u_e * dep_e + up_s * v_s * dep_s -
                  up_n * v_n * dep_n;

    const T strain_n = divide(v.vn_n - v.vn_c, v.e2t_n) * (v.ht_n + v.st_n);
    const T strain_s = divide(v.vn_c - v.vn_s, v.e2t_c) * (v.ht_c + v.st_c);
    T vis = (strain_n - strain_s) * v.e1v_c +
            (shear_e - shear_w) * v.e2v_c * T(0.5);
    vis = p.visc * vis;

    const T cor = -T(0.5) * (v.fcor * (umean_e + umean_w)) * v.area *
                  (v.hv_c + v.sv_c);
    const T hpg = -p.g * (v.hv_c + v.sv_c) * v.e1v_c * (v.st_n - v.st_c);

    *out = divide(divide(v.vn_c * (v.hv_c + v.sv_c) +
                             divide(p.rdt * (adv + vis + cor + hpg), v.area),
                         v.hv_c + v.ssha),
                  p.fric);
    return true;
}

template <typename T>
__device__ __forceinline__ bool momentum_v_pt(const StepIn<T> &f,
                                              const StepScalars<T> &p, int c,
                                              int W, T *out)
{
    const int *__restrict__ tm = f.tmask;
    const int e = c + 1, w = c - 1, s = c - W, n = c + W;
    const int nw = n - 1, ne = n + 1;
    MomVVals<T> v;
    v.m_c = tm[c];
    v.m_e = tm[e];
    v.m_n = tm[n];
    if (v.m_c + v.m_e <= 0 || v.m_c <= 0 || v.m_n <= 0) return false;
    v.m_w = tm[w];
    v.m_nw = tm[nw];
    v.m_ne = tm[ne];
    v.vn_c = f.vn[c];
    v.vn_n = f.vn[n];
    v.vn_s = f.vn[s];
    v.vn_w = f.vn[w];
    v.vn_e = f.vn[e];
    v.un_w = f.un[w];
    v.un_nw = f.un[nw];
    v.un_c = f.un[c];
    v.un_n = f.un[n];
    v.e2u_w = f.dy_u[w];
    v.e2u_nw = f.dy_u[nw];
    v.e2u_c = f.dy_u[c];
    v.e2u_n = f.dy_u[n];
    v.hu_w = f.hu[w];
    v.su_w = f.sshn_u[w];
    v.hu_nw = f.hu[nw];
    v.su_nw = f.sshn_u[nw];
    v.hu_c = f.hu[c];
    v.su_c = f.sshn_u[c];
    v.hu_n = f.hu[n];
    v.su_n = f.sshn_u[n];
    v.e1t_c = f.dx_t[c];
    v.e1t_n = f.dx_t[n];
    v.e2t_c = f.dy_t[c];
    v.e2t_n = f.dy_t[n];
    v.ht_c = f.ht[c];
    v.ht_n = f.ht[n];
    v.st_c = f.sshn_t[c];
    v.st_n = f.sshn_t[n];
    v.e1v_c = f.dx_v[c];
    v.e1v_w = f.dx_v[w];
    v.e1v_e = f.dx_v[e];
    v.hv_c = f.hv[c];
    v.sv_c = f.sshn_v[c];
    v.hv_w = f.hv[w];
    v.sv_w = f.sshn_v[w];
    v.hv_e = f.hv[e];
    v.sv_e = f.sshn_v[e];
    v.e2v_c = f.dy_v[c];
    v.fcor = f.fcor_v[c];
    v.area = f.area_v[c];
    v.ssha = f.ssha_v[c];
    return momentum_v_val(v, p, out);
}

// time_update_kern.c:101-112 (U faces) and :143-153 (V faces): face ssh
// from the new T-cell ssh on both sides. m_c, m_p: tmask of the two cells;
// once their sum is positive and their product is not, exactly one of them
// is dry. False = no write.
template <typename T>
__device__ __forceinline__ bool next_ssh_pt(T ssh_c, T ssh_p, T area_c,
                                            T area_p, T area_face, int m_c,
                                            int m_p, T *out)
{
    if (m_c + m_p <= 0) return false;
    if (m_c * m_p > 0) {
        const T vol = area_c * ssh_c + area_p * ssh_p;
        *out = divide(T(0.5) * vol, area_face);
    } else if (m_c <= 0) {
        *out = ssh_p;
    } else {
        *out = ssh_c;
    }
    return true;
}

}  // namespace nemo
