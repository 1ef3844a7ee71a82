// Mixed-precision NEMOLite2D step (CDNA4 / gfx950): FP32 storage and
// arithmetic for velocities, face ssh, depths and metrics; the T-cell ssh is
// a compensated pair (hi, lo) of FP32 words whose update is accumulated in
// FP64, and the tidal elevation is evaluated in FP64 on the host. Bytes per
// interior update: 21 x 4 (fields, metrics) + 4 (tmask) + 5 x 4 (writes)
// + 4 + 4 (the lo word read and written) = 116, against 212 in FP64.
//
// The arithmetic is nemo_point.h's *_val functions instantiated for float,
// except continuity, whose flux divergence is formed in FP32 (the same
// expression tree as continuity_kern.c:110-116) and then added to the
// (hi, lo) pair in FP64. The kernel walks rows like k_fused_wave (61 output
// columns per 64-lane wavefront, neighbours by cross-lane permute).
#include "nemo_hip_common.h"
#include "nemo_mixed.h"

namespace nemo {
namespace {

__device__ __forceinline__ bool in_rect(int i, int j, int x0, int x1, int y0,
                                        int y1)
{
    return i >= x0 && i <= x1 && j >= y0 && j <= y1;
}

__device__ __forceinline__ float from_below(float x) { return __shfl_up(x, 1, 64); }
__device__ __forceinline__ float from_above(float x) { return __shfl_down(x, 1, 64); }
__device__ __forceinline__ int from_below(int x) { return __shfl_up(x, 1, 64); }
__device__ __forceinline__ int from_above(int x) { return __shfl_down(x, 1, 64); }

constexpr int OUT = 61;

// FP32 flux divergence of T(i,j) times rdt/area (continuity_kern.c:110-116
// without the "sshn +").
__device__ __forceinline__ float ssh_increment(const ContVals<float> &v, float rdt)
{
    const float flux_e = (v.su_e + v.hu_e) * v.un_e;
    const float flux_w = (v.su_w + v.hu_w) * v.un_w;
    const float flux_n = (v.sv_n + v.hv_n) * v.vn_n;
    const float flux_s = (v.sv_s + v.hv_s) * v.vn_s;
    return divide((flux_w - flux_e + flux_s - flux_n) * rdt, v.area);
}

struct Ssh {
    float hi, lo;
};

__device__ __forceinline__ Ssh split(double s)
{
    const float hi = (float)s;
    return Ssh{hi, (float)(s - (double)hi)};
}

template <int R>
__global__ __launch_bounds__(64) void k_mixed_wave(MixedIn f, StepScalars<float> p,
                                                   MixedOut o, int W, Bounds b,
                                                   int nstrip,
                                                   const double *__restrict__ tide_tab,
                                                   int tide_k, double tide_val)
{
    using T = float;
    const int nblk = (int)gridDim.x, bid = (int)blockIdx.x;
    const int xcd = bid & 7, k = bid >> 3;
    const int per = nblk >> 3, extra = nblk & 7;
    const int logical = xcd * per + min(xcd, extra) + k;
    const int band = logical / nstrip, strip = logical - band * nstrip;

    const int xmax = b.xe + 1, ymax = b.ye + 1;
    const int lane = (int)threadIdx.x;
    const int i = b.xs - 1 + strip * OUT + lane - 1;
    const bool owner = lane >= 1 && lane <= OUT && i <= xmax;
    const int ci = min(max(i, 0), xmax);
    const int jlo = b.ys - 1 + band * R;
    const int jhi = min(jlo + R, ymax + 1);
    auto rowoff = [&](int j) { return (unsigned)(min(max(j, 0), ymax) * W); };
    const Ssh tide = split(tide_tab ? tide_tab[tide_k] : tide_val);

    T un[3], vn[3], hu[3], su[3], hv[3], sv[3], e1v[3], e2u[3];
    T ht[2], st[2], e1t[2], e2t[2], at[2];
    T st_lo1;  // lo word of T(i, j+1)
    int tm[4];

    auto load3 = [&](int j, int s) {
        const unsigned c = rowoff(j) + ci;
        un[s] = ld32(f.un, c);
        vn[s] = ld32(f.vn, c);
        hu[s] = ld32(f.hu, c);
        su[s] = ld32(f.sshn_u, c);
        hv[s] = ld32(f.hv, c);
        sv[s] = ld32(f.sshn_v, c);
        e1v[s] = ld32(f.dx_v, c);
        e2u[s] = ld32(f.dy_u, c);
    };
    auto load2 = [&](int j, int s) {
        const unsigned c = rowoff(j) + ci;
        ht[s] = ld32(f.ht, c);
        st[s] = ld32(f.ssh_hi, c);
        e1t[s] = ld32(f.dx_t, c);
        e2t[s] = ld32(f.dy_t, c);
        at[s] = ld32(f.area_t, c);
    };
    // new (hi, lo) ssh of T(i, r) from 3-slot s / s-1, 2-slot s2, tm t-1..t+1
    auto new_ssh = [&](int s, int s2, int t, float lo) {
        ContVals<T> v;
        v.sshn = st[s2];
        v.su_e = su[s];
        v.hu_e = hu[s];
        v.un_e = un[s];
        v.su_w = from_below(su[s]);
        v.hu_w = from_below(hu[s]);
        v.un_w = from_below(un[s]);
        v.sv_n = sv[s];
        v.hv_n = hv[s];
        v.vn_n = vn[s];
        v.sv_s = sv[s - 1];
        v.hv_s = hv[s - 1];
        v.vn_s = vn[s - 1];
        v.area = at[s2];
        const int m_w = from_below(tm[t]), m_e = from_above(tm[t]);
        const float inc = ssh_increment(v, p.rdt);
        if (tide_applies(tm[t], tm[t - 1], tm[t + 1], m_e, m_w)) return tide;
        return split((double)st[s2] + (double)lo + (double)inc);
    };

    load3(jlo - 1, 0);
    load3(jlo, 1);
    load2(jlo, 0);
    tm[0] = ld32(f.tmask, rowoff(jlo - 1) + ci);
    tm[1] = ld32(f.tmask, rowoff(jlo) + ci);
    tm[2] = ld32(f.tmask, rowoff(jlo + 1) + ci);
    Ssh ssh_c = new_ssh(1, 0, 1, ld32(f.ssh_lo, rowoff(jlo) + ci));

    for (int j = jlo; j < jhi; j++) {
        load3(j + 1, 2);
        load2(j + 1, 1);
        st_lo1 = ld32(f.ssh_lo, rowoff(j + 1) + ci);
        tm[3] = ld32(f.tmask, rowoff(j + 2) + ci);
        const unsigned c = rowoff(j) + ci;
        const T dxu = ld32(f.dx_u, c), dyv = ld32(f.dy_v, c), fcu = ld32(f.fcor_u, c), fcv = ld32(f.fcor_v, c);
        const T aru = ld32(f.area_u, c), arv = ld32(f.area_v, c), sau = ld32(f.ssha_u, c), sav = ld32(f.ssha_v, c);

        const Ssh ssh_n = new_ssh(2, 1, 2, st_lo1);
        const float ssh_e = from_above(ssh_c.hi);

        if (owner && in_rect(i, j, b.xs, b.xe, b.ys, b.ye)) {
            st32(o.ssh_hi, c, ssh_c.hi);
            st32(o.ssh_lo, c, ssh_c.lo);
        }
        const int tm_e0 = from_above(tm[1]);
        const int tm_e1 = from_above(tm[2]);
        {
            MomUVals<T> v;
            v.m_c = tm[1];
            v.m_e = tm_e0;
            v.m_s = tm[0];
            v.m_se = from_above(tm[0]);
            v.m_n = tm[2];
            v.m_ne = tm_e1;
            v.un_c = un[1];
            v.un_e = from_above(un[1]);
            v.un_w = from_below(un[1]);
            v.un_s = un[0];
            v.un_n = un[2];
            v.vn_s = vn[0];
            v.vn_se = from_above(vn[0]);
            v.vn_c = vn[1];
            v.vn_e = from_above(vn[1]);
            v.e1v_s = e1v[0];
            v.e1v_se = from_above(e1v[0]);
            v.e1v_c = e1v[1];
            v.e1v_e = from_above(e1v[1]);
            v.hv_s = hv[0];
            v.sv_s = sv[0];
            v.hv_se = from_above(hv[0]);
            v.sv_se = from_above(sv[0]);
            v.hv_c = hv[1];
            v.sv_c = sv[1];
            v.hv_e = from_above(hv[1]);
            v.sv_e = from_above(sv[1]);
            v.e2t_c = e2t[0];
            v.e2t_e = from_above(e2t[0]);
            v.e1t_c = e1t[0];
            v.e1t_e = from_above(e1t[0]);
            v.ht_c = ht[0];
            v.ht_e = from_above(ht[0]);
            v.st_c = st[0];
            v.st_e = from_above(st[0]);
            v.hu_c = hu[1];
            v.su_c = su[1];
            v.hu_s = hu[0];
            v.su_s = su[0];
            v.hu_n = hu[2];
            v.su_n = su[2];
            v.e2u_c = e2u[1];
            v.e2u_s = e2u[0];
            v.e2u_n = e2u[2];
            v.e1u_c = dxu;
            v.fcor = fcu;
            v.area = aru;
            v.ssha = sau;
            T val = T(0);
            bool wrote = in_rect(i, j, b.xs, b.xe - 1, b.ys, b.ye) &&
                         momentum_u_val(v, p, &val);
            if (!wrote && in_rect(i, j, b.xs - 1, b.xe, b.ys - 1, b.ye + 1) &&
                tm[1] * tm_e0 == 0) {
                val = T(0);
                wrote = true;
            }
            if (owner && wrote) st32(o.u, c, val);
        }
        {
            MomVVals<T> v;
            v.m_c = tm[1];
            v.m_e = tm_e0;
            v.m_n = tm[2];
            v.m_w = from_below(tm[1]);
            v.m_nw = from_below(tm[2]);
            v.m_ne = tm_e1;
            v.vn_c = vn[1];
            v.vn_n = vn[2];
            v.vn_s = vn[0];
            v.vn_w = from_below(vn[1]);
            v.vn_e = from_above(vn[1]);
            v.un_w = from_below(un[1]);
            v.un_nw = from_below(un[2]);
            v.un_c = un[1];
            v.un_n = un[2];
            v.e2u_w = from_below(e2u[1]);
            v.e2u_nw = from_below(e2u[2]);
            v.e2u_c = e2u[1];
            v.e2u_n = e2u[2];
            v.hu_w = from_below(hu[1]);
            v.su_w = from_below(su[1]);
            v.hu_nw = from_below(hu[2]);
            v.su_nw = from_below(su[2]);
            v.hu_c = hu[1];
            v.su_c = su[1];
            v.hu_n = hu[2];
            v.su_n = su[2];
            v.e1t_c = e1t[0];
            v.e1t_n = e1t[1];
            v.e2t_c = e2t[0];
            v.e2t_n = e2t[1];
            v.ht_c = ht[0];
            v.ht_n = ht[1];
            v.st_c = st[0];
            v.st_n = st[1];
            v.e1v_c = e1v[1];
            v.e1v_w = from_below(e1v[1]);
            v.e1v_e = from_above(e1v[1]);
            v.hv_c = hv[1];
            v.sv_c = sv[1];
            v.hv_w = from_below(hv[1]);
            v.sv_w = from_below(sv[1]);
            v.hv_e = from_above(hv[1]);
            v.sv_e = from_above(sv[1]);
            v.e2v_c = dyv;
            v.fcor = fcv;
            v.area = arv;
            v.ssha = sav;
            T val = T(0);
            bool wrote = in_rect(i, j, b.xs, b.xe, b.ys, b.ye - 1) &&
                         momentum_v_val(v, p, &val);
            if (!wrote && in_rect(i, j, b.xs - 1, b.xe + 1, b.ys - 1, b.ye) &&
                tm[1] * tm[2] == 0) {
                val = T(0);
                wrote = true;
            }
            if (owner && wrote) st32(o.v, c, val);
        }
        {
            const T at_e0 = from_above(at[0]);
            T val;
            if (owner && in_rect(i, j, b.xs, b.xe - 1, b.ys, b.ye) &&
                next_ssh_pt(ssh_c.hi, ssh_e, at[0], at_e0, aru, tm[1], tm_e0, &val))
                st32(o.sshu, c, val);
            if (owner && in_rect(i, j, b.xs, b.xe, b.ys, b.ye - 1) &&
                next_ssh_pt(ssh_c.hi, ssh_n.hi, at[0], at[1], arv, tm[1], tm[2], &val))
                st32(o.sshv, c, val);
        }

#define NEMO_SHIFT3(a) \
    a[0] = a[1];       \
    a[1] = a[2];
#define NEMO_SHIFT2(a) a[0] = a[1];
        NEMO_SHIFT3(un) NEMO_SHIFT3(vn) NEMO_SHIFT3(hu) NEMO_SHIFT3(su)
        NEMO_SHIFT3(hv) NEMO_SHIFT3(sv) NEMO_SHIFT3(e1v) NEMO_SHIFT3(e2u)
        NEMO_SHIFT2(ht) NEMO_SHIFT2(st) NEMO_SHIFT2(e1t) NEMO_SHIFT2(e2t) NEMO_SHIFT2(at)
#undef NEMO_SHIFT3
#undef NEMO_SHIFT2
        tm[0] = tm[1];
        tm[1] = tm[2];
        tm[2] = tm[3];
        ssh_c = ssh_n;
    }
}

// Flather over the same segment lists as the FP64 path, FP32 values.
__global__ __launch_bounds__(64) void k_flather_f32(FlatherList lu, FlatherList lv,
                                                    float *__restrict__ ua,
                                                    float *__restrict__ va,
                                                    const float *__restrict__ su,
                                                    const float *__restrict__ sv)
{
    int t = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const FlatherList *L = &lu;
    float *vel = ua;
    const float *ssh = su;
    if (t >= lu.nseg) {
        t -= lu.nseg;
        if (t >= lv.nseg) return;
        L = &lv;
        vel = va;
        ssh = sv;
    }
    const int k1 = L->seg_ptr[t + 1];
    for (int k = L->seg_ptr[t]; k < k1; k++) {
        const int c = L->idx[k], off = L->off[k];
        const float v_nb = off ? vel[c + off] : 0.0f;
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

void launch_mixed_step(hipStream_t s, int rows, const MixedIn &f,
                       const StepScalars<float> &p, const MixedOut &o, int W,
                       Bounds b, const double *tide_tab, int tide_k, double tide_val)
{
    const int ncol = b.xe - b.xs + 3, nrow = b.ye - b.ys + 3;
    const int nstrip = (ncol + OUT - 1) / OUT;
    switch (rows) {
#define NEMO_MIXED_CASE(RR)                                                       \
    case RR: {                                                                    \
        const int nband = (nrow + RR - 1) / RR;                                   \
        k_mixed_wave<RR><<<dim3((unsigned)(nstrip * nband)), dim3(64), 0, s>>>( \
            f, p, o, W, b, nstrip, tide_tab, tide_k, tide_val);                   \
        break;                                                                    \
    }
        NEMO_MIXED_CASE(16)
        NEMO_MIXED_CASE(32)
        NEMO_MIXED_CASE(64)
        NEMO_MIXED_CASE(128)
#undef NEMO_MIXED_CASE
    default: NEMO_FATAL("mixed step: unsupported rows per band %d", rows);
    }
    NEMO_LAUNCH_CHECK();
}

void launch_flather_f32(hipStream_t s, const FlatherList &lu, const FlatherList &lv,
                        float *ua, float *va, const float *su, const float *sv)
{
    const int n = lu.nseg + lv.nseg;
    if (n == 0) return;
    k_flather_f32<<<dim3((unsigned)((n + 63) / 64)), dim3(64), 0, s>>>(lu, lv, ua, va,
                                                                      su, sv);
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
