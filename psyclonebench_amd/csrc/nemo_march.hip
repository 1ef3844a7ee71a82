// Fused NEMOLite2D step, FP64, as a row march (CDNA4 / gfx950).
//
// Each lane owns one column i of the M x N extent and walks down R rows of a
// band; a block is MX consecutive columns (MX/64 wavefronts side by side).
// At row j a lane needs its stencil at rows j-1, j, j+1 and columns i-1, i,
// i+1 (i+2 for one tmask entry). Every value enters registers once, when its
// row is j+1 (aligned loads for column i, shifted loads for i-1 / i+1 that
// hit the lines the aligned loads of the neighbouring lanes just brought
// into L1), and is carried down the window as j advances, so each array is
// read from HBM once per band plus one halo row at each end.
//
// Per row the lane computes the new ssh of T(i,j+1) and T(i+1,j+1)
// (continuity + bc_ssh; the i+1 copy is what next_sshu needs, recomputed
// rather than exchanged across lanes), then momentum + bc_solid for U(i,j)
// and V(i,j) and next_sshu/v for U(i,j) and V(i,j). No LDS, no barriers.
//
// Block order is XCD-aware: the dispatcher hands block b to XCD b % 8, and
// block b is mapped to a logical tile so that each XCD owns one contiguous
// range of tiles, strips fastest: neighbouring strips (which share the
// cache lines of their halo columns) run on the same XCD at the same time.
//
// The arithmetic is nemo_point.h's *_val functions, identical to the other
// forms, so results are bitwise those of the reference.
#include "nemo_hip_common.h"
#include "nemo_step.h"
#include "nemo_wave.h"

namespace nemo {
namespace {

__device__ __forceinline__ bool in_rect(int i, int j, int x0, int x1, int y0,
                                        int y1)
{
    return i >= x0 && i <= x1 && j >= y0 && j <= y1;
}

// Everything one lane loads for row r: its own column i, the neighbours
// i-1 / i+1 (i+2 for one tmask entry), the tmask of row r+1, and the values
// of row r-1 that only the point itself reads.
template <typename T>
struct RowRaw {
    T un, vn, hu, su, hv, sv, e1v, e2u, ht, st, e1t, e2t, at;      // (i, r)
    T un_w, vn_w, hu_w, su_w, hv_w, sv_w, e1v_w, e2u_w;              // (i-1, r)
    T un_e, vn_e, hu_e, su_e, hv_e, sv_e, e1v_e, ht_e, st_e, e1t_e, e2t_e,
        at_e;                                                        // (i+1, r)
    int tm_w, tm_ee;                                                 // (i-1|i+2, r)
    int tm_n, tm_ne;                                                 // (i|i+1, r+1)
    T dxu, dyv, fcu, fcv, aru, arv, sau, sav;                        // (i, r-1)
};

template <typename T>
__device__ __forceinline__ void load_raw(RowRaw<T> &q, const StepIn<T> &f,
                                         long r, long rn, long rp, int i, int cw,
                                         int ce, int cee)
{
    const long c = r + i, w = r + cw, e = r + ce;
    q.un = f.un[c];
    q.vn = f.vn[c];
    q.hu = f.hu[c];
    q.su = f.sshn_u[c];
    q.hv = f.hv[c];
    q.sv = f.sshn_v[c];
    q.e1v = f.dx_v[c];
    q.e2u = f.dy_u[c];
    q.ht = f.ht[c];
    q.st = f.sshn_t[c];
    q.e1t = f.dx_t[c];
    q.e2t = f.dy_t[c];
    q.at = f.area_t[c];
    q.un_w = f.un[w];
    q.vn_w = f.vn[w];
    q.hu_w = f.hu[w];
    q.su_w = f.sshn_u[w];
    q.hv_w = f.hv[w];
    q.sv_w = f.sshn_v[w];
    q.e1v_w = f.dx_v[w];
    q.e2u_w = f.dy_u[w];
    q.un_e = f.un[e];
    q.vn_e = f.vn[e];
    q.hu_e = f.hu[e];
    q.su_e = f.sshn_u[e];
    q.hv_e = f.hv[e];
    q.sv_e = f.sshn_v[e];
    q.e1v_e = f.dx_v[e];
    q.ht_e = f.ht[e];
    q.st_e = f.sshn_t[e];
    q.e1t_e = f.dx_t[e];
    q.e2t_e = f.dy_t[e];
    q.at_e = f.area_t[e];
    q.tm_w = f.tmask[w];
    q.tm_ee = f.tmask[r + cee];
    q.tm_n = f.tmask[rn + i];
    q.tm_ne = f.tmask[rn + ce];
    const long cp = rp + i;
    q.dxu = f.dx_u[cp];
    q.dyv = f.dy_v[cp];
    q.fcu = f.fcor_u[cp];
    q.fcv = f.fcor_v[cp];
    q.aru = f.area_u[cp];
    q.arv = f.area_v[cp];
    q.sau = f.ssha_u[cp];
    q.sav = f.ssha_v[cp];
}

// PF = prefetch: the loads of row j+2 are issued before the arithmetic of
// row j, so a lane always has one row of loads in flight while it computes.
template <int MX, int R, bool PF>
__global__ __launch_bounds__(MX) void k_fused_march(
    StepIn<double> f, StepScalars<double> p, StepOut<double> o, int W, Bounds b,
    int nstrip, const double *__restrict__ tide_tab, int tide_k, double tide_val)
{
    using T = double;
    // ---- tile of this block (XCD-aware order) ----
    const int nblk = (int)gridDim.x, bid = (int)blockIdx.x;
    const int xcd = bid & 7, k = bid >> 3;
    const int per = nblk >> 3, extra = nblk & 7;
    const int logical = xcd * per + min(xcd, extra) + k;
    const int band = logical / nstrip, strip = logical - band * nstrip;

    const int xmax = b.xe + 1, ymax = b.ye + 1;  // last column / row of M x N
    const int i = b.xs - 1 + strip * MX + (int)threadIdx.x;
    if (i > xmax) return;  // no barriers below: idle lanes may leave
    const int jlo = b.ys - 1 + band * R;
    const int jhi = min(jlo + R, ymax + 1);

    // clamped columns and rows: values of clamped loads are never used
    // (every stencil of a written point lies inside M x N)
    const int cw = max(i - 1, 0), ce = min(i + 1, xmax), cee = min(i + 2, xmax);
    auto rowoff = [&](int j) { return (long)min(max(j, 0), ymax) * W; };
    auto fetch = [&](RowRaw<T> &q, int r) {
        load_raw(q, f, rowoff(r), rowoff(r + 1), rowoff(r - 1), i, cw, ce, cee);
    };
    const T tide = tide_tab ? tide_tab[tide_k] : tide_val;

    // ---- register window ----
    // 3-slot: rows j-1, j, j+1; 2-slot: rows j, j+1; tm: rows j-1 .. j+2
    T un[3], vn[3], hu[3], su[3], hv[3], sv[3], e1v[3], e2u[3];
    T vn_e[3], hv_e[3], sv_e[3], e1v_e[3];
    T ht[2], st[2], e1t[2], e2t[2], at[2];
    T un_w[2], un_e[2], vn_w[2], hu_w[2], su_w[2], hv_w[2], sv_w[2];
    T e1v_w[2], e2u_w[2], ht_e[2], st_e[2], e1t_e[2], e2t_e[2], at_e[2];
    int tm[4], tm_e[4], tm_w[2];
    // row j+1 values used once (continuity at i+1) and row-j point values
    T hu_e1, su_e1, dxu, dyv, fcu, fcv, aru, arv, sau, sav;
    int tm_ee1;

    auto put = [&](const RowRaw<T> &q, int s3, int s2) {
        un[s3] = q.un;
        vn[s3] = q.vn;
        hu[s3] = q.hu;
        su[s3] = q.su;
        hv[s3] = q.hv;
        sv[s3] = q.sv;
        e1v[s3] = q.e1v;
        e2u[s3] = q.e2u;
        vn_e[s3] = q.vn_e;
        hv_e[s3] = q.hv_e;
        sv_e[s3] = q.sv_e;
        e1v_e[s3] = q.e1v_e;
        if (s2 >= 0) {
            ht[s2] = q.ht;
            st[s2] = q.st;
            e1t[s2] = q.e1t;
            e2t[s2] = q.e2t;
            at[s2] = q.at;
            un_w[s2] = q.un_w;
            un_e[s2] = q.un_e;
            vn_w[s2] = q.vn_w;
            hu_w[s2] = q.hu_w;
            su_w[s2] = q.su_w;
            hv_w[s2] = q.hv_w;
            sv_w[s2] = q.sv_w;
            e1v_w[s2] = q.e1v_w;
            e2u_w[s2] = q.e2u_w;
            ht_e[s2] = q.ht_e;
            st_e[s2] = q.st_e;
            e1t_e[s2] = q.e1t_e;
            e2t_e[s2] = q.e2t_e;
            at_e[s2] = q.at_e;
            tm_w[s2] = q.tm_w;
        }
        hu_e1 = q.hu_e;
        su_e1 = q.su_e;
        tm_ee1 = q.tm_ee;
        dxu = q.dxu;
        dyv = q.dyv;
        fcu = q.fcu;
        fcv = q.fcv;
        aru = q.aru;
        arv = q.arv;
        sau = q.sau;
        sav = q.sav;
    };

    // new ssh of T(i, r) and T(i+1, r) from 3-slot s3 (row r) / s3-1 (row
    // r-1), 2-slot s2 (row r) and tmask slots t-1 .. t+1 (rows r-1 .. r+1)
    auto new_ssh_pair = [&](int s3, int s2, int t, T *sc, T *se) {
        ContVals<T> v;
        v.sshn = st[s2];
        v.su_e = su[s3];
        v.hu_e = hu[s3];
        v.un_e = un[s3];
        v.su_w = su_w[s2];
        v.hu_w = hu_w[s2];
        v.un_w = un_w[s2];
        v.sv_n = sv[s3];
        v.hv_n = hv[s3];
        v.vn_n = vn[s3];
        v.sv_s = sv[s3 - 1];
        v.hv_s = hv[s3 - 1];
        v.vn_s = vn[s3 - 1];
        v.area = at[s2];
        const T x = continuity_val(v, p.rdt);
        *sc = tide_applies(tm[t], tm[t - 1], tm[t + 1], tm_e[t], tm_w[s2]) ? tide : x;

        ContVals<T> q;
        q.sshn = st_e[s2];
        q.su_e = su_e1;
        q.hu_e = hu_e1;
        q.un_e = un_e[s2];
        q.su_w = su[s3];
        q.hu_w = hu[s3];
        q.un_w = un[s3];
        q.sv_n = sv_e[s3];
        q.hv_n = hv_e[s3];
        q.vn_n = vn_e[s3];
        q.sv_s = sv_e[s3 - 1];
        q.hv_s = hv_e[s3 - 1];
        q.vn_s = vn_e[s3 - 1];
        q.area = at_e[s2];
        const T y = continuity_val(q, p.rdt);
        *se = tide_applies(tm_e[t], tm_e[t - 1], tm_e[t + 1], tm_ee1, tm[t]) ? tide : y;
    };

    // ---- prologue: rows jlo-1 (3-slot 0) and jlo (3-slot 1, 2-slot 0) ----
    RowRaw<T> raw;
    fetch(raw, jlo - 1);
    put(raw, 0, -1);
    fetch(raw, jlo);
    put(raw, 1, 0);
    {
        const long rm = rowoff(jlo - 1);
        tm[0] = f.tmask[rm + i];
        tm_e[0] = f.tmask[rm + ce];
        const long r0 = rowoff(jlo);
        tm[1] = f.tmask[r0 + i];
        tm_e[1] = f.tmask[r0 + ce];
        tm[2] = raw.tm_n;
        tm_e[2] = raw.tm_ne;
    }
    T ssh_c, ssh_e;  // new ssh of T(i, j), T(i+1, j)
    new_ssh_pair(1, 0, 1, &ssh_c, &ssh_e);
    if (PF) fetch(raw, jlo + 1);

    for (int j = jlo; j < jhi; j++) {
        // ---- row j+1 into 3-slot 2 / 2-slot 1, tmask row j+2 into slot 3 ----
        if (!PF) fetch(raw, j + 1);
        put(raw, 2, 1);
        tm[3] = raw.tm_n;
        tm_e[3] = raw.tm_ne;
        if (PF) fetch(raw, j + 2);  // in flight during the arithmetic below
        const long c = rowoff(j) + i;

        // ---- new ssh of T(i, j+1), T(i+1, j+1) ----
        T ssh_n, ssh_ne;
        new_ssh_pair(2, 1, 2, &ssh_n, &ssh_ne);

        // ---- T(i,j) ----
        if (in_rect(i, j, b.xs, b.xe, b.ys, b.ye)) o.ssh[c] = ssh_c;

        // ---- U(i,j): momentum_u, else bc_solid_u ----
        {
            bool wrote = false;
            T val = T(0);
            if (in_rect(i, j, b.xs, b.xe - 1, b.ys, b.ye)) {
                MomUVals<T> v;
                v.m_c = tm[1];
                v.m_e = tm_e[1];
                v.m_s = tm[0];
                v.m_se = tm_e[0];
                v.m_n = tm[2];
                v.m_ne = tm_e[2];
                v.un_c = un[1];
                v.un_e = un_e[0];
                v.un_w = un_w[0];
                v.un_s = un[0];
                v.un_n = un[2];
                v.vn_s = vn[0];
                v.vn_se = vn_e[0];
                v.vn_c = vn[1];
                v.vn_e = vn_e[1];
                v.e1v_s = e1v[0];
                v.e1v_se = e1v_e[0];
                v.e1v_c = e1v[1];
                v.e1v_e = e1v_e[1];
                v.hv_s = hv[0];
                v.sv_s = sv[0];
                v.hv_se = hv_e[0];
                v.sv_se = sv_e[0];
                v.hv_c = hv[1];
                v.sv_c = sv[1];
                v.hv_e = hv_e[1];
                v.sv_e = sv_e[1];
                v.e2t_c = e2t[0];
                v.e2t_e = e2t_e[0];
                v.e1t_c = e1t[0];
                v.e1t_e = e1t_e[0];
                v.ht_c = ht[0];
                v.ht_e = ht_e[0];
                v.st_c = st[0];
                v.st_e = st_e[0];
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
                wrote = momentum_u_val(v, p, &val);
            }
            if (!wrote && in_rect(i, j, b.xs - 1, b.xe, b.ys - 1, b.ye + 1) &&
                tm[1] * tm_e[1] == 0) {
                val = T(0);
                wrote = true;
            }
            if (wrote) o.u[c] = val;
        }

        // ---- V(i,j): momentum_v, else bc_solid_v ----
        {
            bool wrote = false;
            T val = T(0);
            if (in_rect(i, j, b.xs, b.xe, b.ys, b.ye - 1)) {
                MomVVals<T> v;
                v.m_c = tm[1];
                v.m_e = tm_e[1];
                v.m_n = tm[2];
                v.m_w = tm_w[0];
                v.m_nw = tm_w[1];
                v.m_ne = tm_e[2];
                v.vn_c = vn[1];
                v.vn_n = vn[2];
                v.vn_s = vn[0];
                v.vn_w = vn_w[0];
                v.vn_e = vn_e[1];
                v.un_w = un_w[0];
                v.un_nw = un_w[1];
                v.un_c = un[1];
                v.un_n = un[2];
                v.e2u_w = e2u_w[0];
                v.e2u_nw = e2u_w[1];
                v.e2u_c = e2u[1];
                v.e2u_n = e2u[2];
                v.hu_w = hu_w[0];
                v.su_w = su_w[0];
                v.hu_nw = hu_w[1];
                v.su_nw = su_w[1];
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
                v.e1v_w = e1v_w[0];
                v.e1v_e = e1v_e[1];
                v.hv_c = hv[1];
                v.sv_c = sv[1];
                v.hv_w = hv_w[0];
                v.sv_w = sv_w[0];
                v.hv_e = hv_e[1];
                v.sv_e = sv_e[1];
                v.e2v_c = dyv;
                v.fcor = fcv;
                v.area = arv;
                v.ssha = sav;
                wrote = momentum_v_val(v, p, &val);
            }
            if (!wrote && in_rect(i, j, b.xs - 1, b.xe + 1, b.ys - 1, b.ye) &&
                tm[1] * tm[2] == 0) {
                val = T(0);
                wrote = true;
            }
            if (wrote) o.v[c] = val;
        }

        // ---- next_sshu(i,j), next_sshv(i,j) from the new ssh ----
        if (in_rect(i, j, b.xs, b.xe - 1, b.ys, b.ye)) {
            T val;
            if (next_ssh_pt(ssh_c, ssh_e, at[0], at_e[0], aru, tm[1], tm_e[1], &val))
                o.sshu[c] = val;
        }
        if (in_rect(i, j, b.xs, b.xe, b.ys, b.ye - 1)) {
            T val;
            if (next_ssh_pt(ssh_c, ssh_n, at[0], at[1], arv, tm[1], tm[2], &val))
                o.sshv[c] = val;
        }

        // ---- advance the window ----
#define NEMO_SHIFT3(a) \
    a[0] = a[1];       \
    a[1] = a[2];
#define NEMO_SHIFT2(a) a[0] = a[1];
        NEMO_SHIFT3(un) NEMO_SHIFT3(vn) NEMO_SHIFT3(hu) NEMO_SHIFT3(su)
        NEMO_SHIFT3(hv) NEMO_SHIFT3(sv) NEMO_SHIFT3(e1v) NEMO_SHIFT3(e2u)
        NEMO_SHIFT3(vn_e) NEMO_SHIFT3(hv_e) NEMO_SHIFT3(sv_e) NEMO_SHIFT3(e1v_e)
        NEMO_SHIFT2(ht) NEMO_SHIFT2(st) NEMO_SHIFT2(e1t) NEMO_SHIFT2(e2t)
        NEMO_SHIFT2(at) NEMO_SHIFT2(un_w) NEMO_SHIFT2(un_e) NEMO_SHIFT2(vn_w)
        NEMO_SHIFT2(hu_w) NEMO_SHIFT2(su_w) NEMO_SHIFT2(hv_w) NEMO_SHIFT2(sv_w)
        NEMO_SHIFT2(e1v_w) NEMO_SHIFT2(e2u_w) NEMO_SHIFT2(ht_e) NEMO_SHIFT2(st_e)
        NEMO_SHIFT2(e1t_e) NEMO_SHIFT2(e2t_e) NEMO_SHIFT2(at_e) NEMO_SHIFT2(tm_w)
#undef NEMO_SHIFT3
#undef NEMO_SHIFT2
        tm[0] = tm[1];
        tm[1] = tm[2];
        tm[2] = tm[3];
        tm_e[0] = tm_e[1];
        tm_e[1] = tm_e[2];
        tm_e[2] = tm_e[3];
        ssh_c = ssh_n;
        ssh_e = ssh_ne;
    }
}

// ---------------------------------------------------------------------------
// Overlapped-wavefront march (the tile body is wave::tile_f64): one block per
// tile, XCD-aware tile order.
template <int R, int MINW>
__global__ __launch_bounds__(64, MINW) void k_fused_wave(
    StepIn<double> f, StepScalars<double> p, StepOut<double> o, int W, Bounds b,
    int nstrip, const double *__restrict__ tide_tab, int tide_k, double tide_val)
{
    const int logical = wave::xcd_tile((int)blockIdx.x, (int)gridDim.x);
    const int band = logical / nstrip, strip = logical - band * nstrip;
    wave::tile_f64<R>(f, p, o, W, 0u, b, strip, band, tide_tab ? tide_tab[tide_k] : tide_val);
}

// The same march with OW owned columns per wavefront, OW a multiple of 8, so
// that every strip's stores start on a 64-byte boundary (OW = 56: the loads
// start 32 bytes before one; OW = 48: on one).
template <int R, int OW>
__global__ __launch_bounds__(64) void k_fused_wave_a(
    StepIn<double> f, StepScalars<double> p, StepOut<double> o, int W, Bounds b,
    int nstrip, const double *__restrict__ tide_tab, int tide_k, double tide_val)
{
    static_assert(OW % 8 == 0, "64-byte-aligned strips");
    const int logical = wave::xcd_tile((int)blockIdx.x, (int)gridDim.x);
    const int band = logical / nstrip, strip = logical - band * nstrip;
    wave::tile_f64<R, OW>(f, p, o, W, 0u, b, strip, band,
                          tide_tab ? tide_tab[tide_k] : tide_val);
}

// The aligned march with the loads issued PF+1 rows ahead (tile_f64's PF);
// MINW wavefronts per SIMD requested from the register allocator.
template <int R, int OW, int MINW, int PF = 1>
__global__ __launch_bounds__(64, MINW) void k_fused_wave_p(
    StepIn<double> f, StepScalars<double> p, StepOut<double> o, int W, Bounds b,
    int nstrip, const double *__restrict__ tide_tab, int tide_k, double tide_val)
{
    static_assert(OW % 8 == 0, "64-byte-aligned strips");
    const int logical = wave::xcd_tile((int)blockIdx.x, (int)gridDim.x);
    const int band = logical / nstrip, strip = logical - band * nstrip;
    wave::tile_f64<R, OW, PF>(f, p, o, W, 0u, b, strip, band,
                              tide_tab ? tide_tab[tide_k] : tide_val);
}

template <int R, int OW, int MINW, int PF = 1>
void launch_wave_prefetch_t(hipStream_t s, const StepIn<double> &f,
                            const StepScalars<double> &p, const StepOut<double> &o, int W,
                            Bounds b, const double *tide_tab, int tide_k, double tide_val)
{
    const int ncol = b.xe - b.xs + 3, nrow = b.ye - b.ys + 3;
    const int nstrip = (ncol + OW - 1) / OW, nband = (nrow + R - 1) / R;
    k_fused_wave_p<R, OW, MINW, PF><<<dim3((unsigned)(nstrip * nband)), dim3(64), 0, s>>>(
        f, p, o, W, b, nstrip, tide_tab, tide_k, tide_val);
    NEMO_LAUNCH_CHECK();
}

// Persistent chunked march: one wavefront per (strip, chunk), a chunk being
// `rows` consecutive rows of the strip. The host sizes the launch to the
// wavefronts the GPU holds at once (the kernel's occupancy times the CU
// count), so every wavefront marches one long run with its window kept (a
// halo of 2 rows per run instead of per short band) and all of them finish
// together. PF and U as in tile_f64_rows.
template <int OW, int PF, int U>
__global__ __launch_bounds__(64, 2) void k_fused_wave_c(
    StepIn<double> f, StepScalars<double> p, StepOut<double> o, int W, Bounds b,
    int nstrip, int rows, const double *__restrict__ tide_tab, int tide_k, double tide_val)
{
    static_assert(OW % 8 == 0, "64-byte-aligned strips");
    const int logical = wave::xcd_tile((int)blockIdx.x, (int)gridDim.x);
    const int chunk = logical / nstrip, strip = logical - chunk * nstrip;
    const int jlo = b.ys - 1 + chunk * rows;
    const int jhi = min(jlo + rows, b.ye + 2);
    wave::tile_f64_rows<1 << 24, OW, PF, U>(f, p, o, W, 0u, b, strip, jlo, jhi,
                                            tide_tab ? tide_tab[tide_k] : tide_val);
}

// Wavefronts of k_fused_wave_c<OW, PF, U> the current device holds at once.
template <int OW, int PF, int U>
int chunk_residency()
{
    constexpr int kMaxDevices = 64;
    static int cache[kMaxDevices] = {};
    int dev = 0;
    NEMO_HIP_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) NEMO_FATAL("device index %d beyond %d", dev, kMaxDevices);
    if (!cache[dev]) {
        int per_cu = 0, ncu = 0;
        NEMO_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &per_cu, k_fused_wave_c<OW, PF, U>, 64, 0));
        NEMO_HIP_CHECK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
        cache[dev] = per_cu * ncu > 0 ? per_cu * ncu : 1;
    }
    return cache[dev];
}

template <int OW, int PF, int U>
void launch_wave_chunk_t(hipStream_t s, const StepIn<double> &f,
                         const StepScalars<double> &p, const StepOut<double> &o, int W,
                         Bounds b, const double *tide_tab, int tide_k, double tide_val)
{
    const int ncol = b.xe - b.xs + 3, nrow = b.ye - b.ys + 3;
    const int nstrip = (ncol + OW - 1) / OW;
    const int want = max(1, chunk_residency<OW, PF, U>() / nstrip);
    const int rows = (nrow + want - 1) / want;
    const int nchunk = (nrow + rows - 1) / rows;
    k_fused_wave_c<OW, PF, U><<<dim3((unsigned)(nstrip * nchunk)), dim3(64), 0, s>>>(
        f, p, o, W, b, nstrip, rows, tide_tab, tide_k, tide_val);
    NEMO_LAUNCH_CHECK();
}

template <int R, int OW>
void launch_wave_aligned_t(hipStream_t s, const StepIn<double> &f,
                           const StepScalars<double> &p, const StepOut<double> &o, int W,
                           Bounds b, const double *tide_tab, int tide_k, double tide_val)
{
    const int ncol = b.xe - b.xs + 3, nrow = b.ye - b.ys + 3;
    const int nstrip = (ncol + OW - 1) / OW, nband = (nrow + R - 1) / R;
    k_fused_wave_a<R, OW><<<dim3((unsigned)(nstrip * nband)), dim3(64), 0, s>>>(
        f, p, o, W, b, nstrip, tide_tab, tide_k, tide_val);
    NEMO_LAUNCH_CHECK();
}

template <int R, int MINW>
void launch_wave_t(hipStream_t s, const StepIn<double> &f,
                   const StepScalars<double> &p, const StepOut<double> &o, int W,
                   Bounds b, const double *tide_tab, int tide_k, double tide_val)
{
    const int ncol = b.xe - b.xs + 3, nrow = b.ye - b.ys + 3;
    const int nstrip = (ncol + wave::OUT - 1) / wave::OUT, nband = (nrow + R - 1) / R;
    k_fused_wave<R, MINW><<<dim3((unsigned)(nstrip * nband)), dim3(64), 0, s>>>(
        f, p, o, W, b, nstrip, tide_tab, tide_k, tide_val);
    NEMO_LAUNCH_CHECK();
}

// kind: 0 register-window march, 1 the same with the row loads issued one
// row ahead, 2 overlapped-wavefront march (61 output columns per block),
// 3 / 4 the overlapped-wavefront march compiled for >= 3 / 4 wavefronts per
// SIMD (the compiler keeps its registers within that budget), 5 / 6 the
// overlapped-wavefront march with 56 / 48 output columns (64-byte-aligned
// stores; with 48 the loads are aligned too), 7 / 8 the 56-column march with
// its loads issued two rows ahead, compiled for >= 2 / 3 wavefronts per SIMD,
// 9 / 10 the persistent chunked 56-column march with / without the loads two
// rows ahead (rows 0: the chunk height is set at launch from the occupancy),
// 11 the 56-column march with its loads issued three rows ahead
struct MarchShape {
    int mx, rows, prefetch;
};

template <int MX, int R, bool PF>
void launch_march_t(hipStream_t s, const StepIn<double> &f,
                    const StepScalars<double> &p, const StepOut<double> &o,
                    int W, Bounds b, const double *tide_tab, int tide_k,
                    double tide_val)
{
    const int ncol = b.xe - b.xs + 3, nrow = b.ye - b.ys + 3;
    const int nstrip = (ncol + MX - 1) / MX, nband = (nrow + R - 1) / R;
    k_fused_march<MX, R, PF><<<dim3((unsigned)(nstrip * nband)), dim3(MX), 0, s>>>(
        f, p, o, W, b, nstrip, tide_tab, tide_k, tide_val);
    NEMO_LAUNCH_CHECK();
}

}  // namespace

// Variants: index -> (block width MX, rows per band R, prefetch).
static const MarchShape kShapes[] = {
    {256, 32, 0}, {256, 64, 0}, {256, 16, 0}, {128, 32, 0}, {128, 64, 0},
    {64, 64, 0},  {64, 128, 0}, {256, 32, 1}, {256, 64, 1}, {256, 16, 1},
    {128, 32, 1}, {64, 64, 1},  {64, 32, 1},  {64, 16, 2},  {64, 32, 2},
    {64, 64, 2},  {64, 128, 2}, {64, 32, 3},  {64, 64, 3},  {64, 32, 4},
    {64, 64, 4},  {64, 16, 5},  {64, 32, 5},  {64, 8, 5},   {64, 16, 6},
    {64, 32, 6},  {64, 4, 5},   {64, 8, 6},   {64, 4, 6},   {64, 2, 5},
    {64, 6, 5},   {64, 4, 7},   {64, 6, 7},   {64, 8, 7},   {64, 12, 7},
    {64, 4, 8},   {64, 8, 8},   {64, 32, 7},  {64, 64, 7},  {64, 0, 9},
    {64, 0, 9},   {64, 0, 10},  {64, 4, 11},  {64, 6, 11},  {64, 8, 11},
};

// Kernel names as the profiler reports them (template arguments included).
static const char *const kNames[] = {
    "k_fused_march<256, 32, false>", "k_fused_march<256, 64, false>",
    "k_fused_march<256, 16, false>", "k_fused_march<128, 32, false>",
    "k_fused_march<128, 64, false>", "k_fused_march<64, 64, false>",
    "k_fused_march<64, 128, false>", "k_fused_march<256, 32, true>",
    "k_fused_march<256, 64, true>",  "k_fused_march<256, 16, true>",
    "k_fused_march<128, 32, true>",  "k_fused_march<64, 64, true>",
    "k_fused_march<64, 32, true>",   "k_fused_wave<16, 1>",
    "k_fused_wave<32, 1>",           "k_fused_wave<64, 1>",
    "k_fused_wave<128, 1>",          "k_fused_wave<32, 3>",
    "k_fused_wave<64, 3>",           "k_fused_wave<32, 4>",
    "k_fused_wave<64, 4>",           "k_fused_wave_a<16, 56>",
    "k_fused_wave_a<32, 56>",        "k_fused_wave_a<8, 56>",
    "k_fused_wave_a<16, 48>",        "k_fused_wave_a<32, 48>",
    "k_fused_wave_a<4, 56>",         "k_fused_wave_a<8, 48>",
    "k_fused_wave_a<4, 48>",         "k_fused_wave_a<2, 56>",
    "k_fused_wave_a<6, 56>",         "k_fused_wave_p<4, 56, 2, 1>",
    "k_fused_wave_p<6, 56, 2, 1>",      "k_fused_wave_p<8, 56, 2, 1>",
    "k_fused_wave_p<12, 56, 2, 1>",     "k_fused_wave_p<4, 56, 3, 1>",
    "k_fused_wave_p<8, 56, 3, 1>",      "k_fused_wave_p<32, 56, 2, 1>",
    "k_fused_wave_p<64, 56, 2, 1>",     "k_fused_wave_c<56, 1, 8>",
    "k_fused_wave_c<56, 1, 4>",      "k_fused_wave_c<56, 0, 4>",
    "k_fused_wave_p<4, 56, 2, 2>",   "k_fused_wave_p<6, 56, 2, 2>",
    "k_fused_wave_p<8, 56, 2, 2>",
};
static_assert(sizeof(kNames) / sizeof(kNames[0]) == sizeof(kShapes) / sizeof(kShapes[0]),
              "one kernel name per shape");

int march_variant_count() { return (int)(sizeof(kShapes) / sizeof(kShapes[0])); }

const char *march_variant_kernel(int variant)
{
    if (variant < 0 || variant >= march_variant_count())
        NEMO_FATAL("unknown fused march variant %d", variant);
    return kNames[variant];
}

void march_variant_shape(int variant, int *mx, int *rows, int *prefetch)
{
    if (variant < 0 || variant >= march_variant_count())
        NEMO_FATAL("unknown fused march variant %d", variant);
    *mx = kShapes[variant].mx;
    *rows = kShapes[variant].rows;
    *prefetch = kShapes[variant].prefetch;
}

void launch_fused_march(hipStream_t s, int variant, const StepIn<double> &f,
                        const StepScalars<double> &p, const StepOut<double> &o,
                        int W, Bounds b, const double *tide_tab, int tide_k,
                        double tide_val)
{
#define NEMO_MARCH_CASE(n, MX, R, PF)                                           \
    case n:                                                                    \
        launch_march_t<MX, R, PF>(s, f, p, o, W, b, tide_tab, tide_k, tide_val); \
        break;
    switch (variant) {
        NEMO_MARCH_CASE(0, 256, 32, false)
        NEMO_MARCH_CASE(1, 256, 64, false)
        NEMO_MARCH_CASE(2, 256, 16, false)
        NEMO_MARCH_CASE(3, 128, 32, false)
        NEMO_MARCH_CASE(4, 128, 64, false)
        NEMO_MARCH_CASE(5, 64, 64, false)
        NEMO_MARCH_CASE(6, 64, 128, false)
        NEMO_MARCH_CASE(7, 256, 32, true)
        NEMO_MARCH_CASE(8, 256, 64, true)
        NEMO_MARCH_CASE(9, 256, 16, true)
        NEMO_MARCH_CASE(10, 128, 32, true)
        NEMO_MARCH_CASE(11, 64, 64, true)
        NEMO_MARCH_CASE(12, 64, 32, true)
    case 13: launch_wave_t<16, 1>(s, f, p, o, W, b, tide_tab, tide_k, tide_val); break;
    case 14: launch_wave_t<32, 1>(s, f, p, o, W, b, tide_tab, tide_k, tide_val); break;
    case 15: launch_wave_t<64, 1>(s, f, p, o, W, b, tide_tab, tide_k, tide_val); break;
    case 16: launch_wave_t<128, 1>(s, f, p, o, W, b, tide_tab, tide_k, tide_val); break;
    case 17: launch_wave_t<32, 3>(s, f, p, o, W, b, tide_tab, tide_k, tide_val); break;
    case 18: launch_wave_t<64, 3>(s, f, p, o, W, b, tide_tab, tide_k, tide_val); break;
    case 19: launch_wave_t<32, 4>(s, f, p, o, W, b, tide_tab, tide_k, tide_val); break;
    case 20: launch_wave_t<64, 4>(s, f, p, o, W, b, tide_tab, tide_k, tide_val); break;
    case 21:
        launch_wave_aligned_t<16, 56>(s, f, p, o, W, b, tide_tab, tide_k, tide_val);
        break;
    case 22:
        launch_wave_aligned_t<32, 56>(s, f, p, o, W, b, tide_tab, tide_k, tide_val);
        break;
    case 23:
        launch_wave_aligned_t<8, 56>(s, f, p, o, W, b, tide_tab, tide_k, tide_val);
        break;
    case 24:
        launch_wave_aligned_t<16, 48>(s, f, p, o, W, b, tide_tab, tide_k, tide_val);
        break;
    case 25:
        launch_wave_aligned_t<32, 48>(s, f, p, o, W, b, tide_tab, tide_k, tide_val);
        break;
    case 26:
        launch_wave_aligned_t<4, 56>(s, f, p, o, W, b, tide_tab, tide_k, tide_val);
        break;
    case 27:
        launch_wave_aligned_t<8, 48>(s, f, p, o, W, b, tide_tab, tide_k, tide_val);
        break;
    case 28:
        launch_wave_aligned_t<4, 48>(s, f, p, o, W, b, tide_tab, tide_k, tide_val);
        break;
    case 29:
        launch_wave_aligned_t<2, 56>(s, f, p, o, W, b, tide_tab, tide_k, tide_val);
        break;
    case 30:
        launch_wave_aligned_t<6, 56>(s, f, p, o, W, b, tide_tab, tide_k, tide_val);
        break;
    case 31:
        launch_wave_prefetch_t<4, 56, 2>(s, f, p, o, W, b, tide_tab, tide_k, tide_val);
        break;
    case 32:
        launch_wave_prefetch_t<6, 56, 2>(s, f, p, o, W, b, tide_tab, tide_k, tide_val);
        break;
    case 33:
        launch_wave_prefetch_t<8, 56, 2>(s, f, p, o, W, b, tide_tab, tide_k, tide_val);
        break;
    case 34:
        launch_wave_prefetch_t<12, 56, 2>(s, f, p, o, W, b, tide_tab, tide_k, tide_val);
        break;
    case 35:
        launch_wave_prefetch_t<4, 56, 3>(s, f, p, o, W, b, tide_tab, tide_k, tide_val);
        break;
    case 36:
        launch_wave_prefetch_t<8, 56, 3>(s, f, p, o, W, b, tide_tab, tide_k, tide_val);
        break;
    case 37:
        launch_wave_prefetch_t<32, 56, 2>(s, f, p, o, W, b, tide_tab, tide_k, tide_val);
        break;
    case 38:
        launch_wave_prefetch_t<64, 56, 2>(s, f, p, o, W, b, tide_tab, tide_k, tide_val);
        break;
    case 39:
        launch_wave_chunk_t<56, 1, 8>(s, f, p, o, W, b, tide_tab, tide_k, tide_val);
        break;
    case 40:
        launch_wave_chunk_t<56, 1, 4>(s, f, p, o, W, b, tide_tab, tide_k, tide_val);
        break;
    case 41:
        launch_wave_chunk_t<56, 0, 4>(s, f, p, o, W, b, tide_tab, tide_k, tide_val);
        break;
    case 42:
        launch_wave_prefetch_t<4, 56, 2, 2>(s, f, p, o, W, b, tide_tab, tide_k, tide_val);
        break;
    case 43:
        launch_wave_prefetch_t<6, 56, 2, 2>(s, f, p, o, W, b, tide_tab, tide_k, tide_val);
        break;
    case 44:
        launch_wave_prefetch_t<8, 56, 2, 2>(s, f, p, o, W, b, tide_tab, tide_k, tide_val);
        break;
    default: NEMO_FATAL("unknown fused march variant %d", variant);
    }
#undef NEMO_MARCH_CASE
}

}  // namespace nemo
