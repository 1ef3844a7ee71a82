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

namespace nemo {
namespace {

__device__ __forceinline__ bool in_rect(int i, int j, int x0, int x1, int y0,
                                        int y1)
{
    return i >= x0 && i <= x1 && j >= y0 && j <= y1;
}

template <int MX, int R>
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

    // clamped columns: values of clamped loads are never used (every stencil
    // of a written point lies inside M x N)
    const int cw = max(i - 1, 0), ce = min(i + 1, xmax), cee = min(i + 2, xmax);
    const T tide = tide_tab ? tide_tab[tide_k] : tide_val;

    // ---- register window ----
    // 3-slot: rows j-1, j, j+1; 2-slot: rows j, j+1; tm: rows j-1 .. j+2
    T un[3], vn[3], hu[3], su[3], hv[3], sv[3], e1v[3], e2u[3];
    T ht[2], st[2], e1t[2], e2t[2], at[2];
    T un_w[2], un_e[2], vn_w[2], hu_w[2], su_w[2], hv_w[2], sv_w[2];
    T e1v_w[2], e2u_w[2], ht_e[2], st_e[2], e1t_e[2], e2t_e[2], at_e[2];
    T vn_e[3], hv_e[3], sv_e[3], e1v_e[3];
    int tm[4], tm_e[4], tm_w[2];

    auto rowoff = [&](int j) { return (long)min(max(j, 0), ymax) * W; };

    // aligned + shifted values of one row into slot s3 (3-slot) / s2 (2-slot)
    auto load_row = [&](int j, int s3, int s2) {
        const long r = rowoff(j);
        const long c = r + i, w = r + cw, e = r + ce;
        un[s3] = f.un[c];
        vn[s3] = f.vn[c];
        hu[s3] = f.hu[c];
        su[s3] = f.sshn_u[c];
        hv[s3] = f.hv[c];
        sv[s3] = f.sshn_v[c];
        e1v[s3] = f.dx_v[c];
        e2u[s3] = f.dy_u[c];
        ht[s2] = f.ht[c];
        st[s2] = f.sshn_t[c];
        e1t[s2] = f.dx_t[c];
        e2t[s2] = f.dy_t[c];
        at[s2] = f.area_t[c];
        un_w[s2] = f.un[w];
        un_e[s2] = f.un[e];
        vn_w[s2] = f.vn[w];
        hu_w[s2] = f.hu[w];
        su_w[s2] = f.sshn_u[w];
        hv_w[s2] = f.hv[w];
        sv_w[s2] = f.sshn_v[w];
        e1v_w[s2] = f.dx_v[w];
        e2u_w[s2] = f.dy_u[w];
        ht_e[s2] = f.ht[e];
        st_e[s2] = f.sshn_t[e];
        e1t_e[s2] = f.dx_t[e];
        e2t_e[s2] = f.dy_t[e];
        at_e[s2] = f.area_t[e];
        vn_e[s3] = f.vn[e];
        hv_e[s3] = f.hv[e];
        sv_e[s3] = f.sshn_v[e];
        e1v_e[s3] = f.dx_v[e];
        tm_w[s2] = f.tmask[w];
    };

    // new ssh of T(i, r) and T(i+1, r) from the slots of rows r-1 (s3-1) and
    // r (s3 / s2), and the tmask rows r-1 .. r+1 (tm slots t-1 .. t+1)
    auto new_ssh_pair = [&](int r, int s3, int s2, int t, T *sc, T *se) {
        const long off = rowoff(r);
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
        T x = continuity_val(v, p.rdt);
        *sc = tide_applies(tm[t], tm[t - 1], tm[t + 1], tm_e[t], tm_w[s2]) ? tide : x;

        ContVals<T> q;
        q.sshn = st_e[s2];
        q.su_e = f.sshn_u[off + ce];
        q.hu_e = f.hu[off + ce];
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
        const int m_ee = f.tmask[off + cee];
        T y = continuity_val(q, p.rdt);
        *se = tide_applies(tm_e[t], tm_e[t - 1], tm_e[t + 1], m_ee, tm[t]) ? tide : y;
    };

    // ---- prologue: rows jlo-1 (slot 0) and jlo (slot 1 / 0) ----
    load_row(jlo - 1, 0, 1);  // the 2-slot part of row jlo-1 is overwritten next
    load_row(jlo, 1, 0);
    {
        const long rm = rowoff(jlo - 1), r0 = rowoff(jlo), r1 = rowoff(jlo + 1);
        tm[0] = f.tmask[rm + i];
        tm[1] = f.tmask[r0 + i];
        tm[2] = f.tmask[r1 + i];
        tm_e[0] = f.tmask[rm + ce];
        tm_e[1] = f.tmask[r0 + ce];
        tm_e[2] = f.tmask[r1 + ce];
        tm_w[0] = f.tmask[r0 + cw];
    }
    // new ssh on row jlo: the pair routine reads 3-slot s3=1 (row jlo) and
    // s3-1=0 (row jlo-1), 2-slot 0 (row jlo), tm slots 0..2
    T ssh_c, ssh_e;
    new_ssh_pair(jlo, 1, 0, 1, &ssh_c, &ssh_e);

    for (int j = jlo; j < jhi; j++) {
        // ---- row j+1 into slot 2 / 1, tmask row j+2 into slot 3 ----
        load_row(j + 1, 2, 1);
        {
            const long r2 = rowoff(j + 2);
            tm[3] = f.tmask[r2 + i];
            tm_e[3] = f.tmask[r2 + ce];
        }
        const long r = rowoff(j);
        const long c = r + i;

        // ---- new ssh of T(i, j+1), T(i+1, j+1) ----
        T ssh_n, ssh_ne;
        new_ssh_pair(j + 1, 2, 1, 2, &ssh_n, &ssh_ne);

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
                v.e1u_c = f.dx_u[c];
                v.fcor = f.fcor_u[c];
                v.area = f.area_u[c];
                v.ssha = f.ssha_u[c];
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
                v.e2v_c = f.dy_v[c];
                v.fcor = f.fcor_v[c];
                v.area = f.area_v[c];
                v.ssha = f.ssha_v[c];
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
            if (next_ssh_pt(ssh_c, ssh_e, at[0], at_e[0], f.area_u[c], tm[1],
                            tm_e[1], &val))
                o.sshu[c] = val;
        }
        if (in_rect(i, j, b.xs, b.xe, b.ys, b.ye - 1)) {
            T val;
            if (next_ssh_pt(ssh_c, ssh_n, at[0], at[1], f.area_v[c], tm[1], tm[2],
                            &val))
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

struct MarchShape {
    int mx, rows;
};

template <int MX, int R>
void launch_march_t(hipStream_t s, const StepIn<double> &f,
                    const StepScalars<double> &p, const StepOut<double> &o,
                    int W, Bounds b, const double *tide_tab, int tide_k,
                    double tide_val)
{
    const int ncol = b.xe - b.xs + 3, nrow = b.ye - b.ys + 3;
    const int nstrip = (ncol + MX - 1) / MX, nband = (nrow + R - 1) / R;
    k_fused_march<MX, R><<<dim3((unsigned)(nstrip * nband)), dim3(MX), 0, s>>>(
        f, p, o, W, b, nstrip, tide_tab, tide_k, tide_val);
    NEMO_LAUNCH_CHECK();
}

}  // namespace

// Variants: index -> (block width MX, rows per band R).
static const MarchShape kShapes[] = {
    {256, 32}, {256, 64}, {256, 16}, {128, 32}, {128, 64}, {64, 64}, {64, 128},
};

int march_variant_count() { return (int)(sizeof(kShapes) / sizeof(kShapes[0])); }

void march_variant_shape(int variant, int *mx, int *rows)
{
    if (variant < 0 || variant >= march_variant_count())
        NEMO_FATAL("unknown fused march variant %d", variant);
    *mx = kShapes[variant].mx;
    *rows = kShapes[variant].rows;
}

void launch_fused_march(hipStream_t s, int variant, const StepIn<double> &f,
                        const StepScalars<double> &p, const StepOut<double> &o,
                        int W, Bounds b, const double *tide_tab, int tide_k,
                        double tide_val)
{
    switch (variant) {
    case 0: launch_march_t<256, 32>(s, f, p, o, W, b, tide_tab, tide_k, tide_val); break;
    case 1: launch_march_t<256, 64>(s, f, p, o, W, b, tide_tab, tide_k, tide_val); break;
    case 2: launch_march_t<256, 16>(s, f, p, o, W, b, tide_tab, tide_k, tide_val); break;
    case 3: launch_march_t<128, 32>(s, f, p, o, W, b, tide_tab, tide_k, tide_val); break;
    case 4: launch_march_t<128, 64>(s, f, p, o, W, b, tide_tab, tide_k, tide_val); break;
    case 5: launch_march_t<64, 64>(s, f, p, o, W, b, tide_tab, tide_k, tide_val); break;
    case 6: launch_march_t<64, 128>(s, f, p, o, W, b, tide_tab, tide_k, tide_val); break;
    default: NEMO_FATAL("unknown fused march variant %d", variant);
    }
}

}  // namespace nemo
