// Overlapped-wavefront tile of the fused NEMOLite2D step, FP64 and mixed
// precision: the bodies shared by the single-domain kernels (nemo_march.hip,
// nemo_mixed.hip) and the ensemble kernels (nemo_ensemble.hip).
//
// Overlapped-wavefront march. One block = one wavefront of 64 lanes over 64
// consecutive columns, of which OW lanes own their column's outputs (by
// default lanes 1..61, OW = 61) and the others are halo (neighbouring strips
// overlap by 64 - OW columns). Every value a lane needs at column i-1 or i+1
// sits in lane l-1 / l+1 of the same wavefront, so it is taken with a DPP
// wave shift rather than loaded or kept in a second register: each lane
// carries only its own column's window, which roughly halves the registers
// of the other march and raises the wavefronts per SIMD. The new ssh of
// T(i+1) comes from lane l+1, which computes it for its own column (the
// lane after the last owner is the last one that needs it, and its stencil
// ends one lane further).
//
// A tile is one strip (OW output columns, 64 lanes) of one band (R rows) of
// one domain; the caller picks the tile and passes the arrays, the element
// offset `base` of the domain's point (0,0) in them, its pitch W and its
// internal bounds. Element offsets times the element size must stay below
// 2^32 (every access is a 32-bit byte offset from a uniform base).
#pragma once

#include "nemo_mixed.h"
#include "nemo_step.h"

namespace nemo {
namespace wave {

constexpr int OUT = 61;

__device__ __forceinline__ bool in_rect(int i, int j, int x0, int x1, int y0,
                                        int y1)
{
    return i >= x0 && i <= x1 && j >= y0 && j <= y1;
}

// Neighbour exchange across the wavefront by DPP wave shifts, a modifier of
// a VALU move (no LDS traffic, no wait): from_below gives lane i the value
// of lane i-1 (wave_shr:1), from_above that of lane i+1 (wave_shl:1). Lane 0
// (resp. 63) receives 0; those lanes never own a column, and no owned value
// depends on what they receive. Called in wavefront-uniform control flow.
__device__ __forceinline__ int from_below(int x)
{
    return __builtin_amdgcn_update_dpp(0, x, 0x138, 0xf, 0xf, true);
}
__device__ __forceinline__ int from_above(int x)
{
    return __builtin_amdgcn_update_dpp(0, x, 0x130, 0xf, 0xf, true);
}
__device__ __forceinline__ float from_below(float x)
{
    return __builtin_bit_cast(float, from_below(__builtin_bit_cast(int, x)));
}
__device__ __forceinline__ float from_above(float x)
{
    return __builtin_bit_cast(float, from_above(__builtin_bit_cast(int, x)));
}
__device__ __forceinline__ double from_below(double x)
{
    const unsigned long long u = __builtin_bit_cast(unsigned long long, x);
    const unsigned lo = (unsigned)from_below((int)(unsigned)u);
    const unsigned hi = (unsigned)from_below((int)(unsigned)(u >> 32));
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}
__device__ __forceinline__ double from_above(double x)
{
    const unsigned long long u = __builtin_bit_cast(unsigned long long, x);
    const unsigned lo = (unsigned)from_above((int)(unsigned)u);
    const unsigned hi = (unsigned)from_above((int)(unsigned)(u >> 32));
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}

// Strips and bands of the extent the fused step walks: columns xs-1 .. xe+1,
// rows ys-1 .. ye+1.
__host__ __device__ inline int strips(const Bounds &b, int ow)
{
    return (b.xe - b.xs + 3 + ow - 1) / ow;
}
__host__ __device__ inline int bands(const Bounds &b, int R) { return (b.ye - b.ys + 3 + R - 1) / R; }

// Block b of a launch of nblk blocks -> logical tile, so that each of the 8
// XCDs (the dispatcher hands block b to XCD b % 8) owns one contiguous range
// of tiles, strips fastest.
__device__ __forceinline__ int xcd_tile(int bid, int nblk)
{
    const int xcd = bid & 7, k = bid >> 3;
    const int per = nblk >> 3, extra = nblk & 7;
    return xcd * per + min(xcd, extra) + k;
}

// One FP64 tile: new ssh (continuity + bc_ssh), momentum + bc_solid for U
// and V, next_sshu/v, stored where the lane owns its column. The wavefront
// owns OW columns, lanes L0 .. L0+OW-1 (L0 = (64 - OW) / 2); the lanes on
// either side load the neighbour columns (the next_ssh of the last owned
// column needs the new ssh of the column after it, hence two lanes above).
// OW = 56 with L0 = 4 starts every strip's stores on a 64-byte boundary
// (56 doubles = 448 bytes), at the price of 8 shared columns instead of 3.
//
// PF = 0: the loads of row j+1 are issued at the top of row j, which needs
// them, so a wavefront waits for them and only other wavefronts overlap
// that wait. PF = k > 0: the window holds k more rows and the loads of row
// j+1+k (and the row-j+k values of the per-point arrays) are issued at the
// top of row j, k rows of work before use, behind a scheduling barrier that
// keeps the compiler from moving them next to their use; the band loop is then fully
// unrolled (bands up to 16 rows) so that the window shifts are register
// renaming, never copies of values still in flight. Longer runs of rows are
// unrolled by U; the shifts at the loop's back edge then wait once per U
// rows for the loads in flight.
//
// tile_f64_rows walks rows jlo .. jhi-1 (at most R of them) of one strip;
// tile_f64 is the band form, rows ys-1 + band*R onwards.
template <int R, int OW = OUT, int PF = 0,
          int U = (PF >= 1 ? (R <= 16 ? R : 8) : (R <= 8 ? R : 1))>
__device__ __forceinline__ void tile_f64_rows(const StepIn<double> &f,
                                              const StepScalars<double> &p,
                                              const StepOut<double> &o, int W,
                                              unsigned base, const Bounds &b, int strip,
                                              int jlo, int jhi, double tide)
{
    constexpr int L0 = (64 - OW) / 2;
    static_assert(L0 >= 1 && L0 + OW + 2 <= 64, "one neighbour lane below, two above");
    static_assert(PF >= 0 && PF <= 2, "prefetch depth 0, 1 or 2");
    using T = double;
    const int xmax = b.xe + 1, ymax = b.ye + 1;
    const int lane = (int)(threadIdx.x & 63);  // lane within the wavefront
    const int i = b.xs - 1 + strip * OW + lane - L0;  // may be < 0 or > xmax
    const bool owner = lane >= L0 && lane < L0 + OW && i <= xmax;
    const int ci = min(max(i, 0), xmax);                 // clamped load column
    auto rowoff = [&](int j) { return base + (unsigned)(min(max(j, 0), ymax) * W); };

    // own-column window: 3-slot rows j-1, j, j+1; 2-slot rows j, j+1;
    // tm rows j-1 .. j+2; with PF = 1 one more row of each in flight, and
    // the per-point values of rows j and j+1
    T un[3 + PF], vn[3 + PF], hu[3 + PF], su[3 + PF], hv[3 + PF], sv[3 + PF],
        e1v[3 + PF], e2u[3 + PF];
    T ht[2 + PF], st[2 + PF], e1t[2 + PF], e2t[2 + PF], at[2 + PF];
    int tm[4 + PF];
    T dxu[1 + PF], dyv[1 + PF], fcu[1 + PF], fcv[1 + PF], aru[1 + PF], arv[1 + PF],
        sau[1 + PF], sav[1 + PF];

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
        st[s] = ld32(f.sshn_t, c);
        e1t[s] = ld32(f.dx_t, c);
        e2t[s] = ld32(f.dy_t, c);
        at[s] = ld32(f.area_t, c);
    };
    // new ssh of T(i, r): 3-slot s (row r) and s-1 (row r-1), 2-slot s2,
    // tm slots t-1..t+1; every lane evaluates it for its own column
    auto new_ssh = [&](int s, int s2, int t) {
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
        const T x = continuity_val(v, p.rdt);
        return tide_applies(tm[t], tm[t - 1], tm[t + 1], m_e, m_w) ? tide : x;
    };

    // per-point values of row r into slot s
    auto load1 = [&](int r, int s) {
        const unsigned c = rowoff(r) + ci;
        dxu[s] = ld32(f.dx_u, c);
        dyv[s] = ld32(f.dy_v, c);
        fcu[s] = ld32(f.fcor_u, c);
        fcv[s] = ld32(f.fcor_v, c);
        aru[s] = ld32(f.area_u, c);
        arv[s] = ld32(f.area_v, c);
        sau[s] = ld32(f.ssha_u, c);
        sav[s] = ld32(f.ssha_v, c);
    };

    // ---- prologue ----
    load3(jlo - 1, 0);
    load3(jlo, 1);
    load2(jlo, 0);
    tm[0] = ld32(f.tmask, rowoff(jlo - 1) + ci);
    tm[1] = ld32(f.tmask, rowoff(jlo) + ci);
    tm[2] = ld32(f.tmask, rowoff(jlo + 1) + ci);
    // with PF = k, rows up to jlo+k are in flight before the first row
#pragma unroll
    for (int k = 1; k <= PF; k++) {
        load3(jlo + k, 1 + k);
        load2(jlo + k, k);
        tm[2 + k] = ld32(f.tmask, rowoff(jlo + 1 + k) + ci);
        load1(jlo + k - 1, k - 1);
    }
    T ssh_c = new_ssh(1, 0, 1);  // T(i, jlo)

    // the row loop is unrolled by U, so the window shifts inside an unrolled
    // group become register renaming instead of moves
#pragma unroll U
    for (int jj = 0; jj < R; jj++) {
        const int j = jlo + jj;
        if (j >= jhi) break;  // wavefront-uniform
        if constexpr (PF >= 1) {
            if (jj + PF < R && j + PF < jhi) {  // rows the run still needs, PF+1 ahead
                load3(j + 1 + PF, 2 + PF);
                load2(j + 1 + PF, 1 + PF);
                tm[3 + PF] = ld32(f.tmask, rowoff(j + 2 + PF) + ci);
                load1(j + PF, PF);
            }
            __builtin_amdgcn_sched_barrier(0);
        } else {
            load3(j + 1, 2);
            load2(j + 1, 1);
            tm[3] = ld32(f.tmask, rowoff(j + 2) + ci);
            load1(j, 0);
        }
        const unsigned c = rowoff(j) + ci;
        const T dxu0 = dxu[0], dyv0 = dyv[0], fcu0 = fcu[0], fcv0 = fcv[0];
        const T aru0 = aru[0], arv0 = arv[0], sau0 = sau[0], sav0 = sav[0];

        const T ssh_n = new_ssh(2, 1, 2);   // T(i, j+1)
        const T ssh_e = from_above(ssh_c);  // T(i+1, j)

        // Every lane runs every phase (the permutes are wavefront-wide);
        // only the stores depend on the lane owning its column. Each phase
        // takes its neighbour values right before use, which keeps them
        // short-lived.
        const bool own_t = owner && in_rect(i, j, b.xs, b.xe, b.ys, b.ye);
        if (own_t) st32(o.ssh, c, ssh_c);

        const int tm_e0 = from_above(tm[1]);
        const int tm_e1 = from_above(tm[2]);
        {  // U(i,j): momentum_u, else bc_solid_u
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
            v.e1u_c = dxu0;
            v.fcor = fcu0;
            v.area = aru0;
            v.ssha = sau0;
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
        {  // V(i,j): momentum_v, else bc_solid_v
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
            v.e2v_c = dyv0;
            v.fcor = fcv0;
            v.area = arv0;
            v.ssha = sav0;
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
        {  // next_sshu(i,j), next_sshv(i,j)
            const T at_e0 = from_above(at[0]);
            T val;
            if (owner && in_rect(i, j, b.xs, b.xe - 1, b.ys, b.ye) &&
                next_ssh_pt(ssh_c, ssh_e, at[0], at_e0, aru0, tm[1], tm_e0, &val))
                st32(o.sshu, c, val);
            if (owner && in_rect(i, j, b.xs, b.xe, b.ys, b.ye - 1) &&
                next_ssh_pt(ssh_c, ssh_n, at[0], at[1], arv0, tm[1], tm[2], &val))
                st32(o.sshv, c, val);
        }

#define NEMO_SHIFT(a, n)                    \
    _Pragma("unroll") for (int k = 0; k + 1 < (n); k++) a[k] = a[k + 1];
        NEMO_SHIFT(un, 3 + PF) NEMO_SHIFT(vn, 3 + PF) NEMO_SHIFT(hu, 3 + PF)
        NEMO_SHIFT(su, 3 + PF) NEMO_SHIFT(hv, 3 + PF) NEMO_SHIFT(sv, 3 + PF)
        NEMO_SHIFT(e1v, 3 + PF) NEMO_SHIFT(e2u, 3 + PF)
        NEMO_SHIFT(ht, 2 + PF) NEMO_SHIFT(st, 2 + PF) NEMO_SHIFT(e1t, 2 + PF)
        NEMO_SHIFT(e2t, 2 + PF) NEMO_SHIFT(at, 2 + PF) NEMO_SHIFT(tm, 4 + PF)
        NEMO_SHIFT(dxu, 1 + PF) NEMO_SHIFT(dyv, 1 + PF) NEMO_SHIFT(fcu, 1 + PF)
        NEMO_SHIFT(fcv, 1 + PF) NEMO_SHIFT(aru, 1 + PF) NEMO_SHIFT(arv, 1 + PF)
        NEMO_SHIFT(sau, 1 + PF) NEMO_SHIFT(sav, 1 + PF)
#undef NEMO_SHIFT
        ssh_c = ssh_n;
    }
}

template <int R, int OW = OUT, int PF = 0>
__device__ __forceinline__ void tile_f64(const StepIn<double> &f,
                                         const StepScalars<double> &p,
                                         const StepOut<double> &o, int W,
                                         unsigned base,
                                         const Bounds &b, int strip, int band,
                                         double tide)
{
    const int jlo = b.ys - 1 + band * R;
    const int jhi = min(jlo + R, b.ye + 2);
    tile_f64_rows<R, OW, PF>(f, p, o, W, base, b, strip, jlo, jhi, tide);
}

struct Ssh {
    float hi, lo;
};

__device__ __forceinline__ Ssh split(double s)
{
    const float hi = (float)s;
    return Ssh{hi, (float)(s - (double)hi)};
}

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

// One mixed-precision tile: as tile_f64 over FP32 arrays, the new ssh as a
// compensated (hi, lo) pair accumulated in FP64.
// Owned columns as in tile_f64: lanes L0 .. L0+OW-1, L0 = (64 - OW) / 2
// (OW = 48 starts every strip's FP32 stores on a 64-byte boundary).
// PF as in tile_f64_rows: PF = k > 0 issues the loads of row j+1+k at the
// top of row j, behind a scheduling barrier, in a fully unrolled band loop.
template <int R, int OW = OUT, int PF = 0>
__device__ __forceinline__ void tile_mixed(const MixedIn &f,
                                           const StepScalars<float> &p,
                                           const MixedOut &o, int W,
                                           unsigned base,
                                           const Bounds &b, int strip, int band,
                                           Ssh tide)
{
    constexpr int L0 = (64 - OW) / 2;
    static_assert(L0 >= 1 && L0 + OW + 2 <= 64, "one neighbour lane below, two above");
    static_assert(PF >= 0 && PF <= 2, "prefetch depth 0, 1 or 2");
    static_assert(PF == 0 || R <= 16, "the prefetching band loop is fully unrolled");
    using T = float;
    const int xmax = b.xe + 1, ymax = b.ye + 1;
    const int lane = (int)(threadIdx.x & 63);  // lane within the wavefront
    const int i = b.xs - 1 + strip * OW + lane - L0;
    const bool owner = lane >= L0 && lane < L0 + OW && i <= xmax;
    const int ci = min(max(i, 0), xmax);
    const int jlo = b.ys - 1 + band * R;
    const int jhi = min(jlo + R, ymax + 1);
    auto rowoff = [&](int j) { return base + (unsigned)(min(max(j, 0), ymax) * W); };

    T un[3 + PF], vn[3 + PF], hu[3 + PF], su[3 + PF], hv[3 + PF], sv[3 + PF],
        e1v[3 + PF], e2u[3 + PF];
    T ht[2 + PF], st[2 + PF], e1t[2 + PF], e2t[2 + PF], at[2 + PF];
    T lo[1 + PF];  // lo words of T(i, j+1) ..
    int tm[4 + PF];
    T dxu[1 + PF], dyv[1 + PF], fcu[1 + PF], fcv[1 + PF], aru[1 + PF], arv[1 + PF],
        sau[1 + PF], sav[1 + PF];

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

    auto load1 = [&](int r, int s) {
        const unsigned c = rowoff(r) + ci;
        dxu[s] = ld32(f.dx_u, c);
        dyv[s] = ld32(f.dy_v, c);
        fcu[s] = ld32(f.fcor_u, c);
        fcv[s] = ld32(f.fcor_v, c);
        aru[s] = ld32(f.area_u, c);
        arv[s] = ld32(f.area_v, c);
        sau[s] = ld32(f.ssha_u, c);
        sav[s] = ld32(f.ssha_v, c);
    };

    load3(jlo - 1, 0);
    load3(jlo, 1);
    load2(jlo, 0);
    tm[0] = ld32(f.tmask, rowoff(jlo - 1) + ci);
    tm[1] = ld32(f.tmask, rowoff(jlo) + ci);
    tm[2] = ld32(f.tmask, rowoff(jlo + 1) + ci);
#pragma unroll
    for (int k = 1; k <= PF; k++) {
        load3(jlo + k, 1 + k);
        load2(jlo + k, k);
        tm[2 + k] = ld32(f.tmask, rowoff(jlo + 1 + k) + ci);
        lo[k - 1] = ld32(f.ssh_lo, rowoff(jlo + k) + ci);
        load1(jlo + k - 1, k - 1);
    }
    Ssh ssh_c = new_ssh(1, 0, 1, ld32(f.ssh_lo, rowoff(jlo) + ci));
    // neighbour values of window slots that the previous row already shifted
    // (slot k of row j is slot k+1 of row j-1): carried, not shifted again
    int c_tm0e = from_above(tm[0]), c_tm1e = from_above(tm[1]), c_tm1w = from_below(tm[1]);
    T c_un1w = from_below(un[1]), c_vn0e = from_above(vn[0]), c_e1v0e = from_above(e1v[0]);
    T c_hv0e = from_above(hv[0]), c_sv0e = from_above(sv[0]), c_e2u1w = from_below(e2u[1]);
    T c_hu1w = from_below(hu[1]), c_su1w = from_below(su[1]);

    // short bands: the row loop is unrolled, so the window shifts at its
    // end become register renaming instead of moves
    constexpr int UNROLL = (PF >= 1 || R <= 8) ? R : 1;
#pragma unroll UNROLL
    for (int jj = 0; jj < R; jj++) {
        const int j = jlo + jj;
        if (j >= jhi) break;  // wavefront-uniform
        if constexpr (PF >= 1) {
            if (jj + PF < R && j + PF < jhi) {  // rows the band still needs
                load3(j + 1 + PF, 2 + PF);
                load2(j + 1 + PF, 1 + PF);
                lo[PF] = ld32(f.ssh_lo, rowoff(j + 1 + PF) + ci);
                tm[3 + PF] = ld32(f.tmask, rowoff(j + 2 + PF) + ci);
                load1(j + PF, PF);
            }
            __builtin_amdgcn_sched_barrier(0);
        } else {
            load3(j + 1, 2);
            load2(j + 1, 1);
            lo[0] = ld32(f.ssh_lo, rowoff(j + 1) + ci);
            tm[3] = ld32(f.tmask, rowoff(j + 2) + ci);
            load1(j, 0);
        }
        const unsigned c = rowoff(j) + ci;
        const T dxu0 = dxu[0], dyv0 = dyv[0], fcu0 = fcu[0], fcv0 = fcv[0];
        const T aru0 = aru[0], arv0 = arv[0], sau0 = sau[0], sav0 = sav[0];

        const Ssh ssh_n = new_ssh(2, 1, 2, lo[0]);
        const float ssh_e = from_above(ssh_c.hi);

        if (owner && in_rect(i, j, b.xs, b.xe, b.ys, b.ye)) {
            st32(o.ssh_hi, c, ssh_c.hi);
            st32(o.ssh_lo, c, ssh_c.lo);
        }
        const int tm_e0 = c_tm1e;
        const int tm_e1 = from_above(tm[2]), tm2w = from_below(tm[2]);
        const T vn1e = from_above(vn[1]), e1v1e = from_above(e1v[1]);
        const T hv1e = from_above(hv[1]), sv1e = from_above(sv[1]);
        const T un2w = from_below(un[2]), e2u2w = from_below(e2u[2]);
        const T hu2w = from_below(hu[2]), su2w = from_below(su[2]);
        {
            MomUVals<T> v;
            v.m_c = tm[1];
            v.m_e = tm_e0;
            v.m_s = tm[0];
            v.m_se = c_tm0e;
            v.m_n = tm[2];
            v.m_ne = tm_e1;
            v.un_c = un[1];
            v.un_e = from_above(un[1]);
            v.un_w = c_un1w;
            v.un_s = un[0];
            v.un_n = un[2];
            v.vn_s = vn[0];
            v.vn_se = c_vn0e;
            v.vn_c = vn[1];
            v.vn_e = vn1e;
            v.e1v_s = e1v[0];
            v.e1v_se = c_e1v0e;
            v.e1v_c = e1v[1];
            v.e1v_e = e1v1e;
            v.hv_s = hv[0];
            v.sv_s = sv[0];
            v.hv_se = c_hv0e;
            v.sv_se = c_sv0e;
            v.hv_c = hv[1];
            v.sv_c = sv[1];
            v.hv_e = hv1e;
            v.sv_e = sv1e;
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
            v.e1u_c = dxu0;
            v.fcor = fcu0;
            v.area = aru0;
            v.ssha = sau0;
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
            v.m_w = c_tm1w;
            v.m_nw = tm2w;
            v.m_ne = tm_e1;
            v.vn_c = vn[1];
            v.vn_n = vn[2];
            v.vn_s = vn[0];
            v.vn_w = from_below(vn[1]);
            v.vn_e = vn1e;
            v.un_w = c_un1w;
            v.un_nw = un2w;
            v.un_c = un[1];
            v.un_n = un[2];
            v.e2u_w = c_e2u1w;
            v.e2u_nw = e2u2w;
            v.e2u_c = e2u[1];
            v.e2u_n = e2u[2];
            v.hu_w = c_hu1w;
            v.su_w = c_su1w;
            v.hu_nw = hu2w;
            v.su_nw = su2w;
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
            v.e1v_e = e1v1e;
            v.hv_c = hv[1];
            v.sv_c = sv[1];
            v.hv_w = from_below(hv[1]);
            v.sv_w = from_below(sv[1]);
            v.hv_e = hv1e;
            v.sv_e = sv1e;
            v.e2v_c = dyv0;
            v.fcor = fcv0;
            v.area = arv0;
            v.ssha = sav0;
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
                next_ssh_pt(ssh_c.hi, ssh_e, at[0], at_e0, aru0, tm[1], tm_e0, &val))
                st32(o.sshu, c, val);
            if (owner && in_rect(i, j, b.xs, b.xe, b.ys, b.ye - 1) &&
                next_ssh_pt(ssh_c.hi, ssh_n.hi, at[0], at[1], arv0, tm[1], tm[2], &val))
                st32(o.sshv, c, val);
        }

#define NEMO_SHIFT(a, n)                    \
    _Pragma("unroll") for (int k = 0; k + 1 < (n); k++) a[k] = a[k + 1];
        NEMO_SHIFT(un, 3 + PF) NEMO_SHIFT(vn, 3 + PF) NEMO_SHIFT(hu, 3 + PF)
        NEMO_SHIFT(su, 3 + PF) NEMO_SHIFT(hv, 3 + PF) NEMO_SHIFT(sv, 3 + PF)
        NEMO_SHIFT(e1v, 3 + PF) NEMO_SHIFT(e2u, 3 + PF)
        NEMO_SHIFT(ht, 2 + PF) NEMO_SHIFT(st, 2 + PF) NEMO_SHIFT(e1t, 2 + PF)
        NEMO_SHIFT(e2t, 2 + PF) NEMO_SHIFT(at, 2 + PF) NEMO_SHIFT(tm, 4 + PF)
        NEMO_SHIFT(lo, 1 + PF)
        NEMO_SHIFT(dxu, 1 + PF) NEMO_SHIFT(dyv, 1 + PF) NEMO_SHIFT(fcu, 1 + PF)
        NEMO_SHIFT(fcv, 1 + PF) NEMO_SHIFT(aru, 1 + PF) NEMO_SHIFT(arv, 1 + PF)
        NEMO_SHIFT(sau, 1 + PF) NEMO_SHIFT(sav, 1 + PF)
#undef NEMO_SHIFT
        ssh_c = ssh_n;
        c_tm0e = tm_e0;
        c_tm1e = tm_e1;
        c_tm1w = tm2w;
        c_un1w = un2w;
        c_vn0e = vn1e;
        c_e1v0e = e1v1e;
        c_hv0e = hv1e;
        c_sv0e = sv1e;
        c_e2u1w = e2u2w;
        c_hu1w = hu2w;
        c_su1w = su2w;
    }
}

}  // namespace wave
}  // namespace nemo
