/*
 * TEST INFRASTRUCTURE ONLY -- the oracle. A plain-C CPU restatement of the
 * reference NEMOLite2D time step, used by tests/ as the checker and by
 * bench.py as the timed CPU baseline. Nothing in the product links it.
 *
 * Every per-point function below evaluates its expressions in the same order
 * as the reference kernel it cites, so that it reproduces the reference
 * bit-for-bit when both are built with -ffp-contract=off (no FMA
 * contraction). Loop bounds and sweep order follow
 * manual_versions/psykal_cpp/time_step.cpp:53-136.
 *
 * Index convention (dl_esm_inf layout): a field is a W x H block, point
 * (ji, jj) at idx = jj*W + ji, 0-based; U(i,j) is the east face and V(i,j)
 * the north face of T(i,j).
 *
 * Parity pins (tests/test_oracle_*.py): bitwise against the reference's own
 * time_step.cpp built into oracle/_ref when /root/reference is present, and
 * against the committed golden fixtures in tests/golden/ everywhere; closed
 * forms (rest state, mass conservation, Flather/next_ssh special cases) for
 * what those leave open.
 */
#include <math.h>
#include <stddef.h>

#include "nemo_oracle.h"

/* Reference literals, boundary_conditions_kern.c:185-186. */
#define REF_AMP_TIDE 0.2
#define REF_OMEGA_TIDE (2.0 * 3.14159 / (12.42 * 3600.0))

static double g_amp_tide = REF_AMP_TIDE;
static double g_omega_tide = REF_OMEGA_TIDE;

void oracle_set_tide(double amp, double omega_tide)
{
    g_amp_tide = amp;
    g_omega_tide = omega_tide;
}

void oracle_reset_tide(void)
{
    g_amp_tide = REF_AMP_TIDE;
    g_omega_tide = REF_OMEGA_TIDE;
}

/* 0.5 +/- copysign(0.5, x): the upwind selectors of the advection terms. */
static inline double up_plus(double x) { return 0.5 + copysign(0.5, x); }
static inline double up_minus(double x) { return 0.5 - copysign(0.5, x); }

/* ---- continuity: KC/continuity_kern.c:90-117 ---------------------------- */
static inline void pt_continuity(int c, int W, double *ssha, const double *sshn,
                                 const double *sshn_u, const double *sshn_v,
                                 const double *hu, const double *hv,
                                 const double *un, const double *vn, double rdt,
                                 const double *e12t)
{
    const int w = c - 1, s = c - W;
    /* volume fluxes through the east, west, north and south faces */
    double f_e = (sshn_u[c] + hu[c]) * un[c];
    double f_w = (sshn_u[w] + hu[w]) * un[w];
    double f_n = (sshn_v[c] + hv[c]) * vn[c];
    double f_s = (sshn_v[s] + hv[s]) * vn[s];
    ssha[c] = sshn[c] + (f_w - f_e + f_s - f_n) * rdt / e12t[c];
}

/* ---- momentum_u: KC/momentum_u_kern.c:150-263 --------------------------- */
static inline void pt_momentum_u(int c, int W, double *ua, const double *un,
                                 const double *vn, const double *hu,
                                 const double *hv, const double *ht,
                                 const double *ssha_u, const double *sshn,
                                 const double *sshn_u, const double *sshn_v,
                                 const int *tm, const double *e1u,
                                 const double *e1v, const double *e1t,
                                 const double *e2u, const double *e2t,
                                 const double *e12u, const double *gphiu,
                                 double rdt, double cbfr, double visc,
                                 double omega, double d2r, double g)
{
    const int e = c + 1, w = c - 1, s = c - W, n = c + W, se = s + 1, ne = n + 1;

    /* only faces with wet T cells on both sides carry momentum */
    if (tm[c] + tm[e] <= 0) return;
    if (tm[c] <= 0 || tm[e] <= 0) return;

    /* transports at the four faces of the u control volume */
    double u_e = 0.5 * (un[c] + un[e]) * e2t[e];
    double dep_e = ht[e] + sshn[e];
    double u_w = 0.5 * (un[c] + un[w]) * e2t[c];
    double dep_w = ht[c] + sshn[c];

    double vs_c = 0.5 * (vn[s] + vn[se]);
    double v_s = 0.5 * vs_c * (e1v[s] + e1v[se]);
    double dep_s = 0.5 * (hv[s] + sshn_v[s] + hv[se] + sshn_v[se]);

    double vn_c = 0.5 * (vn[c] + vn[e]);
    double v_n = 0.5 * vn_c * (e1v[c] + e1v[e]);
    double dep_n = 0.5 * (hv[c] + sshn_v[c] + hv[e] + sshn_v[e]);

    /* first-order upwind advection; a dry south/north neighbour drops the
     * off-face term */
    double q_w = up_minus(u_w) * un[c] + up_plus(u_w) * un[w];
    double q_e = up_plus(u_e) * un[c] + up_minus(u_e) * un[e];
    int dry_s = (tm[s] <= 0 || tm[se] <= 0);
    int dry_n = (tm[n] <= 0 || tm[ne] <= 0);
    double q_s = dry_s ? up_minus(v_s) * un[c]
                       : up_minus(v_s) * un[c] + up_plus(v_s) * un[s];
    double q_n = dry_n ? up_plus(v_n) * un[c]
                       : up_plus(v_n) * un[c] + up_minus(v_n) * un[n];
    double adv = q_w * u_w * dep_w - q_e * u_e * dep_e + q_s * v_s * dep_s -
                 q_n * v_n * dep_n;

    /* Laplacian viscosity with slip walls */
    double gx_e = (un[e] - un[c]) / e1t[e] * (ht[e] + sshn[e]);
    double gx_w = (un[c] - un[w]) / e1t[c] * (ht[c] + sshn[c]);
    double gy_s = dry_s ? 0.0
                        : (un[c] - un[s]) / (e2u[c] + e2u[s]) *
                              (hu[c] + sshn_u[c] + hu[s] + sshn_u[s]);
    double gy_n = dry_n ? 0.0
                        : (un[n] - un[c]) / (e2u[c] + e2u[n]) *
                              (hu[c] + sshn_u[c] + hu[n] + sshn_u[n]);
    double vis = (gx_e - gx_w) * e2u[c] + (gy_n - gy_s) * e1u[c] * 0.5;
    vis = visc * vis;

    /* Coriolis and pressure gradient */
    double cor = 0.5 * (2. * omega * sin(gphiu[c] * d2r) * (vs_c + vn_c)) *
                 e12u[c] * (hu[c] + sshn_u[c]);
    double hpg = -g * (hu[c] + sshn_u[c]) * e2u[c] * (sshn[e] - sshn[c]);

    /* implicit linear bottom friction */
    ua[c] = (un[c] * (hu[c] + sshn_u[c]) +
             rdt * (adv + vis + cor + hpg) / e12u[c]) /
            (hu[c] + ssha_u[c]) / (1.0 + cbfr * rdt);
}

/* ---- momentum_v: KC/momentum_v_kern.c:125-242 --------------------------- */
static inline void pt_momentum_v(int c, int W, double *va, const double *un,
                                 const double *vn, const double *hu,
                                 const double *hv, const double *ht,
                                 const double *ssha_v, const double *sshn,
                                 const double *sshn_u, const double *sshn_v,
                                 const int *tm, const double *e1v,
                                 const double *e1t, const double *e2u,
                                 const double *e2v, const double *e2t,
                                 const double *e12v, const double *gphiv,
                                 double rdt, double cbfr, double visc,
                                 double omega, double d2r, double g)
{
    const int e = c + 1, w = c - 1, s = c - W, n = c + W, nw = n - 1, ne = n + 1;

    /* NB: the first test pairs T(i) with T(i+1), not T(j+1); kept as in
     * momentum_v_kern.c:155 */
    if (tm[c] + tm[e] <= 0) return;
    if (tm[c] <= 0 || tm[n] <= 0) return;

    double v_n = 0.5 * (vn[c] + vn[n]) * e1t[n];
    double dep_n = ht[n] + sshn[n];
    double v_s = 0.5 * (vn[c] + vn[s]) * e1t[c];
    double dep_s = ht[c] + sshn[c];

    double uw_c = 0.5 * (un[w] + un[nw]);
    double u_w = 0.5 * uw_c * (e2u[w] + e2u[nw]);
    double dep_w = 0.50 * (hu[w] + sshn_u[w] + hu[nw] + sshn_u[nw]);

    double ue_c = 0.5 * (un[c] + un[n]);
    double u_e = 0.5 * ue_c * (e2u[c] + e2u[n]);
    double dep_e = 0.50 * (hu[c] + sshn_u[c] + hu[n] + sshn_u[n]);

    double q_s = up_minus(v_s) * vn[c] + up_plus(v_s) * vn[s];
    double q_n = up_plus(v_n) * vn[c] + up_minus(v_n) * vn[n];
    int dry_w = (tm[w] <= 0 || tm[nw] <= 0);
    int dry_e = (tm[e] <= 0 || tm[ne] <= 0);
    double q_w = dry_w ? up_minus(u_w) * vn[c]
                       : up_minus(u_w) * vn[c] + up_plus(u_w) * vn[w];
    double q_e = dry_e ? up_plus(u_e) * vn[c]
                       : up_plus(u_e) * vn[c] + up_minus(u_e) * vn[e];
    double adv = q_w * u_w * dep_w - q_e * u_e * dep_e + q_s * v_s * dep_s -
                 q_n * v_n * dep_n;

    double gy_n = (vn[n] - vn[c]) / e2t[n] * (ht[n] + sshn[n]);
    double gy_s = (vn[c] - vn[s]) / e2t[c] * (ht[c] + sshn[c]);
    double gx_w = dry_w ? 0.0
                        : (vn[c] - vn[w]) / (e1v[c] + e1v[w]) *
                              (hv[c] + sshn_v[c] + hv[w] + sshn_v[w]);
    double gx_e = dry_e ? 0.0
                        : (vn[e] - vn[c]) / (e1v[c] + e1v[e]) *
                              (hv[c] + sshn_v[c] + hv[e] + sshn_v[e]);
    double vis = (gy_n - gy_s) * e1v[c] + (gx_e - gx_w) * e2v[c] * 0.5;
    vis = visc * vis;

    double cor = -0.5 * (2. * omega * sin(gphiv[c] * d2r) * (ue_c + uw_c)) *
                 e12v[c] * (hv[c] + sshn_v[c]);
    double hpg = -g * (hv[c] + sshn_v[c]) * e1v[c] * (sshn[n] - sshn[c]);

    va[c] = (vn[c] * (hv[c] + sshn_v[c]) +
             rdt * (adv + vis + cor + hpg) / e12v[c]) /
            (hv[c] + ssha_v[c]) / (1.0 + cbfr * rdt);
}

/* ---- bc_ssh: KC/boundary_conditions_kern.c:178-203 ---------------------- */
static inline void pt_bc_ssh(int c, int W, double tide, double *ssha,
                             const int *tm)
{
    if (tm[c] <= 0) return;
    /* a wet cell next to an open (-1) cell takes the tidal elevation */
    if (tm[c - W] < 0 || tm[c + W] < 0 || tm[c + 1] < 0 || tm[c - 1] < 0)
        ssha[c] = tide;
}

/* ---- bc_solid_u / bc_solid_v: :220-228, :244-252 ------------------------ */
static inline void pt_solid(int c, int stride, double *vel, const int *tm)
{
    if (tm[c] * tm[c + stride] == 0) vel[c] = 0.0;
}

/* ---- bc_flather_u / bc_flather_v: :271-293, :313-333 --------------------
 * `stride` is 1 for u (neighbours along i) and W for v (along j). */
static inline void pt_flather(int c, int stride, double *vel, const double *h,
                              const double *ssh, const int *tm, double g)
{
    const int p = c + stride, m = c - stride;
    if (tm[c] + tm[p] <= -1) return;
    if (tm[c] < 0)
        vel[c] = vel[p] + sqrt(g / h[c]) * (ssh[c] - ssh[p]);
    else if (tm[p] < 0)
        vel[c] = vel[m] + sqrt(g / h[c]) * (ssh[c] - ssh[m]);
}

/* ---- next_sshu / next_sshv: KC/time_update_kern.c:90-114, 132-155 -------- */
static inline void pt_next_ssh(int c, int stride, double *ssh_face,
                               const double *sshn, const int *tm,
                               const double *e12t, const double *e12f)
{
    const int p = c + stride;
    if (tm[c] + tm[p] <= 0) return;
    if (tm[c] * tm[p] > 0) {
        double vol = e12t[c] * sshn[c] + e12t[p] * sshn[p];
        ssh_face[c] = 0.5 * vol / e12f[c];
    } else if (tm[c] <= 0) {
        ssh_face[c] = sshn[p];
    } else if (tm[p] <= 0) {
        ssh_face[c] = sshn[c];
    }
}

/* ---- exported sweeps ---------------------------------------------------- */
#define FOR_REGION(x0, x1, y0, y1)                                             \
    for (int jj = (y0); jj <= (y1); jj++)                                      \
        for (int ji = (x0); ji <= (x1); ji++)

void oracle_continuity(int x0, int x1, int y0, int y1, int W, double *ssha,
                       const double *sshn, const double *sshn_u,
                       const double *sshn_v, const double *hu, const double *hv,
                       const double *un, const double *vn, double rdt,
                       const double *e12t)
{
    FOR_REGION(x0, x1, y0, y1)
    pt_continuity(jj * W + ji, W, ssha, sshn, sshn_u, sshn_v, hu, hv, un, vn,
                  rdt, e12t);
}

void oracle_momentum_u(int x0, int x1, int y0, int y1, int W, double *ua,
                       const double *un, const double *vn, const double *hu,
                       const double *hv, const double *ht, const double *ssha_u,
                       const double *sshn, const double *sshn_u,
                       const double *sshn_v, const int *tmask, const double *e1u,
                       const double *e1v, const double *e1t, const double *e2u,
                       const double *e2t, const double *e12u,
                       const double *gphiu, double rdt, double cbfr,
                       double visc, double omega, double d2r, double g)
{
    FOR_REGION(x0, x1, y0, y1)
    pt_momentum_u(jj * W + ji, W, ua, un, vn, hu, hv, ht, ssha_u, sshn, sshn_u,
                  sshn_v, tmask, e1u, e1v, e1t, e2u, e2t, e12u, gphiu, rdt,
                  cbfr, visc, omega, d2r, g);
}

void oracle_momentum_v(int x0, int x1, int y0, int y1, int W, double *va,
                       const double *un, const double *vn, const double *hu,
                       const double *hv, const double *ht, const double *ssha_v,
                       const double *sshn, const double *sshn_u,
                       const double *sshn_v, const int *tmask, const double *e1v,
                       const double *e1t, const double *e2u, const double *e2v,
                       const double *e2t, const double *e12v,
                       const double *gphiv, double rdt, double cbfr,
                       double visc, double omega, double d2r, double g)
{
    FOR_REGION(x0, x1, y0, y1)
    pt_momentum_v(jj * W + ji, W, va, un, vn, hu, hv, ht, ssha_v, sshn, sshn_u,
                  sshn_v, tmask, e1v, e1t, e2u, e2v, e2t, e12v, gphiv, rdt,
                  cbfr, visc, omega, d2r, g);
}

static double tide_value(int istep, double rdt)
{
    double rtime = istep * rdt;
    return g_amp_tide * sin(g_omega_tide * rtime);
}

void oracle_bc_ssh(int x0, int x1, int y0, int y1, int W, int istep,
                   double *ssha, const int *tmask, double rdt)
{
    double tide = tide_value(istep, rdt);
    FOR_REGION(x0, x1, y0, y1)
    pt_bc_ssh(jj * W + ji, W, tide, ssha, tmask);
}

void oracle_bc_solid_u(int x0, int x1, int y0, int y1, int W, double *ua,
                       const int *tmask)
{
    FOR_REGION(x0, x1, y0, y1)
    pt_solid(jj * W + ji, 1, ua, tmask);
}

void oracle_bc_solid_v(int x0, int x1, int y0, int y1, int W, double *va,
                       const int *tmask)
{
    FOR_REGION(x0, x1, y0, y1)
    pt_solid(jj * W + ji, W, va, tmask);
}

void oracle_bc_flather_u(int x0, int x1, int y0, int y1, int W, double *ua,
                         const double *hu, const double *sshn_u,
                         const int *tmask, double g)
{
    FOR_REGION(x0, x1, y0, y1)
    pt_flather(jj * W + ji, 1, ua, hu, sshn_u, tmask, g);
}

void oracle_bc_flather_v(int x0, int x1, int y0, int y1, int W, double *va,
                         const double *hv, const double *sshn_v,
                         const int *tmask, double g)
{
    FOR_REGION(x0, x1, y0, y1)
    pt_flather(jj * W + ji, W, va, hv, sshn_v, tmask, g);
}

void oracle_next_sshu(int x0, int x1, int y0, int y1, int W, double *sshn_u,
                      const double *sshn, const int *tmask, const double *e12t,
                      const double *e12u)
{
    FOR_REGION(x0, x1, y0, y1)
    pt_next_ssh(jj * W + ji, 1, sshn_u, sshn, tmask, e12t, e12u);
}

void oracle_next_sshv(int x0, int x1, int y0, int y1, int W, double *sshn_v,
                      const double *sshn, const int *tmask, const double *e12t,
                      const double *e12v)
{
    FOR_REGION(x0, x1, y0, y1)
    pt_next_ssh(jj * W + ji, W, sshn_v, sshn, tmask, e12t, e12v);
}

double oracle_checksum(const double *x, int W, int x0, int x1, int y0, int y1)
{
    double s = 0.0;
    FOR_REGION(x0, x1, y0, y1)
    s += fabs(x[jj * W + ji]);
    return s;
}

/* ---- the step ----------------------------------------------------------- *
 * One body serves both the serial and the OpenMP restatement: PAR expands to
 * an OpenMP parallel-for over rows in the OpenMP build (time_step_omp.cpp
 * puts one on every sweep except bc_flather_v, :117-124), to nothing in the
 * serial one. Every sweep writes only its own point, so row-parallel order
 * does not change the result (flather_v aside, which stays serial). */
#define PRAGMA(x) _Pragma(#x)
#define STEP_BODY(PAR)                                                         \
    const int W = width;                                                       \
    const double tide = tide_value(istep, rdt);                                \
    PAR for (int jj = ys; jj <= ye; jj++) for (int ji = xs; ji <= xe; ji++)    \
        pt_continuity(jj * W + ji, W, ssha_t, sshn_t, sshn_u, sshn_v, hu, hv,  \
                      un, vn, rdt, area_t);                                    \
    PAR for (int jj = ys; jj <= ye; jj++) for (int ji = xs; ji <= xe - 1; ji++)\
        pt_momentum_u(jj * W + ji, W, ua, un, vn, hu, hv, ht, ssha_u, sshn_t,  \
                      sshn_u, sshn_v, tmask, dx_u, dx_v, dx_t, dy_u, dy_t,     \
                      area_u, gphiu, rdt, cbfr, visc, omega, d2r, g);          \
    PAR for (int jj = ys; jj <= ye - 1; jj++) for (int ji = xs; ji <= xe; ji++)\
        pt_momentum_v(jj * W + ji, W, va, un, vn, hu, hv, ht, ssha_v, sshn_t,  \
                      sshn_u, sshn_v, tmask, dx_v, dx_t, dy_u, dy_v, dy_t,     \
                      area_v, gphiv, rdt, cbfr, visc, omega, d2r, g);          \
    PAR for (int jj = ys; jj <= ye; jj++) for (int ji = xs; ji <= xe; ji++)    \
        pt_bc_ssh(jj * W + ji, W, tide, ssha_t, tmask);                        \
    PAR for (int jj = ys - 1; jj <= ye + 1; jj++)                              \
        for (int ji = xs - 1; ji <= xe; ji++)                                  \
            pt_solid(jj * W + ji, 1, ua, tmask);                               \
    PAR for (int jj = ys - 1; jj <= ye; jj++)                                  \
        for (int ji = xs - 1; ji <= xe + 1; ji++)                              \
            pt_solid(jj * W + ji, W, va, tmask);                               \
    PAR for (int jj = ys - 1; jj <= ye + 1; jj++)                              \
        for (int ji = xs - 1; ji <= xe; ji++)                                  \
            pt_flather(jj * W + ji, 1, ua, hu, sshn_u, tmask, g);              \
    for (int jj = ys - 1; jj <= ye; jj++)                                      \
        for (int ji = xs - 1; ji <= xe + 1; ji++)                              \
            pt_flather(jj * W + ji, W, va, hv, sshn_v, tmask, g);              \
    /* copy region: rows ys-1 .. ye (time_step.cpp:115 uses < ye+1) */         \
    PAR for (int jj = ys - 1; jj < ye + 1; jj++)                               \
        for (int ji = xs - 1; ji <= xe + 1; ji++) {                            \
            const int k = jj * W + ji;                                         \
            un[k] = ua[k];                                                     \
            vn[k] = va[k];                                                     \
            sshn_t[k] = ssha_t[k];                                             \
        }                                                                      \
    PAR for (int jj = ys; jj <= ye; jj++) for (int ji = xs; ji <= xe - 1; ji++)\
        pt_next_ssh(jj * W + ji, 1, sshn_u, sshn_t, tmask, area_t, area_u);    \
    PAR for (int jj = ys; jj <= ye - 1; jj++) for (int ji = xs; ji <= xe; ji++)\
        pt_next_ssh(jj * W + ji, W, sshn_v, sshn_t, tmask, area_t, area_v);

void oracle_invoke_time_step(ORACLE_STEP_ARGS)
{
    STEP_BODY()
}

void oracle_invoke_time_step_omp(ORACLE_STEP_ARGS)
{
    STEP_BODY(PRAGMA(omp parallel for))
}
