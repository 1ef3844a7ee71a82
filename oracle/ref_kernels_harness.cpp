/*
 * TEST INFRASTRUCTURE ONLY. Built by oracle/Makefile into oracle/_ref/ when
 * the reference tree is present, never shipped in the product.
 *
 * Exposes each reference per-point kernel (kernels/c_family/*.c, compiled
 * from where they lie in the reference tree, not copied) as a region sweep
 * with the same argument list as the oracle's oracle_<kernel>() functions,
 * so that tests can compare the oracle kernel by kernel against the
 * reference itself. The whole step comes from the reference's own
 * manual_versions/psykal_cpp/time_step.cpp, built separately.
 */
#include "continuity_kern.c"
#include "momentum_u_kern.c"
#include "momentum_v_kern.c"
#include "boundary_conditions_kern.c"
#include "time_update_kern.c"

#define SWEEP(x0, x1, y0, y1)                                                  \
    for (int jj = (y0); jj <= (y1); jj++)                                      \
        for (int ji = (x0); ji <= (x1); ji++)

#define D(p) const_cast<double *>(p)
#define I(p) const_cast<int *>(p)

extern "C" {

void ref_continuity(int x0, int x1, int y0, int y1, int W, double *ssha,
                    const double *sshn, const double *sshn_u,
                    const double *sshn_v, const double *hu, const double *hv,
                    const double *un, const double *vn, double rdt,
                    const double *e12t)
{
    SWEEP(x0, x1, y0, y1)
    continuity_code(ji, jj, W, ssha, D(sshn), D(sshn_u), D(sshn_v), D(hu),
                    D(hv), D(un), D(vn), rdt, D(e12t));
}

void ref_momentum_u(int x0, int x1, int y0, int y1, int W, double *ua,
                    const double *un, const double *vn, const double *hu,
                    const double *hv, const double *ht, const double *ssha_u,
                    const double *sshn, const double *sshn_u,
                    const double *sshn_v, const int *tmask, const double *e1u,
                    const double *e1v, const double *e1t, const double *e2u,
                    const double *e2t, const double *e12u, const double *gphiu,
                    double rdt, double cbfr, double visc, double omega,
                    double d2r, double g)
{
    SWEEP(x0, x1, y0, y1)
    momentum_u_code(ji, jj, W, ua, D(un), D(vn), D(hu), D(hv), D(ht),
                    D(ssha_u), D(sshn), D(sshn_u), D(sshn_v), I(tmask), D(e1u),
                    D(e1v), D(e1t), D(e2u), D(e2t), D(e12u), D(gphiu), rdt,
                    cbfr, visc, omega, d2r, g);
}

void ref_momentum_v(int x0, int x1, int y0, int y1, int W, double *va,
                    const double *un, const double *vn, const double *hu,
                    const double *hv, const double *ht, const double *ssha_v,
                    const double *sshn, const double *sshn_u,
                    const double *sshn_v, const int *tmask, const double *e1v,
                    const double *e1t, const double *e2u, const double *e2v,
                    const double *e2t, const double *e12v, const double *gphiv,
                    double rdt, double cbfr, double visc, double omega,
                    double d2r, double g)
{
    SWEEP(x0, x1, y0, y1)
    momentum_v_code(ji, jj, W, va, D(un), D(vn), D(hu), D(hv), D(ht),
                    D(ssha_v), D(sshn), D(sshn_u), D(sshn_v), I(tmask), D(e1v),
                    D(e1t), D(e2u), D(e2v), D(e2t), D(e12v), D(gphiv), rdt,
                    cbfr, visc, omega, d2r, g);
}

void ref_bc_ssh(int x0, int x1, int y0, int y1, int W, int istep, double *ssha,
                const int *tmask, double rdt)
{
    SWEEP(x0, x1, y0, y1)
    bc_ssh_code(ji, jj, W, istep, ssha, I(tmask), rdt);
}

void ref_bc_solid_u(int x0, int x1, int y0, int y1, int W, double *ua,
                    const int *tmask)
{
    SWEEP(x0, x1, y0, y1)
    bc_solid_u_code(ji, jj, W, ua, I(tmask));
}

void ref_bc_solid_v(int x0, int x1, int y0, int y1, int W, double *va,
                    const int *tmask)
{
    SWEEP(x0, x1, y0, y1)
    bc_solid_v_code(ji, jj, W, va, I(tmask));
}

void ref_bc_flather_u(int x0, int x1, int y0, int y1, int W, double *ua,
                      const double *hu, const double *sshn_u, const int *tmask,
                      double g)
{
    SWEEP(x0, x1, y0, y1)
    bc_flather_u_code(ji, jj, W, ua, D(hu), D(sshn_u), I(tmask), g);
}

void ref_bc_flather_v(int x0, int x1, int y0, int y1, int W, double *va,
                      const double *hv, const double *sshn_v, const int *tmask,
                      double g)
{
    SWEEP(x0, x1, y0, y1)
    bc_flather_v_code(ji, jj, W, va, D(hv), D(sshn_v), I(tmask), g);
}

void ref_next_sshu(int x0, int x1, int y0, int y1, int W, double *sshn_u,
                   const double *sshn, const int *tmask, const double *e12t,
                   const double *e12u)
{
    SWEEP(x0, x1, y0, y1)
    next_sshu_code(ji, jj, W, sshn_u, sshn, tmask, e12t, e12u);
}

void ref_next_sshv(int x0, int x1, int y0, int y1, int W, double *sshn_v,
                   const double *sshn, const int *tmask, const double *e12t,
                   const double *e12v)
{
    SWEEP(x0, x1, y0, y1)
    next_sshv_code(ji, jj, W, sshn_v, D(sshn), I(tmask), D(e12t), D(e12v));
}

} /* extern "C" */
