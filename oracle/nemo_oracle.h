/*
 * TEST INFRASTRUCTURE ONLY. The oracle is a plain-C CPU restatement of the
 * reference NEMOLite2D time step. Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it, and only as the checker; the
 * product path (psyclonebench_amd, libnemo_hip.so) never calls it.
 *
 * Restated from:
 *   benchmarks/nemo/nemolite2d/manual_versions/psykal_cpp/time_step.cpp:10-137
 *   benchmarks/nemo/nemolite2d/manual_versions/psykal_cpp/time_step_omp.cpp:13-152
 *   benchmarks/nemo/nemolite2d/kernels/c_family/{continuity,momentum_u,
 *     momentum_v,boundary_conditions,time_update}_kern.c
 */
#ifndef NEMO_ORACLE_H
#define NEMO_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Same argument list and meaning as the reference's c_invoke_time_step
 * (psykal_cpp/time_step.cpp:10-51): 13 fields, 12 grid arrays, istep,
 * 0-based internal bounds, row pitch, 6 physical scalars. */
#define ORACLE_STEP_ARGS                                                      \
    double *ssha_t, double *sshn_t, double *sshn_u, double *sshn_v,           \
    double *hu, double *hv, double *un, double *vn, double *ua, double *ht,   \
    double *ssha_u, double *va, double *ssha_v, int *tmask, double *area_t,   \
    double *area_u, double *area_v, double *dx_u, double *dx_v, double *dx_t, \
    double *dy_u, double *dy_v, double *dy_t, double *gphiu, double *gphiv,   \
    int istep, int xs, int xe, int ys, int ye, int width, double rdt,         \
    double cbfr, double visc, double omega, double d2r, double g

/* Serial step: time_step.cpp loop order and bounds. */
void oracle_invoke_time_step(ORACLE_STEP_ARGS);
/* OpenMP step: time_step_omp.cpp (one parallel-for per sweep, flather_v
 * serial). Compiled without OpenMP it runs serially with the same result. */
void oracle_invoke_time_step_omp(ORACLE_STEP_ARGS);

/* Tidal forcing used by bc_ssh: ssha = amp * sin(omega_tide * istep * rdt).
 * Defaults are the reference literals 0.2 and 2*3.14159/(12.42*3600)
 * (boundary_conditions_kern.c:185-186). */
void oracle_set_tide(double amp, double omega_tide);
void oracle_reset_tide(void);

/* Individual sweeps over an explicit 0-based inclusive region, one call of
 * the per-point kernel per (ji, jj) in jj-outer, ji-inner order. */
void oracle_continuity(int x0, int x1, int y0, int y1, int width,
                       double *ssha, const double *sshn, const double *sshn_u,
                       const double *sshn_v, const double *hu, const double *hv,
                       const double *un, const double *vn, double rdt,
                       const double *e12t);
void oracle_momentum_u(int x0, int x1, int y0, int y1, int width, double *ua,
                       const double *un, const double *vn, const double *hu,
                       const double *hv, const double *ht, const double *ssha_u,
                       const double *sshn, const double *sshn_u,
                       const double *sshn_v, const int *tmask, const double *e1u,
                       const double *e1v, const double *e1t, const double *e2u,
                       const double *e2t, const double *e12u,
                       const double *gphiu, double rdt, double cbfr,
                       double visc, double omega, double d2r, double g);
void oracle_momentum_v(int x0, int x1, int y0, int y1, int width, double *va,
                       const double *un, const double *vn, const double *hu,
                       const double *hv, const double *ht, const double *ssha_v,
                       const double *sshn, const double *sshn_u,
                       const double *sshn_v, const int *tmask, const double *e1v,
                       const double *e1t, const double *e2u, const double *e2v,
                       const double *e2t, const double *e12v,
                       const double *gphiv, double rdt, double cbfr,
                       double visc, double omega, double d2r, double g);
void oracle_bc_ssh(int x0, int x1, int y0, int y1, int width, int istep,
                   double *ssha, const int *tmask, double rdt);
void oracle_bc_solid_u(int x0, int x1, int y0, int y1, int width, double *ua,
                       const int *tmask);
void oracle_bc_solid_v(int x0, int x1, int y0, int y1, int width, double *va,
                       const int *tmask);
void oracle_bc_flather_u(int x0, int x1, int y0, int y1, int width, double *ua,
                         const double *hu, const double *sshn_u,
                         const int *tmask, double g);
void oracle_bc_flather_v(int x0, int x1, int y0, int y1, int width, double *va,
                         const double *hv, const double *sshn_v,
                         const int *tmask, double g);
void oracle_next_sshu(int x0, int x1, int y0, int y1, int width, double *sshn_u,
                      const double *sshn, const int *tmask, const double *e12t,
                      const double *e12u);
void oracle_next_sshv(int x0, int x1, int y0, int y1, int width, double *sshn_v,
                      const double *sshn, const int *tmask, const double *e12t,
                      const double *e12v);

/* Sum of |x| over the inclusive region, jj outer / ji inner. */
double oracle_checksum(const double *x, int width, int x0, int x1, int y0,
                       int y1);

#ifdef __cplusplus
}
#endif
#endif
