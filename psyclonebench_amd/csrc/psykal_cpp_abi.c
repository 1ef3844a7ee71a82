/*
 * libnemo_hip_cpp.so: the host-pointer step under the reference's own
 * symbol name, for the psykal_cpp Fortran shim, which binds
 * c_invoke_time_step with the signature of
 * manual_versions/psykal_cpp/time_step.cpp:10-51 (and passes one extra
 * trailing int, the field height, which this signature, like the
 * reference's, does not declare; time_step_mod.f90:86-87).
 * Forwards to c_invoke_time_step_host in libnemo_hip.so.
 */
#define NEMO_PSYKAL_CPP_ONLY
#include "nemolite2d_hip.h"

NEMO_API void c_invoke_time_step(double *ssha_t, double *sshn_t, double *sshn_u,
                        double *sshn_v, double *hu, double *hv, double *un,
                        double *vn, double *ua, double *ht, double *ssha_u,
                        double *va, double *ssha_v, int *tmask, double *area_t,
                        double *area_u, double *area_v, double *dx_u,
                        double *dx_v, double *dx_t, double *dy_u, double *dy_v,
                        double *dy_t, double *gphiu, double *gphiv, int istep,
                        int internal_xstart, int internal_xstop,
                        int internal_ystart, int internal_ystop, int width,
                        double rdt, double cbfr, double visc, double omega,
                        double d2r, double g)
{
    c_invoke_time_step_host(ssha_t, sshn_t, sshn_u, sshn_v, hu, hv, un, vn, ua,
                            ht, ssha_u, va, ssha_v, tmask, area_t, area_u,
                            area_v, dx_u, dx_v, dx_t, dy_u, dy_v, dy_t, gphiu,
                            gphiv, istep, internal_xstart, internal_xstop,
                            internal_ystart, internal_ystop, width, rdt, cbfr,
                            visc, omega, d2r, g);
}
