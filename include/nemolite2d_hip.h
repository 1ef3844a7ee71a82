/*
 * libnemo_hip.so -- NEMOLite2D time step on one AMD Instinct MI355X (gfx950)
 * behind the C ABI of the reference's manual PSy layer.
 *
 * Plain C: pointers, ints and doubles only. Errors print the failing call
 * and exit(1), the reference's convention (psykal_c_opencl/opencl_utils.c:
 * 218-226, check_status).
 *
 * Arrays follow the dl_esm_inf layout: each is `width` doubles per row,
 * point (ji, jj) at jj*width + ji (0-based); tmask is int32 with the same
 * pitch. Bounds are the 0-based internal T region (internal%xstart-1 ...).
 *
 * Entry points and the reference interface each one replaces:
 *
 *   c_invoke_time_step           device-resident PSy layer,
 *                                manual_versions/psykal_c_opencl/time_step.c:
 *                                100-169 (Fortran interface
 *                                psykal_c_opencl/time_step_mod.f90:11-70)
 *   c_read_from_device           psykal_c_opencl/time_step.c:788-797
 *   c_write_to_device            psykal_c_opencl/time_step.c:801-810
 *   c_invoke_time_step_host      host-pointer PSy layer,
 *                                manual_versions/psykal_cpp/time_step.cpp:
 *                                10-51 (libnemo_hip_cpp.so re-exports it
 *                                under the reference name c_invoke_time_step
 *                                for the psykal_cpp/time_step_mod.f90 shim)
 *   c_invoke_time_steps          no reference counterpart: n steps of
 *                                c_invoke_time_step replayed as one hipGraph
 *   nemo_hip_*                   no reference counterpart: configuration,
 *                                timing and teardown
 */
#ifndef NEMOLITE2D_HIP_H
#define NEMOLITE2D_HIP_H

#include <stdbool.h>

#ifndef NEMO_API
#define NEMO_API __attribute__((visibility("default")))
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* The 25 host arrays, in the reference's argument order
 * (psykal_cpp/time_step.cpp:12-37). */
#define NEMO_HOST_ARRAYS                                                      \
    double *ssha_t, double *sshn_t, double *sshn_u, double *sshn_v,           \
    double *hu, double *hv, double *un, double *vn, double *ua, double *ht,   \
    double *ssha_u, double *va, double *ssha_v, int *tmask, double *area_t,   \
    double *area_u, double *area_v, double *dx_u, double *dx_v, double *dx_t, \
    double *dy_u, double *dy_v, double *dy_t, double *gphiu, double *gphiv

/* The 25 device-handle slots, passed by reference in the same order
 * (psykal_c_opencl/time_step.c:138-162, cl_mem* there). */
#define NEMO_DEVICE_SLOTS                                                     \
    void **ssha_t_device, void **sshn_t_device, void **sshn_u_device,         \
    void **sshn_v_device, void **hu_device, void **hv_device,                 \
    void **un_device, void **vn_device, void **ua_device, void **ht_device,   \
    void **ssha_u_device, void **va_device, void **ssha_v_device,             \
    void **tmask_device, void **area_t_device, void **area_u_device,          \
    void **area_v_device, void **dx_u_device, void **dx_v_device,             \
    void **dx_t_device, void **dy_u_device, void **dy_v_device,               \
    void **dy_t_device, void **gphiu_device, void **gphiv_device

/*
 * Device-resident step (psykal_c_opencl/time_step.c:100-169).
 * First call for a given set of slots: allocates device buffers, uploads
 * all 25 host arrays (total_size elements each) and stores the handles in
 * the slots. Every call: one time step on the device, asynchronous on the
 * library's stream; the host arrays are not read again and not updated --
 * fetch results with c_read_from_device(*slot, host, ...).
 * After a call, the slots of each after/now pair (ssha_t/sshn_t, ua/un,
 * va/vn) hold one handle: the reference copies after->now at the end of the
 * step, here the pair shares one buffer instead. The handles change from
 * call to call (ping-pong); always read the slots after the call.
 */
#ifndef NEMO_PSYKAL_CPP_ONLY
NEMO_API void c_invoke_time_step(NEMO_HOST_ARRAYS, NEMO_DEVICE_SLOTS, int istep,
                        int internal_xstart, int internal_xstop,
                        int internal_ystart, int internal_ystop, int width,
                        int total_size, double rdt, double cbfr, double visc,
                        double omega, double d2r, double g);
#endif

/* nsteps consecutive steps istep, istep+1, ... as one replay of a captured
 * hipGraph (captured on first use of a given nsteps). Same arguments and
 * slot semantics as c_invoke_time_step. */
NEMO_API void c_invoke_time_steps(int nsteps, NEMO_HOST_ARRAYS, NEMO_DEVICE_SLOTS,
                         int istep, int internal_xstart, int internal_xstop,
                         int internal_ystart, int internal_ystop, int width,
                         int total_size, double rdt, double cbfr, double visc,
                         double omega, double d2r, double g);

/* Blocking copy of nx*ny doubles from device handle `from` to host `to`,
 * after all queued steps have finished. As in the reference, the copy is
 * the contiguous block from the start of the buffer; startx/starty and
 * blocking are accepted and ignored (the copy is always blocking). */
NEMO_API void c_read_from_device(void *from, void *to, int startx, int starty, int nx,
                        int ny, bool blocking);

/* Blocking copy of nx*ny doubles from host `from` to device handle `to`
 * (the reference body reads instead of writing; this writes). */
NEMO_API void c_write_to_device(void *from, void *to, int startx, int starty, int nx,
                       int ny, bool blocking);

/* Host-pointer step (psykal_cpp/time_step.cpp:10-51): uploads the arrays,
 * runs the unfused exact kernel sequence, downloads ssha_t, sshn_t, sshn_u,
 * sshn_v, un, vn, ua, va, and returns with the host arrays updated. */
NEMO_API void c_invoke_time_step_host(NEMO_HOST_ARRAYS, int istep, int internal_xstart,
                             int internal_xstop, int internal_ystart,
                             int internal_ystop, int width, double rdt,
                             double cbfr, double visc, double omega,
                             double d2r, double g);

/* Allocation-free step. No counterpart in the reference: it is for callers
 * that cannot allocate device memory (code under HIP graph capture, tight
 * serving loops). One step with the semantics of c_invoke_time_step_host
 * (psykal_cpp/time_step.cpp:53-136, every sweep in place, now <- after over
 * the reference's copy region), but the 25 array arguments are DEVICE
 * pointers the caller owns, with total_size elements each (the reference
 * layout), and the step is queued on `stream` (a hipStream_t; NULL = the
 * null stream) without waiting for it. The library allocates nothing: the
 * per-point Coriolis factors, evaluated on the device at every call, go to
 * a static device scratch array of 32 MiB per GPU, so total_size is at most
 * 2097152 (2^21) elements; larger grids end with a message and exit(1).
 * Every call on one GPU shares that array: calls that overlap on one GPU
 * (other streams, other threads) race. The tide is nemo_hip_set_tide's. */
NEMO_API void c_invoke_time_step_noalloc(NEMO_HOST_ARRAYS, int istep,
                                         int internal_xstart, int internal_xstop,
                                         int internal_ystart, int internal_ystop,
                                         int width, int total_size, double rdt,
                                         double cbfr, double visc, double omega,
                                         double d2r, double g, void *stream);

/* ---- configuration, timing, teardown (no reference counterpart) ------- */

/* Tidal forcing of bc_ssh: ssha = amp * sin(omega_tide * istep * rdt).
 * Default: the reference literals amp = 0.2, omega_tide = 2*3.14159/
 * (12.42*3600) (boundary_conditions_kern.c:185-186). */
NEMO_API void nemo_hip_set_tide(double amp, double omega_tide);

/* Execution form for contexts created afterwards:
 *  0 auto    fused step when the mask allows it (no wet cell on the last
 *            T row, which the fused form does not copy), unfused otherwise
 *  1 unfused one kernel per reference sweep, in place (exact for any input)
 *  2 fused   fused step; exits with an error when the mask does not allow it
 */
NEMO_API void nemo_hip_set_mode(int mode);

/* Form in use by the context whose first slot is ssha_t_device:
 * 1 unfused, 2 fused, 0 no such context. */
NEMO_API int nemo_hip_context_mode(void **ssha_t_device);

/* Waits for all queued work of the library. */
NEMO_API void nemo_hip_synchronize(void);

/* Average device time in ms of one fused step kernel over the fused steps
 * of the last c_invoke_time_steps call made with timing on, measured with
 * HIP events on the library stream around each fused launch (timing on
 * runs the steps as plain launches, not as a graph). -1 if none. */
NEMO_API void nemo_hip_set_kernel_timing(int on);
NEMO_API double nemo_hip_fused_kernel_ms(void);

/* Kernel shape of the fused form: -1 = 64x4-cell tiles with an LDS ssh
 * tile; 0 .. nemo_hip_fused_variant_count()-1 = row-march shapes (block
 * width x rows per band). All shapes compute bitwise the same result;
 * default 31 (4-row bands of 56 columns with the loads issued two rows
 * ahead, the fastest measured at 4096^2).
 * Applies to the next launches of every context. */
NEMO_API int nemo_hip_fused_variant_count(void);
NEMO_API void nemo_hip_set_fused_variant(int variant);
NEMO_API int nemo_hip_fused_variant(void);
/* Block width (columns), rows per block, and the march kind of a variant:
 * 0 = row march, 1 = row march with the row loads issued one row ahead,
 * 2/3/4 = overlapped-wavefront march with at least 1/3/4 waves per SIMD
 * requested from the register allocator, 5 / 6 = overlapped-wavefront
 * march with 56 / 48 output columns per wavefront (64-byte-aligned
 * stores), 7 / 8 = the 56-column march with its loads issued two rows
 * ahead, compiled for at least 2 / 3 waves per SIMD, 9 / 10 = the persistent
 * chunked 56-column march with / without the loads two rows ahead (rows = 0:
 * one chunk of rows per wavefront the GPU holds at once, set at launch),
 * 11 = the 56-column march with its loads issued three rows ahead. */
NEMO_API void nemo_hip_fused_variant_shape(int variant, int *block_width,
                                           int *rows, int *prefetch);
/* Name of the kernel a variant launches, as profilers report it. */
NEMO_API const char *nemo_hip_fused_variant_kernel(int variant);

/* ---- mixed-precision form (no reference counterpart) -------------------
 * FP32 storage and arithmetic for velocities, face ssh, depths and metrics;
 * the T-cell ssh is a compensated pair of FP32 words updated with FP64
 * accumulation; the tidal elevation amp*sin(omega_tide*istep*rdt) is FP64.
 * Takes the same 25 FP64 host arrays as c_invoke_time_step (converted on
 * upload; the T mask is kept in one byte a point, so its values must lie in
 * -64..63) and returns an opaque handle. Requires the same mask condition as
 * the fused form (exits otherwise). */
NEMO_API void *nemo_mp_create(NEMO_HOST_ARRAYS, int internal_xstart,
                              int internal_xstop, int internal_ystart,
                              int internal_ystop, int width, int total_size,
                              double rdt, double cbfr, double visc,
                              double omega, double d2r, double g,
                              double tide_amp, double tide_omega);
/* nsteps steps istep, istep+1, ... as one hipGraph replay (asynchronous). */
NEMO_API void nemo_mp_steps(void *handle, int nsteps, int istep);
/* The same steps as plain launches, each timed with HIP events; returns the
 * mean duration in ms of the mixed step kernel. */
NEMO_API double nemo_mp_kernel_ms(void *handle, int nsteps, int istep);
/* Rows per band of the kernel: 4, 8, 16, 32 (default), 48, 64, 80, 96 or
 * 128. */
NEMO_API void nemo_mp_set_rows(void *handle, int rows);
/* Output columns per wavefront: 61 (default; strips overlap by 3 columns),
 * 56 (every strip's stores start on a 32-byte boundary) or 48 (on a 64-byte
 * boundary). */
NEMO_API void nemo_mp_set_columns(void *handle, int cols);
/* 1 / 2: the kernel that issues its loads two / three rows ahead (rows per
 * band 4 or 8 only); 0 (default): the loads of a row issued where the row
 * starts. All compute the same bits. */
NEMO_API void nemo_mp_set_prefetch(void *handle, int prefetch);
/* Blocking read-back as FP64 (total_size values): field 0 sshn_t (hi+lo),
 * 1 un, 2 vn, 3 sshn_u, 4 sshn_v. */
NEMO_API void nemo_mp_read(void *handle, int field, double *out);
NEMO_API void nemo_mp_destroy(void *handle);

/* ---- batched ensemble (no reference counterpart) -----------------------
 * ndom independent domains, each in its own dl_esm_inf layout, advanced
 * together: one launch per step walks a tile -> domain work list (one
 * wavefront per tile, or a persistent grid), then one launch runs every
 * domain's Flather segments.
 *   arrays      ndom x 25 host pointers, each row in the c_invoke_time_step
 *               argument order (tmask as int*); read during creation only
 *   geom        ndom x 6 ints: xs, xe, ys, ye (0-based internal bounds),
 *               width, total_size
 *   tide_amp, tide_omega   per domain: bc_ssh = amp*sin(omega*istep*rdt)
 *   precision   64: the FP64 step (bitwise the single-domain fused form);
 *               32: the mixed form (bitwise the single-domain nemo_mp_*;
 *               mask values in -64..63)
 * Every domain's mask must allow the fused form (no wet cell on row ye+1).
 * One field's arena (every domain, offsets rounded to 64 elements) must stay
 * below 4 GiB. Returns a handle; errors exit(1). */
NEMO_API void *nemo_ens_create(int ndom, void *const *arrays, const int *geom,
                               const double *tide_amp, const double *tide_omega,
                               double rdt, double cbfr, double visc, double omega,
                               double d2r, double g, int precision);
/* Launch shapes: rows per band (tile height), wavefronts per workgroup,
 * workgroups per launch (0: one one-wavefront workgroup per tile; else a
 * persistent launch of that many); nemo_ens_shape_columns gives the tile
 * width (owned columns per wavefront: 61, or 56 for 64-byte-aligned stores,
 * or 0: per domain, whichever of the two its width runs faster with) and
 * nemo_ens_shape_prefetch the rows of loads the tile issues ahead of the row
 * that needs them, beyond the next one (0, 1, or 2 with one workgroup per
 * tile). A new ensemble runs one wavefront per 4-row tile, per-domain
 * columns, prefetch 2, until nemo_ens_set_shape or nemo_ens_autotune. */
NEMO_API int nemo_ens_shape_count(void *handle);
NEMO_API void nemo_ens_shape_info(void *handle, int shape, int *rows,
                                  int *waves_per_block, int *blocks);
NEMO_API int nemo_ens_shape_columns(void *handle, int shape);
NEMO_API int nemo_ens_shape_prefetch(void *handle, int shape);
NEMO_API void nemo_ens_set_shape(void *handle, int shape);
NEMO_API int nemo_ens_shape(void *handle);
/* Times every shape on steps_per_shape steps of this ensemble (HIP events
 * around each step kernel; mean ms per shape into ms_out if not NULL),
 * restores the state, selects and returns the fastest shape. */
NEMO_API int nemo_ens_autotune(void *handle, int steps_per_shape, double *ms_out);
/* nsteps steps istep, istep+1, ... (one hipGraph replay; asynchronous). */
NEMO_API void nemo_ens_steps(void *handle, int nsteps, int istep);
/* The same as plain launches; returns the mean step-kernel time in ms. */
NEMO_API double nemo_ens_kernel_ms(void *handle, int nsteps, int istep);
NEMO_API long long nemo_ens_interior_points(void *handle);
/* Blocking read-back of one domain as FP64 (its total_size values): field
 * 0 sshn_t, 1 un, 2 vn, 3 sshn_u, 4 sshn_v. */
NEMO_API void nemo_ens_read(void *handle, int domain, int field, double *out);
NEMO_API void nemo_ens_destroy(void *handle);

/* ---- learned coarse-grid closure (no reference counterpart) -----------
 * A 4-layer 5x5 convolutional network (4 -> 32 -> 32 -> 32 -> 3 channels,
 * GELU in tanh form) on the coarse grid's interior T points (nx x ny, nx a
 * multiple of 32), bf16 MFMA with FP32 accumulation; inputs u, v, ssh, depth, outputs
 * corrections to u, v, ssh. Parameters are FP32, nemo_cl_param_count() of
 * them, layer by layer: weights [25 taps][cout][cin] (cin 16 for layer 0 of
 * which 4 are used, cout 16 for layer 3 of which 3 are used) then bias.
 * Coarse-state arguments are device pointers of a coarse model created with
 * c_invoke_time_step (the buffers its slots hold: ssh = sshn_t, u = un,
 * v = vn, sshu = sshn_u, sshv = sshn_v, then ht, tmask, area_t/u/v), pitch
 * coarse_width; every call is queued on the library stream. */
NEMO_API void *nemo_cl_create(int nx, int ny, int coarse_width, int coarse_height,
                              const float *init_params);
NEMO_API int nemo_cl_param_count(void);
NEMO_API void nemo_cl_layer_info(int layer, int *cin, int *cout, int *offset);
NEMO_API void nemo_cl_get_params(void *handle, float *out);
NEMO_API void nemo_cl_set_params(void *handle, const float *in);
/* network input c = (x_c - in_shift[c]) / in_scale[c] for c = u, v, ssh,
 * depth; correction k = output_k * out_scale[k] for k = u, v, ssh */
NEMO_API void nemo_cl_set_norm(void *handle, const double *in_shift,
                               const double *in_scale, const double *out_scale);
/* activity gate: the correction of a point is multiplied by
 * 1 - exp(-(a/eps)^2), a = |u| + |v| + |ssh| in normalised input units, so
 * points at rest get none (eps <= 0: no gate; default 0) */
NEMO_API void nemo_cl_set_gate(void *handle, double eps);
/* ssh cells the correction changes and the training loss counts: 1 every wet
 * interior cell, including the ones the step clamps to the tide (wet cells
 * with an open 4-neighbour, whose next step starts from the corrected value
 * and clamps them again); 0 (default) only the cells the step does not
 * clamp */
NEMO_API void nemo_cl_set_ssh_cells(void *handle, int all_wet);
/* kind 0: SGD with momentum beta1; kind 1: Adam */
NEMO_API void nemo_cl_set_optimizer(void *handle, int kind, double lr, double beta1,
                                    double beta2, double eps);
/* correction kept when finite and E_after <= (1 + rel) E_before + abs_tol,
 * E = sum over interior cells of g ssh^2 + ht (u^2 + v^2) */
NEMO_API void nemo_cl_set_guard(void *handle, double rel, double abs_tol, double g);
/* the guard also rejects a correction with any applied component beyond
 * max_out in normalised output units, however the energy changes (a few
 * large local corrections fit a global energy budget and can wreck the next
 * steps); 0 (default): no such bound */
NEMO_API void nemo_cl_set_guard_limit(void *handle, double max_out);
/* box filter of a fine state (device pointers, pitch fine_width) by
 * `factor` onto the coarse grid, kept in the handle */
NEMO_API void nemo_cl_filter(void *handle, const double *fine_ssh, const double *fine_u,
                             const double *fine_v, int fine_width, int factor);
NEMO_API void nemo_cl_read_filtered(void *handle, int field, double *out);
/* coarse state := filtered fine state (wet interior cells and faces; face
 * ssh recomputed as next_sshu/v do) */
NEMO_API void nemo_cl_set_state(void *handle, double *ssh, double *u, double *v,
                                double *sshu, double *sshv, const double *ht,
                                const int *tmask, const double *area_t,
                                const double *area_u, const double *area_v);
/* one optimizer step on the pair (coarse state after a plain step, the
 * filtered fine state of the same time): loss = mean squared error of the
 * normalised correction */
NEMO_API void nemo_cl_train_step(void *handle, double *ssh, double *u, double *v,
                                 double *sshu, double *sshv, const double *ht,
                                 const int *tmask, const double *area_t,
                                 const double *area_u, const double *area_v);
NEMO_API double nemo_cl_last_loss(void *handle);
/* gradients scaled to a global L2 norm of at most max_norm before the
 * optimizer step (0: no clipping) */
NEMO_API void nemo_cl_set_clip(void *handle, double max_norm);
/* weights of the u, v and ssh squared errors in the training loss
 * (default 1, 1, 1); >= 0 */
NEMO_API void nemo_cl_set_loss_weights(void *handle, double wu, double wv, double wssh);
/* out[7]: mean squared error of the last loss per channel (u, v, ssh),
 * gradient norm of the last optimizer step, and the energies before / after
 * the last guarded correction and its non-finite count */
NEMO_API void nemo_cl_last_stats(void *handle, double *out);
/* applies the network correction to the coarse state; with guard != 0 the
 * energy check decides on the device whether it is kept (else the plain
 * step stands) */
NEMO_API void nemo_cl_correct(void *handle, int guard, double *ssh, double *u, double *v,
                              double *sshu, double *sshv, const double *ht,
                              const int *tmask, const double *area_t,
                              const double *area_u, const double *area_v);
/* nsteps coarse steps istep, istep+1, ... of the coarse model whose ssha_t
 * slot is ssha_t_device (stepped once before with c_invoke_time_step, fused
 * form), each followed by the correction of nemo_cl_correct, as one hipGraph
 * replay: the same results as alternating c_invoke_time_step and
 * nemo_cl_correct. The model's slots are republished. A graph is captured
 * on first use of (nsteps, parity, guard) and again after the
 * normalisation, gate or guard settings change. */
NEMO_API void nemo_cl_steps(void *handle, void **ssha_t_device, int nsteps, int istep,
                            int guard);
NEMO_API int nemo_cl_guard_trips(void *handle);
/* HIP events around the forward pass of later calls; ms of the last one */
NEMO_API void nemo_cl_set_timing(void *handle, int on);
NEMO_API double nemo_cl_forward_ms(void *handle);
/* Launch form of the 32 -> 32 convolutions and their input gradients, for
 * every closure in the process: 0 two 32-pixel tiles per wavefront pass,
 * 4 wavefronts per workgroup; 1 (default) one tile, 8 wavefronts per
 * workgroup. Both compute the same bits. Graphs nemo_cl_steps captured
 * under the other form are not replayed. No reference counterpart. */
NEMO_API void nemo_cl_set_conv_form(int form);
NEMO_API int nemo_cl_conv_form(void);
/* test hooks: forward / loss gradient of normalised host inputs
 * [ny*nx][4] and targets [ny*nx][3]; outputs [ny*nx][3], gradients in the
 * parameter layout; returns the loss */
NEMO_API void nemo_cl_forward_host(void *handle, const float *in4, float *out3);
NEMO_API double nemo_cl_grad_host(void *handle, const float *in4, const float *target3,
                                  float *grads);
/* the same followed by one optimizer step */
NEMO_API double nemo_cl_train_host(void *handle, const float *in4, const float *target3);
NEMO_API void nemo_cl_destroy(void *handle);

/* Releases the device buffers, graphs and lists of one context (identified
 * by its ssha_t slot), or of all contexts. Slots become dangling. */
NEMO_API void nemo_hip_release(void **ssha_t_device);
NEMO_API void nemo_hip_release_all(void);

/* Device name and HBM size into buf; returns the device count (0 when no
 * GPU is visible; never exits). */
NEMO_API int nemo_hip_device_info(char *buf, int buflen);

#ifdef __cplusplus
}
#endif
#endif
