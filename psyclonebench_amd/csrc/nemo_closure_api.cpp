// C ABI of the learned coarse-grid closure (no counterpart in the
// reference). A handle owns the network (FP32 master parameters, bf16
// copies, optimizer moments), the activation and gradient tensors of one
// coarse grid, and the box-filtered fine state. The coarse and fine models
// themselves are ordinary device-resident contexts of this library: the
// caller passes the device pointers its slots hold (the current buffers of
// each after/now pair) to every call. All work runs on the library stream,
// in order with the steps.
#include <cmath>
#include <cstring>
#include <memory>
#include <vector>

#include "nemo_closure.h"
#include "nemo_hip_common.h"
#include "nemo_host.h"
#include "nemolite2d_hip.h"

namespace nemo {
namespace cl {
namespace {

struct ClosureContext {
    int nx = 0, ny = 0, Wc = 0, Hc = 0;  // coarse interior, coarse pitch and rows
    Norm nm{};
    float *params = nullptr, *grads = nullptr, *mom1 = nullptr, *mom2 = nullptr;
    __bf16 *wf[NLAYER] = {}, *wx[NLAYER] = {};
    __bf16 *in0 = nullptr, *act[NLAYER - 1] = {}, *pre[NLAYER - 1] = {}, *dpre[NLAYER] = {};
    float *pred = nullptr, *target = nullptr, *gate = nullptr;
    float *dw_partial = nullptr, *b_partial = nullptr, *loss_partial = nullptr;
    double *fssh = nullptr, *fu = nullptr, *fv = nullptr;
    double *guard_partial = nullptr;
    int *flag = nullptr, *trips = nullptr;
    int opt_kind = 1;
    float lr = 1e-3f, b1 = 0.9f, b2 = 0.999f, eps = 1e-8f, max_norm = 0.0f;
    LossWeights loss_w{{1.0f, 1.0f, 1.0f}};
    float *norm_scale = nullptr;  // device: gradient norm, clip factor
    long long step = 0;
    double guard_rel = 0.05, guard_abs = 0.0, g = 9.80665;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    bool timing = false;
    double last_ms = -1.0;
    // changes whenever a value the correction kernels take as an argument
    // changes (normalisation, gate, guard): graphs of nemo_cl_steps captured
    // under an older generation are not replayed
    unsigned long long generation = 0;

    size_t padded(int c) const { return (size_t)(ny + 4) * (size_t)(nx + 4) * (size_t)c; }
    float *bias(int l) const { return params + layer_offset(l) + layer_wsize(l); }

    ~ClosureContext()
    {
        void *all[] = {params, grads, mom1, mom2, in0, pred, target, dw_partial, b_partial,
                       loss_partial, fssh, fu, fv, guard_partial, flag, trips, norm_scale,
                       gate};
        for (void *p : all)
            if (p) (void)hipFree(p);
        for (int l = 0; l < NLAYER; l++) {
            if (wf[l]) (void)hipFree(wf[l]);
            if (wx[l]) (void)hipFree(wx[l]);
            if (dpre[l]) (void)hipFree(dpre[l]);
        }
        for (int l = 0; l < NLAYER - 1; l++) {
            if (act[l]) (void)hipFree(act[l]);
            if (pre[l]) (void)hipFree(pre[l]);
        }
        if (ev0) (void)hipEventDestroy(ev0);
        if (ev1) (void)hipEventDestroy(ev1);
    }
};

template <typename T>
T *zeros(size_t n, hipStream_t s)
{
    T *p = dev_alloc<T>(n);
    NEMO_HIP_CHECK(hipMemsetAsync(p, 0, n * sizeof(T), s));
    return p;
}

ClosureContext &ctx(void *h)
{
    if (!h) NEMO_FATAL("null closure handle");
    return *static_cast<ClosureContext *>(h);
}

// one counter for every handle, so a generation is never reused
unsigned long long next_generation()
{
    static unsigned long long g = 0;
    return ++g;
}

void refresh_bf16(ClosureContext &C, hipStream_t s)
{
    for (int l = 0; l < NLAYER; l++)
        launch_weights_bf16(s, C.params + layer_offset(l), layer_cin(l), layer_cout(l),
                            C.wf[l], C.wx[l]);
}

CoarseState state(ClosureContext &C, double *ssh, double *u, double *v, double *sshu,
                  double *sshv, const double *ht, const int *tmask, const double *area_t,
                  const double *area_u, const double *area_v)
{
    return CoarseState{ssh, u, v, sshu, sshv, ht, tmask, area_t, area_u, area_v, C.Wc};
}

// Forward pass from C.in0; training keeps the pre-activations; gated: the
// output is multiplied by the activity gate k_pack computed with the input;
// timed: HIP events around it (C.timing, outside captured graphs).
// guard: the output layer also writes the energy guard's partial sums
void forward(ClosureContext &C, hipStream_t s, bool training, bool gated, bool timed,
             const GuardSums *guard = nullptr)
{
    if (timed) NEMO_HIP_CHECK(hipEventRecord(C.ev0, s));
    const __bf16 *x = C.in0;
    for (int l = 0; l < NLAYER; l++) {
        ConvArgs a{};
        a.in = x;
        a.w = C.wf[l];
        a.bias = C.bias(l);
        a.nx = C.nx;
        a.ny = C.ny;
        if (l < NLAYER - 1) {
            a.out = C.act[l];
            a.pre = training ? C.pre[l] : nullptr;
            launch_conv(s, layer_cin(l), layer_cout(l), EPI_HIDDEN, a);
            x = C.act[l];
        } else {
            a.outf = C.pred;
            a.gate = gated ? C.gate : nullptr;
            if (guard)
                launch_conv_out_guarded(s, a, *guard);
            else
                launch_conv(s, layer_cin(l), layer_cout(l), EPI_OUTPUT, a);
        }
    }
    if (timed) NEMO_HIP_CHECK(hipEventRecord(C.ev1, s));
}

// The in-step correction of the coarse state c: pack the input, forward
// pass, energy guard (guard != 0), corrected fields, face ssh.
void enqueue_correction(ClosureContext &C, hipStream_t s, int guard, const CoarseState &c,
                        bool timed)
{
    launch_pack(s, c, C.nm, C.in0, C.gate, C.nx, C.ny);
    const GuardSums gs{C.nm, c, C.g, C.guard_partial};
    forward(C, s, false, true, timed, guard ? &gs : nullptr);
    if (guard)
        launch_guard_decide(s, C.guard_partial, guard_blocks(C.nx, C.ny), C.guard_rel,
                            C.guard_abs, C.flag, C.trips);
    const int *flag = guard ? C.flag : nullptr;
    launch_set_fields(s, 1, nullptr, nullptr, nullptr, C.pred, C.nm, flag, c, C.nx, C.ny);
    launch_face_ssh(s, flag, c, C.nx, C.ny);
}

struct CorrectionHook {
    ClosureContext *C;
    int guard;
};

void correction_hook(void *user, hipStream_t s, double *ssh, double *u, double *v,
                     double *sshu, double *sshv, const double *ht, const int *tmask,
                     const double *area_t, const double *area_u, const double *area_v)
{
    const CorrectionHook &h = *static_cast<const CorrectionHook *>(user);
    enqueue_correction(*h.C, s, h.guard,
                       CoarseState{ssh, u, v, sshu, sshv, ht, tmask, area_t, area_u, area_v,
                                   h.C->Wc},
                       false);
}

// Loss gradient already in C.dpre[NLAYER-1]; fills C.grads.
void backward(ClosureContext &C, hipStream_t s)
{
    for (int l = NLAYER - 1; l >= 0; l--) {
        const int ci = layer_cin(l), co = layer_cout(l);
        const __bf16 *x = l == 0 ? C.in0 : C.act[l - 1];
        launch_dw(s, co, ci, C.dpre[l], x, C.dw_partial, C.nx, C.ny);
        launch_reduce(s, C.dw_partial, dw_blocks(C.nx, C.ny), layer_wsize(l),
                      C.grads + layer_offset(l));
        launch_bias_partial(s, co, C.dpre[l], C.b_partial, C.nx, C.ny);
        launch_reduce(s, C.b_partial, bias_blocks(), co,
                      C.grads + layer_offset(l) + layer_wsize(l));
        if (l > 0) {  // gradient of the layer's input, times gelu' of the layer below
            ConvArgs a{};
            a.in = C.dpre[l];
            a.w = C.wx[l];
            a.out = C.dpre[l - 1];
            a.gpre = C.pre[l - 1];
            a.nx = C.nx;
            a.ny = C.ny;
            launch_conv(s, co, ci, EPI_DX, a);
        }
    }
}

void optimizer_step(ClosureContext &C, hipStream_t s)
{
    C.step++;
    const float bc1 = 1.0f - std::pow(C.b1, (float)C.step);
    const float bc2 = 1.0f - std::pow(C.b2, (float)C.step);
    launch_optimizer(s, C.opt_kind, C.params, C.grads, C.mom1, C.mom2, param_count(), C.lr,
                     C.b1, C.b2, C.eps, bc1, bc2, C.max_norm, C.norm_scale);
    refresh_bf16(C, s);
}

// host float -> bf16, round to nearest even (as the device conversion)
__bf16 to_bf16(float f)
{
    unsigned u;
    std::memcpy(&u, &f, 4);
    if ((u & 0x7f800000u) == 0x7f800000u) {
        u = (u & 0x007fffffu) ? (u | 0x00400000u) : u;  // keep inf, quiet NaN
    } else {
        u += 0x7fffu + ((u >> 16) & 1u);
    }
    const unsigned short h = (unsigned short)(u >> 16);
    __bf16 b;
    std::memcpy(&b, &h, 2);
    return b;
}

// normalised host inputs [ny*nx][4] -> C.in0
void upload_inputs(ClosureContext &C, const float *in4, hipStream_t s)
{
    std::vector<__bf16> t(C.padded(IN0C), to_bf16(0.0f));
    for (int y = 0; y < C.ny; y++)
        for (int x = 0; x < C.nx; x++)
            for (int k = 0; k < IN0C; k++)
                t[((size_t)(y + 2) * (C.nx + 4) + x + 2) * IN0C + k] =
                    to_bf16(in4[((size_t)y * C.nx + x) * 4 + k]);
    NEMO_HIP_CHECK(hipMemcpyAsync(C.in0, t.data(), t.size() * sizeof(__bf16),
                                  hipMemcpyHostToDevice, s));
    NEMO_HIP_CHECK(hipStreamSynchronize(s));
}

}  // namespace
}  // namespace cl
}  // namespace nemo

using namespace nemo;
using namespace nemo::cl;

extern "C" {

void *nemo_cl_create(int nx, int ny, int coarse_width, int coarse_height,
                     const float *init_params)
{
    if (nx <= 0 || ny <= 0 || nx % 32 != 0)
        NEMO_FATAL("closure: coarse interior %d x %d (nx must be a multiple of 32)", nx, ny);
    if (coarse_width < nx + 2 || coarse_height < ny + 2)
        NEMO_FATAL("closure: coarse arrays %d x %d too small for %d x %d", coarse_width,
                   coarse_height, nx, ny);
    auto C = std::make_unique<ClosureContext>();
    C->nx = nx;
    C->ny = ny;
    C->Wc = coarse_width;
    C->Hc = coarse_height;
    for (int k = 0; k < 4; k++) C->nm.in_scale[k] = 1.0f;
    for (int k = 0; k < 3; k++) C->nm.out_scale[k] = 1.0f;
    hipStream_t s = lib_stream();
    const int np = param_count();
    C->params = dev_alloc<float>(np);
    NEMO_HIP_CHECK(hipMemcpyAsync(C->params, init_params, np * sizeof(float),
                                  hipMemcpyHostToDevice, s));
    C->grads = zeros<float>(np, s);
    C->mom1 = zeros<float>(np, s);
    C->mom2 = zeros<float>(np, s);
    for (int l = 0; l < NLAYER; l++) {
        C->wf[l] = zeros<__bf16>(layer_wsize(l), s);
        C->wx[l] = zeros<__bf16>(layer_wsize(l), s);
        C->dpre[l] = zeros<__bf16>(C->padded(layer_cout(l)), s);
    }
    C->in0 = zeros<__bf16>(C->padded(IN0C), s);
    for (int l = 0; l < NLAYER - 1; l++) {
        C->act[l] = zeros<__bf16>(C->padded(HID), s);
        C->pre[l] = zeros<__bf16>(C->padded(HID), s);
    }
    const size_t npx = (size_t)nx * ny;
    C->pred = zeros<float>(npx * 4, s);
    C->gate = zeros<float>(npx, s);
    C->target = zeros<float>(npx * 4, s);
    C->dw_partial = zeros<float>((size_t)dw_blocks(nx, ny) * TAPS * HID * HID, s);
    C->b_partial = zeros<float>((size_t)bias_blocks() * HID, s);
    C->loss_partial = zeros<float>((size_t)loss_blocks() * 3, s);
    C->norm_scale = zeros<float>(2, s);
    const size_t nc = (size_t)coarse_width * coarse_height;
    C->fssh = zeros<double>(nc, s);
    C->fu = zeros<double>(nc, s);
    C->fv = zeros<double>(nc, s);
    C->guard_partial = zeros<double>((size_t)guard_blocks(nx, ny) * 3, s);
    C->flag = zeros<int>(1, s);
    C->trips = zeros<int>(1, s);
    NEMO_HIP_CHECK(hipEventCreate(&C->ev0));
    NEMO_HIP_CHECK(hipEventCreate(&C->ev1));
    C->generation = next_generation();
    refresh_bf16(*C, s);
    NEMO_HIP_CHECK(hipStreamSynchronize(s));
    return C.release();
}

int nemo_cl_param_count(void) { return param_count(); }

void nemo_cl_layer_info(int layer, int *cin, int *cout, int *offset)
{
    if (layer < 0 || layer >= NLAYER) NEMO_FATAL("closure: no layer %d", layer);
    *cin = layer_cin(layer);
    *cout = layer_cout(layer);
    *offset = layer_offset(layer);
}

void nemo_cl_get_params(void *h, float *out)
{
    ClosureContext &C = ctx(h);
    NEMO_HIP_CHECK(hipStreamSynchronize(lib_stream()));
    NEMO_HIP_CHECK(hipMemcpy(out, C.params, param_count() * sizeof(float),
                             hipMemcpyDeviceToHost));
}

void nemo_cl_set_params(void *h, const float *in)
{
    ClosureContext &C = ctx(h);
    hipStream_t s = lib_stream();
    NEMO_HIP_CHECK(hipMemcpyAsync(C.params, in, param_count() * sizeof(float),
                                  hipMemcpyHostToDevice, s));
    refresh_bf16(C, s);
    NEMO_HIP_CHECK(hipStreamSynchronize(s));
}

void nemo_cl_set_norm(void *h, const double *in_shift, const double *in_scale,
                      const double *out_scale)
{
    ClosureContext &C = ctx(h);
    for (int k = 0; k < 4; k++) {
        C.nm.in_shift[k] = (float)in_shift[k];
        C.nm.in_scale[k] = (float)in_scale[k];
    }
    for (int k = 0; k < 3; k++) C.nm.out_scale[k] = (float)out_scale[k];
    C.generation = next_generation();
}

void nemo_cl_set_optimizer(void *h, int kind, double lr, double beta1, double beta2,
                           double eps)
{
    ClosureContext &C = ctx(h);
    if (kind != 0 && kind != 1) NEMO_FATAL("closure optimizer kind %d", kind);
    C.opt_kind = kind;
    C.lr = (float)lr;
    C.b1 = (float)beta1;
    C.b2 = (float)beta2;
    C.eps = (float)eps;
}

void nemo_cl_set_guard_limit(void *h, double max_out)
{
    ClosureContext &C = ctx(h);
    if (!(max_out >= 0.0)) NEMO_FATAL("closure: guard limit must be >= 0 (got %g)", max_out);
    C.nm.max_out = (float)max_out;
    C.generation = next_generation();
}

void nemo_cl_set_ssh_cells(void *h, int all_wet)
{
    ClosureContext &C = ctx(h);
    C.nm.ssh_wet = all_wet ? 1 : 0;
    C.generation = next_generation();
}

void nemo_cl_set_gate(void *h, double eps)
{
    ClosureContext &C = ctx(h);
    C.nm.gate_eps = (float)eps;
    C.generation = next_generation();
}

void nemo_cl_set_guard(void *h, double rel, double abs_tol, double g)
{
    ClosureContext &C = ctx(h);
    C.guard_rel = rel;
    C.guard_abs = abs_tol;
    C.g = g;
    C.generation = next_generation();
}

void nemo_cl_filter(void *h, const double *fine_ssh, const double *fine_u,
                    const double *fine_v, int fine_width, int factor)
{
    ClosureContext &C = ctx(h);
    launch_filter(lib_stream(), fine_ssh, fine_u, fine_v, fine_width, factor, C.fssh, C.fu,
                  C.fv, C.Wc, C.nx, C.ny);
}

void nemo_cl_read_filtered(void *h, int field, double *out)
{
    ClosureContext &C = ctx(h);
    const double *src = field == 0 ? C.fssh : field == 1 ? C.fu : field == 2 ? C.fv : nullptr;
    if (!src) NEMO_FATAL("closure: filtered field %d", field);
    NEMO_HIP_CHECK(hipStreamSynchronize(lib_stream()));
    NEMO_HIP_CHECK(hipMemcpy(out, src, (size_t)C.Wc * C.Hc * sizeof(double),
                             hipMemcpyDeviceToHost));
}

void nemo_cl_set_state(void *h, double *ssh, double *u, double *v, double *sshu,
                       double *sshv, const double *ht, const int *tmask,
                       const double *area_t, const double *area_u, const double *area_v)
{
    ClosureContext &C = ctx(h);
    const CoarseState c = state(C, ssh, u, v, sshu, sshv, ht, tmask, area_t, area_u, area_v);
    hipStream_t s = lib_stream();
    launch_set_fields(s, 0, C.fssh, C.fu, C.fv, nullptr, C.nm, nullptr, c, C.nx, C.ny);
    launch_face_ssh(s, nullptr, c, C.nx, C.ny);
}

void nemo_cl_train_step(void *h, double *ssh, double *u, double *v, double *sshu,
                        double *sshv, const double *ht, const int *tmask,
                        const double *area_t, const double *area_u, const double *area_v)
{
    ClosureContext &C = ctx(h);
    const CoarseState c = state(C, ssh, u, v, sshu, sshv, ht, tmask, area_t, area_u, area_v);
    hipStream_t s = lib_stream();
    launch_pack(s, c, C.nm, C.in0, C.gate, C.nx, C.ny);
    forward(C, s, true, true, false);
    launch_target(s, C.fssh, C.fu, C.fv, c, C.nm, C.target, C.nx, C.ny);
    launch_loss(s, C.pred, C.target, C.dpre[NLAYER - 1], C.loss_partial, C.nx, C.ny, tmask,
                C.Wc, C.nm.ssh_wet, C.gate, C.loss_w);
    backward(C, s);
    optimizer_step(C, s);
}

double nemo_cl_last_loss(void *h)
{
    ClosureContext &C = ctx(h);
    std::vector<float> p((size_t)loss_blocks() * 3);
    NEMO_HIP_CHECK(hipStreamSynchronize(lib_stream()));
    NEMO_HIP_CHECK(hipMemcpy(p.data(), C.loss_partial, p.size() * sizeof(float),
                             hipMemcpyDeviceToHost));
    double sum = 0.0;
    for (float x : p) sum += x;
    return sum / (3.0 * C.nx * (double)C.ny);
}

void nemo_cl_set_clip(void *h, double max_norm) { ctx(h).max_norm = (float)max_norm; }

void nemo_cl_set_loss_weights(void *h, double wu, double wv, double wssh)
{
    if (!(wu >= 0.0 && wv >= 0.0 && wssh >= 0.0))
        NEMO_FATAL("closure: loss weights must be >= 0 (got %g, %g, %g)", wu, wv, wssh);
    ctx(h).loss_w = LossWeights{{(float)wu, (float)wv, (float)wssh}};
}

void nemo_cl_last_stats(void *h, double *out)
{
    ClosureContext &C = ctx(h);
    std::vector<float> p((size_t)loss_blocks() * 3);
    float ns[2];
    const int gblk = guard_blocks(C.nx, C.ny);
    std::vector<double> gp((size_t)gblk * 3);
    NEMO_HIP_CHECK(hipStreamSynchronize(lib_stream()));
    NEMO_HIP_CHECK(hipMemcpy(p.data(), C.loss_partial, p.size() * sizeof(float),
                             hipMemcpyDeviceToHost));
    NEMO_HIP_CHECK(hipMemcpy(ns, C.norm_scale, sizeof(ns), hipMemcpyDeviceToHost));
    NEMO_HIP_CHECK(hipMemcpy(gp.data(), C.guard_partial, gp.size() * sizeof(double),
                             hipMemcpyDeviceToHost));
    const double npx = (double)C.nx * C.ny;
    for (int c = 0; c < 3; c++) {
        double s = 0.0;
        for (int b = 0; b < loss_blocks(); b++) s += p[(size_t)b * 3 + c];
        out[c] = s / npx;
    }
    out[3] = ns[0];
    for (int r = 0; r < 3; r++) {
        double s = 0.0;
        for (int b = 0; b < gblk; b++) s += gp[(size_t)b * 3 + r];
        out[4 + r] = s;
    }
}

void nemo_cl_correct(void *h, int guard, double *ssh, double *u, double *v, double *sshu,
                     double *sshv, const double *ht, const int *tmask,
                     const double *area_t, const double *area_u, const double *area_v)
{
    ClosureContext &C = ctx(h);
    const CoarseState c = state(C, ssh, u, v, sshu, sshv, ht, tmask, area_t, area_u, area_v);
    enqueue_correction(C, lib_stream(), guard, c, C.timing);
}

void nemo_cl_steps(void *h, void **ssha_t_device, int nsteps, int istep, int guard)
{
    ClosureContext &C = ctx(h);
    CorrectionHook user{&C, guard ? 1 : 0};
    // the key names everything the captured kernels were built with
    const StepHook hook{&correction_hook, &user,
                        4ull * C.generation + 2ull * (unsigned)conv_form() + (guard ? 1 : 0)};
    psy_steps_with_hook(ssha_t_device, nsteps, istep, hook);
}

int nemo_cl_guard_trips(void *h)
{
    ClosureContext &C = ctx(h);
    int t = 0;
    NEMO_HIP_CHECK(hipStreamSynchronize(lib_stream()));
    NEMO_HIP_CHECK(hipMemcpy(&t, C.trips, sizeof(int), hipMemcpyDeviceToHost));
    return t;
}

void nemo_cl_set_timing(void *h, int on) { ctx(h).timing = on != 0; }

void nemo_cl_set_conv_form(int form) { set_conv_form(form); }

int nemo_cl_conv_form(void) { return conv_form(); }

double nemo_cl_forward_ms(void *h)
{
    ClosureContext &C = ctx(h);
    NEMO_HIP_CHECK(hipEventSynchronize(C.ev1));
    float ms = 0.f;
    NEMO_HIP_CHECK(hipEventElapsedTime(&ms, C.ev0, C.ev1));
    return ms;
}

void nemo_cl_forward_host(void *h, const float *in4, float *out3)
{
    ClosureContext &C = ctx(h);
    hipStream_t s = lib_stream();
    upload_inputs(C, in4, s);
    forward(C, s, false, false, C.timing);
    const size_t npx = (size_t)C.nx * C.ny;
    std::vector<float> p(npx * 4);
    NEMO_HIP_CHECK(hipStreamSynchronize(s));
    NEMO_HIP_CHECK(hipMemcpy(p.data(), C.pred, p.size() * sizeof(float), hipMemcpyDeviceToHost));
    for (size_t k = 0; k < npx; k++)
        for (int c = 0; c < 3; c++) out3[k * 3 + c] = p[k * 4 + c];
}

static void host_pair(ClosureContext &C, const float *in4, const float *target3,
                      hipStream_t s)
{
    upload_inputs(C, in4, s);
    const size_t npx = (size_t)C.nx * C.ny;
    std::vector<float> t(npx * 4, 0.0f);
    for (size_t k = 0; k < npx; k++)
        for (int c = 0; c < 3; c++) t[k * 4 + c] = target3[k * 3 + c];
    NEMO_HIP_CHECK(hipMemcpyAsync(C.target, t.data(), t.size() * sizeof(float),
                                  hipMemcpyHostToDevice, s));
    forward(C, s, true, false, false);
    launch_loss(s, C.pred, C.target, C.dpre[NLAYER - 1], C.loss_partial, C.nx, C.ny,
                nullptr, 0, 0, nullptr, C.loss_w);
    backward(C, s);
    NEMO_HIP_CHECK(hipStreamSynchronize(s));  // t goes out of scope
}

double nemo_cl_grad_host(void *h, const float *in4, const float *target3, float *grads)
{
    ClosureContext &C = ctx(h);
    hipStream_t s = lib_stream();
    host_pair(C, in4, target3, s);
    NEMO_HIP_CHECK(hipMemcpy(grads, C.grads, param_count() * sizeof(float),
                             hipMemcpyDeviceToHost));
    return nemo_cl_last_loss(h);
}

double nemo_cl_train_host(void *h, const float *in4, const float *target3)
{
    ClosureContext &C = ctx(h);
    hipStream_t s = lib_stream();
    host_pair(C, in4, target3, s);
    optimizer_step(C, s);
    return nemo_cl_last_loss(h);
}

void nemo_cl_destroy(void *h)
{
    if (!h) return;
    NEMO_HIP_CHECK(hipStreamSynchronize(lib_stream()));
    delete static_cast<ClosureContext *>(h);
}

}  // extern "C"
