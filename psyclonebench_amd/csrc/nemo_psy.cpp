// C-ABI PSy layer for the NEMOLite2D step on one MI355X: device buffers,
// first-call upload, ping-pong state, Flather segment lists, hipGraph
// replay. The entry points and the reference interfaces they replace are
// listed in include/nemolite2d_hip.h.
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <tuple>
#include <utility>
#include <vector>

#include "nemo_hip_common.h"
#include "nemo_host.h"
#include "nemo_step.h"
#include "nemolite2d_hip.h"

namespace nemo {

hipStream_t lib_stream()
{
    static hipStream_t s = [] {
        hipStream_t t;
        NEMO_HIP_CHECK(hipStreamCreateWithFlags(&t, hipStreamNonBlocking));
        return t;
    }();
    return s;
}

namespace {

// Reference tide literals, boundary_conditions_kern.c:185-186.
double g_amp_tide = 0.2;
double g_omega_tide = 2.0 * 3.14159 / (12.42 * 3600.0);
int g_mode = 0;
// -1: tile kernel; >= 0: row-march variant. The default is the fastest
// measured at 4096^2: 4-row bands of 56-column, 64-byte-aligned strips with
// the loads issued two rows ahead.
int g_fused_variant = 31;
bool g_kernel_timing = false;
double g_fused_ms = -1.0;

}  // namespace

// bc_ssh value for one step: amp * sin(omega_tide * (istep*rdt)), evaluated
// on the host with libm exactly as boundary_conditions_kern.c:187-192 does.
double tide_at(int istep, double rdt)
{
    const double rtime = istep * rdt;
    return g_amp_tide * std::sin(g_omega_tide * rtime);
}

namespace {


// One model's device state.
struct Context {
    Geometry G{};
    Physics P{};
    void *buf[A_COUNT] = {};
    double *fcor_u = nullptr, *fcor_v = nullptr;
    double *spare_su = nullptr, *spare_sv = nullptr;
    DeviceFlather fl_u, fl_v;
    int mode = 1;          // 1 unfused, 2 fused
    int parity = 0;        // fused: which buffer set holds the current state
    bool paired = false;   // fused: both sets agree on never-written points
    double *d_tide = nullptr;
    double *h_tide = nullptr;  // pinned
    int tide_cap = 0;
    // n, parity, variant, hook key (0: plain steps)
    std::map<std::tuple<int, int, int, unsigned long long>, hipGraphExec_t> graphs;
    void **slots[A_COUNT] = {};  // the caller's slots of the last call
    std::vector<hipEvent_t> ev;

    double *d(Arr a) const { return static_cast<double *>(buf[a]); }
    int *mask() const { return static_cast<int *>(buf[A_TMASK]); }

    // the five prognostic pairs: set 0 = the reference's *n arrays
    double *T(int s) const { return s ? d(A_SSHA_T) : d(A_SSHN_T); }
    double *U(int s) const { return s ? d(A_UA) : d(A_UN); }
    double *V(int s) const { return s ? d(A_VA) : d(A_VN); }
    double *SU(int s) const { return s ? spare_su : d(A_SSHN_U); }
    double *SV(int s) const { return s ? spare_sv : d(A_SSHN_V); }

    StepIn<double> inputs(int s) const
    {
        StepIn<double> f;
        f.sshn_t = T(s);
        f.sshn_u = SU(s);
        f.sshn_v = SV(s);
        f.un = U(s);
        f.vn = V(s);
        f.ht = d(A_HT);
        f.hu = d(A_HU);
        f.hv = d(A_HV);
        f.ssha_u = d(A_SSHA_U);
        f.ssha_v = d(A_SSHA_V);
        f.area_t = d(A_AREA_T);
        f.area_u = d(A_AREA_U);
        f.area_v = d(A_AREA_V);
        f.dx_u = d(A_DX_U);
        f.dx_v = d(A_DX_V);
        f.dx_t = d(A_DX_T);
        f.dy_u = d(A_DY_U);
        f.dy_v = d(A_DY_V);
        f.dy_t = d(A_DY_T);
        f.fcor_u = fcor_u;
        f.fcor_v = fcor_v;
        f.tmask = mask();
        return f;
    }
    StepScalars<double> scalars() const
    {
        return StepScalars<double>{P.rdt, P.visc, P.g, 1.0 + P.cbfr * P.rdt};
    }

    ~Context()
    {
        for (auto &kv : graphs) (void)hipGraphExecDestroy(kv.second);
        for (hipEvent_t e : ev) (void)hipEventDestroy(e);
        for (void *p : buf)
            if (p) (void)hipFree(p);
        if (fcor_u) (void)hipFree(fcor_u);
        if (fcor_v) (void)hipFree(fcor_v);
        if (spare_su) (void)hipFree(spare_su);
        if (spare_sv) (void)hipFree(spare_sv);
        if (d_tide) (void)hipFree(d_tide);
        if (h_tide) (void)hipHostFree(h_tide);
        fl_u.release();
        fl_v.release();
    }
};

std::map<void **, std::unique_ptr<Context>> g_contexts;       // device flavour
std::map<double *, std::unique_ptr<Context>> g_host_contexts;  // host flavour


void upload_all(Context &C, double *const host[A_COUNT], hipStream_t s)
{
    const size_t n = (size_t)C.G.total;
    for (int a = 0; a < A_COUNT; a++) {
        const size_t bytes = n * (a == A_TMASK ? sizeof(int) : sizeof(double));
        NEMO_HIP_CHECK(hipMemcpyAsync(C.buf[a], host[a], bytes,
                                      hipMemcpyHostToDevice, s));
    }
    const int *tm = reinterpret_cast<const int *>(host[A_TMASK]);
    std::vector<double> fu = coriolis(host[A_GPHIU], C.G.total, C.P.omega, C.P.d2r);
    std::vector<double> fv = coriolis(host[A_GPHIV], C.G.total, C.P.omega, C.P.d2r);
    NEMO_HIP_CHECK(hipMemcpyAsync(C.fcor_u, fu.data(), n * sizeof(double),
                                  hipMemcpyHostToDevice, s));
    NEMO_HIP_CHECK(hipMemcpyAsync(C.fcor_v, fv.data(), n * sizeof(double),
                                  hipMemcpyHostToDevice, s));
    C.fl_u.upload(build_flather(tm, host[A_HU], C.G, C.P.g, true), s);
    C.fl_v.upload(build_flather(tm, host[A_HV], C.G, C.P.g, false), s);
    NEMO_HIP_CHECK(hipStreamSynchronize(s));
}

std::unique_ptr<Context> make_context(double *const host[A_COUNT],
                                      const Geometry &G, const Physics &P,
                                      int mode_request)
{
    check_geometry(G);
    // The fused form's mask condition and its 32-bit addressing are checked
    // before any device call.
    const int *tm = reinterpret_cast<const int *>(host[A_TMASK]);
    const bool addr_ok = offsets_fit_32(G, sizeof(double));
    if (mode_request == 2 && !addr_ok)
        NEMO_FATAL("fused form requested but total_size %d x 8 bytes exceeds the 4 GiB "
                   "its kernels address with 32-bit offsets", G.total);
    const bool mask_ok = mode_request != 1 && addr_ok && fused_allowed(tm, G);
    if (mode_request == 2 && !mask_ok)
        NEMO_FATAL("fused form requested but the mask has wet cells on row %d",
                   G.ye + 1);
    auto C = std::make_unique<Context>();
    C->G = G;
    C->P = P;
    const size_t n = (size_t)G.total;
    for (int a = 0; a < A_COUNT; a++)
        C->buf[a] = a == A_TMASK ? (void *)dev_alloc<int>(n)
                                 : (void *)dev_alloc<double>(n);
    C->fcor_u = dev_alloc<double>(n);
    C->fcor_v = dev_alloc<double>(n);
    hipStream_t s = lib_stream();
    upload_all(*C, host, s);

    C->mode = (mode_request == 1 || !mask_ok || !addr_ok) ? 1 : 2;
    if (C->mode == 2) {
        C->spare_su = dev_alloc<double>(n);
        C->spare_sv = dev_alloc<double>(n);
        NEMO_HIP_CHECK(hipMemcpyAsync(C->spare_su, C->d(A_SSHN_U), n * sizeof(double),
                                      hipMemcpyDeviceToDevice, s));
        NEMO_HIP_CHECK(hipMemcpyAsync(C->spare_sv, C->d(A_SSHN_V), n * sizeof(double),
                                      hipMemcpyDeviceToDevice, s));
        NEMO_HIP_CHECK(hipStreamSynchronize(s));
    }
    return C;
}

// ---- one step, unfused: the reference sweep sequence, in place ----------
void step_unfused(Context &C, int istep, hipStream_t s)
{
    const Geometry &G = C.G;
    const int W = G.W;
    const StepIn<double> f = C.inputs(0);
    const StepScalars<double> p = C.scalars();
    const Region interior{G.xs, G.xe, G.ys, G.ye};
    launch_continuity(s, f, C.d(A_SSHA_T), W, interior, C.P.rdt);
    launch_momentum_u(s, f, p, C.d(A_UA), W, Region{G.xs, G.xe - 1, G.ys, G.ye});
    launch_momentum_v(s, f, p, C.d(A_VA), W, Region{G.xs, G.xe, G.ys, G.ye - 1});
    launch_bc_ssh(s, C.mask(), C.d(A_SSHA_T), W, interior, tide_at(istep, C.P.rdt));
    launch_bc_solid(s, C.mask(), C.d(A_UA), W, 1,
                    Region{G.xs - 1, G.xe, G.ys - 1, G.ye + 1});
    launch_bc_solid(s, C.mask(), C.d(A_VA), W, W,
                    Region{G.xs - 1, G.xe + 1, G.ys - 1, G.ye});
    launch_flather(s, C.fl_u.view(), C.fl_v.view(), C.d(A_UA), C.d(A_VA), C.d(A_UA),
                   C.d(A_VA), C.d(A_SSHN_U), C.d(A_SSHN_V));
    launch_copy3(s, C.d(A_UN), C.d(A_UA), C.d(A_VN), C.d(A_VA), C.d(A_SSHN_T),
                 C.d(A_SSHA_T), W, Region{G.xs - 1, G.xe + 1, G.ys - 1, G.ye});
    launch_next_ssh(s, C.mask(), C.d(A_SSHN_T), C.d(A_AREA_T), C.d(A_AREA_U),
                    C.d(A_SSHN_U), W, 1, Region{G.xs, G.xe - 1, G.ys, G.ye});
    launch_next_ssh(s, C.mask(), C.d(A_SSHN_T), C.d(A_AREA_T), C.d(A_AREA_V),
                    C.d(A_SSHN_V), W, W, Region{G.xs, G.xe, G.ys, G.ye - 1});
}

// ---- one step, fused: state set `par` -> set 1-par -----------------------
void launch_fused_pair(Context &C, int par, hipStream_t s, const double *tide_tab,
                       int tide_k, double tide_val, hipEvent_t e0 = nullptr,
                       hipEvent_t e1 = nullptr)
{
    const int q = 1 - par;
    const StepOut<double> o{C.T(q), C.U(q), C.V(q), C.SU(q), C.SV(q)};
    const Bounds bd{C.G.xs, C.G.xe, C.G.ys, C.G.ye};
    if (e0) NEMO_HIP_CHECK(hipEventRecord(e0, s));
    if (g_fused_variant < 0)
        launch_fused_step(s, C.inputs(par), C.scalars(), o, C.G.W, bd, tide_tab,
                          tide_k, tide_val);
    else
        launch_fused_march(s, g_fused_variant, C.inputs(par), C.scalars(), o,
                           C.G.W, bd, tide_tab, tide_k, tide_val);
    if (e1) NEMO_HIP_CHECK(hipEventRecord(e1, s));
    // the previous step's velocities: set `par` once the sets are paired; on
    // the first step the after arrays themselves, as the reference has them
    const int prev = C.paired ? par : q;
    launch_flather(s, C.fl_u.view(), C.fl_v.view(), C.U(q), C.V(q), C.U(prev), C.V(prev),
                   C.SU(par), C.SV(par));
}

void run_hook(const Context &C, const StepHook *hook, hipStream_t s)
{
    if (!hook) return;
    const int p = C.parity;
    hook->enqueue(hook->user, s, C.T(p), C.U(p), C.V(p), C.SU(p), C.SV(p), C.d(A_HT),
                  C.mask(), C.d(A_AREA_T), C.d(A_AREA_U), C.d(A_AREA_V));
}

void step_fused(Context &C, int istep, hipStream_t s, const StepHook *hook = nullptr)
{
    launch_fused_pair(C, C.parity, s, nullptr, 0, tide_at(istep, C.P.rdt));
    C.parity = 1 - C.parity;
    run_hook(C, hook, s);
    if (!C.paired) {
        // After the first step both sets must agree wherever no sweep
        // writes, so that later steps may write into either set.
        const size_t bytes = (size_t)C.G.total * sizeof(double);
        const int cur = C.parity, old = 1 - cur;
        NEMO_HIP_CHECK(hipMemcpyAsync(C.T(old), C.T(cur), bytes, hipMemcpyDeviceToDevice, s));
        NEMO_HIP_CHECK(hipMemcpyAsync(C.U(old), C.U(cur), bytes, hipMemcpyDeviceToDevice, s));
        NEMO_HIP_CHECK(hipMemcpyAsync(C.V(old), C.V(cur), bytes, hipMemcpyDeviceToDevice, s));
        NEMO_HIP_CHECK(hipMemcpyAsync(C.SU(old), C.SU(cur), bytes, hipMemcpyDeviceToDevice, s));
        NEMO_HIP_CHECK(hipMemcpyAsync(C.SV(old), C.SV(cur), bytes, hipMemcpyDeviceToDevice, s));
        C.paired = true;
    }
}

void ensure_tide_table(Context &C, int n)
{
    if (n <= C.tide_cap) return;
    if (C.d_tide) NEMO_HIP_CHECK(hipFree(C.d_tide));
    if (C.h_tide) NEMO_HIP_CHECK(hipHostFree(C.h_tide));
    C.d_tide = dev_alloc<double>((size_t)n);
    NEMO_HIP_CHECK(hipHostMalloc((void **)&C.h_tide, (size_t)n * sizeof(double),
                                 hipHostMallocDefault));
    C.tide_cap = n;
    // graphs captured against the old table are stale
    for (auto &kv : C.graphs) NEMO_HIP_CHECK(hipGraphExecDestroy(kv.second));
    C.graphs.clear();
}

hipGraphExec_t fused_graph(Context &C, int n, int par, hipStream_t s,
                           const StepHook *hook)
{
    auto key = std::make_tuple(n, par, g_fused_variant, hook ? hook->key : 0ull);
    auto it = C.graphs.find(key);
    if (it != C.graphs.end()) return it->second;
    hipGraph_t graph;
    const int parity = C.parity;
    NEMO_HIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    for (int k = 0; k < n; k++) {
        launch_fused_pair(C, (par + k) & 1, s, C.d_tide, k, 0.0);
        C.parity = (par + k + 1) & 1;  // the set this step wrote, for the hook
        run_hook(C, hook, s);
    }
    C.parity = parity;
    NEMO_HIP_CHECK(hipStreamEndCapture(s, &graph));
    hipGraphExec_t exec;
    NEMO_HIP_CHECK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
    NEMO_HIP_CHECK(hipGraphDestroy(graph));
    C.graphs[key] = exec;
    return exec;
}

void run_steps(Context &C, int nsteps, int istep, hipStream_t s,
               const StepHook *hook = nullptr)
{
    if (nsteps <= 0) return;
    if (C.mode == 1) {
        if (hook) NEMO_FATAL("steps with a hook need the fused form");
        for (int k = 0; k < nsteps; k++) step_unfused(C, istep + k, s);
        return;
    }
    if (!C.paired) {  // first step sets up the pairing
        step_fused(C, istep, s, hook);
        istep++;
        nsteps--;
        if (nsteps == 0) return;
    }
    ensure_tide_table(C, nsteps);
    NEMO_HIP_CHECK(hipStreamSynchronize(s));  // h_tide may still be in flight
    for (int k = 0; k < nsteps; k++) C.h_tide[k] = tide_at(istep + k, C.P.rdt);
    NEMO_HIP_CHECK(hipMemcpyAsync(C.d_tide, C.h_tide, (size_t)nsteps * sizeof(double),
                                  hipMemcpyHostToDevice, s));
    if (g_kernel_timing) {
        if ((int)C.ev.size() < 2 * nsteps) {
            while ((int)C.ev.size() < 2 * nsteps) {
                hipEvent_t e;
                NEMO_HIP_CHECK(hipEventCreate(&e));
                C.ev.push_back(e);
            }
        }
        const int parity = C.parity;
        for (int k = 0; k < nsteps; k++) {
            launch_fused_pair(C, (parity + k) & 1, s, C.d_tide, k, 0.0, C.ev[2 * k],
                              C.ev[2 * k + 1]);
            C.parity = (parity + k + 1) & 1;
            run_hook(C, hook, s);
        }
        C.parity = parity;
        NEMO_HIP_CHECK(hipStreamSynchronize(s));
        double sum = 0.0;
        for (int k = 0; k < nsteps; k++) {
            float ms = 0.f;
            NEMO_HIP_CHECK(hipEventElapsedTime(&ms, C.ev[2 * k], C.ev[2 * k + 1]));
            sum += ms;
        }
        g_fused_ms = sum / nsteps;
    } else {
        NEMO_HIP_CHECK(hipGraphLaunch(fused_graph(C, nsteps, C.parity, s, hook), s));
    }
    C.parity = (C.parity + nsteps) & 1;
}

// Points the caller's slots at the current buffers.
void publish(const Context &C, void **slot[A_COUNT])
{
    for (int a = 0; a < A_COUNT; a++) *slot[a] = C.buf[a];
    if (C.mode == 2) {
        const int p = C.parity;
        *slot[A_SSHA_T] = *slot[A_SSHN_T] = C.T(p);
        *slot[A_UA] = *slot[A_UN] = C.U(p);
        *slot[A_VA] = *slot[A_VN] = C.V(p);
        *slot[A_SSHN_U] = C.SU(p);
        *slot[A_SSHN_V] = C.SV(p);
    }
}

// The read / write callbacks copy nx*ny doubles between two buffers; a
// negative extent or a null buffer is a caller error, reported like the
// others (message, exit(1)).
void check_copy(const char *who, const void *from, const void *to, int nx, int ny)
{
    if (nx < 0 || ny < 0) NEMO_FATAL("%s: negative extent %d x %d", who, nx, ny);
    if ((long)nx * ny > 0 && (!from || !to))
        NEMO_FATAL("%s: null buffer (from %p, to %p)", who, from, to);
}

}  // namespace
}  // namespace nemo

using namespace nemo;

#define HOST_LIST                                                              \
    {ssha_t, sshn_t, sshn_u, sshn_v, hu, hv, un, vn, ua, ht, ssha_u, va,        \
     ssha_v, reinterpret_cast<double *>(tmask), area_t, area_u, area_v, dx_u,   \
     dx_v, dx_t, dy_u, dy_v, dy_t, gphiu, gphiv}
#define SLOT_LIST                                                              \
    {ssha_t_device, sshn_t_device, sshn_u_device, sshn_v_device, hu_device,     \
     hv_device, un_device, vn_device, ua_device, ht_device, ssha_u_device,     \
     va_device, ssha_v_device, tmask_device, area_t_device, area_u_device,     \
     area_v_device, dx_u_device, dx_v_device, dx_t_device, dy_u_device,        \
     dy_v_device, dy_t_device, gphiu_device, gphiv_device}

static bool same_geometry(const Geometry &a, const Geometry &b)
{
    return a.W == b.W && a.total == b.total && a.xs == b.xs && a.xe == b.xe &&
           a.ys == b.ys && a.ye == b.ye;
}

static bool same_physics(const Physics &a, const Physics &b)
{
    return a.rdt == b.rdt && a.cbfr == b.cbfr && a.visc == b.visc && a.omega == b.omega &&
           a.d2r == b.d2r && a.g == b.g;
}

// New physical scalars for an existing context: the Coriolis factors and the
// Flather coefficients derive from them, and captured graphs carry them as
// kernel arguments. The inputs they derive from (gphiu/v, tmask, hu/hv) are
// read back from the device, where the caller's last upload or
// c_write_to_device left them; the step never writes them.
static void refresh_physics(Context &C, const Physics &P)
{
    hipStream_t s = lib_stream();
    NEMO_HIP_CHECK(hipStreamSynchronize(s));
    const size_t n = (size_t)C.G.total;
    std::vector<double> gu(n), gv(n), hu(n), hv(n);
    std::vector<int> tm(n);
    NEMO_HIP_CHECK(hipMemcpy(gu.data(), C.buf[A_GPHIU], n * sizeof(double), hipMemcpyDeviceToHost));
    NEMO_HIP_CHECK(hipMemcpy(gv.data(), C.buf[A_GPHIV], n * sizeof(double), hipMemcpyDeviceToHost));
    NEMO_HIP_CHECK(hipMemcpy(hu.data(), C.buf[A_HU], n * sizeof(double), hipMemcpyDeviceToHost));
    NEMO_HIP_CHECK(hipMemcpy(hv.data(), C.buf[A_HV], n * sizeof(double), hipMemcpyDeviceToHost));
    NEMO_HIP_CHECK(hipMemcpy(tm.data(), C.buf[A_TMASK], n * sizeof(int), hipMemcpyDeviceToHost));
    C.P = P;
    const std::vector<double> fu = coriolis(gu.data(), C.G.total, P.omega, P.d2r);
    const std::vector<double> fv = coriolis(gv.data(), C.G.total, P.omega, P.d2r);
    NEMO_HIP_CHECK(hipMemcpy(C.fcor_u, fu.data(), n * sizeof(double), hipMemcpyHostToDevice));
    NEMO_HIP_CHECK(hipMemcpy(C.fcor_v, fv.data(), n * sizeof(double), hipMemcpyHostToDevice));
    C.fl_u.upload(build_flather(tm.data(), hu.data(), C.G, P.g, true), s);
    C.fl_v.upload(build_flather(tm.data(), hv.data(), C.G, P.g, false), s);
    for (auto &kv : C.graphs) NEMO_HIP_CHECK(hipGraphExecDestroy(kv.second));
    C.graphs.clear();
}

static Context &device_context(double *const host[A_COUNT], void **slot[A_COUNT],
                               const Geometry &G, const Physics &P)
{
    auto it = g_contexts.find(slot[A_SSHA_T]);
    if (it == g_contexts.end())
        it = g_contexts.emplace(slot[A_SSHA_T], make_context(host, G, P, g_mode)).first;
    Context &C = *it->second;
    if (!same_geometry(C.G, G))
        NEMO_FATAL("geometry changed between calls for the same device slots");
    if (!same_physics(C.P, P)) refresh_physics(C, P);
    return C;
}

namespace nemo {

void psy_steps_with_hook(void **ssha_t_slot, int nsteps, int istep, const StepHook &hook)
{
    auto it = g_contexts.find(ssha_t_slot);
    if (it == g_contexts.end())
        NEMO_FATAL("steps with a hook: no device context for these slots (step the model "
                   "once with c_invoke_time_step first)");
    Context &C = *it->second;
    if (C.mode != 2) NEMO_FATAL("steps with a hook need the fused form");
    if (nsteps < 0) NEMO_FATAL("nsteps must be >= 0, got %d", nsteps);
    if (g_kernel_timing) g_fused_ms = -1.0;
    run_steps(C, nsteps, istep, lib_stream(), &hook);
    publish(C, C.slots);
}

}  // namespace nemo

extern "C" {

void c_invoke_time_step(NEMO_HOST_ARRAYS, NEMO_DEVICE_SLOTS, int istep,
                        int internal_xstart, int internal_xstop,
                        int internal_ystart, int internal_ystop, int width,
                        int total_size, double rdt, double cbfr, double visc,
                        double omega, double d2r, double g)
{
    c_invoke_time_steps(1, ssha_t, sshn_t, sshn_u, sshn_v, hu, hv, un, vn, ua,
                        ht, ssha_u, va, ssha_v, tmask, area_t, area_u, area_v,
                        dx_u, dx_v, dx_t, dy_u, dy_v, dy_t, gphiu, gphiv,
                        ssha_t_device, sshn_t_device, sshn_u_device,
                        sshn_v_device, hu_device, hv_device, un_device,
                        vn_device, ua_device, ht_device, ssha_u_device,
                        va_device, ssha_v_device, tmask_device, area_t_device,
                        area_u_device, area_v_device, dx_u_device, dx_v_device,
                        dx_t_device, dy_u_device, dy_v_device, dy_t_device,
                        gphiu_device, gphiv_device, istep, internal_xstart,
                        internal_xstop, internal_ystart, internal_ystop, width,
                        total_size, rdt, cbfr, visc, omega, d2r, g);
}

void c_invoke_time_steps(int nsteps, NEMO_HOST_ARRAYS, NEMO_DEVICE_SLOTS,
                         int istep, int internal_xstart, int internal_xstop,
                         int internal_ystart, int internal_ystop, int width,
                         int total_size, double rdt, double cbfr, double visc,
                         double omega, double d2r, double g)
{
    double *host[A_COUNT] = HOST_LIST;
    void **slot[A_COUNT] = SLOT_LIST;
    const Geometry G{internal_xstart, internal_xstop, internal_ystart,
                     internal_ystop, width, total_size};
    const Physics P{rdt, cbfr, visc, omega, d2r, g};
    Context &C = device_context(host, slot, G, P);
    for (int a = 0; a < A_COUNT; a++) C.slots[a] = slot[a];
    if (nsteps < 0) NEMO_FATAL("nsteps must be >= 0, got %d", nsteps);
    if (g_kernel_timing && C.mode == 2) g_fused_ms = -1.0;
    hipStream_t s = lib_stream();
    if (nsteps == 1 && C.mode == 2 && C.paired && !g_kernel_timing) {
        step_fused(C, istep, s);  // plain launches for single steps
    } else {
        run_steps(C, nsteps, istep, s);
    }
    publish(C, slot);
}

void c_read_from_device(void *from, void *to, int startx, int starty, int nx,
                        int ny, bool blocking)
{
    (void)startx;
    (void)starty;
    (void)blocking;
    check_copy("c_read_from_device", from, to, nx, ny);
    NEMO_HIP_CHECK(hipStreamSynchronize(lib_stream()));
    NEMO_HIP_CHECK(hipMemcpy(to, from, (size_t)nx * ny * sizeof(double),
                             hipMemcpyDeviceToHost));
}

void c_write_to_device(void *from, void *to, int startx, int starty, int nx,
                       int ny, bool blocking)
{
    (void)startx;
    (void)starty;
    (void)blocking;
    check_copy("c_write_to_device", from, to, nx, ny);
    NEMO_HIP_CHECK(hipStreamSynchronize(lib_stream()));
    NEMO_HIP_CHECK(hipMemcpy(to, from, (size_t)nx * ny * sizeof(double),
                             hipMemcpyHostToDevice));
}

void c_invoke_time_step_host(NEMO_HOST_ARRAYS, int istep, int internal_xstart,
                             int internal_xstop, int internal_ystart,
                             int internal_ystop, int width, double rdt,
                             double cbfr, double visc, double omega,
                             double d2r, double g)
{
    double *host[A_COUNT] = HOST_LIST;
    // the host flavour carries no size: the sweeps reach row ystop+1
    const int total = width * (internal_ystop + 2);
    const Geometry G{internal_xstart, internal_xstop, internal_ystart,
                     internal_ystop, width, total};
    const Physics P{rdt, cbfr, visc, omega, d2r, g};
    hipStream_t s = lib_stream();
    // A context is reused only for the same geometry: its Flather segment
    // lists depend on every bound. upload_all rebuilds the derived arrays
    // from this call's host arrays and scalars.
    auto it = g_host_contexts.find(ssha_t);
    if (it != g_host_contexts.end() && !same_geometry(it->second->G, G)) {
        NEMO_HIP_CHECK(hipStreamSynchronize(s));
        g_host_contexts.erase(it);
        it = g_host_contexts.end();
    }
    if (it == g_host_contexts.end()) {
        it = g_host_contexts.emplace(ssha_t, make_context(host, G, P, 1)).first;
    } else {
        it->second->P = P;
        upload_all(*it->second, host, s);
    }
    Context &C = *it->second;
    step_unfused(C, istep, s);
    const size_t bytes = (size_t)total * sizeof(double);
    const Arr outs[] = {A_SSHA_T, A_SSHN_T, A_SSHN_U, A_SSHN_V,
                        A_UN, A_VN, A_UA, A_VA};
    for (Arr a : outs)
        NEMO_HIP_CHECK(hipMemcpyAsync(host[a], C.buf[a], bytes,
                                      hipMemcpyDeviceToHost, s));
    NEMO_HIP_CHECK(hipStreamSynchronize(s));
}

void nemo_hip_set_tide(double amp, double omega_tide)
{
    g_amp_tide = amp;
    g_omega_tide = omega_tide;
}

void nemo_hip_set_mode(int mode)
{
    if (mode < 0 || mode > 2) NEMO_FATAL("unknown mode %d", mode);
    g_mode = mode;
}

int nemo_hip_context_mode(void **ssha_t_device)
{
    auto it = g_contexts.find(ssha_t_device);
    return it == g_contexts.end() ? 0 : it->second->mode;
}

void nemo_hip_synchronize(void) { NEMO_HIP_CHECK(hipStreamSynchronize(lib_stream())); }

void nemo_hip_set_kernel_timing(int on) { g_kernel_timing = on != 0; }

int nemo_hip_fused_variant_count(void) { return march_variant_count(); }

void nemo_hip_set_fused_variant(int variant)
{
    if (variant < -1 || variant >= march_variant_count())
        NEMO_FATAL("unknown fused variant %d", variant);
    g_fused_variant = variant;
}

int nemo_hip_fused_variant(void) { return g_fused_variant; }

void nemo_hip_fused_variant_shape(int variant, int *block_width, int *rows,
                                  int *prefetch)
{
    if (variant == -1) {
        *block_width = 64;
        *rows = 4;
        *prefetch = 0;
        return;
    }
    march_variant_shape(variant, block_width, rows, prefetch);
}

const char *nemo_hip_fused_variant_kernel(int variant)
{
    return variant == -1 ? "k_fused_step" : march_variant_kernel(variant);
}

double nemo_hip_fused_kernel_ms(void) { return g_fused_ms; }

void nemo_hip_release(void **ssha_t_device)
{
    NEMO_HIP_CHECK(hipStreamSynchronize(lib_stream()));
    g_contexts.erase(ssha_t_device);
}

void nemo_hip_release_all(void)
{
    NEMO_HIP_CHECK(hipStreamSynchronize(lib_stream()));
    g_contexts.clear();
    g_host_contexts.clear();
}

int nemo_hip_device_info(char *buf, int buflen)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) {
        if (buf && buflen > 0) std::snprintf(buf, (size_t)buflen, "no HIP device");
        return 0;
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, 0) != hipSuccess) return 0;
    if (buf && buflen > 0)
        std::snprintf(buf, (size_t)buflen, "%s %s %zu MiB %d CUs", prop.name,
                      prop.gcnArchName, prop.totalGlobalMem >> 20,
                      prop.multiProcessorCount);
    return n;
}

}  // extern "C"
