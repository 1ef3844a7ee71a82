// C ABI of the mixed-precision step (no counterpart in the reference): an
// opaque handle owns FP32 copies of the 25 arrays (ssh as a compensated
// hi/lo pair), the Flather segment lists and captured step graphs. The state
// ping-pongs between two buffer sets like the FP64 fused form.
#include <cmath>
#include <map>
#include <memory>
#include <tuple>
#include <vector>

#include "nemo_hip_common.h"
#include "nemo_host.h"
#include "nemo_mixed.h"
#include "nemolite2d_hip.h"

namespace nemo {
namespace {

struct MixedContext {
    Geometry G{};
    Physics P{};
    double amp = 0.0, omega_tide = 0.0;
    int rows = 32;
    int cols = 61;                   // output columns per wavefront
    int prefetch = 0;                // 1 / 2: loads issued two / three rows ahead (rows 4 or 8)
    float *cst[A_COUNT] = {};        // constant arrays (FP32); tmask in tm
    signed char *tm = nullptr;       // T mask, one byte a point
    float *fcor_u = nullptr, *fcor_v = nullptr;
    float *t_hi[2] = {}, *t_lo[2] = {}, *u[2] = {}, *v[2] = {}, *su[2] = {}, *sv[2] = {};
    DeviceFlather fl_u, fl_v;
    int parity = 0;
    bool paired = false;
    double *d_tide = nullptr, *h_tide = nullptr;
    int tide_cap = 0;
    // n, parity, rows, cols, prefetch
    std::map<std::tuple<int, int, int, int, int>, hipGraphExec_t> graphs;
    std::vector<hipEvent_t> ev;

    MixedIn inputs(int s) const
    {
        MixedIn f;
        f.ssh_hi = t_hi[s];
        f.ssh_lo = t_lo[s];
        f.sshn_u = su[s];
        f.sshn_v = sv[s];
        f.un = u[s];
        f.vn = v[s];
        f.ht = cst[A_HT];
        f.hu = cst[A_HU];
        f.hv = cst[A_HV];
        f.ssha_u = cst[A_SSHA_U];
        f.ssha_v = cst[A_SSHA_V];
        f.area_t = cst[A_AREA_T];
        f.area_u = cst[A_AREA_U];
        f.area_v = cst[A_AREA_V];
        f.dx_u = cst[A_DX_U];
        f.dx_v = cst[A_DX_V];
        f.dx_t = cst[A_DX_T];
        f.dy_u = cst[A_DY_U];
        f.dy_v = cst[A_DY_V];
        f.dy_t = cst[A_DY_T];
        f.fcor_u = fcor_u;
        f.fcor_v = fcor_v;
        f.tmask = tm;
        return f;
    }
    StepScalars<float> scalars() const
    {
        return StepScalars<float>{(float)P.rdt, (float)P.visc, (float)P.g,
                                  (float)(1.0 + P.cbfr * P.rdt)};
    }
    double tide_at(int istep) const
    {
        const double rtime = istep * P.rdt;  // FP64 tidal phase
        return amp * std::sin(omega_tide * rtime);
    }

    ~MixedContext()
    {
        for (auto &kv : graphs) (void)hipGraphExecDestroy(kv.second);
        for (hipEvent_t e : ev) (void)hipEventDestroy(e);
        for (float *p : cst)
            if (p) (void)hipFree(p);
        float *more[] = {fcor_u, fcor_v, t_hi[0], t_hi[1], t_lo[0], t_lo[1], u[0], u[1],
                         v[0], v[1], su[0], su[1], sv[0], sv[1]};
        for (float *p : more)
            if (p) (void)hipFree(p);
        if (tm) (void)hipFree(tm);
        if (d_tide) (void)hipFree(d_tide);
        if (h_tide) (void)hipHostFree(h_tide);
        fl_u.release();
        fl_v.release();
    }
};

float *upload_f32(const double *h, size_t n, hipStream_t s)
{
    std::vector<float> tmp(n);
    for (size_t k = 0; k < n; k++) tmp[k] = (float)h[k];
    float *d = dev_alloc<float>(n);
    NEMO_HIP_CHECK(hipMemcpyAsync(d, tmp.data(), n * sizeof(float), hipMemcpyHostToDevice, s));
    NEMO_HIP_CHECK(hipStreamSynchronize(s));
    return d;
}

void launch_pair(MixedContext &C, int par, hipStream_t s, const double *tide_tab, int k,
                 double tide_val, hipEvent_t e0 = nullptr, hipEvent_t e1 = nullptr)
{
    const int q = 1 - par;
    const MixedOut o{C.t_hi[q], C.t_lo[q], C.u[q], C.v[q], C.su[q], C.sv[q]};
    if (e0) NEMO_HIP_CHECK(hipEventRecord(e0, s));
    launch_mixed_step(s, C.rows, C.cols, C.prefetch, C.inputs(par), C.scalars(), o, C.G.W,
                      Bounds{C.G.xs, C.G.xe, C.G.ys, C.G.ye}, tide_tab, k, tide_val);
    if (e1) NEMO_HIP_CHECK(hipEventRecord(e1, s));
    const int prev = C.paired ? par : q;  // the previous step's velocities
    launch_flather_f32(s, C.fl_u.view(), C.fl_v.view(), C.u[q], C.v[q], C.u[prev],
                       C.v[prev], C.su[par], C.sv[par]);
}

void first_step(MixedContext &C, int istep, hipStream_t s)
{
    launch_pair(C, C.parity, s, nullptr, 0, C.tide_at(istep));
    C.parity = 1 - C.parity;
    const size_t bytes = (size_t)C.G.total * sizeof(float);
    const int cur = C.parity, old = 1 - cur;
    float **sets[] = {C.t_hi, C.t_lo, C.u, C.v, C.su, C.sv};
    for (float **a : sets)
        NEMO_HIP_CHECK(hipMemcpyAsync(a[old], a[cur], bytes, hipMemcpyDeviceToDevice, s));
    C.paired = true;
}

void fill_tide(MixedContext &C, int n, int istep, hipStream_t s)
{
    if (n > C.tide_cap) {
        if (C.d_tide) NEMO_HIP_CHECK(hipFree(C.d_tide));
        if (C.h_tide) NEMO_HIP_CHECK(hipHostFree(C.h_tide));
        C.d_tide = dev_alloc<double>((size_t)n);
        NEMO_HIP_CHECK(hipHostMalloc((void **)&C.h_tide, (size_t)n * sizeof(double),
                                     hipHostMallocDefault));
        C.tide_cap = n;
        for (auto &kv : C.graphs) NEMO_HIP_CHECK(hipGraphExecDestroy(kv.second));
        C.graphs.clear();
    }
    NEMO_HIP_CHECK(hipStreamSynchronize(s));
    for (int k = 0; k < n; k++) C.h_tide[k] = C.tide_at(istep + k);
    NEMO_HIP_CHECK(hipMemcpyAsync(C.d_tide, C.h_tide, (size_t)n * sizeof(double),
                                  hipMemcpyHostToDevice, s));
}

}  // namespace
}  // namespace nemo

using namespace nemo;

extern "C" {

void *nemo_mp_create(NEMO_HOST_ARRAYS, int internal_xstart, int internal_xstop,
                     int internal_ystart, int internal_ystop, int width, int total_size,
                     double rdt, double cbfr, double visc, double omega, double d2r,
                     double g, double tide_amp, double tide_omega)
{
    double *host[A_COUNT] = {ssha_t, sshn_t, sshn_u, sshn_v, hu, hv, un, vn, ua, ht,
                             ssha_u, va, ssha_v, reinterpret_cast<double *>(tmask),
                             area_t, area_u, area_v, dx_u, dx_v, dx_t, dy_u, dy_v,
                             dy_t, gphiu, gphiv};
    const Geometry G{internal_xstart, internal_xstop, internal_ystart, internal_ystop,
                     width, total_size};
    check_geometry(G);
    if (!offsets_fit_32(G, sizeof(float)))
        NEMO_FATAL("mixed form: total_size %d x 4 bytes exceeds the 4 GiB its kernels "
                   "address with 32-bit offsets", total_size);
    if (!fused_allowed(tmask, G))
        NEMO_FATAL("mixed form: the mask has wet cells on row %d", G.ye + 1);
    const std::vector<signed char> tm8 = mask_bytes(tmask, (size_t)total_size);
    auto C = std::make_unique<MixedContext>();
    C->G = G;
    C->P = Physics{rdt, cbfr, visc, omega, d2r, g};
    C->amp = tide_amp;
    C->omega_tide = tide_omega;
    hipStream_t s = lib_stream();
    const size_t n = (size_t)total_size;
    const Arr consts[] = {A_HT, A_HU, A_HV, A_SSHA_U, A_SSHA_V, A_AREA_T, A_AREA_U,
                          A_AREA_V, A_DX_U, A_DX_V, A_DX_T, A_DY_U, A_DY_V, A_DY_T};
    for (Arr a : consts) C->cst[a] = upload_f32(host[a], n, s);
    C->tm = dev_alloc<signed char>(n);
    NEMO_HIP_CHECK(hipMemcpyAsync(C->tm, tm8.data(), n, hipMemcpyHostToDevice, s));
    NEMO_HIP_CHECK(hipStreamSynchronize(s));  // the copy reads tm8 (pageable host memory)
    std::vector<double> fu = coriolis(gphiu, total_size, omega, d2r);
    std::vector<double> fv = coriolis(gphiv, total_size, omega, d2r);
    C->fcor_u = upload_f32(fu.data(), n, s);
    C->fcor_v = upload_f32(fv.data(), n, s);
    // prognostic sets: set 0 from the *n arrays, set 1 from the *a arrays
    // (face ssh: both from sshn_u / sshn_v)
    std::vector<float> hi(n), lo(n);
    for (int set = 0; set < 2; set++) {
        const double *t = set ? ssha_t : sshn_t;
        for (size_t k = 0; k < n; k++) {
            hi[k] = (float)t[k];
            lo[k] = (float)(t[k] - (double)hi[k]);
        }
        C->t_hi[set] = dev_alloc<float>(n);
        C->t_lo[set] = dev_alloc<float>(n);
        NEMO_HIP_CHECK(hipMemcpyAsync(C->t_hi[set], hi.data(), n * sizeof(float),
                                      hipMemcpyHostToDevice, s));
        NEMO_HIP_CHECK(hipMemcpyAsync(C->t_lo[set], lo.data(), n * sizeof(float),
                                      hipMemcpyHostToDevice, s));
        NEMO_HIP_CHECK(hipStreamSynchronize(s));
        C->u[set] = upload_f32(set ? ua : un, n, s);
        C->v[set] = upload_f32(set ? va : vn, n, s);
        C->su[set] = upload_f32(sshn_u, n, s);
        C->sv[set] = upload_f32(sshn_v, n, s);
    }
    C->fl_u.upload(build_flather(tmask, hu, G, g, true), s);
    C->fl_v.upload(build_flather(tmask, hv, G, g, false), s);
    return C.release();
}

void nemo_mp_set_rows(void *h, int rows)
{
    if (rows != 4 && rows != 8 && rows != 16 && rows != 32 && rows != 48 && rows != 64 &&
        rows != 80 && rows != 96 && rows != 128)
        NEMO_FATAL("mixed form: rows per band must be 4, 8, 16, 32, 48, 64, 80, 96 or 128 "
                   "(got %d)", rows);
    MixedContext &C = *static_cast<MixedContext *>(h);
    if (C.prefetch && rows != 4 && rows != 8)
        NEMO_FATAL("mixed form: the prefetching kernel takes 4 or 8 rows per band (got %d)",
                   rows);
    C.rows = rows;
}

void nemo_mp_set_columns(void *h, int cols)
{
    if (cols != 61 && cols != 56 && cols != 48)
        NEMO_FATAL("mixed form: output columns per wavefront must be 61, 56 or 48 (got %d)",
                   cols);
    static_cast<MixedContext *>(h)->cols = cols;
}

void nemo_mp_set_prefetch(void *h, int prefetch)
{
    MixedContext &C = *static_cast<MixedContext *>(h);
    if (prefetch < 0 || prefetch > 2)
        NEMO_FATAL("mixed form: prefetch must be 0, 1 or 2 (got %d)", prefetch);
    if (prefetch && C.rows != 4 && C.rows != 8)
        NEMO_FATAL("mixed form: the prefetching kernel takes 4 or 8 rows per band "
                   "(rows are %d)", C.rows);
    C.prefetch = prefetch;
}

static void mp_run(MixedContext &C, int nsteps, int istep, bool timed, double *ms)
{
    hipStream_t s = lib_stream();
    if (nsteps <= 0) return;
    if (!C.paired) {
        first_step(C, istep, s);
        istep++;
        if (--nsteps == 0) return;
    }
    fill_tide(C, nsteps, istep, s);
    if (timed) {
        while ((int)C.ev.size() < 2 * nsteps) {
            hipEvent_t e;
            NEMO_HIP_CHECK(hipEventCreate(&e));
            C.ev.push_back(e);
        }
        for (int k = 0; k < nsteps; k++)
            launch_pair(C, (C.parity + k) & 1, s, C.d_tide, k, 0.0, C.ev[2 * k],
                        C.ev[2 * k + 1]);
        NEMO_HIP_CHECK(hipStreamSynchronize(s));
        double sum = 0.0;
        for (int k = 0; k < nsteps; k++) {
            float t = 0.f;
            NEMO_HIP_CHECK(hipEventElapsedTime(&t, C.ev[2 * k], C.ev[2 * k + 1]));
            sum += t;
        }
        *ms = sum / nsteps;
    } else {
        auto key = std::make_tuple(nsteps, C.parity, C.rows, C.cols, C.prefetch);
        auto it = C.graphs.find(key);
        if (it == C.graphs.end()) {
            hipGraph_t graph;
            NEMO_HIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            for (int k = 0; k < nsteps; k++)
                launch_pair(C, (C.parity + k) & 1, s, C.d_tide, k, 0.0);
            NEMO_HIP_CHECK(hipStreamEndCapture(s, &graph));
            hipGraphExec_t exec;
            NEMO_HIP_CHECK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
            NEMO_HIP_CHECK(hipGraphDestroy(graph));
            it = C.graphs.emplace(key, exec).first;
        }
        NEMO_HIP_CHECK(hipGraphLaunch(it->second, s));
    }
    C.parity = (C.parity + nsteps) & 1;
}

void nemo_mp_steps(void *h, int nsteps, int istep)
{
    mp_run(*static_cast<MixedContext *>(h), nsteps, istep, false, nullptr);
}

double nemo_mp_kernel_ms(void *h, int nsteps, int istep)
{
    double ms = -1.0;
    mp_run(*static_cast<MixedContext *>(h), nsteps, istep, true, &ms);
    return ms;
}

void nemo_mp_read(void *h, int field, double *out)
{
    MixedContext &C = *static_cast<MixedContext *>(h);
    hipStream_t s = lib_stream();
    NEMO_HIP_CHECK(hipStreamSynchronize(s));
    const int p = C.parity;
    const size_t n = (size_t)C.G.total;
    const float *src = nullptr, *lo = nullptr;
    switch (field) {
    case 0: src = C.t_hi[p]; lo = C.t_lo[p]; break;
    case 1: src = C.u[p]; break;
    case 2: src = C.v[p]; break;
    case 3: src = C.su[p]; break;
    case 4: src = C.sv[p]; break;
    default: NEMO_FATAL("nemo_mp_read: unknown field %d", field);
    }
    std::vector<float> a(n), b(lo ? n : 0);
    NEMO_HIP_CHECK(hipMemcpy(a.data(), src, n * sizeof(float), hipMemcpyDeviceToHost));
    if (lo) NEMO_HIP_CHECK(hipMemcpy(b.data(), lo, n * sizeof(float), hipMemcpyDeviceToHost));
    for (size_t k = 0; k < n; k++) out[k] = lo ? (double)a[k] + (double)b[k] : (double)a[k];
}

void nemo_mp_destroy(void *h)
{
    NEMO_HIP_CHECK(hipStreamSynchronize(lib_stream()));
    delete static_cast<MixedContext *>(h);
}

}  // extern "C"
