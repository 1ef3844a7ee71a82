// C ABI of the batched ensemble (no counterpart in the reference): many
// independent NEMOLite2D domains of uneven size, bathymetry and tide,
// advanced together by one launch per step (plus one Flather
// launch over every domain's boundary segments).
//
// Memory: one arena per field holding every domain back to back, each at an
// element offset rounded to 64 (rows stay 512-B aligned in FP64); the state
// ping-pongs between two sets of arenas like the single-domain fused form.
// Precision 64 runs the FP64 step (bitwise the single-domain fused form of
// each domain); precision 32 runs the mixed form (FP32 fields, compensated
// ssh, FP64 tide), bitwise the single-domain mixed form of each domain.
//
// Launch shapes (rows per band x columns per wavefront x prefetch depth x
// launch form: one wavefront per tile, or persistent with so many wavefronts
// per workgroup and resident wavefronts per CU) are enumerated at creation;
// nemo_ens_autotune times each
// on the ensemble itself with HIP events and keeps the fastest, restoring the
// state it advanced.
#include <cmath>
#include <cstring>
#include <limits>
#include <map>
#include <memory>
#include <tuple>
#include <vector>

#include "nemo_ensemble.h"
#include "nemo_hip_common.h"
#include "nemo_host.h"
#include "nemo_wave.h"
#include "nemolite2d_hip.h"

namespace nemo {
namespace {

constexpr int kProg = 6;  // prognostic arenas per set: t (hi), t lo, u, v, su, sv
// Arena k starts k * 4352 bytes (17 x 256) into its allocation. hipMalloc
// hands out large allocations at the same large alignment, so without the
// shift the ~21 arrays a tile reads at one element offset all start on the
// same L2 channel and HBM bank; the shift spreads them (1-2% faster step on
// the 256-domain ensemble, both precisions). A multiple of 256 keeps every
// row's 64-byte alignment.
constexpr size_t kArenaStagger = 4352;
// tile shapes of the launch list: rows per band x owned columns per wavefront
constexpr int kEnsRows[] = {4, 8, 16, 32, 64};
constexpr int kEnsCols[] = {61, 56, 0};  // 0: per domain, 56 or 61

struct EnsContext {
    int prec = 64;
    int ndom = 0;
    Physics P{};
    std::vector<Geometry> geo;
    std::vector<EnsDomain> hdom;
    std::vector<double> amp, omega_tide;
    size_t total = 0;        // elements per arena
    long long interior = 0;  // interior T points over all domains
    int *tm = nullptr;
    signed char *tm8 = nullptr;  // the mask in one byte a point (FP32 ensembles)
    // constant arenas by C-ABI index (A_HT, ..., A_DY_T) + Coriolis factors
    void *cst[A_COUNT] = {};
    void *fcor_u = nullptr, *fcor_v = nullptr;
    // prognostic arenas: prog[set][k], k = t, t_lo (FP32 only), u, v, su, sv
    void *prog[2][kProg] = {};
    DeviceFlather fl_u, fl_v;
    int parity = 0;
    bool paired = false;
    std::vector<EnsShape> shapes;
    int shape = 0;
    std::map<std::pair<int, int>, std::pair<EnsTile *, int>> lists;  // (rows, cols) -> list
    double *d_tide = nullptr, *h_tide = nullptr;
    int tide_cap = 0;                                 // steps
    std::map<std::tuple<int, int, int, int>, hipGraphExec_t> graphs;  // n, parity, shape, row0
    std::vector<hipEvent_t> ev;
    double last_ms = -1.0;
    // arena allocations as hipMalloc returned them; arena k starts
    // k * kArenaStagger bytes into its allocation
    std::vector<void *> raw;

    size_t esize() const { return prec == 64 ? sizeof(double) : sizeof(float); }

    StepIn<double> in64(int s) const
    {
        StepIn<double> f;
        f.sshn_t = static_cast<const double *>(prog[s][0]);
        f.un = static_cast<const double *>(prog[s][2]);
        f.vn = static_cast<const double *>(prog[s][3]);
        f.sshn_u = static_cast<const double *>(prog[s][4]);
        f.sshn_v = static_cast<const double *>(prog[s][5]);
        f.ht = static_cast<const double *>(cst[A_HT]);
        f.hu = static_cast<const double *>(cst[A_HU]);
        f.hv = static_cast<const double *>(cst[A_HV]);
        f.ssha_u = static_cast<const double *>(cst[A_SSHA_U]);
        f.ssha_v = static_cast<const double *>(cst[A_SSHA_V]);
        f.area_t = static_cast<const double *>(cst[A_AREA_T]);
        f.area_u = static_cast<const double *>(cst[A_AREA_U]);
        f.area_v = static_cast<const double *>(cst[A_AREA_V]);
        f.dx_u = static_cast<const double *>(cst[A_DX_U]);
        f.dx_v = static_cast<const double *>(cst[A_DX_V]);
        f.dx_t = static_cast<const double *>(cst[A_DX_T]);
        f.dy_u = static_cast<const double *>(cst[A_DY_U]);
        f.dy_v = static_cast<const double *>(cst[A_DY_V]);
        f.dy_t = static_cast<const double *>(cst[A_DY_T]);
        f.fcor_u = static_cast<const double *>(fcor_u);
        f.fcor_v = static_cast<const double *>(fcor_v);
        f.tmask = tm;
        return f;
    }
    StepOut<double> out64(int s) const
    {
        return StepOut<double>{static_cast<double *>(prog[s][0]),
                               static_cast<double *>(prog[s][2]),
                               static_cast<double *>(prog[s][3]),
                               static_cast<double *>(prog[s][4]),
                               static_cast<double *>(prog[s][5])};
    }
    MixedIn in32(int s) const
    {
        MixedIn f;
        f.ssh_hi = static_cast<const float *>(prog[s][0]);
        f.ssh_lo = static_cast<const float *>(prog[s][1]);
        f.un = static_cast<const float *>(prog[s][2]);
        f.vn = static_cast<const float *>(prog[s][3]);
        f.sshn_u = static_cast<const float *>(prog[s][4]);
        f.sshn_v = static_cast<const float *>(prog[s][5]);
        f.ht = static_cast<const float *>(cst[A_HT]);
        f.hu = static_cast<const float *>(cst[A_HU]);
        f.hv = static_cast<const float *>(cst[A_HV]);
        f.ssha_u = static_cast<const float *>(cst[A_SSHA_U]);
        f.ssha_v = static_cast<const float *>(cst[A_SSHA_V]);
        f.area_t = static_cast<const float *>(cst[A_AREA_T]);
        f.area_u = static_cast<const float *>(cst[A_AREA_U]);
        f.area_v = static_cast<const float *>(cst[A_AREA_V]);
        f.dx_u = static_cast<const float *>(cst[A_DX_U]);
        f.dx_v = static_cast<const float *>(cst[A_DX_V]);
        f.dx_t = static_cast<const float *>(cst[A_DX_T]);
        f.dy_u = static_cast<const float *>(cst[A_DY_U]);
        f.dy_v = static_cast<const float *>(cst[A_DY_V]);
        f.dy_t = static_cast<const float *>(cst[A_DY_T]);
        f.fcor_u = static_cast<const float *>(fcor_u);
        f.fcor_v = static_cast<const float *>(fcor_v);
        f.tmask = tm8;
        return f;
    }
    MixedOut out32(int s) const
    {
        return MixedOut{static_cast<float *>(prog[s][0]), static_cast<float *>(prog[s][1]),
                        static_cast<float *>(prog[s][2]), static_cast<float *>(prog[s][3]),
                        static_cast<float *>(prog[s][4]), static_cast<float *>(prog[s][5])};
    }

    ~EnsContext()
    {
        for (auto &kv : graphs) (void)hipGraphExecDestroy(kv.second);
        for (hipEvent_t e : ev) (void)hipEventDestroy(e);
        for (void *p : raw) (void)hipFree(p);
        for (auto &kv : lists) (void)hipFree(kv.second.first);
        void *more[] = {d_tide};
        for (void *p : more)
            if (p) (void)hipFree(p);
        if (h_tide) (void)hipHostFree(h_tide);
        fl_u.release();
        fl_v.release();
    }
};

// Constant arrays of the step (C-ABI indices); gphiu/gphiv enter through the
// Coriolis factors.
const Arr kConsts[] = {A_HT,     A_HU,     A_HV,   A_SSHA_U, A_SSHA_V,
                       A_AREA_T, A_AREA_U, A_AREA_V, A_DX_U, A_DX_V,
                       A_DX_T,   A_DY_U,   A_DY_V, A_DY_T};

// Copies host FP64 array `h` of domain d into arena `dst` (FP64 or FP32).
void put(EnsContext &C, void *dst, int d, const double *h, hipStream_t s,
         std::vector<float> &tmp)
{
    const size_t n = (size_t)C.geo[d].total;
    char *base = static_cast<char *>(dst) + (size_t)C.hdom[d].off * C.esize();
    if (C.prec == 64) {
        NEMO_HIP_CHECK(hipMemcpyAsync(base, h, n * sizeof(double), hipMemcpyHostToDevice, s));
    } else {
        tmp.resize(n);
        for (size_t k = 0; k < n; k++) tmp[k] = (float)h[k];
        NEMO_HIP_CHECK(hipMemcpyAsync(base, tmp.data(), n * sizeof(float),
                                      hipMemcpyHostToDevice, s));
        NEMO_HIP_CHECK(hipStreamSynchronize(s));  // tmp is reused
    }
}

// T-cell ssh of domain d into set `set`: FP64 as is; FP32 as the (hi, lo)
// split of the mixed form.
void put_ssh(EnsContext &C, int set, int d, const double *h, hipStream_t s)
{
    const size_t n = (size_t)C.geo[d].total;
    if (C.prec == 64) {
        char *base = static_cast<char *>(C.prog[set][0]) + (size_t)C.hdom[d].off * 8;
        NEMO_HIP_CHECK(hipMemcpyAsync(base, h, n * sizeof(double), hipMemcpyHostToDevice, s));
        return;
    }
    std::vector<float> hi(n), lo(n);
    for (size_t k = 0; k < n; k++) {
        hi[k] = (float)h[k];
        lo[k] = (float)(h[k] - (double)hi[k]);
    }
    const size_t o = (size_t)C.hdom[d].off * sizeof(float);
    NEMO_HIP_CHECK(hipMemcpyAsync(static_cast<char *>(C.prog[set][0]) + o, hi.data(),
                                  n * sizeof(float), hipMemcpyHostToDevice, s));
    NEMO_HIP_CHECK(hipMemcpyAsync(static_cast<char *>(C.prog[set][1]) + o, lo.data(),
                                  n * sizeof(float), hipMemcpyHostToDevice, s));
    NEMO_HIP_CHECK(hipStreamSynchronize(s));
}

void *arena(EnsContext &C, size_t esize, hipStream_t s)
{
    const size_t shift = C.raw.size() * kArenaStagger;
    void *p = nullptr;
    NEMO_HIP_CHECK(hipMalloc(&p, C.total * esize + shift));
    C.raw.push_back(p);
    void *a = static_cast<char *>(p) + shift;
    NEMO_HIP_CHECK(hipMemsetAsync(a, 0, C.total * esize, s));
    return a;
}

void build_shapes(EnsContext &C)
{
    hipDeviceProp_t prop;
    int dev = 0;
    NEMO_HIP_CHECK(hipGetDevice(&dev));
    NEMO_HIP_CHECK(hipGetDeviceProperties(&prop, dev));
    const int cus = prop.multiProcessorCount;
    for (int cols : kEnsCols)
        for (int rows : kEnsRows)
            for (int pf : {0, 1, 2}) {
                // one wavefront per tile
                const EnsShape grid{rows, 1, 0, cols, pf};
                if (ens_shape_supported(grid)) C.shapes.push_back(grid);
                // persistent, at a third, two thirds and all of the kernel's
                // residency
                if (!ens_shape_supported(EnsShape{rows, 1, 8, cols, pf})) continue;
                const int occ = ens_max_waves_per_cu(C.prec, rows, cols, pf);
                if (occ < 1) NEMO_FATAL("ensemble step kernel does not fit on a CU");
                const int levels[3] = {std::max(1, occ / 3), std::max(1, (2 * occ) / 3), occ};
                for (int wpb : {1, 2, 4})
                    for (int lv : levels) {
                        const int per_cu = std::max(1, lv / wpb);
                        int blocks = cus * per_cu;
                        blocks = (blocks + 7) / 8 * 8;
                        C.shapes.push_back(EnsShape{rows, wpb, blocks, cols, pf});
                    }
            }
}

// One tile list per (rows, columns) that some launch shape uses.
void build_lists(EnsContext &C, hipStream_t s)
{
    std::vector<std::vector<EnsTile>> host;  // alive until the copies are done
    for (const EnsShape &sh : C.shapes) {
        const int rows = sh.rows, cols = sh.cols;
        if (C.lists.count({rows, cols})) continue;
        std::vector<EnsTile> L;
        std::vector<int> ow(C.ndom), ns(C.ndom), nb(C.ndom);
        for (int d = 0; d < C.ndom; d++) {
            const EnsDomain &e = C.hdom[d];
            const Bounds b{e.xs, e.xe, e.ys, e.ye};
            ow[d] = cols ? cols : ens_domain_cols(b);
            ns[d] = wave::strips(b, ow[d]);
            nb[d] = wave::bands(b, rows);
        }
        auto push = [&](int d, int band, int strip) {
            const EnsDomain &e = C.hdom[d];
            L.push_back(EnsTile{e.off, e.W, e.xs, e.xe, e.ys, e.ye, strip, band, d, ow[d],
                                {0, 0}});
        };
        for (int d = 0; d < C.ndom; d++)
            for (int band = 0; band < nb[d]; band++)
                for (int strip = 0; strip < ns[d]; strip++) push(d, band, strip);
        if (L.size() > (size_t)std::numeric_limits<int>::max())
            NEMO_FATAL("ensemble: %zu tiles of %d rows exceed one launch", L.size(), rows);
        EnsTile *dl = dev_alloc<EnsTile>(L.size());
        NEMO_HIP_CHECK(hipMemcpyAsync(dl, L.data(), L.size() * sizeof(EnsTile),
                                      hipMemcpyHostToDevice, s));
        C.lists[{rows, cols}] = {dl, (int)L.size()};
        host.push_back(std::move(L));
    }
    NEMO_HIP_CHECK(hipStreamSynchronize(s));
}

void build_flather_lists(EnsContext &C, void *const *arrays, hipStream_t s)
{
    HostFlather U, V;
    for (int d = 0; d < C.ndom; d++) {
        const int *tm = static_cast<const int *>(arrays[d * A_COUNT + A_TMASK]);
        const double *hu = static_cast<const double *>(arrays[d * A_COUNT + A_HU]);
        const double *hv = static_cast<const double *>(arrays[d * A_COUNT + A_HV]);
        const int off = (int)C.hdom[d].off;
        for (int dir = 0; dir < 2; dir++) {
            HostFlather L = build_flather(tm, dir ? hv : hu, C.geo[d], C.P.g, dir == 0);
            HostFlather &A = dir ? V : U;
            const int base = (int)A.idx.size();
            for (size_t k = 1; k < L.seg_ptr.size(); k++) A.seg_ptr.push_back(base + L.seg_ptr[k]);
            for (int c : L.idx) A.idx.push_back(c + off);
            A.off.insert(A.off.end(), L.off.begin(), L.off.end());
            A.prev.insert(A.prev.end(), L.prev.begin(), L.prev.end());
            A.coef.insert(A.coef.end(), L.coef.begin(), L.coef.end());
        }
    }
    C.fl_u.upload(U, s);
    C.fl_v.upload(V, s);
}

double tide_value(const EnsContext &C, int d, int istep)
{
    const double rtime = istep * C.P.rdt;  // FP64 tidal phase
    return C.amp[d] * std::sin(C.omega_tide[d] * rtime);
}

// Rows 0..n-1 of the device tide table = steps istep .. istep+n-1, one value
// per domain.
void fill_tide(EnsContext &C, int n, int istep, hipStream_t s)
{
    if (n > C.tide_cap) {
        if (C.d_tide) NEMO_HIP_CHECK(hipFree(C.d_tide));
        if (C.h_tide) NEMO_HIP_CHECK(hipHostFree(C.h_tide));
        const size_t cnt = (size_t)n * C.ndom;
        C.d_tide = dev_alloc<double>(cnt);
        NEMO_HIP_CHECK(hipHostMalloc((void **)&C.h_tide, cnt * sizeof(double),
                                     hipHostMallocDefault));
        C.tide_cap = n;
        for (auto &kv : C.graphs) NEMO_HIP_CHECK(hipGraphExecDestroy(kv.second));
        C.graphs.clear();
    }
    NEMO_HIP_CHECK(hipStreamSynchronize(s));  // h_tide may still be in flight
    for (int k = 0; k < n; k++)
        for (int d = 0; d < C.ndom; d++)
            C.h_tide[(size_t)k * C.ndom + d] = tide_value(C, d, istep + k);
    NEMO_HIP_CHECK(hipMemcpyAsync(C.d_tide, C.h_tide, (size_t)n * C.ndom * sizeof(double),
                                  hipMemcpyHostToDevice, s));
}

// One step of every domain: set `par` -> set 1-par, tide row `row`.
void launch_pair(EnsContext &C, int par, int row, hipStream_t s, hipEvent_t e0 = nullptr,
                 hipEvent_t e1 = nullptr)
{
    const int q = 1 - par;
    const EnsShape &sh = C.shapes[C.shape];
    const auto &L = C.lists.at({sh.rows, sh.cols});
    const double *tide = C.d_tide + (size_t)row * C.ndom;
    if (e0) NEMO_HIP_CHECK(hipEventRecord(e0, s));
    if (C.prec == 64) {
        const StepScalars<double> p{C.P.rdt, C.P.visc, C.P.g, 1.0 + C.P.cbfr * C.P.rdt};
        launch_ens_f64(s, sh, C.in64(par), p, C.out64(q), L.first, L.second, tide);
    } else {
        const StepScalars<float> p{(float)C.P.rdt, (float)C.P.visc, (float)C.P.g,
                                   (float)(1.0 + C.P.cbfr * C.P.rdt)};
        launch_ens_mixed(s, sh, C.in32(par), p, C.out32(q), L.first, L.second, tide);
    }
    if (e1) NEMO_HIP_CHECK(hipEventRecord(e1, s));
    const int prev = C.paired ? par : q;  // the previous step's velocities
    if (C.prec == 64)
        launch_flather(s, C.fl_u.view(), C.fl_v.view(), static_cast<double *>(C.prog[q][2]),
                       static_cast<double *>(C.prog[q][3]),
                       static_cast<const double *>(C.prog[prev][2]),
                       static_cast<const double *>(C.prog[prev][3]),
                       static_cast<const double *>(C.prog[par][4]),
                       static_cast<const double *>(C.prog[par][5]));
    else
        launch_flather_f32(s, C.fl_u.view(), C.fl_v.view(), static_cast<float *>(C.prog[q][2]),
                           static_cast<float *>(C.prog[q][3]),
                           static_cast<const float *>(C.prog[prev][2]),
                           static_cast<const float *>(C.prog[prev][3]),
                           static_cast<const float *>(C.prog[par][4]),
                           static_cast<const float *>(C.prog[par][5]));
}

void copy_set(EnsContext &C, int from, int to, hipStream_t s)
{
    const size_t bytes = C.total * C.esize();
    for (int k = 0; k < kProg; k++)
        if (C.prog[from][k])
            NEMO_HIP_CHECK(hipMemcpyAsync(C.prog[to][k], C.prog[from][k], bytes,
                                          hipMemcpyDeviceToDevice, s));
}

hipGraphExec_t graph(EnsContext &C, int n, int row0, hipStream_t s)
{
    auto key = std::make_tuple(n, C.parity, C.shape, row0);
    auto it = C.graphs.find(key);
    if (it != C.graphs.end()) return it->second;
    hipGraph_t g;
    NEMO_HIP_CHECK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    for (int k = 0; k < n; k++) launch_pair(C, (C.parity + k) & 1, row0 + k, s);
    NEMO_HIP_CHECK(hipStreamEndCapture(s, &g));
    hipGraphExec_t exec;
    NEMO_HIP_CHECK(hipGraphInstantiate(&exec, g, nullptr, nullptr, 0));
    NEMO_HIP_CHECK(hipGraphDestroy(g));
    C.graphs[key] = exec;
    return exec;
}

// n steps from istep; timed: plain launches with events around every step
// kernel (mean in last_ms), else one graph replay.
void run(EnsContext &C, int n, int istep, bool timed)
{
    if (n < 0) NEMO_FATAL("nsteps must be >= 0, got %d", n);
    if (n == 0) return;
    hipStream_t s = lib_stream();
    fill_tide(C, n, istep, s);
    int row0 = 0;
    if (!C.paired) {
        // First step, then both sets agree wherever no sweep writes.
        launch_pair(C, C.parity, 0, s);
        C.parity ^= 1;
        copy_set(C, C.parity, 1 - C.parity, s);
        C.paired = true;
        row0 = 1;
    }
    const int m = n - row0;
    if (m == 0) return;
    if (timed) {
        while ((int)C.ev.size() < 2 * m) {
            hipEvent_t e;
            NEMO_HIP_CHECK(hipEventCreate(&e));
            C.ev.push_back(e);
        }
        for (int k = 0; k < m; k++)
            launch_pair(C, (C.parity + k) & 1, row0 + k, s, C.ev[2 * k], C.ev[2 * k + 1]);
        NEMO_HIP_CHECK(hipStreamSynchronize(s));
        double sum = 0.0;
        for (int k = 0; k < m; k++) {
            float ms = 0.f;
            NEMO_HIP_CHECK(hipEventElapsedTime(&ms, C.ev[2 * k], C.ev[2 * k + 1]));
            sum += ms;
        }
        C.last_ms = sum / m;
    } else {
        NEMO_HIP_CHECK(hipGraphLaunch(graph(C, m, row0, s), s));
    }
    C.parity = (C.parity + m) & 1;
}

EnsContext &ctx(void *h)
{
    if (!h) NEMO_FATAL("null ensemble handle");
    return *static_cast<EnsContext *>(h);
}

}  // namespace
}  // namespace nemo

using namespace nemo;

extern "C" {

void *nemo_ens_create(int ndom, void *const *arrays, const int *geom,
                      const double *tide_amp, const double *tide_omega, double rdt,
                      double cbfr, double visc, double omega, double d2r, double g,
                      int precision)
{
    if (ndom <= 0) NEMO_FATAL("ensemble needs at least one domain, got %d", ndom);
    if (precision != 64 && precision != 32)
        NEMO_FATAL("ensemble precision must be 64 or 32, got %d", precision);
    auto C = std::make_unique<EnsContext>();
    C->prec = precision;
    C->ndom = ndom;
    C->P = Physics{rdt, cbfr, visc, omega, d2r, g};
    size_t off = 0;
    for (int d = 0; d < ndom; d++) {
        const int *q = geom + 6 * d;
        const Geometry G{q[0], q[1], q[2], q[3], q[4], q[5]};
        check_geometry(G);
        const int *tm = static_cast<const int *>(arrays[d * A_COUNT + A_TMASK]);
        if (!fused_allowed(tm, G))
            NEMO_FATAL("ensemble domain %d: the mask has wet cells on row %d", d, G.ye + 1);
        C->geo.push_back(G);
        EnsDomain e{};
        e.off = (unsigned)off;
        e.W = G.W;
        e.xs = G.xs;
        e.xe = G.xe;
        e.ys = G.ys;
        e.ye = G.ye;
        C->hdom.push_back(e);
        C->amp.push_back(tide_amp[d]);
        C->omega_tide.push_back(tide_omega[d]);
        C->interior += (long long)(G.xe - G.xs + 1) * (G.ye - G.ys + 1);
        off += ((size_t)G.total + 63) / 64 * 64;
        if (off * C->esize() > 0xFFFFFFFFull)
            NEMO_FATAL("ensemble arenas exceed 4 GiB per field at domain %d "
                       "(32-bit offsets); split the ensemble", d);
    }
    C->total = off;
    hipStream_t s = lib_stream();
    const size_t es = C->esize();
    for (Arr a : kConsts) C->cst[a] = arena(*C, es, s);
    C->fcor_u = arena(*C, es, s);
    C->fcor_v = arena(*C, es, s);
    C->tm = static_cast<int *>(arena(*C, sizeof(int), s));
    if (precision == 32) C->tm8 = static_cast<signed char *>(arena(*C, 1, s));
    for (int set = 0; set < 2; set++)
        for (int k = 0; k < kProg; k++)
            if (k != 1 || precision == 32) C->prog[set][k] = arena(*C, es, s);

    std::vector<float> tmp;
    for (int d = 0; d < ndom; d++) {
        auto host = [&](Arr a) { return static_cast<const double *>(arrays[d * A_COUNT + a]); };
        const Geometry &G = C->geo[d];
        for (Arr a : kConsts) put(*C, C->cst[a], d, host(a), s, tmp);
        NEMO_HIP_CHECK(hipMemcpyAsync(C->tm + C->hdom[d].off, arrays[d * A_COUNT + A_TMASK],
                                      (size_t)G.total * sizeof(int), hipMemcpyHostToDevice, s));
        std::vector<signed char> tm8;
        if (C->tm8) {
            tm8 = mask_bytes(static_cast<const int *>(arrays[d * A_COUNT + A_TMASK]),
                             (size_t)G.total);
            NEMO_HIP_CHECK(hipMemcpyAsync(C->tm8 + C->hdom[d].off, tm8.data(),
                                          (size_t)G.total, hipMemcpyHostToDevice, s));
        }
        std::vector<double> fu = coriolis(host(A_GPHIU), G.total, omega, d2r);
        std::vector<double> fv = coriolis(host(A_GPHIV), G.total, omega, d2r);
        put(*C, C->fcor_u, d, fu.data(), s, tmp);
        put(*C, C->fcor_v, d, fv.data(), s, tmp);
        // set 0 from the *n arrays, set 1 from the *a arrays; face ssh from
        // sshn_u / sshn_v in both
        put_ssh(*C, 0, d, host(A_SSHN_T), s);
        put_ssh(*C, 1, d, host(A_SSHA_T), s);
        put(*C, C->prog[0][2], d, host(A_UN), s, tmp);
        put(*C, C->prog[1][2], d, host(A_UA), s, tmp);
        put(*C, C->prog[0][3], d, host(A_VN), s, tmp);
        put(*C, C->prog[1][3], d, host(A_VA), s, tmp);
        for (int set = 0; set < 2; set++) {
            put(*C, C->prog[set][4], d, host(A_SSHN_U), s, tmp);
            put(*C, C->prog[set][5], d, host(A_SSHN_V), s, tmp);
        }
        NEMO_HIP_CHECK(hipStreamSynchronize(s));  // fu / fv / tm8 go out of scope
    }
    build_flather_lists(*C, arrays, s);
    build_shapes(*C);
    build_lists(*C, s);
    // default: one wavefront per tile, 4 rows, 61 or 56 columns per domain,
    // loads three rows ahead, in both precisions (the fastest shape on a
    // 256-domain ensemble of sides 64..1024 on MI355X)
    for (size_t k = 0; k < C->shapes.size(); k++) {
        const EnsShape &sh = C->shapes[k];
        if (sh.blocks == 0 && sh.rows == 4 && sh.cols == 0 && sh.prefetch == 2)
            C->shape = (int)k;
    }
    NEMO_HIP_CHECK(hipStreamSynchronize(s));
    return C.release();
}

int nemo_ens_shape_count(void *h) { return (int)ctx(h).shapes.size(); }

void nemo_ens_shape_info(void *h, int shape, int *rows, int *waves_per_block, int *blocks)
{
    EnsContext &C = ctx(h);
    if (shape < 0 || shape >= (int)C.shapes.size())
        NEMO_FATAL("unknown ensemble shape %d", shape);
    *rows = C.shapes[shape].rows;
    *waves_per_block = C.shapes[shape].waves_per_block;
    *blocks = C.shapes[shape].blocks;
}

void nemo_ens_set_shape(void *h, int shape)
{
    EnsContext &C = ctx(h);
    if (shape < 0 || shape >= (int)C.shapes.size())
        NEMO_FATAL("unknown ensemble shape %d", shape);
    C.shape = shape;
}

int nemo_ens_shape(void *h) { return ctx(h).shape; }

int nemo_ens_shape_columns(void *h, int shape)
{
    EnsContext &C = ctx(h);
    if (shape < 0 || shape >= (int)C.shapes.size())
        NEMO_FATAL("unknown ensemble shape %d", shape);
    return C.shapes[shape].cols;
}

int nemo_ens_shape_prefetch(void *h, int shape)
{
    EnsContext &C = ctx(h);
    if (shape < 0 || shape >= (int)C.shapes.size())
        NEMO_FATAL("unknown ensemble shape %d", shape);
    return C.shapes[shape].prefetch;
}

void nemo_ens_steps(void *h, int nsteps, int istep) { run(ctx(h), nsteps, istep, false); }

double nemo_ens_kernel_ms(void *h, int nsteps, int istep)
{
    EnsContext &C = ctx(h);
    C.last_ms = -1.0;
    run(C, nsteps, istep, true);
    return C.last_ms;
}

int nemo_ens_autotune(void *h, int steps_per_shape, double *ms_out)
{
    EnsContext &C = ctx(h);
    if (steps_per_shape < 1) NEMO_FATAL("autotune needs >= 1 step per shape");
    hipStream_t s = lib_stream();
    // snapshot both state sets and the flags; every shape advances the state
    const size_t bytes = C.total * C.esize();
    void *snap[2][kProg] = {};
    for (int set = 0; set < 2; set++)
        for (int k = 0; k < kProg; k++)
            if (C.prog[set][k]) {
                NEMO_HIP_CHECK(hipMalloc(&snap[set][k], bytes));
                NEMO_HIP_CHECK(hipMemcpyAsync(snap[set][k], C.prog[set][k], bytes,
                                              hipMemcpyDeviceToDevice, s));
            }
    const int parity = C.parity;
    const bool paired = C.paired;
    if (!C.paired) run(C, 1, 1, false);
    int best = 0;
    double best_ms = 1e300;
    for (int k = 0; k < (int)C.shapes.size(); k++) {
        C.shape = k;
        run(C, 1, 1, true);  // first use of the shape: code load
        run(C, steps_per_shape, 1, true);
        const double ms = C.last_ms;
        if (ms_out) ms_out[k] = ms;
        if (ms < best_ms) {
            best_ms = ms;
            best = k;
        }
    }
    for (int set = 0; set < 2; set++)
        for (int k = 0; k < kProg; k++)
            if (snap[set][k]) {
                NEMO_HIP_CHECK(hipMemcpyAsync(C.prog[set][k], snap[set][k], bytes,
                                              hipMemcpyDeviceToDevice, s));
            }
    NEMO_HIP_CHECK(hipStreamSynchronize(s));
    for (auto &set : snap)
        for (void *p : set)
            if (p) NEMO_HIP_CHECK(hipFree(p));
    C.parity = parity;
    C.paired = paired;
    C.shape = best;
    return best;
}

long long nemo_ens_interior_points(void *h) { return ctx(h).interior; }

void nemo_ens_read(void *h, int domain, int field, double *out)
{
    EnsContext &C = ctx(h);
    if (domain < 0 || domain >= C.ndom) NEMO_FATAL("unknown ensemble domain %d", domain);
    static const int kField[] = {0, 2, 3, 4, 5};  // sshn_t, un, vn, sshn_u, sshn_v
    if (field < 0 || field > 4) NEMO_FATAL("unknown ensemble field %d", field);
    hipStream_t s = lib_stream();
    NEMO_HIP_CHECK(hipStreamSynchronize(s));
    const size_t n = (size_t)C.geo[domain].total;
    const size_t o = (size_t)C.hdom[domain].off;
    const void *src = C.prog[C.parity][kField[field]];
    if (C.prec == 64) {
        NEMO_HIP_CHECK(hipMemcpy(out, static_cast<const double *>(src) + o,
                                 n * sizeof(double), hipMemcpyDeviceToHost));
        return;
    }
    std::vector<float> a(n), lo;
    NEMO_HIP_CHECK(hipMemcpy(a.data(), static_cast<const float *>(src) + o,
                             n * sizeof(float), hipMemcpyDeviceToHost));
    if (field == 0) {
        lo.resize(n);
        NEMO_HIP_CHECK(hipMemcpy(lo.data(), static_cast<const float *>(C.prog[C.parity][1]) + o,
                                 n * sizeof(float), hipMemcpyDeviceToHost));
    }
    for (size_t k = 0; k < n; k++)
        out[k] = field == 0 ? (double)a[k] + (double)lo[k] : (double)a[k];
}

void nemo_ens_destroy(void *h)
{
    if (!h) return;
    NEMO_HIP_CHECK(hipStreamSynchronize(lib_stream()));
    delete static_cast<EnsContext *>(h);
}

}  // extern "C"
