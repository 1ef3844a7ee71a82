// Host-side NEMOLite2D model set-up (mask, metrics, depth, initial state,
// checksum) in the dl_esm_inf layout. Declarations and the reference file
// each function restates: include/nemolite2d_setup.h.
#include "nemolite2d_setup.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

inline size_t at(int W, int ji, int jj) { return (size_t)jj * W + ji; }

// SplitMix64 (Steele, Lea, Flood 2014): state += golden gamma, then the
// variant-13 finaliser.
struct SplitMix64 {
    uint64_t s;
    uint64_t next()
    {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    // [0, 1) with 53 random bits
    double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }
};

// Half-sample-symmetric reflection into [0, n): ... c b a | a b c ...
inline int reflect(int k, int n)
{
    while (k < 0 || k >= n) {
        if (k < 0) k = -k - 1;
        if (k >= n) k = 2 * n - k - 1;
    }
    return k;
}

}  // namespace

extern "C" {

void nemo_grid_dims(int M, int N, int align, int *W, int *H)
{
    if (align < 1) align = 1;
    *W = ((M + 1 + align - 1) / align) * align;
    *H = N + 1;
}

int nemo_setup_tmask(int mode, int M, int N, int W, int H, int32_t *tmask)
{
    if (mode < NEMO_MASK_CLOSED || mode > NEMO_MASK_WEST_OPEN) return -1;
    std::fill(tmask, tmask + (size_t)W * H, 0);
    // internal region (1-based 2..M-1 x 2..N-1) is wet
    for (int jj = 1; jj <= N - 2; jj++)
        for (int ji = 1; ji <= M - 2; ji++) tmask[at(W, ji, jj)] = 1;
    // east column (all rows, model_mod.f90:191-199), then the south row (all
    // columns, :215-224): -1 where the south edge is open, else 0. The south
    // rule comes last, so it owns both southern corners: the south-east
    // corner of an east-open mask is 0.
    if (mode == NEMO_MASK_EAST_OPEN || mode == NEMO_MASK_SOUTH_EAST)
        for (int jj = 0; jj < N; jj++) tmask[at(W, M - 1, jj)] = -1;
    const int south = (mode == NEMO_MASK_SOUTH_OPEN || mode == NEMO_MASK_SOUTH_EAST) ? -1 : 0;
    if (mode != NEMO_MASK_WEST_OPEN)
        for (int ji = 0; ji < M; ji++) tmask[at(W, ji, 0)] = south;
    if (mode == NEMO_MASK_WEST_OPEN)
        for (int jj = 0; jj < N; jj++) tmask[at(W, 0, jj)] = -1;
    return 0;
}

void nemo_setup_metrics(int W, int H, double dx, double dy, double gphi,
                        double *area_t, double *area_u, double *area_v,
                        double *dx_u, double *dx_v, double *dx_t, double *dy_u,
                        double *dy_v, double *dy_t, double *gphiu,
                        double *gphiv)
{
    const size_t n = (size_t)W * H;
    const double area = dx * dy;
    std::fill(area_t, area_t + n, area);
    std::fill(area_u, area_u + n, area);
    std::fill(area_v, area_v + n, area);
    std::fill(dx_u, dx_u + n, dx);
    std::fill(dx_v, dx_v + n, dx);
    std::fill(dx_t, dx_t + n, dx);
    std::fill(dy_u, dy_u + n, dy);
    std::fill(dy_v, dy_v + n, dy);
    std::fill(dy_t, dy_t + n, dy);
    std::fill(gphiu, gphiu + n, gphi);
    std::fill(gphiv, gphiv + n, gphi);
}

void nemo_setup_depth_uniform(int W, int H, double depth, double *ht,
                              double *hu, double *hv)
{
    const size_t n = (size_t)W * H;
    std::fill(ht, ht + n, depth);
    std::fill(hu, hu + n, depth);
    std::fill(hv, hv + n, depth);
}

void nemo_uniforms(uint64_t seed, int64_t n, double *out)
{
    SplitMix64 rng{seed};
    for (int64_t k = 0; k < n; k++) out[k] = rng.uniform();
}

void nemo_setup_bathymetry(uint64_t seed, int M, int N, int W, int H,
                           double sigma, double mean, double relief,
                           double min_depth, int threads, double *ht,
                           double *hu, double *hv)
{
#ifdef _OPENMP
    const int nthr = threads > 0 ? threads : omp_get_max_threads();
#else
    const int nthr = 1;
    (void)threads;
    (void)nthr;
#endif
    const int nx = M - 2, ny = N - 2;  // internal T points
    const size_t np = (size_t)nx * ny;

    // white noise, two normals per uniform pair (Box-Muller)
    std::vector<double> a(np), b(np);
    SplitMix64 rng{seed};
    for (size_t k = 0; k < np; k += 2) {
        double u1 = 1.0 - rng.uniform();  // (0, 1]
        double u2 = rng.uniform();
        double r = std::sqrt(-2.0 * std::log(u1));
        double th = 2.0 * M_PI * u2;
        a[k] = r * std::cos(th);
        if (k + 1 < np) a[k + 1] = r * std::sin(th);
    }

    // separable Gaussian taps, truncated at 4 sigma, normalised to sum 1
    const int rad = std::max(1, (int)std::lround(4.0 * sigma));
    std::vector<double> w(2 * rad + 1);
    double wsum = 0.0;
    for (int k = -rad; k <= rad; k++) {
        w[k + rad] = std::exp(-0.5 * (double)k * k / (sigma * sigma));
        wsum += w[k + rad];
    }
    for (double &x : w) x /= wsum;

    // along i (rows), a -> b
#pragma omp parallel for num_threads(nthr) schedule(static)
    for (int j = 0; j < ny; j++) {
        const double *src = &a[(size_t)j * nx];
        double *dst = &b[(size_t)j * nx];
        for (int i = 0; i < nx; i++) {
            double s = 0.0;
            if (i - rad >= 0 && i + rad < nx) {
                for (int k = -rad; k <= rad; k++) s += w[k + rad] * src[i + k];
            } else {
                for (int k = -rad; k <= rad; k++)
                    s += w[k + rad] * src[reflect(i + k, nx)];
            }
            dst[i] = s;
        }
    }
    // along j (columns), b -> a; i innermost for contiguous access
#pragma omp parallel for num_threads(nthr) schedule(static)
    for (int j = 0; j < ny; j++) {
        double *dst = &a[(size_t)j * nx];
        std::fill(dst, dst + nx, 0.0);
        for (int k = -rad; k <= rad; k++) {
            const double wk = w[k + rad];
            const double *src = &b[(size_t)reflect(j + k, ny) * nx];
            for (int i = 0; i < nx; i++) dst[i] += wk * src[i];
        }
    }

    // normalise to [-1, 1]
    double lo = a[0], hi = a[0];
    for (size_t k = 1; k < np; k++) {
        lo = std::min(lo, a[k]);
        hi = std::max(hi, a[k]);
    }
    const double span = hi > lo ? hi - lo : 1.0;

    // ht on every point: nearest internal value
#pragma omp parallel for num_threads(nthr) schedule(static)
    for (int jj = 0; jj < H; jj++) {
        const int j = std::min(std::max(jj - 1, 0), ny - 1);
        for (int ji = 0; ji < W; ji++) {
            const int i = std::min(std::max(ji - 1, 0), nx - 1);
            const double x = 2.0 * (a[(size_t)j * nx + i] - lo) / span - 1.0;
            ht[at(W, ji, jj)] = std::max(mean + relief * x, min_depth);
        }
    }
#pragma omp parallel for num_threads(nthr) schedule(static)
    for (int jj = 0; jj < H; jj++) {
        for (int ji = 0; ji < W; ji++) {
            const size_t c = at(W, ji, jj);
            hu[c] = ji + 1 < W ? 0.5 * (ht[c] + ht[c + 1]) : ht[c];
            hv[c] = jj + 1 < H ? 0.5 * (ht[c] + ht[c + W]) : ht[c];
        }
    }
}

void nemo_initialise(int M, int N, int W, int H, const double *area_t,
                     const double *area_u, const double *area_v,
                     const double *sshn_t, double *sshn_u, double *sshn_v,
                     double *un, double *vn)
{
    // initialisation_mod.f90:35-43 in 0-based form: face i averages T cells
    // min(i+1, M-3) and i
    for (int i = 0; i <= M - 2; i++) {
        for (int j = 0; j <= N - 2; j++) {
            const int t1 = std::min(i + 1, M - 3), t2 = i;
            const double vol = area_t[at(W, t1, j)] * sshn_t[at(W, t1, j)] +
                               area_t[at(W, t2, j)] * sshn_t[at(W, t2, j)];
            sshn_u[at(W, i, j)] = 0.5 * vol / area_u[at(W, i, j)];
        }
    }
    // :47-55, the same along j
    for (int j = 0; j <= N - 2; j++) {
        for (int i = 0; i <= M - 2; i++) {
            const int t1 = std::min(j + 1, N - 3), t2 = j;
            const double vol = area_t[at(W, i, t1)] * sshn_t[at(W, i, t1)] +
                               area_t[at(W, i, t2)] * sshn_t[at(W, i, t2)];
            sshn_v[at(W, i, j)] = 0.5 * vol / area_v[at(W, i, j)];
        }
    }
    // un, vn are zeroed over the whole allocation (:58, :61)
    const size_t n = (size_t)W * H;
    std::memset(un, 0, n * sizeof(double));
    std::memset(vn, 0, n * sizeof(double));
}

double nemo_field_checksum(const double *x, int W, int x0, int x1, int y0,
                           int y1)
{
    double s = 0.0;
    for (int jj = y0; jj <= y1; jj++)
        for (int ji = x0; ji <= x1; ji++) s += std::fabs(x[at(W, ji, jj)]);
    return s;
}

}  // extern "C"
