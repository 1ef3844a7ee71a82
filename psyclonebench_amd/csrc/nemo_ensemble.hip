// Batched ensemble step (CDNA4 / gfx950): one launch per step advances
// every domain of the ensemble.
//
// Work list: every domain is cut into tiles of one overlapped-wavefront strip
// (OW output columns: 61, or 56 so that every strip's stores start on a
// 64-byte boundary, or per domain whichever of the two suits its width) by
// one band (R rows); the list holds one self-contained descriptor per tile
// (domain geometry, strip, band, width), domain by domain. Two launch forms run wave::tile_f64 /
// tile_mixed on the domain's arrays of each tile:
//
//  * persistent: a fixed number of workgroups (a multiple of 8) of 1, 2 or
//    4 wavefronts; every wavefront takes tiles from the list with a stride,
//    so uneven domain sizes cost nothing but the partial tiles at their
//    edges;
//  * one wavefront per tile: a grid of one-wavefront workgroups, one per
//    tile of the list, as the single-domain kernels launch, so that the
//    dispatcher starts a new tile wherever a wavefront slot frees up.
//
// Either form runs the tile body with the loads of the next row issued at
// the top of the row that needs them (PF = 0) or one row earlier (PF = 1).
//
// XCD locality: the dispatcher hands workgroup b to XCD b % 8. The list is
// cut into 8 contiguous chunks and the workgroups of XCD x take tiles of
// chunk x only, so neighbouring strips of a domain (which share the cache
// lines of their overlap columns) are processed under one L2.
#include "nemo_ensemble.h"
#include "nemo_hip_common.h"
#include "nemo_wave.h"

namespace nemo {
namespace {

// Tiles [first, last) of this wavefront's XCD chunk, and its stride. Every
// quantity is wavefront-uniform, so the list and domain reads are scalar.
struct Walk {
    int first, last, stride;
};

__device__ __forceinline__ Walk walk(int ntiles)
{
    const int wpb = (int)(blockDim.x >> 6);
    const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int bid = (int)blockIdx.x, nblk = (int)gridDim.x;
    const int xcd = bid & 7, k = bid >> 3, per = nblk >> 3;  // nblk % 8 == 0
    const int chunk = (ntiles + 7) >> 3;
    const int t0 = xcd * chunk;
    return Walk{t0 + k * wpb + wave, min(t0 + chunk, ntiles), per * wpb};
}

// Tile `tl` of the FP64 step; OW = 0 takes the tile width from the tile
// (wavefront-uniform branch between the 61- and 56-column bodies).
template <int R, int OW, int PF>
__device__ __forceinline__ void ens_tile_f64(const StepIn<double> &f,
                                             const StepScalars<double> &p,
                                             const StepOut<double> &o, const EnsTile &tl,
                                             double tide)
{
    if constexpr (OW == 0) {
        if (tl.cols == 61)
            ens_tile_f64<R, 61, PF>(f, p, o, tl, tide);
        else
            ens_tile_f64<R, 56, PF>(f, p, o, tl, tide);
    } else {
        wave::tile_f64<R, OW, PF>(f, p, o, tl.W, tl.off, Bounds{tl.xs, tl.xe, tl.ys, tl.ye},
                                  tl.strip, tl.band, tide);
    }
}

template <int R, int OW, int PF>
__device__ __forceinline__ void ens_tile_mixed(const MixedIn &f, const StepScalars<float> &p,
                                               const MixedOut &o, const EnsTile &tl,
                                               double tide)
{
    if constexpr (OW == 0) {
        if (tl.cols == 61)
            ens_tile_mixed<R, 61, PF>(f, p, o, tl, tide);
        else
            ens_tile_mixed<R, 56, PF>(f, p, o, tl, tide);
    } else {
        wave::tile_mixed<R, OW, PF>(f, p, o, tl.W, tl.off,
                                    Bounds{tl.xs, tl.xe, tl.ys, tl.ye}, tl.strip, tl.band,
                                    wave::split(tide));
    }
}

template <int R, int OW, int PF>
__global__ __launch_bounds__(256) void k_ens_f64(StepIn<double> f, StepScalars<double> p,
                                                 StepOut<double> o,
                                                 const EnsTile *__restrict__ tiles,
                                                 int ntiles,
                                                 const double *__restrict__ tide)
{
    const Walk w = walk(ntiles);
    for (int t = w.first; t < w.last; t += w.stride) {
        const EnsTile tl = tiles[t];
        ens_tile_f64<R, OW, PF>(f, p, o, tl, tide[tl.dom]);
    }
}

// One wavefront per tile; the grid has exactly ntiles workgroups.
template <int R, int OW, int PF>
__global__ __launch_bounds__(64) void k_ens_f64_grid(StepIn<double> f, StepScalars<double> p,
                                                     StepOut<double> o,
                                                     const EnsTile *__restrict__ tiles,
                                                     const double *__restrict__ tide)
{
    const EnsTile tl = tiles[wave::xcd_tile((int)blockIdx.x, (int)gridDim.x)];
    ens_tile_f64<R, OW, PF>(f, p, o, tl, tide[tl.dom]);
}

template <int R, int OW, int PF>
__global__ __launch_bounds__(256) void k_ens_mixed(MixedIn f, StepScalars<float> p,
                                                   MixedOut o,
                                                   const EnsTile *__restrict__ tiles,
                                                   int ntiles,
                                                   const double *__restrict__ tide)
{
    const Walk w = walk(ntiles);
    for (int t = w.first; t < w.last; t += w.stride) {
        const EnsTile tl = tiles[t];
        ens_tile_mixed<R, OW, PF>(f, p, o, tl, tide[tl.dom]);
    }
}

template <int R, int OW, int PF>
__global__ __launch_bounds__(64) void k_ens_mixed_grid(MixedIn f, StepScalars<float> p,
                                                       MixedOut o,
                                                       const EnsTile *__restrict__ tiles,
                                                       const double *__restrict__ tide)
{
    const EnsTile tl = tiles[wave::xcd_tile((int)blockIdx.x, (int)gridDim.x)];
    ens_tile_mixed<R, OW, PF>(f, p, o, tl, tide[tl.dom]);
}

void check_shape(const EnsShape &sh)
{
    if (sh.blocks < 0 || sh.blocks % 8)
        NEMO_FATAL("ensemble launch: %d workgroups is not 0 (one per tile) or a positive "
                   "multiple of 8", sh.blocks);
    if (sh.waves_per_block != 1 && sh.waves_per_block != 2 && sh.waves_per_block != 4)
        NEMO_FATAL("ensemble launch: %d wavefronts per workgroup", sh.waves_per_block);
    if (sh.blocks == 0 && sh.waves_per_block != 1)
        NEMO_FATAL("ensemble launch: one workgroup per tile takes 1 wavefront, not %d",
                   sh.waves_per_block);
    if (sh.cols != 61 && sh.cols != 56 && sh.cols != 0)
        NEMO_FATAL("ensemble launch: %d columns per wavefront", sh.cols);
    if (sh.prefetch < 0 || sh.prefetch > 2)
        NEMO_FATAL("ensemble launch: prefetch depth %d", sh.prefetch);
}

// rows x columns x prefetch depth of the tile -> one kernel instantiation
// per launch form
#define NEMO_ENS_PERSISTENT(X)                                                         \
    X(4, 61, 0) X(8, 61, 0) X(16, 61, 0) X(32, 61, 0) X(64, 61, 0) X(4, 56, 0)         \
    X(8, 56, 0) X(16, 56, 0) X(32, 56, 0) X(64, 56, 0) X(4, 61, 1) X(8, 61, 1)         \
    X(16, 61, 1) X(4, 56, 1) X(8, 56, 1) X(16, 56, 1)
#define NEMO_ENS_GRID(X)                                                               \
    X(4, 61, 0) X(8, 61, 0) X(16, 61, 0) X(4, 56, 0) X(8, 56, 0) X(16, 56, 0)          \
    X(4, 61, 1) X(8, 61, 1) X(16, 61, 1) X(4, 56, 1) X(8, 56, 1) X(16, 56, 1)          \
    X(4, 0, 0) X(8, 0, 0) X(4, 0, 1) X(8, 0, 1) X(4, 61, 2) X(8, 61, 2) X(4, 56, 2)     \
    X(8, 56, 2) X(4, 0, 2) X(8, 0, 2)

void unsupported(const EnsShape &sh)
{
    NEMO_FATAL("ensemble launch: no %s kernel for %d rows x %d columns, prefetch %d",
               sh.blocks ? "persistent" : "one-wavefront-per-tile", sh.rows, sh.cols,
               sh.prefetch);
}

}  // namespace

bool ens_shape_supported(const EnsShape &sh)
{
#define NEMO_ENS_HAS(R, OW, PF)                                                        \
    if (sh.rows == R && sh.cols == OW && sh.prefetch == PF) return true;
    if (sh.blocks) {
        NEMO_ENS_PERSISTENT(NEMO_ENS_HAS)
    } else {
        NEMO_ENS_GRID(NEMO_ENS_HAS)
    }
#undef NEMO_ENS_HAS
    return false;
}

void launch_ens_f64(hipStream_t s, const EnsShape &sh, const StepIn<double> &in,
                    const StepScalars<double> &p, const StepOut<double> &out,
                    const EnsTile *tiles, int ntiles, const double *tide)
{
    check_shape(sh);
    if (ntiles <= 0) return;
    const dim3 grid((unsigned)sh.blocks), block((unsigned)(64 * sh.waves_per_block));
#define NEMO_ENS_F64(R, OW, PF)                                                        \
    if (sh.rows == R && sh.cols == OW && sh.prefetch == PF) {                          \
        k_ens_f64<R, OW, PF><<<grid, block, 0, s>>>(in, p, out, tiles, ntiles, tide);  \
        NEMO_LAUNCH_CHECK();                                                           \
        return;                                                                        \
    }
#define NEMO_ENS_F64_GRID(R, OW, PF)                                                   \
    if (sh.rows == R && sh.cols == OW && sh.prefetch == PF) {                          \
        k_ens_f64_grid<R, OW, PF><<<dim3((unsigned)ntiles), dim3(64), 0, s>>>(         \
            in, p, out, tiles, tide);                                                  \
        NEMO_LAUNCH_CHECK();                                                           \
        return;                                                                        \
    }
    if (sh.blocks) {
        NEMO_ENS_PERSISTENT(NEMO_ENS_F64)
    } else {
        NEMO_ENS_GRID(NEMO_ENS_F64_GRID)
    }
#undef NEMO_ENS_F64
#undef NEMO_ENS_F64_GRID
    unsupported(sh);
}

void launch_ens_mixed(hipStream_t s, const EnsShape &sh, const MixedIn &in,
                      const StepScalars<float> &p, const MixedOut &out,
                      const EnsTile *tiles, int ntiles, const double *tide)
{
    check_shape(sh);
    if (ntiles <= 0) return;
    const dim3 grid((unsigned)sh.blocks), block((unsigned)(64 * sh.waves_per_block));
#define NEMO_ENS_MIXED(R, OW, PF)                                                      \
    if (sh.rows == R && sh.cols == OW && sh.prefetch == PF) {                          \
        k_ens_mixed<R, OW, PF><<<grid, block, 0, s>>>(in, p, out, tiles, ntiles,       \
                                                      tide);                           \
        NEMO_LAUNCH_CHECK();                                                           \
        return;                                                                        \
    }
#define NEMO_ENS_MIXED_GRID(R, OW, PF)                                                 \
    if (sh.rows == R && sh.cols == OW && sh.prefetch == PF) {                          \
        k_ens_mixed_grid<R, OW, PF><<<dim3((unsigned)ntiles), dim3(64), 0, s>>>(       \
            in, p, out, tiles, tide);                                                  \
        NEMO_LAUNCH_CHECK();                                                           \
        return;                                                                        \
    }
    if (sh.blocks) {
        NEMO_ENS_PERSISTENT(NEMO_ENS_MIXED)
    } else {
        NEMO_ENS_GRID(NEMO_ENS_MIXED_GRID)
    }
#undef NEMO_ENS_MIXED
#undef NEMO_ENS_MIXED_GRID
    unsupported(sh);
}

int ens_domain_cols(const Bounds &b)
{
    // 61-column strips cost about 27/25 of 56-column ones
    return 27 * wave::strips(b, 61) < 25 * wave::strips(b, 56) ? 61 : 56;
}

int ens_max_waves_per_cu(int precision, int rows, int cols, int prefetch)
{
    int blocks = 0;
#define NEMO_ENS_OCC(R, OW, PF)                                                        \
    if (rows == R && cols == OW && prefetch == PF) {                                   \
        const void *k = precision == 64 ? reinterpret_cast<const void *>(&k_ens_f64<R, OW, PF>) \
                                        : reinterpret_cast<const void *>(&k_ens_mixed<R, OW, PF>); \
        NEMO_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, k, 64, 0)); \
        return blocks;                                                                 \
    }
    NEMO_ENS_PERSISTENT(NEMO_ENS_OCC)
#undef NEMO_ENS_OCC
    NEMO_FATAL("ensemble launch: no persistent kernel for %d rows x %d columns, prefetch %d",
               rows, cols, prefetch);
}

}  // namespace nemo
