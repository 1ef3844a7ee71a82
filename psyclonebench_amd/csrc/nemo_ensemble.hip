// Batched ensemble step (CDNA4 / gfx950): one persistent launch per step
// advances every domain of the ensemble.
//
// Work list: every domain is cut into tiles of one overlapped-wavefront strip
// (OW output columns: 61, or 56 so that every strip's stores start on a
// 64-byte boundary) by one band (R rows); the list holds (domain, strip,
// band) per tile, domain by domain. The launch has a fixed number of
// workgroups (a multiple of 8) of 1, 2 or 4 wavefronts; every wavefront takes
// tiles from the list with a stride and runs wave::tile_f64 / tile_mixed on
// the domain's arrays, so uneven domain sizes cost nothing but the partial
// tiles at their edges.
//
// XCD locality: the dispatcher hands workgroup b to XCD b % 8. The list is
// cut into 8 contiguous chunks and the workgroups of XCD x walk chunk x only,
// so neighbouring strips of a domain (which share the cache lines of their
// 3 overlap columns) are processed under one L2.
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

template <int R, int OW>
__global__ __launch_bounds__(256) void k_ens_f64(StepIn<double> f, StepScalars<double> p,
                                                 StepOut<double> o,
                                                 const EnsDomain *__restrict__ dom,
                                                 const EnsTile *__restrict__ tiles,
                                                 int ntiles,
                                                 const double *__restrict__ tide)
{
    const Walk w = walk(ntiles);
    for (int t = w.first; t < w.last; t += w.stride) {
        const EnsTile tl = tiles[t];
        const EnsDomain d = dom[tl.dom];
        wave::tile_f64<R, OW>(f, p, o, d.W, d.off, Bounds{d.xs, d.xe, d.ys, d.ye},
                              tl.strip, tl.band, tide[tl.dom]);
    }
}

template <int R, int OW>
__global__ __launch_bounds__(256) void k_ens_mixed(MixedIn f, StepScalars<float> p,
                                                   MixedOut o,
                                                   const EnsDomain *__restrict__ dom,
                                                   const EnsTile *__restrict__ tiles,
                                                   int ntiles,
                                                   const double *__restrict__ tide)
{
    const Walk w = walk(ntiles);
    for (int t = w.first; t < w.last; t += w.stride) {
        const EnsTile tl = tiles[t];
        const EnsDomain d = dom[tl.dom];
        wave::tile_mixed<R, OW>(f, p, o, d.W, d.off, Bounds{d.xs, d.xe, d.ys, d.ye},
                                tl.strip, tl.band, wave::split(tide[tl.dom]));
    }
}

void check_shape(const EnsShape &sh)
{
    if (sh.blocks <= 0 || sh.blocks % 8)
        NEMO_FATAL("ensemble launch: %d workgroups is not a positive multiple of 8",
                   sh.blocks);
    if (sh.waves_per_block != 1 && sh.waves_per_block != 2 && sh.waves_per_block != 4)
        NEMO_FATAL("ensemble launch: %d wavefronts per workgroup", sh.waves_per_block);
    if (sh.cols != 61 && sh.cols != 56)
        NEMO_FATAL("ensemble launch: %d columns per wavefront", sh.cols);
}

// rows x columns of the tile -> one kernel instantiation
#define NEMO_ENS_SHAPES(X) \
    X(4, 61) X(8, 61) X(16, 61) X(32, 61) X(64, 61) X(4, 56) X(8, 56) X(16, 56) X(32, 56) X(64, 56)

}  // namespace

void launch_ens_f64(hipStream_t s, const EnsShape &sh, const StepIn<double> &in,
                    const StepScalars<double> &p, const StepOut<double> &out,
                    const EnsDomain *dom, const EnsTile *tiles, int ntiles,
                    const double *tide)
{
    check_shape(sh);
    const dim3 grid((unsigned)sh.blocks), block((unsigned)(64 * sh.waves_per_block));
#define NEMO_ENS_F64(R, OW)                                                                \
    if (sh.rows == R && sh.cols == OW) {                                                   \
        k_ens_f64<R, OW><<<grid, block, 0, s>>>(in, p, out, dom, tiles, ntiles, tide);    \
        NEMO_LAUNCH_CHECK();                                                               \
        return;                                                                            \
    }
    NEMO_ENS_SHAPES(NEMO_ENS_F64)
#undef NEMO_ENS_F64
    NEMO_FATAL("ensemble launch: unsupported tile %d rows x %d columns", sh.rows, sh.cols);
}

void launch_ens_mixed(hipStream_t s, const EnsShape &sh, const MixedIn &in,
                      const StepScalars<float> &p, const MixedOut &out,
                      const EnsDomain *dom, const EnsTile *tiles, int ntiles,
                      const double *tide)
{
    check_shape(sh);
    const dim3 grid((unsigned)sh.blocks), block((unsigned)(64 * sh.waves_per_block));
#define NEMO_ENS_MIXED(R, OW)                                                              \
    if (sh.rows == R && sh.cols == OW) {                                                   \
        k_ens_mixed<R, OW><<<grid, block, 0, s>>>(in, p, out, dom, tiles, ntiles, tide);  \
        NEMO_LAUNCH_CHECK();                                                               \
        return;                                                                            \
    }
    NEMO_ENS_SHAPES(NEMO_ENS_MIXED)
#undef NEMO_ENS_MIXED
    NEMO_FATAL("ensemble launch: unsupported tile %d rows x %d columns", sh.rows, sh.cols);
}

int ens_max_waves_per_cu(int precision)
{
    int blocks = 0;
    if (precision == 64)
        NEMO_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &blocks, reinterpret_cast<const void *>(&k_ens_f64<32, 61>), 64, 0));
    else
        NEMO_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(
            &blocks, reinterpret_cast<const void *>(&k_ens_mixed<32, 61>), 64, 0));
    return blocks;
}

}  // namespace nemo
