// Batched ensemble of independent NEMOLite2D domains (nemo_ensemble.hip):
// every domain's arrays live in one arena per field, each domain at its own
// element offset with its own pitch and internal bounds; one launch per step
// walks a tile -> domain work list.
#pragma once

#include "nemo_mixed.h"
#include "nemo_step.h"

namespace nemo {

// One domain of the ensemble: element offset of its (0,0) point in every
// field arena, row pitch, 0-based internal bounds. Arena byte offsets stay
// below 2^32 (the kernels address with 32-bit offsets).
struct EnsDomain {
    unsigned off;
    int W, xs, xe, ys, ye;
};

// One tile of the work list: strip `strip` (of `cols` owned columns, 56 or
// 61) by band `band` of domain `dom`, with that domain's arena offset, pitch
// and internal bounds copied in, so that a wavefront reads one descriptor
// (two independent scalar loads) and no domain table before it issues its
// first row of loads. Tiles are listed domain by domain, bands outer, strips
// inner.
struct alignas(16) EnsTile {
    unsigned off;
    int W, xs, xe, ys, ye;
    int strip, band, dom, cols;
    int pad[2];
};
static_assert(sizeof(EnsTile) == 48, "tile descriptor: 12 dwords");

// Launch shape of the step kernel.
struct EnsShape {
    int rows;              // rows per band (tile height): 4, 8, 16, 32 or 64
    int waves_per_block;   // wavefronts per workgroup: 1, 2 or 4
    int blocks;            // persistent: workgroups in the launch (a multiple
                           // of 8); 0: one one-wavefront workgroup per tile
    int cols;              // owned columns per wavefront (tile width): 61 or
                           // 56, or 0: per domain, whichever of the two its
                           // width runs faster with (one wavefront per tile
                           // only)
    int prefetch;          // rows of loads issued ahead of the row that needs
                           // them, beyond the next: 0, 1 or 2 (2: one
                           // wavefront per tile only)
};

// True when a kernel instantiation exists for the shape's launch form, rows,
// columns and prefetch depth.
bool ens_shape_supported(const EnsShape &shape);

// FP64 step of every domain: state `in` -> `out` (arena base pointers).
// tide: the bc_ssh elevation of each domain for this step.
void launch_ens_f64(hipStream_t s, const EnsShape &shape, const StepIn<double> &in,
                    const StepScalars<double> &p, const StepOut<double> &out,
                    const EnsTile *tiles, int ntiles, const double *tide);

// Mixed-precision step of every domain (FP32 fields, compensated ssh).
void launch_ens_mixed(hipStream_t s, const EnsShape &shape, const MixedIn &in,
                      const StepScalars<float> &p, const MixedOut &out,
                      const EnsTile *tiles, int ntiles, const double *tide);

// Tile width of domain `b` for shapes of cols = 0: 61 where the domain needs
// enough fewer 61-column strips than 56-column ones to make up for the
// slower 61-column tile (a strip costs about 8% more, both precisions,
// measured on the ensemble workload), else 56.
int ens_domain_cols(const Bounds &b);

// Resident wavefronts per CU of the persistent step kernel of these rows,
// columns and prefetch depth (from its register use), for sizing its launch.
int ens_max_waves_per_cu(int precision, int rows, int cols, int prefetch);

}  // namespace nemo
