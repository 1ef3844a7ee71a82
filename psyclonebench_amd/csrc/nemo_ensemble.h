// Batched ensemble of independent NEMOLite2D domains (nemo_ensemble.hip):
// every domain's arrays live in one arena per field, each domain at its own
// element offset with its own pitch and internal bounds; one persistent
// launch per step walks a tile -> domain work list.
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
    int pad[2];
};

// One tile of the work list: domain, strip (61 columns), band (R rows).
// Tiles are listed domain by domain, bands outer, strips inner.
struct EnsTile {
    int dom, strip, band, pad;
};

// Launch shape of the persistent step kernel.
struct EnsShape {
    int rows;              // rows per band (tile height): 4, 8, 16, 32 or 64
    int waves_per_block;   // wavefronts per workgroup: 1, 2 or 4
    int blocks;            // workgroups in the launch (a multiple of 8)
    int cols;              // owned columns per wavefront (tile width): 61 or 56
};

// FP64 step of every domain: state `in` -> `out` (arena base pointers).
// tide: the bc_ssh elevation of each domain for this step.
void launch_ens_f64(hipStream_t s, const EnsShape &shape, const StepIn<double> &in,
                    const StepScalars<double> &p, const StepOut<double> &out,
                    const EnsDomain *dom, const EnsTile *tiles, int ntiles,
                    const double *tide);

// Mixed-precision step of every domain (FP32 fields, compensated ssh).
void launch_ens_mixed(hipStream_t s, const EnsShape &shape, const MixedIn &in,
                      const StepScalars<float> &p, const MixedOut &out,
                      const EnsDomain *dom, const EnsTile *tiles, int ntiles,
                      const double *tide);

// Resident wavefronts per CU of the step kernels (from their register use),
// for sizing persistent launches.
int ens_max_waves_per_cu(int precision);

}  // namespace nemo
