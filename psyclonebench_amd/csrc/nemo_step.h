// Launchers for the NEMOLite2D step kernels (defined in nemo_step.hip).
// Host code in nemo_psy.cpp drives them; all pointers are device pointers.
#pragma once

#include <hip/hip_runtime.h>

#include "nemo_point.h"

namespace nemo {

// Inclusive 0-based rectangle.
struct Region {
    int x0, x1, y0, y1;
};

// Internal T region bounds as passed through the C ABI (0-based).
struct Bounds {
    int xs, xe, ys, ye;
};

// Writable outputs of one fused step (ping-pong targets).
template <typename T>
struct StepOut {
    T *ssh;   // new sshn_t (== ssha_t)
    T *u;     // new un (== ua)
    T *v;     // new vn (== va)
    T *sshu;  // new sshn_u
    T *sshv;  // new sshn_v
};

// Flather faces grouped in segments that must run in order: one segment per
// row for U faces, one per column for V faces (bc_flather_u/v read the
// neighbour face the serial sweep may already have updated). Entry k of a
// segment updates vel[idx[k]] from vel[idx[k] + off[k]]; off == 0 marks a
// neighbour outside the allocation (read as 0). prev[k] != 0 marks a
// neighbour that is itself a Flather face later in the sweep: the serial
// sweep reads it before updating it, so it reads the value that face held
// at the end of the previous step, which a caller that ping-pongs its
// buffers keeps in the previous step's set (vel_prev), not in `vel`.
struct FlatherList {
    const int *seg_ptr;   // nseg + 1
    const int *idx;       // face index
    const int *off;       // +-1 (U) or +-W (V), or 0
    const int *prev;      // 1: read the neighbour from vel_prev
    const double *coef;   // sqrt(g / h) at the face
    int nseg;
};

// --- unfused kernels, one per reference kernel ---------------------------
void launch_continuity(hipStream_t s, const StepIn<double> &f, double *ssha,
                       int W, Region r, double rdt);
void launch_momentum_u(hipStream_t s, const StepIn<double> &f,
                       const StepScalars<double> &p, double *ua, int W,
                       Region r);
void launch_momentum_v(hipStream_t s, const StepIn<double> &f,
                       const StepScalars<double> &p, double *va, int W,
                       Region r);
void launch_bc_ssh(hipStream_t s, const int *tmask, double *ssha, int W,
                   Region r, double tide);
void launch_bc_solid(hipStream_t s, const int *tmask, double *vel, int W,
                     int stride, Region r);
// ua_prev / va_prev: the velocities at the end of the previous step (the
// previous ping-pong set, or ua / va themselves when they are updated in
// place or on a first step that writes into the after arrays).
void launch_flather(hipStream_t s, const FlatherList &lu, const FlatherList &lv,
                    double *ua, double *va, const double *ua_prev,
                    const double *va_prev, const double *sshn_u,
                    const double *sshn_v);
void launch_copy3(hipStream_t s, double *un, const double *ua, double *vn,
                  const double *va, double *sshn_t, const double *ssha_t,
                  int W, Region r);
void launch_next_ssh(hipStream_t s, const int *tmask, const double *sshn_t,
                     const double *area_t, const double *area_face,
                     double *ssh_face, int W, int stride, Region r);

// --- fused step -----------------------------------------------------------
// One launch computes the new ssh (continuity + bc_ssh), the post-solid
// u/v (momentum + bc_solid) and the new face ssh (next_sshu/v) for the whole
// M x N extent, reading `f` and writing `o`. Flather runs afterwards
// (launch_flather) on o.u / o.v with the old face ssh.
// The tide value is tide_tab[tide_k] when tide_tab is non-null, else
// tide_val.
void launch_fused_step(hipStream_t s, const StepIn<double> &f,
                       const StepScalars<double> &p, const StepOut<double> &o,
                       int W, Bounds b, const double *tide_tab, int tide_k,
                       double tide_val);

// Row-march form of the fused step (nemo_march.hip): same inputs, outputs
// and arithmetic as launch_fused_step; `variant` picks the block width and
// rows per band (0 .. march_variant_count()-1).
void launch_fused_march(hipStream_t s, int variant, const StepIn<double> &f,
                        const StepScalars<double> &p, const StepOut<double> &o,
                        int W, Bounds b, const double *tide_tab, int tide_k,
                        double tide_val);
int march_variant_count();
void march_variant_shape(int variant, int *mx, int *rows, int *prefetch);
const char *march_variant_kernel(int variant);

// Small helpers.
void launch_fill_int(hipStream_t s, int *p, int n, int value);

}  // namespace nemo
