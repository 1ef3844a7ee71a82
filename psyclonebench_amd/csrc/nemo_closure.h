// Learned coarse-grid closure (no reference counterpart): a 4-layer 5x5
// convolutional network on the coarse grid's interior T points, run with
// bf16 MFMA (FP32 accumulate), trained on the GPU, and the plumbing between
// the fine and the coarse NEMOLite2D states (nemo_closure.hip).
//
// Activation tensors are bf16 on a padded grid with a 2-pixel zero border:
// padded pixel P = (y+2)*(nx+4) + x+2 for interior pixel (y, x), x fastest.
// Hidden activations and gradients are stored in planes of 16 channels
// (channel c at (c / 16) * plane + P * 16 + c % 16, plane = (ny+4)(nx+4)*16);
// the network input holds its IN0C channels at P * IN0C + c. The border is
// never written, so the 5x5 stencil needs no bounds tests.
#pragma once

#include <hip/hip_runtime.h>

namespace nemo {
namespace cl {

constexpr int TAPS = 25;   // 5 x 5 kernel
constexpr int HID = 32;    // hidden channels
constexpr int CIN0 = 16;   // input channels of the layer-0 weights (4 used: u, v, ssh, depth)
constexpr int IN0C = 4;    // channels of the stored network input
constexpr int COUT4 = 16;  // output channels as computed (3 used: du, dv, dssh)
constexpr int NLAYER = 4;

// Layer l: weights [TAPS][cout][cin] then bias[cout], FP32 master copy.
constexpr int layer_cin(int l) { return l == 0 ? CIN0 : HID; }
constexpr int layer_cout(int l) { return l == NLAYER - 1 ? COUT4 : HID; }
constexpr int layer_wsize(int l) { return TAPS * layer_cin(l) * layer_cout(l); }
constexpr int layer_psize(int l) { return layer_wsize(l) + layer_cout(l); }
constexpr int layer_offset(int l) { return l == 0 ? 0 : layer_offset(l - 1) + layer_psize(l - 1); }
constexpr int param_count() { return layer_offset(NLAYER); }

// Coarse dl_esm_inf state seen by the closure: T-cell ssh, face velocities
// and face ssh (current buffers of the coarse model), depth, mask, areas;
// interior T points (1..nx, 1..ny), pitch W.
struct CoarseState {
    double *ssh, *u, *v, *sshu, *sshv;
    const double *ht;
    const int *tmask;
    const double *area_t, *area_u, *area_v;
    int W;
};

struct Norm {
    float in_scale[4], in_shift[4];  // input channel c: (x - shift) / scale
    float out_scale[3];              // correction = network output * scale
    // Activity gate: the network output of a point is multiplied by
    // 1 - exp(-(a / gate_eps)^2), a = |u| + |v| + |ssh| in input units, so
    // that a point at rest gets no correction (gate_eps <= 0: no gate).
    float gate_eps;
    // ssh cells corrected and learned: 1 every wet cell, 0 only the cells
    // the step does not clamp to the tide (wet, no open 4-neighbour)
    int ssh_wet;
    // the energy guard also rejects a correction with any applied component
    // beyond max_out in output units (0: no such bound)
    float max_out;
};

// ---- network kernels -------------------------------------------------------
// epilogues of the convolution
enum Epilogue { EPI_HIDDEN = 0, EPI_OUTPUT = 1, EPI_DX = 2 };

struct ConvArgs {
    const __bf16 *in;     // padded NHWC, cin channels
    const __bf16 *w;      // [TAPS][cout][cin]
    const float *bias;    // cout (EPI_HIDDEN, EPI_OUTPUT)
    __bf16 *out;          // padded NHWC, cout channels (EPI_HIDDEN act, EPI_DX grad)
    __bf16 *pre;          // padded NHWC pre-activation (EPI_HIDDEN), may be null
    const __bf16 *gpre;   // padded NHWC pre-activation multiplying the gradient (EPI_DX)
    float *outf;          // [ny*nx][4] FP32 network output (EPI_OUTPUT)
    const float *gate;    // [ny*nx] multiplier of the output (EPI_OUTPUT), may be null
    int nx, ny;
};

void launch_conv(hipStream_t s, int cin, int cout, Epilogue epi, const ConvArgs &a);
// Launch form of the 16/32 -> 32 convolutions (process-wide): 0, two tiles
// per wavefront pass (4 wavefronts a workgroup); 1 (default), one tile (8
// wavefronts). Both give the same bits.
void set_conv_form(int form);
int conv_form();
// dW partial sums: partial[block][TAPS][co][ci] over the pixels of each of
// the dw_blocks(nx, ny) workgroups
int dw_blocks(int nx, int ny);
void launch_dw(hipStream_t s, int co, int ci, const __bf16 *dpre, const __bf16 *in,
               float *partial, int nx, int ny);
int bias_blocks();
void launch_bias_partial(hipStream_t s, int co, const __bf16 *dpre, float *partial,
                         int nx, int ny);
// out[i] = sum over b < nblk of partial[b*n + i]
void launch_reduce(hipStream_t s, const float *partial, int nblk, int n, float *out);
// Adam (kind 1) or SGD with momentum (kind 0) over the FP32 master params,
// the gradient first scaled to a global norm of at most max_norm (if > 0);
// norm_scale (2 floats, device) receives the gradient norm and that factor
void launch_optimizer(hipStream_t s, int kind, float *w, const float *g, float *m,
                      float *v, int n, float lr, float b1, float b2, float eps,
                      float bc1, float bc2, float max_norm, float *norm_scale);
// bf16 weights of layer l: forward [TAPS][cout][cin] and the flipped,
// transposed copy for the input-gradient convolution [TAPS][cin][cout]
void launch_weights_bf16(hipStream_t s, const float *w, int cin, int cout,
                         __bf16 *fwd, __bf16 *flip);
// coarse state -> padded network input (IN0C channels) and the activity gate
void launch_pack(hipStream_t s, const CoarseState &c, const Norm &nm, __bf16 *in0,
                 float *gate, int nx, int ny);
// target[(y*nx+x)*4 + k] = (F_k - C_k) / out_scale_k, k = u, v, ssh
void launch_target(hipStream_t s, const double *fssh, const double *fu, const double *fv,
                   const CoarseState &c, const Norm &nm, float *target, int nx, int ny);
// Per-channel weights (u, v, ssh) of the squared error in the loss.
struct LossWeights {
    float w[3];
};

// Weighted MSE loss over the 3 used channels, sum_c w_c sum_p d_c^2 / (3 npx):
// gradient into dpre4 (COUT4 channels, 16-channel plane) and per-block
// partial sums of the unweighted squared error per channel. With a mask
// (pitch W), the ssh of cells not corrected (Norm::ssh_wet) is left out.
int loss_blocks();
void launch_loss(hipStream_t s, const float *pred, const float *target, __bf16 *dpre4,
                 float *partial, int nx, int ny, const int *tmask, int W, int ssh_wet,
                 const float *gate, const LossWeights &lw);

// ---- fine <-> coarse plumbing ----------------------------------------------
// Box filter of the fine state onto the coarse grid (factor f): T cells by
// f x f average, U faces by the average of the f fine faces on the coarse
// face, V faces likewise. Fine interior cell i maps to coarse (i-1)/f + 1.
void launch_filter(hipStream_t s, const double *fssh, const double *fu, const double *fv,
                   int Wf, int factor, double *cssh, double *cu, double *cv, int Wc,
                   int nx, int ny);
// Writes ssh / u / v of the coarse state where the step always writes them
// (wet interior cells, faces between two wet interior cells): mode 0 sets
// them to (sssh, su, sv); mode 1 adds the network correction pred * scale.
// Only when *flag != 0 (flag may be null: always).
void launch_set_fields(hipStream_t s, int mode, const double *sssh, const double *su,
                       const double *sv, const float *pred, const Norm &nm,
                       const int *flag, const CoarseState &c, int nx, int ny);
// Face ssh from the T-cell ssh (next_sshu/v over their regions), when *flag.
void launch_face_ssh(hipStream_t s, const int *flag, const CoarseState &c, int nx, int ny);
// Energy guard: E = sum over interior cells of g*ssh^2 + ht*(u^2 + v^2),
// before and after the correction, and a non-finite count. The guarded output
// layer (launch_conv_out_guarded) writes them as per-workgroup partial sums
// (partial[block][3], guard_blocks(nx, ny) workgroups) beside the network
// output; launch_guard_decide adds them in workgroup order and sets
// flag = finite && E_after <= (1 + rel) * E_before + abs, counting trips.
// A component beyond Norm::max_out counts as non-finite.
struct GuardSums {
    Norm nm;
    CoarseState c;
    double g;
    double *partial;
};
int guard_blocks(int nx, int ny);
// the output layer (32 -> 16 channels, EPI_OUTPUT) with the guard's sums
void launch_conv_out_guarded(hipStream_t s, const ConvArgs &a, const GuardSums &gs);
void launch_guard_decide(hipStream_t s, const double *partial, int nblk, double rel,
                         double abs_tol, int *flag, int *trips);

}  // namespace cl
}  // namespace nemo
