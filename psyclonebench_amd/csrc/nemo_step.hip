// NEMOLite2D step kernels for CDNA4 (gfx950), FP64.
//
// Two execution forms of the manual PSy layer's invoke_time_step
// (manual_versions/psykal_cpp/time_step.cpp:53-136):
//
//  * unfused: one kernel per reference sweep, in the reference order, each
//    updating the reference's arrays in place. Exact for any input, used
//    for the host-pointer entry point and as the correctness ladder;
//  * fused: one kernel computes everything of a step except Flather for
//    the whole grid, from an old-state buffer set into a new one
//    (ping-pong), followed by a small Flather kernel over the open-boundary
//    faces only.
//
// Built with -ffp-contract=off so the double arithmetic rounds exactly like
// the reference C.
#include "nemo_step.h"
#include "nemo_hip_common.h"

namespace nemo {

namespace {

constexpr int BX = 64;  // one wavefront across a row
constexpr int BY = 4;

inline dim3 region_grid(Region r)
{
    const int nx = r.x1 - r.x0 + 1, ny = r.y1 - r.y0 + 1;
    return dim3((unsigned)((nx + BX - 1) / BX), (unsigned)((ny + BY - 1) / BY));
}

inline bool empty(Region r) { return r.x1 < r.x0 || r.y1 < r.y0; }

// (ji, jj) of this thread inside region r; false when outside.
__device__ __forceinline__ bool region_point(Region r, int *ji, int *jj)
{
    *ji = r.x0 + (int)(blockIdx.x * BX + threadIdx.x);
    *jj = r.y0 + (int)(blockIdx.y * BY + threadIdx.y);
    return *ji <= r.x1 && *jj <= r.y1;
}

// ---------------------------------------------------------------- unfused

__global__ __launch_bounds__(BX *BY) void k_continuity(StepIn<double> f,
                                                       double *__restrict__ ssha,
                                                       int W, Region r,
                                                       double rdt)
{
    int ji, jj;
    if (!region_point(r, &ji, &jj)) return;
    const int c = jj * W + ji;
    ssha[c] = continuity_pt(f, c, W, rdt);
}

__global__ __launch_bounds__(BX *BY) void k_momentum_u(StepIn<double> f,
                                                       StepScalars<double> p,
                                                       double *__restrict__ ua,
                                                       int W, Region r)
{
    int ji, jj;
    if (!region_point(r, &ji, &jj)) return;
    const int c = jj * W + ji;
    double v;
    if (momentum_u_pt(f, p, c, W, &v)) ua[c] = v;
}

__global__ __launch_bounds__(BX *BY) void k_momentum_v(StepIn<double> f,
                                                       StepScalars<double> p,
                                                       double *__restrict__ va,
                                                       int W, Region r)
{
    int ji, jj;
    if (!region_point(r, &ji, &jj)) return;
    const int c = jj * W + ji;
    double v;
    if (momentum_v_pt(f, p, c, W, &v)) va[c] = v;
}

__global__ __launch_bounds__(BX *BY) void k_bc_ssh(const int *__restrict__ tm,
                                                   double *__restrict__ ssha,
                                                   int W, Region r,
                                                   double tide)
{
    int ji, jj;
    if (!region_point(r, &ji, &jj)) return;
    const int c = jj * W + ji;
    if (tide_cell(tm, c, W)) ssha[c] = tide;
}

// bc_solid_u (stride 1) / bc_solid_v (stride W):
// boundary_conditions_kern.c:224-226, :248-250
__global__ __launch_bounds__(BX *BY) void k_bc_solid(const int *__restrict__ tm,
                                                     double *__restrict__ vel,
                                                     int W, int stride,
                                                     Region r)
{
    int ji, jj;
    if (!region_point(r, &ji, &jj)) return;
    const int c = jj * W + ji;
    if (tm[c] * tm[c + stride] == 0) vel[c] = 0.0;
}

__global__ __launch_bounds__(BX *BY) void k_copy3(
    double *__restrict__ un, const double *__restrict__ ua,
    double *__restrict__ vn, const double *__restrict__ va,
    double *__restrict__ sshn_t, const double *__restrict__ ssha_t, int W,
    Region r)
{
    int ji, jj;
    if (!region_point(r, &ji, &jj)) return;
    const int c = jj * W + ji;
    un[c] = ua[c];
    vn[c] = va[c];
    sshn_t[c] = ssha_t[c];
}

__global__ __launch_bounds__(BX *BY) void k_next_ssh(
    const int *__restrict__ tm, const double *__restrict__ sshn,
    const double *__restrict__ area_t, const double *__restrict__ area_face,
    double *__restrict__ ssh_face, int W, int stride, Region r)
{
    int ji, jj;
    if (!region_point(r, &ji, &jj)) return;
    const int c = jj * W + ji, q = c + stride;
    double v;
    if (next_ssh_pt(sshn[c], sshn[q], area_t[c], area_t[q], area_face[c], tm[c],
                    tm[q], &v))
        ssh_face[c] = v;
}

// Flather over the segment lists: one thread walks one segment in order,
// which is the order of the reference's serial sweep within that row (U) or
// column (V); different segments touch disjoint faces.
// boundary_conditions_kern.c:282-292 (U), :322-332 (V).
__global__ __launch_bounds__(64) void k_flather(FlatherList lu, FlatherList lv,
                                                double *ua, double *va,
                                                const double *ua_prev,
                                                const double *va_prev,
                                                const double *__restrict__ su,
                                                const double *__restrict__ sv)
{
    int t = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const FlatherList *L = &lu;
    double *vel = ua;
    const double *vel_prev = ua_prev;
    const double *ssh = su;
    if (t >= lu.nseg) {
        t -= lu.nseg;
        if (t >= lv.nseg) return;
        L = &lv;
        vel = va;
        vel_prev = va_prev;
        ssh = sv;
    }
    const int k1 = L->seg_ptr[t + 1];
    for (int k = L->seg_ptr[t]; k < k1; k++) {
        const int c = L->idx[k], o = L->off[k];
        const double v_nb = o ? (L->prev[k] ? vel_prev[c + o] : vel[c + o]) : 0.0;
        const double s_nb = o ? ssh[c + o] : 0.0;
        vel[c] = v_nb + L->coef[k] * (ssh[c] - s_nb);
    }
}

__global__ void k_fill_int(int *p, int n, int value)
{
    const int i = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (i < n) p[i] = value;
}

// ------------------------------------------------------------------ fused
//
// Block = 64 x 4 T cells; thread (tx, ty) owns T(i,j), U(i,j) and V(i,j).
// Phase 1: new ssh of the block's cells plus a one-cell halo to the east
//   and north (the only neighbours next_sshu/v read) into LDS; the block's
//   own cells also go to HBM.
// Phase 2: momentum_u/v + bc_solid for the thread's faces.
// Phase 3 (after the barrier): next_sshu/v from LDS.
// Every old-state value is read from HBM once per block; neighbours come
// from L1/L2.
constexpr int FX = 64;
constexpr int FY = 4;

__device__ __forceinline__ bool in_rect(int i, int j, int x0, int x1, int y0,
                                        int y1)
{
    return i >= x0 && i <= x1 && j >= y0 && j <= y1;
}

__device__ __forceinline__ double new_ssh(const StepIn<double> &f, int c, int W,
                                          double rdt, double tide)
{
    const double v = continuity_pt(f, c, W, rdt);
    return tide_cell(f.tmask, c, W) ? tide : v;
}

__global__ __launch_bounds__(FX *FY) void k_fused_step(
    StepIn<double> f, StepScalars<double> p, StepOut<double> o, int W,
    Bounds b, const double *__restrict__ tide_tab, int tide_k, double tide_val)
{
    __shared__ double s_ssh[FY + 1][FX + 1];

    const int tx = (int)threadIdx.x, ty = (int)threadIdx.y;
    const int i0 = b.xs - 1 + (int)blockIdx.x * FX;
    const int j0 = b.ys - 1 + (int)blockIdx.y * FY;
    const int ji = i0 + tx, jj = j0 + ty;
    const double tide = tide_tab ? tide_tab[tide_k] : tide_val;
    const int *__restrict__ tm = f.tmask;

    // phase 1: new ssh, own cell + halo
    double ssh_c = 0.0;
    if (in_rect(ji, jj, b.xs, b.xe, b.ys, b.ye)) {
        ssh_c = new_ssh(f, jj * W + ji, W, p.rdt, tide);
        o.ssh[jj * W + ji] = ssh_c;
    }
    s_ssh[ty][tx] = ssh_c;
    if (ty == 0) {
        const int jh = j0 + FY;
        s_ssh[FY][tx] = in_rect(ji, jh, b.xs, b.xe, b.ys, b.ye)
                            ? new_ssh(f, jh * W + ji, W, p.rdt, tide)
                            : 0.0;
    }
    if (ty == 1 && tx < FY) {
        const int ih = i0 + FX, jh = j0 + tx;
        s_ssh[tx][FX] = in_rect(ih, jh, b.xs, b.xe, b.ys, b.ye)
                            ? new_ssh(f, jh * W + ih, W, p.rdt, tide)
                            : 0.0;
    }

    // phase 2: faces (the M x N extent, i.e. up to xe+1 / ye+1)
    const int c = jj * W + ji;
    if (ji <= b.xe + 1 && jj <= b.ye + 1) {
        double v;
        bool wrote = false;
        if (in_rect(ji, jj, b.xs, b.xe - 1, b.ys, b.ye))
            wrote = momentum_u_pt(f, p, c, W, &v);
        if (!wrote && in_rect(ji, jj, b.xs - 1, b.xe, b.ys - 1, b.ye + 1) &&
            tm[c] * tm[c + 1] == 0) {
            v = 0.0;
            wrote = true;
        }
        if (wrote) o.u[c] = v;

        wrote = false;
        if (in_rect(ji, jj, b.xs, b.xe, b.ys, b.ye - 1))
            wrote = momentum_v_pt(f, p, c, W, &v);
        if (!wrote && in_rect(ji, jj, b.xs - 1, b.xe + 1, b.ys - 1, b.ye) &&
            tm[c] * tm[c + W] == 0) {
            v = 0.0;
            wrote = true;
        }
        if (wrote) o.v[c] = v;
    }

    __syncthreads();

    // phase 3: face ssh from the new cell ssh
    if (in_rect(ji, jj, b.xs, b.xe - 1, b.ys, b.ye)) {
        double v;
        if (next_ssh_pt(ssh_c, s_ssh[ty][tx + 1], f.area_t[c], f.area_t[c + 1],
                        f.area_u[c], tm[c], tm[c + 1], &v))
            o.sshu[c] = v;
    }
    if (in_rect(ji, jj, b.xs, b.xe, b.ys, b.ye - 1)) {
        double v;
        if (next_ssh_pt(ssh_c, s_ssh[ty + 1][tx], f.area_t[c], f.area_t[c + W],
                        f.area_v[c], tm[c], tm[c + W], &v))
            o.sshv[c] = v;
    }
}

}  // namespace

// ---------------------------------------------------------------- launchers

void launch_continuity(hipStream_t s, const StepIn<double> &f, double *ssha,
                       int W, Region r, double rdt)
{
    if (empty(r)) return;
    k_continuity<<<region_grid(r), dim3(BX, BY), 0, s>>>(f, ssha, W, r, rdt);
    NEMO_LAUNCH_CHECK();
}

void launch_momentum_u(hipStream_t s, const StepIn<double> &f,
                       const StepScalars<double> &p, double *ua, int W,
                       Region r)
{
    if (empty(r)) return;
    k_momentum_u<<<region_grid(r), dim3(BX, BY), 0, s>>>(f, p, ua, W, r);
    NEMO_LAUNCH_CHECK();
}

void launch_momentum_v(hipStream_t s, const StepIn<double> &f,
                       const StepScalars<double> &p, double *va, int W,
                       Region r)
{
    if (empty(r)) return;
    k_momentum_v<<<region_grid(r), dim3(BX, BY), 0, s>>>(f, p, va, W, r);
    NEMO_LAUNCH_CHECK();
}

void launch_bc_ssh(hipStream_t s, const int *tmask, double *ssha, int W,
                   Region r, double tide)
{
    if (empty(r)) return;
    k_bc_ssh<<<region_grid(r), dim3(BX, BY), 0, s>>>(tmask, ssha, W, r, tide);
    NEMO_LAUNCH_CHECK();
}

void launch_bc_solid(hipStream_t s, const int *tmask, double *vel, int W,
                     int stride, Region r)
{
    if (empty(r)) return;
    k_bc_solid<<<region_grid(r), dim3(BX, BY), 0, s>>>(tmask, vel, W, stride,
                                                       r);
    NEMO_LAUNCH_CHECK();
}

void launch_flather(hipStream_t s, const FlatherList &lu, const FlatherList &lv,
                    double *ua, double *va, const double *ua_prev,
                    const double *va_prev, const double *sshn_u,
                    const double *sshn_v)
{
    const int n = lu.nseg + lv.nseg;
    if (n == 0) return;
    k_flather<<<dim3((unsigned)((n + 63) / 64)), dim3(64), 0, s>>>(
        lu, lv, ua, va, ua_prev, va_prev, sshn_u, sshn_v);
    NEMO_LAUNCH_CHECK();
}

void launch_copy3(hipStream_t s, double *un, const double *ua, double *vn,
                  const double *va, double *sshn_t, const double *ssha_t,
                  int W, Region r)
{
    if (empty(r)) return;
    k_copy3<<<region_grid(r), dim3(BX, BY), 0, s>>>(un, ua, vn, va, sshn_t,
                                                    ssha_t, W, r);
    NEMO_LAUNCH_CHECK();
}

void launch_next_ssh(hipStream_t s, const int *tmask, const double *sshn_t,
                     const double *area_t, const double *area_face,
                     double *ssh_face, int W, int stride, Region r)
{
    if (empty(r)) return;
    k_next_ssh<<<region_grid(r), dim3(BX, BY), 0, s>>>(
        tmask, sshn_t, area_t, area_face, ssh_face, W, stride, r);
    NEMO_LAUNCH_CHECK();
}

void launch_fused_step(hipStream_t s, const StepIn<double> &f,
                       const StepScalars<double> &p, const StepOut<double> &o,
                       int W, Bounds b, const double *tide_tab, int tide_k,
                       double tide_val)
{
    // cover T/U/V points (xs-1 .. xe+1) x (ys-1 .. ye+1)
    const int nx = b.xe - b.xs + 3, ny = b.ye - b.ys + 3;
    const dim3 grid((unsigned)((nx + FX - 1) / FX), (unsigned)((ny + FY - 1) / FY));
    k_fused_step<<<grid, dim3(FX, FY), 0, s>>>(f, p, o, W, b, tide_tab, tide_k,
                                               tide_val);
    NEMO_LAUNCH_CHECK();
}

void launch_fill_int(hipStream_t s, int *p, int n, int value)
{
    if (n <= 0) return;
    k_fill_int<<<dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s>>>(p, n,
                                                                     value);
    NEMO_LAUNCH_CHECK();
}

}  // namespace nemo
