// Allocation-free form of the NEMOLite2D step: c_invoke_time_step_noalloc.
//
// One step with the semantics of the host-pointer PSy layer
// (manual_versions/psykal_cpp/time_step.cpp:53-136: every sweep in order, in
// place, now <- after copied over the reference's copy region) on 25 device
// arrays that the caller owns, for callers that cannot allocate device
// memory (code under HIP graph capture, tight serving loops). It runs the
// unfused kernels of nemo_step.hip. Two things the allocating forms keep in
// buffers of their own are arranged differently:
//
//  * the Coriolis factors (2*omega)*sin(gphi*d2r) are evaluated on the
//    device into `scratch` at every call (the allocating forms evaluate them
//    once on the host with libm; the device sin may differ from it in the
//    last bit, far inside the 1e-12 relative bound of the FP64 step);
//  * the Flather faces are walked straight from the mask, one thread per U
//    row and per V column in the reference's serial order, instead of over
//    host-built segment lists.
//
// `scratch` is the only device memory beyond the caller's arrays. It is a
// file-scope array: each GPU holds one copy from the library's load until
// the process ends. Host code never passes the array's name to a kernel (in
// host code it names a host-side stand-in); it passes the device address
// that hipGetSymbolAddress returns for the current device.
#include "nemo_hip_common.h"
#include "nemo_host.h"
#include "nemo_step.h"
#include "nemolite2d_hip.h"

namespace nemo {

// Two FP64 arrays of at most this many elements each: 32 MiB in all.
constexpr int kNoallocMaxElems = 1 << 21;

static __device__ __align__(256) double scratch[2 * (size_t)kNoallocMaxElems];

namespace {

constexpr int kMaxDevices = 64;
double *g_scratch_ptr[kMaxDevices] = {};

// Device address of this device's scratch array, looked up once per device
// with that device current.
double *scratch_ptr()
{
    int dev = 0;
    NEMO_HIP_CHECK(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices)
        NEMO_FATAL("c_invoke_time_step_noalloc: device index %d beyond %d", dev,
                   kMaxDevices);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        NEMO_HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) NEMO_FATAL("c_invoke_time_step_noalloc: no device address for scratch");
        g_scratch_ptr[dev] = static_cast<double *>(p);
    }
    return g_scratch_ptr[dev];
}

// fcor = (2*omega)*sin(gphi*d2r) at every element, the expression of
// momentum_u_kern.c:251 / momentum_v_kern.c:230.
__global__ __launch_bounds__(256) void k_coriolis(const double *__restrict__ gphiu,
                                                  const double *__restrict__ gphiv,
                                                  double *__restrict__ fcor_u,
                                                  double *__restrict__ fcor_v, int n,
                                                  double omega, double d2r)
{
    const int k = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    if (k >= n) return;
    fcor_u[k] = 2. * omega * sin(gphiu[k] * d2r);
    fcor_v[k] = 2. * omega * sin(gphiv[k] * d2r);
}

// bc_flather_u / bc_flather_v (boundary_conditions_kern.c:271-293,
// :313-333) at face c with neighbours along `stride`; a neighbour outside
// the allocation reads as 0, as in the segment-list kernel.
__device__ __forceinline__ void flather_face(int c, int stride, int total,
                                             const int *__restrict__ tm,
                                             double *__restrict__ vel,
                                             const double *__restrict__ h,
                                             const double *__restrict__ ssh, double g)
{
    const int mc = tm[c], mp = tm[c + stride];
    if (mc + mp <= -1) return;
    int d;
    if (mc < 0)
        d = stride;
    else if (mp < 0)
        d = -stride;
    else
        return;
    const int q = c + d;
    const bool in = q >= 0 && q < total;
    const double v_nb = in ? vel[q] : 0.0;
    const double s_nb = in ? ssh[q] : 0.0;
    vel[c] = v_nb + sqrt(g / h[c]) * (ssh[c] - s_nb);
}

// Thread t < nu walks U row ys-1+t over ji = xs-1 .. xe; the next nv
// threads walk V column xs-1+(t-nu) over jj = ys-1 .. ye. Within a row (a
// column) a face may read the one the sweep wrote just before it, so each
// walk keeps the reference's order; rows and columns are independent.
__global__ __launch_bounds__(64) void k_flather_walk(const int *__restrict__ tm,
                                                     double *__restrict__ ua,
                                                     double *__restrict__ va,
                                                     const double *__restrict__ hu,
                                                     const double *__restrict__ hv,
                                                     const double *__restrict__ su,
                                                     const double *__restrict__ sv,
                                                     int W, int total, Bounds b, double g)
{
    const int t = (int)(blockIdx.x * blockDim.x + threadIdx.x);
    const int nu = b.ye - b.ys + 3, nv = b.xe - b.xs + 3;
    if (t < nu) {
        const int row = (b.ys - 1 + t) * W;
        for (int ji = b.xs - 1; ji <= b.xe; ji++)
            flather_face(row + ji, 1, total, tm, ua, hu, su, g);
    } else if (t < nu + nv) {
        const int ji = b.xs - 1 + (t - nu);
        for (int jj = b.ys - 1; jj <= b.ye; jj++)
            flather_face(jj * W + ji, W, total, tm, va, hv, sv, g);
    }
}

}  // namespace
}  // namespace nemo

using namespace nemo;

extern "C" void c_invoke_time_step_noalloc(NEMO_HOST_ARRAYS, int istep, int internal_xstart,
                                           int internal_xstop, int internal_ystart,
                                           int internal_ystop, int width, int total_size,
                                           double rdt, double cbfr, double visc,
                                           double omega, double d2r, double g,
                                           void *stream)
{
    const Geometry G{internal_xstart, internal_xstop, internal_ystart, internal_ystop,
                     width, total_size};
    check_geometry(G);
    if (total_size > kNoallocMaxElems)
        NEMO_FATAL("c_invoke_time_step_noalloc: total_size %d exceeds the %d elements "
                   "of its scratch arrays", total_size, kNoallocMaxElems);
    double *const fcor_u = scratch_ptr();
    double *const fcor_v = fcor_u + kNoallocMaxElems;
    hipStream_t s = static_cast<hipStream_t>(stream);
    // hipGetLastError also returns what earlier calls of other code on this
    // thread left behind (an event query's "not ready", say); clear it so
    // that the launch checks below report this call's launches only.
    (void)hipGetLastError();

    const int n = total_size;
    k_coriolis<<<dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s>>>(
        gphiu, gphiv, fcor_u, fcor_v, n, omega, d2r);
    NEMO_LAUNCH_CHECK();

    StepIn<double> f;
    f.sshn_t = sshn_t;
    f.sshn_u = sshn_u;
    f.sshn_v = sshn_v;
    f.un = un;
    f.vn = vn;
    f.ht = ht;
    f.hu = hu;
    f.hv = hv;
    f.ssha_u = ssha_u;
    f.ssha_v = ssha_v;
    f.area_t = area_t;
    f.area_u = area_u;
    f.area_v = area_v;
    f.dx_u = dx_u;
    f.dx_v = dx_v;
    f.dx_t = dx_t;
    f.dy_u = dy_u;
    f.dy_v = dy_v;
    f.dy_t = dy_t;
    f.fcor_u = fcor_u;
    f.fcor_v = fcor_v;
    f.tmask = tmask;
    const StepScalars<double> p{rdt, visc, g, 1.0 + cbfr * rdt};
    const int W = width;
    const int xs = G.xs, xe = G.xe, ys = G.ys, ye = G.ye;
    const Region interior{xs, xe, ys, ye};

    // the sweep order and regions of time_step.cpp:53-136
    launch_continuity(s, f, ssha_t, W, interior, rdt);
    launch_momentum_u(s, f, p, ua, W, Region{xs, xe - 1, ys, ye});
    launch_momentum_v(s, f, p, va, W, Region{xs, xe, ys, ye - 1});
    launch_bc_ssh(s, tmask, ssha_t, W, interior, tide_at(istep, rdt));
    launch_bc_solid(s, tmask, ua, W, 1, Region{xs - 1, xe, ys - 1, ye + 1});
    launch_bc_solid(s, tmask, va, W, W, Region{xs - 1, xe + 1, ys - 1, ye});
    const int walkers = (ye - ys + 3) + (xe - xs + 3);
    k_flather_walk<<<dim3((unsigned)((walkers + 63) / 64)), dim3(64), 0, s>>>(
        tmask, ua, va, hu, hv, sshn_u, sshn_v, W, total_size, Bounds{xs, xe, ys, ye}, g);
    NEMO_LAUNCH_CHECK();
    launch_copy3(s, un, ua, vn, va, sshn_t, ssha_t, W, Region{xs - 1, xe + 1, ys - 1, ye});
    launch_next_ssh(s, tmask, sshn_t, area_t, area_u, sshn_u, W, 1,
                    Region{xs, xe - 1, ys, ye});
    launch_next_ssh(s, tmask, sshn_t, area_t, area_v, sshn_v, W, W,
                    Region{xs, xe, ys, ye - 1});
}
