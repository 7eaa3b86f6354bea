// The per-sweep call without device allocation.
//
// vg_euler_time_step_2d_noalloc is vg_euler_time_step_2d (the reference's
// computeX|YSLICTimeStep2D / computeX|YFORCETimeStep2D on a caller-built ghosted
// snapshot) for callers that cannot allocate device memory: code under HIP graph
// capture, tight serving loops, harnesses that forbid hipMalloc.  Inputs and
// outputs are the caller's device arrays; the intermediate structure-of-arrays
// planes and the time-step clock live in `scratch`, a file-scope device array of
// fixed size.  Each device holds its own copy of the array until the process
// ends; its device address is looked up once per device with
// hipGetSymbolAddress and used for every launch.  Calls on one device share the
// array, so calls that overlap on one device race; every call rewrites the parts
// it reads (the clock by a kernel, the planes by the layout conversion) before
// reading them.
#include <mutex>

#include "vg_grid.h"
#include "vg_phys_euler.h"

namespace {

using vg::EulerPhys;

// Ghosted cells, (ny + 2G) x (nx + 2G), one call may sweep: 2^20 (a 1020 x 1020
// interior with SLIC ghosts).  Two grids of six planes: 96 MiB, plus the clock.
constexpr long long kMaxCells = 1ll << 20;
constexpr long long kClockDoubles = 32;   // vg::Clock, padded to 256 bytes
constexpr long long kScratchDoubles = kClockDoubles + 2ll * EulerPhys::NF * kMaxCells;
static_assert(sizeof(vg::Clock) <= kClockDoubles * sizeof(double), "clock does not fit its scratch slot");

static __device__ __align__(256) double scratch[kScratchDoubles];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_ptr[kMaxDevices] = {};   // per device: the device address of `scratch`

// The device address of `scratch` on the current device (looked up once per
// device, with that device current).
int scratch_address(int device, double **out)
{
    VG_CHECK(device >= 0 && device < kMaxDevices, "device index out of range for the no-allocation entry point");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        double *ptr = nullptr;
        hipError_t e = hipGetSymbolAddress((void **)&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || ptr == nullptr) {
            vg::set_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
            return VG_ERR_HIP;
        }
        g_scratch_ptr[device] = ptr;
    }
    *out = g_scratch_ptr[device];
    return VG_OK;
}

// The sweep's clock: the caller's time step, nothing else running.
__global__ void noalloc_clock_kernel(vg::Clock *c, double dt)
{
    vg::Clock h{};
    h.dt = dt;
    *c = h;
}

// Interior cells (rows G .. G + ny - 1, columns G .. G + nx - 1 of the ghosted
// planes) to the caller's array of structures.
template <int NF>
__global__ __launch_bounds__(256) void interior_to_aos_kernel(const double *__restrict__ soa, double *__restrict__ aos,
                                                              long long plane, int ny, int nx, int gx, int G)
{
    const long long cells = (long long)ny * nx;
    for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < cells; c += (long long)gridDim.x * blockDim.x) {
        const long long i = c / nx, j = c % nx;
        const long long g = (i + G) * gx + (j + G);
#pragma unroll
        for (int f = 0; f < NF; f++) aos[c * NF + f] = soa[f * plane + g];
    }
}

template <int ORDER, int MODE>
int launch_noalloc(hipStream_t s, double *base, double *cells, const double *with_boundary, int ny, int nx, int dir,
                   double dx, double dt, double bias, int limiter, const EulerPhys::Params &m)
{
    using P = EulerPhys;
    constexpr int NF = P::NF, G = (ORDER == 2) ? 2 : 1;
    constexpr int BSX = vg::Tiles<P>::bsx, RX = vg::Tiles<P>::rx, BSY = vg::Tiles<P>::bsy, HY = vg::Tiles<P>::hy;
    const int gy = ny + 2 * G, gx = nx + 2 * G;
    const long long gcells = (long long)gy * gx;
    const long long plane = (gcells + 63) / 64 * 64;
    vg::Clock *clk = reinterpret_cast<vg::Clock *>(base);
    double *a = base + kClockDoubles, *b = a + NF * plane;
    const int blocks = (int)std::min<long long>((gcells + 255) / 256, 4096);
    noalloc_clock_kernel<<<1, 1, 0, s>>>(clk, dt);
    vg::aos_to_soa_kernel<NF><<<blocks, 256, 0, s>>>(with_boundary, a, plane, gcells);
    vg::SweepArgs<P> args;
    args.in = a;
    args.out = b;
    args.plane = plane;
    args.ny = gy;
    args.nx = gx;
    args.bc = VG_BC_TRANSMISSIVE;
    args.limiter = limiter;
    args.dx = dx;
    args.bias = bias;
    args.dt_scale = 1.0;
    args.dim = 2;
    args.fuse_max = 0;
    args.m = m;
    args.clk = clk;
    args.partials = nullptr;
    args.fixed_params = 1;
    if (dir == 0) {
        dim3 grid((gx + (BSX - 2 * G) - 1) / (BSX - 2 * G), (gy + RX - 1) / RX);
        vg::sweep_x_kernel<P, ORDER, MODE, BSX, RX><<<grid, BSX, 0, s>>>(args);
    } else {
        dim3 grid((gx + BSY - 1) / BSY, (gy + HY - 1) / HY);
        vg::sweep_y_kernel<P, ORDER, MODE, BSY, HY><<<grid, BSY, 0, s>>>(args);
    }
    const long long icells = (long long)ny * nx;
    interior_to_aos_kernel<NF><<<(int)std::min<long long>((icells + 255) / 256, 4096), 256, 0, s>>>(b, cells, plane, ny,
                                                                                                    nx, gx, G);
    return vg::launch_status("euler_time_step_noalloc");
}

}  // namespace

extern "C" int vg_euler_time_step_2d_noalloc(vg_ctx *ctx, double *cells, const double *cells_with_boundary, int ny,
                                             int nx, int dir, double dx, double dt, double bias, int limiter,
                                             int order, const vg_euler_params *mp, int mode)
{
    vg::clear_error();
    VG_CHECK(ctx != nullptr, "null context");
    int rc = vg::check_time_step_args(2, ny, nx, dir, dx, dt, order, mode, cells, cells_with_boundary, mp);
    if (rc) return rc;
    const int G = (order == 2) ? 2 : 1;
    const long long gcells = (long long)(ny + 2 * G) * (nx + 2 * G);
    VG_CHECK(gcells <= kMaxCells, "grid too large for the no-allocation entry point: (ny + 2G) * (nx + 2G) must be at "
                                  "most 1048576 cells");
    rc = vg::bind(ctx);
    if (rc) return rc;
    double *base = nullptr;
    rc = scratch_address(ctx->device, &base);
    if (rc) return rc;
    EulerPhys::Params m;
    m.gamma = mp->gamma;
    m.p_inf = mp->p_inf;
    m.cv = mp->cv;
    m.e_form = mp->e_form;
    m.t_ign = mp->t_ign;
    m.k_rate = mp->k_rate;
    m.gm1 = mp->gamma - 1;
    m.gpinf = mp->gamma * mp->p_inf;
    m.inv_gm1 = 1.0 / (mp->gamma - 1);
    hipStream_t s = ctx->stream;
    if (order == 2)
        rc = mode == VG_MODE_EXACT ? launch_noalloc<2, 0>(s, base, cells, cells_with_boundary, ny, nx, dir, dx, dt, bias,
                                                          limiter, m)
                                   : launch_noalloc<2, 1>(s, base, cells, cells_with_boundary, ny, nx, dir, dx, dt, bias,
                                                          limiter, m);
    else
        rc = mode == VG_MODE_EXACT ? launch_noalloc<1, 0>(s, base, cells, cells_with_boundary, ny, nx, dir, dx, dt, bias,
                                                          limiter, m)
                                   : launch_noalloc<1, 1>(s, base, cells, cells_with_boundary, ny, nx, dir, dx, dt, bias,
                                                          limiter, m);
    if (rc) return rc;
    // Synchronous, except under stream capture (a captured stream cannot wait).
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    VG_HIP(hipStreamIsCapturing(s, &cs));
    if (cs == hipStreamCaptureStatusNone) VG_HIP(hipStreamSynchronize(s));
    return VG_OK;
}
