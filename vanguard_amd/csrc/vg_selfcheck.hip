// Self-check of the EXACT division sequences (vg_quot.h) against the
// compiler's IEEE division, on the device.  No counterpart in the reference:
// the check behind the claim that the shortened sequences round like the IEEE
// division wherever the solvers use them.
#include <cstdint>

#include "vg_quot.h"
#include "vg_runtime.h"

namespace {

__global__ __launch_bounds__(256) void division_check_kernel(const double *a, const double *b, long long n,
                                                             double *shared_q, double *in_range_q, double *ieee_q)
{
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double x = a[i], y = b[i];
    const vg::SharedDivisor d = vg::shared_divisor(y);
    shared_q[i] = vg::quot(x, d);
    in_range_q[i] = vg::div_in_range(x, y);
    ieee_q[i] = x / y;
}

}  // namespace

extern "C" int vg_division_selfcheck(vg_ctx *ctx, int64_t n, const double *a, const double *b, double *shared_q,
                                     double *in_range_q, double *ieee_q)
{
    vg::clear_error();
    VG_CHECK(ctx != nullptr && a && b && shared_q && in_range_q && ieee_q, "null argument");
    VG_CHECK(n >= 1 && n <= (int64_t(1) << 26), "n must be in 1 .. 2^26");
    int rc = vg::bind(ctx);
    if (rc) return rc;
    const size_t bytes = sizeof(double) * (size_t)n;
    double *buf = nullptr;
    VG_HIP(hipMalloc(&buf, 5 * bytes));
    double *da = buf, *db = buf + n, *d0 = buf + 2 * n, *d1 = buf + 3 * n, *d2 = buf + 4 * n;
    hipStream_t s = ctx->stream;
    hipError_t e = hipMemcpyAsync(da, a, bytes, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) e = hipMemcpyAsync(db, b, bytes, hipMemcpyHostToDevice, s);
    if (e == hipSuccess) {
        division_check_kernel<<<(unsigned)((n + 255) / 256), 256, 0, s>>>(da, db, n, d0, d1, d2);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipMemcpyAsync(shared_q, d0, bytes, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipMemcpyAsync(in_range_q, d1, bytes, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipMemcpyAsync(ieee_q, d2, bytes, hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    hipFree(buf);
    if (e != hipSuccess) {
        vg::set_error(std::string("vg_division_selfcheck: ") + hipGetErrorString(e));
        return VG_ERR_HIP;
    }
    return VG_OK;
}
