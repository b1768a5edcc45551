// Auxiliary device work: the FP64 operator-norm power iteration
// (radon_normest, radon_tgv_primal_dual_2d.py 2D:576-588) and the
// feasibility mask (Make_function_feasible, 2D:476-486).
#include "common.h"

#include <algorithm>
#include <cmath>
#include <vector>

namespace gr {

namespace {

// Forward projection of one slice in FP64, thread per (bin, angle); the same
// hat-weight gather as the FP32 kernel, with four candidates so that any
// rounding of the float constants is tolerated.
__global__ void fwd64_kernel(const double* __restrict__ img, double* __restrict__ sino,
                             const FwdAngle* __restrict__ angs, int nx, int ny, int nd, int na)
{
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= nd * na) return;
    const int aa = gid / nd, i = gid - aa * nd;
    const FwdAngle A = angs[aa];
    const int len_step = A.cls == 0 ? nx : ny;
    const int len_cand = A.cls == 0 ? ny : nx;
    const double ms = A.mstep, mc = A.mcand, off = A.off;
    double acc = 0.0;
    for (int l = 0; l < len_step; ++l) {
        const double r = (double)l * ms + off - (double)i;
        const double q = -r / mc;
        const int b = (int)ceil(q - 1.0 / fabs(mc)) - 1;
        for (int k = 0; k < 4; ++k) {
            const int c = b + k;
            if (c < 0 || c >= len_cand) continue;
            const double wgt = 1.0 - fabs((double)c * mc + r);
            if (wgt <= 0.0) continue;
            const int x = A.cls == 0 ? l : c;
            const int y = A.cls == 0 ? c : l;
            acc += wgt * img[(size_t)y * nx + x];
        }
    }
    sino[(size_t)A.id * nd + i] = acc;
}

// Backprojection of one slice in FP64, thread per pixel (2D:436-462).
__global__ void adj64_kernel(const double* __restrict__ sino, double* __restrict__ img,
                             const float4* __restrict__ ofs, int nx, int ny, int nd, int na)
{
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    if (gid >= nx * ny) return;
    const int y = gid / nx, x = gid - y * nx;
    double acc = 0.0;
    for (int j = 0; j < na; ++j) {
        const float4 o = ofs[j];
        const double t = (double)x * o.x + (double)y * o.y + (double)o.z;
        if (t > -1.0 && t < (double)nd) {
            const double fl = floor(t);
            const double p = t - fl;
            const int i0 = (int)fl;
            if (i0 >= 0) acc += (1.0 - p) * sino[(size_t)j * nd + i0];
            if (i0 <= nd - 2) acc += p * sino[(size_t)j * nd + i0 + 1];
        }
    }
    img[gid] = acc;
}

// Single-block FP64 reduction: mode 0 sum(x), 1 sum(x*x), 2 sum(x*y).
__global__ __launch_bounds__(1024) void reduce64_kernel(const double* __restrict__ x,
                                                        const double* __restrict__ y, int n,
                                                        int mode, double* out)
{
    __shared__ double red[1024];
    double acc = 0.0;
    for (int i = threadIdx.x; i < n; i += 1024) {
        const double a = x[i];
        acc += mode == 0 ? a : (mode == 1 ? a * a : a * y[i]);
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 512; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = red[0];
}

// x /= (*n) (mode 0) or x /= sqrt(*n) (mode 1)
__global__ void scale64_kernel(double* x, int n, const double* nrm, int mode)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double d = mode == 0 ? *nrm : sqrt(*nrm);
    x[i] = x[i] / d;
}

__global__ void feasible_kernel(float* f, const float* r1, size_t per_slice, size_t total)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    if (r1[i % per_slice] <= 0.f) f[i] = 0.f;
}

__global__ void fill_kernel(float* x, size_t n, float v)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = v;
}

}  // namespace

int launch_normest(const Geometry& g, const double* x0_host, int iters, int mode, double tol,
                   double* out, hipStream_t st)
{
    const int npx = g.nx * g.ny, nsb = g.nd * g.na;
    double *x = nullptr, *s = nullptr, *y = nullptr, *red = nullptr;
    auto cleanup = [&]() {
        if (x) (void)hipFree(x);
        if (s) (void)hipFree(s);
        if (y) (void)hipFree(y);
        if (red) (void)hipFree(red);
    };
    if (hipMalloc(&x, sizeof(double) * npx) != hipSuccess ||
        hipMalloc(&y, sizeof(double) * npx) != hipSuccess ||
        hipMalloc(&s, sizeof(double) * nsb) != hipSuccess ||
        hipMalloc(&red, sizeof(double) * 2) != hipSuccess) {
        cleanup();
        set_error("normest: device allocation failed");
        return GR_ERR_OOM;
    }
    std::vector<double> h0(npx);
    if (x0_host) {
        std::copy(x0_host, x0_host + npx, h0.begin());
    } else {
        for (int i = 0; i < npx; ++i) h0[i] = 1.0 + 0.25 * std::sin(0.37 * i);
    }
    int rc = GR_OK;
    auto chk = [&](hipError_t e, const char* w) {
        if (e != hipSuccess && rc == GR_OK) rc = hip_status(e, w);
    };
    chk(hipMemcpyAsync(x, h0.data(), sizeof(double) * npx, hipMemcpyHostToDevice, st),
        "normest copy");
    const int tb = 256;
    const float4* ofs = reinterpret_cast<const float4*>(g.ofs_dev);
    double result = 0.0;
    if (mode == GR_NORMEST_REFERENCE) {
        for (int it = 0; it < iters && rc == GR_OK; ++it) {
            hipLaunchKernelGGL(reduce64_kernel, dim3(1), dim3(1024), 0, st, x, x, npx, 0, red);
            hipLaunchKernelGGL(scale64_kernel, dim3((npx + tb - 1) / tb), dim3(tb), 0, st, x, npx,
                               red, 0);
            hipLaunchKernelGGL(fwd64_kernel, dim3((nsb + tb - 1) / tb), dim3(tb), 0, st, x, s,
                               g.fwd_dev, g.nx, g.ny, g.nd, g.na);
            hipLaunchKernelGGL(adj64_kernel, dim3((npx + tb - 1) / tb), dim3(tb), 0, st, s, x, ofs,
                               g.nx, g.ny, g.nd, g.na);
            chk(hipGetLastError(), "normest launch");
        }
        double n = 0.0;
        chk(hipMemcpyAsync(&n, red, sizeof(double), hipMemcpyDeviceToHost, st), "normest read");
        chk(hipStreamSynchronize(st), "normest sync");
        result = std::sqrt(n);
    } else {
        double prev = 0.0;
        for (int it = 0; it < iters && rc == GR_OK; ++it) {
            hipLaunchKernelGGL(reduce64_kernel, dim3(1), dim3(1024), 0, st, x, x, npx, 1, red);
            hipLaunchKernelGGL(scale64_kernel, dim3((npx + tb - 1) / tb), dim3(tb), 0, st, x, npx,
                               red, 1);
            hipLaunchKernelGGL(fwd64_kernel, dim3((nsb + tb - 1) / tb), dim3(tb), 0, st, x, s,
                               g.fwd_dev, g.nx, g.ny, g.nd, g.na);
            hipLaunchKernelGGL(adj64_kernel, dim3((npx + tb - 1) / tb), dim3(tb), 0, st, s, y, ofs,
                               g.nx, g.ny, g.nd, g.na);
            hipLaunchKernelGGL(reduce64_kernel, dim3(1), dim3(1024), 0, st, x, y, npx, 2, red + 1);
            chk(hipGetLastError(), "normest launch");
            double lam = 0.0;
            chk(hipMemcpyAsync(&lam, red + 1, sizeof(double), hipMemcpyDeviceToHost, st),
                "normest read");
            chk(hipStreamSynchronize(st), "normest sync");
            std::swap(x, y);
            result = std::sqrt(std::max(lam, 0.0));
            if (it > 0 && std::fabs(lam - prev) <= tol * std::fabs(lam)) break;
            prev = lam;
        }
    }
    cleanup();
    if (rc != GR_OK) return rc;
    *out = result;
    return GR_OK;
}

int launch_feasible(const Geometry& g, float* f, int nslices, hipStream_t st)
{
    const size_t npx = (size_t)g.nx * g.ny, per = (size_t)g.nd * g.na;
    float *ones = nullptr, *r1 = nullptr;
    if (hipMalloc(&ones, npx * 4) != hipSuccess || hipMalloc(&r1, per * 4) != hipSuccess) {
        if (ones) (void)hipFree(ones);
        set_error("make_feasible: device allocation failed");
        return GR_ERR_OOM;
    }
    const int tb = 256;
    hipLaunchKernelGGL(fill_kernel, dim3((unsigned)((npx + tb - 1) / tb)), dim3(tb), 0, st, ones,
                       npx, 1.0f);
    ProbDev P{};
    P.dim = 2; P.nx = g.nx; P.ny = g.ny; P.nz = 1; P.nch = 1; P.nd = g.nd; P.na = g.na;
    P.ofs = reinterpret_cast<const float4*>(g.ofs_dev);
    P.fwd = g.fwd_dev;
    P.fwd_band = g.fwd_band;
    P.ubar = ones;
    P.ku = r1;
    int rc = launch_fwd(P, g, 1, FWD_WRITE, st);
    if (rc == GR_OK) {
        const size_t total = per * (size_t)nslices;
        hipLaunchKernelGGL(feasible_kernel, dim3((unsigned)((total + tb - 1) / tb)), dim3(tb), 0,
                           st, f, r1, per, total);
        hipError_t e = hipGetLastError();
        if (e == hipSuccess) e = hipStreamSynchronize(st);
        if (e != hipSuccess) rc = hip_status(e, "make_feasible");
    }
    (void)hipFree(ones);
    (void)hipFree(r1);
    return rc;
}

}  // namespace gr
