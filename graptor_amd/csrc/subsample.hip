// Pieces of the two-resolution solver (radon_tgv_primal_dual_3dsubsample.py
// = 3DS:): channels measured on a coarser detector / slice grid are modelled
// as a box-average interpolation S of the fine-grid projection, applied per
// angle over the (bin, slice) plane; its adjoint S^T returns the coarse
// sinogram dual to the fine grid.  Also the sinogram prox as a standalone
// pass and an FP64 sum of squares (power iteration of the joint operator).
#include "common.h"

#include <algorithm>

namespace gr {

namespace {

// out[c][z][a][x] = sum_j val[j] * in[c][z_j][a][x_j] over the CSR row
// x + nx_out * z, whose column k = x_j + nx_in * z_j (3DS:403-428).
__global__ void csr_apply_kernel(const int* __restrict__ rowstart, const int* __restrict__ index,
                                 const float* __restrict__ val, int nx_out, int nz_out, int nx_in,
                                 int nz_in, int na, int nch, const float* __restrict__ in,
                                 float* __restrict__ out)
{
    const size_t total = (size_t)nx_out * na * nz_out * nch;
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const int x = (int)(i % nx_out);
    size_t q = i / nx_out;
    const int a = (int)(q % na);
    q /= na;
    const int z = (int)(q % nz_out);
    const int c = (int)(q / nz_out);
    const int row = x + nx_out * z;
    const size_t chan_in = (size_t)c * nz_in;
    float sum = 0.f;
    for (int j = rowstart[row]; j < rowstart[row + 1]; ++j) {
        const int k = index[j];
        const int zi = k / nx_in, xi = k - zi * nx_in;
        sum += val[j] * in[((chan_in + zi) * na + a) * nx_in + xi];
    }
    out[i] = sum;
}

// lam <- prox(lam + sigma_p ku) per bin (KL / L2 closed forms, 3DS:229-256)
__global__ void prox_kernel(int disc, float sigma_p, const float* __restrict__ mfac,
                            size_t per_channel, size_t total, const float* __restrict__ ku,
                            const float* __restrict__ f, float* __restrict__ lam)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= total) return;
    const float m = mfac[i / per_channel];
    if (disc == GR_DISC_KL) {
        const float sv = lam[i] + sigma_p * ku[i];
        const float z = sv - m;
        const float d = z * z + 4.f * sigma_p * f[i] * m;
        lam[i] = sv - 0.5f * (z + sqrtf(d));
    } else {
        lam[i] = (lam[i] + sigma_p * (ku[i] - f[i])) * m;
    }
}

constexpr int SQ_THREADS = 256;

__global__ __launch_bounds__(SQ_THREADS) void sumsq_partial_kernel(const float* __restrict__ x,
                                                                  size_t n,
                                                                  double* __restrict__ part)
{
    __shared__ double red[SQ_THREADS];
    double acc = 0.0;
    for (size_t i = (size_t)blockIdx.x * SQ_THREADS + threadIdx.x; i < n;
         i += (size_t)gridDim.x * SQ_THREADS) {
        const double v = x[i];
        acc += v * v;
    }
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = SQ_THREADS / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(SQ_THREADS) void sum_partials_kernel(const double* __restrict__ part,
                                                                 int n, double* __restrict__ out)
{
    __shared__ double red[SQ_THREADS];
    double acc = 0.0;
    for (int i = threadIdx.x; i < n; i += SQ_THREADS) acc += part[i];
    red[threadIdx.x] = acc;
    __syncthreads();
    for (int w = SQ_THREADS / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) *out = red[0];
}

}  // namespace

}  // namespace gr

using namespace gr;

static hipStream_t sub_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

extern "C" {

int gr_csr_apply(const int* rowstart, const int* index, const float* val, int nx_out, int nz_out,
                 int nx_in, int nz_in, int na, int nch, const float* in, float* out, void* stream)
{
    if (!rowstart || !index || !val || !in || !out || nx_out <= 0 || nz_out <= 0 || nx_in <= 0 ||
        nz_in <= 0 || na <= 0 || nch <= 0) {
        set_error("csr_apply: invalid arguments");
        return GR_ERR_INVALID;
    }
    const size_t total = (size_t)nx_out * na * nz_out * nch;
    hipLaunchKernelGGL(csr_apply_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0,
                       sub_stream(stream), rowstart, index, val, nx_out, nz_out, nx_in, nz_in, na,
                       nch, in, out);
    GR_LAUNCH_CHECK("csr_apply_kernel");
    return GR_OK;
}

int gr_prox(int discrepancy, float sigma_p, const float* mfac, int nch, size_t per_channel,
            const float* ku, const float* f, float* lam, void* stream)
{
    if (!mfac || !ku || !f || !lam || nch <= 0 || per_channel == 0 ||
        (discrepancy != GR_DISC_KL && discrepancy != GR_DISC_L2)) {
        set_error("prox: invalid arguments");
        return GR_ERR_INVALID;
    }
    const size_t total = per_channel * (size_t)nch;
    hipLaunchKernelGGL(prox_kernel, dim3((unsigned)((total + 255) / 256)), dim3(256), 0,
                       sub_stream(stream), discrepancy, sigma_p, mfac, per_channel, total, ku, f,
                       lam);
    GR_LAUNCH_CHECK("prox_kernel");
    return GR_OK;
}

int gr_sumsq_f64(const float* x, size_t n, double* out, void* stream)
{
    if (!x || !out || n == 0) {
        set_error("sumsq_f64: invalid arguments");
        return GR_ERR_INVALID;
    }
    const hipStream_t st = sub_stream(stream);
    const int nb = (int)std::min<size_t>(1024, (n + SQ_THREADS - 1) / SQ_THREADS);
    double* part = nullptr;
    GR_HIP(hipMalloc(&part, (nb + 1) * sizeof(double)));
    hipLaunchKernelGGL(sumsq_partial_kernel, dim3(nb), dim3(SQ_THREADS), 0, st, x, n, part);
    hipLaunchKernelGGL(sum_partials_kernel, dim3(1), dim3(SQ_THREADS), 0, st, part, nb, part + nb);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(out, part + nb, sizeof(double), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    (void)hipFree(part);
    if (e != hipSuccess) return hip_status(e, "sumsq_f64");
    return GR_OK;
}

}  // extern "C"
