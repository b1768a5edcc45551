// Parallel-beam projector pair for gfx950: forward projection (ray gather
// through LDS bands, fused with the KL/L2 sinogram prox) and backprojection
// (pixel tiles gathering from LDS-staged sinogram segments), plus the FP64
// operator-norm power iteration and the feasibility mask.
//
// Discretisation (radon_tgv_primal_dual_2d.py 2D:391-462): with detector
// coordinate t(x, y) = x*cos + y*sin + off, bin i of the forward projection
// sums hat(t - i) * img over pixels, hat(d) = max(0, 1 - |d|), and the
// backprojection interpolates the sinogram linearly at t.  The two are exact
// transposes of each other.
#include "common.h"

#include <algorithm>
#include <cmath>
#include <vector>

namespace gr {

namespace {

constexpr int FWD_THREADS = 256;
constexpr int FWD_MAX_ITEMS = 8;     // (angle, bin) items per thread
constexpr int FWD_PADL = 4;          // zero candidates left of position 0
constexpr int ADJ_THREADS = 256;
constexpr int ADJ_TX = 64;           // tile width = one wavefront of columns
constexpr int ADJ_MAXR = 8;          // rows per thread (tile_y <= 32)

// Map hardware block order (dealt round-robin over the 8 XCDs) to a work
// order in which consecutive work ids run on the same XCD and share its L2.
__device__ __forceinline__ int xcd_work_id(int bid, int nb)
{
    const int xcd = bid & 7, local = bid >> 3;
    const int per = nb >> 3, rem = nb & 7;
    return (xcd < rem) ? xcd * (per + 1) + local
                       : rem * (per + 1) + (xcd - rem) * per + local;
}

template <int K>
struct Pack {
    float v[K];
};

template <int K>
__device__ __forceinline__ Pack<K> lds_load(const float* sh, int pos)
{
    Pack<K> r;
    if constexpr (K == 4) {
        const float4 t = *reinterpret_cast<const float4*>(sh + 4 * pos);
        r.v[0] = t.x; r.v[1] = t.y; r.v[2] = t.z; r.v[3] = t.w;
    } else if constexpr (K == 2) {
        const float2 t = *reinterpret_cast<const float2*>(sh + 2 * pos);
        r.v[0] = t.x; r.v[1] = t.y;
    } else {
        r.v[0] = sh[pos];
    }
    return r;
}

template <int K>
__device__ __forceinline__ void lds_store(float* sh, int pos, const float (&v)[K])
{
    if constexpr (K == 4) {
        *reinterpret_cast<float4*>(sh + 4 * pos) = make_float4(v[0], v[1], v[2], v[3]);
    } else if constexpr (K == 2) {
        *reinterpret_cast<float2*>(sh + 2 * pos) = make_float2(v[0], v[1]);
    } else {
        sh[pos] = v[0];
    }
}

// --------------------------------------------------------------------------
// Forward projection.  A work-group owns K slices (sharing all geometry)
// and one chunk of (angle, bin) items whose angles all step along the same
// axis.  It streams the K slices through LDS in bands of `band` lines of
// that axis; per line every item reads the <= 3 candidates on the other
// axis whose hat weight can be positive.  After the last band each item
// holds its complete ray sum and applies the sinogram epilogue.
// --------------------------------------------------------------------------
template <int K>
__global__ __launch_bounds__(FWD_THREADS) void radon_fwd_kernel(
    const float* __restrict__ img, float* __restrict__ sino,
    float* __restrict__ lam, const float* __restrict__ f,
    const float* __restrict__ mfac, const FwdAngle* __restrict__ angs,
    const FwdChunk* __restrict__ chunks, int nchunks, int nx, int ny, int nd,
    int na, int nslices, int nz_per_ch, int band, int pitch, int mode, float sigp)
{
    extern __shared__ float sh[];
    const int work = xcd_work_id(blockIdx.x, gridDim.x);
    const int chunk_id = work % nchunks;
    const int s0 = (work / nchunks) * K;
    const FwdChunk ch = chunks[chunk_id];
    const int tid = threadIdx.x;
    const int len_step = ch.cls == 0 ? nx : ny;
    const int len_cand = ch.cls == 0 ? ny : nx;
    const int nitems = ch.na * ch.nb;

    float acc[FWD_MAX_ITEMS][K];
    float fi[FWD_MAX_ITEMS], ms[FWD_MAX_ITEMS], mc[FWD_MAX_ITEMS];
    float of[FWD_MAX_ITEMS], iv[FWD_MAX_ITEMS], hh[FWD_MAX_ITEMS];
#pragma unroll
    for (int it = 0; it < FWD_MAX_ITEMS; ++it) {
        const int item = tid + it * FWD_THREADS;
        const int a = item < nitems ? item / ch.nb : 0;
        const FwdAngle A = angs[ch.a0 + a];
        fi[it] = (float)(ch.b0 + (item < nitems ? item - a * ch.nb : 0));
        ms[it] = A.mstep; mc[it] = A.mcand; of[it] = A.off;
        iv[it] = A.invc;  hh[it] = A.hc;
#pragma unroll
        for (int k = 0; k < K; ++k) acc[it][k] = 0.f;
    }

    const int lds_floats = band * pitch * K;
    for (int e = tid; e < lds_floats; e += FWD_THREADS) sh[e] = 0.f;

    const float cand_hi = (float)len_cand;
    const size_t slice_stride = (size_t)nx * ny;
    for (int l0 = 0; l0 < len_step; l0 += band) {
        __syncthreads();
        // stage lines [l0, l0+band) of K slices, candidates contiguous
        if (ch.cls == 1) {  // lines are rows, candidates are x
            for (int e = tid; e < band * len_cand; e += FWD_THREADS) {
                const int li = e / len_cand, x = e - li * len_cand;
                const int y = l0 + li;
                float v[K];
#pragma unroll
                for (int k = 0; k < K; ++k) {
                    const int s = s0 + k;
                    v[k] = (y < ny && s < nslices)
                               ? img[(size_t)s * slice_stride + (size_t)y * nx + x]
                               : 0.f;
                }
                lds_store<K>(sh, li * pitch + FWD_PADL + x, v);
            }
        } else {  // lines are columns, candidates are y
            for (int e = tid; e < band * len_cand; e += FWD_THREADS) {
                const int y = e / band, li = e - y * band;
                const int x = l0 + li;
                float v[K];
#pragma unroll
                for (int k = 0; k < K; ++k) {
                    const int s = s0 + k;
                    v[k] = (x < nx && s < nslices)
                               ? img[(size_t)s * slice_stride + (size_t)y * nx + x]
                               : 0.f;
                }
                lds_store<K>(sh, li * pitch + FWD_PADL + y, v);
            }
        }
        __syncthreads();

#pragma unroll
        for (int it = 0; it < FWD_MAX_ITEMS; ++it) {
            if (tid + it * FWD_THREADS >= nitems) continue;
            for (int st = 0; st < band; ++st) {
                const float xs = (float)(l0 + st);
                const float r = fmaf(xs, ms[it], of[it]) - fi[it];
                float bf = ceilf(fmaf(-r, iv[it], -hh[it]));
                bf = fminf(fmaxf(bf, -3.f), cand_hi);
                const float d0 = fmaf(bf, mc[it], r);
                const float d1 = d0 + mc[it];
                const float d2 = d1 + mc[it];
                const float w0 = fmaxf(0.f, 1.f - fabsf(d0));
                const float w1 = fmaxf(0.f, 1.f - fabsf(d1));
                const float w2 = fmaxf(0.f, 1.f - fabsf(d2));
                const int pos = st * pitch + FWD_PADL + (int)bf;
                const Pack<K> v0 = lds_load<K>(sh, pos);
                const Pack<K> v1 = lds_load<K>(sh, pos + 1);
                const Pack<K> v2 = lds_load<K>(sh, pos + 2);
#pragma unroll
                for (int k = 0; k < K; ++k) {
                    float a = fmaf(w0, v0.v[k], acc[it][k]);
                    a = fmaf(w1, v1.v[k], a);
                    acc[it][k] = fmaf(w2, v2.v[k], a);
                }
            }
        }
    }

    // epilogue: store and/or apply the sinogram prox (2D:316-338)
#pragma unroll
    for (int it = 0; it < FWD_MAX_ITEMS; ++it) {
        const int item = tid + it * FWD_THREADS;
        if (item >= nitems) continue;
        const int a = item / ch.nb;
        const int i = ch.b0 + item - a * ch.nb;
        const int id = angs[ch.a0 + a].id;
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const int s = s0 + k;
            if (s >= nslices) continue;
            const size_t idx = ((size_t)s * na + id) * nd + i;
            const float val = acc[it][k];
            if (mode & FWD_WRITE) sino[idx] = val;
            if (mode & FWD_KL) {
                const float m = mfac[s / nz_per_ch];
                const float sv = lam[idx] + sigp * val;
                const float z = sv - m;
                const float d = z * z + 4.f * sigp * f[idx] * m;
                lam[idx] = sv - 0.5f * (z + sqrtf(d));
            } else if (mode & FWD_L2) {
                const float m = mfac[s / nz_per_ch];
                lam[idx] = (lam[idx] + sigp * (val - f[idx])) * m;
            }
        }
    }
}

// --------------------------------------------------------------------------
// Backprojection.  A work-group owns a 64 x tile_y pixel tile of K slices.
// Per chunk of angles it stages, for each angle, only the sinogram segment
// the tile projects onto (bins outside [0, nd) staged as zeros, which
// reproduces the reference's edge guards), then every thread interpolates
// for its tile_y/4 pixels.
// --------------------------------------------------------------------------
template <int K>
__global__ __launch_bounds__(ADJ_THREADS) void radon_adj_kernel(
    const float* __restrict__ sino, float* __restrict__ img,
    const float4* __restrict__ ofs, int nx, int ny, int nd, int na, int nslices,
    int tile_y, int segw, int achunk, int tiles_x, int ntiles)
{
    extern __shared__ float sh[];
    float* seg = sh;                                  // achunk * segw * K
    float4* par = reinterpret_cast<float4*>(sh + (size_t)achunk * segw * K);
    int* lo_s = reinterpret_cast<int*>(par + achunk);

    const int work = xcd_work_id(blockIdx.x, gridDim.x);
    const int tile = work % ntiles;
    const int s0 = (work / ntiles) * K;
    const int x0 = (tile % tiles_x) * ADJ_TX;
    const int y0 = (tile / tiles_x) * tile_y;
    const int tid = threadIdx.x;
    const int lane = tid & 63, wv = tid >> 6;
    const int R = tile_y >> 2;
    const float xf = (float)(x0 + lane);

    float acc[ADJ_MAXR][K];
#pragma unroll
    for (int r = 0; r < ADJ_MAXR; ++r)
#pragma unroll
        for (int k = 0; k < K; ++k) acc[r][k] = 0.f;

    const size_t sino_slice = (size_t)na * nd;
    for (int a0 = 0; a0 < na; a0 += achunk) {
        const int A = min(achunk, na - a0);
        __syncthreads();
        if (tid < A) {
            const float4 o = ofs[a0 + tid];
            const float tmin = o.z + (float)x0 * o.x + (float)y0 * o.y +
                               fminf(0.f, (float)(ADJ_TX - 1) * o.x) +
                               fminf(0.f, (float)(tile_y - 1) * o.y);
            const int lo = (int)floorf(tmin) - 1;
            par[tid] = make_float4(o.x, o.y, o.z - (float)lo, 0.f);
            lo_s[tid] = lo;
        }
        __syncthreads();
        for (int e = tid; e < A * segw; e += ADJ_THREADS) {
            const int al = e / segw, b = e - al * segw;
            const int gb = lo_s[al] + b;
            float v[K];
#pragma unroll
            for (int k = 0; k < K; ++k) {
                const int s = s0 + k;
                v[k] = (gb >= 0 && gb < nd && s < nslices)
                           ? sino[(size_t)s * sino_slice + (size_t)(a0 + al) * nd + gb]
                           : 0.f;
            }
            lds_store<K>(seg, al * segw + b, v);
        }
        __syncthreads();
        for (int al = 0; al < A; ++al) {
            const float4 pa = par[al];
            const float tb = fmaf(xf, pa.x, pa.z);
            const int base = al * segw;
#pragma unroll
            for (int r = 0; r < ADJ_MAXR; ++r) {
                if (r >= R) break;
                const float t = fmaf((float)(y0 + wv + 4 * r), pa.y, tb);
                const float fl = floorf(t);
                const float p = t - fl;
                const float q = 1.f - p;
                const int i0 = base + (int)fl;
                const Pack<K> v0 = lds_load<K>(seg, i0);
                const Pack<K> v1 = lds_load<K>(seg, i0 + 1);
#pragma unroll
                for (int k = 0; k < K; ++k)
                    acc[r][k] = fmaf(p, v1.v[k], fmaf(q, v0.v[k], acc[r][k]));
            }
        }
    }

    const int x = x0 + lane;
    if (x >= nx) return;
#pragma unroll
    for (int r = 0; r < ADJ_MAXR; ++r) {
        if (r >= R) break;
        const int y = y0 + wv + 4 * r;
        if (y >= ny) continue;
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const int s = s0 + k;
            if (s < nslices) img[(size_t)s * nx * ny + (size_t)y * nx + x] = acc[r][k];
        }
    }
}

// --------------------------------------------------------------------------
// FP64 kernels for the operator-norm power iteration (one slice).
// --------------------------------------------------------------------------
__global__ void fwd64_kernel(const double* __restrict__ img, double* __restrict__ sino,
                             const FwdAngle* __restrict__ angs, int nx, int ny, int nd,
                             int na)
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
        // four candidates tolerate any rounding of the float constants
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
                                                        const double* __restrict__ y,
                                                        int n, int mode, double* out)
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

__global__ void feasible_kernel(float* f, const float* r1, size_t per_slice,
                                size_t total)
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

// --------------------------------------------------------------------------
// host launchers
// --------------------------------------------------------------------------
static int fwd_pitch(const Geometry& g)
{
    int p = std::max(g.nx, g.ny) + FWD_PADL + 4;
    if ((p & 1) == 0) ++p;  // odd pitch: conflict-free transposed stores
    return p;
}

template <int K>
static int fwd_launch_k(const Geometry& g, const float* img, float* sino, float* lam,
                        const float* f, const float* mfac, int nslices, int nz_per_ch,
                        int mode, float sigp, hipStream_t st)
{
    const int pitch = fwd_pitch(g);
    const int groups = (nslices + K - 1) / K;
    // chunk table: fewest items per thread that still gives >= 1024 groups
    int t = 0;
    while (t < 3 && (long long)groups * (g.tab_off[t + 1] - g.tab_off[t]) < 1024) ++t;
    const int nchunks = g.tab_off[t + 1] - g.tab_off[t];
    const FwdChunk* chunks = g.chunks_dev + g.tab_off[t];
    int band = g.fwd_band;
    if (band <= 0) {  // largest multiple of 4 keeping the LDS stage near 50 KB
        band = (int)(50 * 1024 / ((size_t)pitch * K * sizeof(float)));
        band = std::max(4, std::min(32, band & ~3));
    }
    const size_t lds = (size_t)band * pitch * K * sizeof(float);
    if (band <= 0 || lds > 160 * 1024) {
        set_error("radon: band %d does not fit LDS (%zu bytes)", band, lds);
        return GR_ERR_INVALID;
    }
    const long long blocks = (long long)groups * nchunks;
    if (blocks <= 0 || blocks > 0x7fffffff) {
        set_error("radon: invalid grid (%lld blocks)", blocks);
        return GR_ERR_INVALID;
    }
    hipLaunchKernelGGL(radon_fwd_kernel<K>, dim3((unsigned)blocks), dim3(FWD_THREADS), lds,
                       st, img, sino, lam, f, mfac, g.fwd_dev, chunks, nchunks, g.nx, g.ny,
                       g.nd, g.na, nslices, nz_per_ch, band, pitch, mode, sigp);
    GR_LAUNCH_CHECK("radon_fwd_kernel");
    return GR_OK;
}

int launch_radon_fwd(const Geometry& g, const float* img, float* sino, float* lam,
                     const float* f, const float* mfac, int nslices, int nz_per_ch,
                     int mode, float sigp, hipStream_t st)
{
    if (nslices <= 0) return GR_OK;
    if (nslices >= 4)
        return fwd_launch_k<4>(g, img, sino, lam, f, mfac, nslices, nz_per_ch, mode, sigp, st);
    if (nslices >= 2)
        return fwd_launch_k<2>(g, img, sino, lam, f, mfac, nslices, nz_per_ch, mode, sigp, st);
    return fwd_launch_k<1>(g, img, sino, lam, f, mfac, nslices, nz_per_ch, mode, sigp, st);
}

static int adj_segw(int tile_y)
{
    const double diag = std::sqrt((double)(ADJ_TX - 1) * (ADJ_TX - 1) +
                                  (double)(tile_y - 1) * (tile_y - 1));
    int w = (int)std::ceil(diag) + 6;
    return (w + 3) & ~3;
}

template <int K>
static int adj_launch_k(const Geometry& g, const float* sino, float* img, int nslices,
                        hipStream_t st)
{
    const int tile_y = g.adj_tile_y;
    const int segw = adj_segw(tile_y);
    const size_t per_angle = (size_t)segw * K * sizeof(float) + sizeof(float4) + sizeof(int);
    int achunk = (int)std::min<size_t>((size_t)g.na, (76 * 1024) / per_angle);
    achunk = std::max(achunk, 1);
    const size_t lds = (size_t)achunk * per_angle;
    const int tiles_x = (g.nx + ADJ_TX - 1) / ADJ_TX;
    const int tiles_y = (g.ny + tile_y - 1) / tile_y;
    const int ntiles = tiles_x * tiles_y;
    const int groups = (nslices + K - 1) / K;
    const long long blocks = (long long)groups * ntiles;
    if (blocks > 0x7fffffff) {
        set_error("radon_ad: too many slices");
        return GR_ERR_INVALID;
    }
    hipLaunchKernelGGL(radon_adj_kernel<K>, dim3((unsigned)blocks), dim3(ADJ_THREADS), lds, st,
                       sino, img, reinterpret_cast<const float4*>(g.ofs_dev), g.nx, g.ny,
                       g.nd, g.na, nslices, tile_y, segw, achunk, tiles_x, ntiles);
    GR_LAUNCH_CHECK("radon_adj_kernel");
    return GR_OK;
}

int launch_radon_adj(const Geometry& g, const float* sino, float* img, int nslices,
                     hipStream_t st)
{
    if (nslices <= 0) return GR_OK;
    if (g.adj_tile_y < 4 || g.adj_tile_y > 4 * ADJ_MAXR || (g.adj_tile_y & 3)) {
        set_error("radon_ad: tile_y %d must be a multiple of 4 in [4, %d]", g.adj_tile_y,
                  4 * ADJ_MAXR);
        return GR_ERR_INVALID;
    }
    if (nslices >= 4) return adj_launch_k<4>(g, sino, img, nslices, st);
    if (nslices >= 2) return adj_launch_k<2>(g, sino, img, nslices, st);
    return adj_launch_k<1>(g, sino, img, nslices, st);
}

int launch_normest(const Geometry& g, const double* x0_host, int iters, int mode,
                   double tol, double* out, hipStream_t st)
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
    double result = 0.0;
    if (mode == GR_NORMEST_REFERENCE) {
        for (int it = 0; it < iters && rc == GR_OK; ++it) {
            hipLaunchKernelGGL(reduce64_kernel, dim3(1), dim3(1024), 0, st, x, x, npx, 0, red);
            hipLaunchKernelGGL(scale64_kernel, dim3((npx + tb - 1) / tb), dim3(tb), 0, st, x,
                               npx, red, 0);
            hipLaunchKernelGGL(fwd64_kernel, dim3((nsb + tb - 1) / tb), dim3(tb), 0, st, x, s,
                               g.fwd_dev, g.nx, g.ny, g.nd, g.na);
            hipLaunchKernelGGL(adj64_kernel, dim3((npx + tb - 1) / tb), dim3(tb), 0, st, s, x,
                               reinterpret_cast<const float4*>(g.ofs_dev), g.nx, g.ny, g.nd,
                               g.na);
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
            hipLaunchKernelGGL(scale64_kernel, dim3((npx + tb - 1) / tb), dim3(tb), 0, st, x,
                               npx, red, 1);
            hipLaunchKernelGGL(fwd64_kernel, dim3((nsb + tb - 1) / tb), dim3(tb), 0, st, x, s,
                               g.fwd_dev, g.nx, g.ny, g.nd, g.na);
            hipLaunchKernelGGL(adj64_kernel, dim3((npx + tb - 1) / tb), dim3(tb), 0, st, s, y,
                               reinterpret_cast<const float4*>(g.ofs_dev), g.nx, g.ny, g.nd,
                               g.na);
            hipLaunchKernelGGL(reduce64_kernel, dim3(1), dim3(1024), 0, st, x, y, npx, 2,
                               red + 1);
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
    int rc = launch_radon_fwd(g, ones, r1, nullptr, nullptr, nullptr, 1, 1, FWD_WRITE, 0.f, st);
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
