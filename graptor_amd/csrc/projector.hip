// Parallel-beam projector pair for gfx950.
//
// Forward projection: ray gather through LDS bands, fused with the KL/L2
// sinogram prox.  Backprojection: pixel tiles gathering from LDS-staged
// sinogram segments, optionally fused with the primal step of the
// primal-dual iteration (u, p updates and their extrapolations), so that
// K^T lam never goes through HBM.
//
// Discretisation (radon_tgv_primal_dual_2d.py 2D:391-462): with detector
// coordinate t(x, y) = x*cos + y*sin + off, bin i of the forward projection
// sums hat(t - i) * img over pixels, hat(d) = max(0, 1 - |d|), and the
// backprojection interpolates the sinogram linearly at t.  The two are exact
// transposes of each other.
#include "common.h"
#include "stencil.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>

namespace gr {

namespace {

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

// LDS entries of K slices at a position: K <= 4 interleaved (one read of
// 4K bytes); K = 8 as two planes of four, hs floats apart, so that a
// 16-byte read per plane keeps the bank pattern of four slices.
template <int K>
__device__ __forceinline__ Pack<K> lds_load(const float* sh, int pos, int hs = 0)
{
    Pack<K> r;
    if constexpr (K == 8) {
        const float4 t = *reinterpret_cast<const float4*>(sh + 4 * pos);
        const float4 u = *reinterpret_cast<const float4*>(sh + 4 * pos + hs);
        r.v[0] = t.x; r.v[1] = t.y; r.v[2] = t.z; r.v[3] = t.w;
        r.v[4] = u.x; r.v[5] = u.y; r.v[6] = u.z; r.v[7] = u.w;
    } else if constexpr (K == 4) {
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
__device__ __forceinline__ void lds_store(float* sh, int pos, const float (&v)[K], int hs = 0)
{
    if constexpr (K == 8) {
        *reinterpret_cast<float4*>(sh + 4 * pos) = make_float4(v[0], v[1], v[2], v[3]);
        *reinterpret_cast<float4*>(sh + 4 * pos + hs) = make_float4(v[4], v[5], v[6], v[7]);
    } else if constexpr (K == 4) {
        *reinterpret_cast<float4*>(sh + 4 * pos) = make_float4(v[0], v[1], v[2], v[3]);
    } else if constexpr (K == 2) {
        *reinterpret_cast<float2*>(sh + 2 * pos) = make_float2(v[0], v[1]);
    } else {
        sh[pos] = v[0];
    }
}

// --------------------------------------------------------------------------
// Forward projection of one work item: K slices s0.. (sharing all geometry)
// times a chunk of (angle, bin) items whose angles step along the same axis.
// The K slices stream through LDS in bands of `band` lines of that axis; per
// line every item reads the <= 3 candidates on the other axis whose hat
// weight can be positive.  After the last band each item holds its complete
// ray sum and applies the sinogram epilogue.
// --------------------------------------------------------------------------
template <int K, int MI>
__device__ __forceinline__ void fwd_body(const ProbDev& P, const FwdItem& I, int mode,
                                         float* sh)
{
    if (mode & FWD_PROX) mode |= P.disc == GR_DISC_KL ? FWD_KL : FWD_L2;
    const int tid = threadIdx.x;
    const int nx = P.nx, ny = P.ny, nd = P.nd, na = P.na;
    const int nslices = P.nz * P.nch;
    const int band = P.fwd_band, pitch = P.fwd_pitch;
    const int len_step = I.cls == 0 ? nx : ny;
    const int len_cand = I.cls == 0 ? ny : nx;
    const int nitems = I.na * I.nb;
    const int s0 = I.s0;
    const float* __restrict__ img = P.ubar;

    // Per item, along the ray at stepping coordinate xs: the offset of the
    // candidate-axis origin from bin i is r = xs * mstep + (off - i); the
    // first candidate whose hat weight can be positive is
    // ceil(-r / mcand - hc) = ceil(xs * ta + tb).
    float acc[MI][K];
    float ms[MI], mc[MI], ro[MI], ta[MI], tb[MI];
#pragma unroll
    for (int it = 0; it < MI; ++it) {
        const int item = tid + it * FWD_THREADS;
        const int a = item < nitems ? item / I.nb : 0;
        const FwdAngle A = P.fwd[I.a0 + a];
        const float fi = (float)(I.b0 + (item < nitems ? item - a * I.nb : 0));
        ms[it] = A.mstep;
        mc[it] = A.mcand;
        ro[it] = A.off - fi;
        ta[it] = -A.mstep * A.invc;
        tb[it] = fmaf(-ro[it], A.invc, -A.hc);
#pragma unroll
        for (int k = 0; k < K; ++k) acc[it][k] = 0.f;
    }

    const int lds_floats = band * pitch * K;
    for (int e = tid; e < lds_floats; e += FWD_THREADS) sh[e] = 0.f;

    const float cand_hi = (float)len_cand;
    const size_t slice_stride = (size_t)nx * ny;
    // staging walk: element e = tid + j * FWD_THREADS of a band, split into
    // (outer, inner) = (e / n_in, e % n_in) by stepping rather than dividing
    // (n_in = candidates per line for rows, band for columns)
    const int n_in = I.cls == 1 ? len_cand : band;
    const int n_out = I.cls == 1 ? band : len_cand;
    const int st_out0 = tid / n_in, st_in0 = tid - st_out0 * n_in;
    const int st_dout = FWD_THREADS / n_in, st_din = FWD_THREADS - st_dout * n_in;
    for (int l0 = 0; l0 < len_step; l0 += band) {
        __syncthreads();
        // stage lines [l0, l0+band) of K slices, candidates contiguous
        if (I.cls == 1) {  // lines are rows, candidates are x
            for (int li = st_out0, x = st_in0; li < n_out;) {
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
                x += st_din;
                li += st_dout;
                if (x >= n_in) {
                    x -= n_in;
                    ++li;
                }
            }
        } else {  // lines are columns, candidates are y
            for (int y = st_out0, li = st_in0; y < n_out;) {
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
                li += st_din;
                y += st_dout;
                if (li >= n_in) {
                    li -= n_in;
                    ++y;
                }
            }
        }
        __syncthreads();

#pragma unroll
        for (int it = 0; it < MI; ++it) {
            if (tid + it * FWD_THREADS >= nitems) continue;
            float xs = (float)l0;  // exact integer-valued step coordinate
            for (int st = 0; st < band; ++st, xs += 1.f) {
                const float r = fmaf(xs, ms[it], ro[it]);
                float bf = ceilf(fmaf(xs, ta[it], tb[it]));
                // clamp into [-3, len_cand] (bf is never NaN: finite operands)
                bf = __builtin_amdgcn_fmed3f(bf, -3.f, cand_hi);
                const float d0 = fmaf(bf, mc[it], r);
                const float d1 = d0 + mc[it];
                const float d2 = d1 + mc[it];
                // hat weights max(0, 1 - |d|) (1 - |d| <= 1, so a clamp to [0, 1])
                const float w0 = __builtin_amdgcn_fmed3f(1.f - fabsf(d0), 0.f, 1.f);
                const float w1 = __builtin_amdgcn_fmed3f(1.f - fabsf(d1), 0.f, 1.f);
                const float w2 = __builtin_amdgcn_fmed3f(1.f - fabsf(d2), 0.f, 1.f);
                // line base is uniform across the wave; the candidate offset is per lane
                const float* line = sh + K * (st * pitch + FWD_PADL);
                const int c0 = (int)bf;
                const Pack<K> v0 = lds_load<K>(line, c0);
                const Pack<K> v1 = lds_load<K>(line, c0 + 1);
                const Pack<K> v2 = lds_load<K>(line, c0 + 2);
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
    for (int it = 0; it < MI; ++it) {
        const int item = tid + it * FWD_THREADS;
        if (item >= nitems) continue;
        const int a = item / I.nb;
        const int i = I.b0 + item - a * I.nb;
        const int id = P.fwd[I.a0 + a].id;
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const int s = s0 + k;
            if (s >= nslices) continue;
            const size_t idx = ((size_t)s * na + id) * nd + i;
            const float val = acc[it][k];
            if (mode & FWD_WRITE) P.ku[idx] = val;
            if (mode & FWD_KL) {
                const float m = P.mfac[s / P.nz];
                const float sv = P.lam[idx] + P.sigma_p * val;
                const float z = sv - m;
                const float d = z * z + 4.f * P.sigma_p * P.f[idx] * m;
                P.lam[idx] = sv - 0.5f * (z + sqrtf(d));
            } else if (mode & FWD_L2) {
                const float m = P.mfac[s / P.nz];
                P.lam[idx] = (P.lam[idx] + P.sigma_p * (val - P.f[idx])) * m;
            }
        }
    }
}

template <int K, int MI>
__global__ __launch_bounds__(FWD_THREADS) void fwd_single_kernel(ProbDev P,
                                                                 const FwdChunk* __restrict__ chunks,
                                                                 int nchunks, int mode)
{
    extern __shared__ float sh[];
    const int work = xcd_work_id(blockIdx.x, gridDim.x);
    const FwdChunk ch = chunks[work % nchunks];
    const FwdItem I{0, ch.cls, ch.a0, ch.na, ch.b0, ch.nb, (work / nchunks) * K, 0};
    fwd_body<K, MI>(P, I, mode, sh);
}

template <int K, int MI>
__global__ __launch_bounds__(FWD_THREADS) void fwd_batch_kernel(const ProbDev* __restrict__ probs,
                                                                const FwdItem* __restrict__ items,
                                                                int mode)
{
    extern __shared__ float sh[];
    // items are sorted largest first; plain block order deals them
    // round-robin over the XCDs, which balances the load
    const FwdItem I = items[blockIdx.x];
    const ProbDev P = probs[I.prob];
    fwd_body<K, MI>(P, I, mode, sh);
}

// --------------------------------------------------------------------------
// Primal step at one voxel-channel (update_u 2D:340-357 / 3D:706-729,
// update_p 2D:359-376 / 3D:734-757, update_extra 2D:637-638):
//   u+ = max(0, u + tau (div v - ktl / L)),  ubar = 2 u+ - u
//   p+ = p + tau (v + div2 w),               pbar = 2 p+ - p
// i indexes u[c][z][y][x] (= slice s = z + nz c), z = s % nz.
// --------------------------------------------------------------------------
template <int DIM, bool TGV, int PART = PRIMAL_UP, bool ZH = false, bool NT = false>
__device__ __forceinline__ void primal_update_r(
    float* __restrict__ u, float* __restrict__ ubar, float* __restrict__ p,
    float* __restrict__ pbar, const float* __restrict__ v, const float* __restrict__ w, int nx,
    int ny, int nz, size_t cs, float hinv, float tau, float norming, size_t i, int x, int y, int z,
    float ktl, const ProbDev* zh)
{
    constexpr bool NU = (PART & PRIMAL_U) != 0, NP = (PART & PRIMAL_P) != 0;
    float un, ub, pn[3], pb[3];
    primal_point<DIM, TGV, NU, NP, true, ZH, NT>(u, p, v, w, ktl, nx, ny, nz, cs, hinv, tau, norming, i, x, y,
                                   z, un, ub, pn, pb, zh);
    // streamed once per iteration (NT: with the non-temporal hint)
    if constexpr (NU) {
        st_stream<NT>(un, &u[i]);
        st_stream<NT>(ub, &ubar[i]);
    }
    if constexpr (TGV && NP) {
#pragma unroll
        for (int k = 0; k < DIM; ++k) {
            st_stream<NT>(pn[k], &p[i + k * cs]);
            st_stream<NT>(pb[k], &pbar[i + k * cs]);
        }
    }
}

template <int DIM, bool TGV, int PART = PRIMAL_UP, bool ZH = false, bool NT = false>
__device__ __forceinline__ void primal_update(const ProbDev& P, size_t i, int x, int y, int z,
                                              float ktl)
{
    const size_t cs = (size_t)P.nx * P.ny * P.nz * P.nch;  // component stride
    primal_update_r<DIM, TGV, PART, ZH, NT>(P.u, P.ubar, P.p, P.pbar, P.v, P.w, P.nx, P.ny, P.nz,
                                            cs, P.hinv, P.tau, P.norming, i, x, y, z, ktl, &P);
}

// Primal step of one voxel-channel per thread with K^T lam read from memory
// (the unfused form: high occupancy for the memory-bound update).
template <int DIM, bool TGV, int PART = PRIMAL_UP, bool ZH = false, bool NT = false>
__device__ __forceinline__ void primal_body(const ProbDev& P, const DualItem& I)
{
    const int x = I.x0 + (int)(threadIdx.x & (DUAL_BX - 1));
    const int y = I.y0 + (int)(threadIdx.x / DUAL_BX);
    const int s = I.z;  // slice z + nz * c
    if (x >= P.nx || y >= P.ny) return;
    const size_t i = (size_t)s * P.nx * P.ny + (size_t)y * P.nx + x;
    const float ktl = (PART & PRIMAL_U) ? ld_stream<NT>(&P.ktl[i]) : 0.f;
    primal_update<DIM, TGV, PART, ZH, NT>(P, i, x, y, s % P.nz, ktl);
}

// PART selects the u half (needs K^T lam), the p half (needs only v and w,
// so it can run beside the projections) or both.
template <int DIM, bool TGV, int PART, bool ZH, bool NT>
__global__ __launch_bounds__(DUAL_BX* DUAL_BY) void primal_single_kernel(ProbDev P, int tiles_x,
                                                                         int tiles_y)
{
    const int b = stencil_block<PRIMAL_GROUP>();
    const int tx = b % tiles_x, rest = b / tiles_x;
    const DualItem I{0, tx * DUAL_BX, (rest % tiles_y) * DUAL_BY, rest / tiles_y};
    primal_body<DIM, TGV, PART, ZH, NT>(P, I);
}

// The u half with all channels of a voxel in one thread (a quarter of the
// work-groups of the per-channel form for 4 channels, each thread with as
// many independent load streams): 0.779 against 0.786 ms per two-stream
// iteration on 256x256x64x4 on one box, 0.791 against 0.798 ms on another,
// same bits.  The p half stays per channel (0.810 ms in this form).
template <int DIM, bool TGV, bool NT>
__global__ __launch_bounds__(DUAL_BX* DUAL_BY) void primal_u_voxel_kernel(ProbDev P, int tiles_x,
                                                                          int tiles_y)
{
    const int b = stencil_block<PRIMAL_GROUP>();
    const int tx = b % tiles_x, rest = b / tiles_x;
    const int x = tx * DUAL_BX + (int)(threadIdx.x & (DUAL_BX - 1));
    const int y = (rest % tiles_y) * DUAL_BY + (int)(threadIdx.x / DUAL_BX);
    const int z = rest / tiles_y;
    if (x >= P.nx || y >= P.ny) return;
    const size_t plane = (size_t)P.nx * P.ny;
    for (int c = 0; c < P.nch; ++c) {
        const size_t i = (size_t)(z + P.nz * c) * plane + (size_t)y * P.nx + x;
        const float ktl = ld_stream<NT>(&P.ktl[i]);
        primal_update<DIM, TGV, PRIMAL_U, false, NT>(P, i, x, y, z, ktl);
    }
}

template <int DIM, bool TGV, int PART>
__global__ __launch_bounds__(DUAL_BX* DUAL_BY) void primal_batch_kernel(
    const ProbDev* __restrict__ probs, const DualItem* __restrict__ items)
{
    const DualItem I = items[blockIdx.x];
    const ProbDev P = probs[I.prob];
    primal_body<DIM, TGV, PART>(P, I);
}

// --------------------------------------------------------------------------
// Backprojection of one tile: 64 x tile_y pixels of K slices s0...  Per
// chunk of angles it stages, for each angle, only the sinogram segment the
// tile projects onto (bins outside [0, nd) staged as zeros, which reproduces
// the reference's edge guards), then every thread interpolates for its
// tile_y/4 pixels, tile_y <= 4 * MAXR (the accumulators a lane holds: 8 rows
// take 126 VGPRs, 4 rows 70).  PRIMAL fuses the primal step into the epilogue.
// --------------------------------------------------------------------------
template <int K, bool PRIMAL, int DIM, bool TGV, int MAXR = ADJ_MAXR>
__device__ __forceinline__ void adj_body(const ProbDev& P, const AdjItem& I, float* sh)
{
    const int nx = P.nx, ny = P.ny, nd = P.nd, na = P.na;
    const int nslices = P.nz * P.nch;
    const int tile_y = P.adj_tile_y, segw = P.adj_segw, achunk = P.adj_achunk;
    float* seg = sh;                                  // achunk * segw * K
    const int hs = achunk * segw * 4;                 // K = 8: the second plane
    float4* par = reinterpret_cast<float4*>(sh + (size_t)achunk * segw * K);
    int* lo_s = reinterpret_cast<int*>(par + achunk);

    const int x0 = I.x0, y0 = I.y0, s0 = I.s0;
    const int tid = threadIdx.x;
    const int lane = tid & 63, wv = tid >> 6;
    const int R = tile_y >> 2;
    const float xf = (float)(x0 + lane);
    const float* __restrict__ sino = P.lam;

    float acc[MAXR][K];
#pragma unroll
    for (int r = 0; r < MAXR; ++r)
#pragma unroll
        for (int k = 0; k < K; ++k) acc[r][k] = 0.f;

    const size_t sino_slice = (size_t)na * nd;
    const int sg_a0 = tid / segw, sg_b0 = tid - sg_a0 * segw;
    const int sg_da = ADJ_THREADS / segw, sg_db = ADJ_THREADS - sg_da * segw;
    for (int a0 = 0; a0 < na; a0 += achunk) {
        const int A = min(achunk, na - a0);
        __syncthreads();
        if (tid < A) {
            const float4 o = P.ofs[a0 + tid];
            const float tmin = o.z + (float)x0 * o.x + (float)y0 * o.y +
                               fminf(0.f, (float)(ADJ_TX - 1) * o.x) +
                               fminf(0.f, (float)(tile_y - 1) * o.y);
            const int lo = (int)floorf(tmin) - 1;
            par[tid] = make_float4(o.x, o.y, o.z - (float)lo, 0.f);
            lo_s[tid] = lo;
        }
        __syncthreads();
        // element e = tid + j * ADJ_THREADS as (angle, bin) = (e / segw, e % segw),
        // stepped rather than divided
        for (int al = sg_a0, b = sg_b0; al < A;) {
            const int gb = lo_s[al] + b;
            float v[K];
#pragma unroll
            for (int k = 0; k < K; ++k) {
                const int s = s0 + k;
                v[k] = (gb >= 0 && gb < nd && s < nslices)
                           ? sino[(size_t)s * sino_slice + (size_t)(a0 + al) * nd + gb]
                           : 0.f;
            }
            lds_store<K>(seg, al * segw + b, v, hs);
            b += sg_db;
            al += sg_da;
            if (b >= segw) {
                b -= segw;
                ++al;
            }
        }
        __syncthreads();
        for (int al = 0; al < A; ++al) {
            const float4 pa = par[al];
            const float tb = fmaf(xf, pa.x, pa.z);
            const int base = al * segw;
#pragma unroll
            for (int r = 0; r < MAXR; ++r) {
                if (r >= R) break;
                // t >= 1 over the whole tile (the segment starts at
                // floor(min t) - 1), so truncation is the floor
                const float t = fmaf((float)(y0 + wv + 4 * r), pa.y, tb);
                const float p = __builtin_amdgcn_fractf(t);
                const float q = 1.f - p;
                const int i0 = base + (int)t;
                const Pack<K> v0 = lds_load<K>(seg, i0, hs);
                const Pack<K> v1 = lds_load<K>(seg, i0 + 1, hs);
#pragma unroll
                for (int k = 0; k < K; ++k)
                    acc[r][k] = fmaf(p, v1.v[k], fmaf(q, v0.v[k], acc[r][k]));
            }
        }
    }

    const int x = x0 + lane;
    if (x >= nx) return;
    const size_t plane = (size_t)nx * ny;
#pragma unroll
    for (int r = 0; r < MAXR; ++r) {
        if (r >= R) break;
        const int y = y0 + wv + 4 * r;
        if (y >= ny) continue;
#pragma unroll
        for (int k = 0; k < K; ++k) {
            const int s = s0 + k;
            if (s >= nslices) continue;
            const size_t i = (size_t)s * plane + (size_t)y * nx + x;
            if constexpr (PRIMAL)
                primal_update<DIM, TGV>(P, i, x, y, s % P.nz, acc[r][k]);
            else
                P.ktl[i] = acc[r][k];
        }
    }
}

template <int K, bool PRIMAL, int DIM, bool TGV, int MAXR>
__global__ __launch_bounds__(ADJ_THREADS) void adj_single_kernel(ProbDev P, int tiles_x,
                                                                 int ntiles)
{
    extern __shared__ float sh[];
    const int work = xcd_work_id(blockIdx.x, gridDim.x);
    const int tile = work % ntiles;
    const AdjItem I{0, (tile % tiles_x) * ADJ_TX, (tile / tiles_x) * P.adj_tile_y,
                    (work / ntiles) * K};
    adj_body<K, PRIMAL, DIM, TGV, MAXR>(P, I, sh);
}

template <int K, bool PRIMAL, int DIM, bool TGV>
__global__ __launch_bounds__(ADJ_THREADS) void adj_batch_kernel(const ProbDev* __restrict__ probs,
                                                                const AdjItem* __restrict__ items)
{
    extern __shared__ float sh[];
    const AdjItem I = items[xcd_work_id(blockIdx.x, gridDim.x)];
    const ProbDev P = probs[I.prob];
    adj_body<K, PRIMAL, DIM, TGV>(P, I, sh);
}

}  // namespace

// --------------------------------------------------------------------------
// sizes
// --------------------------------------------------------------------------
int fwd_table_choice(const Geometry& g, long long groups, double wg_per_cu)
{
    const double target = wg_per_cu * 256;  // work-groups per CU on 256 CUs
    int best = 0;
    double best_d = 1e30;
    for (int t = 0; t < 4; ++t) {
        const double wg = (double)groups * (g.tab_off[t + 1] - g.tab_off[t]);
        const double d = std::fabs(std::log(std::max(wg, 1.0) / target));
        if (d < best_d - 1e-9) {
            best_d = d;
            best = t;
        }
    }
    return best;
}

int fwd_pitch_of(int nx, int ny)
{
    int p = std::max(nx, ny) + FWD_PADL + 4;
    if ((p & 1) == 0) ++p;  // odd pitch: conflict-free transposed stores
    return p;
}

int fwd_band_auto(int pitch, int K, int lds_kb)
{
    // lines per band within lds_kb of LDS per work-group (36 KB: four per CU)
    int b = (int)((size_t)lds_kb * 1024 / ((size_t)pitch * K * sizeof(float)));
    return std::max(4, std::min(32, b & ~3));
}

int adj_segw_of(int tile_y)
{
    const double diag = std::sqrt((double)(ADJ_TX - 1) * (ADJ_TX - 1) +
                                  (double)(tile_y - 1) * (tile_y - 1));
    int w = (int)std::ceil(diag) + 6;
    return (w + 3) & ~3;
}

static int adj_lds_budget()
{
    static int kb = -1;
    if (kb < 0) {
        const char* e = getenv("GRAPTOR_ADJ_LDS_KB");
        kb = e ? std::max(4, std::min(150, atoi(e))) : 36;
    }
    return kb;
}

int adj_achunk_of(int na, int segw, int K)
{
    const size_t per_angle = (size_t)segw * K * sizeof(float) + sizeof(float4) + sizeof(int);
    int a = (int)std::min<size_t>((size_t)na, (adj_lds_budget() * 1024) / per_angle);
    return std::max(a, 1);
}

size_t fwd_lds_bytes(const ProbDev& P, int K)
{
    return (size_t)P.fwd_band * P.fwd_pitch * K * sizeof(float);
}

size_t adj_lds_bytes(const ProbDev& P, int K)
{
    return (size_t)P.adj_achunk * ((size_t)P.adj_segw * K * sizeof(float) + sizeof(float4) +
                                   sizeof(int));
}

// three slices take a group of four with one padded slot: one pass instead of
// a pair and a single
int slice_group(int nslices) { return nslices >= 3 ? 4 : (nslices == 2 ? 2 : 1); }

// --------------------------------------------------------------------------
// single-problem launchers
// --------------------------------------------------------------------------
template <int K, int MI>
static void fwd_single_launch(const ProbDev& P, const FwdChunk* chunks, int nchunks, int groups,
                              size_t lds, int mode, hipStream_t st)
{
    hipLaunchKernelGGL((fwd_single_kernel<K, MI>), dim3((unsigned)(groups * nchunks)),
                       dim3(FWD_THREADS), lds, st, P, chunks, nchunks, mode);
}

template <int K>
static int fwd_single_k(const ProbDev& Pin, const Geometry& g, int nslices, int mode,
                        hipStream_t st)
{
    ProbDev P = Pin;
    const int groups = (nslices + K - 1) / K;
    // chunk table: items per thread (8, 4, 2, 1) whose grid is closest to
    // three work-groups per CU, or 1.5 beside the dual step, where fewer,
    // larger work-groups leave it more of the CUs (measured: 0.827 against
    // 0.856 ms per two-stream iteration on 256x256x64x4, although the
    // forward projection alone is slower); GRAPTOR_FWD_MI = 8, 4, 2 or 1
    // forces the table
    int t = fwd_table_choice(g, groups, (mode & FWD_BESIDE) ? 1.5 : 3.0);
    static const int forced = [] {
        const char* e = getenv("GRAPTOR_FWD_MI");
        const int v = e ? atoi(e) : 0;
        return v == 8 ? 0 : v == 4 ? 1 : v == 2 ? 2 : v == 1 ? 3 : -1;
    }();
    if (forced >= 0) t = forced;
    const int nchunks = g.tab_off[t + 1] - g.tab_off[t];
    const FwdChunk* chunks = g.chunks_dev + g.tab_off[t];
    P.fwd_pitch = fwd_pitch_of(P.nx, P.ny);
    if (P.fwd_band <= 0) P.fwd_band = fwd_band_auto(P.fwd_pitch, K);
    if ((mode & FWD_BESIDE) && (mode & FWD_AUTOBAND)) {
        // beside the dual step, bands as tall as an even spread of the grid
        // over the 256 CUs allows: the LDS of a work-group then caps the
        // work-groups a CU takes, leaving every CU room for the dual step
        // (256^2 images: 512 work-groups, 2 per CU, 16-line bands: 0.786
        // against 0.804 ms per iteration with 8; 512^2: grids of thousands
        // keep 4 per CU and 4 lines, 18.6 against 19.2 ms with 8)
        const long long per_cu =
            std::min(4LL, std::max(1LL, ((long long)groups * nchunks + 255) / 256));
        P.fwd_band = fwd_band_auto(P.fwd_pitch, K, (int)(160 / per_cu) - 1);
    }
    const size_t lds = fwd_lds_bytes(P, K);
    if (lds > 160 * 1024) {
        set_error("radon: band %d does not fit LDS (%zu bytes)", P.fwd_band, lds);
        return GR_ERR_INVALID;
    }
    const long long blocks = (long long)groups * nchunks;
    if (blocks <= 0 || blocks > 0x7fffffff) {
        set_error("radon: invalid grid (%lld blocks)", blocks);
        return GR_ERR_INVALID;
    }
    switch (t) {
    case 0: fwd_single_launch<K, 8>(P, chunks, nchunks, groups, lds, mode, st); break;
    case 1: fwd_single_launch<K, 4>(P, chunks, nchunks, groups, lds, mode, st); break;
    case 2: fwd_single_launch<K, 2>(P, chunks, nchunks, groups, lds, mode, st); break;
    default: fwd_single_launch<K, 1>(P, chunks, nchunks, groups, lds, mode, st); break;
    }
    GR_LAUNCH_CHECK("fwd_single_kernel");
    return GR_OK;
}

int launch_fwd(const ProbDev& P, const Geometry& g, int nslices, int mode, hipStream_t st)
{
    if (nslices <= 0) return GR_OK;
    switch (slice_group(nslices)) {
    case 4: return fwd_single_k<4>(P, g, nslices, mode, st);
    case 2: return fwd_single_k<2>(P, g, nslices, mode, st);
    default: return fwd_single_k<1>(P, g, nslices, mode, st);
    }
}

template <int K, bool PRIMAL, int DIM, bool TGV>
static void adj_single_launch(const ProbDev& P, int blocks, int tiles_x, int ntiles, size_t lds,
                              hipStream_t st)
{
    // the backprojection alone takes accumulators for the rows of its tile
    // (tiles of <= 16 rows: 7 instead of 4 waves per SIMD); the fused primal
    // epilogue keeps the 8-row form
    if constexpr (!PRIMAL) {
        if (P.adj_tile_y <= 16) {
            hipLaunchKernelGGL((adj_single_kernel<K, PRIMAL, DIM, TGV, 4>),
                               dim3((unsigned)blocks), dim3(ADJ_THREADS), lds, st, P, tiles_x,
                               ntiles);
            return;
        }
    }
    hipLaunchKernelGGL((adj_single_kernel<K, PRIMAL, DIM, TGV, ADJ_MAXR>), dim3((unsigned)blocks),
                       dim3(ADJ_THREADS), lds, st, P, tiles_x, ntiles);
}

template <int K>
static int adj_single_k(const ProbDev& Pin, int nslices, bool primal, hipStream_t st)
{
    ProbDev P = Pin;
    if (P.adj_tile_y < 4 || P.adj_tile_y > 4 * ADJ_MAXR || (P.adj_tile_y & 3)) {
        set_error("radon_ad: tile_y %d must be a multiple of 4 in [4, %d]", P.adj_tile_y,
                  4 * ADJ_MAXR);
        return GR_ERR_INVALID;
    }
    P.adj_segw = adj_segw_of(P.adj_tile_y);
    P.adj_achunk = adj_achunk_of(P.na, P.adj_segw, K);
    const size_t lds = adj_lds_bytes(P, K);
    const int tiles_x = (P.nx + ADJ_TX - 1) / ADJ_TX;
    const int tiles_y = (P.ny + P.adj_tile_y - 1) / P.adj_tile_y;
    const int ntiles = tiles_x * tiles_y;
    const int groups = (nslices + K - 1) / K;
    const long long blocks = (long long)groups * ntiles;
    if (blocks <= 0 || blocks > 0x7fffffff) {
        set_error("radon_ad: invalid grid (%lld blocks)", blocks);
        return GR_ERR_INVALID;
    }
    const int b = (int)blocks;
    if (!primal) {
        adj_single_launch<K, false, 2, false>(P, b, tiles_x, ntiles, lds, st);
    } else if constexpr (K <= 4) {  // the fused primal epilogue takes up to 4 slices
        if (P.dim == 2) {
            if (P.tgv) adj_single_launch<K, true, 2, true>(P, b, tiles_x, ntiles, lds, st);
            else adj_single_launch<K, true, 2, false>(P, b, tiles_x, ntiles, lds, st);
        } else {
            if (P.tgv) adj_single_launch<K, true, 3, true>(P, b, tiles_x, ntiles, lds, st);
            else adj_single_launch<K, true, 3, false>(P, b, tiles_x, ntiles, lds, st);
        }
    }
    GR_LAUNCH_CHECK("adj_single_kernel");
    return GR_OK;
}

int launch_adj(const ProbDev& P, int nslices, bool primal, hipStream_t st, bool beside)
{
    if (nslices <= 0) return GR_OK;
    // eight slices per work-group: each interpolation weight serves twice the
    // multiply-adds (256^2 x 256 slices, 60 angles, 16-row tiles: 0.111
    // against 0.131 ms alone, same bits).  Not with 32-row tiles (8 rows x 8
    // slices of accumulators per lane: 2.92 against 2.47 ms on 512^2 x 1024
    // slices, 90 angles), nor beside the dual step (0.7874 against 0.7860 ms
    // per two-stream iteration).
    if (nslices >= 8 && !primal && !beside && P.adj_tile_y <= 16)
        return adj_single_k<8>(P, nslices, primal, st);
    switch (slice_group(nslices)) {
    case 4: return adj_single_k<4>(P, nslices, primal, st);
    case 2: return adj_single_k<2>(P, nslices, primal, st);
    default: return adj_single_k<1>(P, nslices, primal, st);
    }
}

template <int DIM, bool TGV, bool ZH, bool NT>
static void primal_single_part_nt(const ProbDev& P, int part, dim3 grd, dim3 blk, int tiles_x,
                                  int tiles_y, hipStream_t st)
{
    if (part == PRIMAL_U)
        hipLaunchKernelGGL((primal_single_kernel<DIM, TGV, PRIMAL_U, ZH, NT>), grd, blk, 0, st,
                           P, tiles_x, tiles_y);
    else if (part == PRIMAL_P)
        hipLaunchKernelGGL((primal_single_kernel<DIM, TGV, PRIMAL_P, ZH, NT>), grd, blk, 0, st,
                           P, tiles_x, tiles_y);
    else
        hipLaunchKernelGGL((primal_single_kernel<DIM, TGV, PRIMAL_UP, ZH, NT>), grd, blk, 0, st,
                           P, tiles_x, tiles_y);
}

template <int DIM, bool TGV, bool ZH>
static void primal_single_part_zh(const ProbDev& P, int part, dim3 grd, dim3 blk, int tiles_x,
                                  int tiles_y, hipStream_t st)
{
    if (stream_hints(P))
        primal_single_part_nt<DIM, TGV, ZH, true>(P, part, grd, blk, tiles_x, tiles_y, st);
    else
        primal_single_part_nt<DIM, TGV, ZH, false>(P, part, grd, blk, tiles_x, tiles_y, st);
}

template <int DIM, bool TGV>
static void primal_u_voxel_launch(const ProbDev& P, dim3 grd, dim3 blk, int tiles_x, int tiles_y,
                                  hipStream_t st)
{
    if (stream_hints(P))
        hipLaunchKernelGGL((primal_u_voxel_kernel<DIM, TGV, true>), grd, blk, 0, st, P, tiles_x,
                           tiles_y);
    else
        hipLaunchKernelGGL((primal_u_voxel_kernel<DIM, TGV, false>), grd, blk, 0, st, P,
                           tiles_x, tiles_y);
}

// a z-slab with halo planes takes the kernels that read them
template <int DIM, bool TGV>
static void primal_single_part(const ProbDev& P, int part, dim3 grd, dim3 blk, int tiles_x,
                               int tiles_y, hipStream_t st)
{
    if (DIM == 3 && (P.halo_ubar_hi || P.halo_v_lo))
        primal_single_part_zh<DIM, TGV, DIM == 3>(P, part, grd, blk, tiles_x, tiles_y, st);
    else
        primal_single_part_zh<DIM, TGV, false>(P, part, grd, blk, tiles_x, tiles_y, st);
}

int launch_primal(const ProbDev& P, hipStream_t st) { return launch_primal_part(P, PRIMAL_UP, st); }

int launch_primal_part(const ProbDev& P, int part, hipStream_t st)
{
    if (part != PRIMAL_U && part != PRIMAL_P && part != PRIMAL_UP) {
        set_error("primal step: invalid part %d", part);
        return GR_ERR_INVALID;
    }
    if (part == PRIMAL_P && !P.tgv) return GR_OK;  // TV: no p
    const int tiles_x = (P.nx + DUAL_BX - 1) / DUAL_BX;
    const int tiles_y = (P.ny + DUAL_BY - 1) / DUAL_BY;
    const long long blocks = (long long)tiles_x * tiles_y * P.nz * P.nch;
    if (blocks > 0x7fffffff) {
        set_error("primal step: grid too large");
        return GR_ERR_INVALID;
    }
    dim3 grd((unsigned)blocks), blk(DUAL_BX * DUAL_BY);
    const long long voxel_blocks = (long long)tiles_x * tiles_y * P.nz;
    if (part == PRIMAL_U && P.nch > 1 && voxel_blocks >= 4096 && !P.halo_ubar_hi &&
        !P.halo_v_lo) {
        // enough work-groups with one per voxel tile: the per-voxel u half
        const dim3 gv((unsigned)voxel_blocks);
        if (P.dim == 2) {
            if (P.tgv) primal_u_voxel_launch<2, true>(P, gv, blk, tiles_x, tiles_y, st);
            else primal_u_voxel_launch<2, false>(P, gv, blk, tiles_x, tiles_y, st);
        } else {
            if (P.tgv) primal_u_voxel_launch<3, true>(P, gv, blk, tiles_x, tiles_y, st);
            else primal_u_voxel_launch<3, false>(P, gv, blk, tiles_x, tiles_y, st);
        }
        GR_LAUNCH_CHECK("primal_u_voxel_kernel");
        return GR_OK;
    }
    if (P.dim == 2) {
        if (P.tgv) primal_single_part<2, true>(P, part, grd, blk, tiles_x, tiles_y, st);
        else primal_single_part<2, false>(P, part, grd, blk, tiles_x, tiles_y, st);
    } else {
        if (P.tgv) primal_single_part<3, true>(P, part, grd, blk, tiles_x, tiles_y, st);
        else primal_single_part<3, false>(P, part, grd, blk, tiles_x, tiles_y, st);
    }
    GR_LAUNCH_CHECK("primal_single_kernel");
    return GR_OK;
}

template <int DIM, bool TGV>
static void primal_batch_part(const ProbDev* probs, const DualItem* items, dim3 grd, dim3 blk,
                              int part, hipStream_t st)
{
    if (part == PRIMAL_U)
        hipLaunchKernelGGL((primal_batch_kernel<DIM, TGV, PRIMAL_U>), grd, blk, 0, st, probs,
                           items);
    else if (part == PRIMAL_P)
        hipLaunchKernelGGL((primal_batch_kernel<DIM, TGV, PRIMAL_P>), grd, blk, 0, st, probs,
                           items);
    else
        hipLaunchKernelGGL((primal_batch_kernel<DIM, TGV, PRIMAL_UP>), grd, blk, 0, st, probs,
                           items);
}

int launch_primal_batch(const ProbDev* probs, const DualItem* items, int nitems, int dim,
                        bool tgv, int part, hipStream_t st)
{
    if (part != PRIMAL_U && part != PRIMAL_P && part != PRIMAL_UP) {
        set_error("primal step: invalid part %d", part);
        return GR_ERR_INVALID;
    }
    if (nitems <= 0 || (part == PRIMAL_P && !tgv)) return GR_OK;  // TV: no p
    dim3 grd((unsigned)nitems), blk(DUAL_BX * DUAL_BY);
    if (dim == 2) {
        if (tgv) primal_batch_part<2, true>(probs, items, grd, blk, part, st);
        else primal_batch_part<2, false>(probs, items, grd, blk, part, st);
    } else {
        if (tgv) primal_batch_part<3, true>(probs, items, grd, blk, part, st);
        else primal_batch_part<3, false>(probs, items, grd, blk, part, st);
    }
    GR_LAUNCH_CHECK("primal_batch_kernel");
    return GR_OK;
}

// --------------------------------------------------------------------------
// batch launchers
// --------------------------------------------------------------------------
int launch_fwd_batch(const ProbDev* probs, const FwdItem* items, int nitems, int K, int mi,
                     size_t lds, int mode, hipStream_t st)
{
    if (nitems <= 0) return GR_OK;
    dim3 grd((unsigned)nitems), blk(FWD_THREADS);
#define GR_FB(KK, MM) \
    hipLaunchKernelGGL((fwd_batch_kernel<KK, MM>), grd, blk, lds, st, probs, items, mode)
#define GR_FB_K(KK)                                  \
    switch (mi) {                                    \
    case 8: GR_FB(KK, 8); break;                     \
    case 4: GR_FB(KK, 4); break;                     \
    case 2: GR_FB(KK, 2); break;                     \
    default: GR_FB(KK, 1); break;                    \
    }
    if (K == 4) { GR_FB_K(4) } else if (K == 2) { GR_FB_K(2) } else { GR_FB_K(1) }
#undef GR_FB_K
#undef GR_FB
    GR_LAUNCH_CHECK("fwd_batch_kernel");
    return GR_OK;
}

int launch_adj_batch(const ProbDev* probs, const AdjItem* items, int nitems, int K, int dim,
                     bool tgv, bool primal, size_t lds, hipStream_t st)
{
    if (nitems <= 0) return GR_OK;
    dim3 grd((unsigned)nitems), blk(ADJ_THREADS);
#define GR_AB(KK, DD, TT) \
    hipLaunchKernelGGL((adj_batch_kernel<KK, true, DD, TT>), grd, blk, lds, st, probs, items)
#define GR_AB_K(KK)                                          \
    if (!primal) hipLaunchKernelGGL((adj_batch_kernel<KK, false, 2, false>), grd, blk, lds, st, probs, items); \
    else if (dim == 2) { if (tgv) GR_AB(KK, 2, true); else GR_AB(KK, 2, false); } \
    else { if (tgv) GR_AB(KK, 3, true); else GR_AB(KK, 3, false); }
    if (K == 4) { GR_AB_K(4) } else if (K == 2) { GR_AB_K(2) } else { GR_AB_K(1) }
#undef GR_AB_K
#undef GR_AB
    GR_LAUNCH_CHECK("adj_batch_kernel");
    return GR_OK;
}

}  // namespace gr
