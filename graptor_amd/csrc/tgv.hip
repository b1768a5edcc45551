// TGV stencil kernels for gfx950.
//
// dual step    : one thread per voxel, all channels: v += sigma (grad ubar - pbar),
//                w += sigma E(pbar), then the coupled projections of v and w
//                (update_v/update_w/update_NormV_*/update_NormW_* of
//                radon_tgv_primal_dual_2d.py 2D:98-304 and _3d.py 3D:290-660
//                fused into one pass).  The primal step lives in projector.hip,
//                fused with the backprojection.
// energy kernels: FP64 block partial sums of the primal energy.
#include "common.h"
#include "proj.h"
#include "stencil.h"

#include <cmath>
#include <cstdlib>
#include <mutex>
#include <vector>

namespace gr {

namespace {

// ---- dual step ------------------------------------------------------------

// Sources of the extrapolated iterates the dual step reads.  MemSrc reads
// the stored ubar / pbar; FusedSrc evaluates the primal step of the
// previous iteration on the fly (own voxel: stored as the new u, ubar, p;
// neighbours: recomputed), so that one kernel performs the primal step of
// iteration k and the dual step of iteration k + 1.
template <int DIM, bool TGV>
struct MemSrc {
    // callers run 64 consecutive x per wave: x +- 1 neighbours come from lanes
    static constexpr bool lanex = true;
    const float* __restrict__ ubar;
    const float* __restrict__ pbar;
    size_t cs;
    __device__ __forceinline__ void own(int, size_t j, int, int, int, float& ub,
                                        float (&pb)[3]) const
    {
        ub = ubar[j];
        if constexpr (TGV) {
#pragma unroll
            for (int k = 0; k < DIM; ++k) pb[k] = pbar[j + k * cs];
        }
    }
    __device__ __forceinline__ float ubn(int, size_t j, int, int, int) const { return ubar[j]; }
    __device__ __forceinline__ void pbn(int, size_t j, int, int, int, float (&pb)[3]) const
    {
#pragma unroll
        for (int k = 0; k < DIM; ++k) pb[k] = pbar[j + k * cs];
    }
};

template <int DIM, bool TGV>
struct FusedSrc {
    static constexpr bool lanex = false;
    // previous iterate (read only in the kernel)
    const float* __restrict__ u;
    const float* __restrict__ p;
    const float* __restrict__ v;
    const float* __restrict__ w;
    const float* __restrict__ ktl;
    // primal results of the own voxel
    float* __restrict__ u_out;
    float* __restrict__ ubar_out;
    float* __restrict__ p_out;
    int nx, ny, nz;
    size_t cs;
    float hinv, tau, norming;
    __device__ __forceinline__ void own(int, size_t j, int x, int y, int z, float& ub,
                                        float (&pb)[3]) const
    {
        float un, pn[3];
        primal_point<DIM, TGV, true, true>(u, p, v, w, ktl[j], nx, ny, nz, cs, hinv, tau, norming,
                                           j, x, y, z, un, ub, pn, pb);
        u_out[j] = un;
        ubar_out[j] = ub;
        if constexpr (TGV) {
#pragma unroll
            for (int k = 0; k < DIM; ++k) p_out[j + k * cs] = pn[k];
        }
    }
    __device__ __forceinline__ float ubn(int, size_t j, int x, int y, int z) const
    {
        float un, ub, pn[3], pb[3];
        primal_point<DIM, TGV, true, false>(u, p, v, w, ktl[j], nx, ny, nz, cs, hinv, tau,
                                            norming, j, x, y, z, un, ub, pn, pb);
        return ub;
    }
    __device__ __forceinline__ void pbn(int, size_t j, int x, int y, int z, float (&pb)[3]) const
    {
        float un, ub, pn[3];
        primal_point<DIM, TGV, false, true>(u, p, v, w, 0.f, nx, ny, nz, cs, hinv, tau, norming,
                                            j, x, y, z, un, ub, pn, pb);
    }
};

// One channel's dual ascent at one voxel, before any projection:
// vv = v + sigma (grad ubar - pbar), ww = w + sigma E(pbar).  Returns the
// weighted squared norm of ww (2D:232, 3D:582; 0 without TGV).  v and w are
// read at the voxel itself only, so with NT (single problems) their loads
// (and the stores of the new v and w) carry the non-temporal hint: streamed
// once, they would otherwise displace the ubar / pbar lines that the
// neighbouring voxels read again (dual step 0.309 -> 0.287 ms, two-stream
// iteration 0.776 -> 0.735 ms on 256x256x64x4 together with the primal
// step's hints, same bits); ubar and pbar are read at the neighbours too (x
// neighbours from the adjacent lanes of a 64-wide row) and keep ordinary
// loads.
template <int DIM, int CPL, bool TGV, class Src, bool ZH, bool NT>
__device__ __forceinline__ float dual_candidate(const ProbDev& P, const Src& src, const float* v,
                                                const float* w, int x, int y, int z, size_t i,
                                                int c, float (&vv)[DIM],
                                                float (&ww)[TGV ? (DIM == 2 ? 3 : 6) : 1])
{
    constexpr int D = DIM;
    constexpr int M = DIM == 2 ? 3 : 6;
    const int nx = P.nx, ny = P.ny, nz = P.nz, nch = P.nch;
    const float hinv = P.hinv, sigma = P.sigma;
    const size_t plane = (size_t)nx * ny;
    const size_t chs = plane * nz;        // channel stride
    const size_t cs = chs * nch;          // component stride
    const size_t ic = i + c * chs;
    float u0;
    float p[3] = {0.f, 0.f, 0.f};
    src.own(c, ic, x, y, z, u0, p);
    float g[3];
    if constexpr (Src::lanex) {
        float un = lane_next(u0);
        if ((threadIdx.x & 63) == 63 && x < nx - 1) un = src.ubn(c, ic + 1, x + 1, y, z);
        g[0] = (x < nx - 1) ? un - u0 : 0.f;
    } else {
        g[0] = (x < nx - 1) ? src.ubn(c, ic + 1, x + 1, y, z) - u0 : 0.f;
    }
    g[1] = (y < ny - 1) ? src.ubn(c, ic + nx, x, y + 1, z) - u0 : 0.f;
    if constexpr (DIM == 3) {
        if constexpr (ZH) {
            // stored ubar; above a slab's last plane the halo plane, if any
            const float* up = z < nz - 1 ? src.ubar + ic + plane
                              : P.halo_ubar_hi
                                  ? P.halo_ubar_hi + (size_t)c * plane + (size_t)y * nx + x
                                  : nullptr;
            g[2] = up ? (*up - u0) * hinv : 0.f;
        } else {
            g[2] = (z < nz - 1) ? (src.ubn(c, ic + plane, x, y, z + 1) - u0) * hinv : 0.f;
        }
    }
#pragma unroll
    for (int k = 0; k < D; ++k)
        vv[k] = ld_stream<NT>(&v[ic + k * cs]) + sigma * (g[k] - p[k]);
    float n2 = 0.f;
    if constexpr (TGV) {
        float ax[3] = {0.f, 0.f, 0.f}, ay[3] = {0.f, 0.f, 0.f}, az[3] = {0.f, 0.f, 0.f};
        float q[3];
        if constexpr (Src::lanex) {
#pragma unroll
            for (int k = 0; k < D; ++k) q[k] = lane_prev(p[k]);
            if ((threadIdx.x & 63) == 0 && x > 0) src.pbn(c, ic - 1, x - 1, y, z, q);
            if (x > 0) {
#pragma unroll
                for (int k = 0; k < D; ++k) ax[k] = p[k] - q[k];
            }
        } else if (x > 0) {
            src.pbn(c, ic - 1, x - 1, y, z, q);
#pragma unroll
            for (int k = 0; k < D; ++k) ax[k] = p[k] - q[k];
        }
        if (y > 0) {
            src.pbn(c, ic - nx, x, y - 1, z, q);
#pragma unroll
            for (int k = 0; k < D; ++k) ay[k] = p[k] - q[k];
        }
        if constexpr (DIM == 3) {
            if constexpr (ZH) {
                // stored pbar; below a slab's first plane the halo plane, if any
                const float* dn = z > 0 ? src.pbar + ic - plane
                                  : P.halo_pbar_lo
                                      ? P.halo_pbar_lo + (size_t)c * plane + (size_t)y * nx + x
                                      : nullptr;
                const size_t kst = z > 0 ? cs : (size_t)nch * plane;
                if (dn) {
#pragma unroll
                    for (int k = 0; k < D; ++k) az[k] = (p[k] - dn[k * kst]) * hinv;
                }
            } else if (z > 0) {
                src.pbn(c, ic - plane, x, y, z - 1, q);
#pragma unroll
                for (int k = 0; k < D; ++k) az[k] = (p[k] - q[k]) * hinv;
            }
        }
        float e[6];
        e[0] = ax[0];
        e[1] = ay[1];
        if constexpr (DIM == 2) {
            e[2] = 0.5f * (ax[1] + ay[0]);
        } else {
            e[2] = az[2];
            e[3] = 0.5f * (ax[1] + ay[0]);
            e[4] = 0.5f * (ax[2] + az[0]);
            e[5] = 0.5f * (ay[2] + az[1]);
        }
#pragma unroll
        for (int m = 0; m < M; ++m)
            ww[m] = ld_stream<NT>(&w[ic + m * cs]) + sigma * e[m];
        // weighted squared norm of this channel's tensor (2D:232, 3D:582)
        if constexpr (DIM == 2) {
            const float od = CPL == GR_COUPLING_UNCORRELATED ? W_OFF * ww[2] : 2.f * ww[2];
            n2 = ww[0] * ww[0] + ww[1] * ww[1] + od * od;
        } else {
            const float o3 = W_OFF * ww[3], o4 = W_OFF * ww[4];
            const float o5 = CPL == GR_COUPLING_UNCORRELATED ? ww[5] : W_OFF * ww[5];
            n2 = ww[0] * ww[0] + ww[1] * ww[1] + ww[2] * ww[2] + o3 * o3 + o4 * o4 + o5 * o5;
        }
    }
    return n2;
}

// Per-channel projections of the uncorrelated coupling (2D:98-128,
// 3D:553-570, 597-613): each channel's v onto its own ball, each channel's
// w (weighted norm n2w) onto its own.
template <int DIM, bool TGV>
__device__ __forceinline__ void project_channel(float (&vv)[DIM],
                                                float (&ww)[TGV ? (DIM == 2 ? 3 : 6) : 1],
                                                float n2w, float ainv0, float ainv1)
{
    constexpr int M = DIM == 2 ? 3 : 6;
    if constexpr (TGV) {
        const float n = sqrtf(n2w) * ainv0;
        const float sc = n > 1.f ? 1.f / n : 1.f;
#pragma unroll
        for (int m = 0; m < M; ++m) ww[m] *= sc;
    }
    float n2 = 0.f;
#pragma unroll
    for (int k = 0; k < DIM; ++k) n2 += vv[k] * vv[k];
    const float n = sqrtf(n2) * ainv1;
    const float sc = n > 1.f ? 1.f / n : 1.f;
#pragma unroll
    for (int k = 0; k < DIM; ++k) vv[k] *= sc;
}

// The 3D nuclear-norm projection of a voxel's v (3D:290-527) as a 3x3 matrix
// from the Gram matrix G = sum_c v_c v_c^T (xx, yy, zz, xy, xz, yz): every
// singular value s of [v_c] maps to min(s, 1/ainv1), i.e. v_c <- Mx^T v_c.
__device__ __forceinline__ void nuclear_matrix3(const float (&G)[6], float ainv1,
                                                float (&Mx)[3][3])
{
    float ev[3], V[3][3], S[3];
    jacobi3<float>(G, ev, V);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const float sg = ev[k] > 0.f ? sqrtf(ev[k]) : 0.f;
        S[k] = sg * ainv1 > 1.f ? 1.f / (ainv1 * sg) : 1.f;
    }
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int q = 0; q < 3; ++q)
            Mx[r][q] = S[0] * V[r][0] * V[q][0] + S[1] * V[r][1] * V[q][1] +
                       S[2] * V[r][2] * V[q][2];
}

// Dual step at one voxel, all channels (update_v/update_w + the coupled
// projections): v = proj(v_in + sigma (grad ubar - pbar), w = proj(w_in +
// sigma E pbar), written to v_out / w_out (the same arrays as v / w for the
// standalone step: each thread reads its own entries before writing them,
// so they are not declared restrict).
template <int DIM, int CPL, bool TGV, int MAXC, class Src, bool ZH = false, bool NT = false>
__device__ __forceinline__ void dual_core(const ProbDev& P, int x, int y, int z, const Src& src,
                                          const float* v, const float* w, float* vout,
                                          float* wout)
{
    constexpr int D = DIM;
    constexpr int M = DIM == 2 ? 3 : 6;
    const int nx = P.nx, ny = P.ny, nz = P.nz, nch = P.nch;
    const float ainv0 = P.ainv0, ainv1 = P.ainv1;
    if (x >= nx || y >= ny) return;
    const size_t plane = (size_t)nx * ny;
    const size_t chs = plane * nz;        // channel stride
    const size_t cs = chs * nch;          // component stride
    const size_t i = (size_t)z * plane + (size_t)y * nx + x;

    float vv[MAXC][D];
    float ww[MAXC][TGV ? M : 1];
    float nv2 = 0.f, nw2 = 0.f;
#pragma unroll
    for (int c = 0; c < MAXC; ++c) {
        if (c >= nch) break;
        const float n2 =
            dual_candidate<DIM, CPL, TGV, Src, ZH, NT>(P, src, v, w, x, y, z, i, c, vv[c], ww[c]);
#pragma unroll
        for (int k = 0; k < D; ++k) nv2 += vv[c][k] * vv[c][k];
        if constexpr (CPL == GR_COUPLING_UNCORRELATED)
            project_channel<DIM, TGV>(vv[c], ww[c], n2, ainv0, ainv1);
        else
            nw2 += n2;
    }

    // coupled projections
    if constexpr (CPL == GR_COUPLING_FROBENIUS) {
        const float n = sqrtf(nv2) * ainv1;
        const float sc = n > 1.f ? 1.f / n : 1.f;
#pragma unroll
        for (int c = 0; c < MAXC; ++c)
#pragma unroll
            for (int k = 0; k < D; ++k) vv[c][k] *= sc;
    } else if constexpr (CPL == GR_COUPLING_NUCLEAR) {
        if constexpr (DIM == 2) {
            float a = 0.f, b = 0.f, c2 = 0.f;
#pragma unroll
            for (int c = 0; c < MAXC; ++c) {
                if (c >= nch) break;
                a += vv[c][0] * vv[c][0];
                b += vv[c][0] * vv[c][1];
                c2 += vv[c][1] * vv[c][1];
            }
            float m00, m01, m11;
            nucl2d(a, b, c2, ainv1, m00, m01, m11);
#pragma unroll
            for (int c = 0; c < MAXC; ++c) {
                const float x0 = vv[c][0], x1 = vv[c][1];
                vv[c][0] = x0 * m00 + x1 * m01;
                vv[c][1] = x0 * m01 + x1 * m11;
            }
        } else {
            float G[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
#pragma unroll
            for (int c = 0; c < MAXC; ++c) {
                if (c >= nch) break;
                G[0] += vv[c][0] * vv[c][0];
                G[1] += vv[c][1] * vv[c][1];
                G[2] += vv[c][2] * vv[c][2];
                G[3] += vv[c][0] * vv[c][1];
                G[4] += vv[c][0] * vv[c][2];
                G[5] += vv[c][1] * vv[c][2];
            }
            float Mx[3][3];
            nuclear_matrix3(G, ainv1, Mx);
#pragma unroll
            for (int c = 0; c < MAXC; ++c) {
                const float a0 = vv[c][0], a1 = vv[c][1], a2 = vv[c][2];
#pragma unroll
                for (int q = 0; q < 3; ++q)
                    vv[c][q] = a0 * Mx[0][q] + a1 * Mx[1][q] + a2 * Mx[2][q];
            }
        }
    }
    float wsc = 1.f;
    if constexpr (TGV && CPL != GR_COUPLING_UNCORRELATED) {
        const float n = sqrtf(nw2) * ainv0;
        wsc = n > 1.f ? 1.f / n : 1.f;
    }
#pragma unroll
    for (int c = 0; c < MAXC; ++c) {
        if (c >= nch) break;
        const size_t ic = i + c * chs;
#pragma unroll
        for (int k = 0; k < D; ++k) st_stream<NT>(vv[c][k], &vout[ic + k * cs]);
        if constexpr (TGV) {
#pragma unroll
            for (int m = 0; m < M; ++m)
                st_stream<NT>(ww[c][m] * wsc, &wout[ic + m * cs]);
        }
    }
}

// Dual step at one voxel for any number of channels: the channels are swept
// twice, once for the coupling statistics (Frobenius: the squared norms of
// all v and w; nuclear: those and the Gram matrix of the v) and once to
// project and write, recomputing each channel's ascent step (its inputs are
// read twice).  Same arithmetic in the same order as dual_core; only one
// channel is held in registers at a time.
template <int DIM, int CPL, bool TGV, class Src, bool ZH = false, bool NT = false>
__device__ __forceinline__ void dual_core_wide(const ProbDev& P, int x, int y, int z,
                                               const Src& src, const float* v, const float* w,
                                               float* vout, float* wout)
{
    constexpr int D = DIM;
    constexpr int M = DIM == 2 ? 3 : 6;
    const int nx = P.nx, ny = P.ny, nch = P.nch;
    const float ainv0 = P.ainv0, ainv1 = P.ainv1;
    if (x >= nx || y >= ny) return;
    const size_t plane = (size_t)nx * ny;
    const size_t chs = plane * P.nz;      // channel stride
    const size_t cs = chs * nch;          // component stride
    const size_t i = (size_t)z * plane + (size_t)y * nx + x;
    float vv[D];
    float ww[TGV ? M : 1];
    float nv2 = 0.f, nw2 = 0.f;
    float G[6] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f};  // 2D: a, b, c of the 2x2 Gram matrix
    if constexpr (CPL != GR_COUPLING_UNCORRELATED) {
        for (int c = 0; c < nch; ++c) {
            nw2 += dual_candidate<DIM, CPL, TGV, Src, ZH, NT>(P, src, v, w, x, y, z, i, c, vv, ww);
#pragma unroll
            for (int k = 0; k < D; ++k) nv2 += vv[k] * vv[k];
            if constexpr (CPL == GR_COUPLING_NUCLEAR) {
                if constexpr (DIM == 2) {
                    G[0] += vv[0] * vv[0];
                    G[1] += vv[0] * vv[1];
                    G[2] += vv[1] * vv[1];
                } else {
                    G[0] += vv[0] * vv[0];
                    G[1] += vv[1] * vv[1];
                    G[2] += vv[2] * vv[2];
                    G[3] += vv[0] * vv[1];
                    G[4] += vv[0] * vv[2];
                    G[5] += vv[1] * vv[2];
                }
            }
        }
    }
    float vsc = 1.f, m00 = 1.f, m01 = 0.f, m11 = 1.f;
    float Mx[3][3] = {{1.f, 0.f, 0.f}, {0.f, 1.f, 0.f}, {0.f, 0.f, 1.f}};
    if constexpr (CPL == GR_COUPLING_FROBENIUS) {
        const float n = sqrtf(nv2) * ainv1;
        vsc = n > 1.f ? 1.f / n : 1.f;
    } else if constexpr (CPL == GR_COUPLING_NUCLEAR) {
        if constexpr (DIM == 2) nucl2d(G[0], G[1], G[2], ainv1, m00, m01, m11);
        else nuclear_matrix3(G, ainv1, Mx);
    }
    float wsc = 1.f;
    if constexpr (TGV && CPL != GR_COUPLING_UNCORRELATED) {
        const float n = sqrtf(nw2) * ainv0;
        wsc = n > 1.f ? 1.f / n : 1.f;
    }
    for (int c = 0; c < nch; ++c) {
        const float n2 =
            dual_candidate<DIM, CPL, TGV, Src, ZH, NT>(P, src, v, w, x, y, z, i, c, vv, ww);
        if constexpr (CPL == GR_COUPLING_UNCORRELATED) {
            project_channel<DIM, TGV>(vv, ww, n2, ainv0, ainv1);
        } else if constexpr (CPL == GR_COUPLING_FROBENIUS) {
#pragma unroll
            for (int k = 0; k < D; ++k) vv[k] *= vsc;
        } else if constexpr (DIM == 2) {
            const float x0 = vv[0], x1 = vv[1];
            vv[0] = x0 * m00 + x1 * m01;
            vv[1] = x0 * m01 + x1 * m11;
        } else {
            const float a0 = vv[0], a1 = vv[1], a2 = vv[2];
#pragma unroll
            for (int q = 0; q < 3; ++q) vv[q] = a0 * Mx[0][q] + a1 * Mx[1][q] + a2 * Mx[2][q];
        }
        const size_t ic = i + c * chs;
#pragma unroll
        for (int k = 0; k < D; ++k) st_stream<NT>(vv[k], &vout[ic + k * cs]);
        if constexpr (TGV) {
#pragma unroll
            for (int m = 0; m < M; ++m)
                st_stream<NT>(ww[m] * wsc, &wout[ic + m * cs]);
        }
    }
}

template <int DIM, int CPL, bool TGV, bool ZH = false, bool NT = false>
__device__ __forceinline__ void dual_body_wide(const ProbDev& P, const DualItem& I)
{
    const int x = I.x0 + (int)(threadIdx.x & (DUAL_BX - 1));
    const int y = I.y0 + (int)(threadIdx.x / DUAL_BX);
    const size_t cs = (size_t)P.nx * P.ny * P.nz * P.nch;
    const MemSrc<DIM, TGV> src{P.ubar, P.pbar, cs};
    dual_core_wide<DIM, CPL, TGV, MemSrc<DIM, TGV>, ZH, NT>(P, x, y, I.z, src, P.v, P.w, P.v,
                                                          P.w);
}

// ZH: the problem is a z-slab with halo planes (gr_tgv_state's halo_*)
template <int DIM, int CPL, bool TGV, int MAXC, bool ZH = false, bool NT = false>
__device__ __forceinline__ void dual_body(const ProbDev& P, const DualItem& I)
{
    const int x = I.x0 + (int)(threadIdx.x & (DUAL_BX - 1));
    const int y = I.y0 + (int)(threadIdx.x / DUAL_BX);
    const size_t cs = (size_t)P.nx * P.ny * P.nz * P.nch;
    const MemSrc<DIM, TGV> src{P.ubar, P.pbar, cs};
    dual_core<DIM, CPL, TGV, MAXC, MemSrc<DIM, TGV>, ZH, NT>(P, x, y, I.z, src, P.v, P.w, P.v,
                                                           P.w);
}

// Work order: x tiles fastest, then rows, then planes, with runs of
// DUAL_GROUP tiles kept on one XCD (plane-major orders, z-fastest orders and
// giving each XCD a contiguous eighth of the volume all measured slower).
// Waves per SIMD requested of the register allocator.  The 3D Frobenius TGV
// step with up to 4 channels needs 66 VGPRs (7 waves); held to 64 it runs 8
// waves with no scratch, which makes it faster beside the forward projection
// on the other stream (0.827 -> 0.813 ms per two-stream iteration on
// 256x256x64x4) at the same speed alone.  Other forms would spill.
template <int DIM, int CPL, bool TGV, int MAXC>
constexpr int dual_min_waves()
{
    return (DIM == 3 && CPL == GR_COUPLING_FROBENIUS && TGV && MAXC == 4) ? 8 : 1;
}

template <int DIM, int CPL, bool TGV, int MAXC, bool ZH, bool NT>
__global__ __launch_bounds__(DUAL_BX* DUAL_BY)
__attribute__((amdgpu_waves_per_eu(dual_min_waves<DIM, CPL, TGV, MAXC>())))
void dual_single_kernel(ProbDev P, int tiles_x, int tiles_y)
{
    const int b = stencil_block<DUAL_GROUP>();
    const int tx = b % tiles_x, rest = b / tiles_x;
    const DualItem I{0, tx * DUAL_BX, (rest % tiles_y) * DUAL_BY, rest / tiles_y};
    dual_body<DIM, CPL, TGV, MAXC, ZH, NT>(P, I);
}

template <int DIM, int CPL, bool TGV, int MAXC>
__global__ __launch_bounds__(DUAL_BX* DUAL_BY) void dual_batch_kernel(
    const ProbDev* __restrict__ probs, const DualItem* __restrict__ items)
{
    const DualItem I = items[blockIdx.x];
    const ProbDev P = probs[I.prob];
    dual_body<DIM, CPL, TGV, MAXC>(P, I);
}

// More than 8 channels: the two-sweep form, one channel in registers at a time.
template <int DIM, int CPL, bool TGV, bool ZH, bool NT>
__global__ __launch_bounds__(DUAL_BX* DUAL_BY) void dual_wide_kernel(ProbDev P, int tiles_x,
                                                                     int tiles_y)
{
    const int b = stencil_block<DUAL_GROUP>();
    const int tx = b % tiles_x, rest = b / tiles_x;
    const DualItem I{0, tx * DUAL_BX, (rest % tiles_y) * DUAL_BY, rest / tiles_y};
    dual_body_wide<DIM, CPL, TGV, ZH, NT>(P, I);
}

template <int DIM, int CPL, bool TGV>
__global__ __launch_bounds__(DUAL_BX* DUAL_BY) void dual_wide_batch_kernel(
    const ProbDev* __restrict__ probs, const DualItem* __restrict__ items)
{
    const DualItem I = items[blockIdx.x];
    const ProbDev P = probs[I.prob];
    dual_body_wide<DIM, CPL, TGV>(P, I);
}

// ---- FP64 energy ----------------------------------------------------------

constexpr int EN_THREADS = 256;
constexpr int EN_REG_BLOCKS = 1024;   // block partials of the regulariser
constexpr int EN_DATA_BLOCKS = 256;   // block partials per channel of the data term

__device__ double block_sum(double v, double* red)
{
    red[threadIdx.x] = v;
    __syncthreads();
    for (int s = EN_THREADS / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    return red[0];
}

// Regulariser a1 |grad u - p|_V* + a0 |E p|_W*, pointwise dual norms of the
// projection balls; one block partial per block.
template <int DIM, int CPL, bool TGV>
__global__ __launch_bounds__(EN_THREADS) void reg_energy_kernel(
    const float* __restrict__ u, const float* __restrict__ p, int nx, int ny, int nz,
    int nch, double hinv, double a0, double a1, double* __restrict__ partial)
{
    __shared__ double red[EN_THREADS];
    constexpr int D = DIM;
    constexpr int M = DIM == 2 ? 3 : 6;
    const size_t plane = (size_t)nx * ny;
    const size_t nvox = plane * nz;
    const size_t chs = nvox, cs = nvox * nch;
    double acc = 0.0;
    for (size_t i = (size_t)blockIdx.x * EN_THREADS + threadIdx.x; i < nvox;
         i += (size_t)gridDim.x * EN_THREADS) {
        const int x = (int)(i % nx);
        const int y = (int)((i / nx) % ny);
        const int z = (int)(i / plane);
        double fro_v = 0.0, fro_w = 0.0, sum_v = 0.0, sum_w = 0.0;
        double G[6] = {0, 0, 0, 0, 0, 0};
        for (int c = 0; c < nch; ++c) {
            const size_t ic = i + c * chs;
            const double u0 = u[ic];
            double q[3];
            q[0] = (x < nx - 1) ? (double)u[ic + 1] - u0 : 0.0;
            q[1] = (y < ny - 1) ? (double)u[ic + nx] - u0 : 0.0;
            q[2] = (DIM == 3 && z < nz - 1) ? ((double)u[ic + plane] - u0) * hinv : 0.0;
            double pc[3] = {0, 0, 0};
            if (TGV)
                for (int k = 0; k < D; ++k) pc[k] = p[ic + k * cs];
            for (int k = 0; k < D; ++k) q[k] -= pc[k];
            double n2 = 0.0;
            for (int k = 0; k < D; ++k) n2 += q[k] * q[k];
            fro_v += n2;
            sum_v += sqrt(n2);
            G[0] += q[0] * q[0];
            G[1] += q[1] * q[1];
            G[2] += q[2] * q[2];
            G[3] += q[0] * q[1];
            G[4] += q[0] * q[2];
            G[5] += q[1] * q[2];
            if (TGV) {
                double ax[3] = {0, 0, 0}, ay[3] = {0, 0, 0}, az[3] = {0, 0, 0};
                for (int k = 0; k < D; ++k) {
                    ax[k] = x > 0 ? pc[k] - (double)p[ic - 1 + k * cs] : 0.0;
                    ay[k] = y > 0 ? pc[k] - (double)p[ic - nx + k * cs] : 0.0;
                    if (DIM == 3)
                        az[k] = z > 0 ? (pc[k] - (double)p[ic - plane + k * cs]) * hinv : 0.0;
                }
                double e[6];
                e[0] = ax[0];
                e[1] = ay[1];
                double cw[6];  // b_k^2 / a_k of the dual norm
                if (DIM == 2) {
                    e[2] = 0.5 * (ax[1] + ay[0]);
                    const double od = CPL == GR_COUPLING_UNCORRELATED
                                          ? (double)W_OFF * W_OFF : 4.0;
                    cw[0] = 1; cw[1] = 1; cw[2] = 4.0 / od;
                } else {
                    e[2] = az[2];
                    e[3] = 0.5 * (ax[1] + ay[0]);
                    e[4] = 0.5 * (ax[2] + az[0]);
                    e[5] = 0.5 * (ay[2] + az[1]);
                    const double r2 = (double)W_OFF * W_OFF;
                    cw[0] = 1; cw[1] = 1; cw[2] = 1;
                    cw[3] = 4.0 / r2; cw[4] = 4.0 / r2;
                    cw[5] = CPL == GR_COUPLING_UNCORRELATED ? 4.0 : 4.0 / r2;
                }
                double m2 = 0.0;
                for (int m = 0; m < M; ++m) m2 += cw[m] * e[m] * e[m];
                fro_w += m2;
                sum_w += sqrt(m2);
            }
        }
        double rv;
        if (CPL == GR_COUPLING_UNCORRELATED) {
            rv = sum_v;
        } else if (CPL == GR_COUPLING_FROBENIUS) {
            rv = sqrt(fro_v);
        } else {
            double ev[3], V[3][3];
            jacobi3<double>(G, ev, V);
            rv = 0.0;
            for (int k = 0; k < (DIM == 2 ? 3 : 3); ++k) rv += ev[k] > 0 ? sqrt(ev[k]) : 0.0;
        }
        double rw = 0.0;
        if (TGV) rw = (CPL == GR_COUPLING_UNCORRELATED) ? sum_w : sqrt(fro_w);
        acc += a1 * rv + a0 * rw;
    }
    const double s = block_sum(acc, red);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// Data term per channel: KL(y, f) = y - f + f log(f / y) (f > 0; +inf when
// y <= 0 < f), y when f == 0; L2: (y - f)^2 / 2.  partial[c * nb + b].
__global__ __launch_bounds__(EN_THREADS) void data_energy_kernel(
    const float* __restrict__ ku, const float* __restrict__ f, size_t per_ch, int disc,
    double* __restrict__ partial)
{
    __shared__ double red[EN_THREADS];
    const int c = blockIdx.y;
    const float* kc = ku + (size_t)c * per_ch;
    const float* fc = f + (size_t)c * per_ch;
    double acc = 0.0;
    for (size_t i = (size_t)blockIdx.x * EN_THREADS + threadIdx.x; i < per_ch;
         i += (size_t)gridDim.x * EN_THREADS) {
        const double yv = kc[i], fv = fc[i];
        if (disc == GR_DISC_KL) {
            if (fv > 0.0)
                acc += yv > 0.0 ? yv - fv + fv * log(fv / yv) : INFINITY;
            else
                acc += yv;
        } else {
            acc += 0.5 * (yv - fv) * (yv - fv);
        }
    }
    const double s = block_sum(acc, red);
    if (threadIdx.x == 0) partial[(size_t)c * gridDim.x + blockIdx.x] = s;
}

}  // namespace

// --------------------------------------------------------------------------
// launchers
// --------------------------------------------------------------------------
bool stream_hints(const ProbDev& P)
{
    const char* e = getenv("GRAPTOR_STREAM_HINTS");
    if (e && (e[0] == '0' || e[0] == '1')) return e[0] == '1';
    const int D = P.dim;
    const int per = 2 + D + 1 + (P.tgv ? 2 * D + D * (D + 1) / 2 : 0);
    const double bytes = (double)P.nx * P.ny * P.nz * P.nch * per * sizeof(float);
    return bytes > STREAM_HINT_MB * 1048576.0;
}

template <int DIM, int CPL, bool TGV, int MC, bool ZH>
static void dual_single_launch_zh(const ProbDev& P, dim3 grd, int tiles_x, int tiles_y,
                                  hipStream_t st)
{
    if (stream_hints(P))
        hipLaunchKernelGGL((dual_single_kernel<DIM, CPL, TGV, MC, ZH, true>), grd,
                           dim3(DUAL_BX * DUAL_BY), 0, st, P, tiles_x, tiles_y);
    else
        hipLaunchKernelGGL((dual_single_kernel<DIM, CPL, TGV, MC, ZH, false>), grd,
                           dim3(DUAL_BX * DUAL_BY), 0, st, P, tiles_x, tiles_y);
}

template <int DIM, int CPL, bool TGV, int MC>
static void dual_single_launch(const ProbDev& P, hipStream_t st)
{
    const int tiles_x = (P.nx + DUAL_BX - 1) / DUAL_BX;
    const int tiles_y = (P.ny + DUAL_BY - 1) / DUAL_BY;
    const dim3 grd((unsigned)((long long)tiles_x * tiles_y * P.nz));
    if (DIM == 3 && (P.halo_ubar_hi || P.halo_v_lo))
        dual_single_launch_zh<DIM, CPL, TGV, MC, DIM == 3>(P, grd, tiles_x, tiles_y, st);
    else
        dual_single_launch_zh<DIM, CPL, TGV, MC, false>(P, grd, tiles_x, tiles_y, st);
}

template <int DIM, int CPL, bool TGV, bool ZH>
static void dual_wide_launch_zh(const ProbDev& P, dim3 grd, int tiles_x, int tiles_y,
                                hipStream_t st)
{
    if (stream_hints(P))
        hipLaunchKernelGGL((dual_wide_kernel<DIM, CPL, TGV, ZH, true>), grd,
                           dim3(DUAL_BX * DUAL_BY), 0, st, P, tiles_x, tiles_y);
    else
        hipLaunchKernelGGL((dual_wide_kernel<DIM, CPL, TGV, ZH, false>), grd,
                           dim3(DUAL_BX * DUAL_BY), 0, st, P, tiles_x, tiles_y);
}

template <int DIM, int CPL, bool TGV>
static void dual_wide_launch(const ProbDev& P, hipStream_t st)
{
    const int tiles_x = (P.nx + DUAL_BX - 1) / DUAL_BX;
    const int tiles_y = (P.ny + DUAL_BY - 1) / DUAL_BY;
    const dim3 grd((unsigned)((long long)tiles_x * tiles_y * P.nz));
    if (DIM == 3 && (P.halo_ubar_hi || P.halo_v_lo))
        dual_wide_launch_zh<DIM, CPL, TGV, DIM == 3>(P, grd, tiles_x, tiles_y, st);
    else
        dual_wide_launch_zh<DIM, CPL, TGV, false>(P, grd, tiles_x, tiles_y, st);
}

// Primal step of iteration k fused with the dual step of iteration k + 1:
// reads the previous iterate (u, p, v, w) from the *_in buffers and K^T lam
// from P.ktl, writes u, ubar (P.ubar), p, v, w of the next state to the
// *_out buffers.  Inputs and outputs are distinct arrays, so neighbours can
// be evaluated from unmodified inputs.
template <int DIM, int CPL, bool TGV, int MAXC>
__global__ __launch_bounds__(DUAL_BX* DUAL_BY) void pd_fused_kernel(ProbDev P, FusedBufs B,
                                                                    int tiles_x, int tiles_y)
{
    const int b = blockIdx.x;
    const int tx = b % tiles_x, rest = b / tiles_x;
    const int x = tx * DUAL_BX + (int)(threadIdx.x & (DUAL_BX - 1));
    const int y = (rest % tiles_y) * DUAL_BY + (int)(threadIdx.x / DUAL_BX);
    const int z = rest / tiles_y;
    const size_t cs = (size_t)P.nx * P.ny * P.nz * P.nch;
    const FusedSrc<DIM, TGV> src{B.u_in, B.p_in, B.v_in, B.w_in, P.ktl, B.u_out, P.ubar,
                                 B.p_out, P.nx, P.ny, P.nz, cs, P.hinv, P.tau, P.norming};
    dual_core<DIM, CPL, TGV, MAXC>(P, x, y, z, src, B.v_in, B.w_in, B.v_out, B.w_out);
}

// LDS-tiled, z-marching form of the fused primal + next-dual step (3D).  A
// work-group owns a 64 x 4 column of voxels over a run of FZ_RUN planes.
// At plane z it has ubar and pbar of plane z in LDS (own values plus the +x/+y
// halo of ubar and the -x/-y halo of pbar, evaluated by the first 136
// threads), ubar of plane z + 1 and pbar of plane z - 1 in registers; it
// evaluates the primal step of plane z + 1 (stored: the new u, ubar, p), then
// the dual step of plane z from LDS and registers.  Each voxel's primal step
// is stored once; the halo and the planes next to a run are recomputed.
constexpr int FZ_RUN = 8;
constexpr int FZ_W = DUAL_BX + 1;
constexpr int FZ_H = DUAL_BY + 1;
// the halo of a tile is loaded with one thread per halo entry; smaller tiles
// (measurement builds) run the 3D fused step with pd_fused_kernel instead
constexpr bool FZ_FITS = 2 * (DUAL_BX + DUAL_BY) <= DUAL_BX * DUAL_BY;

template <int MAXC, bool TGV>
struct FzPlane {
    float u[MAXC][FZ_H][FZ_W];              // [ly][lx] own, [ly][64] +x, [4][lx] +y
    float p[MAXC][TGV ? 3 : 1][FZ_H][FZ_W]; // [ly+1][lx+1] own, [ly+1][0] -x, [0][lx+1] -y
};

template <int MAXC, bool TGV>
struct FzSrc {
    static constexpr bool lanex = false;
    const FzPlane<MAXC, TGV>* sh;
    float un[MAXC];       // ubar at (x, y, z + 1)
    float pp[MAXC][3];    // pbar at (x, y, z - 1)
    int lx, ly, x, z;
    __device__ __forceinline__ void own(int c, size_t, int, int, int, float& ub,
                                        float (&pb)[3]) const
    {
        ub = sh->u[c][ly][lx];
        if constexpr (TGV) {
#pragma unroll
            for (int k = 0; k < 3; ++k) pb[k] = sh->p[c][k][ly + 1][lx + 1];
        }
    }
    __device__ __forceinline__ float ubn(int c, size_t, int xx, int, int zz) const
    {
        if (zz != z) return un[c];
        return xx != x ? sh->u[c][ly][lx + 1] : sh->u[c][ly + 1][lx];
    }
    __device__ __forceinline__ void pbn(int c, size_t, int xx, int, int zz, float (&pb)[3]) const
    {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            if (zz != z) pb[k] = pp[c][k];
            else if (xx != x) pb[k] = sh->p[c][TGV ? k : 0][ly + 1][lx];
            else pb[k] = sh->p[c][TGV ? k : 0][ly][lx + 1];
        }
    }
};

template <int CPL, bool TGV, int MAXC>
__global__ __launch_bounds__(DUAL_BX* DUAL_BY) void pd_fused_z_kernel(ProbDev P, FusedBufs B,
                                                                      int tiles_x, int tiles_y)
{
    __shared__ FzPlane<MAXC, TGV> sh;
    const int tid = threadIdx.x;
    const int lx = tid & (DUAL_BX - 1), ly = tid / DUAL_BX;
    int b = blockIdx.x;
    const int tx = b % tiles_x;
    b /= tiles_x;
    const int ty = b % tiles_y;
    const int zr = b / tiles_y;
    const int x0 = tx * DUAL_BX, y0 = ty * DUAL_BY;
    const int x = x0 + lx, y = y0 + ly;
    const int nx = P.nx, ny = P.ny, nz = P.nz, nch = P.nch;
    const int zs = zr * FZ_RUN, ze = min(nz, zs + FZ_RUN);
    const bool in = x < nx && y < ny;
    const size_t plane = (size_t)nx * ny, chs = plane * nz, cs = chs * nch;
    const FusedSrc<3, TGV> fs{B.u_in, B.p_in, B.v_in, B.w_in, P.ktl, B.u_out, P.ubar,
                              B.p_out, nx, ny, nz, cs, P.hinv, P.tau, P.norming};
    auto at = [&](int xx, int yy, int zz, int c) {
        return (size_t)c * chs + (size_t)zz * plane + (size_t)yy * nx + xx;
    };
    // halo of plane zz: ubar at +x / +y, pbar at -x / -y of the tile
    auto halo = [&](int zz) {
        if (tid < DUAL_BY) {
            const int hx = x0 + DUAL_BX, hy = y0 + tid;
            if (hx < nx && hy < ny) {
#pragma unroll
                for (int c = 0; c < MAXC; ++c)
                    if (c < nch) sh.u[c][tid][DUAL_BX] = fs.ubn(c, at(hx, hy, zz, c), hx, hy, zz);
            }
        } else if (tid < DUAL_BY + DUAL_BX) {
            const int r = tid - DUAL_BY, hx = x0 + r, hy = y0 + DUAL_BY;
            if (hx < nx && hy < ny) {
#pragma unroll
                for (int c = 0; c < MAXC; ++c)
                    if (c < nch) sh.u[c][DUAL_BY][r] = fs.ubn(c, at(hx, hy, zz, c), hx, hy, zz);
            }
        } else if (TGV && tid < 2 * DUAL_BY + DUAL_BX) {
            const int r = tid - DUAL_BY - DUAL_BX, hx = x0 - 1, hy = y0 + r;
            if (hx >= 0 && hy < ny) {
#pragma unroll
                for (int c = 0; c < MAXC; ++c) {
                    if (c >= nch) break;
                    float q[3];
                    fs.pbn(c, at(hx, hy, zz, c), hx, hy, zz, q);
#pragma unroll
                    for (int k = 0; k < 3; ++k) sh.p[c][TGV ? k : 0][r + 1][0] = q[k];
                }
            }
        } else if (TGV && tid < 2 * DUAL_BY + 2 * DUAL_BX) {
            const int r = tid - 2 * DUAL_BY - DUAL_BX, hx = x0 + r, hy = y0 - 1;
            if (hy >= 0 && hx < nx) {
#pragma unroll
                for (int c = 0; c < MAXC; ++c) {
                    if (c >= nch) break;
                    float q[3];
                    fs.pbn(c, at(hx, hy, zz, c), hx, hy, zz, q);
#pragma unroll
                    for (int k = 0; k < 3; ++k) sh.p[c][TGV ? k : 0][0][r + 1] = q[k];
                }
            }
        }
    };

    FzSrc<MAXC, TGV> src;
    src.sh = &sh;
    src.lx = lx;
    src.ly = ly;
    src.x = x;
#pragma unroll
    for (int c = 0; c < MAXC; ++c) {
        src.un[c] = 0.f;
#pragma unroll
        for (int k = 0; k < 3; ++k) src.pp[c][k] = 0.f;
    }
    // prologue: primal step of plane zs (stored), pbar of plane zs - 1
    if (in) {
#pragma unroll
        for (int c = 0; c < MAXC; ++c) {
            if (c >= nch) break;
            float ub, pb[3] = {0.f, 0.f, 0.f};
            fs.own(c, at(x, y, zs, c), x, y, zs, ub, pb);
            sh.u[c][ly][lx] = ub;
            if constexpr (TGV) {
#pragma unroll
                for (int k = 0; k < 3; ++k) sh.p[c][k][ly + 1][lx + 1] = pb[k];
                if (zs > 0) fs.pbn(c, at(x, y, zs - 1, c), x, y, zs - 1, src.pp[c]);
            }
        }
    }
    halo(zs);
    for (int z = zs; z < ze; ++z) {
        float pn[MAXC][3];
        const bool next_own = z + 1 < ze;
        if (in && z + 1 < nz) {
#pragma unroll
            for (int c = 0; c < MAXC; ++c) {
                if (c >= nch) break;
                const size_t j = at(x, y, z + 1, c);
                if (next_own) fs.own(c, j, x, y, z + 1, src.un[c], pn[c]);
                else src.un[c] = fs.ubn(c, j, x, y, z + 1);
            }
        }
        __syncthreads();  // plane z complete in LDS
        if (in) {
            src.z = z;
            dual_core<3, CPL, TGV, MAXC>(P, x, y, z, src, B.v_in, B.w_in, B.v_out, B.w_out);
        }
        __syncthreads();  // plane z no longer read
        if (next_own) {
            if (in) {
#pragma unroll
                for (int c = 0; c < MAXC; ++c) {
                    if (c >= nch) break;
                    sh.u[c][ly][lx] = src.un[c];
                    if constexpr (TGV) {
#pragma unroll
                        for (int k = 0; k < 3; ++k) {
                            src.pp[c][k] = sh.p[c][k][ly + 1][lx + 1];
                            sh.p[c][k][ly + 1][lx + 1] = pn[c][k];
                        }
                    }
                }
            }
            halo(z + 1);
        }
    }
}

template <int CPL, bool TGV, int MAXC>
static void fused_z_launch(const ProbDev& P, const FusedBufs& B, hipStream_t st)
{
    const int tiles_x = (P.nx + DUAL_BX - 1) / DUAL_BX;
    const int tiles_y = (P.ny + DUAL_BY - 1) / DUAL_BY;
    const int runs = (P.nz + FZ_RUN - 1) / FZ_RUN;
    hipLaunchKernelGGL((pd_fused_z_kernel<CPL, TGV, MAXC>),
                       dim3((unsigned)(tiles_x * tiles_y * runs)), dim3(DUAL_BX * DUAL_BY), 0, st,
                       P, B, tiles_x, tiles_y);
}

template <int DIM, int CPL, bool TGV, int MAXC>
static void fused_launch(const ProbDev& P, const FusedBufs& B, hipStream_t st)
{
    if constexpr (DIM == 3 && FZ_FITS) {
        fused_z_launch<CPL, TGV, MAXC>(P, B, st);
        return;
    }
    const int tiles_x = (P.nx + DUAL_BX - 1) / DUAL_BX;
    const int tiles_y = (P.ny + DUAL_BY - 1) / DUAL_BY;
    hipLaunchKernelGGL((pd_fused_kernel<DIM, CPL, TGV, MAXC>),
                       dim3((unsigned)(tiles_x * tiles_y * P.nz)), dim3(DUAL_BX * DUAL_BY), 0, st,
                       P, B, tiles_x, tiles_y);
}

template <int DIM, int CPL, bool TGV>
static void fused_c(const ProbDev& P, const FusedBufs& B, hipStream_t st)
{
    if (P.nch <= 1) fused_launch<DIM, CPL, TGV, 1>(P, B, st);
    else if (P.nch <= 2) fused_launch<DIM, CPL, TGV, 2>(P, B, st);
    else if (P.nch <= 4) fused_launch<DIM, CPL, TGV, 4>(P, B, st);
    else fused_launch<DIM, CPL, TGV, 8>(P, B, st);
}

template <int DIM, int CPL>
static void fused_t(const ProbDev& P, const FusedBufs& B, hipStream_t st)
{
    if (P.tgv) fused_c<DIM, CPL, true>(P, B, st);
    else fused_c<DIM, CPL, false>(P, B, st);
}

template <int DIM>
static void fused_d(const ProbDev& P, const FusedBufs& B, hipStream_t st)
{
    if (P.coupling == GR_COUPLING_UNCORRELATED) fused_t<DIM, GR_COUPLING_UNCORRELATED>(P, B, st);
    else if (P.coupling == GR_COUPLING_FROBENIUS) fused_t<DIM, GR_COUPLING_FROBENIUS>(P, B, st);
    else fused_t<DIM, GR_COUPLING_NUCLEAR>(P, B, st);
}

int launch_fused_pd(const ProbDev& P, const FusedBufs& B, hipStream_t st)
{
    if (P.nch > 8) {
        set_error("fused step: %d channels exceed the 8 the fused kernel holds per voxel", P.nch);
        return GR_ERR_UNSUPPORTED;
    }
    const long long blocks = (long long)((P.nx + DUAL_BX - 1) / DUAL_BX) *
                             ((P.ny + DUAL_BY - 1) / DUAL_BY) * P.nz;
    if (blocks > 0x7fffffff) {
        set_error("fused step: grid too large");
        return GR_ERR_INVALID;
    }
    if (P.dim == 2) fused_d<2>(P, B, st);
    else fused_d<3>(P, B, st);
    GR_LAUNCH_CHECK("pd_fused_kernel");
    return GR_OK;
}

template <int DIM, int CPL, bool TGV>
static void dual_single_c(const ProbDev& P, hipStream_t st)
{
    if (P.nch <= 1) dual_single_launch<DIM, CPL, TGV, 1>(P, st);
    else if (P.nch <= 2) dual_single_launch<DIM, CPL, TGV, 2>(P, st);
    else if (P.nch <= 4) dual_single_launch<DIM, CPL, TGV, 4>(P, st);
    else if (P.nch <= 8) dual_single_launch<DIM, CPL, TGV, 8>(P, st);
    else dual_wide_launch<DIM, CPL, TGV>(P, st);
}

template <int DIM, int CPL>
static void dual_single_t(const ProbDev& P, hipStream_t st)
{
    if (P.tgv) dual_single_c<DIM, CPL, true>(P, st);
    else dual_single_c<DIM, CPL, false>(P, st);
}

template <int DIM>
static void dual_single_d(const ProbDev& P, hipStream_t st)
{
    if (P.coupling == GR_COUPLING_UNCORRELATED) dual_single_t<DIM, GR_COUPLING_UNCORRELATED>(P, st);
    else if (P.coupling == GR_COUPLING_FROBENIUS) dual_single_t<DIM, GR_COUPLING_FROBENIUS>(P, st);
    else dual_single_t<DIM, GR_COUPLING_NUCLEAR>(P, st);
}

int launch_dual(const ProbDev& P, hipStream_t st)
{
    const long long blocks = (long long)((P.nx + DUAL_BX - 1) / DUAL_BX) *
                             ((P.ny + DUAL_BY - 1) / DUAL_BY) * P.nz;
    if (blocks > 0x7fffffff) {
        set_error("dual step: grid too large");
        return GR_ERR_INVALID;
    }
    if (P.dim == 2) dual_single_d<2>(P, st);
    else dual_single_d<3>(P, st);
    GR_LAUNCH_CHECK("dual_single_kernel");
    return GR_OK;
}

int launch_dual_batch(const ProbDev* probs, const DualItem* items, int nitems, int dim,
                      int coupling, bool tgv, int maxc, hipStream_t st)
{
    if (nitems <= 0) return GR_OK;
    dim3 grd((unsigned)nitems), blk(DUAL_BX * DUAL_BY);
#define GR_DB(DD, CC, TT, MM) \
    hipLaunchKernelGGL((dual_batch_kernel<DD, CC, TT, MM>), grd, blk, 0, st, probs, items)
#define GR_DB_M(DD, CC, TT)                                                                 \
    do {                                                                                    \
        if (maxc <= 4) GR_DB(DD, CC, TT, 4);                                                \
        else if (maxc <= 8) GR_DB(DD, CC, TT, 8);                                           \
        else hipLaunchKernelGGL((dual_wide_batch_kernel<DD, CC, TT>), grd, blk, 0, st, probs, \
                                items);                                                     \
    } while (0)
#define GR_DB_T(DD, CC) \
    do { if (tgv) GR_DB_M(DD, CC, true); else GR_DB_M(DD, CC, false); } while (0)
#define GR_DB_C(DD)                                                        \
    do {                                                                   \
        if (coupling == GR_COUPLING_UNCORRELATED) GR_DB_T(DD, GR_COUPLING_UNCORRELATED); \
        else if (coupling == GR_COUPLING_FROBENIUS) GR_DB_T(DD, GR_COUPLING_FROBENIUS);  \
        else GR_DB_T(DD, GR_COUPLING_NUCLEAR);                             \
    } while (0)
    if (dim == 2) GR_DB_C(2); else GR_DB_C(3);
#undef GR_DB_C
#undef GR_DB_T
#undef GR_DB_M
#undef GR_DB
    GR_LAUNCH_CHECK("dual_batch_kernel");
    return GR_OK;
}

template <int DIM, int CPL, bool TGV>
static void reg_launch(const ProbDev& P, double a0, double a1, double* partial, int nb,
                       hipStream_t st)
{
    hipLaunchKernelGGL((reg_energy_kernel<DIM, CPL, TGV>), dim3(nb), dim3(EN_THREADS), 0, st,
                       P.u, P.p, P.nx, P.ny, P.nz, P.nch, (double)P.hinv, a0, a1, partial);
}

// block partials of the regulariser: EN_REG_BLOCKS doubles
static void reg_partials(const ProbDev& P, double a0, double a1, double* partial, hipStream_t st)
{
    const int nb = EN_REG_BLOCKS;
#define GR_REG(DM, CP)                                                   \
    do {                                                                 \
        if (P.tgv) reg_launch<DM, CP, true>(P, a0, a1, partial, nb, st); \
        else reg_launch<DM, CP, false>(P, a0, a1, partial, nb, st);      \
    } while (0)
    if (P.dim == 2) {
        if (P.coupling == GR_COUPLING_UNCORRELATED) GR_REG(2, GR_COUPLING_UNCORRELATED);
        else if (P.coupling == GR_COUPLING_FROBENIUS) GR_REG(2, GR_COUPLING_FROBENIUS);
        else GR_REG(2, GR_COUPLING_NUCLEAR);
    } else {
        if (P.coupling == GR_COUPLING_UNCORRELATED) GR_REG(3, GR_COUPLING_UNCORRELATED);
        else if (P.coupling == GR_COUPLING_FROBENIUS) GR_REG(3, GR_COUPLING_FROBENIUS);
        else GR_REG(3, GR_COUPLING_NUCLEAR);
    }
#undef GR_REG
}

// block partials of the data term: EN_DATA_BLOCKS doubles per channel
static void data_partials(const float* ku, const float* f, int nch, size_t per_ch, int disc,
                          double* partial, hipStream_t st)
{
    hipLaunchKernelGGL(data_energy_kernel, dim3(EN_DATA_BLOCKS, nch), dim3(EN_THREADS), 0, st, ku,
                       f, per_ch, disc, partial);
}

// The host's final sums, shared by both entry points so that they agree bit
// for bit: the regulariser's partials in block order; the data term as
// sum_c mu_c (partials of channel c in block order).
static double sum_reg(const double* h)
{
    double s = 0.0;
    for (int b = 0; b < EN_REG_BLOCKS; ++b) s += h[b];
    return s;
}

static double sum_data(const double* chan_sums, const double* mu, int nch)
{
    double s = 0.0;
    for (int c = 0; c < nch; ++c) s += mu[c] * chan_sums[c];
    return s;
}

int launch_energy(const ProbDev& P, double a0, double a1, double* out_reg, hipStream_t st)
{
    double* partial = nullptr;
    if (hipMalloc(&partial, sizeof(double) * EN_REG_BLOCKS) != hipSuccess) {
        set_error("energy: device allocation failed");
        return GR_ERR_OOM;
    }
    reg_partials(P, a0, a1, partial, st);
    std::vector<double> h(EN_REG_BLOCKS);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess)
        e = hipMemcpyAsync(h.data(), partial, sizeof(double) * EN_REG_BLOCKS,
                           hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    (void)hipFree(partial);
    if (e != hipSuccess) return hip_status(e, "energy");
    *out_reg = sum_reg(h.data());
    return GR_OK;
}

int launch_data_energy(const float* ku, const float* f, const double* mu, int nch,
                       size_t per_ch, int disc, double* out, hipStream_t st)
{
    const int nb = EN_DATA_BLOCKS;
    double* partial = nullptr;
    if (hipMalloc(&partial, sizeof(double) * nb * nch) != hipSuccess) {
        set_error("energy: device allocation failed");
        return GR_ERR_OOM;
    }
    data_partials(ku, f, nch, per_ch, disc, partial, st);
    std::vector<double> h((size_t)nb * nch);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess)
        e = hipMemcpyAsync(h.data(), partial, sizeof(double) * nb * nch, hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    (void)hipFree(partial);
    if (e != hipSuccess) return hip_status(e, "data energy");
    std::vector<double> sc(nch, 0.0);
    for (int c = 0; c < nch; ++c)
        for (int b = 0; b < nb; ++b) sc[c] += h[(size_t)c * nb + b];
    *out = sum_data(sc.data(), mu, nch);
    return GR_OK;
}

// ---- allocation-free energy -------------------------------------------------
//
// The block partials and their per-channel sums live in `scratch`, a fixed
// array of this library on every device it runs on (held until the process
// ends).  Its host-side name is no device address: the address is taken once
// per device with hipGetSymbolAddress, with that device current, and only that
// pointer is handed to kernels and copies.  Every slot a call reads is written
// earlier in the same call (each block writes its partial, one thread per sum
// writes a sum), so nothing needs clearing; calls that overlap on one device
// share the array and race.
namespace {

constexpr size_t EN_NOALLOC_REG = 0;                                      // EN_REG_BLOCKS partials
constexpr size_t EN_NOALLOC_DATA = EN_NOALLOC_REG + EN_REG_BLOCKS;        // nb per channel
constexpr size_t EN_NOALLOC_SUMS =
    EN_NOALLOC_DATA + (size_t)EN_DATA_BLOCKS * GR_ENERGY_NOALLOC_MAX_CHANNELS;  // 1 + nch sums
constexpr size_t EN_NOALLOC_DOUBLES = EN_NOALLOC_SUMS + 1 + GR_ENERGY_NOALLOC_MAX_CHANNELS;

}  // namespace

// 9,249 doubles (72.3 KiB) per device
static __device__ __align__(256) double scratch[EN_NOALLOC_DOUBLES];

namespace {

// sums[0] = regulariser partials in block order; sums[1 + c] = data partials
// of channel c in block order (the order of the host sums above)
__global__ __launch_bounds__(64) void energy_sums_kernel(const double* __restrict__ reg,
                                                         const double* __restrict__ data, int nb,
                                                         int nch, double* __restrict__ sums)
{
    const int t = threadIdx.x;
    if (t == 0) {
        double s = 0.0;
        for (int b = 0; b < EN_REG_BLOCKS; ++b) s += reg[b];
        sums[0] = s;
    } else if (t <= nch) {
        const double* d = data + (size_t)(t - 1) * nb;
        double s = 0.0;
        for (int b = 0; b < nb; ++b) s += d[b];
        sums[t] = s;
    }
}

constexpr int EN_MAX_DEVICES = 64;

double* noalloc_scratch(int* rc)
{
    static std::mutex mtx;
    static double* ptr[EN_MAX_DEVICES] = {};
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) {
        *rc = hip_status(e, "energy_noalloc: current device");
        return nullptr;
    }
    if (dev < 0 || dev >= EN_MAX_DEVICES) {
        set_error("energy_noalloc: device %d beyond the %d this library tracks", dev,
                  EN_MAX_DEVICES);
        *rc = GR_ERR_UNSUPPORTED;
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(mtx);
    if (!ptr[dev]) {
        void* p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            *rc = hip_status(e != hipSuccess ? e : hipErrorInvalidSymbol,
                             "energy_noalloc: address of the scratch array");
            return nullptr;
        }
        ptr[dev] = static_cast<double*>(p);
    }
    return ptr[dev];
}

}  // namespace

int launch_energy_noalloc(const ProbDev& P, const double* mu, double a0, double a1,
                          size_t per_ch, double* out_data, double* out_reg, hipStream_t st)
{
    if (P.nch > GR_ENERGY_NOALLOC_MAX_CHANNELS) {
        set_error("energy_noalloc: %d channels exceed the %d its fixed scratch holds", P.nch,
                  GR_ENERGY_NOALLOC_MAX_CHANNELS);
        return GR_ERR_INVALID;
    }
    int rc = GR_OK;
    double* base = noalloc_scratch(&rc);
    if (!base) return rc;
    double* reg = base + EN_NOALLOC_REG;
    double* data = base + EN_NOALLOC_DATA;
    double* sums = base + EN_NOALLOC_SUMS;
    data_partials(P.ku, P.f, P.nch, per_ch, P.disc, data, st);
    reg_partials(P, a0, a1, reg, st);
    hipLaunchKernelGGL(energy_sums_kernel, dim3(1), dim3(64), 0, st, reg, data, EN_DATA_BLOCKS,
                       P.nch, sums);
    double h[1 + GR_ENERGY_NOALLOC_MAX_CHANNELS];
    hipError_t e = hipGetLastError();
    if (e == hipSuccess)
        e = hipMemcpyAsync(h, sums, sizeof(double) * (1 + P.nch), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return hip_status(e, "energy_noalloc");
    *out_reg = h[0];
    *out_data = sum_data(h + 1, mu, P.nch);
    return GR_OK;
}

}  // namespace gr
