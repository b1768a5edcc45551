// Primal update of one voxel-channel, shared by every kernel that performs
// it (standalone primal step, backprojection epilogue, and the fused
// primal + next-dual kernel, which also evaluates it at neighbours to form
// the extrapolated iterates there).  One definition keeps the arithmetic of
// all of them the same.
//
// u+ = max(0, u + tau (div v - norming * ktl))          (3D:706-729)
// p+ = p + tau (v + div2 w)                              (3D:734-757)
// ubar = 2 u+ - u, pbar = 2 p+ - p                       (3D:1075-1076)
// div: backward differences with the reference's boundary rows; div2: the
// adjoint of the backward-difference symmetrised gradient under the pairing
// that counts off-diagonals twice.  i indexes u[c][z][y][x]; cs is the
// component stride of v, w, p.
#pragma once

#include "common.h"

namespace gr {

// Values of the neighbouring lanes of a wave (wave-wide DPP shifts by one
// lane); the lane at the wave's edge receives 0 and must take its value from
// memory.
__device__ __forceinline__ float lane_prev(float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x138 /* wave_shr:1 */,
                                                      0xf, 0xf, false));
}
__device__ __forceinline__ float lane_next(float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x130 /* wave_shl:1 */,
                                                      0xf, 0xf, false));
}

// LANEX: the 64 lanes of the calling wave hold 64 consecutive x of one row
// (lane = x - x0, every lane below the last active one active), so the x - 1
// and x + 1 neighbours of v and w come from the neighbouring lanes and only
// the wave's edge lanes read them from memory.
template <int DIM, bool TGV, bool NEED_U, bool NEED_P, bool LANEX = false, bool ZH = false,
          bool NT = false>
__device__ __forceinline__ void primal_point(
    const float* __restrict__ u, const float* __restrict__ p, const float* __restrict__ v,
    const float* __restrict__ w, float ktl, int nx, int ny, int nz, size_t cs, float hinv,
    float tau, float norming, size_t i, int x, int y, int z, float& un, float& ub,
    float (&pn)[3], float (&pb)[3], const ProbDev* zh = nullptr)
{
    const size_t plane = (size_t)nx * ny;
    // slab of a larger volume (zh's halo planes): planes exist beyond z = 0 /
    // z = nz - 1, which then are not the volume's boundary
    const bool has_lo = ZH && zh->halo_v_lo, has_hi = ZH && zh->halo_ubar_hi;
    // offset of (component comp, this channel, y, x) in a halo plane set
    const size_t hc = ZH ? (i / plane) / (size_t)nz : 0;  // channel of i = [c][z][y][x]
    auto halo_off = [&](int comp) {
        return ((size_t)comp * zh->nch + hc) * plane + (size_t)y * nx + x;
    };
    float vc[3];
#pragma unroll
    for (int k = 0; k < DIM; ++k) vc[k] = v[i + k * cs];
    const int lane = (int)(threadIdx.x & 63);
    if constexpr (NEED_U) {
        float d0 = (x == nx - 1) ? 0.f : vc[0];
        if constexpr (LANEX) {
            float vm = lane_prev(vc[0]);
            if (lane == 0 && x > 0) vm = v[i - 1];
            if (x > 0) d0 -= vm;
        } else {
            if (x > 0) d0 -= v[i - 1];
        }
        float d1 = (y == ny - 1) ? 0.f : vc[1];
        if (y > 0) d1 -= v[i + cs - nx];
        float dv = d0 + d1;
        if constexpr (DIM == 3) {
            float d2 = (z == nz - 1 && !has_hi) ? 0.f : vc[2];
            if constexpr (ZH) {
                const float* dn = z > 0 ? v + i + 2 * cs - plane
                                  : has_lo ? zh->halo_v_lo + halo_off(0) : nullptr;
                if (dn) d2 -= *dn;
            } else {
                if (z > 0) d2 -= v[i + 2 * cs - plane];
            }
            dv += d2 * hinv;
        }
        const float uo = ld_stream<NT>(&u[i]);  // own voxel only
        const float t = uo + tau * (dv - norming * ktl);
        un = t < 0.f ? 0.f : t;
        ub = 2.f * un - uo;
    }
    if constexpr (TGV && NEED_P) {
        float q[3] = {0.f, 0.f, 0.f};
        constexpr int M = DIM == 2 ? 3 : 6;
        float wc[M];
#pragma unroll
        for (int m = 0; m < M; ++m) wc[m] = w[i + m * cs];
        if constexpr (DIM == 2) {
            // x-part uses (w_xx, w_xy), y-part (w_xy, w_yy)
            float x0 = (x == 0) ? 0.f : -wc[0], x1 = (x == 0) ? 0.f : -wc[2];
            if constexpr (LANEX) {
                float n0 = lane_next(wc[0]), n1 = lane_next(wc[2]);
                if (lane == 63 && x < nx - 1) { n0 = w[i + 1]; n1 = w[i + 2 * cs + 1]; }
                if (x < nx - 1) { x0 += n0; x1 += n1; }
            } else {
                if (x < nx - 1) { x0 += w[i + 1]; x1 += w[i + 2 * cs + 1]; }
            }
            float y0 = (y == 0) ? 0.f : -wc[2], y1 = (y == 0) ? 0.f : -wc[1];
            if (y < ny - 1) { y0 += w[i + 2 * cs + nx]; y1 += w[i + cs + nx]; }
            q[0] = x0 + y0;
            q[1] = x1 + y1;
        } else {
            // tensor order (xx, yy, zz, xy, xz, yz): x-part (0,3,4), y-part
            // (3,1,5), z-part (4,5,2)
            const int xi[3] = {0, 3, 4}, yi[3] = {3, 1, 5}, zi[3] = {4, 5, 2};
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                float xs = (x == 0) ? 0.f : -wc[xi[k]];
                if constexpr (LANEX) {
                    float nb = lane_next(wc[xi[k]]);
                    if (lane == 63 && x < nx - 1) nb = w[i + xi[k] * cs + 1];
                    if (x < nx - 1) xs += nb;
                } else {
                    if (x < nx - 1) xs += w[i + xi[k] * cs + 1];
                }
                float ys = (y == 0) ? 0.f : -wc[yi[k]];
                if (y < ny - 1) ys += w[i + yi[k] * cs + nx];
                float zs = (z == 0 && !has_lo) ? 0.f : -wc[zi[k]];
                if constexpr (ZH) {
                    const float* up = z < nz - 1 ? w + i + zi[k] * cs + plane
                                      : has_hi ? zh->halo_w_hi + halo_off(k) : nullptr;
                    if (up) zs += *up;
                } else {
                    if (z < nz - 1) zs += w[i + zi[k] * cs + plane];
                }
                q[k] = xs + ys + zs * hinv;
            }
        }
#pragma unroll
        for (int k = 0; k < DIM; ++k) {
            const float po = ld_stream<NT>(&p[i + k * cs]);
            pn[k] = po + tau * (vc[k] + q[k]);
            pb[k] = 2.f * pn[k] - po;
        }
    }
}

}  // namespace gr
