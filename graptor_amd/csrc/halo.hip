// Halo planes of a z-slab: the planes one slab of a z-sharded volume sends to
// its neighbours, packed into the layout of the receiver's gr_tgv_state halo
// fields ([component][channel][y][x], contiguous).  A rank sends the packed
// buffer as one message and receives straight into its halo fields, which
// the dual and primal kernels read in place, so there is no unpack step.
// No counterpart in the reference, which solves slabs independently
// (RC:408-458).
#include "common.h"

#include <algorithm>
#include <cstdint>
#include <utility>

namespace gr {

namespace {

// One packed message: planes `comp(k)` of `src` at slice z, for every
// channel, to out[k][c][plane].  Components are 4-bit fields of `comps`
// (component k at bits 4k..4k+3), so the kernel indexes no array.
struct HaloMsg {
    const float* src;
    float* out;
    int comps;
    int ncomp;
    int z;
};

constexpr int HALO_THREADS = 256;

// blockIdx.y = message slot (k, c) over both messages; blockIdx.x strides
// over the plane.  V4: the plane is copied as float4 (plane % 4 == 0 and
// every base pointer 16-byte aligned, checked by the host).
template <bool V4>
__global__ void __launch_bounds__(HALO_THREADS)
halo_pack_kernel(HaloMsg a, HaloMsg b, int nch, int nz, size_t plane, size_t cs)
{
    const int slot = (int)blockIdx.y;
    const int na = a.ncomp * nch;
    const bool first = slot < na;
    const float* src = first ? a.src : b.src;
    float* out = first ? a.out : b.out;
    const int comps = first ? a.comps : b.comps;
    const int z = first ? a.z : b.z;
    const int s = first ? slot : slot - na;
    const int k = s / nch, c = s - k * nch;
    const int comp = (comps >> (4 * k)) & 15;
    const float* from = src + (size_t)comp * cs + ((size_t)c * nz + z) * plane;
    float* to = out + (size_t)s * plane;
    const size_t stride = (size_t)gridDim.x * HALO_THREADS;
    if constexpr (V4) {
        const size_t n4 = plane / 4;
        const float4* f4 = reinterpret_cast<const float4*>(from);
        float4* t4 = reinterpret_cast<float4*>(to);
        for (size_t i = (size_t)blockIdx.x * HALO_THREADS + threadIdx.x; i < n4; i += stride)
            t4[i] = f4[i];
    } else {
        for (size_t i = (size_t)blockIdx.x * HALO_THREADS + threadIdx.x; i < plane; i += stride)
            to[i] = from[i];
    }
}

bool aligned16(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; }

}  // namespace

}  // namespace gr

using namespace gr;

extern "C" int gr_tgv_pack_halos(const gr_tgv_params* p, const gr_tgv_state* s, int phase,
                                 float* to_below, float* to_above, void* stream)
{
    if (!p || !s || p->dim != 3 || p->nx <= 0 || p->ny <= 0 || p->nz <= 0 || p->nch <= 0 ||
        p->nch > 4096 || (phase != GR_HALO_DUAL && phase != GR_HALO_PRIMAL)) {
        set_error("tgv_pack_halos: invalid parameters (3D only, phase GR_HALO_DUAL / _PRIMAL)");
        return GR_ERR_INVALID;
    }
    const bool tgv = p->regularisation == GR_REG_TGV;
    HaloMsg lo{}, hi{};
    if (phase == GR_HALO_DUAL) {
        // to the slab below: ubar of plane 0 (its halo_ubar_hi);
        // to the slab above: pbar of the last plane (its halo_pbar_lo)
        lo = HaloMsg{s->ubar, to_below, 0, 1, 0};
        hi = HaloMsg{s->pbar, to_above, 0x210, 3, p->nz - 1};
        if (to_below && !s->ubar) { set_error("tgv_pack_halos: ubar missing"); return GR_ERR_INVALID; }
        if (to_above && (!tgv || !s->pbar)) {
            set_error("tgv_pack_halos: the dual-phase message upwards is pbar (TGV only)");
            return GR_ERR_INVALID;
        }
    } else {
        // to the slab below: w_xz, w_yz, w_zz of plane 0 (its halo_w_hi);
        // to the slab above: v_z of the last plane (its halo_v_lo)
        lo = HaloMsg{s->w, to_below, 0x254, 3, 0};
        hi = HaloMsg{s->v, to_above, 2, 1, p->nz - 1};
        if (to_below && (!tgv || !s->w)) {
            set_error("tgv_pack_halos: the primal-phase message downwards is w (TGV only)");
            return GR_ERR_INVALID;
        }
        if (to_above && !s->v) { set_error("tgv_pack_halos: v missing"); return GR_ERR_INVALID; }
    }
    if (!to_below) lo.ncomp = 0;
    if (!to_above) hi.ncomp = 0;
    const int slots = (lo.ncomp + hi.ncomp) * p->nch;
    if (slots == 0) return GR_OK;
    const size_t plane = (size_t)p->nx * p->ny;
    const size_t cs = plane * p->nz * p->nch;
    // an empty message goes second, so a slot index never reaches it
    if (lo.ncomp == 0) std::swap(lo, hi);
    const bool v4 = plane % 4 == 0 && aligned16(lo.src) && aligned16(lo.out) &&
                    (hi.ncomp == 0 || (aligned16(hi.src) && aligned16(hi.out)));
    const size_t per_block = (size_t)HALO_THREADS * (v4 ? 4 : 1);
    const unsigned bx = (unsigned)std::min<size_t>(1024, (plane + per_block - 1) / per_block);
    const hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    if (v4)
        hipLaunchKernelGGL(halo_pack_kernel<true>, dim3(bx, (unsigned)slots), dim3(HALO_THREADS),
                           0, st, lo, hi, p->nch, p->nz, plane, cs);
    else
        hipLaunchKernelGGL(halo_pack_kernel<false>, dim3(bx, (unsigned)slots), dim3(HALO_THREADS),
                           0, st, lo, hi, p->nch, p->nz, plane, cs);
    GR_LAUNCH_CHECK("halo_pack_kernel");
    return GR_OK;
}
