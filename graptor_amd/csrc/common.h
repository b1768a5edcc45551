// Shared declarations of the graptor_amd HIP library (gfx950 / CDNA4).
#pragma once
#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include <vector>

#include "../../include/graptor_amd.h"

namespace gr {

// Status plumbing: every extern "C" entry point returns an int status and
// records a message retrievable through gr_last_error().
void set_error(const char* fmt, ...);
int hip_status(hipError_t e, const char* where);

#define GR_HIP(call)                                                        \
    do {                                                                    \
        hipError_t e_ = (call);                                             \
        if (e_ != hipSuccess) return ::gr::hip_status(e_, #call);           \
    } while (0)

#define GR_LAUNCH_CHECK(where)                                              \
    do {                                                                    \
        hipError_t e_ = hipGetLastError();                                  \
        if (e_ != hipSuccess) return ::gr::hip_status(e_, where);           \
    } while (0)

// Per-angle constants of the gather forward projector.  For an angle whose
// rays run closer to the x axis (|sin| >= |cos|) the kernel steps along x and
// picks candidate rows; otherwise it steps along y and picks candidate
// columns.  t(x, y) = x*cos + y*sin + off is the detector coordinate.
struct FwdAngle {
    float mstep;  // coefficient of the stepping coordinate in t
    float mcand;  // coefficient of the candidate coordinate in t
    float off;    // offset term of t
    float invc;   // 1 / mcand
    float hc;     // 1 / |mcand|: half-width of the hat support in candidates
    int cls;      // 0: step x / candidates y, 1: step y / candidates x
    int id;       // angle index in the caller's table
    int pad;
};

// One work chunk of the forward projector: angles [a0, a0+na) of one class
// (indices into the class-sorted FwdAngle array) times bins [b0, b0+nb).
struct FwdChunk {
    int cls, a0, na, b0, nb, pad0, pad1, pad2;
};

// Everything a kernel needs to know about one reconstruction problem (or one
// standalone projector call): sizes, derived scalars, geometry tables and
// device state.  Passed by value for single problems and through a device
// table for batches.
struct ProbDev {
    int dim, nx, ny, nz, nch, nd, na;
    int coupling, tgv, disc;
    float hinv, sigma, tau, sigma_p, norming, ainv0, ainv1;
    int fwd_band, fwd_pitch, adj_tile_y, adj_segw, adj_achunk, fuse_primal;
    const float4* ofs;     // na x (cos, sin, off, 1/cos)
    const FwdAngle* fwd;   // na, sorted by class
    // state; for standalone projector calls: ubar = image in, ku = sinogram
    // out (forward), lam = sinogram in, ktl = image out (backprojection)
    float *u, *ubar, *p, *pbar, *v, *w, *lam, *ktl, *ku;
    const float *f, *mfac;
    // z halo of a slab (3D; NULL: boundary plane), planes [c][y][x]:
    // ubar above, pbar below [k][c], v_z below, (w_xz, w_yz, w_zz) above [j][c]
    const float *halo_ubar_hi, *halo_pbar_lo, *halo_v_lo, *halo_w_hi;
};

// Buffers of the fused primal + next-dual step: the previous state is read
// from *_in, the next written to *_out (u, p, v, w alternate between the
// caller's arrays and a second set).
struct FusedBufs {
    const float *u_in, *p_in, *v_in, *w_in;
    float *u_out, *p_out, *v_out, *w_out;
};

struct FwdItem { int prob, cls, a0, na, b0, nb, s0, pad; };
struct AdjItem { int prob, x0, y0, s0; };
// Stencil items: a 64 x 4 tile of plane z (dual step) or of slice
// z + nz * channel (primal step).
struct DualItem { int prob, x0, y0, z; };

// Device-side geometry (owned by gr_geometry).
struct Geometry {
    int nx, ny, nd, na;
    float* ofs_dev;          // na x float4 (cos, sin, off, 1/cos) as given
    FwdAngle* fwd_dev;       // na, sorted by class
    FwdChunk* chunks_dev;    // four chunk tables back to back
    int tab_off[5];          // table t = chunks [tab_off[t], tab_off[t+1]),
                             // built for 8, 4, 2, 1 items per thread
    std::vector<FwdChunk> chunks_host;
    int fwd_band;            // lines per LDS band (0: automatic)
    int adj_tile_y;          // rows per backprojection tile
    int fuse_primal;         // 1: primal step in the backprojection epilogue
};

// Modes of the forward projector epilogue.
// FWD_PROX applies each problem's own discrepancy (batches).
// FWD_BESIDE: the launch shares the GPU with the dual step of the same
// iteration (two-stream schedule): fewer, larger work-groups.
// FWD_AUTOBAND: the geometry's band height is automatic, so a launch beside
// the dual step may size it for its grid.
enum { FWD_WRITE = 1, FWD_KL = 2, FWD_L2 = 4, FWD_PROX = 8, FWD_BESIDE = 16, FWD_AUTOBAND = 32 };

// Loads / stores of arrays a stencil kernel streams once, with the
// non-temporal hint when NT: single problems whose state outgrows the
// Infinity Cache (their arrays would otherwise displace the lines of the
// neighbour reads; stream_hints decides), not problems or batches of small
// problems whose arrays stay in that cache from one phase to the next (the
// hint there: 1.135 against 1.109 ms per batch iteration).
template <bool NT>
__device__ __forceinline__ float ld_stream(const float* p)
{
    if constexpr (NT) return __builtin_nontemporal_load(p);
    else return *p;
}
template <bool NT>
__device__ __forceinline__ void st_stream(float v, float* p)
{
    if constexpr (NT) __builtin_nontemporal_store(v, p);
    else *p = v;
}

constexpr int FWD_THREADS = 256;
constexpr int FWD_PADL = 4;          // zero candidates left of position 0
constexpr int ADJ_THREADS = 256;
constexpr int ADJ_TX = 64;           // backprojection tile width (one wave)
constexpr int ADJ_MAXR = 8;          // rows per thread (tile_y <= 32)
// stencil tile (threads per work-group = DUAL_BX * DUAL_BY); the macros
// allow other shapes to be built for measurement
#ifndef GR_DUAL_BX
#define GR_DUAL_BX 64
#endif
#ifndef GR_DUAL_BY
#define GR_DUAL_BY 4
#endif
constexpr int DUAL_BX = GR_DUAL_BX, DUAL_BY = GR_DUAL_BY;
// the stencils take x +- 1 neighbours from the adjacent lanes of a wave
// (wave-wide shifts with memory loads at lanes 0 and 63): a tile row must be
// exactly one wave
static_assert(DUAL_BX == 64, "stencil tiles are one 64-lane wave wide");

#ifndef GR_DUAL_GROUP
#define GR_DUAL_GROUP 8
#endif
#ifndef GR_PRIMAL_GROUP
#define GR_PRIMAL_GROUP 4
#endif
constexpr int DUAL_GROUP = GR_DUAL_GROUP, PRIMAL_GROUP = GR_PRIMAL_GROUP;
// Work id of a single-problem stencil work-group.  Hardware dispatch deals
// work-groups round-robin over the 8 XCDs; here runs of G consecutive tiles
// (neighbours in x, then y) are kept on one XCD so that the cache lines their
// neighbour reads share are fetched into one L2, while the 8 XCDs still work
// on adjacent runs (dense, concurrently active addresses).  Grids that are
// not a multiple of 8 runs keep hardware order.
template <int G>
__device__ __forceinline__ int stencil_block()
{
    const int b = blockIdx.x;
    if (G <= 1 || gridDim.x % (8 * G) != 0) return b;
    const int xcd = b & 7, local = b >> 3;
    return ((local / G) * 8 + xcd) * G + local % G;
}

// Kernel-side sizes derived from a problem and its tuning.
int fwd_pitch_of(int nx, int ny);
int fwd_table_choice(const Geometry& g, long long groups, double wg_per_cu = 3.0);
int fwd_band_auto(int pitch, int K, int lds_kb = 36);
int adj_segw_of(int tile_y);
int adj_achunk_of(int na, int segw, int K);
size_t fwd_lds_bytes(const ProbDev& P, int K);
size_t adj_lds_bytes(const ProbDev& P, int K);
int slice_group(int nslices);  // K = 4 (3 or more slices), 2 or 1 slices sharing geometry

// ---- launchers (single problem: descriptor by value) ---------------------
int launch_fwd(const ProbDev& P, const Geometry& g, int nslices, int mode, hipStream_t st);
// adjoint only (primal == false: writes K^T lam into P.ktl) or fused with
// the primal step of the iteration (primal == true); beside: the launch
// shares the GPU with the dual step (two-stream schedule)
int launch_adj(const ProbDev& P, int nslices, bool primal, hipStream_t st, bool beside = false);
// Whether the single-problem stencil kernels stream with the non-temporal
// hint: the stencil state (u, ubar, p, pbar, v, w, K^T lam) above
// STREAM_HINT_MB; GRAPTOR_STREAM_HINTS = 0 or 1 forces the choice.  Per
// two-stream iteration, 4 channels, hints off / on: 3D 256x256x8 (144 MB)
// 0.191 / 0.207 ms, x16 (288 MB) 0.289 / 0.291, x24 (432 MB) 0.410 / 0.389,
// x64 (1152 MB) 0.779 / 0.741; 2D 1024^2 (192 MB) 1.677 / 1.698, 2048^2
// (768 MB) 13.57 / 13.41.
constexpr double STREAM_HINT_MB = 320;
bool stream_hints(const ProbDev& P);
int launch_dual(const ProbDev& P, hipStream_t st);
int launch_fused_pd(const ProbDev& P, const FusedBufs& B, hipStream_t st);
int launch_primal(const ProbDev& P, hipStream_t st);
// the u half (PRIMAL_U), the p half (PRIMAL_P) or both (PRIMAL_UP)
enum { PRIMAL_U = 1, PRIMAL_P = 2, PRIMAL_UP = 3 };
int launch_primal_part(const ProbDev& P, int part, hipStream_t st);

// ---- launchers (batches: descriptor table + item table) -------------------
int launch_fwd_batch(const ProbDev* probs_dev, const FwdItem* items, int nitems, int K,
                     int mi, size_t lds, int mode, hipStream_t st);
int launch_adj_batch(const ProbDev* probs_dev, const AdjItem* items, int nitems, int K,
                     int dim, bool tgv, bool primal, size_t lds, hipStream_t st);
int launch_primal_batch(const ProbDev* probs_dev, const DualItem* items, int nitems, int dim,
                        bool tgv, int part, hipStream_t st);
int launch_dual_batch(const ProbDev* probs_dev, const DualItem* items, int nitems, int dim,
                      int coupling, bool tgv, int maxc, hipStream_t st);

// ---- auxiliary (aux.hip) ---------------------------------------------------
int launch_normest(const Geometry& g, const double* x0_host, int iters, int mode,
                   double tol, double* out, hipStream_t st);
int launch_feasible(const Geometry& g, float* f, int nslices, hipStream_t st);
int launch_energy(const ProbDev& P, double a0, double a1, double* out_reg, hipStream_t st);
int launch_energy_noalloc(const ProbDev& P, const double* mu, double a0, double a1,
                          size_t per_ch, double* out_data, double* out_reg, hipStream_t st);
int launch_data_energy(const float* ku, const float* f, const double* mu, int nch,
                       size_t per_ch, int disc, double* out, hipStream_t st);

}  // namespace gr
