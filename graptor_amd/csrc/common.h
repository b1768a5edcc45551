// Shared declarations of the graptor_amd HIP library (gfx950 / CDNA4).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stddef.h>

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

// Device-side geometry (owned by gr_geometry).
struct Geometry {
    int nx, ny, nd, na;
    float* ofs_dev;          // na x float4 (cos, sin, off, 1/cos) as given
    FwdAngle* fwd_dev;       // na, sorted by class
    FwdChunk* chunks_dev;    // four chunk tables back to back
    int tab_off[5];          // table t = chunks [tab_off[t], tab_off[t+1]),
                             // built for 8, 4, 2, 1 items per thread
    float* ofs_host;         // copy of the table (host)
    int fwd_band;            // lines per LDS band in the forward kernel
    int adj_tile_y;          // rows per backprojection tile
};

// Launch helpers implemented in projector.hip.
int launch_radon_fwd(const Geometry& g, const float* img, float* sino,
                     float* lam, const float* f, const float* mfac, int nslices,
                     int nz_per_ch, int mode, float sigp, hipStream_t st);
int launch_radon_adj(const Geometry& g, const float* sino, float* img, int nslices,
                     hipStream_t st);
int launch_normest(const Geometry& g, const double* x0_host, int iters, int mode,
                   double tol, double* out, hipStream_t st);
int launch_feasible(const Geometry& g, float* f, int nslices, hipStream_t st);

// Modes of the fused forward projector epilogue.
enum { FWD_WRITE = 1, FWD_KL = 2, FWD_L2 = 4 };

// TGV stencil launchers implemented in tgv.hip.
struct TgvDims {
    int dim, nx, ny, nz, nch;
    int coupling, tgv;
    float h;
};
int launch_dual(const TgvDims& d, float* v, float* w, const float* ubar,
                const float* pbar, float sigma, float ainv0, float ainv1,
                hipStream_t st);
int launch_primal(const TgvDims& d, float* u, float* ubar, float* p, float* pbar,
                  const float* v, const float* w, const float* ktl, float tau,
                  float norming, hipStream_t st);
int launch_energy(const TgvDims& d, const float* u, const float* p, double a0,
                  double a1, double* out_reg, hipStream_t st);
int launch_data_energy(const float* ku, const float* f, const double* mu,
                       int nch, size_t per_ch, int disc, double* out, hipStream_t st);

}  // namespace gr
