/*
 * graptor_amd: C ABI of the MI355X-native primal-dual TGV reconstruction path.
 *
 * Plain C: caller-owned device pointers (e.g. torch.Tensor.data_ptr()),
 * sizes and a hipStream_t passed as void*.  Every function returns an int
 * status (GR_OK == 0) and never throws; gr_last_error() describes the last
 * failure of the calling thread.  Device arrays use the reference's
 * pyopencl order='F' layouts (first index fastest):
 *   image stack  u[c][z][y][x]            (2D: nz == 1)
 *   vector field v[k][c][z][y][x]         k < dim     (component-major)
 *   tensor field w[k][c][z][y][x]         k < 3 (2D: xx, yy, xy)
 *                                         k < 6 (3D: xx, yy, zz, xy, xz, yz)
 *   sinograms    f[c][z][angle][bin]
 *
 * Each declaration names the reference interface it replaces
 * (radon_tgv_primal_dual_2d.py = 2D:, radon_tgv_primal_dual_3d.py = 3D:).
 */
#ifndef GRAPTOR_AMD_H
#define GRAPTOR_AMD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    GR_OK = 0,
    GR_ERR_INVALID = 1,     /* bad argument or shape */
    GR_ERR_OOM = 2,         /* device allocation failed (maps to MemoryError) */
    GR_ERR_HIP = 3,         /* any other HIP runtime failure */
    GR_ERR_UNSUPPORTED = 4  /* valid request the library does not implement */
};

enum { GR_COUPLING_UNCORRELATED = 0, GR_COUPLING_FROBENIUS = 1, GR_COUPLING_NUCLEAR = 2 };
enum { GR_DISC_KL = 0, GR_DISC_L2 = 1 };
enum { GR_REG_TV = 0, GR_REG_TGV = 1 };
enum { GR_NORMEST_REFERENCE = 0, GR_NORMEST_POWER = 1 };

/* Message of the last failing call on this thread ("" if none). */
const char *gr_last_error(void);
/* ABI revision of this library. */
int gr_abi_version(void);

/* Replaces radon_struct() 2D:503-535 / 3D:947-979.  Host only: angles
 * (float64, radians, na of them) -> ofs_out[na][4] = (cos/dw, sin/dw,
 * offset, 1/cos) in float32, with near-vertical rays snapped as 2D:518-522. */
int gr_radon_struct(const double *angles, int na, int nx, int ny, int nd,
                    double detector_width, double detector_shift, float *ofs_out);

/* Device-side form of r_struct (2D:531-535): the angle table plus the
 * projector's per-angle constants and work chunks.  ofs is a host table
 * ofs[na][4] as produced by gr_radon_struct. */
typedef struct gr_geometry gr_geometry;
int gr_geometry_create(const float *ofs, int na, int nx, int ny, int nd,
                       gr_geometry **out);
int gr_geometry_destroy(gr_geometry *g);
/* Tile shapes of the projector kernels (autotuner hook); 0 keeps a value.
 * Defaults: forward bands sized automatically; backprojection tiles of 16
 * rows for images up to 256 px, 32 rows above. */
int gr_geometry_set_tuning(gr_geometry *g, int fwd_band, int adj_tile_y);
int gr_geometry_get_tuning(const gr_geometry *g, int *fwd_band, int *adj_tile_y);
/* 1: run the primal step inside the backprojection kernel (K^T lam never
 * stored); 0 (default): backprojection writes st->ktl and a separate
 * primal kernel reads it. */
int gr_geometry_set_fusion(gr_geometry *g, int fuse_primal);

/* Replaces radon() 2D:544-550 / 3D:988-994 (kernel 2D:391-434):
 * sino[s][angle][bin] = R img[s][y][x] for s < nslices. */
int gr_radon(const gr_geometry *g, float *sino, const float *img, int nslices,
             void *stream);

/* Replaces radon_ad() 2D:559-565 / 3D:1002-1008 (kernel 2D:436-462):
 * img[s] = R^T sino[s]. */
int gr_radon_ad(const gr_geometry *g, float *img, const float *sino, int nslices,
                void *stream);

/* Replaces radon_normest() 2D:576-588 / 3D:1018-1030, computed in FP64 on
 * the device over one nx*ny slice.  mode GR_NORMEST_REFERENCE: `iters`
 * sum-normalised rounds from the host start image x0 (nx*ny doubles), as the
 * reference; GR_NORMEST_POWER: L2-normalised power iteration from x0 (or a
 * positive default when x0 is NULL) until the relative change is <= tol or
 * `iters` rounds.  *out = estimate of ||R||_2.  Synchronises the stream. */
int gr_normest(const gr_geometry *g, const double *x0, int iters, int mode,
               double tol, double *out, void *stream);

/* Replaces Make_function_feasible() 2D:476-486 / 3D:860-870: zero every bin
 * of the device sinogram stack f (nslices slices) where R 1 <= 0. */
int gr_make_feasible(const gr_geometry *g, float *f, int nslices, void *stream);

/* Parameters of the iteration of tgv_radon() 2D:710-851 / 3D:1135-1292,
 * already scaled the way 2D:752-762 / 3D:1180-1189 derive them. */
typedef struct gr_tgv_params {
    int dim;             /* 2 or 3 */
    int nx, ny, nz, nch; /* grid per slice, slices per channel, channels */
    int coupling;        /* GR_COUPLING_* */
    int discrepancy;     /* GR_DISC_* */
    int regularisation;  /* GR_REG_* */
    float stepsize;      /* z spacing h (3D:631) */
    float sigma, tau;    /* dual / primal steps */
    float sigma_p;       /* sigma / L, the sinogram dual step */
    float norming;       /* 1 / L */
    float alpha_inv0;    /* 1 / (alpha0 / L^2) */
    float alpha_inv1;    /* 1 / (alpha1 / L^2) */
} gr_tgv_params;

/* Caller-owned device state.  u, ubar: nch*nz*ny*nx; p, pbar, v: dim
 * components; w: 3 (2D) or 6 (3D) components; lam, f, ktl: sinogram /
 * image stacks; mfac: nch floats (mu for KL, mu/(sigma+mu) for L2); ku may
 * be NULL unless ku_on_last or gr_tgv_energy needs it.  u2, v2, w2 (shaped
 * as u, v, w) are optional second buffers: when given (w2 may be NULL for
 * TV), gr_tgv_iterate runs the primal step of each iteration and the dual
 * step of the next as one kernel, alternating the state between the two
 * sets (pbar serves as the second p); the state ends in u, p, v, w as
 * without them and the values are the same.  That fused form holds at most 8
 * channels (more give GR_ERR_UNSUPPORTED); every other path takes any
 * channel count (the dual step holds up to 8 channels of a voxel in
 * registers and sweeps larger counts twice). */
typedef struct gr_tgv_state {
    float *u, *ubar, *p, *pbar, *v, *w;
    float *lam, *ktl, *ku;
    const float *f, *mfac;
    float *u2, *v2, *w2;
    /* Optional z halo (3D) for a z-slab of a larger volume, as one rank of a
     * slab-sharded solve would hold it; NULL: the slab's first / last plane
     * is the volume's boundary.  Planes are [c][y][x] of the neighbouring
     * slab's adjacent plane: halo_ubar_hi = ubar above (C planes),
     * halo_pbar_lo = pbar below ([k][c], 3 C planes), halo_v_lo = v_z below
     * (C planes), halo_w_hi = w_xz, w_yz, w_zz above ([j][c], 3 C planes).
     * No counterpart in the reference (it solves slabs independently,
     * RC:408-458); the dual step reads halo_ubar_hi / halo_pbar_lo, the
     * primal step halo_v_lo / halo_w_hi.  Not supported with u2/v2/w2 or a
     * geometry that fuses the primal step. */
    const float *halo_ubar_hi, *halo_pbar_lo, *halo_v_lo, *halo_w_hi;
} gr_tgv_state;

/* Runs n_iter iterations of the loop 2D:767-799 / 3D:1194-1221 on `stream`:
 * dual step with the coupled projections; forward projection fused with the
 * KL/L2 prox; backprojection; primal step fused with the extrapolation
 * (inside the backprojection kernel when the geometry's fusion is on, in
 * which case st->ktl may be NULL).  With ku_on_last,
 * the last iteration also stores R(ubar) into ku, the sinogram the reference
 * returns (2D:840-841).  A state with z halo planes is rejected with
 * GR_ERR_UNSUPPORTED (its halos must be exchanged inside every iteration:
 * slabs run through gr_tgv_dual_step / gr_tgv_primal_step). */
int gr_tgv_iterate(const gr_geometry *g, const gr_tgv_params *prm,
                   const gr_tgv_state *st, int n_iter, int ku_on_last,
                   void *stream);

/* gr_tgv_iterate with a HIP event pair around each kernel of every
 * iteration (measurement only; no counterpart in the reference).  ms_out[0..4]
 * receive the summed kernel times in milliseconds of the standalone dual
 * steps, the forward projections + prox, the backprojections (+ primal step
 * when the geometry fuses it), the standalone primal steps and the fused
 * primal + next-dual steps.  Synchronises the stream. */
int gr_tgv_profile(const gr_geometry *g, const gr_tgv_params *prm,
                   const gr_tgv_state *st, int n_iter, double *ms_out, void *stream);

/* One dual step (update_v/update_w + the coupled projections, 3D:615-660,
 * 529-613) of the loop, alone: for solvers that compose the iteration from
 * their own operators (the two-resolution solver, 3DS:1067-1072). */
int gr_tgv_dual_step(const gr_geometry *g, const gr_tgv_params *prm, const gr_tgv_state *st,
                     void *stream);
/* One primal step with extrapolation (update_u/update_p/update_extra,
 * 3DS:1098-1106) reading K^T lam from st->ktl. */
int gr_tgv_primal_step(const gr_geometry *g, const gr_tgv_params *prm, const gr_tgv_state *st,
                       void *stream);

/* Halo messages of a z-slab (3D; no counterpart in the reference, which
 * solves slabs independently, RC:408-458): packs the planes this slab sends
 * to its neighbours, each as one contiguous buffer laid out as the
 * receiver's gr_tgv_state halo field, so a rank sends it as one message and
 * receives into its own halo field with no unpack step.
 *   GR_HALO_DUAL (before the dual step): to_below <- ubar of plane 0
 *     (the lower slab's halo_ubar_hi, nch planes); to_above <- pbar of plane
 *     nz-1 (the upper slab's halo_pbar_lo, 3 nch planes, TGV only).
 *   GR_HALO_PRIMAL (before the primal step): to_below <- w_xz, w_yz, w_zz of
 *     plane 0 (halo_w_hi, 3 nch planes, TGV only); to_above <- v_z of plane
 *     nz-1 (halo_v_lo, nch planes).
 * A NULL destination (a slab at the volume's boundary) is skipped; one
 * launch on `stream` packs both. */
enum { GR_HALO_DUAL = 0, GR_HALO_PRIMAL = 1 };
int gr_tgv_pack_halos(const gr_tgv_params *prm, const gr_tgv_state *st, int phase,
                      float *to_below, float *to_above, void *stream);

/* Primal energy at (u, p) reduced in FP64 on the device (no counterpart in
 * the reference's hot path; its Gap_TGV_KL 3D:882-927 is unused):
 *   KL: sum_c mu_c KL(R u_c, f_c) + TGV_{alpha/L}(u; p)
 *   L2: sum_c mu_c/2 |R u_c - f_c|^2 + TGV_alpha(u; p)
 * out[0] = data term, out[1] = regulariser, out[2] = sum.  Uses st->ku as
 * scratch for R u.  Synchronises the stream. */
int gr_tgv_energy(const gr_geometry *g, const gr_tgv_params *prm,
                  const gr_tgv_state *st, const double *mu, double alpha0,
                  double alpha1, double L, double *out, void *stream);

/* gr_tgv_energy without any device allocation (no counterpart in the
 * reference): for callers that cannot allocate device memory, such as tight
 * serving or convergence-monitoring loops and harnesses that forbid
 * hipMalloc.  Same arguments, same results bit for bit; the FP64 partial sums
 * live in a fixed array of the library on each device (72 KiB, held until the
 * process ends), so at most GR_ENERGY_NOALLOC_MAX_CHANNELS channels are
 * accepted (more: GR_ERR_INVALID).  Calls that overlap on one device share
 * that array and race; serialise them.  Synchronises the stream. */
enum { GR_ENERGY_NOALLOC_MAX_CHANNELS = 32 };
int gr_tgv_energy_noalloc(const gr_geometry *g, const gr_tgv_params *prm,
                          const gr_tgv_state *st, const double *mu, double alpha0,
                          double alpha1, double L, double *out, void *stream);

/* Batches (no counterpart in the reference, which solves one problem per
 * call): nprob independent problems with their own geometry, sizes,
 * coupling, discrepancy and state, advanced together with one launch per
 * kernel class and phase (at most a few launches per phase for the whole
 * batch).  The plan keeps device copies of the descriptors; the state
 * buffers stay caller-owned and must outlive the plan.  A problem with z
 * halo planes gives GR_ERR_UNSUPPORTED. */
typedef struct gr_plan gr_plan;
int gr_plan_create(int nprob, const gr_geometry *const *geoms,
                   const gr_tgv_params *prms, const gr_tgv_state *states,
                   gr_plan **out);
int gr_plan_destroy(gr_plan *plan);
/* n_iter iterations of every problem (ku_on_last as in gr_tgv_iterate).
 * From 2^20 voxel-channels in total an iteration runs as three forks of the
 * caller's stream (forward projection beside the dual step, backprojection
 * beside the p half of the primal step, then the u half), each joined back
 * before the next; smaller plans run the four phases in sequence. */
int gr_plan_iterate(gr_plan *plan, int n_iter, int ku_on_last, void *stream);
/* The same through a hipGraph of `block` iterations, captured once on a
 * private stream and replayed on `stream` n_iter / block times; the
 * remaining iterations run as ordinary launches. */
int gr_plan_iterate_graph(gr_plan *plan, int n_iter, int block, void *stream);
/* Per-phase kernel times (ms_out[4], as gr_tgv_profile) over n_iter
 * iterations.  Synchronises the stream. */
int gr_plan_profile(gr_plan *plan, int n_iter, double *ms_out, void *stream);
/* Re-reads every geometry's tile shapes (gr_geometry_set_tuning) into the
 * plan; used by the tile autotuner.  Drops a captured graph. */
int gr_plan_retune(gr_plan *plan, const gr_geometry *const *geoms);

/* ------------------------------------------------------------------------
 * Two-resolution solver pieces (radon_tgv_primal_dual_3dsubsample.py = 3DS:)
 * ------------------------------------------------------------------------ */

/* Replaces Sparsemultiplication 3DS:403-428: out[c][z][a][x] = sum over the
 * CSR row (x + nx_out * z) of val * in[c][z'][a][x'] with column
 * x' + nx_in * z'; rowstart has nx_out * nz_out + 1 entries, every column is
 * < nx_in * nz_in (the caller guarantees it).  Applies the box-average
 * interpolation between detector/slice grids and its adjoint. */
int gr_csr_apply(const int *rowstart, const int *index, const float *val, int nx_out, int nz_out,
                 int nx_in, int nz_in, int na, int nch, const float *in, float *out, void *stream);
/* Replaces update_lambda_KL / _L2 3DS:229-256 as a standalone pass:
 * lam = prox(lam + sigma_p ku) with per-channel mfac over nch blocks of
 * per_channel bins. */
int gr_prox(int discrepancy, float sigma_p, const float *mfac, int nch, size_t per_channel,
            const float *ku, const float *f, float *lam, void *stream);
/* *out = sum of x[i]^2 in FP64 (the norm of the power iteration 3DS:550).
 * Synchronises the stream. */
int gr_sumsq_f64(const float *x, size_t n, double *out, void *stream);

/* ------------------------------------------------------------------------
 * Learned warm start: BF16 MFMA kernels of a small U-Net (no counterpart in
 * the reference, whose solver always starts from u = 0, 2D:730 / 3D:1161).
 *
 * Activations are NHWC bf16 with a per-pixel stride `ld*` (channels
 * allocated) and a channel offset `c*0`, so a tensor may be a channel slice
 * of a wider buffer.  Weights are W[co][tap][ci] (ci fastest), tap = 3*ky+kx
 * for 3x3 convolutions with zero padding 1, or one tap for 1x1.  Products run
 * on v_mfma_f32_32x32x16_bf16 with FP32 accumulation.
 * ------------------------------------------------------------------------ */

/* y[.., cy0+co] = act(bias[co] + sum_{tap,ci} W[co][tap][ci] x[shifted, cx0+ci]);
 * cin % 16 == 0, cout % 32 == 0, ldx/cx0 multiples of 8; bias may be NULL;
 * relu != 0 applies max(0, .); out_f32 != 0 writes float instead of bf16.
 * The data gradient is this call with gr_cnn_weight_flip'ed weights. */
int gr_cnn_conv_fwd(const void *x, int ldx, int cx0, const void *w, const float *bias, void *y,
                    int ldy, int cy0, int n, int H, int W, int cin, int cout, int taps, int relu,
                    int out_f32, void *stream);
/* dw[co][tap][ci] += sum_pixels dy[p][co] x[p shifted by tap][ci] (FP32, atomic
 * split-K over chunks of rows_per_chunk image rows). */
int gr_cnn_conv_wgrad(const void *x, int ldx, int cx0, const void *dy, int ldd, int cd0,
                      float *dw, int n, int H, int W, int cin, int cout, int taps,
                      int rows_per_chunk, void *stream);
/* db[co] += sum_pixels dy[p][co]; cout divides 256 */
int gr_cnn_bias_grad(const void *dy, int ldd, int cd0, size_t npix, int cout, float *db,
                     void *stream);
/* wt[ci][taps-1-tap][co] = w[co][tap][ci] (bf16) */
int gr_cnn_weight_flip(const void *w, void *wt, int cin, int cout, int taps, void *stream);
/* dy *= (y > 0), in place */
int gr_cnn_relu_bwd(void *dy, int ldd, int cd0, const void *y, int ldy, int cy0, size_t npix,
                    int c, void *stream);
/* backward == 0: y = 2x2 average pool of x (H, W of x).  backward != 0: y (fine,
 * H x W) = 1/4 x (coarse) broadcast, added to y when accumulate != 0. */
int gr_cnn_pool(const void *x, int ldx, int cx0, void *y, int ldy, int cy0, int n, int H, int W,
                int c, int backward, int accumulate, void *stream);
/* backward == 0: y = nearest 2x upsample of x (H, W of x).  backward != 0:
 * y (coarse, H x W) = sum of x (fine) over each 2x2 block. */
int gr_cnn_upsample(const void *x, int ldx, int cx0, void *y, int ldy, int cy0, int n, int H,
                    int W, int c, int backward, void *stream);
/* y[p][cy0+ch] = bf16(scale * x[p][ch]) for dense float x with c_in channels */
int gr_cnn_pack_bf16(const float *x, int c_in, void *y, int ldy, int cy0, size_t npix,
                     float scale, void *stream);
/* g[p][ch] = bf16(gscale (pred[p][ch] - target[p][ch])); when loss is not
 * NULL, *loss += sum of squares */
int gr_cnn_mse_grad(const float *pred, int ldp, const float *target, int c, size_t npix, void *g,
                    int ldg, float gscale, float *loss, void *stream);
/* Adam step (bias-corrected; wd adds wd * p to the gradient) on FP32 master
 * weights; refreshes the bf16 copy pb when non-NULL. */
int gr_cnn_adam(float *p, const float *g, float *m, float *v, void *pb, size_t n, float lr,
                float b1, float b2, float eps, int step, float wd, void *stream);

#ifdef __cplusplus
}
#endif
#endif /* GRAPTOR_AMD_H */
