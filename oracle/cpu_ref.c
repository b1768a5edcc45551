/*
 * FP32 C/OpenMP restatement of Graptor's primal-dual TGV kernels.
 *
 * TEST INFRASTRUCTURE ONLY: the FP32 stand-in for the reference's OpenCL
 * path in parity tests, and the CPU baseline timed by bench.py.  The product
 * (graptor_amd) never links or loads it.
 *
 * It follows the kernels of radon_tgv_primal_dual_2d.py (2D:95-463) and
 * radon_tgv_primal_dual_3d.py (3D:96-848) pass by pass: the same eleven
 * passes per iteration in the same order (3D:1197-1221), the same
 * array-of-structures layouts (float2 / 16-byte float3 / float8 elements,
 * first index fastest), the same hypot chains and float constants.  It
 * differs deliberately in two places, both recorded in the design notes:
 *  - the 2D coupled projections run once per pixel (the reference launches
 *    them C times per pixel, 2D:643, racing on the same data);
 *  - 3D nuclear coupling is computed from the mathematics (Jacobi
 *    eigen-decomposition of the 3x3 Gram matrix) because the reference's
 *    helpers return pointers to dead locals (3D:219-220, 282-283).
 */
#include <math.h>
#include <stddef.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define SQRT2_REF 1.4142f

enum { CPL_UNCORR = 0, CPL_FROB = 1, CPL_NUCL = 2 };
enum { DISC_KL = 0, DISC_L2 = 1 };

/* ---------------------------------------------------------------------- */
/* projector: radon 2D:391-434, radon_ad 2D:436-462                        */
/* ---------------------------------------------------------------------- */
void cref_radon(float *sino, const float *img, const float *ofs, int X, int Y,
                int I, int J, int nslices)
{
#pragma omp parallel for collapse(2) schedule(static)
    for (int s = 0; s < nslices; ++s)
        for (int j = 0; j < J; ++j) {
            const float oc = ofs[4 * j + 0], os = ofs[4 * j + 1];
            const float oo = ofs[4 * j + 2], oi = ofs[4 * j + 3];
            const float *im0 = img + (size_t)X * Y * s;
            for (int i = 0; i < I; ++i) {
                float acc = 0.0f;
                const float *im = im0;
                for (int y = 0; y < Y; ++y, im += X) {
                    int xlo, xhi;
                    const float d = y * os + oo;
                    if (oc == 0.0f) {
                        if (d > i - 1 && d < i + 1) {
                            xlo = 0;
                            xhi = X - 1;
                        } else {
                            continue;
                        }
                    } else if (oc > 0.0f) {
                        xlo = (int)((i - 1 - d) * oi);
                        xhi = (int)((i + 1 - d) * oi);
                    } else {
                        xlo = (int)((i + 1 - d) * oi);
                        xhi = (int)((i - 1 - d) * oi);
                    }
                    if (xlo < 0) xlo = 0;
                    if (xhi > X - 1) xhi = X - 1;
                    for (int x = xlo; x <= xhi; ++x) {
                        const float wgt = 1.0f - fabsf(x * oc + d - i);
                        if (wgt > 0.0f) acc += wgt * im[x];
                    }
                }
                sino[(size_t)j * I + i + (size_t)I * J * s] = acc;
            }
        }
}

void cref_radon_ad(float *img, const float *sino, const float *ofs, int X,
                   int Y, int I, int J, int nslices)
{
#pragma omp parallel for collapse(2) schedule(static)
    for (int s = 0; s < nslices; ++s)
        for (int y = 0; y < Y; ++y)
            for (int x = 0; x < X; ++x) {
                const float *sn = sino + (size_t)I * J * s;
                float acc = 0.0f;
                for (int j = 0; j < J; ++j, sn += I) {
                    const float t = x * ofs[4 * j] + y * ofs[4 * j + 1] + ofs[4 * j + 2];
                    if (t > -1.0f && t < (float)I) {
                        const float fl = floorf(t);
                        const float p = t - fl;
                        if (fl >= 0.0f) acc += (1.0f - p) * sn[(int)fl];
                        if (fl <= I - 2) acc += p * sn[(int)fl + 1];
                    }
                }
                img[(size_t)y * X + x + (size_t)X * Y * s] = acc;
            }
}

/* ---------------------------------------------------------------------- */
/* solver state                                                            */
/* ---------------------------------------------------------------------- */
typedef struct {
    int dim, N, nd, na, nz, C, S; /* S = nz*C slices */
    int coupling, disc, tgv;
    float h;
    size_t n; /* voxels per channel-stack: N*N*S */
    int pv;   /* floats per p/v element: 2 (2D) or 4 (3D) */
    int pw;   /* floats per w element: 4 (2D) or 8 (3D) */
} st_t;

/* update_v 2D:265-284 / 3D:615-634 */
static void pass_update_v(const st_t *g, float *v, const float *u, const float *p,
                          float sigma)
{
    const int N = g->N, Nz = g->nz;
#pragma omp parallel for collapse(2) schedule(static)
    for (int s = 0; s < g->S; ++s)
        for (int y = 0; y < N; ++y) {
            const int z = s % Nz;
            for (int x = 0; x < N; ++x) {
                const size_t i = (size_t)N * y + x + (size_t)N * N * s;
                float g0 = -u[i], g1 = -u[i], g2 = -u[i];
                if (x < N - 1) g0 += u[i + 1]; else g0 = 0.0f;
                if (y < N - 1) g1 += u[i + N]; else g1 = 0.0f;
                float *vv = v + g->pv * i;
                const float *pp = p + g->pv * i;
                vv[0] = vv[0] + sigma * (g0 - pp[0]);
                vv[1] = vv[1] + sigma * (g1 - pp[1]);
                if (g->dim == 3) {
                    if (z < Nz - 1) g2 += u[i + (size_t)N * N]; else g2 = 0.0f;
                    g2 /= g->h;
                    vv[2] = vv[2] + sigma * (g2 - pp[2]);
                }
            }
        }
}

/* update_w 2D:286-304 / 3D:637-660 (TV: update_w_empty) */
static void pass_update_w(const st_t *g, float *w, const float *p, float sigma)
{
    const int N = g->N, Nz = g->nz, pv = g->pv;
#pragma omp parallel for collapse(2) schedule(static)
    for (int s = 0; s < g->S; ++s)
        for (int y = 0; y < N; ++y) {
            const int z = s % Nz;
            for (int x = 0; x < N; ++x) {
                const size_t i = (size_t)N * y + x + (size_t)N * N * s;
                float *ww = w + g->pw * i;
                if (!g->tgv) {
                    for (int k = 0; k < g->pw; ++k) ww[k] = 0.0f;
                    continue;
                }
                const float *pc = p + pv * i;
                float ax[3] = {0, 0, 0}, ay[3] = {0, 0, 0}, az[3] = {0, 0, 0};
                const ptrdiff_t sx = pv, sy = (ptrdiff_t)pv * N;
                for (int k = 0; k < g->dim; ++k) {
                    ax[k] = (x > 0) ? pc[k] - pc[k - sx] : 0.0f;
                    ay[k] = (y > 0) ? pc[k] - pc[k - sy] : 0.0f;
                }
                if (g->dim == 2) {
                    ww[0] = ww[0] + sigma * ax[0];
                    ww[1] = ww[1] + sigma * ay[1];
                    ww[2] = ww[2] + sigma * (0.5f * (ax[1] + ay[0]));
                } else {
                    for (int k = 0; k < 3; ++k) {
                        az[k] = (z > 0) ? pc[k] - pc[k - sy * N] : 0.0f;
                        az[k] /= g->h;
                    }
                    ww[0] = ww[0] + sigma * ax[0];
                    ww[1] = ww[1] + sigma * ay[1];
                    ww[2] = ww[2] + sigma * az[2];
                    ww[3] = ww[3] + sigma * (0.5f * (ax[1] + ay[0]));
                    ww[4] = ww[4] + sigma * (0.5f * (ax[2] + az[0]));
                    ww[5] = ww[5] + sigma * (0.5f * (ay[2] + az[1]));
                }
            }
        }
}

/* 2x2 symmetric Gram route of update_NormV_Nucl, 2D:124-204 */
static void nucl2d_matrix(float a, float b, float c, float alphainv, float M[3])
{
    const float eps = 0.00001f;
    float S1, S2;
    const float o = fminf(a, c);
    if ((b * b) / (o * o) > eps && o * o > 0) {
        const float d = a + c, e = a - c;
        float f = 4 * b * b + e * e;
        f = (f < 0) ? 0.0f : sqrtf(f);
        const float sg2 = (d - f > 0) ? sqrtf((d - f) * 0.5f) : 0.0f;
        const float sg1 = (d + f > 0) ? sqrtf((d + f) * 0.5f) : 0.0f;
        const float gg = (e + f) * 0.5f, hh = (e - f) * 0.5f;
        float v0 = gg * a + b * b, v1 = gg * b + c * b;
        float v2 = hh * a + b * b, v3 = hh * b + b * c;
        const float n1 = hypotf(v0, v1), n2 = hypotf(v2, v3);
        if (n1 > 0) { v0 /= n1; v1 /= n1; }
        if (n2 > 0) { v2 /= n2; v3 /= n2; }
        S1 = (sg1 * alphainv > 1) ? 1.0f / (sg1 * alphainv) : 1.0f;
        S2 = (sg2 * alphainv > 1) ? 1.0f / (sg2 * alphainv) : 1.0f;
        M[0] = v0 * v0 * S1 + v2 * v2 * S2;
        M[2] = v1 * v1 * S1 + v3 * v3 * S2;
        M[1] = S1 * v0 * v1 + S2 * v2 * v3;
    } else {
        const float sg1 = sqrtf(a), sg2 = sqrtf(c);
        S1 = (alphainv * sg1 > 1) ? 1.0f / (alphainv * sg1) : 1.0f;
        S2 = (alphainv * sg2 > 1) ? 1.0f / (alphainv * sg2) : 1.0f;
        M[0] = S1;
        M[2] = S2;
        M[1] = 0.0f;
    }
}

/* Symmetric 3x3 eigen-decomposition by cyclic Jacobi rotations.
 * A = {a00, a11, a22, a01, a02, a12}; on return ev[3] eigenvalues and
 * V (column k = eigenvector k, row-major 3x3). */
void cref_jacobi3(const float A[6], float ev[3], float V[9])
{
    float a[3][3] = {{A[0], A[3], A[4]}, {A[3], A[1], A[5]}, {A[4], A[5], A[2]}};
    float v[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    for (int sweep = 0; sweep < 6; ++sweep) {
        for (int pq = 0; pq < 3; ++pq) {
            const int P = (pq == 2) ? 1 : 0, Q = (pq == 0) ? 1 : 2;
            const float apq = a[P][Q];
            if (apq == 0.0f) continue;
            const float theta = (a[Q][Q] - a[P][P]) / (2.0f * apq);
            const float t = copysignf(1.0f, theta) / (fabsf(theta) + sqrtf(theta * theta + 1.0f));
            const float cs = 1.0f / sqrtf(t * t + 1.0f), sn = t * cs;
            for (int k = 0; k < 3; ++k) { /* A <- A J */
                const float akp = a[k][P], akq = a[k][Q];
                a[k][P] = cs * akp - sn * akq;
                a[k][Q] = sn * akp + cs * akq;
            }
            for (int k = 0; k < 3; ++k) { /* A <- J^T A */
                const float apk = a[P][k], aqk = a[Q][k];
                a[P][k] = cs * apk - sn * aqk;
                a[Q][k] = sn * apk + cs * aqk;
            }
            a[P][Q] = a[Q][P] = 0.0f;
            for (int k = 0; k < 3; ++k) {
                const float vkp = v[k][P], vkq = v[k][Q];
                v[k][P] = cs * vkp - sn * vkq;
                v[k][Q] = sn * vkp + cs * vkq;
            }
        }
    }
    for (int k = 0; k < 3; ++k) ev[k] = a[k][k];
    for (int r = 0; r < 3; ++r)
        for (int k = 0; k < 3; ++k) V[3 * r + k] = v[r][k];
}

/* update_NormV_* 2D:98-220 / 3D:290-565 */
static void pass_norm_v(const st_t *g, float *v, float alphainv)
{
    const size_t plane = (size_t)g->N * g->N * g->nz; /* one channel */
    const int C = g->C, pv = g->pv, dim = g->dim;
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < plane; ++i) {
        if (g->coupling == CPL_UNCORR) {
            for (int c = 0; c < C; ++c) {
                float *e = v + pv * (i + plane * c);
                float n = (dim == 2) ? hypotf(e[0], e[1])
                                     : hypotf(hypotf(e[0], e[1]), e[2]);
                n *= alphainv;
                if (n > 1.0f)
                    for (int k = 0; k < dim; ++k) e[k] /= n;
            }
        } else if (g->coupling == CPL_FROB) {
            float n = 0.0f;
            for (int c = 0; c < C; ++c) {
                const float *e = v + pv * (i + plane * c);
                n = (dim == 2) ? hypotf(hypotf(e[0], e[1]), n)
                               : hypotf(hypotf(hypotf(e[0], e[1]), e[2]), n);
            }
            n *= alphainv;
            if (n > 1.0f)
                for (int c = 0; c < C; ++c) {
                    float *e = v + pv * (i + plane * c);
                    for (int k = 0; k < dim; ++k) e[k] /= n;
                }
        } else if (dim == 2) {
            float a = 0, b = 0, cc = 0, M[3];
            for (int c = 0; c < C; ++c) {
                const float *e = v + pv * (i + plane * c);
                a += e[0] * e[0];
                b += e[0] * e[1];
                cc += e[1] * e[1];
            }
            nucl2d_matrix(a, b, cc, alphainv, M);
            for (int c = 0; c < C; ++c) {
                float *e = v + pv * (i + plane * c);
                const float x0 = e[0], x1 = e[1];
                e[0] = x0 * M[0] + x1 * M[1];
                e[1] = x0 * M[1] + x1 * M[2];
            }
        } else {
            float G[6] = {0, 0, 0, 0, 0, 0}, ev[3], V[9], S[3], M[6];
            for (int c = 0; c < C; ++c) {
                const float *e = v + pv * (i + plane * c);
                G[0] += e[0] * e[0];
                G[1] += e[1] * e[1];
                G[2] += e[2] * e[2];
                G[3] += e[0] * e[1];
                G[4] += e[0] * e[2];
                G[5] += e[1] * e[2];
            }
            cref_jacobi3(G, ev, V);
            for (int k = 0; k < 3; ++k) {
                const float sg = ev[k] > 0 ? sqrtf(ev[k]) : 0.0f;
                S[k] = (sg * alphainv > 1.0f) ? 1.0f / (alphainv * sg) : 1.0f;
            }
            /* M = V diag(S) V^T, stored (00, 11, 22, 01, 02, 12) */
            const int rr[6] = {0, 1, 2, 0, 0, 1}, cc2[6] = {0, 1, 2, 1, 2, 2};
            for (int m = 0; m < 6; ++m) {
                float acc = 0.0f;
                for (int k = 0; k < 3; ++k)
                    acc += S[k] * V[3 * rr[m] + k] * V[3 * cc2[m] + k];
                M[m] = acc;
            }
            for (int c = 0; c < C; ++c) {
                float *e = v + pv * (i + plane * c);
                const float A0 = e[0], B0 = e[1], C0 = e[2];
                e[0] = A0 * M[0] + B0 * M[3] + C0 * M[4];
                e[1] = A0 * M[3] + B0 * M[1] + C0 * M[5];
                e[2] = A0 * M[4] + B0 * M[5] + C0 * M[2];
            }
        }
    }
}

/* update_NormW_* 2D:223-262 / 3D:573-613 */
static void pass_norm_w(const st_t *g, float *w, float alphainv)
{
    if (!g->tgv) return;
    const size_t plane = (size_t)g->N * g->N * g->nz;
    const int C = g->C, pw = g->pw, dim = g->dim;
    const int m = (dim == 2) ? 3 : 6;
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < plane; ++i) {
        if (g->coupling == CPL_UNCORR) {
            for (int c = 0; c < C; ++c) {
                float *e = w + pw * (i + plane * c);
                float n;
                if (dim == 2)
                    n = hypotf(hypotf(e[0], e[1]), SQRT2_REF * e[2]);
                else
                    n = hypotf(hypotf(hypotf(e[0], e[1]), hypotf(e[2], SQRT2_REF * e[3])),
                               hypotf(SQRT2_REF * e[4], e[5]));
                n *= alphainv;
                if (n > 1.0f)
                    for (int k = 0; k < m; ++k) e[k] /= n;
            }
        } else {
            float n = 0.0f;
            for (int c = 0; c < C; ++c) {
                const float *e = w + pw * (i + plane * c);
                if (dim == 2)
                    n = hypotf(hypotf(e[0], e[1]), hypotf(2.f * e[2], n));
                else
                    n = hypotf(hypotf(hypotf(e[0], e[1]), hypotf(e[2], SQRT2_REF * e[3])),
                               hypotf(SQRT2_REF * hypotf(e[4], e[5]), n));
            }
            n *= alphainv;
            if (n > 1.0f)
                for (int c = 0; c < C; ++c) {
                    float *e = w + pw * (i + plane * c);
                    for (int k = 0; k < m; ++k) e[k] /= n;
                }
        }
    }
}

/* update_lambda_KL 2D:316-328 / update_lambda_L2 2D:330-338 */
static void pass_lambda(const st_t *g, float *lam, const float *ku, const float *f,
                        float sigp, const float *mu)
{
    const size_t per = (size_t)g->nd * g->na;
#pragma omp parallel for collapse(2) schedule(static)
    for (int s = 0; s < g->S; ++s)
        for (size_t k = 0; k < per; ++k) {
            const int c = s / g->nz;
            const size_t i = k + per * s;
            if (g->disc == DISC_KL) {
                const float sv = lam[i] + sigp * ku[i];
                const float z = sv - mu[c];
                const float d = z * z + 4 * sigp * f[i] * mu[c];
                lam[i] = sv - 0.5f * (z + sqrtf(d));
            } else {
                lam[i] = (lam[i] + sigp * (ku[i] - f[i])) * mu[c];
            }
        }
}

/* update_u 2D:340-357 / 3D:706-729 */
static void pass_update_u(const st_t *g, float *unew, const float *u, const float *v,
                          const float *ktl, float tau, float norming)
{
    const int N = g->N, Nz = g->nz, pv = g->pv;
#pragma omp parallel for collapse(2) schedule(static)
    for (int s = 0; s < g->S; ++s)
        for (int y = 0; y < N; ++y) {
            const int z = s % Nz;
            for (int x = 0; x < N; ++x) {
                const size_t i = (size_t)N * y + x + (size_t)N * N * s;
                const float *vv = v + pv * i;
                float d0 = vv[0], d1 = vv[1], d2 = 0.0f;
                if (x == N - 1) d0 = 0.0f;
                if (x > 0) d0 -= vv[0 - pv];
                if (y == N - 1) d1 = 0.0f;
                if (y > 0) d1 -= vv[1 - (ptrdiff_t)pv * N];
                float val;
                if (g->dim == 3) {
                    d2 = vv[2];
                    if (z == Nz - 1) d2 = 0.0f;
                    if (z > 0) d2 -= vv[2 - (ptrdiff_t)pv * N * N];
                    d2 /= g->h;
                    val = u[i] + tau * (d0 + d1 + d2 - norming * ktl[i]);
                } else {
                    val = u[i] + tau * (d0 + d1 - norming * ktl[i]);
                }
                unew[i] = (val < 0) ? 0.0f : val;
            }
        }
}

/* update_p 2D:359-376 / 3D:734-757 (TV: update_p_empty) */
static void pass_update_p(const st_t *g, float *pnew, const float *p, const float *v,
                          const float *w, float tau)
{
    const int N = g->N, Nz = g->nz, pv = g->pv, pw = g->pw;
#pragma omp parallel for collapse(2) schedule(static)
    for (int s = 0; s < g->S; ++s)
        for (int y = 0; y < N; ++y) {
            const int z = s % Nz;
            for (int x = 0; x < N; ++x) {
                const size_t i = (size_t)N * y + x + (size_t)N * N * s;
                float *pn = pnew + pv * i;
                if (!g->tgv) {
                    for (int k = 0; k < pv; ++k) pn[k] = 0.0f;
                    continue;
                }
                const float *wc = w + pw * i;
                const float *pc = p + pv * i;
                const float *vc = v + pv * i;
                if (g->dim == 2) {
                    float x0 = -wc[0], x1 = -wc[2], y0 = -wc[2], y1 = -wc[1];
                    if (x == 0) x0 = x1 = 0.0f;
                    if (x < N - 1) { x0 += wc[pw + 0]; x1 += wc[pw + 2]; }
                    if (y == 0) y0 = y1 = 0.0f;
                    if (y < N - 1) { y0 += wc[(size_t)pw * N + 2]; y1 += wc[(size_t)pw * N + 1]; }
                    pn[0] = pc[0] + tau * (vc[0] + x0 + y0);
                    pn[1] = pc[1] + tau * (vc[1] + x1 + y1);
                } else {
                    float xs[3] = {-wc[0], -wc[3], -wc[4]};
                    float ys[3] = {-wc[3], -wc[1], -wc[5]};
                    float zs[3] = {-wc[4], -wc[5], -wc[2]};
                    if (x == 0) xs[0] = xs[1] = xs[2] = 0.0f;
                    if (x < N - 1) {
                        const float *q = wc + pw;
                        xs[0] += q[0]; xs[1] += q[3]; xs[2] += q[4];
                    }
                    if (y == 0) ys[0] = ys[1] = ys[2] = 0.0f;
                    if (y < N - 1) {
                        const float *q = wc + (size_t)pw * N;
                        ys[0] += q[3]; ys[1] += q[1]; ys[2] += q[5];
                    }
                    if (z == 0) zs[0] = zs[1] = zs[2] = 0.0f;
                    if (z < Nz - 1) {
                        const float *q = wc + (size_t)pw * N * N;
                        zs[0] += q[4]; zs[1] += q[5]; zs[2] += q[2];
                    }
                    for (int k = 0; k < 3; ++k) {
                        zs[k] /= g->h;
                        pn[k] = pc[k] + tau * (vc[k] + xs[k] + ys[k] + zs[k]);
                    }
                }
            }
        }
}

/* update_extra: u = 2 u_ - u over every element (2D:637-638) */
static void pass_extra(float *u, const float *unew, size_t n)
{
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < n; ++i) u[i] = 2.0f * unew[i] - u[i];
}

/*
 * Full solve (tgv_radon 2D:710-851, 3D:1135-1292).  Layouts of the inputs
 * and outputs are the compact ones of the oracle:
 *   f, ku_out: (C, Nz, na, nd);  u0, u_out: (C, Nz, N, N);
 *   p_out: (d, C, Nz, N, N) or NULL.
 * f must already be zeroed where R 1 <= 0.  L is the operator-norm value,
 * mu has C entries.  Returns 0 on success, 1 on bad arguments, 2 on OOM.
 */
int cref_tgv_solve(int dim, int nd, int na, int nz, int nch, const float *ofs,
                   const float *f, double L, double alpha0, double alpha1,
                   const double *mu, int maxiter, int coupling, int discrepancy,
                   int tgv, double stepsize, const float *u0, float *u_out,
                   float *ku_out, float *p_out)
{
    if ((dim != 2 && dim != 3) || nd <= 1 || na <= 0 || nz <= 0 || nch <= 0 ||
        nch > 64 || L <= 0)
        return 1;
    if (dim == 2 && nz != 1) return 1;
    st_t g;
    g.dim = dim; g.N = nd; g.nd = nd; g.na = na; g.nz = nz; g.C = nch;
    g.S = nz * nch; g.coupling = coupling; g.disc = discrepancy; g.tgv = tgv;
    g.h = (float)stepsize;
    g.n = (size_t)nd * nd * g.S;
    g.pv = (dim == 2) ? 2 : 4;
    g.pw = (dim == 2) ? 4 : 8;
    const size_t nsino = (size_t)nd * na * g.S;
    float *U = calloc(g.n, 4), *U_ = calloc(g.n, 4), *KT = calloc(g.n, 4);
    float *P = calloc(g.n * g.pv, 4), *P_ = calloc(g.n * g.pv, 4);
    float *V = calloc(g.n * g.pv, 4), *W = calloc(g.n * g.pw, 4);
    float *LAM = calloc(nsino, 4), *KU = calloc(nsino, 4);
    if (!U || !U_ || !KT || !P || !P_ || !V || !W || !LAM || !KU) {
        free(U); free(U_); free(KT); free(P); free(P_); free(V); free(W);
        free(LAM); free(KU);
        return 2;
    }
    if (u0) {
        memcpy(U, u0, g.n * 4);
        memcpy(U_, u0, g.n * 4);
    }
    const double lsqr = (dim == 2) ? 0.5 * (18.0 + sqrt(33.0)) : 17.0;
    const double sigma = 1.0 / sqrt(lsqr), tau = sigma;
    const double at0 = alpha0 / (L * L), at1 = alpha1 / (L * L);
    float muf[64];
    for (int c = 0; c < nch; ++c)
        muf[c] = (float)((discrepancy == DISC_KL) ? mu[c] : mu[c] / (sigma + mu[c]));
    const float fs = (float)sigma, ft = (float)tau, fsp = (float)(sigma / L);
    const float fnorm = (float)(1.0 / L);
    const float ai0 = (float)(1.0 / at0), ai1 = (float)(1.0 / at1);
    for (int it = 0; it < maxiter; ++it) {
        pass_update_v(&g, V, U_, P_, fs);
        pass_update_w(&g, W, P_, fs);
        pass_norm_v(&g, V, ai1);
        pass_norm_w(&g, W, ai0);
        cref_radon(KU, U_, ofs, nd, nd, nd, na, g.S);
        pass_lambda(&g, LAM, KU, f, fsp, muf);
        cref_radon_ad(KT, LAM, ofs, nd, nd, nd, na, g.S);
        /* U_ <- new u from U; then U <- 2 U_ - U; swap (3D:1211-1221) */
        pass_update_u(&g, U_, U, V, KT, ft, fnorm);
        pass_update_p(&g, P_, P, V, W, ft);
        pass_extra(U, U_, g.n);
        pass_extra(P, P_, g.n * g.pv);
        float *t = U; U = U_; U_ = t;
        t = P; P = P_; P_ = t;
    }
    if (u_out) memcpy(u_out, U, g.n * 4);
    if (ku_out) memcpy(ku_out, KU, nsino * 4);
    if (p_out) {
        const int d = dim;
        for (int k = 0; k < d; ++k)
            for (size_t i = 0; i < g.n; ++i) p_out[(size_t)k * g.n + i] = P[g.pv * i + k];
    }
    free(U); free(U_); free(KT); free(P); free(P_); free(V); free(W);
    free(LAM); free(KU);
    return 0;
}
