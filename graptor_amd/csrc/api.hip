// extern "C" entry points of libgraptor_amd (declared in include/graptor_amd.h).
#include "common.h"

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <new>
#include <vector>

struct gr_geometry {
    gr::Geometry g;
};

namespace gr {

static thread_local char g_err[512] = "";

void set_error(const char* fmt, ...)
{
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

int hip_status(hipError_t e, const char* where)
{
    set_error("%s: %s", where, hipGetErrorString(e));
    return e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation ? GR_ERR_OOM : GR_ERR_HIP;
}

}  // namespace gr

using namespace gr;

namespace {

constexpr int kFwdThreads = 256;

// Split each angle class into chunks of at most `cap` (angle, bin) items.
std::vector<FwdChunk> make_chunks(const std::vector<FwdAngle>& sorted, int nd, int cap)
{
    std::vector<FwdChunk> out;
    int start = 0;
    const int na = (int)sorted.size();
    while (start < na) {
        int end = start;
        while (end < na && sorted[end].cls == sorted[start].cls) ++end;
        const int ncls = end - start;
        if (nd <= cap) {
            const int per = std::max(1, cap / nd);
            const int nch = (ncls + per - 1) / per;
            const int base = ncls / nch, extra = ncls % nch;  // even split
            int a = start;
            for (int c = 0; c < nch; ++c) {
                const int cnt = base + (c < extra ? 1 : 0);
                out.push_back(FwdChunk{sorted[start].cls, a, cnt, 0, nd, 0, 0, 0});
                a += cnt;
            }
        } else {
            for (int a = start; a < end; ++a)
                for (int b = 0; b < nd; b += cap)
                    out.push_back(FwdChunk{sorted[a].cls, a, 1, b, std::min(cap, nd - b), 0, 0, 0});
        }
        start = end;
    }
    return out;
}

int stream_of(void* s, hipStream_t* out)
{
    *out = reinterpret_cast<hipStream_t>(s);
    return GR_OK;
}

bool valid_params(const gr_geometry* g, const gr_tgv_params* p)
{
    if (!g || !p) return false;
    if (p->dim != 2 && p->dim != 3) return false;
    if (p->nx != g->g.nx || p->ny != g->g.ny) return false;
    if (p->nz <= 0 || p->nch <= 0) return false;
    if (p->dim == 2 && p->nz != 1) return false;
    if (p->coupling < 0 || p->coupling > 2) return false;
    if (p->discrepancy < 0 || p->discrepancy > 1) return false;
    if (p->stepsize <= 0.f) return false;
    return true;
}

}  // namespace

extern "C" {

const char* gr_last_error(void) { return gr::g_err; }

int gr_abi_version(void) { return 1; }

int gr_radon_struct(const double* angles, int na, int nx, int ny, int nd, double dw, double shift,
                    float* ofs_out)
{
    if (!angles || !ofs_out || na <= 0 || nx <= 0 || ny <= 0 || dw == 0.0) {
        set_error("radon_struct: invalid arguments");
        return GR_ERR_INVALID;
    }
    if (nd <= 0) nd = (int)std::ceil(std::hypot((double)nx, (double)ny));
    const double mx = (nx - 1) / 2.0, my = (ny - 1) / 2.0, md = (nd - 1.0) / 2.0;
    for (int j = 0; j < na; ++j) {
        double cx = std::cos(angles[j]) / dw, sy = std::sin(angles[j]) / dw;
        double ci = 1.0 / cx;
        if (std::fabs(ci) > 10.0 * nd) {  // snap near-vertical rays (2D:518-522)
            cx = 0.0;
            sy = std::nearbyint(std::sin(angles[j])) / dw;
            ci = 0.0;
        }
        const double off = md - cx * mx - sy * my + shift / dw;
        ofs_out[4 * j + 0] = (float)cx;
        ofs_out[4 * j + 1] = (float)sy;
        ofs_out[4 * j + 2] = (float)off;
        ofs_out[4 * j + 3] = (float)ci;
    }
    return GR_OK;
}

int gr_geometry_create(const float* ofs, int na, int nx, int ny, int nd, gr_geometry** out)
{
    if (!ofs || !out || na <= 0 || nx <= 1 || ny <= 1 || nd <= 1) {
        set_error("geometry_create: invalid arguments");
        return GR_ERR_INVALID;
    }
    *out = nullptr;
    gr_geometry* gg = new (std::nothrow) gr_geometry();
    if (!gg) {
        set_error("geometry_create: host allocation failed");
        return GR_ERR_OOM;
    }
    Geometry& g = gg->g;
    g.nx = nx; g.ny = ny; g.nd = nd; g.na = na;
    g.fwd_band = 0;
    g.adj_tile_y = 16;
    std::vector<FwdAngle> fa(na);
    for (int j = 0; j < na; ++j) {
        const double c = ofs[4 * j], s = ofs[4 * j + 1], o = ofs[4 * j + 2];
        FwdAngle a{};
        a.cls = std::fabs(s) >= std::fabs(c) ? 0 : 1;
        const double mc = a.cls == 0 ? s : c;
        if (mc == 0.0) {
            delete gg;
            set_error("geometry_create: degenerate angle %d", j);
            return GR_ERR_INVALID;
        }
        a.mstep = (float)(a.cls == 0 ? c : s);
        a.mcand = (float)mc;
        a.off = (float)o;
        a.invc = (float)(1.0 / (double)a.mcand);
        a.hc = (float)(1.0 / std::fabs((double)a.mcand));
        a.id = j;
        fa[j] = a;
    }
    std::stable_sort(fa.begin(), fa.end(),
                     [](const FwdAngle& x, const FwdAngle& y) { return x.cls < y.cls; });
    std::vector<FwdChunk> all;
    const int ipt[4] = {8, 4, 2, 1};
    for (int t = 0; t < 4; ++t) {
        g.tab_off[t] = (int)all.size();
        auto c = make_chunks(fa, nd, kFwdThreads * ipt[t]);
        all.insert(all.end(), c.begin(), c.end());
    }
    g.tab_off[4] = (int)all.size();
    g.ofs_host = (float*)std::malloc(sizeof(float) * 4 * na);
    std::memcpy(g.ofs_host, ofs, sizeof(float) * 4 * na);
    hipError_t e = hipMalloc(&g.ofs_dev, sizeof(float) * 4 * na);
    if (e == hipSuccess) e = hipMalloc(&g.fwd_dev, sizeof(FwdAngle) * na);
    if (e == hipSuccess) e = hipMalloc(&g.chunks_dev, sizeof(FwdChunk) * all.size());
    if (e == hipSuccess)
        e = hipMemcpy(g.ofs_dev, ofs, sizeof(float) * 4 * na, hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(g.fwd_dev, fa.data(), sizeof(FwdAngle) * na, hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(g.chunks_dev, all.data(), sizeof(FwdChunk) * all.size(),
                      hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        int rc = hip_status(e, "geometry_create");
        gr_geometry_destroy(gg);
        return rc;
    }
    *out = gg;
    return GR_OK;
}

int gr_geometry_destroy(gr_geometry* gg)
{
    if (!gg) return GR_OK;
    Geometry& g = gg->g;
    if (g.ofs_dev) (void)hipFree(g.ofs_dev);
    if (g.fwd_dev) (void)hipFree(g.fwd_dev);
    if (g.chunks_dev) (void)hipFree(g.chunks_dev);
    std::free(g.ofs_host);
    delete gg;
    return GR_OK;
}

int gr_geometry_set_tuning(gr_geometry* gg, int fwd_band, int adj_tile_y)
{
    if (!gg) {
        set_error("set_tuning: null geometry");
        return GR_ERR_INVALID;
    }
    if (fwd_band < 0 || fwd_band > 256 || adj_tile_y < 0 ||
        (adj_tile_y && (adj_tile_y < 4 || adj_tile_y > 32 || (adj_tile_y & 3)))) {
        set_error("set_tuning: fwd_band in [1,256], adj_tile_y multiple of 4 in [4,32]");
        return GR_ERR_INVALID;
    }
    if (fwd_band) gg->g.fwd_band = fwd_band;
    if (adj_tile_y) gg->g.adj_tile_y = adj_tile_y;
    return GR_OK;
}

int gr_geometry_get_tuning(const gr_geometry* gg, int* fwd_band, int* adj_tile_y)
{
    if (!gg) {
        set_error("get_tuning: null geometry");
        return GR_ERR_INVALID;
    }
    if (fwd_band) *fwd_band = gg->g.fwd_band;
    if (adj_tile_y) *adj_tile_y = gg->g.adj_tile_y;
    return GR_OK;
}

int gr_radon(const gr_geometry* gg, float* sino, const float* img, int nslices, void* stream)
{
    if (!gg || !sino || !img || nslices < 0) {
        set_error("radon: invalid arguments");
        return GR_ERR_INVALID;
    }
    hipStream_t st;
    stream_of(stream, &st);
    return launch_radon_fwd(gg->g, img, sino, nullptr, nullptr, nullptr, nslices, 1,
                                   FWD_WRITE, 0.f, st);
}

int gr_radon_ad(const gr_geometry* gg, float* img, const float* sino, int nslices, void* stream)
{
    if (!gg || !sino || !img || nslices < 0) {
        set_error("radon_ad: invalid arguments");
        return GR_ERR_INVALID;
    }
    hipStream_t st;
    stream_of(stream, &st);
    return launch_radon_adj(gg->g, sino, img, nslices, st);
}

int gr_normest(const gr_geometry* gg, const double* x0, int iters, int mode, double tol,
               double* out, void* stream)
{
    if (!gg || !out || iters <= 0 || (mode != GR_NORMEST_REFERENCE && mode != GR_NORMEST_POWER) ||
        (mode == GR_NORMEST_REFERENCE && !x0)) {
        set_error("normest: invalid arguments");
        return GR_ERR_INVALID;
    }
    hipStream_t st;
    stream_of(stream, &st);
    return launch_normest(gg->g, x0, iters, mode, tol, out, st);
}

int gr_make_feasible(const gr_geometry* gg, float* f, int nslices, void* stream)
{
    if (!gg || !f || nslices < 0) {
        set_error("make_feasible: invalid arguments");
        return GR_ERR_INVALID;
    }
    hipStream_t st;
    stream_of(stream, &st);
    return launch_feasible(gg->g, f, nslices, st);
}

int gr_tgv_iterate(const gr_geometry* gg, const gr_tgv_params* p, const gr_tgv_state* s,
                   int n_iter, int ku_on_last, void* stream)
{
    if (!valid_params(gg, p) || !s || n_iter < 0) {
        set_error("tgv_iterate: invalid parameters");
        return GR_ERR_INVALID;
    }
    const bool tgv = p->regularisation == GR_REG_TGV;
    if (!s->u || !s->ubar || !s->v || !s->lam || !s->ktl || !s->f || !s->mfac ||
        (tgv && (!s->p || !s->pbar || !s->w)) || (ku_on_last && !s->ku)) {
        set_error("tgv_iterate: missing state buffer");
        return GR_ERR_INVALID;
    }
    hipStream_t st;
    stream_of(stream, &st);
    TgvDims d{p->dim, p->nx, p->ny, p->nz, p->nch, p->coupling, tgv ? 1 : 0, p->stepsize};
    const int nsl = p->nz * p->nch;
    const int pmode = p->discrepancy == GR_DISC_KL ? FWD_KL : FWD_L2;
    for (int it = 0; it < n_iter; ++it) {
        int rc = launch_dual(d, s->v, s->w, s->ubar, s->pbar, p->sigma, p->alpha_inv0,
                             p->alpha_inv1, st);
        if (rc) return rc;
        const int mode = pmode | ((ku_on_last && it == n_iter - 1) ? FWD_WRITE : 0);
        rc = launch_radon_fwd(gg->g, s->ubar, s->ku, s->lam, s->f, s->mfac, nsl, p->nz, mode,
                                     p->sigma_p, st);
        if (rc) return rc;
        rc = launch_radon_adj(gg->g, s->lam, s->ktl, nsl, st);
        if (rc) return rc;
        rc = launch_primal(d, s->u, s->ubar, s->p, s->pbar, s->v, s->w, s->ktl, p->tau,
                           p->norming, st);
        if (rc) return rc;
    }
    return GR_OK;
}

int gr_tgv_profile(const gr_geometry* gg, const gr_tgv_params* p, const gr_tgv_state* s,
                   int n_iter, double* ms_out, void* stream)
{
    if (!valid_params(gg, p) || !s || n_iter <= 0 || !ms_out) {
        set_error("tgv_profile: invalid parameters");
        return GR_ERR_INVALID;
    }
    const bool tgv = p->regularisation == GR_REG_TGV;
    if (!s->u || !s->ubar || !s->v || !s->lam || !s->ktl || !s->f || !s->mfac ||
        (tgv && (!s->p || !s->pbar || !s->w))) {
        set_error("tgv_profile: missing state buffer");
        return GR_ERR_INVALID;
    }
    hipStream_t st;
    stream_of(stream, &st);
    TgvDims d{p->dim, p->nx, p->ny, p->nz, p->nch, p->coupling, tgv ? 1 : 0, p->stepsize};
    const int nsl = p->nz * p->nch;
    const int pmode = p->discrepancy == GR_DISC_KL ? FWD_KL : FWD_L2;
    hipEvent_t ev[5];
    for (int k = 0; k < 5; ++k) GR_HIP(hipEventCreate(&ev[k]));
    double acc[4] = {0, 0, 0, 0};
    int rc = GR_OK;
    for (int it = 0; it < n_iter && rc == GR_OK; ++it) {
        (void)hipEventRecord(ev[0], st);
        rc = launch_dual(d, s->v, s->w, s->ubar, s->pbar, p->sigma, p->alpha_inv0,
                         p->alpha_inv1, st);
        (void)hipEventRecord(ev[1], st);
        if (!rc)
            rc = launch_radon_fwd(gg->g, s->ubar, s->ku, s->lam, s->f, s->mfac, nsl, p->nz,
                                  pmode, p->sigma_p, st);
        (void)hipEventRecord(ev[2], st);
        if (!rc) rc = launch_radon_adj(gg->g, s->lam, s->ktl, nsl, st);
        (void)hipEventRecord(ev[3], st);
        if (!rc)
            rc = launch_primal(d, s->u, s->ubar, s->p, s->pbar, s->v, s->w, s->ktl, p->tau,
                               p->norming, st);
        (void)hipEventRecord(ev[4], st);
        hipError_t e = hipEventSynchronize(ev[4]);
        if (e != hipSuccess && !rc) rc = hip_status(e, "tgv_profile");
        for (int k = 0; k < 4 && !rc; ++k) {
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, ev[k], ev[k + 1]);
            acc[k] += ms;
        }
    }
    for (int k = 0; k < 5; ++k) (void)hipEventDestroy(ev[k]);
    for (int k = 0; k < 4; ++k) ms_out[k] = acc[k];
    return rc;
}

int gr_tgv_energy(const gr_geometry* gg, const gr_tgv_params* p, const gr_tgv_state* s,
                  const double* mu, double alpha0, double alpha1, double L, double* out,
                  void* stream)
{
    if (!valid_params(gg, p) || !s || !s->u || !s->ku || !s->f || !mu || !out || L <= 0) {
        set_error("tgv_energy: invalid arguments");
        return GR_ERR_INVALID;
    }
    const bool tgv = p->regularisation == GR_REG_TGV;
    if (tgv && !s->p) {
        set_error("tgv_energy: p required for TGV");
        return GR_ERR_INVALID;
    }
    hipStream_t st;
    stream_of(stream, &st);
    const int nsl = p->nz * p->nch;
    int rc = launch_radon_fwd(gg->g, s->u, s->ku, nullptr, nullptr, nullptr, nsl, p->nz,
                                     FWD_WRITE, 0.f, st);
    if (rc) return rc;
    double data = 0.0, reg = 0.0;
    const size_t per_ch = (size_t)gg->g.nd * gg->g.na * p->nz;
    rc = launch_data_energy(s->ku, s->f, mu, p->nch, per_ch, p->discrepancy, &data, st);
    if (rc) return rc;
    const double scale = p->discrepancy == GR_DISC_KL ? 1.0 / L : 1.0;
    TgvDims d{p->dim, p->nx, p->ny, p->nz, p->nch, p->coupling, tgv ? 1 : 0, p->stepsize};
    rc = launch_energy(d, s->u, s->p, alpha0 * scale, alpha1 * scale, &reg, st);
    if (rc) return rc;
    out[0] = data;
    out[1] = reg;
    out[2] = data + reg;
    return GR_OK;
}

}  // extern "C"

