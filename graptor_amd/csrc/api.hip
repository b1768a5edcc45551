// extern "C" entry points of libgraptor_amd (declared in include/graptor_amd.h).
#include "common.h"

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <new>
#include <tuple>
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

hipStream_t stream_of(void* s) { return reinterpret_cast<hipStream_t>(s); }

bool valid_params(const gr_geometry* g, const gr_tgv_params* p)
{
    if (!g || !p) return false;
    if (p->dim != 2 && p->dim != 3) return false;
    if (p->nx != g->g.nx || p->ny != g->g.ny) return false;
    if (p->nz <= 0 || p->nch <= 0) return false;
    if (p->dim == 2 && p->nz != 1) return false;
    if (p->coupling < 0 || p->coupling > 2) return false;
    if (p->discrepancy < 0 || p->discrepancy > 1) return false;
    if (p->regularisation < 0 || p->regularisation > 1) return false;
    if (p->stepsize <= 0.f) return false;
    return true;
}

// A halo side is all of its planes (the p / w ones only with TGV), 3D only.
bool valid_halos(const gr_geometry* g, const gr_tgv_params* p, const gr_tgv_state* s)
{
    const bool tgv = p->regularisation == GR_REG_TGV;
    const bool lo = s->halo_v_lo != nullptr, hi = s->halo_ubar_hi != nullptr;
    if ((lo || hi) && p->dim != 3) return false;
    if (!lo && s->halo_pbar_lo) return false;
    if (!hi && s->halo_w_hi) return false;
    if (tgv && lo && !s->halo_pbar_lo) return false;
    if (tgv && hi && !s->halo_w_hi) return false;
    if ((lo || hi) && (g->g.fuse_primal || s->u2)) return false;
    return true;
}

bool has_halos(const gr_tgv_state* s)
{
    return s->halo_ubar_hi || s->halo_pbar_lo || s->halo_v_lo || s->halo_w_hi;
}

bool valid_state(const gr_geometry* g, const gr_tgv_params* p, const gr_tgv_state* s,
                 bool need_ku)
{
    const bool tgv = p->regularisation == GR_REG_TGV;
    if (!s || !s->u || !s->ubar || !s->v || !s->lam || !s->f || !s->mfac) return false;
    if (!g->g.fuse_primal && !s->ktl) return false;
    if (tgv && (!s->p || !s->pbar || !s->w)) return false;
    if (need_ku && !s->ku) return false;
    return valid_halos(g, p, s);
}

// Descriptor of one problem with the geometry's tuning applied for its
// slice-group size K.
ProbDev make_prob(const gr_geometry* gg, const gr_tgv_params* p, const gr_tgv_state* s)
{
    const Geometry& g = gg->g;
    ProbDev P{};
    P.dim = p->dim; P.nx = p->nx; P.ny = p->ny; P.nz = p->nz; P.nch = p->nch;
    P.nd = g.nd; P.na = g.na;
    P.coupling = p->coupling;
    P.tgv = p->regularisation == GR_REG_TGV ? 1 : 0;
    P.disc = p->discrepancy;
    P.hinv = 1.f / p->stepsize;
    P.sigma = p->sigma; P.tau = p->tau; P.sigma_p = p->sigma_p; P.norming = p->norming;
    P.ainv0 = p->alpha_inv0; P.ainv1 = p->alpha_inv1;
    const int K = slice_group(p->nz * p->nch);
    P.fwd_pitch = fwd_pitch_of(g.nx, g.ny);
    P.fwd_band = g.fwd_band > 0 ? g.fwd_band : fwd_band_auto(P.fwd_pitch, K);
    P.adj_tile_y = g.adj_tile_y;
    P.adj_segw = adj_segw_of(P.adj_tile_y);
    P.adj_achunk = adj_achunk_of(g.na, P.adj_segw, K);
    P.fuse_primal = g.fuse_primal;
    P.ofs = reinterpret_cast<const float4*>(g.ofs_dev);
    P.fwd = g.fwd_dev;
    P.u = s->u; P.ubar = s->ubar; P.p = s->p; P.pbar = s->pbar; P.v = s->v; P.w = s->w;
    P.lam = s->lam; P.ktl = s->ktl; P.ku = s->ku; P.f = s->f; P.mfac = s->mfac;
    P.halo_ubar_hi = s->halo_ubar_hi; P.halo_pbar_lo = s->halo_pbar_lo;
    P.halo_v_lo = s->halo_v_lo; P.halo_w_hi = s->halo_w_hi;
    return P;
}

// A side stream and fork/join events per host thread and device, for
// running the dual step and the p half of the primal step beside the
// projections of the same iteration.  Created on first use and kept until
// the process ends.  Not used while the caller's stream is being captured
// (captured iterations keep the one-stream order).
struct Lane {
    hipStream_t side = nullptr;
    hipEvent_t fork = nullptr, dual_done = nullptr, join = nullptr;
};

static Lane* lane_for(hipStream_t st)
{
    static thread_local Lane lanes[64];
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(st, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone)
        return nullptr;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return nullptr;
    Lane& L = lanes[dev];
    if (L.join) return &L;
    if (hipStreamCreateWithFlags(&L.side, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&L.fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&L.dual_done, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&L.join, hipEventDisableTiming) != hipSuccess) {
        (void)hipGetLastError();
        L = Lane{};
        return nullptr;
    }
    return &L;
}

// One launch with an optional HIP event pair accumulated into *acc.
struct Timer {
    hipEvent_t a = nullptr, b = nullptr;
    bool on = false;
    hipStream_t st;
    Timer(bool enable, hipStream_t s) : on(enable), st(s)
    {
        if (on) {
            (void)hipEventCreate(&a);
            (void)hipEventCreate(&b);
        }
    }
    ~Timer()
    {
        if (on) {
            (void)hipEventDestroy(a);
            (void)hipEventDestroy(b);
        }
    }
    template <class F>
    int run(double* acc, F&& f)
    {
        if (on) (void)hipEventRecord(a, st);
        int rc = f();
        if (on && rc == GR_OK) {
            (void)hipEventRecord(b, st);
            hipError_t e = hipEventSynchronize(b);
            if (e != hipSuccess) return hip_status(e, "tgv_profile");
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, a, b);
            *acc += ms;
        }
        return rc;
    }
};

// Iterations of one problem.  Unfused: dual, forward + prox, backprojection,
// primal per iteration.  With the alternate buffers of the fused form (and
// no primal step inside the backprojection) the primal step of iteration k
// and the dual step of k + 1 run as one kernel:
//   dual_0, fwd_0, adj_0, [fused_k, fwd_k+1, adj_k+1] for k < n-1, primal_n-1
// alternating the state between the caller's arrays and the alternates; an
// unfused first iteration makes the number of fused steps even, so the state
// ends in the caller's arrays.  Both forms compute the same values.
// ms_out (optional, 5 entries): summed ms of dual, forward + prox,
// backprojection, primal and fused launches.
int iterate_single(const gr_geometry* gg, const ProbDev& P, int n_iter, int ku_on_last,
                   hipStream_t st, double* ms_out, const gr_tgv_state* s)
{
    const int nsl = P.nz * P.nch;
    const int pmode = P.disc == GR_DISC_KL ? FWD_KL : FWD_L2;
    double scratch[5] = {0, 0, 0, 0, 0};
    double* ms = ms_out ? ms_out : scratch;
    for (int k = 0; k < 5; ++k) ms[k] = 0.0;
    Timer t(ms_out != nullptr, st);
    auto fwd_mode = [&](int it) {
        return pmode | ((ku_on_last && it == n_iter - 1) ? FWD_WRITE : 0);
    };
    const bool fused = s && s->u2 && s->v2 && (!P.tgv || s->w2) && !P.fuse_primal &&
                       n_iter >= 2;
    int rc = GR_OK;
    int it = 0;
    // unfused iterations: all of them, or one to make the fused count even
    const int plain = fused ? ((n_iter - 1) % 2) : n_iter;
    // Two streams per iteration k (untimed runs):
    //   side: dual_k, then the p half of primal_k (needs only v, w)
    //   main: forward + prox_k, backprojection_k, then (after dual_k) the u
    //         half of primal_k
    // dual_k+1 waits for the u half of primal_k (it reads ubar_k+1); the
    // halves write disjoint arrays, the projections only read ubar and lam.
    Lane* lane = (!ms_out && !P.fuse_primal) ? lane_for(st) : nullptr;
    auto hip_rc = [](hipError_t e, const char* where) {
        return e == hipSuccess ? GR_OK : hip_status(e, where);
    };
    for (; lane && it < plain && rc == GR_OK; ++it) {
        rc = hip_rc(hipEventRecord(lane->fork, st), "fork");
        if (!rc) rc = hip_rc(hipStreamWaitEvent(lane->side, lane->fork, 0), "fork");
        if (!rc) rc = launch_dual(P, lane->side);
        if (!rc) rc = hip_rc(hipEventRecord(lane->dual_done, lane->side), "dual event");
        if (!rc) rc = launch_primal_part(P, PRIMAL_P, lane->side);
        if (!rc)
            rc = launch_fwd(P, gg->g, nsl,
                            fwd_mode(it) | FWD_BESIDE | (gg->g.fwd_band > 0 ? 0 : FWD_AUTOBAND),
                            st);
        if (!rc) rc = launch_adj(P, nsl, false, st, true);
        if (!rc) rc = hip_rc(hipStreamWaitEvent(st, lane->dual_done, 0), "dual wait");
        if (!rc) rc = launch_primal_part(P, PRIMAL_U, st);
    }
    if (lane) {  // the caller's stream ends after everything of the side stream, also on error
        const int jr = hip_rc(hipEventRecord(lane->join, lane->side), "join");
        const int jw = jr ? jr : hip_rc(hipStreamWaitEvent(st, lane->join, 0), "join");
        if (!rc) rc = jw;
    }
    for (; it < plain && rc == GR_OK; ++it) {
        rc = t.run(&ms[0], [&] { return launch_dual(P, st); });
        if (!rc) rc = t.run(&ms[1], [&] { return launch_fwd(P, gg->g, nsl, fwd_mode(it), st); });
        if (!rc) rc = t.run(&ms[2], [&] { return launch_adj(P, nsl, P.fuse_primal != 0, st); });
        if (!rc && !P.fuse_primal) rc = t.run(&ms[3], [&] { return launch_primal(P, st); });
    }
    if (rc || it >= n_iter) return rc;
    // fused form over the remaining n_iter - it iterations
    FusedBufs A{P.u, P.p, P.v, P.w, s->u2, P.pbar, s->v2, s->w2};
    FusedBufs Bk{s->u2, P.pbar, s->v2, s->w2, P.u, P.p, P.v, P.w};
    rc = t.run(&ms[0], [&] { return launch_dual(P, st); });
    if (!rc) rc = t.run(&ms[1], [&] { return launch_fwd(P, gg->g, nsl, fwd_mode(it), st); });
    if (!rc) rc = t.run(&ms[2], [&] { return launch_adj(P, nsl, false, st); });
    for (int k = 0; it + 1 < n_iter && rc == GR_OK; ++k) {
        const FusedBufs& B = (k % 2 == 0) ? A : Bk;
        rc = t.run(&ms[4], [&] { return launch_fused_pd(P, B, st); });
        ++it;
        if (!rc) rc = t.run(&ms[1], [&] { return launch_fwd(P, gg->g, nsl, fwd_mode(it), st); });
        if (!rc) rc = t.run(&ms[2], [&] { return launch_adj(P, nsl, false, st); });
    }
    if (!rc) rc = t.run(&ms[3], [&] { return launch_primal(P, st); });
    return rc;
}

}  // namespace

// --------------------------------------------------------------------------
// batch plans
// --------------------------------------------------------------------------
struct gr_plan {
    int nprob = 0;
    std::vector<ProbDev> probs;
    ProbDev* probs_dev = nullptr;
    struct FwdLaunch { int K, mi; size_t lds; std::vector<FwdItem> items; FwdItem* dev; };
    struct AdjLaunch { int K, dim, tgv, fuse; size_t lds; std::vector<AdjItem> items; AdjItem* dev; };
    struct PrimalLaunch { int dim, tgv; std::vector<DualItem> items; DualItem* dev; };
    struct DualLaunch { int dim, coupling, tgv, maxc; std::vector<DualItem> items; DualItem* dev; };
    std::vector<FwdLaunch> fwd;
    std::vector<AdjLaunch> adj;
    std::vector<DualLaunch> dual;
    std::vector<PrimalLaunch> primal;
    hipStream_t cap_stream = nullptr;
    hipGraphExec_t graph = nullptr;
    int graph_block = 0;
    // the launches of one phase touch disjoint problems: up to PLAN_LANES of
    // them run side by side, forked from the caller's stream and joined back
    // (also inside a captured graph, as parallel branches)
    static constexpr int PLAN_LANES = 4;
    hipStream_t side[PLAN_LANES - 1] = {};
    hipEvent_t fork_ev = nullptr;
    hipEvent_t join_ev[PLAN_LANES - 1] = {};
};

namespace {

void plan_free_device(gr_plan* pl)
{
    if (pl->probs_dev) (void)hipFree(pl->probs_dev);
    pl->probs_dev = nullptr;
    for (auto& l : pl->fwd) if (l.dev) (void)hipFree(l.dev);
    for (auto& l : pl->adj) if (l.dev) (void)hipFree(l.dev);
    for (auto& l : pl->dual) if (l.dev) (void)hipFree(l.dev);
    for (auto& l : pl->primal) if (l.dev) (void)hipFree(l.dev);
    if (pl->graph) (void)hipGraphExecDestroy(pl->graph);
    pl->graph = nullptr;
    pl->graph_block = 0;
}

template <typename T>
int upload(const std::vector<T>& v, T** out)
{
    *out = nullptr;
    if (v.empty()) return GR_OK;
    GR_HIP(hipMalloc(out, sizeof(T) * v.size()));
    GR_HIP(hipMemcpy(*out, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice));
    return GR_OK;
}

int plan_build(gr_plan* pl, const gr_geometry* const* geoms)
{
    plan_free_device(pl);
    pl->fwd.clear();
    pl->adj.clear();
    pl->dual.clear();
    pl->primal.clear();
    const int n = pl->nprob;
    // forward chunk tables: a common target of (bin, line) pairs per
    // work-group (the batch's total over ~3 work-groups per CU); each problem
    // takes the items-per-thread table whose chunks come closest to it
    double total_work = 0.0;
    for (int i = 0; i < n; ++i) {
        const ProbDev& P = pl->probs[i];
        const int S = P.nz * P.nch, K = slice_group(S);
        total_work += (double)((S + K - 1) / K) * P.na * P.nd * std::max(P.nx, P.ny);
    }
    const double target = std::max(1.0, total_work / 768.0);
    std::vector<int> tsel(n, 0);
    for (int i = 0; i < n; ++i) {
        const ProbDev& P = pl->probs[i];
        const Geometry& g = geoms[i]->g;
        double best = 1e30;
        for (int tt = 0; tt < 4; ++tt) {
            const int nc = g.tab_off[tt + 1] - g.tab_off[tt];
            const double per = (double)P.na * P.nd * std::max(P.nx, P.ny) / std::max(nc, 1);
            const double d = std::fabs(std::log(per / target));
            if (d < best - 1e-9) {
                best = d;
                tsel[i] = tt;
            }
        }
    }
    std::map<std::tuple<int, int>, size_t> fwd_idx;
    std::map<std::tuple<int, int, int, int>, size_t> adj_idx;
    std::map<std::tuple<int, int>, size_t> primal_idx;
    std::map<std::tuple<int, int, int, int>, size_t> dual_idx;
    for (int i = 0; i < n; ++i) {
        const ProbDev& P = pl->probs[i];
        const Geometry& g = geoms[i]->g;
        const int S = P.nz * P.nch;
        const int K = slice_group(S);
        // forward items: slice groups x chunks (group-major)
        const int t = tsel[i], mi = 8 >> t;
        auto fk = std::make_tuple(K, mi);
        if (!fwd_idx.count(fk)) {
            fwd_idx[fk] = pl->fwd.size();
            pl->fwd.push_back({K, mi, 0, {}, nullptr});
        }
        auto& fl = pl->fwd[fwd_idx[fk]];
        fl.lds = std::max(fl.lds, fwd_lds_bytes(P, K));
        for (int s0 = 0; s0 < S; s0 += K)
            for (int c = g.tab_off[t]; c < g.tab_off[t + 1]; ++c) {
                const FwdChunk& ch = g.chunks_host[c];
                fl.items.push_back(FwdItem{i, ch.cls, ch.a0, ch.na, ch.b0, ch.nb, s0, 0});
            }
        // backprojection + primal items: tiles x slice groups
        auto ak = std::make_tuple(K, P.dim, P.tgv, P.fuse_primal);
        if (!adj_idx.count(ak)) {
            adj_idx[ak] = pl->adj.size();
            pl->adj.push_back({K, P.dim, P.tgv, P.fuse_primal, 0, {}, nullptr});
        }
        auto& al = pl->adj[adj_idx[ak]];
        al.lds = std::max(al.lds, adj_lds_bytes(P, K));
        for (int s0 = 0; s0 < S; s0 += K)
            for (int y0 = 0; y0 < P.ny; y0 += P.adj_tile_y)
                for (int x0 = 0; x0 < P.nx; x0 += ADJ_TX) al.items.push_back(AdjItem{i, x0, y0, s0});
        // unfused primal items: 64 x 4 tiles of every slice of every channel
        if (!P.fuse_primal) {
            auto pk = std::make_tuple(P.dim, P.tgv);
            if (!primal_idx.count(pk)) {
                primal_idx[pk] = pl->primal.size();
                pl->primal.push_back({P.dim, P.tgv, {}, nullptr});
            }
            auto& plp = pl->primal[primal_idx[pk]];
            for (int sl = 0; sl < S; ++sl)
                for (int y0 = 0; y0 < P.ny; y0 += DUAL_BY)
                    for (int x0 = 0; x0 < P.nx; x0 += DUAL_BX)
                        plp.items.push_back(DualItem{i, x0, y0, sl});
        }
        // dual items: 64 x 4 tiles of every plane
        // 4 / 8: channels held in registers; 9: any larger count, swept twice
        const int maxc = P.nch <= 4 ? 4 : P.nch <= 8 ? 8 : 9;
        auto dk = std::make_tuple(P.dim, P.coupling, P.tgv, maxc);
        if (!dual_idx.count(dk)) {
            dual_idx[dk] = pl->dual.size();
            pl->dual.push_back({P.dim, P.coupling, P.tgv, maxc, {}, nullptr});
        }
        auto& dl = pl->dual[dual_idx[dk]];
        for (int z = 0; z < P.nz; ++z)
            for (int y0 = 0; y0 < P.ny; y0 += DUAL_BY)
                for (int x0 = 0; x0 < P.nx; x0 += DUAL_BX) dl.items.push_back(DualItem{i, x0, y0, z});
    }
    // largest chunks first, so the smallest fill the tail of the launch
    for (auto& l : pl->fwd) {
        auto work = [&](const FwdItem& it) {
            const ProbDev& P = pl->probs[it.prob];
            return (double)it.na * it.nb * (it.cls == 0 ? P.nx : P.ny);
        };
        std::stable_sort(l.items.begin(), l.items.end(),
                         [&](const FwdItem& a, const FwdItem& b) { return work(a) > work(b); });
    }
    for (auto& l : pl->fwd)
        if (l.lds > 160 * 1024) {
            set_error("plan: forward LDS stage of %zu bytes too large", l.lds);
            return GR_ERR_INVALID;
        }
    int rc = upload(pl->probs, &pl->probs_dev);
    for (auto& l : pl->fwd) if (!rc) rc = upload(l.items, &l.dev);
    for (auto& l : pl->adj) if (!rc) rc = upload(l.items, &l.dev);
    for (auto& l : pl->dual) if (!rc) rc = upload(l.items, &l.dev);
    for (auto& l : pl->primal) if (!rc) rc = upload(l.items, &l.dev);
    return rc;
}

int plan_lanes(gr_plan* pl)
{
    if (pl->fork_ev) return GR_OK;
    for (auto& s : pl->side) GR_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    for (auto& e : pl->join_ev) GR_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    GR_HIP(hipEventCreateWithFlags(&pl->fork_ev, hipEventDisableTiming));
    return GR_OK;
}

// Phases of a batch iteration: 0 dual step, 1 forward projection + prox,
// 2 backprojection, 3 primal step.  The iteration runs them in three forks,
// each an LDS / VALU-bound projection beside an HBM-bound stencil step that
// does not touch its arrays:
// - PHASE_DUAL_FWD: the forward projection (reads u-bar, writes lambda) and
//   the dual step (reads u-bar / p-bar, writes v / w);
// - PHASE_ADJ_P: the backprojection (reads lambda, writes K^T lambda; with
//   the fused primal step the whole primal step, for problems that have no
//   primal launch) and the p half of the primal step (reads v / w, writes
//   p / p-bar);
// - PHASE_U: the u half (reads v and K^T lambda, writes u / u-bar).
constexpr int PHASE_DUAL_FWD = 4, PHASE_ADJ_P = 5, PHASE_U = 6;

int plan_phase(gr_plan* pl, int phase, int fwd_mode, hipStream_t st)
{
    // (phase, launch index, primal part) in launch order, with the lane of each
    struct Job { int ph; size_t i; int part; int lane; };
    std::vector<Job> jobs;
    auto add = [&](int ph, size_t n) {
        for (size_t i = 0; i < n; ++i) jobs.push_back({ph, i, PRIMAL_UP, 0});
    };
    if (phase == PHASE_DUAL_FWD) {
        add(1, pl->fwd.size());
        add(0, pl->dual.size());
    } else if (phase == PHASE_ADJ_P) {
        add(2, pl->adj.size());
        for (size_t i = 0; i < pl->primal.size(); ++i)
            if (pl->primal[i].tgv) jobs.push_back({3, i, PRIMAL_P, 0});
    } else if (phase == PHASE_U) {
        for (size_t i = 0; i < pl->primal.size(); ++i)
            jobs.push_back({3, i, pl->primal[i].tgv ? PRIMAL_U : PRIMAL_UP, 0});
    } else {
        add(phase, phase == 0 ? pl->dual.size()
                 : phase == 1 ? pl->fwd.size()
                 : phase == 2 ? pl->adj.size()
                              : pl->primal.size());
    }
    const size_t nl = jobs.size();
    const int lanes = (int)std::min<size_t>(nl, gr_plan::PLAN_LANES);
    {
        const size_t nf = phase == PHASE_DUAL_FWD ? pl->fwd.size()
                        : phase == PHASE_ADJ_P    ? pl->adj.size()
                                                  : nl;
        const bool split = nf > 0 && nf < nl && (int)nl > lanes;
        for (size_t k = 0; k < nl; ++k)
            // more launches than lanes: the projections share all lanes but
            // the last, and the (shorter) stencil launches queue on that one
            jobs[k].lane = !split ? (int)(k % (size_t)std::max(lanes, 1))
                         : k < nf  ? (int)(k % (size_t)(lanes - 1))
                                   : lanes - 1;
    }
    if (lanes > 1) {
        int rc = plan_lanes(pl);
        if (rc) return rc;
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        GR_HIP(hipStreamIsCapturing(st, &cs));
        if (cs == hipStreamCaptureStatusInvalidated) {
            set_error("plan: the capturing stream's capture was invalidated");
            return GR_ERR_HIP;
        }
        GR_HIP(hipEventRecord(pl->fork_ev, st));
        for (int j = 1; j < lanes; ++j) {
            GR_HIP(hipStreamWaitEvent(pl->side[j - 1], pl->fork_ev, 0));
            if (cs == hipStreamCaptureStatusActive) {
                // a side stream must have joined the capture through the fork
                // event (recorded inside it); anything else would leave work
                // or a dependency outside the graph
                hipStreamCaptureStatus ss = hipStreamCaptureStatusNone;
                GR_HIP(hipStreamIsCapturing(pl->side[j - 1], &ss));
                if (ss != hipStreamCaptureStatusActive) {
                    for (int i = 1; i < j; ++i) {  // join the lanes already forked
                        (void)hipEventRecord(pl->join_ev[i - 1], pl->side[i - 1]);
                        (void)hipStreamWaitEvent(st, pl->join_ev[i - 1], 0);
                    }
                    set_error("plan: side stream %d did not join the capture", j);
                    return GR_ERR_UNSUPPORTED;
                }
            }
        }
    }
    int rc = GR_OK;
    for (size_t k = 0; k < nl && !rc; ++k) {
        const int lane = jobs[k].lane, ph = jobs[k].ph, part = jobs[k].part;
        const size_t i = jobs[k].i;
        const hipStream_t s = lane == 0 ? st : pl->side[lane - 1];
        if (ph == 0) {
            const auto& l = pl->dual[i];
            rc = launch_dual_batch(pl->probs_dev, l.dev, (int)l.items.size(), l.dim, l.coupling,
                                   l.tgv != 0, l.maxc, s);
        } else if (ph == 1) {
            const auto& l = pl->fwd[i];
            rc = launch_fwd_batch(pl->probs_dev, l.dev, (int)l.items.size(), l.K, l.mi, l.lds,
                                  fwd_mode, s);
        } else if (ph == 2) {
            const auto& l = pl->adj[i];
            rc = launch_adj_batch(pl->probs_dev, l.dev, (int)l.items.size(), l.K, l.dim,
                                  l.tgv != 0, l.fuse != 0, l.lds, s);
        } else {
            const auto& l = pl->primal[i];
            rc = launch_primal_batch(pl->probs_dev, l.dev, (int)l.items.size(), l.dim,
                                     l.tgv != 0, part, s);
        }
    }
    // join every forked lane, also after a failed launch (a capture must end joined)
    for (int j = 1; j < lanes; ++j) {
        hipError_t e = hipEventRecord(pl->join_ev[j - 1], pl->side[j - 1]);
        if (e == hipSuccess) e = hipStreamWaitEvent(st, pl->join_ev[j - 1], 0);
        if (e != hipSuccess && !rc) rc = hip_status(e, "plan: joining side streams");
    }
    return rc;
}

// Below this many voxel-channels a batch iteration keeps the four phases in
// sequence: its kernels are latency-bound, and the extra fork and launch of
// the three-fork order cost more than the overlap gains (183x183, one
// channel, graph replay: 0.0887 against 0.0821 ms per iteration; 64 2D
// problems, 11.5 M voxel-channels: 1.033 against 1.108 ms).
constexpr double PLAN_OVERLAP_MIN = 1 << 20;

int plan_iter(gr_plan* pl, int n_iter, int ku_on_last, hipStream_t st)
{
    double vc = 0.0;
    for (const auto& P : pl->probs) vc += (double)P.nx * P.ny * P.nz * P.nch;
    const bool overlap = vc >= PLAN_OVERLAP_MIN;
    for (int it = 0; it < n_iter; ++it) {
        const int mode = FWD_PROX | ((ku_on_last && it == n_iter - 1) ? FWD_WRITE : 0);
        int rc = GR_OK;
        if (overlap) {
            rc = plan_phase(pl, PHASE_DUAL_FWD, mode, st);
            if (!rc) rc = plan_phase(pl, PHASE_ADJ_P, mode, st);
            if (!rc) rc = plan_phase(pl, PHASE_U, mode, st);
        } else {
            for (int ph = 0; ph < 4 && !rc; ++ph) rc = plan_phase(pl, ph, mode, st);
        }
        if (rc) return rc;
    }
    return GR_OK;
}

ProbDev standalone(const gr_geometry* gg, int nslices)
{
    const Geometry& g = gg->g;
    ProbDev P{};
    P.dim = 2; P.nx = g.nx; P.ny = g.ny; P.nz = 1; P.nch = nslices; P.nd = g.nd; P.na = g.na;
    P.ofs = reinterpret_cast<const float4*>(g.ofs_dev);
    P.fwd = g.fwd_dev;
    P.fwd_band = g.fwd_band;
    P.adj_tile_y = g.adj_tile_y;
    return P;
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
    // backprojection tiles: 16 rows up to 256 px (more work-groups to share the
    // GPU with the dual step: 0.801 against 0.810 ms per two-stream iteration
    // on 256x256x64x4), 32 above (18.54 against 18.69 ms on 512x512x256x4)
    g.adj_tile_y = std::max(nx, ny) <= 256 ? 16 : 32;
    g.fuse_primal = 0;
    g.ofs_dev = nullptr; g.fwd_dev = nullptr; g.chunks_dev = nullptr;
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
    const int ipt[4] = {8, 4, 2, 1};
    for (int t = 0; t < 4; ++t) {
        g.tab_off[t] = (int)g.chunks_host.size();
        auto c = make_chunks(fa, nd, FWD_THREADS * ipt[t]);
        g.chunks_host.insert(g.chunks_host.end(), c.begin(), c.end());
    }
    g.tab_off[4] = (int)g.chunks_host.size();
    hipError_t e = hipMalloc(&g.ofs_dev, sizeof(float) * 4 * na);
    if (e == hipSuccess) e = hipMalloc(&g.fwd_dev, sizeof(FwdAngle) * na);
    if (e == hipSuccess) e = hipMalloc(&g.chunks_dev, sizeof(FwdChunk) * g.chunks_host.size());
    if (e == hipSuccess)
        e = hipMemcpy(g.ofs_dev, ofs, sizeof(float) * 4 * na, hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(g.fwd_dev, fa.data(), sizeof(FwdAngle) * na, hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(g.chunks_dev, g.chunks_host.data(), sizeof(FwdChunk) * g.chunks_host.size(),
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
        (adj_tile_y && (adj_tile_y < 4 || adj_tile_y > 4 * ADJ_MAXR || (adj_tile_y & 3)))) {
        set_error("set_tuning: fwd_band in [1,256], adj_tile_y multiple of 4 in [4,32]");
        return GR_ERR_INVALID;
    }
    if (fwd_band) gg->g.fwd_band = fwd_band;
    if (adj_tile_y) gg->g.adj_tile_y = adj_tile_y;
    return GR_OK;
}

int gr_geometry_set_fusion(gr_geometry* gg, int fuse_primal)
{
    if (!gg || fuse_primal < 0 || fuse_primal > 1) {
        set_error("set_fusion: invalid arguments");
        return GR_ERR_INVALID;
    }
    gg->g.fuse_primal = fuse_primal;
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
    ProbDev P = standalone(gg, nslices);
    P.ubar = const_cast<float*>(img);
    P.ku = sino;
    return launch_fwd(P, gg->g, nslices, FWD_WRITE, stream_of(stream));
}

int gr_radon_ad(const gr_geometry* gg, float* img, const float* sino, int nslices, void* stream)
{
    if (!gg || !sino || !img || nslices < 0) {
        set_error("radon_ad: invalid arguments");
        return GR_ERR_INVALID;
    }
    ProbDev P = standalone(gg, nslices);
    P.lam = const_cast<float*>(sino);
    P.ktl = img;
    return launch_adj(P, nslices, false, stream_of(stream));
}

int gr_normest(const gr_geometry* gg, const double* x0, int iters, int mode, double tol,
               double* out, void* stream)
{
    if (!gg || !out || iters <= 0 || (mode != GR_NORMEST_REFERENCE && mode != GR_NORMEST_POWER) ||
        (mode == GR_NORMEST_REFERENCE && !x0)) {
        set_error("normest: invalid arguments");
        return GR_ERR_INVALID;
    }
    return launch_normest(gg->g, x0, iters, mode, tol, out, stream_of(stream));
}

int gr_make_feasible(const gr_geometry* gg, float* f, int nslices, void* stream)
{
    if (!gg || !f || nslices < 0) {
        set_error("make_feasible: invalid arguments");
        return GR_ERR_INVALID;
    }
    return launch_feasible(gg->g, f, nslices, stream_of(stream));
}

int gr_tgv_iterate(const gr_geometry* gg, const gr_tgv_params* p, const gr_tgv_state* s,
                   int n_iter, int ku_on_last, void* stream)
{
    if (!valid_params(gg, p) || n_iter < 0 || !valid_state(gg, p, s, ku_on_last != 0)) {
        set_error("tgv_iterate: invalid parameters or missing state buffer");
        return GR_ERR_INVALID;
    }
    if (has_halos(s)) {
        // a slab's halo planes must be refreshed between the dual and the
        // primal step of every iteration, which only the caller can do
        set_error("tgv_iterate: z halo planes need an exchange inside every iteration; "
                  "run slabs through gr_tgv_dual_step / gr_tgv_primal_step");
        return GR_ERR_UNSUPPORTED;
    }
    if (n_iter == 0) return GR_OK;
    const ProbDev P = make_prob(gg, p, s);
    return iterate_single(gg, P, n_iter, ku_on_last, stream_of(stream), nullptr, s);
}

int gr_tgv_dual_step(const gr_geometry* gg, const gr_tgv_params* p, const gr_tgv_state* s,
                     void* stream)
{
    if (!valid_params(gg, p) || !s || !s->ubar || !s->v ||
        (p->regularisation == GR_REG_TGV && (!s->pbar || !s->w)) || !valid_halos(gg, p, s)) {
        set_error("tgv_dual_step: invalid parameters, missing state buffer or halo planes");
        return GR_ERR_INVALID;
    }
    const ProbDev P = make_prob(gg, p, s);
    return launch_dual(P, stream_of(stream));
}

int gr_tgv_primal_step(const gr_geometry* gg, const gr_tgv_params* p, const gr_tgv_state* s,
                       void* stream)
{
    if (!valid_params(gg, p) || !s || !s->u || !s->ubar || !s->v || !s->ktl ||
        (p->regularisation == GR_REG_TGV && (!s->p || !s->pbar || !s->w)) ||
        !valid_halos(gg, p, s)) {
        set_error("tgv_primal_step: invalid parameters, missing state buffer or halo planes");
        return GR_ERR_INVALID;
    }
    const ProbDev P = make_prob(gg, p, s);
    return launch_primal(P, stream_of(stream));
}

int gr_tgv_profile(const gr_geometry* gg, const gr_tgv_params* p, const gr_tgv_state* s,
                   int n_iter, double* ms_out, void* stream)
{
    if (!valid_params(gg, p) || n_iter <= 0 || !ms_out || !valid_state(gg, p, s, false)) {
        set_error("tgv_profile: invalid parameters or missing state buffer");
        return GR_ERR_INVALID;
    }
    if (has_halos(s)) {
        set_error("tgv_profile: z halo planes need an exchange inside every iteration; "
                  "run slabs through gr_tgv_dual_step / gr_tgv_primal_step");
        return GR_ERR_UNSUPPORTED;
    }
    const ProbDev P = make_prob(gg, p, s);
    return iterate_single(gg, P, n_iter, 0, stream_of(stream), ms_out, s);
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
    hipStream_t st = stream_of(stream);
    const int nsl = p->nz * p->nch;
    ProbDev P = make_prob(gg, p, s);
    ProbDev Pf = P;
    Pf.ubar = s->u;  // project u itself
    int rc = launch_fwd(Pf, gg->g, nsl, FWD_WRITE, st);
    if (rc) return rc;
    double data = 0.0, reg = 0.0;
    const size_t per_ch = (size_t)gg->g.nd * gg->g.na * p->nz;
    rc = launch_data_energy(s->ku, s->f, mu, p->nch, per_ch, p->discrepancy, &data, st);
    if (rc) return rc;
    const double scale = p->discrepancy == GR_DISC_KL ? 1.0 / L : 1.0;
    rc = launch_energy(P, alpha0 * scale, alpha1 * scale, &reg, st);
    if (rc) return rc;
    out[0] = data;
    out[1] = reg;
    out[2] = data + reg;
    return GR_OK;
}

int gr_tgv_energy_noalloc(const gr_geometry* gg, const gr_tgv_params* p, const gr_tgv_state* s,
                          const double* mu, double alpha0, double alpha1, double L, double* out,
                          void* stream)
{
    if (!valid_params(gg, p) || !s || !s->u || !s->ku || !s->f || !mu || !out || L <= 0) {
        set_error("tgv_energy_noalloc: invalid arguments");
        return GR_ERR_INVALID;
    }
    if (p->regularisation == GR_REG_TGV && !s->p) {
        set_error("tgv_energy_noalloc: p required for TGV");
        return GR_ERR_INVALID;
    }
    if (p->nch > GR_ENERGY_NOALLOC_MAX_CHANNELS) {
        set_error("tgv_energy_noalloc: %d channels exceed the %d its fixed scratch holds",
                  p->nch, GR_ENERGY_NOALLOC_MAX_CHANNELS);
        return GR_ERR_INVALID;
    }
    hipStream_t st = stream_of(stream);
    ProbDev P = make_prob(gg, p, s);
    ProbDev Pf = P;
    Pf.ubar = s->u;  // project u itself into ku
    int rc = launch_fwd(Pf, gg->g, p->nz * p->nch, FWD_WRITE, st);
    if (rc) return rc;
    const double scale = p->discrepancy == GR_DISC_KL ? 1.0 / L : 1.0;
    const size_t per_ch = (size_t)gg->g.nd * gg->g.na * p->nz;
    double data = 0.0, reg = 0.0;
    rc = launch_energy_noalloc(P, mu, alpha0 * scale, alpha1 * scale, per_ch, &data, &reg, st);
    if (rc) return rc;
    out[0] = data;
    out[1] = reg;
    out[2] = data + reg;
    return GR_OK;
}

int gr_plan_create(int nprob, const gr_geometry* const* geoms, const gr_tgv_params* prms,
                   const gr_tgv_state* states, gr_plan** out)
{
    if (nprob <= 0 || !geoms || !prms || !states || !out) {
        set_error("plan_create: invalid arguments");
        return GR_ERR_INVALID;
    }
    *out = nullptr;
    for (int i = 0; i < nprob; ++i) {
        if (!valid_params(geoms[i], &prms[i]) || !valid_state(geoms[i], &prms[i], &states[i], false)) {
            set_error("plan_create: problem %d has invalid parameters or state", i);
            return GR_ERR_INVALID;
        }
        if (has_halos(&states[i])) {
            set_error("plan_create: problem %d has z halo planes (slabs run singly)", i);
            return GR_ERR_UNSUPPORTED;
        }
    }
    gr_plan* pl = new (std::nothrow) gr_plan();
    if (!pl) {
        set_error("plan_create: host allocation failed");
        return GR_ERR_OOM;
    }
    pl->nprob = nprob;
    for (int i = 0; i < nprob; ++i) pl->probs.push_back(make_prob(geoms[i], &prms[i], &states[i]));
    int rc = plan_build(pl, geoms);
    if (rc) {
        gr_plan_destroy(pl);
        return rc;
    }
    *out = pl;
    return GR_OK;
}

int gr_plan_destroy(gr_plan* pl)
{
    if (!pl) return GR_OK;
    plan_free_device(pl);
    if (pl->cap_stream) (void)hipStreamDestroy(pl->cap_stream);
    for (auto s : pl->side) if (s) (void)hipStreamDestroy(s);
    for (auto e : pl->join_ev) if (e) (void)hipEventDestroy(e);
    if (pl->fork_ev) (void)hipEventDestroy(pl->fork_ev);
    delete pl;
    return GR_OK;
}

int gr_plan_iterate(gr_plan* pl, int n_iter, int ku_on_last, void* stream)
{
    if (!pl || n_iter < 0) {
        set_error("plan_iterate: invalid arguments");
        return GR_ERR_INVALID;
    }
    if (ku_on_last)
        for (const auto& P : pl->probs)
            if (!P.ku) {
                set_error("plan_iterate: ku_on_last needs every problem's ku buffer");
                return GR_ERR_INVALID;
            }
    return plan_iter(pl, n_iter, ku_on_last, stream_of(stream));
}

int gr_plan_iterate_graph(gr_plan* pl, int n_iter, int block, void* stream)
{
    if (!pl || n_iter < 0 || block <= 0) {
        set_error("plan_iterate_graph: invalid arguments");
        return GR_ERR_INVALID;
    }
    hipStream_t st = stream_of(stream);
    {
        // the graph is captured on a private stream and replayed on `stream`;
        // replaying into a stream that is itself being captured would nest a
        // captured graph launch, which this entry point does not do
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        GR_HIP(hipStreamIsCapturing(st, &cs));
        if (cs != hipStreamCaptureStatusNone) {
            set_error("plan_iterate_graph: the caller's stream is being captured; use "
                      "gr_plan_iterate inside a capture");
            return GR_ERR_UNSUPPORTED;
        }
    }
    if (pl->graph && pl->graph_block != block) {
        (void)hipGraphExecDestroy(pl->graph);
        pl->graph = nullptr;
    }
    if (!pl->graph && n_iter >= block) {
        if (!pl->cap_stream)
            GR_HIP(hipStreamCreateWithFlags(&pl->cap_stream, hipStreamNonBlocking));
        GR_HIP(hipStreamBeginCapture(pl->cap_stream, hipStreamCaptureModeRelaxed));
        int rc = plan_iter(pl, block, 0, pl->cap_stream);
        hipGraph_t graph = nullptr;
        hipError_t e = hipStreamEndCapture(pl->cap_stream, &graph);
        if (rc) {
            if (graph) (void)hipGraphDestroy(graph);
            return rc;
        }
        if (e != hipSuccess) return hip_status(e, "plan_iterate_graph capture");
        e = hipGraphInstantiate(&pl->graph, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (e != hipSuccess) {
            pl->graph = nullptr;
            return hip_status(e, "plan_iterate_graph instantiate");
        }
        pl->graph_block = block;
    }
    int done = 0;
    if (pl->graph)
        for (; done + block <= n_iter; done += block) GR_HIP(hipGraphLaunch(pl->graph, st));
    return plan_iter(pl, n_iter - done, 0, st);
}

int gr_plan_profile(gr_plan* pl, int n_iter, double* ms_out, void* stream)
{
    if (!pl || n_iter <= 0 || !ms_out) {
        set_error("plan_profile: invalid arguments");
        return GR_ERR_INVALID;
    }
    hipStream_t st = stream_of(stream);
    hipEvent_t ev[5];
    for (int k = 0; k < 5; ++k) GR_HIP(hipEventCreate(&ev[k]));
    for (int k = 0; k < 4; ++k) ms_out[k] = 0.0;
    int rc = GR_OK;
    for (int it = 0; it < n_iter && !rc; ++it) {
        for (int ph = 0; ph < 4 && !rc; ++ph) {
            (void)hipEventRecord(ev[ph], st);
            rc = plan_phase(pl, ph, FWD_PROX, st);
        }
        (void)hipEventRecord(ev[4], st);
        hipError_t e = hipEventSynchronize(ev[4]);
        if (e != hipSuccess && !rc) rc = hip_status(e, "plan_profile");
        for (int k = 0; k < 4 && !rc; ++k) {
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, ev[k], ev[k + 1]);
            ms_out[k] += ms;
        }
    }
    for (int k = 0; k < 5; ++k) (void)hipEventDestroy(ev[k]);
    return rc;
}

int gr_plan_retune(gr_plan* pl, const gr_geometry* const* geoms)
{
    if (!pl || !geoms) {
        set_error("plan_retune: invalid arguments");
        return GR_ERR_INVALID;
    }
    for (int i = 0; i < pl->nprob; ++i) {
        ProbDev& P = pl->probs[i];
        const Geometry& g = geoms[i]->g;
        const int K = slice_group(P.nz * P.nch);
        P.fwd_band = g.fwd_band > 0 ? g.fwd_band : fwd_band_auto(P.fwd_pitch, K);
        P.adj_tile_y = g.adj_tile_y;
        P.adj_segw = adj_segw_of(P.adj_tile_y);
        P.adj_achunk = adj_achunk_of(g.na, P.adj_segw, K);
        P.fuse_primal = g.fuse_primal;
    }
    return plan_build(pl, geoms);
}

}  // extern "C"
