// Continuous-batching server for Poisson / backward-Euler heat solves.
//
// Requests are grouped by operator (grid shape, spacing, BCs, diagonal
// shift). Each group owns one resident multigrid hierarchy with B request
// slots whose vectors are stored slot after slot, so every smoothing,
// restriction, prolongation and Krylov kernel of a V-cycle runs all of the
// group's slots in one launch (blockIdx.y = slot, a device mask selects the
// slots taking part). Groups stay resident in HBM after their slots drain
// and are reused by later requests of the same shape; an empty group is
// evicted only when a new shape needs its memory.
//
// Each slot runs FP32 flexible CG with FP64 reliable updates: the FP64
// solution x and residual r = b - A x are refreshed on the device whenever the
// FP32 recurrence residual has dropped by a fixed factor (the Krylov direction
// continues), and the device stops a slot once its true FP64 residual meets
// the tolerance. One scheduling round
//   1. admits queued requests into free slots while the memory budget allows,
//   2. starts the newly admitted slots of each group (batched: ||b||, the true
//      residual, the first preconditioned direction),
//   3. replays the group's captured hipGraph of two batched iterations (1-3
//      times by problem size); stopped slots skip the work inside the launches,
//   4. reads every slot's scalars with one copy per group, after one wait for
//      all groups, and retires the slots the device has stopped.
// Requests that arrive while others run are admitted in the next round, so
// their work overlaps on the device with requests already in flight.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <deque>
#include <map>
#include <memory>

#include "gmg.hpp"

using namespace opf;

namespace {

enum Phase { FREE = 0, NEWREQ, ITER };

struct Pending {
    uint64_t id;
    opf_srv_request req;
    std::vector<double> rhs;
    double t_submit;
};

struct Slot {
    int phase = FREE;
    uint64_t id = 0;
    double t_submit = 0, t_admit = 0;
    double* out_x = nullptr;
};

struct Group {
    opf_poisson_desc d{};
    Hierarchy H;
    int B = 1;
    long long N = 0;
    double* x64 = nullptr;
    double* b64 = nullptr;
    int* hnew = nullptr;    // pinned row of new-slot flags, copied to H.mk[MK_NEW]
    std::vector<Slot> slots;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipStream_t st = nullptr;  // the group's own stream: groups overlap on the device
    int replays = 1;           // graph replays (2 inner iterations each) per round
    size_t bytes = 0;
    double last_used = 0;
    int busy() const {
        int c = 0;
        for (auto& s : slots) c += s.phase != FREE;
        return c;
    }
};

bool same_op(const opf_poisson_desc& a, const opf_poisson_desc& b) {
    if (a.dim != b.dim || a.sigma != b.sigma) return false;
    for (int q = 0; q < 3; ++q) {
        if (a.n[q] != b.n[q] || a.h[q] != b.h[q]) return false;
        if (a.bc[q][0] != b.bc[q][0] || a.bc[q][1] != b.bc[q][1]) return false;
    }
    return true;
}

}  // namespace

struct opf_srv {
    opf_ctx* ctx = nullptr;
    opf_srv_config cfg{};
    std::deque<Pending> queue;
    std::vector<std::unique_ptr<Group>> groups;
    std::vector<opf_srv_result> done;
    std::vector<hipStream_t> done_streams;  // stream carrying each done result's out_x copy
    uint64_t next_id = 1;
    size_t used = 0;
    int inflight = 0;
    std::chrono::steady_clock::time_point t0;
    double now() const { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); }
};

namespace {

int slots_for(const opf_srv& s, long long N) {
    int B = N <= (1 << 18) ? 8 : N <= (1 << 21) ? 4 : N <= (1 << 23) ? 2 : 1;
    return std::max(1, std::min(B, s.cfg.max_inflight));
}

size_t group_bytes(const opf_poisson_desc& d, int precision, int B, long long N) {
    return opf_gmg_footprint(&d, precision) * (size_t)B + (size_t)B * N * 16 + 4096;
}

void free_group(Group* g) {
    if (g->st) hipStreamSynchronize(g->st);
    if (g->exec) hipGraphExecDestroy(g->exec);
    if (g->graph) hipGraphDestroy(g->graph);
    gmg_free(g->H);
    if (g->x64) hipFree(g->x64);
    if (g->b64) hipFree(g->b64);
    if (g->hnew) hipHostFree(g->hnew);
    if (g->st) hipStreamDestroy(g->st);
}

// Inner iterations per round by problem size only (never by the mix, so a
// request's iteration sequence, and its result, do not depend on other requests)
int replays_for(long long N) { return N <= (1 << 16) ? 3 : N <= (1 << 20) ? 2 : 1; }

int capture(opf_srv* s, Group* g);

int make_group(opf_srv* s, const opf_poisson_desc& d, Group** out) {
    int per[3] = {0, 0, 0};
    double alpha[3][2];
    GLevelHost L{};
    for (int q = 0; q < 3; ++q) L.n[q] = q < d.dim ? d.n[q] : 1;
    L.N = (long long)L.n[0] * L.n[1] * L.n[2];
    L.var = false;
    L.sigma = d.sigma;
    for (int q = 0; q < 3; ++q) {
        alpha[q][0] = alpha[q][1] = 1.0;
        L.a[q] = q < d.dim ? 1.0 / (d.h[q] * d.h[q]) : 0.0;
        if (q >= d.dim) continue;
        for (int side = 0; side < 2; ++side) {
            OPF_CHECK(d.bc[q][side] == OPF_BC_DIRC || d.bc[q][side] == OPF_BC_NEUM || d.bc[q][side] == OPF_BC_PERIODIC,
                      OPF_ERR_UNSUPPORTED, "serve: bc unsupported");
            alpha[q][side] = d.bc[q][side] == OPF_BC_DIRC ? -1.0 : 1.0;
        }
        per[q] = d.bc[q][0] == OPF_BC_PERIODIC;
    }
    auto g = std::make_unique<Group>();
    g->d = d;
    g->N = L.N;
    g->B = slots_for(*s, L.N);
    g->replays = replays_for(L.N);
    OPF_HIP(hipStreamCreate(&g->st));  // default flags (faster dispatch than a non-blocking stream)
    int prec = s->cfg.precision == OPF_PREC_MIXED ? OPF_PREC_MIXED : OPF_PREC_FP64;
    int rc = gmg_build(s->ctx, g->H, L, per, alpha, prec, g->B);
    if (rc != OPF_OK) {
        free_group(g.get());
        return rc;
    }
    if (g->H.singular) {
        free_group(g.get());
        set_error("serve: singular operators (no Dirichlet face, sigma = 0) are not served");
        return OPF_ERR_UNSUPPORTED;
    }
    const size_t BN = (size_t)g->B * g->N;
    if (hipMalloc(&g->x64, 8 * BN) != hipSuccess || hipMalloc(&g->b64, 8 * BN) != hipSuccess ||
        hipHostMalloc(&g->hnew, sizeof(int) * g->B, hipHostMallocDefault) != hipSuccess) {
        free_group(g.get());
        set_error("serve: out of device memory");
        return OPF_ERR_NOMEM;
    }
    g->slots.assign(g->B, Slot{});
    g->bytes = g->H.bytes + 16 * BN;
    {
        // capture the group's iteration graph now, while its stream is idle
        int rc2 = capture(s, g.get());
        if (rc2 != OPF_OK) {
            free_group(g.get());
            return rc2;
        }
    }
    *out = g.get();
    s->used += g->bytes;
    s->groups.push_back(std::move(g));
    return OPF_OK;
}

int max_iter(const opf_srv* s) { return s->cfg.max_iter > 0 ? s->cfg.max_iter : 200; }

// two batched iterations (the p ping-pong returns to its start)
void iterate2(opf_srv* s, Group* g, hipStream_t st) {
    for (int pc = 0; pc < 2; ++pc) {
        if (g->H.precision == OPF_PREC_MIXED) ph_rel_iter<float>(g->H, st, g->x64, g->b64, pc, max_iter(s));
        else ph_rel_iter<double>(g->H, st, g->x64, g->b64, pc, max_iter(s));
    }
}

// capture iterate2 as the group's graph
int capture(opf_srv* s, Group* g) {
    if (!s->cfg.batched || g->exec) return OPF_OK;
    hipStream_t st = g->st;
    OPF_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    iterate2(s, g, st);
    OPF_HIP(hipStreamEndCapture(st, &g->graph));
    OPF_HIP(hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0));
    return OPF_OK;
}

int try_admit(opf_srv* s) {
    for (auto it = s->queue.begin(); it != s->queue.end();) {
        if (s->inflight >= s->cfg.max_inflight) break;
        Pending& p = *it;
        Group* g = nullptr;
        for (auto& gp : s->groups)
            if (same_op(gp->d, p.req.desc) && gp->busy() < gp->B) {
                g = gp.get();
                break;
            }
        if (!g) {
            long long N = 1;
            for (int q = 0; q < p.req.desc.dim; ++q) N *= p.req.desc.n[q];
            int B = slots_for(*s, N);
            size_t need = group_bytes(p.req.desc, s->cfg.precision, B, N);
            // evict idle groups (least recently used first) until the new one fits
            while (s->used + need > s->cfg.mem_budget) {
                int victim = -1;
                double oldest = 1e300;
                for (size_t q = 0; q < s->groups.size(); ++q)
                    if (s->groups[q]->busy() == 0 && s->groups[q]->last_used < oldest) {
                        oldest = s->groups[q]->last_used;
                        victim = (int)q;
                    }
                if (victim < 0) break;
                s->used -= s->groups[victim]->bytes;
                free_group(s->groups[victim].get());  // waits for its stream (pending out_x copies too)
                for (auto& ds : s->done_streams)
                    if (ds == s->groups[victim]->st) ds = nullptr;
                s->groups.erase(s->groups.begin() + victim);
            }
            if (s->used + need > s->cfg.mem_budget) {
                OPF_CHECK(!s->groups.empty() || s->inflight > 0, OPF_ERR_NOMEM,
                          "serve: request needs %zu bytes, budget %zu", need, s->cfg.mem_budget);
                ++it;  // does not fit now: later requests may
                continue;
            }
            OPF_TRY(make_group(s, p.req.desc, &g));
        }
        int k = 0;
        while (g->slots[k].phase != FREE) ++k;
        Slot& sl = g->slots[k];
        sl = Slot{};
        sl.phase = NEWREQ;
        sl.id = p.id;
        sl.t_submit = p.t_submit;
        sl.t_admit = s->now();
        sl.out_x = p.req.out_x;
        hipStream_t st = g->st;
        double* bk = g->b64 + (size_t)k * g->N;
        OPF_HIP(hipMemsetAsync(g->x64 + (size_t)k * g->N, 0, 8 * g->N, st));
        if (p.rhs.empty()) {
            // synthetic request: its bumps are generated in place on the device
            OPF_TRY(gen_bumps_grid_dev(st, p.req.desc.dim, p.req.desc.n, p.req.seed, p.req.nbumps, p.req.noise, bk));
        } else {
            OPF_HIP(hipMemcpyAsync(bk, p.rhs.data(), 8 * g->N, hipMemcpyHostToDevice, st));
            OPF_HIP(hipStreamSynchronize(st));  // the host RHS buffer is released below
        }
        ++s->inflight;
        it = s->queue.erase(it);
    }
    return OPF_OK;
}

int round(opf_srv* s, int* nfinished) {
    *nfinished = 0;
    OPF_TRY(try_admit(s));
    const double tol = s->cfg.tol > 0 ? s->cfg.tol : 1e-8;
    // ---- enqueue one round for every group with work, each on its own stream:
    // start the new slots -> graph replays -> scalars to the host
    std::vector<Group*> active;
    for (auto& gp : s->groups) {
        Group* g = gp.get();
        bool fresh = false, busy = false;
        for (int k = 0; k < g->B; ++k) {
            g->hnew[k] = g->slots[k].phase == NEWREQ;
            fresh |= g->hnew[k] != 0;
            busy |= g->slots[k].phase != FREE;
        }
        if (!busy) continue;
        active.push_back(g);
        hipStream_t st = g->st;
        if (fresh) {
            OPF_HIP(hipMemcpyAsync(g->H.mk + MK_NEW * g->B, g->hnew, sizeof(int) * g->B, hipMemcpyHostToDevice, st));
            // admitted slots start from x = 0 (zeroed at admission)
            if (g->H.precision == OPF_PREC_MIXED)
                ph_rel_start<float>(g->H, st, g->x64, g->b64, tol * tol, rel_delta2<float>(), true);
            else
                ph_rel_start<double>(g->H, st, g->x64, g->b64, tol * tol, rel_delta2<double>(), true);
        }
        for (int r = 0; r < g->replays; ++r) {
            if (s->cfg.batched) OPF_HIP(hipGraphLaunch(g->exec, st));
            else iterate2(s, g, st);
        }
        OPF_HIP(hipMemcpyAsync(g->H.hscal, g->H.scal, sizeof(double) * GMG_SSTRIDE * g->B, hipMemcpyDeviceToHost, st));
    }
    for (Group* g : active) OPF_HIP(hipStreamSynchronize(g->st));
    // ---- retire the slots the device has stopped
    for (Group* g : active) {
        hipStream_t st = g->st;
        for (int k = 0; k < g->B; ++k) {
            Slot& sl = g->slots[k];
            if (sl.phase == FREE) continue;
            const double* sc = g->H.hscal + (size_t)k * GMG_SSTRIDE;
            if (sc[S_DONE] == 0.0) {
                sl.phase = ITER;
                continue;
            }
            // finished: report, copy out (completes before the result is polled), free the slot
            const double bb = sc[S_BB], rr = sc[S_RR];
            opf_srv_result r{};
            r.id = sl.id;
            r.status = OPF_OK;
            r.state.niter = (int)sc[S_ITS];
            r.state.nrefine = (int)sc[S_NREP];
            r.state.abserr = std::sqrt(rr);
            r.state.relerr = bb > 0 ? std::sqrt(rr / bb) : 0.0;
            r.state.converged = r.state.relerr <= tol * (1 + 1e-12) || bb == 0;
            r.t_submit = sl.t_submit;
            r.t_admit = sl.t_admit;
            if (sl.out_x)
                OPF_HIP(hipMemcpyAsync(sl.out_x, g->x64 + (size_t)k * g->N, 8 * g->N, hipMemcpyDeviceToHost, st));
            r.t_done = s->now();
            s->done.push_back(r);
            s->done_streams.push_back(sl.out_x ? st : nullptr);
            sl.phase = FREE;
            --s->inflight;
            ++*nfinished;
            g->last_used = s->now();
        }
    }
    OPF_HIP(hipGetLastError());
    return OPF_OK;
}

}  // namespace

extern "C" int opf_gen_bumps_grid(const opf_poisson_desc* d, uint64_t seed, int nbumps, double noise, double* out) {
    OPF_CHECK(d && out && d->dim >= 1 && d->dim <= 3, OPF_ERR_ARG, "opf_gen_bumps_grid: bad args");
    gen_bumps_grid(d->dim, d->n, seed, nbumps, noise, out);
    return OPF_OK;
}

extern "C" int opf_srv_create(opf_ctx* ctx, const opf_srv_config* cfg, opf_srv** out) {
    OPF_CHECK(ctx && cfg && out && cfg->max_inflight >= 1 && cfg->mem_budget > 0, OPF_ERR_ARG,
              "opf_srv_create: bad args");
    auto* s = new opf_srv();
    s->ctx = ctx;
    s->cfg = *cfg;
    s->t0 = std::chrono::steady_clock::now();
    *out = s;
    return OPF_OK;
}

extern "C" int opf_srv_submit(opf_srv* s, const opf_srv_request* req, uint64_t* id) {
    OPF_CHECK(s && req && id, OPF_ERR_ARG, "opf_srv_submit: null arg");
    OPF_CHECK(req->desc.dim == 2 || req->desc.dim == 3, OPF_ERR_ARG, "opf_srv_submit: dim 2 or 3");
    Pending p;
    p.id = s->next_id++;
    p.req = *req;
    long long N = 1;
    for (int q = 0; q < req->desc.dim; ++q) {
        OPF_CHECK(req->desc.n[q] >= 2, OPF_ERR_ARG, "opf_srv_submit: n >= 2");
        N *= req->desc.n[q];
    }
    if (req->rhs) {
        p.rhs.assign(req->rhs, req->rhs + N);
    } else {
        OPF_CHECK(req->nbumps >= 0 && req->nbumps <= 32, OPF_ERR_ARG, "opf_srv_submit: 0..32 bumps");
    }
    p.t_submit = req->t_submit >= 0 ? req->t_submit : s->now();
    s->queue.push_back(std::move(p));
    *id = s->queue.back().id;
    return OPF_OK;
}

extern "C" int opf_srv_step(opf_srv* s, int* nfinished) {
    OPF_CHECK(s && nfinished, OPF_ERR_ARG, "opf_srv_step: null arg");
    return round(s, nfinished);
}

extern "C" int opf_srv_poll(opf_srv* s, opf_srv_result* out, int max, int* n) {
    OPF_CHECK(s && n && (out || max == 0), OPF_ERR_ARG, "opf_srv_poll: null arg");
    int k = std::min<int>(max, (int)s->done.size());
    for (int q = 0; q < k; ++q) {
        if (s->done_streams[q]) OPF_HIP(hipStreamSynchronize(s->done_streams[q]));  // its out_x copy
        out[q] = s->done[q];
    }
    s->done.erase(s->done.begin(), s->done.begin() + k);
    s->done_streams.erase(s->done_streams.begin(), s->done_streams.begin() + k);
    *n = k;
    return OPF_OK;
}

extern "C" int opf_srv_pending(opf_srv* s, int* queued, int* inflight) {
    OPF_CHECK(s && queued && inflight, OPF_ERR_ARG, "opf_srv_pending: null arg");
    *queued = (int)s->queue.size();
    *inflight = s->inflight;
    return OPF_OK;
}

extern "C" double opf_srv_clock(opf_srv* s) { return s ? s->now() : 0.0; }

extern "C" int opf_srv_destroy(opf_srv* s) {
    if (!s) return OPF_OK;
    for (auto& g : s->groups) free_group(g.get());
    delete s;
    return OPF_OK;
}
