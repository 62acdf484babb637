// ensemble.cpp -- batched integration of independent small systems
// (sol_ensemble_* in include/solaris_amd.h).
//
// Every system occupies a 64-aligned range of the structure-of-arrays
// storage and one ResSystem descriptor; the workgroup-resident kernel
// (resident.hip) integrates one system per workgroup with the caller
// loop's clipping.  Systems are grouped in buckets of 64 bodies so that a
// launch's block shape fits its systems; per bucket the block size is
// picked by timing the first rounds (the autotuner below) unless
// SOLARIS_AMD_ENSEMBLE_BLOCK fixes it.  Once tuned, the launches of one
// round are captured in a HIP graph and replayed until every system is done.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "sol_internal.h"
#include "resident.h"

using namespace sol;

namespace {
thread_local std::string g_ens_create_error;
}  // namespace

struct sol_ensemble {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    int nsys = 0;
    std::vector<sol_nbodies> nb;
    std::vector<int> slot_of, sys_of;   // external index <-> slot (bucket order)
    std::vector<int64_t> base;          // per slot
    int64_t tot = 0;
    std::vector<double> ipe;            // interactions per evaluation, per slot
    // host mirrors
    std::vector<double> y0_h, mass_h;
    std::vector<ResSystem> sys_h;
    bool dirty = true, stale = false;
    // device
    ResSystem *sys_d = nullptr;
    double *y0 = nullptr, *y = nullptr, *acc0 = nullptr, *yscale = nullptr, *errv = nullptr, *fst = nullptr;
    double *mass = nullptr, *rm3 = nullptr, *nnd = nullptr;
    int32_t *nni = nullptr;
    // a bucket: slots [first, first + count), largest system nmax; block is
    // the chosen block size, cand the sizes still to time
    struct Bucket { int first, count, nmax, block; double best_ms; std::vector<int> cand; };
    std::vector<Bucket> buckets;
    hipGraphExec_t graph = nullptr;
    int graph_integ = -1, graph_steps = 0;

    int fail(int code, const std::string &m) { err = m; return code; }
};

#define ECHECK(e, call)                                                                          \
    do {                                                                                         \
        hipError_t r_ = (call);                                                                  \
        if (r_ != hipSuccess) return (e)->fail(SOL_ERR_DEVICE, std::string(#call " failed: ") + hipGetErrorString(r_)); \
    } while (0)

namespace {

int nmassive(const sol_nbodies &b) { return b.central_body + b.giant_planet + b.rocky_planet + b.proto_planet; }

// interactions per evaluation with the reference's astrocentric ranges
double interactions(const sol_nbodies &b)
{
    const int nm = nmassive(b), n = b.total;
    const int ea = nm + b.super_planetesimal, eb = nm;
    double s = 0.0;
    const int na = nm - 1, nbk = n - nm;
    if (na > 0) s += (double)na * (1.0 + std::max(ea - 1, 0)) - (double)std::max(0, std::min(nm, ea) - 1);
    if (nbk > 0) s += (double)nbk * (1.0 + std::max(eb - 1, 0));
    return s;
}

template <typename T>
int ealloc(sol_ensemble *e, T **p, size_t n)
{
    void *q = nullptr;
    ECHECK(e, hipMalloc(&q, std::max<size_t>(n, 1) * sizeof(T)));
    ECHECK(e, hipMemset(q, 0, std::max<size_t>(n, 1) * sizeof(T)));
    *p = static_cast<T *>(q);
    return SOL_OK;
}

ResArrays arrays(sol_ensemble *e)
{
    ResArrays A{};
    A.tot = e->tot;
    A.y0 = e->y0; A.y = e->y; A.acc0 = e->acc0; A.yscale = e->yscale; A.err = e->errv; A.fst = e->fst;
    A.mass = e->mass; A.rm3 = e->rm3; A.nnd = e->nnd; A.nni = e->nni;
    return A;
}

int upload(sol_ensemble *e)
{
    if (!e->dirty) return SOL_OK;
    ECHECK(e, hipMemcpyAsync(e->y0, e->y0_h.data(), e->y0_h.size() * sizeof(double), hipMemcpyHostToDevice, e->stream));
    ECHECK(e, hipMemcpyAsync(e->mass, e->mass_h.data(), e->mass_h.size() * sizeof(double), hipMemcpyHostToDevice,
                             e->stream));
    ECHECK(e, hipMemcpyAsync(e->sys_d, e->sys_h.data(), e->sys_h.size() * sizeof(ResSystem), hipMemcpyHostToDevice,
                             e->stream));
    ECHECK(e, hipStreamSynchronize(e->stream));
    e->dirty = false;
    return SOL_OK;
}

int download(sol_ensemble *e)
{
    if (!e->stale) return SOL_OK;
    ECHECK(e, hipMemcpyAsync(e->y0_h.data(), e->y0, e->y0_h.size() * sizeof(double), hipMemcpyDeviceToHost, e->stream));
    ECHECK(e, hipStreamSynchronize(e->stream));
    e->stale = false;
    return SOL_OK;
}

bool tuning(const sol_ensemble *e)
{
    for (const auto &b : e->buckets)
        if (!b.cand.empty()) return true;
    return false;
}

// One round: every bucket's systems advance up to `steps` accepted steps.
// While a bucket still has untimed block sizes, its launch uses the next
// one and is timed (the autotuner); the fastest size is kept.  The block
// size only changes how a system's bodies are spread over threads and the
// order of the in-block sums.
int launch_round(sol_ensemble *e, int ik, int steps)
{
    const ResArrays A = arrays(e);
    for (auto &b : e->buckets) {
        if (b.count == 0) continue;
        if (!b.cand.empty()) {
            const int block = b.cand.back();
            b.cand.pop_back();
            hipEvent_t t0, t1;
            ECHECK(e, hipEventCreate(&t0));
            ECHECK(e, hipEventCreate(&t1));
            (void)hipEventRecord(t0, e->stream);
            launch_resident(e->stream, ik, e->sys_d + b.first, b.count, A, steps, block, b.nmax, 0);
            ECHECK(e, hipGetLastError());
            (void)hipEventRecord(t1, e->stream);
            (void)hipEventSynchronize(t1);
            float ms = 0.0f;
            (void)hipEventElapsedTime(&ms, t0, t1);
            (void)hipEventDestroy(t0);
            (void)hipEventDestroy(t1);
            if (b.best_ms <= 0.0 || ms < b.best_ms) { b.best_ms = ms; b.block = block; }
        } else {
            launch_resident(e->stream, ik, e->sys_d + b.first, b.count, A, steps, b.block, b.nmax, 0);
            ECHECK(e, hipGetLastError());
        }
    }
    return SOL_OK;
}

}  // namespace

extern "C" {

const char *sol_ensemble_last_error(const sol_ensemble *e) { return e ? e->err.c_str() : g_ens_create_error.c_str(); }

int sol_ensemble_create(sol_ensemble **out, int32_t nsys, const sol_nbodies *nbs, int32_t device)
{
    if (!out || !nbs || nsys <= 0) { g_ens_create_error = "bad argument"; return SOL_ERR_ARGUMENT; }
    *out = nullptr;
    for (int k = 0; k < nsys; ++k) {
        const sol_nbodies &b = nbs[k];
        const int sum = b.central_body + b.giant_planet + b.rocky_planet + b.proto_planet + b.super_planetesimal +
                        b.planetesimal + b.test_particle;
        if (b.total < 2 || sum != b.total || b.central_body != 1 || b.total > kResidentMaxBodies) {
            g_ens_create_error = "system " + std::to_string(k) + ": counts must add up to total, one central body, "
                                 "2 <= total <= " + std::to_string(kResidentMaxBodies);
            return SOL_ERR_ARGUMENT;
        }
    }
    if (hipSetDevice(device) != hipSuccess) { g_ens_create_error = "hipSetDevice failed"; return SOL_ERR_DEVICE; }
    {
        // the resident kernel's LDS (64 B per body of the largest system)
        // must fit a workgroup's share, with 2 KB left for its fixed arrays
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, device) != hipSuccess) {
            g_ens_create_error = "hipGetDeviceProperties failed";
            return SOL_ERR_DEVICE;
        }
        int nmax = 0;
        for (int k = 0; k < nsys; ++k) nmax = std::max(nmax, nbs[k].total);
        if (resident_lds_bytes(nmax) + 2048 > prop.sharedMemPerBlock) {
            g_ens_create_error = "systems of " + std::to_string(nmax) + " bodies need " +
                                 std::to_string(resident_lds_bytes(nmax)) + " B of LDS per workgroup; the device has " +
                                 std::to_string(prop.sharedMemPerBlock);
            return SOL_ERR_ARGUMENT;
        }
    }
    sol_ensemble *e = new sol_ensemble();
    e->device = device;
    e->nsys = nsys;
    e->nb.assign(nbs, nbs + nsys);
    // bucket order: by size class of 64 bodies, stable
    std::vector<int> order(nsys);
    for (int k = 0; k < nsys; ++k) order[k] = k;
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int b) { return (nbs[a].total - 1) / 64 < (nbs[b].total - 1) / 64; });
    e->slot_of.assign(nsys, 0);
    e->sys_of = order;
    e->base.assign(nsys, 0);
    e->ipe.assign(nsys, 0.0);
    int64_t off = 0;
    for (int s = 0; s < nsys; ++s) {
        const int k = order[s];
        e->slot_of[k] = s;
        e->base[s] = off;
        e->ipe[s] = interactions(nbs[k]);
        off += (nbs[k].total + 63) / 64 * 64;
    }
    e->tot = off;
    for (int s = 0; s < nsys;) {
        const int cls = (nbs[order[s]].total - 1) / 64;
        int t = s, nmax = 0;
        while (t < nsys && (nbs[order[t]].total - 1) / 64 == cls) { nmax = std::max(nmax, nbs[order[t]].total); ++t; }
        sol_ensemble::Bucket bk{s, t - s, nmax, resident_block_size(nmax), 0.0, {}};
        // the exact fit (one body per thread up to 640 bodies) and the powers
        // of two below it, at most 4 bodies per thread (the resident
        // kernel's widest form)
        const int fit = resident_block_size(nmax);
        for (int c : {64, 128, 256, 512})
            if (c < fit && 4 * c >= nmax) bk.cand.push_back(c);
        bk.cand.push_back(fit);
        e->buckets.push_back(bk);
        s = t;
    }
    if (const char *fix = std::getenv("SOLARIS_AMD_ENSEMBLE_BLOCK")) {
        const int b = std::max(64, std::min(512, std::atoi(fix) / 64 * 64));
        for (auto &bk : e->buckets) { bk.block = b; bk.cand.clear(); }
    }
    const size_t tot = (size_t)e->tot;
    int r = SOL_OK;
    if (hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess) {
        g_ens_create_error = "hipStreamCreate failed";
        delete e;
        return SOL_ERR_DEVICE;
    }
    if ((r = ealloc(e, &e->sys_d, nsys)) || (r = ealloc(e, &e->y0, 6 * tot)) || (r = ealloc(e, &e->y, 6 * tot)) ||
        (r = ealloc(e, &e->acc0, 6 * tot)) || (r = ealloc(e, &e->yscale, 6 * tot)) ||
        (r = ealloc(e, &e->errv, 6 * tot)) || (r = ealloc(e, &e->fst, 12 * 6 * tot)) ||
        (r = ealloc(e, &e->mass, tot)) || (r = ealloc(e, &e->rm3, tot)) || (r = ealloc(e, &e->nnd, tot)) ||
        (r = ealloc(e, &e->nni, tot))) {
        g_ens_create_error = e->err;
        sol_ensemble_destroy(e);
        return r;
    }
    e->y0_h.assign(6 * tot, 0.0);
    e->mass_h.assign(tot, 0.0);
    e->sys_h.assign(nsys, ResSystem{});
    for (int s = 0; s < nsys; ++s) {
        const sol_nbodies &b = nbs[order[s]];
        ResSystem &S = e->sys_h[s];
        S.n = b.total;
        S.nm = nmassive(b);
        S.end_a = S.nm + b.super_planetesimal;
        S.end_b = S.nm;
        S.base = (int32_t)e->base[s];
        S.status = RES_DONE;   // until set
        S.clip = 1;
    }
    *out = e;
    return SOL_OK;
}

int sol_ensemble_destroy(sol_ensemble *e)
{
    if (!e) return SOL_OK;
    (void)hipSetDevice(e->device);
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    if (e->graph) (void)hipGraphExecDestroy(e->graph);
    void *ptrs[] = {e->sys_d, e->y0, e->y, e->acc0, e->yscale, e->errv, e->fst, e->mass, e->rm3, e->nnd, e->nni};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
    return SOL_OK;
}

int sol_ensemble_set_system(sol_ensemble *e, int32_t k, const sol_body_params *p, const double *y0,
                            const sol_timeline *tl)
{
    if (!e || !p || !y0 || !tl || k < 0 || k >= e->nsys) return SOL_ERR_ARGUMENT;
    if (download(e)) return SOL_ERR_DEVICE;
    const int s = e->slot_of[k];
    const int n = e->nb[k].total;
    const int64_t b = e->base[s], tot = e->tot;
    if (p->type)
        for (int i = 1; i < n; ++i)
            if (p->type[i] < p->type[i - 1]) return e->fail(SOL_ERR_ARGUMENT, "bodies are not sorted by type");
    for (int i = 0; i < n; ++i) {
        e->mass_h[b + i] = p->mass ? p->mass[i] : 0.0;
        for (int c = 0; c < 6; ++c) e->y0_h[c * tot + b + i] = y0[6 * (size_t)i + c];
    }
    ResSystem &S = e->sys_h[s];
    S.status = RES_RUNNING;
    S.t = tl->time; S.h_next = tl->hNext; S.h_did = tl->hDid; S.length = tl->length; S.output = tl->output;
    S.elapsed = tl->elapsedTime; S.last_save = tl->lastSave; S.last_nsteps = tl->lastNSteps;
    S.millenium = tl->millenium;
    S.steps = 0; S.attempts = 0;
    e->dirty = true;
    return SOL_OK;
}

int sol_ensemble_run(sol_ensemble *e, int32_t integrator, double eps, int64_t max_steps, sol_run_stats *stats)
{
    if (!e) return SOL_ERR_ARGUMENT;
    const int ik = integrator == SOL_RUNGE_KUTTA_FEHLBERG78 ? INTEG_RKF78 : integrator == SOL_RUNGE_KUTTA4 ? INTEG_RK4
                 : integrator == SOL_DORMAND_PRINCE ? INTEG_DP : -1;
    if (ik < 0) return e->fail(SOL_ERR_UNSUPPORTED, "integrator not provided");
    ECHECK(e, hipSetDevice(e->device));
    for (auto &S : e->sys_h) S.eps = eps;
    e->dirty = true;
    if (int r = upload(e)) return r;
    std::vector<int64_t> s0(e->nsys), a0(e->nsys);
    for (int s = 0; s < e->nsys; ++s) { s0[s] = e->sys_h[s].steps; a0[s] = e->sys_h[s].attempts; }
    // steps per launch: about 0.1 s of the largest bucket's work
    int nmax = 0;
    for (auto &b : e->buckets) nmax = std::max(nmax, b.nmax);
    const double per_step = 13.0 * (double)nmax * nmax * (double)e->nsys / 2.0e11 + 2.0e-5;
    int chunk = (int)std::max<int64_t>(1, std::min<int64_t>(4096, (int64_t)(0.1 / per_step)));
    if (max_steps > 0) chunk = (int)std::min<int64_t>(chunk, max_steps);
    hipEvent_t t0, t1;
    ECHECK(e, hipEventCreate(&t0));
    ECHECK(e, hipEventCreate(&t1));
    ECHECK(e, hipEventRecord(t0, e->stream));
    int64_t done_steps = 0;
    for (;;) {
        int k = chunk;
        if (max_steps > 0) k = (int)std::min<int64_t>(k, max_steps - done_steps);
        if (tuning(e) || max_steps > 0) {
            if (int r = launch_round(e, ik, k)) return r;
        } else {
            if (!e->graph || e->graph_integ != ik || e->graph_steps != k) {
                if (e->graph) { (void)hipGraphExecDestroy(e->graph); e->graph = nullptr; }
                hipGraph_t g;
                ECHECK(e, hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
                int r = launch_round(e, ik, k);
                hipError_t ce = hipStreamEndCapture(e->stream, &g);
                if (r) return r;
                if (ce != hipSuccess) return e->fail(SOL_ERR_DEVICE, "graph capture failed");
                ECHECK(e, hipGraphInstantiate(&e->graph, g, nullptr, nullptr, 0));
                (void)hipGraphDestroy(g);
                e->graph_integ = ik;
                e->graph_steps = k;
            }
            ECHECK(e, hipGraphLaunch(e->graph, e->stream));
        }
        done_steps += k;
        ECHECK(e, hipMemcpyAsync(e->sys_h.data(), e->sys_d, e->sys_h.size() * sizeof(ResSystem), hipMemcpyDeviceToHost,
                                 e->stream));
        ECHECK(e, hipStreamSynchronize(e->stream));
        const bool any = std::any_of(e->sys_h.begin(), e->sys_h.end(), [](const ResSystem &S) { return S.status == RES_RUNNING; });
        if (!any) break;
        if (max_steps > 0 && done_steps >= max_steps) break;
    }
    ECHECK(e, hipEventRecord(t1, e->stream));
    ECHECK(e, hipEventSynchronize(t1));
    float ms = 0.0f;
    (void)hipEventElapsedTime(&ms, t0, t1);
    (void)hipEventDestroy(t0);
    (void)hipEventDestroy(t1);
    e->stale = true;
    if (stats) {
        const int per = ik == INTEG_RKF78 ? 12 : ik == INTEG_RK4 ? 3 : 8;
        *stats = sol_run_stats{};
        for (int s = 0; s < e->nsys; ++s) {
            const int64_t st = e->sys_h[s].steps - s0[s], at = e->sys_h[s].attempts - a0[s];
            stats->accepted_steps += st;
            stats->attempts += at;
            const int64_t ev = st + at * per;
            stats->force_evaluations += ev;
            stats->interactions += (double)ev * e->ipe[s];
        }
        stats->device_ms = ms;
    }
    return SOL_OK;
}

int sol_ensemble_get_system(sol_ensemble *e, int32_t k, double *y0, sol_timeline *tl, int32_t *status,
                            int64_t *steps)
{
    if (!e || k < 0 || k >= e->nsys) return SOL_ERR_ARGUMENT;
    ECHECK(e, hipSetDevice(e->device));
    if (int r = download(e)) return r;
    const int s = e->slot_of[k];
    const int n = e->nb[k].total;
    const int64_t b = e->base[s], tot = e->tot;
    if (y0)
        for (int i = 0; i < n; ++i)
            for (int c = 0; c < 6; ++c) y0[6 * (size_t)i + c] = e->y0_h[c * tot + b + i];
    const ResSystem &S = e->sys_h[s];
    if (tl) {
        tl->time = S.t; tl->hNext = S.h_next; tl->hDid = S.h_did; tl->length = S.length; tl->output = S.output;
        tl->elapsedTime = S.elapsed; tl->lastSave = S.last_save; tl->lastNSteps = S.last_nsteps;
        tl->millenium = S.millenium;
    }
    if (status) *status = S.status;
    if (steps) *steps = S.steps;
    return SOL_OK;
}

}  // extern "C"
