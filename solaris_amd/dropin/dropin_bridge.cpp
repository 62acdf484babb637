// dropin_bridge.cpp -- see dropin_bridge.h.
#include "dropin_bridge.h"

#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "Acceleration.h"
#include "BodyData.h"
#include "Nebula.h"
#include "TimeLine.h"

namespace solaris_dropin {
namespace {

struct Entry {
    sol_context *ctx = nullptr;
    int n = -1;
    int rm3_cap = 0;          // length of the Acceleration::rm3 array this bridge allocated
    bool barycentric = false;
    bool has_gas = false;
    sol_gas_component gas{};
    const Nebula *gas_from = nullptr;
};

std::mutex g_mu;
std::map<const Acceleration *, Entry> g_table;
thread_local std::string g_err;

int fail(const char *what, sol_context *ctx)
{
    g_err = std::string(what) + ": " + sol_last_error(ctx);
    return 1;
}

int device_index()
{
    const char *s = std::getenv("SOLARIS_AMD_DEVICE");
    return s ? std::atoi(s) : 0;
}

sol_nbodies counts(BodyData *bd)
{
    sol_nbodies nb{};
    nb.central_body = bd->nBodies.centralBody;
    nb.giant_planet = bd->nBodies.giantPlanet;
    nb.rocky_planet = bd->nBodies.rockyPlanet;
    nb.proto_planet = bd->nBodies.protoPlanet;
    nb.super_planetesimal = bd->nBodies.superPlanetsimal;
    nb.planetesimal = bd->nBodies.planetsimal;
    nb.test_particle = bd->nBodies.testParticle;
    nb.total = bd->nBodies.total;
    return nb;
}

sol_gas_component gas_of(const Nebula *nb)
{
    const GasComponent &g = nb->gasComponent;
    sol_gas_component s{};
    s.alpha = g.alpha;
    s.mean_molecular_weight = g.meanMolecularWeight;
    s.particle_diameter = g.particleDiameter;
    s.type = (int32_t)g.type;
    s.time_scale = g.timeScale;
    s.t0 = g.t0;
    s.t1 = g.t1;
    s.inner_edge = g.innerEdge;
    s.a = g.a;
    s.eta = {g.eta.c, g.eta.index};
    s.tau = {g.tau.c, g.tau.index};
    s.scale_height = {g.scaleHeight.c, g.scaleHeight.index};
    s.density = {g.density.c, g.density.index};
    s.mean_free_path = {g.meanFreePath.c, g.meanFreePath.index};
    return s;
}

sol_timeline to_sol(const TimeLine *t)
{
    sol_timeline s{};
    s.start = t->start; s.length = t->length; s.output = t->output; s.save = t->save;
    s.time = t->time; s.hNext = t->hNext; s.hDid = t->hDid; s.elapsedTime = t->elapsedTime;
    s.lastSave = t->lastSave; s.lastNSteps = t->lastNSteps; s.millenium = t->millenium;
    return s;
}

void from_sol(const sol_timeline &s, TimeLine *t)
{
    t->time = s.time; t->hNext = s.hNext; t->hDid = s.hDid;
}

}  // namespace

void Register(Acceleration *acc, bool barycentric)
{
    std::lock_guard<std::mutex> lk(g_mu);
    Entry &e = g_table[acc];
    e.barycentric = barycentric;
}

void Forget(Acceleration *acc)
{
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_table.find(acc);
    if (it == g_table.end()) return;
    if (it->second.ctx) sol_destroy(it->second.ctx);
    g_table.erase(it);
}

const char *LastError() { return g_err.c_str(); }

sol_context *Prepare(Acceleration *acc)
{
    BodyData *bd = acc->bodyData;
    if (!bd) { g_err = "Acceleration has no BodyData"; return nullptr; }
    std::lock_guard<std::mutex> lk(g_mu);
    Entry &e = g_table[acc];
    const int n = bd->nBodies.total;
    if (!e.ctx || e.n != n) {
        if (e.ctx) sol_destroy(e.ctx);
        e.ctx = nullptr;
        sol_nbodies nb = counts(bd);
        int r = sol_create(&e.ctx, &nb, e.barycentric ? 1 : 0, device_index());
        if (r) { fail("sol_create", nullptr); e.ctx = nullptr; return nullptr; }
        e.n = n;
        e.has_gas = false;
        e.gas_from = nullptr;
        if (acc->nebula) {
            e.gas = gas_of(acc->nebula);
            if (sol_set_nebula(e.ctx, &e.gas)) { fail("sol_set_nebula", e.ctx); return nullptr; }
            e.has_gas = true;
            e.gas_from = acc->nebula;
        }
    } else {
        // the gas is re-sent only when it changed: a new upload clears the
        // cached drag and migration accelerations
        const bool want = acc->nebula != nullptr;
        sol_gas_component g{};
        if (want) g = gas_of(acc->nebula);
        if (want != e.has_gas || (want && (acc->nebula != e.gas_from || std::memcmp(&g, &e.gas, sizeof g) != 0))) {
            if (sol_set_nebula(e.ctx, want ? &g : nullptr)) { fail("sol_set_nebula", e.ctx); return nullptr; }
            e.gas = g;
            e.has_gas = want;
            e.gas_from = acc->nebula;
        }
    }
    sol_body_params p{};
    p.mass = bd->mass; p.radius = bd->radius; p.density = bd->density; p.cd = bd->cD;
    p.gamma_stokes = bd->gammaStokes; p.gamma_epstein = bd->gammaEpstein; p.mig_stop_at = bd->migStopAt;
    std::vector<int32_t> type(bd->type, bd->type + n), mig(bd->migType, bd->migType + n);
    p.type = type.data();
    p.mig_type = mig.data();
    if (sol_set_bodies(e.ctx, &p)) { fail("sol_set_bodies", e.ctx); return nullptr; }
    if (sol_set_evaluate(e.ctx, acc->evaluateGasDrag, acc->evaluateTypeIMigration, acc->evaluateTypeIIMigration)) {
        fail("sol_set_evaluate", e.ctx);
        return nullptr;
    }
    return e.ctx;
}

int FetchSideOutputs(Acceleration *acc, sol_context *ctx)
{
    BodyData *bd = acc->bodyData;
    const int n = bd->nBodies.total;
    // Acceleration::rm3 is allocated on first use, as the reference does
    // (Solaris/Acceleration.cpp:65-69).  The reference never grows it; the
    // bridge reallocates it if a caller adds bodies, so that writing n
    // values stays in bounds.
    {
        std::lock_guard<std::mutex> lk(g_mu);
        Entry &e = g_table[acc];
        if (!acc->rm3 || e.rm3_cap < n) {
            delete[] acc->rm3;
            acc->rm3 = new double[n];
            std::memset(acc->rm3, 0, n * sizeof(double));
            e.rm3_cap = n;
        }
    }
    std::vector<int32_t> nn(n), mig(n);
    if (sol_get_side_outputs(ctx, acc->rm3, nn.data(), bd->distanceOfNN, mig.data()))
        return fail("sol_get_side_outputs", ctx);
    for (int i = 0; i < n; ++i) { bd->indexOfNN[i] = nn[i]; bd->migType[i] = mig[i]; }
    int32_t d, m1, m2;
    sol_get_evaluate(ctx, &d, &m1, &m2);
    acc->evaluateGasDrag = d != 0;
    acc->evaluateTypeIMigration = m1 != 0;
    acc->evaluateTypeIIMigration = m2 != 0;
    return 0;
}

int RunDriver(int integrator, double epsilon, BodyData *bd, Acceleration *acc, TimeLine *tl)
{
    sol_context *ctx = Prepare(acc);
    if (!ctx) return 1;
    if (sol_set_state(ctx, bd->y0)) return fail("sol_set_state", ctx);
    sol_timeline t = to_sol(tl);
    double h = 0.0;
    bd->time = tl->time;
    const int r = sol_driver(ctx, integrator, epsilon, &t, &h);
    from_sol(t, tl);
    bd->h = h;
    if (FetchSideOutputs(acc, ctx)) return 1;
    // the reference leaves the drag switch on and the migration switches off
    acc->evaluateGasDrag = true;
    acc->evaluateTypeIMigration = false;
    acc->evaluateTypeIIMigration = false;
    // RungeKutta4 never writes BodyData::error and only RungeKuttaFehlberg78
    // writes yscale (Solaris/RungeKutta4.cpp:58-130, DormandPrince.cpp:471-483)
    const bool has_error = integrator != SOL_RUNGE_KUTTA4;
    const bool has_yscale = integrator == SOL_RUNGE_KUTTA_FEHLBERG78;
    if (r) {
        // no swap on failure; y and error hold the last attempt
        sol_get_array(ctx, SOL_ARRAY_Y, bd->y);
        if (has_error) sol_get_array(ctx, SOL_ARRAY_ERROR, bd->error);
        return fail("Driver", ctx);
    }
    bd->time = tl->time;
    if (sol_get_array(ctx, SOL_ARRAY_Y0, bd->y) || sol_get_array(ctx, SOL_ARRAY_ACCEL, bd->accel) ||
        (has_yscale && sol_get_array(ctx, SOL_ARRAY_YSCALE, bd->yscale)) ||
        (has_error && sol_get_array(ctx, SOL_ARRAY_ERROR, bd->error)))
        return fail("sol_get_array", ctx);
    std::swap(bd->y0, bd->y);
    return 0;
}

int RunStep(int integrator, BodyData *bd, Acceleration *acc)
{
    sol_context *ctx = Prepare(acc);
    if (!ctx) return 1;
    if (sol_set_state(ctx, bd->y0) || sol_set_accel(ctx, bd->accel)) return fail("upload", ctx);
    if (sol_step(ctx, integrator, bd->time, bd->h)) return fail("Step", ctx);
    if (sol_get_array(ctx, SOL_ARRAY_Y, bd->y) ||
        (integrator != SOL_RUNGE_KUTTA4 && sol_get_array(ctx, SOL_ARRAY_ERROR, bd->error)))
        return fail("sol_get_array", ctx);
    return FetchSideOutputs(acc, ctx);
}

}  // namespace solaris_dropin
