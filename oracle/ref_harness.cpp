/*
 * ref_harness.cpp -- C entry points around the compiled reference.
 *
 * TEST INFRASTRUCTURE ONLY.  oracle/Makefile compiles the reference's own
 * Solaris/*.cpp where they lie under /root/reference together with this
 * file into oracle/_ref/libsolref.so.  This file contains no reference
 * code: it constructs the reference's BodyData, Nebula, Acceleration and
 * integrator objects from plain arrays and calls their public methods, so
 * that the oracle restatement (cpu_ref.c) and the GPU library can be
 * compared with the real thing on identical inputs.
 *
 * Objects used (all public members):
 *   BodyData        Solaris/BodyData.h:8-51, Allocate Solaris/BodyData.cpp:66
 *   Nebula          Solaris/Nebula.h, GasComponent Solaris/GasComponent.h
 *   Acceleration    Solaris/Acceleration.h:14-61
 *   RungeKuttaFehlberg78 / RungeKutta4 / DormandPrince drivers
 *   Calculate::Integrals Solaris/Calculate.cpp:43-63
 */
#include <cstring>
#include <cmath>
#include <cstdlib>
#include <algorithm>

#include "Acceleration.h"
#include "BodyData.h"
#include "Calculate.h"
#include "DormandPrince.h"
#include "Nebula.h"
#include "RungeKutta4.h"
#include "RungeKuttaFehlberg78.h"
#include "TimeLine.h"

#include "../include/solaris_amd.h"

struct rf_system {
    BodyData bd;
    Nebula *neb;
    Acceleration *acc;
    RungeKuttaFehlberg78 rkf;
    RungeKutta4 rk4;
    DormandPrince dp;
    int barycentric;
};

extern "C" {

rf_system *rf_create(const sol_nbodies *nb, int32_t barycentric)
{
    rf_system *s = new rf_system();
    s->bd.nBodies.centralBody = nb->central_body;
    s->bd.nBodies.giantPlanet = nb->giant_planet;
    s->bd.nBodies.rockyPlanet = nb->rocky_planet;
    s->bd.nBodies.protoPlanet = nb->proto_planet;
    s->bd.nBodies.superPlanetsimal = nb->super_planetesimal;
    s->bd.nBodies.planetsimal = nb->planetesimal;
    s->bd.nBodies.testParticle = nb->test_particle;
    s->bd.nBodies.total = nb->total;
    if (s->bd.Allocate() != 0) { delete s; return 0; }
    int n = nb->total;
    for (int i = 0; i < n; i++) {
        s->bd.id[i] = i; s->bd.indexOfNN[i] = -1; s->bd.distanceOfNN[i] = 0.0;
    }
    std::memset(s->bd.y0, 0, sizeof(double)*6*n);
    std::memset(s->bd.y, 0, sizeof(double)*6*n);
    std::memset(s->bd.accel, 0, sizeof(double)*6*n);
    std::memset(s->bd.yscale, 0, sizeof(double)*6*n);
    std::memset(s->bd.error, 0, sizeof(double)*6*n);
    s->neb = 0;
    s->barycentric = barycentric;
    s->acc = new Acceleration(RUNGE_KUTTA_FEHLBERG78, barycentric != 0, &s->bd, 0);
    return s;
}

void rf_destroy(rf_system *s)
{
    if (!s) return;
    delete s->acc;
    delete s->neb;
    delete s;
}

void rf_set_bodies(rf_system *s, const sol_body_params *p)
{
    int n = s->bd.nBodies.total;
    for (int i = 0; i < n; i++) {
        s->bd.mass[i] = p->mass ? p->mass[i] : 0.0;
        s->bd.radius[i] = p->radius ? p->radius[i] : 0.0;
        s->bd.density[i] = p->density ? p->density[i] : 0.0;
        s->bd.cD[i] = p->cd ? p->cd[i] : 0.0;
        s->bd.gammaStokes[i] = p->gamma_stokes ? p->gamma_stokes[i] : 0.0;
        s->bd.gammaEpstein[i] = p->gamma_epstein ? p->gamma_epstein[i] : 0.0;
        s->bd.migStopAt[i] = p->mig_stop_at ? p->mig_stop_at[i] : 0.0;
        s->bd.type[i] = p->type ? p->type[i] : 0;
        s->bd.migType[i] = p->mig_type ? p->mig_type[i] : 0;
    }
}

/* Recreates the Acceleration object so its lazily allocated caches start
 * empty, as for a fresh simulation. */
void rf_set_nebula(rf_system *s, const sol_gas_component *g)
{
    delete s->acc;
    delete s->neb;
    s->neb = 0;
    if (g) {
        s->neb = new Nebula();
        GasComponent &c = s->neb->gasComponent;
        c.alpha = g->alpha;
        c.meanMolecularWeight = g->mean_molecular_weight;
        c.particleDiameter = g->particle_diameter;
        c.type = (GasDecreaseType)g->type;
        c.timeScale = g->time_scale;
        c.t0 = g->t0;
        c.t1 = g->t1;
        c.innerEdge = g->inner_edge;
        c.a = g->a;
        c.eta = PowerLaw(g->eta.c, g->eta.index);
        c.tau = PowerLaw(g->tau.c, g->tau.index);
        c.scaleHeight = PowerLaw(g->scale_height.c, g->scale_height.index);
        c.density = PowerLaw(g->density.c, g->density.index);
        c.meanFreePath = PowerLaw(g->mean_free_path.c, g->mean_free_path.index);
    }
    s->acc = new Acceleration(RUNGE_KUTTA_FEHLBERG78, s->barycentric != 0, &s->bd, s->neb);
}

/* The reference constructor's GasComponent, for cross-checking defaults. */
void rf_default_gas(sol_gas_component *g)
{
    GasComponent c;
    std::memset(g, 0, sizeof(*g));
    g->alpha = c.alpha; g->mean_molecular_weight = c.meanMolecularWeight;
    g->particle_diameter = c.particleDiameter; g->type = c.type;
    g->time_scale = c.timeScale; g->t0 = c.t0; g->t1 = c.t1;
    g->inner_edge = c.innerEdge; g->a = c.a;
    g->eta.c = c.eta.c; g->eta.index = c.eta.index;
    g->tau.c = c.tau.c; g->tau.index = c.tau.index;
    g->scale_height.c = c.scaleHeight.c; g->scale_height.index = c.scaleHeight.index;
    g->density.c = c.density.c; g->density.index = c.density.index;
    g->mean_free_path.c = c.meanFreePath.c; g->mean_free_path.index = c.meanFreePath.index;
}

void rf_set_evaluate(rf_system *s, int32_t d, int32_t m1, int32_t m2)
{
    s->acc->evaluateGasDrag = d != 0;
    s->acc->evaluateTypeIMigration = m1 != 0;
    s->acc->evaluateTypeIIMigration = m2 != 0;
}

void rf_set_mig_type(rf_system *s, const int32_t *m)
{
    for (int i = 0; i < s->bd.nBodies.total; i++) s->bd.migType[i] = m[i];
}

int rf_compute(rf_system *s, double t, const double *y, double *dydt)
{
    int n6 = 6*s->bd.nBodies.total;
    double *yc = new double[n6];
    std::memcpy(yc, y, sizeof(double)*n6);
    int r = s->acc->Compute(t, yc, dydt);
    delete[] yc;
    return r;
}

void rf_get_side_outputs(rf_system *s, double *rm3, int32_t *nn, double *dist, int32_t *mig)
{
    int n = s->bd.nBodies.total;
    for (int i = 0; i < n; i++) {
        if (rm3) rm3[i] = s->acc->rm3 ? s->acc->rm3[i] : 0.0;
        if (nn) nn[i] = s->bd.indexOfNN[i];
        if (dist) dist[i] = s->bd.distanceOfNN[i];
        if (mig) mig[i] = s->bd.migType[i];
    }
}

void rf_set_state(rf_system *s, const double *y0)
{
    std::memcpy(s->bd.y0, y0, sizeof(double)*6*s->bd.nBodies.total);
}

void rf_set_accel(rf_system *s, const double *a)
{
    std::memcpy(s->bd.accel, a, sizeof(double)*6*s->bd.nBodies.total);
}

void rf_get_array(rf_system *s, int32_t which, double *out)
{
    const double *src = s->bd.y0;
    switch (which) {
    case SOL_ARRAY_Y0: src = s->bd.y0; break;
    case SOL_ARRAY_Y: src = s->bd.y; break;
    case SOL_ARRAY_ACCEL: src = s->bd.accel; break;
    case SOL_ARRAY_YSCALE: src = s->bd.yscale; break;
    case SOL_ARRAY_ERROR: src = s->bd.error; break;
    }
    std::memcpy(out, src, sizeof(double)*6*s->bd.nBodies.total);
}

static void to_tl(const sol_timeline *a, TimeLine *b)
{
    b->start = a->start; b->length = a->length; b->output = a->output;
    b->save = a->save; b->time = a->time; b->millenium = a->millenium;
    b->hNext = a->hNext; b->hDid = a->hDid; b->elapsedTime = a->elapsedTime;
    b->lastSave = a->lastSave; b->lastNSteps = a->lastNSteps;
}

static void from_tl(const TimeLine *b, sol_timeline *a)
{
    a->start = b->start; a->length = b->length; a->output = b->output;
    a->save = b->save; a->time = b->time; a->millenium = b->millenium;
    a->hNext = b->hNext; a->hDid = b->hDid; a->elapsedTime = b->elapsedTime;
    a->lastSave = b->lastSave; a->lastNSteps = b->lastNSteps;
}

int rf_driver(rf_system *s, int32_t integrator, double eps, sol_timeline *tl, double *body_h)
{
    TimeLine t;
    to_tl(tl, &t);
    int r = 1;
    switch (integrator) {
    case SOL_RUNGE_KUTTA_FEHLBERG78:
        s->rkf.epsilon = eps; r = s->rkf.Driver(&s->bd, s->acc, &t); break;
    case SOL_RUNGE_KUTTA4:
        s->rk4.epsilon = eps; r = s->rk4.Driver(&s->bd, s->acc, &t); break;
    case SOL_DORMAND_PRINCE:
        s->dp.epsilon = eps; r = s->dp.Driver(&s->bd, s->acc, &t); break;
    default: return SOL_ERR_UNSUPPORTED;
    }
    from_tl(&t, tl);
    if (body_h) *body_h = s->bd.h;
    return r;
}

int rf_step(rf_system *s, int32_t integrator, double time, double h)
{
    s->bd.time = time;
    s->bd.h = h;
    switch (integrator) {
    case SOL_RUNGE_KUTTA_FEHLBERG78: return s->rkf.Step(&s->bd, s->acc);
    case SOL_RUNGE_KUTTA4: return s->rk4.Step(&s->bd, s->acc);
    case SOL_DORMAND_PRINCE: return s->dp.Step2(&s->bd, s->acc);
    default: return SOL_ERR_UNSUPPORTED;
    }
}

double rf_error_max(rf_system *s, int32_t integrator, double eps)
{
    int nv = 6*s->bd.nBodies.total;
    if (integrator == SOL_DORMAND_PRINCE) return s->dp.GetErrorMax(nv, s->bd.error);
    s->rkf.epsilon = eps;
    return s->rkf.GetErrorMax(nv, s->bd.error, s->bd.yscale);
}

void rf_integrals(rf_system *s, double *out16)
{
    Calculate::Integrals(&s->bd);
    for (int i = 0; i < 16; i++) out16[i] = s->bd.integrals[i];
}

/* Reference helpers for single-function checks. */
void rf_gas_velocity(const sol_gas_component *g, double mu, const double *r3, double *v3)
{
    GasComponent c;
    c.eta = PowerLaw(g->eta.c, g->eta.index);
    Vector r(r3[0], r3[1], r3[2]);
    Vector v = c.gas_velocity(mu, &r);
    v3[0] = v.x; v3[1] = v.y; v3[2] = v.z;
}

double rf_gas_density_at(const sol_gas_component *g, const double *r3)
{
    GasComponent c;
    c.innerEdge = g->inner_edge;
    c.scaleHeight = PowerLaw(g->scale_height.c, g->scale_height.index);
    c.density = PowerLaw(g->density.c, g->density.index);
    Vector r(r3[0], r3[1], r3[2]);
    return c.gas_density_at(&r);
}

}  /* extern "C" */
