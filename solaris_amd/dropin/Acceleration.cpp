// Acceleration.cpp -- the reference's force-model class
// (Solaris/Acceleration.h:14-61) backed by the GPU library.  Compute() has
// the reference's meaning (Solaris/Acceleration.cpp:60-81): it writes the
// derivative [v, a] of every body for state y at time t, refreshes rm3 and
// the nearest-neighbour arrays, applies the gas drag and the cached
// migration accelerations under the evaluate* switches, and returns 0, or
// 1 on failure.
#include <cstring>

#include "Acceleration.h"
#include "BodyData.h"
#include "dropin_bridge.h"

Acceleration::Acceleration(IntegratorType iType, bool baryCentric, BodyData *bD, Nebula *n)
{
    _integratorType = iType;
    _baryCentric = baryCentric;
    evaluateGasDrag = true;
    evaluateTypeIMigration = true;
    evaluateTypeIIMigration = true;
    bodyData = bD;
    nebula = n;
    rm3 = 0;
    // the drag and migration caches live on the GPU
    accelGasDrag = 0;
    accelMigrationTypeI = 0;
    accelMigrationTypeII = 0;
    solaris_dropin::Register(this, baryCentric);
}

Acceleration::~Acceleration()
{
    solaris_dropin::Forget(this);
    delete[] rm3;
}

int Acceleration::Compute(double t, double *y, double *totalAccel)
{
    sol_context *ctx = solaris_dropin::Prepare(this);
    if (!ctx) return 1;
    if (sol_compute(ctx, t, y, totalAccel) != SOL_OK) return 1;
    return solaris_dropin::FetchSideOutputs(this, ctx);
}

int Acceleration::ComputeAstroCentric(double t, double *y, double *totalAccel)
{
    return Compute(t, y, totalAccel);
}

int Acceleration::ComputeBaryCentric(double t, double *y, double *totalAccel)
{
    return Compute(t, y, totalAccel);
}
