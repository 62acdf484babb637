// dropin_bridge.h -- connects the reference-shaped classes (Acceleration,
// the integrators, BodyData, Nebula, TimeLine) to the C ABI of
// include/solaris_amd.h.  It reads and writes only the public members the
// reference declares, so the drop-in files compile against the reference's
// own headers as well as against solaris_amd/dropin/include.
#ifndef SOLARIS_DROPIN_BRIDGE_H_
#define SOLARIS_DROPIN_BRIDGE_H_

#include "solaris_amd.h"

class Acceleration;
class BodyData;
class TimeLine;

namespace solaris_dropin {

// Called by the Acceleration constructor: remembers the frame, which the
// reference keeps in a private member.
void Register(Acceleration *acc, bool barycentric);

// The GPU context behind an Acceleration object, created on first use and
// rebuilt when the body count changes.  Per-body parameters, migType and
// the gas are uploaded before every force evaluation or driver call, since
// the caller may edit BodyData between calls (Simulator::CheckEvent merges
// and removes bodies).  Returns nullptr on failure (message in LastError()).
sol_context *Prepare(Acceleration *acc);

// Copies the side outputs of the last evaluation back into the caller's
// arrays: Acceleration::rm3 (allocated on first use like the reference,
// Solaris/Acceleration.cpp:65-69), BodyData::indexOfNN / distanceOfNN and
// the migType the migration kernels may have switched off.
int FetchSideOutputs(Acceleration *acc, sol_context *ctx);

// Runs one Driver call of `integrator` and reproduces its effects on the
// caller's objects: BodyData::time and h, the TimeLine fields, the new
// state in BodyData::y followed by the y0/y pointer swap, accel, yscale,
// error, the side outputs and the evaluate* switches the reference leaves
// behind (RungeKuttaFehlberg78.cpp:74-133).
int RunDriver(int integrator, double epsilon, BodyData *bd, Acceleration *acc, TimeLine *tl);

// One Step from BodyData::y0, accel, time and h into y and error.
int RunStep(int integrator, BodyData *bd, Acceleration *acc);

// Message of the last failure of this library in the calling thread.
const char *LastError();
void Forget(Acceleration *acc);

}  // namespace solaris_dropin

#endif
