// RungeKuttaFehlberg78.cpp -- the reference's RKF7(8) driver class
// (Solaris/RungeKuttaFehlberg78.h) on the GPU.  Driver() keeps the
// reference's contract (RungeKuttaFehlberg78.cpp:66-140): it reads
// TimeLine::time and hNext, retries with a shrunk step until the scaled
// error maximum is below 1, then advances time, sets hDid and hNext, leaves
// the new state in BodyData::y0 and the previous one in BodyData::y.
// Returns 1 on step-size underflow or a device failure.
#include <cmath>

#include "RungeKuttaFehlberg78.h"
#include "dropin_bridge.h"

RungeKuttaFehlberg78::RungeKuttaFehlberg78()
{
    name = "Runge-Kutta-Fehlberg 7(8) with step-size control (MI355X)";
    reference = "E. Fehlberg, NASA Technical Report R-381, 1972";
    accuracy = -10.0;
    epsilon = std::pow(10.0, accuracy);
}

int RungeKuttaFehlberg78::Driver(BodyData *bodyData, Acceleration *acceleration, TimeLine *timeLine)
{
    return solaris_dropin::RunDriver(SOL_RUNGE_KUTTA_FEHLBERG78, epsilon, bodyData, acceleration, timeLine);
}

int RungeKuttaFehlberg78::Step(BodyData *bodyData, Acceleration *acceleration)
{
    return solaris_dropin::RunStep(SOL_RUNGE_KUTTA_FEHLBERG78, bodyData, acceleration);
}

// Largest |yerr / yscale| over the n entries, divided by epsilon.
double RungeKuttaFehlberg78::GetErrorMax(const int n, const double *yerr, const double *yscale)
{
    double m = 0.0;
    for (int i = 0; i < n; i++) {
        const double q = std::fabs(yerr[i] / yscale[i]);
        if (q > m) m = q;
    }
    return m / epsilon;
}
