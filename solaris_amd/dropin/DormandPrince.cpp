// DormandPrince.cpp -- the reference's Dormand-Prince RKN 7(6) driver
// class (Solaris/DormandPrince.h) on the GPU.  Driver() retries at most
// ten times with 0.9 h (eps/err)^(1/7) and fails when that is exceeded
// (Solaris/DormandPrince.cpp:126-170).  Step() and Step2() both run the
// scheme the reference's Driver uses (Step2, :246-491).
#include <cmath>

#include "DormandPrince.h"
#include "dropin_bridge.h"

DormandPrince::DormandPrince()
{
    name = "Dormand-Prince RKN 7(6) (MI355X)";
    reference = "J. R. Dormand, P. J. Prince, Celestial Mechanics 18 (1978) 223-232";
    accuracy = -10.0;
    epsilon = std::pow(10.0, accuracy);
}

int DormandPrince::Driver(BodyData *bodyData, Acceleration *acceleration, TimeLine *timeLine)
{
    return solaris_dropin::RunDriver(SOL_DORMAND_PRINCE, epsilon, bodyData, acceleration, timeLine);
}

int DormandPrince::Step(BodyData *bodyData, Acceleration *acceleration)
{
    return Step2(bodyData, acceleration);
}

int DormandPrince::Step2(BodyData *bodyData, Acceleration *acceleration)
{
    return solaris_dropin::RunStep(SOL_DORMAND_PRINCE, bodyData, acceleration);
}

// Largest |yerr| over the n entries (absolute, no scaling).
double DormandPrince::GetErrorMax(int n, const double *yerr)
{
    double m = 0.0;
    for (int i = 0; i < n; i++) {
        const double q = std::fabs(yerr[i]);
        if (q > m) m = q;
    }
    return m;
}
