// RungeKutta4.cpp -- the reference's classical RK4 driver class
// (Solaris/RungeKutta4.h) on the GPU; fixed step: Driver() advances by
// TimeLine::hNext and keeps it (Solaris/RungeKutta4.cpp:20-56).
#include <cmath>

#include "RungeKutta4.h"
#include "dropin_bridge.h"

RungeKutta4::RungeKutta4()
{
    name = "Classical Runge-Kutta 4 (MI355X)";
    reference = "";
    accuracy = -10.0;
    epsilon = std::pow(10.0, accuracy);
}

int RungeKutta4::Driver(BodyData *bodyData, Acceleration *acceleration, TimeLine *timeLine)
{
    return solaris_dropin::RunDriver(SOL_RUNGE_KUTTA4, epsilon, bodyData, acceleration, timeLine);
}

int RungeKutta4::Step(BodyData *bodyData, Acceleration *acceleration)
{
    return solaris_dropin::RunStep(SOL_RUNGE_KUTTA4, bodyData, acceleration);
}
