// support.cpp -- constructors and helpers of the plain data classes that
// the drop-in classes read (BodyData, NBodies, PowerLaw, GasComponent,
// Nebula, TimeLine), for builds that do not link the reference's own
// implementations of them.  Defaults are the reference's
// (Solaris/BodyData.cpp, Solaris/NBodies.cpp, Solaris/GasComponent.cpp:9-34,
// Solaris/Nebula.cpp, Solaris/TimeLine.cpp).
#include <cmath>
#include <cstring>

#include "BodyData.h"
#include "GasComponent.h"
#include "NBodies.h"
#include "Nebula.h"
#include "PowerLaw.h"
#include "TimeLine.h"

NBodies::NBodies()
    : centralBody(0), giantPlanet(0), rockyPlanet(0), protoPlanet(0), superPlanetsimal(0), planetsimal(0),
      testParticle(0), total(0), removed(0), threshold(10)
{
}

int NBodies::NOfMassive() { return centralBody + giantPlanet + rockyPlanet + protoPlanet; }
int NBodies::NOfPlAndSpl() { return superPlanetsimal + planetsimal; }
int NBodies::NOfVar() { return 6 * total; }

BodyData::BodyData()
    : time(0.0), h(0.0), id(0), type(0), migType(0), migStopAt(0), indexOfNN(0), distanceOfNN(0), mass(0),
      radius(0), density(0), cD(0), gammaStokes(0), gammaEpstein(0), y0(0), y(0), yBetterEst(0), yscale(0),
      accel(0), error(0)
{
    std::memset(bc, 0, sizeof bc);
    std::memset(kineticEnergy, 0, sizeof kineticEnergy);
    std::memset(potentialEnergy, 0, sizeof potentialEnergy);
    std::memset(totalEnergy, 0, sizeof totalEnergy);
    std::memset(integrals, 0, sizeof integrals);
}

BodyData::~BodyData() { Free(); }

void BodyData::Free()
{
    delete[] id; delete[] type; delete[] migType; delete[] migStopAt; delete[] indexOfNN; delete[] distanceOfNN;
    delete[] mass; delete[] radius; delete[] density; delete[] cD; delete[] gammaStokes; delete[] gammaEpstein;
    delete[] y0; delete[] y; delete[] yBetterEst; delete[] yscale; delete[] accel; delete[] error;
    id = type = migType = indexOfNN = 0;
    migStopAt = distanceOfNN = mass = radius = density = cD = gammaStokes = gammaEpstein = 0;
    y0 = y = yBetterEst = yscale = accel = error = 0;
}

// Returns 1 when there is nothing to allocate (Solaris/BodyData.cpp:68-72).
int BodyData::Allocate()
{
    const int n = nBodies.total;
    if (n <= 0) return 1;
    id = new int[n](); type = new int[n](); migType = new int[n]();
    migStopAt = new double[n](); indexOfNN = new int[n](); distanceOfNN = new double[n]();
    mass = new double[n](); radius = new double[n](); density = new double[n](); cD = new double[n]();
    gammaStokes = new double[n](); gammaEpstein = new double[n]();
    const int nv = 6 * n;
    y0 = new double[nv](); y = new double[nv](); yBetterEst = new double[nv](); yscale = new double[nv]();
    accel = new double[nv](); error = new double[nv]();
    return 0;
}

PowerLaw::PowerLaw() : c(0.0), index(0.0) {}
PowerLaw::PowerLaw(double co, double idx) : c(co), index(idx) {}
double PowerLaw::Evaluate(double x) { return c * std::pow(x, index); }

GasComponent::GasComponent()
{
    const double pi = 3.14159265358979323846;
    const double gram_to_solar = 1.0 / (1.0e3 * 1.98911e30);
    const double meter_to_au = 1.0 / 1.495978707e11;
    const double gcm3 = gram_to_solar / ((1.0e-2 * meter_to_au) * (1.0e-2 * meter_to_au) * (1.0e-2 * meter_to_au));
    const double proton_mass = 1.672621777e-27 * (1.0 / 1.98911e30);
    innerEdge = 10.0 * (1.0 / 215.094);
    alpha = 2.0e-3;
    meanMolecularWeight = 2.3;
    particleDiameter = 3.0e-10;
    timeScale = 0.0;
    t0 = 0.0;
    t1 = 0.0;
    type = CONSTANT;
    eta = PowerLaw(0.0019, 0.5);
    tau = PowerLaw(2.0 / 3.0, 2.0);
    scaleHeight = PowerLaw(0.02, 1.25);
    density = PowerLaw(1.4e-9 * gcm3, -2.75);
    a = density.Evaluate(innerEdge) / ((innerEdge * innerEdge) * (innerEdge * innerEdge));
    const double d = particleDiameter * meter_to_au;
    meanFreePath = PowerLaw(meanMolecularWeight * proton_mass / (std::sqrt(2.0) * pi * (d * d) * density.c),
                            -density.index);
}

double GasComponent::ReductionFactor(const double t)
{
    switch (type) {
    case CONSTANT: return 1.0;
    case LINEAR:
        if (t <= t0) return 1.0;
        if (t > t0 && t <= t1) return 1.0 - (t - t0) / (t1 - t0);
        return 0.0;
    case EXPONENTIAL: return std::exp(-t / timeScale);
    }
    return 1.0;
}

Nebula::Nebula() : name("MMSN"), massFactor(1.0), gasToDustRatio(100.0), snowLine(2.7) {}

TimeLine::TimeLine()
    : start(0.0), length(0.0), output(0.0), save(0.0), time(0.0), millenium(0), hNext(0.0), hDid(0.0),
      elapsedTime(0.0), lastSave(0.0), lastNSteps(0.0), startTimeDefined(false)
{
}
