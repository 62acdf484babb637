// The arrays the force model and the integrators share, with the
// reference's member names and meaning (Solaris/BodyData.h:8-51):
// AoS states of 6 doubles per body, per-body parameter arrays.
#ifndef BODYDATA_H_
#define BODYDATA_H_
#include "NBodies.h"
#include "Phase.h"
#include "Vector.h"
class BodyData {
public:
    BodyData();
    ~BodyData();
    int Allocate();
    void Free();
    NBodies nBodies;
    double time;
    double h;
    int *id;
    int *type;
    int *migType;
    double *migStopAt;
    int *indexOfNN;
    double *distanceOfNN;
    double *mass;
    double *radius;
    double *density;
    double *cD;
    double *gammaStokes;
    double *gammaEpstein;
    double *y0;
    double *y;
    double *yBetterEst;
    double *yscale;
    double *accel;
    double *error;
    double bc[6];
    Phase phaseOfBC;
    Vector angularMomentum[2];
    double kineticEnergy[2];
    double potentialEnergy[2];
    double totalEnergy[2];
    double integrals[16];
};
#endif
