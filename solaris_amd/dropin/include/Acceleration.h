// The force model's public interface (Solaris/Acceleration.h:14-61):
// constructor, Compute, the public rm3 array and the evaluate* switches.
// Compute runs on the GPU through the C ABI of include/solaris_amd.h.
#ifndef ACCELERATION_H_
#define ACCELERATION_H_
#include "IntegratorType.h"
class BodyData;
class Nebula;
class Acceleration {
public:
    Acceleration(IntegratorType iType, bool baryCentric, BodyData *bD, Nebula *n);
    ~Acceleration();
    int Compute(double t, double *y, double *totalAccel);
    int ComputeAstroCentric(double t, double *y, double *totalAccel);
    int ComputeBaryCentric(double t, double *y, double *totalAccel);
    BodyData *bodyData;
    Nebula *nebula;
    double *accelGasDrag;
    double *accelMigrationTypeI;
    double *accelMigrationTypeII;
    double *rm3;
    bool evaluateGasDrag;
    bool evaluateTypeIMigration;
    bool evaluateTypeIIMigration;
private:
    IntegratorType _integratorType;
    bool _baryCentric;
};
#endif
