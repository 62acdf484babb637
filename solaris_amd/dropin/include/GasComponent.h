// Gas disk parameters read by the force model (Solaris/GasComponent.h):
// same public fields; the constructor sets the reference's defaults.
#ifndef GASCOMPONENT_H_
#define GASCOMPONENT_H_
#include "GasDecreaseType.h"
#include "PowerLaw.h"
class GasComponent {
public:
    GasComponent();
    double alpha;
    double meanMolecularWeight;
    double particleDiameter;
    GasDecreaseType type;
    double timeScale;
    double t0;
    double t1;
    double innerEdge;
    double a;
    PowerLaw eta;
    PowerLaw tau;
    PowerLaw scaleHeight;
    PowerLaw density;
    PowerLaw meanFreePath;
    PowerLaw temperature;
    double ReductionFactor(const double t);
};
#endif
