// The nebula as the force model sees it: its gas component
// (Solaris/Nebula.h; the solids component plays no part in the forces).
#ifndef NEBULA_H_
#define NEBULA_H_
#include <string>
#include "GasComponent.h"
class Nebula {
public:
    Nebula();
    std::string name;
    std::string description;
    double massFactor;
    double gasToDustRatio;
    double snowLine;
    GasComponent gasComponent;
};
#endif
