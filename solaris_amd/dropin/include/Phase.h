// Position/velocity pair with the reference's public fields (Solaris/Phase.h).
#ifndef PHASE_H_
#define PHASE_H_
#include "Vector.h"
class Phase {
public:
    Phase() : bodyId(0) {}
    int bodyId;
    Vector position;
    Vector velocity;
};
#endif
