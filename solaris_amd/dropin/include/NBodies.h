// Body counts by type (Solaris/NBodies.h); the counts define the
// interaction ranges, so bodies are type-sorted.
#ifndef NBODIES_H_
#define NBODIES_H_
class NBodies {
public:
    NBodies();
    int NOfMassive();
    int NOfPlAndSpl();
    int NOfVar();
    int centralBody, giantPlanet, rockyPlanet, protoPlanet, superPlanetsimal, planetsimal, testParticle, total;
    int removed, threshold;
};
#endif
