// Body and migration type enumerations with the reference's values
// (Solaris/Body.h:14-39).  Only the enums the force model reads.
#ifndef BODY_H_
#define BODY_H_
enum BodyType { UndefinedBodyType, CentralBody, GiantPlanet, RockyPlanet, ProtoPlanet,
                SuperPlanetesimal, Planetesimal, TestParticle, NOfBodyType };
enum MigrationType { No, TypeI, TypeII };
#endif
