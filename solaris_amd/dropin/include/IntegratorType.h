// Integrator selector, same enumerators and values as the reference's
// Solaris/IntegratorType.h:12-17.
#ifndef INTEGRATORTYPE_H_
#define INTEGRATORTYPE_H_
enum IntegratorType { DORMAND_PRINCE, RUNGE_KUTTA4, RUNGE_KUTTA56, RUNGE_KUTTA_FEHLBERG78 };
#endif
