// Gas decay law selector (Solaris/GasDecreaseType.h).
#ifndef GASDECREASETYPE_H_
#define GASDECREASETYPE_H_
enum GasDecreaseType { CONSTANT, LINEAR, EXPONENTIAL };
#endif
