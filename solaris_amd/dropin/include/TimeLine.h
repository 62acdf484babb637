// Integration time bookkeeping (Solaris/TimeLine.h:6-44).
#ifndef TIMELINE_H_
#define TIMELINE_H_
class TimeLine {
public:
    TimeLine();
    bool Forward() { return length >= 0 ? true : false; }
    double start;
    double length;
    double output;
    double save;
    double time;
    int millenium;
    double hNext;
    double hDid;
    double elapsedTime;
    double lastSave;
    double lastNSteps;
    bool startTimeDefined;
};
#endif
