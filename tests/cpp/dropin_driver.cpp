// dropin_driver.cpp -- drives the drop-in classes the way the reference's
// Simulator::Integrate does (Solaris/Simulator.cpp:121-178): a BodyData
// filled from a file, an Acceleration with or without a Nebula, and
// repeated Driver calls of the chosen integrator.
//
//   dropin_driver <input.bin> <output.bin> <steps> <integrator>
//
// input.bin (little-endian): int32 counts[8] (central, giant, rocky, proto,
// super-planetesimal, planetesimal, test particle, total); f64 arrays mass,
// radius, density, cd, gammaStokes, gammaEpstein, migStopAt [total]; int32
// type, migType [total]; f64 y0[6*total]; int32 has_gas; f64 gas[21]
// (alpha, mmw, particle diameter, type, timeScale, t0, t1, innerEdge, a,
// eta, tau, scaleHeight, density, meanFreePath as c,index pairs);
// f64 hNext, epsilon.
// output.bin: int32 status, int32 steps done; f64 hDid[steps], f64 time;
// f64 y0[6*total]; f64 accel[6*total]; f64 rm3[total]; int32 indexOfNN
// [total]; int32 migType[total]; f64 a Compute() of the final state
// [6*total].
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "Acceleration.h"
#include "BodyData.h"
#include "DormandPrince.h"
#include "Nebula.h"
#include "RungeKutta4.h"
#include "RungeKuttaFehlberg78.h"
#include "TimeLine.h"
#include "dropin_bridge.h"

template <typename T>
static bool rd(FILE *f, T *p, size_t n) { return fread(p, sizeof(T), n, f) == n; }

int main(int argc, char **argv)
{
    if (argc != 5) {
        fprintf(stderr, "usage: %s input.bin output.bin steps integrator\n", argv[0]);
        return 2;
    }
    FILE *f = fopen(argv[1], "rb");
    if (!f) return 2;
    int c[8];
    if (!rd(f, c, 8)) return 2;
    BodyData bd;
    bd.nBodies.centralBody = c[0]; bd.nBodies.giantPlanet = c[1]; bd.nBodies.rockyPlanet = c[2];
    bd.nBodies.protoPlanet = c[3]; bd.nBodies.superPlanetsimal = c[4]; bd.nBodies.planetsimal = c[5];
    bd.nBodies.testParticle = c[6]; bd.nBodies.total = c[7];
    if (bd.Allocate() != 0) return 2;
    const int n = c[7];
    bool ok = rd(f, bd.mass, n) && rd(f, bd.radius, n) && rd(f, bd.density, n) && rd(f, bd.cD, n) &&
              rd(f, bd.gammaStokes, n) && rd(f, bd.gammaEpstein, n) && rd(f, bd.migStopAt, n) &&
              rd(f, bd.type, n) && rd(f, bd.migType, n) && rd(f, bd.y0, 6 * n);
    int has_gas = 0;
    double g[21], hnext, eps;
    ok = ok && rd(f, &has_gas, 1) && rd(f, g, 21) && rd(f, &hnext, 1) && rd(f, &eps, 1);
    fclose(f);
    if (!ok) { fprintf(stderr, "short input\n"); return 2; }
    for (int i = 0; i < n; ++i) bd.id[i] = i;

    Nebula *neb = 0;
    if (has_gas) {
        neb = new Nebula();
        GasComponent &gc = neb->gasComponent;
        gc.alpha = g[0]; gc.meanMolecularWeight = g[1]; gc.particleDiameter = g[2];
        gc.type = (GasDecreaseType)(int)g[3]; gc.timeScale = g[4]; gc.t0 = g[5]; gc.t1 = g[6];
        gc.innerEdge = g[7]; gc.a = g[8];
        gc.eta = PowerLaw(g[9], g[10]); gc.tau = PowerLaw(g[11], g[12]);
        gc.scaleHeight = PowerLaw(g[13], g[14]); gc.density = PowerLaw(g[15], g[16]);
        gc.meanFreePath = PowerLaw(g[17], g[18]);
    }
    const int integ = atoi(argv[4]);
    Acceleration acc(integ == 3 ? RUNGE_KUTTA_FEHLBERG78 : integ == 1 ? RUNGE_KUTTA4 : DORMAND_PRINCE, false, &bd, neb);
    RungeKuttaFehlberg78 rkf;
    RungeKutta4 rk4;
    DormandPrince dp;
    rkf.epsilon = rk4.epsilon = dp.epsilon = eps;
    TimeLine tl;
    tl.hNext = hnext;
    tl.length = 1e30;
    const int steps = atoi(argv[3]);
    std::vector<double> hdid;
    int status = 0;
    for (int k = 0; k < steps; ++k) {
        int r = integ == 3 ? rkf.Driver(&bd, &acc, &tl) : integ == 1 ? rk4.Driver(&bd, &acc, &tl)
                                                        : dp.Driver(&bd, &acc, &tl);
        if (r) { status = r; fprintf(stderr, "Driver failed: %s\n", solaris_dropin::LastError()); break; }
        hdid.push_back(tl.hDid);
    }
    std::vector<double> dydt(6 * n);
    std::vector<double> rm3(n);
    if (!status) {
        for (int i = 0; i < n; ++i) rm3[i] = acc.rm3 ? acc.rm3[i] : 0.0;
        acc.evaluateTypeIMigration = acc.evaluateTypeIIMigration = false;
        if (acc.Compute(tl.time, bd.y0, dydt.data())) status = 1;
    }
    FILE *o = fopen(argv[2], "wb");
    if (!o) return 2;
    int done = (int)hdid.size();
    fwrite(&status, 4, 1, o);
    fwrite(&done, 4, 1, o);
    fwrite(hdid.data(), 8, done, o);
    fwrite(&tl.time, 8, 1, o);
    fwrite(bd.y0, 8, 6 * n, o);
    fwrite(bd.accel, 8, 6 * n, o);
    fwrite(rm3.data(), 8, n, o);
    fwrite(bd.indexOfNN, 4, n, o);
    fwrite(bd.migType, 4, n, o);
    fwrite(dydt.data(), 8, 6 * n, o);
    fclose(o);
    delete neb;
    return status;
}
