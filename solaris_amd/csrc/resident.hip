// resident.hip -- whole Driver loops of small systems inside one workgroup.
//
// One workgroup integrates one system (n <= kResidentMaxBodies bodies):
// the RKF78 / RK4 / Dormand-Prince Driver (Solaris/RungeKuttaFehlberg78.cpp:
// 66-140, Solaris/RungeKutta4.cpp:20-56, Solaris/DormandPrince.cpp:126-170),
// including the retry loop and the step-size arithmetic, and - when asked -
// the caller loop's step clipping and small-number flush
// (Simulator::DecisionMaking, Solaris/Simulator.cpp:159-245), all on the
// device.  A launch runs at most max_steps accepted steps per system and
// returns; the host relaunches until every system is done.  Systems are
// independent: a failing system (step-size underflow, too many retries)
// stops with its own status and the others go on.
//
// Within a system each thread owns bodies i = tid, tid + blockDim, ...
// The positions of the stage being evaluated are staged in LDS as
// (x, y, z, k^2 m) records; every sink reads them as a broadcast stream.
// Stage derivatives live in global memory ([stage][6][TOT] over all
// systems' bodies, TOT = sum of per-system padded counts) and are read
// back only by the thread that wrote them.  Gravity only (no gas): systems
// with a nebula use the multi-kernel path.
#include <float.h>

#include "sol_internal.h"
#include "resident.h"

namespace sol {
namespace {

constexpr int kRkfD = 0, kRk4A = 156, kRk4B = 168, kRk4C = 172, kDpA = 176, kDpB = 248, kDpBdh = 257, kDpC = 266;

__device__ __forceinline__ double wsum(double v)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
}

__device__ __forceinline__ double wmax(double v)
{
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) v = fmax(v, __shfl_xor(v, o, kWave));
    return v;
}

// Block-wide sum of v[0..5] (fixed order); result in out[0..5] for every
// thread.  A wave whose v are all +0.0 by construction (wave_zero, a
// wave-uniform flag) skips its shuffle reduction: its partial is +0.0
// either way.
__device__ void block_sum6_all(double v[6], double *red /* [16][6] */, double *out /* [6] */,
                               bool wave_zero = false)
{
    const int wid = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
    const int nw = blockDim.x / kWave;
    if (!wave_zero)
        for (int k = 0; k < 6; ++k) v[k] = wsum(v[k]);
    if (lane == 0)
        for (int k = 0; k < 6; ++k) red[wid * 6 + k] = v[k];
    __syncthreads();
    if (threadIdx.x < 6) {
        double s = red[threadIdx.x];
        for (int q = 1; q < nw; ++q) s += red[q * 6 + threadIdx.x];
        out[threadIdx.x] = s;
    }
    __syncthreads();
}

__device__ double block_max_all(double v, double *red, double *out)
{
    const int wid = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
    const int nw = blockDim.x / kWave;
    v = wmax(v);
    if (lane == 0) red[wid] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double m = red[0];
        for (int q = 1; q < nw; ++q) m = fmax(m, red[q]);
        *out = m;
    }
    __syncthreads();
    return *out;
}

struct Ctx {
    const ResSystem *S;
    const ResArrays *A;
    int n, base, tot;
    double4 *sJ;        // LDS: positions + k^2 m of the stage
    double *red;        // LDS scratch
    double *ind;        // LDS: indirect sums [2][3]
    const double *coef;
};

__device__ __forceinline__ double &at(double *p, const Ctx &c, int comp, int i) { return p[(size_t)comp * c.tot + c.base + i]; }
__device__ __forceinline__ double cat(const double *p, const Ctx &c, int comp, int i) { return p[(size_t)comp * c.tot + c.base + i]; }

// Stage derivative component comp of stage l (0 = f0 in ACC0).
__device__ __forceinline__ double F(const Ctx &c, int l, int comp, int i)
{
    if (l == 0) return cat(c.A->acc0, c, comp, i);
    return c.A->fst[(size_t)(l - 1) * 6 * c.tot + (size_t)comp * c.tot + c.base + i];
}

__device__ __forceinline__ double &Fw(const Ctx &c, int l, int comp, int i)
{
    if (l == 0) return c.A->acc0[(size_t)comp * c.tot + c.base + i];
    return c.A->fst[(size_t)(l - 1) * 6 * c.tot + (size_t)comp * c.tot + c.base + i];
}

// Loads the positions of state `y` (6 x n, own bodies) into LDS and forms
// the indirect-term sums of both source ranges (AC).  y is given per body
// through a functor get(i, comp).
template <typename Get>
__device__ void stage_positions(const Ctx &c, Get get)
{
    const ResSystem &S = *c.S;
    // only the bodies below the larger source end enter the indirect sums
    const int end_max = S.end_a > S.end_b ? S.end_a : S.end_b;
    double v[6] = {0, 0, 0, 0, 0, 0};
    for (int i = threadIdx.x; i < c.n; i += blockDim.x) {
        const double x = get(i, 0), y = get(i, 1), z = get(i, 2);
        const double gm = kGauss2 * c.A->mass[c.base + i];
        c.sJ[i] = make_double4(x, y, z, gm);
        if (!S.barycentric && i >= 1 && i < end_max) {
            const double r2 = x * x + y * y + z * z;
            const double r = sqrt(r2);
            const double rm3 = 1.0 / (r2 * r);
            const double cx = gm * rm3 * x, cy = gm * rm3 * y, cz = gm * rm3 * z;
            if (i < S.end_a) { v[0] += cx; v[1] += cy; v[2] += cz; }
            if (i < S.end_b) { v[3] += cx; v[4] += cy; v[5] += cz; }
        }
    }
    // a wave whose first body is past the source end holds no source
    // (its bodies are wave_first + q * blockDim, all >= end_max)
    const bool wave_zero = S.barycentric || (int)(threadIdx.x & ~(kWave - 1)) >= end_max;
    block_sum6_all(v, c.red, c.ind, wave_zero);   // includes the barrier that publishes sJ
}

// Acceleration of body i from the LDS positions (GravityAC / GravityBC
// semantics); nn: track the nearest source.
__device__ void gravity_one(const Ctx &c, int i, bool nn, double &ax, double &ay, double &az,
                            double &rm3_out, int &nni, double &nnd)
{
    const ResSystem &S = *c.S;
    const double4 p = c.sJ[i];
    double best = DBL_MAX;
    int bidx = -1;
    double px = 0.0, py = 0.0, pz = 0.0;
    int jb, je;
    if (S.barycentric) { jb = 0; je = S.nm; }
    else { jb = 1; je = (i < S.nm) ? S.end_a : S.end_b; }
    if (!S.barycentric && i == 0) je = jb;   // the star is not integrated (AC)
    for (int j = jb; j < je; ++j) {
        if (j == i) continue;
        const double4 q = c.sJ[j];
        const double dx = q.x - p.x, dy = q.y - p.y, dz = q.z - p.z;
        const double r2 = dx * dx + dy * dy + dz * dz;
        const double y0 = __builtin_amdgcn_rsq(r2);
        const double tt = y0 * y0;
        const double e = __builtin_fma(-r2, tt, 1.0);
        const double y3 = tt * y0;
        const double mq = q.w * y3;
        const double f = __builtin_fma(mq, e * __builtin_fma(1.875, e, 1.5), mq);
        px = __builtin_fma(f, dx, px);
        py = __builtin_fma(f, dy, py);
        pz = __builtin_fma(f, dz, pz);
        if (nn) {
            if (S.barycentric ? (r2 <= best) : (r2 < best)) { best = r2; bidx = j; }
        }
    }
    rm3_out = 0.0;
    if (S.barycentric) {
        ax = px; ay = py; az = pz;
    } else if (i == 0) {
        ax = ay = az = 0.0;
    } else {
        const double r2 = p.x * p.x + p.y * p.y + p.z * p.z;
        const double r = sqrt(r2);
        const double rm3 = 1.0 / (r2 * r);
        rm3_out = rm3;
        const double mu = kGauss2 * (c.A->mass[c.base] + c.A->mass[c.base + i]);
        const int end = (i < S.nm) ? S.end_a : S.end_b;
        const int sgi = (i < S.nm) ? 0 : 3;
        double ix = c.ind[sgi], iy = c.ind[sgi + 1], iz = c.ind[sgi + 2];
        if (i < end) { ix -= p.w * rm3 * p.x; iy -= p.w * rm3 * p.y; iz -= p.w * rm3 * p.z; }
        ax = -mu * rm3 * p.x + (px - ix);
        ay = -mu * rm3 * p.y + (py - iy);
        az = -mu * rm3 * p.z + (pz - iz);
    }
    nni = bidx;
    nnd = bidx >= 0 ? sqrt(best) : 0.0;
}

// Evaluates all own bodies into stage slot l (velocity half from vel(i,c)).
template <typename Vel>
__device__ void evaluate(const Ctx &c, int l, bool last, Vel vel)
{
    for (int i = threadIdx.x; i < c.n; i += blockDim.x) {
        double ax, ay, az, rm3;
        int nni;
        double nnd;
        gravity_one(c, i, last, ax, ay, az, rm3, nni, nnd);
        const bool star_ac = !c.S->barycentric && i == 0;
        Fw(c, l, 0, i) = star_ac ? 0.0 : vel(i, 0);
        Fw(c, l, 1, i) = star_ac ? 0.0 : vel(i, 1);
        Fw(c, l, 2, i) = star_ac ? 0.0 : vel(i, 2);
        Fw(c, l, 3, i) = ax;
        Fw(c, l, 4, i) = ay;
        Fw(c, l, 5, i) = az;
        if (last) {
            if (!c.S->barycentric) c.A->rm3[c.base + i] = rm3;
            c.A->nni[c.base + i] = nni;
            c.A->nnd[c.base + i] = nnd;
        }
    }
    __syncthreads();   // sJ may be overwritten by the next stage
}

// Next-stage state of body i, component comp (first-order tableaux).  The
// stage derivatives are read first, all together, so that their latencies
// overlap; the sum then takes the nonzero coefficients in the tableau's
// order, as the reference's hand-written stage sums do.
__device__ double combine1(const Ctx &c, const double *D, int sn, int i, int comp, double h)
{
    double f[13];
#pragma unroll
    for (int l = 0; l < 13; ++l) f[l] = l < sn ? F(c, l, comp, i) : 0.0;
    const double y0 = cat(c.A->y0, c, comp, i);
    double acc = 0.0;
    bool first = true;
#pragma unroll
    for (int l = 0; l < 13; ++l) {
        if (l >= sn) break;
        const double d = D[l];
        if (d == 0.0) continue;
        const double v = d * f[l];
        acc = first ? v : acc + v;
        first = false;
    }
    return y0 + h * acc;
}

// Dormand-Prince RKN stage state (component j < 3 position, j + 3 velocity).
__device__ void combine_dp(const Ctx &c, int sn, int i, int j, double h, double &pos, double &vel)
{
    const double *A = c.coef + kDpA + sn * 8;
    double f[9];
#pragma unroll
    for (int l = 0; l < 9; ++l) f[l] = l < sn ? F(c, l, j + 3, i) : 0.0;
    const double p0 = cat(c.A->y0, c, j, i), v0 = cat(c.A->y0, c, j + 3, i);
    double var = A[0] * f[0];
#pragma unroll
    for (int l = 1; l < 9; ++l) {
        if (l >= sn) break;
        if (A[l] == 0.0) continue;
        var += A[l] * f[l];
    }
    pos = p0 + c.coef[kDpC + sn] * h * v0 + h * h * (var);
    vel = v0 + h * (var);
}

template <int INTEG>
__global__ __launch_bounds__(1024) void k_resident(ResSystem *systems, ResArrays A, int max_steps,
                                                    const double *__restrict__ coef)
{
    extern __shared__ double4 sJ[];
    __shared__ double red[16 * 6];
    __shared__ double ind[6];
    __shared__ double sh[4];
    ResSystem &S = systems[blockIdx.x];
    if (S.status != RES_RUNNING) return;
    Ctx c;
    c.S = &S; c.A = &A; c.n = S.n; c.base = S.base; c.tot = A.tot; c.sJ = sJ; c.red = red; c.ind = ind;
    c.coef = coef;
    const int nst = INTEG == INTEG_RKF78 ? 12 : INTEG == INTEG_RK4 ? 3 : 8;

    for (int step = 0; step < max_steps; ++step) {
        const double t0 = S.t;
        double h = S.h_next;
        // f0 at y0 (the Driver's first evaluation)
        stage_positions(c, [&](int i, int k) { return cat(A.y0, c, k, i); });
        evaluate(c, 0, false, [&](int i, int k) { return cat(A.y0, c, k + 3, i); });
        if (INTEG == INTEG_RKF78) {
            for (int i = threadIdx.x; i < c.n; i += blockDim.x)
                for (int k = 0; k < 6; ++k)
                    at(A.yscale, c, k, i) = fabs(cat(A.y0, c, k, i)) + fabs(h * F(c, 0, k, i)) + 1.0e-30;
        }
        int iter = 0;
        double errmax = 0.0;
        int outcome = 0;   // 0 accept, 1 retry, 2 underflow, 3 maxiter
        for (;;) {
            ++iter;
            if (INTEG == INTEG_DP) h = S.h_next;
            for (int s = 1; s <= nst; ++s) {
                // stage state from y0 and F_0..F_{s-1} (all own bodies)
                if (INTEG == INTEG_DP) {
                    for (int i = threadIdx.x; i < c.n; i += blockDim.x)
                        for (int j = 0; j < 3; ++j) {
                            double pos, vel;
                            combine_dp(c, s, i, j, h, pos, vel);
                            Fw(c, s, j, i) = vel;       // velocity half of F_s
                            at(A.y, c, j, i) = pos;     // positions of stage s (Y is free until the final)
                        }
                    stage_positions(c, [&](int i, int k) { return cat(A.y, c, k, i); });
                } else {
                    const double *D = coef + (INTEG == INTEG_RKF78 ? kRkfD + s * 12 : kRk4A + s * 3);
                    for (int i = threadIdx.x; i < c.n; i += blockDim.x) {
                        for (int k = 0; k < 3; ++k) at(A.y, c, k, i) = combine1(c, D, s, i, k, h);
                        for (int k = 3; k < 6; ++k) Fw(c, s, k - 3, i) = combine1(c, D, s, i, k, h);
                    }
                    stage_positions(c, [&](int i, int k) { return cat(A.y, c, k, i); });
                }
                evaluate(c, s, s == nst, [&](int i, int k) { return F(c, s, k, i); });
            }
            // final state and error (all own bodies)
            double ratio = 0.0;
            for (int i = threadIdx.x; i < c.n; i += blockDim.x) {
                if (INTEG == INTEG_RKF78) {
                    const double b0 = 41.0 / 840.0, b5 = 34.0 / 105.0, b6 = 9.0 / 35.0, b8 = 9.0 / 280.0,
                                 b10 = 41.0 / 840.0;
                    for (int k = 0; k < 6; ++k) {
                        const double F0 = F(c, 0, k, i), F10 = F(c, 10, k, i);
                        at(A.y, c, k, i) = cat(A.y0, c, k, i) +
                            h * (b0 * F0 + b5 * F(c, 5, k, i) + b6 * (F(c, 6, k, i) + F(c, 7, k, i)) +
                                 b8 * (F(c, 8, k, i) + F(c, 9, k, i)) + b10 * F10);
                        const double e = h * fabs(F0 + F10 - F(c, 11, k, i) - F(c, 12, k, i)) * 41.0 / 840.0;
                        at(A.err, c, k, i) = e;
                        ratio = fmax(ratio, fabs(e / cat(A.yscale, c, k, i)));
                    }
                } else if (INTEG == INTEG_RK4) {
                    const double *B = coef + kRk4B;
                    for (int k = 0; k < 6; ++k) {
                        at(A.y, c, k, i) = cat(A.y0, c, k, i) +
                            h * (B[0] * F(c, 0, k, i) + B[1] * F(c, 1, k, i) + B[2] * F(c, 2, k, i) +
                                 B[3] * F(c, 3, k, i));
                    }
                } else {
                    const double *B = coef + kDpB, *Bd = coef + kDpBdh;
                    const double h2 = h * h;
                    for (int j = 0; j < 3; ++j) {
                        const double F0 = F(c, 0, j + 3, i), F4 = F(c, 4, j + 3, i), F5 = F(c, 5, j + 3, i);
                        const double F6 = F(c, 6, j + 3, i), F7 = F(c, 7, j + 3, i), F8 = F(c, 8, j + 3, i);
                        const double p0 = cat(A.y0, c, j, i), v0 = cat(A.y0, c, j + 3, i);
                        at(A.y, c, j, i) = p0 + h * v0 + h2 * (B[0] * F0 + B[4] * F4 + B[5] * F5 + B[6] * F6 +
                                                               B[7] * F7 + B[8] * F8);
                        const double e = h2 * fabs(F7 - F8) / 20.0;
                        at(A.err, c, j, i) = e;
                        at(A.y, c, j + 3, i) = v0 + h * (Bd[0] * F0 + Bd[4] * F4 + Bd[5] * F5 + Bd[6] * F6 +
                                                         Bd[7] * F7);
                        at(A.err, c, j + 3, i) = 0.0;
                        ratio = fmax(ratio, e);
                    }
                }
            }
            const double raw = block_max_all(ratio, red, &sh[0]);
            if (threadIdx.x == 0) S.attempts += 1;
            // the step decision (identical in every thread)
            if (INTEG == INTEG_RKF78) {
                errmax = raw / S.eps;
                if (errmax < 1.0) { outcome = 0; break; }
                const double htemp = 0.9 * h * pow(errmax, -0.25);
                h = fabs(htemp) > fabs(0.1 * h) ? htemp : 0.1 * h;
                if (t0 + h == t0) { outcome = 2; break; }
            } else if (INTEG == INTEG_RK4) {
                outcome = 0;
                break;
            } else {
                errmax = raw;
                const double hn = errmax < 1.0e-20 ? 2.0 * h : 0.9 * h * pow(S.eps / errmax, 1.0 / 7.0);
                if (threadIdx.x == 0) { S.h_did = h; S.h_next = hn; }
                __syncthreads();
                if (!(errmax > S.eps && iter <= 10)) { outcome = iter > 10 ? 3 : 0; break; }
            }
        }
        __syncthreads();
        if (outcome == 2 || outcome == 3) {
            if (threadIdx.x == 0) {
                S.status = outcome == 2 ? RES_UNDERFLOW : RES_MAXITER;
                S.body_h = h;
            }
            return;
        }
        // accept: time, hDid, hNext (RungeKuttaFehlberg78.cpp:124-133 and the others)
        double h_next;
        if (INTEG == INTEG_RKF78) {
            h_next = errmax > 1.89e-4 ? (0.9 * h * pow(errmax, -0.2)) : (5.0 * h);
        } else if (INTEG == INTEG_RK4) {
            h_next = h;
        } else {
            h_next = S.h_next;
            h = S.h_did;
        }
        // swap y0 and y (contents), own bodies
        for (int i = threadIdx.x; i < c.n; i += blockDim.x)
            for (int k = 0; k < 6; ++k) {
                const double a = cat(A.y0, c, k, i);
                at(A.y0, c, k, i) = cat(A.y, c, k, i);
                at(A.y, c, k, i) = a;
            }
        __syncthreads();
        // CheckEvent's thresholds on the last stage's side outputs
        // (Solaris/Simulator.cpp:631-644, :688-695), when asked for
        int event = 0;
        if (S.clip && (S.stop_mask & SOL_STOP_EVENT) && (S.e3 > 0.0 || S.h3 > 0.0 || S.coll > 0.0)) {
            for (int i = threadIdx.x; i < c.n; i += blockDim.x) {
                const double r3 = A.rm3[c.base + i];
                if (i >= 1 && S.e3 > 0.0 && r3 < S.e3) event = 1;
                if (i >= 1 && S.h3 > 0.0 && r3 > S.h3) event = 1;
                const int j = A.nni[c.base + i];
                if (S.coll > 0.0 && j >= 0 && j < c.n &&
                    S.coll * (A.radius[c.base + i] + A.radius[c.base + j]) > A.nnd[c.base + i])
                    event = 1;
            }
        }
        event = __syncthreads_or(event);
        if (threadIdx.x == 0) {
            bool stop = false;
            int reason = 0, flush = 0;
            S.h_did = h;
            S.t = t0 + h;
            S.h_next = h_next;
            S.body_h = INTEG == INTEG_DP ? h : h_next;
            S.steps += 1;
            if (S.clip) {
                // Simulator::DecisionMaking, Solaris/Simulator.cpp:183-245
                S.elapsed += h;
                S.last_save += h;
                S.last_nsteps += h;
                if (event) reason |= SOL_STOP_EVENT;
                const double actual = 1000.0 * 365.25 * S.millenium + S.t;
                if (fabs(actual) >= fabs(S.length)) {
                    stop = true;
                } else {
                    if (fabs(actual + S.h_next) > fabs(S.length)) S.h_next = S.length - actual;
                    if (fabs(S.last_save) >= fabs(S.output)) {
                        S.last_save = 0.0;
                        reason |= SOL_STOP_OUTPUT;
                    }
                    if (fabs(S.last_save + S.h_next) > fabs(S.output)) S.h_next = S.output - S.last_save;
                }
                // Counter::succededStep % CheckForSM (:154-159)
                S.flush_count = (S.flush_count + 1) % 100;
                flush = S.flush_count == 0;
                S.reason = reason | (stop ? SOL_STOP_END : 0);
                reason &= S.stop_mask;
            }
            sh[1] = stop ? 1.0 : reason ? 2.0 : 0.0;
            sh[2] = flush ? 1.0 : 0.0;
        }
        __syncthreads();
        const double next = sh[1];
        if (sh[2] != 0.0) {
            for (int i = threadIdx.x; i < c.n; i += blockDim.x)
                for (int k = 0; k < 6; ++k) {
                    if (fabs(cat(A.y0, c, k, i)) < 1.0e-50) at(A.y0, c, k, i) = 0.0;
                    if (fabs(cat(A.y, c, k, i)) < 1.0e-50) at(A.y, c, k, i) = 0.0;
                }
        }
        __syncthreads();
        if (next != 0.0) {
            if (threadIdx.x == 0) S.status = next == 1.0 ? RES_DONE : RES_STOPPED;
            return;
        }
    }
}

}  // namespace

int resident_block_size(int n)
{
    int b = ((n + kWave - 1) / kWave) * kWave;
    if (b < kWave) b = kWave;
    if (b > 1024) b = 1024;
    return b;
}

size_t resident_lds_bytes(int nmax) { return (size_t)nmax * sizeof(double4); }

void launch_resident(hipStream_t st, int integ, ResSystem *systems, int nsys, const ResArrays &A, int max_steps,
                     const double *coef, int block, int nmax)
{
    const size_t lds = resident_lds_bytes(nmax);
    switch (integ) {
    case INTEG_RKF78:
        hipLaunchKernelGGL(k_resident<INTEG_RKF78>, dim3(nsys), dim3(block), lds, st, systems, A, max_steps, coef);
        break;
    case INTEG_RK4:
        hipLaunchKernelGGL(k_resident<INTEG_RK4>, dim3(nsys), dim3(block), lds, st, systems, A, max_steps, coef);
        break;
    default:
        hipLaunchKernelGGL(k_resident<INTEG_DP>, dim3(nsys), dim3(block), lds, st, systems, A, max_steps, coef);
        break;
    }
}

}  // namespace sol
