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

#include <type_traits>

#include "sol_internal.h"
#include "resident.h"
#include "tableaux.h"

namespace sol {
namespace {

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

// pow of the step-size control, once per attempt.  Out of line: inlined,
// its polynomial constants are hoisted out of the step loop and spilled.
__device__ __attribute__((noinline)) double step_pow(double x, double y) { return pow(x, y); }

// Per-launch constants of the block's system, held in registers (the
// ResSystem fields would otherwise be re-read from memory after every
// barrier).
struct Ctx {
    const ResArrays *A;
    int n, base, nm, end_a, end_b, end_max;
    bool bc;
    int64_t tot;
    double4 *sJ;        // LDS: positions + k^2 m of the stage
    double *red;        // LDS scratch
};

__device__ __forceinline__ double &at(double *p, const Ctx &c, int comp, int i) { return p[comp * c.tot + c.base + i]; }
__device__ __forceinline__ double cat(const double *p, const Ctx &c, int comp, int i) { return p[comp * c.tot + c.base + i]; }

// The own bodies' accesses go through a block-uniform array pointer plus a
// 32-bit per-lane byte offset, vb = 8 (k tot + base + i) for component k of
// body i, which the global load/store instructions add themselves: no
// per-lane 64-bit address has to be formed or kept for each array and
// stage.
__device__ __forceinline__ double ld(const double *p, uint32_t vb) { return *(const double *)((const char *)p + vb); }
__device__ __forceinline__ void st(double *p, uint32_t vb, double v) { *(double *)((char *)p + vb) = v; }

// Derivatives of stage l (0 = f0 in ACC0).
__device__ __forceinline__ double *stage(const Ctx &c, int l)
{
    return l == 0 ? c.A->acc0 : c.A->fst + (l - 1) * 6 * c.tot;
}

// Byte offsets of the six components of own body i.  The empty asm makes i
// opaque, so the offsets are formed here, next to their loads and stores,
// and not hoisted out of the step loop as one live 64-bit address per
// array, stage and component (which the 128 registers of a 1024-thread
// block cannot hold).
__device__ __forceinline__ void offsets(const Ctx &c, int i, uint32_t vb[6])
{
    asm volatile("" : "+v"(i));
#pragma unroll
    for (int k = 0; k < 6; ++k) vb[k] = (uint32_t)((k * c.tot + c.base + i) * 8);
}

// Compile-time loop: fn(std::integral_constant<int, L>) for L in [B, E).
template <int B, int E, typename Fn>
__device__ __forceinline__ void static_for(Fn &&fn)
{
    if constexpr (B < E) {
        fn(std::integral_constant<int, B>{});
        static_for<B + 1, E>(fn);
    }
}

template <int INTEG>
constexpr int n_stages() { return INTEG == INTEG_RKF78 ? 12 : INTEG == INTEG_RK4 ? 3 : 8; }

// Entry l of stage s of the first-order tableaux (RKF78 D, RK4 a).
template <int INTEG>
constexpr double dcoef(int s, int l) { return INTEG == INTEG_RKF78 ? tab::kRkfD[s][l] : tab::kRk4A[s][l]; }

// Own bodies of this thread: i = threadIdx.x + q * blockDim.x, q < BPT.
__device__ __forceinline__ int own(int q) { return threadIdx.x + q * blockDim.x; }

// Publishes the stage positions of the own bodies to LDS and forms the
// indirect-term sums of both source ranges (AC) in ind (sum over the first
// source range in ind[0..2], over the second in ind[3..5]); the barrier
// inside makes sJ visible to every sink.
template <int BPT>
__device__ void publish(const Ctx &c, const double (&pos)[BPT][3], const double (&gm)[BPT], double (&ind)[6])
{
    double v[6] = {0, 0, 0, 0, 0, 0};
#pragma unroll
    for (int q = 0; q < BPT; ++q) {
        const int i = own(q);
        if (i >= c.n) break;
        const double x = pos[q][0], y = pos[q][1], z = pos[q][2];
        c.sJ[i] = make_double4(x, y, z, gm[q]);
        // only the bodies below the larger source end enter the indirect sums
        if (!c.bc && i >= 1 && i < c.end_max) {
            const double r2 = x * x + y * y + z * z;
            const double r = sqrt(r2);
            const double rm3 = 1.0 / (r2 * r);
            const double cx = gm[q] * rm3 * x, cy = gm[q] * rm3 * y, cz = gm[q] * rm3 * z;
            if (i < c.end_a) { v[0] += cx; v[1] += cy; v[2] += cz; }
            if (i < c.end_b) { v[3] += cx; v[4] += cy; v[5] += cz; }
        }
    }
    // Only the waves whose first body is below the source end hold sources
    // (a wave's bodies are wave_first + q * blockDim); the others' partials
    // are +0.0 and are left out.  The source waves reduce their partials and
    // publish them; after the one barrier (which also publishes sJ) every
    // thread adds the published partials in wave order itself.
    const int nsw = c.bc ? 0 : min((int)(blockDim.x / kWave), (c.end_max + kWave - 1) / kWave);
    const int wid = threadIdx.x / kWave;
    if (wid < nsw) {
        for (int k = 0; k < 6; ++k) v[k] = wsum(v[k]);
        if ((threadIdx.x & (kWave - 1)) == 0)
            for (int k = 0; k < 6; ++k) c.red[wid * 6 + k] = v[k];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        double s = nsw > 0 ? c.red[k] : 0.0;
        for (int w = 1; w < nsw; ++w) s += c.red[w * 6 + k];
        ind[k] = s;
    }
}

// Acceleration of body i from the LDS positions (GravityAC / GravityBC
// semantics, Solaris/Acceleration.cpp:248-329, :530-634); mu = k^2 (m_0 +
// m_i); nn: track the nearest source.
__device__ void gravity_one(const Ctx &c, int i, double mu, const double (&ind)[6], bool nn, double &ax,
                            double &ay, double &az, double &rm3_out, int &nni, double &nnd)
{
    const double4 p = c.sJ[i];
    double best = DBL_MAX;
    int bidx = -1;
    double px = 0.0, py = 0.0, pz = 0.0;
    int jb, je;
    if (c.bc) { jb = 0; je = c.nm; }
    else { jb = 1; je = (i < c.nm) ? c.end_a : c.end_b; }
    if (!c.bc && i == 0) je = jb;   // the star is not integrated (AC)
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
            if (c.bc ? (r2 <= best) : (r2 < best)) { best = r2; bidx = j; }
        }
    }
    rm3_out = 0.0;
    if (c.bc) {
        ax = px; ay = py; az = pz;
    } else if (i == 0) {
        ax = ay = az = 0.0;
    } else {
        const double r2 = p.x * p.x + p.y * p.y + p.z * p.z;
        const double r = sqrt(r2);
        const double rm3 = 1.0 / (r2 * r);
        rm3_out = rm3;
        const int end = (i < c.nm) ? c.end_a : c.end_b;
        const int sgi = (i < c.nm) ? 0 : 3;
        double ix = ind[sgi], iy = ind[sgi + 1], iz = ind[sgi + 2];
        if (i < end) { ix -= p.w * rm3 * p.x; iy -= p.w * rm3 * p.y; iz -= p.w * rm3 * p.z; }
        ax = -mu * rm3 * p.x + (px - ix);
        ay = -mu * rm3 * p.y + (py - iy);
        az = -mu * rm3 * p.z + (pz - iz);
    }
    nni = bidx;
    nnd = bidx >= 0 ? sqrt(best) : 0.0;
}

// Evaluates the own bodies into stage slot l: the velocity half from vel,
// the acceleration into acc (and slot l); the last stage also writes the
// side outputs.  Ends with the barrier after which sJ may be overwritten.
template <int BPT>
__device__ void evaluate(const Ctx &c, int l, bool last, const double (&vel)[BPT][3], const double (&mu)[BPT],
                         const double (&ind)[6], double (&acc)[BPT][3])
{
#pragma unroll
    for (int q = 0; q < BPT; ++q) {
        const int i = own(q);
        if (i >= c.n) break;
        double ax, ay, az, rm3, nnd;
        int nni;
        gravity_one(c, i, mu[q], ind, last, ax, ay, az, rm3, nni, nnd);
        const bool star_ac = !c.bc && i == 0;
        uint32_t vb[6];
        offsets(c, i, vb);
        double *f = stage(c, l);
        st(f, vb[0], star_ac ? 0.0 : vel[q][0]);
        st(f, vb[1], star_ac ? 0.0 : vel[q][1]);
        st(f, vb[2], star_ac ? 0.0 : vel[q][2]);
        st(f, vb[3], ax);
        st(f, vb[4], ay);
        st(f, vb[5], az);
        acc[q][0] = ax; acc[q][1] = ay; acc[q][2] = az;
        if (last) {
            if (!c.bc) c.A->rm3[c.base + i] = rm3;
            c.A->nni[c.base + i] = nni;
            c.A->nnd[c.base + i] = nnd;
        }
    }
    __syncthreads();
}

// State of stage S (the reference's yTemp before the S-th evaluation) for
// the own bodies: positions into pos, velocities into vel.  The stage
// derivatives with nonzero coefficients and y0 are read first, all
// together, so that their latencies overlap.
//
// First-order tableaux (RKF78, RK4): y0 + h * sum, the sum over the nonzero
// coefficients in the tableau's order, as the reference's hand-written
// stage sums (RungeKuttaFehlberg78.cpp:147-240, RungeKutta4.cpp:58-130).
// Dormand-Prince RKN (DormandPrince.cpp:246-491): positions
// p0 + c_S h v0 + h^2 var and velocities v0 + h var, var = sum_l a_{S,l} F_l
// over the acceleration halves.
template <int INTEG, int S, int BPT>
__device__ __forceinline__ void stage_state(const Ctx &c, double h, double (&pos)[BPT][3], double (&vel)[BPT][3])
{
#pragma unroll
    for (int q = 0; q < BPT; ++q) {
        const int i = own(q);
        if (i >= c.n) break;
        uint32_t vb[6];
        offsets(c, i, vb);
        if constexpr (INTEG == INTEG_DP) {
            double p0[3], v0[3], f[3][S];
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                p0[j] = ld(c.A->y0, vb[j]);
                v0[j] = ld(c.A->y0, vb[j + 3]);
                static_for<0, S>([&](auto L) {
                    if constexpr (L == 0 || tab::kDpA[S][L] != 0.0) f[j][L] = ld(stage(c, L), vb[j + 3]);
                });
            }
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                double var = tab::kDpA[S][0] * f[j][0];
                static_for<1, S>([&](auto L) {
                    if constexpr (tab::kDpA[S][L] != 0.0) var += tab::kDpA[S][L] * f[j][L];
                });
                pos[q][j] = p0[j] + tab::kDpC[S] * h * v0[j] + h * h * (var);
                vel[q][j] = v0[j] + h * (var);
            }
        } else {
            // positions, then velocities: three components' loads at a time
#pragma unroll
            for (int half = 0; half < 2; ++half) {
                double y0[3], f[3][S];
#pragma unroll
                for (int kk = 0; kk < 3; ++kk) {
                    const uint32_t o = vb[3 * half + kk];
                    y0[kk] = ld(c.A->y0, o);
                    static_for<0, S>([&](auto L) {
                        if constexpr (dcoef<INTEG>(S, L) != 0.0) f[kk][L] = ld(stage(c, L), o);
                    });
                }
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int kk = 0; kk < 3; ++kk) {
                    double acc = 0.0;
                    bool first = true;
                    static_for<0, S>([&](auto L) {
                        constexpr double d = dcoef<INTEG>(S, L);
                        if constexpr (d != 0.0) {
                            const double v = d * f[kk][L];
                            acc = first ? v : acc + v;
                            first = false;
                        }
                    });
                    const double out = y0[kk] + h * acc;
                    if (half == 0) pos[q][kk] = out;
                    else vel[q][kk] = out;
                }
            }
        }
    }
}

// stage_state for the run-time stage number s (1 <= s <= n_stages).
template <int INTEG, int BPT, int S = 1>
__device__ __forceinline__ void stage_state_at(int s, const Ctx &c, double h, double (&pos)[BPT][3],
                                               double (&vel)[BPT][3])
{
    if constexpr (S <= n_stages<INTEG>()) {
        if (s == S) stage_state<INTEG, S, BPT>(c, h, pos, vel);
        else stage_state_at<INTEG, BPT, S + 1>(s, c, h, pos, vel);
    }
}

// Final state into Y and the error estimate into ERR for the own bodies;
// returns this thread's contribution to the error maximum (RKF78:
// max |err / yscale|, Dormand-Prince: max position error).
template <int INTEG, int BPT>
__device__ double final_state(const Ctx &c, double h)
{
    double ratio = 0.0;
#pragma unroll
    for (int q = 0; q < BPT; ++q) {
        const int i = own(q);
        if (i >= c.n) break;
        uint32_t vb[6];
        offsets(c, i, vb);
        const double *y0 = c.A->y0;
        if constexpr (INTEG == INTEG_RKF78) {
            // RungeKuttaFehlberg78.cpp:231-248
            constexpr double b0 = 41.0 / 840.0, b5 = 34.0 / 105.0, b6 = 9.0 / 35.0, b8 = 9.0 / 280.0,
                             b10 = 41.0 / 840.0;
            double y[6], e[6];
#pragma unroll
            for (int half = 0; half < 2; ++half) {
                double f[3][11], y0v[3], sc[3];
#pragma unroll
                for (int kk = 0; kk < 3; ++kk) {
                    const uint32_t o = vb[3 * half + kk];
                    f[kk][0] = ld(stage(c, 0), o);
#pragma unroll
                    for (int l = 5; l <= 12; ++l) f[kk][l - 4] = ld(stage(c, l), o);
                    y0v[kk] = ld(y0, o);
                    sc[kk] = ld(c.A->yscale, o);
                }
                __builtin_amdgcn_sched_barrier(0);
#pragma unroll
                for (int kk = 0; kk < 3; ++kk) {
                    const int k = 3 * half + kk;
                    const double F0 = f[kk][0], F5 = f[kk][1], F6 = f[kk][2], F7 = f[kk][3], F8 = f[kk][4];
                    const double F9 = f[kk][5], F10 = f[kk][6], F11 = f[kk][7], F12 = f[kk][8];
                    y[k] = y0v[kk] + h * (b0 * F0 + b5 * F5 + b6 * (F6 + F7) + b8 * (F8 + F9) + b10 * F10);
                    e[k] = h * fabs(F0 + F10 - F11 - F12) * 41.0 / 840.0;
                    ratio = fmax(ratio, fabs(e[k] / sc[kk]));
                }
            }
#pragma unroll
            for (int k = 0; k < 6; ++k) { st(c.A->y, vb[k], y[k]); st(c.A->err, vb[k], e[k]); }
        } else if constexpr (INTEG == INTEG_RK4) {
            double y[6];
#pragma unroll
            for (int k = 0; k < 6; ++k) {
                const uint32_t o = vb[k];
                y[k] = ld(y0, o) + h * (tab::kRk4B[0] * ld(stage(c, 0), o) + tab::kRk4B[1] * ld(stage(c, 1), o) +
                                        tab::kRk4B[2] * ld(stage(c, 2), o) + tab::kRk4B[3] * ld(stage(c, 3), o));
            }
#pragma unroll
            for (int k = 0; k < 6; ++k) st(c.A->y, vb[k], y[k]);
        } else {
            // DormandPrince.cpp:455-491
            constexpr const double *B = tab::kDpB, *Bd = tab::kDpBdh;
            const double h2 = h * h;
            double p[3], v[3], e[3];
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                const uint32_t o = vb[j + 3];
                const double F0 = ld(stage(c, 0), o), F4 = ld(stage(c, 4), o), F5 = ld(stage(c, 5), o);
                const double F6 = ld(stage(c, 6), o), F7 = ld(stage(c, 7), o), F8 = ld(stage(c, 8), o);
                const double p0 = ld(y0, vb[j]), v0 = ld(y0, o);
                p[j] = p0 + h * v0 + h2 * (B[0] * F0 + B[4] * F4 + B[5] * F5 + B[6] * F6 + B[7] * F7 + B[8] * F8);
                e[j] = h2 * fabs(F7 - F8) / 20.0;
                v[j] = v0 + h * (Bd[0] * F0 + Bd[4] * F4 + Bd[5] * F5 + Bd[6] * F6 + Bd[7] * F7);
                ratio = fmax(ratio, e[j]);
            }
#pragma unroll
            for (int j = 0; j < 3; ++j) {
                st(c.A->y, vb[j], p[j]);
                st(c.A->err, vb[j], e[j]);
                st(c.A->y, vb[j + 3], v[j]);
                st(c.A->err, vb[j + 3], 0.0);
            }
        }
    }
    return ratio;
}

// One workgroup per system; BPT bodies per thread at most
// (n <= BPT * blockDim.x); blocks of BPT = 4 have at most 512 threads.
template <int INTEG, int BPT>
__global__ __launch_bounds__(BPT >= 4 ? 512 : 1024) void k_resident(ResSystem *systems, ResArrays A, int max_steps)
{
    extern __shared__ double4 sJ[];
    __shared__ double red[16 * 6];
    __shared__ double sh[4];
    ResSystem &S = systems[blockIdx.x];
    if (S.status != RES_RUNNING) return;
    Ctx c;
    c.A = &A; c.n = S.n; c.base = S.base; c.nm = S.nm; c.end_a = S.end_a; c.end_b = S.end_b;
    c.end_max = c.end_a > c.end_b ? c.end_a : c.end_b;
    c.bc = S.barycentric != 0;
    c.tot = A.tot; c.sJ = sJ; c.red = red;
    constexpr int nst = n_stages<INTEG>();

    // k^2 m_i and k^2 (m_0 + m_i) of the own bodies, for the whole launch
    double gm[BPT], mu[BPT];
#pragma unroll
    for (int q = 0; q < BPT; ++q) {
        const int i = own(q) < c.n ? own(q) : 0;
        gm[q] = kGauss2 * A.mass[c.base + i];
        mu[q] = kGauss2 * (A.mass[c.base] + A.mass[c.base + i]);
    }

    for (int step = 0; step < max_steps; ++step) {
        const double t0 = S.t;
        double h = S.h_next;
        // f0 at y0 (the Driver's first evaluation), and yscale
        {
            double y0[BPT][6], pos[BPT][3], vel[BPT][3], acc[BPT][3], ind[6];
#pragma unroll
            for (int q = 0; q < BPT; ++q) {
                uint32_t vb[6];
                offsets(c, own(q) < c.n ? own(q) : 0, vb);
#pragma unroll
                for (int k = 0; k < 6; ++k) y0[q][k] = ld(A.y0, vb[k]);
#pragma unroll
                for (int k = 0; k < 3; ++k) { pos[q][k] = y0[q][k]; vel[q][k] = y0[q][k + 3]; }
            }
            publish<BPT>(c, pos, gm, ind);
            evaluate<BPT>(c, 0, false, vel, mu, ind, acc);
            if (INTEG == INTEG_RKF78) {
#pragma unroll
                for (int q = 0; q < BPT; ++q) {
                    const int i = own(q);
                    if (i >= c.n) break;
                    const bool star_ac = !c.bc && i == 0;
                    uint32_t vb[6];
                    offsets(c, i, vb);
#pragma unroll
                    for (int k = 0; k < 6; ++k) {
                        const double f = star_ac ? 0.0 : k < 3 ? vel[q][k] : acc[q][k - 3];
                        st(A.yscale, vb[k], fabs(y0[q][k]) + fabs(h * f) + 1.0e-30);
                    }
                }
            }
        }
        int iter = 0;
        double errmax = 0.0;
        int outcome = 0;   // 0 accept, 1 retry, 2 underflow, 3 maxiter
        for (;;) {
            ++iter;
            if (INTEG == INTEG_DP) h = S.h_next;
            for (int s = 1; s <= nst; ++s) {
                double pos[BPT][3], vel[BPT][3], acc[BPT][3], ind[6];
                stage_state_at<INTEG, BPT>(s, c, h, pos, vel);
                publish<BPT>(c, pos, gm, ind);
                evaluate<BPT>(c, s, s == nst, vel, mu, ind, acc);
            }
            const double ratio = final_state<INTEG, BPT>(c, h);
            const double raw = block_max_all(ratio, red, &sh[0]);
            if (threadIdx.x == 0) S.attempts += 1;
            // the step decision (identical in every thread)
            if (INTEG == INTEG_RKF78) {
                errmax = raw / S.eps;
                if (errmax < 1.0) { outcome = 0; break; }
                const double htemp = 0.9 * h * step_pow(errmax, -0.25);
                h = fabs(htemp) > fabs(0.1 * h) ? htemp : 0.1 * h;
                if (t0 + h == t0) { outcome = 2; break; }
            } else if (INTEG == INTEG_RK4) {
                outcome = 0;
                break;
            } else {
                errmax = raw;
                const double hn = errmax < 1.0e-20 ? 2.0 * h : 0.9 * h * step_pow(S.eps / errmax, 1.0 / 7.0);
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
            h_next = errmax > 1.89e-4 ? (0.9 * h * step_pow(errmax, -0.2)) : (5.0 * h);
        } else if (INTEG == INTEG_RK4) {
            h_next = h;
        } else {
            h_next = S.h_next;
            h = S.h_did;
        }
        // swap y0 and y (contents), own bodies
#pragma unroll
        for (int q = 0; q < BPT; ++q) {
            const int i = own(q);
            if (i >= c.n) break;
            uint32_t vb[6];
            offsets(c, i, vb);
            double a[6], b[6];
#pragma unroll
            for (int k = 0; k < 6; ++k) { a[k] = ld(A.y0, vb[k]); b[k] = ld(A.y, vb[k]); }
#pragma unroll
            for (int k = 0; k < 6; ++k) { st(A.y0, vb[k], b[k]); st(A.y, vb[k], a[k]); }
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

namespace {

template <int INTEG>
void launch_bpt(hipStream_t st, ResSystem *systems, int nsys, const ResArrays &A, int max_steps, int block,
                int bpt, size_t lds)
{
    switch (bpt) {
    case 1: hipLaunchKernelGGL((k_resident<INTEG, 1>), dim3(nsys), dim3(block), lds, st, systems, A, max_steps); break;
    case 2: hipLaunchKernelGGL((k_resident<INTEG, 2>), dim3(nsys), dim3(block), lds, st, systems, A, max_steps); break;
    default: hipLaunchKernelGGL((k_resident<INTEG, 4>), dim3(nsys), dim3(block), lds, st, systems, A, max_steps); break;
    }
}

}  // namespace

// Bodies per thread: the smallest of 1, 2, 4 that covers nmax.  A block too
// small for 4 per thread is widened to the smallest that is not, and one of
// more than 512 threads that would need 4 to 1024 threads (2 per thread
// cover kResidentMaxBodies).
void launch_resident(hipStream_t st, int integ, ResSystem *systems, int nsys, const ResArrays &A, int max_steps,
                     int block, int nmax)
{
    const size_t lds = resident_lds_bytes(nmax);
    if (block * 4 < nmax) block = ((nmax + 3) / 4 + kWave - 1) / kWave * kWave;
    if (block > 512 && nmax > 2 * block) block = 1024;
    const int bpt = nmax <= block ? 1 : nmax <= 2 * block ? 2 : 4;
    switch (integ) {
    case INTEG_RKF78: launch_bpt<INTEG_RKF78>(st, systems, nsys, A, max_steps, block, bpt, lds); break;
    case INTEG_RK4: launch_bpt<INTEG_RK4>(st, systems, nsys, A, max_steps, block, bpt, lds); break;
    default: launch_bpt<INTEG_DP>(st, systems, nsys, A, max_steps, block, bpt, lds); break;
    }
}

}  // namespace sol
