// pairs.hip -- direct-sum pair accelerations on gfx950.
//
// Replaces the j-loop of Acceleration::GravityAC (Solaris/Acceleration.cpp:
// 294-317) and of GravityBC_* (:558-580, :604-627).  One wavefront owns 64
// sinks (one per lane); the sources of a chunk are read as a wave-uniform
// stream of 32-byte J records, which the compiler turns into scalar loads
// (s_load_dwordx16: two bodies per load) so the pair arithmetic runs with
// SGPR operands and no LDS or barrier.  The pair sum is
//   sum_j k^2 m_j (r_j - r_i) / |r_j - r_i|^3
// with |d|^-3 from v_rsq_f64 plus one second-order correction:
//   y0 = rsq(r2), e = 1 - r2 y0^2, |d|^-3 = y0^3 (1 + 3/2 e + 15/8 e^2),
// 16 FP64 VALU operations and one v_rsq_f64 per pair.  The indirect and
// Kepler terms are added per sink in the epilogue (stages.hip).
#include <float.h>

#include <type_traits>

#include "sol_internal.h"

namespace sol {
namespace {

__device__ __forceinline__ void decode_item(const PairPlan &p, int64_t it, int &s, int &tile, int &chunk)
{
    s = (p.nseg > 1 && it >= p.seg[1].item_begin) ? 1 : 0;
    const int64_t local = it - p.seg[s].item_begin;
    tile = (int)(local / p.seg[s].nchunks);
    chunk = (int)(local - (int64_t)tile * p.seg[s].nchunks);
}

// NN: track the nearest source (BodyData::indexOfNN/distanceOfNN).
// BC: ties go to the larger index (the reference scans j downwards).
// SOFT: Plummer softening eps2 on the force (not on the NN distance).
template <bool NN, bool BC, bool SOFT>
__global__ __launch_bounds__(256) void k_pairs_fp64(PairPlan plan, const double4 *__restrict__ J,
                                                     double *__restrict__ part, double *__restrict__ nnr2,
                                                     int32_t *__restrict__ nnidx, double eps2,
                                                     double c15, double c1875)
{
    const int lane = threadIdx.x & (kWave - 1);
    const int w = blockIdx.x * (kBlock / kWave) + __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    if (w >= plan.nwaves) return;
    const int64_t lo = ((int64_t)w * plan.total_items) / plan.nwaves;
    const int64_t hi = ((int64_t)(w + 1) * plan.total_items) / plan.nwaves;

    int slot = 0, cur = -1;
    int sb = 0, se = 0, i = 0;
    double xi = 0.0, yi = 0.0, zi = 0.0;
    double ax = 0.0, ay = 0.0, az = 0.0;
    double best = DBL_MAX;
    int bidx = -1;

    auto flush = [&](int k) {
        const int64_t base = ((int64_t)w * plan.slots_per_wave + k) * (3 * kWave);
        part[base + lane] = ax;
        part[base + kWave + lane] = ay;
        part[base + 2 * kWave + lane] = az;
        if (NN) {
            const int64_t nb = ((int64_t)w * plan.slots_per_wave + k) * kWave;
            nnr2[nb + lane] = best;
            nnidx[nb + lane] = bidx;
        }
    };

    for (int64_t it = lo; it < hi; ++it) {
        int s, t, c;
        decode_item(plan, it, s, t, c);
        const PairSeg &sg = plan.seg[s];
        const int g = sg.tile_begin + t;
        if (g != cur) {
            if (cur >= 0) { flush(slot); ++slot; }
            cur = g;
            ax = ay = az = 0.0;
            best = DBL_MAX; bidx = -1;
            sb = sg.sink_begin + t * kWave;
            se = min(sb + kWave, sg.sink_end);
            i = min(sb + lane, se - 1);      // spare lanes repeat the last sink
            const double4 me = J[i];
            xi = me.x; yi = me.y; zi = me.z;
        }
        const int jb = sg.src_begin + c * plan.chunk;
        const int je = min(jb + plan.chunk, sg.src_end);
        if (jb >= se || sb >= je) {
            // the chunk holds none of this wave's sinks: no self pair
#pragma unroll 4
            for (int j = jb; j < je; ++j) {
                const double4 q = J[j];
                const double dx = q.x - xi, dy = q.y - yi, dz = q.z - zi;
                const double r2 = dx * dx + dy * dy + dz * dz;
                const double rs = SOFT ? r2 + eps2 : r2;
                const double y0 = __builtin_amdgcn_rsq(rs);
                const double tt = y0 * y0;
                const double e = __builtin_fma(-rs, tt, 1.0);
                const double y3 = tt * y0;
                const double p = __builtin_fma(c1875, e, c15);
                const double qq = e * p;
                const double mq = q.w * y3;
                const double f = __builtin_fma(mq, qq, mq);
                ax = __builtin_fma(f, dx, ax);
                ay = __builtin_fma(f, dy, ay);
                az = __builtin_fma(f, dz, az);
                if (NN) {
                    if (BC ? (r2 <= best) : (r2 < best)) { best = r2; bidx = j; }
                }
            }
        } else {
            for (int j = jb; j < je; ++j) {
                const double4 q = J[j];
                const bool self = (j == i);
                const double dx = q.x - xi, dy = q.y - yi, dz = q.z - zi;
                const double r2 = dx * dx + dy * dy + dz * dz;
                const double rs = self ? 1.0 : (SOFT ? r2 + eps2 : r2);
                const double m = self ? 0.0 : q.w;
                const double y0 = __builtin_amdgcn_rsq(rs);
                const double tt = y0 * y0;
                const double e = __builtin_fma(-rs, tt, 1.0);
                const double y3 = tt * y0;
                const double p = __builtin_fma(c1875, e, c15);
                const double qq = e * p;
                const double mq = m * y3;
                const double f = __builtin_fma(mq, qq, mq);
                ax = __builtin_fma(f, dx, ax);
                ay = __builtin_fma(f, dy, ay);
                az = __builtin_fma(f, dz, az);
                if (NN && !self) {
                    if (BC ? (r2 <= best) : (r2 < best)) { best = r2; bidx = j; }
                }
            }
        }
    }
    if (cur >= 0) flush(slot);
}

template <bool NN, bool BC, bool SOFT>
void launch_fp64(hipStream_t st, const PairPlan &plan, const double4 *J, double *part,
                 double *nnr2, int32_t *nnidx, double eps2)
{
    const int blocks = (plan.nwaves + (kBlock / kWave) - 1) / (kBlock / kWave);
    hipLaunchKernelGGL((k_pairs_fp64<NN, BC, SOFT>), dim3(blocks), dim3(kBlock), 0, st,
                       plan, J, part, nnr2, nnidx, eps2, 1.5, 1.875);
}

// Mixed precision: FP32 pair terms (differences, |d|^-3 from v_rsq_f32 and
// the products), summed in FP32 over one chunk of at most `chunk` sources
// and folded into FP64 accumulators per chunk, so that the accumulation
// error grows with the chunk length, not with N.  J32 holds the FP32
// copies (x, y, z, k^2 m) of the same positions.
//
// Close pairs are evaluated in FP64 from J.  Rounding the absolute
// positions to FP32 moves d = r_j - r_i by up to ~2^-24 (|r_i| + |r_j|), a
// relative error of 2^-24 |r_i| / |d| that grows without bound as the pair
// closes in; so a pair with |d| < near_frac |r_i| (FP32 test, a wave-level
// branch taken only when some lane has such a pair) takes the FP64 path
// instead.  With near_frac = 3e-3 the FP32 terms keep |dd|/|d| below
// ~4e-5, and a close pair's term, which can rival the star's, is exact to
// FP64 rounding.
//
// The nearest neighbour is chosen on the FP32 squared distance; its
// distance is recomputed in FP64 by the epilogue.
template <bool NN, bool BC, bool SOFT>
__global__ __launch_bounds__(256) void k_pairs_mixed(PairPlan plan, const float4 *__restrict__ J32,
                                                      const double4 *__restrict__ J,
                                                      double *__restrict__ part, double *__restrict__ nnr2,
                                                      int32_t *__restrict__ nnidx, float eps2, double eps2d,
                                                      float near_frac2, double c15, double c1875)
{
    const int lane = threadIdx.x & (kWave - 1);
    const int w = blockIdx.x * (kBlock / kWave) + __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    if (w >= plan.nwaves) return;
    const int64_t lo = ((int64_t)w * plan.total_items) / plan.nwaves;
    const int64_t hi = ((int64_t)(w + 1) * plan.total_items) / plan.nwaves;

    int slot = 0, cur = -1;
    int sb = 0, se = 0, i = 0;
    float xi = 0.f, yi = 0.f, zi = 0.f, thr2 = 0.f;
    double xd = 0.0, yd = 0.0, zd = 0.0;
    double ax = 0.0, ay = 0.0, az = 0.0;
    float best = FLT_MAX;
    int bidx = -1;

    auto flush = [&](int k) {
        const int64_t base = ((int64_t)w * plan.slots_per_wave + k) * (3 * kWave);
        part[base + lane] = ax;
        part[base + kWave + lane] = ay;
        part[base + 2 * kWave + lane] = az;
        if (NN) {
            const int64_t nb = ((int64_t)w * plan.slots_per_wave + k) * kWave;
            nnr2[nb + lane] = (double)best;
            nnidx[nb + lane] = bidx;
        }
    };
    // FP64 pair term of source j, added to the FP64 accumulators
    auto near_term = [&](int j) {
        const double4 q = J[j];
        const double dx = q.x - xd, dy = q.y - yd, dz = q.z - zd;
        const double r2 = dx * dx + dy * dy + dz * dz;
        const double rs = SOFT ? r2 + eps2d : r2;
        const double y0 = __builtin_amdgcn_rsq(rs);
        const double tt = y0 * y0;
        const double e = __builtin_fma(-rs, tt, 1.0);
        const double mq = q.w * (tt * y0);
        const double f = __builtin_fma(mq, e * __builtin_fma(c1875, e, c15), mq);
        ax = __builtin_fma(f, dx, ax);
        ay = __builtin_fma(f, dy, ay);
        az = __builtin_fma(f, dz, az);
    };

    for (int64_t it = lo; it < hi; ++it) {
        int s, t, c;
        decode_item(plan, it, s, t, c);
        const PairSeg &sg = plan.seg[s];
        const int g = sg.tile_begin + t;
        if (g != cur) {
            if (cur >= 0) { flush(slot); ++slot; }
            cur = g;
            ax = ay = az = 0.0;
            best = FLT_MAX; bidx = -1;
            sb = sg.sink_begin + t * kWave;
            se = min(sb + kWave, sg.sink_end);
            i = min(sb + lane, se - 1);
            const float4 me = J32[i];
            xi = me.x; yi = me.y; zi = me.z;
            const double4 md = J[i];
            xd = md.x; yd = md.y; zd = md.z;
            thr2 = near_frac2 * __builtin_fmaf(zi, zi, __builtin_fmaf(yi, yi, xi * xi));
        }
        const int jb = sg.src_begin + c * plan.chunk;
        const int je = min(jb + plan.chunk, sg.src_end);
        const bool has_self = !(jb >= se || sb >= je);
        float fx = 0.f, fy = 0.f, fz = 0.f;
        bool lane_near = false;   // this lane met a close pair in this chunk
        if (!has_self) {
#pragma unroll 4
            for (int j = jb; j < je; ++j) {
                const float4 q = J32[j];
                const float dx = q.x - xi, dy = q.y - yi, dz = q.z - zi;
                const float r2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
                const float y0 = __builtin_amdgcn_rsqf(SOFT ? r2 + eps2 : r2);
                const bool near = r2 < thr2;
                lane_near |= near;
                const float y0m = near ? 0.0f : y0;
                const float f = (q.w * y0m) * (y0 * y0);
                fx = __builtin_fmaf(f, dx, fx);
                fy = __builtin_fmaf(f, dy, fy);
                fz = __builtin_fmaf(f, dz, fz);
                if (NN) {
                    if (BC ? (r2 <= best) : (r2 < best)) { best = r2; bidx = j; }
                }
            }
        } else {
            for (int j = jb; j < je; ++j) {
                const float4 q = J32[j];
                const bool self = (j == i);
                const float dx = q.x - xi, dy = q.y - yi, dz = q.z - zi;
                const float r2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
                const float rs = self ? 1.0f : (SOFT ? r2 + eps2 : r2);
                const float y0 = __builtin_amdgcn_rsqf(rs);
                const bool near = !self && r2 < thr2;
                lane_near |= near;
                const float y0m = (self || near) ? 0.0f : y0;
                const float f = (q.w * y0m) * (y0 * y0);
                fx = __builtin_fmaf(f, dx, fx);
                fy = __builtin_fmaf(f, dy, fy);
                fz = __builtin_fmaf(f, dz, fz);
                if (NN && !self) {
                    if (BC ? (r2 <= best) : (r2 < best)) { best = r2; bidx = j; }
                }
            }
        }
        ax += (double)fx;
        ay += (double)fy;
        az += (double)fz;
        if (__any(lane_near)) {
            // second pass over the chunk for its close pairs, in FP64 (rare:
            // taken by the waves whose sinks have a source this close)
            for (int j = jb; j < je; ++j) {
                const float4 q = J32[j];
                const float dx = q.x - xi, dy = q.y - yi, dz = q.z - zi;
                const float r2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
                if (j != i && r2 < thr2) near_term(j);
            }
        }
    }
    if (cur >= 0) flush(slot);
}

// Pairs closer than this fraction of the sink's distance from the origin
// are evaluated in FP64 by the mixed kernel.
constexpr float kNearFrac = 3.0e-3f;

template <bool NN, bool BC, bool SOFT>
void launch_mixed(hipStream_t st, const PairPlan &plan, const float4 *J32, const double4 *J, double *part,
                  double *nnr2, int32_t *nnidx, double eps2)
{
    const int blocks = (plan.nwaves + (kBlock / kWave) - 1) / (kBlock / kWave);
    hipLaunchKernelGGL((k_pairs_mixed<NN, BC, SOFT>), dim3(blocks), dim3(kBlock), 0, st,
                       plan, J32, J, part, nnr2, nnidx, (float)eps2, eps2, kNearFrac * kNearFrac, 1.5, 1.875);
}

// ---- symmetric pair sum over the mutual block ------------------------------
// One wave per range of group-pair items (sym_plan.h).  The wave holds the
// 256 bodies of the row group, four per lane, and streams the column group
// as four packets of 64 bodies.  At step k (0..63) lane l pairs its four
// row bodies with packet body (l + k) mod 64: the packet's positions and
// k^2 m are read from the wave's own LDS slice, where the packet is stored
// twice in a row so that entry l + k needs no wrap, and the packet's force
// sums travel with it, rotating one lane per step (DPP wave_rol:1); after
// 64 steps every row body has met every packet body and the sums are back
// on their lanes.  Each step evaluates four pairs per lane and forms both
// accelerations from each, so a pair of the block costs one reciprocal
// square root and 20 FP64 operations for two interactions, plus 6 rotated
// dwords and two 16-byte LDS reads per step of four pairs.  In a diagonal
// item (row == column) the row tile t meets packet j only for t < j, and
// for t == j at steps 1..32 (step 32 on lanes 0..31): each unordered pair
// once, no self pair.
// Bodies past the end of the block are far sentinels of zero mass at
// +1e30 AU (row side) and -1e30 AU (column side), so that no pair has zero
// distance; their sums are never read.
//
// The LDS slices are private to their waves, and the four waves of a block
// run different item ranges, so there is no block barrier: a wave's LDS
// operations execute in order, and a wavefront-scope fence keeps the
// compiler from moving a packet's reads across the stores of the next.

__device__ __forceinline__ double rot_lane(double v)
{
    const int2 p = __builtin_bit_cast(int2, v);
    int2 r;
    r.x = __builtin_amdgcn_mov_dpp(p.x, 0x134, 0xF, 0xF, false);   // wave_rol:1, lane l <- lane l+1
    r.y = __builtin_amdgcn_mov_dpp(p.y, 0x134, 0xF, 0xF, false);
    return __builtin_bit_cast(double, r);
}

template <bool SOFT>
__device__ __forceinline__ void sym_pair(double x, double y, double z, double gma, double4 q, double eps2,
                                         double c15, double c1875, double &fx, double &fy, double &fz,
                                         double &gx, double &gy, double &gz)
{
    const double dx = q.x - x, dy = q.y - y, dz = q.z - z;
    const double r2 = dx * dx + dy * dy + dz * dz;
    const double rs = SOFT ? r2 + eps2 : r2;
    const double y0 = __builtin_amdgcn_rsq(rs);
    const double tt = y0 * y0;
    const double e = __builtin_fma(-rs, tt, 1.0);
    const double y3 = tt * y0;
    const double w = __builtin_fma(y3, e * __builtin_fma(c1875, e, c15), y3);   // |d|^-3
    const double fa = q.w * w, fb = gma * w;
    fx = __builtin_fma(fa, dx, fx);
    fy = __builtin_fma(fa, dy, fy);
    fz = __builtin_fma(fa, dz, fz);
    gx = __builtin_fma(-fb, dx, gx);
    gy = __builtin_fma(-fb, dy, gy);
    gz = __builtin_fma(-fb, dz, gz);
}

// 246 VGPRs, two waves per SIMD: bounding it to three (168 VGPRs, four
// values spilled outside the step loop) measured 1.1% slower at C3
template <bool SOFT>
__global__ __launch_bounds__(256) void k_pairs_sym_fp64(SymPlan sp, const double4 *__restrict__ J,
                                                        double *__restrict__ apart, double *__restrict__ bpart,
                                                        double eps2, double c15, double c1875)
{
    __shared__ double4 pk_lds[kBlock / kWave][2 * kWave];
    const int lane = threadIdx.x & (kWave - 1);
    const int wib = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const int w = blockIdx.x * (kBlock / kWave) + wib;
    if (w >= sp.nwaves) return;
    const int64_t lo = sym_wave_lo(sp, w), hi = sym_wave_lo(sp, w + 1);
    if (lo >= hi) return;
    int a0, h0;
    sym_decode(lo, a0, h0);
    int hcur = -1;
    double bs[4][3];                         // column-side sums of the current column group
    double4 *pk = pk_lds[wib];

    auto flush_col = [&]() {
        double *q = bpart + ((int64_t)w * sp.slots_per_wave + (hcur - h0)) * (3 * kSymGroup);
#pragma unroll
        for (int t = 0; t < 4; ++t)
#pragma unroll
            for (int c = 0; c < 3; ++c) q[c * kSymGroup + t * kWave + lane] = bs[t][c];
    };

    for (int64_t it = lo; it < hi; ++it) {
        int ag, hg;
        sym_decode(it, ag, hg);
        if (hg != hcur) {
            if (hcur >= 0) flush_col();
            hcur = hg;
#pragma unroll
            for (int t = 0; t < 4; ++t) bs[t][0] = bs[t][1] = bs[t][2] = 0.0;
        }
        double px[4], py[4], pz[4], pm[4], fx[4], fy[4], fz[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const int ia = 1 + ag * kSymGroup + t * kWave + lane;
            const double4 v = ia < sp.M ? J[ia] : make_double4(1.0e30, 0.0, 0.0, 0.0);
            px[t] = v.x; py[t] = v.y; pz[t] = v.z; pm[t] = v.w;
            fx[t] = fy[t] = fz[t] = 0.0;
        }
        const bool diag = (ag == hg);
        for (int j = 0; j < 4; ++j) {
            const int ib = 1 + hg * kSymGroup + j * kWave + lane;
            const double4 q0 = ib < sp.M ? J[ib] : make_double4(-1.0e30, 0.0, 0.0, 0.0);
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            pk[lane] = q0;
            pk[lane + kWave] = q0;
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            double gx = 0.0, gy = 0.0, gz = 0.0;
            // one step: the four row bodies against packet body q, then the
            // packet's sums move one lane on
            auto step = [&](const double4 &q, int k, auto diag_c) {
                constexpr bool D = decltype(diag_c)::value;
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    if (!D || t < j || (t == j && k >= 1 && (k < 32 || (k == 32 && lane < 32))))
                        sym_pair<SOFT>(px[t], py[t], pz[t], pm[t], q, eps2, c15, c1875, fx[t], fy[t], fz[t],
                                       gx, gy, gz);
                }
                gx = rot_lane(gx); gy = rot_lane(gy); gz = rot_lane(gz);
            };
            // two steps per iteration, each step's packet body read one step
            // ahead (entry lane + k + 2 <= 127 stays in the slice)
            auto steps = [&](auto diag_c) {
                double4 qa = pk[lane];
                for (int k = 0; k < kWave; k += 2) {
                    const double4 qb = pk[lane + k + 1];
                    step(qa, k, diag_c);
                    qa = pk[lane + k + 2];
                    step(qb, k + 1, diag_c);
                }
            };
            if (diag) steps(std::true_type{}); else steps(std::false_type{});
            bs[j][0] += gx;
            bs[j][1] += gy;
            bs[j][2] += gz;
        }
        double *r = apart + it * (3 * kSymGroup);
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            r[t * kWave + lane] = fx[t];
            r[kSymGroup + t * kWave + lane] = fy[t];
            r[2 * kSymGroup + t * kWave + lane] = fz[t];
        }
    }
    if (hcur >= 0) flush_col();
}

__device__ __forceinline__ float rot_lane_f(float v)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_mov_dpp(__builtin_bit_cast(int, v), 0x134, 0xF, 0xF, false));
}

// Mixed-precision form of k_pairs_sym_fp64: FP32 pair terms from J32 (and
// FP32 sums over the 256 bodies an item pairs with each body, folded into
// FP64), with the close pairs evaluated in FP64.  A pair is close when
// |d|^2 < near_frac2 |r_a|^2, r_a the row body (one decision for both
// bodies; a close pair has |r_b| = |r_a| to within |d|, so this is the
// one-directional kernel's test, |d| < near_frac |r_i|, from either side).
// The softening is folded into the squared distance (rs = eps2 + dx^2 +
// dy^2 + dz^2) and into the threshold (th = eps2 + near_frac2 |r_a|^2), so
// the test is one compare.  The FP32 pass leaves a close pair out and
// records it in a wave mask (a ballot, on the scalar unit).  The pair's
// accumulation sits under a divergent branch on that compare, so leaving it
// out costs an exec-mask update on the scalar unit rather than a select on
// the vector unit (3.7% less time per launch at 65,536 bodies); after the 64
// steps of a packet a wave with a close pair replays the steps, recomputing
// the same rs and the same test, and adds the FP64 terms of its close pairs
// (positions from J) to FP64 sums that travel with the packet.  As in the
// FP64 kernel the packet's (x, y, z, k^2 m) are read from the wave's LDS
// slice, one 16-byte read per step, and only its three FP32 sums rotate.
// Per pair: 17 FP32 VALU operations and one v_rsq_f32.
// At least four waves per SIMD (128 VGPRs; six values spilled outside the
// step loop): 1.5% faster at C3 than the unbounded three waves
template <bool SOFT>
__global__ __launch_bounds__(256, 4) void k_pairs_sym_mixed(SymPlan sp, const float4 *__restrict__ J32,
                                                         const double4 *__restrict__ J, double *__restrict__ apart,
                                                         double *__restrict__ bpart, float eps2, double eps2d,
                                                         float near_frac2, double c15, double c1875)
{
    __shared__ float4 pk_lds[kBlock / kWave][2 * kWave];
    const int lane = threadIdx.x & (kWave - 1);
    const int wib = __builtin_amdgcn_readfirstlane(threadIdx.x / kWave);
    const int w = blockIdx.x * (kBlock / kWave) + wib;
    if (w >= sp.nwaves) return;
    const int64_t lo = sym_wave_lo(sp, w), hi = sym_wave_lo(sp, w + 1);
    if (lo >= hi) return;
    int a0, h0;
    sym_decode(lo, a0, h0);
    int hcur = -1;
    double bs[4][3];
    const float e2 = SOFT ? eps2 : 0.0f;
    float4 *pk = pk_lds[wib];

    auto flush_col = [&]() {
        double *q = bpart + ((int64_t)w * sp.slots_per_wave + (hcur - h0)) * (3 * kSymGroup);
#pragma unroll
        for (int t = 0; t < 4; ++t)
#pragma unroll
            for (int c = 0; c < 3; ++c) q[c * kSymGroup + t * kWave + lane] = bs[t][c];
    };
    // diagonal items: row tile t meets packet j for t < j, and for t == j at
    // steps 1..32 (wave-uniform tests); at step 32 of t == j only lanes 0..31
    // take the pair, which is masked arithmetically, not branched on, so that
    // every lane reaches the packet rotations and the FP32 pass has no
    // divergent branch (its close-pair mask stays wave-uniform)
    auto skip = [&](bool diag, int t, int j, int k) { return diag && (t > j || (t == j && (k < 1 || k > 32))); };
    auto masked = [&](bool diag, int t, int j, int k) { return diag && t == j && k == 32 && lane >= 32; };

    for (int64_t it = lo; it < hi; ++it) {
        int ag, hg;
        sym_decode(it, ag, hg);
        if (hg != hcur) {
            if (hcur >= 0) flush_col();
            hcur = hg;
#pragma unroll
            for (int t = 0; t < 4; ++t) bs[t][0] = bs[t][1] = bs[t][2] = 0.0;
        }
        float px[4], py[4], pz[4], pm[4], th[4];
        double ax[4], ay[4], az[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            const int ia = 1 + ag * kSymGroup + t * kWave + lane;
            const float4 v = ia < sp.M ? J32[ia] : make_float4(1.0e18f, 0.0f, 0.0f, 0.0f);
            px[t] = v.x; py[t] = v.y; pz[t] = v.z; pm[t] = v.w;
            th[t] = __builtin_fmaf(near_frac2, __builtin_fmaf(v.z, v.z, __builtin_fmaf(v.y, v.y, v.x * v.x)), e2);
            ax[t] = ay[t] = az[t] = 0.0;
        }
        const bool diag = (ag == hg);
        for (int j = 0; j < 4; ++j) {
            const int ib = 1 + hg * kSymGroup + j * kWave + lane;
            const float4 v = ib < sp.M ? J32[ib] : make_float4(-1.0e18f, 0.0f, 0.0f, 0.0f);
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            pk[lane] = v;
            pk[lane + kWave] = v;
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            float fx[4] = {0.f, 0.f, 0.f, 0.f}, fy[4] = {0.f, 0.f, 0.f, 0.f}, fz[4] = {0.f, 0.f, 0.f, 0.f};
            float gx = 0.f, gy = 0.f, gz = 0.f;
            uint64_t wave_far = ~0ull;   // lanes that met no close pair in this packet's pass
            // one FP32 step against packet body q; the diagonal test is
            // resolved at compile time so that the off-diagonal loop has no branch
            auto step = [&](const float4 &q, int k, auto diag_c) {
                constexpr bool D = decltype(diag_c)::value;
#pragma unroll
                for (int t = 0; t < 4; ++t) {
                    if (D && skip(true, t, j, k)) continue;
                    const bool off = D && masked(true, t, j, k);
                    const float dx = q.x - px[t], dy = q.y - py[t], dz = q.z - pz[t];
                    const float rs = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, __builtin_fmaf(dx, dx, e2)));
                    const float y0 = __builtin_amdgcn_rsqf(rs);
                    const bool far = !(rs < th[t]);
                    wave_far &= __ballot(far);
                    if (far && !off) {
                        const float wv = y0 * (y0 * y0);
                        const float fa = q.w * wv, fb = pm[t] * wv;
                        fx[t] = __builtin_fmaf(fa, dx, fx[t]);
                        fy[t] = __builtin_fmaf(fa, dy, fy[t]);
                        fz[t] = __builtin_fmaf(fa, dz, fz[t]);
                        gx = __builtin_fmaf(-fb, dx, gx);
                        gy = __builtin_fmaf(-fb, dy, gy);
                        gz = __builtin_fmaf(-fb, dz, gz);
                    }
                }
                gx = rot_lane_f(gx); gy = rot_lane_f(gy); gz = rot_lane_f(gz);
            };
            // two steps per iteration, each packet body read one step ahead
            // (entry lane + k + 2 <= 127 stays in the slice)
            auto fp32_pass = [&](auto diag_c) {
                float4 qa = pk[lane];
                for (int k = 0; k < kWave; k += 2) {
                    const float4 qb = pk[lane + k + 1];
                    step(qa, k, diag_c);
                    qa = pk[lane + k + 2];
                    step(qb, k + 1, diag_c);
                }
            };
            if (diag) fp32_pass(std::true_type{}); else fp32_pass(std::false_type{});
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                ax[t] += (double)fx[t];
                ay[t] += (double)fy[t];
                az[t] += (double)fz[t];
            }
            double cx = (double)gx, cy = (double)gy, cz = (double)gz;
            if (wave_far != ~0ull) {
                // replay the packet's steps for the close pairs, in FP64
                double hx = 0.0, hy = 0.0, hz = 0.0;
                for (int k = 0; k < kWave; ++k) {
                    const float4 q = pk[lane + k];
#pragma unroll
                    for (int t = 0; t < 4; ++t) {
                        if (skip(diag, t, j, k)) continue;
                        const float dx = q.x - px[t], dy = q.y - py[t], dz = q.z - pz[t];
                        const float rs =
                            __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, __builtin_fmaf(dx, dx, e2)));
                        if (!masked(diag, t, j, k) && rs < th[t]) {
                            const double4 A = J[1 + ag * kSymGroup + t * kWave + lane];
                            const double4 B = J[1 + hg * kSymGroup + j * kWave + ((lane + k) & (kWave - 1))];
                            const double ex = B.x - A.x, ey = B.y - A.y, ez = B.z - A.z;
                            const double s2 = ex * ex + ey * ey + ez * ez;
                            const double r2d = SOFT ? s2 + eps2d : s2;
                            const double y0 = __builtin_amdgcn_rsq(r2d);
                            const double tt = y0 * y0;
                            const double e = __builtin_fma(-r2d, tt, 1.0);
                            const double y3 = tt * y0;
                            const double wv = __builtin_fma(y3, e * __builtin_fma(c1875, e, c15), y3);
                            const double fa = B.w * wv, fb = A.w * wv;
                            ax[t] = __builtin_fma(fa, ex, ax[t]);
                            ay[t] = __builtin_fma(fa, ey, ay[t]);
                            az[t] = __builtin_fma(fa, ez, az[t]);
                            hx = __builtin_fma(-fb, ex, hx);
                            hy = __builtin_fma(-fb, ey, hy);
                            hz = __builtin_fma(-fb, ez, hz);
                        }
                    }
                    hx = rot_lane(hx); hy = rot_lane(hy); hz = rot_lane(hz);
                }
                cx += hx; cy += hy; cz += hz;
            }
            bs[j][0] += cx;
            bs[j][1] += cy;
            bs[j][2] += cz;
        }
        double *r = apart + it * (3 * kSymGroup);
#pragma unroll
        for (int t = 0; t < 4; ++t) {
            r[t * kWave + lane] = ax[t];
            r[kSymGroup + t * kWave + lane] = ay[t];
            r[2 * kSymGroup + t * kWave + lane] = az[t];
        }
    }
    if (hcur >= 0) flush_col();
}

}  // namespace

void launch_pairs_sym_mixed(hipStream_t st, const SymPlan &sp, const float4 *J32, const double4 *J, double *apart,
                            double *bpart, double eps2)
{
    if (sp.M <= 2) return;
    const int blocks = (sp.nwaves + (kBlock / kWave) - 1) / (kBlock / kWave);
    const float nf2 = kNearFrac * kNearFrac;
    if (eps2 != 0.0)
        hipLaunchKernelGGL(k_pairs_sym_mixed<true>, dim3(blocks), dim3(kBlock), 0, st, sp, J32, J, apart, bpart,
                           (float)eps2, eps2, nf2, 1.5, 1.875);
    else
        hipLaunchKernelGGL(k_pairs_sym_mixed<false>, dim3(blocks), dim3(kBlock), 0, st, sp, J32, J, apart, bpart,
                           (float)eps2, eps2, nf2, 1.5, 1.875);
}

void launch_pairs_sym(hipStream_t st, const SymPlan &sp, const double4 *J, double *apart, double *bpart,
                      double eps2)
{
    if (sp.M <= 2) return;
    const int blocks = (sp.nwaves + (kBlock / kWave) - 1) / (kBlock / kWave);
    if (eps2 != 0.0)
        hipLaunchKernelGGL(k_pairs_sym_fp64<true>, dim3(blocks), dim3(kBlock), 0, st, sp, J, apart, bpart, eps2,
                           1.5, 1.875);
    else
        hipLaunchKernelGGL(k_pairs_sym_fp64<false>, dim3(blocks), dim3(kBlock), 0, st, sp, J, apart, bpart, eps2,
                           1.5, 1.875);
}

void launch_pairs(hipStream_t st, const PairPlan &plan, const double4 *J, double *part,
                  double *nnr2, int32_t *nnidx, bool nn, bool bc, bool mixed, double eps2,
                  const float4 *J32)
{
    if (plan.total_items == 0) return;
    const bool soft = eps2 != 0.0;
    if (mixed) {
#define SOL_MIXED(NN, BC, SOFT) launch_mixed<NN, BC, SOFT>(st, plan, J32, J, part, nnr2, nnidx, eps2)
        if (nn) {
            if (bc) { if (soft) SOL_MIXED(true, true, true); else SOL_MIXED(true, true, false); }
            else    { if (soft) SOL_MIXED(true, false, true); else SOL_MIXED(true, false, false); }
        } else {
            if (bc) { if (soft) SOL_MIXED(false, true, true); else SOL_MIXED(false, true, false); }
            else    { if (soft) SOL_MIXED(false, false, true); else SOL_MIXED(false, false, false); }
        }
#undef SOL_MIXED
        return;
    }
#define SOL_PAIRS(NN, BC, SOFT) launch_fp64<NN, BC, SOFT>(st, plan, J, part, nnr2, nnidx, eps2)
    if (nn) {
        if (bc) { if (soft) SOL_PAIRS(true, true, true); else SOL_PAIRS(true, true, false); }
        else    { if (soft) SOL_PAIRS(true, false, true); else SOL_PAIRS(true, false, false); }
    } else {
        if (bc) { if (soft) SOL_PAIRS(false, true, true); else SOL_PAIRS(false, true, false); }
        else    { if (soft) SOL_PAIRS(false, false, true); else SOL_PAIRS(false, false, false); }
    }
#undef SOL_PAIRS
}

}  // namespace sol
