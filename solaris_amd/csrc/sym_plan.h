// sym_plan.h -- work plan of the symmetric pair sum (pairs.hip, k_pairs_sym_fp64).
//
// The mutual block is bodies [1, M): every body of it is a source of every
// other (the self-interacting bodies of GravityAC, Solaris/Acceleration.cpp:
// 285-317).  Each unordered pair of the block is evaluated once and both
// accelerations are formed from it.  The block is cut into groups of 256
// bodies; a work item is an unordered pair of groups (row a <= column h),
// numbered column by column: item = h (h + 1) / 2 + a.  The items are split
// into nwaves contiguous ranges of equal length (to within one).
//
// Partial sums:
//   row side    (the bodies of group a) -- one record per item, in
//               apart[item][3][256];
//   column side (the bodies of group h) -- summed by a wave over the
//               consecutive items of one column, one record per (wave,
//               column touched) in bpart[wave * slots_per_wave + (h - h_first)]
//               [3][256], h_first the column of the wave's first item.
// A body in group g thus finds its partials in items (g, h), h >= g, and in
// the records of the waves that cover column g.
//
// Shared by the kernels and the host (plan construction, tests).
#pragma once

#include <math.h>
#include <stdint.h>

#if defined(__HIPCC__)
#define SOL_HD __host__ __device__
#else
#define SOL_HD
#endif

namespace sol {

constexpr int kSymGroup = 256;     // bodies per group: four tiles of 64, one body of each per lane

struct SymPlan {
    int32_t M;                     // mutual block [1, M); 0 when the symmetric sum is off
    int32_t ngroups;
    int64_t nitems;                // ngroups (ngroups + 1) / 2
    int32_t nwaves;
    int32_t slots_per_wave;        // column records per wave
};

SOL_HD inline int64_t sym_col_start(int64_t h) { return h * (h + 1) / 2; }

// item -> (row a, column h), 0 <= a <= h
SOL_HD inline void sym_decode(int64_t item, int &a, int &h)
{
    int64_t hh = (int64_t)((sqrt(8.0 * (double)item + 1.0) - 1.0) * 0.5);
    while (hh > 0 && sym_col_start(hh) > item) --hh;
    while (sym_col_start(hh + 1) <= item) ++hh;
    h = (int)hh;
    a = (int)(item - sym_col_start(hh));
}

SOL_HD inline int64_t sym_wave_lo(const SymPlan &p, int64_t w) { return (w * p.nitems) / p.nwaves; }

// the wave whose range [lo, hi) holds item x
SOL_HD inline int64_t sym_wave_of(const SymPlan &p, int64_t x)
{
    const int64_t w = ((x + 1) * (int64_t)p.nwaves - 1) / p.nitems;
    return w < p.nwaves - 1 ? w : p.nwaves - 1;
}

// Fills ngroups, nitems and slots_per_wave for mutual block [1, M) and
// `waves` waves (at most one wave per item).
inline void sym_plan_make(SymPlan &p, int M, int64_t waves)
{
    p = SymPlan{};
    if (M <= 2) return;
    p.M = M;
    p.ngroups = (M - 1 + kSymGroup - 1) / kSymGroup;
    p.nitems = (int64_t)p.ngroups * (p.ngroups + 1) / 2;
    p.nwaves = (int32_t)(waves < p.nitems ? waves : p.nitems);
    if (p.nwaves < 1) p.nwaves = 1;
    int spw = 1;
    for (int64_t w = 0; w < p.nwaves; ++w) {
        const int64_t lo = sym_wave_lo(p, w), hi = sym_wave_lo(p, w + 1);
        if (hi <= lo) continue;
        int a0, h0, a1, h1;
        sym_decode(lo, a0, h0);
        sym_decode(hi - 1, a1, h1);
        if (h1 - h0 + 1 > spw) spw = h1 - h0 + 1;
    }
    p.slots_per_wave = spw;
}

}  // namespace sol
