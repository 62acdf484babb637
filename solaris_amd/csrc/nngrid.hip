// nngrid.hip -- exact nearest-neighbour search over a uniform grid.
//
// The side outputs BodyData::indexOfNN / distanceOfNN of GravityAC and
// GravityBC (Solaris/Acceleration.cpp:296-312, :558-627): for every sink,
// the nearest body of its source range (the reference's type-based ranges,
// as the pair plan's segments give them), the sink itself excluded.  Ties
// go to the smaller index (AC) or to the larger (BC), as the reference's
// ascending scans with "<" and "<=" leave them.
//
// Instead of testing all N^2 pairs, the sources are counting-sorted into a
// uniform grid of cubic cells sized for about two sources per cell, and
// each sink searches the cells ring by ring (Chebyshev shells around its
// own cell) until no unvisited cell can hold a source nearer than its best
// one.  The squared distance is formed as in the all-pairs search
// (d = r_j - r_i, d.x^2 + d.y^2 + d.z^2) and candidates are compared on
// (r^2, index), so the result is that of the all-pairs scan whatever order
// the cells are visited in.
//
// Launches per search: bounding box partials, grid set-up (one block:
// the cell size by bisection, counts cleared), count, scan (one block),
// scatter, search.  All are O(N).
#include <float.h>

#include <algorithm>

#include "sol_internal.h"

namespace sol {
namespace {

constexpr int kGB = 256;          // block of the per-body kernels
constexpr int kScanBlock = 1024;  // the one-block scan

__global__ __launch_bounds__(kGB) void k_grid_bbox(const double4 *__restrict__ J, int s0, int s1,
                                                   double *__restrict__ part)
{
    __shared__ double sm[6][kGB];
    double mn[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, mx[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
    for (int j = s0 + blockIdx.x * kGB + threadIdx.x; j < s1; j += gridDim.x * kGB) {
        const double4 q = J[j];
        mn[0] = fmin(mn[0], q.x); mn[1] = fmin(mn[1], q.y); mn[2] = fmin(mn[2], q.z);
        mx[0] = fmax(mx[0], q.x); mx[1] = fmax(mx[1], q.y); mx[2] = fmax(mx[2], q.z);
    }
    for (int d = 0; d < 3; ++d) { sm[d][threadIdx.x] = mn[d]; sm[3 + d][threadIdx.x] = mx[d]; }
    __syncthreads();
    for (int s = kGB / 2; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s)
            for (int d = 0; d < 3; ++d) {
                sm[d][threadIdx.x] = fmin(sm[d][threadIdx.x], sm[d][threadIdx.x + s]);
                sm[3 + d][threadIdx.x] = fmax(sm[3 + d][threadIdx.x], sm[3 + d][threadIdx.x + s]);
            }
        __syncthreads();
    }
    if (threadIdx.x < 6) part[blockIdx.x * 6 + threadIdx.x] = sm[threadIdx.x][0];
}

__device__ __forceinline__ int64_t grid_cells(const double e[3], double h)
{
    int64_t c = 1;
    for (int d = 0; d < 3; ++d) c *= e[d] > 0.0 ? (int64_t)ceil(e[d] / h) : 1;
    return c;
}

// Reduces the box, picks the cell size h (the smallest of a geometric
// sweep of candidates for which the grid over the box has at most
// `target` cells, target = nsrc / 2 capped at max_cells; two sweeps of kGB
// candidates each) and clears the counts.
__global__ __launch_bounds__(kGB) void k_grid_setup(const double *__restrict__ part, int nb, int nsrc, int max_cells,
                                                    GridParams *__restrict__ gp, int32_t *__restrict__ count)
{
    __shared__ double red[6][kGB];
    __shared__ int first_ok[kGB];
    __shared__ double hlo_sh, hhi_sh;
    __shared__ int nc_sh;
    const int t = threadIdx.x;
    double mn[3] = {DBL_MAX, DBL_MAX, DBL_MAX}, mx[3] = {-DBL_MAX, -DBL_MAX, -DBL_MAX};
    for (int b = t; b < nb; b += kGB)
        for (int d = 0; d < 3; ++d) {
            mn[d] = fmin(mn[d], part[b * 6 + d]);
            mx[d] = fmax(mx[d], part[b * 6 + 3 + d]);
        }
    for (int d = 0; d < 3; ++d) { red[d][t] = mn[d]; red[3 + d][t] = mx[d]; }
    __syncthreads();
    for (int s = kGB / 2; s > 0; s >>= 1) {
        if (t < s)
            for (int d = 0; d < 3; ++d) {
                red[d][t] = fmin(red[d][t], red[d][t + s]);
                red[3 + d][t] = fmax(red[3 + d][t], red[3 + d][t + s]);
            }
        __syncthreads();
    }
    double lo[3], e[3], emax = 0.0;
    for (int d = 0; d < 3; ++d) {
        lo[d] = nsrc > 0 ? red[d][0] : 0.0;
        e[d] = nsrc > 0 ? red[3 + d][0] - red[d][0] : 0.0;
        emax = fmax(emax, e[d]);
    }
    const bool one_cell = !(emax > 0.0) || !(emax < DBL_MAX);   // one point, or a non-finite box
    if (one_cell)
        for (int d = 0; d < 3; ++d) e[d] = 0.0;
    const double target = fmin(fmax(1.0, 0.5 * nsrc), (double)max_cells);
    if (t == 0) { hlo_sh = one_cell ? 1.0 : emax / (target + 1.0); hhi_sh = one_cell ? 1.0 : emax; }
    __syncthreads();
    // cells(h) is non-increasing in h and cells(emax) = 1 <= target
    for (int sweep = 0; sweep < 2 && !one_cell; ++sweep) {
        const double hlo = hlo_sh, hhi = hhi_sh;
        const double h = t == kGB - 1 ? hhi : hlo * pow(hhi / hlo, (double)t / (kGB - 1));
        first_ok[t] = grid_cells(e, h) <= (int64_t)target ? t : kGB;
        __syncthreads();
        for (int s = kGB / 2; s > 0; s >>= 1) {
            if (t < s) first_ok[t] = min(first_ok[t], first_ok[t + s]);
            __syncthreads();
        }
        const int k = first_ok[0];   // <= kGB - 1: the last candidate is hhi
        __syncthreads();
        if (t == k) { hhi_sh = h; }
        if (t == k - 1) { hlo_sh = h; }
        __syncthreads();
    }
    if (t == 0) {
        GridParams g;
        const double h = hhi_sh;
        int64_t nc = 1;
        double amax = 0.0;
        for (int d = 0; d < 3; ++d) {
            g.lo[d] = lo[d];
            g.dim[d] = e[d] > 0.0 ? (int)ceil(e[d] / h) : 1;
            if (g.dim[d] < 1) g.dim[d] = 1;
            nc *= g.dim[d];
            amax = fmax(amax, fmax(fabs(lo[d]), fabs(lo[d] + e[d])));
        }
        g.h = h;
        g.ncells = (int)nc;
        // slack for the termination test: cell assignment rounds at the
        // level of the coordinates' magnitude
        g.slack = 1e-12 * (amax + h);
        *gp = g;
        nc_sh = g.ncells;
    }
    __syncthreads();
    for (int k = t; k <= nc_sh; k += kGB) count[k] = 0;
}

__device__ __forceinline__ void cell_coords(const GridParams &g, double x, double y, double z, int c[3])
{
    const double p[3] = {x, y, z};
    for (int d = 0; d < 3; ++d) {
        int k = (int)floor((p[d] - g.lo[d]) / g.h);
        c[d] = k < 0 ? 0 : k >= g.dim[d] ? g.dim[d] - 1 : k;
    }
}

__global__ __launch_bounds__(kGB) void k_grid_count(const double4 *__restrict__ J, int s0, int s1,
                                                    const GridParams *__restrict__ gp, int32_t *__restrict__ count,
                                                    int32_t *__restrict__ cellof, int32_t *__restrict__ rank)
{
    const int j = s0 + blockIdx.x * kGB + threadIdx.x;
    if (j >= s1) return;
    const GridParams g = *gp;
    const double4 q = J[j];
    int c[3];
    cell_coords(g, q.x, q.y, q.z, c);
    const int cell = (c[2] * g.dim[1] + c[1]) * g.dim[0] + c[0];
    cellof[j - s0] = cell;
    rank[j - s0] = atomicAdd(&count[cell], 1);
}

__device__ __forceinline__ int wave_incl_scan(int v)
{
    const int lane = threadIdx.x & (kWave - 1);
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        const int u = __shfl_up(v, o, kWave);
        if (lane >= o) v += u;
    }
    return v;
}

// Exclusive scan of count[0..ncells) into start[0..ncells] in one block,
// tiles of 4 consecutive counts per thread.
__global__ __launch_bounds__(kScanBlock) void k_grid_scan(const GridParams *__restrict__ gp,
                                                          const int32_t *__restrict__ count,
                                                          int32_t *__restrict__ start)
{
    constexpr int nw = kScanBlock / kWave;
    __shared__ int32_t wsum[nw];
    const int nc = gp->ncells;
    const int wid = threadIdx.x / kWave, lane = threadIdx.x & (kWave - 1);
    int carry = 0;
    for (int base = 0; base < nc; base += 4 * kScanBlock) {
        const int k0 = base + 4 * threadIdx.x;
        int v[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) v[q] = k0 + q < nc ? count[k0 + q] : 0;
        const int local = v[0] + v[1] + v[2] + v[3];
        const int incl = wave_incl_scan(local);
        if (lane == kWave - 1) wsum[wid] = incl;
        __syncthreads();
        if (wid == 0) {
            int x = lane < nw ? wsum[lane] : 0;
            x = wave_incl_scan(x);
            if (lane < nw) wsum[lane] = x;
        }
        __syncthreads();
        int run = carry + (wid > 0 ? wsum[wid - 1] : 0) + incl - local;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            if (k0 + q < nc) start[k0 + q] = run;
            run += v[q];
        }
        carry += wsum[nw - 1];
        __syncthreads();
    }
    if (threadIdx.x == 0) start[nc] = carry;
}

__global__ __launch_bounds__(kGB) void k_grid_scatter(const double4 *__restrict__ J, int s0, int s1,
                                                      const int32_t *__restrict__ start,
                                                      const int32_t *__restrict__ cellof,
                                                      const int32_t *__restrict__ rank, double4 *__restrict__ spos)
{
    const int j = s0 + blockIdx.x * kGB + threadIdx.x;
    if (j >= s1) return;
    const int k = start[cellof[j - s0]] + rank[j - s0];
    const double4 q = J[j];
    spos[k] = make_double4(q.x, q.y, q.z, (double)j);
}

// (r2, j) lexicographic minimum with the reference's tie rule.
template <bool BC>
__device__ __forceinline__ bool better(double r2, int j, double best, int bidx)
{
    return BC ? (r2 < best || (r2 == best && j > bidx)) : (r2 < best || (r2 == best && (bidx < 0 || j < bidx)));
}

// Four lanes per sink, each taking every fourth source of the cells
// visited; after each shell the four minima are combined.  The sinks are
// taken in cell order (the sorted sources first, then the sinks outside
// the source range), so that the lanes of a wave search neighbouring cells.
template <bool BC>
__global__ __launch_bounds__(kGB) void k_grid_search(PairPlan plan, const double4 *__restrict__ J, int n, int s0,
                                                     int nsrc, const GridParams *__restrict__ gp,
                                                     const int32_t *__restrict__ start,
                                                     const double4 *__restrict__ spos, double *__restrict__ out_r2,
                                                     int32_t *__restrict__ out_idx)
{
    const int gtid = blockIdx.x * kGB + threadIdx.x;
    const int t = gtid >> 2, sub = gtid & 3;
    if (t >= n) return;   // whole groups of four
    int i;
    double4 p;
    if (t < nsrc) {
        p = spos[t];
        i = (int)p.w;
    } else {
        const int u = t - nsrc;
        i = u < s0 ? u : u + nsrc;
        p = J[i];
    }
    int sb = 0, se = 0;
    for (int s = 0; s < plan.nseg; ++s)
        if (i >= plan.seg[s].sink_begin && i < plan.seg[s].sink_end) {
            sb = plan.seg[s].src_begin;
            se = plan.seg[s].src_end;
            break;
        }
    double best = DBL_MAX;
    int bidx = -1;
    if (se > sb) {
        const GridParams g = *gp;
        int c[3];
        cell_coords(g, p.x, p.y, p.z, c);
        const double pc[3] = {p.x, p.y, p.z};
        auto scan = [&](int kb, int ke) {
#pragma unroll 2
            for (int k = kb + sub; k < ke; k += 4) {
                const double4 q = spos[k];
                const int j = (int)q.w;
                if (j < sb || j >= se || j == i) continue;
                const double dx = q.x - p.x, dy = q.y - p.y, dz = q.z - p.z;
                const double r2 = dx * dx + dy * dy + dz * dz;
                if (better<BC>(r2, j, best, bidx)) { best = r2; bidx = j; }
            }
        };
        for (int r = 0;; ++r) {
            const int z0 = max(c[2] - r, 0), z1 = min(c[2] + r, g.dim[2] - 1);
            const int y0 = max(c[1] - r, 0), y1 = min(c[1] + r, g.dim[1] - 1);
            const int x0 = max(c[0] - r, 0), x1 = min(c[0] + r, g.dim[0] - 1);
            for (int z = z0; z <= z1; ++z)
                for (int y = y0; y <= y1; ++y) {
                    const int row = (z * g.dim[1] + y) * g.dim[0];
                    if (abs(z - c[2]) == r || abs(y - c[1]) == r) {
                        scan(start[row + x0], start[row + x1 + 1]);   // a face: the whole row of the cube
                    } else {
                        if (c[0] - r >= 0) scan(start[row + c[0] - r], start[row + c[0] - r + 1]);
                        if (c[0] + r < g.dim[0]) scan(start[row + c[0] + r], start[row + c[0] + r + 1]);
                    }
                }
            // combine the four lanes' minima
#pragma unroll
            for (int o = 1; o < 4; o <<= 1) {
                const double ob = __shfl_xor(best, o, kWave);
                const int oi = __shfl_xor(bidx, o, kWave);
                if (oi >= 0 && better<BC>(ob, oi, best, bidx)) { best = ob; bidx = oi; }
            }
            // distance from the sink to the nearest cell outside the shells
            // visited so far (none beyond a grid face)
            double bound = DBL_MAX;
            for (int d = 0; d < 3; ++d) {
                if (c[d] - r > 0) bound = fmin(bound, pc[d] - (g.lo[d] + (c[d] - r) * g.h));
                if (c[d] + r + 1 < g.dim[d]) bound = fmin(bound, g.lo[d] + (c[d] + r + 1) * g.h - pc[d]);
            }
            if (bound == DBL_MAX) break;   // the whole grid has been visited
            const double be = bound - g.slack;
            if (bidx >= 0 && be > 0.0 && best * (1.0 + 1e-9) < be * be) break;
        }
    }
    if (sub == 0) {
        out_r2[i] = best;
        out_idx[i] = bidx;
    }
}

}  // namespace

size_t nn_grid_max_cells(int n) { return (size_t)std::max(64, n); }

void launch_nn_grid(hipStream_t st, const NNGrid &G, const PairPlan &plan, const double4 *J, int n, bool bc)
{
    // the union of the segments' source ranges is one prefix of the bodies
    int s0 = n, s1 = 0;
    for (int s = 0; s < plan.nseg; ++s) {
        if (plan.seg[s].src_end <= plan.seg[s].src_begin) continue;
        s0 = std::min(s0, plan.seg[s].src_begin);
        s1 = std::max(s1, plan.seg[s].src_end);
    }
    if (s1 < s0) s1 = s0;
    const int nsrc = s1 - s0;
    const int nb = std::max(1, std::min(kGridBboxBlocks, (nsrc + kGB - 1) / kGB));
    hipLaunchKernelGGL(k_grid_bbox, dim3(nb), dim3(kGB), 0, st, J, s0, s1, G.bbox);
    hipLaunchKernelGGL(k_grid_setup, dim3(1), dim3(kGB), 0, st, G.bbox, nb, nsrc, G.max_cells, G.params, G.count);
    if (nsrc > 0) {
        const int gb = (nsrc + kGB - 1) / kGB;
        hipLaunchKernelGGL(k_grid_count, dim3(gb), dim3(kGB), 0, st, J, s0, s1, G.params, G.count, G.cellof, G.rank);
        hipLaunchKernelGGL(k_grid_scan, dim3(1), dim3(kScanBlock), 0, st, G.params, G.count, G.start);
        hipLaunchKernelGGL(k_grid_scatter, dim3(gb), dim3(kGB), 0, st, J, s0, s1, G.start, G.cellof, G.rank, G.spos);
    } else {
        hipLaunchKernelGGL(k_grid_scan, dim3(1), dim3(kScanBlock), 0, st, G.params, G.count, G.start);
    }
    const int sb = (4 * n + kGB - 1) / kGB;
    if (bc)
        hipLaunchKernelGGL(k_grid_search<true>, dim3(sb), dim3(kGB), 0, st, plan, J, n, s0, nsrc, G.params, G.start,
                           G.spos, G.r2, G.idx);
    else
        hipLaunchKernelGGL(k_grid_search<false>, dim3(sb), dim3(kGB), 0, st, plan, J, n, s0, nsrc, G.params, G.start,
                           G.spos, G.r2, G.idx);
}

}  // namespace sol
