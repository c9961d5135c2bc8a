// walk.h -- device helpers shared by the tree walks: wave-wide reductions,
// correctly rounded small powers, and the leaf-group tree walk.
#pragma once

#include "serenx_internal.h"

namespace serenx {

extern "C" __device__ int __ockl_wfred_min_i32(int);
extern "C" __device__ int __ockl_wfred_max_i32(int);

// minimum / maximum over the active lanes of the wave (DPP reductions in the
// device library)
__device__ __forceinline__ int wave_min_int(int v) { return __ockl_wfred_min_i32(v); }
__device__ __forceinline__ int wave_max_int(int v) { return __ockl_wfred_max_i32(v); }

__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

// The dispatcher deals the workgroups of a launch to the 8 XCDs round-robin
// (workgroup i runs on XCD i % 8), and each XCD has its own L2.  This returns
// the logical block index under which XCD x receives the x-th contiguous
// eighth of the blocks, so that spatially adjacent leaf groups -- which walk
// the same cells and read the same particles -- share one L2.  A bijection of
// [0, gridDim.x) for any grid size.
constexpr int NXCD = 8;
__device__ __forceinline__ int xcd_block() {
    const int nb = gridDim.x, i = blockIdx.x;
    const int q = nb / NXCD, r = nb % NXCD, xcd = i % NXCD;
    return xcd * q + min(xcd, r) + i / NXCD;
}

// x^3 and x^4 correctly rounded (double-double products with exact FMA
// residuals, one final rounding).  The reference evaluates invh**3.0 and
// invh**4.0 with libm pow, which is correctly rounded for these arguments
// except in rare hard cases.
__device__ __forceinline__ double cr_pow3(double x) {
    const double p = x * x;
    const double e1 = __fma_rn(x, x, -p);
    const double q = p * x;
    const double e2 = __fma_rn(p, x, -q);
    const double t = __fma_rn(e1, x, e2);
    return q + t;
}

__device__ __forceinline__ double cr_pow4(double x) {
    const double p = x * x;
    const double e1 = __fma_rn(x, x, -p);
    const double q = p * p;
    const double e2 = __fma_rn(p, p, -q);
    const double t = __fma_rn(2.0 * p, e1, e2);
    return q + t;
}

// x**0.333333333 (the reference's ONETHIRD, macros.h:18) correctly rounded,
// as libm pow returns it on the host.  With y = 0.333333333 = 1/3 - delta,
//   x^y = cbrt(x) * exp(-delta ln x).
// cbrt(x) is taken to ~2^-100 relative by one Newton step whose residual
// c^3 - x is formed exactly with FMA products; |delta ln x| < 2e-8, so its
// exponential needs only three Taylor terms, and a double-precision ln x
// leaves ~2^-79 relative error in the sum.  One final rounding follows.
__device__ __forceinline__ double ref_pow_third(double x) {
    if (!(x > 0.0) || __builtin_isinf(x)) return pow(x, ONETHIRD);
    constexpr double THIRD_HI = 0.3333333333333333;       // nearest double to 1/3
    constexpr double THIRD_LO = 1.850371707708594e-17;    // 1/3 - THIRD_HI
    constexpr double DELTA = (THIRD_HI - ONETHIRD) + THIRD_LO;
    const double c0 = cbrt(x);
    const double p = c0 * c0;
    const double e1 = __fma_rn(c0, c0, -p);
    const double q = p * c0;
    const double e2 = __fma_rn(p, c0, -q);
    const double res = (q - x) + __fma_rn(e1, c0, e2);   // c0^3 - x
    const double clo = -res / (3.0 * p);
    const double d = -DELTA * log(x);
    const double g = d + d * d * (0.5 + d * (1.0 / 6.0));
    return c0 + __fma_rn(c0, g, clo);
}

// Leaf-group walk over a DFS-ordered forest.  Every lane of the wave owns one
// target and walks the tree [start, end) of its own cloud; the wave visits,
// in DFS order, the union of the cells its lanes visit on their own
// stackless walks (nextcell / ifopen = c + 1).  A lane is
// active at cell c when its own walk would be at c (resume == c); it then
// tests c and returns WALK_DESCEND (open a branch: resume = c + 1),
// WALK_LEAF (open a leaf) or WALK_SKIP (cell used or rejected), the latter two
// moving past the subtree (resume = next).  Each lane therefore sees exactly
// its own interaction set in its own order, as a per-particle walk would.
//
// visit(node, here) runs on every lane in wave-uniform control flow, so the
// node's fields come in one batch of scalar loads at the top of the visit
// rather than behind a divergent branch; it must act (accumulate) only where
// here is set, and its return value is ignored on the other lanes.
//
// Opened leaves are handled in wave-uniform control flow: leaf(node, mine) is
// called once per leaf any lane opened, with mine set on the lanes that did,
// so the leaf's particles (contiguous in tree order) can be read once through
// the scalar cache for the whole wave.
//
// Returns the number of cells the wave visited; lane_visits (wave-uniform)
// adds the number of lanes active at each visited cell (their ratio is the
// walk's SIMD utilisation).
enum { WALK_SKIP = 0, WALK_DESCEND = 1, WALK_LEAF = 2 };

template <class Node, class Visit, class Leaf>
__device__ __forceinline__ unsigned group_walk(const Node *__restrict__ nodes, int nc,
                                               bool walking, int start, int end,
                                               unsigned &lane_visits, Visit &&visit,
                                               Leaf &&leaf) {
    int resume = walking ? start : nc;
    int c = uniform(wave_min_int(resume));
    unsigned visits = 0;
    while (c < nc) {
        // read the cell in wave-uniform control flow from a uniform address so
        // that it comes through the scalar cache into SGPRs
        const Node nd = nodes[c];
        visits++;
        const bool here = resume == c;
        int act = visit(nd, here);
        lane_visits += (unsigned)__popcll(__ballot(here));
        if (here) {
            resume = (act == WALK_DESCEND) ? c + 1 : nd.next;
            if (resume >= end) resume = nc;   // past the end of the lane's own tree
        } else {
            act = WALK_SKIP;
        }
        if (__any(act == WALK_LEAF)) leaf(nd, act == WALK_LEAF);
        if (__any(act == WALK_DESCEND))
            c = c + 1;
        else
            c = uniform(wave_min_int(resume));
    }
    return visits;
}

// Adds per-wave walk statistics into stats[k0], stats[k0 + 1] from lane 0:
// the wave's visit count and its (wave-uniform) sum of lane visits.
__device__ __forceinline__ void wave_stats(unsigned long long *stats, int k0, unsigned visits,
                                           unsigned lane_visits) {
    if ((threadIdx.x & 63) == 0) {
        unsigned long long *row = ctr_row(stats);
        atomicAdd(row + k0, (unsigned long long)visits);
        atomicAdd(row + k0 + 1, (unsigned long long)lane_visits);
    }
}

// number of lanes of the wave for which pred holds (wave-uniform)
__device__ __forceinline__ unsigned lanes(bool pred) { return (unsigned)__popcll(__ballot(pred)); }

}  // namespace serenx
