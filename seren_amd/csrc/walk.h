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

// whether pred holds on any lane of the wave
__device__ __forceinline__ bool any_lane(bool pred) { return __builtin_amdgcn_ballot_w64(pred) != 0; }

// mask of the lanes of the wave for which pred holds, from the condition's
// own lane mask (HIP's __ballot takes an int, which costs a select and a
// compare per call)
__device__ __forceinline__ unsigned long long ballot(bool pred) {
    return __builtin_amdgcn_ballot_w64(pred);
}

// The dispatcher deals the workgroups of a launch to the 8 XCDs round-robin
// (workgroup i runs on XCD i % 8), and each XCD has its own L2.  This returns
// the logical block index under which XCD x receives the x-th contiguous
// eighth of the blocks, so that spatially adjacent leaf groups -- which walk
// the same cells and read the same particles -- share one L2.  A bijection of
// [0, gridDim.x) for any grid size.
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
// as libm pow returns it on the host, for 2^-1000 <= x <= 2^1000 (other x go
// to the device pow).  With y = 0.333333333 = 1/3 - delta,
//   x^y = cbrt(x) * exp(-delta ln x).
// * cube root: x = m 2^(3k) with m in [1, 8).  The hardware FP32 log2 and
//   exp2 give 2^(log2(m)/3) to ~2^-22; one Newton step c <- c - (c^3 - m) /
//   (3 c^2) in FP64 takes it to ~2^-44, and a second one on x, whose residual
//   c^3 - x is formed exactly with FMA products, to ~2^-88.
// * ln x = e ln 2 + 2 atanh(s), s = (f - 1)/(f + 1) with f in [sqrt(1/2),
//   sqrt(2)), from the odd series of atanh to s^17 (|s| < 0.172: truncation
//   below 1e-15).  |delta ln x| < 3e-7, so exp(-delta ln x) - 1 needs three
//   Taylor terms, and the error of ln x enters the result below 2^-70
//   relative.
// One final rounding follows.  It stays a call: inlined into the h-solve
// iteration kernel it raised that kernel from 84 to 123 VGPRs (3 waves per
// SIMD instead of 6 on a latency-bound walk).
static __device__ __noinline__ double ref_pow_third(double x) {
    if (!(x >= 0x1p-1000 && x <= 0x1p1000)) return pow(x, ONETHIRD);
    constexpr double THIRD_HI = 0.3333333333333333;       // nearest double to 1/3
    constexpr double THIRD_LO = 1.850371707708594e-17;    // 1/3 - THIRD_HI
    constexpr double DELTA = (THIRD_HI - ONETHIRD) + THIRD_LO;
    constexpr double LN2 = 0.6931471805599453;            // nearest double to ln 2
    int ex;
    const double fm = frexp(x, &ex);                      // x = fm 2^ex, fm in [1/2, 1)
    // cube root
    const int t = ex - 1;
    const int r = ((t % 3) + 3) % 3;
    const int k = (t - r) / 3;                            // x = (2 fm 2^r) 2^(3k)
    const double m = ldexp(2.0 * fm, r);
    double c = (double)__builtin_amdgcn_exp2f(__builtin_amdgcn_logf((float)m) * (1.0f / 3.0f));
    c = c - (c * c * c - m) / (3.0 * (c * c));
    const double c0 = ldexp(c, k);
    const double p = c0 * c0;
    const double e1 = __fma_rn(c0, c0, -p);
    const double q = p * c0;
    const double e2 = __fma_rn(p, c0, -q);
    const double res = (q - x) + __fma_rn(e1, c0, e2);   // c0^3 - x
    const double clo = -res / (3.0 * p);
    // ln x
    double f = fm;
    int e = ex;
    if (f < 0.7071067811865476) {
        f = 2.0 * f;
        e = e - 1;
    }
    const double s = (f - 1.0) / (f + 1.0);
    const double s2 = s * s;
    double poly = 1.0 / 17.0;
    poly = poly * s2 + 1.0 / 15.0;
    poly = poly * s2 + 1.0 / 13.0;
    poly = poly * s2 + 1.0 / 11.0;
    poly = poly * s2 + 1.0 / 9.0;
    poly = poly * s2 + 1.0 / 7.0;
    poly = poly * s2 + 1.0 / 5.0;
    poly = poly * s2 + 1.0 / 3.0;
    poly = poly * s2 + 1.0;
    const double lnx = (double)e * LN2 + 2.0 * s * poly;
    const double d = -DELTA * lnx;
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
        lane_visits += (unsigned)__popcll(ballot(here));
        if (here) {
            resume = (act == WALK_DESCEND) ? c + 1 : nd.next;
            if (resume >= end) resume = nc;   // past the end of the lane's own tree
        } else {
            act = WALK_SKIP;
        }
        if (ballot(act == WALK_LEAF)) leaf(nd, act == WALK_LEAF);
        // the next cell is the lowest resume pointer.  A lane descending makes
        // it c + 1.  Otherwise no lane's pointer lies inside c's subtree
        // (c, next): a lane reaches such a cell only through c itself, since
        // subtrees are contiguous and nest, so a lane that was at c and moved
        // on to c's next (not past its tree's end) holds the minimum; only
        // when none did is the wave minimum reduced.
        if (ballot(act == WALK_DESCEND))
            c = c + 1;
        else if (ballot(here) & ballot(resume == nd.next))
            c = nd.next;
        else
            c = uniform(wave_min_int(resume));
    }
    return visits;
}

// Work queue of a persistent walk.  A group is up to 64 consecutive particles
// of one cloud in tree order (one wave's targets, as in the one-wave-per-group
// launch).  The groups are split into NXCD contiguous ranges of the tree
// order, one per XCD
// (group g belongs to range 8 g / ngroups), and order[] lists each range's
// groups by decreasing cost of their previous walk, so that the most
// expensive groups start first (longest-processing-time order).  A wave of
// block b takes groups from range b % NXCD -- the blocks of a launch are
// dealt to the XCDs round-robin, so a range's groups, which are spatially
// close, share one XCD's L2 -- and steals from the following ranges when its
// own is empty.  Every call either returns a group or moves to the next
// range, so every wave leaves the loop after at most ngroups + NXCD calls.
struct GroupQueue {
    int *ctr;            // NXCD counters (one set per walk phase), zero at the launch
    const int *order;    // ngroups group ids
    int ngroups;
};

__device__ __forceinline__ int queue_lo(int q, int ngroups) { return (q * ngroups + NXCD - 1) / NXCD; }

// the next group of the calling wave (wave-uniform), -1 when none is left;
// q counts the ranges this wave has exhausted
__device__ __forceinline__ int next_group(const GroupQueue &Q, int home, int &q) {
    while (q < NXCD) {
        const int r = (home + q) % NXCD;
        const int lo = queue_lo(r, Q.ngroups), hi = queue_lo(r + 1, Q.ngroups);
        int t = 0;
        if ((threadIdx.x & 63) == 0) t = atomicAdd(Q.ctr + r, 1);
        t = uniform(__shfl(t, 0, 64));
        if (lo + t < hi) return Q.order[lo + t];
        q++;
    }
    return -1;
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
__device__ __forceinline__ unsigned lanes(bool pred) { return (unsigned)__popcll(ballot(pred)); }

}  // namespace serenx
