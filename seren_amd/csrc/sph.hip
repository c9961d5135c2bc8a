// sph.hip -- grad-h smoothing-length / density solve, neighbour lists,
// isothermal thermal update and grad-h hydro pair forces on gfx950.
//
// Each wave owns 64 consecutive particles in tree order (a compact leaf
// group) and walks the hydro tree once per pass for all of them with
// group_walk; every lane keeps its own state machine of all_sph_gradh.
// Sums run over the walk's leaves in DFS order, which is the order of the
// reference's potential-neighbour list, and use the reference's operation
// order (built without FMA contraction), so h, rho, Omega, zeta and the
// neighbour lists come out identical to the reference.
#include <cstring>
#include <rocprim/rocprim.hpp>

#include "walk.h"

namespace serenx {

namespace {

constexpr int TB = 256;
inline int nblk(long n, int tb = TB) { return (int)((n + tb - 1) / tb); }

// lane modes of the h solve
enum { M_DONE = 0, M_ITER = 1, M_FINAL = 2 };

// Compact 32-byte records for the per-lane candidate loops of the h solve,
// in the record buffer (which the hydro pass later rewrites): rec[i] =
// (x, y, z, m) and rec[n + i] = (vx, vy, vz, 0).  Four particles share a
// cache line, so a leaf's candidates are two lines, and the loops' working
// set stays in L2.
__global__ void k_pack_hsolve(int n, const double *__restrict__ x, const double *__restrict__ y,
                              const double *__restrict__ z, const double *__restrict__ vx,
                              const double *__restrict__ vy, const double *__restrict__ vz,
                              const double *__restrict__ m, double4 *__restrict__ rec) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    rec[i] = make_double4(x[i], y[i], z[i], m[i]);
    rec[(size_t)n + i] = make_double4(vx[i], vy[i], vz[i], 0.0);
}

__global__ void k_pack_xyzh(int n, const double *__restrict__ x, const double *__restrict__ y,
                            const double *__restrict__ z, const double *__restrict__ h,
                            double4 *__restrict__ rec) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) rec[i] = make_double4(x[i], y[i], z[i], h[i]);
}

// Squared distance from the point at offset (dx, dy, dz) from a node's box
// centre to the node's (grown) bounding box: a lower bound of the distance to
// any particle of the subtree.
__device__ __forceinline__ double box_dist2(const HNode &nd, double dx, double dy, double dz) {
    const double ex = fmax(fabs(dx) - (double)nd.half[0], 0.0);
    const double ey = fmax(fabs(dy) - (double)nd.half[1], 0.0);
    const double ez = fmax(fabs(dz) - (double)nd.half[2], 0.0);
    return ex * ex + ey * ey + ez * ez;
}

// HNode visit of BHhydrowalk_hgather.F90:40 for a search radius R (the
// reference's bounding-sphere test, so the leaves it yields are the
// reference's potential list, whose size the h solve reads)
__device__ __forceinline__ int hgather_visit(const HNode &nd, double rpx, double rpy, double rpz,
                                             double R) {
    const double dx = nd.r[0] - rpx, dy = nd.r[1] - rpy, dz = nd.r[2] - rpz;
    const double drsqd = dx * dx + dy * dy + dz * dz;
    const double lim = (R + nd.rmax) * (R + nd.rmax);
    if (!(drsqd <= lim)) return WALK_SKIP;
    return nd.leaf == 0 ? WALK_DESCEND : WALK_LEAF;
}

// Whether a leaf of that list can hold a particle within R: with fresh boxes
// (box) a leaf whose box lies farther than R holds none, and its particles
// are left out of the candidate lists (they cannot contribute, so the
// candidates that remain come in the reference's order)
__device__ __forceinline__ bool leaf_near(const HNode &nd, double rpx, double rpy, double rpz,
                                          double R, bool box) {
    if (!box) return true;
    return box_dist2(nd, nd.r[0] - rpx, nd.r[1] - rpy, nd.r[2] - rpz) <= R * R * (1.0 + 1.0e-12);
}

// Loops every summing lane over its own potential-neighbour list, leaf by
// leaf in DFS order, calling use(j, load(j)) for each listed particle j != p.
// Each lane walks its own stream of candidates and issues the load of the
// next candidate (and the next list slot) before using the current one, so
// the record's memory latency overlaps the current candidate's table lookup
// and arithmetic.  The order of the candidates is the list order.
template <class Load, class Use>
__device__ __forceinline__ void list_loop(const int *__restrict__ wl, int nslot, bool sum, int p,
                                          Load &&load, Use &&use) {
    int sl = 0, jcur = 0, jend = 0;
    // moves to the next listed particle; false when the list is done
    auto advance = [&]() -> bool {
        while (jcur >= jend) {
            if (sl >= nslot) return false;
            const int d = wl[sl * WAVE];
            sl++;
            jcur = d >> 4;
            jend = jcur + (d & 15);
        }
        return true;
    };
    bool have = sum && advance();
    decltype(load(0)) next{};
    if (have) next = load(jcur);
    while (any_lane(have)) {
        const auto cur = next;
        const int jc = jcur;
        const bool hc = have;
        if (have) {
            jcur++;
            have = advance();
            if (have) next = load(jcur);
        }
        if (hc && jc != p) use(jc, cur);
    }
}

// all_sph_gradh.F90:103-309, the h iteration, for every active particle.
//
// Potential-neighbour lists (BHhydrowalk_hgather) are recorded once per
// change of search radius: a group walk stores, for every lane, the leaves
// its own walk opens, in DFS order, as (first << 4) | count in HL_SLOTS
// slots per lane (slot-major, so the lanes of a wave read one contiguous row
// per step).  The iterations loop over each lane's own list, which is the
// reference's pp_potlist in the reference's order, so every sum is formed
// exactly as the reference forms it.  A wave with a lane whose list does not
// fit sums over the tree walk itself for that radius (same leaves, same
// order).  On exit every lane leaves rhotemp in rho, h_high in hhigh and its
// list length in nlist (-1: list incomplete) for k_hsolve_final; its list
// stays in hlist (a lane joins no walk after it has converged).
//
// WAVES > 0 holds the kernel to WAVES waves per SIMD (64 VGPRs at 8, with a
// few spills outside the candidate loops): more waves in flight hide more of
// the walks' memory latency, at the cost of issue slots taken from kernels
// running beside it.  0 leaves the register budget to the compiler
// (92 VGPRs, 5 waves).
template <int WAVES>
__global__ void __launch_bounds__(256)
__attribute__((amdgpu_waves_per_eu(WAVES > 0 ? WAVES : 1, WAVES > 0 ? WAVES : 8))) k_hsolve_iter(
    KGeom kg, bool box, int n, int nc, const HNode *__restrict__ hn, const double4 *__restrict__ rec,
    const int *__restrict__ cid, const int2 *__restrict__ root_range,
    const double *__restrict__ roots, int *__restrict__ hlist, int hl_cap,
    const uint8_t *__restrict__ active, const double *__restrict__ ktab,
    double h_fac, int pp_gather, double *__restrict__ h,
    double *__restrict__ rho, double *__restrict__ hhigh, int *__restrict__ nlist,
    unsigned long long *__restrict__ wstat) {
    const int p = xcd_block() * blockDim.x + threadIdx.x;
    const bool valid = p < n && active[p];
    int *wl = hlist + (size_t)(p / WAVE) * (HL_SLOTS * WAVE) + (threadIdx.x & (WAVE - 1));
    const double *w0 = ktab + T_W0 * KT;
    const double *wh = ktab + T_WH * KT;
    const double *h4 = ktab + KT_H4_OFF;   // (w0, wh, w4, wg) per index
    double rpx = 0, rpy = 0, rpz = 0, mp = 0, hp = 1, rextent = 0.0;
    int2 rr = make_int2(nc, nc);
    if (valid) {
        const double4 a = rec[p];
        rpx = a.x; rpy = a.y; rpz = a.z; mp = a.w;
        hp = h[p];
        const int k = cid[p];
        rr = root_range[k];
        rextent = roots[11 * k + 4];   // the particle's own cloud (bounding_box.F90:39)
    }
    double h_high = 0.0, h_old = hp, h_lower = 0.0, h_upper = rextent;
    double rhotemp = 0.0;
    int iteration = 0;
    bool iterating = valid;
    unsigned passes = 0, vis = 0, lvis = 0, walks = 0, fallback = 0;
    int nslot = 0, cnt = 0, slot_max = 0;
    bool have_list = false, wave_ovf = false, my_ovf = false;

    while (any_lane(iterating)) {
        passes++;
        bool newlist = false;
        if (iterating && hp > h_high) {
            h_high = hp * HMULT;
            newlist = true;
        }
        const double R = kg.range * h_high;
        const double invhp = 1.0 / hp;
        const double hfactor_p = cr_pow3(invhp);
        const double hrangesqd = kg.rangesqd * hp * hp;
        double srho = mp * w0[0] * hfactor_p;                  // :209-210
        double som = mp * hfactor_p * invhp * wh[0];
        auto visit = [&](const HNode &nd, bool) { return hgather_visit(nd, rpx, rpy, rpz, R); };
        // :214-231 for one potential neighbour
        auto pair = [&](double xj, double yj, double zj, double mj) {
            const double ex = xj - rpx, ey = yj - rpy, ez = zj - rpz;
            const double dsq = ex * ex + ey * ey + ez * ez;
            if (dsq >= hrangesqd) return;                       // :223
            const double drmag = sqrt(dsq) + SMALL_NUMBER;
            const int kern = (int)(kg.half * drmag * invhp);
            if (kern >= KERNTOT) return;                        // :227
            const double2 w = *reinterpret_cast<const double2 *>(h4 + 4 * (size_t)kern);
            srho = srho + mj * hfactor_p * w.x;
            som = som + mj * hfactor_p * invhp * w.y;
        };

        if (!have_list || any_lane(newlist)) {
            // record every iterating lane's list (unchanged radii give the same
            // lists); converged lanes keep theirs
            if (iterating) {
                nslot = 0;
                cnt = 0;
            }
            bool lovf = false;
            vis += group_walk(hn, nc, iterating, rr.x, rr.y, lvis, visit,
                              [&](const HNode &nd, bool mine) {
                if (mine) {
                    cnt += nd.leaf;                       // the reference's pp_pot
                    if (leaf_near(nd, rpx, rpy, rpz, R, box)) {
                        if (nslot < hl_cap)
                            wl[nslot * WAVE] = (nd.first << 4) | nd.leaf;
                        else
                            lovf = true;
                        nslot++;
                    }
                }
            });
            if (iterating) {
                slot_max = max(slot_max, nslot);
                my_ovf = lovf;
            }
            have_list = true;
            wave_ovf = any_lane(lovf);
            walks++;
        }
        // too few potential neighbours: grow h and gather again (:185-189)
        const bool grow = iterating && newlist && cnt < pp_gather;
        const bool sum = iterating && !grow;
        if (!wave_ovf) {
            list_loop(wl, nslot, sum, p, [&](int j) { return rec[j]; },
                      [&](int, const double4 &a) { pair(a.x, a.y, a.z, a.w); });
        } else {
            vis += group_walk(hn, nc, sum, rr.x, rr.y, lvis, visit,
                              [&](const HNode &nd, bool mine) {
                const bool use = mine && leaf_near(nd, rpx, rpy, rpz, R, box);
                for (int i = 0; i < nd.leaf; i++) {
                    const int j = nd.first + i;
                    const double4 a = rec[j];
                    if (use && j != p) pair(a.x, a.y, a.z, a.w);
                }
            });
            fallback++;
        }

        if (grow) {
            h_high = 0.0;
            hp = hp * HMULT;
        } else if (iterating) {
            rhotemp = srho;
            const double hnew = h_fac * ref_pow_third(mp / rhotemp);
            if (fabs(hp - hnew) / hp < H_CON && iteration >= 1) {  // :246
                iterating = false;
            } else if (iteration < GRADH_ITERATION_MAX) {           // :252
                hp = hnew;
                iteration++;
            } else if (iteration == GRADH_ITERATION_MAX) {          // :256
                h_upper = rextent;
                h_lower = 0.0;
                hp = 0.5 * (h_lower + h_upper);
                iteration++;
            } else if (iteration < 2 * GRADH_ITERATION_MAX) {       // :262
                double fb;
                if (rhotemp > SMALL_NUMBER)
                    fb = cr_pow3(hp) - cr_pow3(h_fac) * mp / rhotemp;
                else
                    fb = 1.0;
                if (fb > 0.0)
                    h_upper = hp;
                else
                    h_lower = hp;
                hp = 0.5 * (h_lower + h_upper);
                iteration++;
            } else {                                                // :276-279
                iterating = false;
            }
        }
    }
    (void)h_old;   // the final h is formed from rhotemp alone (:317)
    if (valid) {
        rho[p] = rhotemp;
        hhigh[p] = h_high;
        nlist[p] = my_ovf ? -1 : nslot;
    }
    // walk statistics: cells visited / lane visits, passes, lane iterations,
    // list walks, fallback passes, longest list
    wave_stats(wstat, 2, vis, lvis);
    unsigned long long it = valid ? (unsigned long long)iteration : 0ull;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) it += __shfl_xor(it, o, 64);
    const int smx = wave_max_int(slot_max);
    if ((threadIdx.x & 63) == 0) {
        unsigned long long *ws = ctr_row(wstat);
        atomicAdd(ws + 4, (unsigned long long)passes);
        atomicMax(ws + 5, (unsigned long long)passes);
        atomicAdd(ws + 6, it);
        atomicAdd(ws + 7, 1ull);
        atomicAdd(ws + 10, (unsigned long long)walks);
        atomicAdd(ws + 11, (unsigned long long)fallback);
        atomicMax(ws + 12, (unsigned long long)smx);
    }
}

// all_sph_gradh.F90:310-489: with h from the converged density, the gather
// sums rho, Omega, div v and zeta over the last potential-neighbour list.
__global__ void __launch_bounds__(256) k_hsolve_final(
    KGeom kg, bool box, int n, int nc, const HNode *__restrict__ hn, const double4 *__restrict__ rec,
    const int *__restrict__ cid, const int2 *__restrict__ root_range,
    const int *__restrict__ hlist, const uint8_t *__restrict__ active,
    const double *__restrict__ ktab, double h_fac, const double *__restrict__ hhigh,
    const int *__restrict__ nlist, double *__restrict__ h, double *__restrict__ zeta,
    double *__restrict__ rho, double *__restrict__ omega, double *__restrict__ divv,
    double *__restrict__ drhodt, int *__restrict__ err, unsigned long long *__restrict__ wstat) {
    const int p = xcd_block() * blockDim.x + threadIdx.x;
    const bool valid = p < n && active[p];
    const int *wl = hlist + (size_t)(p / WAVE) * (HL_SLOTS * WAVE) + (threadIdx.x & (WAVE - 1));
    const double *w0 = ktab + T_W0 * KT;
    const double *wh = ktab + T_WH * KT;
    const double *wg = ktab + T_WG * KT;
    const double *h4 = ktab + KT_H4_OFF;   // (w0, wh, w4, wg) per index
    const double4 *vrec = rec + n;
    double rpx = 0, rpy = 0, rpz = 0, mp = 0, vpx = 0, vpy = 0, vpz = 0, R = 0, hp = 1;
    int nslot = 0;
    if (valid) {
        const double4 a = rec[p], b = vrec[p];
        rpx = a.x; rpy = a.y; rpz = a.z; mp = a.w;
        vpx = b.x; vpy = b.y; vpz = b.z;
        hp = h_fac * ref_pow_third(mp / rho[p]);                 // :317
        R = kg.range * hhigh[p];
        nslot = nlist[p];
        // a final h beyond the list's radius can reach particles past R in
        // leaves the box test left out: gather over the walk instead
        if (box && kg.range * hp > R) nslot = -1;
    }
    const double invhp = 1.0 / hp;
    const double hfactor_p = cr_pow3(invhp);
    const double hrangesqd = kg.rangesqd * hp * hp;
    // self contributions (:323-346)
    double srho = mp * w0[0] * hfactor_p;
    double som = mp * hfactor_p * invhp * wh[0];
    double sdiv = 0.0;
    double szeta = mp * invhp * invhp * wg[0];
    // :356-381 for one potential neighbour
    auto pair = [&](double xj, double yj, double zj, double mj, int j) {
        const double ex = xj - rpx, ey = yj - rpy, ez = zj - rpz;
        const double dsq = ex * ex + ey * ey + ez * ez;
        if (dsq > hrangesqd) return;                            // :364
        const double drmag = sqrt(dsq) + SMALL_NUMBER;
        const int kern = (int)(kg.half * drmag * invhp);
        if (kern > KERNTOT) return;                             // :368
        const double4 vj = vrec[j];
        const double dvx = vj.x - vpx, dvy = vj.y - vpy, dvz = vj.z - vpz;
        const double dvdr = dvx * ex + dvy * ey + dvz * ez;
        const double4 w = *reinterpret_cast<const double4 *>(h4 + 4 * (size_t)kern);
        srho = srho + mj * hfactor_p * w.x;
        som = som + mj * hfactor_p * invhp * w.y;
        sdiv = sdiv - mj * dvdr * w.z * hfactor_p * invhp / drmag;
        szeta = szeta + mj * invhp * invhp * w.w;
    };
    list_loop(wl, nslot, valid && nslot >= 0, p, [&](int j) { return rec[j]; },
              [&](int j, const double4 &a) { pair(a.x, a.y, a.z, a.w, j); });
    unsigned vis = 0, lvis = 0;
    const bool walk = valid && nslot < 0;
    if (any_lane(walk)) {
        // lists that did not fit: gather over the tree walk at the same radius
        const int2 rr = valid ? root_range[cid[p]] : make_int2(nc, nc);
        vis = group_walk(
            hn, nc, walk, rr.x, rr.y, lvis,
            [&](const HNode &nd, bool) { return hgather_visit(nd, rpx, rpy, rpz, R); },
            [&](const HNode &nd, bool mine) {
                const bool use = mine && leaf_near(nd, rpx, rpy, rpz, fmax(R, kg.range * hp), box);
                for (int i = 0; i < nd.leaf; i++) {
                    const int j = nd.first + i;
                    const double4 a = rec[j];
                    if (use && j != p) pair(a.x, a.y, a.z, a.w, j);
                }
            });
        wave_stats(wstat, 2, vis, lvis);
        if ((threadIdx.x & 63) == 0) atomicAdd(ctr_row(wstat) + 11, 1ull);
    }
    if (valid) {
        const double om = 1.0 + ((hp * som) / (3.0 * srho));
        const double dv = sdiv / srho;
        rho[p] = srho;
        omega[p] = om;
        divv[p] = dv;
        drhodt[p] = -srho * dv / om;
        h[p] = hp;
        zeta[p] = -(hp * szeta) / (3.0 * srho * om);
        if (hp < 0.0) atomicOr(err, 1);
    }
}

constexpr int OVF_PER_WAVE = 1;

// get_neib.F90:46-118: gather-or-scatter neighbours in walk (DFS) order.
// Rows of pp_limit ids; particles with more neighbours only get their count
// here (the reference does not store their list either, :115-118) and are
// listed completely by the OVERFLOW_PASS instance.
template <bool OVERFLOW_PASS>
__global__ void __launch_bounds__(256) k_neib(KGeom kg, bool box, int n, const int *__restrict__ ovf_idx,
                                              int nc,
                                              const HNode *__restrict__ hn,
                                              const double4 *__restrict__ rec,
                                              const int *__restrict__ cid,
                                              const int2 *__restrict__ root_range,
                                              const uint8_t *__restrict__ active, int pp_limit,
                                              int *__restrict__ pptot, int *__restrict__ pplist,
                                              const int *__restrict__ ovf_off,
                                              int *__restrict__ ovf_list,
                                              int *__restrict__ ovf_count,
                                              unsigned long long *__restrict__ wstat,
                                              int *__restrict__, int *__restrict__, int) {
    // the overflow pass runs over the list of overflow particles (n of them),
    // OVF_PER_WAVE to a wave: their walks are long, and a wave walks the union
    // of its lanes' walks, so a few lanes per wave keep many waves in flight
    const int t = xcd_block() * blockDim.x + threadIdx.x;
    const int ot = (t / WAVE) * OVF_PER_WAVE + (t % WAVE);
    const bool ovf_lane = (t % WAVE) < OVF_PER_WAVE && ot < n;
    const int p = OVERFLOW_PASS ? (ovf_lane ? ovf_idx[ot] : 0) : t;
    const bool valid = OVERFLOW_PASS ? ovf_lane : (p < n && active[p]);
    double rpx = 0, rpy = 0, rpz = 0, hp = 0;
    if (valid) {
        const double4 a = rec[p];
        rpx = a.x; rpy = a.y; rpz = a.z;
        hp = a.w;
    }
    const double hrangesqd_p = kg.rangesqd * hp * hp;
    const double R = kg.range * hp;
    int tot = 0;
    int *row = nullptr;
    int lim = pp_limit;   // entries this lane may store
    if (valid) {
        row = OVERFLOW_PASS ? ovf_list + ovf_off[p] : pplist + (size_t)p * pp_limit;
        if (OVERFLOW_PASS) lim = pptot[p];
    }
    unsigned lvis = 0;
    const int2 rr = valid ? root_range[cid[p]] : make_int2(nc, nc);
    const unsigned vis = group_walk(
        hn, nc, valid, rr.x, rr.y, lvis,
        [&](const HNode &nd, bool) -> int {
            const double dx = nd.r[0] - rpx, dy = nd.r[1] - rpy, dz = nd.r[2] - rpz;
            const double drsqd = dx * dx + dy * dy + dz * dz;
            const double m1 = (R + nd.rmax) * (R + nd.rmax);
            const double m2 = (nd.hrangemax + nd.rmax) * (nd.hrangemax + nd.rmax);
            if (!(drsqd <= m1 || drsqd <= m2)) return WALK_SKIP;   // BHhydro_walk.F90:43
            // and the box test at the larger of the two radii (hgather_visit)
            const double Rb = fmax(R, nd.hrangemax);
            if (box && box_dist2(nd, dx, dy, dz) > Rb * Rb * (1.0 + 1.0e-12)) return WALK_SKIP;
            return nd.leaf == 0 ? WALK_DESCEND : WALK_LEAF;
        },
        [&](const HNode &nd, bool mine) {
            const int k = nd.leaf, f0 = nd.first;
            for (int i = 0; i < k; i++) {
                const int j = f0 + i;
                const double4 aj = rec[j];   // one scalar load per candidate
                const double xj = aj.x, yj = aj.y, zj = aj.z, hj = aj.w;
                if (!mine || j == p) continue;
                const double hrangesqd_pp = kg.rangesqd * hj * hj;
                const double ex = xj - rpx, ey = yj - rpy, ez = zj - rpz;
                const double dsq = ex * ex + ey * ey + ez * ez;
                if (dsq <= hrangesqd_p || dsq <= hrangesqd_pp) {
                    if (tot < lim) row[tot] = j;
                    tot++;
                }
            }
        });
    if (valid && !OVERFLOW_PASS) {
        pptot[p] = tot;
        if (tot > pp_limit) atomicAdd(ovf_count, 1);
    }
    if (!OVERFLOW_PASS) wave_stats(wstat, 8, vis, lvis);
}

// The cell test of the group neighbour walk for one lane (own: the cell is in
// the lane's cloud).  With fresh boxes, the box test at the larger of the two
// radii alone: every neighbour j of p in the cell's subtree lies in the box
// and within max(KERNRANGE h_p, hrangemax) of p, so the cells it passes are
// a superset of those holding neighbours.  With stale boxes,
// BHhydro_walk.F90:36-43 for this lane.
__device__ __forceinline__ bool neib_pass(const HNode &nd, bool box, bool own, double R,
                                          double rpx, double rpy, double rpz) {
    const double dx = nd.r[0] - rpx, dy = nd.r[1] - rpy, dz = nd.r[2] - rpz;
    if (box) {
        const double Rb = fmax(R, nd.hrangemax);
        return own && box_dist2(nd, dx, dy, dz) <= Rb * Rb * (1.0 + 1.0e-12);
    }
    const double drsqd = dx * dx + dy * dy + dz * dz;
    const double m1 = (R + nd.rmax) * (R + nd.rmax);
    const double m2 = (nd.hrangemax + nd.rmax) * (nd.hrangemax + nd.rmax);
    return own && (drsqd <= m1 || drsqd <= m2);
}

// The complete row of overflow particle p from the leaves its first-pass
// wave (particles 64 g .. 64 g + 63) opened, in DFS order: lane l takes
// leaf e0 + l, applies p's own cell test and the pair test to its
// particles, and the lanes' matches are placed by a prefix sum of their
// counts, so the row holds the candidates in leaf order and, within a leaf,
// in particle order -- the order of p's own walk.  The recorded leaves
// include every leaf p's own test passes (the test of a leaf implies the
// test of each ancestor, so the wave reached it through p's own path), and
// only those pass here: the row is the one the walk would write.
__device__ void neib_ovf_scan(KGeom kg, bool box, int p, int g, int nl,
                              const int *__restrict__ nbl, const HNode *__restrict__ hn,
                              const double4 *__restrict__ rec, const int *__restrict__ cid,
                              const int2 *__restrict__ root_range, const int *__restrict__ pptot,
                              const int *__restrict__ ovf_off, int *__restrict__ ovf_list) {
    const int lane = (int)(threadIdx.x & (WAVE - 1));
    const double4 a = rec[p];
    const double rpx = a.x, rpy = a.y, rpz = a.z, hp = a.w;
    const double hrangesqd_p = kg.rangesqd * hp * hp;
    const double R = kg.range * hp;
    const int2 rr = root_range[cid[p]];
    int *row = ovf_list + ovf_off[p];
    const int lim = pptot[p];
    int base = 0;
    for (int e0 = 0; e0 < nl; e0 += WAVE) {
        const int e = e0 + lane;
        unsigned hit = 0;   // bit i: particle i of the lane's leaf is a neighbour
        int first = 0, cnt = 0;
        if (e < nl) {
            const int c = nbl[(size_t)g * NB_CAP + e];
            const HNode nd = hn[c];
            if (nd.leaf > 0 && neib_pass(nd, box, c >= rr.x && c < rr.y, R, rpx, rpy, rpz)) {
                first = nd.first;
                for (int i = 0; i < nd.leaf; i++) {
                    const int j = first + i;
                    if (j == p) continue;
                    const double4 aj = rec[j];
                    const double hrangesqd_pp = kg.rangesqd * aj.w * aj.w;
                    const double qx = aj.x - rpx, qy = aj.y - rpy, qz = aj.z - rpz;
                    const double dsq = qx * qx + qy * qy + qz * qz;
                    if (dsq <= hrangesqd_p || dsq <= hrangesqd_pp) {
                        hit |= 1u << i;
                        cnt++;
                    }
                }
            }
        }
        int incl = cnt;   // inclusive prefix sum over the lanes
        for (int d = 1; d < WAVE; d <<= 1) {
            const int u = __shfl_up(incl, d, WAVE);
            if (lane >= d) incl += u;
        }
        int o = base + incl - cnt;
        for (int i = 0; hit != 0; i++, hit >>= 1)
            if (hit & 1u) {
                if (o < lim) row[o] = first + i;
                o++;
            }
        base += __shfl(incl, WAVE - 1, WAVE);
    }
}

// get_neib.F90:46-118 with one walk per wave instead of one per particle.
// Every particle j with r^2 <= KERNRANGESQD h_p^2 or r^2 <= KERNRANGESQD
// h_j^2 lies in a leaf that the reference's BHhydro_walk opens for p: every
// ancestor's bounding sphere holds j, so |r_p - r_c| <= |r_p - r_j| + rmax_c,
// and hrangemax_c >= KERNRANGE h_j.  The reference keeps exactly those
// particles, in DFS order.  Any walk that opens a superset of those leaves,
// in DFS order, therefore yields the reference's lists once each candidate
// passes the reference's pair test (formed as get_neib.F90:89-95 forms it).
// This walk visits a cell's children when any lane's own test of
// BHhydro_walk.F90:36-43 (and, with fresh boxes, the box test) passes the
// cell, so every cell on any lane's own path is visited; a lane takes the
// candidates of the leaves its own test passes.  Unlike group_walk it keeps
// no per-lane resume state: the visit is one test per lane and a ballot.
template <bool OVERFLOW_PASS>
__global__ void __launch_bounds__(256) k_neib_group(KGeom kg, bool box, int n,
                                                    const int *__restrict__ ovf_idx, int nc,
                                                    const HNode *__restrict__ hn,
                                                    const double4 *__restrict__ rec,
                                                    const int *__restrict__ cid,
                                                    const int2 *__restrict__ root_range,
                                                    const uint8_t *__restrict__ active,
                                                    int pp_limit, int *__restrict__ pptot,
                                                    int *__restrict__ pplist,
                                                    const int *__restrict__ ovf_off,
                                                    int *__restrict__ ovf_list,
                                                    int *__restrict__ ovf_count,
                                                    unsigned long long *__restrict__ wstat,
                                                    int *__restrict__ nbl,
                                                    int *__restrict__ nbl_len, int nb_cap) {
    const int t = xcd_block() * blockDim.x + threadIdx.x;
    const int wv = t / WAVE;
    if (OVERFLOW_PASS) {
        // one overflow particle per wave: from the leaf list of its
        // first-pass wave when that list is complete, else by its own walk
        if (wv >= n) return;
        const int po = uniform(ovf_idx[wv]);
        const int g = po / WAVE;
        const int nl = uniform(nbl_len[g]);
        if (nl <= nb_cap) {
            neib_ovf_scan(kg, box, po, g, nl, nbl, hn, rec, cid, root_range, pptot, ovf_off,
                          ovf_list);
            return;
        }
    }
    const int ot = (t / WAVE) * OVF_PER_WAVE + (t % WAVE);
    const bool ovf_lane = (t % WAVE) < OVF_PER_WAVE && ot < n;
    const int p = OVERFLOW_PASS ? (ovf_lane ? ovf_idx[ot] : 0) : t;
    const bool valid = OVERFLOW_PASS ? ovf_lane : (p < n && active[p]);
    double rpx = 0, rpy = 0, rpz = 0, hp = 0;
    if (valid) {
        const double4 a = rec[p];
        rpx = a.x; rpy = a.y; rpz = a.z;
        hp = a.w;
    }
    if (ballot(valid) == 0) return;
    const double hrangesqd_p = kg.rangesqd * hp * hp;
    const double R = kg.range * hp;
    int tot = 0;
    int *row = nullptr;
    int lim = pp_limit;   // entries this lane may store
    if (valid) {
        row = OVERFLOW_PASS ? ovf_list + ovf_off[p] : pplist + (size_t)p * pp_limit;
        if (OVERFLOW_PASS) lim = pptot[p];
    }
    const int2 rr = valid ? root_range[cid[p]] : make_int2(nc, nc);
    int c = uniform(wave_min_int(rr.x));
    const int cend = uniform(wave_max_int(valid ? rr.y : 0));
    unsigned vis = 0, lvis = 0;
    int nl = 0;   // leaves opened (wave-uniform), recorded for the overflow pass
    while (c < cend) {
        const HNode nd = hn[c];
        vis++;
        const bool pass = neib_pass(nd, box, valid && c >= rr.x && c < rr.y, R, rpx, rpy, rpz);
        lvis += lanes(pass);
        if (ballot(pass) == 0) {
            c = nd.next;
            continue;
        }
        if (nd.leaf == 0) {
            c = c + 1;
            continue;
        }
        if (!OVERFLOW_PASS) {
            if ((threadIdx.x & (WAVE - 1)) == 0 && nl < nb_cap) nbl[(size_t)wv * NB_CAP + nl] = c;
            nl++;
        }
        const int k = nd.leaf, f0 = nd.first;
        for (int i = 0; i < k; i++) {
            const int j = f0 + i;
            const double4 aj = rec[j];   // one scalar load per candidate
            if (!pass || j == p) continue;
            const double hrangesqd_pp = kg.rangesqd * aj.w * aj.w;
            const double qx = aj.x - rpx, qy = aj.y - rpy, qz = aj.z - rpz;
            const double dsq = qx * qx + qy * qy + qz * qz;
            if (dsq <= hrangesqd_p || dsq <= hrangesqd_pp) {
                if (tot < lim) row[tot] = j;
                tot++;
            }
        }
        c = nd.next;
    }
    if (valid && !OVERFLOW_PASS) {
        pptot[p] = tot;
        if (tot > pp_limit) atomicAdd(ovf_count, 1);
    }
    if (!OVERFLOW_PASS) {
        if ((threadIdx.x & (WAVE - 1)) == 0) nbl_len[wv] = nl;
        wave_stats(wstat, 8, vis, lvis);
    }
}

__global__ void k_ovf_sizes(int n, int pp_limit, const int *__restrict__ pptot,
                            const uint8_t *__restrict__ active, int *__restrict__ sz,
                            int *__restrict__ is_ovf) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const bool o = active[i] && pptot[i] > pp_limit;
    sz[i] = o ? pptot[i] : 0;
    is_ovf[i] = o ? 1 : 0;
}

// overflow particles in tree order: ovf_idx[pos[i]] = i
__global__ void k_ovf_compact(int n, const int *__restrict__ is_ovf, const int *__restrict__ pos,
                              int *__restrict__ ovf_idx) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && is_ovf[i]) ovf_idx[pos[i]] = i;
}

// thermal.F90:52-60 with the sound speed given: P = cs^2 rho, c = cs
__global__ void k_thermal_iso(int n, double cs, const uint8_t *__restrict__ active,
                              const double *__restrict__ rho, double *__restrict__ press,
                              double *__restrict__ sound) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !active[i]) return;
    press[i] = cs * cs * rho[i];
    sound[i] = cs;
}

// Per-particle record of everything a hydro pair reads from the neighbour,
// 128 bytes (one cache line) per particle, so that each neighbour costs one
// line instead of one line per field.  The per-particle factors are formed
// with the reference's operations (hydro_gradh.F90:90-96, :178, :195-196),
// so the pair terms use the same values bit for bit:
//   [0] x y z h   [1] vx vy vz m   [2] rho pfactor sound invh   [3] hfactor
__global__ void k_pack_hydro(int n, const double *__restrict__ x, const double *__restrict__ y,
                             const double *__restrict__ z, const double *__restrict__ h,
                             const double *__restrict__ vx, const double *__restrict__ vy,
                             const double *__restrict__ vz, const double *__restrict__ m,
                             const double *__restrict__ rho, const double *__restrict__ omega,
                             const double *__restrict__ press, const double *__restrict__ sound,
                             double4 *__restrict__ rec) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double hi = h[i], rh = rho[i];
    const double invh = 1.0 / hi;
    const double pf = press[i] / rh / rh / omega[i];
    double4 *r = rec + 4 * (size_t)i;
    r[0] = make_double4(x[i], y[i], z[i], hi);
    r[1] = make_double4(vx[i], vy[i], vz[i], m[i]);
    r[2] = make_double4(rh, pf, sound[i], invh);
    r[3] = make_double4(cr_pow4(invh), 0.0, 0.0, 0.0);
}

// hydro_gradh.F90:87-297 (AB viscosity, beta := 2 alpha).  ENERGY: the
// INTERNAL_ENERGY terms, du/dt = sum 0.5 m_j visc W_mean (dv.dr) / |dr|
// (:258-259) - P div v / (Omega rho) (:326-332)
template <bool ENERGY>
__global__ void __launch_bounds__(256) k_hydro(
    KGeom kg, int n, const double4 *__restrict__ rec, const uint8_t *__restrict__ active,
    const int *__restrict__ pptot, const int *__restrict__ pplist, const int *__restrict__ ovf_off,
    const int *__restrict__ ovf_list, int pp_limit, const double *__restrict__ ktab, double alpha,
    double *__restrict__ ax, double *__restrict__ ay, double *__restrict__ az,
    const double *__restrict__ press, const double *__restrict__ divv,
    const double *__restrict__ omega, double *__restrict__ dudt) {
    const int p = xcd_block() * blockDim.x + threadIdx.x;
    if (p >= n || !active[p]) return;
    const double *w1 = ktab + T_W1 * KT;
    const double4 *rp = rec + 4 * (size_t)p;
    const double4 r0 = rp[0], r1 = rp[1], r2 = rp[2], r3 = rp[3];
    const double rpx = r0.x, rpy = r0.y, rpz = r0.z, hp = r0.w;
    const double vpx = r1.x, vpy = r1.y, vpz = r1.z;
    const double rho_p = r2.x, pfactor_p = r2.y, sound_p = r2.z, invhp = r2.w;
    const double hfactor_p = r3.x;
    const double beta = 2.0 * alpha;
    double ahx = 0.0, ahy = 0.0, ahz = 0.0, adx = 0.0, ady = 0.0, adz = 0.0;
    double dudt_diss = 0.0;
    const int nn = pptot[p];
    const int *lst = (nn <= pp_limit) ? pplist + (size_t)p * pp_limit : ovf_list + ovf_off[p];
    for (int i = 0; i < nn; i++) {
        const double4 *rq = rec + 4 * (size_t)lst[i];
        const double4 q0 = rq[0], q1 = rq[1], q2 = rq[2];
        double wmean = 0.0;
        const double hpp = q0.w, mpp = q1.w, rho_pp = q2.x;
        double ux = q0.x - rpx, uy = q0.y - rpy, uz = q0.z - rpz;
        double drmag = ux * ux + uy * uy + uz * uz;
        if (drmag >= kg.rangesqd * hp * hp && drmag >= kg.rangesqd * hpp * hpp) continue;
        drmag = sqrt(drmag) + SMALL_NUMBER;
        const double dvx = q1.x - vpx, dvy = q1.y - vpy, dvz = q1.z - vpz;
        const double dvdr = dvx * ux + dvy * uy + dvz * uz;
        const double invdrmag = 1.0 / drmag;
        ux = ux * invdrmag;
        uy = uy * invdrmag;
        uz = uz * invdrmag;
        if (drmag < kg.range * hp) {
            int kern = (int)(kg.half * drmag * invhp);
            kern = min(kern, KERNTOT);
            const double wk = w1[kern];
            wmean = 0.5 * hfactor_p * wk;
            const double f = mpp * pfactor_p * hfactor_p * wk;
            ahx = ahx + f * ux;
            ahy = ahy + f * uy;
            ahz = ahz + f * uz;
        }
        if (drmag < kg.range * hpp) {
            const double invhpp = q2.w, hfactor_pp = rq[3].x, pfactor_pp = q2.y;
            int kern = (int)(kg.half * drmag * invhpp);
            kern = min(kern, KERNTOT);
            const double wk = w1[kern];
            wmean = wmean + 0.5 * hfactor_pp * wk;
            const double f = mpp * pfactor_pp * hfactor_pp * wk;
            ahx = ahx + f * ux;
            ahy = ahy + f * uy;
            ahz = ahz + f * uz;
        }
        const double rhomean = 0.5 * (rho_p + rho_pp);
        if (dvdr < 0.0) {
            const double vsignal = 0.5 * (sound_p + q2.z);
            const double mu = 0.5 * (hp + hpp) * dvdr /
                              (drmag * drmag + 0.25 * ETA_SQD * (hp + hpp) * (hp + hpp));
            const double visc = (-alpha * vsignal * mu + beta * mu * mu) / rhomean;
            const double f = mpp * visc * wmean;
            adx = adx + f * ux;
            ady = ady + f * uy;
            adz = adz + f * uz;
            if (ENERGY) dudt_diss = dudt_diss + 0.5 * mpp * visc * wmean * dvdr * invdrmag;
        }
    }
    ax[p] = ahx + adx;
    ay[p] = ahy + ady;
    az[p] = ahz + adz;
    if (ENERGY) dudt[p] = (0.0 + dudt_diss) - press[p] * divv[p] / omega[p] / rho_p;
}

// Mixed-precision hydro record, 64 bytes per particle: the position in FP64
// (the offsets are formed in FP64, then rounded), everything else in FP32,
// with the per-particle factors of k_pack_hydro formed in FP64 first.
struct __attribute__((aligned(64))) HydroRecF {
    double x, y, z;
    float h, vx, vy, vz, m, rho, pf, sound, invh, hfac;
};
static_assert(sizeof(HydroRecF) == 64, "HydroRecF layout");

__global__ void k_pack_hydro_mixed(int n, const double *__restrict__ x,
                                   const double *__restrict__ y, const double *__restrict__ z,
                                   const double *__restrict__ h, const double *__restrict__ vx,
                                   const double *__restrict__ vy, const double *__restrict__ vz,
                                   const double *__restrict__ m, const double *__restrict__ rho,
                                   const double *__restrict__ omega,
                                   const double *__restrict__ press,
                                   const double *__restrict__ sound,
                                   HydroRecF *__restrict__ rec) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double hi = h[i], rh = rho[i];
    const double invh = 1.0 / hi;
    HydroRecF r;
    r.x = x[i];
    r.y = y[i];
    r.z = z[i];
    r.h = (float)hi;
    r.vx = (float)vx[i];
    r.vy = (float)vy[i];
    r.vz = (float)vz[i];
    r.m = (float)m[i];
    r.rho = (float)rh;
    r.pf = (float)(press[i] / rh / rh / omega[i]);
    r.sound = (float)sound[i];
    r.invh = (float)invh;
    r.hfac = (float)cr_pow4(invh);
    rec[i] = r;
}

// hydro_gradh.F90:87-297 as k_hydro computes it, in FP32 on the
// FP64-formed offsets (mixed precision): half the record bytes of the FP64
// kernel, whose neighbour gathers miss L2 and bound it.  |r| and 1/|r| come
// from one reciprocal square root, the divisions from reciprocals; the
// accelerations are summed in FP32 over the particle's neighbours (tens of
// terms) and stored in FP64.  ENERGY: du/dt as k_hydro, the P div v term in
// FP64.
template <bool ENERGY>
__global__ void __launch_bounds__(256) k_hydro_mixed(
    KGeom kg, int n, const HydroRecF *__restrict__ rec, const uint8_t *__restrict__ active,
    const int *__restrict__ pptot, const int *__restrict__ pplist, const int *__restrict__ ovf_off,
    const int *__restrict__ ovf_list, int pp_limit, const float *__restrict__ w1, float alpha,
    double *__restrict__ ax, double *__restrict__ ay, double *__restrict__ az,
    const double *__restrict__ press, const double *__restrict__ divv,
    const double *__restrict__ omega, const double *__restrict__ rho,
    double *__restrict__ dudt) {
#pragma clang fp contract(fast)
    const int p = xcd_block() * blockDim.x + threadIdx.x;
    if (p >= n || !active[p]) return;
    const HydroRecF s = rec[p];
    const double rpx = s.x, rpy = s.y, rpz = s.z;
    const float hp = s.h, vpx = s.vx, vpy = s.vy, vpz = s.vz;
    const float rho_p = s.rho, pfactor_p = s.pf, sound_p = s.sound, invhp = s.invh,
                hfactor_p = s.hfac;
    const float rangesqd = (float)kg.rangesqd, range = (float)kg.range, half = (float)kg.half;
    const float beta = 2.0f * alpha;
    const float lim_p = rangesqd * hp * hp, rng_p = range * hp;
    float ahx = 0.0f, ahy = 0.0f, ahz = 0.0f, adx = 0.0f, ady = 0.0f, adz = 0.0f;
    float dudt_diss = 0.0f;
    const int nn = pptot[p];
    const int *lst = (nn <= pp_limit) ? pplist + (size_t)p * pp_limit : ovf_list + ovf_off[p];
    for (int i = 0; i < nn; i++) {
        const HydroRecF q = rec[lst[i]];
        const float hpp = q.h, mpp = q.m;
        float ux = (float)(q.x - rpx), uy = (float)(q.y - rpy), uz = (float)(q.z - rpz);
        const float dr2 = ux * ux + uy * uy + uz * uz;
        if (dr2 >= lim_p && dr2 >= rangesqd * hpp * hpp) continue;
        const float ir = __builtin_amdgcn_rsqf(dr2 + 1.0e-30f);
        const float drmag = dr2 * ir;
        const float dvx = q.vx - vpx, dvy = q.vy - vpy, dvz = q.vz - vpz;
        const float dvdr = dvx * ux + dvy * uy + dvz * uz;
        ux = ux * ir;
        uy = uy * ir;
        uz = uz * ir;
        float wmean = 0.0f;
        if (drmag < rng_p) {
            const float wk = w1[min((int)(half * drmag * invhp), KERNTOT)];
            wmean = 0.5f * hfactor_p * wk;
            const float f = mpp * pfactor_p * hfactor_p * wk;
            ahx = ahx + f * ux;
            ahy = ahy + f * uy;
            ahz = ahz + f * uz;
        }
        if (drmag < range * hpp) {
            const float wk = w1[min((int)(half * drmag * q.invh), KERNTOT)];
            wmean = wmean + 0.5f * q.hfac * wk;
            const float f = mpp * q.pf * q.hfac * wk;
            ahx = ahx + f * ux;
            ahy = ahy + f * uy;
            ahz = ahz + f * uz;
        }
        if (dvdr < 0.0f) {
            const float hs = hp + hpp;
            const float vsignal = 0.5f * (sound_p + q.sound);
            const float mu = 0.5f * hs * dvdr *
                             __builtin_amdgcn_rcpf(drmag * drmag + 0.25f * (float)ETA_SQD * hs * hs);
            const float visc = (-alpha * vsignal * mu + beta * mu * mu) *
                               __builtin_amdgcn_rcpf(0.5f * (rho_p + q.rho));
            const float f = mpp * visc * wmean;
            adx = adx + f * ux;
            ady = ady + f * uy;
            adz = adz + f * uz;
            if (ENERGY) dudt_diss = dudt_diss + 0.5f * mpp * visc * wmean * dvdr * ir;
        }
    }
    ax[p] = (double)(ahx + adx);
    ay[p] = (double)(ahy + ady);
    az[p] = (double)(ahz + adz);
    if (ENERGY)
        dudt[p] = (0.0 + (double)dudt_diss) - press[p] * divv[p] / omega[p] / rho[p];
}

// thermal.F90:108-113 (INTERNAL_ENERGY, ideal gas): P = (gamma - 1) rho u,
// c = sqrt(P / rho), for every particle: update_thermal_properties.F90:76-81
// loops over all particles once any is active (its active-list loop is
// commented out), so inactive particles take P and c from their advanced u.
// The temperature is not used on the force path.
__global__ void k_thermal_energy(int n, double gamma, const double *__restrict__ rho,
                                 const double *__restrict__ u, double *__restrict__ press,
                                 double *__restrict__ sound) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double rho_p = rho[i];
    const double press_p = (gamma - 1.0) * rho_p * u[i];
    press[i] = press_p;
    sound[i] = sqrt(press_p / rho_p);
}

// ---- standard SPH (SPH=STANDARD, HFIND=NUMBER) ---------------------------
//
// h_gather.F90:13-203 for every active particle.  The reference brackets h
// between h_low and h_high = HMULT h_low until pp_low < pp_gather <= pp_tot
// (the counts of other particles with r^2 below KERNRANGESQD h_low^2 and
// h_high^2), sorts the r^2 of the band between them and takes element
// pp_gather - pp_low: h = INVKERNRANGE sqrt(r^2 of the pp_gather-th nearest
// other particle), which does not depend on the path the bracketing took.
// Here each lane brackets the same way over its own potential-neighbour leaf
// list (recorded by a group walk at radius KERNRANGE h_search whenever h_high
// outgrows h_search, as the reference re-walks), narrows a band of more than
// BAND_MAX values by bisection in r^2, copies the band to its slice of the
// band buffer and sorts it there.  r^2 is formed as distance2.F90 forms it,
// so h is the reference's bit for bit.  A lane that cannot bracket within
// HG_ITER_MAX passes (fewer than pp_gather other particles) sets *err.
constexpr int HG_ITER_MAX = 400;

__global__ void __launch_bounds__(256) k_hgather(
    KGeom kg, bool box, int n, int nc, const HNode *__restrict__ hn, const double4 *__restrict__ rec,
    const int *__restrict__ cid, const int2 *__restrict__ root_range, int *__restrict__ hlist,
    int hl_cap, double *__restrict__ band, const uint8_t *__restrict__ active, int pp_gather,
    double *__restrict__ h, int *__restrict__ err, unsigned long long *__restrict__ wstat) {
    const int p = xcd_block() * blockDim.x + threadIdx.x;
    const bool valid = p < n && active[p];
    const int lane = threadIdx.x & (WAVE - 1);
    int *wl = hlist + (size_t)(p / WAVE) * (HL_SLOTS * WAVE) + lane;
    double *bl = band + (size_t)(p / WAVE) * (BAND_MAX * WAVE) + lane;
    double rpx = 0, rpy = 0, rpz = 0, h_old = 1.0;
    int2 rr = make_int2(nc, nc);
    if (valid) {
        const double4 a = rec[p];
        rpx = a.x; rpy = a.y; rpz = a.z;
        h_old = h[p];
        rr = root_range[cid[p]];
    }
    double h_high = HMULT * h_old, h_low = h_old / HMULT, h_search = 0.0;
    bool iterating = valid;
    int nslot = 0, cnt = 0, pp_low = 0, pp_tot = 0, passes = 0;
    bool my_ovf = false, wave_ovf = false, have_list = false, failed = false;
    unsigned vis = 0, lvis = 0, walked = 0;   // walked: passes over the tree walk
    auto dist2 = [&](const double4 &a) {   // distance2.F90, no contraction
        const double dx = a.x - rpx, dy = a.y - rpy, dz = a.z - rpz;
        return dx * dx + dy * dy + dz * dz;
    };
    // one pass over the lane's candidates: f(j, r^2) for each other particle
    auto for_each = [&](bool on, auto &&f) {
        if (!wave_ovf) {
            list_loop(wl, nslot, on, p, [&](int j) { return rec[j]; },
                      [&](int j, const double4 &a) { f(j, dist2(a)); });
        } else {
            const double R = kg.range * h_search;
            walked++;
            vis += group_walk(hn, nc, on, rr.x, rr.y, lvis,
                              [&](const HNode &nd, bool) {
                                  return hgather_visit(nd, rpx, rpy, rpz, R);
                              },
                              [&](const HNode &nd, bool mine) {
                const bool use = mine && leaf_near(nd, rpx, rpy, rpz, R, box);
                for (int i = 0; i < nd.leaf; i++) {
                    const int j = nd.first + i;
                    const double4 a = rec[j];
                    if (use && j != p) f(j, dist2(a));
                }
            });
        }
    };

    // bracketing (h_gather.F90:57-175)
    while (any_lane(iterating)) {
        passes++;
        bool newlist = false;
        if (iterating && h_high > h_search) {
            h_search = HMULT * h_high;
            newlist = true;
        }
        if (!have_list || any_lane(newlist)) {
            if (iterating) {
                nslot = 0;
                cnt = 0;
            }
            bool lovf = false;
            const double R = kg.range * h_search;
            vis += group_walk(hn, nc, iterating, rr.x, rr.y, lvis,
                              [&](const HNode &nd, bool) {
                                  return hgather_visit(nd, rpx, rpy, rpz, R);
                              },
                              [&](const HNode &nd, bool mine) {
                if (mine) {
                    cnt += nd.leaf;                       // the reference's pp_pot
                    if (leaf_near(nd, rpx, rpy, rpz, R, box)) {
                        if (nslot < hl_cap)
                            wl[nslot * WAVE] = (nd.first << 4) | nd.leaf;
                        else
                            lovf = true;
                        nslot++;
                    }
                }
            });
            if (iterating) my_ovf = lovf;
            have_list = true;
            wave_ovf = any_lane(my_ovf);   // lanes that stopped keep their lists
        }
        // too few potential neighbours: grow and walk again (:93-99)
        const bool grow = iterating && newlist && cnt <= pp_gather;
        const bool count = iterating && !grow;
        const double hr2_high = kg.rangesqd * h_high * h_high;
        const double hr2_low = kg.rangesqd * h_low * h_low;
        int lo = 0, tot = 0;
        for_each(count, [&](int, double r2) {
            if (r2 < hr2_high) {
                tot++;
                if (r2 < hr2_low) lo++;
            }
        });
        if (grow) {
            h_search = 0.0;
            h_low = h_high;
            h_high = HMULT * h_high;
        } else if (count) {
            pp_low = lo;
            pp_tot = tot;
            if (tot < pp_gather) {                    // :143-147
                h_low = h_high;
                h_high = HMULT * h_high;
            } else if (lo >= pp_gather) {             // :150-153
                h_high = h_low;
                h_low = h_low / HMULT;
            } else {                                  // :156-158
                iterating = false;
            }
        }
        if (iterating && passes >= HG_ITER_MAX) {
            iterating = false;
            failed = true;
        }
    }

    // band [lo2, hi2) of r^2 with count(r^2 < lo2) = pp_low < pp_gather <=
    // count(r^2 < hi2) = pp_tot; narrow it to BAND_MAX values at most
    double lo2 = kg.rangesqd * h_low * h_low, hi2 = kg.rangesqd * h_high * h_high;
    bool narrowing = valid && !failed && pp_tot - pp_low > BAND_MAX;
    // tie: the band [lo2, hi2) holds one double only (lo2 and hi2 adjacent),
    // so every band value is lo2 -- more than BAND_MAX particles at exactly
    // that r^2, as in a lattice shell -- and the pp_gather-th is lo2 itself
    bool tie = false;
    int nb = 0;
    while (any_lane(narrowing)) {
        const double mid = 0.5 * (lo2 + hi2);
        const bool split = mid > lo2 && mid < hi2;
        int c = 0;
        for_each(narrowing && split, [&](int, double r2) {
            if (r2 < mid) c++;
        });
        if (narrowing && !split) {
            narrowing = false;
            tie = true;
        } else if (narrowing) {
            if (c >= pp_gather) {
                hi2 = mid;
                pp_tot = c;
            } else {
                lo2 = mid;
                pp_low = c;
            }
            if (pp_tot - pp_low <= BAND_MAX) narrowing = false;
            if (++nb >= HG_ITER_MAX) {
                narrowing = false;
                failed = true;
            }
        }
    }
    // copy the band and sort it (insertion_sort.F90:4-30)
    if (tie) h[p] = kg.invrange * sqrt(lo2);
    const bool fill = valid && !failed && !tie;
    int m = 0;
    for_each(fill, [&](int, double r2) {
        if (r2 < hi2 && !(r2 < lo2) && m < BAND_MAX) {
            bl[m * WAVE] = r2;
            m++;
        }
    });
    if (fill) {
        for (int j = 1; j < m; j++) {
            const double raux = bl[j * WAVE];
            int i = j - 1;
            for (; i >= 0; i--) {
                const double ri = bl[i * WAVE];
                if (ri <= raux) break;
                bl[(i + 1) * WAVE] = ri;
            }
            bl[(i + 1) * WAVE] = raux;
        }
        const int k = pp_gather - pp_low - 1;
        if (k >= 0 && k < m)
            h[p] = kg.invrange * sqrt(bl[k * WAVE]);
        else
            failed = true;
    }
    if (failed) atomicOr(err, 1);
    wave_stats(wstat, 2, vis, lvis);
    if (walked && (threadIdx.x & 63) == 0) atomicAdd(ctr_row(wstat) + 11, (unsigned long long)walked);
}

// all_sph.F90:11-167 (standard SPH) for every active particle: rho, div v
// and drho/dt over the gather neighbours (r < KERNRANGE h_p) of its list, in
// list order.  rec holds [2i] x y z m and [2i+1] vx vy vz 0.
__global__ void __launch_bounds__(256) k_all_sph(
    KGeom kg, int n, const double4 *__restrict__ rec, const uint8_t *__restrict__ active,
    const int *__restrict__ pptot, const int *__restrict__ pplist, const int *__restrict__ ovf_off,
    const int *__restrict__ ovf_list, int pp_limit, const double *__restrict__ ktab,
    const double *__restrict__ h, double *__restrict__ rho, double *__restrict__ divv,
    double *__restrict__ drhodt) {
    const int p = xcd_block() * blockDim.x + threadIdx.x;
    if (p >= n || !active[p]) return;
    const double *w0 = ktab + T_W0 * KT, *w4 = ktab + T_W4 * KT;
    const double4 a = rec[2 * (size_t)p], b = rec[2 * (size_t)p + 1];
    const double rpx = a.x, rpy = a.y, rpz = a.z;
    const double vpx = b.x, vpy = b.y, vpz = b.z;
    const double hp = h[p], invhp = 1.0 / hp, hfactor = cr_pow3(invhp);
    const double range_p = kg.range * hp;
    double rhotemp = a.w * w0[0] * hfactor;
    double div_v_p = 0.0;
    const int nn = pptot[p];
    const int *lst = (nn <= pp_limit) ? pplist + (size_t)p * pp_limit : ovf_list + ovf_off[p];
    for (int i = 0; i < nn; i++) {
        const int j = lst[i];
        const double4 q = rec[2 * (size_t)j], qv = rec[2 * (size_t)j + 1];
        const double dvx = qv.x - vpx, dvy = qv.y - vpy, dvz = qv.z - vpz;
        const double mpp = q.w;
        const double dx = q.x - rpx, dy = q.y - rpy, dz = q.z - rpz;
        const double drsqd = dx * dx + dy * dy + dz * dz;
        const double dvdr = dvx * dx + dvy * dy + dvz * dz;
        const double drmag = sqrt(drsqd) + SMALL_NUMBER;
        if (drmag >= range_p) continue;                        // all_sph.F90:103
        const double invdrmag = 1.0 / drmag;
        const int kern = min((int)(kg.half * drmag * invhp), KERNTOT);
        rhotemp = rhotemp + mpp * hfactor * w0[kern];
        div_v_p = div_v_p - mpp * dvdr * w4[kern] * hfactor * invhp * invdrmag;
    }
    rho[p] = rhotemp;
    div_v_p = div_v_p / rhotemp;
    divv[p] = div_v_p;
    drhodt[p] = -rhotemp * div_v_p;
}

__global__ void k_pack_sph(int n, const double *__restrict__ x, const double *__restrict__ y,
                           const double *__restrict__ z, const double *__restrict__ vx,
                           const double *__restrict__ vy, const double *__restrict__ vz,
                           const double *__restrict__ m, double4 *__restrict__ rec) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    rec[2 * (size_t)i] = make_double4(x[i], y[i], z[i], m[i]);
    rec[2 * (size_t)i + 1] = make_double4(vx[i], vy[i], vz[i], 0.0);
}

// Hydro record of standard SPH (hydro.F90:83-98, :180-188): [0] x y z h
// [1] vx vy vz m  [2] rho pfactor sound invh  [3] hfactor P/(rho rho), with
// pfactor = P / rho / rho, invh = 1 / h and hfactor = invh**3 formed as the
// reference forms them.
__global__ void k_pack_hydro_std(int n, const double *__restrict__ x, const double *__restrict__ y,
                                 const double *__restrict__ z, const double *__restrict__ h,
                                 const double *__restrict__ vx, const double *__restrict__ vy,
                                 const double *__restrict__ vz, const double *__restrict__ m,
                                 const double *__restrict__ rho, const double *__restrict__ press,
                                 const double *__restrict__ sound, double4 *__restrict__ rec) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double4 *r = rec + 4 * (size_t)i;
    const double hi = h[i], invh = 1.0 / hi, rho_i = rho[i], p_i = press[i];
    r[0] = make_double4(x[i], y[i], z[i], hi);
    r[1] = make_double4(vx[i], vy[i], vz[i], m[i]);
    r[2] = make_double4(rho_i, p_i / rho_i / rho_i, sound[i], invh);
    r[3] = make_double4(cr_pow3(invh), p_i / (rho_i * rho_i), 0.0, 0.0);
}

// hydro.F90:11-323 (standard SPH, AB viscosity): the kernel gradient
// averaged over h_p and h_pp (wmean) weighs both the pressure term
// (P_p/rho_p^2 + P_pp/rho_pp^2) and the viscous term.  ENERGY adds du/dt,
// the same pair terms projected on dv.dr / |dr| (hydro.F90:206-208, :262-264).
template <bool ENERGY>
__global__ void __launch_bounds__(256) k_hydro_std(
    KGeom kg, int n, const double4 *__restrict__ rec, const uint8_t *__restrict__ active,
    const int *__restrict__ pptot, const int *__restrict__ pplist, const int *__restrict__ ovf_off,
    const int *__restrict__ ovf_list, int pp_limit, const double *__restrict__ ktab, double alpha,
    double *__restrict__ ax, double *__restrict__ ay, double *__restrict__ az,
    double *__restrict__ dudt_out) {
    const int p = xcd_block() * blockDim.x + threadIdx.x;
    if (p >= n || !active[p]) return;
    const double *w1 = ktab + T_W1 * KT;
    const double4 *rp = rec + 4 * (size_t)p;
    const double4 r0 = rp[0], r1 = rp[1], r2 = rp[2], r3 = rp[3];
    const double rpx = r0.x, rpy = r0.y, rpz = r0.z, hp = r0.w;
    const double vpx = r1.x, vpy = r1.y, vpz = r1.z;
    const double rho_p = r2.x, pfactor_p = r2.y, sound_p = r2.z, invhp = r2.w;
    const double hfactor_p = r3.x;
    double ahx = 0.0, ahy = 0.0, ahz = 0.0, adx = 0.0, ady = 0.0, adz = 0.0, dudt = 0.0;
    const int nn = pptot[p];
    const int *lst = (nn <= pp_limit) ? pplist + (size_t)p * pp_limit : ovf_list + ovf_off[p];
    for (int i = 0; i < nn; i++) {
        const double4 *rq = rec + 4 * (size_t)lst[i];
        const double4 q0 = rq[0], q1 = rq[1], q2 = rq[2];
        const double hpp = q0.w, mpp = q1.w, rho_pp = q2.x;
        const double dx = q0.x - rpx, dy = q0.y - rpy, dz = q0.z - rpz;
        const double drsqd = dx * dx + dy * dy + dz * dz;
        if (drsqd >= kg.rangesqd * hp * hp && drsqd >= kg.rangesqd * hpp * hpp) continue;
        const double drmag = sqrt(drsqd) + SMALL_NUMBER;
        const double invdrmag = 1.0 / drmag;
        const double ux = dx * invdrmag, uy = dy * invdrmag, uz = dz * invdrmag;
        const double dvx = q1.x - vpx, dvy = q1.y - vpy, dvz = q1.z - vpz;
        const double dvdr = dvx * dx + dvy * dy + dvz * dz;
        double wmean = 0.0;
        if (drmag < kg.range * hp) {
            const int kern = min((int)(kg.half * drmag * invhp), KERNTOT);
            wmean = 0.5 * hfactor_p * invhp * w1[kern];
        }
        if (drmag < kg.range * hpp) {
            const double invhpp = q2.w, hfactor_pp = rq[3].x;
            const int kern = min((int)(kg.half * drmag * invhpp), KERNTOT);
            wmean = wmean + 0.5 * hfactor_pp * invhpp * w1[kern];
        }
        const double pf = pfactor_p + rq[3].y;
        ahx = ahx + mpp * wmean * pf * ux;
        ahy = ahy + mpp * wmean * pf * uy;
        ahz = ahz + mpp * wmean * pf * uz;
        if (ENERGY) dudt = dudt + 0.5 * mpp * wmean * dvdr * invdrmag * pf;
        const double rhomean = 0.5 * (rho_p + rho_pp);
        if (dvdr < 0.0) {                                      // hydro.F90:221-224
            const double beta = 2 * alpha;
            const double vsignal = 0.5 * (sound_p + q2.z);
            const double mu = 0.5 * (hp + hpp) * dvdr /
                              (drsqd + 0.25 * ETA_SQD * (hp + hpp) * (hp + hpp));
            const double visc = (-alpha * vsignal * mu + beta * mu * mu) / rhomean;
            adx = adx + mpp * visc * wmean * ux;
            ady = ady + mpp * visc * wmean * uy;
            adz = adz + mpp * visc * wmean * uz;
            if (ENERGY) dudt = dudt + 0.5 * mpp * visc * wmean * dvdr * invdrmag;
        }
    }
    ax[p] = ahx + adx;
    ay[p] = ahy + ady;
    az[p] = ahz + adz;
    if (ENERGY) dudt_out[p] = dudt;
}

// sph_update.F90:107-109 (HFIND=CONSTANT): parray(SMOO,1:ptot) = hmin
__global__ void k_fill_h(int n, double hmin, double *__restrict__ h) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) h[i] = hmin;
}

__global__ void k_pow_third(int n, const double *__restrict__ x, double *__restrict__ y) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) y[i] = ref_pow_third(x[i]);
}

}  // namespace

int eval_pow_third(Ctx &c, int n, const double *x, double *y) {
    hipLaunchKernelGGL(k_pow_third, dim3(nblk(n)), dim3(TB), 0, c.stream, n, x, y);
    SX_HIP(hipGetLastError());
    return SERENX_OK;
}

int sph_hsolve(Ctx &c) {
    if (c.n < c.pp_gather)
        return set_error(SERENX_EFEW, "fewer particles than pp_gather: the grad-h solve "
                                      "cannot reach its neighbour number");
    SX_TRY(zero_words(c.flag, 1, c.stream));
    double4 *rec = reinterpret_cast<double4 *>(c.hrec);
    hipLaunchKernelGGL(k_pack_hsolve, dim3(nblk(c.n)), dim3(TB), 0, c.stream, c.n, c.x, c.y, c.z,
                       c.vx, c.vy, c.vz, c.m, rec);
    SX_HIP(hipGetLastError());
    // h_high per particle in the permutation scratch, list lengths in tmpi
    // the h solve at 8 waves per SIMD in mixed precision, where it is on the
    // critical path; beside the FP64 walk (issue-bound) at the compiler's 5
    const bool w8 = c.hsolve_waves == 8 || (c.hsolve_waves < 0 && c.opt.precision == SERENX_MIXED);
    hipLaunchKernelGGL(w8 ? k_hsolve_iter<8> : k_hsolve_iter<0>, dim3(nblk(c.n)), dim3(TB), 0, c.stream, c.kg, c.box_ok, c.n, c.nc, c.hn,
                       rec, c.cid, c.root_range, c.root, c.hlist, c.hl_cap, c.active, c.ktab,
                       c.opt.h_fac,
                       c.pp_gather, c.h, c.rho, c.tmpd, c.tmpi, c.wstat);
    SX_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_hsolve_final, dim3(nblk(c.n)), dim3(TB), 0, c.stream, c.kg, c.box_ok, c.n, c.nc, c.hn,
                       rec, c.cid, c.root_range, c.hlist, c.active, c.ktab, c.opt.h_fac, c.tmpd, c.tmpi, c.h, c.zeta,
                       c.rho, c.omega, c.divv, c.drhodt, c.flag, c.wstat);
    SX_HIP(hipGetLastError());
    SX_HIP(hipMemcpyAsync(c.hostpin + 8, c.flag, sizeof(int), hipMemcpyDeviceToHost, c.stream));
    return SERENX_OK;
}

int sph_hgather(Ctx &c) {
    if (c.opt.hfind == SERENX_HFIND_CONSTANT) {
        hipLaunchKernelGGL(k_fill_h, dim3(nblk(c.n)), dim3(TB), 0, c.stream, c.n, c.opt.hmin,
                           c.h);
        SX_HIP(hipGetLastError());
        c.hostpin[8] = 0;   // no h solve ran: sph_neighbours reads slot 8 as its error flag
        return SERENX_OK;
    }
    if (c.n <= c.pp_gather)
        return set_error(SERENX_EFEW, "h_gather needs more than pp_gather particles");
    const size_t need = ((size_t)c.cap + TB - 1) / TB * TB;   // whole blocks, as hlist
    if (c.band_cap < need) {
        c.graph_epoch++;
        if (c.band) (void)hipFree(c.band);
        c.band = nullptr;
        c.band_cap = 0;
        if (hipMalloc((void **)&c.band, sizeof(double) * need * BAND_MAX) != hipSuccess)
            return set_error(SERENX_ENOMEM, "hipMalloc failed for the h_gather band buffer");
        c.band_cap = need;
    }
    SX_TRY(zero_words(c.flag, 1, c.stream));
    double4 *rec = reinterpret_cast<double4 *>(c.hrec);
    hipLaunchKernelGGL(k_pack_hsolve, dim3(nblk(c.n)), dim3(TB), 0, c.stream, c.n, c.x, c.y, c.z,
                       c.vx, c.vy, c.vz, c.m, rec);
    hipLaunchKernelGGL(k_hgather, dim3(nblk(c.n)), dim3(TB), 0, c.stream, c.kg, c.box_ok, c.n, c.nc, c.hn,
                       rec, c.cid, c.root_range, c.hlist, c.hl_cap, c.band, c.active,
                       c.pp_gather, c.h,
                       c.flag, c.wstat);
    SX_HIP(hipGetLastError());
    SX_HIP(hipMemcpyAsync(c.hostpin + 8, c.flag, sizeof(int), hipMemcpyDeviceToHost, c.stream));
    SX_HIP(hipStreamSynchronize(c.stream));
    if (c.hostpin[8] != 0)
        return set_error(SERENX_EFEW, "h_gather could not bracket pp_gather neighbours "
                                      "(a cloud with too few particles)");
    c.hostpin[8] = 0;   // sph_neighbours reads it as the h-solve error flag
    return SERENX_OK;
}

int sph_density(Ctx &c) {
    double4 *rec = reinterpret_cast<double4 *>(c.hrec);
    hipLaunchKernelGGL(k_pack_sph, dim3(nblk(c.n)), dim3(TB), 0, c.stream, c.n, c.x, c.y, c.z,
                       c.vx, c.vy, c.vz, c.m, rec);
    hipLaunchKernelGGL(k_all_sph, dim3(nblk(c.n)), dim3(TB), 0, c.stream, c.kg, c.n, rec,
                       c.active, c.pptot, c.pplist, c.ovf_off, c.ovf_list, c.pp_limit, c.ktab,
                       c.h, c.rho, c.divv, c.drhodt);
    SX_HIP(hipGetLastError());
    return SERENX_OK;
}

// The neighbour rows of sph_neighbours, launches only: rows of up to
// pp_limit ids, every particle's count, and the number of particles above
// pp_limit copied to pinned slot 9 (read by sph_neib_finish once the stream
// has got there).
int sph_neib_rows(Ctx &c) {
    hipStream_t s = c.stream;
    double4 *rec = reinterpret_cast<double4 *>(c.hrec);
    // [0] x y z h with the solved h
    hipLaunchKernelGGL(k_pack_xyzh, dim3(nblk(c.n)), dim3(TB), 0, s, c.n, c.x, c.y, c.z, c.h, rec);
    SX_TRY(zero_words(c.ovf_count, 1, s));
    hipLaunchKernelGGL(c.neib_group ? k_neib_group<false> : k_neib<false>, dim3(nblk(c.n)),
                       dim3(TB), 0, s, c.kg, c.box_ok, c.n, c.ovf_idx, c.nc, c.hn,
                       rec, c.cid, c.root_range, c.active, c.pp_limit, c.pptot,
                       c.pplist, c.ovf_off, c.ovf_list, c.ovf_count, c.wstat, c.nbl, c.nbl_len, c.nb_cap);
    SX_HIP(hipGetLastError());
    SX_HIP(hipMemcpyAsync(c.hostpin + 9, c.ovf_count, sizeof(int), hipMemcpyDeviceToHost, s));
    return SERENX_OK;
}

int sph_neighbours(Ctx &c) {
    SX_TRY(sph_neib_rows(c));
    SX_HIP(hipStreamSynchronize(c.stream));
    return sph_neib_finish(c);
}

// The host part of sph_neighbours, after the stream has run sph_neib_rows:
// the h solve's status (slot 8) and the complete lists of the particles
// above pp_limit (the rows get_neib_on_fly.F90 would walk for).
int sph_neib_finish(Ctx &c) {
    hipStream_t s = c.stream;
    double4 *rec = reinterpret_cast<double4 *>(c.hrec);
    if (c.hostpin[8] != 0)
        return set_error(SERENX_EHNEG, "Negative smoothing lengths (all_sph_gradh)");
    if (c.hostpin[9] > 0) {
        // complete lists for particles above pp_limit, packed by offset
        // sizes (flag) and overflow marks (incl; both tree-build scratch)
        hipLaunchKernelGGL(k_ovf_sizes, dim3(nblk(c.n)), dim3(TB), 0, s, c.n, c.pp_limit, c.pptot,
                           c.active, c.flag, c.incl);
        SX_TRY(zero_words(c.flag + c.n, 1, s));
        size_t tb = c.sorttmp_bytes;
        SX_HIP(rocprim::exclusive_scan(c.sorttmp, tb, c.flag, c.ovf_off, 0, (size_t)c.n + 1,
                                       rocprim::plus<int>(), s));
        tb = c.sorttmp_bytes;
        SX_HIP(rocprim::exclusive_scan(c.sorttmp, tb, c.incl, c.cellof, 0, (size_t)c.n,
                                       rocprim::plus<int>(), s));
        hipLaunchKernelGGL(k_ovf_compact, dim3(nblk(c.n)), dim3(TB), 0, s, c.n, c.incl, c.cellof,
                           c.ovf_idx);
        SX_HIP(hipMemcpyAsync(c.hostpin + 10, c.ovf_off + c.n, sizeof(int),
                              hipMemcpyDeviceToHost, s));
        SX_HIP(hipStreamSynchronize(s));
        const size_t need = (size_t)c.hostpin[10];
        if (need > c.ovf_cap) {
            c.graph_epoch++;   // a captured hydro pass holds the old address
            if (c.ovf_list) (void)hipFree(c.ovf_list);
            c.ovf_list = nullptr;
            c.ovf_cap = need + need / 4 + 1024;
            if (hipMalloc(&c.ovf_list, sizeof(int) * c.ovf_cap) != hipSuccess)
                return set_error(SERENX_ENOMEM, "hipMalloc failed for overflow neighbour lists");
        }
        const int nov = c.hostpin[9];   // overflow particles, counted by k_neib<false>
        const long ovf_threads = (long)((nov + OVF_PER_WAVE - 1) / OVF_PER_WAVE) * WAVE;
        hipLaunchKernelGGL(c.neib_group ? k_neib_group<true> : k_neib<true>, dim3(nblk(ovf_threads)), dim3(TB), 0, s, c.kg, c.box_ok, nov, c.ovf_idx, c.nc,
                           c.hn, rec, c.cid, c.root_range, c.active, c.pp_limit,
                           c.pptot, c.pplist, c.ovf_off, c.ovf_list, c.ovf_count, c.wstat,
                           c.nbl, c.nbl_len, c.nb_cap);
        SX_HIP(hipGetLastError());
    }
    return SERENX_OK;
}

int sph_thermal(Ctx &c) {
    if (c.opt.thermal == SERENX_THERMAL_ENERGY_EQN) {
        hipLaunchKernelGGL(k_thermal_energy, dim3(nblk(c.n)), dim3(TB), 0, c.stream, c.n,
                           c.opt.gamma, c.rho, c.u, c.press, c.sound);
        SX_HIP(hipGetLastError());
        return SERENX_OK;
    }
    return sph_thermal_iso(c, c.opt.sound_speed);
}

int sph_thermal_iso(Ctx &c, double cs) {
    hipLaunchKernelGGL(k_thermal_iso, dim3(nblk(c.n)), dim3(TB), 0, c.stream, c.n, cs, c.active,
                       c.rho, c.press, c.sound);
    SX_HIP(hipGetLastError());
    return SERENX_OK;
}

int hydro_forces(Ctx &c) {
    double4 *rec = reinterpret_cast<double4 *>(c.hrec);
    if (c.opt.sph == SERENX_SPH_STANDARD) {
        hipLaunchKernelGGL(k_pack_hydro_std, dim3(nblk(c.n)), dim3(TB), 0, c.stream, c.n, c.x,
                           c.y, c.z, c.h, c.vx, c.vy, c.vz, c.m, c.rho, c.press, c.sound, rec);
        if (c.opt.thermal == SERENX_THERMAL_ENERGY_EQN)
            hipLaunchKernelGGL(k_hydro_std<true>, dim3(nblk(c.n)), dim3(TB), 0, c.stream, c.kg,
                               c.n, rec, c.active, c.pptot, c.pplist, c.ovf_off, c.ovf_list,
                               c.pp_limit, c.ktab, c.opt.alpha, c.ax, c.ay, c.az, c.dudt);
        else
            hipLaunchKernelGGL(k_hydro_std<false>, dim3(nblk(c.n)), dim3(TB), 0, c.stream, c.kg,
                               c.n, rec, c.active, c.pptot, c.pplist, c.ovf_off, c.ovf_list,
                               c.pp_limit, c.ktab, c.opt.alpha, c.ax, c.ay, c.az, c.dudt);
        SX_HIP(hipGetLastError());
        return SERENX_OK;
    }
    if (c.opt.precision == SERENX_MIXED) {
        // mixed precision: FP32 pair terms on 64-byte records
        HydroRecF *rf = reinterpret_cast<HydroRecF *>(c.hrec);
        const float *w1f = reinterpret_cast<const float *>(c.ktab + KT_F32_OFF) + T_W1 * KT;
        hipLaunchKernelGGL(k_pack_hydro_mixed, dim3(nblk(c.n)), dim3(TB), 0, c.stream, c.n, c.x,
                           c.y, c.z, c.h, c.vx, c.vy, c.vz, c.m, c.rho, c.omega, c.press, c.sound,
                           rf);
        if (c.opt.thermal == SERENX_THERMAL_ENERGY_EQN)
            hipLaunchKernelGGL(k_hydro_mixed<true>, dim3(nblk(c.n)), dim3(TB), 0, c.stream, c.kg,
                               c.n, rf, c.active, c.pptot, c.pplist, c.ovf_off, c.ovf_list,
                               c.pp_limit, w1f, (float)c.opt.alpha, c.ax, c.ay, c.az, c.press,
                               c.divv, c.omega, c.rho, c.dudt);
        else
            hipLaunchKernelGGL(k_hydro_mixed<false>, dim3(nblk(c.n)), dim3(TB), 0, c.stream, c.kg,
                               c.n, rf, c.active, c.pptot, c.pplist, c.ovf_off, c.ovf_list,
                               c.pp_limit, w1f, (float)c.opt.alpha, c.ax, c.ay, c.az, c.press,
                               c.divv, c.omega, c.rho, c.dudt);
        SX_HIP(hipGetLastError());
        return SERENX_OK;
    }
    hipLaunchKernelGGL(k_pack_hydro, dim3(nblk(c.n)), dim3(TB), 0, c.stream, c.n, c.x, c.y, c.z,
                       c.h, c.vx, c.vy, c.vz, c.m, c.rho, c.omega, c.press, c.sound, rec);
    SX_HIP(hipGetLastError());
    if (c.opt.thermal == SERENX_THERMAL_ENERGY_EQN)
        hipLaunchKernelGGL(k_hydro<true>, dim3(nblk(c.n)), dim3(TB), 0, c.stream, c.kg, c.n, rec,
                           c.active, c.pptot, c.pplist, c.ovf_off, c.ovf_list, c.pp_limit, c.ktab,
                           c.opt.alpha, c.ax, c.ay, c.az, c.press, c.divv, c.omega, c.dudt);
    else
        hipLaunchKernelGGL(k_hydro<false>, dim3(nblk(c.n)), dim3(TB), 0, c.stream, c.kg, c.n, rec,
                           c.active, c.pptot, c.pplist, c.ovf_off, c.ovf_list, c.pp_limit, c.ktab,
                           c.opt.alpha, c.ax, c.ay, c.az, c.press, c.divv, c.omega, c.dudt);
    SX_HIP(hipGetLastError());
    return SERENX_OK;
}

}  // namespace serenx
