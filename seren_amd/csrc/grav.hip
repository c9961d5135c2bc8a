// grav.hip -- Barnes-Hut gravity walk with monopole/quadrupole/octupole cell
// terms and the grad-h kernel-softened near field, plus the FP64
// kernel-softened direct sum, on gfx950.
//
// A wave walks the gravity tree for 64 consecutive (tree-ordered) particles
// with group_walk.  At each cell every active lane applies its own MAC
// (BHgrav_accel.F90:154-158): accepted cells add a multipole term, opened
// leaves add kernel-softened pair terms (gravity_gradh.F90), opened branches
// descend.  Each lane therefore sees exactly the reference particle's
// interaction set; MAC tests use the reference's arithmetic (no contraction)
// so that accept/open decisions are identical.  The reference sums all cell
// terms first and the near-field pairs at the end (BHgrav_accel.F90:272-290);
// here cells and pairs go to two accumulators added at the end, so results
// agree with the reference to rounding.
//
// Mixed precision (SERENX_MIXED): the cell offset rp - rc is formed in FP64
// and the multipole and pair terms are evaluated in FP32 on that
// cell-relative offset; cell terms are summed in FP32 batches added to FP64
// sums, pair terms in FP32 per leaf, each leaf's sum added to
// Neumaier-compensated FP32 sums.  Its walk uses a group MAC (see
// group_accepts).
#include <algorithm>
#include <cstring>
#include <rocprim/rocprim.hpp>

#include "walk.h"

namespace serenx {

namespace {

constexpr int TB = 256;
inline int nblk(long n, int tb = TB) { return (int)((n + tb - 1) / tb); }

// BHgrav_accel.F90:97-116: the walk's MAC factor from the previous
// evaluation's |a| (GADGET, GADGET2) or potential (EIGEN); zero while |a| is
// not yet known
template <int MACT>
__device__ __forceinline__ double mac_afactor(double agravmag, double gpot) {
    if (!(agravmag > SMALL_NUMBER)) return 0.0;
    if (MACT == SERENX_MAC_GADGET) return pow(agravmag, -ONETHIRD);
    if (MACT == SERENX_MAC_GADGET2) return 1.0 / sqrt(agravmag);
    return pow(1.0 / gpot, 2.0 * ONETHIRD);
}

// the kernel-range constants in FP32
struct KGeomF {
    float half, range, invrange;
};

// The gravity walks take the kernel range as a template argument (QUINTIC:
// the quintic's range 3, else M4's 2), so that the constants are immediates:
// as kernel arguments they cost the walks 2-6 VGPRs and a wave of occupancy.
template <bool QUINTIC>
__host__ __device__ constexpr KGeom walk_geom() {
    return kernel_geom(QUINTIC ? SERENX_KERNEL_QUINTIC : SERENX_KERNEL_M4);
}

// dx^2 + dy^2 + dz^2 in the reference's operation order with no fused
// multiply-add, also when inlined into code compiled with contraction on
__device__ __forceinline__ double dist2_unfused(double dx, double dy, double dz) {
#pragma clang fp contract(off)
    return dx * dx + dy * dy + dz * dz;
}

// 1/d for d > 0 to a few ulp: the hardware estimate refined by two Newton
// steps y <- y + y (1 - d y) (each squares the relative error).
__device__ __forceinline__ double rcp_fp64(double d) {
#pragma clang fp contract(fast)
    double y = __builtin_amdgcn_rcp(d);
#pragma unroll
    for (int k = 0; k < 2; k++) y = __fma_rn(y, __fma_rn(-d, y, 1.0), y);
    return y;
}

// gravity_gradh.F90:42-86 for one pair, (dx, dy, dz) = r_j - r_p; t4 holds
// (w4, w5, w6, 0) per table index; rangehj = KERNRANGE h_j (the reference's
// product), invhj = 1/h_j.  The kernel index uses the reference's drmag =
// sqrt(r^2) + 1e-20; 1/drmag comes from rcp_fp64.
// Returns bit 0: r inside the kernel range of h_p, bit 1: of h_j.
__device__ __forceinline__ unsigned pair_fp64(const KGeom &kg, const double4 *__restrict__ t4,
                                               double invhp,
                                               double zo_p, double dx, double dy, double dz,
                                               double mj, double rangehj, double invhj,
                                               double zetaj,
                                               double &ax, double &ay, double &az,
                                               double &pot) {
#pragma clang fp contract(fast)
    // r^2 unfused, as the reference forms it: standard SPH's h_gather puts
    // the pp_gather-th neighbour at r = KERNRANGE h_p exactly, and a fused
    // sum can move r across that boundary
    const double drsqd = dist2_unfused(dx, dy, dz);
    const double drmag = sqrt(drsqd) + SMALL_NUMBER;
    const double invdrmag = rcp_fp64(drmag);
    const double id2 = invdrmag * invdrmag;
    double grav, dpot;
    unsigned inside = 0;
    if (invdrmag > kg.invrange * invhp) {
        inside = 1;
        const int kern = min((int)(kg.half * drmag * invhp), KERNTOT);
        const double4 w = t4[kern];
        grav = invhp * invhp * (w.y + zo_p * w.x * invhp * invhp);
        dpot = invhp * w.z;
    } else {
        grav = id2;
        dpot = invdrmag;
    }
    if (drmag < rangehj) {
        inside |= 2;
        const int kern = min((int)(kg.half * drmag * invhj), KERNTOT);
        const double4 w = t4[kern];
        grav = grav + invhj * invhj * (w.y + zetaj * w.x * invhj * invhj);
        dpot = dpot + invhj * w.z;
    } else {
        grav = grav + id2;
        dpot = dpot + invdrmag;
    }
    const double f = 0.5 * mj * invdrmag * grav;
    ax = ax + f * dx;
    ay = ay + f * dy;
    az = az + f * dz;
    pot = pot + 0.5 * mj * dpot;
    return inside;
}

// Gravity source record of particle j, 64 bytes in a buffer of its own (the
// walk runs beside the neighbour and hydro kernels, which use theirs):
// [0] x y z m  [1] h 1/h zeta KERNRANGE*h
__global__ void k_pack_grav(int n, double kernrange, const double *__restrict__ x,
                            const double *__restrict__ y,
                            const double *__restrict__ z, const double *__restrict__ m,
                            const double *__restrict__ h, const double *__restrict__ zeta,
                            double4 *__restrict__ rec) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double4 *r = rec + 2 * (size_t)i;
    r[0] = make_double4(x[i], y[i], z[i], m[i]);
    r[1] = make_double4(h[i], 1.0 / h[i], zeta[i], kernrange * h[i]);
}

// Gravity source records of the mixed walk, 32 bytes per particle, written
// per leaf of the gravity tree (one thread per node; leaves partition the
// particles):
//   [0] x - x_f, y - y_f, z - z_f (FP32), m / 2   [1] KERNRANGE h, 1/h, 1/h^2, zeta/h^2
// where (x_f, y_f, z_f) is the leaf's first particle.  A pair term then
// takes the offset r_j - r_p as (r_j - r_f) + (r_f - r_p): the second term is
// formed in FP64 once per leaf and lane, so no FP64 work remains per pair.
__global__ void k_pack_grav_mixed(int nc, double kernrange, const GNodeF *__restrict__ gn,
                                  const double *__restrict__ x, const double *__restrict__ y,
                                  const double *__restrict__ z, const double *__restrict__ m,
                                  const double *__restrict__ h, const double *__restrict__ zeta,
                                  float4 *__restrict__ rec) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    const int k = gn[c].leaf, f = gn[c].first;
    if (k <= 0) return;
    const double xf = x[f], yf = y[f], zf = z[f];
    for (int i = 0; i < k; i++) {
        const int j = f + i;
        const double hj = h[j], ih = 1.0 / hj;
        rec[2 * (size_t)j] = make_float4((float)(x[j] - xf), (float)(y[j] - yf),
                                         (float)(z[j] - zf), (float)(0.5 * m[j]));
        rec[2 * (size_t)j + 1] = make_float4((float)(kernrange * hj), (float)ih,
                                             (float)(ih * ih), (float)(zeta[j] * ih * ih));
    }
}

// a += a_grav for the active particles: the one addition sph_grav_forces.F90
// makes (a holds the hydro acceleration)
__global__ void k_add_grav(int n, const uint8_t *__restrict__ active,
                           const double *__restrict__ gx, const double *__restrict__ gy,
                           const double *__restrict__ gz, double *__restrict__ ax,
                           double *__restrict__ ay, double *__restrict__ az) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !active[i]) return;
    ax[i] = ax[i] + gx[i];
    ay[i] = ay[i] + gy[i];
    az[i] = az[i] + gz[i];
}

// 1/sqrt(a) for a > 0 to about one ulp: the hardware estimate (relative
// error below 2^-23) refined by one third-order step
//   y <- y (1 + e/2 + 3e^2/8),  e = 1 - a y^2,
// whose truncation error (5/16) e^3 is below 2^-66.
__device__ __forceinline__ double rsqrt_fp64(double a) {
#pragma clang fp contract(fast)
    const double y = __builtin_amdgcn_rsq(a);
    const double e = __fma_rn(-a * y, y, 1.0);
    return __fma_rn(y * e, __fma_rn(0.375, e, 0.5), y);
}

// Multipole terms of one accepted cell at offset (x, y, z) = r_p - r_cell:
// the monopole, quadrupole and octupole terms of BHgrav_accel.F90:159-222,
// evaluated in factored form.  With Q the traceless quadrupole,
// S(r) = sum_k s_k m_k(r) the octupole cubic and H = grad(S) / 2,
//   a   = -m r / r^3 + Q r / r^5 - 5/2 (r.Q.r) r / r^7
//         + H / r^7 - 7/3 (r.H) r / r^9,
//   phi = m / r + (r.Q.r) / (2 r^5) + (r.H) / (3 r^7),
// term for term the reference's expressions (its octupole force terms are
// the components of H); S is formed as 2 (r.H) / 3 (Euler's identity for a
// homogeneous cubic).  q = [xx xy yy zx zy zz]; oc holds H's coefficients on
// the monomials (xx xy xz yy zz yz), stocked per node.
template <int OCT, class T>
__device__ __forceinline__ void cell_terms(T x, T y, T z, T ir, T m, const T *q, const T *oc,
                                           T &ax, T &ay, T &az, T &pot) {
#pragma clang fp contract(fast)
    const T ir2 = ir * ir;
    const T ir3 = ir2 * ir;
    const T ir5 = ir3 * ir2;
    const T ir7 = ir5 * ir2;
    const T Qx = q[0] * x + q[1] * y + q[3] * z;
    const T Qy = q[1] * x + q[2] * y + q[4] * z;
    const T Qz = q[3] * x + q[4] * y + q[5] * z;
    const T t5 = (x * Qx + y * Qy + z * Qz) * ir5;   // (r.Q.r) / r^5
    T crad = -(m * ir3) - T(2.5) * t5 * ir2;
    T phi = m * ir + T(0.5) * t5;
    T gx = ax + Qx * ir5, gy = ay + Qy * ir5, gz = az + Qz * ir5;
    if (OCT) {
        const T xx = x * x, yy = y * y, zz = z * z;
        const T xy = x * y, xz = x * z, yz = y * z;
        const T Hx = oc[0] * xx + oc[1] * xy + oc[2] * xz + oc[3] * yy + oc[4] * zz + oc[5] * yz;
        const T Hy = oc[6] * xx + oc[7] * xy + oc[8] * xz + oc[9] * yy + oc[10] * zz + oc[11] * yz;
        const T Hz =
            oc[12] * xx + oc[13] * xy + oc[14] * xz + oc[15] * yy + oc[16] * zz + oc[17] * yz;
        const T rh = (x * Hx + y * Hy + z * Hz) * ir7;   // (r.H) / r^7
        crad = crad - T(7.0 / 3.0) * rh * ir2;
        phi = phi + T(1.0 / 3.0) * rh;
        gx = gx + Hx * ir7;
        gy = gy + Hy * ir7;
        gz = gz + Hz * ir7;
    }
    ax = gx + crad * x;
    ay = gy + crad * y;
    az = gz + crad * z;
    pot = pot + phi;
}

// The same terms from the moments themselves (the FP32 walk, whose nodes keep
// 10 moments instead of 18 coefficients to stay at two cache lines):
// G = grad S = 2 H, S formed as (r.G) / 3.
//   s = [x^3 x^2y y^2x y^3 z^2x z^2y z^3 x^2z y^2z xyz]
template <int OCT, class T>
__device__ __forceinline__ void cell_terms_s(T x, T y, T z, T ir, T m, const T *q, const T *s,
                                             T &ax, T &ay, T &az, T &pot) {
#pragma clang fp contract(fast)
    const T ir2 = ir * ir;
    const T ir3 = ir2 * ir;
    const T ir5 = ir3 * ir2;
    const T ir7 = ir5 * ir2;
    const T Qx = q[0] * x + q[1] * y + q[3] * z;
    const T Qy = q[1] * x + q[2] * y + q[4] * z;
    const T Qz = q[3] * x + q[4] * y + q[5] * z;
    const T qs = x * Qx + y * Qy + z * Qz;
    T crad = -(m * ir3) - T(2.5) * qs * ir7;
    T phi = m * ir + T(0.5) * qs * ir5;
    T fx = Qx * ir5, fy = Qy * ir5, fz = Qz * ir5;
    if (OCT) {
        const T xx = x * x, yy = y * y, zz = z * z;
        const T xy = x * y, xz = x * z, yz = y * z;
        const T xx3 = T(3) * xx, yy3 = T(3) * yy, zz3 = T(3) * zz;
        const T xy2 = xy + xy, xz2 = xz + xz, yz2 = yz + yz;
        const T Gx = s[0] * xx3 + s[1] * xy2 + s[7] * xz2 + s[2] * yy + s[4] * zz + s[9] * yz;
        const T Gy = s[1] * xx + s[2] * xy2 + s[3] * yy3 + s[5] * zz + s[9] * xz + s[8] * yz2;
        const T Gz = s[7] * xx + s[8] * yy + s[6] * zz3 + s[4] * xz2 + s[5] * yz2 + s[9] * xy;
        const T rg = x * Gx + y * Gy + z * Gz;   // 3 S
        const T h7 = T(0.5) * ir7;
        crad = crad - T(3.5 / 3.0) * rg * ir7 * ir2;
        phi = phi + T(0.5 / 3.0) * rg * ir7;
        fx = fx + h7 * Gx;
        fy = fy + h7 * Gy;
        fz = fz + h7 * Gz;
    }
    ax = ax + (fx + crad * x);
    ay = ay + (fy + crad * y);
    az = az + (fz + crad * z);
    pot = pot + phi;
}

template <int OCT>
__device__ __forceinline__ void cell_fp64(const GNode &g, const double dr[3], double drsqd,
                                          double &ax, double &ay, double &az, double &pot) {
    const double ir = rsqrt_fp64(drsqd + SMALL_NUMBER);
    cell_terms<OCT, double>(dr[0], dr[1], dr[2], ir, g.m, g.q, g.oc, ax, ay, az, pot);
}

// Near-field list of a group: the leaves its walk opened, in DFS order, as
// (first particle << 4) | particle count, with the mask of the lanes that
// opened each.  Phase 1 of the walk writes it, phase 2 reads it.
struct NearList {
    int *leaf;                    // ngroups x stride entries
    unsigned long long *mask;     // ngroups x stride lane masks
    int *len;                     // per group: entries found (> cap: list incomplete)
    int cap;                      // entries a group may store (<= stride)
    int stride;                   // entries per group in leaf[] and mask[]
};

__device__ __forceinline__ void near_append(const NearList &L, int g, int &len, int first,
                                            int count, unsigned long long mask) {
    if ((threadIdx.x & 63) == 0 && len < L.cap) {
        L.leaf[(size_t)g * L.stride + len] = (first << 4) | count;
        L.mask[(size_t)g * L.stride + len] = mask;
    }
    len++;
}

// The FP64 walk in two phases, so that the traversal and the cell terms --
// which need only the positions, the tree stocked with the previous h (Q2:
// BHgrav_stock runs before sph_update) and the previous |a| -- run beside
// the h solve, and only the kernel-softened pair terms wait for the new h
// and zeta.  The reference sums the cell terms of the walk first and the
// near-field pairs of its list afterwards (BHgrav_accel.F90:225-226,
// 272-290); so does this split, in the same orders.
//
// Phase 1: at each cell every active lane applies its own MAC
// (BHgrav_accel.F90:154-158); accepted cells add a multipole term, opened
// leaves go to the group's near-field list, opened branches descend.
template <int OCT, int MACT>
__global__ void __launch_bounds__(256) k_grav_cells_fp64(
    int n, int nc, const GNode *__restrict__ gn, const double *__restrict__ x,
    const double *__restrict__ y, const double *__restrict__ z, const int *__restrict__ cid,
    const int2 *__restrict__ root_range, const uint8_t *__restrict__ active,
    const double *__restrict__ gpot, const double *__restrict__ agm, double *__restrict__ cx,
    double *__restrict__ cy, double *__restrict__ cz, double *__restrict__ cp, NearList near,
    unsigned long long *__restrict__ icount, unsigned long long *__restrict__ wstat,
    GroupQueue queue, int *__restrict__ cost, const int2 *__restrict__ grp) {
    // the walk is the critical path while the SPH kernels run beside it: its
    // waves take issue priority over theirs on shared SIMDs
    __builtin_amdgcn_s_setprio(2);
    unsigned ncell = 0, vis = 0, lvis = 0;   // wave totals
    const int home = blockIdx.x % NXCD;
    int qdone = 0;
    for (int g = next_group(queue, home, qdone); g >= 0; g = next_group(queue, home, qdone)) {
        const int2 gr = grp[g];
        const int p = gr.x + (int)(threadIdx.x & (WAVE - 1));
        const bool valid = p < gr.y && active[p];
        double rpx = 0, rpy = 0, rpz = 0, afactor = 0.0;
        if (valid) {
            rpx = x[p]; rpy = y[p]; rpz = z[p];
            if (MACT != SERENX_MAC_GEOMETRIC) afactor = mac_afactor<MACT>(agm[p], gpot[p]);
        }
        double gx = 0, gy = 0, gz = 0, gp = 0;
        int len = 0;
        const int2 rr = valid ? root_range[cid[p]] : make_int2(nc, nc);
        const unsigned gv = group_walk(
            gn, nc, valid, rr.x, rr.y, lvis,
            [&](const GNode &nd, bool here) -> int {
                const double dr[3] = {rpx - nd.r[0], rpy - nd.r[1], rpz - nd.r[2]};
                const double drsqd = dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2];
                bool accept;
                if (MACT == SERENX_MAC_GEOMETRIC)
                    accept = drsqd > nd.dminsqd;
                else
                    accept = drsqd > nd.mac * afactor && drsqd > nd.dminsqd;
                accept = accept && nd.leaf != 1;                   // :128-136, no MAC test
                ncell += lanes(here && accept);
                if (here && accept) cell_fp64<OCT>(nd, dr, drsqd, gx, gy, gz, gp);
                if (accept) return WALK_SKIP;
                return nd.leaf > 0 ? WALK_LEAF : WALK_DESCEND;     // :242-258
            },
            [&](const GNode &nd, bool mine) {
                near_append(near, g, len, nd.first, nd.leaf, ballot(mine));
            });
        vis += gv;
        if ((threadIdx.x & 63) == 0) {
            cost[g] = (int)gv;
            near.len[g] = len;
        }
        if (valid) {
            cx[p] = gx;
            cy[p] = gy;
            cz[p] = gz;
            cp[p] = gp;
        }
    }
    if ((threadIdx.x & 63) == 0) atomicAdd(ctr_row(icount) + 0, (unsigned long long)ncell);
    wave_stats(wstat, 0, vis, lvis);
}

// Phase 2: the pair terms of gravity_gradh.F90 over each lane's opened
// leaves, in list order, added to the cell sums of phase 1.  A group whose
// list did not fit repeats its walk (the same MAC decisions: they read the
// stocked tree and the previous |a|, which phase 2 overwrites only after)
// and takes the pairs from the leaves as it opens them.  One wave per group
// at a time; the kernels below take the groups from the walk's queue.
template <int MACT, bool QUINTIC>
__device__ __forceinline__ void pairs_group_fp64(
    int g, int n, int nc, const GNode *__restrict__ gn, const double4 *__restrict__ rec,
    const int *__restrict__ cid, const int2 *__restrict__ root_range,
    const uint8_t *__restrict__ active, const double4 *__restrict__ t4,
    double *__restrict__ ax, double *__restrict__ ay, double *__restrict__ az,
    const double *__restrict__ cp, double *__restrict__ gpot, double *__restrict__ agm,
    NearList near, unsigned long long *__restrict__ icount, const int2 *__restrict__ grp,
    int ngrp,
    unsigned &npp_out, unsigned &nin_p_out, unsigned &nin_j_out) {
    constexpr KGeom kg = walk_geom<QUINTIC>();
    const int lane = (int)(threadIdx.x & (WAVE - 1));
    const int2 gr = grp[g];
    const int p = gr.x + lane;
    const bool valid = p < gr.y && active[p];
    if (ballot(valid) == 0) return;
    double rpx = 0, rpy = 0, rpz = 0, invhp = 1.0, zo_p = 0.0;
    double gx = 0, gy = 0, gz = 0, gp = 0, afactor = 0.0;
    if (valid) {
        const double4 a = rec[2 * (size_t)p], b = rec[2 * (size_t)p + 1];
        rpx = a.x; rpy = a.y; rpz = a.z;
        invhp = b.y;
        zo_p = b.z;
        gx = ax[p]; gy = ay[p]; gz = az[p]; gp = cp[p];
        if (MACT != SERENX_MAC_GEOMETRIC) afactor = mac_afactor<MACT>(agm[p], gpot[p]);
    }
    unsigned npp = 0, nin_p = 0, nin_j = 0;
    auto leaf_pairs = [&](int first, int count, bool mine) {
        for (int i = 0; i < count; i++) {
            const int j = first + i;
            const double4 a = rec[2 * (size_t)j], b = rec[2 * (size_t)j + 1];
            const bool act = mine && j != p;
            unsigned in_ = 0;
            if (act)
                in_ = pair_fp64(kg, t4, invhp, zo_p, a.x - rpx, a.y - rpy, a.z - rpz, a.w, b.w,
                                b.y, b.z, gx, gy, gz, gp);
            npp += lanes(act);
            nin_p += lanes(in_ & 1);
            nin_j += lanes(in_ & 2);
        }
    };
    const int len = uniform(near.len[g]);
    if (len <= near.cap) {
        for (int e = 0; e < len; e++) {
            const int d = uniform(near.leaf[(size_t)g * near.stride + e]);
            const unsigned long long m = near.mask[(size_t)g * near.stride + e];
            leaf_pairs(d >> 4, d & 15, valid && ((m >> lane) & 1ull));
        }
    } else {
        const int2 rr = valid ? root_range[cid[p]] : make_int2(nc, nc);
        unsigned lv = 0;
        (void)group_walk(
            gn, nc, valid, rr.x, rr.y, lv,
            [&](const GNode &nd, bool) -> int {
                const double dr[3] = {rpx - nd.r[0], rpy - nd.r[1], rpz - nd.r[2]};
                const double drsqd = dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2];
                bool accept;
                if (MACT == SERENX_MAC_GEOMETRIC)
                    accept = drsqd > nd.dminsqd;
                else
                    accept = drsqd > nd.mac * afactor && drsqd > nd.dminsqd;
                accept = accept && nd.leaf != 1;
                if (accept) return WALK_SKIP;
                return nd.leaf > 0 ? WALK_LEAF : WALK_DESCEND;
            },
            [&](const GNode &nd, bool mine) { leaf_pairs(nd.first, nd.leaf, mine); });
    }
    if (valid) {
        ax[p] = gx;   // the gravity acceleration; k_add_grav adds it to a
        ay[p] = gy;
        az[p] = gz;
        gpot[p] = gp;
        if (MACT != SERENX_MAC_GEOMETRIC) agm[p] = sqrt(gx * gx + gy * gy + gz * gz);
    }
    npp_out += npp;
    nin_p_out += nin_p;
    nin_j_out += nin_j;
}

template <int MACT, bool QUINTIC>
__global__ void __launch_bounds__(256) k_grav_pairs_fp64(
    int n, int nc, const GNode *__restrict__ gn, const double4 *__restrict__ rec,
    const int *__restrict__ cid, const int2 *__restrict__ root_range,
    const uint8_t *__restrict__ active, const double4 *__restrict__ t4,
    double *__restrict__ ax, double *__restrict__ ay, double *__restrict__ az,
    const double *__restrict__ cp, double *__restrict__ gpot, double *__restrict__ agm,
    NearList near, unsigned long long *__restrict__ icount, const int2 *__restrict__ grp,
    int ngrp,
    GroupQueue queue) {
    // persistent over the walk's queue: the groups of each XCD range
    // longest-first (by the previous walk's cost), as phase 1 takes them
    __builtin_amdgcn_s_setprio(2);
    unsigned npp = 0, nin_p = 0, nin_j = 0;   // wave totals
    const int home = blockIdx.x % NXCD;
    int qdone = 0;
    for (int g = next_group(queue, home, qdone); g >= 0; g = next_group(queue, home, qdone))
        pairs_group_fp64<MACT, QUINTIC>(g, n, nc, gn, rec, cid, root_range, active, t4, ax, ay,
                                        az, cp, gpot, agm, near, icount, grp, ngrp, npp, nin_p,
                                        nin_j);
    if ((threadIdx.x & 63) == 0) {
        unsigned long long *ic = ctr_row(icount);
        atomicAdd(ic + 1, (unsigned long long)npp);
        atomicAdd(ic + 2, (unsigned long long)nin_p);
        atomicAdd(ic + 3, (unsigned long long)nin_j);
    }
}

// ---- mixed precision ------------------------------------------------------
// Neumaier compensated FP32 sum
struct CSum {
    float s = 0.0f, c = 0.0f;
    __device__ __forceinline__ void add(float v) {
        const float t = s + v;
        c += (fabsf(s) >= fabsf(v)) ? ((s - t) + v) : ((v - t) + s);
        s = t;
    }
    __device__ __forceinline__ double value() const { return (double)s + (double)c; }
};

// The target's side of a mixed pair term, per lane: 1/h_p, 1/h_p^2,
// zeta_p/h_p^2, the kernel-range limit on 1/r (KERNRANGE h_p)^-1 and the
// table scale KERNTOT / (KERNRANGE h_p)
struct PairTarget {
    float ih, ih2, zi2, lim, hk;
};

// gravity_gradh.F90:42-86 for one pair in FP32 on the offset (dx, dy, dz) =
// r_j - r_p, with the source record of k_pack_grav_mixed (hm = m_j / 2, rh =
// KERNRANGE h_j, ih = 1/h_j, ih2 = 1/h_j^2, zi2 = zeta_j/h_j^2).  |r| and 1/|r|
// come from one reciprocal square root.  The softened terms are selected per
// lane under wave-uniform branches (taken when any lane is inside the
// range), so that the wave runs no divergent branches; a lane with act
// unset adds nothing.  The terms
// are added to FP32 sums (the caller's per-leaf partial sums).  mact is the
// wave's mask of the lanes with act set; the wave counts the pairs inside
// the kernel range of h_p (nin_p) and of h_j (nin_j) from the masks of the
// two comparisons themselves.
__device__ __forceinline__ void pair_f32(const KGeomF &kg, const float4 *__restrict__ t4,
                                         const PairTarget &tp, bool act,
                                         unsigned long long mact, float dx, float dy,
                                         float dz, float hm, float rh, float ih, float ih2,
                                         float zi2, float &ax, float &ay, float &az, float &pot,
                                         unsigned &nin_p, unsigned &nin_j) {
#pragma clang fp contract(fast)
    const float d2 = dx * dx + dy * dy + dz * dz;
    const float ir = __builtin_amdgcn_rsqf(d2 + 1.0e-30f);
    const float dr = d2 * ir;
    const float ir2 = ir * ir;
    const bool in_p = act && ir > tp.lim;
    const bool in_j = act && dr < rh;
    const unsigned long long mp = ballot(ir > tp.lim) & mact, mj = ballot(dr < rh) & mact;
    nin_p += (unsigned)__popcll(mp);
    nin_j += (unsigned)__popcll(mj);
    float grav = ir2, dpot = ir;
    // the kernel-softened parts only when a lane needs them (wave-uniform
    // branches); unsigned table indices (at most KERNTOT) give 32-bit offsets
    if (mp) {
        const float4 wp = t4[(unsigned)min((int)(tp.hk * dr), KERNTOT)];
        grav = in_p ? tp.ih2 * (wp.y + tp.zi2 * wp.x) : grav;
        dpot = in_p ? tp.ih * wp.z : dpot;
    }
    float gj = ir2, pj = ir;
    if (mj) {
        const float4 wj = t4[(unsigned)min((int)(kg.half * dr * ih), KERNTOT)];
        gj = in_j ? ih2 * (wj.y + zi2 * wj.x) : gj;
        pj = in_j ? ih * wj.z : pj;
    }
    grav = grav + gj;
    dpot = dpot + pj;
    const float f = act ? hm * ir * grav : 0.0f;
    ax = ax + f * dx;
    ay = ay + f * dy;
    az = az + f * dz;
    pot = pot + (act ? hm * dpot : 0.0f);
}

// FP32 multipole terms on the FP32 cell offset, added to an FP32 batch sum
template <int OCT>
__device__ __forceinline__ void cell_fp32(const GNodeF &g, const float dr[3], float &ax,
                                          float &ay, float &az, float &pot) {
#pragma clang fp contract(fast)
    const float drsqd = dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2];
    const float ir = __builtin_amdgcn_rsqf(drsqd + 1.0e-30f);
    cell_terms_s<OCT, float>(dr[0], dr[1], dr[2], ir, g.m, g.q, g.s, ax, ay, az, pot);
}

// Cell terms are summed in FP32 over batches of CELL_BATCH cell visits of the
// wave and each batch is added to FP64 sums: a batch sum carries at most a
// few FP32 roundings per term, and the conversions and FP64 additions are
// paid once per batch instead of once per cell.
constexpr int CELL_BATCH = 8;

// The mixed walk uses a group MAC: a wave of 64 tree-ordered particles
// accepts a cell only when the MAC of BHgrav_accel.F90:154-158 accepts it for
// every one of its particles, and otherwise opens it for all of them (leaves
// give pair terms to all).  Every accepted cell therefore satisfies each
// particle's own MAC, so the truncation error is at most the per-particle
// walk's; the walk is wave-uniform (no per-lane resume state), and no cell
// term is issued for a cell that only some lanes use.  The FP64 walk keeps
// the reference's per-particle interaction sets.  As the FP64 walk, it runs
// in two phases: cells (beside the h solve), then pairs.
//
// A group lies in one cloud (its walk covers that cloud's DFS range), so
// the lanes taking part are the same at every cell: vmask, the valid lanes.
// The test forms each comparison's lane mask directly and combines the
// masks in scalar registers.
template <int MACT>
__device__ __forceinline__ bool group_accepts(const GNodeF &nd, unsigned long long vmask,
                                              double rpx, double rpy, double rpz, double afactor,
                                              double &d0, double &d1, double &d2) {
    d0 = rpx - nd.r[0];
    d1 = rpy - nd.r[1];
    d2 = rpz - nd.r[2];
    const double drsqd = d0 * d0 + d1 * d1 + d2 * d2;
    unsigned long long ok = ballot(drsqd > nd.dminsqd);
    if (MACT != SERENX_MAC_GEOMETRIC) ok &= ballot(drsqd > nd.mac * afactor);
    return nd.leaf != 1 && (vmask & ~ok) == 0;       // :128-136; every lane accepts
}

template <int OCT, int MACT>
__global__ void __launch_bounds__(256) k_grav_cells_mixed(
    int n, int nc, const GNodeF *__restrict__ gn, const double *__restrict__ x,
    const double *__restrict__ y, const double *__restrict__ z, const int *__restrict__ cid,
    const int2 *__restrict__ root_range, const uint8_t *__restrict__ active,
    const double *__restrict__ gpot, const double *__restrict__ agm, double *__restrict__ ox,
    double *__restrict__ oy, double *__restrict__ oz, double *__restrict__ op, NearList near,
    unsigned long long *__restrict__ icount, unsigned long long *__restrict__ wstat,
    GroupQueue queue, int *__restrict__ cost, const int2 *__restrict__ grp) {
    __builtin_amdgcn_s_setprio(2);
    unsigned ncell = 0, vis = 0, lvis = 0;   // wave totals
    const int home = blockIdx.x % NXCD;
    int qdone = 0;
    for (int g = next_group(queue, home, qdone); g >= 0; g = next_group(queue, home, qdone)) {
        const int2 gr = grp[g];
        const int p = gr.x + (int)(threadIdx.x & (WAVE - 1));
        const bool valid = p < gr.y && active[p];
        double rpx = 0, rpy = 0, rpz = 0, afactor = 0.0;
        if (valid) {
            rpx = x[p]; rpy = y[p]; rpz = z[p];
            if (MACT != SERENX_MAC_GEOMETRIC) afactor = mac_afactor<MACT>(agm[p], gpot[p]);
        }
        double cx = 0.0, cy = 0.0, cz = 0.0, cp = 0.0;   // cell terms
        float bx = 0.0f, by = 0.0f, bz = 0.0f, bp = 0.0f; // current batch of cell terms
        int nbatch = 0, len = 0;                          // wave-uniform
        unsigned gv = 0;
        // the walk covers the tree of the group's cloud (groups never straddle
        // clouds); every valid lane takes part at every cell, and the cell
        // terms of the other lanes are formed and never stored
        const unsigned long long vmask = ballot(valid);
        const unsigned nvalid = (unsigned)__popcll(vmask);
        const int2 rr = valid ? root_range[cid[p]] : make_int2(nc, nc);
        int c = uniform(wave_min_int(valid ? rr.x : nc));
        const int cend = uniform(wave_max_int(valid ? rr.y : 0));
        while (c < cend) {
            const GNodeF nd = gn[c];
            gv++;
            lvis += nvalid;
            double d0, d1, d2;
            if (group_accepts<MACT>(nd, vmask, rpx, rpy, rpz, afactor, d0, d1, d2)) {
                ncell += nvalid;
                const float dr[3] = {(float)d0, (float)d1, (float)d2};
                cell_fp32<OCT>(nd, dr, bx, by, bz, bp);
                if (++nbatch == CELL_BATCH) {
                    cx += (double)bx; cy += (double)by; cz += (double)bz; cp += (double)bp;
                    bx = by = bz = bp = 0.0f;
                    nbatch = 0;
                }
                c = nd.next;
            } else if (nd.leaf > 0) {
                near_append(near, g, len, nd.first, nd.leaf, vmask);
                c = nd.next;
            } else {
                c = c + 1;
            }
        }
        cx += (double)bx; cy += (double)by; cz += (double)bz; cp += (double)bp;
        vis += gv;
        if ((threadIdx.x & 63) == 0) {
            cost[g] = (int)gv;
            near.len[g] = len;
        }
        if (valid) {
            ox[p] = cx;
            oy[p] = cy;
            oz[p] = cz;
            op[p] = cp;
        }
    }
    if ((threadIdx.x & 63) == 0) atomicAdd(ctr_row(icount) + 0, (unsigned long long)ncell);
    wave_stats(wstat, 0, vis, lvis);
}

// leaves of a near list staged in LDS at once by the mixed pair phase
constexpr int PL_STAGE = 32;

template <int MACT, bool QUINTIC>
__device__ __forceinline__ void pairs_group_mixed(
    int g, int n, int nc, const GNodeF *__restrict__ gn, const float4 *__restrict__ rec,
    float4 *__restrict__ stage,
    const double *__restrict__ x, const double *__restrict__ y, const double *__restrict__ z,
    const double *__restrict__ h, const double *__restrict__ zeta,
    const int *__restrict__ cid, const int2 *__restrict__ root_range,
    const uint8_t *__restrict__ active, const float4 *__restrict__ t4,
    double *__restrict__ ax, double *__restrict__ ay, double *__restrict__ az,
    const double *__restrict__ cpot, double *__restrict__ gpot, double *__restrict__ agm,
    NearList near, unsigned long long *__restrict__ icount, const int2 *__restrict__ grp,
    int ngrp,
    unsigned &npp_out, unsigned &nin_p_out, unsigned &nin_j_out) {
    constexpr KGeom kg64 = walk_geom<QUINTIC>();
    constexpr KGeomF kg{(float)kg64.half, (float)kg64.range, (float)kg64.invrange};
    const int lane = (int)(threadIdx.x & (WAVE - 1));
    const int2 gr = grp[g];
    const int p = gr.x + lane;
    const bool valid = p < gr.y && active[p];
    if (ballot(valid) == 0) return;
    double rpx = 0, rpy = 0, rpz = 0, afactor = 0.0;
    PairTarget tp{1.0f, 1.0f, 0.0f, 0.0f, 0.0f};
    if (valid) {
        rpx = x[p]; rpy = y[p]; rpz = z[p];
        const double ih = 1.0 / h[p];
        tp.ih = (float)ih;
        tp.ih2 = (float)(ih * ih);
        tp.zi2 = (float)(zeta[p] * ih * ih);
        tp.lim = (float)(kg64.invrange * ih);
        tp.hk = (float)(kg64.half * ih);
        if (MACT != SERENX_MAC_GEOMETRIC) afactor = mac_afactor<MACT>(agm[p], gpot[p]);
    }
    CSum px, py, pz, pq;   // pair terms, one compensated addition per leaf
    unsigned npp = 0, nin_p = 0, nin_j = 0;
    auto leaf_pairs = [&](int first, int count, bool mine) {
        // r_f - r_p in FP64, once per leaf (r_f: the leaf's first particle)
        const float ox = (float)(x[first] - rpx), oy = (float)(y[first] - rpy),
                    oz = (float)(z[first] - rpz);
        float bx = 0.0f, by = 0.0f, bz = 0.0f, bp = 0.0f;
        const unsigned long long mm = ballot(mine);
        // the next source's record is loaded (scalar loads) before the
        // current pair is formed; after the leaf's last particle that is
        // particle first + count <= n, within the buffer's n + 1 records
        float4 a = rec[2 * (size_t)first], b = rec[2 * (size_t)first + 1];
        for (int i = 0; i < count; i++) {
            const int j = first + i;
            const float4 an = rec[2 * (size_t)(j + 1)], bn = rec[2 * (size_t)(j + 1) + 1];
            // the lanes taking the pair: those that opened the leaf, less the
            // one whose particle is j
            const unsigned sl = (unsigned)(j - gr.x);
            const unsigned long long mact = mm & ~(sl < WAVE ? 1ull << sl : 0ull);
            npp += (unsigned)__popcll(mact);
            pair_f32(kg, t4, tp, mine && j != p, mact, a.x + ox, a.y + oy, a.z + oz, a.w, b.x,
                     b.y, b.z, b.w, bx, by, bz, bp, nin_p, nin_j);
            a = an;
            b = bn;
        }
        if (mine) {
            px.add(bx);
            py.add(by);
            pz.add(bz);
            pq.add(bp);
        }
    };
    // the same terms with the leaf's records read from the wave's LDS stage
    // (src: the leaf's LEAFMAX slots of two float4)
    auto leaf_pairs_staged = [&](const float4 *src, int first, int count, bool mine) {
        const float ox = (float)(x[first] - rpx), oy = (float)(y[first] - rpy),
                    oz = (float)(z[first] - rpz);
        float bx = 0.0f, by = 0.0f, bz = 0.0f, bp = 0.0f;
        const unsigned long long mm = ballot(mine);
        for (int i = 0; i < count; i++) {
            const int j = first + i;
            const float4 a = src[2 * i], b = src[2 * i + 1];
            const unsigned sl = (unsigned)(j - gr.x);
            const unsigned long long mact = mm & ~(sl < WAVE ? 1ull << sl : 0ull);
            npp += (unsigned)__popcll(mact);
            pair_f32(kg, t4, tp, mine && j != p, mact, a.x + ox, a.y + oy, a.z + oz, a.w, b.x,
                     b.y, b.z, b.w, bx, by, bz, bp, nin_p, nin_j);
        }
        if (mine) {
            px.add(bx);
            py.add(by);
            pz.add(bz);
            pq.add(bp);
        }
    };
    const int len = uniform(near.len[g]);
    if (len <= near.cap) {
        // the near list in batches of PL_STAGE leaves: lane l copies the
        // records of the batch's leaf l into the wave's LDS stage (all the
        // batch's loads in flight at once), then the wave forms the pairs
        // leaf by leaf in list order from LDS
        for (int e0 = 0; e0 < len; e0 += PL_STAGE) {
            const int nb = min(PL_STAGE, len - e0);
            int dl = 0;
            if (lane < nb) {
                dl = near.leaf[(size_t)g * near.stride + e0 + lane];
                const int f = dl >> 4, k = dl & 15;
                float4 *dst = stage + 2 * LEAFMAX * lane;
                for (int i = 0; i < k; i++) {
                    dst[2 * i] = rec[2 * (size_t)(f + i)];
                    dst[2 * i + 1] = rec[2 * (size_t)(f + i) + 1];
                }
            }
            // the wave's LDS writes before its reads (lanes read other
            // lanes' slots)
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
            for (int e = 0; e < nb; e++) {
                const int d = __builtin_amdgcn_readlane(dl, e);
                const unsigned long long m = near.mask[(size_t)g * near.stride + e0 + e];
                leaf_pairs_staged(stage + 2 * LEAFMAX * e, d >> 4, d & 15,
                                  valid && ((m >> lane) & 1ull));
            }
            // every read of this batch before the next batch's writes
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
    } else {
        // the list did not fit: the same wave-uniform walk again, pairs only
        const int2 rr = valid ? root_range[cid[p]] : make_int2(nc, nc);
        int c = uniform(wave_min_int(valid ? rr.x : nc));
        const int cend = uniform(wave_max_int(valid ? rr.y : 0));
        const unsigned long long vmask = ballot(valid);
        while (c < cend) {
            const GNodeF nd = gn[c];
            double d0, d1, d2;
            if (group_accepts<MACT>(nd, vmask, rpx, rpy, rpz, afactor, d0, d1, d2)) {
                c = nd.next;
            } else if (nd.leaf > 0) {
                leaf_pairs(nd.first, nd.leaf, valid);
                c = nd.next;
            } else {
                c = c + 1;
            }
        }
    }
    if (valid) {
        const double gx = ax[p] + px.value(), gy = ay[p] + py.value(), gz = az[p] + pz.value();
        ax[p] = gx;   // the gravity acceleration; k_add_grav adds it to a
        ay[p] = gy;
        az[p] = gz;
        gpot[p] = cpot[p] + pq.value();
        if (MACT != SERENX_MAC_GEOMETRIC) agm[p] = sqrt(gx * gx + gy * gy + gz * gz);
    }
    npp_out += npp;
    nin_p_out += nin_p;
    nin_j_out += nin_j;
}

template <int MACT, bool QUINTIC>
__global__ void __launch_bounds__(256) k_grav_pairs_mixed(
    int n, int nc, const GNodeF *__restrict__ gn, const float4 *__restrict__ rec,
    const double *__restrict__ x, const double *__restrict__ y, const double *__restrict__ z,
    const double *__restrict__ h, const double *__restrict__ zeta,
    const int *__restrict__ cid, const int2 *__restrict__ root_range,
    const uint8_t *__restrict__ active, const float4 *__restrict__ t4,
    double *__restrict__ ax, double *__restrict__ ay, double *__restrict__ az,
    const double *__restrict__ cpot, double *__restrict__ gpot, double *__restrict__ agm,
    NearList near, unsigned long long *__restrict__ icount, const int2 *__restrict__ grp,
    int ngrp,
    GroupQueue queue) {
    // persistent over the walk's queue: the groups of each XCD range
    // longest-first (by the previous walk's cost), as phase 1 takes them
    __builtin_amdgcn_s_setprio(2);
    // each wave's LDS stage: PL_STAGE leaves of LEAFMAX two-float4 records
    __shared__ float4 stage_all[(256 / WAVE) * PL_STAGE * LEAFMAX * 2];
    float4 *stage = stage_all + (threadIdx.x / WAVE) * (PL_STAGE * LEAFMAX * 2);
    unsigned npp = 0, nin_p = 0, nin_j = 0;   // wave totals
    const int home = blockIdx.x % NXCD;
    int qdone = 0;
    for (int g = next_group(queue, home, qdone); g >= 0; g = next_group(queue, home, qdone))
        pairs_group_mixed<MACT, QUINTIC>(g, n, nc, gn, rec, stage, x, y, z, h, zeta, cid,
                                         root_range, active, t4, ax, ay, az, cpot, gpot, agm,
                                         near, icount, grp, ngrp, npp, nin_p, nin_j);
    if ((threadIdx.x & 63) == 0) {
        unsigned long long *ic = ctr_row(icount);
        atomicAdd(ic + 1, (unsigned long long)npp);
        atomicAdd(ic + 2, (unsigned long long)nin_p);
        atomicAdd(ic + 3, (unsigned long long)nin_j);
    }
}

// Work-queue order of the walk's groups: keys (range, ~cost) so that an
// ascending sort lists every XCD range's groups by decreasing cost of their
// previous walk; with no cost yet the keys are the ranges alone (a stable
// sort keeps tree order within each)
__global__ void k_queue_keys(int ngroups, const int *__restrict__ cost, int have_cost,
                             uint64_t *__restrict__ keys, int *__restrict__ vals) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= ngroups) return;
    const uint64_t range = (uint64_t)(((long long)NXCD * g) / ngroups);
    const uint32_t c = have_cost ? (uint32_t)max(cost[g], 0) : 0u;
    keys[g] = (range << 32) | (uint64_t)(0xFFFFFFFFu - c);
    vals[g] = g;
}

// the queue in tree order (each XCD range's groups in order)
__global__ void k_queue_identity(int ngroups, int *__restrict__ order) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g < ngroups) order[g] = g;
}

__global__ void k_zero_ints(int n, int *__restrict__ a) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = 0;
}

// direct_sph_gravity.F90:63-79: sum of gravity_gradh over all other
// particles, FP64, sources staged through LDS
__global__ void __launch_bounds__(256) k_direct(KGeom kg, int n, int nt,
                                                const int *__restrict__ targets,
                                                const double *__restrict__ x,
                                                const double *__restrict__ y,
                                                const double *__restrict__ z,
                                                const double *__restrict__ m,
                                                const double *__restrict__ h,
                                                const double *__restrict__ zeta,
                                                const int *__restrict__ cid,
                                                const int *__restrict__ cloud_off,
                                                const double4 *__restrict__ t4,
                                                double *__restrict__ acc, double *__restrict__ pot) {
    __shared__ double sx[256], sy[256], sz[256], sm[256], sh[256], sih[256], szt[256];
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = t < nt;
    const int p = valid ? targets[t] : 0;
    const double rpx = x[p], rpy = y[p], rpz = z[p], invhp = 1.0 / h[p], zo_p = zeta[p];
    // sources: the target's own cloud
    const int lo = cloud_off[cid[p]], hi = cloud_off[cid[p] + 1];
    double gx = 0, gy = 0, gz = 0, gp = 0;
    for (int base = 0; base < n; base += 256) {
        const int j = base + threadIdx.x;
        if (j < n) {
            sx[threadIdx.x] = x[j];
            sy[threadIdx.x] = y[j];
            sz[threadIdx.x] = z[j];
            sm[threadIdx.x] = m[j];
            sh[threadIdx.x] = kg.range * h[j];
            sih[threadIdx.x] = 1.0 / h[j];
            szt[threadIdx.x] = zeta[j];
        }
        __syncthreads();
        const int lim = min(256, n - base);
        if (valid) {
            for (int k = 0; k < lim; k++) {
                if (base + k == p || base + k < lo || base + k >= hi) continue;
                (void)pair_fp64(kg, t4, invhp, zo_p, sx[k] - rpx, sy[k] - rpy, sz[k] - rpz, sm[k],
                                sh[k], sih[k], szt[k], gx, gy, gz, gp);
            }
        }
        __syncthreads();
    }
    if (valid) {
        acc[3 * t] = gx;
        acc[3 * t + 1] = gy;
        acc[3 * t + 2] = gz;
        pot[t] = gp;
    }
}

// Blocks per CU of the walk launch: the knob's value (default: as many as
// fit).  Fewer blocks leave wave slots to the SPH kernels that run beside
// the walk; the device's vector ALUs are the bound both share, and the best
// split depends on the workload (seren_amd/autotune.py times a few
// candidates on it).  A walk with nothing beside it (serenx_grav_forces)
// takes all that fit.
int walk_cap(const Ctx &c) { return c.walk_alone ? 64 : c.walk_queue; }

// Blocks per CU of the pair phase: its own knob when set.  The pair phase
// runs beside the neighbour and hydro kernels, which need less of the
// device than the h solve beside the cell phase.
int pair_cap(const Ctx &c) {
    if (c.walk_alone) return 64;
    return c.pair_blocks > 0 ? c.pair_blocks : c.walk_queue;
}

// Blocks of the walk launch: with the work queue, at most cap per CU and as
// many as the device holds at once (each wave then takes group after
// group); without it, one wave per group.
template <class K>
int walk_blocks(Ctx &c, K kernel, int cap) {
    const int per_wave = (c.gq.ngroups + 3) / 4;
    if (c.walk_queue == 0) return std::max(per_wave, 1);
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, TB, 0) != hipSuccess ||
        per_cu <= 0)
        per_cu = 1;
    per_cu = std::min(per_cu, cap);
    return std::max(1, std::min(per_wave, per_cu * c.ncu));
}

// The lists share pl_entries slots: each of the ngrp groups gets an equal
// share, at most PL_CAP.  One cloud has at most n / 64 + 1 groups and gets
// PL_CAP entries per group; a forest of many small clouds, whose groups walk
// small trees, gets fewer.
NearList near_list(Ctx &c) {
    const size_t share = c.pl_entries / (size_t)std::max(c.ngrp, 1);
    const int stride = (int)std::min<size_t>(PL_CAP, share);
    return NearList{c.pl_leaf, c.pl_mask, c.pl_len, std::min(c.near_cap, stride), stride};
}

// phase 1 of the walk (traversal and cell terms), persistent over the
// queue, on stream s with at most cap blocks per CU
template <int OCT, int MACT>
void launch_cells(Ctx &c, hipStream_t s, int cap) {
    GroupQueue q{c.gq.ctr, c.gq.order, c.gq.ngroups};
    if (c.opt.precision == SERENX_MIXED) {
        auto k = k_grav_cells_mixed<OCT, MACT>;
        hipLaunchKernelGGL(k, dim3(walk_blocks(c, k, cap)), dim3(TB), 0, s, c.n, c.nc, c.gnf,
                           c.x, c.y, c.z, c.cid, c.root_range, c.active, c.gpot, c.agm, c.gax,
                           c.gay, c.gaz, c.gpc, near_list(c), c.icount, c.wstat, q, c.gq.cost,
                           c.grp);
    } else {
        auto k = k_grav_cells_fp64<OCT, MACT>;
        hipLaunchKernelGGL(k, dim3(walk_blocks(c, k, cap)), dim3(TB), 0, s, c.n, c.nc, c.gn,
                           c.x, c.y, c.z, c.cid, c.root_range, c.active, c.gpot, c.agm, c.gax,
                           c.gay, c.gaz, c.gpc, near_list(c), c.icount, c.wstat, q, c.gq.cost,
                           c.grp);
    }
}

// phase 2 (pair terms), persistent over the same queue order (its second
// set of counters), on stream s with at most cap blocks per CU
template <int MACT, bool QUINTIC>
void launch_pairs(Ctx &c, const double4 *rec, hipStream_t s, int cap) {
    GroupQueue q{c.gq.ctr + NXCD, c.gq.order, c.gq.ngroups};
    if (c.opt.precision == SERENX_MIXED) {
        auto k = k_grav_pairs_mixed<MACT, QUINTIC>;
        hipLaunchKernelGGL(k, dim3(walk_blocks(c, k, cap)), dim3(TB), 0, s, c.n, c.nc, c.gnf,
                           reinterpret_cast<const float4 *>(rec), c.x, c.y, c.z, c.h, c.zeta,
                           c.cid, c.root_range, c.active,
                           reinterpret_cast<const float4 *>(c.ktab + KT_G4F_OFF), c.gax, c.gay,
                           c.gaz, c.gpc, c.gpot, c.agm, near_list(c), c.icount, c.grp, c.ngrp, q);
    } else {
        auto k = k_grav_pairs_fp64<MACT, QUINTIC>;
        hipLaunchKernelGGL(k, dim3(walk_blocks(c, k, cap)), dim3(TB), 0, s, c.n, c.nc, c.gn,
                           rec, c.cid, c.root_range, c.active,
                           reinterpret_cast<const double4 *>(c.ktab + KT_G4_OFF), c.gax, c.gay,
                           c.gaz, c.gpc, c.gpot, c.agm, near_list(c), c.icount, c.grp, c.ngrp, q);
    }
}

// sorts the queue order from the costs of the last walk, on the gravity
// stream
int sort_queue(Ctx &c) {
    Ctx::WalkQueue &q = c.gq;
    const int ng = q.ngroups;
    hipLaunchKernelGGL(k_queue_keys, dim3(nblk(ng)), dim3(TB), 0, c.gstream, ng, q.cost,
                       (int)(q.have_cost && c.walk_queue != 0), q.keys, q.vals);
    size_t tb = q.tmp_bytes;
    SX_HIP(rocprim::radix_sort_pairs(q.tmp, tb, q.keys, q.keys2, q.vals, q.order, (unsigned)ng,
                                     0, 35, c.gstream));
    return SERENX_OK;
}

// the queue order of this walk (from the previous walk's costs) and zeroed
// counters, on the gravity stream.  The order is normally sorted at the end
// of the previous walk (grav_pairs), where the sort runs beside the hydro
// kernels; here it is sorted only when that order is missing or stale.
int prepare_queue(Ctx &c) {
    Ctx::WalkQueue &q = c.gq;
    const int ng = c.ngrp;
    if (ng != q.ngroups) {   // another group layout
        q.have_cost = false;
        q.order_ready = false;
    }
    q.ngroups = ng;
    if (c.capturing) {
        // a walk recorded into a graph takes its groups in tree order: the
        // sort library clears its buffers with memset nodes, which graphs
        // here must not hold (they have given wrong results on replays)
        hipLaunchKernelGGL(k_queue_identity, dim3(nblk(ng)), dim3(TB), 0, c.gstream, ng, q.order);
    } else if (!q.order_ready) {
        SX_TRY(sort_queue(c));
    }
    q.order_ready = false;   // used by this walk
    hipLaunchKernelGGL(k_zero_ints, dim3(1), dim3(64), 0, c.gstream, 2 * NXCD, q.ctr);
    SX_HIP(hipGetLastError());
    return SERENX_OK;
}

// the cell phase on the gravity stream: queue prepared, then the walk
template <int OCT, int MACT>
int launch_grav_cells(Ctx &c) {
    SX_TRY(prepare_queue(c));
    launch_cells<OCT, MACT>(c, c.gstream, walk_cap(c));
    return SERENX_OK;
}

int cells_dispatch(Ctx &c) {
    const bool oct = c.opt.multipole != SERENX_QUADRUPOLE;
    switch (c.opt.mac) {
    case SERENX_MAC_GEOMETRIC:
        return oct ? launch_grav_cells<1, SERENX_MAC_GEOMETRIC>(c)
                   : launch_grav_cells<0, SERENX_MAC_GEOMETRIC>(c);
    case SERENX_MAC_GADGET:
        return oct ? launch_grav_cells<1, SERENX_MAC_GADGET>(c)
                   : launch_grav_cells<0, SERENX_MAC_GADGET>(c);
    case SERENX_MAC_EIGEN:
        return oct ? launch_grav_cells<1, SERENX_MAC_EIGEN>(c)
                   : launch_grav_cells<0, SERENX_MAC_EIGEN>(c);
    default:
        return oct ? launch_grav_cells<1, SERENX_MAC_GADGET2>(c)
                   : launch_grav_cells<0, SERENX_MAC_GADGET2>(c);
    }
}

// the pair phase on the gravity stream: source records packed, then the
// walk
template <int MACT>
int launch_grav_pairs(Ctx &c) {
    double4 *rec = reinterpret_cast<double4 *>(c.grec);
    if (c.opt.precision == SERENX_MIXED)
        hipLaunchKernelGGL(k_pack_grav_mixed, dim3(nblk(c.nc)), dim3(TB), 0, c.gstream, c.nc,
                           c.kg.range, c.gnf, c.x, c.y, c.z, c.m, c.h, c.zeta,
                           reinterpret_cast<float4 *>(c.grec));
    else
        hipLaunchKernelGGL(k_pack_grav, dim3(nblk(c.n)), dim3(TB), 0, c.gstream, c.n, c.kg.range,
                           c.x, c.y, c.z, c.m, c.h, c.zeta, rec);
    if (c.kg.range == walk_geom<true>().range)
        launch_pairs<MACT, true>(c, rec, c.gstream, pair_cap(c));
    else
        launch_pairs<MACT, false>(c, rec, c.gstream, pair_cap(c));
    return SERENX_OK;
}

int pairs_dispatch(Ctx &c) {
    switch (c.opt.mac) {
    case SERENX_MAC_GEOMETRIC: return launch_grav_pairs<SERENX_MAC_GEOMETRIC>(c);
    case SERENX_MAC_GADGET: return launch_grav_pairs<SERENX_MAC_GADGET>(c);
    case SERENX_MAC_EIGEN: return launch_grav_pairs<SERENX_MAC_EIGEN>(c);
    default: return launch_grav_pairs<SERENX_MAC_GADGET2>(c);
    }
}

}  // namespace

// The gravity walk runs on the context's gravity stream in two phases.
// grav_launch, queued after the tree is stocked, forks the gravity stream
// and runs phase 1 (traversal and cell terms: they read the tree, the
// positions and the previous |a| and potential) beside the h solve.
// grav_pairs, queued after the h solve, makes the gravity stream wait for
// it and runs phase 2 (pair terms, which need the new h and zeta), beside
// the neighbour lists, thermal and hydro forces.  grav_join adds the walk's
// acceleration to a on the main stream.
int grav_launch(Ctx &c) {
    SX_HIP(hipEventRecord(c.evfork, c.stream));
    SX_HIP(hipStreamWaitEvent(c.gstream, c.evfork, 0));
    SX_HIP(hipEventRecord(c.evg[0], c.gstream));
    SX_TRY(zero_words(c.icount, 2 * CTR_WORDS, c.gstream));
    SX_TRY(cells_dispatch(c));
    SX_HIP(hipGetLastError());
    SX_HIP(hipEventRecord(c.evg[1], c.gstream));
    c.grav_timed = false;
    return SERENX_OK;
}

int grav_pairs(Ctx &c) {
    SX_HIP(hipEventRecord(c.evfork2, c.stream));
    SX_HIP(hipStreamWaitEvent(c.gstream, c.evfork2, 0));
    SX_HIP(hipEventRecord(c.evg[2], c.gstream));
    SX_TRY(pairs_dispatch(c));
    SX_HIP(hipGetLastError());
    SX_HIP(hipEventRecord(c.evg[3], c.gstream));
    c.gq.have_cost = true;
    c.grav_timed = true;
    // the next walk's order, from this walk's costs: it runs beside the
    // main stream's hydro kernels, and the join waits for it, so that no
    // work of the walk outlives the evaluation
    if (!c.capturing && c.gq.ngroups == c.ngrp) {
        SX_TRY(sort_queue(c));
        SX_HIP(hipGetLastError());
        c.gq.order_ready = true;
    }
    SX_HIP(hipEventRecord(c.evjoin, c.gstream));
    return SERENX_OK;
}

int grav_join(Ctx &c) {
    SX_HIP(hipStreamWaitEvent(c.stream, c.evjoin, 0));
    hipLaunchKernelGGL(k_add_grav, dim3(nblk(c.n)), dim3(TB), 0, c.stream, c.n, c.active, c.gax,
                       c.gay, c.gaz, c.ax, c.ay, c.az);
    SX_HIP(hipGetLastError());
    return SERENX_OK;
}

int direct_gravity(Ctx &c, int nt, const int *dtargets, double *dacc, double *dpot) {
    hipLaunchKernelGGL(k_direct, dim3(nblk(nt)), dim3(256), 0, c.stream, c.kg, c.n, nt, dtargets, c.x,
                       c.y, c.z, c.m, c.h, c.zeta, c.cid, c.cloud_off,
                       reinterpret_cast<const double4 *>(c.ktab + KT_G4_OFF), dacc, dpot);
    SX_HIP(hipGetLastError());
    return SERENX_OK;
}

}  // namespace serenx
