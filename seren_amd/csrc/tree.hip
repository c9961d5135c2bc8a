// tree.hip -- Barnes-Hut octree build, particle reordering and stocking on
// gfx950.
//
// The reference builds its octree level by level, comparing each particle
// with the centre of its cell at every level (BHgrav_build.F90:94-291), and
// numbers new cells serially in (parent, octant) order.  Here the same
// comparisons are done per particle up front: following the FP64 centre
// recurrence rc(l+1) = rc(l) +/- cs(l+1), cs(l+1) = cs(l)/2 of
// BHgrav_build.F90:95,121-125,212-228, each particle gets its octant path to
// LMAX = 30 levels as a 90-bit key (levels 1-21 in khi, 22-30 in klo).  One
// radix sort of the keys groups every cell's particles contiguously, in the
// reference's (parent, octant) = BFS order at every level.  Cells are then
// emitted level by level as runs of equal key prefix inside non-leaf parents;
// a run with <= LEAFMAX particles is a leaf.  Within a leaf the particles are
// put back in their previous array order, which is the reference's plist
// order, and the concatenation of the leaves is the reference's tree order
// (BH_reorder_particles.F90:51-78).
//
// Cells are stored in DFS pre-order for the walks (ifopen = c + 1, nextcell =
// first cell after the subtree) and keep their BFS id for export.
#include <cmath>
#include <cstring>
#include <rocprim/rocprim.hpp>

#include "serenx_internal.h"

namespace serenx {

namespace {

constexpr int TB = 256;

inline int nblk(long n, int tb = TB) { return (int)((n + tb - 1) / tb); }

template <class T>
int dalloc(T **p, size_t count) {
    if (*p) {
        (void)hipFree(*p);
        *p = nullptr;
    }
    if (count == 0) count = 1;
    hipError_t e = hipMalloc((void **)p, sizeof(T) * count);
    if (e != hipSuccess) return set_error(SERENX_ENOMEM, "hipMalloc failed in tree buffers");
    return SERENX_OK;
}

// ---- bounding boxes (bounding_box.F90:11-43), one per cloud -------------
constexpr int BB_CHUNK = 8192;

// min / max of the positions of one chunk of one cloud (one block per chunk)
__global__ void __launch_bounds__(256) k_bbox_chunks(const int *__restrict__ chunks,
                                                     const double *__restrict__ x,
                                                     const double *__restrict__ y,
                                                     const double *__restrict__ z,
                                                     double *__restrict__ part) {
    __shared__ double s[6][256];
    const int ch = blockIdx.x, b = chunks[3 * ch + 1], e = chunks[3 * ch + 2];
    double mn0 = BIG_NUMBER, mn1 = BIG_NUMBER, mn2 = BIG_NUMBER;
    double mx0 = -BIG_NUMBER, mx1 = -BIG_NUMBER, mx2 = -BIG_NUMBER;
    for (int i = b + (int)threadIdx.x; i < e; i += blockDim.x) {
        const double a = x[i], bb = y[i], c = z[i];
        mn0 = fmin(mn0, a); mx0 = fmax(mx0, a);
        mn1 = fmin(mn1, bb); mx1 = fmax(mx1, bb);
        mn2 = fmin(mn2, c); mx2 = fmax(mx2, c);
    }
    const int t = threadIdx.x;
    s[0][t] = mn0; s[1][t] = mn1; s[2][t] = mn2;
    s[3][t] = mx0; s[4][t] = mx1; s[5][t] = mx2;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (t < w) {
            for (int k = 0; k < 3; k++) s[k][t] = fmin(s[k][t], s[k][t + w]);
            for (int k = 3; k < 6; k++) s[k][t] = fmax(s[k][t], s[k][t + w]);
        }
        __syncthreads();
    }
    if (t < 6) part[6 * ch + t] = s[t][0];
}

// per cloud (one block): root cell centre 0.5 (rmin + rmax), half-size
// 0.5 max extent (BHgrav_build.F90:65-67); rextent = max extent
// (bounding_box.F90:39)
__global__ void __launch_bounds__(64) k_bbox_clouds(const int *__restrict__ chunk_first,
                                                    const double *__restrict__ part,
                                                    double *__restrict__ root) {
    __shared__ double s[6][64];
    const int k = blockIdx.x, t = threadIdx.x;
    double v[6] = {BIG_NUMBER, BIG_NUMBER, BIG_NUMBER, -BIG_NUMBER, -BIG_NUMBER, -BIG_NUMBER};
    for (int ch = chunk_first[k] + t; ch < chunk_first[k + 1]; ch += 64) {
        for (int q = 0; q < 3; q++) v[q] = fmin(v[q], part[6 * ch + q]);
        for (int q = 3; q < 6; q++) v[q] = fmax(v[q], part[6 * ch + q]);
    }
    for (int q = 0; q < 6; q++) s[q][t] = v[q];
    __syncthreads();
    for (int w = 32; w > 0; w >>= 1) {
        if (t < w) {
            for (int q = 0; q < 3; q++) s[q][t] = fmin(s[q][t], s[q][t + w]);
            for (int q = 3; q < 6; q++) s[q][t] = fmax(s[q][t], s[q][t + w]);
        }
        __syncthreads();
    }
    if (t != 0) return;
    const double mn[3] = {s[0][0], s[1][0], s[2][0]}, mx[3] = {s[3][0], s[4][0], s[5][0]};
    const double e0 = mx[0] - mn[0], e1 = mx[1] - mn[1], e2 = mx[2] - mn[2];
    double ext = e0;
    if (e1 > ext) ext = e1;
    if (e2 > ext) ext = e2;
    double *r = root + 11 * k;
    for (int q = 0; q < 3; q++) r[q] = 0.5 * (mn[q] + mx[q]);
    r[3] = 0.5 * ext;
    r[4] = ext;
    for (int q = 0; q < 3; q++) {
        r[5 + q] = mn[q];
        r[8 + q] = mx[q];
    }
}

__global__ void k_fill_cid(int ncloud, const int *__restrict__ cloud_off, int *__restrict__ cid) {
    const int k = blockIdx.y;
    if (k >= ncloud) return;
    for (int i = cloud_off[k] + blockIdx.x * blockDim.x + threadIdx.x; i < cloud_off[k + 1];
         i += gridDim.x * blockDim.x)
        cid[i] = k;
}

// ---- octant keys ---------------------------------------------------------
__global__ void __launch_bounds__(256) k_keys(int n, const double *__restrict__ x,
                                              const double *__restrict__ y,
                                              const double *__restrict__ z,
                                              const double *__restrict__ roots,
                                              const int *__restrict__ cid,
                                              uint64_t *__restrict__ khi,
                                              uint64_t *__restrict__ klo,
                                              int *__restrict__ idx) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const double px = x[i], py = y[i], pz = z[i];
    const double *root = roots + 11 * cid[i];
    double cx = root[0], cy = root[1], cz = root[2], cs = root[3];
    uint64_t hi = 0, lo = 0;
#pragma unroll
    for (int l = 1; l <= LMAX; l++) {
        cs = 0.5 * cs;
        const bool bx = px > cx, by = py > cy, bz = pz > cz;
        const uint64_t d = (uint64_t)bx | ((uint64_t)by << 1) | ((uint64_t)bz << 2);
        cx = bx ? cx + cs : cx - cs;
        cy = by ? cy + cs : cy - cs;
        cz = bz ? cz + cs : cz - cs;
        if (l <= 21)
            hi |= d << (3 * (21 - l));
        else
            lo |= d << (3 * (30 - l));
    }
    khi[i] = hi;
    klo[i] = lo;
    idx[i] = i;
}

__global__ void k_gather_u64(int n, const int *__restrict__ idx, const uint64_t *__restrict__ src,
                             uint64_t *__restrict__ dst) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[idx[i]];
}

// the cloud of every element of a key-sorted order (vals: particle indices)
// and its position, keys of the stable sort by cloud that follows
__global__ void k_cloud_keys(int n, const int *__restrict__ vals, const int *__restrict__ cid,
                             int *__restrict__ ck, int *__restrict__ pos) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    ck[i] = cid[vals[i]];
    pos[i] = i;
}

// keys and values reordered by perm, into k2/v2
__global__ void k_permute_kv(int n, const int *__restrict__ perm, const uint64_t *__restrict__ k,
                             const int *__restrict__ v, uint64_t *__restrict__ k2,
                             int *__restrict__ v2) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int j = perm[i];
    k2[i] = k[j];
    v2[i] = v[j];
}

__global__ void k_copy_kv(int n, const uint64_t *__restrict__ k2, const int *__restrict__ v2,
                          uint64_t *__restrict__ k, int *__restrict__ v) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    k[i] = k2[i];
    v[i] = v2[i];
}

// any run of > LEAFMAX equal 21-level keys in one cloud needs the deeper 9
// levels sorted
__global__ void k_deep_check(int n, const uint64_t *__restrict__ khi, const int *__restrict__ cid,
                             int *__restrict__ flag) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i + LEAFMAX < n && khi[i] == khi[i + LEAFMAX] && cid[i] == cid[i + LEAFMAX]) *flag = 1;
}

// common key prefix of sorted neighbours i-1, i in levels (0..30); zero at
// the first particle of every cloud (its root starts there)
__global__ void k_cpl(int n, const uint64_t *__restrict__ khi, const uint64_t *__restrict__ klo,
                      const int *__restrict__ cid, uint8_t *__restrict__ cpl) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (i == 0 || cid[i] != cid[i - 1]) {
        cpl[i] = 0;
        return;
    }
    uint64_t a = khi[i - 1] ^ khi[i];
    int lv;
    if (a) {
        lv = (__clzll(a) - 1) / 3;       // bits 62..0 hold levels 1..21
    } else {
        uint64_t b = klo[i - 1] ^ klo[i];
        lv = b ? 21 + (__clzll(b) - 37) / 3 : 30;  // bits 26..0 hold levels 22..30
    }
    cpl[i] = (uint8_t)lv;
}

// level 0: one root cell per cloud, cell id = cloud id
__global__ void k_init_levels(int n, int ncloud, const int *__restrict__ cloud_off,
                              const int *__restrict__ cid, uint8_t *pact, int *cellof, int *ncs,
                              int8_t *firstlev, int *c_start, int *c_end, int *c_parent,
                              int *c_leaf, int *c_level) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i <= n) {
        const bool first = (i < n) && (i == 0 || cid[i] != cid[i - 1]);
        ncs[i] = first ? 1 : 0;
        firstlev[i] = first ? 0 : -1;
    }
    if (i < n) {
        pact[i] = 1;
        cellof[i] = cid[i];
    }
    if (i < ncloud) {
        c_start[i] = cloud_off[i];
        c_end[i] = cloud_off[i + 1];
        c_parent[i] = -1;
        c_leaf[i] = 0;  // a root is always divided (BHgrav_build.F90:74)
        c_level[i] = 0;
    }
}

// DFS range of each cloud's tree, [root, root.next)
__global__ void k_root_range(int ncloud, const int *__restrict__ bfs2dfs,
                             const GNode *__restrict__ gn, int2 *__restrict__ rr) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= ncloud) return;
    const int d = bfs2dfs[k];
    rr[k] = make_int2(d, gn[d].next);
}

// child-start flags at level L: an active particle starts a cell when its
// L-prefix differs from its predecessor's
__global__ void k_mark(int n, int L, const uint8_t *__restrict__ pact,
                       const uint8_t *__restrict__ cpl, int *__restrict__ f) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    f[i] = (pact[i] && cpl[i] < L) ? 1 : 0;
}

__global__ void k_emit(int n, int L, int base, const uint8_t *__restrict__ pact,
                       const int *__restrict__ f, const int *__restrict__ incl,
                       int *__restrict__ cellof, int *__restrict__ c_start,
                       int *__restrict__ c_end, int *__restrict__ c_parent,
                       int *__restrict__ c_level, int *__restrict__ ncs,
                       int8_t *__restrict__ firstlev) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !pact[i]) return;
    const int c = base + incl[i] - 1;
    if (f[i]) {
        c_start[c] = i;
        c_parent[c] = cellof[i];
        c_level[c] = L;
        if (ncs[i] == 0) firstlev[i] = (int8_t)L;
        ncs[i] += 1;
    }
    cellof[i] = c;
    // the last particle of a run closes the cell (single writer)
    if (i + 1 == n || !pact[i + 1] || f[i + 1]) c_end[c] = i + 1;
}

__global__ void k_leafcells(int base, int cnt, const int *__restrict__ c_start,
                            const int *__restrict__ c_end, int *__restrict__ c_leaf) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= cnt) return;
    const int c = base + j;
    const int k = c_end[c] - c_start[c];
    c_leaf[c] = (k <= LEAFMAX) ? k : 0;
}

__global__ void k_deactivate(int n, uint8_t *__restrict__ pact, const int *__restrict__ cellof,
                             const int *__restrict__ c_leaf) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && pact[i] && c_leaf[cellof[i]] > 0) pact[i] = 0;
}

// DFS pre-order index of every cell: cells starting at sorted position s form
// a chain of consecutive levels, so dfs = off[s] + (L - firstlev[s]); the
// cell after a subtree is the first cell starting at its end.
__global__ void k_dfs(int n, int nc, const int *__restrict__ c_start, const int *__restrict__ c_end,
                      const int *__restrict__ c_level, const int *__restrict__ c_leaf,
                      const int *__restrict__ off, const int8_t *__restrict__ firstlev,
                      int *__restrict__ bfs2dfs, int *__restrict__ dfs2bfs, GNode *__restrict__ gn,
                      HNode *__restrict__ hn) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    const int s = c_start[c], e = c_end[c], L = c_level[c];
    const int d = off[s] + (L - firstlev[s]);
    bfs2dfs[c] = d;
    dfs2bfs[d] = c;
    const int nxt = (e < n) ? off[e] : nc;
    gn[d].next = nxt;
    gn[d].leaf = c_leaf[c];
    gn[d].first = s;
    gn[d].count = e - s;
    hn[d].next = nxt;
    hn[d].leaf = c_leaf[c];
    hn[d].first = s;
}

__global__ void k_dfs_parent(int nc, const int *__restrict__ c_parent, const int *__restrict__ bfs2dfs,
                             int *__restrict__ dfs_parent) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    const int p = c_parent[c];
    dfs_parent[bfs2dfs[c]] = (p >= 0) ? bfs2dfs[p] : -1;
}

// leaf particles back into their previous array order (the plist order of
// BHgrav_build.F90:195-204, which follows the linked list in array order)
__global__ void k_leafsort(int nc, const int *__restrict__ c_start, const int *__restrict__ c_leaf,
                           int *__restrict__ idx) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    const int k = c_leaf[c];
    if (k <= 1) return;
    int v[LEAFMAX];
    const int s = c_start[c];
    for (int i = 0; i < k; i++) v[i] = idx[s + i];
    for (int i = 1; i < k; i++) {
        int t = v[i], j = i - 1;
        while (j >= 0 && v[j] > t) {
            v[j + 1] = v[j];
            j--;
        }
        v[j + 1] = t;
    }
    for (int i = 0; i < k; i++) idx[s + i] = v[i];
}

template <class T>
__global__ void k_gather(int n, const int *__restrict__ perm, const T *__restrict__ src,
                         T *__restrict__ dst) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[perm[i]];
}

// Exact parallel-axis shift of a child's traceless octupole to the parent's
// COM (multipole SERENX_OCTUPOLE_EXACT).  With d the child's COM offset, Q'
// its traceless quadrupole and m its mass,
//   O_ijk = O'_ijk + m [15 d_i d_j d_k - 3 d^2 (d_i delta_jk + 2 perms)]
//         + 5 (d_i Q'_jk + 2 perms) - 2 ((Q'd)_i delta_jk + 2 perms).
// s[] holds the reference's units (BHgrav_stock.F90:194-206): s0 = O_xxx/3,
// s1 = O_xxy, s2 = O_xyy, s3 = O_yyy/3, s4 = O_xzz, s5 = O_yzz, s6 = O_zzz/3,
// s7 = O_xxz, s8 = O_yyz, and a stored cell carries s9 = 2 O_xyz (the walk's
// expansion has coefficient 1 on s9 xyz where the series has 2 O_xyz/3).
// The reference's branch shift (:263-297) halves the quadrupole terms and
// takes 25 for 5 on the xyz term; SERENX_OCTUPOLE keeps that form.
__device__ __forceinline__ void octupole_shift_exact(double *sc, const double *cs,
                                                     const double *q, double mcc,
                                                     const double *dr, double drsqd) {
    const double qzz = -(q[0] + q[2]);
    const double qd0 = q[0] * dr[0] + q[1] * dr[1] + q[3] * dr[2];
    const double qd1 = q[1] * dr[0] + q[2] * dr[1] + q[4] * dr[2];
    const double qd2 = q[3] * dr[0] + q[4] * dr[1] + qzz * dr[2];
    for (int d = 0; d < 9; d++) sc[d] = sc[d] + cs[d];
    sc[9] = sc[9] + 0.5 * cs[9];
    sc[0] = sc[0] + mcc * dr[0] * (5.0 * dr[0] * dr[0] - 3.0 * drsqd) +
            (5.0 * dr[0] * q[0] - 2.0 * qd0);
    sc[1] = sc[1] + mcc * dr[1] * (15.0 * dr[0] * dr[0] - 3.0 * drsqd) +
            (5.0 * (2.0 * dr[0] * q[1] + dr[1] * q[0]) - 2.0 * qd1);
    sc[2] = sc[2] + mcc * dr[0] * (15.0 * dr[1] * dr[1] - 3.0 * drsqd) +
            (5.0 * (2.0 * dr[1] * q[1] + dr[0] * q[2]) - 2.0 * qd0);
    sc[3] = sc[3] + mcc * dr[1] * (5.0 * dr[1] * dr[1] - 3.0 * drsqd) +
            (5.0 * dr[1] * q[2] - 2.0 * qd1);
    sc[4] = sc[4] + mcc * dr[0] * (15.0 * dr[2] * dr[2] - 3.0 * drsqd) +
            (5.0 * (2.0 * dr[2] * q[3] + dr[0] * qzz) - 2.0 * qd0);
    sc[5] = sc[5] + mcc * dr[1] * (15.0 * dr[2] * dr[2] - 3.0 * drsqd) +
            (5.0 * (2.0 * dr[2] * q[4] + dr[1] * qzz) - 2.0 * qd1);
    sc[6] = sc[6] + mcc * dr[2] * (5.0 * dr[2] * dr[2] - 3.0 * drsqd) +
            (5.0 * dr[2] * qzz - 2.0 * qd2);
    sc[7] = sc[7] + mcc * dr[2] * (15.0 * dr[0] * dr[0] - 3.0 * drsqd) +
            (5.0 * (2.0 * dr[0] * q[3] + dr[2] * q[0]) - 2.0 * qd2);
    sc[8] = sc[8] + mcc * dr[2] * (15.0 * dr[1] * dr[1] - 3.0 * drsqd) +
            (5.0 * (2.0 * dr[1] * q[4] + dr[2] * q[2]) - 2.0 * qd2);
    sc[9] = sc[9] + 15.0 * mcc * dr[0] * dr[1] * dr[2] +
            5.0 * (dr[0] * q[4] + dr[1] * q[3] + dr[2] * q[1]);
}

// ---- stocking: BHgrav_stock.F90:53-354 and BHhydro_stock.F90:40-159 -------
// One thread per cell of one level, deepest level first.  Arithmetic follows
// the reference's operation order (built without FMA contraction) so that
// moments, COM and opening distances are bit-identical.
__global__ void __launch_bounds__(256) k_stock(
    int nc, int base, int cnt, const int *__restrict__ bfs2dfs, GNode *__restrict__ gn,
    double *__restrict__ gs,
    HNode *__restrict__ hn, double *__restrict__ bbmin, double *__restrict__ bbmax,
    double *__restrict__ hmaxs, const double *__restrict__ x, const double *__restrict__ y,
    const double *__restrict__ z, const double *__restrict__ m, const double *__restrict__ h,
    int octupole, int mac, double thetamaxsqd, double abserror, double daux, double kernrange) {
    int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= cnt) return;
    const int c = bfs2dfs[base + j];
    GNode &g = gn[c];
    const int leaf = g.leaf, first = g.first, next = g.next;
    double mc = 0.0, rc[3] = {0.0, 0.0, 0.0}, dr[3], drsqd;
    double bmn[3] = {BIG_NUMBER, BIG_NUMBER, BIG_NUMBER};
    double bmx[3] = {-BIG_NUMBER, -BIG_NUMBER, -BIG_NUMBER};
    double hmax = 0.0;
    double qc[5] = {0, 0, 0, 0, 0};
    double sc[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    double gm, gr[3];
    if (leaf == 1) {
        const int p = first;
        rc[0] = x[p]; rc[1] = y[p]; rc[2] = z[p];
        gr[0] = rc[0]; gr[1] = rc[1]; gr[2] = rc[2];
        gm = m[p];
        hmax = h[p];
        for (int k = 0; k < 3; k++) bmn[k] = bmx[k] = rc[k];
    } else if (leaf > 0) {
        for (int i = 0; i < leaf; i++) {
            const int p = first + i;
            const double pr[3] = {x[p], y[p], z[p]};
            const double mp = m[p];
            mc = mc + mp;
            hmax = fmax(hmax, h[p]);
            for (int k = 0; k < 3; k++) rc[k] = rc[k] + mp * pr[k];
            for (int k = 0; k < 3; k++) {
                bmx[k] = fmax(bmx[k], pr[k]);
                bmn[k] = fmin(bmn[k], pr[k]);
            }
        }
        for (int k = 0; k < 3; k++) rc[k] = rc[k] / mc;
        for (int i = 0; i < leaf; i++) {
            const int p = first + i;
            dr[0] = x[p] - rc[0]; dr[1] = y[p] - rc[1]; dr[2] = z[p] - rc[2];
            const double mp = m[p];
            drsqd = dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2];
            qc[0] = qc[0] + mp * (3.0 * dr[0] * dr[0] - drsqd);
            qc[1] = qc[1] + mp * (3.0 * dr[0] * dr[1]);
            qc[2] = qc[2] + mp * (3.0 * dr[1] * dr[1] - drsqd);
            qc[3] = qc[3] + mp * (3.0 * dr[2] * dr[0]);
            qc[4] = qc[4] + mp * (3.0 * dr[2] * dr[1]);
            if (octupole) {
                sc[0] = sc[0] + mp * dr[0] * (5.0 * dr[0] * dr[0] - 3.0 * drsqd);
                sc[1] = sc[1] + mp * dr[1] * (15.0 * dr[0] * dr[0] - 3.0 * drsqd);
                sc[2] = sc[2] + mp * dr[0] * (15.0 * dr[1] * dr[1] - 3.0 * drsqd);
                sc[3] = sc[3] + mp * dr[1] * (5.0 * dr[1] * dr[1] - 3.0 * drsqd);
                sc[4] = sc[4] + mp * dr[0] * (15.0 * dr[2] * dr[2] - 3.0 * drsqd);
                sc[5] = sc[5] + mp * dr[1] * (15.0 * dr[2] * dr[2] - 3.0 * drsqd);
                sc[6] = sc[6] + mp * dr[2] * (5.0 * dr[2] * dr[2] - 3.0 * drsqd);
                sc[7] = sc[7] + mp * dr[2] * (15.0 * dr[0] * dr[0] - 3.0 * drsqd);
                sc[8] = sc[8] + mp * dr[2] * (15.0 * dr[1] * dr[1] - 3.0 * drsqd);
                sc[9] = sc[9] + 15.0 * mp * dr[0] * dr[1] * dr[2];
            }
        }
        gm = mc;
        gr[0] = rc[0]; gr[1] = rc[1]; gr[2] = rc[2];
    } else {
        // branch: children are c+1 and the chain of next pointers up to next
        for (int cc = c + 1; cc != next && cc < nc; cc = gn[cc].next) {
            const GNode &k = gn[cc];
            mc = mc + k.m;
            for (int d = 0; d < 3; d++) rc[d] = rc[d] + k.m * k.r[d];
            hmax = fmax(hmaxs[cc], hmax);
            for (int d = 0; d < 3; d++) {
                bmx[d] = fmax(bbmax[3 * cc + d], bmx[d]);
                bmn[d] = fmin(bbmin[3 * cc + d], bmn[d]);
            }
        }
        if (mc > 0.0)
            for (int d = 0; d < 3; d++) rc[d] = rc[d] / mc;
        for (int cc = c + 1; cc != next && cc < nc; cc = gn[cc].next) {
            const GNode &k = gn[cc];
            const double mcc = k.m;
            const double *q = k.q;
            for (int d = 0; d < 5; d++) qc[d] = qc[d] + q[d];
            for (int d = 0; d < 3; d++) dr[d] = k.r[d] - rc[d];
            drsqd = dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2];
            qc[0] = qc[0] + mcc * (3.0 * dr[0] * dr[0] - drsqd);
            qc[1] = qc[1] + mcc * (3.0 * dr[0] * dr[1]);
            qc[2] = qc[2] + mcc * (3.0 * dr[1] * dr[1] - drsqd);
            qc[3] = qc[3] + mcc * (3.0 * dr[2] * dr[0]);
            qc[4] = qc[4] + mcc * (3.0 * dr[2] * dr[1]);
            const double *ks = gs + 10 * (size_t)cc;
            if (octupole == 2) {
                octupole_shift_exact(sc, ks, q, mcc, dr, drsqd);
            } else if (octupole) {
                // parallel-axis shift exactly as BHgrav_stock.F90:263-297
                for (int d = 0; d < 10; d++) sc[d] = sc[d] + ks[d];
                sc[0] = sc[0] + mcc * dr[0] * (5.0 * dr[0] * dr[0] - 3.0 * drsqd) +
                        1.5 * dr[0] * q[0] - dr[1] * q[1] - dr[2] * q[3];
                sc[1] = sc[1] + mcc * dr[1] * (15.0 * dr[0] * dr[0] - 3.0 * drsqd) +
                        4.0 * dr[0] * q[1] + 2.5 * dr[1] * q[0] - dr[1] * q[2] - dr[2] * q[4];
                sc[2] = sc[2] + mcc * dr[0] * (15.0 * dr[1] * dr[1] - 3.0 * drsqd) +
                        4.0 * dr[1] * q[1] + 2.5 * dr[0] * q[2] - dr[0] * q[0] - dr[2] * q[3];
                sc[3] = sc[3] + mcc * dr[1] * (5.0 * dr[1] * dr[1] - 3.0 * drsqd) +
                        1.5 * dr[1] * q[2] - dr[0] * q[1] - dr[2] * q[4];
                sc[4] = sc[4] + mcc * dr[0] * (15.0 * dr[2] * dr[2] - 3.0 * drsqd) +
                        4.0 * dr[2] * q[3] - 2.5 * dr[0] * (q[0] + q[2]) - dr[0] * q[0] -
                        dr[1] * q[1];
                sc[5] = sc[5] + mcc * dr[1] * (15.0 * dr[2] * dr[2] - 3.0 * drsqd) +
                        4.0 * dr[2] * q[4] - 2.5 * dr[1] * (q[0] + q[2]) - dr[0] * q[1] -
                        dr[1] * q[2];
                sc[6] = sc[6] + mcc * dr[2] * (5.0 * dr[2] * dr[2] - 3.0 * drsqd) -
                        1.5 * dr[2] * (q[0] + q[2]) - dr[0] * q[3] - dr[1] * q[4];
                sc[7] = sc[7] + mcc * dr[2] * (15.0 * dr[0] * dr[0] - 3.0 * drsqd) +
                        4.0 * dr[0] * q[3] + 2.5 * dr[2] * q[0] - dr[1] * q[4] +
                        dr[2] * (q[0] + q[2]);
                sc[8] = sc[8] + mcc * dr[2] * (15.0 * dr[1] * dr[1] - 3.0 * drsqd) +
                        4.0 * dr[1] * q[4] + 2.5 * dr[2] * q[2] - dr[0] * q[3] +
                        dr[2] * (q[0] + q[2]);
                sc[9] = sc[9] + 15.0 * mcc * dr[0] * dr[1] * dr[2] +
                        25.0 * (dr[0] * q[4] + dr[1] * q[3] + dr[2] * q[1]);
            }
        }
        gm = mc;
        gr[0] = rc[0]; gr[1] = rc[1]; gr[2] = rc[2];
    }
    // MAC quantities (BHgrav_stock.F90:317-348); rc is the COM, or the
    // particle position for a one-particle leaf
    for (int k = 0; k < 3; k++) dr[k] = fmax(bmx[k] - rc[k], rc[k] - bmn[k]);
    const double diagonalsqd = dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2];
    const double rsize = sqrt(diagonalsqd) + kernrange * hmax;
    double dminsqd, macv = 0.0;
    if (mac == SERENX_MAC_GEOMETRIC) {
        dminsqd = (thetamaxsqd > 0.0) ? diagonalsqd / thetamaxsqd : BIG_NUMBER;
    } else if (mac == SERENX_MAC_GADGET) {
        dminsqd = diagonalsqd;
        macv = pow(mc * diagonalsqd * diagonalsqd / abserror, ONETHIRD);
    } else if (mac == SERENX_MAC_GADGET2) {
        dminsqd = diagonalsqd;
        macv = sqrt(mc * diagonalsqd / abserror);
    } else {
        // eigenvalue_mac.F90: largest eigenvalue of the traceless quadrupole
        dminsqd = diagonalsqd;
        const double pq = qc[0] * qc[2] - (qc[0] + qc[2]) * (qc[0] + qc[2]) - qc[1] * qc[1] -
                          qc[3] * qc[3] - qc[4] * qc[4];
        if (pq >= 0.0) {
            macv = 0.0;
        } else {
            const double lambda = 2.0 * sqrt(-pq / 3.0);
            macv = pow(0.5 * lambda / abserror, TWOTHIRDS_DP);
        }
    }
    dminsqd = fmax(dminsqd, rsize * rsize);
    if (dminsqd < daux) dminsqd = daux;

    for (int k = 0; k < 3; k++) g.r[k] = gr[k];
    g.m = gm;
    g.dminsqd = dminsqd;
    g.mac = macv;
    for (int k = 0; k < 5; k++) g.q[k] = qc[k];
    g.q[5] = -(qc[0] + qc[2]);
    double sv[10];
    for (int k = 0; k < 10; k++) sv[k] = octupole ? sc[k] : 0.0;
    if (octupole == 2) sv[9] = 2.0 * sc[9];   // the walk's xyz coefficient is 1
    for (int k = 0; k < 10; k++) gs[10 * (size_t)c + k] = sv[k];
    // H = grad(S) / 2 with S = s0 x^3 + s1 x^2y + s2 y^2x + s3 y^3 + s4 z^2x
    //   + s5 z^2y + s6 z^3 + s7 x^2z + s8 y^2z + s9 xyz, on (xx xy xz yy zz yz)
    const double oc[18] = {1.5 * sv[0], sv[1], sv[7], 0.5 * sv[2], 0.5 * sv[4], 0.5 * sv[9],
                           0.5 * sv[1], sv[2], 0.5 * sv[9], 1.5 * sv[3], 0.5 * sv[5], sv[8],
                           0.5 * sv[7], 0.5 * sv[9], sv[4], 0.5 * sv[8], 1.5 * sv[6], sv[5]};
    for (int k = 0; k < 18; k++) g.oc[k] = oc[k];
    for (int k = 0; k < 3; k++) {
        bbmin[3 * c + k] = bmn[k];
        bbmax[3 * c + k] = bmx[k];
    }
    hmaxs[c] = hmax;

    // hydro node: bounding-box centre and half-diagonal
    HNode &hh = hn[c];
    if (leaf == 1) {
        for (int k = 0; k < 3; k++) hh.r[k] = bmn[k];
        hh.rmax = 0.0;
        for (int k = 0; k < 3; k++) hh.half[k] = (float)(1.0e-9 * fabs(bmn[k]) + 1.0e-300);
    } else {
        double hc[3], half[3], dsq = 0.0;
        for (int k = 0; k < 3; k++) {
            hc[k] = 0.5 * (bmx[k] + bmn[k]);
            half[k] = bmx[k] - hc[k];
            dsq = dsq + half[k] * half[k];
        }
        for (int k = 0; k < 3; k++) hh.r[k] = hc[k];
        hh.rmax = sqrt(dsq);
        // the walks' box test: half-extents grown by 1e-9 of the coordinates
        // and 1e-6 relative (covering the rounding of the centre, the
        // subtraction and the FP32 conversion)
        for (int k = 0; k < 3; k++)
            hh.half[k] = (float)((half[k] + 1.0e-9 * (fabs(hc[k]) + half[k])) * (1.0 + 1.0e-6));
    }
}

// macscale multiplies the multipole-MAC term for the mixed walk's own
// tolerance (1 for the options' abserror)
__global__ void k_fill_f32_nodes(int nc, const GNode *__restrict__ gn,
                                 const double *__restrict__ gs, double macscale,
                                 GNodeF *__restrict__ gf) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    const GNode &g = gn[c];
    GNodeF &f = gf[c];
    for (int k = 0; k < 3; k++) f.r[k] = g.r[k];
    f.dminsqd = g.dminsqd;
    f.mac = g.mac * macscale;
    f.m = (float)g.m;
    for (int k = 0; k < 6; k++) f.q[k] = (float)g.q[k];
    for (int k = 0; k < 10; k++) f.s[k] = (float)gs[10 * (size_t)c + k];
    f.next = g.next;
    f.leaf = g.leaf;
    f.first = g.first;
    f.pad = 0;
}

// ---- BHhydro_hguess.F90:108-155 -----------------------------------------
__global__ void k_hguess(int nc, int thresh, double invkernrange, const int *__restrict__ c_leaf,
                         const int *__restrict__ c_start, const int *__restrict__ c_end,
                         const int *__restrict__ c_parent, const int *__restrict__ c_level,
                         const int *__restrict__ bfs2dfs, const HNode *__restrict__ hn,
                         double *__restrict__ h) {
    int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    const int k = c_leaf[c];
    if (k <= 0) return;
    int a = c, l = c_level[c];
    for (;;) {
        if ((c_end[a] - c_start[a]) > thresh || l == 0) break;
        a = c_parent[a];
        l--;
    }
    const double hg = invkernrange * hn[bfs2dfs[a]].rmax;
    const int s = c_start[c];
    for (int i = 0; i < k; i++) h[s + i] = hg;
}


// BHhydro_update_hmax.F90:29-88 in one pass.  Branches start at zero; one
// thread per leaf writes the leaf's maximum of KERNRANGE h and raises its
// ancestors to it with 64-bit atomic maxima (non-negative doubles order as
// their bit patterns), stopping at the first ancestor that already holds as
// much: the thread that put that value there carries it further up.  Every
// cell ends at the maximum over its subtree, as a pass over the levels from
// the deepest up gives it: a maximum does not depend on the order.
__global__ void k_hmax_zero(int nc, HNode *__restrict__ hn) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < nc && hn[c].leaf == 0) hn[c].hrangemax = 0.0;
}

__global__ void k_hmax_up(int nc, HNode *__restrict__ hn, const int *__restrict__ dfs_parent,
                          const double *__restrict__ h, double kernrange) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    const int leaf = hn[c].leaf, first = hn[c].first;
    if (leaf <= 0) return;
    double hr = 0.0;
    if (leaf == 1) {
        hr = kernrange * h[first];
    } else {
        for (int i = 0; i < leaf; i++) hr = fmax(hr, kernrange * h[first + i]);
    }
    hn[c].hrangemax = hr;
    const unsigned long long v = (unsigned long long)__double_as_longlong(hr);
    for (int p = dfs_parent[c]; p >= 0; p = dfs_parent[p]) {
        unsigned long long *slot = reinterpret_cast<unsigned long long *>(&hn[p].hrangemax);
        if (atomicMax(slot, v) >= v) break;
    }
}

}  // namespace

// --------------------------------------------------------------------------
int ensure_tree_buffers(Ctx &c) {
    const int n = c.cap;
    const size_t nn = (size_t)n + 1;
    SX_TRY(dalloc(&c.khi, nn));
    SX_TRY(dalloc(&c.klo, nn));
    SX_TRY(dalloc(&c.khi2, nn));
    SX_TRY(dalloc(&c.klo2, nn));
    SX_TRY(dalloc(&c.idx, nn));
    SX_TRY(dalloc(&c.idx2, nn));
    SX_TRY(dalloc(&c.cpl, nn));
    SX_TRY(dalloc(&c.pact, nn));
    SX_TRY(dalloc(&c.flag, nn));
    SX_TRY(dalloc(&c.incl, nn));
    SX_TRY(dalloc(&c.cellof, nn));
    SX_TRY(dalloc(&c.ncs, nn));
    SX_TRY(dalloc(&c.off, nn));
    SX_TRY(dalloc(&c.firstlev, nn));
    // a forest holds at most one cloud per pp_gather particles (set_clouds)
    c.max_clouds = forest_max_clouds(n, c.pp_gather);
    c.max_chunks = n / BB_CHUNK + c.max_clouds + 1;
    SX_TRY(dalloc(&c.root, 11 * (size_t)c.max_clouds + 16));
    SX_TRY(dalloc(&c.root_range, (size_t)c.max_clouds));
    SX_TRY(dalloc(&c.cloud_off, (size_t)c.max_clouds + 1));
    SX_TRY(dalloc(&c.chunk_first, (size_t)c.max_clouds + 1));
    SX_TRY(dalloc(&c.chunks, 3 * (size_t)c.max_chunks));
    SX_TRY(dalloc(&c.bbpart, 6 * (size_t)c.max_chunks));
    SX_TRY(dalloc(&c.cid, nn));
    SX_TRY(dalloc(&c.fs_ck, nn));
    SX_TRY(dalloc(&c.fs_ck2, nn));
    SX_TRY(dalloc(&c.fs_pos, nn));
    SX_TRY(dalloc(&c.fs_perm, nn));
    SX_TRY(dalloc(&c.fs_v, nn));
    SX_TRY(dalloc(&c.fs_k, nn));
    c.ccap = 2 * n + 2 + c.max_clouds;
    const size_t cc = (size_t)c.ccap;
    SX_TRY(dalloc(&c.c_start, cc));
    SX_TRY(dalloc(&c.c_end, cc));
    SX_TRY(dalloc(&c.c_parent, cc));
    SX_TRY(dalloc(&c.c_leaf, cc));
    SX_TRY(dalloc(&c.c_level, cc));
    SX_TRY(dalloc(&c.bfs2dfs, cc));
    SX_TRY(dalloc(&c.dfs2bfs, cc));
    SX_TRY(dalloc(&c.dfs_parent, cc));
    SX_TRY(dalloc(&c.bbmin, 3 * cc));
    SX_TRY(dalloc(&c.bbmax, 3 * cc));
    SX_TRY(dalloc(&c.hmaxs, cc));
    SX_TRY(dalloc(&c.gn, cc));
    SX_TRY(dalloc(&c.gs, cc * 10));
    SX_TRY(dalloc(&c.gnf, cc));
    SX_TRY(dalloc(&c.hn, cc));
    // temporary storage for sort and scans
    size_t b1 = 0, b2 = 0, b3 = 0, b4 = 0;
    SX_HIP(rocprim::radix_sort_pairs(nullptr, b1, c.khi, c.khi2, c.idx, c.idx2, (unsigned)n, 0,
                                     63, c.stream));
    SX_HIP(rocprim::radix_sort_pairs(nullptr, b4, c.fs_ck, c.fs_ck2, c.fs_pos, c.fs_perm,
                                     (unsigned)n, 0, 16, c.stream));
    SX_HIP(rocprim::inclusive_scan(nullptr, b2, c.flag, c.incl, (size_t)n, rocprim::plus<int>(),
                                   c.stream));
    SX_HIP(rocprim::exclusive_scan(nullptr, b3, c.ncs, c.off, 0, (size_t)n + 1,
                                   rocprim::plus<int>(), c.stream));
    size_t need = std::max(std::max(b1, b4), std::max(b2, b3));
    if (c.sorttmp) (void)hipFree(c.sorttmp);
    c.sorttmp = nullptr;
    if (hipMalloc(&c.sorttmp, need + 256) != hipSuccess)
        return set_error(SERENX_ENOMEM, "hipMalloc failed for sort scratch");
    c.sorttmp_bytes = need + 256;
    // pinned readback slots start defined: slot 8 is the h-solve error flag
    // that every sph_update path writes before sph_neighbours reads it
    if (!c.hostpin) {
        SX_HIP(hipHostMalloc((void **)&c.hostpin, 64 * sizeof(int)));
        std::memset(c.hostpin, 0, 64 * sizeof(int));
    }
    if (!c.hostpin_d) {
        SX_HIP(hipHostMalloc((void **)&c.hostpin_d, 16 * sizeof(double)));
        std::memset(c.hostpin_d, 0, 16 * sizeof(double));
    }
    return SERENX_OK;
}

int permute_particles(Ctx &c, const int *perm) {
    const int n = c.n;
    // (u and du_dt: reorder_array.F90:98-104)
    double **arrs[] = {&c.x, &c.y, &c.z, &c.m, &c.h, &c.zeta, &c.vx, &c.vy, &c.vz,
                       &c.rho, &c.omega, &c.press, &c.sound, &c.divv, &c.drhodt,
                       &c.ax, &c.ay, &c.az, &c.gpot, &c.agm, &c.u, &c.dudt};
    for (double **a : arrs) {
        hipLaunchKernelGGL(k_gather<double>, dim3(nblk(n)), dim3(TB), 0, c.stream, n, perm,
                           (const double *)*a, c.tmpd);
        std::swap(*a, c.tmpd);
    }
    hipLaunchKernelGGL(k_gather<int>, dim3(nblk(n)), dim3(TB), 0, c.stream, n, perm,
                       (const int *)c.porig, c.tmpi);
    std::swap(c.porig, c.tmpi);
    hipLaunchKernelGGL(k_gather<uint8_t>, dim3(nblk(n)), dim3(TB), 0, c.stream, n, perm,
                       (const uint8_t *)c.active, c.pact);
    std::swap(c.active, c.pact);
    SX_HIP(hipGetLastError());
    SX_TRY(permute_serve(c, perm));
    return permute_step(c, perm);
}

int set_clouds(Ctx &c, int ncloud, const int *counts) {
    if (ncloud < 1 || ncloud > c.max_clouds || ncloud > 65535)
        return set_error(SERENX_EINVAL, "cloud count out of range for the particle capacity");
    c.cloud_off_h.assign(ncloud + 1, 0);
    std::vector<int> chunks, first(ncloud + 1, 0);
    for (int k = 0; k < ncloud; k++) {
        // (a single cloud is checked when its h solve runs, as before)
        if (ncloud > 1 && counts[k] < c.pp_gather)
            return set_error(SERENX_EFEW, "a cloud has fewer particles than pp_gather: the "
                                          "grad-h solve cannot reach its neighbour number");
        c.cloud_off_h[k + 1] = c.cloud_off_h[k] + counts[k];
        first[k] = (int)chunks.size() / 3;
        for (int b = c.cloud_off_h[k]; b < c.cloud_off_h[k + 1]; b += BB_CHUNK) {
            chunks.push_back(k);
            chunks.push_back(b);
            chunks.push_back(std::min(b + BB_CHUNK, c.cloud_off_h[k + 1]));
        }
    }
    first[ncloud] = (int)chunks.size() / 3;
    if (c.cloud_off_h[ncloud] != c.n)
        return set_error(SERENX_EINVAL, "cloud counts do not add up to the particle count");
    if (first[ncloud] > c.max_chunks) return set_error(SERENX_EINVAL, "too many clouds");
    c.ncloud = ncloud;
    c.nchunk = first[ncloud];
    c.graph_epoch++;
    SX_HIP(hipMemcpy(c.cloud_off, c.cloud_off_h.data(), sizeof(int) * (ncloud + 1),
                     hipMemcpyHostToDevice));
    SX_HIP(hipMemcpy(c.chunk_first, first.data(), sizeof(int) * (ncloud + 1),
                     hipMemcpyHostToDevice));
    SX_HIP(hipMemcpy(c.chunks, chunks.data(), sizeof(int) * chunks.size(),
                     hipMemcpyHostToDevice));
    // the gravity walk's groups: 64 particles at a time from each cloud's start
    std::vector<int2> groups;
    for (int k = 0; k < ncloud; k++)
        for (int b = c.cloud_off_h[k]; b < c.cloud_off_h[k + 1]; b += WAVE)
            groups.push_back(make_int2(b, std::min(b + WAVE, c.cloud_off_h[k + 1])));
    if (groups.size() > walk_max_groups(c.cap, c.pp_gather))
        return set_error(SERENX_EINVAL, "walk groups beyond capacity");
    SX_HIP(hipMemcpy(c.grp, groups.data(), sizeof(int2) * groups.size(), hipMemcpyHostToDevice));
    c.ngrp = (int)groups.size();
    c.gq.have_cost = false;   // another group layout
    c.gq.order_ready = false;
    hipLaunchKernelGGL(k_fill_cid, dim3(64, ncloud), dim3(TB), 0, c.stream, ncloud, c.cloud_off,
                       c.cid);
    SX_HIP(hipGetLastError());
    SX_HIP(hipStreamSynchronize(c.stream));
    return SERENX_OK;
}

int build_tree(Ctx &c) {
    const int n = c.n;
    hipStream_t s = c.stream;
    c.graph_epoch++;   // a new topology, and the particle arrays trade places
    c.tree_ok = false;
    c.sph_ok = false;
    // bounding box and root cell of every cloud
    hipLaunchKernelGGL(k_bbox_chunks, dim3(c.nchunk), dim3(256), 0, s, c.chunks, c.x, c.y, c.z,
                       c.bbpart);
    hipLaunchKernelGGL(k_bbox_clouds, dim3(c.ncloud), dim3(64), 0, s, c.chunk_first, c.bbpart,
                       c.root);
    hipLaunchKernelGGL(k_keys, dim3(nblk(n)), dim3(TB), 0, s, n, c.x, c.y, c.z, c.root, c.cid,
                       c.khi, c.klo, c.idx);
    SX_HIP(hipGetLastError());
    // sort by the 21-level key within each cloud; redo with all 30 levels
    // only when needed.  A forest sorts all particles by key, then by cloud:
    // both radix sorts are stable, so every cloud comes out in the order of
    // sorting it alone (ties in input order), in the cloud's own place.  (A
    // segmented sort gives each cloud to one workgroup: tens of ms for a
    // cloud of a million particles.)
    const int cbits = std::max(1, 32 - __builtin_clz((unsigned)c.ncloud));
    auto sort_keys = [&](uint64_t *kin, uint64_t *kout, int *vin, int *vout,
                         int bits) -> hipError_t {
        size_t bytes = c.sorttmp_bytes;
        hipError_t e = rocprim::radix_sort_pairs(c.sorttmp, bytes, kin, kout, vin, vout,
                                                 (unsigned)n, 0, bits, s);
        if (e != hipSuccess || c.ncloud == 1) return e;
        hipLaunchKernelGGL(k_cloud_keys, dim3(nblk(n)), dim3(TB), 0, s, n, vout, c.cid, c.fs_ck,
                           c.fs_pos);
        bytes = c.sorttmp_bytes;
        e = rocprim::radix_sort_pairs(c.sorttmp, bytes, c.fs_ck, c.fs_ck2, c.fs_pos, c.fs_perm,
                                      (unsigned)n, 0, cbits, s);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_permute_kv, dim3(nblk(n)), dim3(TB), 0, s, n, c.fs_perm, kout, vout,
                           c.fs_k, c.fs_v);
        hipLaunchKernelGGL(k_copy_kv, dim3(nblk(n)), dim3(TB), 0, s, n, c.fs_k, c.fs_v, kout,
                           vout);
        return hipGetLastError();
    };
    SX_HIP(sort_keys(c.khi, c.khi2, c.idx, c.idx2, 63));
    SX_TRY(zero_words(c.flag, 1, s));
    hipLaunchKernelGGL(k_deep_check, dim3(nblk(n)), dim3(TB), 0, s, n, c.khi2, c.cid, c.flag);
    SX_HIP(hipMemcpyAsync(c.hostpin, c.flag, sizeof(int), hipMemcpyDeviceToHost, s));
    SX_HIP(hipMemcpyAsync(c.hostpin_d, c.root, 11 * sizeof(double), hipMemcpyDeviceToHost, s));
    SX_HIP(hipStreamSynchronize(s));
    c.rextent = c.hostpin_d[4];
    const bool deep = c.hostpin[0] != 0;
    if (deep) {
        // stable LSD order: deep 9 levels first, then the top 21 levels
        SX_HIP(sort_keys(c.klo, c.klo2, c.idx, c.idx2, 27));
        hipLaunchKernelGGL(k_gather_u64, dim3(nblk(n)), dim3(TB), 0, s, n, c.idx2, c.khi, c.khi2);
        SX_HIP(sort_keys(c.khi2, c.khi, c.idx2, c.idx, 63));
        // khi = sorted keys, idx = order; gather the deep keys in that order
        hipLaunchKernelGGL(k_gather_u64, dim3(nblk(n)), dim3(TB), 0, s, n, c.idx, c.klo, c.klo2);
        std::swap(c.khi, c.khi2);  // khi2 <- sorted hi keys
    } else {
        std::swap(c.idx, c.idx2);  // idx <- sorted order
        hipLaunchKernelGGL(k_gather_u64, dim3(nblk(n)), dim3(TB), 0, s, n, c.idx, c.klo, c.klo2);
    }
    hipLaunchKernelGGL(k_cpl, dim3(nblk(n)), dim3(TB), 0, s, n, c.khi2, c.klo2, c.cid, c.cpl);
    hipLaunchKernelGGL(k_init_levels, dim3(nblk(std::max(n + 1, c.ncloud))), dim3(TB), 0, s, n,
                       c.ncloud, c.cloud_off, c.cid, c.pact, c.cellof, c.ncs, c.firstlev,
                       c.c_start, c.c_end, c.c_parent, c.c_leaf, c.c_level);
    SX_HIP(hipGetLastError());
    c.level_off[0] = 0;
    c.level_cnt[0] = c.ncloud;
    int base = c.ncloud, ltot = 0;
    size_t tb = 0;
    for (int L = 1; L <= LMAX + 1; L++) {
        hipLaunchKernelGGL(k_mark, dim3(nblk(n)), dim3(TB), 0, s, n, L, c.pact, c.cpl, c.flag);
        tb = c.sorttmp_bytes;
        SX_HIP(rocprim::inclusive_scan(c.sorttmp, tb, c.flag, c.incl, (size_t)n,
                                       rocprim::plus<int>(), s));
        SX_HIP(hipMemcpyAsync(c.hostpin, c.incl + (n - 1), sizeof(int), hipMemcpyDeviceToHost, s));
        SX_HIP(hipStreamSynchronize(s));
        const int cnt = c.hostpin[0];
        if (cnt == 0) break;
        if (L == LMAX + 1)
            return set_error(SERENX_ELMAX, "Reached maximum level (more than 8 particles "
                                           "share a level-30 cell)");
        if (base + cnt > c.ccap) return set_error(SERENX_ENOMEM, "cell capacity exceeded");
        hipLaunchKernelGGL(k_emit, dim3(nblk(n)), dim3(TB), 0, s, n, L, base, c.pact, c.flag,
                           c.incl, c.cellof, c.c_start, c.c_end, c.c_parent, c.c_level, c.ncs,
                           c.firstlev);
        hipLaunchKernelGGL(k_leafcells, dim3(nblk(cnt)), dim3(TB), 0, s, base, cnt, c.c_start,
                           c.c_end, c.c_leaf);
        hipLaunchKernelGGL(k_deactivate, dim3(nblk(n)), dim3(TB), 0, s, n, c.pact, c.cellof,
                           c.c_leaf);
        SX_HIP(hipGetLastError());
        c.level_off[L] = base;
        c.level_cnt[L] = cnt;
        base += cnt;
        ltot = L;
    }
    c.ltot = ltot;
    c.nc = base;
    for (int L = ltot + 1; L <= LMAX + 1; L++) c.level_off[L] = base, c.level_cnt[L] = 0;
    // DFS layout
    SX_TRY(zero_words(c.ncs + n, 1, s));
    tb = c.sorttmp_bytes;
    SX_HIP(rocprim::exclusive_scan(c.sorttmp, tb, c.ncs, c.off, 0, (size_t)n + 1,
                                   rocprim::plus<int>(), s));
    hipLaunchKernelGGL(k_dfs, dim3(nblk(c.nc)), dim3(TB), 0, s, n, c.nc, c.c_start, c.c_end,
                       c.c_level, c.c_leaf, c.off, c.firstlev, c.bfs2dfs, c.dfs2bfs, c.gn, c.hn);
    hipLaunchKernelGGL(k_dfs_parent, dim3(nblk(c.nc)), dim3(TB), 0, s, c.nc, c.c_parent,
                       c.bfs2dfs, c.dfs_parent);
    hipLaunchKernelGGL(k_leafsort, dim3(nblk(c.nc)), dim3(TB), 0, s, c.nc, c.c_start, c.c_leaf,
                       c.idx);
    hipLaunchKernelGGL(k_root_range, dim3(nblk(c.ncloud)), dim3(TB), 0, s, c.ncloud, c.bfs2dfs,
                       c.gn, c.root_range);
    SX_HIP(hipGetLastError());
    SX_TRY(permute_particles(c, c.idx));
    c.tree_ok = true;
    return SERENX_OK;
}

// Factor on the stocked multipole-MAC term (BHgrav_stock.F90:329-338) that
// turns abserror into the mixed walk's tolerance: the term is proportional to
// abserror^(-1/3) (GADGET), abserror^(-1/2) (GADGET2) or abserror^(-2/3)
// (EIGEN).  The geometric MAC has no tolerance to trade.
static double mixed_mac_scale(const Ctx &c) {
    const double am = c.mixed_abserror, a0 = c.opt.abserror;
    if (!(am > 0.0) || !(a0 > 0.0) || am == a0) return 1.0;
    const double r = a0 / am;
    switch (c.opt.mac) {
    case SERENX_MAC_GADGET: return std::cbrt(r);
    case SERENX_MAC_GADGET2: return std::sqrt(r);
    case SERENX_MAC_EIGEN: return std::pow(r, 2.0 / 3.0);
    default: return 1.0;
    }
}

int stock_tree(Ctx &c) {
    hipStream_t s = c.stream;
    const int oct = c.opt.multipole == SERENX_OCTUPOLE_EXACT ? 2
                    : c.opt.multipole == SERENX_OCTUPOLE ? 1 : 0;
    for (int L = c.ltot; L >= 0; L--) {
        const int cnt = c.level_cnt[L];
        if (cnt == 0) continue;
        hipLaunchKernelGGL(k_stock, dim3(nblk(cnt)), dim3(TB), 0, s, c.nc, c.level_off[L], cnt,
                           c.bfs2dfs, c.gn, c.gs, c.hn, c.bbmin, c.bbmax, c.hmaxs, c.x, c.y, c.z, c.m,
                           c.h, oct, c.opt.mac, c.opt.thetamaxsqd, c.opt.abserror, c.daux,
                           c.kg.range);
    }
    if (c.opt.precision == SERENX_MIXED)
        hipLaunchKernelGGL(k_fill_f32_nodes, dim3(nblk(c.nc)), dim3(TB), 0, s, c.nc, c.gn, c.gs,
                           mixed_mac_scale(c), c.gnf);
    c.box_ok = true;   // the hydro boxes hold the current positions
    SX_HIP(hipGetLastError());
    return SERENX_OK;
}

int refresh_f32_nodes(Ctx &c) {
    if (!c.tree_ok || c.nc <= 0) return SERENX_OK;
    hipLaunchKernelGGL(k_fill_f32_nodes, dim3(nblk(c.nc)), dim3(TB), 0, c.stream, c.nc, c.gn,
                       c.gs, mixed_mac_scale(c), c.gnf);
    SX_HIP(hipGetLastError());
    return SERENX_OK;
}

int hguess(Ctx &c) {
    const int thresh = (int)(0.5 * (double)c.pp_gather);
    hipLaunchKernelGGL(k_hguess, dim3(nblk(c.nc)), dim3(TB), 0, c.stream, c.nc, thresh, c.kg.invrange, c.c_leaf,
                       c.c_start, c.c_end, c.c_parent, c.c_level, c.bfs2dfs, c.hn, c.h);
    SX_HIP(hipGetLastError());
    return SERENX_OK;
}

int update_hmax(Ctx &c) {
    if (c.nc <= 0) return SERENX_OK;
    hipLaunchKernelGGL(k_hmax_zero, dim3(nblk(c.nc)), dim3(TB), 0, c.stream, c.nc, c.hn);
    hipLaunchKernelGGL(k_hmax_up, dim3(nblk(c.nc)), dim3(TB), 0, c.stream, c.nc, c.hn,
                       c.dfs_parent, c.h, c.kg.range);
    SX_HIP(hipGetLastError());
    return SERENX_OK;
}

}  // namespace serenx
