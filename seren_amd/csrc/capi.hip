// capi.hip -- the C ABI of include/serenx.h: contexts, the host-state
// (drop-in) mode that mirrors SEREN's driver subroutines, the device-resident
// mode, tree export and the kernel tables.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <rocprim/rocprim.hpp>

#include "serenx_internal.h"

namespace serenx {

static thread_local std::string g_last_error;

namespace {
__global__ void k_zero_words(size_t nwords, uint32_t *__restrict__ p) {
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < nwords;
         i += (size_t)gridDim.x * blockDim.x)
        p[i] = 0u;
}
}  // namespace

int zero_words(void *p, size_t nwords, hipStream_t s) {
    if (nwords == 0) return SERENX_OK;
    const int blocks = (int)std::min<size_t>((nwords + 255) / 256, 1024);
    hipLaunchKernelGGL(k_zero_words, dim3(blocks), dim3(256), 0, s, nwords,
                       static_cast<uint32_t *>(p));
    SX_HIP(hipGetLastError());
    return SERENX_OK;
}

int set_error(int code, const std::string &msg) {
    g_last_error = msg;
    return code;
}

int hip_check(hipError_t e, const char *what) {
    char buf[512];
    snprintf(buf, sizeof(buf), "%s failed: %s", what, hipGetErrorString(e));
    return set_error(SERENX_EHIP, buf);
}

// x**n for a small integer constant n as the reference compiler expands it,
// a left-to-right product ((x*x)*x)*...
static double ipow_seq(double x, int n) {
    double y = x;
    for (int i = 1; i < n; i++) y = y * x;
    return y;
}

// kernel.F90:130-236, the quintic spline (GRAD_H_SPH, NDIM = 3), entry i at
// s = (3 i + 1) / KERNTOT; w2 and w3 are zero for this kernel
static void quintic_tables(double *tab, bool tc) {
    const double invkerntot = 1.0 / (double)KERNTOT;
    const int kern1 = (int)(ONETHIRD * (double)KERNTOT);          // kernel.F90:40-41
    const int kern2 = (int)(2.0 * ONETHIRD * (double)KERNTOT);
    const int kerntc = (int)(0.75929848074 * (double)KERNTOT / 3.0);   // :43
    const double nd = 3.0;
    double *w0 = tab + T_W0 * KT, *w1 = tab + T_W1 * KT, *w2 = tab + T_W2 * KT,
           *w3 = tab + T_W3 * KT, *w4 = tab + T_W4 * KT, *w5 = tab + T_W5 * KT,
           *w6 = tab + T_W6 * KT, *wh = tab + T_WH * KT, *wg = tab + T_WG * KT;
    auto P = ipow_seq;
    for (int i = 0; i <= KERNTOT; i++) {
        const double s = (3.0 * (double)i + 1.0) * invkerntot;
        w2[i] = 0.0;
        w3[i] = 0.0;
        if (i < kern1) {
            const double g = -(120.0 * s) + 120.0 * P(s, 3) - 50.0 * P(s, 4);
            w1[i] = (tc && i < kerntc) ? -55.2040 : g;
            w0[i] = 66.0 - 60.0 * s * s + 30.0 * P(s, 4) - 10.0 * P(s, 5);
            w4[i] = g;
            w5[i] = 22.0 * s - 12.0 * P(s, 3) + (30.0 / 7.0) * P(s, 5) - (5.0 / 4.0) * P(s, 6);
            w6[i] = -(11.0 * s * s) + 3.0 * P(s, 4) - (5.0 / 7.0) * P(s, 6) +
                    (5.0 / 28.0) * P(s, 7) + (478.0 / 14.0);
            wh[i] = -(66.0 * nd) + 60.0 * (nd + 2.0) * s * s - 30.0 * (nd + 4.0) * P(s, 4) +
                    10.0 * (nd + 5.0) * P(s, 5);
            wg[i] = 33.0 * s * s - 15.0 * P(s, 4) + 5.0 * P(s, 6) - (10.0 / 7.0) * P(s, 7) -
                    (478.0 / 14.0);
        } else if (i < kern2) {
            w0[i] = 51.0 + 75.0 * s - 210.0 * s * s + 150.0 * P(s, 3) - 45.0 * P(s, 4) +
                    5.0 * P(s, 5);
            const double g = 75.0 - 420.0 * s + 450.0 * s * s - 180.0 * P(s, 3) + 25.0 * P(s, 4);
            w1[i] = g;
            w4[i] = g;
            w5[i] = 17.0 * s + (75.0 / 4.0) * P(s, 2) - 42.0 * P(s, 3) + 25.0 * P(s, 4) -
                    (45.0 / 7.0) * P(s, 5) + (5.0 / 8.0) * P(s, 6) + (5.0 / 56.0) / P(s, 2);
            w6[i] = -((17.0 / 2.0) * s * s) - (25.0 / 4.0) * P(s, 3) + (21.0 / 2.0) * P(s, 4) -
                    5.0 * P(s, 5) + (15.0 / 14.0) * P(s, 6) - (5.0 / 56.0) * P(s, 7) +
                    (473.0 / 14.0) + (5.0 / 56.0) / s;
            wh[i] = -(51.0 * nd) - 75.0 * (nd + 1.0) * s + 210.0 * (nd + 2.0) * P(s, 2) -
                    150.0 * (nd + 3.0) * P(s, 3) + 45.0 * (nd + 4.0) * P(s, 4) -
                    5.0 * (nd + 5.0) * P(s, 5);
            wg[i] = (51.0 / 2.0) * s * s + 25.0 * P(s, 3) - (105.0 / 2.0) * P(s, 4) +
                    30.0 * P(s, 5) - (15.0 / 2.0) * P(s, 6) + (5.0 / 7.0) * P(s, 7) -
                    (473.0 / 14.0);
        } else {
            w0[i] = 243.0 - 405.0 * s + 270.0 * s * s - 90.0 * P(s, 3) + 15.0 * P(s, 4) - P(s, 5);
            const double g = -405.0 + 540.0 * s - 270.0 * s * s + 60.0 * P(s, 3) - 5.0 * P(s, 4);
            w1[i] = g;
            w4[i] = g;
            w5[i] = 81.0 * s - (405.0 / 4.0) * P(s, 2) + 54.0 * P(s, 3) - 15.0 * P(s, 4) +
                    (15.0 / 7.0) * P(s, 5) - (1.0 / 8.0) * P(s, 6) - (507.0 / 56.0) / P(s, 2);
            w6[i] = -((81.0 / 2.0) * s * s) + (135.0 / 4.0) * P(s, 3) - (27.0 / 2.0) * P(s, 4) +
                    3.0 * P(s, 5) - (5.0 / 14.0) * P(s, 6) + (1.0 / 56.0) * P(s, 7) +
                    (729.0 / 14.0) - (507.0 / 56.0) / s;
            wh[i] = -(243.0 * nd) + 405.0 * (nd + 1.0) * s - 270.0 * (nd + 2.0) * P(s, 2) +
                    90.0 * (nd + 3.0) * P(s, 3) - 15.0 * (nd + 4.0) * P(s, 4) +
                    (nd + 5.0) * P(s, 5);
            wg[i] = (243.0 / 2.0) * s * s - 135.0 * P(s, 3) + (135.0 / 2.0) * P(s, 4) -
                    18.0 * P(s, 5) + (5.0 / 2.0) * P(s, 6) - (1.0 / 7.0) * P(s, 7) -
                    (729.0 / 14.0);
        }
        // normalisation, NDIM = 3 (kernel.F90:220-236): left to right
        w0[i] = w0[i] * INVPI * 3.0 / 359.0;
        w1[i] = w1[i] * INVPI * 3.0 / 359.0;
        w4[i] = w4[i] * INVPI * 3.0 / 359.0;
        w5[i] = w5[i] * 12.0 / 359.0;
        w6[i] = w6[i] * 12.0 / 359.0;
        wh[i] = wh[i] * INVPI * 3.0 / 359.0;
        wg[i] = -wg[i] * 12.0 / 359.0;
    }
}

// kernel.F90:14-127 (M4, GRAD_H_SPH, NDIM = 3), or the quintic; tab[j*KT + i]
void make_kernel_tables(double *tab, int kernel) {
    if (kernel == SERENX_KERNEL_QUINTIC || kernel == SERENX_KERNEL_QUINTIC_TC) {
        quintic_tables(tab, kernel == SERENX_KERNEL_QUINTIC_TC);
        return;
    }
    const double invkerntot = 1.0 / (double)KERNTOT;
    const int kern1 = 5000;                              // int(HALFKERNTOT)
    const int kerntc = (int)((double)KERNTOT / 3.0);   // kernel.F90:37
    const double nd = 3.0;
    double *w0 = tab + T_W0 * KT, *w1 = tab + T_W1 * KT, *w2 = tab + T_W2 * KT,
           *w3 = tab + T_W3 * KT, *w4 = tab + T_W4 * KT, *w5 = tab + T_W5 * KT,
           *w6 = tab + T_W6 * KT, *wh = tab + T_WH * KT, *wg = tab + T_WG * KT;
    for (int i = 0; i <= KERNTOT; i++) {
        const double s = (2.0 * (double)i + 1.0) * invkerntot;
        const double saux = 2.0 - s;
        if (i < kern1) {
            if (kernel == SERENX_KERNEL_M4_TC && i < kerntc)
                w1[i] = -1.0;   // Thomas-Couchman gradient (kernel.F90:56-62)
            else
                w1[i] = -(3.0 * s) + 2.25 * s * s;
            w0[i] = 1.0 - 1.5 * s * s + 0.75 * ipow_seq(s, 3);
            w2[i] = (4.0 / 3.0) * ipow_seq(s, 3) - 1.2 * ipow_seq(s, 5) + 0.5 * ipow_seq(s, 6);
            w3[i] = (7.0 / 5.0) * s - (2.0 / 3.0) * ipow_seq(s, 3) + 0.3 * ipow_seq(s, 5) -
                    0.1 * ipow_seq(s, 6);
            w4[i] = -(3.0 * s) + 2.25 * s * s;
            w5[i] = (4.0 / 3.0) * s - 1.2 * ipow_seq(s, 3) + 0.5 * ipow_seq(s, 4);
            w6[i] = 7.0 / 5.0 - (2.0 / 3.0) * ipow_seq(s, 2) + 0.3 * ipow_seq(s, 4) -
                    0.1 * ipow_seq(s, 5);
            wh[i] = -nd + 1.5 * (nd + 2.0) * s * s - 0.75 * (nd + 3.0) * ipow_seq(s, 3);
            wg[i] = 1.4 - 2.0 * s * s + 1.5 * ipow_seq(s, 4) - 0.6 * ipow_seq(s, 5);
        } else {
            w0[i] = 0.25 * ipow_seq(saux, 3);
            w1[i] = -(0.75 * saux * saux);
            w2[i] = (8.0 / 3.0) * ipow_seq(s, 3) - 3.0 * ipow_seq(s, 4) + 1.2 * ipow_seq(s, 5) -
                    (1.0 / 6.0) * ipow_seq(s, 6) - (1.0 / 15.0);
            w3[i] = -(1.0 / 15.0) + (8.0 / 5.0) * s - (4.0 / 3.0) * ipow_seq(s, 3) +
                    ipow_seq(s, 4) - 0.3 * ipow_seq(s, 5) + (1.0 / 30.0) * ipow_seq(s, 6);
            w4[i] = -(0.75 * saux * saux);
            w5[i] = (8.0 / 3.0) * s - 3.0 * ipow_seq(s, 2) + 1.2 * ipow_seq(s, 3) -
                    (1.0 / 6.0) * ipow_seq(s, 4) - 1.0 / (15.0 * ipow_seq(s, 2));
            w6[i] = -(1.0 / (15.0 * s)) + 8.0 / 5.0 - (4.0 / 3.0) * ipow_seq(s, 2) +
                    ipow_seq(s, 3) - 0.3 * ipow_seq(s, 4) + (1.0 / 30.0) * ipow_seq(s, 5);
            wh[i] = -2.0 * nd + 3.0 * (nd + 1.0) * s - 1.50 * (nd + 2.0) * s * s +
                    0.25 * (nd + 3.0) * ipow_seq(s, 3);
            wg[i] = 1.6 - 4.0 * s * s + 4.0 * ipow_seq(s, 3) - 1.5 * ipow_seq(s, 4) +
                    0.2 * ipow_seq(s, 5);
        }
        w0[i] = w0[i] * INVPI;
        w1[i] = w1[i] * INVPI;
        w4[i] = w4[i] * INVPI;
        wh[i] = wh[i] * INVPI;
    }
}

namespace {

constexpr int TB = 256;
inline int nblk(long n) { return (int)((n + TB - 1) / TB); }

__global__ void k_init_state(int n, int *porig, uint8_t *active, double *rho, double *omega,
                             double *zeta, double *ax, double *ay, double *az, double *gpot,
                             double *agm, double *divv, double *press, double *sound,
                             double *drhodt) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    porig[i] = i + 1;
    active[i] = 1;
    rho[i] = 1.0;     // initialize_sph_variables_1.F90:68-72
    omega[i] = 1.0;
    zeta[i] = 0.0;
    ax[i] = ay[i] = az[i] = 0.0;
    gpot[i] = 0.0;
    agm[i] = 0.0;
    divv[i] = 0.0;
    press[i] = sound[i] = drhodt[i] = 0.0;
}

__global__ void k_iota1(int n, int *porig) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) porig[i] = i + 1;
}

__global__ void k_zero_active(int n, const uint8_t *active, double *ax, double *ay, double *az) {
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && active[i]) ax[i] = ay[i] = az[i] = 0.0;
}

template <class T>
int dalloc(T **p, size_t count) {
    if (*p) {
        (void)hipFree(*p);
        *p = nullptr;
    }
    if (count == 0) count = 1;
    if (hipMalloc((void **)p, sizeof(T) * count) != hipSuccess)
        return set_error(SERENX_ENOMEM, "hipMalloc failed for particle arrays");
    return SERENX_OK;
}

int ensure_capacity(Ctx &c, int n) {
    c.graph_epoch++;   // buffers may move
    // leaf descriptors of the h solve pack a particle index in 27 bits
    if (n >= (1 << 27))
        return set_error(SERENX_EINVAL, "more than 2^27 - 1 particles on one device");
    if (n <= c.cap && c.x) return SERENX_OK;
    c.cap = n;
    const size_t nn = (size_t)n + 1;
    double **arrs[] = {&c.x, &c.y, &c.z, &c.m, &c.h, &c.zeta, &c.vx, &c.vy, &c.vz,
                       &c.rho, &c.omega, &c.press, &c.sound, &c.divv, &c.drhodt,
                       &c.ax, &c.ay, &c.az, &c.gpot, &c.agm, &c.tmpd, &c.u, &c.dudt};
    for (double **a : arrs) SX_TRY(dalloc(a, nn));
    SX_HIP(hipMemset(c.u, 0, sizeof(double) * nn));
    SX_HIP(hipMemset(c.dudt, 0, sizeof(double) * nn));
    SX_TRY(dalloc(&c.porig, nn));
    SX_TRY(dalloc(&c.tmpi, nn));
    SX_TRY(dalloc(&c.active, nn));
    SX_TRY(dalloc(&c.pptot, nn));
    SX_TRY(dalloc(&c.pplist, nn * (size_t)c.pp_limit));
    SX_TRY(dalloc(&c.ovf_off, nn));
    SX_TRY(dalloc(&c.ovf_idx, nn));
    SX_TRY(dalloc(&c.nbl, particle_waves(n) * NB_CAP));
    SX_TRY(dalloc(&c.nbl_len, particle_waves(n)));
    SX_TRY(dalloc(&c.hrec, nn * 16));
    SX_TRY(dalloc(&c.grec, nn * 8));
    SX_TRY(dalloc(&c.gax, nn));
    SX_TRY(dalloc(&c.gay, nn));
    SX_TRY(dalloc(&c.gaz, nn));
    SX_TRY(dalloc(&c.gpc, nn));
    // per group of the gravity walk (at most one partial group per cloud)
    const size_t ng = walk_max_groups(n, c.pp_gather);
    SX_TRY(dalloc(&c.grp, ng));
    c.pl_entries = near_list_entries(n);
    SX_TRY(dalloc(&c.pl_leaf, c.pl_entries));
    SX_TRY(dalloc(&c.pl_mask, c.pl_entries));
    SX_TRY(dalloc(&c.pl_len, ng));
    SX_TRY(dalloc(&c.hlist, (((size_t)n + 255) / 256) * 4 * HL_SLOTS * WAVE));
    // work queue of the gravity walk: one entry per group
    Ctx::WalkQueue &q = c.gq;
    SX_TRY(dalloc(&q.ctr, (size_t)(2 * NXCD)));   // phase 1 and phase 2 counters
    SX_TRY(dalloc(&q.cost, ng));
    SX_TRY(dalloc(&q.order, ng));
    SX_TRY(dalloc(&q.vals, ng));
    SX_TRY(dalloc(&q.keys, ng));
    SX_TRY(dalloc(&q.keys2, ng));
    size_t tb = 0;
    SX_HIP(rocprim::radix_sort_pairs(nullptr, tb, q.keys, q.keys2, q.vals, q.order, (unsigned)ng,
                                     0, 35, c.stream));
    if (q.tmp) (void)hipFree(q.tmp);
    q.tmp = nullptr;
    if (hipMalloc(&q.tmp, tb + 256) != hipSuccess)
        return set_error(SERENX_ENOMEM, "hipMalloc failed for the walk queue's sort space");
    q.tmp_bytes = tb + 256;
    q.ngroups = 0;
    q.have_cost = false;
    SX_TRY(ensure_tree_buffers(c));
    return SERENX_OK;
}

void free_all(Ctx &c) {
    step_free(c);
    serve_free(c);
    void *ptrs[] = {c.x, c.y, c.z, c.m, c.h, c.zeta, c.vx, c.vy, c.vz, c.rho, c.omega,
                    c.press, c.sound, c.divv, c.drhodt, c.ax, c.ay, c.az, c.gpot, c.agm,
                    c.u, c.dudt,
                    c.tmpd, c.hrec, c.hlist, c.porig, c.tmpi, c.active, c.pptot, c.pplist, c.ovf_off,
                    c.ovf_idx, c.nbl, c.nbl_len,
                    c.ovf_list, c.ovf_count, c.khi, c.klo, c.khi2, c.klo2, c.idx, c.idx2,
                    c.cpl, c.pact, c.flag, c.incl, c.cellof, c.ncs, c.off, c.firstlev,
                    c.root, c.bbpart, c.cloud_off, c.cid, c.chunks, c.chunk_first,
                    c.root_range, c.sorttmp, c.c_start, c.c_end, c.c_parent, c.c_leaf,
                    c.c_level, c.bfs2dfs, c.dfs2bfs, c.dfs_parent, c.bbmin, c.bbmax,
                    c.hmaxs, c.gn, c.gs, c.gnf, c.hn, c.ktab, c.icount, c.wstat, c.band,
                    c.grec, c.gax, c.gay, c.gaz, c.gpc, c.grp, c.pl_leaf, c.pl_mask, c.pl_len, c.gq.ctr, c.gq.cost, c.gq.order,
                    c.gq.vals, c.gq.keys, c.gq.keys2, c.gq.tmp, c.fs_ck, c.fs_ck2, c.fs_pos,
                    c.fs_perm, c.fs_v, c.fs_k};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    if (c.hostpin) (void)hipHostFree(c.hostpin);
    if (c.hostpin_d) (void)hipHostFree(c.hostpin_d);
    for (auto &e : c.ev)
        if (e) (void)hipEventDestroy(e);
    for (auto &e : c.evg)
        if (e) (void)hipEventDestroy(e);
    if (c.evfork) (void)hipEventDestroy(c.evfork);
    if (c.evfork2) (void)hipEventDestroy(c.evfork2);
    if (c.evjoin) (void)hipEventDestroy(c.evjoin);
    graphs_free(c);
    if (c.gstream && c.gstream != c.stream) (void)hipStreamDestroy(c.gstream);
    if (c.stream) (void)hipStreamDestroy(c.stream);
}

int check_options(const serenx_options &o) {
    if (o.multipole != SERENX_QUADRUPOLE && o.multipole != SERENX_OCTUPOLE &&
        o.multipole != SERENX_OCTUPOLE_EXACT)
        return set_error(SERENX_EINVAL,
                         "multipole must be SERENX_QUADRUPOLE, SERENX_OCTUPOLE or "
                         "SERENX_OCTUPOLE_EXACT");
    if (o.mac < SERENX_MAC_GEOMETRIC || o.mac > SERENX_MAC_EIGEN)
        return set_error(SERENX_EINVAL, "mac must be GEOMETRIC, GADGET, GADGET2 or EIGEN");
    if (o.precision != SERENX_FP64 && o.precision != SERENX_MIXED)
        return set_error(SERENX_EINVAL, "precision must be SERENX_FP64 or SERENX_MIXED");
    if (o.kernel < SERENX_KERNEL_M4 || o.kernel > SERENX_KERNEL_QUINTIC_TC)
        return set_error(SERENX_EINVAL, "kernel must be one of SERENX_KERNEL_*");
    if (!(o.h_fac > 0.0)) return set_error(SERENX_EINVAL, "h_fac must be positive");
    if (o.mac != SERENX_MAC_GEOMETRIC && !(o.abserror > 0.0))
        return set_error(SERENX_EINVAL, "abserror must be positive for multipole MACs");
    if (o.thermal != SERENX_THERMAL_ISOTHERMAL && o.thermal != SERENX_THERMAL_ENERGY_EQN)
        return set_error(SERENX_EINVAL,
                         "thermal must be SERENX_THERMAL_ISOTHERMAL or SERENX_THERMAL_ENERGY_EQN");
    if (o.thermal == SERENX_THERMAL_ENERGY_EQN && !(o.gamma > 1.0))
        return set_error(SERENX_EINVAL, "gamma must exceed 1 for the energy equation");
    if (o.sph != SERENX_SPH_GRADH && o.sph != SERENX_SPH_STANDARD)
        return set_error(SERENX_EINVAL, "sph must be SERENX_SPH_GRADH or SERENX_SPH_STANDARD");
    if (o.serial != 0 && o.serial != 1) return set_error(SERENX_EINVAL, "serial must be 0 or 1");
    if (o.hfind != SERENX_HFIND_NUMBER && o.hfind != SERENX_HFIND_CONSTANT)
        return set_error(SERENX_EINVAL, "hfind must be SERENX_HFIND_NUMBER or SERENX_HFIND_CONSTANT");
    if (o.hfind == SERENX_HFIND_CONSTANT && o.sph != SERENX_SPH_STANDARD)
        return set_error(SERENX_EINVAL, "SERENX_HFIND_CONSTANT needs SERENX_SPH_STANDARD");
    if (o.hfind == SERENX_HFIND_CONSTANT && !(o.hmin > 0.0))
        return set_error(SERENX_EINVAL, "SERENX_HFIND_CONSTANT needs hmin > 0");
    if (o.pp_gather < 0 || (o.sph == SERENX_SPH_GRADH && o.pp_gather != 0))
        return set_error(SERENX_EINVAL,
                         "pp_gather must be 0 for grad-h SPH (h_fac sets it) and >= 0 otherwise");
    return SERENX_OK;
}

// device bytes held by a context, from the allocation sizes
size_t ctx_device_bytes(const Ctx &c) {
    const size_t nn = (size_t)c.cap + 1, cc = (size_t)c.ccap;
    size_t b = 0;
    b += nn * (24 * 8 + 16 * 8 + 12 * 8);                   // SoA, tmpd, hydro/gravity records
    b += walk_max_groups(c.cap, c.pp_gather) * (size_t)(4 + 8 + 6 * 8);  // walk groups
    b += near_list_entries(c.cap) * 12;                              // near-field lists
    b += particle_waves(c.cap) * (NB_CAP + 1) * 4;                   // neighbour-walk leaves
    b += nn * (5 * 4 + 1) + nn * (size_t)c.pp_limit * 4;    // ints, active, neighbour rows
    b += (((size_t)c.cap + 255) / 256) * 4 * HL_SLOTS * WAVE * 4;   // h-solve lists
    b += nn * (4 * 8 + 2 * 4 + 2 + 5 * 4 + 1);              // tree build scratch
    b += cc * (sizeof(GNode) + 10 * 8 + sizeof(GNodeF) + sizeof(HNode) + 8 * 4 + 7 * 8);
    b += c.sorttmp_bytes + c.ovf_cap * 4 + sizeof(double) * KT_TOTAL;
    if (c.st.on) b += nn * (11 * 8 + 8 + 4);                 // step state
    if (c.sv.on) b += nn * 3 * 8 + (size_t)c.ncloud * 25;    // serving state
    b += c.band_cap * BAND_MAX * 8;                         // standard SPH band
    return b;
}

Ctx *as_ctx(serenx_ctx *p) { return reinterpret_cast<Ctx *>(p); }
const Ctx *as_ctx(const serenx_ctx *p) { return reinterpret_cast<const Ctx *>(p); }

}  // namespace

int run_tree_update(Ctx &c) {
    SX_TRY(build_tree(c));
    SX_TRY(stock_tree(c));
    return SERENX_OK;
}

// sph_update.F90:55-101 (grad-h) and :106-152 (standard SPH), in two parts
// so that the gravity walk, which needs only h (not the hydro tree's
// hrangemax), can start between them
int run_sph_h(Ctx &c) {
    if (!c.tree_ok) return set_error(SERENX_ESTATE, "sph_update called before tree_update");
    c.sph_ok = false;
    if (c.opt.sph == SERENX_SPH_STANDARD)
        SX_TRY(sph_hgather(c));
    else
        SX_TRY(sph_hsolve(c));
    return SERENX_OK;
}

// the neighbour lists (after update_hmax)
int run_sph_lists(Ctx &c) {
    SX_TRY(sph_neighbours(c));
    if (c.opt.sph == SERENX_SPH_STANDARD) SX_TRY(sph_density(c));
    c.sph_ok = true;
    return SERENX_OK;
}

int run_sph_update(Ctx &c) {
    SX_TRY(run_sph_h(c));
    SX_TRY(update_hmax(c));
    SX_TRY(run_sph_lists(c));
    return SERENX_OK;
}

int run_hydro(Ctx &c) {
    if (!c.sph_ok) return set_error(SERENX_ESTATE, "hydro forces need sph_update first");
    hipLaunchKernelGGL(k_zero_active, dim3(nblk(c.n)), dim3(TB), 0, c.stream, c.n, c.active, c.ax,
                       c.ay, c.az);
    SX_TRY(hydro_forces(c));
    return SERENX_OK;
}

int run_grav(Ctx &c) {
    if (!c.tree_ok) return set_error(SERENX_ESTATE, "gravity needs tree_update first");
    GravGuard guard(c.gstream);
    guard.armed = true;
    c.walk_alone = true;   // nothing runs beside this walk
    const int rc1 = grav_launch(c);
    const int rc2 = rc1 == SERENX_OK ? grav_pairs(c) : rc1;
    c.walk_alone = false;
    SX_TRY(rc2);
    SX_TRY(grav_join(c));
    guard.armed = false;
    return SERENX_OK;
}

namespace {

// ---- host-state transfers -------------------------------------------------
// parray(6,ptot) column-major (DATATOT = ZETA with GRAVITY, macros.h:53-57):
// element (k, p) at parray[6 p + k].  Standard SPH has no zeta: column 5 is
// neither read nor written, and the device keeps zeta = 0.
bool skip_col(const Ctx &c, int k) { return k == 5 && c.opt.sph == SERENX_SPH_STANDARD; }

int up_parray(Ctx &c, const int *cols, int ncols) {
    const int n = c.n;
    double *dst[6] = {c.x, c.y, c.z, c.m, c.h, c.zeta};
    c.hbuf.resize(n);
    for (int t = 0; t < ncols; t++) {
        const int k = cols[t];
        if (skip_col(c, k)) continue;
        for (int p = 0; p < n; p++) c.hbuf[p] = c.hs.parray[6 * (size_t)p + k];
        SX_HIP(hipMemcpy(dst[k], c.hbuf.data(), sizeof(double) * n, hipMemcpyHostToDevice));
    }
    return SERENX_OK;
}

int down_parray(Ctx &c, const int *cols, int ncols) {
    const int n = c.n;
    double *src[6] = {c.x, c.y, c.z, c.m, c.h, c.zeta};
    c.hbuf.resize(n);
    for (int t = 0; t < ncols; t++) {
        const int k = cols[t];
        if (skip_col(c, k)) continue;
        SX_HIP(hipMemcpy(c.hbuf.data(), src[k], sizeof(double) * n, hipMemcpyDeviceToHost));
        for (int p = 0; p < n; p++) c.hs.parray[6 * (size_t)p + k] = c.hbuf[p];
    }
    return SERENX_OK;
}

int up_vec3(Ctx &c, const double *h3, double *dx, double *dy, double *dz) {
    const int n = c.n;
    double *d[3] = {dx, dy, dz};
    c.hbuf.resize(n);
    for (int k = 0; k < 3; k++) {
        for (int p = 0; p < n; p++) c.hbuf[p] = h3[3 * (size_t)p + k];
        SX_HIP(hipMemcpy(d[k], c.hbuf.data(), sizeof(double) * n, hipMemcpyHostToDevice));
    }
    return SERENX_OK;
}

int down_vec3(Ctx &c, double *h3, const double *dx, const double *dy, const double *dz) {
    const int n = c.n;
    const double *d[3] = {dx, dy, dz};
    c.hbuf.resize(n);
    for (int k = 0; k < 3; k++) {
        SX_HIP(hipMemcpy(c.hbuf.data(), d[k], sizeof(double) * n, hipMemcpyDeviceToHost));
        for (int p = 0; p < n; p++) h3[3 * (size_t)p + k] = c.hbuf[p];
    }
    return SERENX_OK;
}

int up_scalar(Ctx &c, const double *hsrc, double *d) {
    if (!hsrc) return SERENX_OK;
    SX_HIP(hipMemcpy(d, hsrc, sizeof(double) * c.n, hipMemcpyHostToDevice));
    return SERENX_OK;
}

int down_scalar(Ctx &c, double *hdst, const double *d) {
    if (!hdst) return SERENX_OK;
    SX_HIP(hipMemcpy(hdst, d, sizeof(double) * c.n, hipMemcpyDeviceToHost));
    return SERENX_OK;
}

int up_accdo(Ctx &c) {
    std::vector<uint8_t> a(c.n, 1);
    if (c.hs.accdo)
        for (int p = 0; p < c.n; p++) a[p] = c.hs.accdo[p] != 0;
    SX_HIP(hipMemcpy(c.active, a.data(), c.n, hipMemcpyHostToDevice));
    return SERENX_OK;
}

int need_bound(Ctx &c) {
    if (!c.bound) return set_error(SERENX_ESTATE, "no host state bound");
    return SERENX_OK;
}

}  // namespace
}  // namespace serenx

using namespace serenx;

extern "C" {

const char *serenx_last_error(void) { return g_last_error.c_str(); }

int serenx_default_options(serenx_options *opt) {
    if (!opt) return set_error(SERENX_EINVAL, "null options");
    std::memset(opt, 0, sizeof(*opt));
    opt->multipole = SERENX_QUADRUPOLE;
    opt->mac = SERENX_MAC_GEOMETRIC;
    opt->precision = SERENX_FP64;
    opt->h_fac = 1.2;
    opt->thetamaxsqd = 0.5;
    opt->abserror = 0.1;
    opt->alpha = 1.0;
    opt->sound_speed = 1.0;
    opt->thermal = SERENX_THERMAL_ISOTHERMAL;
    opt->gamma = 1.4;                    // default_parameters.F90:173
    return SERENX_OK;
}

int serenx_create(const serenx_options *opt, int device, serenx_ctx **out) {
    if (!opt || !out) return set_error(SERENX_EINVAL, "null argument");
    *out = nullptr;
    SX_TRY(check_options(*opt));
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
        return set_error(SERENX_EHIP, "no HIP device available");
    if (device < 0 || device >= ndev) return set_error(SERENX_EINVAL, "bad device index");
    SX_HIP(hipSetDevice(device));
    // the allocation-free diagnostics' scratch address on this device (once)
    SX_TRY(diag_scratch_setup(device));
    Ctx *c = new Ctx();
    c->opt = *opt;
    c->device = device;
    c->kg = kernel_geom(opt->kernel);
    c->pp_gather = (opt->sph == SERENX_SPH_STANDARD && opt->pp_gather > 0)
                       ? opt->pp_gather
                       : (int)((4.0 * PI_REF / 3.0) * ipow_seq(c->kg.range * opt->h_fac, 3));
    c->pp_limit = 2 * c->pp_gather;
    c->daux = 10.0 * std::pow(1.7976931348623157e308, -0.2);
    int ncu = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess &&
        ncu > 0)
        c->ncu = ncu;
    // two streams at the default priority: the main one and the gravity walk's
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e == hipSuccess) {
        if (opt->serial)
            c->gstream = c->stream;
        else
            e = hipStreamCreateWithFlags(&c->gstream, hipStreamNonBlocking);
    }
    for (auto &ev : c->ev)
        if (e == hipSuccess) e = hipEventCreate(&ev);
    for (auto &ev : c->evg)
        if (e == hipSuccess) e = hipEventCreate(&ev);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c->evfork, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c->evfork2, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c->evjoin, hipEventDisableTiming);
    if (e == hipSuccess) e = hipMalloc((void **)&c->ktab, sizeof(double) * KT_TOTAL);
    if (e == hipSuccess) e = hipMalloc((void **)&c->icount, CTR_WORDS * sizeof(unsigned long long));
    if (e == hipSuccess) e = hipMemset(c->icount, 0, CTR_WORDS * sizeof(unsigned long long));
    if (e == hipSuccess) e = hipMalloc((void **)&c->wstat, CTR_WORDS * sizeof(unsigned long long));
    if (e == hipSuccess) e = hipMemset(c->wstat, 0, CTR_WORDS * sizeof(unsigned long long));
    if (e == hipSuccess) e = hipMalloc((void **)&c->ovf_count, sizeof(int));
    if (e != hipSuccess) {
        free_all(*c);
        delete c;
        return hip_check(e, "serenx_create");
    }
    make_kernel_tables(c->host_ktab, c->opt.kernel);
    {
        std::vector<double> dev(KT_TOTAL, 0.0);
        const double *t = c->host_ktab;
        std::copy(t, t + NTAB * KT, dev.begin());
        float *tf = reinterpret_cast<float *>(dev.data() + KT_F32_OFF);
        for (int i = 0; i < NTAB * KT; i++) tf[i] = (float)t[i];
        double *g4 = dev.data() + KT_G4_OFF;
        float *g4f = reinterpret_cast<float *>(dev.data() + KT_G4F_OFF);
        double *h4 = dev.data() + KT_H4_OFF;
        for (int k = 0; k < KT; k++) {
            const double w[3] = {t[T_W4 * KT + k], t[T_W5 * KT + k], t[T_W6 * KT + k]};
            for (int q = 0; q < 3; q++) {
                g4[4 * k + q] = w[q];
                g4f[4 * k + q] = (float)w[q];
            }
            h4[4 * k + 0] = t[T_W0 * KT + k];
            h4[4 * k + 1] = t[T_WH * KT + k];
            h4[4 * k + 2] = t[T_W4 * KT + k];
            h4[4 * k + 3] = t[T_WG * KT + k];
        }
        e = hipMemcpy(c->ktab, dev.data(), sizeof(double) * KT_TOTAL, hipMemcpyHostToDevice);
    }
    if (e != hipSuccess) {
        free_all(*c);
        delete c;
        return hip_check(e, "serenx_create tables");
    }
    *out = reinterpret_cast<serenx_ctx *>(c);
    return SERENX_OK;
}

int serenx_destroy(serenx_ctx *ctx) {
    if (!ctx) return SERENX_OK;
    Ctx *c = as_ctx(ctx);
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    if (c->gstream) (void)hipStreamSynchronize(c->gstream);
    free_all(*c);
    delete c;
    return SERENX_OK;
}

int serenx_neighbour_params(const serenx_ctx *ctx, int32_t *pp_gather, int32_t *pp_limit) {
    if (!ctx) return set_error(SERENX_EINVAL, "null context");
    const Ctx *c = as_ctx(ctx);
    if (pp_gather) *pp_gather = c->pp_gather;
    if (pp_limit) *pp_limit = c->pp_limit;
    return SERENX_OK;
}

// ---- host-state mode ------------------------------------------------------
int serenx_bind_host_state(serenx_ctx *ctx, const serenx_host_state *st) {
    if (!ctx || !st) return set_error(SERENX_EINVAL, "null argument");
    Ctx &c = *as_ctx(ctx);
    if (st->ptot <= 0 || !st->parray || !st->v || !st->a)
        return set_error(SERENX_EINVAL, "host state needs ptot > 0, parray, v and a");
    if (st->pp_limit != c.pp_limit)
        return set_error(SERENX_EINVAL, "pp_limit of the host state differs from 2*pp_gather");
    SX_HIP(hipSetDevice(c.device));
    // re-binding the same particle count keeps the device tree and lists: a
    // host may re-bind before every call (its arrays can move between steps)
    const bool same = c.bound && st->ptot == c.n;
    SX_TRY(ensure_capacity(c, st->ptot));
    c.hs = *st;
    c.n = st->ptot;
    c.bound = true;
    if (!same || c.ncloud != 1) {
        const int32_t one = c.n;
        SX_TRY(set_clouds(c, 1, &one));
    }
    if (same) return SERENX_OK;
    c.tree_ok = c.sph_ok = false;
    hipLaunchKernelGGL(k_init_state, dim3(nblk(c.n)), dim3(TB), 0, c.stream, c.n, c.porig,
                       c.active, c.rho, c.omega, c.zeta, c.ax, c.ay, c.az, c.gpot, c.agm, c.divv,
                       c.press, c.sound, c.drhodt);
    SX_HIP(hipStreamSynchronize(c.stream));
    return SERENX_OK;
}

int serenx_tree_update(serenx_ctx *ctx, int32_t *dummylist) {
    if (!ctx) return set_error(SERENX_EINVAL, "null context");
    Ctx &c = *as_ctx(ctx);
    SX_TRY(need_bound(c));
    SX_HIP(hipSetDevice(c.device));
    const int cols[] = {0, 1, 2, 3, 4};
    SX_TRY(up_parray(c, cols, 5));
    hipLaunchKernelGGL(k_iota1, dim3(nblk(c.n)), dim3(TB), 0, c.stream, c.n, c.porig);
    SX_TRY(run_tree_update(c));
    SX_HIP(hipStreamSynchronize(c.stream));
    if (dummylist)
        SX_HIP(hipMemcpy(dummylist, c.porig, sizeof(int) * c.n, hipMemcpyDeviceToHost));
    if (c.hs.rextent) *c.hs.rextent = c.rextent;
    return SERENX_OK;
}

int serenx_tree_stock(serenx_ctx *ctx) {
    if (!ctx) return set_error(SERENX_EINVAL, "null context");
    Ctx &c = *as_ctx(ctx);
    SX_TRY(need_bound(c));
    if (!c.tree_ok) return set_error(SERENX_ESTATE, "tree_stock called before tree_update");
    SX_HIP(hipSetDevice(c.device));
    const int cols[] = {0, 1, 2, 3, 4};
    SX_TRY(up_parray(c, cols, 5));
    SX_TRY(stock_tree(c));
    SX_HIP(hipStreamSynchronize(c.stream));
    return SERENX_OK;
}

int serenx_hguess(serenx_ctx *ctx) {
    if (!ctx) return set_error(SERENX_EINVAL, "null context");
    Ctx &c = *as_ctx(ctx);
    SX_TRY(need_bound(c));
    if (!c.tree_ok) return set_error(SERENX_ESTATE, "hguess called before tree_update");
    SX_HIP(hipSetDevice(c.device));
    SX_TRY(hguess(c));
    SX_HIP(hipStreamSynchronize(c.stream));
    const int cols[] = {4};
    SX_TRY(down_parray(c, cols, 1));
    return SERENX_OK;
}

int serenx_sph_update(serenx_ctx *ctx) {
    if (!ctx) return set_error(SERENX_EINVAL, "null context");
    Ctx &c = *as_ctx(ctx);
    SX_TRY(need_bound(c));
    SX_HIP(hipSetDevice(c.device));
    const int cols[] = {0, 1, 2, 3, 4};
    SX_TRY(up_parray(c, cols, 5));
    c.box_ok = false;   // the host's positions may have moved since its last stock
    SX_TRY(up_vec3(c, c.hs.v, c.vx, c.vy, c.vz));
    SX_TRY(up_accdo(c));
    SX_TRY(run_sph_update(c));
    SX_HIP(hipStreamSynchronize(c.stream));
    const int outc[] = {4, 5};
    SX_TRY(down_parray(c, outc, 2));
    SX_TRY(down_scalar(c, c.hs.rho, c.rho));
    SX_TRY(down_scalar(c, c.hs.omega, c.omega));
    SX_TRY(down_scalar(c, c.hs.div_v, c.divv));
    SX_TRY(down_scalar(c, c.hs.drhodt, c.drhodt));
    if (c.hs.pptot) {
        std::vector<int> tot(c.n);
        SX_HIP(hipMemcpy(tot.data(), c.pptot, sizeof(int) * c.n, hipMemcpyDeviceToHost));
        std::vector<int> lst((size_t)c.n * c.pp_limit);
        SX_HIP(hipMemcpy(lst.data(), c.pplist, sizeof(int) * lst.size(), hipMemcpyDeviceToHost));
        std::vector<uint8_t> act(c.n);
        SX_HIP(hipMemcpy(act.data(), c.active, c.n, hipMemcpyDeviceToHost));
        for (int p = 0; p < c.n; p++) {
            if (!act[p]) continue;
            c.hs.pptot[p] = tot[p];
            if (c.hs.pplist && tot[p] <= c.pp_limit)
                for (int i = 0; i < tot[p]; i++)
                    c.hs.pplist[(size_t)p * c.pp_limit + i] = lst[(size_t)p * c.pp_limit + i] + 1;
        }
    }
    return SERENX_OK;
}

int serenx_hydro_forces(serenx_ctx *ctx) {
    if (!ctx) return set_error(SERENX_EINVAL, "null context");
    Ctx &c = *as_ctx(ctx);
    SX_TRY(need_bound(c));
    SX_HIP(hipSetDevice(c.device));
    const int cols[] = {0, 1, 2, 3, 4};
    SX_TRY(up_parray(c, cols, 5));
    SX_TRY(up_vec3(c, c.hs.v, c.vx, c.vy, c.vz));
    SX_TRY(up_vec3(c, c.hs.a, c.ax, c.ay, c.az));
    SX_TRY(up_scalar(c, c.hs.rho, c.rho));
    SX_TRY(up_scalar(c, c.hs.omega, c.omega));
    SX_TRY(up_scalar(c, c.hs.press, c.press));
    SX_TRY(up_scalar(c, c.hs.sound, c.sound));
    const bool energy = c.opt.thermal == SERENX_THERMAL_ENERGY_EQN;
    if (energy) {   // du/dt reads div v of p; inactive particles keep their du/dt
        if (!c.hs.du_dt)
            return set_error(SERENX_EINVAL, "the energy equation needs du_dt in the host state");
        SX_TRY(up_scalar(c, c.hs.div_v, c.divv));
        SX_TRY(up_scalar(c, c.hs.du_dt, c.dudt));
    }
    SX_TRY(up_accdo(c));
    SX_TRY(run_hydro(c));
    SX_HIP(hipStreamSynchronize(c.stream));
    SX_TRY(down_vec3(c, c.hs.a, c.ax, c.ay, c.az));
    if (energy) SX_TRY(down_scalar(c, c.hs.du_dt, c.dudt));
    return SERENX_OK;
}

int serenx_grav_forces(serenx_ctx *ctx) {
    if (!ctx) return set_error(SERENX_EINVAL, "null context");
    Ctx &c = *as_ctx(ctx);
    SX_TRY(need_bound(c));
    SX_HIP(hipSetDevice(c.device));
    const int cols[] = {0, 1, 2, 3, 4, 5};
    SX_TRY(up_parray(c, cols, 6));
    SX_TRY(up_vec3(c, c.hs.a, c.ax, c.ay, c.az));
    if (c.opt.mac != SERENX_MAC_GEOMETRIC && c.hs.agravmag) SX_TRY(up_scalar(c, c.hs.agravmag, c.agm));
    // the EIGEN MAC reads the previous potential (BHgrav_accel.F90:109-110)
    if (c.opt.mac == SERENX_MAC_EIGEN && c.hs.gpot) SX_TRY(up_scalar(c, c.hs.gpot, c.gpot));
    SX_TRY(up_accdo(c));
    SX_TRY(run_grav(c));
    SX_HIP(hipStreamSynchronize(c.stream));
    SX_TRY(down_vec3(c, c.hs.a, c.ax, c.ay, c.az));
    SX_TRY(down_scalar(c, c.hs.gpot, c.gpot));
    if (c.opt.mac != SERENX_MAC_GEOMETRIC) SX_TRY(down_scalar(c, c.hs.agravmag, c.agm));
    return SERENX_OK;
}

// ---- device-resident mode --------------------------------------------------
int serenx_load(serenx_ctx *ctx, int32_t n, const double *pos, const double *vel,
                const double *mass, const double *h) {
    return serenx_load_forest(ctx, 1, &n, pos, vel, mass, h);
}

int serenx_load_forest(serenx_ctx *ctx, int32_t ncloud, const int32_t *counts,
                       const double *pos, const double *vel, const double *mass,
                       const double *h) {
    if (!ctx || !counts || !pos || !vel || !mass || ncloud <= 0)
        return set_error(SERENX_EINVAL, "bad load args");
    long long total = 0;
    for (int k = 0; k < ncloud; k++) {
        if (counts[k] <= 0) return set_error(SERENX_EINVAL, "empty cloud");
        total += counts[k];
    }
    if (total >= (1LL << 27))
        return set_error(SERENX_EINVAL, "more than 2^27 - 1 particles on one device");
    const int n = (int)total;
    Ctx &c = *as_ctx(ctx);
    SX_HIP(hipSetDevice(c.device));
    step_free(c);
    serve_free(c);
    SX_TRY(ensure_capacity(c, n));
    c.n = n;
    SX_TRY(set_clouds(c, ncloud, counts));
    c.bound = false;
    c.tree_ok = c.sph_ok = false;
    SX_TRY(up_vec3(c, pos, c.x, c.y, c.z));
    SX_TRY(up_vec3(c, vel, c.vx, c.vy, c.vz));
    SX_HIP(hipMemcpy(c.m, mass, sizeof(double) * n, hipMemcpyHostToDevice));
    if (h)
        SX_HIP(hipMemcpy(c.h, h, sizeof(double) * n, hipMemcpyHostToDevice));
    else
        SX_HIP(hipMemset(c.h, 0, sizeof(double) * n));
    hipLaunchKernelGGL(k_init_state, dim3(nblk(n)), dim3(TB), 0, c.stream, n, c.porig, c.active,
                       c.rho, c.omega, c.zeta, c.ax, c.ay, c.az, c.gpot, c.agm, c.divv, c.press,
                       c.sound, c.drhodt);
    SX_HIP(hipStreamSynchronize(c.stream));
    return SERENX_OK;
}

int serenx_set_agravmag(serenx_ctx *ctx, const double *agravmag) {
    if (!ctx || !agravmag) return set_error(SERENX_EINVAL, "null argument");
    Ctx &c = *as_ctx(ctx);
    SX_HIP(hipSetDevice(c.device));
    SX_HIP(hipMemcpy(c.agm, agravmag, sizeof(double) * c.n, hipMemcpyHostToDevice));
    return SERENX_OK;
}

int serenx_force_eval(serenx_ctx *ctx, int32_t flags) {
    if (!ctx) return set_error(SERENX_EINVAL, "null context");
    Ctx &c = *as_ctx(ctx);
    if (c.n <= 0) return set_error(SERENX_ESTATE, "no particles loaded");
    SX_HIP(hipSetDevice(c.device));
    return run_force_eval(c, flags);
}

}  // extern "C"

namespace serenx {

namespace {

// Part A of a restock evaluation: stock the tree of the last build with the
// current positions (tree_update.F90's steps between rebuilds), the h
// solve, hmax and the neighbour rows; the walk's cell phase beside them on
// the gravity stream, joined at the end.  Launches only.
int eval_part_a(Ctx &c, int flags) {
    const bool grav = !(flags & SERENX_EVAL_NO_GRAV);
    SX_TRY(zero_words(c.wstat, 2 * CTR_WORDS, c.stream));
    SX_HIP(hipEventRecord(c.ev[0], c.stream));
    SX_TRY(stock_tree(c));
    SX_HIP(hipEventRecord(c.ev[1], c.stream));
    c.grav_timed = false;
    if (grav) SX_TRY(grav_launch(c));
    SX_TRY(run_sph_h(c));
    SX_TRY(update_hmax(c));
    SX_TRY(sph_neib_rows(c));
    if (grav) {
        SX_HIP(hipEventRecord(c.evjoin, c.gstream));
        SX_HIP(hipStreamWaitEvent(c.stream, c.evjoin, 0));
    }
    return SERENX_OK;
}

// Part B: thermal and hydro forces, the walk's pair phase beside them, and
// a_grav added to a.  Launches only.
int eval_part_b(Ctx &c, int flags) {
    const bool grav = !(flags & SERENX_EVAL_NO_GRAV);
    if (c.opt.sph == SERENX_SPH_STANDARD) SX_TRY(sph_density(c));
    c.sph_ok = true;
    SX_TRY(sph_thermal(c));
    SX_HIP(hipEventRecord(c.ev[2], c.stream));
    if (grav) SX_TRY(grav_pairs(c));
    if (flags & SERENX_EVAL_NO_HYDRO) {
        hipLaunchKernelGGL(k_zero_active, dim3(nblk(c.n)), dim3(TB), 0, c.stream, c.n, c.active,
                           c.ax, c.ay, c.az);
    } else {
        SX_TRY(run_hydro(c));
    }
    SX_HIP(hipEventRecord(c.ev[3], c.stream));
    if (grav) SX_TRY(grav_join(c));
    SX_HIP(hipEventRecord(c.ev[4], c.stream));
    return SERENX_OK;
}

// Records part (0: A, 1: B) into *exec.  Host state the launches set is set
// by the recording as by an eager run; a failed recording leaves no graph.
int capture_part(Ctx &c, int part, int flags, hipGraphExec_t *exec) {
    if (*exec) (void)hipGraphExecDestroy(*exec);
    *exec = nullptr;
    SX_HIP(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
    c.capturing = true;
    int rc = part == 0 ? eval_part_a(c, flags) : eval_part_b(c, flags);
    c.capturing = false;
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(c.stream, &g);
    if (rc == SERENX_OK && e != hipSuccess)
        rc = set_error(SERENX_EHIP, std::string("graph capture failed: ") + hipGetErrorString(e));
    if (rc == SERENX_OK) {
        const hipError_t ei = hipGraphInstantiate(exec, g, nullptr, nullptr, 0);
        if (ei != hipSuccess) {
            *exec = nullptr;
            rc = set_error(SERENX_EHIP, std::string("graph instantiation failed: ") +
                                            hipGetErrorString(ei));
        }
    }
    if (g) (void)hipGraphDestroy(g);
    return rc;
}

// A restock evaluation from two graphs: A (recorded once per graph epoch
// and flags), the host's look at the h-solve status and the rows above
// pp_limit (whose eager pass may reallocate their list, a new epoch for B),
// then B.  The graphs run the same kernels with the same arguments as the
// eager path, except the walk's queue order (tree order; the order does not
// change results).
int run_eval_graph(Ctx &c, int flags) {
    Ctx::Graphs &G = c.gr;
    if (!G.a || G.epoch_a != c.graph_epoch || G.flags_a != flags) {
        SX_TRY(capture_part(c, 0, flags, &G.a));
        G.epoch_a = c.graph_epoch;
        G.flags_a = flags;
        G.captures++;
    }
    SX_HIP(hipGraphLaunch(G.a, c.stream));
    c.gq.order_ready = false;   // graph A leaves the queue in tree order
    SX_HIP(hipStreamSynchronize(c.stream));
    SX_TRY(sph_neib_finish(c));
    if (!G.b || G.epoch_b != c.graph_epoch || G.flags_b != flags) {
        SX_TRY(capture_part(c, 1, flags, &G.b));
        G.epoch_b = c.graph_epoch;
        G.flags_b = flags;
        G.captures++;
    }
    SX_HIP(hipGraphLaunch(G.b, c.stream));
    G.launches++;
    c.grav_timed = !(flags & SERENX_EVAL_NO_GRAV);
    c.box_ok = true;
    c.sph_ok = true;
    return SERENX_OK;
}

}  // namespace

void graphs_free(Ctx &c) {
    if (c.gr.a) (void)hipGraphExecDestroy(c.gr.a);
    if (c.gr.b) (void)hipGraphExecDestroy(c.gr.b);
    c.gr.a = c.gr.b = nullptr;
    c.gr.epoch_a = c.gr.epoch_b = -1;
}

int run_force_eval(Ctx &c, int flags) {
    if (flags & (SERENX_EVAL_RESTOCK | SERENX_EVAL_GRAPH)) {
        if (!c.tree_ok)
            return set_error(SERENX_ESTATE, "a restock evaluation needs a tree built by an "
                                            "earlier evaluation");
        if (flags & SERENX_EVAL_HGUESS)
            return set_error(SERENX_EINVAL, "SERENX_EVAL_HGUESS needs a tree build");
        if ((flags & SERENX_EVAL_GRAPH) && c.opt.sph == SERENX_SPH_GRADH &&
            c.opt.hfind != SERENX_HFIND_CONSTANT) {
            GravGuard guard(c.gstream);
            guard.armed = true;
            SX_TRY(run_eval_graph(c, flags & ~SERENX_EVAL_GRAPH));
            guard.armed = false;
            return SERENX_OK;
        }
        GravGuard guard(c.gstream);
        guard.armed = true;
        SX_TRY(eval_part_a(c, flags));
        SX_HIP(hipStreamSynchronize(c.stream));
        SX_TRY(sph_neib_finish(c));
        SX_TRY(eval_part_b(c, flags));
        guard.armed = false;
        return SERENX_OK;
    }
    SX_TRY(zero_words(c.wstat, 2 * CTR_WORDS, c.stream));
    SX_HIP(hipEventRecord(c.ev[0], c.stream));
    SX_TRY(run_tree_update(c));
    if (flags & SERENX_EVAL_HGUESS) SX_TRY(hguess(c));
    SX_HIP(hipEventRecord(c.ev[1], c.stream));
    // sph_update (h), then the gravity walk on its own stream beside the
    // neighbour lists, thermal and hydro forces (sph_setup's order otherwise)
    const bool grav = !(flags & SERENX_EVAL_NO_GRAV);
    c.grav_timed = false;
    // the walk's cell phase beside the h solve, its pair phase beside the
    // neighbour lists, thermal and hydro forces
    GravGuard guard(c.gstream);
    if (grav) {
        guard.armed = true;
        SX_TRY(grav_launch(c));
    }
    SX_TRY(run_sph_h(c));
    // the pair phase needs the new h only; the per-level hmax launches (for
    // the neighbour walk) follow it, so the walk's stream does not wait
    if (grav) SX_TRY(grav_pairs(c));
    SX_TRY(update_hmax(c));
    SX_TRY(run_sph_lists(c));
    SX_TRY(sph_thermal(c));
    SX_HIP(hipEventRecord(c.ev[2], c.stream));
    if (flags & SERENX_EVAL_NO_HYDRO) {
        hipLaunchKernelGGL(k_zero_active, dim3(nblk(c.n)), dim3(TB), 0, c.stream, c.n, c.active,
                           c.ax, c.ay, c.az);
    } else {
        SX_TRY(run_hydro(c));
    }
    SX_HIP(hipEventRecord(c.ev[3], c.stream));
    if (grav) SX_TRY(grav_join(c));
    guard.armed = false;
    SX_HIP(hipEventRecord(c.ev[4], c.stream));
    return SERENX_OK;
}

}  // namespace serenx

extern "C" {

int serenx_graph_info(const serenx_ctx *ctx, int64_t *captures, int64_t *launches) {
    if (!ctx || !captures || !launches) return set_error(SERENX_EINVAL, "null argument");
    const Ctx &c = *as_ctx(ctx);
    *captures = c.gr.captures;
    *launches = c.gr.launches;
    return SERENX_OK;
}

int serenx_sync(serenx_ctx *ctx) {
    if (!ctx) return set_error(SERENX_EINVAL, "null context");
    Ctx &c = *as_ctx(ctx);
    SX_HIP(hipStreamSynchronize(c.stream));
    return SERENX_OK;
}

int serenx_last_timings(const serenx_ctx *ctx, double *ms4) {
    if (!ctx || !ms4) return set_error(SERENX_EINVAL, "null argument");
    const Ctx &c = *as_ctx(ctx);
    SX_HIP(hipEventSynchronize(c.ev[4]));
    for (int k = 0; k < 3; k++) {
        float ms = 0.0f;
        SX_HIP(hipEventElapsedTime(&ms, c.ev[k], c.ev[k + 1]));
        ms4[k] = ms;
    }
    // gravity: the walk's own span on the gravity stream (it overlaps the
    // SPH and hydro phases)
    float g1 = 0.0f, g2 = 0.0f;
    if (c.grav_timed) {
        SX_HIP(hipEventElapsedTime(&g1, c.evg[0], c.evg[1]));
        SX_HIP(hipEventElapsedTime(&g2, c.evg[2], c.evg[3]));
    }
    ms4[3] = g1 + g2;
    return SERENX_OK;
}

int serenx_walk_timings(const serenx_ctx *ctx, double *ms2) {
    if (!ctx || !ms2) return set_error(SERENX_EINVAL, "null argument");
    const Ctx &c = *as_ctx(ctx);
    ms2[0] = ms2[1] = 0.0;
    if (!c.grav_timed) return SERENX_OK;
    SX_HIP(hipEventSynchronize(c.evg[3]));
    float g1 = 0.0f, g2 = 0.0f;
    SX_HIP(hipEventElapsedTime(&g1, c.evg[0], c.evg[1]));
    SX_HIP(hipEventElapsedTime(&g2, c.evg[2], c.evg[3]));
    ms2[0] = g1;
    ms2[1] = g2;
    return SERENX_OK;
}

int serenx_interaction_counts(serenx_ctx *ctx, int64_t *counts4) {
    if (!ctx || !counts4) return set_error(SERENX_EINVAL, "null argument");
    Ctx &c = *as_ctx(ctx);
    std::vector<unsigned long long> rows(CTR_WORDS);
    unsigned long long v[4];
    SX_HIP(hipStreamSynchronize(c.stream));
    SX_HIP(hipMemcpy(rows.data(), c.icount, CTR_WORDS * sizeof(unsigned long long),
                     hipMemcpyDeviceToHost));
    ctr_fold(rows.data(), v, 4);
    for (int k = 0; k < 4; k++) counts4[k] = (int64_t)v[k];
    return SERENX_OK;
}

int serenx_walk_stats(serenx_ctx *ctx, int64_t *stats16) {
    if (!ctx || !stats16) return set_error(SERENX_EINVAL, "null argument");
    Ctx &c = *as_ctx(ctx);
    std::vector<unsigned long long> rows(CTR_WORDS);
    unsigned long long v[16];
    SX_HIP(hipStreamSynchronize(c.stream));
    SX_HIP(hipMemcpy(rows.data(), c.wstat, CTR_WORDS * sizeof(unsigned long long),
                     hipMemcpyDeviceToHost));
    ctr_fold(rows.data(), v, 16, 0, WSTAT_MAX_MASK);
    for (int k = 0; k < 16; k++) stats16[k] = (int64_t)v[k];
    return SERENX_OK;
}

int serenx_fetch(serenx_ctx *ctx, int32_t *porig, double *pos, double *vel, double *mass,
                 double *h, double *zeta, double *rho, double *omega, double *div_v,
                 double *drhodt, double *press, double *sound, double *acc, double *gpot,
                 double *agravmag, int32_t *pptot, int32_t *pplist) {
    if (!ctx) return set_error(SERENX_EINVAL, "null context");
    Ctx &c = *as_ctx(ctx);
    SX_HIP(hipSetDevice(c.device));
    SX_HIP(hipStreamSynchronize(c.stream));
    const size_t nb = sizeof(double) * c.n;
    if (porig) SX_HIP(hipMemcpy(porig, c.porig, sizeof(int) * c.n, hipMemcpyDeviceToHost));
    if (pos) SX_TRY(down_vec3(c, pos, c.x, c.y, c.z));
    if (vel) SX_TRY(down_vec3(c, vel, c.vx, c.vy, c.vz));
    if (acc) SX_TRY(down_vec3(c, acc, c.ax, c.ay, c.az));
    struct {
        double *hp;
        const double *dp;
    } sc[] = {{mass, c.m},     {h, c.h},       {zeta, c.zeta},   {rho, c.rho},
              {omega, c.omega}, {div_v, c.divv}, {drhodt, c.drhodt}, {press, c.press},
              {sound, c.sound}, {gpot, c.gpot}, {agravmag, c.agm}};
    for (auto &e : sc)
        if (e.hp) SX_HIP(hipMemcpy(e.hp, e.dp, nb, hipMemcpyDeviceToHost));
    if (pptot) SX_HIP(hipMemcpy(pptot, c.pptot, sizeof(int) * c.n, hipMemcpyDeviceToHost));
    if (pplist) {
        const size_t cnt = (size_t)c.n * c.pp_limit;
        SX_HIP(hipMemcpy(pplist, c.pplist, sizeof(int) * cnt, hipMemcpyDeviceToHost));
        std::vector<int> tot(c.n);
        SX_HIP(hipMemcpy(tot.data(), c.pptot, sizeof(int) * c.n, hipMemcpyDeviceToHost));
        for (int p = 0; p < c.n; p++) {
            int *row = pplist + (size_t)p * c.pp_limit;
            const int k = tot[p] <= c.pp_limit ? tot[p] : 0;
            for (int i = 0; i < c.pp_limit; i++) row[i] = (i < k) ? row[i] + 1 : 0;
        }
    }
    return SERENX_OK;
}

int serenx_direct_gravity(serenx_ctx *ctx, int32_t nt, const int32_t *targets, double *acc,
                          double *pot) {
    if (!ctx || !targets || !acc || !pot || nt <= 0) return set_error(SERENX_EINVAL, "bad args");
    Ctx &c = *as_ctx(ctx);
    if (c.n <= 0) return set_error(SERENX_ESTATE, "no particles loaded");
    for (int i = 0; i < nt; i++)
        if (targets[i] < 0 || targets[i] >= c.n)
            return set_error(SERENX_EINVAL, "direct-gravity target out of range");
    SX_HIP(hipSetDevice(c.device));
    int *dt = nullptr;
    double *da = nullptr, *dp = nullptr;
    SX_HIP(hipMalloc(&dt, sizeof(int) * nt));
    SX_HIP(hipMalloc(&da, sizeof(double) * 3 * nt));
    SX_HIP(hipMalloc(&dp, sizeof(double) * nt));
    int rc = SERENX_OK;
    if (hipMemcpy(dt, targets, sizeof(int) * nt, hipMemcpyHostToDevice) != hipSuccess)
        rc = set_error(SERENX_EHIP, "copy of targets failed");
    if (rc == SERENX_OK) rc = direct_gravity(c, nt, dt, da, dp);
    if (rc == SERENX_OK && hipStreamSynchronize(c.stream) != hipSuccess)
        rc = set_error(SERENX_EHIP, "direct gravity failed");
    if (rc == SERENX_OK) {
        (void)hipMemcpy(acc, da, sizeof(double) * 3 * nt, hipMemcpyDeviceToHost);
        (void)hipMemcpy(pot, dp, sizeof(double) * nt, hipMemcpyDeviceToHost);
    }
    (void)hipFree(dt);
    (void)hipFree(da);
    (void)hipFree(dp);
    return rc;
}

int serenx_set_energy(serenx_ctx *ctx, const double *u) {
    if (!ctx || !u) return set_error(SERENX_EINVAL, "null argument");
    Ctx &c = *as_ctx(ctx);
    if (c.n <= 0) return set_error(SERENX_ESTATE, "no particles loaded");
    SX_HIP(hipSetDevice(c.device));
    SX_HIP(hipStreamSynchronize(c.stream));
    SX_HIP(hipMemcpy(c.u, u, sizeof(double) * c.n, hipMemcpyHostToDevice));
    return SERENX_OK;
}

int serenx_fetch_energy(serenx_ctx *ctx, double *u, double *du_dt) {
    if (!ctx) return set_error(SERENX_EINVAL, "null context");
    Ctx &c = *as_ctx(ctx);
    SX_HIP(hipSetDevice(c.device));
    SX_HIP(hipStreamSynchronize(c.stream));
    if (u) SX_HIP(hipMemcpy(u, c.u, sizeof(double) * c.n, hipMemcpyDeviceToHost));
    if (du_dt) SX_HIP(hipMemcpy(du_dt, c.dudt, sizeof(double) * c.n, hipMemcpyDeviceToHost));
    return SERENX_OK;
}

int serenx_pow_third(serenx_ctx *ctx, int32_t n, const double *x, double *y) {
    if (!ctx || !x || !y || n < 0) return set_error(SERENX_EINVAL, "bad args");
    if (n == 0) return SERENX_OK;
    Ctx &c = *as_ctx(ctx);
    SX_HIP(hipSetDevice(c.device));
    double *d = nullptr;
    SX_HIP(hipMalloc(&d, sizeof(double) * 2 * (size_t)n));
    int rc = SERENX_OK;
    if (hipMemcpy(d, x, sizeof(double) * n, hipMemcpyHostToDevice) != hipSuccess)
        rc = set_error(SERENX_EHIP, "copy of x failed");
    if (rc == SERENX_OK) rc = eval_pow_third(c, n, d, d + n);
    if (rc == SERENX_OK && hipStreamSynchronize(c.stream) != hipSuccess)
        rc = set_error(SERENX_EHIP, "pow_third failed");
    if (rc == SERENX_OK &&
        hipMemcpy(y, d + n, sizeof(double) * n, hipMemcpyDeviceToHost) != hipSuccess)
        rc = set_error(SERENX_EHIP, "copy of y failed");
    (void)hipFree(d);
    return rc;
}

int serenx_tree_sizes(const serenx_ctx *ctx, int32_t *ncells, int32_t *ltot) {
    if (!ctx) return set_error(SERENX_EINVAL, "null context");
    const Ctx &c = *as_ctx(ctx);
    if (ncells) *ncells = c.nc;
    if (ltot) *ltot = c.ltot;
    return SERENX_OK;
}

int serenx_export_tree(const serenx_ctx *ctx, int32_t *leaf, int32_t *plist, int32_t *nextcell,
                       int32_t *ifopen, double *r, double *m, double *dminsqd, double *mac,
                       double *q, double *s, double *hr, double *rmax, double *hrangemax,
                       int32_t *first_cell, int32_t *last_cell) {
    if (!ctx) return set_error(SERENX_EINVAL, "null context");
    Ctx &c = *const_cast<Ctx *>(as_ctx(ctx));
    if (!c.tree_ok) return set_error(SERENX_ESTATE, "no tree built");
    SX_HIP(hipSetDevice(c.device));
    // hrangemax from the current h: a step with no active particles leaves
    // it to the next force evaluation, which recomputes it before use
    SX_TRY(update_hmax(c));
    SX_HIP(hipStreamSynchronize(c.stream));
    const int nc = c.nc;
    std::vector<int> b2d(nc), d2b(nc), cleaf(nc);
    std::vector<GNode> g(nc);
    std::vector<HNode> hh(nc);
    SX_HIP(hipMemcpy(b2d.data(), c.bfs2dfs, sizeof(int) * nc, hipMemcpyDeviceToHost));
    SX_HIP(hipMemcpy(d2b.data(), c.dfs2bfs, sizeof(int) * nc, hipMemcpyDeviceToHost));
    SX_HIP(hipMemcpy(cleaf.data(), c.c_leaf, sizeof(int) * nc, hipMemcpyDeviceToHost));
    SX_HIP(hipMemcpy(g.data(), c.gn, sizeof(GNode) * nc, hipMemcpyDeviceToHost));
    std::vector<double> gs((size_t)nc * 10);
    SX_HIP(hipMemcpy(gs.data(), c.gs, sizeof(double) * gs.size(), hipMemcpyDeviceToHost));
    SX_HIP(hipMemcpy(hh.data(), c.hn, sizeof(HNode) * nc, hipMemcpyDeviceToHost));
    const int sentinel = 2 * c.n + 1;   // cmax_grav + 1 (BHgrav_build.F90:75)
    for (int b = 0; b < nc; b++) {
        const int d = b2d[b];
        const GNode &gg = g[d];
        if (leaf) leaf[b] = cleaf[b];
        if (plist)
            for (int i = 0; i < LEAFMAX; i++)
                plist[LEAFMAX * b + i] = (i < cleaf[b]) ? gg.first + i + 1 : 0;
        if (nextcell) nextcell[b] = (gg.next >= nc) ? sentinel : d2b[gg.next];
        if (ifopen) ifopen[b] = (cleaf[b] == 0 && d + 1 < nc) ? d2b[d + 1] : 0;
        for (int k = 0; k < 3; k++) {
            if (r) r[3 * b + k] = gg.r[k];
            if (hr) hr[3 * b + k] = hh[d].r[k];
        }
        if (m) m[b] = gg.m;
        if (dminsqd) dminsqd[b] = gg.dminsqd;
        if (mac) mac[b] = gg.mac;
        if (q)
            for (int k = 0; k < 5; k++) q[5 * b + k] = gg.q[k];
        if (s)
            for (int k = 0; k < 10; k++)
                s[10 * b + k] = (k == 9 && c.opt.multipole == SERENX_OCTUPOLE_EXACT)
                                    ? 0.5 * gs[10 * (size_t)d + k] : gs[10 * (size_t)d + k];
        if (rmax) rmax[b] = hh[d].rmax;
        if (hrangemax) hrangemax[b] = hh[d].hrangemax;
    }
    for (int L = 0; L <= LMAX; L++) {
        const bool has = L <= c.ltot && c.level_cnt[L] > 0;
        if (first_cell) first_cell[L] = has ? c.level_off[L] : 0;
        if (last_cell) last_cell[L] = has ? c.level_off[L] + c.level_cnt[L] - 1 : 0;
    }
    return SERENX_OK;
}

void serenx_default_step_options(serenx_step_options *opt) {
    if (!opt) return;
    *opt = serenx_step_options{};
    opt->integrator = SERENX_INT_RK;     // Makefile:45 SPH_INTEGRATION = RK
    opt->nlevels = 1;                    // default_parameters.F90:66
    opt->nbuildstep = 8;                 // default_parameters.F90:117
    opt->accel_mult = 0.5;               // default_parameters.F90:63
    opt->courant_mult = 0.2;             // default_parameters.F90:62
    opt->levels = SERENX_LEVELS_RESTRICTED;   // Makefile:100 TIMESTEP = RESTRICTED
    opt->dt_fixed = 1.0;                      // default_parameters.F90:67
    opt->neib_check = SERENX_NEIB_CHECK_END;  // Makefile:101 CHECK_NEIB_TIMESTEP = 1
}

int serenx_step_begin(serenx_ctx *ctx, const serenx_step_options *opt) {
    if (!ctx || !opt) return set_error(SERENX_EINVAL, "null argument");
    Ctx &c = *as_ctx(ctx);
    SX_HIP(hipSetDevice(c.device));
    return step_begin(c, *opt);
}

int serenx_step(serenx_ctx *ctx, int32_t count) {
    if (!ctx || count < 0) return set_error(SERENX_EINVAL, "bad step arguments");
    Ctx &c = *as_ctx(ctx);
    SX_HIP(hipSetDevice(c.device));
    return step_run(c, count);
}

int serenx_cloud_offsets(const serenx_ctx *ctx, int32_t *offsets) {
    if (!ctx || !offsets) return set_error(SERENX_EINVAL, "null argument");
    const Ctx &c = *as_ctx(ctx);
    for (size_t k = 0; k < c.cloud_off_h.size(); k++) offsets[k] = c.cloud_off_h[k];
    return SERENX_OK;
}

int serenx_step_get_info(const serenx_ctx *ctx, serenx_step_info *info) {
    if (!ctx || !info) return set_error(SERENX_EINVAL, "null argument");
    const Ctx &c = *as_ctx(ctx);
    if (!c.st.on) return set_error(SERENX_ESTATE, "no step state (serenx_step_begin)");
    const Ctx::Step &s = c.st;
    *info = serenx_step_info{};
    info->n = s.n;
    info->nsteps = s.nsteps;
    info->nresync = s.nresync;
    info->level_max = s.level_max;
    info->level_step = s.level_step;
    info->nbuild = s.nbuild;
    info->nstock = s.nstock;
    info->nactive = s.nactive;
    info->time = s.time;
    info->timestep = s.timestep;
    info->dt_max = s.dt_max;
    return SERENX_OK;
}

int serenx_fetch_step(serenx_ctx *ctx, int32_t *nlevel, int64_t *nlast, double *laststep,
                      int32_t *active, double *r_old, double *v_old, double *v_half,
                      int32_t *nminneib) {
    if (!ctx) return set_error(SERENX_EINVAL, "null context");
    Ctx &c = *as_ctx(ctx);
    SX_HIP(hipSetDevice(c.device));
    return step_fetch(c, nlevel, nlast, laststep, active, r_old, v_old, v_half, nminneib);
}

int serenx_set_precision(serenx_ctx *ctx, int32_t precision) {
    if (!ctx) return set_error(SERENX_EINVAL, "null context");
    if (precision != SERENX_FP64 && precision != SERENX_MIXED)
        return set_error(SERENX_EINVAL, "precision must be SERENX_FP64 or SERENX_MIXED");
    Ctx &c = *as_ctx(ctx);
    SX_HIP(hipSetDevice(c.device));
    const bool to_mixed = precision == SERENX_MIXED && c.opt.precision != SERENX_MIXED;
    if (precision != c.opt.precision) c.graph_epoch++;
    c.opt.precision = precision;
    if (to_mixed) SX_TRY(refresh_f32_nodes(c));
    return SERENX_OK;
}

int serenx_diagnostics(serenx_ctx *ctx, double *out) {
    if (!ctx || !out) return set_error(SERENX_EINVAL, "null argument");
    Ctx &c = *as_ctx(ctx);
    if (c.n <= 0) return set_error(SERENX_ESTATE, "no particles loaded");
    SX_HIP(hipSetDevice(c.device));
    double *buf = nullptr;
    const size_t count = (size_t)DIAG_MAX_BLOCKS * DIAG_SUMS + DIAG_VALUES;
    SX_HIP(hipMalloc(&buf, sizeof(double) * count));
    int rc = diagnostics(c, buf, buf + (size_t)DIAG_MAX_BLOCKS * DIAG_SUMS);
    if (rc == SERENX_OK &&
        hipMemcpyAsync(out, buf + (size_t)DIAG_MAX_BLOCKS * DIAG_SUMS,
                       sizeof(double) * DIAG_VALUES, hipMemcpyDeviceToHost, c.stream) != hipSuccess)
        rc = set_error(SERENX_EHIP, "copy of the diagnostics failed");
    if (rc == SERENX_OK && hipStreamSynchronize(c.stream) != hipSuccess)
        rc = set_error(SERENX_EHIP, "diagnostics failed");
    (void)hipFree(buf);
    return rc;
}

int serenx_serve_begin(serenx_ctx *ctx, double accel_mult, double courant_mult) {
    if (!ctx) return set_error(SERENX_EINVAL, "null context");
    Ctx &c = *as_ctx(ctx);
    SX_HIP(hipSetDevice(c.device));
    return serve_begin(c, accel_mult, courant_mult);
}

int serenx_serve_step(serenx_ctx *ctx, int32_t ncloud, const uint8_t *select) {
    if (!ctx || !select) return set_error(SERENX_EINVAL, "null argument");
    Ctx &c = *as_ctx(ctx);
    SX_HIP(hipSetDevice(c.device));
    return serve_step(c, ncloud, select);
}

int serenx_serve_info(serenx_ctx *ctx, int32_t ncloud, double *dt, double *time) {
    if (!ctx) return set_error(SERENX_EINVAL, "null context");
    Ctx &c = *as_ctx(ctx);
    SX_HIP(hipSetDevice(c.device));
    return serve_info(c, ncloud, dt, time);
}

int serenx_set_tuning(serenx_ctx *ctx, int32_t key, double value) {
    if (!ctx) return set_error(SERENX_EINVAL, "null context");
    Ctx &c = *as_ctx(ctx);
    c.graph_epoch++;   // recorded graphs hold launch shapes the knobs choose
    switch (key) {
    case SERENX_TUNE_HSOLVE_SLOTS:
        if (!(value >= 1.0 && value <= (double)HL_SLOTS) || value != (double)(int)value)
            return set_error(SERENX_EINVAL, "h-solve list slots must be an integer in 1..96");
        c.hl_cap = (int)value;
        return SERENX_OK;
    case SERENX_TUNE_MIXED_ABSERROR:
        if (!(value >= 0.0) || !std::isfinite(value))
            return set_error(SERENX_EINVAL, "mixed-walk MAC tolerance must be >= 0");
        c.mixed_abserror = value;
        return SERENX_OK;
    case SERENX_TUNE_NEIB_GROUP:
        if (value != 0.0 && value != 1.0)
            return set_error(SERENX_EINVAL, "neighbour walk mode must be 0 or 1");
        c.neib_group = value != 0.0;
        return SERENX_OK;
    case SERENX_TUNE_WALK_QUEUE:
        if (!(value >= 0.0 && value <= 64.0) || value != (double)(int)value)
            return set_error(SERENX_EINVAL, "walk queue blocks per CU must be an integer in 0..64");
        c.walk_queue = (int)value;
        c.gq.order_ready = false;   // its keys depend on the queue's use
        return SERENX_OK;
    case SERENX_TUNE_HSOLVE_WAVES:
        if (value != -1.0 && value != 0.0 && value != 8.0)
            return set_error(SERENX_EINVAL, "h-solve waves per SIMD must be -1, 0 or 8");
        c.hsolve_waves = (int)value;
        return SERENX_OK;
    case SERENX_TUNE_NEAR_CAP:
        if (!(value >= 1.0 && value <= (double)PL_CAP) || value != (double)(int)value)
            return set_error(SERENX_EINVAL, "near-list capacity must be an integer in 1..4096");
        c.near_cap = (int)value;
        return SERENX_OK;
    case SERENX_TUNE_PAIR_BLOCKS:
        if (!(value >= 0.0 && value <= 64.0) || value != (double)(int)value)
            return set_error(SERENX_EINVAL, "pair-phase blocks per CU must be an integer in 0..64");
        c.pair_blocks = (int)value;
        return SERENX_OK;
    case SERENX_TUNE_NEIB_LEAVES:
        if (!(value >= 0.0 && value <= (double)NB_CAP) || value != (double)(int)value)
            return set_error(SERENX_EINVAL, "neighbour-walk leaf list length must be an integer in 0..512");
        c.nb_cap = (int)value;
        return SERENX_OK;
    default:
        return set_error(SERENX_EINVAL, "unknown tuning key");
    }
}

int serenx_get_tuning(const serenx_ctx *ctx, int32_t key, double *value) {
    if (!ctx || !value) return set_error(SERENX_EINVAL, "null argument");
    const Ctx &c = *as_ctx(ctx);
    switch (key) {
    case SERENX_TUNE_HSOLVE_SLOTS: *value = c.hl_cap; return SERENX_OK;
    case SERENX_TUNE_MIXED_ABSERROR:
        *value = c.mixed_abserror > 0.0 ? c.mixed_abserror : c.opt.abserror;
        return SERENX_OK;
    case SERENX_TUNE_WALK_QUEUE: *value = c.walk_queue; return SERENX_OK;
    case SERENX_TUNE_NEIB_GROUP: *value = c.neib_group ? 1.0 : 0.0; return SERENX_OK;
    case SERENX_TUNE_HSOLVE_WAVES: *value = c.hsolve_waves; return SERENX_OK;
    case SERENX_TUNE_NEAR_CAP: *value = c.near_cap; return SERENX_OK;
    case SERENX_TUNE_NEIB_LEAVES: *value = c.nb_cap; return SERENX_OK;
    case SERENX_TUNE_PAIR_BLOCKS: *value = c.pair_blocks; return SERENX_OK;
    default: return set_error(SERENX_EINVAL, "unknown tuning key");
    }
}

int serenx_memory(const serenx_ctx *ctx, int64_t *free_bytes, int64_t *total_bytes,
                  int64_t *ctx_bytes) {
    if (!ctx) return set_error(SERENX_EINVAL, "null context");
    const Ctx &c = *as_ctx(ctx);
    SX_HIP(hipSetDevice(c.device));
    size_t fr = 0, tot = 0;
    SX_HIP(hipMemGetInfo(&fr, &tot));
    if (free_bytes) *free_bytes = (int64_t)fr;
    if (total_bytes) *total_bytes = (int64_t)tot;
    if (ctx_bytes) *ctx_bytes = (int64_t)ctx_device_bytes(c);
    return SERENX_OK;
}

}  // extern "C"
