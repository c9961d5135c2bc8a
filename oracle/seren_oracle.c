/*
 * seren_oracle.c -- CPU restatement of SEREN's per-step force evaluation.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker the HIP path is compared
 * against (tests/, __graft_entry__.smoke(), bench.py's cpu_baseline leg).  The
 * product library never links, loads or calls it.
 *
 * It restates, in plain C99 with FP64 arithmetic in the reference's operation
 * order (compiled with -ffp-contract=off so no FMA contraction happens), the
 * option set GRAD_H_SPH + AB viscosity + isothermal + KS gravity +
 * BH tree + REORDER_PARTICLES + NEIGHBOUR_LISTS, NDIM=3, DOUBLE precision,
 * with QUADRUPOLE or OCTUPOLE moments, the GEOMETRIC, GADGET, GADGET2 or
 * EIGEN MAC, and the M4 or quintic kernel, each with or without the
 * Thomas-Couchman gradient (TC_KERNEL).
 *
 * Arrays are 1-based like the Fortran module arrays they mirror: index 0 of
 * every particle array is unused; tree cell ids start at 0 (root).
 *
 * Pinning (tests/test_oracle_*.py):
 *   kernel tables       bit-exact vs the flang-built reference dump + closed
 *                       forms (normalisation, wh identity)
 *   tree build/reorder  bit-exact vs reference (topology, BFS ids, plist)
 *   stock (m, COM, Q, S, dminsqd, mac, bbox)  bit-exact vs reference
 *   hguess, h, rho, omega, zeta, div_v        bit-exact vs reference
 *   pptot / pplist      bit-exact vs reference + brute-force O(N^2) set test
 *   hydro accel         bit-exact vs reference + momentum conservation
 *   gravity accel/pot   bit-exact vs reference + FP64 direct-sum agreement
 *   direct sum          vs reference + point-mass closed form
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define NDIM 3
#define LEAFMAX 8
#define LMAX 30
#define NCHILD 8
#define KERNTOT 10000
#define SMALL_NUMBER 1.0e-20
#define BIG_NUMBER 9.9e20
#define HMULT 1.1
#define H_CON 0.02
#define GRADH_ITERATION_MAX 20
#define LISTSIZE 10000
#define GLISTSIZE 10000
#define ETA_SQD 0.01
#define PI_ 3.1415926536
#define INVPI 0.318309886
#define ONETHIRD 0.333333333
#define TWOTHIRDS_DP 0.666666666666666666

enum { MAC_GEOMETRIC = 0, MAC_GADGET = 1, MAC_GADGET2 = 2, MAC_EIGEN = 3 };

typedef struct {
    int leaf, plist[LEAFMAX], nextcell, ifopen;
    double r[3], m, dminsqd, mac, q[5], s[10];
} gnode;
typedef struct {
    int leaf, nextcell, ifopen, plist[LEAFMAX];
    double r[3], rmax, hrangemax;
} hnode;
typedef struct {
    int pfirst, plast, leaf, nextcell, ifopen, plist[LEAFMAX], childof[NCHILD];
    double r[3];
} tnode;
typedef struct {
    double hmax, bbmin[3], bbmax[3];
} snode;

typedef struct oracle_ctx {
    int n, pp_gather, pp_limit, octupole, mac, kernel;
    int sph_std; /* SPH=STANDARD with HFIND=NUMBER (else GRAD_H_SPH) */
    double hmin_const; /* > 0: HFIND=CONSTANT, every h is this value */
    /* kernel range constants, macros.h:101-118 */
    double halfkerntot, kernrange, kernrangesqd, invkernrange;
    int energy;          /* THERMAL=ENERGY_EQN (INTERNAL_ENERGY) */
    double gamma;
    double *u, *du_dt;   /* 1-based, allocated by oracle_set_energy */
    double h_fac, thetamaxsqd, abserror, alpha, beta, rextent;
    int *porig;
    double (*parray)[6]; /* x y z m h zeta */
    double (*v)[3], (*a)[3];
    double *gpot, *agravmag, *rho, *omega, *press, *sound, *div_v, *drhodt;
    int *pptot, *pplist; /* pplist[(p-1)*pp_limit + (i-1)] */
    int cmax_grav, cmax_hydro, ctot_grav, ltot_grav, ctot_hydro, ltot_hydro;
    int first_cell[LMAX + 1], last_cell[LMAX + 1];
    gnode *BHgrav;
    hnode *BHhydro;
    tnode *BHtemp;
    snode *BHstock;
    int *cellof, *BHnextptcl, *whichchild;
    double w0[KERNTOT + 1], w1[KERNTOT + 1], w2[KERNTOT + 1], w3[KERNTOT + 1],
        w4[KERNTOT + 1], w5[KERNTOT + 1], w6[KERNTOT + 1], wh[KERNTOT + 1],
        wg[KERNTOT + 1];
    int err; /* 0 ok, 1 LMAX reached, 2 too few particles */
} oracle_ctx;

/* Exact parallel-axis shift of a child's traceless octupole (the library's
 * SERENX_OCTUPOLE_EXACT mode; no counterpart in the reference, whose shift
 * BHgrav_stock.F90:263-297 halves the quadrupole terms and takes 25 for 5 on
 * the xyz term).  From the moments' definitions,
 *   O_ijk = O'_ijk + m [15 d_i d_j d_k - 3 d^2 (d_i delta_jk + 2 perms)]
 *         + 5 (d_i Q'_jk + 2 perms) - 2 ((Q'd)_i delta_jk + 2 perms),
 * in the reference's s-units (s0 = O_xxx/3, s1 = O_xxy, s2 = O_xyy,
 * s3 = O_yyy/3, s4 = O_xzz, s5 = O_yzz, s6 = O_zzz/3, s7 = O_xxz, s8 = O_yyz)
 * with a stored cell holding s9 = 2 O_xyz, which makes the walk's expansion
 * (coefficient 1 on s9 xyz) the exact series.  Parity unpinned against the
 * reference (it has no such mode); pinned against moments summed directly
 * over each cell's particles and the far-field error order. */
static void oct_shift_exact(double *sc, const double *cs, const double *q, double mcc,
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

/* x**n for a small integer constant n.  The reference compiler expands
 * integer powers as a left-to-right product ((x*x)*x)*...; the kernel tables
 * only reproduce bit-for-bit with that association. */
static double ipow(double x, int n) {
    double y = x;
    for (int i = 1; i < n; i++) y = y * x;
    return n ? y : 1.0;
}

/* kernel.F90:14-127 (M4, GRAD_H_SPH, NDIM=3) */
static void kernel_m4(oracle_ctx *c) {
    const double invkerntot = 1.0 / (double)KERNTOT;
    const int kern1 = 5000; /* int(HALFKERNTOT) */
    const int kerntc = (int)((double)KERNTOT / 3.0); /* kernel.F90:37 */
    const double ndim = 3.0;
    for (int i = 0; i <= KERNTOT; i++) {
        double s = (c->kernrange * (double)i + 1.0) * invkerntot;
        double saux = c->kernrange - s;
        if (i < kern1) {
            if (c->kernel == 1 && i < kerntc) /* TC_KERNEL, kernel.F90:56-62 */
                c->w1[i] = -1.0;
            else
                c->w1[i] = -(3.0 * s) + 2.25 * s * s;
            c->w0[i] = 1.0 - 1.5 * s * s + 0.75 * ipow(s, 3);
            c->w2[i] = (4.0 / 3.0) * ipow(s, 3) - 1.2 * ipow(s, 5) + 0.5 * ipow(s, 6);
            c->w3[i] = (7.0 / 5.0) * s - (2.0 / 3.0) * ipow(s, 3) + 0.3 * ipow(s, 5) -
                       0.1 * ipow(s, 6);
            c->w4[i] = -(3.0 * s) + 2.25 * s * s;
            c->w5[i] = (4.0 / 3.0) * s - 1.2 * ipow(s, 3) + 0.5 * ipow(s, 4);
            c->w6[i] = 7.0 / 5.0 - (2.0 / 3.0) * ipow(s, 2) + 0.3 * ipow(s, 4) -
                       0.1 * ipow(s, 5);
            c->wh[i] = -ndim + 1.5 * (ndim + 2.0) * s * s - 0.75 * (ndim + 3.0) * ipow(s, 3);
            c->wg[i] = 1.4 - 2.0 * s * s + 1.5 * ipow(s, 4) - 0.6 * ipow(s, 5);
        } else {
            c->w0[i] = 0.25 * ipow(saux, 3);
            c->w1[i] = -(0.75 * saux * saux);
            c->w2[i] = (8.0 / 3.0) * ipow(s, 3) - 3.0 * ipow(s, 4) + 1.2 * ipow(s, 5) -
                       (1.0 / 6.0) * ipow(s, 6) - (1.0 / 15.0);
            c->w3[i] = -(1.0 / 15.0) + (8.0 / 5.0) * s - (4.0 / 3.0) * ipow(s, 3) +
                       ipow(s, 4) - 0.3 * ipow(s, 5) + (1.0 / 30.0) * ipow(s, 6);
            c->w4[i] = -(0.75 * saux * saux);
            c->w5[i] = (8.0 / 3.0) * s - 3.0 * ipow(s, 2) + 1.2 * ipow(s, 3) -
                       (1.0 / 6.0) * ipow(s, 4) - 1.0 / (15.0 * ipow(s, 2));
            c->w6[i] = -(1.0 / (15.0 * s)) + 8.0 / 5.0 - (4.0 / 3.0) * ipow(s, 2) +
                       ipow(s, 3) - 0.3 * ipow(s, 4) + (1.0 / 30.0) * ipow(s, 5);
            c->wh[i] = -2.0 * ndim + 3.0 * (ndim + 1.0) * s - 1.50 * (ndim + 2.0) * s * s +
                       0.25 * (ndim + 3.0) * ipow(s, 3);
            c->wg[i] = 1.6 - 4.0 * s * s + 4.0 * ipow(s, 3) - 1.5 * ipow(s, 4) +
                       0.2 * ipow(s, 5);
        }
        c->w0[i] = c->w0[i] * INVPI;
        c->w1[i] = c->w1[i] * INVPI;
        c->w4[i] = c->w4[i] * INVPI;
        c->wh[i] = c->wh[i] * INVPI;
    }
}

/* kernel.F90:130-236 (QUINTIC, GRAD_H_SPH, NDIM=3): three polynomial
 * pieces split at entries int(ONETHIRD KERNTOT) and int(2 ONETHIRD KERNTOT);
 * w2 = w3 = 0; the TC gradient is the constant -55.2040 below
 * int(0.75929848074 KERNTOT / 3). */
static void kernel_quintic(oracle_ctx *c) {
    const double invkerntot = 1.0 / (double)KERNTOT;
    const int kern1 = (int)(ONETHIRD * (double)KERNTOT);
    const int kern2 = (int)(2.0 * ONETHIRD * (double)KERNTOT);
    const int kerntc = (int)(0.75929848074 * (double)KERNTOT / 3.0);
    const double d = 3.0; /* NDIMPR */
    for (int i = 0; i <= KERNTOT; i++) {
        const double s = (3.0 * (double)i + 1.0) * invkerntot;
        const double s2 = ipow(s, 2), s3 = ipow(s, 3), s4 = ipow(s, 4), s5 = ipow(s, 5),
                     s6 = ipow(s, 6), s7 = ipow(s, 7);
        double w0, w1, w4, w5, w6, wh, wg;
        if (i < kern1) {
            w4 = -(120.0 * s) + 120.0 * s3 - 50.0 * s4;
            w1 = (c->kernel == 3 && i < kerntc) ? -55.2040 : w4;
            w0 = 66.0 - 60.0 * s * s + 30.0 * s4 - 10.0 * s5;
            w5 = 22.0 * s - 12.0 * s3 + (30.0 / 7.0) * s5 - (5.0 / 4.0) * s6;
            w6 = -(11.0 * s * s) + 3.0 * s4 - (5.0 / 7.0) * s6 + (5.0 / 28.0) * s7 +
                 (478.0 / 14.0);
            wh = -(66.0 * d) + 60.0 * (d + 2.0) * s * s - 30.0 * (d + 4.0) * s4 +
                 10.0 * (d + 5.0) * s5;
            wg = 33.0 * s * s - 15.0 * s4 + 5.0 * s6 - (10.0 / 7.0) * s7 - (478.0 / 14.0);
        } else if (i < kern2) {
            w0 = 51.0 + 75.0 * s - 210.0 * s * s + 150.0 * s3 - 45.0 * s4 + 5.0 * s5;
            w4 = 75.0 - 420.0 * s + 450.0 * s * s - 180.0 * s3 + 25.0 * s4;
            w1 = w4;
            w5 = 17.0 * s + (75.0 / 4.0) * s2 - 42.0 * s3 + 25.0 * s4 - (45.0 / 7.0) * s5 +
                 (5.0 / 8.0) * s6 + (5.0 / 56.0) / s2;
            w6 = -((17.0 / 2.0) * s * s) - (25.0 / 4.0) * s3 + (21.0 / 2.0) * s4 - 5.0 * s5 +
                 (15.0 / 14.0) * s6 - (5.0 / 56.0) * s7 + (473.0 / 14.0) + (5.0 / 56.0) / s;
            wh = -(51.0 * d) - 75.0 * (d + 1.0) * s + 210.0 * (d + 2.0) * s2 -
                 150.0 * (d + 3.0) * s3 + 45.0 * (d + 4.0) * s4 - 5.0 * (d + 5.0) * s5;
            wg = (51.0 / 2.0) * s * s + 25.0 * s3 - (105.0 / 2.0) * s4 + 30.0 * s5 -
                 (15.0 / 2.0) * s6 + (5.0 / 7.0) * s7 - (473.0 / 14.0);
        } else {
            w0 = 243.0 - 405.0 * s + 270.0 * s * s - 90.0 * s3 + 15.0 * s4 - s5;
            w4 = -405.0 + 540.0 * s - 270.0 * s * s + 60.0 * s3 - 5.0 * s4;
            w1 = w4;
            w5 = 81.0 * s - (405.0 / 4.0) * s2 + 54.0 * s3 - 15.0 * s4 + (15.0 / 7.0) * s5 -
                 (1.0 / 8.0) * s6 - (507.0 / 56.0) / s2;
            w6 = -((81.0 / 2.0) * s * s) + (135.0 / 4.0) * s3 - (27.0 / 2.0) * s4 + 3.0 * s5 -
                 (5.0 / 14.0) * s6 + (1.0 / 56.0) * s7 + (729.0 / 14.0) - (507.0 / 56.0) / s;
            wh = -(243.0 * d) + 405.0 * (d + 1.0) * s - 270.0 * (d + 2.0) * s2 +
                 90.0 * (d + 3.0) * s3 - 15.0 * (d + 4.0) * s4 + (d + 5.0) * s5;
            wg = (243.0 / 2.0) * s * s - 135.0 * s3 + (135.0 / 2.0) * s4 - 18.0 * s5 +
                 (5.0 / 2.0) * s6 - (1.0 / 7.0) * s7 - (729.0 / 14.0);
        }
        /* normalisation for NDIM = 3, evaluated left to right */
        c->w0[i] = w0 * INVPI * 3.0 / 359.0;
        c->w1[i] = w1 * INVPI * 3.0 / 359.0;
        c->w2[i] = 0.0;
        c->w3[i] = 0.0;
        c->w4[i] = w4 * INVPI * 3.0 / 359.0;
        c->w5[i] = w5 * 12.0 / 359.0;
        c->w6[i] = w6 * 12.0 / 359.0;
        c->wh[i] = wh * INVPI * 3.0 / 359.0;
        c->wg[i] = -wg * 12.0 / 359.0;
    }
}

static void oracle_kernel(oracle_ctx *c) {
    if (c->kernel >= 2)
        kernel_quintic(c);
    else
        kernel_m4(c);
}

#define PX(p, k) (c->parray[(p)][(k)])
#define MASS 3
#define SMOO 4
#define ZETA 5

/* ------------------------------------------------------------------------ */
void oracle_destroy(oracle_ctx *c) {
    if (!c) return;
    free(c->u); free(c->du_dt);
    free(c->porig); free(c->parray); free(c->v); free(c->a); free(c->gpot);
    free(c->agravmag); free(c->rho); free(c->omega); free(c->press);
    free(c->sound); free(c->div_v); free(c->drhodt); free(c->pptot);
    free(c->pplist); free(c->BHgrav); free(c->BHhydro); free(c->BHtemp);
    free(c->BHstock); free(c->cellof); free(c->BHnextptcl); free(c->whichchild);
    free(c);
}

/* ref setup: allocate_memory.F90:43-271, types.F90, initialize_seren_variables_1
 * .F90:30-39 (pp_gather), initialize_sph_variables_1.F90:33-110, kernel.F90 */
oracle_ctx *oracle_create(int n, const double *pos, const double *mass,
                          const double *h, const double *vel, int octupole,
                          int mac, double h_fac, double thetamaxsqd,
                          double abserror, double alpha, int kernel, int sph) {
    oracle_ctx *c = (oracle_ctx *)calloc(1, sizeof(oracle_ctx));
    c->n = n;
    c->sph_std = sph;
    c->kernel = kernel; /* 0 M4, 1 M4 TC, 2 QUINTIC, 3 QUINTIC TC */
    if (kernel >= 2) {
        c->halfkerntot = 3333.0;
        c->kernrange = 3.0;
        c->kernrangesqd = 9.0;
        c->invkernrange = 0.3333333333333;
    } else {
        c->halfkerntot = 5000.0;
        c->kernrange = 2.0;
        c->kernrangesqd = 4.0;
        c->invkernrange = 0.5;
    }
    c->octupole = octupole;
    c->mac = mac;
    c->h_fac = h_fac;
    c->thetamaxsqd = thetamaxsqd;
    c->abserror = abserror;
    c->alpha = alpha;
    c->beta = 2.0 * alpha;
    c->pp_gather = (int)((4.0 * PI_ / 3.0) * ipow(c->kernrange * h_fac, 3));
    c->pp_limit = 2 * c->pp_gather;
    c->porig = (int *)calloc(n + 1, sizeof(int));
    c->parray = calloc(n + 1, sizeof(*c->parray));
    c->v = calloc(n + 1, sizeof(*c->v));
    c->a = calloc(n + 1, sizeof(*c->a));
    c->gpot = calloc(n + 1, sizeof(double));
    c->agravmag = calloc(n + 1, sizeof(double));
    c->rho = calloc(n + 1, sizeof(double));
    c->omega = calloc(n + 1, sizeof(double));
    c->press = calloc(n + 1, sizeof(double));
    c->sound = calloc(n + 1, sizeof(double));
    c->div_v = calloc(n + 1, sizeof(double));
    c->drhodt = calloc(n + 1, sizeof(double));
    c->pptot = calloc(n + 1, sizeof(int));
    c->pplist = calloc((size_t)n * c->pp_limit + 1, sizeof(int));
    c->cmax_grav = 2 * n;
    c->cmax_hydro = 2 * n;
    c->BHgrav = calloc(c->cmax_grav + 1, sizeof(gnode));
    c->BHhydro = calloc(c->cmax_hydro + 1, sizeof(hnode));
    c->BHtemp = calloc(c->cmax_hydro + 1, sizeof(tnode));
    c->BHstock = calloc(c->cmax_hydro + 1, sizeof(snode));
    c->cellof = calloc(n + 1, sizeof(int));
    c->BHnextptcl = calloc(n + 2, sizeof(int));
    c->whichchild = calloc(n + 1, sizeof(int));
    for (int p = 1; p <= n; p++) {
        c->porig[p] = p;
        for (int k = 0; k < 3; k++) {
            PX(p, k) = pos[3 * (p - 1) + k];
            c->v[p][k] = vel[3 * (p - 1) + k];
        }
        PX(p, MASS) = mass[p - 1];
        PX(p, SMOO) = h[p - 1];
        PX(p, ZETA) = 0.0;
        c->rho[p] = 1.0;
        c->omega[p] = 1.0;
    }
    oracle_kernel(c);
    return c;
}

int oracle_error(const oracle_ctx *c) { return c->err; }

/* bounding_box.F90:11-43 */
static void bounding_box(oracle_ctx *c, int pstart, int pend, double rmax[3],
                         double rmin[3]) {
    for (int k = 0; k < 3; k++) {
        rmax[k] = -BIG_NUMBER;
        rmin[k] = BIG_NUMBER;
    }
    for (int p = pstart; p <= pend; p++)
        for (int k = 0; k < 3; k++) {
            rmax[k] = fmax(rmax[k], PX(p, k));
            rmin[k] = fmin(rmin[k], PX(p, k));
        }
    double e = rmax[0] - rmin[0];
    for (int k = 1; k < 3; k++)
        if (rmax[k] - rmin[k] > e) e = rmax[k] - rmin[k];
    c->rextent = e;
}

/* BHgrav_build.F90:12-309 */
static int BHgrav_build(oracle_ctx *c) {
    int n = c->n;
    int *templist = (int *)malloc(sizeof(int) * (c->cmax_grav + 1));
    for (int p = 1; p <= n; p++) {
        c->cellof[p] = 0;
        c->BHnextptcl[p] = p + 1;
    }
    for (int cc = 0; cc <= c->cmax_grav; cc++) {
        for (int k = 0; k < NCHILD; k++) c->BHtemp[cc].childof[k] = -2;
        c->BHtemp[cc].pfirst = -1;
        c->BHtemp[cc].plast = -1;
        c->BHtemp[cc].leaf = 0;
    }
    c->BHtemp[0].pfirst = 1;
    c->BHtemp[0].plast = n;
    for (int l = 0; l <= LMAX; l++) c->first_cell[l] = c->last_cell[l] = 0;
    double rmax[3], rmin[3];
    bounding_box(c, 1, n, rmax, rmin);
    double ext = rmax[0] - rmin[0];
    for (int k = 1; k < 3; k++)
        if (rmax[k] - rmin[k] > ext) ext = rmax[k] - rmin[k];
    double cellsize = 0.5 * ext;
    for (int k = 0; k < 3; k++) c->BHtemp[0].r[k] = 0.5 * (rmin[k] + rmax[k]);
    int l = 0;
    c->ctot_grav = 0;
    c->BHtemp[0].leaf = 0;
    c->BHtemp[0].nextcell = c->cmax_grav + 1;
    c->BHgrav[0].leaf = 0;
    c->BHgrav[0].nextcell = c->cmax_grav + 1;
    for (int i = 0; i < LEAFMAX; i++) c->BHgrav[0].plist[i] = 0;
    int nlist = 1;
    templist[0] = 0;
    for (;;) {
        cellsize = 0.5 * cellsize;
        for (int j = 0; j < nlist; j++) { /* :107-142 */
            int cc = templist[j];
            int p = c->BHtemp[cc].pfirst;
            double *rc = c->BHtemp[cc].r;
            for (;;) {
                int ckid = 1;
                if (PX(p, 0) > rc[0]) ckid += 1;
                if (PX(p, 1) > rc[1]) ckid += 2;
                if (PX(p, 2) > rc[2]) ckid += 4;
                c->whichchild[p] = ckid;
                c->BHtemp[cc].childof[ckid - 1] = -1;
                if (p == c->BHtemp[cc].plast) break;
                p = c->BHnextptcl[p];
            }
        }
        for (int j = 0; j < nlist; j++) { /* :149-157 */
            int cc = templist[j];
            for (int k = 0; k < NCHILD; k++)
                if (c->BHtemp[cc].childof[k] == -1) {
                    c->ctot_grav++;
                    c->BHtemp[cc].childof[k] = c->ctot_grav;
                }
        }
        for (int j = 0; j < nlist; j++) { /* :165-255 */
            int cc = templist[j];
            int nptclin[NCHILD] = {0};
            int p = c->BHtemp[cc].pfirst;
            for (;;) {
                int ckid = c->BHtemp[cc].childof[c->whichchild[p] - 1];
                c->cellof[p] = ckid;
                if (c->BHtemp[ckid].pfirst == -1)
                    c->BHtemp[ckid].pfirst = p;
                else
                    c->BHnextptcl[c->BHtemp[ckid].plast] = p;
                c->BHtemp[ckid].plast = p;
                nptclin[c->whichchild[p] - 1]++;
                if (p == c->BHtemp[cc].plast) break;
                p = c->BHnextptcl[p];
            }
            int nkids = 0, clist[NCHILD];
            double rc[3] = {c->BHtemp[cc].r[0], c->BHtemp[cc].r[1], c->BHtemp[cc].r[2]};
            for (int k = 1; k <= NCHILD; k++) {
                if (nptclin[k - 1] == 0) continue;
                int ckid = c->BHtemp[cc].childof[k - 1];
                if (nptclin[k - 1] > 0 && nptclin[k - 1] <= LEAFMAX) {
                    c->BHtemp[ckid].leaf = nptclin[k - 1];
                    int i = 0;
                    int q = c->BHtemp[ckid].pfirst;
                    for (;;) {
                        c->BHtemp[ckid].plist[i++] = q;
                        if (q == c->BHtemp[ckid].plast) break;
                        q = c->BHnextptcl[q];
                    }
                } else {
                    c->BHtemp[ckid].leaf = 0;
                }
                c->BHtemp[ckid].r[0] = (k == 1 || k == 3 || k == 5 || k == 7)
                                           ? rc[0] - cellsize : rc[0] + cellsize;
                c->BHtemp[ckid].r[1] = (k == 1 || k == 2 || k == 5 || k == 6)
                                           ? rc[1] - cellsize : rc[1] + cellsize;
                c->BHtemp[ckid].r[2] = (k < 5) ? rc[2] - cellsize : rc[2] + cellsize;
                clist[nkids++] = ckid;
            }
            c->BHtemp[cc].ifopen = clist[0];
            for (int k = 0; k < nkids - 1; k++) c->BHtemp[clist[k]].nextcell = clist[k + 1];
            c->BHtemp[clist[nkids - 1]].nextcell = c->BHtemp[cc].nextcell;
            c->BHgrav[cc].ifopen = c->BHtemp[cc].ifopen;
            for (int k = 0; k < nkids; k++) {
                int ck = clist[k];
                c->BHgrav[ck].leaf = c->BHtemp[ck].leaf;
                c->BHgrav[ck].nextcell = c->BHtemp[ck].nextcell;
                c->BHgrav[ck].ifopen = c->BHtemp[ck].ifopen;
                memcpy(c->BHgrav[ck].plist, c->BHtemp[ck].plist, sizeof(int) * LEAFMAX);
            }
        }
        l++;
        c->first_cell[l] = c->last_cell[l - 1] + 1;
        c->last_cell[l] = c->ctot_grav;
        nlist = 0;
        int alldone = 1;
        for (int cc = c->first_cell[l]; cc <= c->last_cell[l]; cc++)
            if (c->BHtemp[cc].leaf == 0) {
                alldone = 0;
                templist[nlist++] = cc;
            }
        if (alldone) break;
        if (l >= LMAX) { /* :289 'Reached maximum level' */
            free(templist);
            c->err = 1;
            return 1;
        }
    }
    c->ltot_grav = l;
    free(templist);
    return 0;
}

/* BH_reorder_particles.F90:51-78 + reorder_array.F90 (all-gas case) */
static void BH_reorder_particles(oracle_ctx *c) {
    int n = c->n;
    int *dummy = (int *)malloc(sizeof(int) * (n + 1));
    for (int p = 1; p <= n; p++) dummy[p] = p;
    int gascount = 0, cc = 0;
    for (;;) {
        gnode *g = &c->BHgrav[cc];
        if (g->leaf > 0) {
            for (int i = 0; i < g->leaf; i++) {
                int p = g->plist[i];
                gascount++;
                dummy[gascount] = p;
                g->plist[i] = gascount;
            }
            cc = g->nextcell;
        } else if (g->leaf == 0) {
            cc = g->ifopen;
        } else {
            cc = g->nextcell;
        }
        if (cc > c->ctot_grav) break;
    }
#define PERMUTE(arr, type, width)                                              \
    do {                                                                       \
        type *tmp = (type *)malloc(sizeof(type) * (width) * (n + 1));          \
        memcpy(tmp, (arr), sizeof(type) * (width) * (n + 1));                  \
        for (int p = 1; p <= n; p++)                                           \
            memcpy(((type *)(arr)) + (size_t)(width)*p,                        \
                   tmp + (size_t)(width)*dummy[p], sizeof(type) * (width));    \
        free(tmp);                                                             \
    } while (0)
    PERMUTE(c->porig, int, 1);
    PERMUTE(c->parray, double, 6);
    PERMUTE(c->v, double, 3);
    PERMUTE(c->a, double, 3);
    PERMUTE(c->gpot, double, 1);
    PERMUTE(c->agravmag, double, 1);
    PERMUTE(c->rho, double, 1);
    PERMUTE(c->drhodt, double, 1);
    PERMUTE(c->press, double, 1);
    PERMUTE(c->sound, double, 1);
    PERMUTE(c->div_v, double, 1);
    PERMUTE(c->omega, double, 1);
    if (c->energy) {   /* reorder_array.F90:98-104 */
        PERMUTE(c->u, double, 1);
        PERMUTE(c->du_dt, double, 1);
    }
#undef PERMUTE
    free(dummy);
}

/* BHgrav_stock.F90:11-358 */
static void BHgrav_stock(oracle_ctx *c) {
    for (int cc = 0; cc <= c->ctot_grav; cc++) {
        c->BHgrav[cc].dminsqd = 0.0;
        c->BHstock[cc].hmax = 0.0;
        for (int k = 0; k < 3; k++) {
            c->BHstock[cc].bbmin[k] = BIG_NUMBER;
            c->BHstock[cc].bbmax[k] = -BIG_NUMBER;
        }
    }
    const double daux = 10.0 * pow(1.7976931348623157e308, -0.2);
    for (int l = c->ltot_grav; l >= 0; l--) {
        int lo = (l == 0) ? 0 : c->first_cell[l], hi = (l == 0) ? 0 : c->last_cell[l];
        for (int ci = lo; ci <= hi; ci++) {
            gnode *g = &c->BHgrav[ci];
            snode *st = &c->BHstock[ci];
            double mc = 0.0, rc[3] = {0, 0, 0}, dr[3], drsqd, qc[5], sc[10];
            if (g->leaf == 1) {
                int p = g->plist[0];
                for (int k = 0; k < 3; k++) rc[k] = PX(p, k);
                for (int k = 0; k < 3; k++) g->r[k] = rc[k];
                g->m = PX(p, MASS);
                for (int k = 0; k < 5; k++) g->q[k] = 0.0;
                for (int k = 0; k < 10; k++) g->s[k] = 0.0;
                st->hmax = PX(p, SMOO);
                for (int k = 0; k < 3; k++) st->bbmin[k] = st->bbmax[k] = rc[k];
            } else if (g->leaf > 0) {
                int npart = g->leaf;
                for (int i = 0; i < npart; i++) {
                    int p = g->plist[i];
                    double mp = PX(p, MASS);
                    mc = mc + mp;
                    st->hmax = fmax(st->hmax, PX(p, SMOO));
                    for (int k = 0; k < 3; k++) rc[k] = rc[k] + mp * PX(p, k);
                    for (int k = 0; k < 3; k++) {
                        st->bbmax[k] = fmax(st->bbmax[k], PX(p, k));
                        st->bbmin[k] = fmin(st->bbmin[k], PX(p, k));
                    }
                }
                for (int k = 0; k < 3; k++) rc[k] = rc[k] / mc;
                memset(sc, 0, sizeof(sc));
                memset(qc, 0, sizeof(qc));
                for (int i = 0; i < npart; i++) {
                    int p = g->plist[i];
                    for (int k = 0; k < 3; k++) dr[k] = PX(p, k) - rc[k];
                    double mp = PX(p, MASS);
                    drsqd = dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2];
                    qc[0] = qc[0] + mp * (3.0 * dr[0] * dr[0] - drsqd);
                    qc[1] = qc[1] + mp * (3.0 * dr[0] * dr[1]);
                    qc[2] = qc[2] + mp * (3.0 * dr[1] * dr[1] - drsqd);
                    qc[3] = qc[3] + mp * (3.0 * dr[2] * dr[0]);
                    qc[4] = qc[4] + mp * (3.0 * dr[2] * dr[1]);
                    if (c->octupole) {
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
                g->m = mc;
                for (int k = 0; k < 3; k++) g->r[k] = rc[k];
                for (int k = 0; k < 5; k++) g->q[k] = qc[k];
                for (int k = 0; k < 10; k++) g->s[k] = c->octupole ? sc[k] : 0.0;
                if (c->octupole == 2) g->s[9] = 2.0 * sc[9];
            } else if (g->leaf == 0) {
                int cfirst = g->ifopen, cend = g->nextcell, cc = cfirst;
                for (;;) {
                    gnode *h = &c->BHgrav[cc];
                    snode *hs = &c->BHstock[cc];
                    if (h->leaf >= 0) {
                        mc = mc + h->m;
                        for (int k = 0; k < 3; k++) rc[k] = rc[k] + h->m * h->r[k];
                        st->hmax = fmax(hs->hmax, st->hmax);
                        for (int k = 0; k < 3; k++) {
                            st->bbmax[k] = fmax(hs->bbmax[k], st->bbmax[k]);
                            st->bbmin[k] = fmin(hs->bbmin[k], st->bbmin[k]);
                        }
                    }
                    cc = h->nextcell;
                    if (cc == cend) break;
                }
                if (mc > 0.0)
                    for (int k = 0; k < 3; k++) rc[k] = rc[k] / mc;
                memset(sc, 0, sizeof(sc));
                memset(qc, 0, sizeof(qc));
                cc = cfirst;
                for (;;) {
                    gnode *h = &c->BHgrav[cc];
                    double mcc = h->m, *q = h->q;
                    for (int k = 0; k < 5; k++) qc[k] = qc[k] + q[k];
                    for (int k = 0; k < 3; k++) dr[k] = h->r[k] - rc[k];
                    drsqd = dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2];
                    qc[0] = qc[0] + mcc * (3.0 * dr[0] * dr[0] - drsqd);
                    qc[1] = qc[1] + mcc * (3.0 * dr[0] * dr[1]);
                    qc[2] = qc[2] + mcc * (3.0 * dr[1] * dr[1] - drsqd);
                    qc[3] = qc[3] + mcc * (3.0 * dr[2] * dr[0]);
                    qc[4] = qc[4] + mcc * (3.0 * dr[2] * dr[1]);
                    if (c->octupole == 2) {
                        oct_shift_exact(sc, h->s, q, mcc, dr, drsqd);
                    } else if (c->octupole) { /* :263-297 (shift kept as written) */
                        for (int k = 0; k < 10; k++) sc[k] = sc[k] + h->s[k];
                        sc[0] = sc[0] + mcc * dr[0] * (5.0 * dr[0] * dr[0] - 3.0 * drsqd) +
                                1.5 * dr[0] * q[0] - dr[1] * q[1] - dr[2] * q[3];
                        sc[1] = sc[1] + mcc * dr[1] * (15.0 * dr[0] * dr[0] - 3.0 * drsqd) +
                                4.0 * dr[0] * q[1] + 2.5 * dr[1] * q[0] - dr[1] * q[2] -
                                dr[2] * q[4];
                        sc[2] = sc[2] + mcc * dr[0] * (15.0 * dr[1] * dr[1] - 3.0 * drsqd) +
                                4.0 * dr[1] * q[1] + 2.5 * dr[0] * q[2] - dr[0] * q[0] -
                                dr[2] * q[3];
                        sc[3] = sc[3] + mcc * dr[1] * (5.0 * dr[1] * dr[1] - 3.0 * drsqd) +
                                1.5 * dr[1] * q[2] - dr[0] * q[1] - dr[2] * q[4];
                        sc[4] = sc[4] + mcc * dr[0] * (15.0 * dr[2] * dr[2] - 3.0 * drsqd) +
                                4.0 * dr[2] * q[3] - 2.5 * dr[0] * (q[0] + q[2]) -
                                dr[0] * q[0] - dr[1] * q[1];
                        sc[5] = sc[5] + mcc * dr[1] * (15.0 * dr[2] * dr[2] - 3.0 * drsqd) +
                                4.0 * dr[2] * q[4] - 2.5 * dr[1] * (q[0] + q[2]) -
                                dr[0] * q[1] - dr[1] * q[2];
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
                    cc = h->nextcell;
                    if (cc == cend) break;
                }
                g->m = mc;
                for (int k = 0; k < 3; k++) g->r[k] = rc[k];
                for (int k = 0; k < 5; k++) g->q[k] = qc[k];
                for (int k = 0; k < 10; k++) g->s[k] = c->octupole ? sc[k] : 0.0;
                if (c->octupole == 2) g->s[9] = 2.0 * sc[9];
            }
            /* MAC quantities :317-348 */
            for (int k = 0; k < 3; k++)
                dr[k] = fmax(st->bbmax[k] - rc[k], rc[k] - st->bbmin[k]);
            double diagonalsqd = dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2];
            double rsize = sqrt(diagonalsqd) + c->kernrange * st->hmax;
            if (c->mac == MAC_GEOMETRIC) {
                g->dminsqd = (c->thetamaxsqd > 0.0) ? diagonalsqd / c->thetamaxsqd : BIG_NUMBER;
                g->mac = 0.0;
            } else if (c->mac == MAC_GADGET) {
                g->dminsqd = diagonalsqd;
                g->mac = pow(mc * diagonalsqd * diagonalsqd / c->abserror, ONETHIRD);
            } else if (c->mac == MAC_GADGET2) {
                g->dminsqd = diagonalsqd;
                g->mac = sqrt(mc * diagonalsqd / c->abserror);
            } else {
                /* eigenvalue_mac.F90:12-33 */
                const double *q = g->q;
                double pq = q[0] * q[2] - (q[0] + q[2]) * (q[0] + q[2]) - q[1] * q[1] -
                            q[3] * q[3] - q[4] * q[4];
                g->dminsqd = diagonalsqd;
                if (pq >= 0.0) {
                    g->mac = 0.0;
                } else {
                    double lambda = 2.0 * sqrt(-pq / 3.0);
                    g->mac = pow(0.5 * lambda / c->abserror, TWOTHIRDS_DP);
                }
            }
            g->dminsqd = fmax(g->dminsqd, rsize * rsize);
            if (g->dminsqd < daux) g->dminsqd = daux;
        }
    }
}

/* copy_BHgrav_to_BHhydro.F90:11-37 */
static void copy_BHgrav_to_BHhydro(oracle_ctx *c) {
    c->ctot_hydro = c->ctot_grav;
    c->ltot_hydro = c->ltot_grav;
    for (int cc = 0; cc <= c->ctot_hydro; cc++) {
        c->BHhydro[cc].leaf = c->BHgrav[cc].leaf;
        c->BHhydro[cc].nextcell = c->BHgrav[cc].nextcell;
        c->BHhydro[cc].ifopen = c->BHgrav[cc].ifopen;
        memcpy(c->BHhydro[cc].plist, c->BHgrav[cc].plist, sizeof(int) * LEAFMAX);
    }
}

/* BHhydro_stock.F90:10-174 */
static void BHhydro_stock(oracle_ctx *c) {
    for (int cc = 0; cc <= c->ctot_hydro; cc++)
        for (int k = 0; k < 3; k++) {
            c->BHstock[cc].bbmin[k] = BIG_NUMBER;
            c->BHstock[cc].bbmax[k] = -BIG_NUMBER;
        }
    for (int l = c->ltot_hydro; l >= 0; l--) {
        int lo = (l == 0) ? 0 : c->first_cell[l], hi = (l == 0) ? 0 : c->last_cell[l];
        for (int ci = lo; ci <= hi; ci++) {
            hnode *hn = &c->BHhydro[ci];
            snode *st = &c->BHstock[ci];
            double rc[3], halfbox[3], diagonalsqd;
            if (hn->leaf == 1) {
                int p = hn->plist[0];
                for (int k = 0; k < 3; k++) rc[k] = PX(p, k);
                for (int k = 0; k < 3; k++) hn->r[k] = rc[k];
                hn->rmax = 0.0;
                for (int k = 0; k < 3; k++) st->bbmin[k] = st->bbmax[k] = rc[k];
            } else if (hn->leaf > 0) {
                for (int i = 0; i < hn->leaf; i++) {
                    int p = hn->plist[i];
                    for (int k = 0; k < 3; k++) {
                        st->bbmax[k] = fmax(st->bbmax[k], PX(p, k));
                        st->bbmin[k] = fmin(st->bbmin[k], PX(p, k));
                    }
                }
                diagonalsqd = 0.0;
                for (int k = 0; k < 3; k++) {
                    rc[k] = 0.5 * (st->bbmax[k] + st->bbmin[k]);
                    halfbox[k] = st->bbmax[k] - rc[k];
                    diagonalsqd = diagonalsqd + halfbox[k] * halfbox[k];
                }
                for (int k = 0; k < 3; k++) hn->r[k] = rc[k];
                hn->rmax = sqrt(diagonalsqd);
            } else if (hn->leaf == 0) {
                int cc = hn->ifopen, cend = hn->nextcell;
                for (;;) {
                    if (c->BHhydro[cc].leaf >= 0)
                        for (int k = 0; k < 3; k++) {
                            st->bbmax[k] = fmax(c->BHstock[cc].bbmax[k], st->bbmax[k]);
                            st->bbmin[k] = fmin(c->BHstock[cc].bbmin[k], st->bbmin[k]);
                        }
                    cc = c->BHhydro[cc].nextcell;
                    if (cc == cend) break;
                }
                diagonalsqd = 0.0;
                for (int k = 0; k < 3; k++) {
                    rc[k] = 0.5 * (st->bbmax[k] + st->bbmin[k]);
                    halfbox[k] = st->bbmax[k] - rc[k];
                    diagonalsqd = diagonalsqd + halfbox[k] * halfbox[k];
                }
                for (int k = 0; k < 3; k++) hn->r[k] = rc[k];
                hn->rmax = sqrt(diagonalsqd);
            }
        }
    }
}

/* tree_update.F90:26-60 (build, reorder, grav stock, copy, hydro stock) */
int oracle_tree_update(oracle_ctx *c) {
    if (BHgrav_build(c)) return c->err;
    BH_reorder_particles(c);
    BHgrav_stock(c);
    copy_BHgrav_to_BHhydro(c);
    BHhydro_stock(c);
    return 0;
}

/* BHhydro_hguess.F90:11-163 */
void oracle_hguess(oracle_ctx *c) {
    int nc = c->ctot_hydro + 1;
    int *levelof = calloc(nc, sizeof(int)), *pcont = calloc(nc, sizeof(int));
    for (int l = c->ltot_hydro; l >= 0; l--) {
        int lo = (l == 0) ? 0 : c->first_cell[l], hi = (l == 0) ? 0 : c->last_cell[l];
        for (int ci = lo; ci <= hi; ci++) {
            levelof[ci] = l;
            hnode *hn = &c->BHhydro[ci];
            if (hn->leaf > 0) {
                pcont[ci] = hn->leaf;
            } else if (hn->leaf == 0) {
                int cc = hn->ifopen, cend = hn->nextcell;
                for (;;) {
                    pcont[ci] += pcont[cc];
                    cc = c->BHhydro[cc].nextcell;
                    if (cc == cend) break;
                }
            }
        }
    }
    int pstack[LMAX + 1];
    double rstack[LMAX + 1];
    int ci = 0;
    const int thresh = (int)(0.5 * (double)c->pp_gather);
    for (;;) {
        int l = levelof[ci];
        hnode *hn = &c->BHhydro[ci];
        pstack[l] = pcont[ci];
        rstack[l] = hn->rmax;
        if (hn->leaf > 0) {
            int laux = l;
            for (;;) {
                if (pstack[laux] > thresh || laux == 0) break;
                laux--;
            }
            double hguess = c->invkernrange * rstack[laux];
            for (int i = 0; i < hn->leaf; i++) PX(hn->plist[i], SMOO) = hguess;
            ci = hn->nextcell;
        } else if (hn->leaf == 0) {
            ci = hn->ifopen;
        } else {
            ci = hn->nextcell;
        }
        if (ci > c->ctot_hydro) break;
    }
    free(levelof);
    free(pcont);
}

/* BHhydrowalk_hgather.F90:11-77 (mode 0) and BHhydro_walk.F90:9-80 (mode 1).
 * Returns pp_pot, or -1 when the list would exceed pp_max. */
static int hydro_walk(const oracle_ctx *c, const double rp[3], double hrange,
                      int pp_max, int *list, int scatter) {
    int ci = 0, pp_pot = 0;
    for (;;) {
        const hnode *hn = &c->BHhydro[ci];
        double dr0 = hn->r[0] - rp[0], dr1 = hn->r[1] - rp[1], dr2 = hn->r[2] - rp[2];
        double drsqd = dr0 * dr0 + dr1 * dr1 + dr2 * dr2;
        double m1 = (hrange + hn->rmax) * (hrange + hn->rmax);
        int pass = drsqd <= m1;
        if (scatter && !pass) {
            double m2 = (hn->hrangemax + hn->rmax) * (hn->hrangemax + hn->rmax);
            pass = drsqd <= m2;
        }
        if (pass) {
            if (hn->leaf > 0) {
                if (pp_pot + hn->leaf > pp_max) return -1;
                for (int i = 0; i < hn->leaf; i++) list[pp_pot++] = hn->plist[i];
                ci = hn->nextcell;
            } else if (hn->leaf == 0) {
                ci = hn->ifopen;
            } else {
                ci = hn->nextcell;
            }
        } else {
            ci = hn->nextcell;
        }
        if (ci > c->cmax_hydro) break;
    }
    return pp_pot;
}

/* all_sph_gradh.F90:11-492 (GRAD_H_SPH, BH_TREE, GRAVITY; no MINIMUM_H) */
static void all_sph_gradh(oracle_ctx *c, int p, int *pp_potlist, int **bigbuf) {
    double rp[3] = {PX(p, 0), PX(p, 1), PX(p, 2)};
    double mp = PX(p, MASS), hp = PX(p, SMOO);
    double vp[3] = {c->v[p][0], c->v[p][1], c->v[p][2]};
    double rhotemp, omega_p, div_v_p, zeta_p, invhp, hfactor_p, hrangesqd;
    int iteration = 0, pp_pot = 0, pp_max = (LISTSIZE < c->n) ? LISTSIZE : c->n;
    int *list = pp_potlist;
    double h_old = hp, h_new = hp, h_high = 0.0, h_upper = c->rextent, h_lower = 0.0;
    invhp = 1.0 / hp;
    for (;;) {
        for (;;) {
            if (hp > h_high) {
                pp_pot = 0;
                h_high = hp * HMULT;
                pp_pot = hydro_walk(c, rp, c->kernrange * h_high, pp_max, list, 0);
            }
            if (pp_pot < 0) {
                h_high = 0.0;
                pp_pot = 0;
                pp_max = c->n;
                if (!*bigbuf) *bigbuf = (int *)malloc(sizeof(int) * (c->n + 1));
                list = *bigbuf;
            } else if (pp_pot < c->pp_gather) {
                h_high = 0.0;
                pp_pot = 0;
                hp = hp * HMULT;
            }
            if (pp_pot > 0) break;
        }
        h_new = hp;
        invhp = 1.0 / hp;
        hfactor_p = pow(invhp, 3.0);
        hrangesqd = c->kernrangesqd * hp * hp;
        rhotemp = mp * c->w0[0] * hfactor_p;
        omega_p = mp * hfactor_p * invhp * c->wh[0];
        for (int i = 0; i < pp_pot; i++) {
            int pp = list[i];
            if (pp == p) continue;
            double dr0 = PX(pp, 0) - rp[0], dr1 = PX(pp, 1) - rp[1], dr2 = PX(pp, 2) - rp[2];
            double drsqd = dr0 * dr0 + dr1 * dr1 + dr2 * dr2;
            if (drsqd >= hrangesqd) continue;
            double drmag = sqrt(drsqd) + SMALL_NUMBER;
            double skern = c->halfkerntot * drmag * invhp;
            int kern = (int)skern;
            if (kern >= KERNTOT) continue;
            double mpp = PX(pp, MASS);
            rhotemp = rhotemp + mpp * hfactor_p * c->w0[kern];
            omega_p = omega_p + mpp * hfactor_p * invhp * c->wh[kern];
        }
        omega_p = 1.0 + ((hp * omega_p) / (3.0 * rhotemp));
        if (fabs(hp - c->h_fac * pow(mp / rhotemp, ONETHIRD)) / hp < H_CON && iteration >= 1)
            break;
        if (iteration < GRADH_ITERATION_MAX) {
            h_new = c->h_fac * pow(mp / rhotemp, ONETHIRD);
        } else if (iteration == GRADH_ITERATION_MAX) {
            h_upper = c->rextent;
            h_lower = 0.0;
            h_new = 0.5 * (h_lower + h_upper);
        } else if (iteration < 2 * GRADH_ITERATION_MAX) {
            double fb;
            if (rhotemp > SMALL_NUMBER)
                fb = pow(hp, 3.0) - (pow(c->h_fac, 3.0)) * mp / rhotemp;
            else
                fb = 1.0;
            if (fb > 0.0)
                h_upper = hp;
            else
                h_lower = hp;
            h_new = 0.5 * (h_lower + h_upper);
        } else {
            h_new = h_old;
            hp = h_new;
            break;
        }
        hp = h_new;
        iteration++;
    }
    /* final gather :313-442 */
    hp = c->h_fac * pow(mp / rhotemp, ONETHIRD);
    invhp = 1.0 / hp;
    hfactor_p = pow(invhp, 3.0);
    hrangesqd = c->kernrangesqd * hp * hp;
    rhotemp = mp * c->w0[0] * hfactor_p;
    omega_p = mp * hfactor_p * invhp * c->wh[0];
    div_v_p = 0.0;
    zeta_p = mp * invhp * invhp * c->wg[0];
    for (int i = 0; i < pp_pot; i++) {
        int pp = list[i];
        if (pp == p) continue;
        double dr[3] = {PX(pp, 0) - rp[0], PX(pp, 1) - rp[1], PX(pp, 2) - rp[2]};
        double drmag = dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2];
        if (drmag > hrangesqd) continue;
        drmag = sqrt(drmag) + SMALL_NUMBER;
        double skern = c->halfkerntot * drmag * invhp;
        int kern = (int)skern;
        if (kern > KERNTOT) continue;
        double mpp = PX(pp, MASS);
        double dv[3] = {c->v[pp][0] - vp[0], c->v[pp][1] - vp[1], c->v[pp][2] - vp[2]};
        double dvdr = dv[0] * dr[0] + dv[1] * dr[1] + dv[2] * dr[2];
        rhotemp = rhotemp + mpp * hfactor_p * c->w0[kern];
        omega_p = omega_p + mpp * hfactor_p * invhp * c->wh[kern];
        div_v_p = div_v_p - mpp * dvdr * c->w4[kern] * hfactor_p * invhp / drmag;
        zeta_p = zeta_p + mpp * invhp * invhp * c->wg[kern];
    }
    omega_p = 1.0 + ((hp * omega_p) / (3.0 * rhotemp));
    div_v_p = div_v_p / rhotemp;
    c->rho[p] = rhotemp;
    c->omega[p] = omega_p;
    c->div_v[p] = div_v_p;
    c->drhodt[p] = -rhotemp * div_v_p / omega_p;
    PX(p, SMOO) = hp;
    PX(p, ZETA) = -(hp * zeta_p) / (3.0 * rhotemp * omega_p);
}

/* BHhydro_update_hmax.F90:11-92 */
static void BHhydro_update_hmax(oracle_ctx *c) {
    for (int l = c->ltot_hydro; l >= 0; l--) {
        int lo = (l == 0) ? 0 : c->first_cell[l], hi = (l == 0) ? 0 : c->last_cell[l];
        for (int ci = lo; ci <= hi; ci++) {
            hnode *hn = &c->BHhydro[ci];
            double hr = 0.0;
            if (hn->leaf == 1) {
                hn->hrangemax = c->kernrange * PX(hn->plist[0], SMOO);
            } else if (hn->leaf > 0) {
                for (int i = 0; i < hn->leaf; i++) hr = fmax(hr, c->kernrange * PX(hn->plist[i], SMOO));
                hn->hrangemax = hr;
            } else if (hn->leaf == 0) {
                int cc = hn->ifopen, cend = hn->nextcell;
                for (;;) {
                    hr = fmax(hr, c->BHhydro[cc].hrangemax);
                    cc = c->BHhydro[cc].nextcell;
                    if (cc == cend) break;
                }
                hn->hrangemax = hr;
            }
        }
    }
}

/* get_neib.F90:10-140; with store=0 it is get_neib_on_fly.F90:11-130 and
 * writes the full list to out (caller sizes it to n). */
static int get_neib(oracle_ctx *c, int p, double hp, int *potlist, int **bigbuf,
                    int *out, int outmax) {
    double rp[3] = {PX(p, 0), PX(p, 1), PX(p, 2)};
    double hrangesqd_p = c->kernrangesqd * hp * hp;
    int pp_max = (LISTSIZE < c->n) ? LISTSIZE : c->n;
    int *list = potlist;
    int pp_pot = hydro_walk(c, rp, c->kernrange * hp, pp_max, list, 1);
    if (pp_pot < 0) {
        if (!*bigbuf) *bigbuf = (int *)malloc(sizeof(int) * (c->n + 1));
        list = *bigbuf;
        pp_pot = hydro_walk(c, rp, c->kernrange * hp, c->n, list, 1);
    }
    int pp_tot = 0;
    for (int i = 0; i < pp_pot; i++) {
        int pp = list[i];
        if (pp == p) continue;
        double hrangesqd_pp = c->kernrangesqd * PX(pp, SMOO) * PX(pp, SMOO);
        double dr0 = PX(pp, 0) - rp[0], dr1 = PX(pp, 1) - rp[1], dr2 = PX(pp, 2) - rp[2];
        double drsqd = dr0 * dr0 + dr1 * dr1 + dr2 * dr2;
        if (drsqd <= hrangesqd_p || drsqd <= hrangesqd_pp) {
            pp_tot++;
            out[(pp_tot < outmax ? pp_tot : outmax) - 1] = pp;
        }
    }
    return pp_tot;
}

/* h_gather.F90:13-203 (HGATHER, BH_TREE, no MINIMUM_H, insertion sort):
 * brackets h between h_low and h_high = HMULT h_low until the pp_gather-th
 * neighbour lies between them, then sorts the band's r^2. */
static double h_gather(oracle_ctx *c, int p, double hguess) {
    const double rp[3] = {PX(p, 0), PX(p, 1), PX(p, 2)};
    const double kr2 = c->kernrangesqd;
    int pp_pot = 0, pp_max = LISTSIZE;
    double h_old = hguess, h_high = HMULT * h_old, h_low = h_old / HMULT, h_search = 0.0;
    double hrangesqd_high = kr2 * h_high * h_high, hrangesqd_low;
    int *hl = (int *)malloc(sizeof(int) * pp_max), *pot = (int *)malloc(sizeof(int) * pp_max);
    double *rs = (double *)malloc(sizeof(double) * pp_max);
    int pp_low = 0, pp_high = 0, pp_tot;
    if (c->pp_gather >= c->n) { /* the reference would search forever */
        c->err = 2;
        free(hl); free(pot); free(rs);
        return hguess;
    }
    for (;;) {
        for (;;) {
            if (h_high > h_search) {
                h_search = HMULT * h_high;
                pp_pot = hydro_walk(c, rp, c->kernrange * h_search, pp_max, pot, 0);
            }
            if (pp_pot < 0) { /* list overflow: retry with room for every particle */
                h_search = 0.0;
                pp_pot = 0;
                pp_max = c->n;
                free(hl); free(pot); free(rs);
                hl = (int *)malloc(sizeof(int) * pp_max);
                pot = (int *)malloc(sizeof(int) * pp_max);
                rs = (double *)malloc(sizeof(double) * pp_max);
            } else if (pp_pot <= c->pp_gather) {
                pp_pot = 0;
                h_search = 0.0;
                h_low = h_high;
                h_high = HMULT * h_high;
                hrangesqd_high = kr2 * h_high * h_high;
            }
            if (pp_pot > 0) break;
        }
        pp_high = 0;
        pp_low = 0;
        hrangesqd_high = kr2 * h_high * h_high;
        hrangesqd_low = kr2 * h_low * h_low;
        for (int i = 0; i < pp_pot; i++) {
            int pp = pot[i];
            if (pp == p) continue;
            double dr0 = PX(pp, 0) - rp[0], dr1 = PX(pp, 1) - rp[1], dr2 = PX(pp, 2) - rp[2];
            double drsqd = dr0 * dr0 + dr1 * dr1 + dr2 * dr2;
            if (drsqd < hrangesqd_high) {
                if (drsqd < hrangesqd_low) {
                    pp_low++;
                } else {
                    hl[pp_high] = pp;
                    rs[pp_high] = drsqd;
                    pp_high++;
                }
            }
            if (pp_high == pp_max) break;
        }
        pp_tot = pp_low + pp_high;
        if (pp_tot < c->pp_gather) {
            h_low = h_high;
            h_high = HMULT * h_high;
            hrangesqd_high = kr2 * h_high * h_high;
        } else if (pp_low >= c->pp_gather) {
            h_high = h_low;
            h_low = h_low / HMULT;
        } else {
            break;
        }
    }
    /* insertion_sort.F90:4-30 */
    for (int j = 1; j < pp_high; j++) {
        int iaux = hl[j];
        double raux = rs[j];
        int i = j - 1;
        for (; i >= 0; i--) {
            if (rs[i] <= raux) break;
            rs[i + 1] = rs[i];
            hl[i + 1] = hl[i];
        }
        rs[i + 1] = raux;
        hl[i + 1] = iaux;
    }
    double h = c->invkernrange * sqrt(rs[c->pp_gather - pp_low - 1]);
    free(hl); free(pot); free(rs);
    return h;
}

/* all_sph.F90:11-167 (standard SPH): rho, div v and drho/dt over the gather
 * neighbours in list order */
static void all_sph(oracle_ctx *c, int p, int *potlist, int **big, int *onfly) {
    double rp[3] = {PX(p, 0), PX(p, 1), PX(p, 2)};
    double hp = PX(p, SMOO), invhp = 1.0 / hp, hfactor = pow(invhp, 3.0);
    double rhotemp = PX(p, MASS) * c->w0[0] * hfactor;
    double vp[3] = {c->v[p][0], c->v[p][1], c->v[p][2]};
    double div_v_p = 0.0;
    int pp_numb = c->pptot[p];
    const int *lst;
    if (pp_numb <= c->pp_limit) {
        lst = &c->pplist[(size_t)(p - 1) * c->pp_limit];
    } else { /* gather_neib_on_fly: the gather subset comes in the same order */
        pp_numb = get_neib(c, p, hp, potlist, big, onfly, c->n);
        lst = onfly;
    }
    for (int i = 0; i < pp_numb; i++) {
        int pp = lst[i];
        double dv[3] = {c->v[pp][0] - vp[0], c->v[pp][1] - vp[1], c->v[pp][2] - vp[2]};
        double mpp = PX(pp, MASS);
        double dr[3] = {PX(pp, 0) - rp[0], PX(pp, 1) - rp[1], PX(pp, 2) - rp[2]};
        double drsqd = dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2];
        double dvdr = dv[0] * dr[0] + dv[1] * dr[1] + dv[2] * dr[2];
        double drmag = sqrt(drsqd) + SMALL_NUMBER;
        if (drmag >= c->kernrange * hp) continue;
        double invdrmag = 1.0 / drmag;
        int kern = (int)(c->halfkerntot * drmag * invhp);
        if (kern > KERNTOT) kern = KERNTOT;
        rhotemp = rhotemp + mpp * hfactor * c->w0[kern];
        div_v_p = div_v_p - mpp * dvdr * c->w4[kern] * hfactor * invhp * invdrmag;
    }
    c->rho[p] = rhotemp;
    div_v_p = div_v_p / rhotemp;
    c->div_v[p] = div_v_p;
    c->drhodt[p] = -rhotemp * div_v_p;
}

static void sph_update_standard(oracle_ctx *c) {
    int n = c->n;
    for (int p = 1; p <= n; p++) { /* each h from its own old value */
        if (c->hmin_const > 0.0) { /* sph_update.F90:107-109, HFIND=CONSTANT */
            PX(p, SMOO) = c->hmin_const;
            continue;
        }
        PX(p, SMOO) = h_gather(c, p, PX(p, SMOO));
        if (c->err) return;
    }
    BHhydro_update_hmax(c);
#pragma omp parallel
    {
        int *buf = (int *)malloc(sizeof(int) * (LISTSIZE + 1));
        int *tmp = (int *)malloc(sizeof(int) * (c->pp_limit + 1));
        int *big = NULL;
#pragma omp for schedule(dynamic, 64)
        for (int p = 1; p <= n; p++) {
            int tot = get_neib(c, p, PX(p, SMOO), buf, &big, tmp, c->pp_limit);
            c->pptot[p] = tot;
            if (tot <= c->pp_limit)
                memcpy(&c->pplist[(size_t)(p - 1) * c->pp_limit], tmp, sizeof(int) * tot);
        }
        free(buf);
        free(tmp);
        free(big);
    }
#pragma omp parallel
    {
        int *buf = (int *)malloc(sizeof(int) * (LISTSIZE + 1));
        int *onfly = (int *)malloc(sizeof(int) * (n + 1));
        int *big = NULL;
#pragma omp for schedule(dynamic, 64)
        for (int p = 1; p <= n; p++) all_sph(c, p, buf, &big, onfly);
        free(buf);
        free(onfly);
        free(big);
    }
}

/* sph_update.F90:10-158 (all particles active) */
void oracle_sph_update(oracle_ctx *c) {
    int n = c->n;
    if (c->sph_std) {
        sph_update_standard(c);
        return;
    }
#pragma omp parallel
    {
        int *buf = (int *)malloc(sizeof(int) * (LISTSIZE + 1));
        int *big = NULL;
#pragma omp for schedule(dynamic, 64)
        for (int p = 1; p <= n; p++) all_sph_gradh(c, p, buf, &big);
        free(buf);
        free(big);
    }
    BHhydro_update_hmax(c);
#pragma omp parallel
    {
        int *buf = (int *)malloc(sizeof(int) * (LISTSIZE + 1));
        int *tmp = (int *)malloc(sizeof(int) * (c->pp_limit + 1));
        int *big = NULL;
#pragma omp for schedule(dynamic, 64)
        for (int p = 1; p <= n; p++) {
            int tot = get_neib(c, p, PX(p, SMOO), buf, &big, tmp, c->pp_limit);
            c->pptot[p] = tot;
            if (tot <= c->pp_limit)
                memcpy(&c->pplist[(size_t)(p - 1) * c->pp_limit], tmp, sizeof(int) * tot);
        }
        free(buf);
        free(tmp);
        free(big);
    }
}

/* thermal.F90:52-60 with injected sound speed (P = cs^2 rho, c = cs) */
void oracle_thermal_iso(oracle_ctx *c, double cs) {
    if (c->energy) {   /* thermal.F90:108-113, ideal gas from u (DIMENSIONLESS) */
        for (int p = 1; p <= c->n; p++) {
            double rho_p = c->rho[p];
            double press_p = (c->gamma - 1.0) * rho_p * c->u[p];
            c->press[p] = press_p;
            c->sound[p] = sqrt(press_p / rho_p);
        }
        return;
    }
    for (int p = 1; p <= c->n; p++) {
        c->press[p] = cs * cs * c->rho[p];
        c->sound[p] = cs;
    }
}

/* THERMAL=ENERGY_EQN: u per particle (current order) and gamma; from then on
 * oracle_thermal_iso applies the ideal-gas branch and hydro forces form
 * du/dt */
void oracle_set_energy(oracle_ctx *c, const double *u, double gamma) {
    if (!c->u) {
        c->u = (double *)calloc(c->n + 1, sizeof(double));
        c->du_dt = (double *)calloc(c->n + 1, sizeof(double));
    }
    c->energy = 1;
    c->gamma = gamma;
    for (int p = 1; p <= c->n; p++) c->u[p] = u[p - 1];
}

/* HFIND=CONSTANT with SPH=STANDARD: sph_update sets every h to hmin */
void oracle_set_hmin(oracle_ctx *c, double hmin) { c->hmin_const = hmin; }

void oracle_get_energy(const oracle_ctx *c, double *u, double *du_dt) {
    for (int p = 1; p <= c->n; p++) {
        u[p - 1] = c->energy ? c->u[p] : 0.0;
        du_dt[p - 1] = c->energy ? c->du_dt[p] : 0.0;
    }
}

/* hydro_gradh.F90:9-339 (AB viscosity, no VISC_TD/BALSARA); with the energy
 * equation du/dt (:119-124, :258-259, :326-332) */
static void hydro_gradh(oracle_ctx *c, int p, int *onfly, int *potlist, int **big) {
    double rp[3] = {PX(p, 0), PX(p, 1), PX(p, 2)};
    double hp = PX(p, SMOO), invhp = 1.0 / hp, hfactor_p = pow(invhp, 4.0);
    double vp[3] = {c->v[p][0], c->v[p][1], c->v[p][2]};
    double sound_p = c->sound[p], rho_p = c->rho[p];
    double pfactor_p = c->press[p] / rho_p / rho_p / c->omega[p];
    double ah[3] = {0, 0, 0}, ad[3] = {0, 0, 0};
    double dudt_diss = 0.0;
    int pp_numb = c->pptot[p];
    const int *lst;
    if (pp_numb <= c->pp_limit) {
        lst = &c->pplist[(size_t)(p - 1) * c->pp_limit];
    } else {
        pp_numb = get_neib(c, p, hp, potlist, big, onfly, c->n);
        lst = onfly;
    }
    const double alpha = c->alpha, beta = 2.0 * c->alpha;
    for (int i = 0; i < pp_numb; i++) {
        int pp = lst[i];
        double wmean = 0.0, rho_pp = c->rho[pp], mpp = PX(pp, MASS), hpp = PX(pp, SMOO);
        double du[3] = {PX(pp, 0) - rp[0], PX(pp, 1) - rp[1], PX(pp, 2) - rp[2]};
        double drmag = du[0] * du[0] + du[1] * du[1] + du[2] * du[2];
        if (drmag >= c->kernrangesqd * hp * hp && drmag >= c->kernrangesqd * hpp * hpp) continue;
        drmag = sqrt(drmag) + SMALL_NUMBER;
        double dv[3] = {c->v[pp][0] - vp[0], c->v[pp][1] - vp[1], c->v[pp][2] - vp[2]};
        double dvdr = dv[0] * du[0] + dv[1] * du[1] + dv[2] * du[2];
        double invdrmag = 1.0 / drmag;
        for (int k = 0; k < 3; k++) du[k] = du[k] * invdrmag;
        double pfactor_pp = c->press[pp] / rho_pp / rho_pp / c->omega[pp];
        if (drmag < c->kernrange * hp) {
            double skern = c->halfkerntot * drmag * invhp;
            int kern = (int)skern;
            if (kern > KERNTOT) kern = KERNTOT;
            wmean = 0.5 * hfactor_p * c->w1[kern];
            for (int k = 0; k < 3; k++)
                ah[k] = ah[k] + mpp * pfactor_p * hfactor_p * c->w1[kern] * du[k];
        }
        if (drmag < c->kernrange * hpp) {
            double invhpp = 1.0 / hpp, hfactor_pp = pow(invhpp, 4.0);
            double skern = c->halfkerntot * drmag * invhpp;
            int kern = (int)skern;
            if (kern > KERNTOT) kern = KERNTOT;
            wmean = wmean + 0.5 * hfactor_pp * c->w1[kern];
            for (int k = 0; k < 3; k++)
                ah[k] = ah[k] + mpp * pfactor_pp * hfactor_pp * c->w1[kern] * du[k];
        }
        double rhomean = 0.5 * (rho_p + rho_pp);
        if (dvdr < 0.0) {
            double alpha_mean = alpha;
            double vsignal = 0.5 * (sound_p + c->sound[pp]);
            double mu = 0.5 * (hp + hpp) * dvdr /
                        (drmag * drmag + 0.25 * ETA_SQD * (hp + hpp) * (hp + hpp));
            double visc_factor = (-alpha_mean * vsignal * mu + beta * mu * mu) / rhomean;
            for (int k = 0; k < 3; k++) ad[k] = ad[k] + mpp * visc_factor * wmean * du[k];
            if (c->energy)
                dudt_diss = dudt_diss + 0.5 * mpp * visc_factor * wmean * dvdr * invdrmag;
        }
    }
    for (int k = 0; k < 3; k++) c->a[p][k] = ah[k] + ad[k];
    if (c->energy) {
        double dudt = 0.0;
        dudt = dudt + dudt_diss;
        dudt = dudt - c->press[p] * c->div_v[p] / c->omega[p] / rho_p;
        c->du_dt[p] = dudt;
    }
}

/* hydro.F90:11-323 (standard SPH, AB viscosity): the kernel gradient
 * averaged over h_p and h_pp (wmean) carries both the pressure and the
 * viscous term; with the energy equation the same pair terms, projected on
 * dv.dr / |dr|, give du/dt (hydro.F90:206-208, :262-264) */
static void hydro_std(oracle_ctx *c, int p, int *onfly, int *potlist, int **big) {
    double rp[3] = {PX(p, 0), PX(p, 1), PX(p, 2)};
    double hp = PX(p, SMOO), invhp = 1.0 / hp, hfactor_p = pow(invhp, 3.0);
    double vp[3] = {c->v[p][0], c->v[p][1], c->v[p][2]};
    double sound_p = c->sound[p], rho_p = c->rho[p];
    double pfactor_p = c->press[p] / rho_p / rho_p;
    double ah[3] = {0, 0, 0}, ad[3] = {0, 0, 0}, dudt = 0.0;
    int pp_numb = c->pptot[p];
    const int *lst;
    if (pp_numb <= c->pp_limit) {
        lst = &c->pplist[(size_t)(p - 1) * c->pp_limit];
    } else {
        pp_numb = get_neib(c, p, hp, potlist, big, onfly, c->n);
        lst = onfly;
    }
    const double alpha = c->alpha;
    for (int i = 0; i < pp_numb; i++) {
        int pp = lst[i];
        double wmean = 0.0, rho_pp = c->rho[pp], mpp = PX(pp, MASS), hpp = PX(pp, SMOO);
        double dr[3] = {PX(pp, 0) - rp[0], PX(pp, 1) - rp[1], PX(pp, 2) - rp[2]};
        double drsqd = dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2];
        if (drsqd >= c->kernrangesqd * hp * hp && drsqd >= c->kernrangesqd * hpp * hpp) continue;
        double drmag = sqrt(drsqd) + SMALL_NUMBER;
        double invdrmag = 1.0 / drmag;
        double du[3] = {dr[0] * invdrmag, dr[1] * invdrmag, dr[2] * invdrmag};
        double dv[3] = {c->v[pp][0] - vp[0], c->v[pp][1] - vp[1], c->v[pp][2] - vp[2]};
        double dvdr = dv[0] * dr[0] + dv[1] * dr[1] + dv[2] * dr[2];
        if (drmag < c->kernrange * hp) {
            int kern = (int)(c->halfkerntot * drmag * invhp);
            if (kern > KERNTOT) kern = KERNTOT;
            wmean = 0.5 * hfactor_p * invhp * c->w1[kern];
        }
        if (drmag < c->kernrange * hpp) {
            double invhpp = 1.0 / hpp, hfactor_pp = pow(invhpp, 3.0);
            int kern = (int)(c->halfkerntot * drmag * invhpp);
            if (kern > KERNTOT) kern = KERNTOT;
            wmean = wmean + 0.5 * hfactor_pp * invhpp * c->w1[kern];
        }
        double pf = pfactor_p + c->press[pp] / (rho_pp * rho_pp);
        for (int k = 0; k < 3; k++) ah[k] = ah[k] + mpp * wmean * pf * du[k];
        if (c->energy) dudt = dudt + 0.5 * mpp * wmean * dvdr * invdrmag * pf;
        double rhomean = 0.5 * (rho_p + rho_pp);
        if (dvdr < 0.0) {
            double beta = 2 * alpha;
            double alpha_mean = alpha;
            double vsignal = 0.5 * (sound_p + c->sound[pp]);
            double mu = 0.5 * (hp + hpp) * dvdr /
                        (drsqd + 0.25 * ETA_SQD * (hp + hpp) * (hp + hpp));
            double visc_factor = (-alpha_mean * vsignal * mu + beta * mu * mu) / rhomean;
            for (int k = 0; k < 3; k++) ad[k] = ad[k] + mpp * visc_factor * wmean * du[k];
            if (c->energy)
                dudt = dudt + 0.5 * mpp * visc_factor * wmean * dvdr * invdrmag;
        }
    }
    for (int k = 0; k < 3; k++) c->a[p][k] = ah[k] + ad[k];
    if (c->energy) c->du_dt[p] = dudt;
}

/* sph_hydro_forces.F90:9-120 */
void oracle_hydro_forces(oracle_ctx *c) {
    int n = c->n;
    for (int p = 1; p <= n; p++) c->a[p][0] = c->a[p][1] = c->a[p][2] = 0.0;
#pragma omp parallel
    {
        int *onfly = NULL, *potlist = (int *)malloc(sizeof(int) * (LISTSIZE + 1)), *big = NULL;
#pragma omp for schedule(dynamic, 64)
        for (int p = 1; p <= n; p++) {
            if (c->pptot[p] > c->pp_limit && !onfly) onfly = (int *)malloc(sizeof(int) * (n + 1));
            if (c->sph_std)
                hydro_std(c, p, onfly, potlist, &big);
            else
                hydro_gradh(c, p, onfly, potlist, &big);
        }
        free(onfly);
        free(potlist);
        free(big);
    }
}

/* gravity_gradh.F90:11-95 */
static inline void gravity_gradh(const oracle_ctx *c, double invhp, double hpp,
                                 double mpp, const double rp[3], const double rpp[3],
                                 double zo_p, double zo_pp, double at[3], double *dpotp) {
    double dr[3] = {rpp[0] - rp[0], rpp[1] - rp[1], rpp[2] - rp[2]};
    double drsqd = dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2];
    double drmag = sqrt(drsqd) + SMALL_NUMBER;
    double invdrmag = 1.0 / drmag, grav, dpot;
    if (invdrmag > c->invkernrange * invhp) {
        int kern = (int)(c->halfkerntot * drmag * invhp);
        if (kern > KERNTOT) kern = KERNTOT;
        grav = invhp * invhp * (c->w5[kern] + zo_p * c->w4[kern] * invhp * invhp);
        dpot = invhp * c->w6[kern];
    } else {
        grav = invdrmag * invdrmag;
        dpot = invdrmag;
    }
    if (drmag < c->kernrange * hpp) {
        double invhpp = 1.0 / hpp;
        int kern = (int)(c->halfkerntot * drmag * invhpp);
        if (kern > KERNTOT) kern = KERNTOT;
        grav = grav + invhpp * invhpp * (c->w5[kern] + zo_pp * c->w4[kern] * invhpp * invhpp);
        dpot = dpot + invhpp * c->w6[kern];
    } else {
        grav = grav + invdrmag * invdrmag;
        dpot = dpot + invdrmag;
    }
    for (int k = 0; k < 3; k++) at[k] = 0.5 * mpp * invdrmag * grav * dr[k];
    *dpotp = 0.5 * mpp * dpot;
}

/* gravity_sph.F90:11-92 (standard SPH near field, no zeta term) */
static inline void gravity_sph(const oracle_ctx *c, double invhp, double hpp, double mpp,
                               const double rp[3], const double rpp[3], double at[3],
                               double *dpotp) {
    double dr[3] = {rpp[0] - rp[0], rpp[1] - rp[1], rpp[2] - rp[2]};
    double drsqd = dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2];
    double drmag = sqrt(drsqd) + SMALL_NUMBER;
    double invdrmag = 1.0 / drmag, grav, dpot;
    if (invdrmag > c->invkernrange * invhp) {
        int kern = (int)(c->halfkerntot * drmag * invhp);
        if (kern > KERNTOT) kern = KERNTOT;
        grav = invdrmag * invhp * invhp * c->w5[kern];
        dpot = invhp * c->w6[kern];
    } else {
        grav = invdrmag * invdrmag * invdrmag;
        dpot = invdrmag;
    }
    if (drmag < c->kernrange * hpp) {
        double invhpp = 1.0 / hpp;
        int kern = (int)(c->halfkerntot * drmag * invhpp);
        if (kern > KERNTOT) kern = KERNTOT;
        grav = grav + invdrmag * invhpp * invhpp * c->w5[kern];
        dpot = dpot + invhpp * c->w6[kern];
    } else {
        grav = grav + invdrmag * invdrmag * invdrmag;
        dpot = dpot + invdrmag;
    }
    for (int k = 0; k < 3; k++) at[k] = 0.5 * mpp * grav * dr[k];
    *dpotp = 0.5 * mpp * dpot;
}

/* BHgrav_accel.F90:21-348 */
static void BHgrav_accel(const oracle_ctx *c, int p, double invhp, const double rp[3],
                         double agravp[3], double *potp, int *treelist) {
    agravp[0] = agravp[1] = agravp[2] = 0.0;
    *potp = 0.0;
    int nlist = 0;
    double zo_p = PX(p, ZETA), afactor = 0.0;
    if (c->mac != MAC_GEOMETRIC) {
        afactor = c->agravmag[p];
        if (afactor > SMALL_NUMBER) {
            if (c->mac == MAC_GADGET)
                afactor = pow(afactor, -ONETHIRD);
            else if (c->mac == MAC_GADGET2)
                afactor = 1.0 / sqrt(afactor);
            else /* EIGEN */
                afactor = pow(1.0 / c->gpot[p], 2.0 * ONETHIRD);
        } else {
            afactor = 0.0;
        }
    }
    int ci = 0;
    for (;;) {
        const gnode *g = &c->BHgrav[ci];
        if (g->leaf == 1) {
            int pp = g->plist[0];
            if (pp != p) treelist[nlist++] = pp;
            ci = g->nextcell;
        } else {
            double dr[3] = {rp[0] - g->r[0], rp[1] - g->r[1], rp[2] - g->r[2]};
            double drsqd = dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2];
            int accept = (c->mac == MAC_GEOMETRIC)
                             ? (drsqd > g->dminsqd)
                             : (drsqd > g->mac * afactor && drsqd > g->dminsqd);
            if (accept) {
                const double *q = g->q, *s = g->s;
                double invdrsqd = 1.0 / (drsqd + SMALL_NUMBER);
                double invdrmag = sqrt(invdrsqd);
                double at[3], dpot;
                for (int k = 0; k < 3; k++) at[k] = -g->m * invdrsqd * invdrmag * dr[k];
                dpot = g->m * invdrmag;
                double invdr5 = invdrsqd * invdrsqd * invdrmag;
                double qscalar = q[0] * dr[0] * dr[0] + q[2] * dr[1] * dr[1] -
                                 (q[0] + q[2]) * dr[2] * dr[2] +
                                 2.0 * (q[1] * dr[0] * dr[1] + q[3] * dr[0] * dr[2] +
                                        q[4] * dr[1] * dr[2]);
                at[0] = at[0] + (q[0] * dr[0] + q[1] * dr[1] + q[3] * dr[2]) * invdr5 -
                        2.5 * qscalar * dr[0] * invdr5 * invdrsqd;
                at[1] = at[1] + (q[1] * dr[0] + q[2] * dr[1] + q[4] * dr[2]) * invdr5 -
                        2.5 * qscalar * dr[1] * invdr5 * invdrsqd;
                at[2] = at[2] + (q[3] * dr[0] + q[4] * dr[1] - (q[0] + q[2]) * dr[2]) * invdr5 -
                        2.5 * qscalar * dr[2] * invdr5 * invdrsqd;
                dpot = dpot + 0.5 * qscalar * invdr5;
                if (c->octupole) {
                    double sscalar = s[0] * dr[0] * dr[0] * dr[0] + s[1] * dr[0] * dr[0] * dr[1] +
                                     s[2] * dr[1] * dr[1] * dr[0] + s[3] * dr[1] * dr[1] * dr[1] +
                                     s[4] * dr[2] * dr[2] * dr[0] + s[5] * dr[2] * dr[2] * dr[1] +
                                     s[6] * dr[2] * dr[2] * dr[2] + s[7] * dr[0] * dr[0] * dr[2] +
                                     s[8] * dr[1] * dr[1] * dr[2] + s[9] * dr[0] * dr[1] * dr[2];
                    at[0] = at[0] +
                            0.5 * (3.0 * s[0] * dr[0] * dr[0] + 2.0 * s[1] * dr[0] * dr[1] +
                                   2.0 * s[7] * dr[0] * dr[2] + s[2] * dr[1] * dr[1] +
                                   s[4] * dr[2] * dr[2] + s[9] * dr[1] * dr[2]) *
                                invdr5 * invdrsqd -
                            3.5 * sscalar * dr[0] * invdr5 * invdrsqd * invdrsqd;
                    at[1] = at[1] +
                            0.5 * (3. * s[3] * dr[1] * dr[1] + 2.0 * s[2] * dr[0] * dr[1] +
                                   2.0 * s[8] * dr[1] * dr[2] + s[1] * dr[0] * dr[0] +
                                   s[5] * dr[2] * dr[2] + s[9] * dr[0] * dr[2]) *
                                invdr5 * invdrsqd -
                            3.5 * sscalar * dr[1] * invdr5 * invdrsqd * invdrsqd;
                    at[2] = at[2] +
                            0.5 * (3. * s[6] * dr[2] * dr[2] + 2.0 * s[4] * dr[0] * dr[2] +
                                   2.0 * s[5] * dr[1] * dr[2] + s[7] * dr[0] * dr[0] +
                                   s[8] * dr[1] * dr[1] + s[9] * dr[0] * dr[1]) *
                                invdr5 * invdrsqd -
                            3.5 * sscalar * dr[2] * invdr5 * invdrsqd * invdrsqd;
                    dpot = dpot + 0.5 * sscalar * invdr5 * invdrsqd;
                }
                for (int k = 0; k < 3; k++) agravp[k] = agravp[k] + at[k];
                *potp = *potp + dpot;
                ci = g->nextcell;
            } else {
                if (g->leaf > 0) {
                    for (int i = 0; i < g->leaf; i++) {
                        int pp = g->plist[i];
                        if (pp == p) continue;
                        treelist[nlist++] = pp;
                    }
                    ci = g->nextcell;
                } else if (g->leaf == 0) {
                    ci = g->ifopen;
                } else {
                    ci = g->nextcell;
                }
            }
        }
        if (nlist > GLISTSIZE - LEAFMAX || (nlist > 0 && ci > c->ctot_grav)) {
            for (int i = 0; i < nlist; i++) {
                int pp = treelist[i];
                double at[3], dpot, rpp[3] = {PX(pp, 0), PX(pp, 1), PX(pp, 2)};
                if (c->sph_std)
                    gravity_sph(c, invhp, PX(pp, SMOO), PX(pp, MASS), rp, rpp, at, &dpot);
                else
                    gravity_gradh(c, invhp, PX(pp, SMOO), PX(pp, MASS), rp, rpp, zo_p,
                                  PX(pp, ZETA), at, &dpot);
                for (int k = 0; k < 3; k++) agravp[k] = agravp[k] + at[k];
                *potp = *potp + dpot;
            }
            nlist = 0;
        }
        if (ci > c->ctot_grav) break;
    }
}

/* sph_grav_forces.F90:89-137 */
void oracle_grav_forces(oracle_ctx *c) {
    int n = c->n;
#pragma omp parallel
    {
        int *treelist = (int *)malloc(sizeof(int) * (GLISTSIZE + 1));
#pragma omp for schedule(dynamic, 64)
        for (int p = 1; p <= n; p++) {
            double ag[3], pot, rp[3] = {PX(p, 0), PX(p, 1), PX(p, 2)};
            BHgrav_accel(c, p, 1.0 / PX(p, SMOO), rp, ag, &pot, treelist);
            for (int k = 0; k < 3; k++) c->a[p][k] = c->a[p][k] + ag[k];
            c->gpot[p] = pot;
            if (c->mac != MAC_GEOMETRIC)
                c->agravmag[p] = sqrt(ag[0] * ag[0] + ag[1] * ag[1] + ag[2] * ag[2]);
        }
        free(treelist);
    }
}

/* direct_sph_gravity.F90:10-107 for a list of (1-based) targets */
void oracle_direct_gravity(const oracle_ctx *c, int nt, const int *targets,
                           double *agrav, double *pot) {
#pragma omp parallel for schedule(dynamic, 16)
    for (int i = 0; i < nt; i++) {
        int p = targets[i];
        double rp[3] = {PX(p, 0), PX(p, 1), PX(p, 2)}, invhp = 1.0 / PX(p, SMOO);
        double ag[3] = {0, 0, 0}, pt = 0.0;
        for (int pp = 1; pp <= c->n; pp++) {
            if (pp == p) continue;
            double at[3], dpot, rpp[3] = {PX(pp, 0), PX(pp, 1), PX(pp, 2)};
            gravity_gradh(c, invhp, PX(pp, SMOO), PX(pp, MASS), rp, rpp, PX(p, ZETA),
                          PX(pp, ZETA), at, &dpot);
            for (int k = 0; k < 3; k++) ag[k] = ag[k] + at[k];
            pt = pt + dpot;
        }
        for (int k = 0; k < 3; k++) agrav[3 * i + k] = ag[k];
        pot[i] = pt;
    }
}

/* ------------------------------------------------------------------------ */
/* Setters / getters with the same shapes as the reference dump harness.     */
void oracle_set_agravmag(oracle_ctx *c, const double *am) {
    for (int p = 1; p <= c->n; p++) c->agravmag[p] = am[p - 1];
}
void oracle_set_h(oracle_ctx *c, const double *h) {
    for (int p = 1; p <= c->n; p++) PX(p, SMOO) = h[p - 1];
}

void oracle_get_params(const oracle_ctx *c, int *ipar, double *rpar) {
    ipar[0] = c->pp_gather;
    ipar[1] = c->pp_limit;
    ipar[2] = c->ctot_grav;
    ipar[3] = c->ltot_grav;
    ipar[4] = c->ctot_hydro;
    ipar[5] = c->ltot_hydro;
    ipar[6] = c->cmax_grav;
    ipar[7] = c->cmax_hydro;
    rpar[0] = c->rextent;
    rpar[1] = c->thetamaxsqd;
    rpar[2] = c->abserror;
    rpar[3] = c->beta;
}

void oracle_get_levels(const oracle_ctx *c, int *first, int *last) {
    for (int l = 0; l <= LMAX; l++) {
        first[l] = c->first_cell[l];
        last[l] = c->last_cell[l];
    }
}

void oracle_get_particles(const oracle_ctx *c, int *porig, double *parray, double *v,
                          double *a, double *gpot, double *rho, double *omega,
                          double *divv, double *drhodt, double *press, double *sound,
                          double *agravmag) {
    for (int p = 1; p <= c->n; p++) {
        int i = p - 1;
        porig[i] = c->porig[p];
        for (int k = 0; k < 6; k++) parray[6 * i + k] = c->parray[p][k];
        for (int k = 0; k < 3; k++) {
            v[3 * i + k] = c->v[p][k];
            a[3 * i + k] = c->a[p][k];
        }
        gpot[i] = c->gpot[p];
        rho[i] = c->rho[p];
        omega[i] = c->omega[p];
        divv[i] = c->div_v[p];
        drhodt[i] = c->drhodt[p];
        press[i] = c->press[p];
        sound[i] = c->sound[p];
        agravmag[i] = (c->mac != MAC_GEOMETRIC) ? c->agravmag[p] : 0.0;
    }
}

void oracle_get_neib(const oracle_ctx *c, int *pptot, int *pplist) {
    for (int p = 1; p <= c->n; p++) {
        pptot[p - 1] = c->pptot[p];
        int m = c->pptot[p] < c->pp_limit ? c->pptot[p] : c->pp_limit;
        int *dst = &pplist[(size_t)(p - 1) * c->pp_limit];
        memset(dst, 0, sizeof(int) * c->pp_limit);
        if (c->pptot[p] <= c->pp_limit)
            memcpy(dst, &c->pplist[(size_t)(p - 1) * c->pp_limit], sizeof(int) * m);
    }
}

void oracle_get_grav_tree(const oracle_ctx *c, int nc, int *leaf, int *plist, int *nextcell,
                          int *ifopen, double *r, double *m, double *dminsqd, double *mac,
                          double *q, double *s, double *hmax) {
    for (int ci = 0; ci < nc; ci++) {
        const gnode *g = &c->BHgrav[ci];
        leaf[ci] = g->leaf;
        for (int i = 0; i < LEAFMAX; i++) plist[LEAFMAX * ci + i] = g->plist[i];
        nextcell[ci] = g->nextcell;
        ifopen[ci] = g->ifopen;
        for (int k = 0; k < 3; k++) r[3 * ci + k] = g->r[k];
        m[ci] = g->m;
        dminsqd[ci] = g->dminsqd;
        mac[ci] = g->mac;
        for (int k = 0; k < 5; k++) q[5 * ci + k] = g->q[k];
        for (int k = 0; k < 10; k++)
            s[10 * ci + k] = (k == 9 && c->octupole == 2) ? 0.5 * g->s[k] : g->s[k];
        hmax[ci] = c->BHstock[ci].hmax;
    }
}

void oracle_get_hydro_tree(const oracle_ctx *c, int nc, double *r, double *rmax,
                           double *hrangemax) {
    for (int ci = 0; ci < nc; ci++) {
        for (int k = 0; k < 3; k++) r[3 * ci + k] = c->BHhydro[ci].r[k];
        rmax[ci] = c->BHhydro[ci].rmax;
        hrangemax[ci] = c->BHhydro[ci].hrangemax;
    }
}

void oracle_get_kernel(const oracle_ctx *c, double *tab) {
    const double *t[9] = {c->w0, c->w1, c->w2, c->w3, c->w4, c->w5, c->w6, c->wh, c->wg};
    for (int j = 0; j < 9; j++)
        for (int i = 0; i <= KERNTOT; i++) tab[(size_t)j * (KERNTOT + 1) + i] = t[j][i];
}
