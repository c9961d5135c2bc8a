/*
 * serenx.h -- C ABI of the MI355X (gfx950) force-evaluation library for
 * SEREN-style self-gravitating grad-h SPH.
 *
 * The library replaces SEREN's per-step force path: the Barnes-Hut tree
 * build + stock, the grad-h smoothing-length/density solve with neighbour
 * lists, the hydro pair forces and the tree-gravity walk.  Two ways to use
 * it share the same kernels:
 *
 *  1. Host-state (drop-in) mode.  The caller binds its own host arrays, laid
 *     out as SEREN's module arrays (column-major, 1-based ids in pplist and in
 *     the reorder list), and calls serenx_tree_update / serenx_hguess /
 *     serenx_sph_update / serenx_hydro_forces / serenx_grav_forces in place of
 *     the reference driver subroutines.  Each call uploads what the reference
 *     routine reads and downloads what it writes.
 *
 *  2. Device-resident mode.  serenx_load copies a cloud to the GPU once;
 *     serenx_force_eval runs one whole force evaluation on the device with no
 *     host transfer; serenx_fetch copies results back.
 *
 * All functions return SERENX_OK (0) or an error code; serenx_last_error()
 * gives the message of the last failure on the calling thread.  One context
 * must be driven from one host thread at a time.
 *
 * Citations are to the SEREN reference sources (src/...).
 */
#ifndef SERENX_H
#define SERENX_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct serenx_ctx serenx_ctx;

/* status codes */
enum {
    SERENX_OK = 0,
    SERENX_EINVAL = 1,   /* bad argument / option */
    SERENX_EHIP = 2,     /* HIP runtime error (no device, launch failure, ...) */
    SERENX_ELMAX = 3,    /* 'Reached maximum level' (BHgrav_build.F90:289) */
    SERENX_ESTATE = 4,   /* call made out of order (e.g. hydro before sph) */
    SERENX_ENOMEM = 5,   /* device allocation failed */
    SERENX_EHNEG = 6,    /* negative smoothing length (all_sph_gradh.F90:486) */
    SERENX_EFEW = 7      /* fewer particles than pp_gather: h solve cannot close */
};

/* MULTIPOLE option (Makefile: MULTIPOLE = QUADRUPOLE | OCTUPOLE).
 * SERENX_OCTUPOLE reproduces BHgrav_stock.F90 exactly, including its branch
 * parallel-axis shift (:263-297), whose quadrupole terms are half the exact
 * ones and whose xyz term takes 25 for 5.  SERENX_OCTUPOLE_EXACT (no
 * counterpart in the reference) stocks exact octupole moments at every level;
 * serenx_export_tree reports them in the same units. */
enum { SERENX_QUADRUPOLE = 2, SERENX_OCTUPOLE = 3, SERENX_OCTUPOLE_EXACT = 4 };
/* MAC option (Makefile: MAC = GEOMETRIC | GADGET | GADGET2) */
enum {
    SERENX_MAC_GEOMETRIC = 0, /* GEOMETRIC_MAC: d^2 > diag^2 / theta^2 */
    SERENX_MAC_GADGET = 1,    /* GADGET_MAC (BHgrav_stock.F90:329-331) */
    SERENX_MAC_GADGET2 = 2,   /* GADGET2_MAC (:332-334) */
    SERENX_MAC_EIGEN = 3      /* EIGEN_MAC (:335-337, eigenvalue_mac.F90), |phi| based */
};
/* arithmetic of the gravity walk: FP64 throughout, or FP32 terms on
 * cell-relative coordinates with compensated FP32 accumulation */
enum { SERENX_FP64 = 0, SERENX_MIXED = 1 };

/* serenx_force_eval flags */
enum {
    SERENX_EVAL_HGUESS = 1,   /* run BHhydro_hguess before the h solve (setup) */
    SERENX_EVAL_NO_GRAV = 2,  /* skip the gravity walk */
    SERENX_EVAL_NO_HYDRO = 4, /* skip the hydro pair forces */
    SERENX_EVAL_RESTOCK = 8,  /* no build: the tree of the last build, restocked
                                 with the current positions (tree_update.F90
                                 between rebuilds) */
    SERENX_EVAL_GRAPH = 16    /* with RESTOCK: run as recorded HIP graphs */
};

/* kernel (Makefile KERNEL, kernel.F90:14-236, macros.h:101-118): M4 cubic
 * spline (range 2h), M4 with the Thomas-Couchman (1992) gradient in the
 * hydro force (M4TC: w1 = -1/pi for s < 2/3), the quintic spline of Morris
 * (1996) (range 3h, table index 3333 r/h), or the quintic with its
 * Thomas-Couchman gradient (QUINTICTC).  The kernel range sets the
 * neighbour radius, so pp_gather = 4 pi/3 (range h_fac)^3. */
enum {
    SERENX_KERNEL_M4 = 0,
    SERENX_KERNEL_M4_TC = 1,
    SERENX_KERNEL_QUINTIC = 2,
    SERENX_KERNEL_QUINTIC_TC = 3
};

/* THERMAL option (Makefile THERMAL): ISOTHERMAL (P = cs^2 rho, c = cs with
 * the given sound speed), or ENERGY_EQN (INTERNAL_ENERGY): the ideal gas
 * P = (gamma - 1) rho u, c = sqrt(P / rho) (thermal.F90:108-113), du/dt from
 * the pressure and viscous-heating terms (hydro_gradh.F90:258-259,326-332),
 * u advanced with the particles and the energy time-step criterion
 * (timestep_size.F90:117-122). */
enum { SERENX_THERMAL_ISOTHERMAL = 0, SERENX_THERMAL_ENERGY_EQN = 1 };

/* SPH option (Makefile SPH): GRAD_H (grad-h SPH: the rho-h iteration with
 * Omega and zeta, all_sph_gradh.F90, hydro_gradh.F90, gravity_gradh.F90) or
 * STANDARD with HFIND=NUMBER (h_gather.F90: h is the distance of the
 * pp_gather-th nearest neighbour over the kernel range; all_sph.F90,
 * hydro.F90, gravity_sph.F90; Omega = 1, zeta = 0).  With
 * THERMAL_ENERGY_EQN, STANDARD forms du/dt as hydro.F90 does: the pressure
 * and viscous pair terms projected on dv.dr / |dr|, with no P div v term. */
enum { SERENX_SPH_GRADH = 0, SERENX_SPH_STANDARD = 1 };

/* HFIND option of SERENX_SPH_STANDARD (Makefile HFIND): NUMBER (h_gather.F90)
 * or CONSTANT (sph_update.F90:107-109: every h is set to hmin before the
 * neighbour search; sph_setup.F90:45-49 makes no h guess). */
enum { SERENX_HFIND_NUMBER = 0, SERENX_HFIND_CONSTANT = 1 };

typedef struct serenx_options {
    int32_t multipole;   /* SERENX_QUADRUPOLE or SERENX_OCTUPOLE */
    int32_t mac;         /* SERENX_MAC_* */
    int32_t precision;   /* SERENX_FP64 or SERENX_MIXED */
    int32_t kernel;      /* SERENX_KERNEL_* */
    double h_fac;        /* grad-h h = h_fac (m/rho)^(1/3); params 'h_fac' (1.2) */
    double thetamaxsqd;  /* geometric MAC theta^2; params 'thetamaxsqd' */
    double abserror;     /* multipole MAC tolerance; params 'abserror' */
    double alpha;        /* AB viscosity alpha (beta := 2 alpha); params 'alpha' */
    double sound_speed;  /* isothermal c_s used by serenx_force_eval */
    int32_t thermal;     /* SERENX_THERMAL_* */
    int32_t sph;         /* SERENX_SPH_* */
    double gamma;        /* ratio of specific heats; params 'gamma' (1.4) */
    int32_t pp_gather;   /* SERENX_SPH_STANDARD: neighbours per h (params
                            'pp_gather'); 0 takes int((4 pi/3)(range h_fac)^3).
                            Must be 0 for grad-h SPH, where h_fac sets it. */
    int32_t serial;      /* 0: the gravity walk runs on a second stream beside the
                            neighbour and hydro kernels; 1: on the context's one
                            stream (for callers that run many contexts at once,
                            whose streams share the device's hardware queues) */
    int32_t hfind;       /* SERENX_HFIND_*; NUMBER unless sph is STANDARD */
    double hmin;         /* HFIND_CONSTANT: the smoothing length of every particle
                            (params hmin), > 0 */
} serenx_options;

/* Fills SEREN's parameter defaults (default_parameters.F90:62-117,173):
 * quadrupole, geometric MAC, h_fac 1.2, thetamaxsqd 0.5, abserror 0.1,
 * alpha 1, FP64, isothermal with sound speed 1 (DIMENSIONLESS isothermal,
 * thermal.F90:52-60), gamma 1.4. */
int serenx_default_options(serenx_options *opt);

int serenx_create(const serenx_options *opt, int device, serenx_ctx **out);
int serenx_destroy(serenx_ctx *ctx);
const char *serenx_last_error(void);

/* pp_gather = int((4 pi/3)(range h_fac)^3) (initialize_seren_variables_1
 * .F90:30-39; for STANDARD SPH the pp_gather option when set),
 * pp_limit = 2 pp_gather */
int serenx_neighbour_params(const serenx_ctx *ctx, int32_t *pp_gather, int32_t *pp_limit);

/* ---------------------------------------------------------------------------
 * Host-state (drop-in) mode
 * ------------------------------------------------------------------------- */

/* Pointers to the caller's arrays, shaped as SEREN's module arrays
 * (particle_module, hydro_module, neighbour_module; modules.F90:292-542).
 * parray(6,ptot) = [x y z m h zeta/omega] (with SERENX_SPH_STANDARD the zeta
 * column is neither read nor written, and omega and rextent may be NULL);
 * v(3,ptot); a(3,ptot); pplist(pp_limit,ptot) holds 1-based ids.
 * agravmag may be NULL under the geometric MAC.  accdo (Fortran default logical, 4 bytes) may be NULL: all
 * particles active. */
typedef struct serenx_host_state {
    int32_t ptot;
    int32_t pp_limit;
    double *parray;
    double *v;
    double *a;
    double *gpot;
    double *agravmag;
    double *rho;
    double *omega;
    double *press;
    double *sound;
    double *div_v;
    double *drhodt;
    int32_t *pptot;
    int32_t *pplist;
    int32_t *accdo;
    double *rextent;     /* scalar out: max extent (bounding_box.F90:39) */
    double *du_dt;       /* SERENX_THERMAL_ENERGY_EQN: du_dt(ptot), in and out of
                            serenx_hydro_forces; may be NULL otherwise */
} serenx_host_state;

/* Binding a state with a different ptot resets the device state; re-binding
 * the same ptot (arrays moved) keeps the tree and neighbour lists. */
int serenx_bind_host_state(serenx_ctx *ctx, const serenx_host_state *st);

/* replaces tree_update (src/main/tree_update.F90:9-92): BHgrav_build,
 * BH_reorder_particles' traversal, BHgrav_stock, copy_BHgrav_to_BHhydro,
 * BHhydro_stock.  Writes the new tree order to dummylist(1:ptot) (1-based old
 * ids, the list BH_reorder_particles.F90:123 passes to reorder_particle_arrays)
 * and the domain extent to *rextent.  Host arrays are not touched: the caller
 * then permutes all its particle arrays with that list (reorder_particle_arrays,
 * reorder_array.F90:9-270), exactly where the reference does. */
int serenx_tree_update(serenx_ctx *ctx, int32_t *dummylist);

/* replaces the stock-only branch of tree_update (tree_update.F90:35,60): the
 * tree keeps its topology and particle order, cells are restocked from the
 * bound parray (BHgrav_stock + BHhydro_stock).  Requires a previous
 * serenx_tree_update on the same particles. */
int serenx_tree_stock(serenx_ctx *ctx);

/* replaces BHhydro_hguess (src/BHtree/BHhydro_hguess.F90:11-163) */
int serenx_hguess(serenx_ctx *ctx);

/* replaces sph_update (src/sph_sim/sph_update.F90:10-158): all_sph_gradh over
 * active particles, BHhydro_update_hmax, get_neib over active particles */
int serenx_sph_update(serenx_ctx *ctx);

/* replaces sph_hydro_forces (src/sph_sim/sph_hydro_forces.F90:9-120):
 * a = a_hydro for active particles (hydro_gradh.F90:9-339) */
int serenx_hydro_forces(serenx_ctx *ctx);

/* replaces sph_grav_forces (src/sph_sim/sph_grav_forces.F90:10-199):
 * a += a_grav, gpot = pot, agravmag = |a_grav| for active particles
 * (BHgrav_accel.F90:21-348 with gravity_gradh.F90:11-95) */
int serenx_grav_forces(serenx_ctx *ctx);

/* ---------------------------------------------------------------------------
 * Device-resident mode
 * ------------------------------------------------------------------------- */

/* Copies one cloud (row-major pos[n][3], vel[n][3], mass[n], h[n]) to the
 * device in its input order (porig = 1..n) and resets the grad-h state as
 * initialize_sph_variables_1.F90:33-110 does.  h may be NULL (zeros). */
int serenx_load(serenx_ctx *ctx, int32_t n, const double *pos, const double *vel,
                const double *mass, const double *h);

/* A forest of ncloud independent clouds in one context (no counterpart in
 * the reference, which holds one cloud): counts[k] particles of cloud k,
 * concatenated in the arrays as for serenx_load.  Every cloud gets its own
 * tree (its own bounding box and root, BHgrav_build.F90:65-67), walks stay
 * inside a particle's own cloud, and each cloud's results are those of
 * loading it alone; one launch per phase serves all clouds.  Tree order
 * permutes particles within their cloud only, so cloud k keeps the
 * particle range given by serenx_cloud_offsets.  Integration steps
 * (serenx_step_begin) need a single cloud. */
int serenx_load_forest(serenx_ctx *ctx, int32_t ncloud, const int32_t *counts,
                       const double *pos, const double *vel, const double *mass,
                       const double *h);

/* offsets[0..ncloud]: particle range of every cloud, in input and tree order */
int serenx_cloud_offsets(const serenx_ctx *ctx, int32_t *offsets);

/* Switches the gravity precision of later evaluations (SERENX_FP64 or
 * SERENX_MIXED); takes effect at the next tree stock (every force
 * evaluation, and every force-evaluating integration step).  Lets a serving
 * loop interleave FP64 guard steps with mixed-precision steps. */
int serenx_set_precision(serenx_ctx *ctx, int32_t precision);

/* FP64 conservation diagnostics of the loaded particles, the sums of
 * diagnostics.F90:85-134 (isothermal or energy-equation thermal energy; no
 * sinks), reduced on the device in a fixed order (deterministic):
 *   out[0] mtot, out[1..3] rcom, out[4..6] vcom, out[7..9] linear momentum,
 *   out[10..12] angular momentum, out[13..15] net force sum m a,
 *   out[16] ketot = sum m v^2 / 2, out[17] gpetot = sum m gpot / 2,
 *   out[18] utot (3/2 sum m P/rho isothermal, sum m u with the energy
 *   equation), out[19] etot = ketot - gpetot + utot.
 * Allocates its partial-sum buffer on the device for the call. */
int serenx_diagnostics(serenx_ctx *ctx, double *out);

/* serenx_diagnostics without allocation.  No counterpart in the reference:
 * for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops).  Writes the same DIAG_VALUES = 20 values,
 * bit for bit, to out_dev -- device memory of at least 20 doubles -- on the
 * context's stream and returns without waiting: they are there once
 * serenx_sync(ctx) returns.  The block partial sums live in a fixed
 * 128 KiB device array of the library (1024 rows of 16 sums, one copy per
 * device, held until the process ends; every particle count fits, because
 * the rows take ever longer slices of particles), so calls on one device
 * that overlap in time (two contexts on different streams) race on it.
 * Calls no hipMalloc, hipFree or other allocating routine. */
int serenx_diagnostics_noalloc(serenx_ctx *ctx, double *out_dev);

/* Batched serving steps (no single counterpart in the reference: each call
 * does for every selected cloud of a forest what sph_integrate.F90 does for
 * one cloud with the LFKDK integrator, one time-step level and adaptive
 * steps, advance_leapfrog_kdk.F90:32-116 with timestep_size.F90:92-114).
 * serenx_serve_begin, after the setup evaluation (serenx_force_eval with
 * SERENX_EVAL_HGUESS) of a cloud or a forest (serenx_load_forest), takes
 * every cloud's first step size: the minimum of timestep_size over its
 * particles.  serenx_serve_step advances the clouds with select[k] != 0
 * (ncloud entries) by one kick-drift-kick step of their own size: kick and
 * drift with the last acceleration, one force evaluation (tree rebuilt) of
 * their particles, the second kick, and their next step sizes.  Other
 * clouds do not move, and no walk leaves its own cloud, so a cloud's state
 * does not depend on which clouds share a call.  Isothermal option only.
 * serenx_serve_info copies each cloud's next step size and its time. */
int serenx_serve_begin(serenx_ctx *ctx, double accel_mult, double courant_mult);
int serenx_serve_step(serenx_ctx *ctx, int32_t ncloud, const uint8_t *select);
int serenx_serve_info(serenx_ctx *ctx, int32_t ncloud, double *dt, double *time);

/* Tuning knobs of a context (no counterpart in the reference: they choose
 * how the kernels run, and the autotuner in seren_amd/autotune.py sets them).
 *  SERENX_TUNE_HSOLVE_SLOTS: potential-neighbour leaves recorded per particle
 *    in the grad-h solve and h_gather lists, 1..96 (default 96).  A particle
 *    whose list does not fit is summed over the tree walk at the same radius
 *    (the same leaves in the same order): results do not change, only speed;
 *    tests lower it to drive that path.
 *  SERENX_TUNE_MIXED_ABSERROR: multipole-MAC tolerance of the mixed-precision
 *    gravity walk (GADGET, GADGET2, EIGEN MACs; 0 = the options' abserror).
 *    The FP64 walk always uses abserror (the reference's interaction sets);
 *    the mixed walk may trade part of its accuracy margin for speed.  Takes
 *    effect at the next tree stock.
 *  SERENX_TUNE_WALK_QUEUE: v >= 1 runs the gravity walk as a persistent
 *    kernel of at most v blocks per CU (default 64: as many as fit) whose
 *    waves take groups of 64 particles from per-XCD work queues, each ordered
 *    by the cost of the group's previous walk (longest first).  Fewer blocks
 *    leave wave slots to the SPH kernels that run beside the walk in force
 *    evaluations and steps; the best value depends on the workload (the
 *    autotuner in seren_amd/autotune.py times a few).  A walk with nothing
 *    beside it (serenx_grav_forces) takes all that fit.  0 launches one wave
 *    per group in tree order.  Results do not change.
 *  SERENX_TUNE_NEIB_GROUP: 1 (default) builds the neighbour lists with one
 *    walk per wave that opens every cell any of its particles could need; 0
 *    with the particles' own walks (BHhydro_walk.F90's test per particle).
 *    The lists are the same.
 *  SERENX_TUNE_HSOLVE_WAVES: waves per SIMD of the grad-h iteration kernel:
 *    8 (registers capped at 64, a few spills), 0 (the compiler's budget,
 *    5 waves), -1 (default: 8 in mixed precision, where the SPH kernels are
 *    the critical path, 0 in FP64, where the h solve runs beside the
 *    issue-bound FP64 walk).  Results do not change.
 *  SERENX_TUNE_NEAR_CAP: near-field leaves the gravity walk's cell phase
 *    records per group of 64 particles for the pair phase, 1..4096 (default
 *    4096).  A group whose list does not fit walks the tree again in the
 *    pair phase with the same MAC decisions (the same leaves in the same
 *    order): results do not change, only speed; tests lower it to drive
 *    that path.
 *  SERENX_TUNE_NEIB_LEAVES: leaves each wave of the neighbour walk records
 *    (in DFS order) for its particles with more than pp_limit neighbours,
 *    0..512 (default 512).  Such a particle's complete row is formed from
 *    that list by the whole wave; when the list did not fit (or at 0) the
 *    particle walks the tree alone.  The rows are the same either way.
 *  SERENX_TUNE_PAIR_BLOCKS: blocks per CU of the walk's pair phase, 1..64,
 *    or 0 (default) for SERENX_TUNE_WALK_QUEUE's value; used when the work
 *    queue is on.  Results do not change.
 * serenx_get_tuning reads the current value. */
enum {
    SERENX_TUNE_HSOLVE_SLOTS = 1,
    SERENX_TUNE_MIXED_ABSERROR = 2,
    SERENX_TUNE_WALK_QUEUE = 3,
    SERENX_TUNE_NEIB_GROUP = 4,
    SERENX_TUNE_HSOLVE_WAVES = 5,
    SERENX_TUNE_NEAR_CAP = 6,
    SERENX_TUNE_NEIB_LEAVES = 7,
    SERENX_TUNE_PAIR_BLOCKS = 8
};
int serenx_set_tuning(serenx_ctx *ctx, int32_t key, double value);
int serenx_get_tuning(const serenx_ctx *ctx, int32_t key, double *value);

/* Free and total device memory (hipMemGetInfo) of the context's device, and
 * the bytes this context holds. */
int serenx_memory(const serenx_ctx *ctx, int64_t *free_bytes, int64_t *total_bytes,
                  int64_t *ctx_bytes);

/* Replaces |a| from a previous evaluation (used by the GADGET/GADGET2 MACs,
 * BHgrav_accel.F90:97-116); values are in the current tree order. */
int serenx_set_agravmag(serenx_ctx *ctx, const double *agravmag);

/* One force evaluation in the sph_setup.F90:42-66 order: tree_update,
 * [hguess], sph_update, isothermal thermal (P = c_s^2 rho, c = c_s),
 * sph_hydro_forces, sph_grav_forces.  Enqueued on the context's stream;
 * returns after the work is queued (serenx_sync waits).
 * SERENX_EVAL_RESTOCK evaluates on the tree of the last build, restocked
 * with the current positions, as tree_update.F90 does between rebuilds
 * (tree_update.F90:24-88; the particle order does not change).  It waits
 * once, after the neighbour rows, for the host to see whether any particle
 * has more than pp_limit neighbours.  Adding SERENX_EVAL_GRAPH runs it as
 * two recorded HIP graphs (grad-h SPH; other options run it eagerly):
 * recorded at the first such call and replayed until a build, a load, a
 * precision or tuning change, or a reallocation makes a new recording;
 * results are identical to the eager restock evaluation. */
int serenx_force_eval(serenx_ctx *ctx, int32_t flags);

/* Recordings and replays of the restock evaluation's graphs so far (no
 * counterpart in the reference). */
int serenx_graph_info(const serenx_ctx *ctx, int64_t *captures, int64_t *launches);
int serenx_sync(serenx_ctx *ctx);

/* Copies the current state to the host, in tree order.  Any pointer may be
 * NULL.  Vectors are row-major [n][3]; pplist is [n][pp_limit], 1-based ids,
 * rows valid where pptot <= pp_limit (get_neib.F90:115-118). */
int serenx_fetch(serenx_ctx *ctx, int32_t *porig, double *pos, double *vel,
                 double *mass, double *h, double *zeta, double *rho, double *omega,
                 double *div_v, double *drhodt, double *press, double *sound,
                 double *acc, double *gpot, double *agravmag, int32_t *pptot,
                 int32_t *pplist);

/* FP64 kernel-softened direct sum over all particles for nt targets (0-based
 * indices in tree order): the semantics of direct_sph_gravity.F90:10-107. */
int serenx_direct_gravity(serenx_ctx *ctx, int32_t nt, const int32_t *targets,
                          double *acc, double *pot);

/* SERENX_THERMAL_ENERGY_EQN, device-resident mode: the specific internal
 * energy u of every particle, in the current particle order (the load order
 * until the first tree build, tree order after it; serenx_fetch's porig
 * gives it), and the fetch of u and du/dt in the same order. */
int serenx_set_energy(serenx_ctx *ctx, const double *u);
int serenx_fetch_energy(serenx_ctx *ctx, double *u, double *du_dt);

/* y[i] = x[i]**0.333333333 as the h solve evaluates it on the device (the
 * reference's (m/rho)**ONETHIRD, all_sph_gradh.F90:253,317, correctly
 * rounded); n values, host arrays.  For checking the device arithmetic. */
int serenx_pow_third(serenx_ctx *ctx, int32_t n, const double *x, double *y);

/* Tree sizes: number of cells and the deepest level (ctot_grav + 1, ltot_grav). */
int serenx_tree_sizes(const serenx_ctx *ctx, int32_t *ncells, int32_t *ltot);

/* Exports the tree in the reference's form, indexed by BFS cell id
 * (BHgrav_build.F90:149-157): leaf, plist[8] (1-based tree-order ids),
 * nextcell (sentinel 2n+1), ifopen (branches), r[3] (COM), m, dminsqd, mac,
 * q[5], s[10], hydro r[3] (bbox centre), rmax, hrangemax, and the first/last
 * cell id of each level (31 entries each).  Any pointer may be NULL. */
int serenx_export_tree(const serenx_ctx *ctx, int32_t *leaf, int32_t *plist,
                       int32_t *nextcell, int32_t *ifopen, double *r, double *m,
                       double *dminsqd, double *mac, double *q, double *s,
                       double *hr, double *rmax, double *hrangemax,
                       int32_t *first_cell, int32_t *last_cell);

/* Per-phase device time of the last serenx_force_eval in milliseconds
 * (tree, sph, hydro, grav), and interaction counts of the last gravity walk
 * summed over particles: counts[0] cell terms, [1] particle-particle terms,
 * [2] pairs inside 2 h_p, [3] pairs inside 2 h_j (gravity_gradh.F90:61,72). */
int serenx_last_timings(const serenx_ctx *ctx, double *ms4);

/* Durations (ms, HIP events on the gravity stream) of the last evaluation's
 * two walk phases: [0] traversal and cell terms, [1] near-field pair terms.
 * Their sum is serenx_last_timings' gravity entry. */
int serenx_walk_timings(const serenx_ctx *ctx, double *ms2);
int serenx_interaction_counts(serenx_ctx *ctx, int64_t *counts4);

/* Walk statistics of the last serenx_force_eval (performance diagnostics):
 * [0] gravity cells visited by waves, [1] gravity lane visits,
 * [2] h-solve cells visited, [3] h-solve lane visits, [4] h-solve passes
 * summed over waves, [5] most passes of one wave, [6] h iterations summed over
 * particles, [7] h-solve waves, [8] neighbour-walk cells visited,
 * [9] neighbour-walk lane visits, [10] h-solve list-recording walks summed over
 * waves, [11] h-solve passes summed over waves whose lists overflowed (summed
 * over the tree walk instead), [12] longest potential-neighbour list in
 * leaves; [13..15] zero. */
int serenx_walk_stats(serenx_ctx *ctx, int64_t *stats16);

/* ---------------------------------------------------------------------------
 * Device-resident integration steps (sph_integrate.F90:41-113)
 * ------------------------------------------------------------------------- */

enum {
    SERENX_INT_RK = 0,    /* advance_runge_kutta.F90 (the reference default) */
    SERENX_INT_LFKDK = 1  /* advance_leapfrog_kdk.F90, with the second-half
                             kick taken over the time since the half step
                             (the reference leaves dt_half unset there) */
};

/* Makefile TIMESTEP: how the largest step is chosen at a resynchronisation
 * (sph_timesteps.F90:84-91) */
enum {
    SERENX_LEVELS_ADAPTIVE = 0,    /* dt_max = dt_min 2^level_max */
    SERENX_LEVELS_RESTRICTED = 1,  /* dt_min rounded down to dt_fixed / 2^k */
    SERENX_LEVELS_FIXED = 2        /* dt_min = dt_fixed / 2^level_max */
};
/* Makefile CHECK_NEIB_TIMESTEP: neighbour time-step limiter
 * (check_neighbour_timesteps.F90, reduce_timesteps.F90) */
enum {
    SERENX_NEIB_CHECK_OFF = 0,
    SERENX_NEIB_CHECK_END = 1,       /* levels lowered at the end of a step */
    SERENX_NEIB_CHECK_IMMEDIATE = 2  /* ... or at once when synchronised
                                        (IMMEDIATE_TIMESTEP_REDUCTION) */
};

typedef struct serenx_step_options {
    int32_t integrator;   /* SERENX_INT_RK or SERENX_INT_LFKDK */
    int32_t nlevels;      /* block time-step levels (parameter nlevels, default 1) */
    int32_t nbuildstep;   /* integer steps between tree builds (default 8) */
    int32_t levels;       /* SERENX_LEVELS_* (Makefile default RESTRICTED) */
    double accel_mult;    /* timestep_size acceleration factor (default 0.5) */
    double courant_mult;  /* timestep_size Courant factor (default 0.2) */
    double dt_fixed;      /* reference step for RESTRICTED / FIXED levels
                             (parameter dt_fixed, default 1.0) */
    int32_t neib_check;   /* SERENX_NEIB_CHECK_* (Makefile default END) */
    int32_t reserved;
} serenx_step_options;

typedef struct serenx_step_info {
    int64_t n;            /* integer time (time_module n) */
    int64_t nsteps;       /* steps taken */
    int64_t nresync;      /* integer time of the next resynchronisation */
    int64_t level_max;    /* deepest occupied time-step level */
    int64_t level_step;   /* level of the smallest (half) step */
    int64_t nbuild;       /* step of the next tree build */
    int64_t nstock;       /* step of the next tree stock */
    int64_t nactive;      /* particles flagged for a force evaluation next step */
    double time;          /* physical time */
    double timestep;      /* real time of the smallest integer step */
    double dt_max;        /* largest step size */
    double reserved;
} serenx_step_info;

void serenx_default_step_options(serenx_step_options *opt);

/* Starts device-resident stepping after the setup evaluation
 * (serenx_force_eval with SERENX_EVAL_HGUESS, as sph_setup.F90 does): the
 * step state of initialize_sph_variables_2.F90:56-70,157-175 and
 * initialize_seren_variables_2.F90:60-61 (r_old = r, v_old = v_half = v, no
 * particle active, n = nresync = 0, next build and stock at step 1). */
int serenx_step_begin(serenx_ctx *ctx, const serenx_step_options *opt);

/* count x sph_integrate: tree build/stock on the reference's cadence
 * (tree_update.F90:24-88), sph_update, thermal and forces for the active
 * particles, sph_timesteps (sph_timesteps.F90:30-337 with timestep_size.F90),
 * sph_advance.  Synchronous. */
int serenx_step(serenx_ctx *ctx, int32_t count);

int serenx_step_get_info(const serenx_ctx *ctx, serenx_step_info *info);

/* Per-particle step state in tree order; any pointer may be NULL.  Vectors
 * row-major [n][3]; active = accdo (1/0); nminneib = the deepest level among
 * the particles that hold this one as a neighbour (zero without the
 * neighbour check). */
int serenx_fetch_step(serenx_ctx *ctx, int32_t *nlevel, int64_t *nlast, double *laststep,
                      int32_t *active, double *r_old, double *v_old, double *v_half,
                      int32_t *nminneib);

#ifdef __cplusplus
}
#endif
#endif /* SERENX_H */
