// serenx_internal.h -- shared definitions of the gfx950 force-evaluation
// library: constants of the reference option set, device node layouts, the
// context, and the launch functions each translation unit provides.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/serenx.h"

namespace serenx {

// Constants of the option set (src/headers/macros.h:12-191)
constexpr int LEAFMAX = 8;
constexpr int LMAX = 30;
constexpr int KERNTOT = 10000;
constexpr double SMALL_NUMBER = 1.0e-20;
constexpr double BIG_NUMBER = 9.9e20;
constexpr double HMULT = 1.1;
constexpr double H_CON = 0.02;
constexpr int GRADH_ITERATION_MAX = 20;
constexpr double ETA_SQD = 0.01;
constexpr double ONETHIRD = 0.333333333;
constexpr double TWOTHIRDS_DP = 0.666666666666666666;
constexpr double PI_REF = 3.1415926536;
constexpr double INVPI = 0.318309886;

// Range constants of the kernel (macros.h:101-118): the table index per unit
// r/h (HALFKERNTOT), the range in units of h (KERNRANGE), its square and the
// reference's rounded inverse (INVKERNRANGE).  Passed to every kernel that
// tests a pair against the kernel range.
struct KGeom {
    double half, range, rangesqd, invrange;
};
__host__ __device__ constexpr KGeom kernel_geom(int kernel) {
    if (kernel == SERENX_KERNEL_QUINTIC || kernel == SERENX_KERNEL_QUINTIC_TC)
        return KGeom{3333.0, 3.0, 9.0, 0.3333333333333};
    return KGeom{5000.0, 2.0, 4.0, 0.5};
}

constexpr int NTAB = 9;            // w0 w1 w2 w3 w4 w5 w6 wh wg
constexpr int KT = KERNTOT + 1;    // entries per table
enum { T_W0 = 0, T_W1, T_W2, T_W3, T_W4, T_W5, T_W6, T_WH, T_WG };
// Device table buffer (in doubles): the NTAB FP64 tables, their FP32 copy,
// then the gravity tables interleaved per index as (w4, w5, w6, 0) in FP64
// and in FP32, so that one softened pair term reads one 32- or 16-byte entry.
constexpr size_t KT_F32_OFF = (size_t)NTAB * KT;
constexpr size_t KT_G4_OFF = ((KT_F32_OFF + ((size_t)NTAB * KT + 1) / 2) + 3) / 4 * 4;
constexpr size_t KT_G4F_OFF = KT_G4_OFF + 4 * (size_t)KT;
// (w0, wh, w4, wg) per index for the h solve: one 16- or 32-byte load per
// candidate
constexpr size_t KT_H4_OFF = (KT_G4F_OFF + 2 * (size_t)KT + 3) / 4 * 4;
constexpr size_t KT_TOTAL = KT_H4_OFF + 4 * (size_t)KT;

constexpr int WAVE = 64;
// XCDs of an MI355X; the dispatcher deals a launch's workgroups to them
// round-robin (workgroup i to XCD i % NXCD)
constexpr int NXCD = 8;
// potential-neighbour list slots (opened leaves) per particle in the h solve
constexpr int HL_SLOTS = 96;
// standard SPH: r^2 values of the h_gather band kept per particle for the
// selection of the pp_gather-th neighbour (more are narrowed by bisection)
constexpr int BAND_MAX = 64;
// near-field list entries per group of the gravity walk (a longer list makes
// the group repeat its walk in the pair phase).  Groups deep in a 1M uniform
// sphere under the relative MAC at 1e-4 open more than 2048 leaves; 4096
// entries (48 KB per group, 768 B per particle) hold them.
constexpr int PL_CAP = 4096;
// leaves a wave of the neighbour walk records for the rows of its particles
// above pp_limit (a longer list makes such a particle walk the tree alone)
constexpr int NB_CAP = 512;

// Counters that every wave of a launch updates (walk statistics, interaction
// counts, the step's reductions) are spread over CTR_ROWS rows of CTR_SLOTS
// 64-bit slots, one 128-byte line per row.  A wave updates the row of its own
// index: device atomics on one address serialise in L2 (about 12 ns each, so
// a single counter updated by the 16k waves of a 1M-particle launch costs
// about 0.2 ms).  The host folds the rows (ctr_fold).
constexpr int CTR_ROWS = 128;
constexpr int CTR_SLOTS = 16;
constexpr size_t CTR_WORDS = (size_t)CTR_ROWS * CTR_SLOTS;

__device__ __forceinline__ unsigned long long *ctr_row(unsigned long long *base) {
    const unsigned w = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    return base + (size_t)(w % CTR_ROWS) * CTR_SLOTS;
}

// Folds counter rows into nslots values: the minimum for the slots in
// min_mask, the maximum for those in max_mask, the sum for the others.
inline void ctr_fold(const unsigned long long *rows, unsigned long long *out, int nslots,
                     unsigned min_mask = 0, unsigned max_mask = 0) {
    for (int k = 0; k < nslots; k++) {
        unsigned long long v = rows[k];
        for (int r = 1; r < CTR_ROWS; r++) {
            const unsigned long long x = rows[(size_t)r * CTR_SLOTS + k];
            if (min_mask >> k & 1u)
                v = x < v ? x : v;
            else if (max_mask >> k & 1u)
                v = x > v ? x : v;
            else
                v += x;
        }
        out[k] = v;
    }
}
// walk-statistics slots that hold maxima (serenx_walk_stats)
constexpr unsigned WSTAT_MAX_MASK = (1u << 5) | (1u << 12);

// Gravity node, DFS pre-order.  Read with wave-uniform indices in the walk, so
// a node is a run of scalar loads; everything a visit needs (MAC data and
// links) sits in the first 64 bytes, one s_load_dwordx16, and the moments
// follow.  next = DFS index of the cell that follows this subtree (nc for the
// last); leaf = particle count of a leaf, 0 for a branch; first = first
// particle (tree order) of the cell.
struct __attribute__((aligned(64))) GNode {
    double r[3];      // centre of mass
    double dminsqd;   // opening distance^2 (BHgrav_stock.F90:317-348)
    double mac;       // multipole MAC term (GADGET/GADGET2)
    double m;
    int next;
    int leaf;
    int first;
    int count;        // particles in the subtree
    double q[6];      // quadrupole [xx xy yy zx zy zz], zz = -(xx + yy)
    // octupole gradient coefficients of the walk: H = grad(S) / 2 with S the
    // octupole cubic, H_a = sum_k oc[6a + k] mono_k over the monomials
    // (xx, xy, xz, yy, zz, yz); the moments s themselves are in Ctx::gs
    double oc[18];
};

// Hydro node (BHhydro_node, modules.F90:929-937), DFS order.
struct __attribute__((aligned(64))) HNode {
    double r[3];      // bounding-box centre
    double rmax;      // half-diagonal of the bounding box
    double hrangemax; // 2 max(h) over the subtree
    int next;
    int leaf;
    int first;
    float half[3];    // bounding-box half-extents, rounded up with a margin: every
                      // particle of the subtree lies within r +- half
};

// Mixed-precision gravity node: FP64 COM (the cell offset is subtracted in
// FP64) and MAC data, FP32 moments; walk data in the first 64 bytes.
struct __attribute__((aligned(64))) GNodeF {
    double r[3];
    double dminsqd;
    double mac;
    float m;
    int next;
    int leaf;
    int first;
    int pad;
    float q[6];       // [xx xy yy zx zy zz]
    float s[10];      // octupole moments (the FP32 walk forms grad S from them)
};

static_assert(offsetof(GNode, q) == 64 && sizeof(GNode) == 256, "GNode layout");
static_assert(offsetof(GNodeF, q) <= 64 && sizeof(GNodeF) == 128, "GNodeF layout");
static_assert(sizeof(HNode) == 64, "HNode layout");

struct DevBuf {
    void *p = nullptr;
    size_t bytes = 0;
};

struct Ctx {
    serenx_options opt{};
    int device = 0;
    hipStream_t stream = nullptr;
    hipStream_t gstream = nullptr;   // gravity walk, beside the SPH kernels
    hipEvent_t evfork = nullptr;     // fork of the gravity stream from the main one
    hipEvent_t evfork2 = nullptr;    // the h solve done: the walk's pair phase may start
    hipEvent_t evg[4] = {};          // walk phase 1 start, end; phase 2 start, end (timed)
    bool grav_timed = false;         // evg holds a walk of the last evaluation
    bool box_ok = false;             // HNode boxes hold the current positions (stocked
                                     // since the positions last changed)
    hipEvent_t ev[8] = {};
    hipEvent_t evjoin = nullptr;     // the gravity stream's work done (cell phase in graph part A)
    double last_ms[4] = {0, 0, 0, 0};

    // Captured restock evaluations (SERENX_EVAL_GRAPH): part A (stock, h
    // solve, neighbour rows, walk cell phase) and part B (thermal, hydro,
    // walk pair phase, join), each valid while graph_epoch and the flags it
    // was captured with are unchanged.  graph_epoch changes whenever a
    // device address or a launch shape a captured graph holds may change (a
    // build, a load, a reallocation, a precision or tuning change).
    struct Graphs {
        hipGraphExec_t a = nullptr, b = nullptr;
        long epoch_a = -1, epoch_b = -1;
        int flags_a = 0, flags_b = 0;
        long long captures = 0, launches = 0;
    } gr;
    long graph_epoch = 0;
    bool capturing = false;          // host code is being recorded into a graph

    int n = 0;          // particles loaded
    int cap = 0;        // particle capacity of the device arrays
    int pp_gather = 0, pp_limit = 0;
    KGeom kg{};         // range constants of opt.kernel
    double daux = 0.0;  // 10 huge^-0.2 (BHgrav_stock.F90:347)

    // particle SoA, tree order
    double *x = nullptr, *y = nullptr, *z = nullptr, *m = nullptr, *h = nullptr,
           *zeta = nullptr, *vx = nullptr, *vy = nullptr, *vz = nullptr,
           *rho = nullptr, *omega = nullptr, *press = nullptr, *sound = nullptr,
           *divv = nullptr, *drhodt = nullptr, *ax = nullptr, *ay = nullptr,
           *az = nullptr, *gpot = nullptr, *agm = nullptr;
    double *u = nullptr, *dudt = nullptr;   // energy equation: u, du/dt
    int *porig = nullptr;
    uint8_t *active = nullptr;   // accdo
    double *hrec = nullptr;      // hydro pair records, 16 doubles per particle
    double *grec = nullptr;      // gravity source records, 8 doubles per particle
    double *gax = nullptr, *gay = nullptr, *gaz = nullptr;   // gravity acceleration
    double *gpc = nullptr;       // potential of the walk's cell terms (phase 1)
    // near-field lists of the walk: per group of 64 particles, up to PL_CAP
    // leaves (first << 4 | count) and the masks of the lanes that opened them
    int *pl_leaf = nullptr, *pl_len = nullptr;
    unsigned long long *pl_mask = nullptr;
    size_t pl_entries = 0;       // slots of pl_leaf and pl_mask
    // groups of the gravity walk: [start, end) of up to 64 consecutive
    // particles of one cloud (each cloud's groups start at its first
    // particle, so a cloud is grouped the same alone and in a forest)
    int2 *grp = nullptr;
    int ngrp = 0;
    int *hlist = nullptr;        // h-solve leaf lists, HL_SLOTS x 64 per wave
    int hl_cap = HL_SLOTS;       // slots used per lane (SERENX_TUNE_HSOLVE_SLOTS)
    bool neib_group = true;      // SERENX_TUNE_NEIB_GROUP: one neighbour walk per wave
    int hsolve_waves = -1;       // SERENX_TUNE_HSOLVE_WAVES: -1 auto, 0, 8
    double mixed_abserror = 0.0; // mixed walk's MAC tolerance (0: opt.abserror)
    double *band = nullptr;      // standard SPH: BAND_MAX x 64 r^2 values per wave
    size_t band_cap = 0;         // particles the band buffer holds
    double *tmpd = nullptr;      // permutation scratch (cap doubles)
    int *tmpi = nullptr;

    // tree build scratch
    uint64_t *khi = nullptr, *klo = nullptr, *khi2 = nullptr, *klo2 = nullptr;
    // forest sort: cloud keys, positions, permutation and gathered keys/values
    int *fs_ck = nullptr, *fs_ck2 = nullptr, *fs_pos = nullptr, *fs_perm = nullptr,
        *fs_v = nullptr;
    uint64_t *fs_k = nullptr;
    int *idx = nullptr, *idx2 = nullptr;
    uint8_t *cpl = nullptr, *pact = nullptr;
    int *flag = nullptr, *incl = nullptr, *cellof = nullptr;
    int *ncs = nullptr, *off = nullptr;
    int8_t *firstlev = nullptr;
    // clouds: a forest of independent trees, one root per cloud at level 0;
    // cloud k owns the particles [cloud_off[k], cloud_off[k+1]) in input and
    // in tree order (the build permutes within clouds only)
    int ncloud = 1, max_clouds = 1, max_chunks = 1, nchunk = 0;   // forest_max_clouds
    std::vector<int> cloud_off_h;
    int *cloud_off = nullptr;    // device, ncloud + 1
    int *cid = nullptr;          // cloud of each particle
    int *chunks = nullptr;       // bounding-box chunks [cloud, begin, end]
    int *chunk_first = nullptr;  // first chunk of each cloud, ncloud + 1
    int2 *root_range = nullptr;  // DFS range [root, root.next) of each cloud's tree
    double *root = nullptr;      // per cloud [rc x y z, cellsize0, rextent, rmin3, rmax3]
    double *bbpart = nullptr;    // bounding-box chunk partials
    void *sorttmp = nullptr;
    size_t sorttmp_bytes = 0;
    int *hostpin = nullptr;      // pinned readback slots
    double *hostpin_d = nullptr;

    // cells by BFS id (capacity 2 cap + 2)
    int ccap = 0;
    int *c_start = nullptr, *c_end = nullptr, *c_parent = nullptr, *c_leaf = nullptr,
        *c_level = nullptr, *bfs2dfs = nullptr, *dfs2bfs = nullptr, *dfs_parent = nullptr;
    double *bbmin = nullptr, *bbmax = nullptr, *hmaxs = nullptr;  // per DFS cell
    GNode *gn = nullptr;
    double *gs = nullptr;        // octupole moments s[10] per node (stock, export)
    GNodeF *gnf = nullptr;
    HNode *hn = nullptr;
    int nc = 0, ltot = 0;
    int level_off[LMAX + 2] = {0};
    int level_cnt[LMAX + 2] = {0};
    double rextent = 0.0;
    bool tree_ok = false, sph_ok = false;

    // neighbour lists: rows of pp_limit ids (tree order, 0-based) + overflow CSR
    int *pptot = nullptr, *pplist = nullptr;
    int *ovf_off = nullptr, *ovf_list = nullptr;  // rows with pptot > pp_limit
    size_t ovf_cap = 0;
    int *ovf_count = nullptr;     // device scalar: number of overflow particles
    int *ovf_idx = nullptr;       // the overflow particles, in tree order
    // leaves each wave of the neighbour walk opened (NB_CAP per wave of 64
    // particles, in DFS order) and their number, for the overflow rows
    int *nbl = nullptr, *nbl_len = nullptr;

    // kernel tables [NTAB][KT] (kernel.F90)
    double *ktab = nullptr;
    double host_ktab[NTAB * KT];

    // work queue of the persistent gravity walk (walk.h: GroupQueue)
    struct WalkQueue {
        int ngroups = 0;
        int *ctr = nullptr;          // 2 NXCD counters (one set per walk phase)
        int *cost = nullptr;         // per group: cells its last walk visited
        int *order = nullptr;        // group ids by range, then decreasing cost
        uint64_t *keys = nullptr, *keys2 = nullptr;
        int *vals = nullptr;
        void *tmp = nullptr;
        size_t tmp_bytes = 0;
        bool have_cost = false;      // cost holds a walk of the current group layout
        bool order_ready = false;    // order was sorted, after the last walk, for the next
    } gq;
    int walk_queue = 64;             // SERENX_TUNE_WALK_QUEUE: blocks per CU (0: no queue)
    int pair_blocks = 0;             // SERENX_TUNE_PAIR_BLOCKS: pair phase's (0: walk_queue's)
    int near_cap = PL_CAP;           // SERENX_TUNE_NEAR_CAP: near-list leaves a group may store
    int nb_cap = NB_CAP;             // SERENX_TUNE_NEIB_LEAVES: leaves a neighbour-walk wave records
    bool walk_alone = false;         // no SPH kernels run beside the walk (gravity only)
    int ncu = 1;                     // compute units of the device

    // interaction counters of the gravity walk
    unsigned long long *icount = nullptr;
    // walk statistics of the last force evaluation (serenx_walk_stats)
    unsigned long long *wstat = nullptr;

    // device-resident stepping (sph_integrate): per-particle state in tree
    // order, permuted with the particles, and the time_module scalars
    struct Step {
        bool on = false;
        serenx_step_options opt{};
        double *rox = nullptr, *roy = nullptr, *roz = nullptr;   // r_old
        double *vox = nullptr, *voy = nullptr, *voz = nullptr;   // v_old
        double *vhx = nullptr, *vhy = nullptr, *vhz = nullptr;   // v_half
        double *laststep = nullptr, *dt = nullptr;
        double *u_old = nullptr;             // energy equation
        long long *nlast = nullptr;
        int *nlevel = nullptr;
        int *nminneib = nullptr;             // neighbour check: deepest level among neighbours
        unsigned long long *red = nullptr;   // device reductions in counter rows [dt_min bits,
                                             // p_update, level max, nactive, nminneib max]
        std::vector<unsigned long long> red_host;
        long long n = 0, nsteps = 0, nresync = 0, level_max = 0, level_step = 0;
        long long nbuild = 0, nstock = 0, nactive = 0;
        double time = 0.0, timestep = 0.0, dt_max = 0.0;
    } st;

    // batched serving steps (serenx_serve_*): one kick-drift-kick step of
    // selected clouds of the forest, each with its own step size.  v_half per
    // particle (tree order, permuted with the particles), per-cloud step size,
    // time and selection on the device.
    struct Serve {
        bool on = false;
        double accel_mult = 0.5, courant_mult = 0.2;
        double *vhx = nullptr, *vhy = nullptr, *vhz = nullptr;
        unsigned long long *dtbits = nullptr;   // per cloud: minimum step size (bits)
        double *dt = nullptr, *time = nullptr;  // per cloud
        uint8_t *sel = nullptr;                 // per cloud: 1 = stepped by the call
        uint8_t *sel_host = nullptr;            // pinned staging of the selection
        int ncloud = 0;
        long long nsteps = 0;
    } sv;

    // host-state binding
    serenx_host_state hs{};
    bool bound = false;
    std::vector<double> hbuf;   // staging for host-state transfers
    std::vector<int> hibuf;
};

// Zeroes nwords 32-bit words at p on stream s with a fill kernel (capi.hip).
// The force path clears its flags and counters this way rather than with
// hipMemsetAsync, whose nodes have given wrong results on graph replays.
int zero_words(void *p, size_t nwords, hipStream_t s);

// error reporting (capi.hip)
int set_error(int code, const std::string &msg);
int hip_check(hipError_t e, const char *what);
#define SX_HIP(call)                                                     \
    do {                                                                 \
        hipError_t e_ = (call);                                          \
        if (e_ != hipSuccess) return ::serenx::hip_check(e_, #call);     \
    } while (0)
#define SX_TRY(call)                                                     \
    do {                                                                 \
        int r_ = (call);                                                 \
        if (r_ != SERENX_OK) return r_;                                  \
    } while (0)

// Waits for the gravity stream when it leaves scope armed: an error return
// between grav_launch and grav_join must not leave the walk running while the
// caller reuses the context (a rebuild rewrites the buffers it reads).
struct GravGuard {
    hipStream_t gs;
    bool armed = false;
    explicit GravGuard(hipStream_t s) : gs(s) {}
    ~GravGuard() {
        if (armed) (void)hipStreamSynchronize(gs);
    }
};

// host-side kernel tables (capi.hip): kernel.F90:14-127, M4, NDIM=3
void make_kernel_tables(double *tab, int kernel);

// diag.hip: FP64 diagnostics.  partial holds DIAG_MAX_BLOCKS x DIAG_SUMS
// doubles (block partial sums), dout DIAG_VALUES doubles, both on the device.
constexpr int DIAG_MAX_BLOCKS = 1024;
constexpr int DIAG_SUMS = 16;
constexpr int DIAG_VALUES = 20;
// Most clouds a forest of n particles may hold: one per pp_gather particles
// (a cloud needs pp_gather particles for its h solve).  The gravity walk has
// at most (n + 63) / 64 + this many groups.
inline int forest_max_clouds(int n, int pp_gather) { return n / (pp_gather > 1 ? pp_gather : 1) + 1; }
inline size_t walk_max_groups(int n, int pp_gather) {
    return ((size_t)n + WAVE - 1) / WAVE + (size_t)forest_max_clouds(n, pp_gather);
}
// near-list slots of a context for n particles: PL_CAP for each group of a
// single cloud (at most n / 64 + 1 groups), shared by all groups otherwise
// waves of a one-thread-per-particle launch over n particles
inline size_t particle_waves(int n) { return (((size_t)n + 255) / 256) * 4; }
inline size_t near_list_entries(int n) { return (((size_t)n + WAVE - 1) / WAVE + 1) * PL_CAP; }

int diag_blocks(int n, int *per_block);
int diagnostics(Ctx &c, double *partial, double *dout);
// the device address of the allocation-free diagnostics scratch on a device:
// resolved once per device by diag_scratch_setup (with that device current),
// null before
int diag_scratch_setup(int device);
double *diag_scratch(int device);

// tree.hip
int ensure_tree_buffers(Ctx &c);
int build_tree(Ctx &c);          // bbox, keys, sort, cells, DFS, permutation
int stock_tree(Ctx &c);          // BHgrav_stock + BHhydro_stock (fused)
int refresh_f32_nodes(Ctx &c);   // mixed-precision node copy of the stocked tree
int hguess(Ctx &c);              // BHhydro_hguess
int update_hmax(Ctx &c);         // BHhydro_update_hmax
int eval_pow_third(Ctx &c, int n, const double *x, double *y);   // device x**ONETHIRD
int permute_particles(Ctx &c, const int *perm);
// cloud layout of the loaded particles (counts[k] particles in cloud k)
int set_clouds(Ctx &c, int ncloud, const int *counts);

// sph.hip
int sph_hsolve(Ctx &c);          // all_sph_gradh over active particles
int sph_neighbours(Ctx &c);      // get_neib over active particles
int sph_neib_rows(Ctx &c);       // sph_neighbours, launch part: rows and overflow count
int sph_neib_finish(Ctx &c);     // sph_neighbours, host part: status, rows above pp_limit
int sph_hgather(Ctx &c);         // standard SPH: h_gather over active particles
int sph_density(Ctx &c);         // standard SPH: all_sph over active particles
int sph_thermal_iso(Ctx &c, double cs);
int sph_thermal(Ctx &c);          // thermal for the context's THERMAL option
int hydro_forces(Ctx &c);        // hydro_gradh over active particles

// step.hip
int step_begin(Ctx &c, const serenx_step_options &opt);
int step_run(Ctx &c, int count);
int step_fetch(Ctx &c, int32_t *nlevel, int64_t *nlast, double *laststep, int32_t *active,
               double *r_old, double *v_old, double *v_half, int32_t *nminneib);
void step_free(Ctx &c);
int permute_step(Ctx &c, const int *perm);
int serve_begin(Ctx &c, double accel_mult, double courant_mult);
int serve_step(Ctx &c, int ncloud, const uint8_t *select);
int serve_info(Ctx &c, int ncloud, double *dt, double *time);
void serve_free(Ctx &c);
int permute_serve(Ctx &c, const int *perm);

// capi.hip: the force path pieces the step driver reuses
int run_tree_update(Ctx &c);
int run_sph_update(Ctx &c);
int run_hydro(Ctx &c);
int run_grav(Ctx &c);
int run_force_eval(Ctx &c, int flags);   // serenx_force_eval's body
void graphs_free(Ctx &c);

// grav.hip
int grav_launch(Ctx &c);         // fork the gravity stream: walk phase 1 (cells), after the tree
int grav_pairs(Ctx &c);          // walk phase 2 (pairs), after the h solve
int grav_join(Ctx &c);           // main stream waits for the walk and adds a_grav to a
int run_sph_h(Ctx &c);           // sph_update, part 1: h of the active particles
int run_sph_lists(Ctx &c);       // sph_update, part 2: hmax, neighbour lists (+ standard density)
int direct_gravity(Ctx &c, int nt, const int *dtargets, double *dacc, double *dpot);

}  // namespace serenx
