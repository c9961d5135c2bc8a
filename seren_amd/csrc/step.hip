// step.hip -- device-resident integration steps on gfx950: the order of
// sph_integrate.F90:41-113 around the force path, the hierarchical block
// time steps of sph_timesteps.F90 with timestep_size.F90, and the particle
// advance of advance_runge_kutta.F90 (default) or advance_leapfrog_kdk.F90.
//
// Per-particle step state (r_old, v_old, v_half, nlast, nlevel, laststep)
// lives on the device in tree order and is permuted with the particles on
// every tree build, as reorder_particle_arrays does.  The integer-time
// scalars of time_module stay on the host; each step reads back three
// reductions (minimum dt, number of updated particles, deepest level).
//
// All arithmetic follows the reference's operation order without FMA
// contraction (the library is compiled with -ffp-contract=off).
#include <cmath>
#include <vector>

#include "serenx_internal.h"

namespace serenx {

namespace {

constexpr int TB = 256;
inline int nblk(long n) { return (int)((n + TB - 1) / TB); }

constexpr double INVLOGETWO = 1.442695041;   // macros.h:15 (1/ln 2 to 10 digits)
constexpr double TVISC_FAC = 0.6;            // macros.h:133
constexpr double SMALL_NUMBER_DP = 1.0e-20;  // macros.h:82

constexpr int TIMESTEP_LEVEL_DIFF_MAX = 2;   // macros.h:185

// reductions, in counter rows (ctr_row): folded by minimum, maximum or sum
enum { R_DTMIN = 0, R_PUPDATE = 1, R_LEVMAX = 2, R_NACTIVE = 3, R_NMAX = 4, R_COUNT = 5 };
constexpr unsigned R_MIN_MASK = 1u << R_DTMIN;
constexpr unsigned R_MAX_MASK = (1u << R_LEVMAX) | (1u << R_NMAX);
constexpr double DT_INIT = 9.9e20;

__global__ void k_red_init(unsigned long long *__restrict__ red) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < (int)CTR_WORDS)
        red[i] = (i % CTR_SLOTS == R_DTMIN)
                     ? (unsigned long long)__double_as_longlong(DT_INIT) : 0ull;
}

// timestep_size.F90:92-122: acceleration and Courant criteria with the AB
// viscosity terms (beta = 2 alpha, hydro_gradh.F90:227), and with the energy
// equation the energy criterion, which the reference applies to
// p < phydroend only (:118), i.e. not to the last particle
__device__ __forceinline__ double timestep_size(double ax, double ay, double az, double hp,
                                                double divv, double vsignal, double alpha_p,
                                                double beta_p, double accel_mult,
                                                double courant_mult, bool energy_p, double up,
                                                double dudt_p) {
    const double div_v_p = fabs(divv);
    const double amag = sqrt(ax * ax + ay * ay + az * az);
    const double tacc = accel_mult * sqrt(hp / (amag + SMALL_NUMBER_DP));
    double dt = tacc;
    const double tcour = courant_mult * hp /
                         (vsignal + hp * div_v_p +
                          TVISC_FAC * (alpha_p * vsignal + beta_p * hp * div_v_p));
    dt = (tcour < dt) ? tcour : dt;
    if (energy_p) {
        const double m = (accel_mult < courant_mult) ? accel_mult : courant_mult;
        const double tenergy = m * up / (fabs(dudt_p) + SMALL_NUMBER_DP);
        dt = (tenergy < dt) ? tenergy : dt;
    }
    return dt;
}

// INVLOGETWO*log(dt_max/dt) truncated, + 1 (sph_timesteps.F90:99, :159)
__device__ __forceinline__ long long level_of(double dt_max, double dt) {
    return (long long)(INVLOGETWO * log(dt_max / dt)) + 1;
}

__device__ __forceinline__ void wave_min_bits(unsigned long long *dst, double v) {
    unsigned long long b = (unsigned long long)__double_as_longlong(v);
    for (int o = 32; o > 0; o >>= 1) {
        const unsigned long long t = __shfl_xor(b, o, 64);
        b = t < b ? t : b;
    }
    if ((threadIdx.x & 63) == 0) atomicMin(ctr_row(dst), b);
}

// sph_timesteps.F90:59-66 and :94-102: every particle's step size, the
// minimum, nlast = 0; then (second kernel) the level of each particle.
__global__ void k_ts_all(int n, const double *__restrict__ ax, const double *__restrict__ ay,
                         const double *__restrict__ az, const double *__restrict__ h,
                         const double *__restrict__ divv, const double *__restrict__ sound,
                         const double *__restrict__ u, const double *__restrict__ dudt,
                         double alpha, double accel_mult, double courant_mult,
                         double *__restrict__ dt, long long *__restrict__ nlast,
                         unsigned long long *__restrict__ red) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    double d = DT_INIT;
    if (p < n) {
        nlast[p] = 0;
        const bool en = u && p < n - 1;
        d = timestep_size(ax[p], ay[p], az[p], h[p], divv[p], sound[p], alpha, 2.0 * alpha,
                          accel_mult, courant_mult, en, en ? u[p] : 0.0, en ? dudt[p] : 0.0);
        dt[p] = d;
    }
    wave_min_bits(red + R_DTMIN, d);
}

__global__ void k_ts_levels(int n, double dt_max, long long level_max,
                            const double *__restrict__ dt, int *__restrict__ nlevel,
                            int *__restrict__ nminneib) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    long long level = level_of(dt_max, dt[p]);
    level = level < level_max ? level : level_max;
    level = level > 0 ? level : 0;
    nlevel[p] = (int)level;
    if (nminneib) nminneib[p] = (int)level;   // :104-106
}

// 2^e divides nint, as mod(n, 2**e) == 0 tests it; a negative e makes the
// reference's 2**e zero and its mod divide by zero, and counts as false here
__device__ __forceinline__ bool synced(long long nint, long long e) {
    return e >= 0 && e < 63 && nint % (1LL << e) == 0;
}

// sph_timesteps.F90:226-240: the neighbour check after a level update.
// level_step is the value before this update, level_max the new one.
__global__ void k_ts_neib(int n, long long nint, long long level_step, long long level_max,
                          const long long *__restrict__ nlast, int *__restrict__ nlevel,
                          int *__restrict__ nminneib) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    long long nm = nminneib[p];
    nm = nm < level_max ? nm : level_max;
    const long long lev = nlevel[p];
    if (nm - lev > TIMESTEP_LEVEL_DIFF_MAX && nlast[p] == nint &&
        synced(nint, level_step - nm + TIMESTEP_LEVEL_DIFF_MAX)) {
        long long nl = nm - TIMESTEP_LEVEL_DIFF_MAX;
        nl = nl < level_max ? nl : level_max;
        nlevel[p] = (int)nl;
        nm = nl;
    } else if (nlast[p] == nint) {
        nm = lev;
    }
    nminneib[p] = (int)nm;
}

// check_neighbour_timesteps.F90:43-66: every active particle raises the
// nminneib of each of its neighbours to its own level (a maximum, so the
// order of the updates does not matter; without sinks nminneib is never -1)
__global__ void __launch_bounds__(256) k_neib_levels(
    int n, const uint8_t *__restrict__ active, const int *__restrict__ nlevel,
    const int *__restrict__ pptot, const int *__restrict__ pplist,
    const int *__restrict__ ovf_off, const int *__restrict__ ovf_list, int pp_limit,
    int *__restrict__ nminneib) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n || !active[p]) return;
    const int levelp = nlevel[p];
    if (levelp == 0) return;   // nminneib >= 0 already
    const int nn = pptot[p];
    const int *lst = (nn <= pp_limit) ? pplist + (size_t)p * pp_limit : ovf_list + ovf_off[p];
    for (int i = 0; i < nn; i++) {
        const int j = lst[i];
        if (levelp > nminneib[j]) atomicMax(nminneib + j, levelp);
    }
}

// reduce_timesteps.F90:28-40 with reduce_particle_timestep.F90:31-66 (after
// the advance): a particle whose neighbours step more than
// TIMESTEP_LEVEL_DIFF_MAX levels deeper restarts its step at a deeper level
// when the integer time allows.  Also the largest nminneib, which must not
// exceed level_max (:30-33).
template <int KDK>
__global__ void k_reduce_steps(int n, long long nint, long long level_step, double timestep,
                               double *__restrict__ x, double *__restrict__ y,
                               double *__restrict__ z, const double *__restrict__ vx,
                               const double *__restrict__ vy, const double *__restrict__ vz,
                               const double *__restrict__ ax, const double *__restrict__ ay,
                               const double *__restrict__ az, double *__restrict__ rox,
                               double *__restrict__ roy, double *__restrict__ roz,
                               double *__restrict__ vox, double *__restrict__ voy,
                               double *__restrict__ voz, double *__restrict__ vhx,
                               double *__restrict__ vhy, double *__restrict__ vhz,
                               long long *__restrict__ nlast, int *__restrict__ nlevel,
                               int *__restrict__ nminneib, double *__restrict__ laststep,
                               uint8_t *__restrict__ active, const double *__restrict__ u,
                               double *__restrict__ u_old,
                               unsigned long long *__restrict__ red) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    unsigned added = 0;
    int nmax = 0;
    if (p < n) {
        const long long nm = nminneib[p];
        nmax = (int)nm;
        if (nm - nlevel[p] > TIMESTEP_LEVEL_DIFF_MAX && nm != -1 && nlast[p] != nint &&
            synced(nint, level_step - nm + TIMESTEP_LEVEL_DIFF_MAX)) {
            const long long lnew = nm - TIMESTEP_LEVEL_DIFF_MAX;
            const double r[3] = {x[p], y[p], z[p]};
            const double v[3] = {vx[p], vy[p], vz[p]};
            rox[p] = r[0]; roy[p] = r[1]; roz[p] = r[2];
            vox[p] = v[0]; voy[p] = v[1]; voz[p] = v[2];
            const double ls = timestep * (double)(nint - nlast[p]);
            laststep[p] = ls;
            if (KDK) {
                const double a[3] = {ax[p], ay[p], az[p]};
                double vh[3];
                for (int k = 0; k < 3; k++) vh[k] = v[k] + 0.5 * a[k] * ls;
                vhx[p] = vh[0]; vhy[p] = vh[1]; vhz[p] = vh[2];
                x[p] = r[0] + vh[0] * ls;
                y[p] = r[1] + vh[1] * ls;
                z[p] = r[2] + vh[2] * ls;
            }
            if (u) u_old[p] = u[p];   // INTERNAL_ENERGY (:40-42)
            nlevel[p] = (int)lnew;
            nlast[p] = nint;
            nminneib[p] = (int)lnew;
            added = active[p] ? 0u : 1u;
            active[p] = 1;   // RUNGE_KUTTA and LEAPFROG_KDK
        }
    }
    unsigned long long cnt = added;
    for (int o = 32; o > 0; o >>= 1) {
        cnt += __shfl_xor(cnt, o, 64);
        nmax = max(nmax, __shfl_xor(nmax, o, 64));
    }
    if ((threadIdx.x & 63) == 0) {
        unsigned long long *row = ctr_row(red);
        atomicAdd(row + R_NACTIVE, cnt);
        atomicMax(row + R_NMAX, (unsigned long long)(nmax > 0 ? nmax : 0));
    }
}

// sph_timesteps.F90:151-170: particles at the end of their step take a new
// level (down by at most one, and only on a synchronised integer time);
// counts the updates and the deepest level over all particles.
__global__ void k_ts_update(int n, long long nint, long long level_step, double dt_max,
                            const double *__restrict__ ax, const double *__restrict__ ay,
                            const double *__restrict__ az, const double *__restrict__ h,
                            const double *__restrict__ divv, const double *__restrict__ sound,
                            const double *__restrict__ u, const double *__restrict__ dudt,
                            double alpha, double accel_mult, double courant_mult,
                            const long long *__restrict__ nlast, int *__restrict__ nlevel,
                            unsigned long long *__restrict__ red) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    unsigned upd = 0;
    int lev = 0;
    if (p < n) {
        lev = nlevel[p];
        if (nint == nlast[p]) {
            upd = 1;
            const long long nlastlevel = lev;
            const bool en = u && p < n - 1;
            const double dt = timestep_size(ax[p], ay[p], az[p], h[p], divv[p], sound[p], alpha,
                                            2.0 * alpha, accel_mult, courant_mult, en,
                                            en ? u[p] : 0.0, en ? dudt[p] : 0.0);
            long long level = level_of(dt_max, dt);
            level = level > 0 ? level : 0;
            if (level < nlastlevel && nlastlevel > 1 &&
                nint % (1LL << (level_step - (nlastlevel - 1))) == 0)
                lev = (int)(nlastlevel - 1);
            else if (level > nlastlevel)
                lev = (int)level;
            nlevel[p] = lev;
        }
    }
    unsigned long long cnt = upd;
    int lm = lev;
    for (int o = 32; o > 0; o >>= 1) {
        cnt += __shfl_xor(cnt, o, 64);
        lm = max(lm, __shfl_xor(lm, o, 64));
    }
    if ((threadIdx.x & 63) == 0) {
        unsigned long long *row = ctr_row(red);
        atomicAdd(row + R_PUPDATE, cnt);
        atomicMax(row + R_LEVMAX, (unsigned long long)lm);
    }
}

// sph_timesteps.F90:261-277: integer times follow a change of level_max
__global__ void k_rescale_nlast(int n, int shift, long long *__restrict__ nlast) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    if (shift > 0)
        nlast[p] = nlast[p] * (1LL << shift);
    else
        nlast[p] = nlast[p] / (1LL << (-shift));
}

// advance_runge_kutta.F90:30-112 (RK) and advance_leapfrog_kdk.F90:32-116
// (KDK, second-half kick over the time since the half step)
template <int KDK>
__global__ void k_advance(int n, long long nint, long long level_step, double timestep,
                          double *__restrict__ x, double *__restrict__ y, double *__restrict__ z,
                          double *__restrict__ vx, double *__restrict__ vy,
                          double *__restrict__ vz, const double *__restrict__ ax,
                          const double *__restrict__ ay, const double *__restrict__ az,
                          double *__restrict__ rox, double *__restrict__ roy,
                          double *__restrict__ roz, double *__restrict__ vox,
                          double *__restrict__ voy, double *__restrict__ voz,
                          double *__restrict__ vhx, double *__restrict__ vhy,
                          double *__restrict__ vhz, long long *__restrict__ nlast,
                          const int *__restrict__ nlevel, double *__restrict__ laststep,
                          uint8_t *__restrict__ active, double *__restrict__ u,
                          double *__restrict__ u_old, const double *__restrict__ dudt,
                          unsigned long long *__restrict__ red) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    unsigned act = 0;
    if (p < n) {
        const double a[3] = {ax[p], ay[p], az[p]};
        double vo[3] = {vox[p], voy[p], voz[p]};
        double vh[3] = {vhx[p], vhy[p], vhz[p]};
        if (KDK && active[p]) {   // :34-38
            for (int k = 0; k < 3; k++) vo[k] = vh[k] + 0.5 * a[k] * laststep[p];
            vox[p] = vo[0]; voy[p] = vo[1]; voz[p] = vo[2];
        }
        const double ro[3] = {rox[p], roy[p], roz[p]};
        const long long dn = nint - nlast[p];
        const long long nfull = 1LL << (level_step - nlevel[p]);
        const long long nhalf = nfull / 2;
        const double dt = timestep * (double)dn;
        double rp[3], vp[3];
        if (dn <= nhalf) {
            for (int k = 0; k < 3; k++) {
                rp[k] = ro[k] + vo[k] * dt;
                vp[k] = vo[k] + a[k] * dt;
            }
            if (dn == nhalf) {
                vhx[p] = vp[0]; vhy[p] = vp[1]; vhz[p] = vp[2];
                if (!KDK) act = 1;
            }
        } else if (!KDK) {
            for (int k = 0; k < 3; k++) {
                rp[k] = ro[k] + vh[k] * dt;
                vp[k] = vo[k] + a[k] * dt;
            }
        } else {
            const double dt_half = timestep * (double)(dn - nhalf);
            for (int k = 0; k < 3; k++) {
                rp[k] = ro[k] + vh[k] * dt;
                vp[k] = vh[k] + a[k] * dt_half;
            }
        }
        x[p] = rp[0]; y[p] = rp[1]; z[p] = rp[2];
        vx[p] = vp[0]; vy[p] = vp[1]; vz[p] = vp[2];
        double up = 0.0;
        if (u) {   // INTERNAL_ENERGY (RK :79-81, KDK :86-89)
            const double uo = u_old[p], du = dudt[p];
            up = uo + du * dt;
            if (KDK && up < SMALL_NUMBER) up = uo * exp(-uo / du);
            u[p] = up;
        }
        if (dn == nfull) {
            act = 1;
            nlast[p] = nint;
            laststep[p] = timestep * (double)nfull;
            rox[p] = rp[0]; roy[p] = rp[1]; roz[p] = rp[2];
            vox[p] = vp[0]; voy[p] = vp[1]; voz[p] = vp[2];
            if (u) u_old[p] = up;
        }
        active[p] = (uint8_t)act;
    }
    unsigned long long cnt = act;
    for (int o = 32; o > 0; o >>= 1) cnt += __shfl_xor(cnt, o, 64);
    if ((threadIdx.x & 63) == 0) atomicAdd(ctr_row(red) + R_NACTIVE, cnt);
}

__global__ void k_step_init(int n, const double *__restrict__ x, const double *__restrict__ y,
                            const double *__restrict__ z, const double *__restrict__ vx,
                            const double *__restrict__ vy, const double *__restrict__ vz,
                            double *__restrict__ rox, double *__restrict__ roy,
                            double *__restrict__ roz, double *__restrict__ vox,
                            double *__restrict__ voy, double *__restrict__ voz,
                            double *__restrict__ vhx, double *__restrict__ vhy,
                            double *__restrict__ vhz, long long *__restrict__ nlast,
                            int *__restrict__ nlevel, double *__restrict__ laststep,
                            uint8_t *__restrict__ active) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    rox[p] = x[p]; roy[p] = y[p]; roz[p] = z[p];
    vox[p] = vx[p]; voy[p] = vy[p]; voz[p] = vz[p];
    vhx[p] = vx[p]; vhy[p] = vy[p]; vhz[p] = vz[p];
    nlast[p] = 0;
    nlevel[p] = 0;
    laststep[p] = 0.0;
    active[p] = 0;
}

template <class T>
__global__ void k_gather_step(int n, const int *__restrict__ perm, const T *__restrict__ src,
                              T *__restrict__ dst) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[perm[i]];
}

template <class T>
int dmalloc(T **p, size_t count) {
    if (hipMalloc((void **)p, sizeof(T) * (count ? count : 1)) != hipSuccess)
        return set_error(SERENX_ENOMEM, "hipMalloc failed for step state");
    return SERENX_OK;
}

int red_init(Ctx &c) {
    hipLaunchKernelGGL(k_red_init, dim3(nblk((long)CTR_WORDS)), dim3(TB), 0, c.stream, c.st.red);
    SX_HIP(hipGetLastError());
    return SERENX_OK;
}

// waits for the stream and folds the reduction rows
int red_read(Ctx &c, unsigned long long *r) {
    Ctx::Step &s = c.st;
    SX_HIP(hipMemcpyAsync(s.red_host.data(), s.red, CTR_WORDS * sizeof(unsigned long long),
                          hipMemcpyDeviceToHost, c.stream));
    SX_HIP(hipStreamSynchronize(c.stream));
    ctr_fold(s.red_host.data(), r, R_COUNT, R_MIN_MASK, R_MAX_MASK);
    return SERENX_OK;
}

// sph_timesteps.F90:30-337 (no sinks)
int step_timesteps(Ctx &c) {
    Ctx::Step &s = c.st;
    const int n = c.n;
    hipStream_t st = c.stream;
    const bool neib = s.opt.neib_check != SERENX_NEIB_CHECK_OFF;
    const bool energy = c.opt.thermal == SERENX_THERMAL_ENERGY_EQN;
    const double *u = energy ? c.u : nullptr, *dudt = energy ? c.dudt : nullptr;
    SX_TRY(red_init(c));
    const double alpha = c.opt.alpha;
    if (s.n == s.nresync) {
        s.n = 0;
        hipLaunchKernelGGL(k_ts_all, dim3(nblk(n)), dim3(TB), 0, st, n, c.ax, c.ay, c.az, c.h,
                           c.divv, c.sound, u, dudt, alpha, s.opt.accel_mult,
                           s.opt.courant_mult, s.dt, s.nlast, s.red);
        SX_HIP(hipGetLastError());
        unsigned long long r[R_COUNT];
        SX_TRY(red_read(c, r));
        double dt_min = __builtin_bit_cast(double, (long long)r[R_DTMIN]);
        s.level_max = s.opt.nlevels - 1;
        // :84-91; 2.0**level is a power of two, so ldexp gives the same quotient
        if (s.opt.levels == SERENX_LEVELS_RESTRICTED) {
            const long long level = (long long)(INVLOGETWO * std::log(s.opt.dt_fixed / dt_min)) + 1;
            dt_min = s.opt.dt_fixed / std::ldexp(1.0, (int)level);
        } else if (s.opt.levels == SERENX_LEVELS_FIXED) {
            dt_min = s.opt.dt_fixed / std::ldexp(1.0, (int)s.level_max);
        }
        s.dt_max = dt_min * std::ldexp(1.0, (int)s.level_max);
        hipLaunchKernelGGL(k_ts_levels, dim3(nblk(n)), dim3(TB), 0, st, n, s.dt_max,
                           s.level_max, s.dt, s.nlevel, neib ? s.nminneib : nullptr);
        SX_HIP(hipGetLastError());
        s.level_step = s.level_max + 1;   // :131-135, RK and KDK take half steps
        s.nresync = 1LL << s.level_step;
        s.timestep = s.dt_max / (double)s.nresync;
    } else {
        const long long level_old = s.level_max;
        hipLaunchKernelGGL(k_ts_update, dim3(nblk(n)), dim3(TB), 0, st, n, s.n, s.level_step,
                           s.dt_max, c.ax, c.ay, c.az, c.h, c.divv, c.sound, u, dudt, alpha,
                           s.opt.accel_mult, s.opt.courant_mult, s.nlast, s.nlevel, s.red);
        SX_HIP(hipGetLastError());
        unsigned long long r[R_COUNT];
        SX_TRY(red_read(c, r));
        if (r[R_PUPDATE] > 0) {
            long long level_max = (long long)r[R_LEVMAX];
            if (level_max <= level_old - 1 && level_old > 1 &&
                s.n % (1LL << (s.level_step - (level_old - 1))) == 0)
                level_max = level_old - 1;
            else if (level_max < level_old)
                level_max = level_old;
            if (neib) {
                hipLaunchKernelGGL(k_ts_neib, dim3(nblk(n)), dim3(TB), 0, st, n, s.n,
                                   s.level_step, level_max, s.nlast, s.nlevel, s.nminneib);
                SX_HIP(hipGetLastError());
            }
            const int shift = (int)(level_max - level_old);
            if (shift != 0) {
                hipLaunchKernelGGL(k_rescale_nlast, dim3(nblk(n)), dim3(TB), 0, st, n, shift,
                                   s.nlast);
                SX_HIP(hipGetLastError());
                if (shift > 0) {
                    s.n *= 1LL << shift;
                    s.nresync *= 1LL << shift;
                } else {
                    s.n /= 1LL << (-shift);
                    s.nresync /= 1LL << (-shift);
                }
            }
            s.level_max = level_max;
            s.level_step = s.level_max + 1;
            s.timestep = s.dt_max / (double)s.nresync;
        }
    }
    s.n = s.n + 1;
    s.nsteps = s.nsteps + 1;
    s.time = s.time + s.timestep;
    return SERENX_OK;
}

// check_neighbour_timesteps.F90 (sph_integrate.F90:90-92)
int step_neib_levels(Ctx &c) {
    Ctx::Step &s = c.st;
    hipLaunchKernelGGL(k_neib_levels, dim3(nblk(c.n)), dim3(TB), 0, c.stream, c.n, c.active,
                       s.nlevel, c.pptot, c.pplist, c.ovf_off, c.ovf_list, c.pp_limit,
                       s.nminneib);
    SX_HIP(hipGetLastError());
    return SERENX_OK;
}

int step_advance(Ctx &c) {
    Ctx::Step &s = c.st;
    const int n = c.n;
    c.box_ok = false;   // positions move: the hydro boxes are stale until the next stock
    hipStream_t st = c.stream;
    const bool energy = c.opt.thermal == SERENX_THERMAL_ENERGY_EQN;
    double *u = energy ? c.u : nullptr, *u_old = energy ? s.u_old : nullptr;
    SX_TRY(red_init(c));
    if (s.opt.integrator == SERENX_INT_LFKDK)
        hipLaunchKernelGGL(k_advance<1>, dim3(nblk(n)), dim3(TB), 0, st, n, s.n, s.level_step,
                           s.timestep, c.x, c.y, c.z, c.vx, c.vy, c.vz, c.ax, c.ay, c.az, s.rox,
                           s.roy, s.roz, s.vox, s.voy, s.voz, s.vhx, s.vhy, s.vhz, s.nlast,
                           s.nlevel, s.laststep, c.active, u, u_old, c.dudt, s.red);
    else
        hipLaunchKernelGGL(k_advance<0>, dim3(nblk(n)), dim3(TB), 0, st, n, s.n, s.level_step,
                           s.timestep, c.x, c.y, c.z, c.vx, c.vy, c.vz, c.ax, c.ay, c.az, s.rox,
                           s.roy, s.roz, s.vox, s.voy, s.voz, s.vhx, s.vhy, s.vhz, s.nlast,
                           s.nlevel, s.laststep, c.active, u, u_old, c.dudt, s.red);
    SX_HIP(hipGetLastError());
    const bool reduce = s.opt.neib_check == SERENX_NEIB_CHECK_IMMEDIATE;
    if (reduce) {   // sph_integrate.F90:103-105
        if (s.opt.integrator == SERENX_INT_LFKDK)
            hipLaunchKernelGGL(k_reduce_steps<1>, dim3(nblk(n)), dim3(TB), 0, st, n, s.n,
                               s.level_step, s.timestep, c.x, c.y, c.z, c.vx, c.vy, c.vz,
                               c.ax, c.ay, c.az, s.rox, s.roy, s.roz, s.vox, s.voy, s.voz,
                               s.vhx, s.vhy, s.vhz, s.nlast, s.nlevel, s.nminneib, s.laststep,
                               c.active, u, u_old, s.red);
        else
            hipLaunchKernelGGL(k_reduce_steps<0>, dim3(nblk(n)), dim3(TB), 0, st, n, s.n,
                               s.level_step, s.timestep, c.x, c.y, c.z, c.vx, c.vy, c.vz,
                               c.ax, c.ay, c.az, s.rox, s.roy, s.roz, s.vox, s.voy, s.voz,
                               s.vhx, s.vhy, s.vhz, s.nlast, s.nlevel, s.nminneib, s.laststep,
                               c.active, u, u_old, s.red);
        SX_HIP(hipGetLastError());
    }
    unsigned long long r[R_COUNT];
    SX_TRY(red_read(c, r));
    s.nactive = (long long)r[R_NACTIVE];
    if (reduce && (long long)r[R_NMAX] > s.level_max)
        return set_error(SERENX_ESTATE, "Problem with nminneib > level_max (reduce_timesteps)");
    return SERENX_OK;
}

// sph_integrate.F90:41-113 for one step
int step_once(Ctx &c) {
    Ctx::Step &s = c.st;
    // phase markers as serenx_force_eval records them (serenx_last_timings)
    // and walk statistics of this step's force evaluation
    SX_TRY(zero_words(c.wstat, 2 * CTR_WORDS, c.stream));
    SX_TRY(zero_words(c.icount, 2 * CTR_WORDS, c.stream));
    SX_HIP(hipEventRecord(c.ev[0], c.stream));
    // tree_update.F90:24-88
    const bool build = s.nbuild == s.nsteps;
    if (build) {
        SX_TRY(build_tree(c));   // build + reorder, permutes the step state too
        s.nstock = s.nsteps;
    }
    if (s.nstock == s.nsteps) SX_TRY(stock_tree(c));
    if (build) s.nbuild = s.nbuild + s.opt.nbuildstep;
    if (s.nstock == s.nsteps) s.nstock = s.nstock + (s.opt.integrator == SERENX_INT_LFKDK ? 2 : 1);
    SX_HIP(hipEventRecord(c.ev[1], c.stream));
    // sph_update, thermal, forces for the active particles; with none active
    // (the first step, the half steps of a one-level leapfrog) they change
    // nothing but the tree's hrangemax, which the next force evaluation
    // recomputes from the same h before its neighbour walk reads it (and
    // serenx_export_tree before it exports it), so such a step skips it
    // the gravity walk runs on its own stream beside the neighbour lists,
    // thermal and hydro forces, from the new h on
    c.grav_timed = false;
    GravGuard guard(c.gstream);
    if (s.nactive > 0) {
        guard.armed = true;
        SX_TRY(grav_launch(c));   // cell phase beside the h solve
        SX_TRY(run_sph_h(c));
        // the pair phase needs the new h only; the per-level hmax launches
        // (for the neighbour walk) follow it, so the walk's stream, the
        // longer of the two, does not wait for them
        SX_TRY(grav_pairs(c));    // pair phase beside the lists, thermal and hydro
        SX_TRY(update_hmax(c));
        SX_TRY(run_sph_lists(c));
        SX_TRY(sph_thermal(c));
    }
    SX_HIP(hipEventRecord(c.ev[2], c.stream));
    if (s.nactive > 0) SX_TRY(run_hydro(c));
    SX_HIP(hipEventRecord(c.ev[3], c.stream));
    if (s.nactive > 0) SX_TRY(grav_join(c));
    guard.armed = false;
    SX_HIP(hipEventRecord(c.ev[4], c.stream));
    SX_TRY(step_timesteps(c));
    // every level 0 (and nminneib >= 0 without sinks): the check changes nothing
    if (s.opt.neib_check != SERENX_NEIB_CHECK_OFF && s.nactive > 0 && s.level_max > 0)
        SX_TRY(step_neib_levels(c));
    SX_TRY(step_advance(c));
    return SERENX_OK;
}

}  // namespace

void step_free(Ctx &c) {
    Ctx::Step &s = c.st;
    void *ptrs[] = {s.rox, s.roy, s.roz, s.vox, s.voy, s.voz, s.vhx, s.vhy, s.vhz,
                    s.laststep, s.dt, s.u_old, s.nlast, s.nlevel, s.nminneib, s.red};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    s = Ctx::Step{};
}

int permute_step(Ctx &c, const int *perm) {
    Ctx::Step &s = c.st;
    if (!s.on) return SERENX_OK;
    const int n = c.n;
    double **arrs[] = {&s.rox, &s.roy, &s.roz, &s.vox, &s.voy, &s.voz, &s.vhx, &s.vhy, &s.vhz,
                       &s.laststep, &s.u_old};
    for (double **a : arrs) {
        hipLaunchKernelGGL(k_gather_step<double>, dim3(nblk(n)), dim3(TB), 0, c.stream, n, perm,
                           (const double *)*a, s.dt);
        std::swap(*a, s.dt);
    }
    // the integer arrays go through the double scratch (same element size
    // for nlast; nlevel uses its lower half)
    long long *tl = reinterpret_cast<long long *>(s.dt);
    hipLaunchKernelGGL(k_gather_step<long long>, dim3(nblk(n)), dim3(TB), 0, c.stream, n, perm,
                       (const long long *)s.nlast, tl);
    s.dt = reinterpret_cast<double *>(s.nlast);
    s.nlast = tl;
    int *ti = reinterpret_cast<int *>(s.dt);
    hipLaunchKernelGGL(k_gather_step<int>, dim3(nblk(n)), dim3(TB), 0, c.stream, n, perm,
                       (const int *)s.nlevel, ti);
    // nlevel's old buffer (n ints) is too small to serve as the double
    // scratch, so copy back rather than swap
    SX_HIP(hipMemcpyAsync(s.nlevel, ti, sizeof(int) * n, hipMemcpyDeviceToDevice, c.stream));
    // nminneib (reorder_array.F90:234-237)
    hipLaunchKernelGGL(k_gather_step<int>, dim3(nblk(n)), dim3(TB), 0, c.stream, n, perm,
                       (const int *)s.nminneib, ti);
    SX_HIP(hipMemcpyAsync(s.nminneib, ti, sizeof(int) * n, hipMemcpyDeviceToDevice, c.stream));
    SX_HIP(hipGetLastError());
    return SERENX_OK;
}

int step_begin(Ctx &c, const serenx_step_options &opt) {
    if (opt.integrator != SERENX_INT_RK && opt.integrator != SERENX_INT_LFKDK)
        return set_error(SERENX_EINVAL, "integrator must be SERENX_INT_RK or SERENX_INT_LFKDK");
    if (opt.nlevels < 1 || opt.nlevels > 40)
        return set_error(SERENX_EINVAL, "nlevels must be in 1..40");
    if (opt.nbuildstep < 1) return set_error(SERENX_EINVAL, "nbuildstep must be positive");
    if (!(opt.accel_mult > 0.0) || !(opt.courant_mult > 0.0))
        return set_error(SERENX_EINVAL, "accel_mult and courant_mult must be positive");
    if (opt.levels < SERENX_LEVELS_ADAPTIVE || opt.levels > SERENX_LEVELS_FIXED)
        return set_error(SERENX_EINVAL, "levels must be SERENX_LEVELS_ADAPTIVE, _RESTRICTED or "
                                        "_FIXED");
    if (opt.levels != SERENX_LEVELS_ADAPTIVE && !(opt.dt_fixed > 0.0))
        return set_error(SERENX_EINVAL, "dt_fixed must be positive");
    if (opt.neib_check < SERENX_NEIB_CHECK_OFF || opt.neib_check > SERENX_NEIB_CHECK_IMMEDIATE)
        return set_error(SERENX_EINVAL, "neib_check must be SERENX_NEIB_CHECK_OFF, _END or "
                                        "_IMMEDIATE");
    if (!c.tree_ok || !c.sph_ok)
        return set_error(SERENX_ESTATE, "step_begin needs the setup force evaluation first");
    if (c.ncloud != 1)
        return set_error(SERENX_EINVAL, "integration steps run on a single cloud (one context "
                                        "per cloud)");
    step_free(c);
    serve_free(c);
    Ctx::Step &s = c.st;
    const size_t n = (size_t)c.cap + 1;
    double **arrs[] = {&s.rox, &s.roy, &s.roz, &s.vox, &s.voy, &s.voz, &s.vhx, &s.vhy, &s.vhz,
                       &s.laststep, &s.dt, &s.u_old};
    for (double **a : arrs) SX_TRY(dmalloc(a, n));
    SX_TRY(dmalloc(&s.nlast, n));
    SX_TRY(dmalloc(&s.nlevel, n));
    SX_TRY(dmalloc(&s.nminneib, n));
    SX_TRY(dmalloc(&s.red, CTR_WORDS));
    s.red_host.assign(CTR_WORDS, 0);
    SX_TRY(zero_words(s.nminneib, n, c.stream));
    s.opt = opt;
    s.on = true;
    hipLaunchKernelGGL(k_step_init, dim3(nblk(c.n)), dim3(TB), 0, c.stream, c.n, c.x, c.y, c.z,
                       c.vx, c.vy, c.vz, s.rox, s.roy, s.roz, s.vox, s.voy, s.voz, s.vhx, s.vhy,
                       s.vhz, s.nlast, s.nlevel, s.laststep, c.active);
    SX_HIP(hipGetLastError());
    // u_old = u (initialize_sph_variables_2.F90:75-77)
    SX_HIP(hipMemcpyAsync(s.u_old, c.u, sizeof(double) * c.n, hipMemcpyDeviceToDevice,
                          c.stream));
    SX_HIP(hipStreamSynchronize(c.stream));
    // initialize_seren_variables_2.F90:60-61, initialize_sph_variables_2.F90:165-175
    s.n = 0;
    s.nresync = s.n;
    s.nsteps = 0;
    s.time = 0.0;
    s.timestep = 0.0;
    s.dt_max = 0.0;
    s.level_max = 0;
    s.level_step = 0;
    s.nactive = 0;
    const long long ahead = (opt.integrator == SERENX_INT_LFKDK) ? 2 : 1;
    s.nbuild = s.nsteps + ahead;
    s.nstock = s.nsteps + ahead;
    return SERENX_OK;
}

int step_run(Ctx &c, int count) {
    if (!c.st.on) return set_error(SERENX_ESTATE, "serenx_step before serenx_step_begin");
    for (int k = 0; k < count; k++) SX_TRY(step_once(c));
    return SERENX_OK;
}

int step_fetch(Ctx &c, int32_t *nlevel, int64_t *nlast, double *laststep, int32_t *active,
               double *r_old, double *v_old, double *v_half, int32_t *nminneib) {
    Ctx::Step &s = c.st;
    if (!s.on) return set_error(SERENX_ESTATE, "no step state (serenx_step_begin)");
    const int n = c.n;
    SX_HIP(hipStreamSynchronize(c.stream));
    if (nlevel) SX_HIP(hipMemcpy(nlevel, s.nlevel, sizeof(int) * n, hipMemcpyDeviceToHost));
    if (nminneib)
        SX_HIP(hipMemcpy(nminneib, s.nminneib, sizeof(int) * n, hipMemcpyDeviceToHost));
    if (nlast) SX_HIP(hipMemcpy(nlast, s.nlast, sizeof(long long) * n, hipMemcpyDeviceToHost));
    if (laststep)
        SX_HIP(hipMemcpy(laststep, s.laststep, sizeof(double) * n, hipMemcpyDeviceToHost));
    if (active) {
        std::vector<uint8_t> a(n);
        SX_HIP(hipMemcpy(a.data(), c.active, n, hipMemcpyDeviceToHost));
        for (int p = 0; p < n; p++) active[p] = a[p];
    }
    double *dst[3] = {r_old, v_old, v_half};
    double *srcs[3][3] = {{s.rox, s.roy, s.roz}, {s.vox, s.voy, s.voz}, {s.vhx, s.vhy, s.vhz}};
    std::vector<double> t(n);
    for (int v = 0; v < 3; v++) {
        if (!dst[v]) continue;
        for (int k = 0; k < 3; k++) {
            SX_HIP(hipMemcpy(t.data(), srcs[v][k], sizeof(double) * n, hipMemcpyDeviceToHost));
            for (int p = 0; p < n; p++) dst[v][3 * (size_t)p + k] = t[p];
        }
    }
    return SERENX_OK;
}

// ---- batched serving steps ------------------------------------------------
//
// One kick-drift-kick leapfrog step of every selected cloud of a forest, each
// with its own step size: what sph_integrate.F90 does for one cloud with the
// LFKDK integrator, one time-step level and adaptive steps (the step size is
// the minimum of timestep_size over the cloud's particles, taken after each
// force evaluation), in the operation order of the step wrapper above, so a
// cloud stepped here follows the trajectory of serenx_step on that cloud
// alone (tree rebuilt every step).  Clouds that are not selected are left
// untouched: their particles are inactive in every walk, and walks never
// leave their own cloud, so a cloud's results do not depend on which other
// clouds share the call.

namespace {

__global__ void k_serve_select(int n, const int *__restrict__ cid,
                               const uint8_t *__restrict__ sel, uint8_t *__restrict__ active) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p < n) active[p] = sel[cid[p]];
}

// first kick and drift with the acceleration of the last evaluation (as
// k_advance over the half step and the full step: v_half = v + a dt/2,
// x = x + v_half dt); v is the predicted full-step velocity v_half + a dt/2
// that the force evaluation sees (k_advance, dn > nhalf)
__global__ void k_serve_kick_drift(int n, const int *__restrict__ cid,
                                   const uint8_t *__restrict__ active,
                                   const double *__restrict__ dtc, double *__restrict__ x,
                                   double *__restrict__ y, double *__restrict__ z,
                                   double *__restrict__ vx, double *__restrict__ vy,
                                   double *__restrict__ vz, const double *__restrict__ ax,
                                   const double *__restrict__ ay, const double *__restrict__ az,
                                   double *__restrict__ vhx, double *__restrict__ vhy,
                                   double *__restrict__ vhz) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n || !active[p]) return;
    const double dt = dtc[cid[p]];
    const double dth = dt / 2.0;                     // the wrapper's timestep
    const double a[3] = {ax[p], ay[p], az[p]};
    const double v[3] = {vx[p], vy[p], vz[p]};
    double vh[3], r[3] = {x[p], y[p], z[p]};
    for (int k = 0; k < 3; k++) {
        vh[k] = v[k] + a[k] * dth;
        r[k] = r[k] + vh[k] * (dth * 2.0);
    }
    x[p] = r[0]; y[p] = r[1]; z[p] = r[2];
    vhx[p] = vh[0]; vhy[p] = vh[1]; vhz[p] = vh[2];
    vx[p] = vh[0] + a[0] * dth;
    vy[p] = vh[1] + a[1] * dth;
    vz[p] = vh[2] + a[2] * dth;
}

// second kick with the new acceleration (advance_leapfrog_kdk.F90:34-38)
__global__ void k_serve_kick(int n, const int *__restrict__ cid,
                             const uint8_t *__restrict__ active, const double *__restrict__ dtc,
                             double *__restrict__ vx, double *__restrict__ vy,
                             double *__restrict__ vz, const double *__restrict__ ax,
                             const double *__restrict__ ay, const double *__restrict__ az,
                             const double *__restrict__ vhx, const double *__restrict__ vhy,
                             const double *__restrict__ vhz) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n || !active[p]) return;
    const double laststep = dtc[cid[p]];
    vx[p] = vhx[p] + 0.5 * ax[p] * laststep;
    vy[p] = vhy[p] + 0.5 * ay[p] * laststep;
    vz[p] = vhz[p] + 0.5 * az[p] * laststep;
}

__global__ void k_serve_dt_init(int ncloud, const uint8_t *__restrict__ sel,
                                unsigned long long *__restrict__ dtbits) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k < ncloud && sel[k]) dtbits[k] = (unsigned long long)__double_as_longlong(DT_INIT);
}

// timestep_size of every active particle, reduced to a minimum per cloud (a
// positive double's bits order as the double).  The lanes of a wave mostly
// belong to one cloud: then one atomic per wave.
__global__ void k_serve_dt(int n, const int *__restrict__ cid, const uint8_t *__restrict__ active,
                           const double *__restrict__ ax, const double *__restrict__ ay,
                           const double *__restrict__ az, const double *__restrict__ h,
                           const double *__restrict__ divv, const double *__restrict__ sound,
                           double alpha, double accel_mult, double courant_mult,
                           unsigned long long *__restrict__ dtbits) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    const bool on = p < n && active[p];
    double d = DT_INIT;
    int k = -1;
    if (on) {
        d = timestep_size(ax[p], ay[p], az[p], h[p], divv[p], sound[p], alpha, 2.0 * alpha,
                          accel_mult, courant_mult, false, 0.0, 0.0);
        k = cid[p];
    }
    const unsigned long long m = __ballot(on);
    if (m == 0) return;
    unsigned long long b = (unsigned long long)__double_as_longlong(d);
    const int k0 = __shfl(k, __builtin_amdgcn_readfirstlane(__ffsll((long long)m) - 1), 64);
    if (__all(!on || k == k0)) {
        for (int o = 32; o > 0; o >>= 1) {
            const unsigned long long t = __shfl_xor(b, o, 64);
            b = t < b ? t : b;
        }
        if ((threadIdx.x & 63) == 0) atomicMin(dtbits + k0, b);
    } else if (on) {
        atomicMin(dtbits + k, b);
    }
}

// the selected clouds advance their time by the step just taken and take
// the new step size
__global__ void k_serve_fold(int ncloud, const uint8_t *__restrict__ sel,
                             const unsigned long long *__restrict__ dtbits,
                             double *__restrict__ dtc, double *__restrict__ tcl, int advance) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= ncloud || !sel[k]) return;
    if (advance) tcl[k] = tcl[k] + dtc[k];
    dtc[k] = __longlong_as_double((long long)dtbits[k]);
}

__global__ void k_zero_doubles(int n, double *__restrict__ a) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = 0.0;
}

__global__ void k_fill_u8(int n, uint8_t v, uint8_t *__restrict__ a) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

}  // namespace

void serve_free(Ctx &c) {
    Ctx::Serve &v = c.sv;
    void *ptrs[] = {v.vhx, v.vhy, v.vhz, v.dtbits, v.dt, v.time, v.sel};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    if (v.sel_host) (void)hipHostFree(v.sel_host);
    v = Ctx::Serve{};
}

int permute_serve(Ctx &c, const int *perm) {
    Ctx::Serve &v = c.sv;
    if (!v.on) return SERENX_OK;
    double **arrs[] = {&v.vhx, &v.vhy, &v.vhz};
    for (double **a : arrs) {
        hipLaunchKernelGGL(k_gather_step<double>, dim3(nblk(c.n)), dim3(TB), 0, c.stream, c.n,
                           perm, (const double *)*a, c.tmpd);
        std::swap(*a, c.tmpd);
    }
    SX_HIP(hipGetLastError());
    return SERENX_OK;
}

// the step sizes of the selected clouds from the current accelerations
static int serve_timesteps(Ctx &c, int advance) {
    Ctx::Serve &v = c.sv;
    hipLaunchKernelGGL(k_serve_dt_init, dim3(nblk(c.ncloud)), dim3(TB), 0, c.stream, c.ncloud,
                       v.sel, v.dtbits);
    hipLaunchKernelGGL(k_serve_dt, dim3(nblk(c.n)), dim3(TB), 0, c.stream, c.n, c.cid, c.active,
                       c.ax, c.ay, c.az, c.h, c.divv, c.sound, c.opt.alpha, v.accel_mult,
                       v.courant_mult, v.dtbits);
    hipLaunchKernelGGL(k_serve_fold, dim3(nblk(c.ncloud)), dim3(TB), 0, c.stream, c.ncloud,
                       v.sel, v.dtbits, v.dt, v.time, advance);
    SX_HIP(hipGetLastError());
    return SERENX_OK;
}

int serve_begin(Ctx &c, double accel_mult, double courant_mult) {
    if (!(accel_mult > 0.0) || !(courant_mult > 0.0))
        return set_error(SERENX_EINVAL, "accel_mult and courant_mult must be positive");
    if (c.opt.thermal != SERENX_THERMAL_ISOTHERMAL)
        return set_error(SERENX_EINVAL, "serving steps support the isothermal option only");
    if (!c.tree_ok || !c.sph_ok)
        return set_error(SERENX_ESTATE, "serve_begin needs the setup force evaluation first");
    step_free(c);
    serve_free(c);
    Ctx::Serve &v = c.sv;
    const size_t n = (size_t)c.cap + 1, nk = (size_t)c.ncloud;
    SX_TRY(dmalloc(&v.vhx, n));
    SX_TRY(dmalloc(&v.vhy, n));
    SX_TRY(dmalloc(&v.vhz, n));
    SX_TRY(dmalloc(&v.dtbits, nk));
    SX_TRY(dmalloc(&v.dt, nk));
    SX_TRY(dmalloc(&v.time, nk));
    SX_TRY(dmalloc(&v.sel, nk));
    SX_HIP(hipHostMalloc((void **)&v.sel_host, nk));
    v.accel_mult = accel_mult;
    v.courant_mult = courant_mult;
    v.ncloud = c.ncloud;
    v.nsteps = 0;
    v.on = true;
    // every cloud selected: the first step sizes from the setup evaluation
    hipLaunchKernelGGL(k_fill_u8, dim3(nblk(c.ncloud)), dim3(TB), 0, c.stream, c.ncloud,
                       (uint8_t)1, v.sel);
    hipLaunchKernelGGL(k_fill_u8, dim3(nblk(c.n)), dim3(TB), 0, c.stream, c.n, (uint8_t)1,
                       c.active);
    hipLaunchKernelGGL(k_zero_doubles, dim3(nblk(c.ncloud)), dim3(TB), 0, c.stream, c.ncloud,
                       v.time);
    SX_TRY(serve_timesteps(c, 0));
    SX_HIP(hipStreamSynchronize(c.stream));
    return SERENX_OK;
}

int serve_step(Ctx &c, int ncloud, const uint8_t *select) {
    Ctx::Serve &v = c.sv;
    if (!v.on) return set_error(SERENX_ESTATE, "serenx_serve_step before serenx_serve_begin");
    if (ncloud != c.ncloud || ncloud != v.ncloud)
        return set_error(SERENX_EINVAL, "selection length differs from the number of clouds");
    bool any = false;
    for (int k = 0; k < ncloud; k++) {
        v.sel_host[k] = select[k] ? 1 : 0;
        any = any || select[k];
    }
    if (!any) return SERENX_OK;
    SX_HIP(hipMemcpyAsync(v.sel, v.sel_host, ncloud, hipMemcpyHostToDevice, c.stream));
    hipLaunchKernelGGL(k_serve_select, dim3(nblk(c.n)), dim3(TB), 0, c.stream, c.n, c.cid, v.sel,
                       c.active);
    hipLaunchKernelGGL(k_serve_kick_drift, dim3(nblk(c.n)), dim3(TB), 0, c.stream, c.n, c.cid,
                       c.active, v.dt, c.x, c.y, c.z, c.vx, c.vy, c.vz, c.ax, c.ay, c.az, v.vhx,
                       v.vhy, v.vhz);
    SX_HIP(hipGetLastError());
    c.box_ok = false;   // positions moved: hydro boxes are stale until the stock
    SX_TRY(run_force_eval(c, 0));
    hipLaunchKernelGGL(k_serve_kick, dim3(nblk(c.n)), dim3(TB), 0, c.stream, c.n, c.cid,
                       c.active, v.dt, c.vx, c.vy, c.vz, c.ax, c.ay, c.az, v.vhx, v.vhy, v.vhz);
    SX_TRY(serve_timesteps(c, 1));
    // the selection buffer is reused by the next call: wait for this one
    SX_HIP(hipStreamSynchronize(c.stream));
    v.nsteps++;
    return SERENX_OK;
}

int serve_info(Ctx &c, int ncloud, double *dt, double *time) {
    Ctx::Serve &v = c.sv;
    if (!v.on) return set_error(SERENX_ESTATE, "no serving state (serenx_serve_begin)");
    if (ncloud != v.ncloud) return set_error(SERENX_EINVAL, "wrong number of clouds");
    SX_HIP(hipStreamSynchronize(c.stream));
    if (dt) SX_HIP(hipMemcpy(dt, v.dt, sizeof(double) * ncloud, hipMemcpyDeviceToHost));
    if (time) SX_HIP(hipMemcpy(time, v.time, sizeof(double) * ncloud, hipMemcpyDeviceToHost));
    return SERENX_OK;
}

}  // namespace serenx
