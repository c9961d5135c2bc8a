// diag.hip -- FP64 conservation diagnostics of the resident particles on
// gfx950: the sums of diagnostics.F90:85-134 (mass, centre of mass and its
// velocity, linear and angular momentum, net force, kinetic, gravitational
// and thermal energy).
//
// Two launches, both deterministic (the same particles in the same tree
// order give the same bits): k_diag_partial reduces a contiguous slice of
// particles per block in LDS and writes DIAG_SUMS partial sums per block;
// k_diag_final adds the block partials in block order.  The partial sums are
// the scratch buffer of the reduction.
#include <algorithm>
#include <mutex>

#include "serenx_internal.h"

namespace serenx {

namespace {

constexpr int TB = 256;

// per-particle terms, summed: m, m r (3), m v (3), m r x v (3), m a (3),
// m v.v, m gpot, m P/rho (isothermal) or m u (energy equation)
enum {
    S_M = 0, S_MR = 1, S_MV = 4, S_ANG = 7, S_F = 10, S_KE = 13, S_GPE = 14, S_U = 15,
    S_COUNT = 16
};

__global__ void __launch_bounds__(TB) k_diag_partial(
    int n, int per_block, const double *__restrict__ x, const double *__restrict__ y,
    const double *__restrict__ z, const double *__restrict__ m, const double *__restrict__ vx,
    const double *__restrict__ vy, const double *__restrict__ vz, const double *__restrict__ ax,
    const double *__restrict__ ay, const double *__restrict__ az, const double *__restrict__ gpot,
    const double *__restrict__ press, const double *__restrict__ rho,
    const double *__restrict__ u, double *__restrict__ partial) {
    __shared__ double red[S_COUNT][TB];
    double s[S_COUNT];
    for (int k = 0; k < S_COUNT; k++) s[k] = 0.0;
    const int lo = blockIdx.x * per_block;
    const int hi = min(n, lo + per_block);
    for (int p = lo + (int)threadIdx.x; p < hi; p += TB) {
        const double mp = m[p];
        const double r[3] = {x[p], y[p], z[p]}, v[3] = {vx[p], vy[p], vz[p]};
        s[S_M] += mp;
        for (int k = 0; k < 3; k++) {
            s[S_MR + k] += mp * r[k];
            s[S_MV + k] += mp * v[k];
        }
        s[S_ANG + 0] += mp * (r[1] * v[2] - r[2] * v[1]);
        s[S_ANG + 1] += mp * (r[2] * v[0] - r[0] * v[2]);
        s[S_ANG + 2] += mp * (r[0] * v[1] - r[1] * v[0]);
        s[S_F + 0] += mp * ax[p];
        s[S_F + 1] += mp * ay[p];
        s[S_F + 2] += mp * az[p];
        s[S_KE] += mp * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        s[S_GPE] += mp * gpot[p];
        s[S_U] += u ? mp * u[p] : mp * (press[p] / rho[p]);
    }
    for (int k = 0; k < S_COUNT; k++) red[k][threadIdx.x] = s[k];
    __syncthreads();
    for (int w = TB / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
            for (int k = 0; k < S_COUNT; k++) red[k][threadIdx.x] += red[k][threadIdx.x + w];
        __syncthreads();
    }
    if ((int)threadIdx.x < S_COUNT)
        partial[(size_t)blockIdx.x * S_COUNT + threadIdx.x] = red[threadIdx.x][0];
}

// adds the block partials in block order and forms the reported values
//   [0] mtot  [1..3] rcom  [4..6] vcom  [7..9] mom  [10..12] angmom
//   [13..15] force  [16] ketot  [17] gpetot  [18] utot  [19] etot
// (diagnostics.F90:211-238: ketot = sum m v^2 / 2, gpetot = sum m gpot / 2,
// utot = 3/2 sum m P/rho isothermal or sum m u, etot = ketot - gpetot + utot)
__global__ void __launch_bounds__(64) k_diag_final(int nblocks, bool energy,
                                                   const double *__restrict__ partial,
                                                   double *__restrict__ out) {
    const int k = threadIdx.x;
    __shared__ double tot[S_COUNT];
    if (k < S_COUNT) {
        double t = 0.0;
        for (int b = 0; b < nblocks; b++) t += partial[(size_t)b * S_COUNT + k];
        tot[k] = t;
    }
    __syncthreads();
    if (k != 0) return;
    const double mtot = tot[S_M];
    out[0] = mtot;
    for (int j = 0; j < 3; j++) {
        out[1 + j] = tot[S_MR + j] / mtot;
        out[4 + j] = tot[S_MV + j] / mtot;
        out[7 + j] = tot[S_MV + j];
        out[10 + j] = tot[S_ANG + j];
        out[13 + j] = tot[S_F + j];
    }
    const double ke = 0.5 * tot[S_KE];
    const double gpe = 0.5 * tot[S_GPE];
    const double ut = energy ? tot[S_U] : 1.5 * tot[S_U];
    out[16] = ke;
    out[17] = gpe;
    out[18] = ut;
    out[19] = ke - gpe + ut;
}

}  // namespace

// The number of blocks of the partial-sum launch for n particles: every
// block reduces per_block consecutive particles.
int diag_blocks(int n, int *per_block) {
    const int nb = std::min(DIAG_MAX_BLOCKS, std::max(1, (n + TB - 1) / TB));
    *per_block = (n + nb - 1) / nb;
    return nb;
}

// Scratch of serenx_diagnostics_noalloc: the block partial sums of
// k_diag_partial, DIAG_MAX_BLOCKS x DIAG_SUMS doubles (128 KiB), one copy
// per device for the life of the process.  In host code its name is not a
// device address: kernels receive the address hipGetSymbolAddress gives for
// the device (diag_scratch).
static __device__ __align__(256) double scratch[(size_t)DIAG_MAX_BLOCKS * DIAG_SUMS];

namespace {
constexpr int SCRATCH_DEVICES = 64;
std::mutex scratch_mu;
double *scratch_ptr[SCRATCH_DEVICES] = {};
}  // namespace

int diag_scratch_setup(int device) {
    if (device < 0 || device >= SCRATCH_DEVICES)
        return set_error(SERENX_EINVAL, "device index beyond the diagnostics scratch table");
    std::lock_guard<std::mutex> lock(scratch_mu);
    if (scratch_ptr[device]) return SERENX_OK;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return set_error(SERENX_EHIP, std::string("hipGetSymbolAddress of the diagnostics "
                                                  "scratch failed: ") + hipGetErrorString(e));
    scratch_ptr[device] = static_cast<double *>(p);
    return SERENX_OK;
}

double *diag_scratch(int device) {
    if (device < 0 || device >= SCRATCH_DEVICES) return nullptr;
    std::lock_guard<std::mutex> lock(scratch_mu);
    return scratch_ptr[device];
}

int diagnostics(Ctx &c, double *partial, double *dout) {
    int per_block = 0;
    const int nb = diag_blocks(c.n, &per_block);
    const bool energy = c.opt.thermal == SERENX_THERMAL_ENERGY_EQN;
    hipLaunchKernelGGL(k_diag_partial, dim3(nb), dim3(TB), 0, c.stream, c.n, per_block, c.x, c.y,
                       c.z, c.m, c.vx, c.vy, c.vz, c.ax, c.ay, c.az, c.gpot, c.press, c.rho,
                       energy ? c.u : nullptr, partial);
    hipLaunchKernelGGL(k_diag_final, dim3(1), dim3(64), 0, c.stream, nb, energy, partial, dout);
    SX_HIP(hipGetLastError());
    return SERENX_OK;
}

}  // namespace serenx

extern "C" int serenx_diagnostics_noalloc(serenx_ctx *ctx, double *out_dev) {
    using namespace serenx;
    if (!ctx || !out_dev) return set_error(SERENX_EINVAL, "null argument");
    Ctx &c = *reinterpret_cast<Ctx *>(ctx);
    if (c.n <= 0) return set_error(SERENX_ESTATE, "no particles loaded");
    SX_HIP(hipSetDevice(c.device));
    double *part = diag_scratch(c.device);
    if (!part)
        return set_error(SERENX_ESTATE, "diagnostics scratch not set up on this device "
                                        "(serenx_create sets it up)");
    int per_block = 0;
    const int nb = diag_blocks(c.n, &per_block);
    if (nb > DIAG_MAX_BLOCKS || per_block <= 0)
        return set_error(SERENX_EINVAL, "particle count beyond the diagnostics scratch "
                                        "(1024 partial-sum rows)");
    // every block writes all of its row before k_diag_final reads the rows
    // [0, nb), so no row needs clearing between calls
    return diagnostics(c, part, out_dev);
}
