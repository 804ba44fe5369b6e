// pmw_stats.hip -- total mass and total energy (compute_stats) on gfx950.
//
// Pass 1: block b takes rows b, b + gridDim.x, ... and its lanes the columns
// t, t + 256, ... of each (coalesced, no index division); each lane
// accumulates its cells in that fixed order, then the block reduces with
// 64-lane shuffles and one LDS exchange between the 4 waves, writing one
// (mass, energy) partial per block.
// Pass 2: one block sums the partials in a fixed order and scales by dx*dz.
// For a given grid size the summation order is fixed, so results are
// reproducible run to run (they differ from NumPy's pairwise order only in
// rounding).
//
// pmw_compute_stats_noalloc runs the same two kernels with the partials in
// `scratch`, a file-scope device array, so that it needs no workspace and
// allocates nothing.
#include <mutex>

#include "pmw_common.h"

using namespace pmw;

namespace {

constexpr int kSThreads = 256;
constexpr int kSMaxBlocks = 1024;

__device__ __forceinline__ double wave_sum(double x) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) x += __shfl_xor(x, off, 64);
  return x;
}

// Block-wide sum of two values; result valid in thread 0.
__device__ __forceinline__ void block_sum2(double& a, double& b) {
  __shared__ double sa[kSThreads / 64], sb[kSThreads / 64];
  a = wave_sum(a);
  b = wave_sum(b);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) {
    sa[wv] = a;
    sb[wv] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    a = ((sa[0] + sa[1]) + sa[2]) + sa[3];
    b = ((sb[0] + sb[1]) + sb[2]) + sb[3];
  }
}

// Per-cell terms of post/stats.py:13-30.
__device__ __forceinline__ void cell_stats(double D, double U, double W, double T,
                                           double hyr, double hyt, double& rho,
                                           double& en) {
  rho = D + hyr;
  const double u = U / rho;
  const double w = W / rho;
  const double th = (T + hyt) / rho;
  // both powers have a positive base and an exponent in (0, 1.5] (gamma and
  // rd/cp): the table-driven power of the flux kernels, correctly rounded in
  // all but ~4e-6 of arguments, instead of the general device pow
  const double p = kC0 * pow_pos(rho * th, kGamma);
  const double temp = th / pow_pos(kP0 / p, kRd / kCp);
  const double ke = rho * (u * u + w * w);
  const double ie = rho * kCv * temp;
  en = ke + ie;
}

__global__ __launch_bounds__(kSThreads) void k_stats_partial(
    const double* __restrict__ s, const double* __restrict__ hyr,
    const double* __restrict__ hyt, Grid g, double* __restrict__ part) {
  double m = 0.0, e = 0.0;
  for (int k = blockIdx.x; k < g.nz; k += gridDim.x) {
    const int pr = k + kHs;
    const double hr = hyr[pr], ht = hyt[pr];
    for (int i = threadIdx.x; i < g.nx; i += blockDim.x) {
      const int pc = i + kHs;
      double rho, en;
      cell_stats(s[g.at(0, pr, pc)], s[g.at(1, pr, pc)], s[g.at(2, pr, pc)],
                 s[g.at(3, pr, pc)], hr, ht, rho, en);
      m += rho;
      e += en;
    }
  }
  block_sum2(m, e);
  if (threadIdx.x == 0) {
    part[blockIdx.x] = m;
    part[gridDim.x + blockIdx.x] = e;
  }
}

__global__ __launch_bounds__(kSThreads) void k_stats_final(
    const double* __restrict__ part, int nb, double dx, double dz,
    double* __restrict__ out2) {
  double m = 0.0, e = 0.0;
  for (int j = threadIdx.x; j < nb; j += blockDim.x) {
    m += part[j];
    e += part[nb + j];
  }
  block_sum2(m, e);
  if (threadIdx.x == 0) {
    out2[0] = m * dx * dz;
    out2[1] = e * dx * dz;
  }
}

int stats_blocks(int nx, int nz) {
  (void)nx;
  return nz < kSMaxBlocks ? nz : kSMaxBlocks;  // one or more rows per block
}

// Partials of pmw_compute_stats_noalloc: one (mass, energy) pair per block of
// pass 1, which never has more than kSMaxBlocks blocks, so every grid fits.
constexpr int kScratchLen = 2 * kSMaxBlocks;  // 16 KiB per GPU
constexpr int kMaxDevices = 64;

}  // namespace

// Each GPU holds its own copy at its own address; host code reaches it only
// through the address hipGetSymbolAddress returns on that GPU (the name
// itself is a host-side stand-in, not a device address).
static __device__ __align__(256) double scratch[kScratchLen];

namespace {

std::mutex g_scratch_mu;
double* g_scratch_dev[kMaxDevices] = {};

// Device address of `scratch` on the current device, looked up once per
// device. Returns PMW_OK, PMW_EINVAL or a hipError_t.
int scratch_on_current_device(double** out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return (int)e;
  if (dev < 0 || dev >= kMaxDevices) return PMW_EINVAL;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return (int)e;
    if (!p) return PMW_EINVAL;
    g_scratch_dev[dev] = (double*)p;
  }
  *out = g_scratch_dev[dev];
  return PMW_OK;
}

int launch_stats(const double* state, const double* hy_dens_cell,
                 const double* hy_dens_theta_cell, int nx, int nz, double dx,
                 double dz, double* part, double* out2, hipStream_t st) {
  const int nb = stats_blocks(nx, nz);
  Grid g(nx, nz);
  k_stats_partial<<<nb, kSThreads, 0, st>>>(state, hy_dens_cell, hy_dens_theta_cell,
                                            g, part);
  PMW_CHECK_LAUNCH();
  k_stats_final<<<1, kSThreads, 0, st>>>(part, nb, dx, dz, out2);
  PMW_CHECK_LAUNCH();
  return PMW_OK;
}

}  // namespace

extern "C" {

long long pmw_stats_workspace_size(int nx, int nz) {
  if (nx < 1 || nz < 1) return 0;
  return 2LL * stats_blocks(nx, nz);
}

int pmw_compute_stats(const double* state, const double* hy_dens_cell,
                      const double* hy_dens_theta_cell, int nx, int nz,
                      double dx, double dz, double* workspace,
                      long long workspace_len, double* out2, void* stream) {
  if (nx < 1 || nz < 1 || !state || !hy_dens_cell || !hy_dens_theta_cell ||
      !workspace || !out2)
    return PMW_EINVAL;
  const int nb = stats_blocks(nx, nz);
  if (workspace_len < 2LL * nb) return PMW_ESPACE;
  return launch_stats(state, hy_dens_cell, hy_dens_theta_cell, nx, nz, dx, dz,
                      workspace, out2, (hipStream_t)stream);
}

int pmw_compute_stats_noalloc(const double* state, const double* hy_dens_cell,
                              const double* hy_dens_theta_cell, int nx, int nz,
                              double dx, double dz, double* out2, void* stream) {
  if (nx < 1 || nz < 1 || !state || !hy_dens_cell || !hy_dens_theta_cell || !out2)
    return PMW_EINVAL;
  const int nb = stats_blocks(nx, nz);
  if (2LL * nb > kScratchLen) return PMW_ESPACE;  // cannot happen: nb <= kSMaxBlocks
  hipStream_t st = (hipStream_t)stream;
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return (int)e;
  if (st) {  // the partials must live on the GPU the stream runs on
    hipDevice_t sdev = -1;
    e = hipStreamGetDevice(st, &sdev);
    if (e != hipSuccess) return (int)e;
    if ((int)sdev != dev) return PMW_EINVAL;
  }
  double* part = nullptr;
  const int rc = scratch_on_current_device(&part);
  if (rc != PMW_OK) return rc;
  // every partial is written by pass 1 before pass 2 reads it: nothing to clear
  return launch_stats(state, hy_dens_cell, hy_dens_theta_cell, nx, nz, dx, dz, part,
                      out2, st);
}

}  // extern "C"
