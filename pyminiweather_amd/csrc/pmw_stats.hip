// pmw_stats.hip -- total mass and total energy (compute_stats) on gfx950.
//
// Pass 1: a fixed-size grid-stride kernel; each lane accumulates its cells in
// a fixed order, then the block reduces with 64-lane shuffles and one LDS
// exchange between the 4 waves, writing one (mass, energy) partial per block.
// Pass 2: one block sums the partials in a fixed order and scales by dx*dz.
// For a given grid size the summation order is fixed, so results are
// reproducible run to run (they differ from NumPy's pairwise order only in
// rounding).
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
  const double p = kC0 * pow(rho * th, kGamma);
  const double temp = th / pow(kP0 / p, kRd / kCp);
  const double ke = rho * (u * u + w * w);
  const double ie = rho * kCv * temp;
  en = ke + ie;
}

__global__ __launch_bounds__(kSThreads) void k_stats_partial(
    const double* __restrict__ s, const double* __restrict__ hyr,
    const double* __restrict__ hyt, Grid g, double* __restrict__ part) {
  const long long n = (long long)g.nz * g.nx;
  double m = 0.0, e = 0.0;
  for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < n;
       idx += (long long)gridDim.x * blockDim.x) {
    const int k = (int)(idx / g.nx), i = (int)(idx % g.nx);
    const int pr = k + kHs, pc = i + kHs;
    double rho, en;
    cell_stats(s[g.at(0, pr, pc)], s[g.at(1, pr, pc)], s[g.at(2, pr, pc)],
               s[g.at(3, pr, pc)], hyr[pr], hyt[pr], rho, en);
    m += rho;
    e += en;
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
  const long long n = (long long)nx * nz;
  long long b = (n + kSThreads - 1) / kSThreads;
  if (b > kSMaxBlocks) b = kSMaxBlocks;
  if (b < 1) b = 1;
  return (int)b;
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
  hipStream_t st = (hipStream_t)stream;
  Grid g(nx, nz);
  k_stats_partial<<<nb, kSThreads, 0, st>>>(state, hy_dens_cell, hy_dens_theta_cell,
                                            g, workspace);
  PMW_CHECK_LAUNCH();
  k_stats_final<<<1, kSThreads, 0, st>>>(workspace, nb, dx, dz, out2);
  PMW_CHECK_LAUNCH();
  return PMW_OK;
}

}  // extern "C"
