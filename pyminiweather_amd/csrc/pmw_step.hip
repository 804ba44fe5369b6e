// pmw_step.hip -- halo, per-function and fused RK-stage kernels for the
// PyMiniWeather time step on gfx950.
//
// Fused stage kernels (the hot path):
//  * x: one 256-thread block per (row, 252-cell tile). The block stages the
//    4 variables' forcing values of its tile plus 2+2 neighbours in LDS
//    (periodic wrap instead of reading halo columns), each thread computes one
//    face flux into LDS, then one cell's tendency and RK update. Blocks whose
//    window covers a halo column also store the wrapped value there, which is
//    exactly set_bc_x.
//  * z: a 256-thread block holds a tile of 8 columns x 32 rows of the
//    forcing state in LDS (the wall halo rows synthesised from the edge rows),
//    computes the 29 face fluxes of its 28 output rows into LDS, then each
//    cell's tendency and RK update; the 4 extra rows per 28 are all the
//    forcing that is read twice (measured against 8x64, 8x48, 4x32, 16x32,
//    16x64, 4x128 and 8x128 tiles, and a column-marching kernel: the small
//    tile was fastest). Extra blocks after the tiles write the z halos of the
//    forcing buffer (set_bc_z) concurrently; no compute block reads them.
// No kernel reads a location another block of the same launch writes, which
// is why `out` may not alias `forcing`.
#include "pmw_common.h"
#include "pmw_fused.h"

using namespace pmw;

namespace {

constexpr int kXThreads = 256;
constexpr int kXCells = kXThreads - 4;  // cells per x tile
constexpr int kZtCols = 8;              // z tile columns
constexpr int kZtRows = 32;             // z tile rows loaded (28 output rows)
constexpr int kNumXcd = 8;              // XCDs of an MI355X

inline bool bad_ptr(const void* p) { return p == nullptr; }

// ------------------------------------------------------------------ halos --

__global__ void k_halo_x(double* __restrict__ s, Grid g, int bc,
                         const int* __restrict__ inj_rows,
                         const double* __restrict__ hyr,
                         const double* __restrict__ hyt) {
  // one thread per (interior row, halo column slot 0..3)
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= g.nz * 4) return;
  const int k = idx >> 2, slot = idx & 3;
  const int prow = k + kHs;
  const int nx = g.nx;
  if (slot >= 2) {
    if (bc == PMW_BC_INJECTION) return;  // right halo left as is (bcs.py:37)
    const int c = nx + slot;             // nx+2, nx+3 <- 2, 3
    for (int v = 0; v < kNv; ++v) s[g.at(v, prow, c)] = s[g.at(v, prow, c - nx)];
    return;
  }
  const int c = slot;  // 0, 1 <- nx, nx+1
  double q[4];
  for (int v = 0; v < kNv; ++v) q[v] = s[g.at(v, prow, c + nx)];
  if (bc == PMW_BC_INJECTION && inj_rows[k]) {
    const double r = q[0] + hyr[prow];
    q[1] = r * 50.0;
    q[3] = r * 298.0 - hyt[prow];
  }
  for (int v = 0; v < kNv; ++v) s[g.at(v, prow, c)] = q[v];
}

// Right x-halo columns nx+2, nx+3 of the interior rows, all 4 variables,
// from `src` to `dst`. With the injection inflow set_bc_x leaves that halo as
// the forcing buffer holds it (bcs.py:37), and the reference's stage 2 reads
// state_tmp's (step.py:122-131); the stage 2 here reads state_work, so
// state_tmp's right halo is copied there first.
__global__ void k_copy_right_halo(const double* __restrict__ src,
                                  double* __restrict__ dst, Grid g) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;  // (var, row, slot)
  if (idx >= kNv * g.nz * 2) return;
  const int slot = idx & 1;
  const int k = (idx >> 1) % g.nz;
  const int v = (idx >> 1) / g.nz;
  const long long a = g.at(v, k + kHs, g.nx + kHs + slot);
  dst[a] = src[a];
}

__global__ void k_halo_z(double* __restrict__ s, Grid g,
                         const double* __restrict__ hy) {
  // one thread per column 0..nx+3; writes rows 0,1,nz+2,nz+3
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= g.ld) return;
  const int lo = kHs, hi = g.nz + kHs - 1;
  for (int v = 0; v < kNv; ++v) {
    const double a = s[g.at(v, lo, c)];
    const double b = s[g.at(v, hi, c)];
    s[g.at(v, 0, c)] = wall_halo(v, a, hy[lo], hy[0]);
    s[g.at(v, 1, c)] = wall_halo(v, a, hy[lo], hy[1]);
    s[g.at(v, g.nz + 2, c)] = wall_halo(v, b, hy[hi], hy[g.nz + 2]);
    s[g.at(v, g.nz + 3, c)] = wall_halo(v, b, hy[hi], hy[g.nz + 3]);
  }
}

// --------------------------------------------------------- per function --

__global__ void k_interp_x(const double* __restrict__ s, double* __restrict__ vals,
                           double* __restrict__ d3, Grid g) {
  const int nf = g.nx + 1;
  const long long n = (long long)kNv * g.nz * nf;
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n) return;
  const int i = (int)(idx % nf);
  const int k = (int)((idx / nf) % g.nz);
  const int v = (int)(idx / ((long long)nf * g.nz));
  const double* r = s + g.at(v, k + kHs, i);
  double a, b;
  recon4(r[0], r[1], r[2], r[3], a, b);
  vals[idx] = a;
  d3[idx] = b;
}

__global__ void k_interp_z(const double* __restrict__ s, double* __restrict__ vals,
                           double* __restrict__ d3, Grid g) {
  const int nfz = g.nz + 1;
  const long long n = (long long)kNv * nfz * g.nx;
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n) return;
  const int i = (int)(idx % g.nx);
  const int k = (int)((idx / g.nx) % nfz);
  const int v = (int)(idx / ((long long)g.nx * nfz));
  const double* c = s + g.at(v, k, i + kHs);
  double a, b;
  recon4(c[0], c[g.ld], c[2 * g.ld], c[3 * g.ld], a, b);
  vals[idx] = a;
  d3[idx] = b;
}

__global__ void k_flux_x(const double* __restrict__ vals, const double* __restrict__ d3,
                         const double* __restrict__ hyr, const double* __restrict__ hyt,
                         double* __restrict__ flux, Grid g, double hv) {
  const int nf = g.nx + 1;
  const long long pv = (long long)g.nz * nf;          // plane of vals/d3
  const long long pf = (long long)(g.nz + 1) * nf;    // plane of flux
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= pv) return;
  const int k = (int)(idx / nf);
  double a[4], b[4], F[4];
  for (int v = 0; v < kNv; ++v) {
    a[v] = vals[v * pv + idx];
    b[v] = d3[v * pv + idx];
  }
  flux_x_face(a, b, hyr[k + kHs], hyt[k + kHs], hv, F);
  for (int v = 0; v < kNv; ++v) flux[v * pf + idx] = F[v];
}

__global__ void k_flux_z(const double* __restrict__ vals, double* __restrict__ d3,
                         const double* __restrict__ hyr, const double* __restrict__ hyt,
                         const double* __restrict__ hyp, double* __restrict__ flux,
                         Grid g, double hv) {
  const int nfz = g.nz + 1;
  const long long pv = (long long)nfz * g.nx;
  const long long pf = (long long)(g.nz + 1) * (g.nx + 1);
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= pv) return;
  const int k = (int)(idx / g.nx);
  const int i = (int)(idx % g.nx);
  const bool wall = (k == 0 || k == g.nz);
  double a[4], b[4], G[4];
  for (int v = 0; v < kNv; ++v) {
    a[v] = vals[v * pv + idx];
    b[v] = d3[v * pv + idx];
  }
  flux_z_face(a, b, hyr[k], hyt[k], hyp[k], hv, wall, G);
  if (wall) d3[idx] = 0.0;  // d3_vals_z[rho'] zeroed in place (interpolate.py:172-173)
  const long long o = (long long)k * (g.nx + 1) + i;
  for (int v = 0; v < kNv; ++v) flux[v * pf + o] = G[v];
}

__global__ void k_tend_x(const double* __restrict__ flux, double* __restrict__ tend,
                         Grid g, double dx) {
  const long long pt = (long long)g.nz * g.nx;
  const long long n = kNv * pt;
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n) return;
  const int i = (int)(idx % g.nx);
  const int k = (int)((idx / g.nx) % g.nz);
  const int v = (int)(idx / pt);
  const long long pf = (long long)(g.nz + 1) * (g.nx + 1);
  const double* f = flux + v * pf + (long long)k * (g.nx + 1) + i;
  tend[idx] = tend_of(f[0], f[1], dx, 1.0 / dx);
}

__global__ void k_tend_z(const double* __restrict__ flux, const double* __restrict__ sf,
                         double* __restrict__ tend, Grid g, double dz) {
  const long long pt = (long long)g.nz * g.nx;
  const long long n = kNv * pt;
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n) return;
  const int i = (int)(idx % g.nx);
  const int k = (int)((idx / g.nx) % g.nz);
  const int v = (int)(idx / pt);
  const long long pf = (long long)(g.nz + 1) * (g.nx + 1);
  const double* f = flux + v * pf + (long long)k * (g.nx + 1) + i;
  double t = tend_of(f[0], f[g.nx + 1], dz, 1.0 / dz);
  if (v == 2) t = t - sf[g.at(0, k + kHs, i + kHs)] * kGrav;
  tend[idx] = t;
}

__global__ void k_add_w(double* __restrict__ tend, const double* __restrict__ src,
                        long long n) {
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n) return;
  tend[2 * n + idx] = tend[2 * n + idx] + src[idx];
}

__global__ void k_rk_update(const double* init, const double* __restrict__ tend,
                            double* out, Grid g, double dt) {
  const long long pt = (long long)g.nz * g.nx;
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= kNv * pt) return;
  const int i = (int)(idx % g.nx);
  const int k = (int)((idx / g.nx) % g.nz);
  const int v = (int)(idx / pt);
  const long long o = g.at(v, k + kHs, i + kHs);
  out[o] = init[o] + dt * tend[idx];
}

// ------------------------------------------------------------ fused x ----
// MODE 0: full RK stage (writes out). MODE 1: tendency only (writes tend).
template <int MODE>
__global__ __launch_bounds__(kXThreads) void k_stage_x(
    const double* init, double* forcing, double* out, double* __restrict__ tend,
    const double* __restrict__ hyr, const double* __restrict__ hyt,
    const double* __restrict__ src_w, int bc, const int* __restrict__ inj_rows,
    Grid g, double dx, double dt, double hv) {
  __shared__ double sv[kNv][kXThreads];
  __shared__ double sf[kNv][kXThreads];
  const int t = threadIdx.x;
  const int k = blockIdx.y;
  const int i0 = blockIdx.x * kXCells;  // first interior cell of the tile
  const int nx = g.nx;
  const int prow = k + kHs;
  const int c = i0 + t;                 // padded column loaded by this thread
  if (c < nx + 4) {
    int src = c;
    bool halo = false;
    if (c < 2) {
      src = c + nx;
      halo = true;
    } else if (c >= nx + 2) {
      if (bc != PMW_BC_INJECTION) {
        src = c - nx;
        halo = true;
      }
    }
    double q[4];
#pragma unroll
    for (int v = 0; v < kNv; ++v) q[v] = forcing[g.at(v, prow, src)];
    if (halo) {
      if (bc == PMW_BC_INJECTION && inj_rows[k]) {
        const double r = q[0] + hyr[prow];
        q[1] = r * 50.0;
        q[3] = r * 298.0 - hyt[prow];
      }
#pragma unroll
      for (int v = 0; v < kNv; ++v) forcing[g.at(v, prow, c)] = q[v];
    }
#pragma unroll
    for (int v = 0; v < kNv; ++v) sv[v][t] = q[v];
  }
  __syncthreads();
  if (t <= kXCells && i0 + t <= nx) {
    double a[4], b[4], F[4];
#pragma unroll
    for (int v = 0; v < kNv; ++v)
      recon4(sv[v][t], sv[v][t + 1], sv[v][t + 2], sv[v][t + 3], a[v], b[v]);
    flux_x_face(a, b, hyr[prow], hyt[prow], hv, F);
#pragma unroll
    for (int v = 0; v < kNv; ++v) sf[v][t] = F[v];
  }
  __syncthreads();
  const int i = i0 + t;
  if (t < kXCells && i < nx) {
    double td[4];
#pragma unroll
    for (int v = 0; v < kNv; ++v) td[v] = tend_of(sf[v][t], sf[v][t + 1], dx, 1.0 / dx);
    const long long cell = (long long)k * nx + i;
    if (src_w) td[2] = td[2] + src_w[cell];
    if (MODE == 0) {
#pragma unroll
      for (int v = 0; v < kNv; ++v) {
        const long long o = g.at(v, prow, i + kHs);
        out[o] = init[o] + dt * td[v];
      }
    } else {
      const long long pt = (long long)g.nz * nx;
#pragma unroll
      for (int v = 0; v < kNv; ++v) tend[v * pt + cell] = td[v];
    }
  }
}

// ------------------------------------------------------------ fused z ----
// Loads interior row r (-2 <= r <= nz+1) of column pcol, synthesising the
// wall halo rows exactly as set_bc_z would have written them.
__device__ __forceinline__ void load_z_row(const double* f, const Grid& g,
                                           const double* __restrict__ hy, int r,
                                           int pcol, double q[4]) {
  if (r >= 0 && r < g.nz) {
#pragma unroll
    for (int v = 0; v < kNv; ++v) q[v] = f[g.at(v, r + kHs, pcol)];
    return;
  }
  const int edge = (r < 0) ? kHs : g.nz + kHs - 1;  // padded edge row
  const int prow = r + kHs;                          // padded halo row
#pragma unroll
  for (int v = 0; v < kNv; ++v)
    q[v] = wall_halo(v, f[g.at(v, edge, pcol)], hy[edge], hy[prow]);
}

// Tiles are ordered so that each XCD runs a contiguous range of them (tiles
// next to each other share cache lines in that XCD's L2).
__device__ __forceinline__ int zt_logical(int b, int per_xcd) {
  return (b % kNumXcd) * per_xcd + b / kNumXcd;
}

template <int MODE>
__global__ __launch_bounds__(kZtCols* kZtRows) void k_stage_z(
    const double* init, double* forcing, double* out, double* __restrict__ tend,
    const double* __restrict__ hy_cell, const double* __restrict__ hyr,
    const double* __restrict__ hyt, const double* __restrict__ hyp,
    const double* __restrict__ src_w, Grid g, double dz, double dt, double hv, int ncolgrp,
    int ntiles, int per_xcd) {
  constexpr int CW = kZtCols, NR = kZtRows, NT = CW * NR, C = NR - 4;
  const int nx = g.nx, nz = g.nz, t = threadIdx.x;
  const int tiles_padded = per_xcd * kNumXcd;
  if ((int)blockIdx.x >= tiles_padded) {
    // halo writer blocks: set_bc_z on the forcing buffer
    const int c = ((int)blockIdx.x - tiles_padded) * NT + t;
    if (c >= g.ld) return;
    const int lo = kHs, hi = nz + kHs - 1;
#pragma unroll
    for (int v = 0; v < kNv; ++v) {
      const double a = forcing[g.at(v, lo, c)];
      const double b = forcing[g.at(v, hi, c)];
      forcing[g.at(v, 0, c)] = wall_halo(v, a, hy_cell[lo], hy_cell[0]);
      forcing[g.at(v, 1, c)] = wall_halo(v, a, hy_cell[lo], hy_cell[1]);
      forcing[g.at(v, nz + 2, c)] = wall_halo(v, b, hy_cell[hi], hy_cell[nz + 2]);
      forcing[g.at(v, nz + 3, c)] = wall_halo(v, b, hy_cell[hi], hy_cell[nz + 3]);
    }
    return;
  }
  const int tile = zt_logical(blockIdx.x, per_xcd);
  if (tile >= ntiles) return;  // whole block, before any barrier
  __shared__ double s[kNv][NT];
  __shared__ double fl[kNv][NT];
  const int col = t % CW, j = t / CW;
  const int i = (tile % ncolgrp) * CW + col;
  const int k = (tile / ncolgrp) * C - 2 + j;  // this slot's row, face and cell
  const bool colok = i < nx;
  const int pcol = i + kHs;
  double q[4] = {0.0, 0.0, 0.0, 0.0};
  if (colok && k >= -2 && k <= nz + 1) load_z_row(forcing, g, hy_cell, k, pcol, q);
#pragma unroll
  for (int v = 0; v < kNv; ++v) s[v][t] = q[v];
  __syncthreads();
  // face k (between rows k-1 and k) from rows k-2..k+1 = tile rows j-2..j+1
  if (colok && j >= 2 && j <= NR - 2 && k >= 0 && k <= nz) {
    double a[4], b[4], G[4];
#pragma unroll
    for (int v = 0; v < kNv; ++v)
      recon4(s[v][t - 2 * CW], s[v][t - CW], s[v][t], s[v][t + CW], a[v], b[v]);
    flux_z_face(a, b, hyr[k], hyt[k], hyp[k], hv, k == 0 || k == nz, G);
#pragma unroll
    for (int v = 0; v < kNv; ++v) fl[v][t] = G[v];
  }
  __syncthreads();
  // cell k from faces k and k+1 (tile rows 2..NR-3: the C output rows)
  if (colok && j >= 2 && j <= NR - 3 && k >= 0 && k < nz) {
    double td[4];
#pragma unroll
    for (int v = 0; v < kNv; ++v) td[v] = tend_of(fl[v][t], fl[v][t + CW], dz, 1.0 / dz);
    td[2] = td[2] - s[0][t] * kGrav;
    const long long cell = (long long)k * nx + i;
    if (src_w) td[2] = td[2] + src_w[cell];
    if (MODE == 0) {
#pragma unroll
      for (int v = 0; v < kNv; ++v) {
        const long long o = g.at(v, k + kHs, pcol);
        out[o] = init[o] + dt * td[v];
      }
    } else {
      const long long pt = (long long)nz * nx;
#pragma unroll
      for (int v = 0; v < kNv; ++v) tend[v * pt + cell] = td[v];
    }
  }
}

inline unsigned blocks_for(long long n, int per) { return (unsigned)((n + per - 1) / per); }

bool grid_ok(int nx, int nz) {
  // the 4-tap stencil needs at least 2 interior cells per direction; the
  // z kernel indexes rows as int
  return nx >= 2 && nz >= 2 && (long long)(nx + 4) * (nz + 4) * 4 < (1LL << 40);
}

int launch_stage_x(int mode, const double* init, double* forcing, double* out,
                   double* tend, const double* hyr, const double* hyt,
                   const double* src_w, int bc, const int* inj_rows, Grid g,
                   double dx, double dt, double hv, hipStream_t st) {
  dim3 grid(blocks_for(g.nx, kXCells), g.nz);
  if (mode == 0)
    k_stage_x<0><<<grid, kXThreads, 0, st>>>(init, forcing, out, tend, hyr, hyt,
                                            src_w, bc, inj_rows, g, dx, dt, hv);
  else
    k_stage_x<1><<<grid, kXThreads, 0, st>>>(init, forcing, out, tend, hyr, hyt,
                                            src_w, bc, inj_rows, g, dx, dt, hv);
  PMW_CHECK_LAUNCH();
  return PMW_OK;
}

int launch_stage_z(int mode, const double* init, double* forcing, double* out,
                   double* tend, const double* hy_cell, const double* hyr,
                   const double* hyt, const double* hyp, const double* src_w,
                   Grid g, double dz, double dt, double hv, hipStream_t st) {
  constexpr int NT = kZtCols * kZtRows;
  const int ncolgrp = (g.nx + kZtCols - 1) / kZtCols;
  const int ntiles = ncolgrp * ((g.nz + kZtRows - 5) / (kZtRows - 4));
  const int per = (ntiles + kNumXcd - 1) / kNumXcd;
  const unsigned nblk = (unsigned)per * kNumXcd + blocks_for(g.ld, NT);  // + halo writers
  if (mode == 0)
    k_stage_z<0><<<nblk, NT, 0, st>>>(init, forcing, out, tend, hy_cell, hyr, hyt, hyp,
                                      src_w, g, dz, dt, hv, ncolgrp, ntiles, per);
  else
    k_stage_z<1><<<nblk, NT, 0, st>>>(init, forcing, out, tend, hy_cell, hyr, hyt, hyp,
                                      src_w, g, dz, dt, hv, ncolgrp, ntiles, per);
  PMW_CHECK_LAUNCH();
  return PMW_OK;
}

}  // namespace

// =============================================================== C ABI ====
extern "C" {

int pmw_abi_version(void) { return 100; }

const char* pmw_error_string(int code) {
  switch (code) {
    case PMW_OK: return "ok";
    case PMW_EINVAL: return "invalid argument (size, null pointer or enum)";
    case PMW_EALIAS: return "output buffer aliases an input it must not";
    case PMW_ESPACE: return "workspace too small";
    default: return code > 0 ? hipGetErrorString((hipError_t)code) : "unknown error";
  }
}

int pmw_set_bc_x(double* state, int nx, int nz, int bc, const int* inj_rows,
                 const double* hy_dens_cell, const double* hy_dens_theta_cell,
                 void* stream) {
  if (!grid_ok(nx, nz) || bad_ptr(state)) return PMW_EINVAL;
  if (bc != PMW_BC_PERIODIC && bc != PMW_BC_INJECTION) return PMW_EINVAL;
  if (bc == PMW_BC_INJECTION && (!inj_rows || !hy_dens_cell || !hy_dens_theta_cell))
    return PMW_EINVAL;
  Grid g(nx, nz);
  k_halo_x<<<blocks_for(4LL * nz, 256), 256, 0, (hipStream_t)stream>>>(
      state, g, bc, inj_rows, hy_dens_cell, hy_dens_theta_cell);
  PMW_CHECK_LAUNCH();
  return PMW_OK;
}

int pmw_set_bc_z(double* state, const double* hy_dens_cell, int nx, int nz,
                 void* stream) {
  if (!grid_ok(nx, nz) || bad_ptr(state) || bad_ptr(hy_dens_cell)) return PMW_EINVAL;
  Grid g(nx, nz);
  k_halo_z<<<blocks_for(nx + 4, 256), 256, 0, (hipStream_t)stream>>>(state, g,
                                                                   hy_dens_cell);
  PMW_CHECK_LAUNCH();
  return PMW_OK;
}

int pmw_interpolate_x(const double* state, double* vals_x, double* d3_vals_x,
                      int nx, int nz, void* stream) {
  if (!grid_ok(nx, nz) || bad_ptr(state) || bad_ptr(vals_x) || bad_ptr(d3_vals_x))
    return PMW_EINVAL;
  Grid g(nx, nz);
  k_interp_x<<<blocks_for(4LL * nz * (nx + 1), 256), 256, 0, (hipStream_t)stream>>>(
      state, vals_x, d3_vals_x, g);
  PMW_CHECK_LAUNCH();
  return PMW_OK;
}

int pmw_interpolate_z(const double* state, double* vals_z, double* d3_vals_z,
                      int nx, int nz, void* stream) {
  if (!grid_ok(nx, nz) || bad_ptr(state) || bad_ptr(vals_z) || bad_ptr(d3_vals_z))
    return PMW_EINVAL;
  Grid g(nx, nz);
  k_interp_z<<<blocks_for(4LL * (nz + 1) * nx, 256), 256, 0, (hipStream_t)stream>>>(
      state, vals_z, d3_vals_z, g);
  PMW_CHECK_LAUNCH();
  return PMW_OK;
}

int pmw_compute_flux_x(const double* vals_x, const double* d3_vals_x,
                       const double* hy_dens_cell,
                       const double* hy_dens_theta_cell, double* flux, int nx,
                       int nz, double hv_coeff, void* stream) {
  if (!grid_ok(nx, nz) || bad_ptr(vals_x) || bad_ptr(d3_vals_x) ||
      bad_ptr(hy_dens_cell) || bad_ptr(hy_dens_theta_cell) || bad_ptr(flux))
    return PMW_EINVAL;
  Grid g(nx, nz);
  k_flux_x<<<blocks_for((long long)nz * (nx + 1), 256), 256, 0, (hipStream_t)stream>>>(
      vals_x, d3_vals_x, hy_dens_cell, hy_dens_theta_cell, flux, g, hv_coeff);
  PMW_CHECK_LAUNCH();
  return PMW_OK;
}

int pmw_compute_flux_z(const double* vals_z, double* d3_vals_z,
                       const double* hy_dens_int,
                       const double* hy_dens_theta_int,
                       const double* hy_pressure_int, double* flux, int nx,
                       int nz, double hv_coeff, void* stream) {
  if (!grid_ok(nx, nz) || bad_ptr(vals_z) || bad_ptr(d3_vals_z) ||
      bad_ptr(hy_dens_int) || bad_ptr(hy_dens_theta_int) ||
      bad_ptr(hy_pressure_int) || bad_ptr(flux))
    return PMW_EINVAL;
  Grid g(nx, nz);
  k_flux_z<<<blocks_for((long long)(nz + 1) * nx, 256), 256, 0, (hipStream_t)stream>>>(
      vals_z, d3_vals_z, hy_dens_int, hy_dens_theta_int, hy_pressure_int, flux, g,
      hv_coeff);
  PMW_CHECK_LAUNCH();
  return PMW_OK;
}

int pmw_compute_tend_x(const double* flux, double* tend, int nx, int nz,
                       double dx, void* stream) {
  if (!grid_ok(nx, nz) || bad_ptr(flux) || bad_ptr(tend)) return PMW_EINVAL;
  Grid g(nx, nz);
  k_tend_x<<<blocks_for(4LL * nz * nx, 256), 256, 0, (hipStream_t)stream>>>(flux, tend,
                                                                          g, dx);
  PMW_CHECK_LAUNCH();
  return PMW_OK;
}

int pmw_compute_tend_z(const double* flux, const double* state_forcing,
                       double* tend, int nx, int nz, double dz, void* stream) {
  if (!grid_ok(nx, nz) || bad_ptr(flux) || bad_ptr(state_forcing) || bad_ptr(tend))
    return PMW_EINVAL;
  Grid g(nx, nz);
  k_tend_z<<<blocks_for(4LL * nz * nx, 256), 256, 0, (hipStream_t)stream>>>(
      flux, state_forcing, tend, g, dz);
  PMW_CHECK_LAUNCH();
  return PMW_OK;
}

int pmw_add_source_w(double* tend, const double* src_w, int nx, int nz,
                     void* stream) {
  if (!grid_ok(nx, nz) || bad_ptr(tend) || bad_ptr(src_w)) return PMW_EINVAL;
  const long long n = (long long)nz * nx;
  k_add_w<<<blocks_for(n, 256), 256, 0, (hipStream_t)stream>>>(tend, src_w, n);
  PMW_CHECK_LAUNCH();
  return PMW_OK;
}

int pmw_rk_update(const double* state_init, const double* tend,
                  double* state_out, int nx, int nz, double dt_stage,
                  void* stream) {
  if (!grid_ok(nx, nz) || bad_ptr(state_init) || bad_ptr(tend) || bad_ptr(state_out))
    return PMW_EINVAL;
  Grid g(nx, nz);
  k_rk_update<<<blocks_for(4LL * nz * nx, 256), 256, 0, (hipStream_t)stream>>>(
      state_init, tend, state_out, g, dt_stage);
  PMW_CHECK_LAUNCH();
  return PMW_OK;
}

int pmw_tendencies_x(double* state_forcing, const double* hy_dens_cell,
                     const double* hy_dens_theta_cell, const double* src_w,
                     double* tend, int nx, int nz, double dx, double hv_coeff,
                     void* stream) {
  if (!grid_ok(nx, nz) || bad_ptr(state_forcing) || bad_ptr(hy_dens_cell) ||
      bad_ptr(hy_dens_theta_cell) || bad_ptr(tend))
    return PMW_EINVAL;
  return launch_stage_x(1, nullptr, state_forcing, nullptr, tend, hy_dens_cell,
                        hy_dens_theta_cell, src_w, PMW_BC_PERIODIC, nullptr,
                        Grid(nx, nz), dx, 0.0, hv_coeff, (hipStream_t)stream);
}

int pmw_tendencies_z(double* state_forcing, const double* hy_dens_cell,
                     const double* hy_dens_int,
                     const double* hy_dens_theta_int,
                     const double* hy_pressure_int, const double* src_w,
                     double* tend, int nx, int nz, double dz, double hv_coeff,
                     void* stream) {
  if (!grid_ok(nx, nz) || bad_ptr(state_forcing) || bad_ptr(hy_dens_cell) ||
      bad_ptr(hy_dens_int) || bad_ptr(hy_dens_theta_int) ||
      bad_ptr(hy_pressure_int) || bad_ptr(tend))
    return PMW_EINVAL;
  return launch_stage_z(1, nullptr, state_forcing, nullptr, tend, hy_dens_cell,
                        hy_dens_int, hy_dens_theta_int, hy_pressure_int, src_w,
                        Grid(nx, nz), dz, 0.0, hv_coeff, (hipStream_t)stream);
}

int pmw_rk_stage(int direction, const double* state_init,
                 double* state_forcing, double* state_out,
                 const double* hy_dens_cell, const double* hy_dens_theta_cell,
                 const double* hy_dens_int, const double* hy_dens_theta_int,
                 const double* hy_pressure_int, const double* src_w, int bc,
                 const int* inj_rows, int nx, int nz, double d,
                 double dt_stage, double hv_coeff, void* stream) {
  if (!grid_ok(nx, nz) || bad_ptr(state_init) || bad_ptr(state_forcing) ||
      bad_ptr(state_out) || bad_ptr(hy_dens_cell))
    return PMW_EINVAL;
  if (state_out == state_forcing) return PMW_EALIAS;
  Grid g(nx, nz);
  if (direction == 0) {
    if (bad_ptr(hy_dens_theta_cell)) return PMW_EINVAL;
    if (bc != PMW_BC_PERIODIC && bc != PMW_BC_INJECTION) return PMW_EINVAL;
    if (bc == PMW_BC_INJECTION && !inj_rows) return PMW_EINVAL;
    return launch_stage_x(0, state_init, state_forcing, state_out, nullptr,
                          hy_dens_cell, hy_dens_theta_cell, src_w, bc, inj_rows, g,
                          d, dt_stage, hv_coeff, (hipStream_t)stream);
  }
  if (direction == 1) {
    if (bad_ptr(hy_dens_int) || bad_ptr(hy_dens_theta_int) || bad_ptr(hy_pressure_int))
      return PMW_EINVAL;
    return launch_stage_z(0, state_init, state_forcing, state_out, nullptr,
                          hy_dens_cell, hy_dens_int, hy_dens_theta_int,
                          hy_pressure_int, src_w, g, d, dt_stage, hv_coeff,
                          (hipStream_t)stream);
  }
  return PMW_EINVAL;
}

namespace {

struct Tuning {
  int mode;   // 0: one fused launch per direction; 1: one launch per RK stage
  int x_nt;   // fused x tiles: 256, 448, 512 threads per row tile; 1256, 1512 dense
  int z_seg;  // fused z: rows per marching segment (8..64) or LDS tile code (8064, 4128, 8096, 8128)
};
constexpr Tuning kDefaultTuning = {0, 512, 8096};

int tuning_from(const int* t, Tuning& out) {
  out = kDefaultTuning;
  if (!t) return PMW_OK;
  out.mode = t[0];
  out.x_nt = t[1];
  out.z_seg = t[2];
  if (out.mode != 0 && out.mode != 1) return PMW_EINVAL;
  if (out.mode == 0 && (!fused_x_tile_ok(out.x_nt) || !fused_z_seg_ok(out.z_seg)))
    return PMW_EINVAL;
  return PMW_OK;
}

// Shared body of the evolve entry points. With `ev` non-null, event ev[2j]
// is recorded before and ev[2j+1] after launch j, dirs[j] = its direction.
int evolve_impl(double* state, double* state_tmp, double* state_work,
                const double* hy_dens_cell, const double* hy_dens_theta_cell,
                const double* hy_dens_int, const double* hy_dens_theta_int,
                const double* hy_pressure_int, const double* src_w, int bc,
                const int* inj_rows, int nx, int nz, double dx, double dz,
                double dt, double hv_coeff_x, double hv_coeff_z,
                int* reverse_direction, int nsteps, hipStream_t st,
                const Tuning& tune, hipEvent_t* ev, int* dirs, int* nlaunch) {
  if (!grid_ok(nx, nz) || bad_ptr(state) || bad_ptr(state_tmp) ||
      bad_ptr(state_work) || bad_ptr(hy_dens_cell) || bad_ptr(hy_dens_theta_cell) ||
      bad_ptr(hy_dens_int) || bad_ptr(hy_dens_theta_int) ||
      bad_ptr(hy_pressure_int) || bad_ptr(reverse_direction) || nsteps < 0)
    return PMW_EINVAL;
  if (bc != PMW_BC_PERIODIC && bc != PMW_BC_INJECTION) return PMW_EINVAL;
  if (bc == PMW_BC_INJECTION && !inj_rows) return PMW_EINVAL;
  if (state == state_tmp || state == state_work || state_tmp == state_work)
    return PMW_EALIAS;
  Grid g(nx, nz);
  // stage dt as the reference computes them: dt/3, dt/2, dt/1 (step.py:111-141)
  const double dts[3] = {dt / 3, dt / 2, dt / 1};
  // the injection inflow is only implemented by the per-stage kernels
  const bool fused = tune.mode == 0 && bc == PMW_BC_PERIODIC;
  int launch = 0;
  auto rec = [&](int k, int dir) -> int {
    if (!ev) return PMW_OK;
    if (k == 0) dirs[launch] = dir;
    hipError_t e = hipEventRecord(ev[2 * launch + k], st);
    return e == hipSuccess ? PMW_OK : (int)e;
  };
  for (int step = 0; step < nsteps; ++step) {
    const int first = *reverse_direction ? 0 : 1;  // 0 = x, 1 = z
    for (int pass = 0; pass < 2; ++pass) {
      const int dir = pass == 0 ? first : 1 - first;
      if (fused) {
        // pass 0: S -> W, pass 1: W -> S; halos of stage 1 always into S
        DirArgs a;
        a.in = gptr(pass == 0 ? state : state_work);
        a.out = gptr(pass == 0 ? state_work : state);
        a.st = gptr(state);
        a.tmp = gptr(state_tmp);
        // state_tmp's interior is rewritten by every direction, and no launch
        // reads it: only the last direction of the call keeps its stores
        a.tmp_interior = pass == 1 && step == nsteps - 1;
        a.hy_cell = gptr(hy_dens_cell);
        a.hyt_cell = gptr(hy_dens_theta_cell);
        a.hyr_int = gptr(hy_dens_int);
        a.hyt_int = gptr(hy_dens_theta_int);
        a.hyp_int = gptr(hy_pressure_int);
        a.src_w = gptr(src_w);
        a.d = dir == 0 ? dx : dz;
        a.id = 1.0 / a.d;
        a.dt1 = dts[0];
        a.dt2 = dts[1];
        a.dt3 = dts[2];
        a.hv = dir == 0 ? hv_coeff_x : hv_coeff_z;
        int rc = rec(0, dir);
        if (rc) return rc;
        rc = dir == 0 ? launch_dir_x(a, g, tune.x_nt, st) : launch_dir_z(a, g, tune.z_seg, st);
        if (rc) return rc;
        if ((rc = rec(1, dir))) return rc;
        ++launch;
        continue;
      }
      // stage 1: S -> W; stage 2: W -> T; stage 3: T -> S (init = S always)
      double* forc[3] = {state, state_work, state_tmp};
      double* outs[3] = {state_work, state_tmp, state};
      for (int s = 0; s < 3; ++s) {
        if (s == 1 && dir == 0 && bc == PMW_BC_INJECTION) {
          const int n = kNv * nz * 2;
          k_copy_right_halo<<<(n + 255) / 256, 256, 0, st>>>(state_tmp, state_work, g);
          PMW_CHECK_LAUNCH();
        }
        int rc = rec(0, dir);
        if (rc) return rc;
        if (dir == 0)
          rc = launch_stage_x(0, state, forc[s], outs[s], nullptr, hy_dens_cell,
                              hy_dens_theta_cell, src_w, bc, inj_rows, g, dx, dts[s],
                              hv_coeff_x, st);
        else
          rc = launch_stage_z(0, state, forc[s], outs[s], nullptr, hy_dens_cell,
                              hy_dens_int, hy_dens_theta_int, hy_pressure_int, src_w,
                              g, dz, dts[s], hv_coeff_z, st);
        if (rc) return rc;
        if ((rc = rec(1, dir))) return rc;
        ++launch;
      }
    }
    *reverse_direction = !*reverse_direction;
  }
  if (nlaunch) *nlaunch = launch;
  return PMW_OK;
}

}  // namespace

int pmw_evolve(double* state, double* state_tmp, double* state_work,
               const double* hy_dens_cell, const double* hy_dens_theta_cell,
               const double* hy_dens_int, const double* hy_dens_theta_int,
               const double* hy_pressure_int, const double* src_w, int bc,
               const int* inj_rows, int nx, int nz, double dx, double dz,
               double dt, double hv_coeff_x, double hv_coeff_z,
               int* reverse_direction, int nsteps, void* stream) {
  return evolve_impl(state, state_tmp, state_work, hy_dens_cell, hy_dens_theta_cell,
                     hy_dens_int, hy_dens_theta_int, hy_pressure_int, src_w, bc,
                     inj_rows, nx, nz, dx, dz, dt, hv_coeff_x, hv_coeff_z,
                     reverse_direction, nsteps, (hipStream_t)stream, kDefaultTuning,
                     nullptr, nullptr, nullptr);
}

int pmw_evolve_ex(double* state, double* state_tmp, double* state_work,
                  const double* hy_dens_cell, const double* hy_dens_theta_cell,
                  const double* hy_dens_int, const double* hy_dens_theta_int,
                  const double* hy_pressure_int, const double* src_w, int bc,
                  const int* inj_rows, int nx, int nz, double dx, double dz,
                  double dt, double hv_coeff_x, double hv_coeff_z,
                  int* reverse_direction, int nsteps, const int* tuning,
                  void* stream, double* kernel_ms) {
  Tuning tune;
  int rc = tuning_from(tuning, tune);
  if (rc) return rc;
  if (!kernel_ms)
    return evolve_impl(state, state_tmp, state_work, hy_dens_cell, hy_dens_theta_cell,
                       hy_dens_int, hy_dens_theta_int, hy_pressure_int, src_w, bc,
                       inj_rows, nx, nz, dx, dz, dt, hv_coeff_x, hv_coeff_z,
                       reverse_direction, nsteps, (hipStream_t)stream, tune, nullptr,
                       nullptr, nullptr);
  if (nsteps < 0 || nsteps > 100000) return PMW_EINVAL;
  const int nl = 6 * nsteps;  // upper bound on launches
  hipEvent_t* ev = new hipEvent_t[2 * (size_t)nl + 2];
  int* dirs = new int[(size_t)nl + 1];
  int made = 0, used = 0;
  rc = PMW_OK;
  for (; made < 2 * nl; ++made) {
    hipError_t e = hipEventCreate(&ev[made]);
    if (e != hipSuccess) { rc = (int)e; break; }
  }
  if (rc == PMW_OK)
    rc = evolve_impl(state, state_tmp, state_work, hy_dens_cell, hy_dens_theta_cell,
                     hy_dens_int, hy_dens_theta_int, hy_pressure_int, src_w, bc,
                     inj_rows, nx, nz, dx, dz, dt, hv_coeff_x, hv_coeff_z,
                     reverse_direction, nsteps, (hipStream_t)stream, tune, ev, dirs,
                     &used);
  kernel_ms[0] = kernel_ms[1] = 0.0;
  if (rc == PMW_OK && used > 0) {
    hipError_t e = hipEventSynchronize(ev[2 * used - 1]);
    if (e != hipSuccess) rc = (int)e;
    for (int j = 0; rc == PMW_OK && j < used; ++j) {
      float ms = 0.f;
      e = hipEventElapsedTime(&ms, ev[2 * j], ev[2 * j + 1]);
      if (e != hipSuccess) { rc = (int)e; break; }
      kernel_ms[dirs[j]] += ms;
    }
  }
  for (int j = 0; j < made; ++j) (void)hipEventDestroy(ev[j]);  // cleanup only
  delete[] ev;
  delete[] dirs;
  return rc;
}

}  // extern "C"
