// pmw_adjoint.hip -- hand-written adjoint (reverse mode) of the time step.
//
// Forward, per direction d and RK stage s (solve/step.py:21-82, 111-141):
//   out = init + a_s * T_d(f),  a_s in {dt/3, dt/2, dt}
// where T_d is halo -> 4-tap reconstruction -> flux -> flux difference
// (+ gravity in z). Its adjoint, for a cotangent lam of `out`:
//   lam_init += lam,   lam_f += a_s * T_d'(f)^T lam.
// T_d'(f)^T lam is computed in two passes:
//  * face pass: for every face, lamF = (lamT(cell right) - lamT(cell left))
//    / d, then the transposed 4x4 flux Jacobian at the face (from the
//    forward face values of f) gives the cotangents of the reconstructed
//    values and of the 3rd differences (8 numbers per face, to workspace);
//  * cell pass: every interior cell gathers the 4-tap transposed
//    reconstruction from the faces that read it, wrapping periodically in x
//    and folding the wall halo rows of z onto the edge rows with the factors
//    set_bc_z uses (1 for rho' and (rho*theta)', hy[halo]/hy[edge] for
//    rho*u, 0 for rho*w); z adds the gravity term -g * lam_w to lam_rho'.
// One RK3 direction: lamS2 = dt T'(S2)^T lam3; lamS1 = dt/2 T'(S1)^T lamS2;
// lamS = dt/3 T'(S)^T lamS1 + lam3 + lamS2 + lamS1 (the init terms).
// Cotangent buffers use the state layout (4, nz+4, nx+4); only the interior
// is read or written. Derivatives are those of the real-number formulas the
// reference evaluates (e.g. rho * (v/rho) is v).
#include "pmw_common.h"

using namespace pmw;

namespace {

constexpr double kW4[4] = {kW4a, kW4b, kW4b, kW4a};
constexpr double kD3[4] = {-1.0, 3.0, -3.0, 1.0};

__device__ __forceinline__ int wrapc(int c, int n) {
  int w = c % n;
  return w < 0 ? w + n : w;
}

// Transposed x-flux Jacobian: cotangents of (vals, d3) from lamF.
__device__ __forceinline__ void adj_flux_x(const double val[4], double hyr, double hyt,
                                           double hv, const double lf[4], double lv[4],
                                           double ld[4]) {
  const double rho = val[0] + hyr;
  const double u = val[1] / rho;
  const double w = val[2] / rho;
  const double t = (val[3] + hyt) / rho;
  const double q = rho * t;
  const double dp = kGamma * (kC0 * pow(q, kGamma)) / q;  // dp/dq
  lv[0] = -(u * u) * lf[1] - (u * w) * lf[2] - (u * t) * lf[3];
  lv[1] = lf[0] + 2.0 * u * lf[1] + w * lf[2] + t * lf[3];
  lv[2] = u * lf[2];
  lv[3] = dp * lf[1] + u * lf[3];
#pragma unroll
  for (int v = 0; v < kNv; ++v) ld[v] = -hv * lf[v];
}

// Transposed z-flux Jacobian; at the wall faces w and d3[rho'] are cut.
__device__ __forceinline__ void adj_flux_z(const double val[4], double hyr, double hyt,
                                           double hv, bool wall, const double lf[4],
                                           double lv[4], double ld[4]) {
  const double rho = val[0] + hyr;
  const double u = val[1] / rho;
  const double w = val[2] / rho;
  const double t = (val[3] + hyt) / rho;
  const double q = rho * t;
  const double dp = kGamma * (kC0 * pow(q, kGamma)) / q;
  if (wall) {
    lv[0] = lv[1] = lv[2] = 0.0;
    lv[3] = dp * lf[2];
    ld[0] = 0.0;
  } else {
    lv[0] = -(u * w) * lf[1] - (w * w) * lf[2] - (w * t) * lf[3];
    lv[1] = w * lf[1];
    lv[2] = lf[0] + u * lf[1] + 2.0 * w * lf[2] + t * lf[3];
    lv[3] = dp * lf[2] + w * lf[3];
    ld[0] = -hv * lf[0];
  }
  ld[1] = -hv * lf[1];
  ld[2] = -hv * lf[2];
  ld[3] = -hv * lf[3];
}

// ---- face passes: 8 cotangents per face into fw[8][faces] ---------------
__global__ void k_adj_face_x(const double* __restrict__ f, const double* __restrict__ lam,
                             double scale, const double* __restrict__ hyr,
                             const double* __restrict__ hyt, double* __restrict__ fw,
                             Grid g, double dx, double hv) {
  const int nf = g.nx + 1;
  const long long nfaces = (long long)g.nz * nf;
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= nfaces) return;
  const int k = (int)(idx / nf), face = (int)(idx % nf);
  const int prow = k + kHs;
  double val[4], d3[4];
#pragma unroll
  for (int v = 0; v < kNv; ++v) {
    double s[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) s[j] = f[g.at(v, prow, wrapc(face - 2 + j, g.nx) + kHs)];
    recon4(s[0], s[1], s[2], s[3], val[v], d3[v]);
  }
  // T(i) = (F(i) - F(i+1)) / dx: face `face` is F(i) of cell face and F(i+1)
  // of cell face-1
  double lf[4];
#pragma unroll
  for (int v = 0; v < kNv; ++v) {
    const double right = face < g.nx ? lam[g.at(v, prow, face + kHs)] : 0.0;
    const double left = face > 0 ? lam[g.at(v, prow, face - 1 + kHs)] : 0.0;
    lf[v] = scale * (right - left) / dx;
  }
  double lv[4], ld[4];
  adj_flux_x(val, hyr[prow], hyt[prow], hv, lf, lv, ld);
#pragma unroll
  for (int v = 0; v < kNv; ++v) {
    fw[v * nfaces + idx] = lv[v];
    fw[(4 + v) * nfaces + idx] = ld[v];
  }
}

// wall halo value of padded row ph for variable v given the edge row value
__device__ __forceinline__ double zrow_value(const double* f, const Grid& g,
                                             const double* hy, int v, int r, int pc) {
  if (r >= 0 && r < g.nz) return f[g.at(v, r + kHs, pc)];
  const int pe = r < 0 ? kHs : g.nz + kHs - 1;
  return wall_halo(v, f[g.at(v, pe, pc)], hy[pe], hy[r + kHs]);
}

__global__ void k_adj_face_z(const double* __restrict__ f, const double* __restrict__ lam,
                             double scale, const double* __restrict__ hy_cell,
                             const double* __restrict__ hyr, const double* __restrict__ hyt,
                             double* __restrict__ fw, Grid g, double dz, double hv) {
  const int nx = g.nx;
  const long long nfaces = (long long)(g.nz + 1) * nx;
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= nfaces) return;
  const int face = (int)(idx / nx), i = (int)(idx % nx);
  const int pc = i + kHs;
  double val[4], d3[4];
#pragma unroll
  for (int v = 0; v < kNv; ++v) {
    double s[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) s[j] = zrow_value(f, g, hy_cell, v, face - 2 + j, pc);
    recon4(s[0], s[1], s[2], s[3], val[v], d3[v]);
  }
  double lf[4];
#pragma unroll
  for (int v = 0; v < kNv; ++v) {
    const double above = face < g.nz ? lam[g.at(v, face + kHs, pc)] : 0.0;
    const double below = face > 0 ? lam[g.at(v, face - 1 + kHs, pc)] : 0.0;
    lf[v] = scale * (above - below) / dz;
  }
  double lv[4], ld[4];
  adj_flux_z(val, hyr[face], hyt[face], hv, face == 0 || face == g.nz, lf, lv, ld);
#pragma unroll
  for (int v = 0; v < kNv; ++v) {
    fw[v * nfaces + idx] = lv[v];
    fw[(4 + v) * nfaces + idx] = ld[v];
  }
}

// ---- cell passes ---------------------------------------------------------
// lam_out = T'^T part + acc1 + acc2 + acc3 (accs nullable; lam_out may alias
// any acc, which is read only at the same cell).
__global__ void k_adj_cell_x(const double* __restrict__ fw, const double* acc1,
                             const double* acc2, const double* acc3, double* lam_out,
                             Grid g) {
  const int nx = g.nx, nf = nx + 1;
  const long long nfaces = (long long)g.nz * nf;
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long long)g.nz * nx) return;
  const int k = (int)(idx / nx), c = (int)(idx % nx);
  double l[4] = {0.0, 0.0, 0.0, 0.0};
  // every copy of cell c among padded columns -2..nx+1 (periodic halo)
  for (int m = -1; m <= 1; ++m) {
    const int cp = c + m * nx;
    if (cp < -2 || cp > nx + 1) continue;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int face = cp + 2 - j;
      if (face < 0 || face > nx) continue;
      const long long fi = (long long)k * nf + face;
#pragma unroll
      for (int v = 0; v < kNv; ++v)
        l[v] += kW4[j] * fw[v * nfaces + fi] + kD3[j] * fw[(4 + v) * nfaces + fi];
    }
  }
#pragma unroll
  for (int v = 0; v < kNv; ++v) {
    const long long o = g.at(v, k + kHs, c + kHs);
    double r = l[v];
    if (acc1) r += acc1[o];
    if (acc2) r += acc2[o];
    if (acc3) r += acc3[o];
    lam_out[o] = r;
  }
}

__device__ __forceinline__ void gather_z_row(const double* fw, long long nfaces, int nx,
                                             int nz, int r, int i, double l[4]) {
  // cotangent of forcing row r (-2 <= r <= nz+1) of column i
#pragma unroll
  for (int v = 0; v < kNv; ++v) l[v] = 0.0;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int face = r + 2 - j;
    if (face < 0 || face > nz) continue;
    const long long fi = (long long)face * nx + i;
#pragma unroll
    for (int v = 0; v < kNv; ++v)
      l[v] += kW4[j] * fw[v * nfaces + fi] + kD3[j] * fw[(4 + v) * nfaces + fi];
  }
}

__global__ void k_adj_cell_z(const double* __restrict__ fw, const double* __restrict__ lam,
                             double scale, const double* __restrict__ hy, const double* acc1,
                             const double* acc2, const double* acc3, double* lam_out,
                             Grid g) {
  const int nx = g.nx, nz = g.nz;
  const long long nfaces = (long long)(nz + 1) * nx;
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long long)nz * nx) return;
  const int k = (int)(idx / nx), i = (int)(idx % nx);
  double l[4], h[4];
  gather_z_row(fw, nfaces, nx, nz, k, i, l);
  // wall halo rows are functions of the edge row (set_bc_z): fold them back
  if (k == 0 || k == nz - 1) {
    const int pe = k + kHs;
    const int r0 = k == 0 ? -2 : nz;
    for (int r = r0; r < r0 + 2; ++r) {
      gather_z_row(fw, nfaces, nx, nz, r, i, h);
      const int ph = r + kHs;
      l[0] += h[0];
      l[1] += h[1] * hy[ph] / hy[pe];
      l[3] += h[3];
    }
  }
  // gravity: T_w -= rho'_forcing * g
  l[0] += -kGrav * (scale * lam[g.at(2, k + kHs, i + kHs)]);
#pragma unroll
  for (int v = 0; v < kNv; ++v) {
    const long long o = g.at(v, k + kHs, i + kHs);
    double r = l[v];
    if (acc1) r += acc1[o];
    if (acc2) r += acc2[o];
    if (acc3) r += acc3[o];
    lam_out[o] = r;
  }
}

inline unsigned nblk(long long n) { return (unsigned)((n + 255) / 256); }

int adj_stage(int dir, const double* f, const double* lam, double scale, const double* acc1,
              const double* acc2, const double* acc3, double* lam_out, double* fw,
              const double* hy_cell, const double* hyt_cell, const double* hyr_int,
              const double* hyt_int, Grid g, double d, double hv, hipStream_t st) {
  if (dir == 0) {
    const long long nfaces = (long long)g.nz * (g.nx + 1);
    k_adj_face_x<<<nblk(nfaces), 256, 0, st>>>(f, lam, scale, hy_cell, hyt_cell, fw, g, d, hv);
    PMW_CHECK_LAUNCH();
    k_adj_cell_x<<<nblk((long long)g.nz * g.nx), 256, 0, st>>>(fw, acc1, acc2, acc3,
                                                              lam_out, g);
  } else {
    const long long nfaces = (long long)(g.nz + 1) * g.nx;
    k_adj_face_z<<<nblk(nfaces), 256, 0, st>>>(f, lam, scale, hy_cell, hyr_int, hyt_int, fw,
                                             g, d, hv);
    PMW_CHECK_LAUNCH();
    k_adj_cell_z<<<nblk((long long)g.nz * g.nx), 256, 0, st>>>(fw, lam, scale, hy_cell, acc1,
                                                              acc2, acc3, lam_out, g);
  }
  PMW_CHECK_LAUNCH();
  return PMW_OK;
}

bool grid_ok(int nx, int nz) { return nx >= 2 && nz >= 2; }

}  // namespace

// Forward stage kernels of pmw_step.hip (one RK stage, fused).
extern "C" int pmw_rk_stage(int direction, const double* state_init, double* state_forcing,
                            double* state_out, const double* hy_dens_cell,
                            const double* hy_dens_theta_cell, const double* hy_dens_int,
                            const double* hy_dens_theta_int, const double* hy_pressure_int,
                            const double* src_w, int bc, const int* inj_rows, int nx, int nz,
                            double d, double dt_stage, double hv_coeff, void* stream);

extern "C" {

long long pmw_adjoint_workspace_size(int nx, int nz) {
  if (!grid_ok(nx, nz)) return 0;
  const long long state = 4LL * (nz + 4) * (nx + 4);
  const long long faces = 8LL * (nz + 1) * (nx + 1);
  return 5 * state + faces;
}

int pmw_adjoint_stage(int direction, const double* state_forcing, const double* lam_in,
                      double scale, const double* acc1, const double* acc2,
                      const double* acc3, double* lam_out, double* face_workspace,
                      const double* hy_dens_cell, const double* hy_dens_theta_cell,
                      const double* hy_dens_int, const double* hy_dens_theta_int, int nx,
                      int nz, double d, double hv_coeff, void* stream) {
  if (!grid_ok(nx, nz) || !state_forcing || !lam_in || !lam_out || !face_workspace ||
      !hy_dens_cell || (direction == 0 && !hy_dens_theta_cell) ||
      (direction == 1 && (!hy_dens_int || !hy_dens_theta_int)))
    return PMW_EINVAL;
  if (direction != 0 && direction != 1) return PMW_EINVAL;
  if (lam_out == lam_in) return PMW_EALIAS;
  return adj_stage(direction, state_forcing, lam_in, scale, acc1, acc2, acc3, lam_out,
                   face_workspace, hy_dens_cell, hy_dens_theta_cell, hy_dens_int,
                   hy_dens_theta_int, Grid(nx, nz), d, hv_coeff, (hipStream_t)stream);
}

int pmw_adjoint_step(double* state, double* lam, double* workspace,
                     long long workspace_len, const double* hy_dens_cell,
                     const double* hy_dens_theta_cell, const double* hy_dens_int,
                     const double* hy_dens_theta_int, const double* hy_pressure_int,
                     const double* src_w, int nx, int nz, double dx, double dz, double dt,
                     double hv_coeff_x, double hv_coeff_z, int reverse_direction,
                     void* stream) {
  if (!grid_ok(nx, nz) || !state || !lam || !workspace || !hy_dens_cell ||
      !hy_dens_theta_cell || !hy_dens_int || !hy_dens_theta_int || !hy_pressure_int)
    return PMW_EINVAL;
  if (workspace_len < pmw_adjoint_workspace_size(nx, nz)) return PMW_ESPACE;
  const long long ns = 4LL * (nz + 4) * (nx + 4);
  double* mid = workspace;        // state after the first direction
  double* s1 = workspace + ns;    // stage-1 state of the direction being reversed
  double* s2 = workspace + 2 * ns;
  double* l1 = workspace + 3 * ns;
  double* l2 = workspace + 4 * ns;
  double* fw = workspace + 5 * ns;
  double* s0 = state;  // its halos are rewritten (set_bc of the forward stages)
  Grid g(nx, nz);
  hipStream_t st = (hipStream_t)stream;
  const double dts[3] = {dt / 3, dt / 2, dt / 1};
  const int first = reverse_direction ? 0 : 1;  // 0 = x, 1 = z
  const int dirs[2] = {first, 1 - first};
  auto fwd = [&](int dir, double* init, double* forc, double* out, double dts_) {
    return pmw_rk_stage(dir, init, forc, out, hy_dens_cell, hy_dens_theta_cell, hy_dens_int,
                        hy_dens_theta_int, hy_pressure_int, src_w, PMW_BC_PERIODIC, nullptr,
                        nx, nz, dir == 0 ? dx : dz, dts_, dir == 0 ? hv_coeff_x : hv_coeff_z,
                        st);
  };
  // the state after the first direction: S -> s1 -> s2 -> mid
  int rc;
  if ((rc = fwd(dirs[0], s0, s0, s1, dts[0]))) return rc;
  if ((rc = fwd(dirs[0], s0, s1, s2, dts[1]))) return rc;
  if ((rc = fwd(dirs[0], s0, s2, mid, dts[2]))) return rc;
  for (int pass = 1; pass >= 0; --pass) {
    const int dir = dirs[pass];
    double* start = pass == 1 ? mid : s0;
    // stage states of this direction
    if ((rc = fwd(dir, start, start, s1, dts[0]))) return rc;
    if ((rc = fwd(dir, start, s1, s2, dts[1]))) return rc;
    const double d = dir == 0 ? dx : dz;
    const double hv = dir == 0 ? hv_coeff_x : hv_coeff_z;
    auto adj = [&](const double* f, const double* lin, double sc, const double* a1,
                   const double* a2, const double* a3, double* lout) {
      return adj_stage(dir, f, lin, sc, a1, a2, a3, lout, fw, hy_dens_cell,
                       hy_dens_theta_cell, hy_dens_int, hy_dens_theta_int, g, d, hv, st);
    };
    if ((rc = adj(s2, lam, dts[2], nullptr, nullptr, nullptr, l2))) return rc;
    if ((rc = adj(s1, l2, dts[1], nullptr, nullptr, nullptr, l1))) return rc;
    if ((rc = adj(start, l1, dts[0], lam, l2, l1, lam))) return rc;
  }
  return PMW_OK;
}

}  // extern "C"
