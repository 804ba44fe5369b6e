// pmw_adjoint.hip -- hand-written adjoint (reverse mode) of the time step.
//
// Forward, per direction d and RK stage s (solve/step.py:21-82, 111-141):
//   out = init + a_s * T_d(f),  a_s in {dt/3, dt/2, dt}
// where T_d is halo -> 4-tap reconstruction -> flux -> flux difference
// (+ gravity in z). Its adjoint, for a cotangent lam of `out`:
//   lam_init += lam,   lam_f += a_s * T_d'(f)^T lam.
// One launch computes a_s * T_d'(f)^T lam for a tile of cells:
//  * every face of the tile (+1 face below, +2 above) forms
//    lamF = a_s (lamT(cell after) - lamT(cell before)) / d and applies the
//    transposed 4x4 flux Jacobian at the face (from the reconstructed
//    forward values of f), giving the cotangents of its 4 reconstructed
//    values and 4 third differences; these 8 numbers per face stay in LDS;
//  * every cell then gathers the transposed 4-tap reconstruction from the 4
//    faces that read it. In x the faces form a ring (face 0 and face nx read
//    the same wrapped cells, so they are one face). In z the wall halo rows
//    are functions of the edge rows (set_bc_z, ics/bcs.py:64-148): their
//    cotangents fold back onto the edge rows with factors 1 for rho' and
//    (rho*theta)', hy[halo]/hy[edge] for rho*u and 0 for rho*w. z adds the
//    gravity term -g * lam_w to lam_rho'.
// One RK3 direction: lamS2 = dt T'(S2)^T lam3; lamS1 = dt/2 T'(S1)^T lamS2;
// lamS = dt/3 T'(S)^T lamS1 + lam3 + lamS2 + lamS1 (the init terms).
// Cotangent buffers use the state layout (4, nz+4, nx+4); only the interior
// is read or written. Derivatives are those of the real-number formulas the
// reference evaluates (e.g. rho * (v/rho) is v).
//
// Tape: pmw_tape_forward advances `members` identical grids (one batched
// launch per direction) and keeps, per step k, six state slots
//   6k: S_k, 6k+1: S1 of the first direction, 6k+2: its S2, 6k+3: the state
//   after the first direction, 6k+4 / 6k+5: S1 / S2 of the second direction,
// with S_{k+1} in slot 6k+6; pmw_tape_adjoint sweeps the steps backwards
// from it, one launch per direction (all three adjoint stages) for all
// members.
#include "pmw_common.h"
#include "pmw_fused.h"

using namespace pmw;

// The adjoint needs no bit-for-bit agreement with the reference's
// arithmetic (only with itself across kernel shapes, which share this
// code): its expressions may contract to fma.
#pragma clang fp contract(fast)

namespace {

// The 4-tap face value alone (the adjoint's faces need no third
// difference), symmetric taps summed first.
__device__ __forceinline__ double recon_val(double s0, double s1, double s2, double s3) {
  return (s0 + s3) * kW4a + (s1 + s2) * kW4b;
}

constexpr double kW4[4] = {kW4a, kW4b, kW4b, kW4a};
constexpr double kD3[4] = {-1.0, 3.0, -3.0, 1.0};
constexpr int kAdjXThreads = 256;          // x: faces per block
constexpr int kAdjXCells = kAdjXThreads - 3;  // x: cells per block
constexpr int kAdjZRows = 8;               // z: cell rows per block (64 columns)
constexpr int kAdjZWaves = 8;              // z: waves per block

__device__ __forceinline__ int wrapc(int c, int n) {
  int w = c % n;
  return w < 0 ? w + n : w;
}

// q^(gamma-1) for dp/dq = gamma C0 q^(gamma-1) (q = rho*theta at a face): a
// derivative, so a few-ulp power.
__device__ __forceinline__ double dpow(double q) { return pow_pos_fast(q, kGamma - 1.0); }

// dpow(q) continued from dpow(q0) = p0 at a nearby q0: q^a = p0 (1 + e)^a,
// e = (q - q0)/q0, by the binomial series to e^4 (truncation below 3e-17
// for |e| <= 2^-10); ok is set false for larger changes (and NaN), where
// the caller evaluates dpow(q) itself. 1/q0 is the hardware reciprocal
// (2^29 ulp) with one Newton step, ~2^-46 relative, which moves the result
// by less than 0.4 |e| 2^-46 < 4e-17. The adjoint tiles' later stages
// reverse the same face.
__device__ __forceinline__ double dpow_near(double q, double q0, double p0, bool& ok) {
  double r = __builtin_amdgcn_rcp(q0);
  r = fma(fma(-q0, r, 1.0), r, r);
  const double e = (q - q0) * r;
  ok = fabs(e) <= 0x1p-10;
  double c = -0x1.54c5337464d59p-5;          // a(a-1)(a-2)(a-3)/24, a = gamma - 1
  c = c * e + 0x1.0628bd48c9ba4p-4;          // a(a-1)(a-2)/6
  c = c * e + -0x1.eba2540426299p-4;         // a(a-1)/2
  c = c * e + 0x1.99e2b8f2ca162p-2;          // a
  return p0 + p0 * (e * c);
}

// Primitive variables of a face and dp/dq (q = rho*theta, p = C0 q^gamma)
// from pw = dpow(q): one division per face.
__device__ __forceinline__ void face_prims(const double val[4], double hyr, double hyt,
                                           double pw, double& u, double& w, double& t,
                                           double& dp) {
  const double rho = val[0] + hyr;
  const double ir = rcp_approx(rho);  // a derivative needs no IEEE quotient
  const double q = val[3] + hyt;
  u = val[1] * ir;
  w = val[2] * ir;
  t = q * ir;
  dp = (kGamma * kC0) * pw;
}

// Transposed x-flux Jacobian: cotangents of (vals, d3) from lamF.
__device__ __forceinline__ void adj_flux_x(const double val[4], double hyr, double hyt,
                                           double hv, const double lf[4], double lv[4],
                                           double ld[4], double pw) {
  double u, w, t, dp;
  face_prims(val, hyr, hyt, pw, u, w, t, dp);
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
                                           double lv[4], double ld[4], double pw) {
  double u, w, t, dp;
  face_prims(val, hyr, hyt, pw, u, w, t, dp);
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

// Arguments of one adjoint stage; per-member buffers are mstride elements
// apart (member = blockIdx.y).
struct AdjArgs {
  const double* f;     // forcing state of the stage
  const double* lam;   // cotangent of the stage output
  const double* acc1;  // nullable terms added to the result (same cell)
  const double* acc2;
  const double* acc3;
  double* out;         // result; may alias an acc, never f or lam
  const double* hy_cell;
  const double* hyt_cell;
  const double* hyr_int;
  const double* hyt_int;
  double scale;   // stage dt
  double sd;      // scale / d
  double hv;
  long long mstride;
};

__device__ __forceinline__ void adj_offset(AdjArgs& a) {
  const long long off = (long long)blockIdx.y * a.mstride;
  a.f += off;
  a.lam += off;
  a.out += off;
  if (a.acc1) a.acc1 += off;
  if (a.acc2) a.acc2 += off;
  if (a.acc3) a.acc3 += off;
}

__device__ __forceinline__ void finish_cell(const AdjArgs& a, const Grid& g, int prow, int pc,
                                            const double l[4]) {
#pragma unroll
  for (int v = 0; v < kNv; ++v) {
    const long long o = g.at(v, prow, pc);
    double r = l[v];
    if (a.acc1) r += a.acc1[o];
    if (a.acc2) r += a.acc2[o];
    if (a.acc3) r += a.acc3[o];
    a.out[o] = r;
  }
}

// ---- x: block = one row, faces c0-1 .. c0+kAdjXThreads-2 (ring, wrapped),
// cells c0 .. c0+kAdjXCells-1
__global__ __launch_bounds__(kAdjXThreads) void k_adj_x(AdjArgs a, Grid g, int ntiles) {
  __shared__ double fw[8][kAdjXThreads];
  adj_offset(a);
  const int nx = g.nx, t = threadIdx.x;
  const int k = blockIdx.x / ntiles, c0 = (blockIdx.x % ntiles) * kAdjXCells;
  const int prow = k + kHs;
  {
    const int face = c0 - 1 + t;  // reads cells face-2 .. face+1
    int col[4];                   // padded columns of those cells
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int c = face - 2 + j;
      col[j] = (c >= 0 && c < nx ? c : wrapc(c, nx)) + kHs;
    }
    double val[4], lf[4], lv[4], ld[4];
#pragma unroll
    for (int v = 0; v < kNv; ++v) {
      double s[4];
#pragma unroll
      for (int j = 0; j < 4; ++j) s[j] = a.f[g.at(v, prow, col[j])];
      val[v] = recon_val(s[0], s[1], s[2], s[3]);
    }
    // T(i) = (F(i) - F(i+1)) / dx: the face is F(i) of cell `face` (col[2])
    // and F(i+1) of cell face-1 (col[1])
#pragma unroll
    for (int v = 0; v < kNv; ++v)
      lf[v] = a.sd * (a.lam[g.at(v, prow, col[2])] - a.lam[g.at(v, prow, col[1])]);
    adj_flux_x(val, a.hy_cell[prow], a.hyt_cell[prow], a.hv, lf, lv, ld,
               dpow(val[3] + a.hyt_cell[prow]));
#pragma unroll
    for (int v = 0; v < kNv; ++v) {
      fw[v][t] = lv[v];
      fw[4 + v][t] = ld[v];
    }
  }
  __syncthreads();
  const int c = c0 + t;
  if (t < kAdjXCells && c < nx) {
    // tap j of face c+2-j reads cell c; that face is slot t+3-j
    double l[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int s = t + 3 - j;
#pragma unroll
      for (int v = 0; v < kNv; ++v) l[v] += kW4[j] * fw[v][s] + kD3[j] * fw[4 + v][s];
    }
    finish_cell(a, g, prow, c + kHs, l);
  }
}

// ---- z: block = 64 columns x kAdjZRows cell rows r0.. ; face slots
// s = 0 .. R+2 hold faces r0-1+s (those in 0..nz)
__device__ __forceinline__ double zrow_value(const double* f, const Grid& g,
                                             const double* hy, int v, int r, int pc) {
  if (r >= 0 && r < g.nz) return f[g.at(v, r + kHs, pc)];
  const int pe = r < 0 ? kHs : g.nz + kHs - 1;
  return wall_halo(v, f[g.at(v, pe, pc)], hy[pe], hy[r + kHs]);
}

template <int R>
__device__ __forceinline__ void gather_z_row(const double (*fw)[R + 3][64], int nz, int r,
                                             int r0, int lane, double l[4]) {
  // cotangent of forcing row r (-2 <= r <= nz+1) of this lane's column
#pragma unroll
  for (int v = 0; v < kNv; ++v) l[v] = 0.0;
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int face = r + 2 - j;
    if (face < 0 || face > nz) continue;
    const int s = face - r0 + 1;
#pragma unroll
    for (int v = 0; v < kNv; ++v) l[v] += kW4[j] * fw[v][s][lane] + kD3[j] * fw[4 + v][s][lane];
  }
}

template <int R>
__global__ __launch_bounds__(64 * kAdjZWaves) void k_adj_z(AdjArgs a, Grid g, int ncol) {
  constexpr int NF = R + 3;
  __shared__ double fw[8][NF][64];
  adj_offset(a);
  const int nx = g.nx, nz = g.nz;
  const int lane = threadIdx.x & 63, wy = threadIdx.x >> 6;
  const int i = (blockIdx.x % ncol) * 64 + lane, r0 = (blockIdx.x / ncol) * R;
  const int pc = i + kHs;
  if (i < nx) {
    for (int s = wy; s < NF; s += kAdjZWaves) {
      const int face = r0 - 1 + s;
      if (face < 0 || face > nz) continue;
      double val[4], lf[4], lv[4], ld[4];
#pragma unroll
      for (int v = 0; v < kNv; ++v) {
        double q[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) q[j] = zrow_value(a.f, g, a.hy_cell, v, face - 2 + j, pc);
        val[v] = recon_val(q[0], q[1], q[2], q[3]);
      }
#pragma unroll
      for (int v = 0; v < kNv; ++v) {
        const double above = face < nz ? a.lam[g.at(v, face + kHs, pc)] : 0.0;
        const double below = face > 0 ? a.lam[g.at(v, face - 1 + kHs, pc)] : 0.0;
        lf[v] = a.sd * (above - below);
      }
      adj_flux_z(val, a.hyr_int[face], a.hyt_int[face], a.hv, face == 0 || face == nz, lf, lv,
                 ld, dpow(val[3] + a.hyt_int[face]));
#pragma unroll
      for (int v = 0; v < kNv; ++v) {
        fw[v][s][lane] = lv[v];
        fw[4 + v][s][lane] = ld[v];
      }
    }
  }
  __syncthreads();
  if (i >= nx) return;
  for (int rr = wy; rr < R; rr += kAdjZWaves) {
    const int k = r0 + rr;
    if (k >= nz) break;
    double l[4], h[4];
    gather_z_row<R>(fw, nz, k, r0, lane, l);
    if (k == 0 || k == nz - 1) {
      // the wall halo rows of this edge row
      const int pe = k + kHs;
      const int rh = k == 0 ? -2 : nz;
      for (int r = rh; r < rh + 2; ++r) {
        gather_z_row<R>(fw, nz, r, r0, lane, h);
        l[0] += h[0];
        l[1] += h[1] * a.hy_cell[r + kHs] / a.hy_cell[pe];
        l[3] += h[3];
      }
    }
    // gravity: T_w -= rho'_forcing * g
    l[0] += -kGrav * (a.scale * a.lam[g.at(2, k + kHs, pc)]);
    finish_cell(a, g, k + kHs, pc, l);
  }
}

int adj_stage(int dir, const double* f, const double* lam, double scale, const double* acc1,
              const double* acc2, const double* acc3, double* lam_out,
              const double* hy_cell, const double* hyt_cell, const double* hyr_int,
              const double* hyt_int, Grid g, double d, double hv, int members,
              long long mstride, hipStream_t st) {
  AdjArgs a{f, lam, acc1, acc2, acc3, lam_out, hy_cell, hyt_cell, hyr_int, hyt_int,
            scale, scale / d, hv, mstride};
  if (dir == 0) {
    const int ntiles = (g.nx + kAdjXCells - 1) / kAdjXCells;
    k_adj_x<<<dim3((unsigned)(ntiles * g.nz), members), kAdjXThreads, 0, st>>>(a, g, ntiles);
  } else {
    const int ncol = (g.nx + 63) / 64;
    const int nrt = (g.nz + kAdjZRows - 1) / kAdjZRows;
    k_adj_z<kAdjZRows><<<dim3((unsigned)(ncol * nrt), members), 64 * kAdjZWaves, 0, st>>>(
        a, g, ncol);
  }
  PMW_CHECK_LAUNCH();
  return PMW_OK;
}

// ---- one RK3 direction in one launch --------------------------------------
// A block holds a line tile: x = one row, NT cells along it (CW = 1); z = CW
// columns x NR = NT/CW rows (slot t = j*CW + col). The three adjoint stages
// run in LDS on a tile with a 7-cell halo (each stage widens the dependency
// by 2 cells, plus 1 for the face before): lamS2 = dt T'(S2)^T lam3 on tile
// slots 3..NR-4, lamS1 on 5..NR-6, lamS on 7..NR-8 (the NR-14 output cells).
// Per cell and direction it reads S, S1, S2 and lam3 once and writes lamS.
struct AdjDirArgs {
  const double* s[3];  // forcing states of RK stages 1..3: S, S1, S2
  const double* lam;   // cotangent of the direction's result
  double* out;         // cotangent of its start state (must not alias lam)
  const double* hy_cell;
  const double* hyt_cell;
  const double* hyr_int;
  const double* hyt_int;
  double dts[3];       // dt/3, dt/2, dt
  double d, id, hv;
  long long mstride;
};

// la holds lam3, then lamS2, then lamS1, each stage's result written in
// place after its faces are done: a stage's faces read only slots the stage
// before wrote (faces 2+2m..NR-2-2m read slots 1+2m..NR-2-2m, inside the
// cells 3+2(m-1)..NR-4-2(m-1) of stage m-1), and slots outside the domain
// keep the zero they were loaded with.
// The face cotangents of the third differences are -hv lamF, so only those
// of the reconstructed values are kept; a cell forms the others from la
// (96 B of LDS per slot).
template <int NT>
struct AdjDirLds {
  double st[kNv][NT];  // forcing state of the stage being reversed
  double fw[kNv][NT];  // face cotangents of the reconstructed values
  double la[kNv][NT];  // lam3, then lamS2, then lamS1
};

// Consecutive tiles (neighbouring columns / row segments) go to the same XCD
// so that the cache lines they share are fetched into one L2.
__device__ __forceinline__ int adj_xcd_logical(int b, int per_xcd) {
  return (b % 8) * per_xcd + b / 8;
}

template <int DIR, int CW, int NT>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(6))) void k_adj_dir(
    AdjDirArgs a, Grid g, int nline, int nreal,
                                                int per_xcd) {
  constexpr int NR = NT / CW, H = 7, C = NR - 2 * H, S = DIR == 0 ? 1 : CW;
  __shared__ AdjDirLds<NT> sm;
  const int blk = adj_xcd_logical(blockIdx.x, per_xcd);
  if (blk >= nreal) return;  // whole block, before any barrier
  const long long off = (long long)blockIdx.y * a.mstride;
  const int nx = g.nx, nz = g.nz, t = threadIdx.x;
  const int j = t / CW, col = t % CW;
  // line and position along it
  int prow = 0, pcol = 0, p;
  bool inside, wall = false;
  int pe = 0;  // z wall slots: padded edge row
  if (DIR == 0) {
    const int k = blk / nline;
    p = (blk % nline) * C - H + j;
    prow = k + kHs;
    pcol = wrapc(p, nx) + kHs;
    inside = true;
  } else {
    const int i = (blk % nline) * CW + col;
    p = (blk / nline) * C - H + j;
    pcol = i + kHs;
    const bool colok = i < nx;
    inside = colok && p >= 0 && p < nz;
    wall = colok && (p == -2 || p == -1 || p == nz || p == nz + 1);
    pe = p < 0 ? kHs : nz + kHs - 1;
    prow = p + kHs;
  }
  const bool own = inside && j >= H && j < H + C && (DIR == 0 ? p < nx : true);
  // lam3 (zero outside the domain), lamS2 cleared
  // acc: the init terms lam3 + lamS2 + lamS1, summed as the stages finish
  double acc[4];
#pragma unroll
  for (int v = 0; v < kNv; ++v) {
    acc[v] = inside ? a.lam[off + g.at(v, prow, pcol)] : 0.0;
    sm.la[v][t] = acc[v];
  }
  // this thread's slot of a forcing state (wall halo rows rebuilt)
  auto fetch_state = [&](const double* src, double q[4]) {
#pragma unroll
    for (int v = 0; v < kNv; ++v) {
      q[v] = 0.0;
      if (inside)
        q[v] = src[off + g.at(v, prow, pcol)];
      else if (wall)
        q[v] = wall_halo(v, src[off + g.at(v, pe, pcol)], a.hy_cell[pe], a.hy_cell[prow]);
    }
  };
  {
    double q[4];
    fetch_state(a.s[2], q);
#pragma unroll
    for (int v = 0; v < kNv; ++v) sm.st[v][t] = q[v];
  }
  // hydrostatic values of this thread's face (z) or row (x): the same for
  // the three stages, loaded once
  double hy_r = 0.0, hy_t = 0.0;
  if (DIR == 0) {
    hy_r = a.hy_cell[prow];
    hy_t = a.hyt_cell[prow];
  } else if (pcol - kHs < nx && p >= 0 && p <= nz) {
    hy_r = a.hyr_int[p];
    hy_t = a.hyt_int[p];
  }
  double q0 = 1.0, p0 = 1.0;  // this face's stage-3 power reference
  __syncthreads();
#pragma unroll
  for (int m = 0; m < 3; ++m) {  // m = 0: stage 3 (forcing S2), 1: stage 2, 2: stage 1
    const int stage = 2 - m;
    double (*lin)[NT] = sm.la;
    const double sc = a.dts[stage];
    const double sd = sc * a.id;
    // the next stage's forcing state, loaded now and stored to LDS once this
    // stage's faces are done, so the loads overlap the face arithmetic
    double nq[4];
    if (m < 2) fetch_state(a.s[stage - 1], nq);
    // faces at slots 2+2m .. NR-2-2m (z: faces outside 0..nz are zero)
    const bool zout = DIR == 1 && (p < 0 || p > nz);
    bool pok = true;
    if (zout) {
#pragma unroll
      for (int q = 0; q < kNv; ++q) sm.fw[q][t] = 0.0;
    } else if (j >= 2 + 2 * m && j <= NR - 2 - 2 * m && (DIR == 0 || col + (blk % nline) * CW < nx)) {
      double val[4], lf[4], lv[4], ld[4];
#pragma unroll
      for (int v = 0; v < kNv; ++v)
        val[v] = recon_val(sm.st[v][t - 2 * S], sm.st[v][t - S], sm.st[v][t], sm.st[v][t + S]);
#pragma unroll
      for (int v = 0; v < kNv; ++v) lf[v] = sd * (lin[v][t] - lin[v][t - S]);
      // the face's power: evaluated at stage 3, continued at stages 2 and 1
      const double q = val[3] + hy_t;
      double pw;
      if (m == 0) {
        pw = dpow(q);
        q0 = q;
        p0 = pw;
      } else {
        pw = dpow_near(q, q0, p0, pok);
      }
      if (DIR == 0)
        adj_flux_x(val, hy_r, hy_t, a.hv, lf, lv, ld, pw);
      else
        adj_flux_z(val, hy_r, hy_t, a.hv, p == 0 || p == nz, lf, lv, ld, pw);
#pragma unroll
      for (int v = 0; v < kNv; ++v) sm.fw[v][t] = lv[v];
    }
    if (!pok) {  // the face again with the power itself (rare)
      double val[4], lf[4], lv[4], ld[4];
#pragma unroll
      for (int v = 0; v < kNv; ++v)
        val[v] = recon_val(sm.st[v][t - 2 * S], sm.st[v][t - S], sm.st[v][t], sm.st[v][t + S]);
#pragma unroll
      for (int v = 0; v < kNv; ++v) lf[v] = sd * (lin[v][t] - lin[v][t - S]);
      const double pw = dpow(val[3] + hy_t);
      if (DIR == 0)
        adj_flux_x(val, hy_r, hy_t, a.hv, lf, lv, ld, pw);
      else
        adj_flux_z(val, hy_r, hy_t, a.hv, p == 0 || p == nz, lf, lv, ld, pw);
#pragma unroll
      for (int v = 0; v < kNv; ++v) sm.fw[v][t] = lv[v];
    }
    __syncthreads();
    // gravity needs lin[w] at this slot: read before lin may be overwritten
    const double lw = lin[2][t];
    // cells at slots 3+2m .. NR-4-2m
    double l[4] = {0.0, 0.0, 0.0, 0.0};
    const bool cell = inside && j >= 3 + 2 * m && j <= NR - 4 - 2 * m;
    if (cell) {
      // d3 cotangent of the face at slot sl (the face between slots sl-S
      // and sl, at position fp): -hv lamF, lamF formed as the face phase
      // forms it; zero outside 0..nz and for rho' at the walls (z)
      auto ldf = [&](int sl, int fp, int v) {
        if (DIR == 1 && (fp < 0 || fp > nz || (v == 0 && (fp == 0 || fp == nz)))) return 0.0;
        return -a.hv * (sd * (lin[v][sl] - lin[v][sl - S]));
      };
      // Away from the walls (all of x) the four faces' d3 cotangents sum
      // to -hv sd times the fourth difference of lambda at the cell
      // (kD3 = -1, 3, -3, 1 over differences of neighbours).
      const bool interior = DIR == 0 || (p >= 2 && p <= nz - 3);
      const double c4 = -a.hv * sd;
#pragma unroll
      for (int v = 0; v < kNv; ++v) {
        double lam[5];  // lin at slots t-2S .. t+2S
#pragma unroll
        for (int i = 0; i < 5; ++i) lam[i] = lin[v][t + (i - 2) * S];
        // transposed 4-tap reconstruction (kW4 = a, b, b, a)
        double w = kW4a * (sm.fw[v][t + 2 * S] + sm.fw[v][t - S]) +
                   kW4b * (sm.fw[v][t + S] + sm.fw[v][t]);
        if (interior) {
          // sum_q kD3[q] (lam[4-q] - lam[3-q]) = 4 (l1 + l3) - (l0 + l4) - 6 l2
          const double d4 = (4.0 * (lam[1] + lam[3]) - (lam[0] + lam[4])) - 6.0 * lam[2];
          l[v] = w + c4 * d4;
        } else {
#pragma unroll
          for (int q = 0; q < 4; ++q) {  // face at slot t+(2-q)S, position p+2-q
            const int fp = p + 2 - q;
            const bool zero = fp < 0 || fp > nz || (v == 0 && (fp == 0 || fp == nz));
            if (!zero) w += kD3[q] * (c4 * (lam[4 - q] - lam[3 - q]));
          }
          l[v] = w;
        }
      }
      if (DIR == 1) {
        if (p == 0 || p == nz - 1) {
          // the wall halo rows of this edge row: positions -2, -1 or nz, nz+1
          const int r0w = p == 0 ? -2 : nz;
          for (int r = r0w; r < r0w + 2; ++r) {
            double h[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
            for (int q = 0; q < 4; ++q) {
              const int f = r + 2 - q;
              if (f < 0 || f > nz) continue;
              const int sl = t + (f - p) * S;
#pragma unroll
              for (int v = 0; v < kNv; ++v) h[v] += kW4[q] * sm.fw[v][sl] + kD3[q] * ldf(sl, f, v);
            }
            l[0] += h[0];
            l[1] += h[1] * a.hy_cell[r + kHs] / a.hy_cell[prow];
            l[3] += h[3];
          }
        }
        l[0] += -kGrav * (sc * lw);
      }
    }
    if (m < 2) {
      // next forcing state; st is free once the faces are done
#pragma unroll
      for (int v = 0; v < kNv; ++v) sm.st[v][t] = nq[v];
      __syncthreads();  // every cell has read its neighbours' la
      if (cell) {
#pragma unroll
        for (int v = 0; v < kNv; ++v) {
          sm.la[v][t] = l[v];  // in place: this stage's faces are done
          acc[v] += l[v];
        }
      }
      __syncthreads();
    } else if (own) {
#pragma unroll
      for (int v = 0; v < kNv; ++v) {
        a.out[off + g.at(v, prow, pcol)] = l[v] + acc[v];
      }
    }
  }
}

template <int DIR, int CW, int NT>
int launch_adj_dir(const AdjDirArgs& a, Grid g, int members, hipStream_t st) {
  constexpr int C = NT / CW - 14;
  int nline;
  long long nreal;
  if (DIR == 0) {
    nline = (g.nx + C - 1) / C;
    nreal = (long long)nline * g.nz;
  } else {
    nline = (g.nx + CW - 1) / CW;
    nreal = (long long)nline * ((g.nz + C - 1) / C);
  }
  if (nreal > (1LL << 31) - 64) return PMW_EINVAL;
  const int per = (int)((nreal + 7) / 8);
  k_adj_dir<DIR, CW, NT><<<dim3((unsigned)per * 8, members), NT, 0, st>>>(a, g, nline,
                                                                          (int)nreal, per);
  PMW_CHECK_LAUNCH();
  return PMW_OK;
}

bool grid_ok(int nx, int nz) {
  // blocks per launch stay below 2^31 and rows index as int
  return nx >= 2 && nz >= 2 && (long long)(nx + 4) * (nz + 4) * 4 < (1LL << 40);
}

long long state_len(int nx, int nz) { return 4LL * (nz + 4) * (nx + 4); }

// ---- halo transposes -------------------------------------------------------
// Periodic x halos: column 0 <- nx, 1 <- nx+1, nx+2 <- 2, nx+3 <- 3 on the
// interior rows. One thread per (variable, interior row) does all four
// columns, so sources shared between them (nx = 2, 3) are summed in order.
__global__ void k_adj_halo_x(double* __restrict__ l, Grid g) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= kNv * g.nz) return;
  const int v = idx / g.nz, k = idx % g.nz, nx = g.nx;
  double* r = l + g.at(v, k + kHs, 0);
  const double a0 = r[0], a1 = r[1], b0 = r[nx + 2], b1 = r[nx + 3];
  r[0] = 0.0;
  r[1] = 0.0;
  r[nx + 2] = 0.0;
  r[nx + 3] = 0.0;
  r[nx] += a0;
  r[nx + 1] += a1;
  r[2] += b0;
  r[3] += b1;
}

// Wall z halos on every column: rows 0, 1 from row 2 and rows nz+2, nz+3
// from row nz+1, each a linear map of the edge value (wall_halo: 1, the
// density ratio, or 0), so the transpose applies the same factor.
__global__ void k_adj_halo_z(double* __restrict__ l, Grid g, const double* __restrict__ hy) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= g.ld) return;
  const int nz = g.nz, lo = kHs, hi = nz + kHs - 1;
  for (int v = 0; v < kNv; ++v) {
    const double h0 = l[g.at(v, 0, c)], h1 = l[g.at(v, 1, c)];
    const double h2 = l[g.at(v, nz + 2, c)], h3 = l[g.at(v, nz + 3, c)];
    l[g.at(v, 0, c)] = 0.0;
    l[g.at(v, 1, c)] = 0.0;
    l[g.at(v, nz + 2, c)] = 0.0;
    l[g.at(v, nz + 3, c)] = 0.0;
    l[g.at(v, lo, c)] += wall_halo(v, h0, hy[lo], hy[0]) + wall_halo(v, h1, hy[lo], hy[1]);
    l[g.at(v, hi, c)] += wall_halo(v, h2, hy[hi], hy[nz + 2]) +
                         wall_halo(v, h3, hy[hi], hy[nz + 3]);
  }
}

}  // namespace

extern "C" {

int pmw_adjoint_set_bc_x(double* lam, int nx, int nz, void* stream) {
  if (!grid_ok(nx, nz) || !lam) return PMW_EINVAL;
  Grid g(nx, nz);
  k_adj_halo_x<<<(unsigned)((kNv * nz + 255) / 256), 256, 0, (hipStream_t)stream>>>(lam, g);
  PMW_CHECK_LAUNCH();
  return PMW_OK;
}

int pmw_adjoint_set_bc_z(double* lam, const double* hy_dens_cell, int nx, int nz,
                         void* stream) {
  if (!grid_ok(nx, nz) || !lam || !hy_dens_cell) return PMW_EINVAL;
  Grid g(nx, nz);
  k_adj_halo_z<<<(unsigned)((nx + 4 + 255) / 256), 256, 0, (hipStream_t)stream>>>(
      lam, g, hy_dens_cell);
  PMW_CHECK_LAUNCH();
  return PMW_OK;
}

int pmw_adjoint_stage(int direction, const double* state_forcing, const double* lam_in,
                      double scale, const double* acc1, const double* acc2,
                      const double* acc3, double* lam_out, const double* hy_dens_cell,
                      const double* hy_dens_theta_cell, const double* hy_dens_int,
                      const double* hy_dens_theta_int, int nx, int nz, double d,
                      double hv_coeff, void* stream) {
  if (!grid_ok(nx, nz) || !state_forcing || !lam_in || !lam_out || !hy_dens_cell ||
      (direction == 0 && !hy_dens_theta_cell) ||
      (direction == 1 && (!hy_dens_int || !hy_dens_theta_int)))
    return PMW_EINVAL;
  if (direction != 0 && direction != 1) return PMW_EINVAL;
  if (lam_out == lam_in || lam_out == state_forcing) return PMW_EALIAS;
  return adj_stage(direction, state_forcing, lam_in, scale, acc1, acc2, acc3, lam_out,
                   hy_dens_cell, hy_dens_theta_cell, hy_dens_int, hy_dens_theta_int,
                   Grid(nx, nz), d, hv_coeff, 1, 0, (hipStream_t)stream);
}

long long pmw_tape_slot_len(int nx, int nz, int members) {
  if (!grid_ok(nx, nz) || members < 1 || members > 65535) return 0;
  return state_len(nx, nz) * members;
}

int pmw_tape_forward(double* tape, long long tape_len, int members,
                     const double* hy_dens_cell, const double* hy_dens_theta_cell,
                     const double* hy_dens_int, const double* hy_dens_theta_int,
                     const double* hy_pressure_int, const double* src_w, int nx, int nz,
                     double dx, double dz, double dt, double hv_coeff_x, double hv_coeff_z,
                     int reverse_direction, int nsteps, int store_stages, const int* tuning,
                     void* stream) {
  const long long slot = pmw_tape_slot_len(nx, nz, members);
  if (slot == 0 || !tape || !hy_dens_cell || !hy_dens_theta_cell || !hy_dens_int ||
      !hy_dens_theta_int || !hy_pressure_int || nsteps < 0 || nsteps > 100000)
    return PMW_EINVAL;
  int x_nt = 512, z_seg = 4128;
  if (tuning) {
    if (tuning[0] != 0 || !fused_x_tile_ok(tuning[1]) || !fused_z_seg_ok(tuning[2]))
      return PMW_EINVAL;
    x_nt = tuning[1];
    z_seg = tuning[2];
  }
  if (tape_len < (6LL * nsteps + 1) * slot) return PMW_ESPACE;
  Grid g(nx, nz);
  hipStream_t st = (hipStream_t)stream;
  DirArgs a;
  a.hy_cell = gptr(hy_dens_cell);
  a.hyt_cell = gptr(hy_dens_theta_cell);
  a.hyr_int = gptr(hy_dens_int);
  a.hyt_int = gptr(hy_dens_theta_int);
  a.hyp_int = gptr(hy_pressure_int);
  a.src_w = gptr(src_w);
  a.dt1 = dt / 3;
  a.dt2 = dt / 2;
  a.dt3 = dt / 1;
  a.mstride = state_len(nx, nz);
  for (int k = 0; k < nsteps; ++k) {
    const int rev = (reverse_direction != 0) ^ (k & 1);
    const int first = rev ? 0 : 1;  // 0 = x, 1 = z
    for (int pass = 0; pass < 2; ++pass) {
      const int dir = pass == 0 ? first : 1 - first;
      if (store_stages) {
        double* base = tape + (6LL * k + 3 * pass) * slot;
        a.in = gptr(base);
        a.st = gptr(base);  // halo refresh of the start state (never read by this launch)
        a.s1 = gptr(base + slot);
        a.tmp = gptr(base + 2 * slot);
        a.out = gptr(base + 3 * slot);
        a.tmp_interior = true;
      } else {
        // end state only: the states alternate between slots 0 and 1 (a
        // working set the size of evolve's buffers instead of a stream of
        // fresh slots), slot 2 takes the stage-3 halo refreshes without the
        // dead stage-2 interior, and the last direction writes slot 6n
        const long long q = 2LL * k + pass;
        a.st = gptr(tape + (q % 2) * slot);  // halo refresh of the start state
        a.in = a.st;
        a.s1 = nullptr;
        a.tmp = gptr(tape + 2 * slot);
        a.out = gptr(tape + (q == 2LL * nsteps - 1 ? 6LL * nsteps : (q + 1) % 2) * slot);
        a.tmp_interior = false;
      }
      a.d = dir == 0 ? dx : dz;
      a.id = 1.0 / a.d;
      a.hv = dir == 0 ? hv_coeff_x : hv_coeff_z;
      const int rc = dir == 0 ? launch_dir_x(a, g, x_nt, st, members)
                              : launch_dir_z(a, g, z_seg, st, members);
      if (rc) return rc;
    }
  }
  return PMW_OK;
}

int pmw_tape_adjoint(const double* tape, long long tape_len, int members, double* lam,
                     double* workspace, long long workspace_len, const double* hy_dens_cell,
                     const double* hy_dens_theta_cell, const double* hy_dens_int,
                     const double* hy_dens_theta_int, int nx, int nz, double dx, double dz,
                     double dt, double hv_coeff_x, double hv_coeff_z, int reverse_direction,
                     int nsteps, const int* tuning, void* stream) {
  const long long slot = pmw_tape_slot_len(nx, nz, members);
  if (slot == 0 || !tape || !lam || !workspace || !hy_dens_cell || !hy_dens_theta_cell ||
      !hy_dens_int || !hy_dens_theta_int || nsteps < 0 || nsteps > 100000)
    return PMW_EINVAL;
  const int x_nt = tuning ? tuning[0] : 256;
  const int z_tile = tuning ? tuning[1] : 8512;
  if ((x_nt != 256 && x_nt != 384) ||
      (z_tile != 4256 && z_tile != 4384 && z_tile != 2256 && z_tile != 8512))
    return PMW_EINVAL;
  if (tape_len < (6LL * nsteps + 1) * slot || workspace_len < slot) return PMW_ESPACE;
  const double* tape_end = tape + (6LL * nsteps + 1) * slot;
  if ((lam >= tape && lam < tape_end) || (workspace >= tape && workspace < tape_end) ||
      (lam >= workspace && lam < workspace + slot) || (workspace >= lam && workspace < lam + slot))
    return PMW_EALIAS;
  Grid g(nx, nz);
  hipStream_t st = (hipStream_t)stream;
  AdjDirArgs a;
  a.hy_cell = hy_dens_cell;
  a.hyt_cell = hy_dens_theta_cell;
  a.hyr_int = hy_dens_int;
  a.hyt_int = hy_dens_theta_int;
  a.dts[0] = dt / 3;
  a.dts[1] = dt / 2;
  a.dts[2] = dt / 1;
  a.mstride = state_len(nx, nz);
  // the cotangent ping-pongs lam -> workspace -> lam within each step
  double* bufs[2] = {lam, workspace};
  for (int k = nsteps - 1; k >= 0; --k) {
    const int rev = (reverse_direction != 0) ^ (k & 1);
    const int first = rev ? 0 : 1;  // 0 = x, 1 = z
    for (int pass = 1; pass >= 0; --pass) {
      const int dir = pass == 0 ? first : 1 - first;
      const double* s0 = tape + (6LL * k + 3 * pass) * slot;
      a.s[0] = s0;
      a.s[1] = s0 + slot;
      a.s[2] = s0 + 2 * slot;
      a.lam = bufs[pass == 1 ? 0 : 1];
      a.out = bufs[pass == 1 ? 1 : 0];
      a.d = dir == 0 ? dx : dz;
      a.id = 1.0 / a.d;
      a.hv = dir == 0 ? hv_coeff_x : hv_coeff_z;
      int rc;
      if (dir == 0)
        rc = x_nt == 256 ? launch_adj_dir<0, 1, 256>(a, g, members, st)
                         : launch_adj_dir<0, 1, 384>(a, g, members, st);
      else
        rc = z_tile == 4256   ? launch_adj_dir<1, 4, 256>(a, g, members, st)
             : z_tile == 4384 ? launch_adj_dir<1, 4, 384>(a, g, members, st)
             : z_tile == 8512 ? launch_adj_dir<1, 8, 512>(a, g, members, st)
                              : launch_adj_dir<1, 2, 256>(a, g, members, st);
      if (rc) return rc;
    }
  }
  return PMW_OK;
}

}  // extern "C"
