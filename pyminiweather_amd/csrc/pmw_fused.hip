// pmw_fused.hip -- one launch per direction: the three RK stages of one
// dimensional-splitting direction fused into a single kernel (gfx950).
//
// For one direction the reference runs (solve/step.py:111-141)
//   S1 = S + dt/3 T(S),  S2 = S + dt/2 T(S1),  S3 = S + dt T(S2)
// with a halo refresh of each forcing state. Each stage widens the stencil
// by 2 cells, so an output tile needs S on the tile +- 6 cells: stage 1 is
// computed on tile +- 4, stage 2 on tile +- 2, stage 3 on the tile. Per
// interior cell and direction this reads S once and writes the result (and,
// in the last direction of a call, the stage-2 state, state_tmp): 64 or 96 B
// instead of the 256 B of three stage launches. Every stage value is computed
// with the same operations as the per-stage kernels, so results are bitwise
// identical to them.
//
// Buffers of a direction: `in` (its start state, read only), `out` (its
// result interior, never aliases `in`), `st` (the reference's `state`: it
// receives the halo refresh of stage 1, computed from `in`) and `tmp` (the
// reference's `state_tmp`: stage-2 interior plus the stage-3 halo refresh).
//
//  * x (k_dir_x<NT>): one block of NT threads per (row, tile of NT-12 cells).
//    S, the running stage state and face fluxes live in LDS; periodic wrap
//    replaces the halo columns at every stage. k_dir_xd<NT>: the same over
//    the rows laid end to end, so that narrow rows fill whole blocks.
//  * z, LDS tiles (k_dir_zt<CW,NT>; k_dir_zt2<CW,NR> with two cells per
//    thread, the default): the x scheme turned on its side, CW columns per
//    block, the wall halo rows of every stage state rebuilt from its edge row.
//    k_dir_zt2d: the same over the column groups laid end to end.
//  * z, marching (k_dir_z<L>): a block is three waves over 64 columns and a
//    segment of L rows; wave s runs RK stage s+1 one row step behind the
//    previous stage, rows handed over through a 2-slot LDS ring with one
//    barrier per step. Wall halo rows are synthesised in registers. Extra
//    lanes past nx write the corner columns of the halo rows.
//  * k_ens_*: the same tiles over the flattened tile list of an ensemble.
// Block ids are remapped so that each XCD runs a contiguous range of tiles.
#include <mutex>
#include <new>
#include <unordered_map>

#include "pmw_common.h"
#include "pmw_fused.h"

using namespace pmw;

namespace {

constexpr int kNumXcd = 8;

__device__ __forceinline__ int xcd_logical(int b, int per_xcd) {
  return (b % kNumXcd) * per_xcd + b / kNumXcd;
}

// Periodic column of cell c; the division only for the few lanes outside
// the row (the tile halos at the row ends).
__device__ __forceinline__ int wrap_cell(int c, int nx) {
  if (c >= 0 && c < nx) return c;
  int w = c % nx;
  return w < 0 ? w + nx : w;
}

// Phase order between the waves of a block by flags in LDS instead of block
// barriers, for tiles whose every phase reads and writes only a wave's own
// cells and the 2 cells on either side of them along the tile's line (in
// the one-cell z tiles a wave holds 8 whole rows of 8 columns): a wave waits
// for its two neighbours alone, so one late wave no longer holds up the
// whole block. wf[1 + w] = the last phase wave w finished; wf[0] and
// wf[NW + 1] stay above every phase. Every wave passes every phase in order,
// so the wave with the lowest phase never waits and the block cannot stall.
// The flags are set before the block's first barrier, which publishes them.
__device__ __forceinline__ void wave_flags_init(int* wf, int nw, int t) {
  if (t < nw + 2)
    __hip_atomic_store(&wf[t], (t == 0 || t == nw + 1) ? 1 << 20 : 0, __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_WORKGROUP);
}

// Phase P (> 0, increasing) of the calling wave is done; returns once both
// neighbouring waves have done it too.
__device__ __forceinline__ void wave_phase_done(int* wf, int P, int t) {
  const int wv = t >> 6;
  __builtin_amdgcn_sched_barrier(0);
  // this wave's LDS writes of phase P are complete before its flag says so
  asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory");
  if ((t & 63) == 0)
    __hip_atomic_store(&wf[1 + wv], P, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  while (true) {
    const int lo = __hip_atomic_load(&wf[wv], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    const int hi = __hip_atomic_load(&wf[2 + wv], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    if (lo >= P && hi >= P) break;
    __builtin_amdgcn_s_sleep(1);
  }
  // the neighbours' LDS writes are read after their flags
  asm volatile("" ::: "memory");
  __builtin_amdgcn_sched_barrier(0);
}

// ------------------------------------------------------------------ x ------
// LDS of one x tile: S on slots 0..NT-1 (cell = i0 - 6 + slot), the running
// stage state (S1, then S2 in place) and the face fluxes (face j between
// slots j-1 and j).
template <int NT>
struct XTileLds {
  double s0[kNv][NT];
  double s1[kNv][NT];
  double fl[kNv][NT];
};

// The three RK stages of one x tile: row k, output cells i0 .. i0+NT-13.
// Called by all NT threads of the block (contains barriers).
template <int NT>
__device__ __forceinline__ void dir_x_tile(const DirArgs& a, const Grid& g, int k, int i0,
                                           XTileLds<NT>& sm) {
  constexpr int H = 6, C = NT - 2 * H;
  auto& s0 = sm.s0;
  auto& s1 = sm.s1;
  auto& fl = sm.fl;
  const int nx = g.nx, prow = k + kHs, t = threadIdx.x;
  const int cell = i0 - H + t;
  const int cw = wrap_cell(cell, nx);
  const long long srow = (long long)k * nx;
  const long long rowoff = (long long)prow * g.ld;         // block-uniform
  const unsigned bw = (unsigned)(cw + kHs) * 8u;            // byte offsets in the row
  const unsigned bc = (unsigned)(cell + kHs) * 8u;          // (used where cell >= -2)
  {
    double q[4];
#pragma unroll
    for (int v = 0; v < kNv; ++v) q[v] = ld_off(a.in + v * g.plane + rowoff, bw);
#pragma unroll
    for (int v = 0; v < kNv; ++v) s0[v][t] = q[v];
    if (cell == -2 || cell == -1 || cell == nx || cell == nx + 1) {
#pragma unroll
      for (int v = 0; v < kNv; ++v) st_off(a.st + v * g.plane + rowoff, bc, q[v]);
    }
  }
  const double hyr = a.hy_cell[prow], hyt = a.hyt_cell[prow];
  const double src = a.src_w ? a.src_w[srow + cw] : 0.0;
  const double id = a.id;
  __syncthreads();
  // face j is computed by thread j in all three stages, which keeps its
  // stage-1 pressure power for the later stages (POW_FIRST / POW_NEXT)
  PowRef pref;
  // ---- stage 1: faces 2..NT-2, cells 2..NT-3, S -> s1
  {
    const int j = t;
    if (j >= 2 && j <= NT - 2) {
      double v4[4], d3[4], F[4];
#pragma unroll
      for (int v = 0; v < kNv; ++v)
        recon4(s0[v][j - 2], s0[v][j - 1], s0[v][j], s0[v][j + 1], v4[v], d3[v]);
      flux_x_face(v4, d3, hyr, hyt, a.hv, F, POW_FIRST, &pref);
#pragma unroll
      for (int v = 0; v < kNv; ++v) fl[v][j] = F[v];
    }
  }
  __syncthreads();
  if (t >= 2 && t <= NT - 3) {
    double td[4];
#pragma unroll
    for (int v = 0; v < kNv; ++v) td[v] = tend_of(fl[v][t], fl[v][t + 1], a.d, id);
    if (a.src_w) td[2] = td[2] + src;
#pragma unroll
    for (int v = 0; v < kNv; ++v) s1[v][t] = s0[v][t] + a.dt1 * td[v];
    if (a.s1 && cell >= i0 && cell < i0 + C && cell < nx) {
#pragma unroll
      for (int v = 0; v < kNv; ++v) st_off(a.s1 + v * g.plane + rowoff, bc, s1[v][t]);
    }
  }
  __syncthreads();
  // ---- stage 2: faces 4..NT-4, cells 4..NT-5, s1 -> s1 (in place)
  {
    const int j = t;
    bool ok = true;
    if (j >= 4 && j <= NT - 4) {
      double v4[4], d3[4], F[4];
#pragma unroll
      for (int v = 0; v < kNv; ++v)
        recon4(s1[v][j - 2], s1[v][j - 1], s1[v][j], s1[v][j + 1], v4[v], d3[v]);
      flux_x_face(v4, d3, hyr, hyt, a.hv, F, POW_TRY, &pref, &ok);
#pragma unroll
      for (int v = 0; v < kNv; ++v) fl[v][j] = F[v];
    }
    if (!ok) {  // the face again with pow_pos (rare)
      double v4[4], d3[4], F[4];
#pragma unroll
      for (int v = 0; v < kNv; ++v)
        recon4(s1[v][j - 2], s1[v][j - 1], s1[v][j], s1[v][j + 1], v4[v], d3[v]);
      flux_x_face(v4, d3, hyr, hyt, a.hv, F);
#pragma unroll
      for (int v = 0; v < kNv; ++v) fl[v][j] = F[v];
    }
  }
  __syncthreads();
  if (t >= 4 && t <= NT - 5) {
    double q[4];
#pragma unroll
    for (int v = 0; v < kNv; ++v) {
      const double td = tend_of(fl[v][t], fl[v][t + 1], a.d, id);
      q[v] = td;
    }
    if (a.src_w) q[2] = q[2] + src;
#pragma unroll
    for (int v = 0; v < kNv; ++v) q[v] = s0[v][t] + a.dt2 * q[v];
#pragma unroll
    for (int v = 0; v < kNv; ++v) s1[v][t] = q[v];
    // state_tmp: own cells, plus its x halos (set_bc_x of stage 3's forcing)
    const bool own = a.tmp_interior && cell >= i0 && cell < i0 + C && cell < nx;
    const bool halo = cell == -2 || cell == -1 || cell == nx || cell == nx + 1;
    if (own || halo) {
#pragma unroll
      for (int v = 0; v < kNv; ++v) st_off(a.tmp + v * g.plane + rowoff, bc, q[v]);
    }
  }
  __syncthreads();
  // ---- stage 3: faces 6..NT-6, cells 6..NT-7 -> out
  {
    const int j = t;
    bool ok = true;
    if (j >= 6 && j <= NT - 6) {
      double v4[4], d3[4], F[4];
#pragma unroll
      for (int v = 0; v < kNv; ++v)
        recon4(s1[v][j - 2], s1[v][j - 1], s1[v][j], s1[v][j + 1], v4[v], d3[v]);
      flux_x_face(v4, d3, hyr, hyt, a.hv, F, POW_TRY, &pref, &ok);
#pragma unroll
      for (int v = 0; v < kNv; ++v) fl[v][j] = F[v];
    }
    if (!ok) {  // the face again with pow_pos (rare)
      double v4[4], d3[4], F[4];
#pragma unroll
      for (int v = 0; v < kNv; ++v)
        recon4(s1[v][j - 2], s1[v][j - 1], s1[v][j], s1[v][j + 1], v4[v], d3[v]);
      flux_x_face(v4, d3, hyr, hyt, a.hv, F);
#pragma unroll
      for (int v = 0; v < kNv; ++v) fl[v][j] = F[v];
    }
  }
  __syncthreads();
  if (t >= 6 && t <= NT - 7 && cell < nx) {
    double td[4];
#pragma unroll
    for (int v = 0; v < kNv; ++v) td[v] = tend_of(fl[v][t], fl[v][t + 1], a.d, id);
    if (a.src_w) td[2] = td[2] + src;
#pragma unroll
    for (int v = 0; v < kNv; ++v)
      st_off(a.out + v * g.plane + rowoff, bc, s0[v][t] + a.dt3 * td[v]);
  }
}

// Member blockIdx.y of a batched launch: offset every per-member buffer.
__device__ __forceinline__ void member_offset(DirArgs& a) {
  const long long off = (long long)blockIdx.y * a.mstride;
  a.in += off;
  a.st += off;
  a.out += off;
  a.tmp += off;
  if (a.s1) a.s1 += off;
}

template <int NT>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(6))) void k_dir_x(
    DirArgs a, Grid g, int ntiles, int nreal, int per_xcd) {
  constexpr int C = NT - 12;
  __shared__ XTileLds<NT> sm;
  const int logical = xcd_logical(blockIdx.x, per_xcd);
  if (logical >= nreal) return;  // whole block, before any barrier
  member_offset(a);
  dir_x_tile<NT>(a, g, logical / ntiles, (logical % ntiles) * C, sm);
}

// ---- x, dense tiles: the extended rows of a grid (cells -6 .. nx+5, E = nx+12
// slots each, the periodic cells the three stages consume included) laid end
// to end and cut into blocks of NT slots every NT-12. A row boundary inside a
// block costs no more than the block boundary does, so narrow rows (200, 400
// cells) fill the blocks instead of leaving a partial tile per row. A slot
// holds a correct stage-s value when it lies 2s slots inside its block and
// inside its own extended row (-6+2s <= cell <= nx+5-2s); values computed
// across a row boundary are never written. Block windows [6, NT-6) partition
// the slots, so every interior cell is written by one block; halo cells are
// written wherever they are correct (the same value). Slot -> (row, cell)
// takes one 32-bit division per thread.
inline long long dense_x_slots(int nx, int nz) { return (long long)nz * (nx + 12); }
inline int dense_x_blocks(int nx, int nz, int nt) {
  return (int)((dense_x_slots(nx, nz) - 13) / (nt - 12)) + 1;
}

template <int NT>
__device__ __forceinline__ void dir_xd_tile(const DirArgs& a, const Grid& g, int b,
                                            XTileLds<NT>& sm) {
  auto& s0 = sm.s0;
  auto& s1 = sm.s1;
  auto& fl = sm.fl;
  const int nx = g.nx, nz = g.nz, t = threadIdx.x;
  const unsigned E = (unsigned)nx + 12u;
  const unsigned gs = (unsigned)b * (unsigned)(NT - 12) + (unsigned)t;
  const int r = (int)(gs / E);
  const int cell = (int)(gs - (unsigned)r * E) - 6;  // -6 .. nx+5
  const bool row_ok = r < nz;
  const int prow = (row_ok ? r : 0) + kHs;
  const int cw = wrap_cell(cell, nx);
  const bool own = row_ok && t >= 6 && t < NT - 6 && cell >= 0 && cell < nx;
  const bool halo = row_ok && (cell == -2 || cell == -1 || cell == nx || cell == nx + 1);
  {
    double q[4] = {0.0, 0.0, 0.0, 0.0};
    if (row_ok) {
#pragma unroll
      for (int v = 0; v < kNv; ++v) q[v] = a.in[g.at(v, prow, cw + kHs)];
    }
#pragma unroll
    for (int v = 0; v < kNv; ++v) s0[v][t] = q[v];
    if (halo) {
#pragma unroll
      for (int v = 0; v < kNv; ++v) a.st[g.at(v, prow, cell + kHs)] = q[v];
    }
  }
  // face j (between slots j-1 and j) is computed by thread j with its row's
  // hydrostatic values; faces across a row boundary feed no correct cell
  const double hyr = a.hy_cell[prow], hyt = a.hyt_cell[prow];
  const double src = (a.src_w && row_ok) ? a.src_w[(long long)r * nx + cw] : 0.0;
  const double id = a.id;
  PowRef pref;  // this face's stage-1 pressure power, continued in stages 2 and 3
  __syncthreads();
#pragma unroll
  for (int stage = 0; stage < 3; ++stage) {
    const int m = 2 + 2 * stage;  // faces m..NT-m, cells m..NT-1-m
    auto& src_s = stage == 0 ? s0 : s1;
    bool ok = true;
    if (t >= m && t <= NT - m) {
      double v4[4], d3[4], F[4];
#pragma unroll
      for (int v = 0; v < kNv; ++v)
        recon4(src_s[v][t - 2], src_s[v][t - 1], src_s[v][t], src_s[v][t + 1], v4[v], d3[v]);
      flux_x_face(v4, d3, hyr, hyt, a.hv, F, stage == 0 ? POW_FIRST : POW_TRY, &pref, &ok);
#pragma unroll
      for (int v = 0; v < kNv; ++v) fl[v][t] = F[v];
    }
    if (!ok) {  // the face again with pow_pos (rare)
      double v4[4], d3[4], F[4];
#pragma unroll
      for (int v = 0; v < kNv; ++v)
        recon4(src_s[v][t - 2], src_s[v][t - 1], src_s[v][t], src_s[v][t + 1], v4[v], d3[v]);
      flux_x_face(v4, d3, hyr, hyt, a.hv, F);
#pragma unroll
      for (int v = 0; v < kNv; ++v) fl[v][t] = F[v];
    }
    __syncthreads();
    if (t >= m && t <= NT - 1 - m) {
      const double dts = stage == 0 ? a.dt1 : stage == 1 ? a.dt2 : a.dt3;
      double q[4];
#pragma unroll
      for (int v = 0; v < kNv; ++v) q[v] = tend_of(fl[v][t], fl[v][t + 1], a.d, id);
      if (a.src_w) q[2] = q[2] + src;
#pragma unroll
      for (int v = 0; v < kNv; ++v) q[v] = s0[v][t] + dts * q[v];
      if (stage < 2) {
#pragma unroll
        for (int v = 0; v < kNv; ++v) s1[v][t] = q[v];  // in place: faces are done
      }
      if (stage == 0 && a.s1 && own) {
#pragma unroll
        for (int v = 0; v < kNv; ++v) a.s1[g.at(v, prow, cell + kHs)] = q[v];
      }
      if (stage == 1 && ((own && a.tmp_interior) || halo)) {
#pragma unroll
        for (int v = 0; v < kNv; ++v) a.tmp[g.at(v, prow, cell + kHs)] = q[v];
      }
      if (stage == 2 && own) {
#pragma unroll
        for (int v = 0; v < kNv; ++v) a.out[g.at(v, prow, cell + kHs)] = q[v];
      }
    }
    if (stage < 2) __syncthreads();
  }
}

template <int NT>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(5))) void k_dir_xd(
    DirArgs a, Grid g, int nreal, int per_xcd) {
  __shared__ XTileLds<NT> sm;
  const int logical = xcd_logical(blockIdx.x, per_xcd);
  if (logical >= nreal) return;  // whole block, before any barrier
  member_offset(a);
  dir_xd_tile<NT>(a, g, logical, sm);
}

// ------------------------------------------------------------------ z ------
// Face k of a z stage from the window rows k-2..k+1 (w[0..3]).
__device__ __forceinline__ void z_face(const double w[4][4], const DirArgs& a, int k,
                                       int nz, double G[4]) {
  double v4[4], d3[4];
#pragma unroll
  for (int v = 0; v < kNv; ++v) recon4(w[0][v], w[1][v], w[2][v], w[3][v], v4[v], d3[v]);
  flux_z_face(v4, d3, a.hyr_int[k], a.hyt_int[k], a.hyp_int[k], a.hv, k == 0 || k == nz, G);
}

__device__ __forceinline__ void shift4(double w[4][4]) {
#pragma unroll
  for (int p = 0; p < 3; ++p)
#pragma unroll
    for (int v = 0; v < kNv; ++v) w[p][v] = w[p + 1][v];
}

// Halo row values of set_bc_z from an edge row e (padded row pe) into halo
// padded row ph.
__device__ __forceinline__ void wall_row(const double e[4], const double* hy, int pe,
                                         int ph, double out[4]) {
#pragma unroll
  for (int v = 0; v < kNv; ++v) out[v] = wall_halo(v, e[v], hy[pe], hy[ph]);
}

// One step of one z stage for one lane. The window slot of input row
// base + p (p = 0..3) is (PH + 1 + p) & 3 where PH = step & 3, so advancing a
// row changes the phase instead of copying registers. Input row r arrives,
// face r-1 and cell r-2 are finished. Returns true and the cell's new value
// in `res` when a cell was finished.
template <int PH, int stage>
__device__ __forceinline__ bool z_stage_step(double w[4][4], double F[4], const DirArgs& a,
                                             const Grid& g, int r, int lo,
                                             int hi, int pc, int i, const double* lds_row,
                                             const double cur[4], double dts, double res[4]) {
  constexpr int S0 = (PH + 1) & 3, S1 = (PH + 2) & 3, S2 = (PH + 3) & 3, S3 = PH & 3;
  const int nz = g.nz;
  const gdouble* hy = a.hy_cell;
  // the new row is formed in nv and stored once, so that every window slot
  // is written through a static index (no merged private-array stores)
  if (r >= lo - 2 && r <= hi + 1 && r < nz + 2) {
    double nv[4];
    if (r >= 0 && r < nz) {
      if (stage == 0) {
#pragma unroll
        for (int v = 0; v < kNv; ++v) nv[v] = cur[v];  // row r, loaded a step ahead
      } else {
#pragma unroll
        for (int v = 0; v < kNv; ++v) nv[v] = lds_row[v * 64];
      }
    } else if (r == nz) {
      wall_row(w[S2], hy, nz + 1, nz + 2, nv);
    } else {  // r == nz + 1, or r < 0 (rows -2, -1: refreshed when row 0 arrives)
      wall_row(w[S1], hy, nz + 1, nz + 3, nv);
    }
#pragma unroll
    for (int v = 0; v < kNv; ++v) w[S3][v] = nv[v];
    if (r == 0) {
      double h1[4], h2[4];
      wall_row(nv, hy, kHs, 0, h1);
      wall_row(nv, hy, kHs, 1, h2);
#pragma unroll
      for (int v = 0; v < kNv; ++v) {
        w[S1][v] = h1[v];
        w[S2][v] = h2[v];
      }
    }
  }
  const int f = r - 1;
  bool done = false;
  // RK init of the cell this step may finish (row r-2; clamped so the load
  // is always in bounds), issued before the face so its latency overlaps it
  double iv[4];
  if (stage > 0) {
    const int cc = min(max(r - 2, 0), nz - 1);
#pragma unroll
    for (int v = 0; v < kNv; ++v) iv[v] = a.in[g.at(v, cc + kHs, pc)];
  }
  if (f >= lo && f <= hi) {
    double v4[4], d3[4], G[4];
#pragma unroll
    for (int v = 0; v < kNv; ++v)
      recon4(w[S0][v], w[S1][v], w[S2][v], w[S3][v], v4[v], d3[v]);
    flux_z_face(v4, d3, a.hyr_int[f], a.hyt_int[f], a.hyp_int[f], a.hv, f == 0 || f == nz, G);
    const int c = f - 1;  // window row 1
    if (c >= lo) {
      const double id = a.id;
      double td[4];
#pragma unroll
      for (int v = 0; v < kNv; ++v) td[v] = tend_of(F[v], G[v], a.d, id);
      td[2] = td[2] - w[S1][0] * kGrav;
      if (a.src_w) td[2] = td[2] + a.src_w[(long long)c * g.nx + i];
      if (stage == 0) {
#pragma unroll
        for (int v = 0; v < kNv; ++v) res[v] = w[S1][v] + dts * td[v];
      } else {
#pragma unroll
        for (int v = 0; v < kNv; ++v) res[v] = iv[v] + dts * td[v];
      }
      done = true;
    }
#pragma unroll
    for (int v = 0; v < kNv; ++v) F[v] = G[v];
  }
  return done;
}

__device__ __forceinline__ void store_row(gdouble* buf, const Grid& g, int prow, int pc,
                                          const double q[4]) {
#pragma unroll
  for (int v = 0; v < kNv; ++v) buf[g.at(v, prow, pc)] = q[v];
}

// Body of one step j for the wave of `stage` (0, 1, 2 = RK stage 1, 2, 3).
template <int PH, int stage>
__device__ __forceinline__ void z_block_step(double w[4][4], double F[4], const DirArgs& a,
                                             const Grid& g, double (*ring)[2][kNv][64],
                                             int j, int lo, int hi, int k0,
                                             int k1, int pc, int i, int lane, bool active,
                                             const double cur[4], double dts) {
  const int nz = g.nz;
  const gdouble* hy = a.hy_cell;
  const int r = j - 3 * stage;  // input row of this stage at step j
  if (active) {
    // rows of S1 / S2 come from the previous stage's ring (written one step ago)
    const double* lds_row = stage > 0 ? &ring[stage - 1][r & 1][0][lane] : nullptr;
    double res[4];
    const bool done =
        z_stage_step<PH, stage>(w, F, a, g, r, lo, hi, pc, i, lds_row, cur, dts, res);
    if (stage == 0 && r == 0 && k0 == 0) {  // set_bc_z of stage 1 into st
      double h[4];
      wall_row(w[PH & 3], hy, kHs, 0, h);
      store_row(a.st, g, 0, pc, h);
      wall_row(w[PH & 3], hy, kHs, 1, h);
      store_row(a.st, g, 1, pc, h);
    }
    if (stage == 0 && r == nz - 1 && k1 == nz && r >= lo - 2) {
      double h[4];
      wall_row(w[PH & 3], hy, nz + 1, nz + 2, h);
      store_row(a.st, g, nz + 2, pc, h);
      wall_row(w[PH & 3], hy, nz + 1, nz + 3, h);
      store_row(a.st, g, nz + 3, pc, h);
    }
    if (done) {
      const int c = r - 2;
      if (stage < 2) {
#pragma unroll
        for (int v = 0; v < kNv; ++v) ring[stage][c & 1][v][lane] = res[v];
      }
      if (stage == 0 && a.s1 && c >= k0 && c < k1) store_row(a.s1, g, c + kHs, pc, res);
      if (stage == 1) {
        if (a.tmp_interior && c >= k0 && c < k1) store_row(a.tmp, g, c + kHs, pc, res);
        if (c == 0 && k0 == 0) {  // set_bc_z of stage 3 into tmp
          double h[4];
          wall_row(res, hy, kHs, 0, h);
          store_row(a.tmp, g, 0, pc, h);
          wall_row(res, hy, kHs, 1, h);
          store_row(a.tmp, g, 1, pc, h);
        }
        if (c == nz - 1 && k1 == nz) {
          double h[4];
          wall_row(res, hy, nz + 1, nz + 2, h);
          store_row(a.tmp, g, nz + 2, pc, h);
          wall_row(res, hy, nz + 1, nz + 3, h);
          store_row(a.tmp, g, nz + 3, pc, h);
        }
      }
      if (stage == 2) store_row(a.out, g, c + kHs, pc, res);
    }
  }
  __syncthreads();
}

// March of one stage's wave over steps j0..j1 (one barrier per step; all
// three waves of the block run the same number of steps).
template <int stage>
__device__ __forceinline__ void z_march(const DirArgs& a, const Grid& g,
                                        double (*ring)[2][kNv][64], int j0, int j1, int lo,
                                        int hi, int k0, int k1, int pc, int i, int lane,
                                        bool active, double dts) {
  double w[4][4], F[4];
#pragma unroll
  for (int p = 0; p < 4; ++p)
#pragma unroll
    for (int v = 0; v < kNv; ++v) w[p][v] = 0.0;
#pragma unroll
  for (int v = 0; v < kNv; ++v) F[v] = 0.0;
  // stage 0 reads S row j at step j: keep the next row's load in flight
  const int nz = g.nz;
  double pre[4] = {0.0, 0.0, 0.0, 0.0};
  if (stage == 0 && active) {
    const int rr = min(max(j0, 0), nz - 1);
#pragma unroll
    for (int v = 0; v < kNv; ++v) pre[v] = a.in[g.at(v, rr + kHs, pc)];
  }
#pragma unroll 1
  for (int j = j0; j <= j1; ++j) {
    // slot p holds row r-3+p once the step's row has arrived in slot 3
#pragma unroll
    for (int p = 0; p < 3; ++p)
#pragma unroll
      for (int v = 0; v < kNv; ++v) w[p][v] = w[p + 1][v];
    double cur[4];
#pragma unroll
    for (int v = 0; v < kNv; ++v) cur[v] = pre[v];
    if (stage == 0 && active) {
      const int rr = min(max(j + 1, 0), nz - 1);
#pragma unroll
      for (int v = 0; v < kNv; ++v) pre[v] = a.in[g.at(v, rr + kHs, pc)];
    }
    z_block_step<3, stage>(w, F, a, g, ring, j, lo, hi, k0, k1, pc, i, lane, active, cur, dts);
  }
}

// z direction: a block is 3 waves over the same 64 columns and segment of L
// output rows; wave s runs RK stage s+1. Stage s+1 consumes the rows stage s
// produced one step earlier through a 2-slot LDS ring, so it lags by 3 rows
// (2 for the stencil, 1 for the hand-over); one barrier per step.
// One z block: 64 columns starting at column 64*colgrp, output rows of
// segment `seg` (L rows). Called by all 192 threads (contains barriers).
template <int L>
__device__ __forceinline__ void dir_z_block(const DirArgs& a, const Grid& g, int colgrp,
                                            int seg, double (*ring)[2][kNv][64]) {
  const int nx = g.nx, nz = g.nz;
  const int lane = threadIdx.x & 63;
  // wave-uniform; readfirstlane makes that visible to the compiler so the
  // stage's bounds and branches are scalar
  const int stage = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int i = colgrp * 64 + lane;  // interior column
  const int k0 = seg * L, k1 = min(k0 + L, nz);
  const gdouble* hy = a.hy_cell;
  if (stage == 0 && i >= nx && i < nx + 4) {
    // corner columns of the halo rows (padded 0, 1, nx+2, nx+3) of st and tmp
    const int q = i - nx;
    const int cc = q < 2 ? q : nx + q;
    const int lo = kHs, hi = nz + kHs - 1;
    for (int pass = 0; pass < 2; ++pass) {
      gdouble* buf = pass == 0 ? a.st : a.tmp;
      if (k0 == 0) {
#pragma unroll
        for (int v = 0; v < kNv; ++v) {
          const double e = buf[g.at(v, lo, cc)];
          buf[g.at(v, 0, cc)] = wall_halo(v, e, hy[lo], hy[0]);
          buf[g.at(v, 1, cc)] = wall_halo(v, e, hy[lo], hy[1]);
        }
      }
      if (k1 == nz) {
#pragma unroll
        for (int v = 0; v < kNv; ++v) {
          const double e = buf[g.at(v, hi, cc)];
          buf[g.at(v, nz + 2, cc)] = wall_halo(v, e, hy[hi], hy[nz + 2]);
          buf[g.at(v, nz + 3, cc)] = wall_halo(v, e, hy[hi], hy[nz + 3]);
        }
      }
    }
  }
  const bool active = i < nx;
  const int pc = i + kHs;
  // output rows of each stage: [a1,b1) S1, [a2,b2) S2, [k0,k1) result
  const int a1 = max(k0 - 4, 0), b1 = min(k1 + 4, nz);
  const int a2 = max(k0 - 2, 0), b2 = min(k1 + 2, nz);
  const int lo = stage == 0 ? a1 : stage == 1 ? a2 : k0;
  const int hi = stage == 0 ? b1 : stage == 1 ? b2 : k1;
  if (stage == 0)
    z_march<0>(a, g, ring, a1 - 2, k1 + 7, lo, hi, k0, k1, pc, i, lane, active, a.dt1);
  else if (stage == 1)
    z_march<1>(a, g, ring, a1 - 2, k1 + 7, lo, hi, k0, k1, pc, i, lane, active, a.dt2);
  else
    z_march<2>(a, g, ring, a1 - 2, k1 + 7, lo, hi, k0, k1, pc, i, lane, active, a.dt3);
}

template <int L>
__global__ __launch_bounds__(192) void k_dir_z(DirArgs a, Grid g, int ncolblk, int nreal,
                                               int per_xcd) {
  __shared__ double ring[2][2][kNv][64];
  const int logical = xcd_logical(blockIdx.x, per_xcd);
  if (logical >= nreal) return;  // whole block, before any barrier
  member_offset(a);
  dir_z_block<L>(a, g, logical % ncolblk, logical / ncolblk, ring);
}

// ------------------------------------------------------------- z tiles --
// LDS of a one-cell-per-thread z tile: the x tile's arrays and the tile
// rows' face hydrostatics (rho, rho*theta, p at the lower face of each tile
// row), read from HBM once per tile; all CW columns of a row read one word.
template <int NT, int NR>
struct ZTileLds : XTileLds<NT> {
  double hz[3][NR];
};

// z direction as LDS tiles, the x kernel's scheme turned on its side: a block
// of NT threads holds CW columns x NR = NT/CW rows (slot t = row*CW + col,
// tile row j = logical row r0-6+j), runs the three stages in LDS (stage 1 on
// tile rows 2..NR-3, stage 2 on 4..NR-5, stage 3 on the NR-12 output rows)
// and rebuilds the wall halo rows of every stage state from its edge row, as
// set_bc_z does between the stages (ics/bcs.py:64-148). Same operations in
// the same order as the marching kernel, so results are bitwise equal.
template <int CW, int NT, bool CONT = true>
__device__ __forceinline__ void dir_zt_tile(const DirArgs& a, const Grid& g, int c0, int r0,
                                            ZTileLds<NT, NT / CW>& sm, bool corners) {
  constexpr int NR = NT / CW, H = 6, C = NR - 2 * H;
  auto& s0 = sm.s0;
  auto& s1 = sm.s1;
  auto& fl = sm.fl;
  auto& hz = sm.hz;
  const int nx = g.nx, nz = g.nz, t = threadIdx.x;
  const int j = t / CW, col = t % CW;
  const int i = c0 + col;     // interior column
  const int k = r0 - H + j;   // logical row of this slot
  const bool colok = i < nx;
  const int pc = i + kHs;
  const gdouble* hy = a.hy_cell;
  const int plo = kHs, phi = nz + kHs - 1;  // padded edge rows
  const bool wlo = k == -2 || k == -1, whi = k == nz || k == nz + 1;
  const bool own = colok && k >= r0 && k < r0 + C && k < nz;
  const bool own_lo = r0 == 0, own_hi = r0 <= nz - 1 && nz - 1 < r0 + C;
  gdouble* const tmp_own = a.tmp_interior ? a.tmp : nullptr;
  static_assert(NT >= NR, "one thread per tile row fills the face hydrostatics");
  if (t < NR) {
    const int kr = r0 - H + t;
    const bool f = kr >= 0 && kr <= nz;
    hz[0][t] = f ? a.hyr_int[kr] : 0.0;
    hz[1][t] = f ? a.hyt_int[kr] : 0.0;
    hz[2][t] = f ? a.hyp_int[kr] : 0.0;
  }
  {
    double q[4] = {0.0, 0.0, 0.0, 0.0};
    if (colok) {
      if (k >= 0 && k < nz) {
#pragma unroll
        for (int v = 0; v < kNv; ++v) q[v] = a.in[g.at(v, k + kHs, pc)];
      } else if (wlo || whi) {
        const int pe = wlo ? plo : phi;
#pragma unroll
        for (int v = 0; v < kNv; ++v) q[v] = wall_halo(v, a.in[g.at(v, pe, pc)], hy[pe], hy[k + kHs]);
        if ((wlo && own_lo) || (whi && own_hi)) {  // set_bc_z of stage 1 into st
#pragma unroll
          for (int v = 0; v < kNv; ++v) a.st[g.at(v, k + kHs, pc)] = q[v];
        }
      }
    }
#pragma unroll
    for (int v = 0; v < kNv; ++v) s0[v][t] = q[v];
  }
  if (corners && t < 8) {
    // corner columns (padded 0, 1, nx+2, nx+3) of the halo rows of st, tmp
    gdouble* buf = t < 4 ? a.st : a.tmp;
    const int qq = t & 3;
    const int cc = qq < 2 ? qq : nx + qq;
#pragma unroll
    for (int v = 0; v < kNv; ++v) {
      const double e = buf[g.at(v, plo, cc)];
      buf[g.at(v, 0, cc)] = wall_halo(v, e, hy[plo], hy[0]);
      buf[g.at(v, 1, cc)] = wall_halo(v, e, hy[plo], hy[1]);
      const double f = buf[g.at(v, phi, cc)];
      buf[g.at(v, nz + 2, cc)] = wall_halo(v, f, hy[phi], hy[nz + 2]);
      buf[g.at(v, nz + 3, cc)] = wall_halo(v, f, hy[phi], hy[nz + 3]);
    }
  }
  const double id = a.id;
  const double src = (a.src_w && colok && k >= 0 && k < nz) ? a.src_w[(long long)k * nx + i] : 0.0;
  PowRef pref;  // CONT: this face's stage-1 power, continued in stages 2 and 3
  __syncthreads();
#pragma unroll
  for (int stage = 0; stage < 3; ++stage) {
    double (*sin)[NT] = stage == 0 ? s0 : s1;
    const int e = 2 + 2 * stage;
    // faces: tile rows e .. NR-e (face k between logical rows k-1 and k)
    bool ok = true;
    if (colok && j >= e && j <= NR - e && k >= 0 && k <= nz) {
      double v4[4], d3[4], G[4];
#pragma unroll
      for (int v = 0; v < kNv; ++v)
        recon4(sin[v][t - 2 * CW], sin[v][t - CW], sin[v][t], sin[v][t + CW], v4[v], d3[v]);
      flux_z_face(v4, d3, hz[0][j], hz[1][j], hz[2][j], a.hv, k == 0 || k == nz, G,
                  !CONT ? POW_PLAIN : stage == 0 ? POW_FIRST : POW_TRY, &pref, &ok);
#pragma unroll
      for (int v = 0; v < kNv; ++v) fl[v][t] = G[v];
    }
    if (CONT && !ok) {  // the face again with pow_pos (rare)
      double v4[4], d3[4], G[4];
#pragma unroll
      for (int v = 0; v < kNv; ++v)
        recon4(sin[v][t - 2 * CW], sin[v][t - CW], sin[v][t], sin[v][t + CW], v4[v], d3[v]);
      flux_z_face(v4, d3, hz[0][j], hz[1][j], hz[2][j], a.hv, k == 0 || k == nz, G);
#pragma unroll
      for (int v = 0; v < kNv; ++v) fl[v][t] = G[v];
    }
    __syncthreads();
    // cells: tile rows e .. NR-e-1
    if (colok && j >= e && j <= NR - e - 1 && k >= 0 && k < nz) {
      const double dts = stage == 0 ? a.dt1 : stage == 1 ? a.dt2 : a.dt3;
      double td[4];
#pragma unroll
      for (int v = 0; v < kNv; ++v) td[v] = tend_of(fl[v][t], fl[v][t + CW], a.d, id);
      td[2] = td[2] - sin[0][t] * kGrav;
      if (a.src_w) td[2] = td[2] + src;
      double q[4];
#pragma unroll
      for (int v = 0; v < kNv; ++v) q[v] = s0[v][t] + dts * td[v];
      if (stage < 2) {
#pragma unroll
        for (int v = 0; v < kNv; ++v) s1[v][t] = q[v];
        // The wall halo rows of the new stage state, written by the edge
        // row's thread: no thread reads them in this phase (an update reads
        // its own cell only), and a tile holding a wall has its edge row in
        // the window, updated at every stage.
#pragma unroll
        for (int side = 0; side < 2; ++side) {
          if (k == (side == 0 ? 0 : nz - 1)) {
            const int pe = side == 0 ? plo : phi;
            const int step = side == 0 ? -CW : CW;
#pragma unroll
            for (int h = 1; h <= 2; ++h) {
              const int kh = side == 0 ? -h : nz - 1 + h;  // the halo row's k
              double w[4];
#pragma unroll
              for (int v = 0; v < kNv; ++v) w[v] = wall_halo(v, q[v], hy[pe], hy[kh + kHs]);
#pragma unroll
              for (int v = 0; v < kNv; ++v) s1[v][t + h * step] = w[v];
              if (stage == 1 && (side == 0 ? own_lo : own_hi)) {  // set_bc_z of stage 3
#pragma unroll
                for (int v = 0; v < kNv; ++v) a.tmp[g.at(v, kh + kHs, pc)] = w[v];
              }
            }
          }
        }
      }
      if (own) {
        gdouble* dst = stage == 0 ? a.s1 : stage == 1 ? tmp_own : a.out;
        if (dst) {
#pragma unroll
          for (int v = 0; v < kNv; ++v) dst[g.at(v, k + kHs, pc)] = q[v];
        }
      }
    }
    if (stage == 2) break;
    __syncthreads();
  }
}

// 6 waves per SIMD (the LDS limit of a 512-thread tile): left alone the
// compiler allocates ~106 VGPRs (4 waves); asked, it fits 72 without spills.
template <int CW, int NT>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(6))) void k_dir_zt(
    DirArgs a, Grid g, int ncolgrp, int nreal, int per_xcd) {
  constexpr int C = NT / CW - 12;
  __shared__ ZTileLds<NT, NT / CW> sm;
  const int logical = xcd_logical(blockIdx.x, per_xcd);
  if (logical >= nreal) return;  // whole block, before any barrier
  member_offset(a);
  dir_zt_tile<CW, NT>(a, g, (logical % ncolgrp) * CW, (logical / ncolgrp) * C, sm,
                      logical == 0);
}

// z tiles of CW columns x NR rows with two cells per thread (tile rows j
// and j + NR/2 of one column) and the stage state updated in place: LDS holds
// the running stage state and the face fluxes only (the RK init S of a
// thread's two cells stays in registers), 64 B per cell, so an 8 x 128 tile
// fits in 64 KB. Same operations in the same order as dir_zt_tile.
template <int NC>
struct ZTile2Lds {
  double s[kNv][NC];
  double fl[kNv][NC];
};

// The same with the tile rows' face hydrostatics (rho, rho*theta, p at the
// lower face of each tile row), read from HBM once per tile instead of once
// per stage and cell; all CW columns of a row read one LDS word (broadcast).
template <int CW, int NR>
struct ZTile2HLds : ZTile2Lds<CW * NR> {
  double hz[3][NR];
  int wf[CW * NR / 64 + 2];  // wave phase flags (one-cell tiles)
};

// CPT cells per thread: 2 (tile rows j and j + NR/2) or 1. WSYNC (one cell
// per thread, 8 columns, so that a wave holds 8 whole tile rows): the phases
// are ordered by wave flags instead of block barriers; every phase touches
// only the wave's rows and 2 rows on either side.
template <int CW, int NR, int CPT = 2, bool WSYNC = false>
__device__ __forceinline__ void dir_zt2_tile(const DirArgs& a, const Grid& g, int c0, int r0,
                                             ZTile2HLds<CW, NR>& sm, bool corners) {
  constexpr int NC = CW * NR, NT = NC / CPT, H = 6, C = NR - 2 * H;
  auto& s = sm.s;
  auto& fl = sm.fl;
  auto& hz = sm.hz;
  const int nx = g.nx, nz = g.nz, t = threadIdx.x;
  const int col = t % CW;
  const int i = c0 + col;
  const bool colok = i < nx;
  const int pc = i + kHs;
  const gdouble* hy = a.hy_cell;
  const int plo = kHs, phi = nz + kHs - 1;
  const bool own_lo = r0 == 0, own_hi = r0 <= nz - 1 && nz - 1 < r0 + C;
  gdouble* const tmp_own = a.tmp_interior ? a.tmp : nullptr;
  static_assert(NT >= NR, "one thread per tile row fills the face hydrostatics");
  static_assert(!WSYNC || (CPT == 1 && CW == 8 && NT % 64 == 0), "waves of 8 whole rows");
  if (WSYNC) wave_flags_init(sm.wf, NT / 64, t);
  // between phases: block barrier, or the neighbouring waves' flags
  auto phase_done = [&](int P) {
    if constexpr (WSYNC)
      wave_phase_done(sm.wf, P, t);
    else
      __syncthreads();
  };
  if (t < NR) {
    const int k = r0 - H + t;
    const bool f = k >= 0 && k <= nz;
    hz[0][t] = f ? a.hyr_int[k] : 0.0;
    hz[1][t] = f ? a.hyt_int[k] : 0.0;
    hz[2][t] = f ? a.hyp_int[k] : 0.0;
  }
  int jj[CPT], kk[CPT], sl[CPT];
  bool wl[CPT], wh[CPT], own[CPT];
  double S[CPT][4], src[CPT];
#pragma unroll
  for (int c = 0; c < CPT; ++c) {
    sl[c] = t + c * NT;
    jj[c] = sl[c] / CW;
    kk[c] = r0 - H + jj[c];
    wl[c] = kk[c] == -2 || kk[c] == -1;
    wh[c] = kk[c] == nz || kk[c] == nz + 1;
    own[c] = colok && kk[c] >= r0 && kk[c] < r0 + C && kk[c] < nz;
    double q[4] = {0.0, 0.0, 0.0, 0.0};
    if (colok) {
      if (kk[c] >= 0 && kk[c] < nz) {
#pragma unroll
        for (int v = 0; v < kNv; ++v) q[v] = a.in[g.at(v, kk[c] + kHs, pc)];
      } else if (wl[c] || wh[c]) {
        const int pe = wl[c] ? plo : phi;
#pragma unroll
        for (int v = 0; v < kNv; ++v)
          q[v] = wall_halo(v, a.in[g.at(v, pe, pc)], hy[pe], hy[kk[c] + kHs]);
        if ((wl[c] && own_lo) || (wh[c] && own_hi)) {  // set_bc_z of stage 1 into st
#pragma unroll
          for (int v = 0; v < kNv; ++v) a.st[g.at(v, kk[c] + kHs, pc)] = q[v];
        }
      }
    }
#pragma unroll
    for (int v = 0; v < kNv; ++v) {
      s[v][sl[c]] = q[v];
      S[c][v] = q[v];
    }
    src[c] = (a.src_w && colok && kk[c] >= 0 && kk[c] < nz)
                 ? a.src_w[(long long)kk[c] * nx + i] : 0.0;
  }
  if (corners && t < 8) {
    // corner columns (padded 0, 1, nx+2, nx+3) of the halo rows of st, tmp
    gdouble* buf = t < 4 ? a.st : a.tmp;
    const int qq = t & 3;
    const int cc = qq < 2 ? qq : nx + qq;
#pragma unroll
    for (int v = 0; v < kNv; ++v) {
      const double e = buf[g.at(v, plo, cc)];
      buf[g.at(v, 0, cc)] = wall_halo(v, e, hy[plo], hy[0]);
      buf[g.at(v, 1, cc)] = wall_halo(v, e, hy[plo], hy[1]);
      const double f = buf[g.at(v, phi, cc)];
      buf[g.at(v, nz + 2, cc)] = wall_halo(v, f, hy[phi], hy[nz + 2]);
      buf[g.at(v, nz + 3, cc)] = wall_halo(v, f, hy[phi], hy[nz + 3]);
    }
  }
  const double id = a.id;
  PowRef pref[CPT];  // each thread computes the faces at its slots in every stage
  __syncthreads();
#pragma unroll
  for (int stage = 0; stage < 3; ++stage) {
    const int e = 2 + 2 * stage;
    bool ok[CPT];
#pragma unroll
    for (int c = 0; c < CPT; ++c) ok[c] = true;
#pragma unroll
    for (int c = 0; c < CPT; ++c) {
      const int j = jj[c], k = kk[c], u = sl[c];
      if (colok && j >= e && j <= NR - e && k >= 0 && k <= nz) {
        double v4[4], d3[4], G[4];
#pragma unroll
        for (int v = 0; v < kNv; ++v)
          recon4(s[v][u - 2 * CW], s[v][u - CW], s[v][u], s[v][u + CW], v4[v], d3[v]);
        flux_z_face(v4, d3, hz[0][j], hz[1][j], hz[2][j], a.hv, k == 0 || k == nz,
                    G, stage == 0 ? POW_FIRST : POW_TRY, &pref[c], &ok[c]);
#pragma unroll
        for (int v = 0; v < kNv; ++v) fl[v][u] = G[v];
      }
    }
#pragma unroll
    for (int c = 0; c < CPT; ++c) {
      if (!ok[c]) {  // the face again with pow_pos (rare)
        const int j = jj[c], k = kk[c], u = sl[c];
        double v4[4], d3[4], G[4];
#pragma unroll
        for (int v = 0; v < kNv; ++v)
          recon4(s[v][u - 2 * CW], s[v][u - CW], s[v][u], s[v][u + CW], v4[v], d3[v]);
        flux_z_face(v4, d3, hz[0][j], hz[1][j], hz[2][j], a.hv, k == 0 || k == nz, G);
#pragma unroll
        for (int v = 0; v < kNv; ++v) fl[v][u] = G[v];
      }
    }
    phase_done(2 * stage + 1);
    const double dts = stage == 0 ? a.dt1 : stage == 1 ? a.dt2 : a.dt3;
#pragma unroll
    for (int c = 0; c < CPT; ++c) {
      const int j = jj[c], k = kk[c], u = sl[c];
      if (colok && j >= e && j <= NR - e - 1 && k >= 0 && k < nz) {
        double td[4];
#pragma unroll
        for (int v = 0; v < kNv; ++v) td[v] = tend_of(fl[v][u], fl[v][u + CW], a.d, id);
        td[2] = td[2] - s[0][u] * kGrav;
        if (a.src_w) td[2] = td[2] + src[c];
        double q[4];
#pragma unroll
        for (int v = 0; v < kNv; ++v) q[v] = S[c][v] + dts * td[v];
        if (stage < 2) {
#pragma unroll
          for (int v = 0; v < kNv; ++v) s[v][u] = q[v];  // in place: faces are done
          // The wall halo rows of the new stage state, written by the edge
          // row's thread: no thread reads them in this phase (every update
          // reads its own cell and faces only), and a tile holding a wall
          // has its edge row in the window, updated at every stage.
#pragma unroll
          for (int side = 0; side < 2; ++side) {
            if (k == (side == 0 ? 0 : nz - 1)) {
              const int pe = side == 0 ? plo : phi;
              const int step = side == 0 ? -CW : CW;
#pragma unroll
              for (int h = 1; h <= 2; ++h) {
                const int kh = side == 0 ? -h : nz - 1 + h;  // the halo row's k
                double w[4];
#pragma unroll
                for (int v = 0; v < kNv; ++v) w[v] = wall_halo(v, q[v], hy[pe], hy[kh + kHs]);
#pragma unroll
                for (int v = 0; v < kNv; ++v) s[v][u + h * step] = w[v];
                if (stage == 1 && (side == 0 ? own_lo : own_hi)) {  // set_bc_z of stage 3
#pragma unroll
                  for (int v = 0; v < kNv; ++v) a.tmp[g.at(v, kh + kHs, pc)] = w[v];
                }
              }
            }
          }
        }
        if (own[c]) {
          gdouble* dst = stage == 0 ? a.s1 : stage == 1 ? tmp_own : a.out;
          if (dst) {
#pragma unroll
            for (int v = 0; v < kNv; ++v) dst[g.at(v, k + kHs, pc)] = q[v];
          }
        }
      }
    }
    if (stage == 2) break;
    phase_done(2 * stage + 2);
  }
}

// ---- z, dense tiles: the extended columns of the CW-column groups (rows
// -2 .. nz+1: the wall halo rows are rebuilt from the edge row at every
// stage, so a column needs no more) laid end to end, slot row g of block b
// being b*(NR-12) + j - 4, cut into tiles of NR rows; the offset puts row 0
// of the first group at the first window row. A slot row is (group, row);
// the two cells of a thread may lie in different groups. Windows [6, NR-6)
// partition the interior rows; the wall halo rows written to st and tmp
// belong to the block whose window holds their edge row. A halo row whose
// edge row lies outside the block is left unrebuilt: no correct cell of the
// block reaches it.
inline long long dense_z_rows(int nx, int nz, int cw) {
  return (long long)((nx + cw - 1) / cw) * (nz + 4);
}
inline int dense_z_blocks(int nx, int nz, int cw, int nr) {
  return (int)((dense_z_rows(nx, nz, cw) - 5) / (nr - 12)) + 1;
}

template <int CW, int NR>
__device__ __forceinline__ void dir_zt2d_tile(const DirArgs& a, const Grid& g, int b,
                                              ZTile2HLds<CW, NR>& sm, bool corners) {
  constexpr int NC = CW * NR, NT = NC / 2, C = NR - 12;
  auto& s = sm.s;
  auto& fl = sm.fl;
  auto& hz = sm.hz;
  const int nx = g.nx, nz = g.nz, t = threadIdx.x;
  const int Ez = nz + 4;
  static_assert(NT >= NR, "one thread per tile row fills the face hydrostatics");
  if (t < NR) {  // a tile row is one row k of one column group
    const int gr = b * C + t - 4;
    const int k = gr >= 0 ? gr % Ez - 2 : -3;
    const bool f = k >= 0 && k <= nz;
    hz[0][t] = f ? a.hyr_int[k] : 0.0;
    hz[1][t] = f ? a.hyt_int[k] : 0.0;
    hz[2][t] = f ? a.hyp_int[k] : 0.0;
  }
  const int col = t % CW;
  const gdouble* hy = a.hy_cell;
  const int plo = kHs, phi = nz + kHs - 1;
  int jj[2], kk[2], sl[2], es[2], pc[2], ii[2];
  bool colok[2], wl[2], wh[2], own[2], ownh[2], esok[2];
  double S[2][4], src[2];
#pragma unroll
  for (int c = 0; c < 2; ++c) {
    sl[c] = t + c * NT;
    jj[c] = sl[c] / CW;
    const int gr = b * C + jj[c] - 4;
    const int grp = gr >= 0 ? gr / Ez : -1;
    kk[c] = gr >= 0 ? gr - grp * Ez - 2 : -3;  // -3: padding before the first group
    ii[c] = grp * CW + col;
    colok[c] = gr >= 0 && ii[c] < nx;
    pc[c] = (colok[c] ? ii[c] : 0) + kHs;
    wl[c] = kk[c] == -2 || kk[c] == -1;
    wh[c] = kk[c] == nz || kk[c] == nz + 1;
    own[c] = colok[c] && jj[c] >= 6 && jj[c] < NR - 6 && kk[c] >= 0 && kk[c] < nz;
    const int je = wl[c] ? jj[c] - kk[c] : jj[c] - (kk[c] - (nz - 1));  // edge row's tile row
    esok[c] = (wl[c] || wh[c]) && je >= 0 && je < NR;
    es[c] = esok[c] ? je * CW + col : 0;
    ownh[c] = colok[c] && (wl[c] || wh[c]) && je >= 6 && je < NR - 6;
    double q[4] = {0.0, 0.0, 0.0, 0.0};
    if (colok[c]) {
      if (kk[c] >= 0 && kk[c] < nz) {
#pragma unroll
        for (int v = 0; v < kNv; ++v) q[v] = a.in[g.at(v, kk[c] + kHs, pc[c])];
      } else if (wl[c] || wh[c]) {
        const int pe = wl[c] ? plo : phi;
#pragma unroll
        for (int v = 0; v < kNv; ++v)
          q[v] = wall_halo(v, a.in[g.at(v, pe, pc[c])], hy[pe], hy[kk[c] + kHs]);
        if (ownh[c]) {  // set_bc_z of stage 1 into st
#pragma unroll
          for (int v = 0; v < kNv; ++v) a.st[g.at(v, kk[c] + kHs, pc[c])] = q[v];
        }
      }
    }
#pragma unroll
    for (int v = 0; v < kNv; ++v) {
      s[v][sl[c]] = q[v];
      S[c][v] = q[v];
    }
    src[c] = (a.src_w && colok[c] && kk[c] >= 0 && kk[c] < nz)
                 ? a.src_w[(long long)kk[c] * nx + ii[c]] : 0.0;
  }
  if (corners && t < 8) {
    // corner columns (padded 0, 1, nx+2, nx+3) of the halo rows of st, tmp
    gdouble* buf = t < 4 ? a.st : a.tmp;
    const int qq = t & 3;
    const int cc = qq < 2 ? qq : nx + qq;
#pragma unroll
    for (int v = 0; v < kNv; ++v) {
      const double e = buf[g.at(v, plo, cc)];
      buf[g.at(v, 0, cc)] = wall_halo(v, e, hy[plo], hy[0]);
      buf[g.at(v, 1, cc)] = wall_halo(v, e, hy[plo], hy[1]);
      const double f = buf[g.at(v, phi, cc)];
      buf[g.at(v, nz + 2, cc)] = wall_halo(v, f, hy[phi], hy[nz + 2]);
      buf[g.at(v, nz + 3, cc)] = wall_halo(v, f, hy[phi], hy[nz + 3]);
    }
  }
  gdouble* const tmp_own = a.tmp_interior ? a.tmp : nullptr;
  const double id = a.id;
  PowRef pref[2];  // each thread computes the faces at its two slots in every stage
  __syncthreads();
#pragma unroll
  for (int stage = 0; stage < 3; ++stage) {
    const int e = 2 + 2 * stage;
    bool ok[2] = {true, true};
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      const int j = jj[c], k = kk[c], u = sl[c];
      if (colok[c] && j >= e && j <= NR - e && k >= 0 && k <= nz) {
        double v4[4], d3[4], G[4];
#pragma unroll
        for (int v = 0; v < kNv; ++v)
          recon4(s[v][u - 2 * CW], s[v][u - CW], s[v][u], s[v][u + CW], v4[v], d3[v]);
        flux_z_face(v4, d3, hz[0][j], hz[1][j], hz[2][j], a.hv, k == 0 || k == nz,
                    G, stage == 0 ? POW_FIRST : POW_TRY, &pref[c], &ok[c]);
#pragma unroll
        for (int v = 0; v < kNv; ++v) fl[v][u] = G[v];
      }
    }
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      if (!ok[c]) {  // the face again with pow_pos (rare)
        const int j = jj[c], k = kk[c], u = sl[c];
        double v4[4], d3[4], G[4];
#pragma unroll
        for (int v = 0; v < kNv; ++v)
          recon4(s[v][u - 2 * CW], s[v][u - CW], s[v][u], s[v][u + CW], v4[v], d3[v]);
        flux_z_face(v4, d3, hz[0][j], hz[1][j], hz[2][j], a.hv, k == 0 || k == nz, G);
#pragma unroll
        for (int v = 0; v < kNv; ++v) fl[v][u] = G[v];
      }
    }
    __syncthreads();
    const double dts = stage == 0 ? a.dt1 : stage == 1 ? a.dt2 : a.dt3;
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      const int j = jj[c], k = kk[c], u = sl[c];
      if (colok[c] && j >= e && j <= NR - e - 1 && k >= 0 && k < nz) {
        double td[4];
#pragma unroll
        for (int v = 0; v < kNv; ++v) td[v] = tend_of(fl[v][u], fl[v][u + CW], a.d, id);
        td[2] = td[2] - s[0][u] * kGrav;
        if (a.src_w) td[2] = td[2] + src[c];
        double q[4];
#pragma unroll
        for (int v = 0; v < kNv; ++v) q[v] = S[c][v] + dts * td[v];
        if (stage < 2) {
#pragma unroll
          for (int v = 0; v < kNv; ++v) s[v][u] = q[v];  // in place: faces are done
        }
        if (own[c]) {
          gdouble* dst = stage == 0 ? a.s1 : stage == 1 ? tmp_own : a.out;
          if (dst) {
#pragma unroll
            for (int v = 0; v < kNv; ++v) dst[g.at(v, k + kHs, pc[c])] = q[v];
          }
        }
      }
    }
    if (stage == 2) break;
    __syncthreads();
#pragma unroll
    for (int c = 0; c < 2; ++c) {
      if (colok[c] && esok[c]) {
        const int pe = wl[c] ? plo : phi;
        double q[4];
#pragma unroll
        for (int v = 0; v < kNv; ++v) q[v] = wall_halo(v, s[v][es[c]], hy[pe], hy[kk[c] + kHs]);
#pragma unroll
        for (int v = 0; v < kNv; ++v) s[v][sl[c]] = q[v];
        if (stage == 1 && ownh[c]) {  // set_bc_z of stage 3
#pragma unroll
          for (int v = 0; v < kNv; ++v) a.tmp[g.at(v, kk[c] + kHs, pc[c])] = q[v];
        }
      }
    }
    __syncthreads();
  }
}

// The dense tiles with one cell per thread (CW x NR threads, a wave holding 8
// whole slot rows): the RK init in registers, the stage state updated in
// place, the wall halo rows of each stage state written by the edge row's
// thread in the update phase, and the waves ordered by neighbour flags. A
// face that reads a wall halo row lies 2 slots inside the next stage's face
// window, so its edge row was updated, and its halo rows written, in this
// block at the stage before; halo rows of edge rows outside that window are
// left unwritten, as no face of the block that is kept reads them.
template <int CW, int NR>
__device__ __forceinline__ void dir_zt1d_tile(const DirArgs& a, const Grid& g, int b,
                                              ZTile2HLds<CW, NR>& sm, bool corners) {
  constexpr int NT = CW * NR, C = NR - 12;
  static_assert(CW == 8 && NT % 64 == 0, "waves of 8 whole slot rows");
  auto& s = sm.s;
  auto& fl = sm.fl;
  auto& hz = sm.hz;
  const int nx = g.nx, nz = g.nz, t = threadIdx.x;
  const int Ez = nz + 4;
  wave_flags_init(sm.wf, NT / 64, t);
  if (t < NR) {  // a tile row is one row k of one column group
    const int gr = b * C + t - 4;
    const int k = gr >= 0 ? gr % Ez - 2 : -3;
    const bool f = k >= 0 && k <= nz;
    hz[0][t] = f ? a.hyr_int[k] : 0.0;
    hz[1][t] = f ? a.hyt_int[k] : 0.0;
    hz[2][t] = f ? a.hyp_int[k] : 0.0;
  }
  const int col = t % CW, u = t, j = t / CW;
  const gdouble* hy = a.hy_cell;
  const int plo = kHs, phi = nz + kHs - 1;
  const int gr = b * C + j - 4;
  const int grp = gr >= 0 ? gr / Ez : -1;
  const int k = gr >= 0 ? gr - grp * Ez - 2 : -3;  // -3: padding before the first group
  const int i = grp * CW + col;
  const bool colok = gr >= 0 && i < nx;
  const int pc = (colok ? i : 0) + kHs;
  const bool wl = k == -2 || k == -1, wh = k == nz || k == nz + 1;
  const bool own = colok && j >= 6 && j < NR - 6 && k >= 0 && k < nz;
  double S[4];
  {
    double q[4] = {0.0, 0.0, 0.0, 0.0};
    if (colok) {
      if (k >= 0 && k < nz) {
#pragma unroll
        for (int v = 0; v < kNv; ++v) q[v] = a.in[g.at(v, k + kHs, pc)];
      } else if (wl || wh) {
        const int pe = wl ? plo : phi;
#pragma unroll
        for (int v = 0; v < kNv; ++v) q[v] = wall_halo(v, a.in[g.at(v, pe, pc)], hy[pe], hy[k + kHs]);
        const int je = wl ? j - k : j - (k - (nz - 1));  // the edge row's tile row
        if (je >= 6 && je < NR - 6) {  // set_bc_z of stage 1 into st
#pragma unroll
          for (int v = 0; v < kNv; ++v) a.st[g.at(v, k + kHs, pc)] = q[v];
        }
      }
    }
#pragma unroll
    for (int v = 0; v < kNv; ++v) {
      s[v][u] = q[v];
      S[v] = q[v];
    }
  }
  const double src = (a.src_w && colok && k >= 0 && k < nz) ? a.src_w[(long long)k * nx + i] : 0.0;
  if (corners && t < 8) {
    // corner columns (padded 0, 1, nx+2, nx+3) of the halo rows of st, tmp
    gdouble* buf = t < 4 ? a.st : a.tmp;
    const int qq = t & 3;
    const int cc = qq < 2 ? qq : nx + qq;
#pragma unroll
    for (int v = 0; v < kNv; ++v) {
      const double e = buf[g.at(v, plo, cc)];
      buf[g.at(v, 0, cc)] = wall_halo(v, e, hy[plo], hy[0]);
      buf[g.at(v, 1, cc)] = wall_halo(v, e, hy[plo], hy[1]);
      const double f = buf[g.at(v, phi, cc)];
      buf[g.at(v, nz + 2, cc)] = wall_halo(v, f, hy[phi], hy[nz + 2]);
      buf[g.at(v, nz + 3, cc)] = wall_halo(v, f, hy[phi], hy[nz + 3]);
    }
  }
  gdouble* const tmp_own = a.tmp_interior ? a.tmp : nullptr;
  const double id = a.id;
  PowRef pref;  // the face at this slot, every stage
  __syncthreads();  // also publishes the flags
#pragma unroll
  for (int stage = 0; stage < 3; ++stage) {
    const int e = 2 + 2 * stage;
    bool ok = true;
    if (colok && j >= e && j <= NR - e && k >= 0 && k <= nz) {
      double v4[4], d3[4], G[4];
#pragma unroll
      for (int v = 0; v < kNv; ++v)
        recon4(s[v][u - 2 * CW], s[v][u - CW], s[v][u], s[v][u + CW], v4[v], d3[v]);
      flux_z_face(v4, d3, hz[0][j], hz[1][j], hz[2][j], a.hv, k == 0 || k == nz, G,
                  stage == 0 ? POW_FIRST : POW_TRY, &pref, &ok);
#pragma unroll
      for (int v = 0; v < kNv; ++v) fl[v][u] = G[v];
    }
    if (!ok) {  // the face again with pow_pos (rare)
      double v4[4], d3[4], G[4];
#pragma unroll
      for (int v = 0; v < kNv; ++v)
        recon4(s[v][u - 2 * CW], s[v][u - CW], s[v][u], s[v][u + CW], v4[v], d3[v]);
      flux_z_face(v4, d3, hz[0][j], hz[1][j], hz[2][j], a.hv, k == 0 || k == nz, G);
#pragma unroll
      for (int v = 0; v < kNv; ++v) fl[v][u] = G[v];
    }
    wave_phase_done(sm.wf, 2 * stage + 1, t);
    const double dts = stage == 0 ? a.dt1 : stage == 1 ? a.dt2 : a.dt3;
    if (colok && j >= e && j <= NR - e - 1 && k >= 0 && k < nz) {
      double td[4];
#pragma unroll
      for (int v = 0; v < kNv; ++v) td[v] = tend_of(fl[v][u], fl[v][u + CW], a.d, id);
      td[2] = td[2] - s[0][u] * kGrav;
      if (a.src_w) td[2] = td[2] + src;
      double q[4];
#pragma unroll
      for (int v = 0; v < kNv; ++v) q[v] = S[v] + dts * td[v];
      if (stage < 2) {
#pragma unroll
        for (int v = 0; v < kNv; ++v) s[v][u] = q[v];  // in place: faces are done
        // the wall halo rows of the new stage state (slots u -+ CW, 2 CW of
        // this column group; no thread reads them in this phase)
#pragma unroll
        for (int side = 0; side < 2; ++side) {
          if (k == (side == 0 ? 0 : nz - 1)) {
            const int pe = side == 0 ? plo : phi;
            const int step = side == 0 ? -CW : CW;
#pragma unroll
            for (int h = 1; h <= 2; ++h) {
              const int kh = side == 0 ? -h : nz - 1 + h;  // the halo row's k
              double w[4];
#pragma unroll
              for (int v = 0; v < kNv; ++v) w[v] = wall_halo(v, q[v], hy[pe], hy[kh + kHs]);
#pragma unroll
              for (int v = 0; v < kNv; ++v) s[v][u + h * step] = w[v];
              if (stage == 1 && own) {  // set_bc_z of stage 3
#pragma unroll
                for (int v = 0; v < kNv; ++v) a.tmp[g.at(v, kh + kHs, pc)] = w[v];
              }
            }
          }
        }
      }
      if (own) {
        gdouble* dst = stage == 0 ? a.s1 : stage == 1 ? tmp_own : a.out;
        if (dst) {
#pragma unroll
          for (int v = 0; v < kNv; ++v) dst[g.at(v, k + kHs, pc)] = q[v];
        }
      }
    }
    if (stage == 2) break;
    wave_phase_done(sm.wf, 2 * stage + 2, t);
  }
}

template <int CW, int NR>
__global__ __launch_bounds__(CW * NR / 2) __attribute__((amdgpu_waves_per_eu(4))) void k_dir_zt2d(DirArgs a, Grid g, int nreal,
                                                          int per_xcd) {
  __shared__ ZTile2HLds<CW, NR> sm;
  const int logical = xcd_logical(blockIdx.x, per_xcd);
  if (logical >= nreal) return;  // whole block, before any barrier
  member_offset(a);
  dir_zt2d_tile<CW, NR>(a, g, logical, sm, logical == 0);
}

template <int CW, int NR>
__global__ __launch_bounds__(CW * NR / 2) __attribute__((amdgpu_waves_per_eu(4))) void k_dir_zt2(
    DirArgs a, Grid g, int ncolgrp, int nreal, int per_xcd) {
  __shared__ ZTile2HLds<CW, NR> sm;
  const int logical = xcd_logical(blockIdx.x, per_xcd);
  if (logical >= nreal) return;  // whole block, before any barrier
  member_offset(a);
  dir_zt2_tile<CW, NR>(a, g, (logical % ncolgrp) * CW, (logical / ncolgrp) * (NR - 12), sm,
                       logical == 0);
}

template <int CW, int NR>
__global__ __launch_bounds__(CW * NR) __attribute__((amdgpu_waves_per_eu(6))) void k_dir_zt1d(
    DirArgs a, Grid g, int nreal, int per_xcd) {
  __shared__ ZTile2HLds<CW, NR> sm;
  const int logical = xcd_logical(blockIdx.x, per_xcd);
  if (logical >= nreal) return;  // whole block, before any barrier
  member_offset(a);
  dir_zt1d_tile<CW, NR>(a, g, logical, sm, logical == 0);
}

// The same tiles with one cell per thread (CW x NR threads): at 8 x 96, two
// 12-wave blocks per CU (50 KB of LDS each) at 6 waves per SIMD, the waves
// ordered by neighbour flags.
template <int CW, int NR>
__global__ __launch_bounds__(CW * NR) __attribute__((amdgpu_waves_per_eu(6))) void k_dir_zt1(
    DirArgs a, Grid g, int ncolgrp, int nreal, int per_xcd) {
  __shared__ ZTile2HLds<CW, NR> sm;
  const int logical = xcd_logical(blockIdx.x, per_xcd);
  if (logical >= nreal) return;  // whole block, before any barrier
  member_offset(a);
  dir_zt2_tile<CW, NR, 1, true>(a, g, (logical % ncolgrp) * CW,
                                (logical / ncolgrp) * (NR - 12), sm, logical == 0);
}

// ------------------------------------------------------------- ensemble --
// Device descriptor of one member of an ensemble (uneven grids).
struct MemberDev {
  double* state;
  double* tmp;
  double* work;
  const double* hyc;   // hy_dens_cell
  const double* hytc;  // hy_dens_theta_cell
  const double* hyri;  // hy_dens_int
  const double* hyti;  // hy_dens_theta_int
  const double* hypi;  // hy_pressure_int
  const double* src;   // w source or nullptr
  int nx, nz;
  double dx, dz, dt1, dt2, dt3, hvx, hvz;
  double idx, idz;       // RN(1/dx), RN(1/dz)
  int x_start, z_start;  // first work item of this member in the x / z lists
  int x_ntiles, z_ncol;  // x tiles per row (dense x tiles: blocks of the member), z column groups
};

__device__ __forceinline__ DirArgs member_args(const MemberDev& m, int dir, int pass,
                                               bool tmp_interior) {
  DirArgs a;
  a.in = gptr(pass == 0 ? m.state : m.work);
  a.out = gptr(pass == 0 ? m.work : m.state);
  a.st = gptr(m.state);
  a.tmp = gptr(m.tmp);
  a.tmp_interior = tmp_interior;
  a.hy_cell = gptr(m.hyc);
  a.hyt_cell = gptr(m.hytc);
  a.hyr_int = gptr(m.hyri);
  a.hyt_int = gptr(m.hyti);
  a.hyp_int = gptr(m.hypi);
  a.src_w = m.src ? gptr(m.src) : nullptr;
  a.d = dir == 0 ? m.dx : m.dz;
  a.id = dir == 0 ? m.idx : m.idz;
  a.dt1 = m.dt1;
  a.dt2 = m.dt2;
  a.dt3 = m.dt3;
  a.hv = dir == 0 ? m.hvx : m.hvz;
  return a;
}

// Member owning work item `item`: binary search over the ascending start
// offsets (block-uniform, scalar loads). The offsets are also kept as two
// packed int arrays after the descriptors (x starts, then z starts): a
// search then touches a few cache lines instead of one descriptor per probe
// (256 descriptors are 39 KB, more than the scalar cache holds).
template <int DIR>
__device__ __forceinline__ int member_of(const MemberDev* m, int n, int item) {
  const PMW_GLOBAL int* start = (const PMW_GLOBAL int*)(m + n) + DIR * n;
  int lo = 0, hi = n - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (start[mid] <= item) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// Ensemble kernels: one launch covers the flattened work list of every
// member (tiles of member q are items x_start[q] .. ); block b takes item
// xcd_logical(b), so neighbouring tiles of a member share an XCD's L2, and
// finds its member by a scalar binary search over the start offsets. (A
// persistent form, each block looping over a chunk of items, kept ~70 more
// VGPRs live and halved the occupancy.) `pass`: bit 0 = direction of the
// step (0: state -> work, 1: work -> state), bit 1 = store state_tmp's
// interior (the last direction of the call only).
template <int NT>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(6))) void k_ens_x(const MemberDev* __restrict__ m, int n, int pass,
                                              int total, int per_xcd) {
  constexpr int C = NT - 12;
  __shared__ XTileLds<NT> sm;
  const int item = xcd_logical(blockIdx.x, per_xcd);
  if (item >= total) return;  // whole block, before any barrier
  const MemberDev& md = m[member_of<0>(m, n, item)];
  const DirArgs a = member_args(md, 0, pass & 1, pass >> 1);
  const int local = item - md.x_start;
  const int row = local / md.x_ntiles;
  if (row >= md.nz) return;  // a plan that is not the prepared one: stay in bounds
  dir_x_tile<NT>(a, Grid(md.nx, md.nz), row, (local % md.x_ntiles) * C, sm);
}

template <int NT>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(5))) void k_ens_xd(const MemberDev* __restrict__ m, int n, int pass,
                                               int total, int per_xcd) {
  __shared__ XTileLds<NT> sm;
  const int item = xcd_logical(blockIdx.x, per_xcd);
  if (item >= total) return;  // whole block, before any barrier
  const MemberDev& md = m[member_of<0>(m, n, item)];
  const DirArgs a = member_args(md, 0, pass & 1, pass >> 1);
  const int local = item - md.x_start;
  if (local >= md.x_ntiles) return;  // as in k_ens_x
  dir_xd_tile<NT>(a, Grid(md.nx, md.nz), local, sm);
}

template <int L>
__global__ __launch_bounds__(192) void k_ens_z(const MemberDev* __restrict__ m, int n, int pass,
                                               int total, int per_xcd) {
  __shared__ double ring[2][2][kNv][64];
  const int item = xcd_logical(blockIdx.x, per_xcd);
  if (item >= total) return;  // whole block, before any barrier
  const MemberDev& md = m[member_of<1>(m, n, item)];
  const DirArgs a = member_args(md, 1, pass & 1, pass >> 1);
  const int local = item - md.z_start;
  if ((long long)(local / md.z_ncol) * L >= md.nz) return;  // as in k_ens_x
  dir_z_block<L>(a, Grid(md.nx, md.nz), local % md.z_ncol, local / md.z_ncol, ring);
}

template <int CW, int NT>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(6))) void k_ens_zt(
    const MemberDev* __restrict__ m, int n, int pass, int total, int per_xcd) {
  constexpr int C = NT / CW - 12;
  __shared__ ZTileLds<NT, NT / CW> sm;
  const int item = xcd_logical(blockIdx.x, per_xcd);
  if (item >= total) return;  // whole block, before any barrier
  const MemberDev& md = m[member_of<1>(m, n, item)];
  const DirArgs a = member_args(md, 1, pass & 1, pass >> 1);
  const int local = item - md.z_start;
  if ((long long)(local / md.z_ncol) * C >= md.nz) return;  // as in k_ens_x
  dir_zt_tile<CW, NT>(a, Grid(md.nx, md.nz), (local % md.z_ncol) * CW,
                      (local / md.z_ncol) * C, sm, local == 0);
}

template <int CW, int NR>
__global__ __launch_bounds__(CW * NR / 2) __attribute__((amdgpu_waves_per_eu(4))) void k_ens_zt2d(const MemberDev* __restrict__ m, int n,
                                                          int pass, int total, int per_xcd) {
  __shared__ ZTile2HLds<CW, NR> sm;
  const int item = xcd_logical(blockIdx.x, per_xcd);
  if (item >= total) return;  // whole block, before any barrier
  const MemberDev& md = m[member_of<1>(m, n, item)];
  const DirArgs a = member_args(md, 1, pass & 1, pass >> 1);
  const int local = item - md.z_start;
  if (local >= md.z_ncol) return;  // dense z: z_ncol holds the member's blocks
  dir_zt2d_tile<CW, NR>(a, Grid(md.nx, md.nz), local, sm, local == 0);
}

template <int CW, int NR>
__global__ __launch_bounds__(CW * NR) __attribute__((amdgpu_waves_per_eu(6))) void k_ens_zt1d(
    const MemberDev* __restrict__ m, int n, int pass, int total, int per_xcd) {
  __shared__ ZTile2HLds<CW, NR> sm;
  const int item = xcd_logical(blockIdx.x, per_xcd);
  if (item >= total) return;  // whole block, before any barrier
  const MemberDev& md = m[member_of<1>(m, n, item)];
  const DirArgs a = member_args(md, 1, pass & 1, pass >> 1);
  const int local = item - md.z_start;
  if (local >= md.z_ncol) return;  // dense z: z_ncol holds the member's blocks
  dir_zt1d_tile<CW, NR>(a, Grid(md.nx, md.nz), local, sm, local == 0);
}

template <int CW, int NR>
__global__ __launch_bounds__(CW * NR) __attribute__((amdgpu_waves_per_eu(6))) void k_ens_zt1(
    const MemberDev* __restrict__ m, int n, int pass, int total, int per_xcd) {
  __shared__ ZTile2HLds<CW, NR> sm;
  const int item = xcd_logical(blockIdx.x, per_xcd);
  if (item >= total) return;  // whole block, before any barrier
  const MemberDev& md = m[member_of<1>(m, n, item)];
  const DirArgs a = member_args(md, 1, pass & 1, pass >> 1);
  const int local = item - md.z_start;
  if ((long long)(local / md.z_ncol) * (NR - 12) >= md.nz) return;  // as in k_ens_x
  dir_zt2_tile<CW, NR, 1, true>(a, Grid(md.nx, md.nz), (local % md.z_ncol) * CW,
                                (local / md.z_ncol) * (NR - 12), sm, local == 0);
}

template <int CW, int NR>
__global__ __launch_bounds__(CW * NR / 2) __attribute__((amdgpu_waves_per_eu(4))) void k_ens_zt2(const MemberDev* __restrict__ m, int n,
                                                         int pass, int total, int per_xcd) {
  __shared__ ZTile2HLds<CW, NR> sm;
  const int item = xcd_logical(blockIdx.x, per_xcd);
  if (item >= total) return;  // whole block, before any barrier
  const MemberDev& md = m[member_of<1>(m, n, item)];
  const DirArgs a = member_args(md, 1, pass & 1, pass >> 1);
  const int local = item - md.z_start;
  if ((long long)(local / md.z_ncol) * (NR - 12) >= md.nz) return;  // as in k_ens_x
  dir_zt2_tile<CW, NR>(a, Grid(md.nx, md.nz), (local % md.z_ncol) * CW,
                       (local / md.z_ncol) * (NR - 12), sm, local == 0);
}

template <int CW, int NR>
int launch_zt2(const DirArgs& a, Grid g, hipStream_t st, int members) {
  constexpr int C = NR - 12;
  const int ncolgrp = (g.nx + CW - 1) / CW;
  const long long nreal = (long long)ncolgrp * ((g.nz + C - 1) / C);
  const int per = (int)((nreal + kNumXcd - 1) / kNumXcd);
  k_dir_zt2<CW, NR><<<dim3(per * kNumXcd, members), CW * NR / 2, 0, st>>>(a, g, ncolgrp,
                                                                          (int)nreal, per);
  PMW_CHECK_LAUNCH();
  return PMW_OK;
}

template <int CW, int NR>
int launch_zt1(const DirArgs& a, Grid g, hipStream_t st, int members) {
  constexpr int C = NR - 12;
  const int ncolgrp = (g.nx + CW - 1) / CW;
  const long long nreal = (long long)ncolgrp * ((g.nz + C - 1) / C);
  const int per = (int)((nreal + kNumXcd - 1) / kNumXcd);
  k_dir_zt1<CW, NR><<<dim3(per * kNumXcd, members), CW * NR, 0, st>>>(a, g, ncolgrp,
                                                                      (int)nreal, per);
  PMW_CHECK_LAUNCH();
  return PMW_OK;
}

template <int CW, int NR>
int launch_zt1d(const DirArgs& a, Grid g, hipStream_t st, int members) {
  const int nreal = dense_z_blocks(g.nx, g.nz, CW, NR);
  const int per = (nreal + kNumXcd - 1) / kNumXcd;
  k_dir_zt1d<CW, NR><<<dim3(per * kNumXcd, members), CW * NR, 0, st>>>(a, g, nreal, per);
  PMW_CHECK_LAUNCH();
  return PMW_OK;
}

template <int CW, int NR>
int launch_zt2d(const DirArgs& a, Grid g, hipStream_t st, int members) {
  const int nreal = dense_z_blocks(g.nx, g.nz, CW, NR);
  const int per = (nreal + kNumXcd - 1) / kNumXcd;
  k_dir_zt2d<CW, NR><<<dim3(per * kNumXcd, members), CW * NR / 2, 0, st>>>(a, g, nreal, per);
  PMW_CHECK_LAUNCH();
  return PMW_OK;
}

template <int NT>
int launch_x(const DirArgs& a, Grid g, hipStream_t st, int members) {
  constexpr int C = NT - 12;
  const int ntiles = (g.nx + C - 1) / C;
  const long long nreal = (long long)ntiles * g.nz;
  const int per = (int)((nreal + kNumXcd - 1) / kNumXcd);
  k_dir_x<NT><<<dim3(per * kNumXcd, members), NT, 0, st>>>(a, g, ntiles, (int)nreal, per);
  PMW_CHECK_LAUNCH();
  return PMW_OK;
}

template <int NT>
int launch_xd(const DirArgs& a, Grid g, hipStream_t st, int members) {
  const int nreal = dense_x_blocks(g.nx, g.nz, NT);
  const int per = (nreal + kNumXcd - 1) / kNumXcd;
  k_dir_xd<NT><<<dim3(per * kNumXcd, members), NT, 0, st>>>(a, g, nreal, per);
  PMW_CHECK_LAUNCH();
  return PMW_OK;
}

template <int CW, int NT>
int launch_zt(const DirArgs& a, Grid g, hipStream_t st, int members) {
  constexpr int C = NT / CW - 12;
  const int ncolgrp = (g.nx + CW - 1) / CW;
  const long long nreal = (long long)ncolgrp * ((g.nz + C - 1) / C);
  const int per = (int)((nreal + kNumXcd - 1) / kNumXcd);
  k_dir_zt<CW, NT><<<dim3(per * kNumXcd, members), NT, 0, st>>>(a, g, ncolgrp, (int)nreal,
                                                                 per);
  PMW_CHECK_LAUNCH();
  return PMW_OK;
}

template <int L>
int launch_z(const DirArgs& a, Grid g, hipStream_t st, int members) {
  const int ncolblk = (g.nx + 4 + 63) / 64;
  const int nseg = (g.nz + L - 1) / L;
  const long long nreal = (long long)ncolblk * nseg;
  const int per = (int)((nreal + kNumXcd - 1) / kNumXcd);
  k_dir_z<L><<<dim3(per * kNumXcd, members), 192, 0, st>>>(a, g, ncolblk, (int)nreal, per);
  PMW_CHECK_LAUNCH();
  return PMW_OK;
}

}  // namespace

namespace pmw {

// x: 256, 448, 512 = threads per row tile; 1000 + 256 or 512 = dense tiles
bool fused_x_tile_ok(int nt) {
  return nt == 256 || nt == 448 || nt == 512 || nt == 1256 || nt == 1512;
}
// z: 8..64 = rows per marching segment; columns*1000 + rows = LDS tiles:
// 8064, 4128 (one cell per thread, RK init in LDS), 8096 (one cell per
// thread, RK init in registers, stage state updated in place), 8128 (two
// cells per thread); 18128 = 8128 as dense tiles (column groups end to end)
bool fused_z_seg_ok(int seg) {
  return seg == 8 || seg == 16 || seg == 24 || seg == 32 || seg == 48 || seg == 64 ||
         seg == 8064 || seg == 4128 || seg == 8128 || seg == 18128 || seg == 8096 || seg == 18096;
}

int launch_dir_x(const DirArgs& a, Grid g, int nt, hipStream_t st, int members) {
  if (members < 1 || members > 65535 || (members > 1 && a.mstride <= 0)) return PMW_EINVAL;
  if (nt > 1000 && dense_x_slots(g.nx, g.nz) > (1LL << 30)) return PMW_EINVAL;
  switch (nt) {
    case 256: return launch_x<256>(a, g, st, members);
    case 448: return launch_x<448>(a, g, st, members);
    case 512: return launch_x<512>(a, g, st, members);
    case 1256: return launch_xd<256>(a, g, st, members);
    case 1512: return launch_xd<512>(a, g, st, members);
    default: return PMW_EINVAL;
  }
}

int launch_dir_z(const DirArgs& a, Grid g, int seg, hipStream_t st, int members) {
  if (members < 1 || members > 65535 || (members > 1 && a.mstride <= 0)) return PMW_EINVAL;
  switch (seg) {
    case 8: return launch_z<8>(a, g, st, members);
    case 16: return launch_z<16>(a, g, st, members);
    case 24: return launch_z<24>(a, g, st, members);
    case 32: return launch_z<32>(a, g, st, members);
    case 48: return launch_z<48>(a, g, st, members);
    case 64: return launch_z<64>(a, g, st, members);
    case 8064: return launch_zt<8, 512>(a, g, st, members);
    case 4128: return launch_zt<4, 512>(a, g, st, members);
    case 8128: return launch_zt2<8, 128>(a, g, st, members);
    case 8096: return launch_zt1<8, 96>(a, g, st, members);
    case 18128:
      if (dense_z_rows(g.nx, g.nz, 8) > (1LL << 30)) return PMW_EINVAL;
      return launch_zt2d<8, 128>(a, g, st, members);
    case 18096:
      if (dense_z_rows(g.nx, g.nz, 8) > (1LL << 30)) return PMW_EINVAL;
      return launch_zt1d<8, 96>(a, g, st, members);
    default: return PMW_EINVAL;
  }
}

}  // namespace pmw

// ============================================================ ensemble ABI ==
namespace {

int ens_x_cells(int nt) { return nt - 12; }

struct EnsPlan {
  int x_total, z_total;
};

int plan_members(const pmw_member* mem, int n, int x_nt, int z_seg, MemberDev* dev,
                 EnsPlan& plan) {
  long long xs = 0, zs = 0;
  for (int q = 0; q < n; ++q) {
    const pmw_member& m = mem[q];
    if (m.nx < 2 || m.nz < 2 || !m.state || !m.state_tmp || !m.state_work ||
        !m.hy_dens_cell || !m.hy_dens_theta_cell || !m.hy_dens_int ||
        !m.hy_dens_theta_int || !m.hy_pressure_int)
      return PMW_EINVAL;
    if (m.state == m.state_tmp || m.state == m.state_work || m.state_tmp == m.state_work)
      return PMW_EALIAS;
    const bool dense = x_nt > 1000;
    if (dense && dense_x_slots(m.nx, m.nz) > (1LL << 30)) return PMW_EINVAL;
    const int xt = dense ? dense_x_blocks(m.nx, m.nz, x_nt - 1000)
                         : (m.nx + ens_x_cells(x_nt) - 1) / ens_x_cells(x_nt);
    int zc, zsg;
    if (z_seg == 18128 || z_seg == 18096) {  // dense tiles: zc = the member's blocks, one "segment"
      if (dense_z_rows(m.nx, m.nz, 8) > (1LL << 30)) return PMW_EINVAL;
      zc = dense_z_blocks(m.nx, m.nz, 8, z_seg % 1000);
      zsg = 1;
    } else if (z_seg >= 1000) {  // LDS tiles: columns*1000 + rows (12 of them halo)
      const int cw = z_seg / 1000, rows = z_seg % 1000 - 12;
      zc = (m.nx + cw - 1) / cw;
      zsg = (m.nz + rows - 1) / rows;
    } else {
      zc = (m.nx + 4 + 63) / 64;
      zsg = (m.nz + z_seg - 1) / z_seg;
    }
    if (dev) {
      MemberDev& d = dev[q];
      d.state = m.state;
      d.tmp = m.state_tmp;
      d.work = m.state_work;
      d.hyc = m.hy_dens_cell;
      d.hytc = m.hy_dens_theta_cell;
      d.hyri = m.hy_dens_int;
      d.hyti = m.hy_dens_theta_int;
      d.hypi = m.hy_pressure_int;
      d.src = m.src_w;
      d.nx = m.nx;
      d.nz = m.nz;
      d.dx = m.dx;
      d.dz = m.dz;
      d.idx = 1.0 / m.dx;
      d.idz = 1.0 / m.dz;
      d.dt1 = m.dt / 3;
      d.dt2 = m.dt / 2;
      d.dt3 = m.dt / 1;
      d.hvx = -kHvBeta * m.dx / (16 * m.dt_full);
      d.hvz = -kHvBeta * m.dz / (16 * m.dt_full);
      d.x_start = (int)xs;
      d.z_start = (int)zs;
      d.x_ntiles = xt;
      d.z_ncol = zc;
    }
    xs += dense ? (long long)xt : (long long)xt * m.nz;  // dense: blocks, else tiles x rows
    zs += (long long)zc * zsg;
    if (xs > (1LL << 30) || zs > (1LL << 30)) return PMW_EINVAL;
  }
  plan.x_total = (int)xs;
  plan.z_total = (int)zs;
  return PMW_OK;
}

int ens_tuning(const int* t, int& x_nt, int& z_seg) {
  x_nt = t ? t[1] : 512;
  z_seg = t ? t[2] : 4128;
  if (t && t[0] != 0) return PMW_EINVAL;  // ensembles run the fused kernels only
  if (!pmw::fused_x_tile_ok(x_nt) || !pmw::fused_z_seg_ok(z_seg)) return PMW_EINVAL;
  return PMW_OK;
}

// What a workspace was prepared for: pmw_ensemble_evolve refuses members or
// a tuning other than those (the device descriptors and work-list offsets
// would not match the launches). Kept on the host, keyed by the workspace's
// address, so that checking it needs no device read (a call stays
// capturable in a HIP graph).
struct EnsPrepared {
  int nmembers, x_nt, z_seg, x_total, z_total;
  unsigned long long members_hash;
  bool operator==(const EnsPrepared& o) const {
    return nmembers == o.nmembers && x_nt == o.x_nt && z_seg == o.z_seg &&
           x_total == o.x_total && z_total == o.z_total &&
           members_hash == o.members_hash;
  }
};

std::mutex g_prepared_mu;
std::unordered_map<const void*, EnsPrepared> g_prepared;

// FNV-1a over the member descriptors (plain pointers, ints and doubles with
// no padding between them).
unsigned long long hash_members(const pmw_member* m, int n) {
  const unsigned char* b = reinterpret_cast<const unsigned char*>(m);
  const size_t len = sizeof(pmw_member) * (size_t)n;
  unsigned long long h = 1469598103934665603ULL;
  for (size_t i = 0; i < len; ++i) {
    h ^= b[i];
    h *= 1099511628211ULL;
  }
  return h;
}

EnsPrepared prepared_for(const pmw_member* m, int n, int x_nt, int z_seg,
                         const EnsPlan& plan) {
  return EnsPrepared{n, x_nt, z_seg, plan.x_total, plan.z_total, hash_members(m, n)};
}

inline int ens_per_xcd(int total) { return (total + kNumXcd - 1) / kNumXcd; }

template <int NT>
int launch_ens_x(const MemberDev* d, int n, int pass, int total, hipStream_t st) {
  const int per = ens_per_xcd(total);
  k_ens_x<NT><<<per * kNumXcd, NT, 0, st>>>(d, n, pass, total, per);
  PMW_CHECK_LAUNCH();
  return PMW_OK;
}

template <int NT>
int launch_ens_xd(const MemberDev* d, int n, int pass, int total, hipStream_t st) {
  const int per = ens_per_xcd(total);
  k_ens_xd<NT><<<per * kNumXcd, NT, 0, st>>>(d, n, pass, total, per);
  PMW_CHECK_LAUNCH();
  return PMW_OK;
}

template <int L>
int launch_ens_z(const MemberDev* d, int n, int pass, int total, hipStream_t st) {
  const int per = ens_per_xcd(total);
  k_ens_z<L><<<per * kNumXcd, 192, 0, st>>>(d, n, pass, total, per);
  PMW_CHECK_LAUNCH();
  return PMW_OK;
}

template <int CW, int NR>
int launch_ens_zt2d(const MemberDev* d, int n, int pass, int total, hipStream_t st) {
  const int per = ens_per_xcd(total);
  k_ens_zt2d<CW, NR><<<per * kNumXcd, CW * NR / 2, 0, st>>>(d, n, pass, total, per);
  PMW_CHECK_LAUNCH();
  return PMW_OK;
}

template <int CW, int NR>
int launch_ens_zt1d(const MemberDev* d, int n, int pass, int total, hipStream_t st) {
  const int per = ens_per_xcd(total);
  k_ens_zt1d<CW, NR><<<per * kNumXcd, CW * NR, 0, st>>>(d, n, pass, total, per);
  PMW_CHECK_LAUNCH();
  return PMW_OK;
}

template <int CW, int NR>
int launch_ens_zt1(const MemberDev* d, int n, int pass, int total, hipStream_t st) {
  const int per = ens_per_xcd(total);
  k_ens_zt1<CW, NR><<<per * kNumXcd, CW * NR, 0, st>>>(d, n, pass, total, per);
  PMW_CHECK_LAUNCH();
  return PMW_OK;
}

template <int CW, int NR>
int launch_ens_zt2(const MemberDev* d, int n, int pass, int total, hipStream_t st) {
  const int per = ens_per_xcd(total);
  k_ens_zt2<CW, NR><<<per * kNumXcd, CW * NR / 2, 0, st>>>(d, n, pass, total, per);
  PMW_CHECK_LAUNCH();
  return PMW_OK;
}

template <int CW, int NT>
int launch_ens_zt(const MemberDev* d, int n, int pass, int total, hipStream_t st) {
  const int per = ens_per_xcd(total);
  k_ens_zt<CW, NT><<<per * kNumXcd, NT, 0, st>>>(d, n, pass, total, per);
  PMW_CHECK_LAUNCH();
  return PMW_OK;
}

int launch_ens(int dir, int x_nt, int z_seg, const MemberDev* d, int n, int pass,
               const EnsPlan& plan, hipStream_t st) {
  if (dir == 0) {
    if (x_nt == 256) return launch_ens_x<256>(d, n, pass, plan.x_total, st);
    if (x_nt == 448) return launch_ens_x<448>(d, n, pass, plan.x_total, st);
    if (x_nt == 1256) return launch_ens_xd<256>(d, n, pass, plan.x_total, st);
    if (x_nt == 1512) return launch_ens_xd<512>(d, n, pass, plan.x_total, st);
    return launch_ens_x<512>(d, n, pass, plan.x_total, st);
  }
  switch (z_seg) {
    case 8: return launch_ens_z<8>(d, n, pass, plan.z_total, st);
    case 16: return launch_ens_z<16>(d, n, pass, plan.z_total, st);
    case 24: return launch_ens_z<24>(d, n, pass, plan.z_total, st);
    case 32: return launch_ens_z<32>(d, n, pass, plan.z_total, st);
    case 48: return launch_ens_z<48>(d, n, pass, plan.z_total, st);
    case 8064: return launch_ens_zt<8, 512>(d, n, pass, plan.z_total, st);
    case 4128: return launch_ens_zt<4, 512>(d, n, pass, plan.z_total, st);
    case 8128: return launch_ens_zt2<8, 128>(d, n, pass, plan.z_total, st);
    case 8096: return launch_ens_zt1<8, 96>(d, n, pass, plan.z_total, st);
    case 18128: return launch_ens_zt2d<8, 128>(d, n, pass, plan.z_total, st);
    case 18096: return launch_ens_zt1d<8, 96>(d, n, pass, plan.z_total, st);
    default: return launch_ens_z<64>(d, n, pass, plan.z_total, st);
  }
}

}  // namespace

extern "C" {

long long pmw_ensemble_workspace_size(int nmembers) {
  if (nmembers < 1) return 0;
  return ((long long)sizeof(MemberDev) + 2 * sizeof(int)) * nmembers;
}

int pmw_ensemble_prepare(const pmw_member* members, int nmembers, const int* tuning,
                         void* workspace, long long workspace_bytes, void* stream) {
  if (!members || nmembers < 1 || !workspace) return PMW_EINVAL;
  if (workspace_bytes < pmw_ensemble_workspace_size(nmembers)) return PMW_ESPACE;
  int x_nt, z_seg;
  int rc = ens_tuning(tuning, x_nt, z_seg);
  if (rc) return rc;
  // descriptors, then the packed x and z start offsets (see member_of)
  const long long bytes = pmw_ensemble_workspace_size(nmembers);
  char* host = new char[bytes];
  MemberDev* dev = reinterpret_cast<MemberDev*>(host);
  for (int q = 0; q < nmembers; ++q) new (&dev[q]) MemberDev();
  EnsPlan plan;
  rc = plan_members(members, nmembers, x_nt, z_seg, dev, plan);
  if (rc == PMW_OK) {
    int* starts = reinterpret_cast<int*>(dev + nmembers);
    for (int q = 0; q < nmembers; ++q) {
      starts[q] = dev[q].x_start;
      starts[nmembers + q] = dev[q].z_start;
    }
    hipStream_t st = (hipStream_t)stream;
    {
      // until the copy has landed the workspace matches no tuning
      std::lock_guard<std::mutex> lock(g_prepared_mu);
      g_prepared.erase(workspace);
    }
    hipError_t e = hipMemcpyAsync(workspace, host, bytes, hipMemcpyHostToDevice, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) rc = (int)e;
    if (rc == PMW_OK) {
      std::lock_guard<std::mutex> lock(g_prepared_mu);
      g_prepared[workspace] = prepared_for(members, nmembers, x_nt, z_seg, plan);
    }
  }
  delete[] host;
  return rc;
}

int pmw_ensemble_evolve(const pmw_member* members, int nmembers, const int* tuning,
                        const void* workspace, int* reverse_direction, int nsteps,
                        void* stream, double* kernel_ms) {
  if (!members || nmembers < 1 || !workspace || !reverse_direction || nsteps < 0 ||
      nsteps > 100000)
    return PMW_EINVAL;
  int x_nt, z_seg;
  int rc = ens_tuning(tuning, x_nt, z_seg);
  if (rc) return rc;
  EnsPlan plan;
  rc = plan_members(members, nmembers, x_nt, z_seg, nullptr, plan);
  if (rc) return rc;
  {
    std::lock_guard<std::mutex> lock(g_prepared_mu);
    const auto it = g_prepared.find(workspace);
    if (it == g_prepared.end() ||
        !(it->second == prepared_for(members, nmembers, x_nt, z_seg, plan)))
      return PMW_EINVAL;  // not prepared, or prepared for other members / tuning
  }
  const MemberDev* d = (const MemberDev*)workspace;
  hipStream_t st = (hipStream_t)stream;
  hipEvent_t ev[2] = {nullptr, nullptr};
  if (kernel_ms) {
    kernel_ms[0] = kernel_ms[1] = 0.0;
    if (hipEventCreate(&ev[0]) != hipSuccess || hipEventCreate(&ev[1]) != hipSuccess) {
      if (ev[0]) (void)hipEventDestroy(ev[0]);
      return PMW_EINVAL;
    }
  }
  for (int step = 0; step < nsteps && rc == PMW_OK; ++step) {
    const int first = *reverse_direction ? 0 : 1;  // 0 = x, 1 = z
    for (int pass = 0; pass < 2 && rc == PMW_OK; ++pass) {
      const int dir = pass == 0 ? first : 1 - first;
      if (kernel_ms) {
        const hipError_t e = hipEventRecord(ev[0], st);
        if (e != hipSuccess) { rc = (int)e; break; }
      }
      // state_tmp's interior: only the call's last direction keeps its stores
      const int keep_tmp = pass == 1 && step == nsteps - 1;
      rc = launch_ens(dir, x_nt, z_seg, d, nmembers, pass | keep_tmp << 1, plan, st);
      if (kernel_ms && rc == PMW_OK) {
        hipError_t e = hipEventRecord(ev[1], st);
        float ms = 0.f;
        if (e == hipSuccess) e = hipEventSynchronize(ev[1]);
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, ev[0], ev[1]);
        if (e != hipSuccess) { rc = (int)e; break; }
        kernel_ms[dir] += ms;
      }
    }
    if (rc == PMW_OK) *reverse_direction = !*reverse_direction;
  }
  if (kernel_ms) {  // measurement only: a failed destroy changes nothing
    (void)hipEventDestroy(ev[0]);
    (void)hipEventDestroy(ev[1]);
  }
  return rc;
}

}  // extern "C"
