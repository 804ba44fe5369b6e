// pmw_fused.h -- launch interface of the fused-direction kernels (internal).
#pragma once

#include "pmw_common.h"

namespace pmw {

// Arguments of one fused direction (three RK stages) of one grid.
struct DirArgs {
  const gdouble* in;   // start state of the direction (interior read; RK init)
  gdouble* st;         // receives the stage-1 halo refresh (reference `state`)
  gdouble* out;        // result interior; must not alias `in`
  gdouble* tmp;        // stage-2 interior + stage-3 halo refresh (`state_tmp`)
  const gdouble* hy_cell;   // hy_dens_cell (x flux rows, z wall halos)
  const gdouble* hyt_cell;  // hy_dens_theta_cell (x flux rows)
  const gdouble* hyr_int;   // hy_dens_int (z faces)
  const gdouble* hyt_int;   // hy_dens_theta_int (z faces)
  const gdouble* hyp_int;   // hy_pressure_int (z faces)
  const gdouble* src_w;     // (nz, nx) w source or nullptr
  double d;                // dx or dz
  double id;               // RN(1/d), set with d (a host division, so no
                           // kernel spends an IEEE division on it)
  double dt1, dt2, dt3;    // dt/3, dt/2, dt
  double hv;               // -hv_beta * d / (16 * dt_full)
  gdouble* s1 = nullptr;   // optional: receives the stage-1 interior
  bool tmp_interior = true;  // false: tmp gets only its halo refresh (its
                             // interior is dead when the next direction of
                             // the same step rewrites every interior cell)
  long long mstride = 0;   // batched launches: elements between members of
                           // in / st / out / tmp / s1 (member = blockIdx.y)
};

bool fused_x_tile_ok(int nt);
bool fused_z_seg_ok(int seg);
// `members` identical grids (same hydrostatic profiles and source) laid out
// a.mstride elements apart are advanced by one launch.
int launch_dir_x(const DirArgs& a, Grid g, int nt, hipStream_t st, int members = 1);
int launch_dir_z(const DirArgs& a, Grid g, int seg, hipStream_t st, int members = 1);

}  // namespace pmw
