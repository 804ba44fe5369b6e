/*
 * pmw.h -- C ABI of libpmw.so, the gfx950 (MI355X) HIP implementation of
 * PyMiniWeather's time-stepping path.
 *
 * Conventions shared by every entry point
 * ---------------------------------------
 * - All arrays are FP64 device pointers (hipMalloc'd or torch-owned), C order.
 * - A state buffer is (4, nz+4, nx+4): variable-major (rho', rho*u, rho*w,
 *   (rho*theta)' in that order, pyminiweather/__init__.py:14-18), then z,
 *   then x fastest, with a 2-cell halo on each side
 *   (pyminiweather/data/fields.py:65-72).
 * - Hydrostatic profiles: *_cell arrays have nz+4 entries, *_int arrays nz+1
 *   (pyminiweather/data/fields.py:80-85).
 * - `stream` is a hipStream_t passed as void* (NULL = the null stream). Every
 *   call is asynchronous on that stream; nothing allocates device memory.
 * - Return value: 0 on success; a negative PMW_E* code for an argument error
 *   (nothing was launched); a positive hipError_t if a launch failed.
 * - Calls on distinct buffers and streams are thread-safe. Calls sharing a
 *   buffer must be ordered by the caller (same stream).
 */
#ifndef PMW_H
#define PMW_H

#ifdef __cplusplus
extern "C" {
#endif

#define PMW_OK 0
#define PMW_EINVAL (-1)   /* bad size, null pointer, or bad enum value      */
#define PMW_EALIAS (-2)   /* output buffer aliases an input it must not     */
#define PMW_ESPACE (-3)   /* caller-provided workspace too small            */

/* Boundary-condition flavour of set_bc_x: periodic, or periodic with the
 * injection inflow on the left (pyminiweather/ics/bcs.py:37-64). */
#define PMW_BC_PERIODIC 0
#define PMW_BC_INJECTION 1

/* Initial-condition kinds (pyminiweather/ics/initial_conditions.py:333-429). */
#define PMW_IC_THERMAL 0
#define PMW_IC_COLLISION 1
#define PMW_IC_GRAVITY 2
#define PMW_IC_DENSITY_CURRENT 3
#define PMW_IC_INJECTION 4

/* ---- library ----------------------------------------------------------- */

/* ABI version (major*100 + minor). No reference counterpart. */
int pmw_abi_version(void);

/* Human-readable text for a return code. No reference counterpart. */
const char* pmw_error_string(int code);

/* ---- halos ------------------------------------------------------------- */

/* Replaces set_bc_x (pyminiweather/ics/bcs.py:8-64): periodic x halos on
 * interior rows 2..nz+1 of `state`, all 4 variables. With bc ==
 * PMW_BC_INJECTION the right halo is left alone and, on rows flagged in
 * inj_rows (nz ints, interior row order), rho*u and (rho*theta)' of the two
 * left halo columns are set from the inflow (bcs.py:41-64); hy_dens_cell and
 * hy_dens_theta_cell are then required. */
int pmw_set_bc_x(double* state, int nx, int nz, int bc, const int* inj_rows,
                 const double* hy_dens_cell, const double* hy_dens_theta_cell,
                 void* stream);

/* Replaces set_bc_z (pyminiweather/ics/bcs.py:67-148): wall halos on rows
 * 0, 1, nz+2, nz+3 over all nx+4 columns: rho*w = 0; rho*u =
 * (rho*u[edge] / hy[edge]) * hy[halo row]; rho' and (rho*theta)' copied from
 * rows 2 and nz+1. */
int pmw_set_bc_z(double* state, const double* hy_dens_cell, int nx, int nz,
                 void* stream);

/* ---- per-function kernels (unfused; fill the reference's intermediates) - */

/* Replaces interpolate_x (pyminiweather/solve/interpolate.py:10-43):
 * vals_x, d3_vals_x are (4, nz, nx+1). Face i uses padded columns i..i+3 of
 * rows 2..nz+1 with the convolution's exact accumulation order. */
int pmw_interpolate_x(const double* state, double* vals_x, double* d3_vals_x,
                      int nx, int nz, void* stream);

/* Replaces interpolate_z (pyminiweather/solve/interpolate.py:46-79):
 * vals_z, d3_vals_z are (4, nz+1, nx); face k uses padded rows k..k+3 of
 * columns 2..nx+1. */
int pmw_interpolate_z(const double* state, double* vals_z, double* d3_vals_z,
                      int nx, int nz, void* stream);

/* Replaces compute_flux_x (pyminiweather/solve/interpolate.py:82-129):
 * writes flux[:, 0:nz, 0:nx+1] of the (4, nz+1, nx+1) flux array.
 * hv_coeff = -hv_beta * dx / (16 * dt_full). */
int pmw_compute_flux_x(const double* vals_x, const double* d3_vals_x,
                       const double* hy_dens_cell,
                       const double* hy_dens_theta_cell, double* flux, int nx,
                       int nz, double hv_coeff, void* stream);

/* Replaces compute_flux_z (pyminiweather/solve/interpolate.py:132-186):
 * writes flux[:, 0:nz+1, 0:nx]; zeroes d3_vals_z[rho'] on the wall faces
 * k = 0 and k = nz in place, as the reference does. */
int pmw_compute_flux_z(const double* vals_z, double* d3_vals_z,
                       const double* hy_dens_int,
                       const double* hy_dens_theta_int,
                       const double* hy_pressure_int, double* flux, int nx,
                       int nz, double hv_coeff, void* stream);

/* Replaces compute_tend_x (pyminiweather/solve/interpolate.py:189-215):
 * tend (4, nz, nx) = -(flux[.., i+1] - flux[.., i]) / dx. */
int pmw_compute_tend_x(const double* flux, double* tend, int nx, int nz,
                       double dx, void* stream);

/* Replaces compute_tend_z (pyminiweather/solve/interpolate.py:218-250):
 * tend = -(flux[k+1] - flux[k]) / dz, then tend[rho*w] -= rho'_forcing * g. */
int pmw_compute_tend_z(const double* flux, const double* state_forcing,
                       double* tend, int nx, int nz, double dz, void* stream);

/* Replaces add_source_terms (pyminiweather/solve/source.py:53-75) for the
 * one active source (GravityBCSourceTerm, source.py:20-50):
 * tend[rho*w] += src_w, with src_w (nz, nx) = wpert * hy_dens_cell
 * precomputed once. */
int pmw_add_source_w(double* tend, const double* src_w, int nx, int nz,
                     void* stream);

/* Replaces the RK update of discrete_step (pyminiweather/solve/step.py:80-82):
 * out[interior] = init[interior] + dt_stage * tend. `out` may alias `init`. */
int pmw_rk_update(const double* state_init, const double* tend,
                  double* state_out, int nx, int nz, double dt_stage,
                  void* stream);

/* ---- fused hot path ---------------------------------------------------- */

/* compute_tendencies_x: halo + reconstruction + flux + tendency in one
 * kernel, tend (4, nz, nx) written, nothing else materialised. Same inputs
 * and arithmetic as set_bc_x + interpolate_x + compute_flux_x +
 * compute_tend_x (+ add_source_terms when src_w != NULL). Writes the x halos
 * of `state_forcing` like set_bc_x. */
int pmw_tendencies_x(double* state_forcing, const double* hy_dens_cell,
                     const double* hy_dens_theta_cell, const double* src_w,
                     double* tend, int nx, int nz, double dx, double hv_coeff,
                     void* stream);

/* compute_tendencies_z: the z counterpart (set_bc_z + interpolate_z +
 * compute_flux_z + compute_tend_z [+ source]). */
int pmw_tendencies_z(double* state_forcing, const double* hy_dens_cell,
                     const double* hy_dens_int,
                     const double* hy_dens_theta_int,
                     const double* hy_pressure_int, const double* src_w,
                     double* tend, int nx, int nz, double dz, double hv_coeff,
                     void* stream);

/* One RK stage in one kernel: replaces discrete_step
 * (pyminiweather/solve/step.py:21-82) for one direction (0 = x, 1 = z):
 * halo of state_forcing, reconstruction, flux, tendency, optional source,
 * and out[interior] = init[interior] + dt_stage * tend.
 * `state_out` must not alias `state_forcing` (PMW_EALIAS): the reference's
 * aliased stage 2 (step.py:122-131) is served by ping-ponging buffers or by
 * pmw_tendencies_* + pmw_rk_update. `state_init` may alias either.
 * hy: x uses hy_dens_cell, hy_dens_theta_cell; z uses hy_dens_cell (halo),
 * hy_dens_int, hy_dens_theta_int, hy_pressure_int. Unused ones may be NULL.
 * d = dx or dz. bc/inj_rows as in pmw_set_bc_x (x only). */
int pmw_rk_stage(int direction, const double* state_init,
                 double* state_forcing, double* state_out,
                 const double* hy_dens_cell, const double* hy_dens_theta_cell,
                 const double* hy_dens_int, const double* hy_dens_theta_int,
                 const double* hy_pressure_int, const double* src_w, int bc,
                 const int* inj_rows, int nx, int nz, double d,
                 double dt_stage, double hv_coeff, void* stream);

/* Replaces evolve (pyminiweather/solve/step.py:85-143): `nsteps` timesteps,
 * each two directions x three RK stages, one fused launch per direction
 * (pmw_evolve_ex documents the kernel choice). Direction order is
 * Z then X when *reverse_direction == 0, X then Z otherwise; the flag is
 * flipped after every step (step.py:143). `state_work` is a third
 * state-shaped buffer (the first direction's result, or the stage-1 output)
 * so that no launch reads and writes the same buffer; on return `state` and
 * `state_tmp` hold exactly what the reference's evolve leaves in them. */
int pmw_evolve(double* state, double* state_tmp, double* state_work,
               const double* hy_dens_cell, const double* hy_dens_theta_cell,
               const double* hy_dens_int, const double* hy_dens_theta_int,
               const double* hy_pressure_int, const double* src_w, int bc,
               const int* inj_rows, int nx, int nz, double dx, double dz,
               double dt, double hv_coeff_x, double hv_coeff_z,
               int* reverse_direction, int nsteps, void* stream);

/* pmw_evolve with explicit kernel choice and optional per-launch timing.
 * `tuning` (HOST, 3 ints, NULL = defaults {0, 512, 8096}):
 *   [0] mode: 0 = one fused launch per direction (all three RK stages,
 *       periodic x only; the injection BC always uses mode 1), 1 = one
 *       launch per RK stage (the pmw_rk_stage kernels);
 *   [1] fused x tiles: 256, 448 or 512 threads per block, one row per block
 *       (tile = that minus 12 cells); 1256 / 1512 = dense tiles of 256 /
 *       512 threads over the rows laid end to end (each row with its 6
 *       periodic cells per side), for rows narrower than a block;
 *   [2] fused z kernel: 8, 16, 24, 32, 48 or 64 = rows per block of the
 *       marching kernel (one wave per RK stage); 8064 / 4128 = LDS tiles of
 *       8 columns x 64 rows / 4 columns x 128 rows (all stages in LDS);
 *       8096 = LDS tiles of 8 columns x 96 rows, one cell per thread, the
 *       RK init in registers and the stage state updated in place;
 *       8128 = LDS tiles of 8 columns x 128 rows, two cells per thread;
 *       18128 / 18096 = the 8128 / 8096 tiles over the column groups
 *       laid end to end (each column with its two wall halo rows per side).
 * Both modes leave bit-identical `state` and `state_tmp`.
 * `kernel_ms` (HOST, 2 doubles, or NULL): if given, a HIP event is recorded
 * around every launch, the call blocks until they finish, and each launch's
 * duration is added to kernel_ms[0] (x launches) or kernel_ms[1] (z). For
 * measurement only (creates and destroys events). No reference counterpart
 * beyond evolve itself. */
int pmw_evolve_ex(double* state, double* state_tmp, double* state_work,
                  const double* hy_dens_cell, const double* hy_dens_theta_cell,
                  const double* hy_dens_int, const double* hy_dens_theta_int,
                  const double* hy_pressure_int, const double* src_w, int bc,
                  const int* inj_rows, int nx, int nz, double dx, double dz,
                  double dt, double hv_coeff_x, double hv_coeff_z,
                  int* reverse_direction, int nsteps, const int* tuning,
                  void* stream, double* kernel_ms);

/* ---- ensembles (batched, uneven grids) --------------------------------- */

/* One member of an ensemble: its own Fields-like buffers and grid. All
 * pointers are device pointers; src_w may be NULL. No reference counterpart
 * (the reference advances one grid per process). */
typedef struct pmw_member {
  double* state;
  double* state_tmp;
  double* state_work;
  const double* hy_dens_cell;
  const double* hy_dens_theta_cell;
  const double* hy_dens_int;
  const double* hy_dens_theta_int;
  const double* hy_pressure_int;
  const double* src_w;
  int nx;
  int nz;
  double dx;
  double dz;
  double dt;       /* this member's timestep: stages use dt/3, dt/2, dt */
  double dt_full;  /* this member's params["dt"]: the hyperviscosity timestep */
} pmw_member;

/* Bytes of device workspace an ensemble of n members needs. */
long long pmw_ensemble_workspace_size(int nmembers);

/* Uploads the member descriptors and work-list offsets for `tuning` (same
 * format as pmw_evolve_ex; mode must be 0) into `workspace` and waits for the
 * copy. Call again whenever members or tuning change. The library records
 * (on the host, keyed by the workspace address) which members and tuning a
 * workspace holds. Rewriting a workspace that a captured HIP graph uses
 * changes what the graph's launches read: give each tuning its own
 * workspace instead. */
int pmw_ensemble_prepare(const pmw_member* members, int nmembers, const int* tuning,
                         void* workspace, long long workspace_bytes, void* stream);

/* evolve (pyminiweather/solve/step.py:85-143) for every member at once:
 * `nsteps` timesteps, each one launch per direction over the concatenated
 * tile lists of all members (the three RK stages fused as in
 * pmw_evolve_ex). `members` and `tuning` must be what `workspace` was last
 * prepared with: otherwise (or if it was never prepared) PMW_EINVAL and
 * nothing is launched. Only kernel launches are issued,
 * so a call can be captured in a HIP graph. Each member ends bit-identical
 * to pmw_evolve on that member alone. kernel_ms as in pmw_evolve_ex (each
 * launch is synchronised when given). */
int pmw_ensemble_evolve(const pmw_member* members, int nmembers, const int* tuning,
                        const void* workspace, int* reverse_direction, int nsteps,
                        void* stream, double* kernel_ms);

/* ---- adjoint (reverse mode) ------------------------------------------- */

/* Transpose of pmw_set_bc_x with the periodic BC (pyminiweather/ics/bcs.py:
 * 8-39), in place on a cotangent buffer `lam` (state layout): on interior
 * rows the cotangents of halo columns 0, 1, nx+2, nx+3 are added to the
 * columns they were copied from (nx, nx+1, 2, 3) and set to zero; every
 * other entry is left as is. No reference counterpart (the reference only
 * steps forward). */
int pmw_adjoint_set_bc_x(double* lam, int nx, int nz, void* stream);

/* Transpose of pmw_set_bc_z (pyminiweather/ics/bcs.py:67-148), in place:
 * on every column the cotangents of rows 0, 1 (nz+2, nz+3) are added to row
 * 2 (nz+1) with the forward copy's factor (1 for rho' and (rho*theta)',
 * hy[halo row] / hy[edge row] for rho*u, 0 for rho*w) and set to zero. No
 * reference counterpart. */
int pmw_adjoint_set_bc_z(double* lam, const double* hy_dens_cell, int nx, int nz,
                         void* stream);

/* Adjoint of one RK stage of discrete_step (pyminiweather/solve/step.py:
 * 21-82) in one direction (0 = x, 1 = z), excluding the init term:
 * lam_out = scale * T'(state_forcing)^T lam_in + acc1 + acc2 + acc3, where
 * T is halo + reconstruction + flux + tendency (+ gravity in z) and scale is
 * the stage dt. Cotangent buffers have the state layout; only interiors are
 * read and written. acc* may be NULL; lam_out may alias an acc but not
 * lam_in or state_forcing. Periodic x. With scale = 1 and no acc terms it is
 * the adjoint of pmw_tendencies_x / pmw_tendencies_z (halo transpose
 * included). No reference counterpart (the reference only steps forward). */
int pmw_adjoint_stage(int direction, const double* state_forcing, const double* lam_in,
                      double scale, const double* acc1, const double* acc2,
                      const double* acc3, double* lam_out, const double* hy_dens_cell,
                      const double* hy_dens_theta_cell, const double* hy_dens_int,
                      const double* hy_dens_theta_int, int nx, int nz, double d,
                      double hv_coeff, void* stream);

/* FP64 entries of one tape slot: `members` states (4, nz+4, nx+4) back to
 * back; 0 for invalid arguments. */
long long pmw_tape_slot_len(int nx, int nz, int members);

/* nsteps evolve steps (pyminiweather/solve/step.py:85-143, periodic x) of
 * `members` identical grids that share the hydrostatic arrays and src_w,
 * one fused launch per direction for all members. Slot 0 of `tape` holds
 * the start states; with store_stages, step k writes slots 6k+1 .. 6k+6:
 * stage 1 and stage 2 of its first direction, the state after it, stage 1
 * and stage 2 of the second direction, and the state after the step.
 * Without, only the end states are kept, in slot 6*nsteps (slots 0..2 serve
 * as scratch on the way). reverse_direction is the reference toggle at step
 * 0 (nonzero = x first); it alternates per step. tape_len >= (6*nsteps+1)
 * slots. tuning as pmw_evolve_ex (fused mode only; NULL = default). The
 * interiors equal pmw_evolve_ex's bit for bit. No reference counterpart. */
int pmw_tape_forward(double* tape, long long tape_len, int members,
                     const double* hy_dens_cell, const double* hy_dens_theta_cell,
                     const double* hy_dens_int, const double* hy_dens_theta_int,
                     const double* hy_pressure_int, const double* src_w, int nx, int nz,
                     double dx, double dz, double dt, double hv_coeff_x, double hv_coeff_z,
                     int reverse_direction, int nsteps, int store_stages, const int* tuning,
                     void* stream);

/* Reverse sweep over a tape written by pmw_tape_forward with store_stages:
 * on entry lam (one slot) is dL/d(states after step nsteps-1), on return
 * dL/d(start states) (interiors). Each step reverses its second direction,
 * then its first; each direction is one launch for all members that runs
 * the three stage adjoints (the arithmetic of pmw_adjoint_stage) in LDS.
 * workspace: 1 slot. tuning: NULL (= {256, 8512}) or {x block (256 | 384
 * threads), z tile (columns*1000 + threads: 4256, 4384, 2256, 8512)}. Every
 * choice gives the same result bit for bit. No reference counterpart. */
int pmw_tape_adjoint(const double* tape, long long tape_len, int members, double* lam,
                     double* workspace, long long workspace_len, const double* hy_dens_cell,
                     const double* hy_dens_theta_cell, const double* hy_dens_int,
                     const double* hy_dens_theta_int, int nx, int nz, double dx, double dz,
                     double dt, double hv_coeff_x, double hv_coeff_z, int reverse_direction,
                     int nsteps, const int* tuning, void* stream);

/* ---- reductions -------------------------------------------------------- */

/* Number of FP64 workspace entries pmw_compute_stats needs for an nx x nz
 * grid. No reference counterpart. */
long long pmw_stats_workspace_size(int nx, int nz);

/* Replaces compute_stats (pyminiweather/post/stats.py:8-35): writes
 * out2[0] = sum(rho) * dx * dz and out2[1] = sum(rho*(u^2+w^2) + rho*cv*T)
 * * dx * dz over the interior (device pointer, 2 doubles). Two-pass,
 * deterministic order. */
int pmw_compute_stats(const double* state, const double* hy_dens_cell,
                      const double* hy_dens_theta_cell, int nx, int nz,
                      double dx, double dz, double* workspace,
                      long long workspace_len, double* out2, void* stream);

/* pmw_compute_stats without a workspace, for callers that cannot allocate
 * device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). No counterpart in the reference. The
 * per-block partials go to a 16 KiB array in the library's own device memory
 * (one per GPU, held until the process ends), which covers every grid size:
 * pass 1 never uses more than 1024 blocks. Allocates nothing; the stream
 * (NULL = null stream) must belong to the current device, else PMW_EINVAL.
 * Same kernels and summation order as pmw_compute_stats, so out2 is bit for
 * bit the same. All calls on one GPU share the array: calls that overlap on
 * one GPU (different streams, or host threads) race and must be serialised
 * by the caller. */
int pmw_compute_stats_noalloc(const double* state, const double* hy_dens_cell,
                              const double* hy_dens_theta_cell, int nx, int nz,
                              double dx, double dz, double* out2, void* stream);

/* ---- initial conditions ------------------------------------------------ */

/* Replaces init (pyminiweather/ics/initial.py:11-105) with the quadrature
 * evaluated on the device: fills `state` (4, nz+4, nx+4) and the five
 * hydrostatic arrays; for PMW_IC_GRAVITY also `src_w` (nz, nx), the
 * w-momentum source wpert * hy_dens_cell of solve/source.py:41-50.
 * `geometry` is a HOST array of 16 doubles: xlen, zlen, dx, dz, then
 * (start, step) pairs of numpy.linspace for the cell left edges in x and z
 * incl. halos (mesh.py:37-49), the vertical cell centres incl. halos
 * (mesh.py:104-110), the vertical cell edges (mesh.py:88-90) and the
 * interior cell centres in x and z (mesh.py:63-69). `bubbles` is a HOST
 * array of nbubbles (<= 4) x (amplitude, x0, z0, xrad, zrad) cos^2 theta'
 * bubbles (NULL = the reference's fixed bubbles for ic_kind,
 * ics/initial_conditions.py:131-132, 197, 266). */
int pmw_init(int ic_kind, double* state, double* hy_dens_cell,
             double* hy_dens_theta_cell, double* hy_dens_int,
             double* hy_dens_theta_int, double* hy_pressure_int,
             double* src_w, int nx, int nz, const double* geometry,
             const double* bubbles, int nbubbles, void* stream);

/* Replaces the last line of init (pyminiweather/ics/initial.py:105):
 * hy_pressure_int[k] = C0 * hy_dens_theta_int[k]^gamma for k = 0..nz, with
 * the device pow, i.e. the same pow the z flux uses to form p - hy_pressure,
 * so a fluid at rest stays exactly at rest. */
int pmw_hydro_pressure(const double* hy_dens_theta_int,
                       double* hy_pressure_int, int nz, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* PMW_H */
