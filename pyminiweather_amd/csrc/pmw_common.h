// pmw_common.h -- constants and pointwise device arithmetic shared by the
// PyMiniWeather HIP kernels (gfx950).
//
// Every expression keeps the reference's floating-point operation order, and
// the library is compiled with -ffp-contract=off, so that a face flux or a
// cell update rounds exactly like the NumPy code it replaces. The only
// expected differences from NumPy are last-bit differences of the powers
// (pow_pos of pmw_pow.h is correctly rounded in all but ~1e-6 of cases,
// glibc's pow, which NumPy calls, in all but 8e-4) and of exp/cos in the
// initial conditions.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "pmw.h"
#include "pmw_pow.h"

namespace pmw {

// Device-global pointers. A pointer read from memory (a descriptor array)
// is generic to the compiler, which then emits flat_* accesses that also
// count against the LDS wait counter; typed in the global address space it
// emits global_* instead. Host code sees plain pointers (same layout).
#if defined(__HIP_DEVICE_COMPILE__)
#define PMW_GLOBAL __attribute__((address_space(1)))
#else
#define PMW_GLOBAL
#endif
using gdouble = PMW_GLOBAL double;
template <typename T>
__host__ __device__ inline PMW_GLOBAL T* gptr(T* p) {
  return (PMW_GLOBAL T*)p;
}

// A state element addressed from a block-uniform base (the start of a row
// or tile rows of one variable plane, kept in scalar registers) by a 32-bit
// byte offset per lane: one global access with a scalar base, instead of a
// 64-bit address computed per lane and access.
// The base goes through readfirstlane so that the compiler keeps it in
// scalar registers instead of folding the variable's plane offset into a
// per-lane 64-bit address.
__device__ __forceinline__ unsigned long long sgpr64(unsigned long long u) {
  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)u);
  const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)(u >> 32));
  return ((unsigned long long)hi << 32) | lo;
}
__device__ __forceinline__ double ld_off(const gdouble* base, unsigned boff) {
  const unsigned long long b = sgpr64((unsigned long long)base);
  return *(const gdouble*)((const PMW_GLOBAL char*)b + boff);
}
__device__ __forceinline__ void st_off(gdouble* base, unsigned boff, double v) {
  const unsigned long long b = sgpr64((unsigned long long)base);
  *(gdouble*)((PMW_GLOBAL char*)b + boff) = v;
}

// pyminiweather/data/constants.py:13-27 (rd is 287 in a float Enum -> 287.0)
constexpr double kHvBeta = 0.05;
constexpr double kP0 = 1.0e5;
constexpr double kC0 = 27.5629410929725921310572974482;
constexpr double kGamma = 1.40027894002789400278940027894;
constexpr double kGrav = 9.8;
constexpr double kCp = 1004.0;
constexpr double kCv = 717.0;
constexpr double kRd = 287.0;
constexpr double kPi = 3.14159265358979323846264338327;
constexpr double kTheta0 = 300.0;
constexpr double kExner0 = 1.0;

// Convolution weights in the order the direct correlation applies them
// (the stored kernels of data/fields.py:94-97, reversed).
constexpr double kW4a = -1.0 / 12;  // taps 0 and 3 of the 4th-order kernel
constexpr double kW4b = 7.0 / 12;   // taps 1 and 2
// 3rd-difference kernel [1,-3,3,-1] reversed: [-1, 3, -3, 1]

constexpr int kHs = 2;
constexpr int kNv = 4;

// Index helpers for a (4, nz+4, nx+4) state.
struct Grid {
  int nx, nz;
  int ld;          // nx + 4
  long long plane; // (nz + 4) * (nx + 4)
  __host__ __device__ Grid(int nx_, int nz_)
      : nx(nx_), nz(nz_), ld(nx_ + 4), plane((long long)(nz_ + 4) * (nx_ + 4)) {}
  __host__ __device__ long long at(int v, int prow, int pcol) const {
    return (long long)v * plane + (long long)prow * ld + pcol;
  }
};

// a / b from y = RN(1/b): q = RN(a*y) is within an ulp of a/b and the
// residual a - q*b is exact under fma, so q + r*y rounds to the IEEE
// quotient (Markstein's correction; valid while a/b, a*y and the residual
// neither overflow nor underflow, as for the fluxes and densities here).
// Three instructions instead of a full division.
__device__ __forceinline__ double div_rn(double a, double b, double y) {
  const double q = a * y;
  const double r = fma(-q, b, a);
  return fma(r, y, q);
}

// 4-tap reconstruction at one face: value and 3rd difference, accumulated
// left to right like scipy.signal.convolve's direct method.
__device__ __forceinline__ void recon4(double s0, double s1, double s2,
                                       double s3, double& val, double& d3) {
  val = ((s0 * kW4a + s1 * kW4b) + s2 * kW4b) + s3 * kW4a;
  d3 = ((s0 * -1.0 + s1 * 3.0) + s2 * -3.0) + s3 * 1.0;
}

// The pressure power (rho*theta)^gamma of a face: POW_PLAIN evaluates
// pow_pos; POW_FIRST also keeps its unrounded value in `ref` (the first RK
// stage of a fused direction); POW_NEXT continues from `ref` (the later
// stages, whose face values differ from the first stage's by a relative
// ~1e-4 or less: pow_pos_near, with pow_pos for larger changes); POW_TRY is
// POW_NEXT without the pow_pos: *ok is set false where the continuation
// could round differently from pow_pos (about 4e-5 of faces), and the
// caller recomputes such a face with POW_PLAIN after its other faces, so
// that the common path has no pow_pos inside it.
enum PowMode { POW_PLAIN = 0, POW_FIRST = 1, POW_NEXT = 2, POW_TRY = 3 };

__device__ __forceinline__ double face_power(double q, int mode, PowRef* ref, bool* ok) {
  if (mode == POW_PLAIN) return pow_pos(q, kGamma);
  double r;
  if (mode == POW_TRY) {
    *ok = pow_pos_near_try(q, kGamma, *ref, r);
    return r;
  }
  if (mode == POW_NEXT && pow_pos_near_try(q, kGamma, *ref, r)) return r;
  // first stage, or a continuation too close to a rounding boundary:
  // pow_pos; only the first stage keeps the reference
  PowRef scratch;
  return pow_pos_ref(q, kGamma, mode == POW_FIRST ? *ref : scratch);
}

// x-direction Euler flux with hyperviscosity (solve/interpolate.py:105-129).
__device__ __forceinline__ void flux_x_face(const double val[4],
                                            const double d3[4], double hyr,
                                            double hyt, double hv,
                                            double F[4], int pmode = POW_PLAIN,
                                            PowRef* ref = nullptr, bool* ok = nullptr) {
  const double rho = val[0] + hyr;
  const double ir = 1.0 / rho;
  const double u = div_rn(val[1], rho, ir);
  const double w = div_rn(val[2], rho, ir);
  const double t = div_rn(val[3] + hyt, rho, ir);
  const double p = kC0 * face_power(rho * t, pmode, ref, ok);
  const double ru = rho * u;
  F[0] = ru - hv * d3[0];
  F[1] = (rho * (u * u) + p) - hv * d3[1];
  F[2] = ru * w - hv * d3[2];
  F[3] = ru * t - hv * d3[3];
}

// z-direction flux (solve/interpolate.py:153-186); `wall` marks faces k = 0
// and k = nz, where w and the rho' 3rd difference are zeroed.
__device__ __forceinline__ void flux_z_face(const double val[4], double d3[4],
                                            double hyr, double hyt,
                                            double hyp, double hv, bool wall,
                                            double G[4], int pmode = POW_PLAIN,
                                            PowRef* ref = nullptr, bool* ok = nullptr) {
  const double rho = val[0] + hyr;
  const double ir = 1.0 / rho;
  const double u = div_rn(val[1], rho, ir);
  double w = div_rn(val[2], rho, ir);
  const double t = div_rn(val[3] + hyt, rho, ir);
  const double p = kC0 * face_power(rho * t, pmode, ref, ok) - hyp;
  if (wall) {
    w = 0.0;
    d3[0] = 0.0;
  }
  const double rw = rho * w;
  G[0] = rw - hv * d3[0];
  G[1] = rw * u - hv * d3[1];
  G[2] = (rho * (w * w) + p) - hv * d3[2];
  G[3] = rw * t - hv * d3[3];
}

// Tendency from two face fluxes (solve/interpolate.py:209-215, 239-245).
// id = RN(1/d).
__device__ __forceinline__ double tend_of(double f_lo, double f_hi, double d, double id) {
  return div_rn(-(f_hi - f_lo), d, id);
}

// Value the reference's set_bc_z leaves in z-halo row `prow` (0, 1, nz+2 or
// nz+3) of variable v, given the same column's value on the adjacent interior
// edge row (bcs.py:92-148).
__device__ __forceinline__ double wall_halo(int v, double edge_val,
                                            double hy_edge, double hy_halo) {
  if (v == 2) return 0.0;
  if (v == 1) return edge_val / hy_edge * hy_halo;
  return edge_val;
}

}  // namespace pmw

#define PMW_CHECK_LAUNCH()                         \
  do {                                             \
    hipError_t e_ = hipGetLastError();             \
    if (e_ != hipSuccess) return (int)e_;          \
  } while (0)
