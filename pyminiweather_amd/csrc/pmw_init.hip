// pmw_init.hip -- initial conditions on the device (ics/initial.py:11-105).
//
// One thread per padded cell evaluates the IC at the 3x3 Gauss-Legendre
// points and accumulates the weighted sums in NumPy's order (inner sum over
// the x points, then over the z points, each ((a+b)+c)). A second kernel fills
// the five 1-D hydrostatic profiles and, for the gravity-wave case, the
// precomputed w-momentum source. Coordinates are rebuilt as numpy.linspace
// builds them: i * step + start, with start/step computed by the caller.
#include "pmw_common.h"

using namespace pmw;

namespace {

constexpr double kQp[3] = {0.112701665379258311482073460022, 0.5,
                           0.887298334620741688517926539980};
constexpr double kQw[3] = {0.277777777777777777777777777779,
                           0.444444444444444444444444444444,
                           0.277777777777777777777777777779};
constexpr double kBv = 0.02;  // Brunt-Vaisala frequency of the gravity IC
constexpr int kMaxBubbles = 4;

struct Geometry {
  double xlen, zlen, dx, dz;
  double xe0, xes, ze0, zes;    // cell left edges incl. halos (mesh.py:37-49)
  double zc0, zcs;              // vertical cell centres incl. halos (mesh.py:104-110)
  double zi0, zis;              // vertical cell edges (mesh.py:88-90)
  double xcc0, xccs, zcc0, zccs;  // interior cell centres (mesh.py:63-69)
};

struct Bubbles {
  int n;
  double p[kMaxBubbles][5];  // amplitude, x0, z0, xrad, zrad
};

// initial_conditions.py:26-52
__device__ __forceinline__ void hydro_theta(double z, double& hr, double& ht) {
  ht = kTheta0;
  const double exner = kExner0 - kGrav * z / (kCp * kTheta0);
  const double p = kP0 * pow(exner, kCp / kRd);
  hr = pow(p / kC0, 1.0 / kGamma) / ht;
}

// initial_conditions.py:54-81
__device__ __forceinline__ void hydro_bvfreq(double z, double& hr, double& ht) {
  ht = kTheta0 * exp(kBv * kBv / kGrav * z);
  const double exner =
      kExner0 - kGrav * kGrav / (kCp * kBv * kBv) * (ht - kTheta0) / (ht * kTheta0);
  const double p = kP0 * pow(exner, kCp / kRd);
  hr = pow(p / kC0, 1.0 / kGamma) / ht;
}

// utils/utils.py:42-49
__device__ __forceinline__ double ellipse_cos2(double x, double z, const double* b) {
  const double a = (x - b[1]) / b[3];
  const double c = (z - b[2]) / b[4];
  const double dist = sqrt(a * a + c * c) * kPi / 2.0;
  if (!(dist <= kPi / 2.0)) return 0.0;
  const double cs = cos(dist);
  return b[0] * (cs * cs);
}

__global__ void k_init_state(int ic, double* __restrict__ s, Grid g, Geometry geo,
                             Bubbles bub) {
  const long long n = (long long)(g.nz + 4) * g.ld;
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n) return;
  const int pr = (int)(idx / g.ld), pc = (int)(idx % g.ld);
  const double xe = (double)pc * geo.xes + geo.xe0;
  const double ze = (double)pr * geo.zes + geo.ze0;
  double acc[4][3];
  for (int a = 0; a < 3; ++a) {  // z quadrature index (outer sum)
    const double z = ze + kQp[a] * geo.dz;
    double hr, ht;
    if (ic == PMW_IC_GRAVITY)
      hydro_bvfreq(z, hr, ht);
    else
      hydro_theta(z, hr, ht);
    double row[4][3];
    for (int b = 0; b < 3; ++b) {  // x quadrature index (inner sum)
      const double x = xe + kQp[b] * geo.dx;
      const double wt = kQw[a] * kQw[b];
      const double r = 0.0, w = 0.0;
      const double u = (ic == PMW_IC_GRAVITY) ? 15.0 : 0.0;
      double t = 0.0;
      for (int j = 0; j < bub.n; ++j) t = t + ellipse_cos2(x, z, bub.p[j]);
      row[0][b] = r * wt;
      row[1][b] = ((r + hr) * u) * wt;
      row[2][b] = ((r + hr) * w) * wt;
      row[3][b] = ((r + hr) * (t + ht) - hr * ht) * wt;
    }
    for (int v = 0; v < 4; ++v) acc[v][a] = (row[v][0] + row[v][1]) + row[v][2];
  }
  for (int v = 0; v < 4; ++v)
    s[g.at(v, pr, pc)] = (acc[v][0] + acc[v][1]) + acc[v][2];
}

__global__ void k_init_hydro(int ic, double* __restrict__ hy_dens_cell,
                             double* __restrict__ hy_dens_theta_cell,
                             double* __restrict__ hy_dens_int,
                             double* __restrict__ hy_dens_theta_int,
                             double* __restrict__ hy_pressure_int, int nz,
                             Geometry geo) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  double hr, ht;
  if (k < nz + 4) {
    const double z = (double)k * geo.zcs + geo.zc0;
    if (ic == PMW_IC_GRAVITY) hydro_bvfreq(z, hr, ht); else hydro_theta(z, hr, ht);
    hy_dens_cell[k] = hr * ((kQw[0] + kQw[1]) + kQw[2]);
    const double th = ht * hr;
    hy_dens_theta_cell[k] = (th * kQw[0] + th * kQw[1]) + th * kQw[2];
  }
  if (k < nz + 1) {
    const double z = (double)k * geo.zis + geo.zi0;
    if (ic == PMW_IC_GRAVITY) hydro_bvfreq(z, hr, ht); else hydro_theta(z, hr, ht);
    hy_dens_int[k] = hr;
    hy_dens_theta_int[k] = hr * ht;
    hy_pressure_int[k] = kC0 * pow_pos(hr * ht, kGamma);
  }
}

// solve/source.py:41-50: wpert on the interior cell centres times
// hy_dens_cell of the row.
__global__ void k_init_source_w(double* __restrict__ src, const double* __restrict__ hyr,
                                int nx, int nz, Geometry geo) {
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long long)nx * nz) return;
  const int k = (int)(idx / nx), i = (int)(idx % nx);
  const double x = (double)i * geo.xccs + geo.xcc0;
  const double z = (double)k * geo.zccs + geo.zcc0;
  const double b[5] = {0.01, geo.xlen / 8, 1000.0, 500.0, 500.0};
  src[idx] = ellipse_cos2(x, z, b) * hyr[k + kHs];
}

__global__ void k_hydro_pressure(const double* __restrict__ hy_dens_theta_int,
                                 double* __restrict__ hy_pressure_int, int n) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < n) hy_pressure_int[k] = kC0 * pow_pos(hy_dens_theta_int[k], kGamma);
}

}  // namespace

extern "C" {

int pmw_hydro_pressure(const double* hy_dens_theta_int, double* hy_pressure_int,
                       int nz, void* stream) {
  if (nz < 1 || !hy_dens_theta_int || !hy_pressure_int) return PMW_EINVAL;
  k_hydro_pressure<<<(unsigned)((nz + 1 + 255) / 256), 256, 0, (hipStream_t)stream>>>(
      hy_dens_theta_int, hy_pressure_int, nz + 1);
  PMW_CHECK_LAUNCH();
  return PMW_OK;
}


int pmw_init(int ic_kind, double* state, double* hy_dens_cell,
             double* hy_dens_theta_cell, double* hy_dens_int,
             double* hy_dens_theta_int, double* hy_pressure_int, double* src_w,
             int nx, int nz, const double* geometry, const double* bubbles,
             int nbubbles, void* stream) {
  if (nx < 1 || nz < 1 || !state || !hy_dens_cell || !hy_dens_theta_cell ||
      !hy_dens_int || !hy_dens_theta_int || !hy_pressure_int || !geometry)
    return PMW_EINVAL;
  if (ic_kind < PMW_IC_THERMAL || ic_kind > PMW_IC_INJECTION) return PMW_EINVAL;
  if (nbubbles < 0 || nbubbles > kMaxBubbles || (nbubbles > 0 && !bubbles))
    return PMW_EINVAL;
  if (ic_kind == PMW_IC_GRAVITY && !src_w) return PMW_EINVAL;
  Geometry geo;
  {
    double* dst = &geo.xlen;
    for (int j = 0; j < 16; ++j) dst[j] = geometry[j];
  }
  Bubbles bub{};
  const double xlen = geo.xlen;
  if (bubbles) {
    bub.n = nbubbles;
    for (int j = 0; j < nbubbles; ++j)
      for (int q = 0; q < 5; ++q) bub.p[j][q] = bubbles[5 * j + q];
  } else {
    // the reference's fixed bubbles (initial_conditions.py:131-132, 197, 266)
    const double thermal[5] = {3.0, xlen / 2, 2000.0, 2000.0, 2000.0};
    const double hot[5] = {20.0, xlen / 2, 2000.0, 2000.0, 2000.0};
    const double cold[5] = {-20.0, xlen / 2, 8000.0, 2000.0, 2000.0};
    const double dc[5] = {-20.0, xlen / 2, 5000.0, 4000.0, 2000.0};
    auto put = [&](const double* p) {
      for (int q = 0; q < 5; ++q) bub.p[bub.n][q] = p[q];
      bub.n++;
    };
    if (ic_kind == PMW_IC_THERMAL) put(thermal);
    if (ic_kind == PMW_IC_COLLISION) { put(hot); put(cold); }
    if (ic_kind == PMW_IC_DENSITY_CURRENT) put(dc);
  }
  hipStream_t st = (hipStream_t)stream;
  Grid g(nx, nz);
  const long long n = (long long)(nz + 4) * (nx + 4);
  k_init_state<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(ic_kind, state, g, geo, bub);
  PMW_CHECK_LAUNCH();
  k_init_hydro<<<(unsigned)((nz + 4 + 255) / 256), 256, 0, st>>>(
      ic_kind, hy_dens_cell, hy_dens_theta_cell, hy_dens_int, hy_dens_theta_int,
      hy_pressure_int, nz, geo);
  PMW_CHECK_LAUNCH();
  if (ic_kind == PMW_IC_GRAVITY) {
    const long long m = (long long)nx * nz;
    k_init_source_w<<<(unsigned)((m + 255) / 256), 256, 0, st>>>(src_w, hy_dens_cell,
                                                                nx, nz, geo);
    PMW_CHECK_LAUNCH();
  }
  return PMW_OK;
}

}  // extern "C"
