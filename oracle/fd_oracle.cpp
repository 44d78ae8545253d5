// CPU restatement of astronomix's finite-difference MHD step (HOW-MHD:
// characteristic WENO-5 fluxes, high-order constrained transport, SSPRK(5,4)).
//
// TEST INFRASTRUCTURE ONLY: the checker of the HIP FD kernels and a CPU
// timing reference. Nothing in the product links or calls it.
//
// Restates, templated on the floating type, the reference of
// leo1200/astronomix v0.5.1 (paths under astronomix/_finite_difference/):
//   * conserved_state_from_primitive_mhd / primitive_state_from_conserved_mhd
//                                   _fluid_equations/_equations.py:65-153
//   * _mhd_flux_x                   _fluid_equations/_fluxes.py:20-69
//   * _eigenvalue_building_blocks / _eigen_lambdas
//                                   _fluid_equations/_eigen.py:27-114, 949-993
//   * _eigenvector_building_blocks  _fluid_equations/_eigen.py:116-337
//   * _eigen_R_col / _eigen_L_row   _fluid_equations/_eigen.py:340-622, 625-911
//   * _weno_flux_x/_y/_z            _interface_fluxes/_weno.py:111-324
//   * constrained_transport_rhs, update_cell_center_fields,
//     initialize_interface_fields   _magnetic_update/_constrained_transport.py:43-213
//   * finite_difference_int6, _interface_field_divergence
//                                   _maths/_differencing.py:8-51
//   * interp_center_to_face, interp_face_to_center, point_values_to_averages
//                                   _maths/_interpolate.py:8-62
//   * _enforce_positivity           _fluid_equations/_enforce_positivity.py:21-59
//   * _ssprk4_with_ct               _time_integrators/_ssprk.py:177-348
//   * _evolve_state_fd              _state_evolution/_evolve_state.py:24-72
//   * _cfl_time_step_fd             _timestep_estimation/_timestep_estimator.py:25-91
//
// The reference evaluates the y and z fluxes by transposing the state and
// swapping (m_x, m_dir), (B_x, B_dir), then applying the x routine. That is
// restated here as a per-direction component permutation of an interface's
// stencil (no transposes). Boundaries are periodic (jnp.roll semantics).
//
// Conserved rows: 0 rho, 1-3 momentum, 4 E, 5-7 B (registry of a 3D MHD
// state); face fields bx, by, bz hold B at (i+1/2, j, k), (i, j+1/2, k),
// (i, j, k+1/2).

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace {

template <typename T>
struct G3 {
  int n[3];
  long nc;
  long st[3];
  T gamma, dx, rhomin, pgmin;
  int positivity;
  long nb(long c, int axis, int s) const {
    const long i = axis == 0 ? c / st[0] : (c / st[axis]) % n[axis];
    long j = (i + s) % n[axis];
    if (j < 0) j += n[axis];
    return c + (j - i) * st[axis];
  }
};

template <typename T>
G3<T> make_g3(int nx, int ny, int nz, double gamma, double dx, double rhomin, double pgmin, int positivity) {
  G3<T> g;
  g.n[0] = nx;
  g.n[1] = ny;
  g.n[2] = nz;
  g.nc = (long)nx * ny * nz;
  g.st[2] = 1;
  g.st[1] = nz;
  g.st[0] = (long)ny * nz;
  g.gamma = (T)gamma;
  g.dx = (T)dx;
  g.rhomin = (T)rhomin;
  g.pgmin = (T)pgmin;
  g.positivity = positivity;
  return g;
}

// component permutation of the direction's frame (the reference's swaps)
const int PERM[3][8] = {{0, 1, 2, 3, 4, 5, 6, 7}, {0, 2, 1, 3, 4, 6, 5, 7}, {0, 3, 2, 1, 4, 7, 6, 5}};

template <typename T>
inline T dss(T x) {  // diff_safe_sqrt, _eigen.py:21-24
  return std::sqrt(std::max(x, T(1e-12)));
}

// _mhd_flux_x on one cell (q in the direction's frame), _fluxes.py:20-69
template <typename T>
void mhd_flux(const T* q, T* f, const G3<T>& g) {
  T rho = q[0];
  const T vx = q[1] / rho, vy = q[2] / rho, vz = q[3] / rho;
  const T Bx = q[5], By = q[6], Bz = q[7];
  const T u2 = (vx * vx + vy * vy) + vz * vz;
  const T b2 = (Bx * Bx + By * By) + Bz * Bz;
  T p = (g.gamma - T(1)) * ((q[4] - (T(0.5) * rho) * u2) - T(0.5) * b2);
  if (g.positivity) {
    rho = std::max(rho, g.rhomin);
    p = std::max(p, g.pgmin);
  }
  const T Ptot = p + T(0.5) * b2;
  const T vB = (vx * Bx + vy * By) + vz * Bz;
  const T E = (p / (g.gamma - T(1)) + (T(0.5) * rho) * u2) + T(0.5) * b2;
  f[0] = rho * vx;
  f[1] = ((rho * (vx * vx)) + Ptot) - Bx * Bx;
  f[2] = ((rho * vx) * vy) - Bx * By;
  f[3] = ((rho * vx) * vz) - Bx * Bz;
  f[4] = (E + Ptot) * vx - vB * Bx;
  f[5] = 0;
  f[6] = By * vx - Bx * vy;
  f[7] = Bz * vx - Bx * vz;
}

// cell-centred floors shared by both building-block routines
template <typename T>
void floored_cell(const T* q, const G3<T>& g, T& rho, T& vx, T& vy, T& vz, T& v2, T& b2, T& p, T& E) {
  rho = q[0];
  vx = q[1] / rho;
  vy = q[2] / rho;
  vz = q[3] / rho;
  v2 = (vx * vx + vy * vy) + vz * vz;
  b2 = (q[5] * q[5] + q[6] * q[6]) + q[7] * q[7];
  E = q[4];
  p = (g.gamma - T(1)) * (E - T(0.5) * (rho * v2 + b2));
  if ((rho < g.rhomin) || (p < g.pgmin)) rho = std::max(rho, g.rhomin);
  if ((rho < g.rhomin) || (p < g.pgmin)) p = std::max(p, g.pgmin);
  if ((rho < g.rhomin) || (p < g.pgmin)) E = p / (g.gamma - T(1)) + T(0.5) * (rho * v2 + b2);
}

// _eigenvalue_building_blocks (cell centred), _eigen.py:27-114 -> 7 lambdas
template <typename T>
void cell_lambdas(const T* q, const G3<T>& g, T* lam) {
  T rho, vx, vy, vz, v2, b2, p, E;
  floored_cell(q, g, rho, vx, vy, vz, v2, b2, p, E);
  const T Bx = q[5];
  const T cs2 = std::max(T(0), g.gamma * std::fabs(p / rho));
  const T s = b2 / rho + cs2;
  const T droot = dss(std::max(T(0), s * s - ((T(4) * (Bx * Bx)) / rho) * cs2));
  const T cf = dss(std::max(T(0), T(0.5) * (s + droot)));
  const T ca = dss(std::max(T(0), (Bx * Bx) / rho));
  const T csl = dss(std::max(T(0), T(0.5) * (s - droot)));
  lam[0] = vx - cf;
  lam[1] = vx - ca;
  lam[2] = vx - csl;
  lam[3] = vx;
  lam[4] = vx + csl;
  lam[5] = vx + ca;
  lam[6] = vx + cf;
}

template <typename T>
struct EB {
  T rho, sqrt_rho, vx, vy, vz, v2, bx, by, bz, bty, btz, sgn_bx, sgn_bt, cs, cs2, cs2inv, cf, ca, csl, af, as, gam0,
      gam1, gam2;
  bool cs_ge_ca;
};

// _eigenvector_building_blocks at the interface between cells a and b,
// _eigen.py:116-337
template <typename T>
EB<T> interface_blocks(const T* qa, const T* qb, const G3<T>& g) {
  T r[2], vx[2], vy[2], vz[2], v2[2], b2[2], p[2], E[2], H[2];
  const T* qs[2] = {qa, qb};
  for (int k = 0; k < 2; ++k) {
    floored_cell(qs[k], g, r[k], vx[k], vy[k], vz[k], v2[k], b2[k], p[k], E[k]);
    H[k] = (E[k] + p[k]) / r[k];
  }
  auto avg = [](T a, T b) { return T(0.5) * (a + b); };
  EB<T> e;
  e.rho = avg(r[0], r[1]);
  e.vx = avg(vx[0], vx[1]);
  e.vy = avg(vy[0], vy[1]);
  e.vz = avg(vz[0], vz[1]);
  e.bx = avg(qa[5], qb[5]);
  e.by = avg(qa[6], qb[6]);
  e.bz = avg(qa[7], qb[7]);
  const T Hi = avg(H[0], H[1]);
  e.v2 = (e.vx * e.vx + e.vy * e.vy) + e.vz * e.vz;
  const T bsq = (e.bx * e.bx + e.by * e.by) + e.bz * e.bz;
  const T b2r = bsq / e.rho;
  const T bx2r = (e.bx * e.bx) / e.rho;
  e.cs2 = (g.gamma - T(1)) * (Hi - T(0.5) * (e.v2 + b2r));
  e.cs = dss(std::max(T(0), e.cs2));
  const T s = b2r + e.cs2;
  const T disc = s * s - (T(4) * bx2r) * e.cs2;
  const T droot = dss(std::max(T(0), disc));
  e.cf = dss(std::max(T(0), T(0.5) * (s + droot)));
  e.ca = dss(std::max(T(0), bx2r));
  e.csl = dss(std::max(T(0), T(0.5) * (s - droot)));
  T bt2 = e.by * e.by + e.bz * e.bz;
  e.sgn_bx = e.bx >= T(0) ? T(1) : T(-1);
  bt2 = std::max(bt2, T(1e-12));
  e.bty = bt2 >= T(1e-12) ? e.by / dss(bt2) : T(1) / dss(T(2));
  e.btz = bt2 >= T(1e-12) ? e.bz / dss(bt2) : T(1) / dss(T(2));
  T denom = e.cf * e.cf - e.csl * e.csl;
  denom = std::max(denom, T(1e-12));
  e.af = denom >= T(1e-12) ? dss(std::max(T(0), e.cs2 - e.csl * e.csl)) / dss(denom) : T(1);
  e.as = denom >= T(1e-12) ? dss(std::max(T(0), e.cf * e.cf - e.cs2)) / dss(denom) : T(1);
  e.sqrt_rho = dss(e.rho);
  e.gam0 = T(1) - g.gamma;
  e.gam1 = T(0.5) * (g.gamma - T(1));
  e.gam2 = (g.gamma - T(2)) / (g.gamma - T(1));
  e.cs2inv = e.cs2 > T(0) ? T(1) / e.cs2 : T(0);
  e.sgn_bt = e.by != T(0) ? (e.by >= T(0) ? T(1) : T(-1)) : (e.bz >= T(0) ? T(1) : T(-1));
  e.cs_ge_ca = e.cs >= e.ca;
  return e;
}

// _eigen_R_col, _eigen.py:340-622; R[8] over conserved rows (B_x row 0)
template <typename T>
void R_col(const EB<T>& e, int s, T* R) {
  for (int v = 0; v < 8; ++v) R[v] = 0;
  const T btv = e.bty * e.vy + e.btz * e.vz;
  switch (s) {
    case 0:
      R[0] = e.af;
      R[1] = e.af * (e.vx - e.cf);
      R[2] = e.af * e.vy + ((e.as * e.csl) * e.bty) * e.sgn_bx;
      R[3] = e.af * e.vz + ((e.as * e.csl) * e.btz) * e.sgn_bx;
      R[6] = ((e.cs * e.as) * e.bty) / e.sqrt_rho;
      R[7] = ((e.cs * e.as) * e.btz) / e.sqrt_rho;
      R[4] = e.af * (((e.cf * e.cf - e.cf * e.vx) + T(0.5) * e.v2) - e.gam2 * e.cs2) + ((e.as * e.csl) * btv) * e.sgn_bx;
      if (!e.cs_ge_ca)
        for (int v = 0; v < 8; ++v) R[v] *= e.sgn_bt;
      break;
    case 1:
      R[2] = -e.btz;
      R[3] = e.bty;
      R[6] = (-e.btz * e.sgn_bx) / e.sqrt_rho;
      R[7] = (e.bty * e.sgn_bx) / e.sqrt_rho;
      R[4] = e.bty * e.vz - e.btz * e.vy;
      break;
    case 2:
      R[0] = e.as;
      R[1] = e.as * (e.vx - e.csl);
      R[2] = e.as * e.vy - ((e.af * e.cf) * e.bty) * e.sgn_bx;
      R[3] = e.as * e.vz - ((e.af * e.cf) * e.btz) * e.sgn_bx;
      R[6] = ((-e.cs * e.af) * e.bty) / e.sqrt_rho;
      R[7] = ((-e.cs * e.af) * e.btz) / e.sqrt_rho;
      R[4] = e.as * (((e.csl * e.csl - e.csl * e.vx) + T(0.5) * e.v2) - e.gam2 * e.cs2) - ((e.af * e.cf) * btv) * e.sgn_bx;
      if (e.cs_ge_ca)
        for (int v = 0; v < 8; ++v) R[v] *= e.sgn_bt;
      break;
    case 3:
      R[0] = 1;
      R[1] = e.vx;
      R[2] = e.vy;
      R[3] = e.vz;
      R[4] = T(0.5) * e.v2;
      break;
    case 4:
      R[0] = e.as;
      R[1] = e.as * (e.vx + e.csl);
      R[2] = e.as * e.vy + ((e.af * e.cf) * e.bty) * e.sgn_bx;
      R[3] = e.as * e.vz + ((e.af * e.cf) * e.btz) * e.sgn_bx;
      R[6] = ((-e.cs * e.af) * e.bty) / e.sqrt_rho;
      R[7] = ((-e.cs * e.af) * e.btz) / e.sqrt_rho;
      R[4] = e.as * (((e.csl * e.csl + e.csl * e.vx) + T(0.5) * e.v2) - e.gam2 * e.cs2) + ((e.af * e.cf) * btv) * e.sgn_bx;
      if (e.cs_ge_ca)
        for (int v = 0; v < 8; ++v) R[v] *= e.sgn_bt;
      break;
    case 5:
      R[2] = -e.btz;
      R[3] = e.bty;
      R[6] = (e.btz * e.sgn_bx) / e.sqrt_rho;
      R[7] = (-e.bty * e.sgn_bx) / e.sqrt_rho;
      R[4] = e.bty * e.vz - e.btz * e.vy;
      break;
    case 6:
      R[0] = e.af;
      R[1] = e.af * (e.vx + e.cf);
      R[2] = e.af * e.vy - ((e.as * e.csl) * e.bty) * e.sgn_bx;
      R[3] = e.af * e.vz - ((e.as * e.csl) * e.btz) * e.sgn_bx;
      R[6] = ((e.cs * e.as) * e.bty) / e.sqrt_rho;
      R[7] = ((e.cs * e.as) * e.btz) / e.sqrt_rho;
      R[4] = e.af * (((e.cf * e.cf + e.cf * e.vx) + T(0.5) * e.v2) - e.gam2 * e.cs2) - ((e.as * e.csl) * btv) * e.sgn_bx;
      if (!e.cs_ge_ca)
        for (int v = 0; v < 8; ++v) R[v] *= e.sgn_bt;
      break;
  }
}

// _eigen_L_row, _eigen.py:625-911; L[8] over conserved rows (B_x entry 0)
template <typename T>
void L_row(const EB<T>& e, int s, T* L) {
  for (int v = 0; v < 8; ++v) L[v] = 0;
  const T btv = e.bty * e.vy + e.btz * e.vz;
  switch (s) {
    case 0:
      L[0] = e.af * (e.gam1 * e.v2 + e.cf * e.vx) - ((e.as * e.csl) * btv) * e.sgn_bx;
      L[1] = e.af * (e.gam0 * e.vx - e.cf);
      L[2] = (e.gam0 * e.af) * e.vy + ((e.as * e.csl) * e.bty) * e.sgn_bx;
      L[3] = (e.gam0 * e.af) * e.vz + ((e.as * e.csl) * e.btz) * e.sgn_bx;
      L[6] = (e.gam0 * e.af) * e.by + ((e.cs * e.as) * e.bty) * e.sqrt_rho;
      L[7] = (e.gam0 * e.af) * e.bz + ((e.cs * e.as) * e.btz) * e.sqrt_rho;
      L[4] = -e.gam0 * e.af;
      for (int v = 0; v < 8; ++v) L[v] = (T(0.5) * L[v]) * e.cs2inv;
      if (!e.cs_ge_ca)
        for (int v = 0; v < 8; ++v) L[v] *= e.sgn_bt;
      break;
    case 1:
      L[0] = e.btz * e.vy - e.bty * e.vz;
      L[2] = -e.btz;
      L[3] = e.bty;
      L[6] = (-e.btz * e.sgn_bx) * e.sqrt_rho;
      L[7] = (e.bty * e.sgn_bx) * e.sqrt_rho;
      for (int v = 0; v < 8; ++v) L[v] = T(0.5) * L[v];
      break;
    case 2:
      L[0] = e.as * (e.gam1 * e.v2 + e.csl * e.vx) + ((e.af * e.cf) * btv) * e.sgn_bx;
      L[1] = (e.gam0 * e.as) * e.vx - e.as * e.csl;
      L[2] = (e.gam0 * e.as) * e.vy - ((e.af * e.cf) * e.bty) * e.sgn_bx;
      L[3] = (e.gam0 * e.as) * e.vz - ((e.af * e.cf) * e.btz) * e.sgn_bx;
      L[6] = (e.gam0 * e.as) * e.by - ((e.cs * e.af) * e.bty) * e.sqrt_rho;
      L[7] = (e.gam0 * e.as) * e.bz - ((e.cs * e.af) * e.btz) * e.sqrt_rho;
      L[4] = -e.gam0 * e.as;
      for (int v = 0; v < 8; ++v) L[v] = (T(0.5) * L[v]) * e.cs2inv;
      if (e.cs_ge_ca)
        for (int v = 0; v < 8; ++v) L[v] *= e.sgn_bt;
      break;
    case 3:
      L[0] = -e.cs2 / e.gam0 - T(0.5) * e.v2;
      L[1] = e.vx;
      L[2] = e.vy;
      L[3] = e.vz;
      L[6] = e.by;
      L[7] = e.bz;
      L[4] = -1;
      for (int v = 0; v < 8; ++v) L[v] = ((-e.gam0) * L[v]) * e.cs2inv;
      break;
    case 4:
      L[0] = e.as * (e.gam1 * e.v2 - e.csl * e.vx) - ((e.af * e.cf) * btv) * e.sgn_bx;
      L[1] = e.as * (e.gam0 * e.vx + e.csl);
      L[2] = (e.gam0 * e.as) * e.vy + ((e.af * e.cf) * e.bty) * e.sgn_bx;
      L[3] = (e.gam0 * e.as) * e.vz + ((e.af * e.cf) * e.btz) * e.sgn_bx;
      L[6] = (e.gam0 * e.as) * e.by - ((e.cs * e.af) * e.bty) * e.sqrt_rho;
      L[7] = (e.gam0 * e.as) * e.bz - ((e.cs * e.af) * e.btz) * e.sqrt_rho;
      L[4] = -e.gam0 * e.as;
      for (int v = 0; v < 8; ++v) L[v] = (T(0.5) * L[v]) * e.cs2inv;
      if (e.cs_ge_ca)
        for (int v = 0; v < 8; ++v) L[v] *= e.sgn_bt;
      break;
    case 5:
      L[0] = e.btz * e.vy - e.bty * e.vz;
      L[2] = -e.btz;
      L[3] = e.bty;
      L[6] = (e.btz * e.sgn_bx) * e.sqrt_rho;
      L[7] = (-e.bty * e.sgn_bx) * e.sqrt_rho;
      for (int v = 0; v < 8; ++v) L[v] = T(0.5) * L[v];
      break;
    case 6:
      L[0] = e.af * (e.gam1 * e.v2 - e.cf * e.vx) + ((e.as * e.csl) * btv) * e.sgn_bx;
      L[1] = e.af * (e.gam0 * e.vx + e.cf);
      L[2] = (e.gam0 * e.af) * e.vy - ((e.as * e.csl) * e.bty) * e.sgn_bx;
      L[3] = (e.gam0 * e.af) * e.vz - ((e.as * e.csl) * e.btz) * e.sgn_bx;
      L[6] = (e.gam0 * e.af) * e.by + ((e.cs * e.as) * e.bty) * e.sqrt_rho;
      L[7] = (e.gam0 * e.af) * e.bz + ((e.cs * e.as) * e.btz) * e.sqrt_rho;
      L[4] = -e.gam0 * e.af;
      for (int v = 0; v < 8; ++v) L[v] = (T(0.5) * L[v]) * e.cs2inv;
      if (!e.cs_ge_ca)
        for (int v = 0; v < 8; ++v) L[v] *= e.sgn_bt;
      break;
  }
}

// WENO interpolant phi with Jiang-Shu weights (_weno.py:185-205)
template <typename T>
T weno_phi(T a, T b, T c, T d) {
  const T eps = T(1e-8);
  const T IS0 = T(13) * ((a - b) * (a - b)) + T(3) * ((a - T(3) * b) * (a - T(3) * b));
  const T IS1 = T(13) * ((b - c) * (b - c)) + T(3) * ((b + c) * (b + c));
  const T IS2 = T(13) * ((c - d) * (c - d)) + T(3) * ((T(3) * c - d) * (T(3) * c - d));
  const T a0 = T(1) / ((eps + IS0) * (eps + IS0));
  const T a1 = T(6) / ((eps + IS1) * (eps + IS1));
  const T a2 = T(3) / ((eps + IS2) * (eps + IS2));
  const T sum = std::max((a0 + a1) + a2, T(1e-14));
  const T w0 = a0 / sum, w2 = a2 / sum;
  return (w0 * ((a - T(2) * b) + c)) / T(3) + ((w2 - T(0.5)) * ((b - T(2) * c) + d)) / T(6);
}

// _weno_flux_x in the frame of direction `dir`: F[v][c] = flux at c+1/2 (_weno.py:111-324)
template <typename T>
void weno_flux(const G3<T>& g, const T* q, int dir, T* F) {
  const long nc = g.nc;
  const int* P = PERM[dir];
  // cell-centred x-frame fluxes and lambdas
  std::vector<T> Fc(8 * nc), lam(7 * nc);
#pragma omp parallel for schedule(static)
  for (long c = 0; c < nc; ++c) {
    T qq[8], f[8], l[7];
    for (int v = 0; v < 8; ++v) qq[v] = q[P[v] * nc + c];
    mhd_flux(qq, f, g);
    cell_lambdas(qq, g, l);
    for (int v = 0; v < 8; ++v) Fc[v * nc + c] = f[v];
    for (int s = 0; s < 7; ++s) lam[s * nc + c] = l[s];
  }
#pragma omp parallel for schedule(static)
  for (long c = 0; c < nc; ++c) {
    long m[6];
    for (int k = 0; k < 6; ++k) m[k] = g.nb(c, dir, k - 2);  // i-2 .. i+3
    T qs[6][8], fs[6][8];
    for (int k = 0; k < 6; ++k)
      for (int v = 0; v < 8; ++v) {
        qs[k][v] = q[P[v] * nc + m[k]];
        fs[k][v] = Fc[v * nc + m[k]];
      }
    T out[8];
    for (int v = 0; v < 8; ++v)
      out[v] = T(1.0 / 12.0) * (((-fs[1][v] + T(7) * fs[2][v]) + T(7) * fs[3][v]) - fs[4][v]);
    const EB<T> e = interface_blocks(qs[2], qs[3], g);
    for (int s = 0; s < 7; ++s) {
      T L[8], R[8];
      L_row(e, s, L);
      R_col(e, s, R);
      T sk[6], qk[6];
      for (int k = 0; k < 6; ++k) {
        T a = 0, b = 0;
        for (int v = 0; v < 8; ++v) {
          a += L[v] * fs[k][v];
          b += L[v] * qs[k][v];
        }
        sk[k] = a;
        qk[k] = b;
      }
      T d[5], dq[5];
      for (int k = 0; k < 5; ++k) {
        d[k] = sk[k + 1] - sk[k];
        dq[k] = qk[k + 1] - qk[k];
      }
      T amx = 0;
      for (int k = 0; k < 6; ++k) amx = std::max(amx, std::fabs(lam[s * nc + m[k]]));
      const T ap = T(0.5) * (d[0] + amx * dq[0]), bp = T(0.5) * (d[1] + amx * dq[1]);
      const T cp = T(0.5) * (d[2] + amx * dq[2]), dp = T(0.5) * (d[3] + amx * dq[3]);
      const T am = T(0.5) * (d[4] - amx * dq[4]), bm = T(0.5) * (d[3] - amx * dq[3]);
      const T cm = T(0.5) * (d[2] - amx * dq[2]), dm = T(0.5) * (d[1] - amx * dq[1]);
      const T Fs = -weno_phi(ap, bp, cp, dp) + weno_phi(am, bm, cm, dm);
      for (int v = 0; v < 8; ++v) out[v] = out[v] + R[v] * Fs;
    }
    for (int v = 0; v < 8; ++v) F[P[v] * nc + c] = out[v];
  }
}

// _interpolate.py / _differencing.py stencils along `axis`
template <typename T>
T c2f(const G3<T>& g, const T* a, long c, int ax) {
  return (((-a[g.nb(c, ax, -1)] + T(9) * a[c]) + T(9) * a[g.nb(c, ax, 1)]) - a[g.nb(c, ax, 2)]) / T(16);
}
template <typename T>
T f2c(const G3<T>& g, const T* f, long c, int ax) {
  return (((((T(3) * f[g.nb(c, ax, -3)] - T(25) * f[g.nb(c, ax, -2)]) + T(150) * f[g.nb(c, ax, -1)]) + T(150) * f[c]) -
           T(25) * f[g.nb(c, ax, 1)]) +
          T(3) * f[g.nb(c, ax, 2)]) /
         T(256);
}
template <typename T>
T d6(const G3<T>& g, const T* f, long c, int ax) {
  const T c1 = T(75.0 / 64.0), c2 = T(-25.0 / 384.0), c3 = T(3.0 / 640.0);
  return (c1 * (f[c] - f[g.nb(c, ax, -1)]) + c2 * (f[g.nb(c, ax, 1)] - f[g.nb(c, ax, -2)])) +
         c3 * (f[g.nb(c, ax, 2)] - f[g.nb(c, ax, -3)]);
}
template <typename T>
T p2a(const G3<T>& g, const T* q, long c, int A, int B) {
  const T sa = ((q[g.nb(c, A, -1)] - T(2) * q[c]) + q[g.nb(c, A, 1)]) / T(24);
  const T sb = ((q[g.nb(c, B, -1)] - T(2) * q[c]) + q[g.nb(c, B, 1)]) / T(24);
  return (q[c] + sa) + sb;
}

// update_cell_center_fields, _constrained_transport.py:160-198 (in place on q)
template <typename T>
void update_cell_center_fields(const G3<T>& g, T* q, const T* bx, const T* by, const T* bz) {
  const long nc = g.nc;
  std::vector<T> nb3(3 * nc);
#pragma omp parallel for schedule(static)
  for (long c = 0; c < nc; ++c) {
    nb3[c] = f2c(g, bx, c, 0);
    nb3[nc + c] = f2c(g, by, c, 1);
    nb3[2 * nc + c] = f2c(g, bz, c, 2);
  }
#pragma omp parallel for schedule(static)
  for (long c = 0; c < nc; ++c) {
    const T o = (q[5 * nc + c] * q[5 * nc + c] + q[6 * nc + c] * q[6 * nc + c]) + q[7 * nc + c] * q[7 * nc + c];
    const T X = nb3[c], Y = nb3[nc + c], Z = nb3[2 * nc + c];
    const T nw = (X * X + Y * Y) + Z * Z;
    q[5 * nc + c] = X;
    q[6 * nc + c] = Y;
    q[7 * nc + c] = Z;
    q[4 * nc + c] = q[4 * nc + c] + T(0.5) * (nw - o);
  }
}

// constrained_transport_rhs, _constrained_transport.py:43-155
template <typename T>
void ct_rhs(const G3<T>& g, const T* q, const T* Fx, const T* Fy, const T* Fz, T dtdx, T* rbx, T* rby, T* rbz) {
  const long nc = g.nc;
  std::vector<T> vx(nc), vy(nc), vz(nc);
  for (long c = 0; c < nc; ++c) {
    vx[c] = q[nc + c] / q[c];
    vy[c] = q[2 * nc + c] / q[c];
    vz[c] = q[3 * nc + c] / q[c];
  }
  const T* Bx = q + 5 * nc;
  const T* By = q + 6 * nc;
  const T* Bz = q + 7 * nc;
  std::vector<T> fyx(nc), fzx(nc), fxy(nc), fzy(nc), fxz(nc), fyz(nc);
  // modified fluxes (Eq. 12-17): flux of B_a through the b-faces
#pragma omp parallel for schedule(static)
  for (long c = 0; c < nc; ++c) {
    const T bxx = c2f(g, Bx, c, 0), byy = c2f(g, By, c, 1), bzz = c2f(g, Bz, c, 2);
    fyx[c] = Fx[6 * nc + c] + bxx * c2f(g, vy.data(), c, 0);  // By flux at x-faces
    fzx[c] = Fx[7 * nc + c] + bxx * c2f(g, vz.data(), c, 0);  // Bz flux at x-faces
    fxy[c] = Fy[5 * nc + c] + byy * c2f(g, vx.data(), c, 1);  // Bx flux at y-faces
    fzy[c] = Fy[7 * nc + c] + byy * c2f(g, vz.data(), c, 1);  // Bz flux at y-faces
    fxz[c] = Fz[5 * nc + c] + bzz * c2f(g, vx.data(), c, 2);  // Bx flux at z-faces
    fyz[c] = Fz[6 * nc + c] + bzz * c2f(g, vy.data(), c, 2);  // By flux at z-faces
  }
  std::vector<T> Oz(nc), Ox(nc), Oy(nc);
#pragma omp parallel for schedule(static)
  for (long c = 0; c < nc; ++c) {
    Oz[c] = c2f(g, fxy.data(), c, 0) - c2f(g, fyx.data(), c, 1);
    Ox[c] = c2f(g, fyz.data(), c, 1) - c2f(g, fzy.data(), c, 2);
    Oy[c] = c2f(g, fzx.data(), c, 2) - c2f(g, fxz.data(), c, 0);
  }
  std::vector<T> Ozb(nc), Oxb(nc), Oyb(nc);
#pragma omp parallel for schedule(static)
  for (long c = 0; c < nc; ++c) {
    Ozb[c] = p2a(g, Oz.data(), c, 0, 1);
    Oxb[c] = p2a(g, Ox.data(), c, 1, 2);
    Oyb[c] = p2a(g, Oy.data(), c, 0, 2);
  }
#pragma omp parallel for schedule(static)
  for (long c = 0; c < nc; ++c) {
    rbx[c] = (-dtdx) * d6(g, Ozb.data(), c, 1) + dtdx * d6(g, Oyb.data(), c, 2);
    rby[c] = (-dtdx) * d6(g, Oxb.data(), c, 2) + dtdx * d6(g, Ozb.data(), c, 0);
    rbz[c] = (-dtdx) * d6(g, Oyb.data(), c, 0) + dtdx * d6(g, Oxb.data(), c, 1);
  }
}

// _enforce_positivity (in place), _enforce_positivity.py:21-59
template <typename T>
void enforce_positivity(const G3<T>& g, T* q) {
  const long nc = g.nc;
#pragma omp parallel for schedule(static)
  for (long c = 0; c < nc; ++c) {
    const T rho = std::max(q[c], g.rhomin);
    const T vx = q[nc + c] / rho, vy = q[2 * nc + c] / rho, vz = q[3 * nc + c] / rho;
    const T b2 = (q[5 * nc + c] * q[5 * nc + c] + q[6 * nc + c] * q[6 * nc + c]) + q[7 * nc + c] * q[7 * nc + c];
    const T v2 = (vx * vx + vy * vy) + vz * vz;
    T p = (g.gamma - T(1)) * ((q[4 * nc + c] - (T(0.5) * rho) * v2) - T(0.5) * b2);
    p = std::max(p, g.pgmin);
    q[c] = rho;
    q[4 * nc + c] = (p / (g.gamma - T(1)) + (T(0.5) * rho) * v2) + T(0.5) * b2;
  }
}

// compute_rhs of _ssprk.py:197-243
template <typename T>
void compute_rhs(const G3<T>& g, const T* qin, const T* bx, const T* by, const T* bz, T k2, T dt, T* rq, T* rbx, T* rby,
                 T* rbz) {
  const long nc = g.nc;
  std::vector<T> q(qin, qin + 8 * nc);
  update_cell_center_fields(g, q.data(), bx, by, bz);
  const T dtdx = (k2 * dt) / g.dx;
  std::vector<T> Fx(8 * nc), Fy(8 * nc), Fz(8 * nc);
  weno_flux(g, q.data(), 0, Fx.data());
  weno_flux(g, q.data(), 1, Fy.data());
  weno_flux(g, q.data(), 2, Fz.data());
  ct_rhs(g, q.data(), Fx.data(), Fy.data(), Fz.data(), dtdx, rbx, rby, rbz);
#pragma omp parallel for schedule(static)
  for (long c = 0; c < nc; ++c) {
    const long xm = g.nb(c, 0, -1), ym = g.nb(c, 1, -1), zm = g.nb(c, 2, -1);
    for (int v = 0; v < 8; ++v) {
      const T dx_ = Fx[v * nc + c] - Fx[v * nc + xm];
      const T dy_ = Fy[v * nc + c] - Fy[v * nc + ym];
      const T dz_ = Fz[v * nc + c] - Fz[v * nc + zm];
      rq[v * nc + c] = (-dtdx) * ((dx_ + dy_) + dz_);
    }
  }
}

// _ssprk4_with_ct, _ssprk.py:177-348 (in place on q and the faces)
template <typename T>
void ssprk4_with_ct(const G3<T>& g, T* q, T* bx, T* by, T* bz, T dt) {
  const long nc = g.nc;
  const T k1[5] = {T(1.0), T(0.44437049406734), T(0.62010185138540), T(0.17807995410773), T(-2.081261929715610e-02)};
  const T k2[5] = {T(0.39175222700392), T(0.36841059262959), T(0.25189177424738), T(0.54497475021237),
                   T(0.22600748319395)};
  const T k3[5] = {T(0.0), T(0.55562950593266), T(0.37989814861460), T(0.82192004589227), T(5.03580947213895e-01)};
  const T ff[5] = {T(-2.081261929715610e-02), T(0.0), T(0.51723167208978), T(-6.518979800418380e-12),
                   T(5.03580947213895e-01)};
  std::vector<T> q0(q, q + 8 * nc), b0(3 * nc), qf(8 * nc), bf(3 * nc), rq(8 * nc), rb(3 * nc), b(3 * nc);
  std::memcpy(&b0[0], bx, sizeof(T) * nc);
  std::memcpy(&b0[nc], by, sizeof(T) * nc);
  std::memcpy(&b0[2 * nc], bz, sizeof(T) * nc);
  b = b0;
  std::vector<T> qc(q0);
  for (long i = 0; i < 8 * nc; ++i) qf[i] = ff[0] * q0[i];
  for (long i = 0; i < 3 * nc; ++i) bf[i] = ff[0] * b0[i];
  for (int s = 0; s < 4; ++s) {
    if (g.positivity) enforce_positivity(g, qc.data());
    compute_rhs(g, qc.data(), &b[0], &b[nc], &b[2 * nc], k2[s], dt, rq.data(), &rb[0], &rb[nc], &rb[2 * nc]);
#pragma omp parallel for schedule(static)
    for (long i = 0; i < 8 * nc; ++i) {
      qc[i] = (k1[s] * q0[i] + k3[s] * qc[i]) + rq[i];
      qf[i] = qf[i] + qc[i] * ff[s + 1];
    }
#pragma omp parallel for schedule(static)
    for (long i = 0; i < 3 * nc; ++i) {
      b[i] = (k1[s] * b0[i] + k3[s] * b[i]) + rb[i];
      bf[i] = bf[i] + b[i] * ff[s + 1];
    }
  }
  compute_rhs(g, qc.data(), &b[0], &b[nc], &b[2 * nc], k2[4], dt, rq.data(), &rb[0], &rb[nc], &rb[2 * nc]);
  for (long i = 0; i < 8 * nc; ++i) qf[i] = qf[i] + rq[i];
  for (long i = 0; i < 3 * nc; ++i) bf[i] = bf[i] + rb[i];
  update_cell_center_fields(g, qf.data(), &bf[0], &bf[nc], &bf[2 * nc]);
  if (g.positivity) enforce_positivity(g, qf.data());
  std::memcpy(q, qf.data(), sizeof(T) * 8 * nc);
  std::memcpy(bx, &bf[0], sizeof(T) * nc);
  std::memcpy(by, &bf[nc], sizeof(T) * nc);
  std::memcpy(bz, &bf[2 * nc], sizeof(T) * nc);
}

// primitive (11 rows) -> conserved (8 rows), _equations.py:65-101
template <typename T>
void cons_from_prim(const G3<T>& g, const T* w, T* q) {
  const long nc = g.nc;
  for (long c = 0; c < nc; ++c) {
    const T rho = w[c];
    const T u2 = (w[nc + c] * w[nc + c] + w[2 * nc + c] * w[2 * nc + c]) + w[3 * nc + c] * w[3 * nc + c];
    const T b2 = (w[5 * nc + c] * w[5 * nc + c] + w[6 * nc + c] * w[6 * nc + c]) + w[7 * nc + c] * w[7 * nc + c];
    q[c] = rho;
    for (int a = 1; a <= 3; ++a) q[a * nc + c] = rho * w[a * nc + c];
    q[4 * nc + c] = (w[4 * nc + c] / (g.gamma - T(1)) + (T(0.5) * rho) * u2) + T(0.5) * b2;
    for (int a = 5; a < 8; ++a) q[a * nc + c] = w[a * nc + c];
  }
}

// conserved -> primitive with floors, _equations.py:104-153
template <typename T>
void prim_from_cons(const G3<T>& g, const T* q, T* w) {
  const long nc = g.nc;
  for (long c = 0; c < nc; ++c) {
    const T rho = q[c];
    const T ux = q[nc + c] / rho, uy = q[2 * nc + c] / rho, uz = q[3 * nc + c] / rho;
    const T u2 = (ux * ux + uy * uy) + uz * uz;
    const T b2 = (q[5 * nc + c] * q[5 * nc + c] + q[6 * nc + c] * q[6 * nc + c]) + q[7 * nc + c] * q[7 * nc + c];
    T p = (g.gamma - T(1)) * ((q[4 * nc + c] - (T(0.5) * rho) * u2) - T(0.5) * b2);
    T r = rho;
    if (g.positivity) {
      r = std::max(r, g.rhomin);
      p = std::max(p, g.pgmin);
    }
    w[c] = r;
    w[nc + c] = ux;
    w[2 * nc + c] = uy;
    w[3 * nc + c] = uz;
    w[4 * nc + c] = p;
    for (int a = 5; a < 8; ++a) w[a * nc + c] = q[a * nc + c];
  }
}

// _evolve_state_fd: state (11, nx, ny, nz) in place
template <typename T>
void evolve_state_fd(const G3<T>& g, T* state, T dt) {
  const long nc = g.nc;
  std::vector<T> q(8 * nc);
  cons_from_prim(g, state, q.data());
  ssprk4_with_ct(g, q.data(), state + 8 * nc, state + 9 * nc, state + 10 * nc, dt);
  prim_from_cons(g, q.data(), state);
}

// _cfl_time_step_fd, _timestep_estimator.py(FD):25-91
template <typename T>
T cfl_dt_fd(const G3<T>& g, const T* state, T C, T dt_max) {
  const long nc = g.nc;
  std::vector<T> q(8 * nc);
  cons_from_prim(g, state, q.data());
  T lmax[3] = {0, 0, 0};
  for (int d = 0; d < 3; ++d) {
    T m = 0;
#pragma omp parallel for schedule(static) reduction(max : m)
    for (long c = 0; c < nc; ++c) {
      T qq[8], l[7];
      for (int v = 0; v < 8; ++v) qq[v] = q[PERM[d][v] * nc + c];
      cell_lambdas(qq, g, l);
      for (int s = 0; s < 7; ++s) m = std::max(m, std::fabs(l[s]));
    }
    lmax[d] = m;
  }
  const T dt = (C * g.dx) / ((lmax[0] + lmax[1]) + lmax[2]);
  return std::min(dt, dt_max);
}

struct FdConfig {
  int nx, ny, nz;
  int positivity;
  int fixed_timestep;
  int num_timesteps;
  int exact_end_time;
  double gamma, dx, C_cfl, t_end, dt_max, rhomin, pgmin;
};

template <typename T>
long time_integration_fd(const FdConfig& c, T* state, long max_steps, double* t_out) {
  G3<T> g = make_g3<T>(c.nx, c.ny, c.nz, c.gamma, c.dx, c.rhomin, c.pgmin, c.positivity);
  T t = 0;
  const T t_end = (T)c.t_end;
  long steps = 0;
  auto one = [&]() {
    T dt = c.fixed_timestep ? t_end / (T)c.num_timesteps : cfl_dt_fd(g, state, (T)c.C_cfl, (T)c.dt_max);
    if (c.exact_end_time) dt = std::min(dt, t_end - t);
    evolve_state_fd(g, state, dt);
    t = t + dt;
    ++steps;
  };
  if (c.fixed_timestep) {
    for (int s = 0; s < c.num_timesteps && (max_steps < 0 || steps < max_steps); ++s) one();
  } else {
    while (t < t_end && (max_steps < 0 || steps < max_steps)) one();
  }
  if (t_out) *t_out = (double)t;
  return steps;
}

}  // namespace

extern "C" {

#define AMX_FD_DEFS(T, SFX)                                                                                    \
  long fd_time_integration_##SFX(const FdConfig* c, T* state, long max_steps, double* t_out) {                 \
    return time_integration_fd<T>(*c, state, max_steps, t_out);                                               \
  }                                                                                                            \
  void fd_evolve_state_##SFX(const FdConfig* c, T* state, double dt) {                                         \
    G3<T> g = make_g3<T>(c->nx, c->ny, c->nz, c->gamma, c->dx, c->rhomin, c->pgmin, c->positivity);            \
    evolve_state_fd<T>(g, state, (T)dt);                                                                       \
  }                                                                                                            \
  double fd_cfl_dt_##SFX(const FdConfig* c, const T* state) {                                                  \
    G3<T> g = make_g3<T>(c->nx, c->ny, c->nz, c->gamma, c->dx, c->rhomin, c->pgmin, c->positivity);            \
    return (double)cfl_dt_fd<T>(g, state, (T)c->C_cfl, (T)c->dt_max);                                          \
  }                                                                                                            \
  /* WENO interface fluxes of a conserved state (8, nx, ny, nz) along dir */                                 \
  void fd_weno_flux_##SFX(const FdConfig* c, const T* q, int dir, T* F) {                                      \
    G3<T> g = make_g3<T>(c->nx, c->ny, c->nz, c->gamma, c->dx, c->rhomin, c->pgmin, c->positivity);            \
    weno_flux<T>(g, q, dir, F);                                                                                \
  }                                                                                                            \
  /* one RHS evaluation: rq (8, ...), rb (3, ...) */                                                           \
  void fd_compute_rhs_##SFX(const FdConfig* c, const T* q, const T* b, double k2, double dt, T* rq, T* rb) {   \
    G3<T> g = make_g3<T>(c->nx, c->ny, c->nz, c->gamma, c->dx, c->rhomin, c->pgmin, c->positivity);            \
    const long nc = g.nc;                                                                                      \
    compute_rhs<T>(g, q, b, b + nc, b + 2 * nc, (T)k2, (T)dt, rq, rb, rb + nc, rb + 2 * nc);                   \
  }                                                                                                            \
  /* interface eigen-system of the x-frame between two cells qa, qb (8 each) */                              \
  void fd_eigensystem_##SFX(const FdConfig* c, const T* qa, const T* qb, T* L, T* R, T* lam) {                 \
    G3<T> g = make_g3<T>(1, 1, 1, c->gamma, c->dx, c->rhomin, c->pgmin, c->positivity);                        \
    const EB<T> e = interface_blocks<T>(qa, qb, g);                                                            \
    for (int s = 0; s < 7; ++s) {                                                                              \
      L_row<T>(e, s, L + 8 * s);                                                                               \
      T col[8];                                                                                                \
      R_col<T>(e, s, col);                                                                                     \
      for (int v = 0; v < 8; ++v) R[7 * v + s] = col[v];                                                       \
    }                                                                                                          \
    cell_lambdas<T>(qa, g, lam);                                                                               \
  }                                                                                                            \
  void fd_mhd_flux_##SFX(const FdConfig* c, const T* q, T* f) {                                                \
    G3<T> g = make_g3<T>(1, 1, 1, c->gamma, c->dx, c->rhomin, c->pgmin, c->positivity);                        \
    mhd_flux<T>(q, f, g);                                                                                      \
  }                                                                                                            \
  /* face divergence of _interface_field_divergence (without the 1/dx) */                                    \
  void fd_face_divergence_##SFX(const FdConfig* c, const T* bx, const T* by, const T* bz, T* out) {            \
    G3<T> g = make_g3<T>(c->nx, c->ny, c->nz, c->gamma, c->dx, c->rhomin, c->pgmin, c->positivity);            \
    for (long i = 0; i < g.nc; ++i) out[i] = (d6(g, bx, i, 0) + d6(g, by, i, 1)) + d6(g, bz, i, 2);            \
  }                                                                                                            \
  void fd_interface_fields_##SFX(const FdConfig* c, const T* B, T* b) {                                        \
    G3<T> g = make_g3<T>(c->nx, c->ny, c->nz, c->gamma, c->dx, c->rhomin, c->pgmin, c->positivity);            \
    const long nc = g.nc;                                                                                      \
    for (long i = 0; i < nc; ++i) {                                                                            \
      b[i] = c2f(g, B, i, 0);                                                                                  \
      b[nc + i] = c2f(g, B + nc, i, 1);                                                                        \
      b[2 * nc + i] = c2f(g, B + 2 * nc, i, 2);                                                                \
    }                                                                                                          \
  }

AMX_FD_DEFS(float, f32)
AMX_FD_DEFS(double, f64)

}  // extern "C"
