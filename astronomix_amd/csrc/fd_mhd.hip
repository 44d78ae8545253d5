// MI355X (gfx950) kernels and C ABI of the finite-difference MHD step:
// characteristic WENO-5 fluxes, HOW-MHD constrained transport and SSPRK(5,4)
// (astronomix/_finite_difference/**, leo1200/astronomix v0.5.1).
//
// Passes of one right-hand-side evaluation (all direction-specific; the
// reference's transposes become index arithmetic):
//   k_prep     positivity floors of the carried stage state and the
//              face-to-centre B with its energy correction
//              (_enforce_positivity.py:21-59, _constrained_transport.py:160-198)
//   k_weno<d>  interface fluxes along d: the cells of a tile are staged in LDS
//              with their flux and wave speeds; each thread builds the
//              interface eigen-system once and runs the 7 characteristic
//              WENO-5 reconstructions (_weno.py:111-324, _eigen.py:116-911);
//              also emits the CT modified fluxes (_constrained_transport.py:81-102)
//   k_ct_edge  edge EMFs by 4th-order interpolation (_constrained_transport.py:107-131)
//   k_ct_avg   point values to averages (_interpolate.py:38-62)
//   k_face     6th-order curl of the EMFs + SSPRK combine of the face fields
//   k_qupd     flux divergence + SSPRK combine of the conserved state
//              (_ssprk.py:224-228, 300-311)
// Boundaries are periodic.

#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <new>
#include <vector>

#include "amx.h"
#include "common.h"
#include "snapshots.h"

namespace amx {
namespace fd {

constexpr int TBZ = 64;  // lanes along the contiguous axis
constexpr int TP = 8;    // waves per WENO block
constexpr int NTH = TP * TBZ;

struct Phys {
  float gamma, gm1, inv_gm1, rhomin, pgmin;
  int positivity;
};

struct Grid {
  int nx, ny, nz;
  long sx, sy;  // strides of x and y (z contiguous)
  long nc;
};

__device__ __forceinline__ long cell(const Grid& g, int x, int y, int z) { return (long)x * g.sx + (long)y * g.sy + z; }

// Cell index with its coordinates: computed once per cell with 32-bit
// arithmetic (grids have < 2^30 cells), so that stencil neighbours need no
// division.
struct Pos {
  unsigned c;
  int x, y, z;
};

__device__ __forceinline__ Pos pos_of(const Grid& g, unsigned c) {
  Pos p;
  p.c = c;
  const unsigned t = c / (unsigned)g.nz;
  p.z = (int)(c - t * (unsigned)g.nz);
  p.x = (int)(t / (unsigned)g.ny);
  p.y = (int)(t - (unsigned)p.x * (unsigned)g.ny);
  return p;
}

// neighbour of p along axis by s (periodic)
__device__ __forceinline__ unsigned nb(const Grid& g, const Pos& p, int axis, int s) {
  if (axis == 2) return p.c - p.z + wrapi(p.z + s, g.nz);
  if (axis == 1) return p.c + (unsigned)((wrapi(p.y + s, g.ny) - p.y) * g.nz);
  return p.c + (unsigned)((wrapi(p.x + s, g.nx) - p.x)) * (unsigned)g.sx;
}

__device__ __forceinline__ float dss(float x) { return fsqrt(fmaxf(x, 1e-12f)); }

// component permutation of the direction's frame (the reference's swaps of
// (m_x, m_d) and (B_x, B_d)); an involution
__device__ __forceinline__ int perm(int dir, int v) {
  if (dir == 1) {
    const int t[8] = {0, 2, 1, 3, 4, 6, 5, 7};
    return t[v];
  }
  if (dir == 2) {
    const int t[8] = {0, 3, 2, 1, 4, 7, 6, 5};
    return t[v];
  }
  return v;
}

// _mhd_flux_x of one cell in the direction's frame (_fluxes.py:20-69)
__device__ __forceinline__ void cell_flux(const float* q, float* f, const Phys& ph) {
  float rho = q[0];
  const float ri = frcp(rho);
  const float vx = q[1] * ri, vy = q[2] * ri, vz = q[3] * ri;
  const float Bx = q[5], By = q[6], Bz = q[7];
  const float u2 = (vx * vx + vy * vy) + vz * vz;
  const float b2 = (Bx * Bx + By * By) + Bz * Bz;
  float p = ph.gm1 * ((q[4] - (0.5f * rho) * u2) - 0.5f * b2);
  if (ph.positivity) {
    rho = fmaxf(rho, ph.rhomin);
    p = fmaxf(p, ph.pgmin);
  }
  const float Pt = p + 0.5f * b2;
  const float vB = (vx * Bx + vy * By) + vz * Bz;
  const float E = (p * ph.inv_gm1 + (0.5f * rho) * u2) + 0.5f * b2;
  f[0] = rho * vx;
  f[1] = ((rho * (vx * vx)) + Pt) - Bx * Bx;
  f[2] = ((rho * vx) * vy) - Bx * By;
  f[3] = ((rho * vx) * vz) - Bx * Bz;
  f[4] = (E + Pt) * vx - vB * Bx;
  f[5] = 0.f;
  f[6] = By * vx - Bx * vy;
  f[7] = Bz * vx - Bx * vz;
}

// cell floors of the building-block routines (_eigen.py:62-74 / 148-160)
struct Cell {
  float rho, vx, vy, vz, v2, b2, p, E;
};

__device__ __forceinline__ Cell floored(const float* q, const Phys& ph) {
  Cell c;
  c.rho = q[0];
  const float ri = frcp(c.rho);
  c.vx = q[1] * ri;
  c.vy = q[2] * ri;
  c.vz = q[3] * ri;
  c.v2 = (c.vx * c.vx + c.vy * c.vy) + c.vz * c.vz;
  c.b2 = (q[5] * q[5] + q[6] * q[6]) + q[7] * q[7];
  c.E = q[4];
  c.p = ph.gm1 * (c.E - 0.5f * (c.rho * c.v2 + c.b2));
  if (c.rho < ph.rhomin || c.p < ph.pgmin) c.rho = fmaxf(c.rho, ph.rhomin);
  if (c.rho < ph.rhomin || c.p < ph.pgmin) c.p = fmaxf(c.p, ph.pgmin);
  if (c.rho < ph.rhomin || c.p < ph.pgmin) c.E = c.p * ph.inv_gm1 + 0.5f * (c.rho * c.v2 + c.b2);
  return c;
}

// (vx, c_fast, c_alfven, c_slow) of one cell (_eigenvalue_building_blocks,
// _eigen.py:27-114); the 7 eigenvalues are vx -/+ these and vx itself
__device__ __forceinline__ void cell_speeds(const float* q, float* sp, const Phys& ph) {
  const Cell c = floored(q, ph);
  const float Bx = q[5];
  const float rinv = frcp(c.rho);
  const float cs2 = fmaxf(0.f, ph.gamma * fabsf(c.p * rinv));
  const float s = c.b2 * rinv + cs2;
  const float droot = dss(fmaxf(0.f, s * s - ((4.f * (Bx * Bx)) * rinv) * cs2));
  sp[0] = c.vx;
  sp[1] = dss(fmaxf(0.f, 0.5f * (s + droot)));
  sp[2] = dss(fmaxf(0.f, (Bx * Bx) * rinv));
  sp[3] = dss(fmaxf(0.f, 0.5f * (s - droot)));
}

// interface eigen-system building blocks (_eigen.py:116-337)
struct EB {
  float sqrt_rho, rsr, vx, vy, vz, v2, by, bz, bty, btz, sgn_bx, sgn_bt, cs, cs2, cs2inv, cf, csl, af, as, gam0, gam1,
      gam2;
  bool cs_ge_ca;
};

__device__ __forceinline__ EB interface_blocks(const float* qa, const float* qb, const Phys& ph) {
  const Cell a = floored(qa, ph), b = floored(qb, ph);
  const float Ha = (a.E + a.p) * frcp(a.rho), Hb = (b.E + b.p) * frcp(b.rho);
  EB e;
  const float rho = 0.5f * (a.rho + b.rho);
  e.vx = 0.5f * (a.vx + b.vx);
  e.vy = 0.5f * (a.vy + b.vy);
  e.vz = 0.5f * (a.vz + b.vz);
  const float bx = 0.5f * (qa[5] + qb[5]);
  e.by = 0.5f * (qa[6] + qb[6]);
  e.bz = 0.5f * (qa[7] + qb[7]);
  const float H = 0.5f * (Ha + Hb);
  e.v2 = (e.vx * e.vx + e.vy * e.vy) + e.vz * e.vz;
  const float bsq = (bx * bx + e.by * e.by) + e.bz * e.bz;
  const float ri = frcp(rho);
  const float b2r = bsq * ri;
  const float bx2r = (bx * bx) * ri;
  e.cs2 = ph.gm1 * (H - 0.5f * (e.v2 + b2r));
  e.cs = dss(fmaxf(0.f, e.cs2));
  const float s = b2r + e.cs2;
  const float droot = dss(fmaxf(0.f, s * s - (4.f * bx2r) * e.cs2));
  e.cf = dss(fmaxf(0.f, 0.5f * (s + droot)));
  const float ca = dss(fmaxf(0.f, bx2r));
  e.csl = dss(fmaxf(0.f, 0.5f * (s - droot)));
  float bt2 = e.by * e.by + e.bz * e.bz;
  e.sgn_bx = bx >= 0.f ? 1.f : -1.f;
  bt2 = fmaxf(bt2, 1e-12f);
  const float rbt = frcp(dss(bt2));
  e.bty = e.by * rbt;
  e.btz = e.bz * rbt;
  const float denom = fmaxf(e.cf * e.cf - e.csl * e.csl, 1e-12f);
  const float rd = frcp(dss(denom));
  e.af = dss(fmaxf(0.f, e.cs2 - e.csl * e.csl)) * rd;
  e.as = dss(fmaxf(0.f, e.cf * e.cf - e.cs2)) * rd;
  e.sqrt_rho = dss(rho);
  e.rsr = frcp(e.sqrt_rho);
  e.gam0 = 1.f - ph.gamma;
  e.gam1 = 0.5f * ph.gm1;
  e.gam2 = (ph.gamma - 2.f) * ph.inv_gm1;
  e.cs2inv = e.cs2 > 0.f ? frcp(e.cs2) : 0.f;
  e.sgn_bt = e.by != 0.f ? (e.by >= 0.f ? 1.f : -1.f) : (e.bz >= 0.f ? 1.f : -1.f);
  e.cs_ge_ca = e.cs >= ca;
  return e;
}

// _eigen_L_row (_eigen.py:625-911): the 7 entries (rows 0,1,2,3,4,6,7 of the
// conserved vector; the B_x entry is 0)
template <int S>
__device__ __forceinline__ void L_row(const EB& e, float* L) {
  const float btv = e.bty * e.vy + e.btz * e.vz;
  if (S == 0 || S == 6) {
    const float sg = S == 0 ? 1.f : -1.f;  // fast -: +cf terms; fast +: -cf terms
    L[0] = e.af * (e.gam1 * e.v2 + sg * (e.cf * e.vx)) - sg * (((e.as * e.csl) * btv) * e.sgn_bx);
    L[1] = e.af * (e.gam0 * e.vx - sg * e.cf);
    L[2] = (e.gam0 * e.af) * e.vy + sg * (((e.as * e.csl) * e.bty) * e.sgn_bx);
    L[3] = (e.gam0 * e.af) * e.vz + sg * (((e.as * e.csl) * e.btz) * e.sgn_bx);
    L[6] = (e.gam0 * e.af) * e.by + ((e.cs * e.as) * e.bty) * e.sqrt_rho;
    L[7] = (e.gam0 * e.af) * e.bz + ((e.cs * e.as) * e.btz) * e.sqrt_rho;
    L[4] = -e.gam0 * e.af;
    const float f = e.cs_ge_ca ? 1.f : e.sgn_bt;
#pragma unroll
    for (int v = 0; v < 8; ++v) L[v] = ((0.5f * L[v]) * e.cs2inv) * f;
  } else if (S == 1 || S == 5) {
    const float sg = S == 1 ? 1.f : -1.f;
    L[0] = 0.5f * (e.btz * e.vy - e.bty * e.vz);
    L[1] = 0.f;
    L[2] = 0.5f * -e.btz;
    L[3] = 0.5f * e.bty;
    L[4] = 0.f;
    L[6] = 0.5f * ((sg * -e.btz * e.sgn_bx) * e.sqrt_rho);
    L[7] = 0.5f * ((sg * e.bty * e.sgn_bx) * e.sqrt_rho);
  } else if (S == 2) {
    L[0] = e.as * (e.gam1 * e.v2 + e.csl * e.vx) + ((e.af * e.cf) * btv) * e.sgn_bx;
    L[1] = (e.gam0 * e.as) * e.vx - e.as * e.csl;
    L[2] = (e.gam0 * e.as) * e.vy - ((e.af * e.cf) * e.bty) * e.sgn_bx;
    L[3] = (e.gam0 * e.as) * e.vz - ((e.af * e.cf) * e.btz) * e.sgn_bx;
    L[6] = (e.gam0 * e.as) * e.by - ((e.cs * e.af) * e.bty) * e.sqrt_rho;
    L[7] = (e.gam0 * e.as) * e.bz - ((e.cs * e.af) * e.btz) * e.sqrt_rho;
    L[4] = -e.gam0 * e.as;
    const float f = e.cs_ge_ca ? e.sgn_bt : 1.f;
#pragma unroll
    for (int v = 0; v < 8; ++v) L[v] = ((0.5f * L[v]) * e.cs2inv) * f;
  } else if (S == 3) {
    const float k = (-e.gam0) * e.cs2inv;
    L[0] = (-e.cs2 * frcp(e.gam0) - 0.5f * e.v2) * k;
    L[1] = e.vx * k;
    L[2] = e.vy * k;
    L[3] = e.vz * k;
    L[6] = e.by * k;
    L[7] = e.bz * k;
    L[4] = -1.f * k;
  } else {  // S == 4
    L[0] = e.as * (e.gam1 * e.v2 - e.csl * e.vx) - ((e.af * e.cf) * btv) * e.sgn_bx;
    L[1] = e.as * (e.gam0 * e.vx + e.csl);
    L[2] = (e.gam0 * e.as) * e.vy + ((e.af * e.cf) * e.bty) * e.sgn_bx;
    L[3] = (e.gam0 * e.as) * e.vz + ((e.af * e.cf) * e.btz) * e.sgn_bx;
    L[6] = (e.gam0 * e.as) * e.by - ((e.cs * e.af) * e.bty) * e.sqrt_rho;
    L[7] = (e.gam0 * e.as) * e.bz - ((e.cs * e.af) * e.btz) * e.sqrt_rho;
    L[4] = -e.gam0 * e.as;
    const float f = e.cs_ge_ca ? e.sgn_bt : 1.f;
#pragma unroll
    for (int v = 0; v < 8; ++v) L[v] = ((0.5f * L[v]) * e.cs2inv) * f;
  }
}

// _eigen_R_col (_eigen.py:340-622): rows 0,1,2,3,4,6,7 (B_x row is 0)
template <int S>
__device__ __forceinline__ void R_col(const EB& e, float* R) {
  const float btv = e.bty * e.vy + e.btz * e.vz;
  if (S == 0 || S == 6) {
    const float sg = S == 0 ? -1.f : 1.f;  // fast -: vx - cf; fast +: vx + cf
    R[0] = e.af;
    R[1] = e.af * (e.vx + sg * e.cf);
    R[2] = e.af * e.vy - sg * (((e.as * e.csl) * e.bty) * e.sgn_bx);
    R[3] = e.af * e.vz - sg * (((e.as * e.csl) * e.btz) * e.sgn_bx);
    R[6] = ((e.cs * e.as) * e.bty) * e.rsr;
    R[7] = ((e.cs * e.as) * e.btz) * e.rsr;
    R[4] = e.af * (((e.cf * e.cf + sg * (e.cf * e.vx)) + 0.5f * e.v2) - e.gam2 * e.cs2) - sg * (((e.as * e.csl) * btv) * e.sgn_bx);
    if (!e.cs_ge_ca) {
#pragma unroll
      for (int v = 0; v < 8; ++v) R[v] *= e.sgn_bt;
    }
  } else if (S == 1 || S == 5) {
    const float sg = S == 1 ? 1.f : -1.f;
    R[0] = 0.f;
    R[1] = 0.f;
    R[2] = -e.btz;
    R[3] = e.bty;
    R[6] = (sg * -e.btz * e.sgn_bx) * e.rsr;
    R[7] = (sg * e.bty * e.sgn_bx) * e.rsr;
    R[4] = e.bty * e.vz - e.btz * e.vy;
  } else if (S == 2 || S == 4) {
    const float sg = S == 2 ? -1.f : 1.f;
    R[0] = e.as;
    R[1] = e.as * (e.vx + sg * e.csl);
    R[2] = e.as * e.vy + sg * (((e.af * e.cf) * e.bty) * e.sgn_bx);
    R[3] = e.as * e.vz + sg * (((e.af * e.cf) * e.btz) * e.sgn_bx);
    R[6] = ((-e.cs * e.af) * e.bty) * e.rsr;
    R[7] = ((-e.cs * e.af) * e.btz) * e.rsr;
    R[4] = e.as * (((e.csl * e.csl + sg * (e.csl * e.vx)) + 0.5f * e.v2) - e.gam2 * e.cs2) + sg * (((e.af * e.cf) * btv) * e.sgn_bx);
    if (e.cs_ge_ca) {
#pragma unroll
      for (int v = 0; v < 8; ++v) R[v] *= e.sgn_bt;
    }
  } else {  // S == 3
    R[0] = 1.f;
    R[1] = e.vx;
    R[2] = e.vy;
    R[3] = e.vz;
    R[4] = 0.5f * e.v2;
    R[6] = 0.f;
    R[7] = 0.f;
  }
}

// WENO interpolant with Jiang-Shu weights (_weno.py:185-205)
__device__ __forceinline__ float weno_phi(float a, float b, float c, float d) {
  const float eps = 1e-8f;
  const float t0 = a - 3.f * b, t2 = 3.f * c - d;
  const float IS0 = 13.f * ((a - b) * (a - b)) + 3.f * (t0 * t0);
  const float IS1 = 13.f * ((b - c) * (b - c)) + 3.f * ((b + c) * (b + c));
  const float IS2 = 13.f * ((c - d) * (c - d)) + 3.f * (t2 * t2);
  const float e0 = eps + IS0, e1 = eps + IS1, e2 = eps + IS2;
  const float a0 = frcp(e0 * e0);
  const float a1 = 6.f * frcp(e1 * e1);
  const float a2 = 3.f * frcp(e2 * e2);
  const float rs = frcp(fmaxf((a0 + a1) + a2, 1e-14f));
  const float w0 = a0 * rs, w2 = a2 * rs;
  return (w0 * ((a - 2.f * b) + c)) * (1.f / 3.f) + ((w2 - 0.5f) * ((b - 2.f * c) + d)) * (1.f / 6.f);
}

struct WenoArgs {
  const float* q[4];   // rho, m_x, m_y, m_z of the stage state (after floors)
  const float* qt[4];  // E, B_x, B_y, B_z after update_cell_center_fields
  float* F[8];         // interface flux along the direction (cell i -> face i+1/2)
  float* modA;         // CT modified flux of the tangential B component in frame row 6
  float* modB;         // ... in frame row 7
  Grid g;
  Phys ph;
};

// Cells staged per block: direction x / y: (TP + 5) positions x 64 lanes;
// direction z: TP rows x (64 + 5) positions.
template <int DIR>
struct Tile {
  static constexpr int NCELL = DIR < 2 ? (TP + 5) * TBZ : TP * (TBZ + 5);
};

// LDS index of window cell k (0..5 = i-2 .. i+3) of thread (w, l)
template <int DIR>
__device__ __forceinline__ int win(int w, int l, int k) {
  return DIR < 2 ? (w + k) * TBZ + l : w * (TBZ + 5) + l + k;
}

// characteristic mode S of the WENO-5 flux at one interface (_weno.py:134-235)
template <int DIR, int S>
__device__ __forceinline__ void weno_mode(const EB& e, const float (*sQ)[Tile<DIR>::NCELL],
                                          const float (*sF)[Tile<DIR>::NCELL], const float (*sS)[Tile<DIR>::NCELL],
                                          int w, int l, float* out) {
  float L[8], R[8];
  L_row<S>(e, L);
  R_col<S>(e, R);
  float sk[6], qk[6], amx = 0.f;
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    const int c = win<DIR>(w, l, k);
    float a = 0.f, b = 0.f;
#pragma unroll
    for (int v = 0; v < 8; ++v) {
      if (v == 5) continue;
      a += L[v] * sF[v][c];
      b += L[v] * sQ[v][c];
    }
    sk[k] = a;
    qk[k] = b;
    const float vx = sS[0][c];
    float lam;
    if (S == 0) lam = vx - sS[1][c];
    else if (S == 1) lam = vx - sS[2][c];
    else if (S == 2) lam = vx - sS[3][c];
    else if (S == 3) lam = vx;
    else if (S == 4) lam = vx + sS[3][c];
    else if (S == 5) lam = vx + sS[2][c];
    else lam = vx + sS[1][c];
    amx = fmaxf(amx, fabsf(lam));
  }
  float d[5], dq[5];
#pragma unroll
  for (int k = 0; k < 5; ++k) {
    d[k] = sk[k + 1] - sk[k];
    dq[k] = qk[k + 1] - qk[k];
  }
  const float ap = 0.5f * (d[0] + amx * dq[0]), bp = 0.5f * (d[1] + amx * dq[1]);
  const float cp = 0.5f * (d[2] + amx * dq[2]), dp = 0.5f * (d[3] + amx * dq[3]);
  const float am = 0.5f * (d[4] - amx * dq[4]), bm = 0.5f * (d[3] - amx * dq[3]);
  const float cm = 0.5f * (d[2] - amx * dq[2]), dm = 0.5f * (d[1] - amx * dq[1]);
  const float Fs = -weno_phi(ap, bp, cp, dp) + weno_phi(am, bm, cm, dm);
#pragma unroll
  for (int v = 0; v < 8; ++v) {
    if (v == 5) continue;
    out[v] = out[v] + R[v] * Fs;
  }
}

// Interface fluxes along DIR (_weno_flux_x/_y/_z): one face per thread.
template <int DIR>
__global__ __launch_bounds__(NTH) void k_weno(const WenoArgs a) {
  constexpr int NC = Tile<DIR>::NCELL;
  __shared__ float sQ[8][NC];
  __shared__ float sF[8][NC];
  __shared__ float sS[4][NC];
  const Grid g = a.g;
  const Phys ph = a.ph;
  const int tid = threadIdx.x, w = tid >> 6, l = tid & 63;
  const int nbz = (g.nz + TBZ - 1) / TBZ;
  int bid = blockIdx.x;
  const int bz = bid % nbz;
  bid /= nbz;
  int x0, y0;
  if (DIR == 0) {  // tile: TP x-positions, fixed y
    y0 = bid % g.ny;
    x0 = (bid / g.ny) * TP;
  } else {         // tile: TP y-positions (or rows), fixed x
    const int nby = (g.ny + TP - 1) / TP;
    y0 = (bid % nby) * TP;
    x0 = bid / nby;
  }
  const int z0 = bz * TBZ;
  // stage the cells of the tile (frame of DIR) with their flux and speeds
  for (int c = tid; c < NC; c += NTH) {
    int x, y, z;
    if (DIR == 0) {
      x = wrapi(x0 - 2 + c / TBZ, g.nx);
      y = y0;
      z = wrapi(z0 + c % TBZ, g.nz);
    } else if (DIR == 1) {
      x = x0;
      y = wrapi(y0 - 2 + c / TBZ, g.ny);
      z = wrapi(z0 + c % TBZ, g.nz);
    } else {
      x = x0;
      y = wrapi(y0 + c / (TBZ + 5), g.ny);
      z = wrapi(z0 - 2 + c % (TBZ + 5), g.nz);
    }
    const long gc = cell(g, x, y, z);
    const float u[8] = {a.q[0][gc], a.q[1][gc], a.q[2][gc], a.q[3][gc], a.qt[0][gc], a.qt[1][gc], a.qt[2][gc], a.qt[3][gc]};
    float r[8], f[8], sp[4];
#pragma unroll
    for (int v = 0; v < 8; ++v) r[v] = u[perm(DIR, v)];
    cell_flux(r, f, ph);
    cell_speeds(r, sp, ph);
#pragma unroll
    for (int v = 0; v < 8; ++v) {
      sQ[v][c] = r[v];
      sF[v][c] = f[v];
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) sS[v][c] = sp[v];
  }
  __syncthreads();
  int xi, yi, zi;
  if (DIR == 0) {
    xi = x0 + w;
    yi = y0;
  } else {
    xi = x0;
    yi = y0 + w;
  }
  zi = z0 + l;
  if (xi >= g.nx || yi >= g.ny || zi >= g.nz) return;
  float qa[8], qb[8];
  const int c2 = win<DIR>(w, l, 2), c3 = win<DIR>(w, l, 3);
#pragma unroll
  for (int v = 0; v < 8; ++v) {
    qa[v] = sQ[v][c2];
    qb[v] = sQ[v][c3];
  }
  const EB e = interface_blocks(qa, qb, ph);
  float out[8];
  {
    const int c1 = win<DIR>(w, l, 1), c4 = win<DIR>(w, l, 4);
#pragma unroll
    for (int v = 0; v < 8; ++v)
      out[v] = (1.0f / 12.0f) * (((-sF[v][c1] + 7.f * sF[v][c2]) + 7.f * sF[v][c3]) - sF[v][c4]);
  }
  weno_mode<DIR, 0>(e, sQ, sF, sS, w, l, out);
  weno_mode<DIR, 1>(e, sQ, sF, sS, w, l, out);
  weno_mode<DIR, 2>(e, sQ, sF, sS, w, l, out);
  weno_mode<DIR, 3>(e, sQ, sF, sS, w, l, out);
  weno_mode<DIR, 4>(e, sQ, sF, sS, w, l, out);
  weno_mode<DIR, 5>(e, sQ, sF, sS, w, l, out);
  weno_mode<DIR, 6>(e, sQ, sF, sS, w, l, out);
  const long gc = cell(g, xi, yi, zi);
#pragma unroll
  for (int v = 0; v < 8; ++v) a.F[perm(DIR, v)][gc] = out[v];
  // CT modified fluxes: frame rows 6 / 7 plus c2f(B_d) c2f(v_t) (_constrained_transport.py:81-102)
  float bn[4], v2[4], v3[4];
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int c = win<DIR>(w, l, k + 1);
    const float ri = frcp(sQ[0][c]);
    bn[k] = sQ[5][c];
    v2[k] = sQ[2][c] * ri;
    v3[k] = sQ[3][c] * ri;
  }
  const float bf = (((-bn[0] + 9.f * bn[1]) + 9.f * bn[2]) - bn[3]) * (1.f / 16.f);
  const float v2f = (((-v2[0] + 9.f * v2[1]) + 9.f * v2[2]) - v2[3]) * (1.f / 16.f);
  const float v3f = (((-v3[0] + 9.f * v3[1]) + 9.f * v3[2]) - v3[3]) * (1.f / 16.f);
  a.modA[gc] = out[6] + bf * v2f;
  a.modB[gc] = out[7] + bf * v3f;
}

// ---------------------------------------------------------------- stencils
__device__ __forceinline__ float c2f(const Grid& g, const float* a, const Pos& p, int ax) {
  return (((-a[nb(g, p, ax, -1)] + 9.f * a[p.c]) + 9.f * a[nb(g, p, ax, 1)]) - a[nb(g, p, ax, 2)]) * (1.f / 16.f);
}

__device__ __forceinline__ float f2c(const Grid& g, const float* f, const Pos& p, int ax) {
  return (((((3.f * f[nb(g, p, ax, -3)] - 25.f * f[nb(g, p, ax, -2)]) + 150.f * f[nb(g, p, ax, -1)]) + 150.f * f[p.c]) -
           25.f * f[nb(g, p, ax, 1)]) +
          3.f * f[nb(g, p, ax, 2)]) *
         (1.f / 256.f);
}

__device__ __forceinline__ float d6(const Grid& g, const float* f, const Pos& p, int ax) {
  const float c1 = 75.f / 64.f, c2 = -25.f / 384.f, c3 = 3.f / 640.f;
  return (c1 * (f[p.c] - f[nb(g, p, ax, -1)]) + c2 * (f[nb(g, p, ax, 1)] - f[nb(g, p, ax, -2)])) +
         c3 * (f[nb(g, p, ax, 2)] - f[nb(g, p, ax, -3)]);
}

__device__ __forceinline__ float p2a(const Grid& g, const float* q, const Pos& p, int A, int B) {
  const float sa = ((q[nb(g, p, A, -1)] - 2.f * q[p.c]) + q[nb(g, p, A, 1)]) * (1.f / 24.f);
  const float sb = ((q[nb(g, p, B, -1)] - 2.f * q[p.c]) + q[nb(g, p, B, 1)]) * (1.f / 24.f);
  return (q[p.c] + sa) + sb;
}

#define FD_LOOP(g)                                                                     \
  for (unsigned c = blockIdx.x * blockDim.x + threadIdx.x; c < (unsigned)(g).nc; \
       c += gridDim.x * blockDim.x)

// _enforce_positivity on the stage state (in place, optional) and
// update_cell_center_fields into qt = (E, B_x, B_y, B_z)
struct PrepArgs {
  float* q[8];
  const float* b[3];
  float* qt[4];
  Grid g;
  Phys ph;
  int floors;
};

__global__ void k_prep(const PrepArgs a) {
  const Grid g = a.g;
  const Phys ph = a.ph;
  FD_LOOP(g) {
    const Pos P = pos_of(g, c);
    (void)P;
    float rho = a.q[0][c], E = a.q[4][c];
    const float Bx = a.q[5][c], By = a.q[6][c], Bz = a.q[7][c];
    const float b2 = (Bx * Bx + By * By) + Bz * Bz;
    if (a.floors) {
      rho = fmaxf(rho, ph.rhomin);
      const float ri = frcp(rho);
      const float vx = a.q[1][c] * ri, vy = a.q[2][c] * ri, vz = a.q[3][c] * ri;
      const float v2 = (vx * vx + vy * vy) + vz * vz;
      const float p = fmaxf(ph.gm1 * ((E - (0.5f * rho) * v2) - 0.5f * b2), ph.pgmin);
      E = (p * ph.inv_gm1 + (0.5f * rho) * v2) + 0.5f * b2;
      a.q[0][c] = rho;
      a.q[4][c] = E;
    }
    const float X = f2c(g, a.b[0], P, 0), Y = f2c(g, a.b[1], P, 1), Z = f2c(g, a.b[2], P, 2);
    const float nb2 = (X * X + Y * Y) + Z * Z;
    a.qt[0][c] = E + 0.5f * (nb2 - b2);
    a.qt[1][c] = X;
    a.qt[2][c] = Y;
    a.qt[3][c] = Z;
  }
}

// edge EMFs (_constrained_transport.py:104-131)
struct EdgeArgs {
  const float* fyx;  // B_y flux at x-faces (modified)
  const float* fzx;  // B_z flux at x-faces
  const float* fxy;  // B_x flux at y-faces
  const float* fzy;  // B_z flux at y-faces
  const float* fyz;  // B_y flux at z-faces
  const float* fxz;  // B_x flux at z-faces
  float* O[3];       // Omega_x, Omega_y, Omega_z
  Grid g;
};

__global__ void k_ct_edge(const EdgeArgs a) {
  const Grid g = a.g;
  FD_LOOP(g) {
    const Pos P = pos_of(g, c);
    (void)P;
    a.O[2][c] = c2f(g, a.fxy, P, 0) - c2f(g, a.fyx, P, 1);
    a.O[0][c] = c2f(g, a.fyz, P, 1) - c2f(g, a.fzy, P, 2);
    a.O[1][c] = c2f(g, a.fzx, P, 2) - c2f(g, a.fxz, P, 0);
  }
}

struct AvgArgs {
  const float* O[3];
  float* Ob[3];
  Grid g;
};

__global__ void k_ct_avg(const AvgArgs a) {
  const Grid g = a.g;
  FD_LOOP(g) {
    const Pos P = pos_of(g, c);
    (void)P;
    a.Ob[2][c] = p2a(g, a.O[2], P, 0, 1);
    a.Ob[0][c] = p2a(g, a.O[0], P, 1, 2);
    a.Ob[1][c] = p2a(g, a.O[1], P, 0, 2);
  }
}

// SSPRK(5,4) coefficients (_ssprk.py:247-272)
struct Coef {
  float k1, k2, k3, ff;  // q0 weight, dt weight, current weight, final-state weight
  int final_stage;
};

struct FaceArgs {
  const float* Ob[3];
  const float* b0[3];
  float* bc[3];
  float* bf[3];
  const float* dtp;
  float inv_dx;
  Coef k;
  Grid g;
};

// 6th-order curl of the averaged EMFs and the stage combine of the faces
__global__ void k_face(const FaceArgs a) {
  const Grid g = a.g;
  const float dtdx = (a.k.k2 * (*a.dtp)) * a.inv_dx;
  FD_LOOP(g) {
    const Pos P = pos_of(g, c);
    (void)P;
    const float r[3] = {(-dtdx) * d6(g, a.Ob[2], P, 1) + dtdx * d6(g, a.Ob[1], P, 2),
                        (-dtdx) * d6(g, a.Ob[0], P, 2) + dtdx * d6(g, a.Ob[2], P, 0),
                        (-dtdx) * d6(g, a.Ob[1], P, 0) + dtdx * d6(g, a.Ob[0], P, 1)};
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      if (a.k.final_stage) {
        a.bf[i][c] = a.bf[i][c] + r[i];
      } else {
        // a zero weight skips its buffer (the first stage: k3 = ff = 0)
        const float bn = a.k.k3 != 0.f ? (a.k.k1 * a.b0[i][c] + a.k.k3 * a.bc[i][c]) + r[i]
                                       : __fmul_rn(a.k.k1, a.b0[i][c]) + r[i];
        a.bc[i][c] = bn;
        if (a.k.ff != 0.f) a.bf[i][c] = a.bf[i][c] + bn * a.k.ff;
      }
    }
  }
}

struct QupdArgs {
  const float* Fx[8];
  const float* Fy[8];
  const float* Fz[8];
  const float* q0[8];
  float* qc[8];
  float* qf[8];
  const float* dtp;
  float inv_dx;
  Coef k;
  Grid g;
};

// flux divergence (_ssprk.py:224-228) and the stage combine (_ssprk.py:300-311)
__global__ void k_qupd(const QupdArgs a) {
  const Grid g = a.g;
  const float dtdx = (a.k.k2 * (*a.dtp)) * a.inv_dx;
  FD_LOOP(g) {
    const Pos P = pos_of(g, c);
    (void)P;
    const unsigned xm = nb(g, P, 0, -1), ym = nb(g, P, 1, -1), zm = nb(g, P, 2, -1);
#pragma unroll
    for (int v = 0; v < 8; ++v) {
      const float rhs = (-dtdx) * (((a.Fx[v][c] - a.Fx[v][xm]) + (a.Fy[v][c] - a.Fy[v][ym])) + (a.Fz[v][c] - a.Fz[v][zm]));
      if (a.k.final_stage) {
        a.qf[v][c] = a.qf[v][c] + rhs;
      } else {
        // a zero weight skips its buffer (the first stage: k3 = ff = 0)
        const float qn = a.k.k3 != 0.f ? (a.k.k1 * a.q0[v][c] + a.k.k3 * a.qc[v][c]) + rhs
                                       : __fmul_rn(a.k.k1, a.q0[v][c]) + rhs;
        a.qc[v][c] = qn;
        if (a.k.ff != 0.f) a.qf[v][c] = a.qf[v][c] + qn * a.k.ff;
      }
    }
  }
}

// primitive (11 rows) -> conserved stage buffers (_equations.py:65-101) and
// the SSPRK accumulators seeded with ff0 * (q0, b0) (_ssprk.py:321-324)
struct InitArgs {
  const float* w[11];
  float* q0[8];
  float* qc[8];
  float* qf[8];
  float* bc[3];
  float* bf[3];
  float ff0;
  Grid g;
  Phys ph;
};

__global__ void k_init(const InitArgs a) {
  const Grid g = a.g;
  const Phys ph = a.ph;
  FD_LOOP(g) {
    const Pos P = pos_of(g, c);
    (void)P;
    const float rho = a.w[0][c];
    const float vx = a.w[1][c], vy = a.w[2][c], vz = a.w[3][c];
    const float Bx = a.w[5][c], By = a.w[6][c], Bz = a.w[7][c];
    const float u2 = (vx * vx + vy * vy) + vz * vz;
    const float b2 = (Bx * Bx + By * By) + Bz * Bz;
    const float q[8] = {rho, rho * vx, rho * vy, rho * vz, (a.w[4][c] * ph.inv_gm1 + (0.5f * rho) * u2) + 0.5f * b2, Bx, By, Bz};
#pragma unroll
    for (int v = 0; v < 8; ++v) {
      a.q0[v][c] = q[v];
      a.qc[v][c] = q[v];
      a.qf[v][c] = a.ff0 * q[v];
    }
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      const float b = a.w[8 + i][c];
      a.bc[i][c] = b;
      a.bf[i][c] = a.ff0 * b;
    }
  }
}

// final update_cell_center_fields + positivity (_ssprk.py:336-347), then
// conserved -> primitive with floors and the faces appended (_evolve_state.py:63-71)
struct FinishArgs {
  const float* qf[8];
  const float* bf[3];
  float* w[11];
  Grid g;
  Phys ph;
};

__global__ void k_finish(const FinishArgs a) {
  const Grid g = a.g;
  const Phys ph = a.ph;
  FD_LOOP(g) {
    const Pos P = pos_of(g, c);
    (void)P;
    float rho = a.qf[0][c], E = a.qf[4][c];
    const float m[3] = {a.qf[1][c], a.qf[2][c], a.qf[3][c]};
    const float ob2 = (a.qf[5][c] * a.qf[5][c] + a.qf[6][c] * a.qf[6][c]) + a.qf[7][c] * a.qf[7][c];
    const float X = f2c(g, a.bf[0], P, 0), Y = f2c(g, a.bf[1], P, 1), Z = f2c(g, a.bf[2], P, 2);
    const float b2 = (X * X + Y * Y) + Z * Z;
    E = E + 0.5f * (b2 - ob2);
    if (ph.positivity) {
      rho = fmaxf(rho, ph.rhomin);
      const float ri = frcp(rho);
      const float v2 = ((m[0] * ri) * (m[0] * ri) + (m[1] * ri) * (m[1] * ri)) + (m[2] * ri) * (m[2] * ri);
      const float p = fmaxf(ph.gm1 * ((E - (0.5f * rho) * v2) - 0.5f * b2), ph.pgmin);
      E = (p * ph.inv_gm1 + (0.5f * rho) * v2) + 0.5f * b2;
    }
    const float ri = frcp(rho);
    const float ux = m[0] * ri, uy = m[1] * ri, uz = m[2] * ri;
    const float u2 = (ux * ux + uy * uy) + uz * uz;
    float p = ph.gm1 * ((E - (0.5f * rho) * u2) - 0.5f * b2);
    float r = rho;
    if (ph.positivity) {
      r = fmaxf(r, ph.rhomin);
      p = fmaxf(p, ph.pgmin);
    }
    a.w[0][c] = r;
    a.w[1][c] = ux;
    a.w[2][c] = uy;
    a.w[3][c] = uz;
    a.w[4][c] = p;
    a.w[5][c] = X;
    a.w[6][c] = Y;
    a.w[7][c] = Z;
    a.w[8][c] = a.bf[0][c];
    a.w[9][c] = a.bf[1][c];
    a.w[10][c] = a.bf[2][c];
  }
}

// step control of the FD loop: dt, t and the slotted CFL partial maxima
struct Ctl {
  float dt;
  float t;
  int steps;
  int pad;
  unsigned cflp[3][NSLOT];
};

struct StepParams {
  float t_end, C, dx, dt_max;
  int fixed, num_timesteps, exact_end, use_cfl;
  int clamp_tp = 0;  // clamp dt to tp - t (snapshot time points, time_integration.py:671-674)
  float tp = 0.f;
};

// dt = min(C dx / (lx + ly + lz), dt_max) (_timestep_estimator.py(FD):89-90),
// or t_end / num_timesteps; clamp to t_end - t (time_integration.py:667-678)
__global__ void k_step_begin(Ctl* c, StepParams p) {
  float dt;
  if (p.fixed) {
    dt = p.t_end / (float)p.num_timesteps;
  } else if (p.use_cfl) {
    float l[3];
    for (int d = 0; d < 3; ++d) l[d] = __uint_as_float(slots_max(c->cflp[d]));
    dt = (p.C * p.dx) / ((l[0] + l[1]) + l[2]);
    dt = p.dt_max < dt ? p.dt_max : dt;
  } else {
    dt = c->dt;
  }
  if (p.clamp_tp) {
    const float rem = p.tp - c->t;
    dt = rem < dt ? rem : dt;
  }
  if (p.exact_end) {
    const float rem = p.t_end - c->t;
    dt = rem < dt ? rem : dt;
  }
  c->dt = dt;
  for (int d = 0; d < 3; ++d)
    for (int i = 0; i < NSLOT; ++i) c->cflp[d][i] = 0u;
}

__global__ void k_set_dt(Ctl* c, const float* dtp) { c->dt = *dtp; }

__global__ void k_step_end(Ctl* c) {
  c->t = c->t + c->dt;
  c->steps += 1;
}

__global__ void k_reset(Ctl* c) {
  c->dt = 0.f;
  c->t = 0.f;
  c->steps = 0;
  for (int d = 0; d < 3; ++d)
    for (int i = 0; i < NSLOT; ++i) c->cflp[d][i] = 0u;
}

__global__ void k_dt_out(const Ctl* c, float* out) { *out = c->dt; }

}  // namespace fd
}  // namespace amx

// ============================================================================
// C ABI
// ============================================================================

using namespace amx;
using namespace amx::fd;

struct amx_fd_ctx {
  amx_fd_config cfg;
  Grid g;
  Phys ph;
  float dx, inv_dx;
  float* ws;
  // buffer offsets (fields of nc floats)
  float *Q0[8], *QC[8], *QF[8], *QT[4], *FX[8], *FY[8], *FZ[8], *MOD[6], *OM[3], *OMB[3], *BC[3], *BF[3];
  Ctl* ctl;
  Ctl* hctl;
  int timing;
  std::vector<hipEvent_t> ev;
  std::vector<int> pend;  // pairs of event indices of WENO launches
  int ev_used;
  double weno_ms;
  long weno_launches;
  void* snapwork;  // snapshot diagnostics partials (allocated on first use)
  amx_step_hook hook;  // between-step hook of the step loop (forcing), or null
  void* hook_user;
};

namespace {

constexpr int FD_FIELDS = 8 * 3 + 4 + 8 * 3 + 6 + 3 + 3 + 3 + 3;  // 70

amx_status fd_validate(const amx_fd_config* c) {
  if (!c) return amx_fail(AMX_ERR_INVALID, "null config%s");
  if (c->nx < 6 || c->ny < 6 || c->nz < 6) return amx_fail(AMX_ERR_INVALID, "FD grid needs >= 6 cells per axis%s");
  if (!(c->grid_spacing > 0)) return amx_fail(AMX_ERR_INVALID, "grid_spacing must be positive%s");
  if ((long)c->nx * c->ny * c->nz >= (1L << 30))
    return amx_fail(AMX_ERR_INVALID, "a grid of 2^30 cells or more is not supported (32-bit cell indices)%s");
  return AMX_OK;
}

dim3 fd_grid(const amx_fd_ctx* c) {
  long b = (c->g.nc + 255) / 256;
  if (b > 256L * 64) b = 256L * 64;
  return dim3((unsigned)b);
}

int fd_tbegin(amx_fd_ctx* c, hipStream_t st) {
  if (!c->timing) return -1;
  if (c->ev_used + 2 > (int)c->ev.size()) {
    for (int i = 0; i < 2; ++i) {
      hipEvent_t e;
      if (hipEventCreate(&e) != hipSuccess) return -1;
      c->ev.push_back(e);
    }
  }
  const int i = c->ev_used;
  c->ev_used += 2;
  (void)hipEventRecord(c->ev[i], st);
  return i;
}

void fd_tend(amx_fd_ctx* c, hipStream_t st, int i) {
  if (i < 0) return;
  (void)hipEventRecord(c->ev[i + 1], st);
  c->pend.push_back(i);
}

void fd_tresolve(amx_fd_ctx* c) {
  for (int i : c->pend) {
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, c->ev[i], c->ev[i + 1]) == hipSuccess) {
      c->weno_ms += ms;
      c->weno_launches += 1;
    }
  }
  c->pend.clear();
  c->ev_used = 0;
}

template <int DIR>
hipError_t launch_weno(amx_fd_ctx* c, float* const* F, float* modA, float* modB, hipStream_t st) {
  WenoArgs a;
  for (int v = 0; v < 4; ++v) {
    a.q[v] = c->QC[v];
    a.qt[v] = c->QT[v];
  }
  for (int v = 0; v < 8; ++v) a.F[v] = F[v];
  a.modA = modA;
  a.modB = modB;
  a.g = c->g;
  a.ph = c->ph;
  const int nbz = (c->g.nz + TBZ - 1) / TBZ;
  long nb;
  if (DIR == 0)
    nb = (long)nbz * c->g.ny * ((c->g.nx + TP - 1) / TP);
  else
    nb = (long)nbz * c->g.nx * ((c->g.ny + TP - 1) / TP);
  const int e = fd_tbegin(c, st);
  hipLaunchKernelGGL(k_weno<DIR>, dim3((unsigned)nb), dim3(NTH), 0, st, a);
  fd_tend(c, st, e);
  return hipGetLastError();
}

// SSPRK(5,4) coefficients (_ssprk.py:319-343)
const float K1[5] = {1.0f, 0.44437049406734f, 0.62010185138540f, 0.17807995410773f, -2.081261929715610e-02f};
const float K2[5] = {0.39175222700392f, 0.36841059262959f, 0.25189177424738f, 0.54497475021237f, 0.22600748319395f};
const float K3[5] = {0.0f, 0.55562950593266f, 0.37989814861460f, 0.82192004589227f, 5.03580947213895e-01f};
const float FF[5] = {-2.081261929715610e-02f, 0.0f, 0.51723167208978f, -6.518979800418380e-12f, 5.03580947213895e-01f};

// one _evolve_state_fd on the (11, nx, ny, nz) state with dt in ctl->dt
amx_status fd_step(amx_fd_ctx* c, float* state, hipStream_t st) {
  const long nc = c->g.nc;
  float* w[11];
  for (int v = 0; v < 11; ++v) w[v] = state + (long)v * nc;
  const dim3 grid = fd_grid(c), blk(256);
  InitArgs ia;
  for (int v = 0; v < 11; ++v) ia.w[v] = w[v];
  for (int v = 0; v < 8; ++v) {
    ia.q0[v] = c->Q0[v];
    ia.qc[v] = c->QC[v];
    ia.qf[v] = c->QF[v];
  }
  for (int i = 0; i < 3; ++i) {
    ia.bc[i] = c->BC[i];
    ia.bf[i] = c->BF[i];
  }
  ia.ff0 = FF[0];
  ia.g = c->g;
  ia.ph = c->ph;
  hipLaunchKernelGGL(k_init, grid, blk, 0, st, ia);
  AMX_HIPCHK(hipGetLastError());
  const float* dtp = &c->ctl->dt;
  for (int s = 0; s < 5; ++s) {
    Coef k;
    k.k1 = K1[s];
    k.k2 = K2[s];
    k.k3 = K3[s];
    k.ff = s < 4 ? FF[s + 1] : 0.f;
    k.final_stage = s == 4;
    PrepArgs pa;
    for (int v = 0; v < 8; ++v) pa.q[v] = c->QC[v];
    for (int i = 0; i < 3; ++i) pa.b[i] = c->BC[i];
    for (int v = 0; v < 4; ++v) pa.qt[v] = c->QT[v];
    pa.g = c->g;
    pa.ph = c->ph;
    pa.floors = (s < 4) && c->ph.positivity;
    hipLaunchKernelGGL(k_prep, grid, blk, 0, st, pa);
    AMX_HIPCHK(hipGetLastError());
    AMX_HIPCHK(launch_weno<0>(c, c->FX, c->MOD[0], c->MOD[1], st));
    AMX_HIPCHK(launch_weno<1>(c, c->FY, c->MOD[2], c->MOD[3], st));
    AMX_HIPCHK(launch_weno<2>(c, c->FZ, c->MOD[4], c->MOD[5], st));
    EdgeArgs ea;
    ea.fyx = c->MOD[0];
    ea.fzx = c->MOD[1];
    ea.fxy = c->MOD[2];
    ea.fzy = c->MOD[3];
    ea.fyz = c->MOD[4];
    ea.fxz = c->MOD[5];
    for (int i = 0; i < 3; ++i) ea.O[i] = c->OM[i];
    ea.g = c->g;
    hipLaunchKernelGGL(k_ct_edge, grid, blk, 0, st, ea);
    AMX_HIPCHK(hipGetLastError());
    AvgArgs av;
    for (int i = 0; i < 3; ++i) {
      av.O[i] = c->OM[i];
      av.Ob[i] = c->OMB[i];
    }
    av.g = c->g;
    hipLaunchKernelGGL(k_ct_avg, grid, blk, 0, st, av);
    AMX_HIPCHK(hipGetLastError());
    FaceArgs fa;
    for (int i = 0; i < 3; ++i) {
      fa.Ob[i] = c->OMB[i];
      fa.b0[i] = w[8 + i];
      fa.bc[i] = c->BC[i];
      fa.bf[i] = c->BF[i];
    }
    fa.dtp = dtp;
    fa.inv_dx = c->inv_dx;
    fa.k = k;
    fa.g = c->g;
    hipLaunchKernelGGL(k_face, grid, blk, 0, st, fa);
    AMX_HIPCHK(hipGetLastError());
    QupdArgs qa;
    for (int v = 0; v < 8; ++v) {
      qa.Fx[v] = c->FX[v];
      qa.Fy[v] = c->FY[v];
      qa.Fz[v] = c->FZ[v];
      qa.q0[v] = c->Q0[v];
      qa.qc[v] = c->QC[v];
      qa.qf[v] = c->QF[v];
    }
    qa.dtp = dtp;
    qa.inv_dx = c->inv_dx;
    qa.k = k;
    qa.g = c->g;
    hipLaunchKernelGGL(k_qupd, grid, blk, 0, st, qa);
    AMX_HIPCHK(hipGetLastError());
  }
  FinishArgs fa;
  for (int v = 0; v < 8; ++v) fa.qf[v] = c->QF[v];
  for (int i = 0; i < 3; ++i) fa.bf[i] = c->BF[i];
  for (int v = 0; v < 11; ++v) fa.w[v] = w[v];
  fa.g = c->g;
  fa.ph = c->ph;
  hipLaunchKernelGGL(k_finish, grid, blk, 0, st, fa);
  AMX_HIPCHK(hipGetLastError());
  return AMX_OK;
}

}  // namespace

extern "C" amx_status amx_fd_workspace_bytes(const amx_fd_config* cfg, size_t* bytes) {
  amx_status s = fd_validate(cfg);
  if (s != AMX_OK) return s;
  if (!bytes) return amx_fail(AMX_ERR_INVALID, "null output%s");
  *bytes = sizeof(float) * (size_t)FD_FIELDS * cfg->nx * cfg->ny * cfg->nz + sizeof(Ctl);
  return AMX_OK;
}

extern "C" amx_status amx_fd_create(const amx_fd_config* cfg, amx_fd_ctx** out) {
  amx_status s = fd_validate(cfg);
  if (s != AMX_OK) return s;
  if (!out) return amx_fail(AMX_ERR_INVALID, "null output%s");
  amx_fd_ctx* c = new (std::nothrow) amx_fd_ctx();
  if (!c) return amx_fail(AMX_ERR_NOMEM, "host allocation failed%s");
  c->cfg = *cfg;
  c->g.nx = cfg->nx;
  c->g.ny = cfg->ny;
  c->g.nz = cfg->nz;
  c->g.sy = cfg->nz;
  c->g.sx = (long)cfg->ny * cfg->nz;
  c->g.nc = (long)cfg->nx * cfg->ny * cfg->nz;
  c->ph.gamma = cfg->gamma;
  c->ph.gm1 = cfg->gamma - 1.0f;
  c->ph.inv_gm1 = 1.0f / c->ph.gm1;
  c->ph.rhomin = cfg->minimum_density;
  c->ph.pgmin = cfg->minimum_pressure;
  c->ph.positivity = cfg->enforce_positivity ? 1 : 0;
  c->dx = (float)cfg->grid_spacing;
  c->inv_dx = (float)(1.0 / cfg->grid_spacing);
  const long nc = c->g.nc;
  if (hipMalloc((void**)&c->ws, sizeof(float) * (size_t)FD_FIELDS * nc) != hipSuccess) {
    delete c;
    return amx_fail(AMX_ERR_NOMEM, "device workspace allocation failed%s");
  }
  float* p = c->ws;
  auto take = [&](float** dst, int n) {
    for (int i = 0; i < n; ++i) {
      dst[i] = p;
      p += nc;
    }
  };
  take(c->Q0, 8);
  take(c->QC, 8);
  take(c->QF, 8);
  take(c->QT, 4);
  take(c->FX, 8);
  take(c->FY, 8);
  take(c->FZ, 8);
  take(c->MOD, 6);
  take(c->OM, 3);
  take(c->OMB, 3);
  take(c->BC, 3);
  take(c->BF, 3);
  if (hipMalloc((void**)&c->ctl, sizeof(Ctl)) != hipSuccess ||
      hipHostMalloc((void**)&c->hctl, sizeof(Ctl), hipHostMallocDefault) != hipSuccess) {
    (void)hipFree(c->ws);
    delete c;
    return amx_fail(AMX_ERR_NOMEM, "control allocation failed%s");
  }
  hipLaunchKernelGGL(k_reset, dim3(1), dim3(1), 0, 0, c->ctl);
  if (hipDeviceSynchronize() != hipSuccess) return amx_fail(AMX_ERR_HIP, "init sync failed%s");
  *out = c;
  return AMX_OK;
}

extern "C" amx_status amx_fd_destroy(amx_fd_ctx* c) {
  if (!c) return AMX_OK;
  (void)hipDeviceSynchronize();
  (void)hipFree(c->ws);
  (void)hipFree(c->ctl);
  (void)hipHostFree(c->hctl);
  if (c->snapwork) (void)hipFree(c->snapwork);
  for (auto e : c->ev) (void)hipEventDestroy(e);
  delete c;
  return AMX_OK;
}

namespace amx {
namespace fd {

struct CflArgs {
  const float* w[8];
  Grid g;
  Phys ph;
  Ctl* ctl;
};

// _cfl_time_step_fd maxima: per direction, max over cells and the 7 modes of
// |lambda| of the conserved state of w (_timestep_estimator.py(FD):25-87)
__global__ void k_cfl(const CflArgs a) {
  __shared__ unsigned sred[3][4];
  const Grid g = a.g;
  const Phys ph = a.ph;
  float mx[3] = {0.f, 0.f, 0.f};
  FD_LOOP(g) {
    const Pos P = pos_of(g, c);
    (void)P;
    const float rho = a.w[0][c];
    const float vx = a.w[1][c], vy = a.w[2][c], vz = a.w[3][c];
    const float Bx = a.w[5][c], By = a.w[6][c], Bz = a.w[7][c];
    const float u2 = (vx * vx + vy * vy) + vz * vz;
    const float b2 = (Bx * Bx + By * By) + Bz * Bz;
    const float u[8] = {rho, rho * vx, rho * vy, rho * vz, (a.w[4][c] * ph.inv_gm1 + (0.5f * rho) * u2) + 0.5f * b2, Bx, By, Bz};
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      float r[8], sp[4];
#pragma unroll
      for (int v = 0; v < 8; ++v) r[v] = u[perm(d, v)];
      cell_speeds(r, sp, ph);
      const float l[7] = {sp[0] - sp[1], sp[0] - sp[2], sp[0] - sp[3], sp[0], sp[0] + sp[3], sp[0] + sp[2], sp[0] + sp[1]};
#pragma unroll
      for (int s = 0; s < 7; ++s) {
        const float m = fabsf(l[s]);
        mx[d] = (m != m || m > mx[d]) ? m : mx[d];
      }
    }
  }
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    const unsigned m = block_max_u<4>(canon_bits(mx[d]), sred[d]);
    if (threadIdx.x == 0 && m) atomicMax(&a.ctl->cflp[d][blockIdx.x & (NSLOT - 1)], m);
  }
}

}  // namespace fd
}  // namespace amx

namespace {

hipError_t fd_cfl(amx_fd_ctx* c, const float* state, hipStream_t st) {
  CflArgs a;
  for (int v = 0; v < 8; ++v) a.w[v] = state + (long)v * c->g.nc;
  a.g = c->g;
  a.ph = c->ph;
  a.ctl = c->ctl;
  hipLaunchKernelGGL(k_cfl, fd_grid(c), dim3(256), 0, st, a);
  return hipGetLastError();
}

StepParams fd_params(const amx_fd_ctx* c, const amx_fv_run* run, int use_cfl) {
  StepParams p;
  p.t_end = run ? run->t_end : 0.f;
  p.C = c->cfg.C_cfl;
  p.dx = c->dx;
  p.dt_max = c->cfg.dt_max;
  p.fixed = run ? run->fixed_timestep : 0;
  p.num_timesteps = run ? run->num_timesteps : 1;
  p.exact_end = run ? run->exact_end_time : 0;
  p.use_cfl = use_cfl;
  return p;
}

}  // namespace

extern "C" amx_status amx_fd_cfl_dt(amx_fd_ctx* c, const float* state, float* d_dt, void* stream) {
  if (!c || !state || !d_dt) return amx_fail(AMX_ERR_INVALID, "null argument%s");
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_reset, dim3(1), dim3(1), 0, st, c->ctl);
  AMX_HIPCHK(hipGetLastError());
  AMX_HIPCHK(fd_cfl(c, state, st));
  hipLaunchKernelGGL(k_step_begin, dim3(1), dim3(1), 0, st, c->ctl, fd_params(c, nullptr, 1));
  hipLaunchKernelGGL(k_dt_out, dim3(1), dim3(1), 0, st, (const Ctl*)c->ctl, d_dt);
  AMX_HIPCHK(hipGetLastError());
  return AMX_OK;
}

extern "C" amx_status amx_fd_evolve_state(amx_fd_ctx* c, float* state, const float* d_dt, void* stream) {
  if (!c || !state || !d_dt) return amx_fail(AMX_ERR_INVALID, "null argument%s");
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_set_dt, dim3(1), dim3(1), 0, st, c->ctl, d_dt);
  AMX_HIPCHK(hipGetLastError());
  return fd_step(c, state, st);
}

namespace {

// Layout of an FD state (registered_variables.py:188-198) for the snapshot
// diagnostics: max |div B| is taken over the face fields.
amx_diag_desc fd_diag_desc(const amx_fd_ctx* c, double box_size) {
  amx_diag_desc d;
  d.num_vars = 11;
  d.dim = 3;
  d.nx = c->cfg.nx;
  d.ny = c->cfg.ny;
  d.nz = c->cfg.nz;
  d.rho_row = 0;
  d.vel_row = 1;
  d.p_row = 4;
  d.mag_row = 8;
  d.div_mode = AMX_DIV_FACES;
  d.gamma = c->cfg.gamma;
  d.grid_spacing = c->cfg.grid_spacing;
  d.box_size = box_size;
  return d;
}

// The FD branch of the step loop (time_integration.py:629-801), with the
// snapshot rows of time_integration.py:539-610, 736-745 when snp is given.
amx_status fd_loop(amx_fd_ctx* c, float* state, const amx_fv_run* run, amx::snap::Driver* snp, amx_fv_stats* stats,
                   hipStream_t st) {
  hipLaunchKernelGGL(k_reset, dim3(1), dim3(1), 0, st, c->ctl);
  AMX_HIPCHK(hipGetLastError());
  StepParams sp = fd_params(c, run, 1);
  // the end-time clamp is off with snapshot time points (time_integration.py:677)
  if (snp && snp->spec->use_timepoints) sp.exact_end = 0;
  long steps = 0;
  float t = 0.f;
  for (;;) {
    if (run->fixed_timestep) {
      if (steps >= run->num_timesteps) break;
    } else if (!(t < run->t_end)) {
      break;
    }
    if (run->max_steps >= 0 && steps >= run->max_steps) break;
    if (snp) {
      amx_status s = snp->begin(t, state, st);
      if (s != AMX_OK) return s;
      sp.clamp_tp = snp->clamp();
      if (sp.clamp_tp) sp.tp = snp->clamp_target();
    }
    if (!run->fixed_timestep) AMX_HIPCHK(fd_cfl(c, state, st));
    hipLaunchKernelGGL(k_step_begin, dim3(1), dim3(1), 0, st, c->ctl, sp);
    AMX_HIPCHK(hipGetLastError());
    if (c->hook && c->hook(c->hook_user, state, &c->ctl->dt, steps, st) != 0)
      return amx_fail(AMX_ERR_INVALID, "the step hook stopped the run%s");
    amx_status s = fd_step(c, state, st);
    if (s != AMX_OK) return s;
    hipLaunchKernelGGL(k_step_end, dim3(1), dim3(1), 0, st, c->ctl);
    AMX_HIPCHK(hipGetLastError());
    AMX_HIPCHK(hipMemcpyAsync(c->hctl, c->ctl, offsetof(Ctl, cflp), hipMemcpyDeviceToHost, st));
    AMX_HIPCHK(hipStreamSynchronize(st));
    fd_tresolve(c);
    t = c->hctl->t;
    ++steps;
    if (snp) {
      s = snp->end(t, state, st);
      if (s != AMX_OK) return s;
    }
  }
  if (snp) snp->finish(steps);
  if (stats) {
    stats->steps = steps;
    stats->t = t;
    stats->fixed_point_iterations = 0;
    stats->max_fixed_point_iterations = 0;
  }
  return AMX_OK;
}

}  // namespace

extern "C" amx_status amx_fd_time_integration(amx_fd_ctx* c, float* state, const amx_fv_run* run, amx_fv_stats* stats,
                                              void* stream) {
  if (!c || !state || !run) return amx_fail(AMX_ERR_INVALID, "null argument%s");
  if (run->fixed_timestep && run->num_timesteps <= 0) return amx_fail(AMX_ERR_INVALID, "num_timesteps must be positive%s");
  return fd_loop(c, state, run, nullptr, stats, (hipStream_t)stream);
}

extern "C" amx_status amx_fd_time_integration_snapshots(amx_fd_ctx* c, float* state, const amx_fv_run* run,
                                                        const amx_snapshot_spec* spec, amx_snapshot_out* out,
                                                        amx_fv_stats* stats, void* stream) {
  if (!c || !state || !run || !spec) return amx_fail(AMX_ERR_INVALID, "null argument%s");
  if (run->fixed_timestep && run->num_timesteps <= 0) return amx_fail(AMX_ERR_INVALID, "num_timesteps must be positive%s");
  amx::snap::Driver d;
  d.spec = spec;
  d.out = out;
  d.desc = fd_diag_desc(c, spec->box_size);
  d.t_end = run->t_end;
  d.work = nullptr;
  amx_status s = amx::snap::validate(spec, out, d.desc);
  if (s != AMX_OK) return s;
  if (!spec->callback) {
    if (!c->snapwork && hipMalloc(&c->snapwork, amx::snap::work_bytes()) != hipSuccess) {
      c->snapwork = nullptr;
      return amx_fail(AMX_ERR_NOMEM, "snapshot workspace allocation failed%s");
    }
    d.work = c->snapwork;
  }
  return fd_loop(c, state, run, &d, stats, (hipStream_t)stream);
}

extern "C" amx_status amx_fd_set_step_hook(amx_fd_ctx* c, amx_step_hook hook, void* user) {
  if (!c) return amx_fail(AMX_ERR_INVALID, "null context%s");
  c->hook = hook;
  c->hook_user = user;
  return AMX_OK;
}

extern "C" amx_status amx_fd_set_timing(amx_fd_ctx* c, int enable) {
  if (!c) return amx_fail(AMX_ERR_INVALID, "null context%s");
  AMX_HIPCHK(hipDeviceSynchronize());
  c->pend.clear();
  c->ev_used = 0;
  c->timing = enable ? 1 : 0;
  c->weno_ms = 0.0;
  c->weno_launches = 0;
  return AMX_OK;
}

extern "C" amx_status amx_fd_get_timing(amx_fd_ctx* c, double* weno_ms, long* weno_launches) {
  if (!c || !weno_ms || !weno_launches) return amx_fail(AMX_ERR_INVALID, "null argument%s");
  AMX_HIPCHK(hipDeviceSynchronize());
  fd_tresolve(c);
  *weno_ms = c->weno_ms;
  *weno_launches = c->weno_launches;
  return AMX_OK;
}

namespace amx {
namespace fd {

// b_d = 4th-order centre-to-face interpolation of B_d along d
// (initialize_interface_fields, _constrained_transport.py:201-213)
struct C2fArgs {
  const float* B[3];
  float* b[3];
  Grid g;
};

__global__ void k_interface_fields(const C2fArgs a) {
  const Grid g = a.g;
  FD_LOOP(g) {
    const Pos P = pos_of(g, c);
    (void)P;
#pragma unroll
    for (int d = 0; d < 3; ++d) a.b[d][c] = c2f(g, a.B[d], P, d);
  }
}

}  // namespace fd
}  // namespace amx

extern "C" amx_status amx_fd_interface_fields(const float* B, float* b, int nx, int ny, int nz, void* stream) {
  if (!B || !b) return amx_fail(AMX_ERR_INVALID, "null argument%s");
  if (nx < 3 || ny < 3 || nz < 3) return amx_fail(AMX_ERR_INVALID, "grid needs >= 3 cells per axis%s");
  if (B == b) return amx_fail(AMX_ERR_INVALID, "B and b must not alias%s");
  C2fArgs a;
  a.g.nx = nx;
  a.g.ny = ny;
  a.g.nz = nz;
  a.g.sy = nz;
  a.g.sx = (long)ny * nz;
  a.g.nc = (long)nx * ny * nz;
  for (int d = 0; d < 3; ++d) {
    a.B[d] = B + d * a.g.nc;
    a.b[d] = b + d * a.g.nc;
  }
  long nb = (a.g.nc + 255) / 256;
  if (nb > 256L * 64) nb = 256L * 64;
  hipLaunchKernelGGL(k_interface_fields, dim3((unsigned)nb), dim3(256), 0, (hipStream_t)stream, a);
  AMX_HIPCHK(hipGetLastError());
  return AMX_OK;
}
