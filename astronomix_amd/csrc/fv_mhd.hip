// MI355X (gfx950) kernels and C ABI of the finite-volume MHD step.
//
// One kernel per global-synchronisation interval of the reference dataflow:
//   * k_stage   one forward-Euler unsplit gas stage (minmod reconstruction +
//               HLL/HLLC fluxes + flux-divergence update), optionally fused with
//               the SSP-RK2 Heun average, the Pang-Wu energy fix on its inputs,
//               the CFL maxima of its output and the copy of the final B.
//               Follows _evolve_gas_state_unsplit_inner / _evolve_gas_state_unsplit
//               (astronomix/_finite_volume/_state_evolution/evolve_state.py:338-469).
//   * k_mag     one fixed-point iterate of the implicit-midpoint magnetic
//               sub-step with its convergence residual
//               (astronomix/_finite_volume/_magnetic_update/_magnetic_field_update.py:82-165).
//   * k_cfl     the unsplit CFL maxima (_timestep_estimator.py:118-128).
//
// Stencil kernels march along the slowest array axis (x in 3D) with a register
// queue of the thread's own column, and keep the current (y, z) plane of a
// tile in LDS with its halo, so each cell value is read from HBM once per pass
// and each interface flux is computed once. Tiles are 64 cells wide along the
// contiguous axis (one wavefront, coalesced) and TA rows tall; the block index
// is swizzled so that neighbouring tiles land on the same XCD (shared L2).
// Periodic boundaries are applied by wrapping indices (the reference's ghost
// cells are periodic copies refreshed before every stencil, boundaries.py:296-319).

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <vector>

#include "amx.h"

namespace amx {

constexpr int TB = 64;      // tile width along the contiguous axis (one wavefront)
constexpr int MAXRES = 1030;  // residual slots (fixed-point cap + 2)

// Device-resident step control: step size, time, fixed-point bookkeeping and
// the CFL maxima (as float bits; all values are >= 0 so unsigned max == float max).
struct Ctl {
  float dt;
  float t;
  int iters;
  int steps;
  unsigned cfl[4];
  unsigned res[MAXRES];
};

struct Phys {
  float gamma;    // params.gamma
  float gm1;      // gamma - 1 evaluated in float, as the reference does
  float inv_gm1;  // 1 / (gamma - 1)
  float inv_dx;   // float(1 / grid_spacing)
};

// gamma is hard-coded to 5/3 in the magnetic energy fix (_magnetic_field_update.py:213,221)
constexpr float G53M1 = 1.6666666f - 1.0f;

__device__ __forceinline__ int wrapi(int i, int n) {
  while (i < 0) i += n;
  while (i >= n) i -= n;
  return i;
}

// limiters.py:13-15: 0.5 (sgn a + sgn b) min(|a|,|b|) is the argument of
// smaller magnitude when both have the same strict sign, and 0 otherwise.
__device__ __forceinline__ float minmod(float a, float b) {
  const bool same = (a > 0.f && b > 0.f) || (a < 0.f && b < 0.f);
  return same ? (fabsf(a) < fabsf(b) ? a : b) : 0.f;
}

// limited gradient times dx/2 (limited_gradients.py:50-60, reconstruction.py:409-412)
__device__ __forceinline__ float half_slope(float wm, float w0, float wp) {
  return 0.5f * minmod(w0 - wm, wp - w0);
}

template <int DIM>
__device__ __forceinline__ float vsq(const float* w) {
  float s = w[1] * w[1] + w[2] * w[2];
  if (DIM == 3) s += w[3] * w[3];
  return s;
}

// conserved_state_from_primitive (_fluid_equations/_equations.py:103-156)
template <int DIM>
__device__ __forceinline__ void cons(const float* w, float* u, const Phys& ph) {
  constexpr int P = DIM + 1;
  u[0] = w[0];
#pragma unroll
  for (int i = 1; i <= DIM; ++i) u[i] = w[0] * w[i];
  u[P] = w[P] * ph.inv_gm1 + (0.5f * w[0]) * vsq<DIM>(w);
}

// primitive_state_from_conserved (_fluid_equations/_equations.py:23-94)
template <int DIM>
__device__ __forceinline__ void prim(const float* u, float* w, const Phys& ph) {
  constexpr int P = DIM + 1;
  const float ri = 1.0f / u[0];
  w[0] = u[0];
#pragma unroll
  for (int i = 1; i <= DIM; ++i) w[i] = u[i] * ri;
  const float s = vsq<DIM>(w) + 1e-20f;
  w[P] = (ph.gm1 * u[0]) * (u[P] * ri - 0.5f * s);
}

// HLL interface flux (_finite_volume/_riemann_solver/hll.py:33-109) with the
// Euler flux of _fluid_equations/_fluxes.py:22-93. AX is the velocity row of
// the flux direction.
template <int DIM, int AX>
__device__ __forceinline__ void flux_hll(const float* wl, const float* wr, float* f, const Phys& ph) {
  constexpr int P = DIM + 1;
  const float rl = wl[0], ul = wl[AX], pl = wl[P];
  const float rr = wr[0], ur = wr[AX], pr = wr[P];
  const float cl = sqrtf(ph.gamma * pl / rl);
  const float cr = sqrtf(ph.gamma * pr / rr);
  const float El = pl * ph.inv_gm1 + (0.5f * rl) * vsq<DIM>(wl);
  const float Er = pr * ph.inv_gm1 + (0.5f * rr) * vsq<DIM>(wr);
  const float sp = fmaxf(fmaxf(ul + cl, ur + cr), 0.f);
  const float sm = fminf(fminf(ul - cl, ur - cr), 0.f);
  const float inv = 1.0f / (sp - sm);
  const float smsp = sm * sp;
  f[0] = ((sp * (ul * rl) - sm * (ur * rr)) + smsp * (rr - rl)) * inv;
#pragma unroll
  for (int i = 1; i <= DIM; ++i) {
    const float ml = rl * wl[i], mr = rr * wr[i];
    float fl = ul * ml, fr = ur * mr;
    if (i == AX) {
      fl += pl;
      fr += pr;
    }
    f[i] = ((sp * fl - sm * fr) + smsp * (mr - ml)) * inv;
  }
  f[P] = ((sp * (ul * (pl + El)) - sm * (ur * (pr + Er))) + smsp * (Er - El)) * inv;
}

// HLLC interface flux, central form (_finite_volume/_riemann_solver/hll.py:113-261).
template <int DIM, int AX>
__device__ __forceinline__ void flux_hllc(const float* wl, const float* wr, float* f, const Phys& ph) {
  constexpr int P = DIM + 1, NV = DIM + 2;
  const float rl = wl[0], ul = wl[AX], pl = wl[P];
  const float rr = wr[0], ur = wr[AX], pr = wr[P];
  const float cl = sqrtf(ph.gamma * pl / rl);
  const float cr = sqrtf(ph.gamma * pr / rr);
  float FL[NV], FR[NV], UL[NV], UR[NV];
  const float El = pl * ph.inv_gm1 + (0.5f * rl) * vsq<DIM>(wl);
  const float Er = pr * ph.inv_gm1 + (0.5f * rr) * vsq<DIM>(wr);
  UL[0] = rl;
  UR[0] = rr;
  FL[0] = ul * rl;
  FR[0] = ur * rr;
#pragma unroll
  for (int i = 1; i <= DIM; ++i) {
    UL[i] = rl * wl[i];
    UR[i] = rr * wr[i];
    FL[i] = ul * UL[i] + (i == AX ? pl : 0.f);
    FR[i] = ur * UR[i] + (i == AX ? pr : 0.f);
  }
  UL[P] = El;
  UR[P] = Er;
  FL[P] = ul * (pl + El);
  FR[P] = ur * (pr + Er);
  const float sql = sqrtf(rl), sqr = sqrtf(rr);
  const float ssum = sql + sqr;
  const float u_hat = (sql * ul + sqr * ur) / ssum;
  const float du = ur - ul;
  const float c_hat2 = ((cl * cl) * sql + (cr * cr) * sqr) / ssum + (0.5f * ((sql * sqr) / (ssum * ssum))) * (du * du);
  const float c_hat = sqrtf(c_hat2);
  const float SL = fminf(ul - cl, u_hat - c_hat);
  const float SR = fmaxf(ur + cr, u_hat + c_hat);
  const float s_star = (((pr - pl) + (rl * ul) * (SL - ul)) - (rr * ur) * (SR - ur)) / (rl * (SL - ul) - rr * (SR - ur));
  const float fl = (SL - ul) / (SL - s_star);
  const float fr = (SR - ur) / (SR - s_star);
  const float as = fabsf(s_star);
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    float usl = UL[v], usr = UR[v];
    if (v == AX) {
      usl = rl * s_star;
      usr = rr * s_star;
    }
    if (v == P) {
      usl = usl + (s_star - ul) * (rl * s_star + pl / (SL - ul));
      usr = usr + (s_star - ur) * (rr * s_star + pr / (SR - ur));
    }
    usl *= fl;
    usr *= fr;
    const float fs = 0.5f * (FL[v] + FR[v]) + 0.5f * ((SL * (usl - UL[v]) + as * (usl - usr)) + SR * (usr - UR[v]));
    float out = SL >= 0.f ? FL[v] : fs;
    f[v] = SR <= 0.f ? FR[v] : out;
  }
}

template <int DIM, int AX, int RS>
__device__ __forceinline__ void riemann(const float* wl, const float* wr, float* f, const Phys& ph) {
  if (RS == AMX_HLLC)
    flux_hllc<DIM, AX>(wl, wr, f, ph);
  else
    flux_hll<DIM, AX>(wl, wr, f, ph);
}

enum { SRC_PLAIN = 0, SRC_MAGFIX = 1 };

// Primitive gas state at one cell. SRC_PLAIN reads the NV gas rows. SRC_MAGFIX
// reads (rho, v0, p) of the gas state entering the magnetic sub-step plus the
// converged velocity vn (slots 5..7) and applies the sub-step's write-back:
// v <- vn and the energy fix with gamma = 5/3 (_magnetic_field_update.py:208-240).
template <int DIM, int MODE>
__device__ __forceinline__ void load_w(const float* const* s, long c, float* w) {
  constexpr int NV = DIM + 2, P = DIM + 1;
  if (MODE == SRC_PLAIN) {
#pragma unroll
    for (int v = 0; v < NV; ++v) w[v] = s[v][c];
  } else {
    const float rho = s[0][c], p = s[P][c];
    const float v0x = s[1][c], v0y = s[2][c], v0z = DIM == 3 ? s[3][c] : 0.f;
    const float vnx = s[5][c], vny = s[6][c], vnz = s[7][c];
    const float n0 = sqrtf((v0x * v0x + v0y * v0y) + v0z * v0z);
    const float nn = sqrtf((vnx * vnx + vny * vny) + vnz * vnz);
    const float hr = 0.5f * rho;
    const float E = p / G53M1 + hr * (n0 * n0);
    const float E2 = (E - hr * (n0 * n0)) + hr * (nn * nn);
    const float e = E2 / rho - 0.5f * (nn * nn);
    w[0] = rho;
    w[1] = vnx;
    w[2] = vny;
    if (DIM == 3) w[3] = vnz;
    w[P] = (G53M1 * rho) * e;
  }
}

struct StageArgs {
  const float* src[8];   // stencil input (see load_w)
  const float* w0[8];    // Heun: stage-0 state at the cell centre
  float* out[5];         // output gas rows
  const float* bsrc[3];  // optional copy bsrc -> bdst (final B of the magnetic sub-step)
  float* bdst[3];
  const float* dtp;      // device step size
  unsigned* cflmax;      // optional CFL maxima of the output
  int nx, na, nb;        // 3D: (nx, ny, nz); 2D: (1, nx, ny)
  int ntb, nta, ncx, lx;
  float dt_scale;        // 0.5: MHD half step dt/2; 1: hydro step
  Phys ph;
};

__device__ __forceinline__ int swizzle_block(int bid, int nblk) {
  // consecutive logical tiles on the same XCD (8 XCDs, round-robin dispatch)
  if ((nblk & 7) == 0) return (bid & 7) * (nblk >> 3) + (bid >> 3);
  return bid;
}

__device__ __forceinline__ unsigned canon_bits(float m) {
  return m != m ? 0x7fc00000u : __float_as_uint(m);
}

__device__ __forceinline__ unsigned wave_max_u(unsigned v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const unsigned w = (unsigned)__shfl_xor((int)v, o, 64);
    v = w > v ? w : v;
  }
  return v;
}

// One unsplit forward-Euler stage (evolve_state.py:338-414), optionally with
// the RK2 Heun average (evolve_state.py:456-469).
template <int DIM, int TA, int SRC, bool HEUN, int W0M, int RS>
__global__ __launch_bounds__(TA * TB) void k_stage(const StageArgs a) {
  constexpr int NV = DIM + 2, P = DIM + 1, NT = TA * TB, WW = TB + 4;
  constexpr int AXA = DIM == 3 ? 2 : 1;  // velocity row of the plane axis A
  constexpr int AXB = DIM == 3 ? 3 : 2;  // velocity row of the plane axis B (contiguous)
  __shared__ float sW[NV][(TA + 4) * WW];
  __shared__ float sDA[NV][(TA + 2) * TB];
  __shared__ float sDB[NV][TA * (TB + 2)];
  __shared__ float sFA[NV][(TA + 1) * TB];
  __shared__ float sFB[NV][TA * (TB + 1)];

  const Phys ph = a.ph;
  const int tid = threadIdx.x, ta = tid / TB, tb = tid % TB;
  const int bid = swizzle_block(blockIdx.x, gridDim.x);
  const int itb = bid % a.ntb;
  const int rest = bid / a.ntb;
  const int ita = rest % a.nta;
  const int icx = rest / a.nta;
  const int a0 = ita * TA, b0 = itb * TB;
  const long sa = a.nb, sx = (long)a.na * a.nb;
  const int ga = wrapi(a0 + ta, a.na), gb = wrapi(b0 + tb, a.nb);
  const bool active = (a0 + ta < a.na) && (b0 + tb < a.nb);
  const float dt = a.dt_scale * (*a.dtp);
  const float cdt = ph.inv_dx;
  int x0 = 0, x1 = 1;
  if (DIM == 3) {
    x0 = icx * a.lx;
    x1 = min(x0 + a.lx, a.nx);
  }

  float q0[NV], q1[NV], d0[NV], Fp[NV];
  if (DIM == 3) {
    float qm2[NV], qm1[NV], dm1[NV], wl[NV], wr[NV];
    const long col = (long)ga * sa + gb;
    load_w<DIM, SRC>(a.src, (long)wrapi(x0 - 2, a.nx) * sx + col, qm2);
    load_w<DIM, SRC>(a.src, (long)wrapi(x0 - 1, a.nx) * sx + col, qm1);
    load_w<DIM, SRC>(a.src, (long)x0 * sx + col, q0);
    load_w<DIM, SRC>(a.src, (long)wrapi(x0 + 1, a.nx) * sx + col, q1);
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      dm1[v] = half_slope(qm2[v], qm1[v], q0[v]);
      d0[v] = half_slope(qm1[v], q0[v], q1[v]);
      wl[v] = qm1[v] + dm1[v];
      wr[v] = q0[v] - d0[v];
    }
    riemann<DIM, 1, RS>(wl, wr, Fp, ph);
  } else {
    load_w<DIM, SRC>(a.src, (long)ga * sa + gb, q0);
  }

  float mx[3] = {0.f, 0.f, 0.f};
  for (int x = x0; x < x1; ++x) {
    const long plane = DIM == 3 ? (long)x * sx : 0;
    float Fn[NV];
    if (DIM == 3) {
      float q2[NV], d1[NV], wl[NV], wr[NV];
      load_w<DIM, SRC>(a.src, (long)wrapi(x + 2, a.nx) * sx + (long)ga * sa + gb, q2);
#pragma unroll
      for (int v = 0; v < NV; ++v) {
        d1[v] = half_slope(q0[v], q1[v], q2[v]);
        wl[v] = q0[v] + d0[v];
        wr[v] = q1[v] - d1[v];
      }
      riemann<DIM, 1, RS>(wl, wr, Fn, ph);
#pragma unroll
      for (int v = 0; v < NV; ++v) d0[v] = d1[v];
      // plane fill (own cell) and halo
#pragma unroll
      for (int v = 0; v < NV; ++v) sW[v][(ta + 2) * WW + tb + 2] = q0[v];
      constexpr int NH = 4 * TB + 4 * TA;
      for (int h = tid; h < NH; h += NT) {
        int r, c;
        if (h < 4 * TB) {
          const int q = h / TB;
          c = h % TB;
          r = q < 2 ? q - 2 : TA + q - 2;
        } else {
          const int j = h - 4 * TB;
          r = j >> 2;
          const int q = j & 3;
          c = q < 2 ? q - 2 : TB + q - 2;
        }
        float w[NV];
        load_w<DIM, SRC>(a.src, plane + (long)wrapi(a0 + r, a.na) * sa + wrapi(b0 + c, a.nb), w);
#pragma unroll
        for (int v = 0; v < NV; ++v) sW[v][(r + 2) * WW + c + 2] = w[v];
      }
      __syncthreads();
      // ---- limited slopes along A and B
      {
        constexpr int NDA = (TA + 2) * TB, ND = NDA + TA * (TB + 2);
        for (int i = tid; i < ND; i += NT) {
          if (i < NDA) {
            const int r = i / TB - 1, c = i % TB;
            const int base = (r + 2) * WW + c + 2;
#pragma unroll
            for (int v = 0; v < NV; ++v)
              sDA[v][(r + 1) * TB + c] = half_slope(sW[v][base - WW], sW[v][base], sW[v][base + WW]);
          } else {
            const int j = i - NDA;
            const int r = j / (TB + 2), c = j % (TB + 2) - 1;
            const int base = (r + 2) * WW + c + 2;
#pragma unroll
            for (int v = 0; v < NV; ++v)
              sDB[v][r * (TB + 2) + c + 1] = half_slope(sW[v][base - 1], sW[v][base], sW[v][base + 1]);
          }
        }
      }
      __syncthreads();
      // ---- interface fluxes along A (r-1/2) and B (c-1/2)
      {
        constexpr int NFA = (TA + 1) * TB, NF = NFA + TA * (TB + 1);
        for (int i = tid; i < NF; i += NT) {
          float wl2[NV], wr2[NV], f[NV];
          if (i < NFA) {
            const int r = i / TB, c = i % TB;
#pragma unroll
            for (int v = 0; v < NV; ++v) {
              wl2[v] = sW[v][(r + 1) * WW + c + 2] + sDA[v][r * TB + c];
              wr2[v] = sW[v][(r + 2) * WW + c + 2] - sDA[v][(r + 1) * TB + c];
            }
            riemann<DIM, AXA, RS>(wl2, wr2, f, ph);
#pragma unroll
            for (int v = 0; v < NV; ++v) sFA[v][r * TB + c] = f[v];
          } else {
            const int j = i - NFA;
            const int r = j / (TB + 1), c = j % (TB + 1);
#pragma unroll
            for (int v = 0; v < NV; ++v) {
              wl2[v] = sW[v][(r + 2) * WW + c + 1] + sDB[v][r * (TB + 2) + c];
              wr2[v] = sW[v][(r + 2) * WW + c + 2] - sDB[v][r * (TB + 2) + c + 1];
            }
            riemann<DIM, AXB, RS>(wl2, wr2, f, ph);
#pragma unroll
            for (int v = 0; v < NV; ++v) sFB[v][r * (TB + 1) + c] = f[v];
          }
        }
      }
      __syncthreads();
      // ---- update of the own cell
      if (active) {
        const long cell = plane + (long)ga * sa + gb;
        float U[NV], W1[NV];
        cons<DIM>(q0, U, ph);
#pragma unroll
        for (int v = 0; v < NV; ++v) U[v] = U[v] + (cdt * (Fp[v] - Fn[v])) * dt;
#pragma unroll
        for (int v = 0; v < NV; ++v) U[v] = U[v] + (cdt * (sFA[v][ta * TB + tb] - sFA[v][(ta + 1) * TB + tb])) * dt;
#pragma unroll
        for (int v = 0; v < NV; ++v)
          U[v] = U[v] + (cdt * (sFB[v][ta * (TB + 1) + tb] - sFB[v][ta * (TB + 1) + tb + 1])) * dt;
        prim<DIM>(U, W1, ph);
        if (HEUN) {
          float W0[NV], U0[NV], U2[NV];
          load_w<DIM, W0M>(a.w0, cell, W0);
          cons<DIM>(W0, U0, ph);
          cons<DIM>(W1, U2, ph);
#pragma unroll
          for (int v = 0; v < NV; ++v) U[v] = 0.5f * (U0[v] + U2[v]);
          prim<DIM>(U, W1, ph);
        }
#pragma unroll
        for (int v = 0; v < NV; ++v) a.out[v][cell] = W1[v];
        if (a.cflmax) {
          const float cs = sqrtf(ph.gamma * W1[P] / W1[0]);
#pragma unroll
          for (int k = 0; k < DIM; ++k) {
            const float m = fabsf(W1[1 + k]) + cs;
            mx[k] = (m != m || m > mx[k]) ? m : mx[k];
          }
        }
        if (a.bsrc[0]) {
#pragma unroll
          for (int k = 0; k < 3; ++k) a.bdst[k][cell] = a.bsrc[k][cell];
        }
      }
#pragma unroll
      for (int v = 0; v < NV; ++v) {
        q0[v] = q1[v];
        q1[v] = q2[v];
        Fp[v] = Fn[v];
      }
    } else {
      // ---------------- 2D: a single plane (axes A = x, B = y)
#pragma unroll
      for (int v = 0; v < NV; ++v) sW[v][(ta + 2) * WW + tb + 2] = q0[v];
      constexpr int NH = 4 * TB + 4 * TA;
      for (int h = tid; h < NH; h += NT) {
        int r, c;
        if (h < 4 * TB) {
          const int q = h / TB;
          c = h % TB;
          r = q < 2 ? q - 2 : TA + q - 2;
        } else {
          const int j = h - 4 * TB;
          r = j >> 2;
          const int q = j & 3;
          c = q < 2 ? q - 2 : TB + q - 2;
        }
        float w[NV];
        load_w<DIM, SRC>(a.src, (long)wrapi(a0 + r, a.na) * sa + wrapi(b0 + c, a.nb), w);
#pragma unroll
        for (int v = 0; v < NV; ++v) sW[v][(r + 2) * WW + c + 2] = w[v];
      }
      __syncthreads();
      {
        constexpr int NDA = (TA + 2) * TB, ND = NDA + TA * (TB + 2);
        for (int i = tid; i < ND; i += NT) {
          if (i < NDA) {
            const int r = i / TB - 1, c = i % TB;
            const int base = (r + 2) * WW + c + 2;
#pragma unroll
            for (int v = 0; v < NV; ++v)
              sDA[v][(r + 1) * TB + c] = half_slope(sW[v][base - WW], sW[v][base], sW[v][base + WW]);
          } else {
            const int j = i - NDA;
            const int r = j / (TB + 2), c = j % (TB + 2) - 1;
            const int base = (r + 2) * WW + c + 2;
#pragma unroll
            for (int v = 0; v < NV; ++v)
              sDB[v][r * (TB + 2) + c + 1] = half_slope(sW[v][base - 1], sW[v][base], sW[v][base + 1]);
          }
        }
      }
      __syncthreads();
      {
        constexpr int NFA = (TA + 1) * TB, NF = NFA + TA * (TB + 1);
        for (int i = tid; i < NF; i += NT) {
          float wl2[NV], wr2[NV], f[NV];
          if (i < NFA) {
            const int r = i / TB, c = i % TB;
#pragma unroll
            for (int v = 0; v < NV; ++v) {
              wl2[v] = sW[v][(r + 1) * WW + c + 2] + sDA[v][r * TB + c];
              wr2[v] = sW[v][(r + 2) * WW + c + 2] - sDA[v][(r + 1) * TB + c];
            }
            riemann<DIM, AXA, RS>(wl2, wr2, f, ph);
#pragma unroll
            for (int v = 0; v < NV; ++v) sFA[v][r * TB + c] = f[v];
          } else {
            const int j = i - NFA;
            const int r = j / (TB + 1), c = j % (TB + 1);
#pragma unroll
            for (int v = 0; v < NV; ++v) {
              wl2[v] = sW[v][(r + 2) * WW + c + 1] + sDB[v][r * (TB + 2) + c];
              wr2[v] = sW[v][(r + 2) * WW + c + 2] - sDB[v][r * (TB + 2) + c + 1];
            }
            riemann<DIM, AXB, RS>(wl2, wr2, f, ph);
#pragma unroll
            for (int v = 0; v < NV; ++v) sFB[v][r * (TB + 1) + c] = f[v];
          }
        }
      }
      __syncthreads();
      if (active) {
        const long cell = (long)ga * sa + gb;
        float U[NV], W1[NV];
        cons<DIM>(q0, U, ph);
#pragma unroll
        for (int v = 0; v < NV; ++v) U[v] = U[v] + (cdt * (sFA[v][ta * TB + tb] - sFA[v][(ta + 1) * TB + tb])) * dt;
#pragma unroll
        for (int v = 0; v < NV; ++v)
          U[v] = U[v] + (cdt * (sFB[v][ta * (TB + 1) + tb] - sFB[v][ta * (TB + 1) + tb + 1])) * dt;
        prim<DIM>(U, W1, ph);
        if (HEUN) {
          float W0[NV], U0[NV], U2[NV];
          load_w<DIM, W0M>(a.w0, cell, W0);
          cons<DIM>(W0, U0, ph);
          cons<DIM>(W1, U2, ph);
#pragma unroll
          for (int v = 0; v < NV; ++v) U[v] = 0.5f * (U0[v] + U2[v]);
          prim<DIM>(U, W1, ph);
        }
#pragma unroll
        for (int v = 0; v < NV; ++v) a.out[v][cell] = W1[v];
        if (a.cflmax) {
          const float cs = sqrtf(ph.gamma * W1[P] / W1[0]);
#pragma unroll
          for (int k = 0; k < DIM; ++k) {
            const float m = fabsf(W1[1 + k]) + cs;
            mx[k] = (m != m || m > mx[k]) ? m : mx[k];
          }
        }
        if (a.bsrc[0]) {
#pragma unroll
          for (int k = 0; k < 3; ++k) a.bdst[k][cell] = a.bsrc[k][cell];
        }
      }
    }
  }
  if (a.cflmax) {
#pragma unroll
    for (int k = 0; k < DIM; ++k) {
      const unsigned m = wave_max_u(canon_bits(mx[k]));
      if ((tid & 63) == 0 && m) atomicMax(&a.cflmax[k], m);
    }
  }
}

struct MagArgs {
  const float* B0[3];
  const float* v0[3];  // v0[2] == nullptr in 2D (vz = 0)
  const float* rho;
  const float* Bk[3];
  const float* vk[3];  // vk[2] == nullptr: vz = 0
  float* Bo[3];
  float* vo[3];
  const float* dtp;
  Ctl* ctl;
  int k;         // 0: initial iterate (Bk = B0, vk = v0); k >= 1: while body k
  int max_iter;
  float tol;
  int euler;     // IMPLICIT_EULER: average = iterate
  int nx, na, nb, ntb, nta, ncx, lx;
  float inv_dx;
};

// Raw iterate and fixed-point inputs of one cell.
struct MagRaw {
  float Bk[3], vk[3], B0[3], v0[3];
};

__device__ __forceinline__ void mag_load(const MagArgs& a, long c, MagRaw& m) {
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    m.B0[i] = a.B0[i][c];
    m.Bk[i] = a.Bk[i][c];
  }
  m.v0[0] = a.v0[0][c];
  m.v0[1] = a.v0[1][c];
  m.v0[2] = a.v0[2] ? a.v0[2][c] : 0.f;
  m.vk[0] = a.vk[0][c];
  m.vk[1] = a.vk[1][c];
  m.vk[2] = a.vk[2] ? a.vk[2][c] : 0.f;
}

// (avgB, E = avgB x avgv) of one cell (_magnetic_field_update.py:84,149-154; _vector_maths.py:5-10)
__device__ __forceinline__ void mag_derive(const MagRaw& m, int euler, float* d) {
  float b[3], v[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    b[i] = euler ? m.Bk[i] : (m.Bk[i] + m.B0[i]) * 0.5f;
    v[i] = euler ? m.vk[i] : (m.vk[i] + m.v0[i]) * 0.5f;
  }
  d[0] = b[0];
  d[1] = b[1];
  d[2] = b[2];
  d[3] = b[1] * v[2] - b[2] * v[1];
  d[4] = b[2] * v[0] - b[0] * v[2];
  d[5] = b[0] * v[1] - b[1] * v[0];
}

// central derivative 0.5 (f+ - f-) / dx (_vector_maths.py:58-100)
__device__ __forceinline__ float cdiff(float fp, float fm, float inv_dx) {
  return (0.5f * (fp - fm)) * inv_dx;
}

// One fixed-point iterate of the implicit-midpoint magnetic sub-step:
// B' = B0 - dt curl(avgB x avgv), v' = v0 - dt (avgB x curl avgB) / rho,
// with the residual max(|Bk - B'|, |vk - v'|) reduced into ctl->res[k]
// (_magnetic_field_update.py:112-165).
template <int DIM, int TA>
__global__ __launch_bounds__(TA * TB) void k_mag(const MagArgs a) {
  constexpr int NT = TA * TB, WW = TB + 2, PL = (TA + 2) * WW;
  if (a.k > 0) {
    const float r = __uint_as_float(a.ctl->res[a.k - 1]);
    if (!(r > a.tol) || a.k - 1 >= a.max_iter) return;
  }
  __shared__ float sM[2][6][PL];
  const int tid = threadIdx.x, ta = tid / TB, tb = tid % TB;
  if (blockIdx.x == 0 && tid == 0) {
    a.ctl->iters = a.k;
    a.ctl->res[a.k + 1] = 0u;
  }
  const int bid = swizzle_block(blockIdx.x, gridDim.x);
  const int itb = bid % a.ntb;
  const int rest = bid / a.ntb;
  const int ita = rest % a.nta;
  const int icx = rest / a.nta;
  const int a0 = ita * TA, b0 = itb * TB;
  const long sa = a.nb, sx = (long)a.na * a.nb;
  const int ga = wrapi(a0 + ta, a.na), gb = wrapi(b0 + tb, a.nb);
  const bool active = (a0 + ta < a.na) && (b0 + tb < a.nb);
  const float dt = *a.dtp;
  const float idx = a.inv_dx;
  const long col = (long)ga * sa + gb;
  int x0 = 0, x1 = 1;
  if (DIM == 3) {
    x0 = icx * a.lx;
    x1 = min(x0 + a.lx, a.nx);
  }
  MagRaw rc, rn;
  float dm[6], dc[6], dp[6];
  if (DIM == 3) {
    MagRaw t;
    mag_load(a, (long)wrapi(x0 - 1, a.nx) * sx + col, t);
    mag_derive(t, a.euler, dm);
    mag_load(a, (long)x0 * sx + col, rc);
    mag_derive(rc, a.euler, dc);
    mag_load(a, (long)wrapi(x0 + 1, a.nx) * sx + col, rn);
    mag_derive(rn, a.euler, dp);
  } else {
    mag_load(a, col, rc);
    mag_derive(rc, a.euler, dc);
  }
  float res = 0.f;
  int buf = 0;
  for (int x = x0; x < x1; ++x) {
    const long plane = DIM == 3 ? (long)x * sx : 0;
    // plane fill: own cell and the halo ring (rows -1, TA; columns -1, TB)
#pragma unroll
    for (int i = 0; i < 6; ++i) sM[buf][i][(ta + 1) * WW + tb + 1] = dc[i];
    constexpr int NH = 2 * TB + 2 * TA;
    for (int h = tid; h < NH; h += NT) {
      int r, c;
      if (h < 2 * TB) {
        r = h < TB ? -1 : TA;
        c = h % TB;
      } else {
        const int j = h - 2 * TB;
        r = j >> 1;
        c = (j & 1) ? TB : -1;
      }
      MagRaw t;
      float d[6];
      mag_load(a, plane + (long)wrapi(a0 + r, a.na) * sa + wrapi(b0 + c, a.nb), t);
      mag_derive(t, a.euler, d);
#pragma unroll
      for (int i = 0; i < 6; ++i) sM[buf][i][(r + 1) * WW + c + 1] = d[i];
    }
    MagRaw rnn;
    float dpp[6];
    if (DIM == 3) {
      // prefetch the next column cell
      mag_load(a, (long)wrapi(x + 2, a.nx) * sx + col, rnn);
    }
    __syncthreads();
    if (active) {
      const long cell = plane + col;
      const int o = (ta + 1) * WW + tb + 1;
      float (*S)[PL] = sM[buf];
      // neighbour values: A +/- 1 rows, B +/- 1 columns
      float cE[3], cB[3];
      if (DIM == 3) {
        cE[0] = cdiff(S[5][o + WW], S[5][o - WW], idx) - cdiff(S[4][o + 1], S[4][o - 1], idx);
        cE[1] = cdiff(S[3][o + 1], S[3][o - 1], idx) - cdiff(dp[5], dm[5], idx);
        cE[2] = cdiff(dp[4], dm[4], idx) - cdiff(S[3][o + WW], S[3][o - WW], idx);
        cB[0] = cdiff(S[2][o + WW], S[2][o - WW], idx) - cdiff(S[1][o + 1], S[1][o - 1], idx);
        cB[1] = cdiff(S[0][o + 1], S[0][o - 1], idx) - cdiff(dp[2], dm[2], idx);
        cB[2] = cdiff(dp[1], dm[1], idx) - cdiff(S[0][o + WW], S[0][o - WW], idx);
      } else {
        cE[0] = cdiff(S[5][o + 1], S[5][o - 1], idx);
        cE[1] = -cdiff(S[5][o + WW], S[5][o - WW], idx);
        cE[2] = cdiff(S[4][o + WW], S[4][o - WW], idx) - cdiff(S[3][o + 1], S[3][o - 1], idx);
        cB[0] = cdiff(S[2][o + 1], S[2][o - 1], idx);
        cB[1] = -cdiff(S[2][o + WW], S[2][o - WW], idx);
        cB[2] = cdiff(S[1][o + WW], S[1][o - WW], idx) - cdiff(S[0][o + 1], S[0][o - 1], idx);
      }
      const float ri = 1.0f / a.rho[cell];
      const float p2[3] = {(dc[1] * cB[2] - dc[2] * cB[1]) * ri, (dc[2] * cB[0] - dc[0] * cB[2]) * ri,
                           (dc[0] * cB[1] - dc[1] * cB[0]) * ri};
      float nb = 0.f, nv = 0.f;
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        const float bn = rc.B0[i] - dt * cE[i];
        const float vn = rc.v0[i] - dt * p2[i];
        a.Bo[i][cell] = bn;
        a.vo[i][cell] = vn;
        const float db = rc.Bk[i] - bn, dv = rc.vk[i] - vn;
        nb += db * db;
        nv += dv * dv;
      }
      const float r = fmaxf(sqrtf(nb), sqrtf(nv));
      res = (r != r || r > res) ? r : res;
    }
    if (DIM == 3) {
      mag_derive(rnn, a.euler, dpp);
#pragma unroll
      for (int i = 0; i < 6; ++i) {
        dm[i] = dc[i];
        dc[i] = dp[i];
        dp[i] = dpp[i];
      }
      rc = rn;
      rn = rnn;
    }
    buf ^= 1;
  }
  const unsigned m = wave_max_u(canon_bits(res));
  if ((tid & 63) == 0 && m) atomicMax(&a.ctl->res[a.k], m);
}

// Unsplit CFL maxima max(|u_a| + c) per axis (_timestep_estimator.py:118-126).
template <int DIM>
__global__ void k_cfl(const float* rho, const float* v0, const float* v1, const float* v2, const float* p,
                      long n, float gamma, unsigned* cflmax) {
  float mx[3] = {0.f, 0.f, 0.f};
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    const float cs = sqrtf(gamma * p[i] / rho[i]);
    const float u[3] = {v0[i], v1[i], DIM == 3 ? v2[i] : 0.f};
#pragma unroll
    for (int k = 0; k < DIM; ++k) {
      const float m = fabsf(u[k]) + cs;
      mx[k] = (m != m || m > mx[k]) ? m : mx[k];
    }
  }
#pragma unroll
  for (int k = 0; k < DIM; ++k) {
    const unsigned m = wave_max_u(canon_bits(mx[k]));
    if ((threadIdx.x & 63) == 0 && m) atomicMax(&cflmax[k], m);
  }
}

struct StepParams {
  float t_end;
  int fixed;
  int num_timesteps;
  int exact_end;
  float C;
  float dx;
  int dim;
  int use_cfl;  // 1: dt from the CFL maxima; 0: dt already in ctl->dt (host value)
};

// dt = C / sum_a(max_a / dx) (_timestep_estimator.py:128), or t_end/num_timesteps
// (time_integration.py:667-668); clamp to t_end - t (time_integration.py:677-678).
__device__ __forceinline__ float cfl_dt_from(const Ctl* c, float C, float dx, int dim) {
  float s = __uint_as_float(c->cfl[0]) / dx + __uint_as_float(c->cfl[1]) / dx;
  if (dim == 3) s = s + __uint_as_float(c->cfl[2]) / dx;
  return (C * 1.0f) / s;
}

__global__ void k_step_begin(Ctl* c, StepParams p) {
  float dt;
  if (p.fixed)
    dt = p.t_end / (float)p.num_timesteps;
  else if (p.use_cfl)
    dt = cfl_dt_from(c, p.C, p.dx, p.dim);
  else
    dt = c->dt;
  if (p.exact_end) {
    const float rem = p.t_end - c->t;
    dt = rem < dt ? rem : dt;
  }
  c->dt = dt;
  c->cfl[0] = c->cfl[1] = c->cfl[2] = c->cfl[3] = 0u;
  c->res[0] = 0u;
  c->iters = 0;
}

__global__ void k_set_dt(Ctl* c, const float* dtp) {
  c->dt = *dtp;
  c->cfl[0] = c->cfl[1] = c->cfl[2] = c->cfl[3] = 0u;
  c->res[0] = 0u;
  c->iters = 0;
}

__global__ void k_step_end(Ctl* c) {
  c->t = c->t + c->dt;
  c->steps += 1;
}

__global__ void k_cfl_out(const Ctl* c, float C, float dx, int dim, float* out) {
  *out = cfl_dt_from(c, C, dx, dim);
}

__global__ void k_reset(Ctl* c) {
  c->dt = 0.f;
  c->t = 0.f;
  c->iters = 0;
  c->steps = 0;
  for (int i = 0; i < 4; ++i) c->cfl[i] = 0u;
  c->res[0] = 0u;
}

}  // namespace amx

// ============================================================================
// C ABI
// ============================================================================

using namespace amx;

namespace {

thread_local char g_err[512] = "";

amx_status fail(amx_status s, const char* fmt, const char* what = "") {
  snprintf(g_err, sizeof(g_err), fmt, what);
  return s;
}

#define HIPCHK(expr)                                                              \
  do {                                                                            \
    hipError_t _e = (expr);                                                       \
    if (_e != hipSuccess) return fail(AMX_ERR_HIP, "HIP error: %s", hipGetErrorString(_e)); \
  } while (0)

constexpr int TA_DEF = 8;

}  // namespace

struct amx_fv_ctx {
  amx_fv_config cfg;
  int NV;
  long ncell;
  Phys ph;
  float dx;
  int na, nb, nxm;  // kernel axes: 3D (ny, nz, nx); 2D (nx, ny, 1)
  int ntb, nta, ncx, lx;
  float* ws;        // G1 (NV fields) | P0 (6 fields) | P1 (6 fields)
  Ctl* ctl;
  Ctl* hctl;        // pinned host mirror
  int last_k;
  // optional HIP-event timing of kernel launches
  int timing;
  struct Rec {
    int e0, e1, cls, k;
  };
  std::vector<hipEvent_t> ev;
  std::vector<Rec> pending;
  int ev_used;
  double ms[2];
  long launches[2];
};

namespace {

enum { CLS_STAGE = 0, CLS_MAG = 1 };

int time_begin(amx_fv_ctx* c, hipStream_t st) {
  if (!c->timing) return -1;
  if (c->ev_used + 2 > (int)c->ev.size()) {
    hipEvent_t e;
    for (int i = 0; i < 2; ++i) {
      if (hipEventCreate(&e) != hipSuccess) return -1;
      c->ev.push_back(e);
    }
  }
  const int i = c->ev_used;
  c->ev_used += 2;
  (void)hipEventRecord(c->ev[i], st);
  return i;
}

void time_end(amx_fv_ctx* c, hipStream_t st, int i, int cls, int k) {
  if (i < 0) return;
  (void)hipEventRecord(c->ev[i + 1], st);
  c->pending.push_back({i, i + 1, cls, k});
}

// Call only after the stream is synchronised: folds completed records into the
// totals; fixed-point launches with k > K exited early and are dropped.
void time_resolve(amx_fv_ctx* c, int K) {
  for (const auto& r : c->pending) {
    if (r.cls == CLS_MAG && r.k > K) continue;
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, c->ev[r.e0], c->ev[r.e1]) == hipSuccess) {
      c->ms[r.cls] += ms;
      c->launches[r.cls] += 1;
    }
  }
  c->pending.clear();
  c->ev_used = 0;
}

}  // namespace

static int env_int(const char* name, int dflt) {
  const char* s = getenv(name);
  return s ? atoi(s) : dflt;
}

static amx_status validate(const amx_fv_config* c) {
  if (!c) return fail(AMX_ERR_INVALID, "null config%s");
  if (c->dim != 2 && c->dim != 3) return fail(AMX_ERR_UNSUPPORTED, "dimensionality must be 2 or 3%s");
  if (c->nx < 2 || c->ny < 2 || (c->dim == 3 && c->nz < 2)) return fail(AMX_ERR_INVALID, "grid too small%s");
  if (c->dim == 2 && c->nz != 1) return fail(AMX_ERR_INVALID, "nz must be 1 in 2D%s");
  if (c->riemann != AMX_HLL && c->riemann != AMX_HLLC) return fail(AMX_ERR_UNSUPPORTED, "riemann solver must be HLL or HLLC%s");
  if (c->magnetic_integrator != AMX_IMPLICIT_MIDPOINT && c->magnetic_integrator != AMX_IMPLICIT_EULER)
    return fail(AMX_ERR_UNSUPPORTED, "unknown magnetic integrator%s");
  if (c->max_fixed_point_iterations < 0 || c->max_fixed_point_iterations > MAXRES - 4)
    return fail(AMX_ERR_INVALID, "max_fixed_point_iterations out of range%s");
  if (!(c->grid_spacing > 0)) return fail(AMX_ERR_INVALID, "grid_spacing must be positive%s");
  return AMX_OK;
}

extern "C" amx_status amx_fv_workspace_bytes(const amx_fv_config* cfg, size_t* bytes) {
  amx_status s = validate(cfg);
  if (s != AMX_OK) return s;
  if (!bytes) return fail(AMX_ERR_INVALID, "null output%s");
  const long n = (long)cfg->nx * cfg->ny * (cfg->dim == 3 ? cfg->nz : 1);
  const int nv = cfg->dim + 2;
  *bytes = sizeof(float) * (size_t)n * (nv + (cfg->mhd ? 12 : 0)) + sizeof(Ctl);
  return AMX_OK;
}

extern "C" amx_status amx_fv_create(const amx_fv_config* cfg, amx_fv_ctx** out) {
  amx_status s = validate(cfg);
  if (s != AMX_OK) return s;
  if (!out) return fail(AMX_ERR_INVALID, "null output%s");
  amx_fv_ctx* c = new (std::nothrow) amx_fv_ctx();
  if (!c) return fail(AMX_ERR_NOMEM, "host allocation failed%s");
  c->cfg = *cfg;
  c->NV = cfg->dim + 2;
  c->ncell = (long)cfg->nx * cfg->ny * (cfg->dim == 3 ? cfg->nz : 1);
  c->ph.gamma = cfg->gamma;
  c->ph.gm1 = cfg->gamma - 1.0f;
  c->ph.inv_gm1 = 1.0f / c->ph.gm1;
  c->ph.inv_dx = (float)(1.0 / cfg->grid_spacing);
  c->dx = (float)cfg->grid_spacing;
  if (cfg->dim == 3) {
    c->nxm = cfg->nx;
    c->na = cfg->ny;
    c->nb = cfg->nz;
  } else {
    c->nxm = 1;
    c->na = cfg->nx;
    c->nb = cfg->ny;
  }
  c->ntb = (c->nb + TB - 1) / TB;
  c->nta = (c->na + TA_DEF - 1) / TA_DEF;
  if (cfg->dim == 3) {
    int lx = env_int("AMX_LX", 0);
    if (lx <= 0) {
      // aim for >= ~2048 blocks (8 per CU) while keeping chunks >= 8 planes
      const long tiles = (long)c->ntb * c->nta;
      lx = (int)((tiles * cfg->nx) / 2048);
      if (lx < 8) lx = 8;
      if (lx > cfg->nx) lx = cfg->nx;
    }
    c->lx = lx;
    c->ncx = (cfg->nx + lx - 1) / lx;
  } else {
    c->lx = 1;
    c->ncx = 1;
  }
  size_t bytes = sizeof(float) * (size_t)c->ncell * (c->NV + (cfg->mhd ? 12 : 0));
  if (hipMalloc((void**)&c->ws, bytes) != hipSuccess) {
    delete c;
    return fail(AMX_ERR_NOMEM, "device workspace allocation failed%s");
  }
  if (hipMalloc((void**)&c->ctl, sizeof(Ctl)) != hipSuccess) {
    (void)hipFree(c->ws);
    delete c;
    return fail(AMX_ERR_NOMEM, "device control allocation failed%s");
  }
  if (hipHostMalloc((void**)&c->hctl, sizeof(Ctl), hipHostMallocDefault) != hipSuccess) {
    (void)hipFree(c->ws);
    (void)hipFree(c->ctl);
    delete c;
    return fail(AMX_ERR_NOMEM, "pinned host allocation failed%s");
  }
  hipLaunchKernelGGL(k_reset, dim3(1), dim3(1), 0, 0, c->ctl);
  if (hipDeviceSynchronize() != hipSuccess) return fail(AMX_ERR_HIP, "init sync failed%s");
  c->last_k = 4;
  *out = c;
  return AMX_OK;
}

extern "C" amx_status amx_fv_destroy(amx_fv_ctx* c) {
  if (!c) return AMX_OK;
  (void)hipDeviceSynchronize();
  (void)hipFree(c->ws);
  (void)hipFree(c->ctl);
  (void)hipHostFree(c->hctl);
  for (auto e : c->ev) (void)hipEventDestroy(e);
  delete c;
  return AMX_OK;
}

extern "C" amx_status amx_fv_set_timing(amx_fv_ctx* c, int enable) {
  if (!c) return fail(AMX_ERR_INVALID, "null context%s");
  HIPCHK(hipDeviceSynchronize());
  c->pending.clear();
  c->ev_used = 0;
  c->timing = enable ? 1 : 0;
  c->ms[0] = c->ms[1] = 0.0;
  c->launches[0] = c->launches[1] = 0;
  return AMX_OK;
}

extern "C" amx_status amx_fv_get_timing(amx_fv_ctx* c, amx_fv_timing* out) {
  if (!c || !out) return fail(AMX_ERR_INVALID, "null argument%s");
  HIPCHK(hipDeviceSynchronize());
  time_resolve(c, 1 << 30);
  out->stage_ms = c->ms[CLS_STAGE];
  out->stage_launches = c->launches[CLS_STAGE];
  out->mag_ms = c->ms[CLS_MAG];
  out->mag_launches = c->launches[CLS_MAG];
  return AMX_OK;
}

namespace {

template <int DIM, int SRC, bool HEUN, int W0M>
hipError_t launch_stage_t(const amx_fv_ctx* c, const StageArgs& a, hipStream_t st) {
  const dim3 grid(c->ntb * c->nta * c->ncx), block(TA_DEF * TB);
  if (c->cfg.riemann == AMX_HLLC)
    hipLaunchKernelGGL((k_stage<DIM, TA_DEF, SRC, HEUN, W0M, AMX_HLLC>), grid, block, 0, st, a);
  else
    hipLaunchKernelGGL((k_stage<DIM, TA_DEF, SRC, HEUN, W0M, AMX_HLL>), grid, block, 0, st, a);
  return hipGetLastError();
}

template <int SRC, bool HEUN, int W0M>
hipError_t launch_stage(const amx_fv_ctx* c, const StageArgs& a, hipStream_t st) {
  if (c->cfg.dim == 3) return launch_stage_t<3, SRC, HEUN, W0M>(c, a, st);
  return launch_stage_t<2, SRC, HEUN, W0M>(c, a, st);
}

StageArgs stage_base(const amx_fv_ctx* c, const float* dtp, float dt_scale) {
  StageArgs a;
  memset(&a, 0, sizeof(a));
  a.dtp = dtp;
  a.nx = c->nxm;
  a.na = c->na;
  a.nb = c->nb;
  a.ntb = c->ntb;
  a.nta = c->nta;
  a.ncx = c->ncx;
  a.lx = c->lx;
  a.dt_scale = dt_scale;
  a.ph = c->ph;
  return a;
}

MagArgs mag_base(const amx_fv_ctx* c, const float* dtp) {
  MagArgs m;
  memset(&m, 0, sizeof(m));
  m.dtp = dtp;
  m.ctl = c->ctl;
  m.max_iter = c->cfg.max_fixed_point_iterations;
  m.tol = c->cfg.fixed_point_tol;
  m.euler = c->cfg.magnetic_integrator == AMX_IMPLICIT_EULER;
  m.nx = c->nxm;
  m.na = c->na;
  m.nb = c->nb;
  m.ntb = c->ntb;
  m.nta = c->nta;
  m.ncx = c->ncx;
  m.lx = c->lx;
  m.inv_dx = c->ph.inv_dx;
  return m;
}

hipError_t launch_mag(const amx_fv_ctx* c, const MagArgs& m, hipStream_t st) {
  const dim3 grid(c->ntb * c->nta * c->ncx), block(TA_DEF * TB);
  if (c->cfg.dim == 3)
    hipLaunchKernelGGL((k_mag<3, TA_DEF>), grid, block, 0, st, m);
  else
    hipLaunchKernelGGL((k_mag<2, TA_DEF>), grid, block, 0, st, m);
  return hipGetLastError();
}

hipError_t launch_cfl(const amx_fv_ctx* c, const float* state, hipStream_t st) {
  const long n = c->ncell;
  const float* f[5];
  for (int v = 0; v < c->NV; ++v) f[v] = state + (long)v * n;
  long blocks = (n + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  unsigned* out = c->ctl->cfl;
  if (c->cfg.dim == 3)
    hipLaunchKernelGGL(k_cfl<3>, dim3(blocks), dim3(256), 0, st, f[0], f[1], f[2], f[3], f[4], n, c->cfg.gamma, out);
  else
    hipLaunchKernelGGL(k_cfl<2>, dim3(blocks), dim3(256), 0, st, f[0], f[1], f[2], (const float*)nullptr, f[3], n,
                       c->cfg.gamma, out);
  return hipGetLastError();
}

// One step on `state` with dt already in ctl->dt. Returns the fixed-point count.
amx_status do_step(amx_fv_ctx* c, float* state, hipStream_t st, int want_cfl, int* k_out) {
  const long n = c->ncell;
  const int NV = c->NV;
  const float* dtp = &c->ctl->dt;
  float* G1 = c->ws;
  float* gas[5];
  for (int v = 0; v < NV; ++v) gas[v] = state + (long)v * n;
  if (!c->cfg.mhd) {
    // SSP-RK2 on the whole state: W1 = E(W0); W = prim(0.5 (U0 + cons(E(W1)))) in place
    StageArgs a1 = stage_base(c, dtp, 1.0f);
    for (int v = 0; v < NV; ++v) {
      a1.src[v] = gas[v];
      a1.out[v] = G1 + (long)v * n;
    }
    int e = time_begin(c, st);
    HIPCHK((launch_stage<SRC_PLAIN, false, SRC_PLAIN>(c, a1, st)));
    time_end(c, st, e, CLS_STAGE, 0);
    StageArgs a2 = stage_base(c, dtp, 1.0f);
    for (int v = 0; v < NV; ++v) {
      a2.src[v] = G1 + (long)v * n;
      a2.w0[v] = gas[v];
      a2.out[v] = gas[v];
    }
    a2.cflmax = want_cfl ? c->ctl->cfl : nullptr;
    e = time_begin(c, st);
    HIPCHK((launch_stage<SRC_PLAIN, true, SRC_PLAIN>(c, a2, st)));
    time_end(c, st, e, CLS_STAGE, 0);
    if (k_out) *k_out = 0;
    return AMX_OK;
  }
  float* Bs[3];
  for (int k = 0; k < 3; ++k) Bs[k] = state + (long)(NV + k) * n;
  float* P[2][6];
  for (int j = 0; j < 2; ++j)
    for (int k = 0; k < 6; ++k) P[j][k] = c->ws + (long)(NV + 6 * j + k) * n;
  const int dim = c->cfg.dim;
  // ---- S_A(dt/2): RK2 with the result in place in the gas rows
  StageArgs a1 = stage_base(c, dtp, 0.5f);
  for (int v = 0; v < NV; ++v) {
    a1.src[v] = gas[v];
    a1.out[v] = G1 + (long)v * n;
  }
  int e = time_begin(c, st);
  HIPCHK((launch_stage<SRC_PLAIN, false, SRC_PLAIN>(c, a1, st)));
  time_end(c, st, e, CLS_STAGE, 0);
  StageArgs a2 = stage_base(c, dtp, 0.5f);
  for (int v = 0; v < NV; ++v) {
    a2.src[v] = G1 + (long)v * n;
    a2.w0[v] = gas[v];
    a2.out[v] = gas[v];
  }
  e = time_begin(c, st);
  HIPCHK((launch_stage<SRC_PLAIN, true, SRC_PLAIN>(c, a2, st)));
  time_end(c, st, e, CLS_STAGE, 0);
  // ---- S_B(dt): fixed point on (B, v) with rho from the gas
  MagArgs m = mag_base(c, dtp);
  for (int k = 0; k < 3; ++k) {
    m.B0[k] = Bs[k];
    m.Bk[k] = Bs[k];
  }
  m.v0[0] = gas[1];
  m.v0[1] = gas[2];
  m.v0[2] = dim == 3 ? gas[3] : nullptr;
  m.vk[0] = m.v0[0];
  m.vk[1] = m.v0[1];
  m.vk[2] = m.v0[2];
  m.rho = gas[0];
  for (int k = 0; k < 3; ++k) {
    m.Bo[k] = P[0][k];
    m.vo[k] = P[0][3 + k];
  }
  m.k = 0;
  e = time_begin(c, st);
  HIPCHK(launch_mag(c, m, st));
  time_end(c, st, e, CLS_MAG, 0);
  const int maxit = c->cfg.max_fixed_point_iterations;
  int kfirst = 1;
  int batch = c->last_k + 1;
  if (batch < 1) batch = 1;
  int K = 0;
  for (;;) {
    const int klast = kfirst + batch - 1 < maxit ? kfirst + batch - 1 : maxit;
    for (int k = kfirst; k <= klast; ++k) {
      MagArgs mk = m;
      mk.k = k;
      for (int i = 0; i < 3; ++i) {
        mk.Bk[i] = P[(k - 1) & 1][i];
        mk.vk[i] = P[(k - 1) & 1][3 + i];
        mk.Bo[i] = P[k & 1][i];
        mk.vo[i] = P[k & 1][3 + i];
      }
      const int ek = time_begin(c, st);
      HIPCHK(launch_mag(c, mk, st));
      time_end(c, st, ek, CLS_MAG, k);
    }
    const size_t nb = offsetof(Ctl, res) + sizeof(unsigned) * (size_t)(klast + 2);
    HIPCHK(hipMemcpyAsync(c->hctl, c->ctl, nb, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    K = c->hctl->iters;
    time_resolve(c, K);
    float r;
    memcpy(&r, &c->hctl->res[K], sizeof(float));
    const bool done = !(r > c->cfg.fixed_point_tol) || K >= maxit;
    if (done || klast >= maxit) break;
    if (K < klast) break;  // a kernel exited early: converged (defensive)
    kfirst = klast + 1;
    batch = 2;
  }
  c->last_k = K;
  // ---- S_A(dt/2) on the magnetically updated gas (v <- vn, energy fix on the fly)
  float* const* Pn = P[K & 1];
  StageArgs b1 = stage_base(c, dtp, 0.5f);
  for (int v = 0; v < NV; ++v) b1.src[v] = gas[v];
  b1.src[5] = Pn[3];
  b1.src[6] = Pn[4];
  b1.src[7] = Pn[5];
  for (int v = 0; v < NV; ++v) b1.out[v] = G1 + (long)v * n;
  e = time_begin(c, st);
  HIPCHK((launch_stage<SRC_MAGFIX, false, SRC_PLAIN>(c, b1, st)));
  time_end(c, st, e, CLS_STAGE, 0);
  StageArgs b2 = stage_base(c, dtp, 0.5f);
  for (int v = 0; v < NV; ++v) {
    b2.src[v] = G1 + (long)v * n;
    b2.w0[v] = gas[v];
    b2.out[v] = gas[v];
  }
  b2.w0[5] = Pn[3];
  b2.w0[6] = Pn[4];
  b2.w0[7] = Pn[5];
  for (int k = 0; k < 3; ++k) {
    b2.bsrc[k] = Pn[k];
    b2.bdst[k] = Bs[k];
  }
  b2.cflmax = want_cfl ? c->ctl->cfl : nullptr;
  e = time_begin(c, st);
  HIPCHK((launch_stage<SRC_PLAIN, true, SRC_MAGFIX>(c, b2, st)));
  time_end(c, st, e, CLS_STAGE, 0);
  if (k_out) *k_out = K;
  return AMX_OK;
}

}  // namespace

extern "C" amx_status amx_fv_cfl_dt(amx_fv_ctx* c, const float* state, float* d_dt, void* stream) {
  if (!c || !state || !d_dt) return fail(AMX_ERR_INVALID, "null argument%s");
  hipStream_t st = (hipStream_t)stream;
  HIPCHK(hipMemsetAsync(c->ctl->cfl, 0, sizeof(c->ctl->cfl), st));
  HIPCHK(launch_cfl(c, state, st));
  hipLaunchKernelGGL(k_cfl_out, dim3(1), dim3(1), 0, st, (const Ctl*)c->ctl, c->cfg.C_cfl, c->dx, c->cfg.dim, d_dt);
  HIPCHK(hipGetLastError());
  return AMX_OK;
}

extern "C" amx_status amx_fv_evolve_state(amx_fv_ctx* c, float* state, const float* d_dt, int* iterations,
                                          void* stream) {
  if (!c || !state || !d_dt) return fail(AMX_ERR_INVALID, "null argument%s");
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_set_dt, dim3(1), dim3(1), 0, st, c->ctl, d_dt);
  HIPCHK(hipGetLastError());
  int K = 0;
  amx_status s = do_step(c, state, st, 0, &K);
  if (s != AMX_OK) return s;
  if (iterations) *iterations = K;
  return AMX_OK;
}

extern "C" amx_status amx_fv_time_integration(amx_fv_ctx* c, float* state, const amx_fv_run* run,
                                              amx_fv_stats* stats, void* stream) {
  if (!c || !state || !run) return fail(AMX_ERR_INVALID, "null argument%s");
  if (run->fixed_timestep && run->num_timesteps <= 0) return fail(AMX_ERR_INVALID, "num_timesteps must be positive%s");
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_reset, dim3(1), dim3(1), 0, st, c->ctl);
  HIPCHK(hipGetLastError());
  if (!run->fixed_timestep) {
    HIPCHK(launch_cfl(c, state, st));
  }
  StepParams sp;
  sp.t_end = run->t_end;
  sp.fixed = run->fixed_timestep;
  sp.num_timesteps = run->num_timesteps;
  sp.exact_end = run->exact_end_time;
  sp.C = c->cfg.C_cfl;
  sp.dx = c->dx;
  sp.dim = c->cfg.dim;
  sp.use_cfl = 1;
  long steps = 0, iters = 0;
  int kmax = 0;
  float t = 0.f;
  for (;;) {
    if (run->fixed_timestep) {
      if (steps >= run->num_timesteps) break;
    } else if (!(t < run->t_end)) {
      break;
    }
    if (run->max_steps >= 0 && steps >= run->max_steps) break;
    hipLaunchKernelGGL(k_step_begin, dim3(1), dim3(1), 0, st, c->ctl, sp);
    HIPCHK(hipGetLastError());
    int K = 0;
    amx_status s = do_step(c, state, st, !run->fixed_timestep, &K);
    if (s != AMX_OK) return s;
    hipLaunchKernelGGL(k_step_end, dim3(1), dim3(1), 0, st, c->ctl);
    HIPCHK(hipGetLastError());
    if (!c->cfg.mhd) {
      // hydro steps have no fixed-point readback: read dt here
      HIPCHK(hipMemcpyAsync(c->hctl, c->ctl, offsetof(Ctl, res), hipMemcpyDeviceToHost, st));
      HIPCHK(hipStreamSynchronize(st));
      time_resolve(c, 1 << 30);
      t = c->hctl->t;
    } else {
      // hctl holds dt and t from the fixed-point readback (before k_step_end)
      t = c->hctl->t + c->hctl->dt;
    }
    iters += K;
    if (K > kmax) kmax = K;
    ++steps;
  }
  HIPCHK(hipMemcpyAsync(c->hctl, c->ctl, offsetof(Ctl, res), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  time_resolve(c, 1 << 30);
  if (stats) {
    stats->steps = steps;
    stats->t = c->hctl->t;
    stats->fixed_point_iterations = iters;
    stats->max_fixed_point_iterations = kmax;
  }
  return AMX_OK;
}

extern "C" const char* amx_status_string(amx_status s) {
  switch (s) {
    case AMX_OK: return "ok";
    case AMX_ERR_INVALID: return "invalid argument";
    case AMX_ERR_HIP: return "HIP runtime error";
    case AMX_ERR_NOMEM: return "out of device memory";
    case AMX_ERR_UNSUPPORTED: return "unsupported configuration";
  }
  return "unknown status";
}

extern "C" const char* amx_last_error(void) { return g_err; }
