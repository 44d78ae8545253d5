// MI355X (gfx950) kernels and C ABI of the finite-volume MHD step.
//
// One kernel per global-synchronisation interval of the reference dataflow:
//   * k_stage   one forward-Euler unsplit gas stage (limited reconstruction +
//               Riemann fluxes + flux-divergence update), optionally fused with
//               the SSP-RK2 Heun average, the Pang-Wu energy fix on its inputs,
//               the CFL maxima of its output and (outside the step loop) the
//               copy of the final B into the state.
//               Follows _evolve_gas_state_unsplit_inner / _evolve_gas_state_unsplit
//               (astronomix/_finite_volume/_state_evolution/evolve_state.py:338-469).
//   * k_lf_alpha, k_am_switch, k_pp_factors  grid-wide or per-cell inputs of
//               the Lax-Friedrichs, hybrid / AM-HLLC and VAN_ALBADA_PP stages.
//   * k_mag     two consecutive fixed-point iterates of the implicit-midpoint
//               magnetic sub-step with their convergence residuals
//               (astronomix/_finite_volume/_magnetic_update/_magnetic_field_update.py:82-165).
//   * k_cfl     the unsplit CFL maxima (_timestep_estimator.py:118-128).
//
// Stencil kernels march along the slowest array axis (x in 3D) with a register
// queue of the thread's own column, and keep the current (y, z) plane of a
// tile in LDS with its halo, so each cell value is read from HBM once per pass
// and each interface flux is computed once. Tiles are 64 cells wide along the
// contiguous axis (one wavefront, coalesced) and TA rows tall; the block index
// is swizzled so that neighbouring tiles land on the same XCD (shared L2).
// Boundaries are applied by mapping indices: the reference pads the state with
// two ghost cells per side and refreshes them before every stencil
// (_geometry/boundaries.py:74-357). Periodic ghosts are wrapped indices; open
// and reflective ghosts (kernels instantiated with BCX) are the edge cell and
// the mirror cell, with one velocity / field row negated at reflective walls.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <cstdint>
#include <functional>
#include <map>
#include <mutex>
#include <tuple>
#include <new>
#include <utility>
#include <vector>

#include "amx.h"
#include "common.h"
#include "snapshots.h"

namespace amx {

constexpr int TB = 64;      // tile width along the contiguous axis (one wavefront)
#ifndef AMX_STAGE_WAVES
#define AMX_STAGE_WAVES 4   // two 8-wave stage blocks per CU: 4 waves/SIMD, <= 128 VGPRs
#endif

// Device-resident step control: step size, time, fixed-point bookkeeping and
// the CFL maxima. Reductions go through NSLOT partial slots (block b adds into
// slot b % NSLOT, one atomic per block) so that thousands of blocks do not
// serialise on one address; the consumer reduces the slots. All reduced values
// are >= 0 floats stored as bits, so unsigned max == float max (NaN, made
// canonical 0x7fc00000, wins as in jnp.max).
// residual slot sets: a pair launch (L, L+1) reads L-2, L-1, accumulates
// into L, L+1 and clears L+2, L+3 for the next pair (all mod NRES)
constexpr int NRES = 6;

struct Ctl {
  float dt;
  float t;
  int iters;   // fixed-point while-body iterations of the current step
  int steps;
  int done;    // fixed point finished (sticky until the next step)
  int live;    // the slot holds a simulation
  int run;     // the slot takes part in the current step
  float t_end; // end time of the slot's simulation
  // queued steps (step loops that read back once per batch of steps): a
  // step whose queued fixed-point launches did not settle sets stall, and
  // the steps queued behind it do nothing until the host has finished it;
  // the while-body iterations of the completed steps are summed on the device
  int stall;
  int iters_sum;
  int iters_max;
  int iters_bmax;  // largest count of the current batch
  unsigned cfl[4];              // reduced CFL maxima per axis
  unsigned cflp[3][NSLOT];      // CFL partial maxima
  unsigned resp[NRES][NSLOT];   // fixed-point residual partials, by launch index % NRES
  unsigned lfp[3][NSLOT];       // Lax-Friedrichs: partial maxima of |u| + c of the stage input
  int bset;                     // iterate set holding B0 of the magnetic sub-step; -1: the state's B rows
  int held;                     // the fixed point had not settled when k_mag_final ran: gated stage kernels return
  int kprev;                    // final launch index of the slot's previous step (k_mag_final): the speculative
                                // pair starting there stores its first iterate, the others leave it out
  int regen_ev;                 // steps of the current batch that ended on a first iterate their pair did not
                                // store (previous final index known): the next batch queues regeneration launches
};

// Iterate set of fixed-point launch k: the three sets other than bset
// (with bset = -1, sets 0, 1, 2 in turn). A loop that keeps the final B in
// its set instead of copying it into the state moves bset there at the end
// of the step, so the next sub-step's launches write around it.
__host__ __device__ inline int iterate_set(int bset, int k) {
  const int ex = bset < 0 ? 3 : bset;
  const int j = k % 3;
  return j < ex ? j : j + 1;
}

struct Phys {
  float gamma;    // params.gamma
  float gm1;      // gamma - 1 evaluated in float, as the reference does
  float inv_gm1;  // 1 / (gamma - 1)
  float inv_dx;   // float(1 / grid_spacing)
};

// gamma is hard-coded to 5/3 in the magnetic energy fix (_magnetic_field_update.py:213,221)
constexpr float G53M1 = 1.6666666f - 1.0f;

// limiters.py:13-15: 0.5 (sgn a + sgn b) min(|a|,|b|) is the argument of
// smaller magnitude when both have the same strict sign, and 0 otherwise --
// i.e. the median of {a, b, 0}: one v_med3_f32.
__device__ __forceinline__ float minmod(float a, float b) { return __builtin_amdgcn_fmed3f(a, b, 0.f); }

// limited gradient times dx/2 (limited_gradients.py:50-60, reconstruction.py:409-412)
__device__ __forceinline__ float half_slope(float wm, float w0, float wp) {
  return 0.5f * minmod(w0 - wm, wp - w0);
}

// Limiter ids (simulation_config.py:29-34 values).
enum { LIM_MINMOD = 0, LIM_OSHER = 1, LIM_DOUBLE_MINMOD = 2, LIM_SUPERBEE = 3, LIM_VAN_ALBADA = 4, LIM_VAN_ALBADA_PP = 5 };

// The other limiters of _calculate_limited_gradients (limited_gradients.py:49-88,
// limiters.py:18-64) on the backward / forward gradients b, f, chosen by a
// block-uniform id; the result is (g dx) / 2 as the reconstruction applies it.
__device__ __forceinline__ float half_slope_x(float wm, float w0, float wp, int lim, float dx) {
  const float b = (w0 - wm) / dx, f = (wp - w0) / dx;
  float g;
  if (lim == LIM_SUPERBEE) {
    // where(a b > 0, minmod(maxmod(a, b), minmod(2a, 2b)), 0): for a, b of one
    // sign s that is s min(max(|a|,|b|), 2 min(|a|,|b|))
    const float aa = fabsf(b), ab = fabsf(f);
    g = b * f > 0.f ? copysignf(fminf(fmaxf(aa, ab), 2.f * fminf(aa, ab)), b) : 0.f;
  } else if (lim == LIM_DOUBLE_MINMOD) {
    // where(a b > 0, minmod3((a + b) / 2, 2a, 2b), 0)
    const float m = (b + f) / 2.f;
    g = b * f > 0.f ? copysignf(fminf(fabsf(m), fminf(fabsf(2.f * b), fabsf(2.f * f))), m) : 0.f;
  } else if (lim == LIM_VAN_ALBADA) {
    // ((f^2 + eps) b + (b^2 + eps) f) / (f^2 + b^2 + 2 eps), eps = 3 dx
    const float eps = 3.f * dx, eps2 = 6.f * dx;
    g = ((f * f + eps) * b + (b * b + eps) * f) / ((f * f + b * b) + eps2);
  } else {  // LIM_OSHER: b max(0, min(1.3, f / (b + 1e-11))) where |b| > 1e-11
    const float r = fabsf(b) > 1e-11f ? f / (b + 1e-11f) : 0.f;
    g = fmaxf(0.f, fminf(1.3f, r)) * b;
  }
  return (g * dx) / 2.f;
}

// Limited half slope: minmod in its exact difference form, or another limiter.
template <bool LIMX>
__device__ __forceinline__ float slope(float wm, float w0, float wp, int lim, float dx) {
  if (LIMX) return half_slope_x(wm, w0, wp, lim, dx);
  return half_slope(wm, w0, wp);
}

// Boundary codes (AMX_BC_*) of the two sides of one axis.
struct AxisBc {
  int lo, hi;
};

// Cell of index j of an axis of n cells (_geometry/boundaries.py:268-357):
// periodic wraps; open repeats the edge cell; reflective mirrors (ghost -1 is
// cell 0, ghost -2 cell 1, ghost n cell n-1) and sets flip, the negation of
// one row. Indices beyond the two ghost layers (rows and columns of a partial
// tile) are clamped into the axis; no result depends on them.
__device__ __forceinline__ int bnd_index(int j, int n, AxisBc b, bool& flip) {
  flip = false;
  if (j < 0) {
    if (b.lo == AMX_BC_PERIODIC) return wrapi(j, n);
    if (b.lo == AMX_BC_REFLECTIVE) {
      flip = true;
      j = -j - 1;
    } else {
      j = 0;
    }
    return j < n ? j : n - 1;
  }
  if (j >= n) {
    if (b.hi == AMX_BC_PERIODIC) return wrapi(j, n);
    if (b.hi == AMX_BC_REFLECTIVE) {
      flip = true;
      j = 2 * n - 1 - j;
    } else {
      j = n - 1;
    }
    return j >= 0 ? j : 0;
  }
  return j;
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
  const float ri = frcp(u[0]);
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
  const float cl = fsqrt(ph.gamma * pl * frcp(rl));
  const float cr = fsqrt(ph.gamma * pr * frcp(rr));
  const float El = pl * ph.inv_gm1 + (0.5f * rl) * vsq<DIM>(wl);
  const float Er = pr * ph.inv_gm1 + (0.5f * rr) * vsq<DIM>(wr);
  const float sp = fmaxf(fmaxf(ul + cl, ur + cr), 0.f);
  const float sm = fminf(fminf(ul - cl, ur - cr), 0.f);
  // (S+ F_L - S- F_R + S- S+ (U_R - U_L)) / (S+ - S-) with the Euler fluxes
  // F = u U + p (e_AX + u e_E): the conserved states enter with the weights
  // kl = S+ (u_L - S-) / (S+ - S-) and kr = S- (S+ - u_R) / (S+ - S-) (both
  // differences are >= the sound speeds, so no cancellation), the pressures
  // with al = S+ / (S+ - S-) and ar = -S- / (S+ - S-)
  const float inv = frcp(sp - sm);
  const float kl = (sp * (ul - sm)) * inv, kr = (sm * (sp - ur)) * inv;
  const float apl = (sp * inv) * pl, apr = -(sm * inv) * pr;
  f[0] = kl * rl + kr * rr;
#pragma unroll
  for (int i = 1; i <= DIM; ++i) {
    f[i] = kl * (rl * wl[i]) + kr * (rr * wr[i]);
    if (i == AX) f[i] += apl + apr;
  }
  f[P] = (kl * El + kr * Er) + (apl * ul + apr * ur);
}

// HLLC interface flux, central form (_finite_volume/_riemann_solver/hll.py:113-261);
// lm: the low-Mach variant HLLC-LM, whose dissipation takes S_L, S_R scaled
// by phi = sin(min(1, Ma / 0.1) pi / 2) (hll.py:206-221); LMD: low-Mach
// dissipation control, the velocity rows of the dissipation scaled by
// min(1, max(|v_L| / c_L, |v_R| / c_R)) (hll.py:230-254).
template <int DIM, int AX, bool LMD>
__device__ __forceinline__ void flux_hllc(const float* wl, const float* wr, float* f, const Phys& ph, bool lm) {
  constexpr int P = DIM + 1, NV = DIM + 2;
  const float rl = wl[0], ul = wl[AX], pl = wl[P];
  const float rr = wr[0], ur = wr[AX], pr = wr[P];
  const float irl = frcp(rl), irr = frcp(rr);
  const float cl = fsqrt(ph.gamma * pl * irl);
  const float cr = fsqrt(ph.gamma * pr * irr);
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
  const float sql = fsqrt(rl), sqr = fsqrt(rr);
  const float ssum = sql + sqr;
  const float iss = frcp(ssum);
  const float u_hat = (sql * ul + sqr * ur) * iss;
  const float du = ur - ul;
  const float c_hat2 = ((cl * cl) * sql + (cr * cr) * sqr) * iss + (0.5f * ((sql * sqr) * (iss * iss))) * (du * du);
  const float c_hat = fsqrt(c_hat2);
  const float SL = fminf(ul - cl, u_hat - c_hat);
  const float SR = fmaxf(ur + cr, u_hat + c_hat);
  const float s_star = (((pr - pl) + (rl * ul) * (SL - ul)) - (rr * ur) * (SR - ur)) * frcp(rl * (SL - ul) - rr * (SR - ur));
  const float fl = (SL - ul) * frcp(SL - s_star);
  const float fr = (SR - ur) * frcp(SR - s_star);
  const float as = fabsf(s_star);
  float SLd = SL, SRd = SR;
  if (lm) {
    const float ma = fmaxf(fabsf(ul / cl), fabsf(ur / cr));
    const float phi = sinf((fminf(1.f, ma / 0.1f) * 3.14159265358979f) / 2.f);
    SLd = SL * phi;
    SRd = SR * phi;
  }
  float fmd = 1.f;
  if (LMD) {
    // get_absolute_velocity (_equations.py:165-197): +1e-20 in 3D only
    const float al = sqrtf(DIM == 3 ? vsq<DIM>(wl) + 1e-20f : vsq<DIM>(wl));
    const float ar = sqrtf(DIM == 3 ? vsq<DIM>(wr) + 1e-20f : vsq<DIM>(wr));
    fmd = fminf(1.f, fmaxf(fabsf(al / cl), fabsf(ar / cr)));
  }
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    float usl = UL[v], usr = UR[v];
    if (v == AX) {
      usl = rl * s_star;
      usr = rr * s_star;
    }
    if (v == P) {
      usl = usl + (s_star - ul) * (rl * s_star + pl * frcp(SL - ul));
      usr = usr + (s_star - ur) * (rr * s_star + pr * frcp(SR - ur));
    }
    usl *= fl;
    usr *= fr;
    float diss = 0.5f * ((SLd * (usl - UL[v]) + as * (usl - usr)) + SRd * (usr - UR[v]));
    if (LMD && v >= 1 && v <= DIM) diss = fmd * diss;
    const float fs = 0.5f * (FL[v] + FR[v]) + diss;
    float out = SL >= 0.f ? FL[v] : fs;
    f[v] = SR <= 0.f ? FR[v] : out;
  }
}

// Lax-Friedrichs flux (_finite_volume/_riemann_solver/_lax_friedrichs.py:21-78):
// 0.5 (F_L + F_R) - 0.5 alpha (U_R - U_L) with alpha = max over the stage
// input of |u| + c along the flux direction (a grid-wide maximum).
template <int DIM, int AX>
__device__ __forceinline__ void flux_lf(const float* wl, const float* wr, float* f, const Phys& ph, float alpha) {
  constexpr int P = DIM + 1;
  const float rl = wl[0], ul = wl[AX], pl = wl[P];
  const float rr = wr[0], ur = wr[AX], pr = wr[P];
  const float El = pl * ph.inv_gm1 + (0.5f * rl) * vsq<DIM>(wl);
  const float Er = pr * ph.inv_gm1 + (0.5f * rr) * vsq<DIM>(wr);
  const float ha = 0.5f * alpha;
  f[0] = 0.5f * (ul * rl + ur * rr) - ha * (rr - rl);
#pragma unroll
  for (int i = 1; i <= DIM; ++i) {
    const float ml = rl * wl[i], mr = rr * wr[i];
    float fl = ul * ml, fr = ur * mr;
    if (i == AX) {
      fl += pl;
      fr += pr;
    }
    f[i] = 0.5f * (fl + fr) - ha * (mr - ml);
  }
  f[P] = 0.5f * (ul * (pl + El) + ur * (pr + Er)) - ha * (Er - El);
}

// Riemann solver ids (simulation_config.py:41-46 values).
// gsw: the compression switch of the hybrid / AM-HLLC solvers at the cell
// right of the interface (HLLC-LM where it is set, hll.py:297-326).
template <int DIM, int AX, int RS>
__device__ __forceinline__ void riemann(const float* wl, const float* wr, float* f, const Phys& ph, float alpha,
                                        bool gsw) {
  if (RS == AMX_HLLC)
    flux_hllc<DIM, AX, false>(wl, wr, f, ph, false);
  else if (RS == AMX_HLLC_LM)
    flux_hllc<DIM, AX, false>(wl, wr, f, ph, true);
  else if (RS == AMX_HYBRID_HLLC)
    flux_hllc<DIM, AX, false>(wl, wr, f, ph, gsw);
  else if (RS == AMX_AM_HLLC)
    flux_hllc<DIM, AX, true>(wl, wr, f, ph, gsw);
  else if (RS == AMX_LAX_FRIEDRICHS)
    flux_lf<DIM, AX>(wl, wr, f, ph, alpha);
  else
    flux_hll<DIM, AX>(wl, wr, f, ph);
}

enum { SRC_PLAIN = 0, SRC_MAGFIX = 1, SRC_PFIX = 2 };

// Primitive gas state at one cell, in two halves so that loads can be issued
// a plane ahead of their use: load_raw fetches the fields, derive_w forms W.
// SRC_PLAIN reads the NV gas rows. SRC_MAGFIX reads (rho, v0, p) of the gas
// state entering the magnetic sub-step plus the converged velocity vn (slots
// 5..7) and applies the sub-step's write-back: v <- vn and the energy fix with
// gamma = 5/3 (_magnetic_field_update.py:208-240). SRC_PFIX (Heun centre
// only) reads the same fixed state ready-made: rho, the converged velocity
// and the fixed pressure that the SRC_MAGFIX stage stored (StageArgs::pfix).
// Element c of a field through a 32-bit byte offset (c < 2^30): the load
// takes the scalar base plus one 32-bit vector offset, no 64-bit address math.
__device__ __forceinline__ float ld32(const float* base, unsigned c) {
  return *(const float*)((const char*)base + (c << 2));
}
__device__ __forceinline__ void st32(float* base, unsigned c, float v) { *(float*)((char*)base + (c << 2)) = v; }
// a store whose data this kernel does not read back (streaming: no reuse in the L2 expected)
__device__ __forceinline__ void st32nt(float* base, unsigned c, float v) {
  __builtin_nontemporal_store(v, (float*)((char*)base + (c << 2)));
}

// s_waitcnt immediate (gfx9 encoding): vmcnt 0, expcnt and lgkmcnt at their
// maxima (no wait on them)
constexpr int VMCNT0 = 0x0F70;

template <int DIM, int MODE>
struct Raw {
  static constexpr int N = MODE != SRC_MAGFIX ? DIM + 2 : 8;
  float f[N];
};

// NT: fields this kernel reads at one cell only (non-temporal loads)
template <int DIM, int MODE, bool NT = false>
__device__ __forceinline__ void load_raw(const float* const* s, unsigned c, Raw<DIM, MODE>& r) {
  auto ld = [&](const float* b) {
    const float* p = (const float*)((const char*)b + (c << 2));
    return NT ? __builtin_nontemporal_load(p) : *p;
  };
  if (MODE != SRC_MAGFIX) {
#pragma unroll
    for (int v = 0; v < DIM + 2; ++v) r.f[v] = ld(s[v]);
  } else {
#pragma unroll
    for (int v = 0; v < 8; ++v) r.f[v] = (DIM == 2 && v == 4) ? 0.f : ld(s[v]);
  }
}

template <int DIM, int MODE>
__device__ __forceinline__ void derive_w(const Raw<DIM, MODE>& r, float* w) {
  constexpr int NV = DIM + 2, P = DIM + 1;
  if (MODE != SRC_MAGFIX) {
#pragma unroll
    for (int v = 0; v < NV; ++v) w[v] = r.f[v];
  } else {
    const float rho = r.f[0], p = r.f[P];
    const float v0x = r.f[1], v0y = r.f[2], v0z = DIM == 3 ? r.f[3] : 0.f;
    const float vnx = r.f[5], vny = r.f[6], vnz = r.f[7];
    const float n0 = fsqrt((v0x * v0x + v0y * v0y) + v0z * v0z);
    const float nn = fsqrt((vnx * vnx + vny * vny) + vnz * vnz);
    const float hr = 0.5f * rho;
    const float E = p * (1.0f / G53M1) + hr * (n0 * n0);
    const float E2 = (E - hr * (n0 * n0)) + hr * (nn * nn);
    const float e = E2 * frcp(rho) - 0.5f * (nn * nn);
    w[0] = rho;
    w[1] = vnx;
    w[2] = vny;
    if (DIM == 3) w[3] = vnz;
    w[P] = (G53M1 * rho) * e;
  }
}

template <int DIM, int MODE>
__device__ __forceinline__ void load_w(const float* const* s, unsigned c, float* w) {
  Raw<DIM, MODE> r;
  load_raw<DIM, MODE>(s, c, r);
  derive_w<DIM, MODE>(r, w);
}

// Buffers are given for slot 0; slot s (blockIdx.y + slot0) adds s * sstride
// floats to every pointer and uses ctl[s] and phs[s]. A single simulation is
// slot 0 with sstride 0.
struct StageArgs {
  const float* src[8];   // stencil input (see load_w)
  const float* w0[8];    // Heun: stage-0 state at the cell centre
  float* out[5];         // output gas rows
  float* bdst[3];        // optional copy of the final B of the magnetic sub-step
  // iterate sets of the fixed point (set 3 only in single-simulation
  // contexts); the final iterate's set is chosen on the device from its launch
  // index K (iterate_set): v_K replaces src[5..7] (SRC_MAGFIX) and the Heun
  // centre's velocity (SRC_PFIX), B_K is the copy source
  const float* selv[4][3];
  const float* selb[4][3];
  Ctl* ctl;
  const Phys* phs;
  long sstride;
  int slot0;
  int want_cfl;          // reduce the CFL maxima of the output into ctl[s].cflp
  int gated;             // return at once while ctl[s].held is set
  int copy_b;            // copy B_K (selb) into bdst
  int nx, na, nb;        // 3D: (nx, ny, nz); 2D: (1, nx, ny)
  int ntb, nta, ncx, lx;
  float dt_scale;        // 0.5: MHD half step dt/2; 1: hydro step
  int limiter;           // LIM_* (the kernels instantiated with LIMX read it)
  float dxf;             // float(grid_spacing): the limiters' gradient scale
  unsigned char* amg;    // hybrid / AM-HLLC: compression switch per cell of the stage input
  long gstride;          // bytes between consecutive slots' switch arrays
  AxisBc bx, ba, bb;     // boundaries of the kernel axes x (3D), A and B (kernels with BCX)
  // VAN_ALBADA_PP: per-cell factors of the stage input (k_pp_factors), in
  // state-axis order alpha_0.., kappa_0.., beta; slot s adds s * ppstride
  float* ppf[7];
  long ppstride;
  float ppq;             // q = 2 / C_cfl with MHD, 1 / C_cfl without
  int pp;                // the kernels instantiated with LIMX scale their slopes by ppf
  // energy-fixed pressure of the magnetic sub-step's output: written by the
  // SRC_MAGFIX stage (when set), read by the W0M == SRC_PFIX Heun centre
  float* pfix;
};

// One unsplit forward-Euler stage (evolve_state.py:338-414), optionally with
// the RK2 Heun average (evolve_state.py:456-469).
//
// Block = (TA + 1) wavefronts of 64 lanes. Waves 0..TA-1 own the TA x 64 tile
// cells of the current plane (lane = contiguous axis B, wave = row along A);
// the extra wave only loads halo and computes interface fluxes, so that the
// (TA+1)*64 A-interfaces and TA*65 B-interfaces of a plane take exactly two
// rounds of the block. In 3D the block marches along x over lx planes with the
// x-stencil of its own column in registers.
//
// BCX: open / reflective boundaries. Ghost cells read the cell that
// bnd_index maps them to, with the normal velocity negated at reflective
// walls (_reflective_*_boundary, boundaries.py:323-357); the flux into the
// ghost layer at an upper edge is computed from those values like any other.
template <int DIM, int TA, int SRC, bool HEUN, int W0M, int RS, bool LIMX, bool BCX>
__global__ __launch_bounds__((TA + 1) * TB) __attribute__((amdgpu_waves_per_eu(AMX_STAGE_WAVES))) void k_stage(const StageArgs a) {
  constexpr int NV = DIM + 2, P = DIM + 1, NT = (TA + 1) * TB, WW = TB + 4;
  constexpr int AXA = DIM == 3 ? 2 : 1;  // velocity row of the plane axis A
  constexpr int AXB = DIM == 3 ? 3 : 2;  // velocity row of the plane axis B (contiguous)
  constexpr int NH = 4 * TB + 4 * TA;    // halo cells of a plane (cross shape)
  constexpr int NFA = (TA + 1) * TB, NF = NFA + TA * (TB + 1);
  static_assert(NFA == NT && NF - NFA <= NT, "flux rounds assume (TA+1)*TB threads");
  static_assert(NH <= NT, "every halo cell of a plane needs its own thread (TA >= 4)");
  // LAG (3D, limiters without plane factors): the cell update runs one
  // plane behind the interface fluxes, and with the plane arrays double
  // buffered one barrier per plane orders both phases
  constexpr bool LAG = DIM == 3 && !LIMX;
  constexpr int NBUF = LAG ? 2 : 1;
  __shared__ float sW[NBUF][NV][(TA + 4) * WW];
  __shared__ float sFA[NBUF][NV][(TA + 1) * TB];
  __shared__ float sFB[NBUF][NV][TA * (TB + 1)];
  constexpr bool AMS = RS == AMX_HYBRID_HLLC || RS == AMX_AM_HLLC;
  static_assert(!(AMS && BCX), "the compression switch of ghost cells is not defined here");
  __shared__ unsigned char sG[NBUF][AMS ? (TA + 4) * WW : 1];
  // VAN_ALBADA_PP factors of the plane cells: alpha_A, kappa_A, alpha_B, kappa_B, beta
  __shared__ float sPF[LIMX ? 5 : 1][LIMX ? (TA + 4) * WW : 1];

  const int slot = a.slot0 + blockIdx.y;
  Ctl* const C = a.ctl + slot;
  if (!C->run) return;  // block-uniform: empty or finished slot
  if (a.gated && C->held) return;  // queued behind a fixed point that needs more launches
  const long so = slot * a.sstride;
  const int par = iterate_set(C->bset, C->iters);  // set of the fixed point's final iterate
  const float* src[8];
  const float* w0p[8];
#pragma unroll
  for (int v = 0; v < 8; ++v) {
    src[v] = a.src[v] ? a.src[v] + so : nullptr;
    w0p[v] = a.w0[v] ? a.w0[v] + so : nullptr;
  }
  if (SRC == SRC_MAGFIX) {
#pragma unroll
    for (int k = 0; k < 3; ++k) src[5 + k] = a.selv[par][k] + so;
  }
  if (W0M == SRC_MAGFIX) {
#pragma unroll
    for (int k = 0; k < 3; ++k) w0p[5 + k] = a.selv[par][k] + so;
  }
  if (W0M == SRC_PFIX) {
#pragma unroll
    for (int k = 0; k < DIM; ++k) w0p[1 + k] = a.selv[par][k] + so;
    w0p[DIM + 1] = a.pfix + so;
  }
  float* outp[5];
#pragma unroll
  for (int v = 0; v < NV; ++v) outp[v] = a.out[v] + so;
  const Phys ph = a.phs[slot];
  const int tid = threadIdx.x, ta = tid / TB, tb = tid % TB;
  const int lim = a.limiter;
  const float ldx = a.dxf;
  // VAN_ALBADA_PP: slopes scaled by the factor of their cell (reconstruction.py:343-365)
  const bool pp = LIMX && a.pp;
  const float* pfp[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};  // plane factors, as sPF
  const float* pfx[2] = {nullptr, nullptr};                             // alpha_x, kappa_x (3D)
  if (pp) {
    constexpr int XA = DIM == 3 ? 1 : 0, XB = DIM == 3 ? 2 : 1;  // state axes of A and B
    const long po = slot * a.ppstride;
    pfp[0] = a.ppf[XA] + po;
    pfp[1] = a.ppf[DIM + XA] + po;
    pfp[2] = a.ppf[XB] + po;
    pfp[3] = a.ppf[DIM + XB] + po;
    pfp[4] = a.ppf[2 * DIM] + po;
    pfx[0] = a.ppf[0] + po;
    pfx[1] = a.ppf[DIM] + po;
  }
  // factor of row v from a cell's (alpha, kappa, beta)
  auto fsel = [&](float fa_, float fk_, float fb_, int v) { return v == 0 ? fa_ : (v == P ? fk_ : fb_); };
  const unsigned char* gsw = AMS ? a.amg + slot * a.gstride : nullptr;
  // Lax-Friedrichs: alpha per velocity row from the stage input's maxima
  float alf[3] = {0.f, 0.f, 0.f};
  if (RS == AMX_LAX_FRIEDRICHS) {
#pragma unroll
    for (int k = 0; k < DIM; ++k) alf[k] = __uint_as_float(wave_max_u(C->lfp[k][tid & 63]));
  }
  const bool owner = ta < TA;
  const int bid = swizzle_block(blockIdx.x, gridDim.x);
  const int itb = bid % a.ntb;
  const int rest = bid / a.ntb;
  const int ita = rest % a.nta;
  const int icx = rest / a.nta;
  const int a0 = ita * TA, b0 = itb * TB;
  // 32-bit cell indices (a slot has < 2^30 cells): global loads then take a
  // scalar base and a 32-bit vector offset
  const unsigned sa = a.nb, sx = (unsigned)a.na * a.nb;
  // plane position of the thread's column; fa / fb: reflective ghost in A / B
  bool fa = false, fb = false;
  const int ga = BCX ? bnd_index(a0 + ta, a.na, a.ba, fa) : wrapi(a0 + ta, a.na);
  const int gb = BCX ? bnd_index(b0 + tb, a.nb, a.bb, fb) : wrapi(b0 + tb, a.nb);
  const unsigned col = (unsigned)ga * sa + gb;
  const bool active = owner && (a0 + ta < a.na) && (b0 + tb < a.nb);
  const float coef = ph.inv_dx * (a.dt_scale * C->dt);
  // plane offset of x index xi (3D), and whether it is a reflective ghost
  auto xplane = [&](int xi, bool& f) -> unsigned {
    f = false;
    return (unsigned)(BCX ? bnd_index(xi, a.nx, a.bx, f) : wrapi(xi, a.nx)) * sx;
  };
  // x-axis factors (3D) of the thread's column at x index xi
  auto ppx = [&](int xi, float& fa_, float& fk_, float& fb_) {
    const unsigned cx = (unsigned)wrapi(xi, a.nx) * sx + col;
    fa_ = ld32(pfx[0], cx);
    fk_ = ld32(pfx[1], cx);
    fb_ = ld32(pfp[4], cx);
  };

  // this thread's halo cell (threads NT-NH .. NT-1), fixed for all planes
  const int h = tid - (NT - NH);
  int hr = 0, hc = 0;
  unsigned hoff = 0;
  bool hfa = false, hfb = false;
  if (h >= 0) {
    if (h < 4 * TB) {
      const int q = h / TB;
      hc = h % TB;
      hr = q < 2 ? q - 2 : TA + q - 2;
    } else {
      const int j = h - 4 * TB;
      hr = j >> 2;
      const int q = j & 3;
      hc = q < 2 ? q - 2 : TB + q - 2;
    }
    if (BCX)
      hoff = (unsigned)bnd_index(a0 + hr, a.na, a.ba, hfa) * sa + bnd_index(b0 + hc, a.nb, a.bb, hfb);
    else
      hoff = (unsigned)wrapi(a0 + hr, a.na) * sa + wrapi(b0 + hc, a.nb);
  }

  int x0 = 0, x1 = 1;
  if (DIM == 3) {
    x0 = icx * a.lx;
    x1 = min(x0 + a.lx, a.nx);
  }
  // Plane queues of the thread's column, as register rings that the x loop,
  // unrolled by four, renames instead of copying: q (4 planes: x .. x+2
  // derived, x+3 in flight), slopes d and x-fluxes F (2 each), the halo
  // cell's and the Heun centre's raw loads (2 each: plane x, plane x+1 in
  // flight).
  float qa[NV], qb[NV], qc[NV], qd[NV], da[NV], db[NV], Fa[NV], Fb[NV];
  if (owner) {
    if (DIM == 3) {
      float qm2[NV], qm1[NV], wl[NV], wr[NV];
      bool f2, f1, fp1;
      load_w<DIM, SRC>(src, xplane(x0 - 2, f2) + col, qm2);
      load_w<DIM, SRC>(src, xplane(x0 - 1, f1) + col, qm1);
      load_w<DIM, SRC>(src, (unsigned)x0 * sx + col, qa);
      load_w<DIM, SRC>(src, xplane(x0 + 1, fp1) + col, qb);
      if (BCX) {
        if (f2) qm2[1] = -qm2[1];
        if (f1) qm1[1] = -qm1[1];
        if (fp1) qb[1] = -qb[1];
      }
      float fa0 = 1.f, fk0 = 1.f, fb0 = 1.f, fam = 1.f, fkm = 1.f, fbm = 1.f;
      if (pp) {
        ppx(x0, fa0, fk0, fb0);
        ppx(x0 - 1, fam, fkm, fbm);
      }
#pragma unroll
      for (int v = 0; v < NV; ++v) {
        da[v] = slope<LIMX>(qm1[v], qa[v], qb[v], lim, ldx);
        float sm = slope<LIMX>(qm2[v], qm1[v], qa[v], lim, ldx);
        if (pp) {
          da[v] = da[v] * fsel(fa0, fk0, fb0, v);
          sm = sm * fsel(fam, fkm, fbm, v);
        }
        wl[v] = qm1[v] + sm;
        wr[v] = qa[v] - da[v];
      }
      riemann<DIM, 1, RS>(wl, wr, Fa, ph, alf[0], AMS && gsw[(unsigned)x0 * sx + col]);
    } else {
      load_w<DIM, SRC>(src, col, qa);
    }
  }

  // loads for plane x are issued one plane ahead of their use (raw fields in
  // registers across the previous plane's flux phase). SRC_PLAIN planes are
  // loaded straight into their q slot; the other forms are derived from a
  // raw prefetch (rq) when the plane becomes x+2.
  constexpr bool QRAW = SRC != SRC_PLAIN;
  Raw<DIM, SRC> rq, ha, hb;  // own column at x+2 (QRAW), halo cell at x / x+1
  Raw<DIM, W0M> ra, rb;      // Heun centre at x / x+1
  bool fq;  // unused: the flip of a prefetched plane is recomputed when it is derived
  if (owner && DIM == 3) {
    if (QRAW) {
      load_raw<DIM, SRC>(src, xplane(x0 + 2, fq) + col, rq);
    } else {
      load_w<DIM, SRC>(src, xplane(x0 + 2, fq) + col, qc);
    }
  }
  if (h >= 0) load_raw<DIM, SRC>(src, (DIM == 3 ? (unsigned)x0 * sx : 0u) + hoff, ha);
  // the 8-field SRC_MAGFIX centre would not be prefetched (register pressure):
  // the Heun stage after the magnetic sub-step takes the 5-field SRC_PFIX form
  constexpr bool PRE0 = HEUN && W0M != SRC_MAGFIX;
  // (LAG: plane x's centre is loaded in plane x, for its update in plane x + 1)
  if (!LAG && PRE0 && active) load_raw<DIM, W0M, true>(w0p, (DIM == 3 ? (unsigned)x0 * sx : 0u) + col, ra);

  float mx[3] = {0.f, 0.f, 0.f};
  // Update of the thread's cell in plane xu from its primitive state q, the
  // x-fluxes through its lower (Fl) and upper (Fr) faces, the plane fluxes
  // FAu / FBu and (Heun) the prefetched centre r0.
  auto update = [&](const int xu, const float (&q)[NV], const float (&Fl)[NV], const float (&Fr)[NV],
                    const Raw<DIM, W0M>& r0, const float (*FAu)[(TA + 1) * TB], const float (*FBu)[TA * (TB + 1)]) {
    const unsigned cell = (DIM == 3 ? (unsigned)xu * sx : 0u) + col;
    float U[NV], W1[NV];
    cons<DIM>(q, U, ph);
    if (DIM == 3) {
#pragma unroll
      for (int v = 0; v < NV; ++v) U[v] = U[v] + (Fl[v] - Fr[v]) * coef;
    }
#pragma unroll
    for (int v = 0; v < NV; ++v) U[v] = U[v] + (FAu[v][ta * TB + tb] - FAu[v][(ta + 1) * TB + tb]) * coef;
#pragma unroll
    for (int v = 0; v < NV; ++v) U[v] = U[v] + (FBu[v][ta * (TB + 1) + tb] - FBu[v][ta * (TB + 1) + tb + 1]) * coef;
    prim<DIM>(U, W1, ph);
    if (HEUN) {
      float W0[NV], U0[NV], U2[NV];
      if (PRE0) {
        derive_w<DIM, W0M>(r0, W0);
      } else {
        load_w<DIM, W0M>(w0p, cell, W0);
      }
      cons<DIM>(W0, U0, ph);
      cons<DIM>(W1, U2, ph);
#pragma unroll
      for (int v = 0; v < NV; ++v) U[v] = 0.5f * (U0[v] + U2[v]);
      prim<DIM>(U, W1, ph);
    }
#pragma unroll
    for (int v = 0; v < NV; ++v) st32nt(outp[v], cell, W1[v]);
    if (SRC == SRC_MAGFIX && a.pfix) st32nt(a.pfix + so, cell, q[P]);
    if (a.want_cfl) {
      const float cs = fsqrt(ph.gamma * W1[P] * frcp(W1[0]));
#pragma unroll
      for (int k = 0; k < DIM; ++k) {
        const float m = fabsf(W1[1 + k]) + cs;
        mx[k] = (m != m || m > mx[k]) ? m : mx[k];
      }
    }
    if (a.copy_b) {
#pragma unroll
      for (int k = 0; k < 3; ++k) st32(a.bdst[k] + so, cell, ld32(a.selb[par][k] + so, cell));
    }
  };
  // plane x: q0..q2 = planes x..x+2 (q2 loaded, or derived here from rq),
  // q3 receives plane x+3; d0 / Fp: slope and x-flux of plane x (from the
  // previous plane), d1 / Fn: those of plane x+1; rh0 / r00: halo and Heun
  // centre of plane x, rh1 / r01 receive plane x+1 (LAG: r00 receives plane
  // x, r01 holds plane x-1, whose cell update runs in this plane)
  auto plane_step = [&](const int x, float (&q0)[NV], float (&q1)[NV], float (&q2)[NV], float (&q3)[NV],
                        float (&d0)[NV], float (&d1)[NV], float (&Fp)[NV], float (&Fn)[NV], Raw<DIM, SRC>& rh0,
                        Raw<DIM, SRC>& rh1, Raw<DIM, W0M>& r00, Raw<DIM, W0M>& r01) {
    const unsigned plane = DIM == 3 ? (unsigned)x * sx : 0u;
    const unsigned nplane = DIM == 3 ? (unsigned)(x + 1) * sx : 0u;
    const bool more = x + 1 < x1;
    const int bb = LAG ? (x & 1) : 0;  // plane buffers of plane x
    float(*const W)[(TA + 4) * WW] = sW[bb];
    float(*const FA)[(TA + 1) * TB] = sFA[bb];
    float(*const FB)[TA * (TB + 1)] = sFB[bb];
    unsigned char* const G = sG[bb];
    // The loads issued one plane ago (q2, rh0, r00) have had the previous
    // plane's flux phase to land: wait for them here, before this plane's
    // prefetch is issued. Without it, a use of q2 / rh0 / r00 behind the new
    // loads would wait for those too (one in-order count of vector memory
    // operations), exposing their latency.
    if (DIM == 3) __builtin_amdgcn_s_waitcnt(VMCNT0);
    float hw[NV];
    if (owner && DIM == 3) {
      if (QRAW) derive_w<DIM, SRC>(rq, q2);
      if (BCX) {
        xplane(x + 2, fq);
        if (fq) q2[1] = -q2[1];
      }
    }
    if (h >= 0) {
      derive_w<DIM, SRC>(rh0, hw);
      if (BCX) {
        if (hfa) hw[AXA] = -hw[AXA];
        if (hfb) hw[AXB] = -hw[AXB];
      }
    }
    if (more) {
      if (owner && DIM == 3) {
        if (QRAW) {
          load_raw<DIM, SRC>(src, xplane(x + 3, fq) + col, rq);
        } else {
          load_w<DIM, SRC>(src, xplane(x + 3, fq) + col, q3);
        }
      }
      if (h >= 0) load_raw<DIM, SRC>(src, nplane + hoff, rh1);
      if (!LAG && PRE0 && active) load_raw<DIM, W0M, true>(w0p, nplane + col, r01);
    }
    if (LAG && PRE0 && active) load_raw<DIM, W0M, true>(w0p, plane + col, r00);
    // x-flux through the upper face of the plane's cell (face x + 1/2)
    auto xflux = [&]() {
      if (owner && DIM == 3) {
        float wl[NV], wr[NV];
        float fa1 = 1.f, fk1 = 1.f, fb1 = 1.f;
        if (pp) ppx(x + 1, fa1, fk1, fb1);
#pragma unroll
        for (int v = 0; v < NV; ++v) {
          d1[v] = slope<LIMX>(q0[v], q1[v], q2[v], lim, ldx);
          if (pp) d1[v] = d1[v] * fsel(fa1, fk1, fb1, v);
          wl[v] = q0[v] + d0[v];
          wr[v] = q1[v] - d1[v];
        }
        riemann<DIM, 1, RS>(wl, wr, Fn, ph, alf[0],
                            AMS && gsw[(unsigned)wrapi(x + 1, a.nx) * sx + col]);
      }
    };
    if (!LAG) xflux();
    if (owner) {
#pragma unroll
      for (int v = 0; v < NV; ++v) W[v][(ta + 2) * WW + tb + 2] = q0[v];
      if (BCX) {
        // a ghost column's cells enter the plane stencils with their flips
        if (fa) W[AXA][(ta + 2) * WW + tb + 2] = -q0[AXA];
        if (fb) W[AXB][(ta + 2) * WW + tb + 2] = -q0[AXB];
      }
      if (AMS) G[(ta + 2) * WW + tb + 2] = gsw[plane + col];
      if (pp) {
#pragma unroll
        for (int i = 0; i < 5; ++i) sPF[i][(ta + 2) * WW + tb + 2] = ld32(pfp[i], plane + col);
      }
    }
    if (h >= 0) {
#pragma unroll
      for (int v = 0; v < NV; ++v) W[v][(hr + 2) * WW + hc + 2] = hw[v];
      if (AMS) G[(hr + 2) * WW + hc + 2] = gsw[plane + hoff];
      if (pp) {
#pragma unroll
        for (int i = 0; i < 5; ++i) sPF[i][(hr + 2) * WW + hc + 2] = ld32(pfp[i], plane + hoff);
      }
    }
    __syncthreads();
    // interface fluxes of the plane; limited slopes fused in (4-point
    // stencils). (TA+1)*64 A-interfaces == one per thread; the TA*65
    // B-interfaces take a second, partial round.
    {
      float wl[NV], wr[NV], f[NV];
      const int r = tid / TB, c = tid % TB;  // interface between rows r-1 and r
      const int o = (r + 2) * WW + c + 2;
#pragma unroll
      for (int v = 0; v < NV; ++v) {
        const float w0 = W[v][o - 2 * WW], w1 = W[v][o - WW], w2 = W[v][o], w3 = W[v][o + WW];
        float s1 = slope<LIMX>(w0, w1, w2, lim, ldx), s2 = slope<LIMX>(w1, w2, w3, lim, ldx);
        if (pp) {
          s1 = s1 * fsel(sPF[0][o - WW], sPF[1][o - WW], sPF[4][o - WW], v);
          s2 = s2 * fsel(sPF[0][o], sPF[1][o], sPF[4][o], v);
        }
        wl[v] = w1 + s1;
        wr[v] = w2 - s2;
      }
      riemann<DIM, AXA, RS>(wl, wr, f, ph, alf[AXA - 1], AMS && G[o]);
#pragma unroll
      for (int v = 0; v < NV; ++v) FA[v][tid] = f[v];
    }
    if (tid < NF - NFA) {
      float wl[NV], wr[NV], f[NV];
      const int r = tid / (TB + 1), c = tid % (TB + 1);  // interface between columns c-1 and c
      const int o = (r + 2) * WW + c + 2;
#pragma unroll
      for (int v = 0; v < NV; ++v) {
        const float w0 = W[v][o - 2], w1 = W[v][o - 1], w2 = W[v][o], w3 = W[v][o + 1];
        float s1 = slope<LIMX>(w0, w1, w2, lim, ldx), s2 = slope<LIMX>(w1, w2, w3, lim, ldx);
        if (pp) {
          s1 = s1 * fsel(sPF[2][o - 1], sPF[3][o - 1], sPF[4][o - 1], v);
          s2 = s2 * fsel(sPF[2][o], sPF[3][o], sPF[4][o], v);
        }
        wl[v] = w1 + s1;
        wr[v] = w2 - s2;
      }
      riemann<DIM, AXB, RS>(wl, wr, f, ph, alf[AXB - 1], AMS && G[o]);
#pragma unroll
      for (int v = 0; v < NV; ++v) FB[v][tid] = f[v];
    }
    if constexpr (LAG) {
      // the cell of plane x - 1: its plane fluxes were written before this
      // plane's barrier; its state is the thread's own entry of that plane's
      // buffer, and its faces are x - 3/2 (still in Fn) and x - 1/2 (Fp)
      if (x > x0 && active) {
        float qm[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) qm[v] = sW[bb ^ 1][v][(ta + 2) * WW + tb + 2];
        update(x - 1, qm, Fn, Fp, r01, sFA[bb ^ 1], sFB[bb ^ 1]);
      }
      xflux();
      return;
    }
    __syncthreads();
    if (active) update(x, q0, Fp, Fn, r00, FA, FB);
  };
  if (DIM == 3) {
    for (int x = x0; x < x1; x += 4) {
      plane_step(x, qa, qb, qc, qd, da, db, Fa, Fb, ha, hb, ra, rb);
      if (x + 1 < x1) plane_step(x + 1, qb, qc, qd, qa, db, da, Fb, Fa, hb, ha, rb, ra);
      if (x + 2 < x1) plane_step(x + 2, qc, qd, qa, qb, da, db, Fa, Fb, ha, hb, ra, rb);
      if (x + 3 < x1) plane_step(x + 3, qd, qa, qb, qc, db, da, Fb, Fa, hb, ha, rb, ra);
    }
    if constexpr (LAG) {
      // the update of the last plane
      __syncthreads();
      if (active) {
        const int xl = x1 - 1, bl = xl & 1;
        float ql[NV];
#pragma unroll
        for (int v = 0; v < NV; ++v) ql[v] = sW[bl][v][(ta + 2) * WW + tb + 2];
        if (((xl - x0) & 1) == 0)
          update(xl, ql, Fa, Fb, ra, sFA[bl], sFB[bl]);
        else
          update(xl, ql, Fb, Fa, rb, sFA[bl], sFB[bl]);
      }
    }
  } else {
    plane_step(0, qa, qb, qc, qd, da, db, Fa, Fb, ha, hb, ra, rb);
  }
  if (a.want_cfl) {
    __shared__ unsigned sred[DIM][TA + 1];
#pragma unroll
    for (int k = 0; k < DIM; ++k) {
      const unsigned m = block_max_u<TA + 1>(canon_bits(mx[k]), sred[k]);
      if (tid == 0 && m) atomicMax(&C->cflp[k][blockIdx.x & (NSLOT - 1)], m);
    }
  }
}

// Lax-Friedrichs alpha of a stage input (_lax_friedrichs.py:61-66): per
// velocity row k, the max over all cells of |u_k| + sqrt(gamma p / rho),
// reduced into the slot's lfp[k] partials (block id mod 64) and read by
// every block of the following stage kernel. k_lf_clear zeroes them first.
constexpr int LF_BLOCKS = 512;

__global__ void k_lf_clear(Ctl* ctl, int slot0, int gated) {
  Ctl* const c = ctl + slot0 + blockIdx.x;
  if (gated && c->held) return;  // the stage it prepares is held too
#pragma unroll
  for (int k = 0; k < 3; ++k) c->lfp[k][threadIdx.x] = 0u;
}

template <int DIM, int SRC>
__global__ __launch_bounds__(256) void k_lf_alpha(const StageArgs a) {
  constexpr int P = DIM + 1;
  __shared__ unsigned sred[4];
  const int slot = a.slot0 + blockIdx.y;
  Ctl* const C = a.ctl + slot;
  // block-uniform; a held (speculatively queued) stage's pre-pass waits for
  // the launch that follows the further fixed-point iterates
  if (!C->run || (a.gated && C->held)) return;
  const long so = slot * a.sstride;
  const int par = iterate_set(C->bset, C->iters);
  const float* src[8];
#pragma unroll
  for (int v = 0; v < 8; ++v) src[v] = a.src[v] ? a.src[v] + so : nullptr;
  if (SRC == SRC_MAGFIX) {
#pragma unroll
    for (int k = 0; k < 3; ++k) src[5 + k] = a.selv[par][k] + so;
  }
  const Phys ph = a.phs[slot];
  const unsigned n = (unsigned)a.nx * (unsigned)a.na * (unsigned)a.nb;
  float mx[3] = {0.f, 0.f, 0.f};
  for (unsigned c = blockIdx.x * 256u + threadIdx.x; c < n; c += gridDim.x * 256u) {
    float w[DIM + 2];
    load_w<DIM, SRC>(src, c, w);
    const float cs = fsqrt(ph.gamma * w[P] * frcp(w[0]));
#pragma unroll
    for (int k = 0; k < DIM; ++k) {
      const float m = fabsf(w[1 + k]) + cs;
      mx[k] = (m != m || m > mx[k]) ? m : mx[k];
    }
  }
#pragma unroll
  for (int k = 0; k < DIM; ++k) {
    const unsigned b = block_max_u<4>(canon_bits(mx[k]), sred);
    if (threadIdx.x == 0 && b) atomicMax(&C->lfp[k][blockIdx.x & (NSLOT - 1)], b);
    __syncthreads();  // sred is reused by the next axis
  }
}

// Compression switch of the hybrid / AM-HLLC solvers (hll.py:281-297): per
// cell of the stage input, g = 1 where div v < -0.05 a / dx with the central
// velocity differences (v[i+1] - v[i-1]) / (2 dx) summed over the axes and a
// the sound speed, else 0.
template <int DIM, int SRC>
__global__ __launch_bounds__(256) void k_am_switch(const StageArgs a) {
  constexpr int P = DIM + 1;
  const int slot = a.slot0 + blockIdx.y;
  Ctl* const C = a.ctl + slot;
  // block-uniform; a held (speculatively queued) stage's pre-pass waits for
  // the launch that follows the further fixed-point iterates
  if (!C->run || (a.gated && C->held)) return;
  const long so = slot * a.sstride;
  const int par = iterate_set(C->bset, C->iters);
  const float* src[8];
#pragma unroll
  for (int v = 0; v < 8; ++v) src[v] = a.src[v] ? a.src[v] + so : nullptr;
  if (SRC == SRC_MAGFIX) {
#pragma unroll
    for (int k = 0; k < 3; ++k) src[5 + k] = a.selv[par][k] + so;
  }
  // velocity rows of the stage input: the converged magnetic iterate's in MAGFIX
  const float* vel[3];
#pragma unroll
  for (int k = 0; k < DIM; ++k) vel[k] = SRC == SRC_MAGFIX ? src[5 + k] : src[1 + k];
  unsigned char* g = a.amg + slot * a.gstride;
  const Phys ph = a.phs[slot];
  const float den = 2.f * a.dxf;  // float32(2 dx): 2 * d in the reference
  const float thr_scale = -0.05f;
  // axes of the state: 3D (nx, na, nb) = (x, y, z); 2D (na, nb) = (x, y)
  const int n0 = DIM == 3 ? a.nx : a.na, n1 = DIM == 3 ? a.na : a.nb, n2 = DIM == 3 ? a.nb : 1;
  const unsigned n = (unsigned)n0 * (unsigned)n1 * (unsigned)n2;
  const unsigned s0 = (unsigned)n1 * (unsigned)n2, s1 = (unsigned)n2;
  for (unsigned c = blockIdx.x * 256u + threadIdx.x; c < n; c += gridDim.x * 256u) {
    const int i2 = (int)(c % (unsigned)n2);
    const unsigned r = c / (unsigned)n2;
    const int i1 = (int)(r % (unsigned)n1);
    const int i0 = (int)(r / (unsigned)n1);
    float w[DIM + 2];
    load_w<DIM, SRC>(src, c, w);
    const float cs = sqrtf((ph.gamma * w[P]) / w[0]);
    const unsigned b0 = c - (unsigned)i0 * s0, b1 = c - (unsigned)i1 * s1, b2 = c - (unsigned)i2;
    float div = ((ld32(vel[0], b0 + (unsigned)(i0 + 1 == n0 ? 0 : i0 + 1) * s0) -
                  ld32(vel[0], b0 + (unsigned)(i0 == 0 ? n0 - 1 : i0 - 1) * s0)) / den);
    div = div + (ld32(vel[1], b1 + (unsigned)(i1 + 1 == n1 ? 0 : i1 + 1) * s1) -
                 ld32(vel[1], b1 + (unsigned)(i1 == 0 ? n1 - 1 : i1 - 1) * s1)) / den;
    if (DIM == 3)
      div = div + (ld32(vel[2], b2 + (unsigned)(i2 + 1 == n2 ? 0 : i2 + 1)) -
                   ld32(vel[2], b2 + (unsigned)(i2 == 0 ? n2 - 1 : i2 - 1))) / den;
    g[c] = div < (thr_scale * cs) / a.dxf ? 1 : 0;
  }
}

// VAN_ALBADA_PP factors of a stage input (reconstruction.py:168-365), one
// thread per cell: the van Albada differences d[a][v] of the cell along
// every axis, then alpha[a] = min(rho / |d[a][rho]|, 1) and kappa[a] (the
// same for p) where |d| > 1e-14, and beta, the velocity scaling, from
// A1 = sum_k (sum_a C[k] alpha[k] d[k][rho] d[a][v_k])^2 (the reference's
// broadcast puts the weights on the component index), A2 = sum_a C[a]
// sum_k d[a][v_k]^2 and q: beta = min(sqrt((q-2)^2 rho p / ((gamma-1)
// (2 A1 + (q-2) rho^2 A2))), 1) where sum_a sum_k d[a][v_k]^2 > 1e-14.
// C = alpha_lax / sum(alpha_lax) from the Lax-Friedrichs maxima of the same
// input (k_lf_alpha). (1 + 1e-14) is 1 in float32. Periodic grids.
template <int DIM, int SRC>
__global__ __launch_bounds__(256) void k_pp_factors(const StageArgs a) {
  constexpr int NV = DIM + 2, P = DIM + 1;
  const int slot = a.slot0 + blockIdx.y;
  Ctl* const C = a.ctl + slot;
  // block-uniform; a held (speculatively queued) stage's pre-pass waits for
  // the launch that follows the further fixed-point iterates
  if (!C->run || (a.gated && C->held)) return;
  const long so = slot * a.sstride;
  const int par = iterate_set(C->bset, C->iters);
  const float* src[8];
#pragma unroll
  for (int v = 0; v < 8; ++v) src[v] = a.src[v] ? a.src[v] + so : nullptr;
  if (SRC == SRC_MAGFIX) {
#pragma unroll
    for (int k = 0; k < 3; ++k) src[5 + k] = a.selv[par][k] + so;
  }
  float* out[2 * DIM + 1];
#pragma unroll
  for (int i = 0; i < 2 * DIM + 1; ++i) out[i] = a.ppf[i] + slot * a.ppstride;
  const Phys ph = a.phs[slot];
  float Cw[3];
  {
    float al[3] = {0.f, 0.f, 0.f};
#pragma unroll
    for (int k = 0; k < DIM; ++k) al[k] = __uint_as_float(wave_max_u(C->lfp[k][threadIdx.x & 63]));
    float sum = al[0] + al[1];
    if (DIM == 3) sum = sum + al[2];
#pragma unroll
    for (int k = 0; k < DIM; ++k) Cw[k] = al[k] / sum;
  }
  const float qm2 = a.ppq - 2.f;
  const float eps = 1e-14f, dx = a.dxf;
  // axes of the state: 3D (nx, na, nb) = (x, y, z); 2D (na, nb) = (x, y)
  const int n0 = DIM == 3 ? a.nx : a.na, n1 = DIM == 3 ? a.na : a.nb, n2 = DIM == 3 ? a.nb : 1;
  const unsigned n = (unsigned)n0 * (unsigned)n1 * (unsigned)n2;
  const unsigned s0 = (unsigned)n1 * (unsigned)n2, s1 = (unsigned)n2;
  for (unsigned c = blockIdx.x * 256u + threadIdx.x; c < n; c += gridDim.x * 256u) {
    const int i2 = (int)(c % (unsigned)n2);
    const unsigned r = c / (unsigned)n2;
    const int i1 = (int)(r % (unsigned)n1);
    const int i0 = (int)(r / (unsigned)n1);
    const int ii[3] = {i0, i1, i2};
    const int nn[3] = {n0, n1, n2};
    const unsigned ss[3] = {s0, s1, 1u};
    float w[NV];
    load_w<DIM, SRC>(src, c, w);
    float d[DIM][NV];
#pragma unroll
    for (int ax = 0; ax < DIM; ++ax) {
      const unsigned base = c - (unsigned)ii[ax] * ss[ax];
      float wm[NV], wp[NV];
      load_w<DIM, SRC>(src, base + (unsigned)(ii[ax] == 0 ? nn[ax] - 1 : ii[ax] - 1) * ss[ax], wm);
      load_w<DIM, SRC>(src, base + (unsigned)(ii[ax] + 1 == nn[ax] ? 0 : ii[ax] + 1) * ss[ax], wp);
#pragma unroll
      for (int v = 0; v < NV; ++v) d[ax][v] = half_slope_x(wm[v], w[v], wp[v], LIM_VAN_ALBADA, dx);
    }
    const float rho = w[0], p = w[P];
    float al[DIM], ka[DIM];
#pragma unroll
    for (int ax = 0; ax < DIM; ++ax) {
      al[ax] = fabsf(d[ax][0]) > eps ? fminf(rho / fabsf(d[ax][0]), 1.f) : 1.f;
      ka[ax] = fabsf(d[ax][P]) > eps ? fminf(p / fabsf(d[ax][P]), 1.f) : 1.f;
    }
    float A1 = 0.f, A2 = 0.f, vsum = 0.f;
#pragma unroll
    for (int k = 0; k < DIM; ++k) {
      const float wk = (Cw[k] * al[k]) * d[k][0];
      float sk = wk * d[0][1 + k];
#pragma unroll
      for (int ax = 1; ax < DIM; ++ax) sk = sk + wk * d[ax][1 + k];
      A1 = k == 0 ? sk * sk : A1 + sk * sk;
    }
#pragma unroll
    for (int ax = 0; ax < DIM; ++ax) {
      float vs = d[ax][1] * d[ax][1];
#pragma unroll
      for (int k = 1; k < DIM; ++k) vs = vs + d[ax][1 + k] * d[ax][1 + k];
      A2 = ax == 0 ? Cw[ax] * vs : A2 + Cw[ax] * vs;
      vsum = ax == 0 ? vs : vsum + vs;
    }
    float beta = 1.f;
    if (vsum > eps) {
      const float num = ((qm2 * qm2) * rho) * p;
      const float den = (ph.gamma - 1.f) * (2.f * A1 + (qm2 * (rho * rho)) * A2);
      beta = fminf(sqrtf(num / den), 1.f);
    }
#pragma unroll
    for (int ax = 0; ax < DIM; ++ax) {
      st32(out[ax], c, al[ax]);
      st32(out[DIM + ax], c, ka[ax]);
    }
    st32(out[2 * DIM], c, beta);
  }
}

struct MagArgs {
  const float* B0[3];
  const float* v0[3];  // v0[2] == nullptr in 2D (vz = 0)
  const float* rho;
  float* PS[4][6];     // iterate sets (B rows 0..2, v rows 3..5); set 3 may be absent
  // slot-local pointers, filled on the device from PS and ctl->bset:
  const float* Bk[3];  // input iterate: output of launch L-1 (the state's B, v for L = 0)
  const float* vk[3];  // vk[2] == nullptr: vz = 0
  float* Bo[2][3];     // outputs of launches L and L+1
  float* vo[2][3];
  Ctl* ctl;            // per-slot control, slot = slot0 + blockIdx.y
  long sstride;        // floats between consecutive slots' buffers
  int slot0;
  int L;               // first launch index of the pair (even)
  int write_first;     // store the pair's first iterate (0: only where L is the slot's kprev)
  int regen;           // regeneration launch: see k_mag
  unsigned skip_mask;  // regeneration: speculative pairs (bit L / 2), whose first iterate depends on kprev
  int max_iter;
  float tol;
  int euler;           // IMPLICIT_EULER: average = iterate
  int nx, na, nb, ntb, nta, ncx, lx;
  float inv_dx;
  AxisBc bx, ba, bb;    // boundaries of the kernel axes x (3D), A and B (kernels with BCX)
};

// The fixed point's stopping rule on an executed pair (L, L+1) with residual
// maxima r0 = res_L, r1 = res_{L+1} (res_k = max(max|B_k - B_{k-1}|,
// max|v_k - v_{k-1}|), launch k making iterate k+1): launch k runs while
// res_{k-1} > tol and k-1 < max_iter (_magnetic_field_update.py:127-141).
// Returns the last launch that the reference would run (L or L+1), or -1 when
// launch L+2 would still run.
__host__ __device__ inline int pair_final(int L, float r0, float r1, float tol, int max_iter) {
  if (!(r0 > tol) || L >= max_iter) return L;
  if (!(r1 > tol) || L + 1 >= max_iter) return L + 1;
  return -1;
}

// Raw iterate and fixed-point inputs of one cell.
struct MagRaw {
  float Bk[3], vk[3], B0[3], v0[3];
};

// FIRST: the pair of launch 0, whose input iterate is (B0, v0) itself: the
// inputs are loaded once and stand for the iterate too. VZ: both z velocity
// rows exist (3D); otherwise a missing row (the state's in 2D) reads as 0.
template <bool FIRST, bool VZ>
__device__ __forceinline__ void mag_load(const MagArgs& a, unsigned c, MagRaw& m) {
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    m.B0[i] = ld32(a.B0[i], c);
    m.Bk[i] = FIRST ? m.B0[i] : ld32(a.Bk[i], c);
  }
  m.v0[0] = ld32(a.v0[0], c);
  m.v0[1] = ld32(a.v0[1], c);
  m.v0[2] = (VZ || a.v0[2]) ? ld32(a.v0[2], c) : 0.f;
  if (FIRST) {
#pragma unroll
    for (int i = 0; i < 3; ++i) m.vk[i] = m.v0[i];
  } else {
    m.vk[0] = ld32(a.vk[0], c);
    m.vk[1] = ld32(a.vk[1], c);
    m.vk[2] = (VZ || a.vk[2]) ? ld32(a.vk[2], c) : 0.f;
  }
}

// (avgB, E = avgB x avgv) of iterate (Bk, vk) (_magnetic_field_update.py:84,149-154; _vector_maths.py:5-10)
__device__ __forceinline__ void mag_derive3(const float* Bk, const float* vk, const float* B0, const float* v0, int euler,
                                            float* d) {
  float b[3], v[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    b[i] = euler ? Bk[i] : (Bk[i] + B0[i]) * 0.5f;
    v[i] = euler ? vk[i] : (vk[i] + v0[i]) * 0.5f;
  }
  d[0] = b[0];
  d[1] = b[1];
  d[2] = b[2];
  d[3] = b[1] * v[2] - b[2] * v[1];
  d[4] = b[2] * v[0] - b[0] * v[2];
  d[5] = b[0] * v[1] - b[1] * v[0];
}

__device__ __forceinline__ void mag_derive(const MagRaw& m, int euler, float* d) {
  mag_derive3(m.Bk, m.vk, m.B0, m.v0, euler, d);
}

// Row `row` of B and v negated: a reflective ghost of the 3-row arrays.
__device__ __forceinline__ void mag_flip(MagRaw& m, int row) {
  m.B0[row] = -m.B0[row];
  m.Bk[row] = -m.Bk[row];
  m.v0[row] = -m.v0[row];
  m.vk[row] = -m.vk[row];
}

// (avgB, E) of the ghost copy of a cell: with row f of B and v negated, avgB_f
// and E_k for k != f change sign (each term of E_k, k != f, has one factor of
// row f); f < 0 is an open ghost, a plain copy.
__device__ __forceinline__ void derived_ghost(const float* d, int f, float* o) {
#pragma unroll
  for (int i = 0; i < 6; ++i) o[i] = d[i];
  if (f >= 0) {
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      if (k == f) o[k] = -d[k];
      else o[3 + k] = -d[3 + k];
    }
  }
}

// central derivative 0.5 (f+ - f-) / dx (_vector_maths.py:58-100)
__device__ __forceinline__ float cdiff(float fp, float fm, float inv_dx) {
  return (0.5f * (fp - fm)) * inv_dx;
}

// curl(E) and curl(avgB) at LDS offset o of a plane (row stride W) from the
// derived fields S[0..5] = (avgB, E); x-neighbours dm, dp (3D)
template <int DIM, int W>
__device__ __forceinline__ void mag_curls(const float* const* S, int o, const float* dm, const float* dp, float idx,
                                          float* cE, float* cB) {
  if (DIM == 3) {
    cE[0] = cdiff(S[5][o + W], S[5][o - W], idx) - cdiff(S[4][o + 1], S[4][o - 1], idx);
    cE[1] = cdiff(S[3][o + 1], S[3][o - 1], idx) - cdiff(dp[5], dm[5], idx);
    cE[2] = cdiff(dp[4], dm[4], idx) - cdiff(S[3][o + W], S[3][o - W], idx);
    cB[0] = cdiff(S[2][o + W], S[2][o - W], idx) - cdiff(S[1][o + 1], S[1][o - 1], idx);
    cB[1] = cdiff(S[0][o + 1], S[0][o - 1], idx) - cdiff(dp[2], dm[2], idx);
    cB[2] = cdiff(dp[1], dm[1], idx) - cdiff(S[0][o + W], S[0][o - W], idx);
  } else {
    cE[0] = cdiff(S[5][o + 1], S[5][o - 1], idx);
    cE[1] = -cdiff(S[5][o + W], S[5][o - W], idx);
    cE[2] = cdiff(S[4][o + W], S[4][o - W], idx) - cdiff(S[3][o + 1], S[3][o - 1], idx);
    cB[0] = cdiff(S[2][o + 1], S[2][o - 1], idx);
    cB[1] = -cdiff(S[2][o + W], S[2][o - W], idx);
    cB[2] = cdiff(S[1][o + W], S[1][o - W], idx) - cdiff(S[0][o + 1], S[0][o - 1], idx);
  }
}

// One iterate: B' = B0 - dt curl E, v' = v0 - dt (avgB x curl avgB) / rho
__device__ __forceinline__ void mag_apply(const float* avgB, const float* cE, const float* cB, const float* B0,
                                          const float* v0, float rho, float dt, float* Bn, float* vn) {
  const float ri = frcp(rho);
  const float p2[3] = {(avgB[1] * cB[2] - avgB[2] * cB[1]) * ri, (avgB[2] * cB[0] - avgB[0] * cB[2]) * ri,
                       (avgB[0] * cB[1] - avgB[1] * cB[0]) * ri};
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    Bn[i] = B0[i] - dt * cE[i];
    vn[i] = v0[i] - dt * p2[i];
  }
}

__device__ __forceinline__ float mag_resid(const float* Bk, const float* vk, const float* Bn, const float* vn) {
  float nb = 0.f, nv = 0.f;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const float db = Bk[i] - Bn[i], dv = vk[i] - vn[i];
    nb += db * db;
    nv += dv * dv;
  }
  return fmaxf(fsqrt(nb), fsqrt(nv));
}

// Two fixed-point launches (L, L+1) of the implicit-midpoint magnetic
// sub-step in one pass (_magnetic_field_update.py:82-165): iterate L+1 from
// iterate L on the tile and a one-cell ring, then iterate L+2 on the tile
// from those, so the pair reads the fixed-point inputs (B0, v0, rho) and the
// input iterate once. The block covers the tile with a two-cell ring; it
// marches along x with the second iterate one plane behind the first. The
// ring cells of the first iterate are recomputed, bit-identically, by the
// neighbouring blocks that own them. Both iterates are written; the residuals
// go to res slots L and L+1. Launch L+1's iterate is speculative: whether it
// is the reference's result is decided by pair_final.
//
// BCX: open / reflective boundaries. The reference refreshes the ghost cells
// of B and v before the first iterate and after each one
// (_magnetic_field_update.py:108-110,123-125,161-163), negating row axis + 1
// at reflective walls (By, vy at x; Bz, vz at y; none at z). Loads of ghost
// cells take the mapped cell with that row negated; the first iterate of the
// ghost cells next to the edge, which the second iterate reads, is the
// refreshed copy of the edge cell's first iterate, written into the ring by
// the edge cell's thread (the ghost's own thread, whose stencil would reach
// past the refreshed layer, writes nothing).
//
// FIRST: instantiated for launch 0 only (not for regeneration launches).
template <int DIM, int TA, int TBM, bool BCX, bool FIRST>
__global__ __launch_bounds__(((TA + 4) * (TBM + 4) + 255) / 256 * 256) void k_mag(const MagArgs a) {
  constexpr int WX = TBM + 4, NC = (TA + 4) * WX;   // ring-2 tile
  constexpr int W1 = TBM + 2, NC1 = (TA + 2) * W1;  // ring-1 tile
  constexpr int NT = (NC + 255) / 256 * 256, NW = NT / 64;
  const int tid = threadIdx.x;
  const int slot = a.slot0 + blockIdx.y;
  Ctl* const C = a.ctl + slot;
  if (!C->run) return;
  int L = a.L;
  bool one = a.regen;  // the launch computes its first iterate only
  if (a.regen) {
    // Regeneration: the fixed point's final launch is read from the residuals
    // of the last executed pair; if it is that pair's first launch and the
    // pair did not store its iterate (a speculative pair not starting at the
    // previous step's final index), the launch is computed again (the same
    // arithmetic on the same inputs, so bit-identical) and only that iterate
    // is stored. Residuals, counters and flags are left as they are.
    const int Ll = C->iters - 1;
    if (Ll < 0 || Ll > 62) return;
    const float r0 = __uint_as_float(wave_max_u(C->resp[Ll % NRES][tid & 63]));
    const float r1 = __uint_as_float(wave_max_u(C->resp[(Ll + 1) % NRES][tid & 63]));
    if (pair_final(Ll, r0, r1, a.tol, a.max_iter) != Ll || !((a.skip_mask >> (Ll >> 1)) & 1u) || Ll == C->kprev)
      return;
    L = Ll;
  } else if (L > 0) {
    if (C->done) return;
    const float r0 = __uint_as_float(wave_max_u(C->resp[(L - 2) % NRES][tid & 63]));
    const float r1 = __uint_as_float(wave_max_u(C->resp[(L - 1) % NRES][tid & 63]));
    if (pair_final(L - 2, r0, r1, a.tol, a.max_iter) >= 0) {
      if (blockIdx.x == 0 && tid == 0) C->done = 1;
      return;
    }
  }
  __shared__ float sK[2][6][NC];
  __shared__ float sK1[2][6][NC1];
  __shared__ unsigned sred[2][NW];
  if (!a.regen && blockIdx.x == 0 && tid < NSLOT) {
    if (tid == 0) C->iters = L + 1;
    C->resp[(L + 2) % NRES][tid] = 0u;
    C->resp[(L + 3) % NRES][tid] = 0u;
  }
  const bool store1 = a.write_first || one || L == C->kprev;
  const long so = slot * a.sstride;
  MagArgs m = a;  // slot-local pointers
  // B0 is in the state's rows or, in a loop that keeps B in the iterate
  // sets, in set bset; launch k writes set iterate_set(bset, k)
  const int bs = C->bset;
  const int sin_ = L > 0 ? iterate_set(bs, L - 1) : 0;
  const int so0 = iterate_set(bs, L), so1 = iterate_set(bs, L + 1);
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    m.B0[i] = (bs < 0 ? a.B0[i] : a.PS[bs][i]) + so;
    m.v0[i] = a.v0[i] ? a.v0[i] + so : nullptr;
    m.Bk[i] = L > 0 ? a.PS[sin_][i] + so : m.B0[i];
    m.vk[i] = L > 0 ? a.PS[sin_][3 + i] + so : m.v0[i];
    m.Bo[0][i] = a.PS[so0][i] + so;
    m.vo[0][i] = a.PS[so0][3 + i] + so;
    m.Bo[1][i] = a.PS[so1][i] + so;
    m.vo[1][i] = a.PS[so1][3 + i] + so;
  }
  m.rho = a.rho + so;
  const int bid = swizzle_block(blockIdx.x, gridDim.x);
  const int itb = bid % a.ntb;
  const int rest = bid / a.ntb;
  const int ita = rest % a.nta;
  const int icx = rest / a.nta;
  const int a0 = ita * TA, b0 = itb * TBM;
  // 32-bit cell offsets (a slot has < 2^30 cells)
  const unsigned sa = a.nb, sx = (unsigned)a.na * a.nb;
  const bool has = tid < NC;
  const int r = has ? tid / WX - 2 : 0, c = has ? tid % WX - 2 : 0;
  const bool ring2 = r == -2 || r == TA + 1 || c == -2 || c == TBM + 1;
  const bool inner = has && !ring2;                       // first iterate here
  const bool tile = inner && r >= 0 && r < TA && c >= 0 && c < TBM;
  const bool own = tile && (a0 + r < a.na) && (b0 + c < a.nb);
  // rows of B and v negated at reflective walls of the kernel axes
  constexpr int RX = 1, RA = DIM == 3 ? 2 : 1, RB = DIM == 3 ? -1 : 2;
  bool fa = false, fb = false;
  const unsigned col = BCX ? (unsigned)bnd_index(a0 + r, a.na, a.ba, fa) * sa + bnd_index(b0 + c, a.nb, a.bb, fb)
                           : (unsigned)wrapi(a0 + r, a.na) * sa + wrapi(b0 + c, a.nb);
  const int o = (r + 2) * WX + (c + 2);
  const int o1 = (r + 1) * W1 + (c + 1);
  // ghost cell of the ring-1 tile on an open / reflective side
  const bool gh = BCX && ((a0 + r < 0 && a.ba.lo != AMX_BC_PERIODIC) || (a0 + r >= a.na && a.ba.hi != AMX_BC_PERIODIC) ||
                          (b0 + c < 0 && a.bb.lo != AMX_BC_PERIODIC) || (b0 + c >= a.nb && a.bb.hi != AMX_BC_PERIODIC));
  // flips of a load of the thread's column at an x plane that is (fx) or is
  // not a reflective ghost
  auto flips = [&](MagRaw& t, bool fx) {
    if (BCX) {
      if (fx) mag_flip(t, RX);
      if (fa) mag_flip(t, RA);
      if constexpr (RB >= 0) {
        if (fb) mag_flip(t, RB);
      }
    }
  };
  auto xoff = [&](int xi, bool& f) -> unsigned {
    f = false;
    return (unsigned)(BCX ? bnd_index(xi, a.nx, a.bx, f) : wrapi(xi, a.nx)) * sx;
  };
  // an edge cell's first iterate e1 copied into the ghost cells beside it
  auto ghost_copies = [&](float (*S1)[NC1], const float* e1) {
    float g[6];
    auto put = [&](int off, int code, int row) {
      derived_ghost(e1, code == AMX_BC_REFLECTIVE ? row : -1, g);
#pragma unroll
      for (int i = 0; i < 6; ++i) S1[i][o1 + off] = g[i];
    };
    if (a0 + r == 0 && a.ba.lo != AMX_BC_PERIODIC) put(-W1, a.ba.lo, RA);
    if (a0 + r == a.na - 1 && a.ba.hi != AMX_BC_PERIODIC) put(W1, a.ba.hi, RA);
    if (b0 + c == 0 && a.bb.lo != AMX_BC_PERIODIC) put(-1, a.bb.lo, RB);
    if (b0 + c == a.nb - 1 && a.bb.hi != AMX_BC_PERIODIC) put(1, a.bb.hi, RB);
  };
  const float dt = C->dt;
  const float idx = a.inv_dx;
  float res0 = 0.f, res1 = 0.f;
  const float* SK[6];
  const float* SK1[6];

  if (DIM == 2) {
    MagRaw rc;
    float d[6], rho = 1.f;
    if (has) {
      mag_load<FIRST, DIM == 3>(m, col, rc);
      flips(rc, false);
      mag_derive(rc, a.euler, d);
#pragma unroll
      for (int i = 0; i < 6; ++i) sK[0][i][o] = d[i];
    }
    if (inner) rho = ld32(m.rho, col);
    __syncthreads();
    float B1[3], v1[3];
    if (inner) {
      float cE[3], cB[3], e1[6];
#pragma unroll
      for (int i = 0; i < 6; ++i) SK[i] = sK[0][i];
      mag_curls<2, WX>(SK, o, nullptr, nullptr, idx, cE, cB);
      mag_apply(d, cE, cB, rc.B0, rc.v0, rho, dt, B1, v1);
      if (own) {
        if (store1) {
#pragma unroll
          for (int i = 0; i < 3; ++i) {
            st32(m.Bo[0][i], col, B1[i]);
            st32(m.vo[0][i], col, v1[i]);
          }
        }
        const float rr = mag_resid(rc.Bk, rc.vk, B1, v1);
        res0 = (rr != rr || rr > res0) ? rr : res0;
      }
      mag_derive3(B1, v1, rc.B0, rc.v0, a.euler, e1);
      if (!gh) {
#pragma unroll
        for (int i = 0; i < 6; ++i) sK1[0][i][o1] = e1[i];
      }
      if (BCX && own) ghost_copies(sK1[0], e1);
    }
    __syncthreads();
    if (own && !one) {
      float cE[3], cB[3], e1[6], B2[3], v2[3];
#pragma unroll
      for (int i = 0; i < 6; ++i) {
        SK1[i] = sK1[0][i];
        e1[i] = sK1[0][i][o1];
      }
      mag_curls<2, W1>(SK1, o1, nullptr, nullptr, idx, cE, cB);
      mag_apply(e1, cE, cB, rc.B0, rc.v0, rho, dt, B2, v2);
#pragma unroll
      for (int i = 0; i < 3; ++i) {
        st32(m.Bo[1][i], col, B2[i]);
        st32(m.vo[1][i], col, v2[i]);
      }
      const float rr = mag_resid(B1, v1, B2, v2);
      res1 = (rr != rr || rr > res1) ? rr : res1;
    }
  } else {
    const int x0 = icx * a.lx, x1 = min(x0 + a.lx, a.nx);
    float dm[6], dc[6], dp[6];
    MagRaw rc, rn;
    float rho_c = 1.f, rho_n = 1.f;
    if (has) {
      MagRaw t;
      bool f2, f1;
      mag_load<FIRST, DIM == 3>(m, xoff(x0 - 2, f2) + col, t);
      flips(t, f2);
      mag_derive(t, a.euler, dm);
      mag_load<FIRST, DIM == 3>(m, xoff(x0 - 1, f1) + col, rc);
      flips(rc, f1);
      mag_derive(rc, a.euler, dc);
      mag_load<FIRST, DIM == 3>(m, (unsigned)x0 * sx + col, rn);
      flips(rn, false);
      mag_derive(rn, a.euler, dp);
    }
    if (inner) {
      bool unused;
      rho_c = ld32(m.rho, xoff(x0 - 1, unused) + col);
      rho_n = ld32(m.rho, (unsigned)x0 * sx + col);
    }
    // second-iterate carries of the own column (tile threads): derived first
    // iterate at x-2, x-1; first iterate and fixed-point inputs at x-1
    float e1m[6], e1c[6], Bp[3], vp[3], B0p[3], v0p[3], rhop = 1.f;
#pragma unroll
    for (int i = 0; i < 6; ++i) e1m[i] = e1c[i] = 0.f;
#pragma unroll
    for (int i = 0; i < 3; ++i) Bp[i] = vp[i] = B0p[i] = v0p[i] = 0.f;
    for (int x = x0 - 1; x <= x1; ++x) {
      const int bx = x & 1;
      // prefetch x + 2
      MagRaw rnn;
      float rho_nn = 1.f;
      bool fnn;  // flips are applied when the prefetched plane is derived
      const unsigned pnn = xoff(x + 2, fnn) + col;
      if (has) mag_load<FIRST, DIM == 3>(m, pnn, rnn);
      if (inner) rho_nn = ld32(m.rho, pnn);
      if (has) {
#pragma unroll
        for (int i = 0; i < 6; ++i) sK[bx][i][o] = dc[i];
      }
      __syncthreads();
      float B1[3], v1[3], e1[6];
      // a launch of the first iterate only needs it at the own cells
      if (one ? own : inner) {
        float cE[3], cB[3];
#pragma unroll
        for (int i = 0; i < 6; ++i) SK[i] = sK[bx][i];
        mag_curls<3, WX>(SK, o, dm, dp, idx, cE, cB);
        mag_apply(dc, cE, cB, rc.B0, rc.v0, rho_c, dt, B1, v1);
        if (own && x >= x0 && x < x1) {
          const unsigned cell = (unsigned)x * sx + col;
          if (store1) {
            // plain stores (with non-temporal ones the pair that keeps its
            // first iterate too, where the previous step ended on it, was
            // slower)
#pragma unroll
            for (int i = 0; i < 3; ++i) {
              st32(m.Bo[0][i], cell, B1[i]);
              st32(m.vo[0][i], cell, v1[i]);
            }
          }
          const float rr = mag_resid(rc.Bk, rc.vk, B1, v1);
          res0 = (rr != rr || rr > res0) ? rr : res0;
        }
        mag_derive3(B1, v1, rc.B0, rc.v0, a.euler, e1);
        if (!one) {
          if (!gh) {
#pragma unroll
            for (int i = 0; i < 6; ++i) sK1[bx][i][o1] = e1[i];
          }
          if (BCX && own) ghost_copies(sK1[bx], e1);
        }
      }
      // No barrier here: the second iterate lags one plane and reads the
      // first iterates of plane x - 1 (sK1[bx ^ 1]), written before this
      // plane's barrier; sK1[bx] is next read at plane x + 1 and next
      // written at plane x + 2, each behind a barrier of a later plane.
      if (own && !one && x - 1 >= x0 && x - 1 < x1) {
        float cE[3], cB[3], B2[3], v2[3];
#pragma unroll
        for (int i = 0; i < 6; ++i) SK1[i] = sK1[bx ^ 1][i];
        if constexpr (BCX) {
          // x neighbours past an open / reflective x edge: ghost copies of this cell
          float xm[6], xp[6];
          derived_ghost(e1m, -1, xm);
          derived_ghost(e1, -1, xp);
          if (x - 1 == 0 && a.bx.lo != AMX_BC_PERIODIC)
            derived_ghost(e1c, a.bx.lo == AMX_BC_REFLECTIVE ? RX : -1, xm);
          if (x - 1 == a.nx - 1 && a.bx.hi != AMX_BC_PERIODIC)
            derived_ghost(e1c, a.bx.hi == AMX_BC_REFLECTIVE ? RX : -1, xp);
          mag_curls<3, W1>(SK1, o1, xm, xp, idx, cE, cB);
        } else {
          mag_curls<3, W1>(SK1, o1, e1m, e1, idx, cE, cB);
        }
        mag_apply(e1c, cE, cB, B0p, v0p, rhop, dt, B2, v2);
        const unsigned cell = (unsigned)(x - 1) * sx + col;
#pragma unroll
        for (int i = 0; i < 3; ++i) {
          // non-temporal stores for launch 0, plain ones for the later
          // pairs (each the faster in A/B runs at 256^3)
          if (FIRST) {
            st32nt(m.Bo[1][i], cell, B2[i]);
            st32nt(m.vo[1][i], cell, v2[i]);
          } else {
            st32(m.Bo[1][i], cell, B2[i]);
            st32(m.vo[1][i], cell, v2[i]);
          }
        }
        const float rr = mag_resid(Bp, vp, B2, v2);
        res1 = (rr != rr || rr > res1) ? rr : res1;
      }
      if (tile) {
#pragma unroll
        for (int i = 0; i < 6; ++i) {
          e1m[i] = e1c[i];
          e1c[i] = e1[i];
        }
#pragma unroll
        for (int i = 0; i < 3; ++i) {
          Bp[i] = B1[i];
          vp[i] = v1[i];
          B0p[i] = rc.B0[i];
          v0p[i] = rc.v0[i];
        }
        rhop = rho_c;
      }
      if (has) {
#pragma unroll
        for (int i = 0; i < 6; ++i) {
          dm[i] = dc[i];
          dc[i] = dp[i];
        }
        flips(rnn, fnn);
        mag_derive(rnn, a.euler, dp);
        rc = rn;
        rn = rnn;
      }
      rho_c = rho_n;
      rho_n = rho_nn;
    }
  }
  const unsigned m0 = block_max_u<NW>(canon_bits(res0), sred[0]);
  const unsigned m1 = block_max_u<NW>(canon_bits(res1), sred[1]);
  if (tid == 0 && !a.regen) {
    if (m0) atomicMax(&C->resp[L % NRES][blockIdx.x & (NSLOT - 1)], m0);
    if (m1) atomicMax(&C->resp[(L + 1) % NRES][blockIdx.x & (NSLOT - 1)], m1);
  }
}

// Final launch index of the fixed point of each running slot (one thread per
// slot) from its last executed pair, recorded in ctl->iters: the stage
// kernels after the sub-step read the iterate of that launch. spec: the call
// is queued before the host has seen the residuals; if the reference's loop
// would still run, it sets held (the gated stage kernels behind it return)
// and leaves iters for the call after the further launches.
//
// A final first iterate that its pair did not store (skip_mask: the
// speculative pairs, which store it only at kprev) is recomputed by the
// regeneration launch queued before this one; a speculative call without
// one queued (regen_queued = 0) holds the step instead, and the host
// queues the regeneration when it finishes the step.
__global__ void k_mag_final(Ctl* ctl, float tol, int max_iter, int spec, unsigned skip_mask, int regen_queued) {
  Ctl* const c = ctl + blockIdx.x;
  if (!c->run) return;
  const int L = c->iters - 1;
  const float r0 = __uint_as_float(slots_max(c->resp[L % NRES]));
  const float r1 = __uint_as_float(slots_max(c->resp[(L + 1) % NRES]));
  const int f = pair_final(L, r0, r1, tol, max_iter);
  const bool unstored = f >= 0 && f == L && L <= 62 && ((skip_mask >> (L >> 1)) & 1u) && L != c->kprev;
  if (spec && unstored && c->kprev >= 0) c->regen_ev += 1;
  if ((f < 0 || (unstored && !regen_queued)) && spec) {
    c->held = 1;
    c->stall = 1;
    return;
  }
  c->iters = f >= 0 ? f : L + 1;
  if (f >= 0) c->kprev = f;
  c->held = 0;
  c->stall = 0;
}

// Unsplit CFL maxima max(|u_a| + c) per axis (_timestep_estimator.py:118-126)
// of the gas rows of slot slot0 + blockIdx.y into its ctl[].cflp.
struct CflArgs {
  const float* w[4];  // rho, v_x, v_y, (v_z): rows of slot 0
  const float* p;
  Ctl* ctl;
  const Phys* phs;
  long sstride;
  long n;
  int slot0;
};

// V4: the rows are 16-byte aligned and n is a multiple of 4; each thread
// reads four cells of every row per load.
template <int DIM, bool V4>
__global__ void k_cfl(const CflArgs a) {
  const int slot = a.slot0 + blockIdx.y;
  Ctl* const C = a.ctl + slot;
  if (!C->live) return;
  const long so = slot * a.sstride;
  const float gamma = a.phs[slot].gamma;
  const float* rho = a.w[0] + so;
  const float* p = a.p + so;
  const float* u0 = a.w[1] + so;
  const float* u1 = a.w[2] + so;
  const float* u2 = DIM == 3 ? a.w[3] + so : nullptr;
  float mx[3] = {0.f, 0.f, 0.f};
  auto cell = [&](float r, float pp, float x, float y, float z) {
    const float cs = fsqrt(gamma * pp * frcp(r));
    const float u[3] = {x, y, z};
#pragma unroll
    for (int k = 0; k < DIM; ++k) {
      const float m = fabsf(u[k]) + cs;
      mx[k] = (m != m || m > mx[k]) ? m : mx[k];
    }
  };
  if (V4) {
    const long n4 = a.n >> 2;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += (long)gridDim.x * blockDim.x) {
      const float4 r = ((const float4*)rho)[i], pp = ((const float4*)p)[i];
      const float4 x = ((const float4*)u0)[i], y = ((const float4*)u1)[i];
      const float4 z = DIM == 3 ? ((const float4*)u2)[i] : make_float4(0.f, 0.f, 0.f, 0.f);
      cell(r.x, pp.x, x.x, y.x, z.x);
      cell(r.y, pp.y, x.y, y.y, z.y);
      cell(r.z, pp.z, x.z, y.z, z.z);
      cell(r.w, pp.w, x.w, y.w, z.w);
    }
  } else {
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < a.n; i += (long)gridDim.x * blockDim.x)
      cell(rho[i], p[i], u0[i], u1[i], DIM == 3 ? u2[i] : 0.f);
  }
  __shared__ unsigned sred[DIM][4];
#pragma unroll
  for (int k = 0; k < DIM; ++k) {
    const unsigned m = block_max_u<4>(canon_bits(mx[k]), sred[k]);
    if (threadIdx.x == 0 && m) atomicMax(&C->cflp[k][blockIdx.x & (NSLOT - 1)], m);
  }
}

struct StepParams {
  int fixed;
  int num_timesteps;
  int exact_end;
  float C;
  float dx;
  int dim;
  int clamp_tp = 0;  // clamp dt to tp - t (snapshot time points, time_integration.py:671-674)
  float tp = 0.f;
  int batch_first = 0;  // first step of a queued batch: restarts the batch's iteration maximum
};

// dt = C / sum_a(max_a / dx) (_timestep_estimator.py:128), or t_end/num_timesteps
// (time_integration.py:667-668); clamp to t_end - t (time_integration.py:677-678).
__device__ __forceinline__ void cfl_reduce(Ctl* c, int dim) {
  for (int k = 0; k < dim; ++k) {
    c->cfl[k] = slots_max(c->cflp[k]);
    for (int i = 0; i < NSLOT; ++i) c->cflp[k][i] = 0u;
  }
}

__device__ __forceinline__ float cfl_dt_from(const Ctl* c, float C, float dx, int dim) {
  float s = __uint_as_float(c->cfl[0]) / dx + __uint_as_float(c->cfl[1]) / dx;
  if (dim == 3) s = s + __uint_as_float(c->cfl[2]) / dx;
  return (C * 1.0f) / s;
}

__device__ __forceinline__ void fixed_point_reset(Ctl* c) {
  for (int i = 0; i < NSLOT; ++i) c->resp[0][i] = c->resp[1][i] = 0u;
  c->iters = 0;
  c->done = 0;
}

__device__ __forceinline__ void zero_cflp(Ctl* c) {
  for (int k = 0; k < 3; ++k)
    for (int i = 0; i < NSLOT; ++i) c->cflp[k][i] = 0u;
}

// One thread per slot (slot = blockIdx.x): the step's dt, and whether the
// slot runs this step (live and not yet at t_end / num_timesteps).
__device__ void step_begin(Ctl* c, const StepParams& p) {
  if (p.batch_first) {
    c->iters_bmax = 0;
    c->regen_ev = 0;
  }
  if (!c->live) {
    c->run = 0;
    return;
  }
  c->run = !c->stall && (p.fixed ? (c->steps < p.num_timesteps) : (c->t < c->t_end));
  if (!c->run) return;
  float dt;
  if (p.fixed) {
    dt = c->t_end / (float)p.num_timesteps;
  } else {
    cfl_reduce(c, p.dim);
    dt = cfl_dt_from(c, p.C, p.dx, p.dim);
  }
  if (p.clamp_tp) {
    const float rem = p.tp - c->t;
    dt = rem < dt ? rem : dt;
  }
  if (p.exact_end) {
    const float rem = c->t_end - c->t;
    dt = rem < dt ? rem : dt;
  }
  c->dt = dt;
  zero_cflp(c);
  fixed_point_reset(c);
}

__global__ void k_step_begin(Ctl* ctl, StepParams p) { step_begin(ctl + blockIdx.x, p); }

// dt given by the caller (one step of _evolve_state_fv)
__global__ void k_set_dt(Ctl* c, const float* dtp) {
  c->dt = *dtp;
  c->run = 1;
  c->held = 0;
  c->stall = 0;
  c->bset = -1;  // one step of a caller's state: B0 is its B rows
  zero_cflp(c);
  fixed_point_reset(c);
}

// keep_b: the final B of the step stays in its iterate set, which holds B0
// of the next step (the state's B rows are written once, by k_copy_bset).
__device__ void step_end(Ctl* c, int keep_b) {
  if (!c->run || c->held) return;  // held: the host finishes the step
  c->t = c->t + c->dt;
  c->steps += 1;
  c->iters_sum += c->iters;
  c->iters_max = c->iters > c->iters_max ? c->iters : c->iters_max;
  c->iters_bmax = c->iters > c->iters_bmax ? c->iters : c->iters_bmax;
  if (keep_b) c->bset = iterate_set(c->bset, c->iters);
}

__global__ void k_step_end(Ctl* ctl, int keep_b) { step_end(ctl + blockIdx.x, keep_b); }

// the end of one queued step and the beginning of the next, in one launch
__global__ void k_step_end_begin(Ctl* ctl, int keep_b, StepParams p) {
  Ctl* const c = ctl + blockIdx.x;
  step_end(c, keep_b);
  step_begin(c, p);
}

// B of slot 0 from its iterate set into the state's B rows (a loop that kept
// B in the sets), then k_bset_clear hands B0 back to the state's rows. A copy
// queued behind the last batch of a loop before the host has read the
// control block (gate) does nothing when a step of the batch stalled: the
// host copies after finishing that step.
struct BsetArgs {
  const float* PB[4][3];  // B rows of the iterate sets
  float* dst[3];
  long n;
  int gate;
};

__global__ void k_copy_bset(const Ctl* ctl, BsetArgs a) {
  const int bs = ctl->bset;
  if (bs < 0 || (a.gate && ctl->stall)) return;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < a.n; i += (long)gridDim.x * blockDim.x) {
#pragma unroll
    for (int k = 0; k < 3; ++k) a.dst[k][i] = a.PB[bs][k][i];
  }
}

__global__ void k_bset_clear(Ctl* ctl, int gate) {
  if (!(gate && ctl->stall)) ctl->bset = -1;
}

// The stalled steps of a queued batch (one slot per block) take part in
// their further launches again, and only they: the steps queued behind a
// stalled one cleared its run (k_step_begin) and changed nothing else; slots
// that completed their steps stop running until the next step begins.
__global__ void k_resume(Ctl* ctl) {
  Ctl* const c = ctl + blockIdx.x;
  c->run = c->stall ? 1 : 0;
}

__global__ void k_cfl_out(Ctl* c, float C, float dx, int dim, float* out) {
  cfl_reduce(c, dim);
  *out = cfl_dt_from(c, C, dx, dim);
}

__global__ void k_zero_cfl(Ctl* c) { zero_cflp(c); }

// (Re)start slot blockIdx.x + slot0 at t = 0 (live = 1), or empty it (live = 0).
// fresh: a new simulation in the slot (a loop restarting on the same
// simulation keeps kprev, the fixed point's last final launch index)
__global__ void k_slot_reset(Ctl* ctl, int slot0, float t_end, int live, int fresh) {
  Ctl* const c = ctl + slot0 + blockIdx.x;
  c->dt = 0.f;
  c->t = 0.f;
  c->steps = 0;
  for (int i = 0; i < 4; ++i) c->cfl[i] = 0u;
  for (int k = 0; k < 3; ++k)
    for (int i = 0; i < NSLOT; ++i) c->cflp[k][i] = 0u;
  for (int k = 0; k < NRES; ++k)
    for (int i = 0; i < NSLOT; ++i) c->resp[k][i] = 0u;
  c->iters = 0;
  c->done = 0;
  c->live = live;
  c->run = live;
  c->t_end = t_end;
  c->bset = -1;
  c->held = 0;
  c->stall = 0;
  c->regen_ev = 0;
  if (fresh) c->kprev = -1;
  c->iters_sum = 0;
  c->iters_max = 0;
  c->iters_bmax = 0;
}

// Streaming copy, 4 floats per lane per iteration.
__global__ void k_copy(float4* __restrict__ dst, const float4* __restrict__ src, long n4) {
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += (long)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

__global__ void k_copy_tail(float* __restrict__ dst, const float* __restrict__ src, long from, long n) {
  const long i = from + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

}  // namespace amx

#include "fv_adjoint.inc"

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

#define HIPCHK(expr) AMX_HIPCHK(expr)

// Tile rows of k_stage. The block is TA + 1 wavefronts; TA = 7 gives 8
// waves, two per SIMD, so the barrier-separated phases of a block are
// balanced over the CU's 4 SIMDs (7- and 9-wave blocks leave one SIMD with
// an extra wave that every barrier waits for: 6-40 % slower at 256^3).
#ifndef AMX_STAGE_TA
#define AMX_STAGE_TA 7
#endif
constexpr int TA_DEF = AMX_STAGE_TA;
// Tile of the pair kernel k_mag: TA_MAG rows x TB_MAG columns; the block
// covers the tile and a two-cell ring, (TA + 4) x (TB + 4) cells rounded up
// to whole 256-thread groups
#ifndef AMX_MAG_TA
#define AMX_MAG_TA 11  // (11 + 4) x 68 = 1020 cells: one 16-wave block, 4 waves per SIMD at <= 128 VGPRs
#endif
#ifndef AMX_MAG_TB
#define AMX_MAG_TB 64
#endif
constexpr int TA_MAG = AMX_MAG_TA;
constexpr int TB_MAG = AMX_MAG_TB;

}  // namespace

// Slots: a context steps `cap` independent simulations of one grid and
// scheme with the same kernels (the slot is blockIdx.y). A single simulation
// (amx_fv_create) is one slot working in place on the caller's state; a pool
// (amx_fv_pool_create) owns cap slots of state + workspace, sstride floats
// apart.
struct amx_fv_ctx {
  amx_fv_config cfg;
  int NV, NS;       // gas rows; state rows (NV + 3 with MHD)
  long ncell;
  float dx;
  int na, nb, nxm;  // kernel axes: 3D (ny, nz, nx); 2D (nx, ny, 1)
  int bcx;          // some axis has open / reflective boundaries
  AxisBc bx, ba, bb;  // boundaries of the kernel axes x, A, B
  int ntb, nta, ntb_mag, nta_mag, ncx, lx, ncx_mag, lx_mag;
  int cap;
  long sstride;     // floats per slot (0 for a single simulation)
  float* mem;       // pool: cap x [state NS | G1 NV | P0 6 | P1 6]; single: [G1 | P0 | P1]
  int own_state;
  Ctl* ctl;         // cap entries
  Ctl* hctl;        // pinned host mirror
  Phys* dph;        // cap entries
  Phys* hph;        // pinned host copies
  int last_k;
  int regen_q;      // device loop: the next batch queues regeneration launches (a step of the last one needed it)
  StepParams sp;    // pool stepping
  int span;         // pool: slots [0, span) take part in the launched step
  int stepping;     // pool: a step is launched and not finished
  hipEvent_t ev_status;
  float* adj;       // reverse-mode workspace (allocated on first use)
  float* adjm;      // MHD reverse mode: state copy, sub-step buffers and the fixed point's iterates
  int adjm_iters;   // iterates adjm holds room for
  unsigned char* amg;  // hybrid / AM-HLLC compression switch, one byte per cell and slot
  float* ppf;       // VAN_ALBADA_PP factors: 2 dim + 1 fields per cell and slot
  amx_step_hook hook;  // between-step hook of the step loop (forcing), or null
  void* hook_user;
  void* snapwork;   // snapshot diagnostics partials (allocated on first use)
  int klast;        // pool: last fixed-point launch of the step in flight
  int keep_b;       // the running loop keeps B in the iterate sets (no per-step copy)
  int use_graphs;   // pool: replay captured HIP graphs of the step halves
  unsigned skip_mask;  // pairs of the current step queued without storing their first iterate
  int batch;        // pool: steps queued by one step_launch (1: a step in two halves)
  int batching;     // pool: the launched step is a queued batch
  std::map<std::tuple<int, int, int>, hipGraphExec_t> g_batch;  // (span, last count, steps)
  std::map<std::pair<int, int>, hipGraphExec_t> g_launch;  // (span, last count)
  std::map<std::pair<int, unsigned>, hipGraphExec_t> g_finish;  // (span, pairs without first iterate)
  std::vector<int> live;
  std::vector<int> fresh;  // loaded since the last status readback
  // optional HIP-event timing of kernel launches (single simulation) on
  // every timing-th step; tstep counts steps since it was set
  int timing;
  long tstep;
  struct Rec {
    int e0, e1, cls, k;
  };
  std::vector<hipEvent_t> ev;
  std::vector<int> ev_free;  // indices into ev of events no record holds
  std::vector<Rec> pending;
  double ms[2];
  long launches[2];
};

struct amx_fv_pool {
  amx_fv_ctx* c;
};

namespace {

enum { CLS_STAGE = 0, CLS_MAG = 1 };

int time_begin(amx_fv_ctx* c, hipStream_t st) {
  if (c->timing <= 0 || c->tstep % c->timing != 0) return -1;
  while (c->ev_free.size() < 2) {
    hipEvent_t e;
    if (hipEventCreate(&e) != hipSuccess) return -1;
    c->ev_free.push_back((int)c->ev.size());
    c->ev.push_back(e);
  }
  const int i = c->ev_free.back();
  c->ev_free.pop_back();
  (void)hipEventRecord(c->ev[i], st);
  return i;
}

void time_end(amx_fv_ctx* c, hipStream_t st, int i, int cls, int k) {
  if (i < 0) return;
  const int j = c->ev_free.back();  // time_begin left at least one
  c->ev_free.pop_back();
  (void)hipEventRecord(c->ev[j], st);
  c->pending.push_back({i, j, cls, k});
}

// Folds the first n pending records, whose events have completed, into the
// totals and frees their events.
void time_resolve_first(amx_fv_ctx* c, size_t n) {
  if (n > c->pending.size()) n = c->pending.size();
  for (size_t i = 0; i < n; ++i) {
    const auto& r = c->pending[i];
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, c->ev[r.e0], c->ev[r.e1]) == hipSuccess) {
      c->ms[r.cls] += ms;
      c->launches[r.cls] += 1;
    }
    c->ev_free.push_back(r.e0);
    c->ev_free.push_back(r.e1);
  }
  c->pending.erase(c->pending.begin(), c->pending.begin() + n);
}

// Call only after the stream is synchronised.
void time_resolve(amx_fv_ctx* c) { time_resolve_first(c, c->pending.size()); }

// Drops the pending records from index `from` on that `drop` selects (launches
// that exited early or were held), freeing their events.
template <class F>
void time_drop(amx_fv_ctx* c, size_t from, F&& drop) {
  size_t w = from;
  for (size_t i = from; i < c->pending.size(); ++i) {
    const auto r = c->pending[i];
    if (drop(r)) {
      c->ev_free.push_back(r.e0);
      c->ev_free.push_back(r.e1);
    } else {
      c->pending[w++] = r;
    }
  }
  c->pending.resize(w);
}

int env_int(const char* name, int dflt) {
  const char* s = getenv(name);
  return s ? atoi(s) : dflt;
}

amx_status validate(const amx_fv_config* c) {
  if (!c) return fail(AMX_ERR_INVALID, "null config%s");
  if (c->dim != 2 && c->dim != 3) return fail(AMX_ERR_UNSUPPORTED, "dimensionality must be 2 or 3%s");
  if (c->nx < 2 || c->ny < 2 || (c->dim == 3 && c->nz < 2)) return fail(AMX_ERR_INVALID, "grid too small%s");
  if (c->dim == 2 && c->nz != 1) return fail(AMX_ERR_INVALID, "nz must be 1 in 2D%s");
  if (c->riemann < AMX_HLL || c->riemann > AMX_AM_HLLC)
    return fail(AMX_ERR_UNSUPPORTED, "riemann solver must be HLL, HLLC, HLLC_LM, LAX_FRIEDRICHS, HYBRID_HLLC or AM_HLLC%s");
  if (c->limiter < AMX_MINMOD || c->limiter > AMX_VAN_ALBADA_PP)
    return fail(AMX_ERR_UNSUPPORTED,
                "limiter must be MINMOD, OSHER, DOUBLE_MINMOD, SUPERBEE, VAN_ALBADA or VAN_ALBADA_PP%s");
  if (c->limiter == AMX_VAN_ALBADA_PP && !(c->C_cfl > 0.0f))
    return fail(AMX_ERR_INVALID, "VAN_ALBADA_PP needs a positive C_cfl (its q = 1 / C_cfl)%s");
  if (c->magnetic_integrator != AMX_IMPLICIT_MIDPOINT && c->magnetic_integrator != AMX_IMPLICIT_EULER)
    return fail(AMX_ERR_UNSUPPORTED, "unknown magnetic integrator%s");
  if (c->max_fixed_point_iterations < 0 || c->max_fixed_point_iterations > 100000)
    return fail(AMX_ERR_INVALID, "max_fixed_point_iterations out of range%s");
  if (!(c->grid_spacing > 0)) return fail(AMX_ERR_INVALID, "grid_spacing must be positive%s");
  if (!(c->gamma > 1.0f)) return fail(AMX_ERR_INVALID, "gamma must exceed 1%s");
  if ((long)c->nx * c->ny * (c->dim == 3 ? c->nz : 1) >= (1L << 30))
    return fail(AMX_ERR_INVALID, "a grid of 2^30 cells or more is not supported (32-bit cell indices)%s");
  int nonper = 0;
  for (int k = 0; k < 6; ++k) {
    const int b = c->boundary[k];
    if (b != AMX_BC_PERIODIC && b != AMX_BC_OPEN && b != AMX_BC_REFLECTIVE)
      return fail(AMX_ERR_INVALID, "boundary codes must be AMX_BC_PERIODIC, AMX_BC_OPEN or AMX_BC_REFLECTIVE%s");
    if (k >= 2 * c->dim && b != AMX_BC_PERIODIC) return fail(AMX_ERR_INVALID, "2D configs have no z boundaries%s");
    nonper |= b != AMX_BC_PERIODIC;
  }
  for (int a = 0; a < 3; ++a)
    if ((c->boundary[2 * a] == AMX_BC_PERIODIC) != (c->boundary[2 * a + 1] == AMX_BC_PERIODIC))
      return fail(AMX_ERR_UNSUPPORTED,
                  "one side of an axis periodic and the other not has no defined ghost cells "
                  "(boundaries.py applies periodic ghosts only when both sides are periodic)%s");
  if (nonper && (c->riemann == AMX_HYBRID_HLLC || c->riemann == AMX_AM_HLLC))
    return fail(AMX_ERR_UNSUPPORTED, "HYBRID_HLLC and AM_HLLC need periodic boundaries here%s");
  if (nonper && c->limiter == AMX_VAN_ALBADA_PP)
    return fail(AMX_ERR_UNSUPPORTED, "VAN_ALBADA_PP needs periodic boundaries here%s");
  return AMX_OK;
}

// G1 (gas rows) | P0 | P1 | P2 (B, v iterates of the fixed point)
// G1 (gas rows) | P0 | P1 | P2 (B, v iterate sets) | PF (fixed pressure) | P3
// (the fourth set: step loops and pools keep B in the sets)
long ws_fields(const amx_fv_config* cfg, int /*own_state*/) {
  return (cfg->dim + 2) + (cfg->mhd ? 25 : 0);
}
long state_fields(const amx_fv_config* cfg) { return (cfg->dim + 2) + (cfg->mhd ? 3 : 0); }
long cells(const amx_fv_config* cfg) { return (long)cfg->nx * cfg->ny * (cfg->dim == 3 ? cfg->nz : 1); }

Phys make_phys(float gamma, double grid_spacing) {
  Phys ph;
  ph.gamma = gamma;
  ph.gm1 = gamma - 1.0f;
  ph.inv_gm1 = 1.0f / ph.gm1;
  ph.inv_dx = (float)(1.0 / grid_spacing);
  return ph;
}

void free_ctx(amx_fv_ctx* c) {
  if (!c) return;
  (void)hipDeviceSynchronize();
  if (c->mem) (void)hipFree(c->mem);
  if (c->adj) (void)hipFree(c->adj);
  if (c->adjm) (void)hipFree(c->adjm);
  if (c->amg) (void)hipFree(c->amg);
  if (c->ppf) (void)hipFree(c->ppf);
  if (c->snapwork) (void)hipFree(c->snapwork);
  if (c->ctl) (void)hipFree(c->ctl);
  if (c->dph) (void)hipFree(c->dph);
  if (c->hctl) (void)hipHostFree(c->hctl);
  if (c->hph) (void)hipHostFree(c->hph);
  if (c->ev_status) (void)hipEventDestroy(c->ev_status);
  for (auto& g : c->g_launch) (void)hipGraphExecDestroy(g.second);
  for (auto& g : c->g_finish) (void)hipGraphExecDestroy(g.second);
  for (auto& g : c->g_batch) (void)hipGraphExecDestroy(g.second);
  for (auto e : c->ev) (void)hipEventDestroy(e);
  delete c;
}

// Planes per block of the 3D pair kernel along x. A block loads its chunk plus
// the 4 planes that the two iterates' x stencils reach past it, and one block
// fills a CU, so a launch runs in rounds of (CU count) blocks. The pick is the
// chunk length (>= 8 planes) whose in-order list schedule of `blocks_per_chunk`
// x chunks blocks onto the CUs, a block costing its loaded planes, ends
// first; past 16 rounds the tail is ignored. At 256^3 on 256 CUs this is 32
// planes: 768 blocks, three full rounds (0.435 ms per launch against 0.468 ms
// for 16-plane chunks in six rounds, 0.52 ms for 64-plane chunks in two).
int pick_mag_chunk(long blocks_per_chunk, int nx) {
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || cus <= 0)
    cus = 256;
  if (nx <= 8) return nx;
  if (cus > 1024) cus = 1024;
  int best = 8, prev = -1;
  double best_t = 0.0;
  double fin_buf[1024];  // the CUs' finish times (no heap use: the allocation-free entry point calls this)
  double* const fin = fin_buf;
  for (int want = 1; want <= (nx + 7) / 8; ++want) {
    const int lx = (nx + want - 1) / want;
    if (lx < 8 || lx == prev) continue;
    prev = lx;
    const int ncx = (nx + lx - 1) / lx;
    const long blocks = blocks_per_chunk * ncx;
    double t;
    if (blocks > 16L * cus) {
      t = (double)blocks_per_chunk * (nx + 4.0 * ncx) / cus;
    } else {
      // min-heap of the CUs' finish times; blocks dispatch in id order (x chunk slowest)
      std::fill(fin, fin + cus, 0.0);
      for (int k = 0; k < ncx; ++k) {
        const double cost = (k + 1 < ncx ? lx : nx - k * lx) + 4.0;
        for (long j = 0; j < blocks_per_chunk; ++j) {
          std::pop_heap(fin, fin + cus, std::greater<double>());
          fin[cus - 1] += cost;
          std::push_heap(fin, fin + cus, std::greater<double>());
        }
      }
      t = *std::max_element(fin, fin + cus);
    }
    if (best_t == 0.0 || t < best_t) {
      best_t = t;
      best = lx;
    }
  }
  return best;
}

// Geometry and launch shapes of a context of cap slots (no allocation).
void init_geometry(amx_fv_ctx* c, const amx_fv_config* cfg, int cap, int own_state) {
  c->cfg = *cfg;
  c->NV = cfg->dim + 2;
  c->NS = (int)state_fields(cfg);
  c->ncell = cells(cfg);
  c->dx = (float)cfg->grid_spacing;
  const AxisBc bcs[3] = {{cfg->boundary[0], cfg->boundary[1]},
                         {cfg->boundary[2], cfg->boundary[3]},
                         {cfg->boundary[4], cfg->boundary[5]}};
  if (cfg->dim == 3) {
    c->nxm = cfg->nx;
    c->na = cfg->ny;
    c->nb = cfg->nz;
    c->bx = bcs[0];
    c->ba = bcs[1];
    c->bb = bcs[2];
  } else {
    c->nxm = 1;
    c->na = cfg->nx;
    c->nb = cfg->ny;
    c->bx = {AMX_BC_PERIODIC, AMX_BC_PERIODIC};
    c->ba = bcs[0];
    c->bb = bcs[1];
  }
  c->bcx = 0;
  for (int k = 0; k < 2 * cfg->dim; ++k) c->bcx |= cfg->boundary[k] != AMX_BC_PERIODIC;
  c->ntb = (c->nb + TB - 1) / TB;
  c->nta = (c->na + TA_DEF - 1) / TA_DEF;
  c->nta_mag = (c->na + TA_MAG - 1) / TA_MAG;
  c->ntb_mag = (c->nb + TB_MAG - 1) / TB_MAG;
  if (cfg->dim == 3) {
    // planes per block along x for the stage kernels: about `blocks` blocks,
    // the chunk the multiple of 4 planes (the kernels' x-march unroll) nearest
    // to that, at least 8 (the kernels reload 4 planes at each chunk start, so
    // longer chunks save traffic). Measured at 256^3 (148 tiles): 20 planes
    // 1.718 ms per step, 16: 1.753, 24 and 30: 1.73, 28: 1.80, 36: 1.85.
    auto pick = [&](long tiles, const char* env, long blocks) {
      int lx = env_int(env, 0);
      if (lx <= 0) {
        const long target = (tiles * cfg->nx + blocks / 2) / blocks;
        lx = (int)(((target + 2) / 4) * 4);
        if (lx < 8) lx = 8;
      }
      return lx > cfg->nx ? cfg->nx : lx;
    };
    c->lx = pick((long)c->ntb * c->nta, "AMX_LX", 1900);
    c->lx_mag = env_int("AMX_LX_MAG", 0);
    // one simulation's blocks: pools run partly filled as often as full
    if (c->lx_mag <= 0) c->lx_mag = pick_mag_chunk((long)c->ntb_mag * c->nta_mag, cfg->nx);
    if (c->lx_mag > cfg->nx) c->lx_mag = cfg->nx;
  } else {
    c->lx = c->lx_mag = 1;
  }
  c->ncx = (c->nxm + c->lx - 1) / c->lx;
  c->ncx_mag = (c->nxm + c->lx_mag - 1) / c->lx_mag;
  c->cap = cap;
  c->own_state = own_state;
  const long per_slot = c->ncell * (ws_fields(cfg, own_state) + (own_state ? c->NS : 0));
  c->sstride = own_state ? per_slot : 0;
  c->last_k = 4;
  c->regen_q = 1;
}

// Geometry, launch shapes and allocations of a context of cap slots.
amx_status make_ctx(const amx_fv_config* cfg, int cap, int own_state, amx_fv_ctx** out) {
  amx_status s = validate(cfg);
  if (s != AMX_OK) return s;
  if (!out) return fail(AMX_ERR_INVALID, "null output%s");
  if (cap < 1 || cap > 65535) return fail(AMX_ERR_INVALID, "capacity must be in 1..65535%s");
  amx_fv_ctx* c = new (std::nothrow) amx_fv_ctx();
  if (!c) return fail(AMX_ERR_NOMEM, "host allocation failed%s");
  init_geometry(c, cfg, cap, own_state);
  const long per_slot = c->ncell * (ws_fields(cfg, own_state) + (own_state ? c->NS : 0));
  if (hipMalloc((void**)&c->mem, sizeof(float) * (size_t)per_slot * cap) != hipSuccess) {
    c->mem = nullptr;
    free_ctx(c);
    return fail(AMX_ERR_NOMEM, "device workspace allocation failed%s");
  }
  if ((cfg->riemann == AMX_HYBRID_HLLC || cfg->riemann == AMX_AM_HLLC) &&
      hipMalloc((void**)&c->amg, (size_t)c->ncell * cap) != hipSuccess) {
    c->amg = nullptr;
    free_ctx(c);
    return fail(AMX_ERR_NOMEM, "device workspace allocation failed%s");
  }
  if (cfg->limiter == AMX_VAN_ALBADA_PP &&
      hipMalloc((void**)&c->ppf, sizeof(float) * (size_t)(2 * cfg->dim + 1) * c->ncell * cap) != hipSuccess) {
    c->ppf = nullptr;
    free_ctx(c);
    return fail(AMX_ERR_NOMEM, "device workspace allocation failed%s");
  }
  if (hipMalloc((void**)&c->ctl, sizeof(Ctl) * cap) != hipSuccess ||
      hipMalloc((void**)&c->dph, sizeof(Phys) * cap) != hipSuccess ||
      hipHostMalloc((void**)&c->hctl, sizeof(Ctl) * cap, hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc((void**)&c->hph, sizeof(Phys) * cap, hipHostMallocDefault) != hipSuccess ||
      hipEventCreateWithFlags(&c->ev_status, hipEventDisableTiming) != hipSuccess) {
    free_ctx(c);
    return fail(AMX_ERR_NOMEM, "control allocation failed%s");
  }
  for (int i = 0; i < cap; ++i) c->hph[i] = make_phys(cfg->gamma, cfg->grid_spacing);
  if (hipMemcpy(c->dph, c->hph, sizeof(Phys) * cap, hipMemcpyHostToDevice) != hipSuccess) {
    free_ctx(c);
    return fail(AMX_ERR_HIP, "parameter upload failed%s");
  }
  // single simulation: slot 0 is live from the start; pool slots start empty
  hipLaunchKernelGGL(k_slot_reset, dim3(cap), dim3(1), 0, 0, c->ctl, 0, 0.f, own_state ? 0 : 1, 1);
  if (hipDeviceSynchronize() != hipSuccess) {
    free_ctx(c);
    return fail(AMX_ERR_HIP, "init sync failed%s");
  }
  c->live.assign(cap, own_state ? 0 : 1);
  c->fresh.assign(cap, 0);
  *out = c;
  return AMX_OK;
}

// Slot-0 addresses of the state rows and the workspace.
struct Bufs {
  float* gas[5];
  float* B[3];
  float* G1;
  float* P[4][6];  // iterate sets: launch k writes P[iterate_set(bset, k)]
  float* PF;       // energy-fixed pressure of the magnetic sub-step's output (MHD)
};

Bufs bufs(const amx_fv_ctx* c, float* state) {
  Bufs b;
  memset(&b, 0, sizeof(b));
  const long n = c->ncell;
  if (c->own_state) state = c->mem;
  for (int v = 0; v < c->NV; ++v) b.gas[v] = state + (long)v * n;
  if (c->cfg.mhd)
    for (int k = 0; k < 3; ++k) b.B[k] = state + (long)(c->NV + k) * n;
  float* ws = c->own_state ? c->mem + (long)c->NS * n : c->mem;
  b.G1 = ws;
  if (c->cfg.mhd)
    for (int j = 0; j < 3; ++j)
      for (int k = 0; k < 6; ++k) b.P[j][k] = ws + (long)(c->NV + 6 * j + k) * n;
  if (c->cfg.mhd) b.PF = ws + (long)(c->NV + 18) * n;
  if (c->cfg.mhd)
    for (int k = 0; k < 6; ++k) b.P[3][k] = ws + (long)(c->NV + 19 + k) * n;
  return b;
}

template <int DIM, int SRC, bool HEUN, int W0M, bool LIMX, bool BCX>
void launch_stage_rs(const amx_fv_ctx* c, const StageArgs& a, dim3 grid, dim3 block, hipStream_t st) {
#define AMX_STAGE(RS) hipLaunchKernelGGL((k_stage<DIM, TA_DEF, SRC, HEUN, W0M, RS, LIMX, BCX>), grid, block, 0, st, a)
  switch (c->cfg.riemann) {
    case AMX_HLLC:
      AMX_STAGE(AMX_HLLC);
      break;
    case AMX_HLLC_LM:
      AMX_STAGE(AMX_HLLC_LM);
      break;
    case AMX_LAX_FRIEDRICHS:
      AMX_STAGE(AMX_LAX_FRIEDRICHS);
      break;
    case AMX_HYBRID_HLLC:
      // validate() rejects the switch solvers with open / reflective boundaries
      if constexpr (!BCX) AMX_STAGE(AMX_HYBRID_HLLC);
      break;
    case AMX_AM_HLLC:
      if constexpr (!BCX) AMX_STAGE(AMX_AM_HLLC);
      break;
    default:
      AMX_STAGE(AMX_HLL);
  }
#undef AMX_STAGE
}

template <int DIM, int SRC, bool HEUN, int W0M>
hipError_t launch_stage_t(const amx_fv_ctx* c, const StageArgs& a, int nspan, hipStream_t st) {
  const bool pp = c->cfg.limiter == AMX_VAN_ALBADA_PP;
  if (c->cfg.riemann == AMX_LAX_FRIEDRICHS || pp) {
    // the grid-wide alpha of every slot's stage input, reduced before the stage
    hipLaunchKernelGGL(k_lf_clear, dim3(nspan), dim3(NSLOT), 0, st, a.ctl, a.slot0, a.gated);
    hipLaunchKernelGGL((k_lf_alpha<DIM, SRC>), dim3(LF_BLOCKS, nspan), dim3(256), 0, st, a);
  }
  if (pp) hipLaunchKernelGGL((k_pp_factors<DIM, SRC>), dim3(LF_BLOCKS, nspan), dim3(256), 0, st, a);
  if (c->cfg.riemann == AMX_HYBRID_HLLC || c->cfg.riemann == AMX_AM_HLLC)
    hipLaunchKernelGGL((k_am_switch<DIM, SRC>), dim3(LF_BLOCKS, nspan), dim3(256), 0, st, a);
  const dim3 grid(c->ntb * c->nta * c->ncx, nspan), block((TA_DEF + 1) * TB);
  const bool lim = c->cfg.limiter != AMX_MINMOD;
  if (c->bcx) {
    if (lim)
      launch_stage_rs<DIM, SRC, HEUN, W0M, true, true>(c, a, grid, block, st);
    else
      launch_stage_rs<DIM, SRC, HEUN, W0M, false, true>(c, a, grid, block, st);
  } else if (lim) {
    launch_stage_rs<DIM, SRC, HEUN, W0M, true, false>(c, a, grid, block, st);
  } else {
    launch_stage_rs<DIM, SRC, HEUN, W0M, false, false>(c, a, grid, block, st);
  }
  return hipGetLastError();
}

template <int SRC, bool HEUN, int W0M>
hipError_t launch_stage(const amx_fv_ctx* c, const StageArgs& a, int nspan, hipStream_t st) {
  if (c->cfg.dim == 3) return launch_stage_t<3, SRC, HEUN, W0M>(c, a, nspan, st);
  return launch_stage_t<2, SRC, HEUN, W0M>(c, a, nspan, st);
}

StageArgs stage_base(const amx_fv_ctx* c, float dt_scale) {
  StageArgs a;
  memset(&a, 0, sizeof(a));
  a.pp = c->cfg.limiter == AMX_VAN_ALBADA_PP;
  a.limiter = a.pp ? (int)LIM_VAN_ALBADA : c->cfg.limiter;  // PP: van Albada slopes, then the factors
  a.ppq = (c->cfg.mhd ? 2.0f : 1.0f) / c->cfg.C_cfl;
  if (a.pp) {
    const long n = c->ncell;
    for (int i = 0; i < 2 * c->cfg.dim + 1; ++i) a.ppf[i] = c->ppf + (long)i * n;
    a.ppstride = (long)(2 * c->cfg.dim + 1) * n;
  }
  a.dxf = (float)c->cfg.grid_spacing;
  a.amg = c->amg;
  a.gstride = c->own_state ? c->ncell : 0;
  a.ctl = c->ctl;
  a.phs = c->dph;
  a.sstride = c->sstride;
  a.slot0 = 0;
  a.nx = c->nxm;
  a.na = c->na;
  a.nb = c->nb;
  a.ntb = c->ntb;
  a.nta = c->nta;
  a.ncx = c->ncx;
  a.lx = c->lx;
  a.dt_scale = dt_scale;
  a.bx = c->bx;
  a.ba = c->ba;
  a.bb = c->bb;
  return a;
}

MagArgs mag_base(const amx_fv_ctx* c) {
  MagArgs m;
  memset(&m, 0, sizeof(m));
  m.ctl = c->ctl;
  m.sstride = c->sstride;
  m.slot0 = 0;
  m.max_iter = c->cfg.max_fixed_point_iterations;
  m.write_first = 1;
  m.regen = 0;
  m.skip_mask = 0u;
  m.tol = c->cfg.fixed_point_tol;
  m.euler = c->cfg.magnetic_integrator == AMX_IMPLICIT_EULER;
  m.nx = c->nxm;
  m.na = c->na;
  m.nb = c->nb;
  m.ntb = c->ntb_mag;
  m.nta = c->nta_mag;
  m.ncx = c->ncx_mag;
  m.lx = c->lx_mag;
  m.inv_dx = (float)(1.0 / c->cfg.grid_spacing);
  m.bx = c->bx;
  m.ba = c->ba;
  m.bb = c->bb;
  return m;
}

hipError_t launch_mag(const amx_fv_ctx* c, const MagArgs& m, int nspan, hipStream_t st) {
  constexpr int NT = ((TA_MAG + 4) * (TB_MAG + 4) + 255) / 256 * 256;
  static_assert(NT <= 1024, "pair-kernel block too large");
  const dim3 grid(c->ntb_mag * c->nta_mag * c->ncx_mag, nspan), block(NT);
  const bool first = m.L == 0 && !m.regen;
  if (c->cfg.dim == 3 && c->bcx) {
    if (first)
      hipLaunchKernelGGL((k_mag<3, TA_MAG, TB_MAG, true, true>), grid, block, 0, st, m);
    else
      hipLaunchKernelGGL((k_mag<3, TA_MAG, TB_MAG, true, false>), grid, block, 0, st, m);
  } else if (c->cfg.dim == 3) {
    if (first)
      hipLaunchKernelGGL((k_mag<3, TA_MAG, TB_MAG, false, true>), grid, block, 0, st, m);
    else
      hipLaunchKernelGGL((k_mag<3, TA_MAG, TB_MAG, false, false>), grid, block, 0, st, m);
  } else if (c->bcx) {
    if (first)
      hipLaunchKernelGGL((k_mag<2, TA_MAG, TB_MAG, true, true>), grid, block, 0, st, m);
    else
      hipLaunchKernelGGL((k_mag<2, TA_MAG, TB_MAG, true, false>), grid, block, 0, st, m);
  } else {
    if (first)
      hipLaunchKernelGGL((k_mag<2, TA_MAG, TB_MAG, false, true>), grid, block, 0, st, m);
    else
      hipLaunchKernelGGL((k_mag<2, TA_MAG, TB_MAG, false, false>), grid, block, 0, st, m);
  }
  return hipGetLastError();
}

// CFL maxima of slots [slot0, slot0 + nspan) of the state whose slot-0 rows start at `state`
hipError_t launch_cfl(const amx_fv_ctx* c, const float* state, int slot0, int nspan, hipStream_t st) {
  const long n = c->ncell;
  CflArgs a;
  memset(&a, 0, sizeof(a));
  for (int v = 0; v <= c->cfg.dim; ++v) a.w[v] = state + (long)v * n;
  a.p = state + (long)(c->NV - 1) * n;
  a.ctl = c->ctl;
  a.phs = c->dph;
  a.sstride = c->sstride;
  a.n = n;
  a.slot0 = slot0;
  bool v4 = (n & 3) == 0 && (a.sstride & 3) == 0;
  for (int v = 0; v <= c->cfg.dim; ++v) v4 = v4 && ((uintptr_t)a.w[v] & 15) == 0;
  v4 = v4 && ((uintptr_t)a.p & 15) == 0;
  long blocks = ((v4 ? n / 4 : n) + 255) / 256;
  if (blocks > 4096) blocks = 4096;
  if (c->cfg.dim == 3) {
    if (v4)
      hipLaunchKernelGGL((k_cfl<3, true>), dim3(blocks, nspan), dim3(256), 0, st, a);
    else
      hipLaunchKernelGGL((k_cfl<3, false>), dim3(blocks, nspan), dim3(256), 0, st, a);
  } else if (v4) {
    hipLaunchKernelGGL((k_cfl<2, true>), dim3(blocks, nspan), dim3(256), 0, st, a);
  } else {
    hipLaunchKernelGGL((k_cfl<2, false>), dim3(blocks, nspan), dim3(256), 0, st, a);
  }
  return hipGetLastError();
}

// Launches the fixed-point pair (L, L+1): input = launch L-1's output (the
// state's B and gas velocity for L = 0); the kernel picks the iterate sets
// from ctl->bset (iterate_set).
hipError_t launch_mag_pair(amx_fv_ctx* c, const Bufs& b, int L, int nspan, hipStream_t st, int write_first = 1,
                           int regen = 0) {
  const int dim = c->cfg.dim;
  MagArgs m = mag_base(c);
  m.write_first = write_first;
  m.regen = regen;
  m.skip_mask = c->skip_mask;
  for (int k = 0; k < 3; ++k) m.B0[k] = b.B[k];
  m.v0[0] = b.gas[1];
  m.v0[1] = b.gas[2];
  m.v0[2] = dim == 3 ? b.gas[3] : nullptr;
  m.rho = b.gas[0];
  for (int j = 0; j < 4; ++j)
    for (int k = 0; k < 6; ++k) m.PS[j][k] = b.P[j][k];
  m.L = L;
  if (regen) return launch_mag(c, m, nspan, st);
  const int e = time_begin(c, st);
  const hipError_t err = launch_mag(c, m, nspan, st);
  time_end(c, st, e, CLS_MAG, L);
  return err;
}

// Pairs (bit L / 2) of the speculative batch L = 0, 2, ... <= upto whose
// first iterate is stored only by the pair starting at the slot's previous
// final launch index (ctl.kprev, decided on the device).
unsigned skip_mask_for(int upto) {
  unsigned m = 0u;
  for (int L = 0; L <= upto && L <= 62; L += 2) m |= 1u << (L >> 1);
  return m;
}

// ---- one step of slots [0, nspan), dt already in ctl[s].dt, ctl[s].run set.
// Phase 1: S_A(dt/2) and the speculative batch of fixed-point launches.
amx_status step_phase1(amx_fv_ctx* c, const Bufs& b, int nspan, hipStream_t st, int want_cfl, int* klast) {
  const int NV = c->NV;
  if (!c->cfg.mhd) {
    // SSP-RK2 on the whole state: W1 = E(W0); W = prim(0.5 (U0 + cons(E(W1)))) in place
    StageArgs a1 = stage_base(c, 1.0f);
    for (int v = 0; v < NV; ++v) {
      a1.src[v] = b.gas[v];
      a1.out[v] = b.G1 + (long)v * c->ncell;
    }
    int e = time_begin(c, st);
    HIPCHK((launch_stage<SRC_PLAIN, false, SRC_PLAIN>(c, a1, nspan, st)));
    time_end(c, st, e, CLS_STAGE, 0);
    StageArgs a2 = stage_base(c, 1.0f);
    for (int v = 0; v < NV; ++v) {
      a2.src[v] = b.G1 + (long)v * c->ncell;
      a2.w0[v] = b.gas[v];
      a2.out[v] = b.gas[v];
    }
    a2.want_cfl = want_cfl;
    e = time_begin(c, st);
    HIPCHK((launch_stage<SRC_PLAIN, true, SRC_PLAIN>(c, a2, nspan, st)));
    time_end(c, st, e, CLS_STAGE, 0);
    *klast = 0;
    return AMX_OK;
  }
  const int dim = c->cfg.dim;
  // ---- S_A(dt/2): RK2 with the result in place in the gas rows
  StageArgs a1 = stage_base(c, 0.5f);
  for (int v = 0; v < NV; ++v) {
    a1.src[v] = b.gas[v];
    a1.out[v] = b.G1 + (long)v * c->ncell;
  }
  int e = time_begin(c, st);
  HIPCHK((launch_stage<SRC_PLAIN, false, SRC_PLAIN>(c, a1, nspan, st)));
  time_end(c, st, e, CLS_STAGE, 0);
  StageArgs a2 = stage_base(c, 0.5f);
  for (int v = 0; v < NV; ++v) {
    a2.src[v] = b.G1 + (long)v * c->ncell;
    a2.w0[v] = b.gas[v];
    a2.out[v] = b.gas[v];
  }
  e = time_begin(c, st);
  HIPCHK((launch_stage<SRC_PLAIN, true, SRC_PLAIN>(c, a2, nspan, st)));
  time_end(c, st, e, CLS_STAGE, 0);
  // ---- S_B(dt): fixed point on (B, v) with rho from the gas, in pairs of
  // launches; the speculative batch covers the last step's launch count
  const int maxit = c->cfg.max_fixed_point_iterations;
  int upto = c->last_k < 1 ? 1 : c->last_k;
  if (upto > maxit) upto = maxit;
  *klast = -1;
  // A pair's first iterate is needed only if it is the final one (the next
  // pair reads the second): each slot's pairs store it only where its
  // previous step ended (ctl.kprev), and step_phase3 regenerates it in the
  // steps where it is final after all (skip_mask: the speculative pairs)
  c->skip_mask = skip_mask_for(upto);
  for (int L = 0; L <= upto; L += 2) {
    HIPCHK(launch_mag_pair(c, b, L, nspan, st, L > 62 || !((c->skip_mask >> (L >> 1)) & 1u)));
    *klast = L + 1;
  }
  return AMX_OK;
}

// Final launch index of a running slot's fixed point from a readback of its
// control block, or -1 if the reference's loop would still run.
int slot_final(const amx_fv_ctx* c, const Ctl& h) {
  const int L = h.iters - 1;  // last executed pair
  unsigned r0 = 0, r1 = 0;
  for (int i = 0; i < NSLOT; ++i) {
    r0 = h.resp[L % NRES][i] > r0 ? h.resp[L % NRES][i] : r0;
    r1 = h.resp[(L + 1) % NRES][i] > r1 ? h.resp[(L + 1) % NRES][i] : r1;
  }
  float f0, f1;
  memcpy(&f0, &r0, sizeof(float));
  memcpy(&f1, &r1, sizeof(float));
  return pair_final(L, f0, f1, c->cfg.fixed_point_tol, c->cfg.max_fixed_point_iterations);
}

// After a readback of ctl[0, nspan) into hctl: 1 when every running slot has
// settled; *kmax = largest final launch index.
int fixed_point_settled(const amx_fv_ctx* c, int nspan, int* kmax) {
  int settled = 1;
  *kmax = 0;
  for (int s = 0; s < nspan; ++s) {
    const Ctl& h = c->hctl[s];
    if (!h.run) continue;
    int f = slot_final(c, h);
    if (f < 0) {
      settled = 0;
      f = h.iters;
    }
    if (f > *kmax) *kmax = f;
  }
  return settled;
}

// Launches further pairs after klast (two pairs per round) with a readback
// after each round until every running slot has settled.
amx_status fixed_point_more(amx_fv_ctx* c, const Bufs& b, int nspan, hipStream_t st, int* klast) {
  const int maxit = c->cfg.max_fixed_point_iterations;
  int kmax = 0;
  while (!fixed_point_settled(c, nspan, &kmax) && *klast < maxit) {
    for (int r = 0; r < 2 && *klast < maxit; ++r) {
      const int L = *klast + 1;
      HIPCHK(launch_mag_pair(c, b, L, nspan, st));
      *klast = L + 1;
    }
    HIPCHK(hipMemcpyAsync(c->hctl, c->ctl, sizeof(Ctl) * nspan, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
  }
  return AMX_OK;
}

// Phase 3: S_A(dt/2) on the magnetically updated gas (v <- v_K, energy fix on
// the fly); the final iterate of every slot is picked on the device.
// regen: queue the regeneration launch (a speculative call without it holds
// a step that ends on an unstored first iterate; see k_mag_final).
amx_status step_phase3(amx_fv_ctx* c, const Bufs& b, int nspan, hipStream_t st, int want_cfl, int spec,
                       int regen = 1) {
  if (!c->cfg.mhd) return AMX_OK;
  const int NV = c->NV;
  // the final iterate, where it is the first of a pair that did not store it
  const int rq = c->skip_mask && (regen || !spec);
  if (rq) HIPCHK(launch_mag_pair(c, b, 0, nspan, st, 1, 1));
  hipLaunchKernelGGL(k_mag_final, dim3(nspan), dim3(1), 0, st, c->ctl, c->cfg.fixed_point_tol,
                     c->cfg.max_fixed_point_iterations, spec, c->skip_mask, rq);
  HIPCHK(hipGetLastError());
  StageArgs b1 = stage_base(c, 0.5f);
  for (int v = 0; v < NV; ++v) {
    b1.src[v] = b.gas[v];
    b1.out[v] = b.G1 + (long)v * c->ncell;
  }
  b1.pfix = b.PF;
  b1.gated = spec;
  for (int j = 0; j < 4; ++j)
    for (int k = 0; k < 3; ++k) {
      b1.selv[j][k] = b.P[j][3 + k];
      b1.selb[j][k] = b.P[j][k];
    }
  int e = time_begin(c, st);
  HIPCHK((launch_stage<SRC_MAGFIX, false, SRC_PLAIN>(c, b1, nspan, st)));
  time_end(c, st, e, CLS_STAGE, spec);
  StageArgs b2 = stage_base(c, 0.5f);
  for (int v = 0; v < NV; ++v) {
    b2.src[v] = b.G1 + (long)v * c->ncell;
    b2.w0[v] = b.gas[v];
    b2.out[v] = b.gas[v];
  }
  for (int j = 0; j < 4; ++j)
    for (int k = 0; k < 3; ++k) {
      b2.selv[j][k] = b.P[j][3 + k];
      b2.selb[j][k] = b.P[j][k];
    }
  // the final B into the state, unless the loop keeps it in its iterate set
  for (int k = 0; k < 3; ++k) b2.bdst[k] = b.B[k];
  b2.copy_b = c->keep_b ? 0 : 1;
  b2.want_cfl = want_cfl;
  b2.pfix = b.PF;
  b2.gated = spec;
  e = time_begin(c, st);
  HIPCHK((launch_stage<SRC_PLAIN, true, SRC_PFIX>(c, b2, nspan, st)));
  time_end(c, st, e, CLS_STAGE, spec);
  return AMX_OK;
}

// One step of a single simulation (slot 0). Blocks once for the fixed-point
// readback; hctl[0] then holds dt and t of the step (t before the update).
amx_status do_step(amx_fv_ctx* c, float* state, hipStream_t st, int want_cfl, int* k_out) {
  const Bufs b = bufs(c, state);
  int klast = 0;
  size_t rec0 = c->pending.size();  // this step's timing records start here
  amx_status s = step_phase1(c, b, 1, st, want_cfl, &klast);
  if (s != AMX_OK) return s;
  int K = 0;
  if (c->cfg.mhd) {
    HIPCHK(hipMemcpyAsync(c->hctl, c->ctl, sizeof(Ctl), hipMemcpyDeviceToHost, st));
    HIPCHK(hipEventRecord(c->ev_status, st));
    // The second half is queued before the residuals reach the host, so the
    // GPU does not wait for it; its stage kernels run only if the launched
    // pairs settled the fixed point (k_mag_final sets held otherwise).
    s = step_phase3(c, b, 1, st, want_cfl, 1);
    if (s != AMX_OK) return s;
    HIPCHK(hipEventSynchronize(c->ev_status));
    // earlier steps' work has completed
    time_resolve_first(c, rec0);
    rec0 = 0;
    if (!fixed_point_settled(c, 1, &K)) {
      time_drop(c, rec0, [](const amx_fv_ctx::Rec& r) { return r.cls == CLS_STAGE && r.k == 1; });
      s = fixed_point_more(c, b, 1, st, &klast);
      if (s != AMX_OK) return s;
      fixed_point_settled(c, 1, &K);
      s = step_phase3(c, b, 1, st, want_cfl, 0);
      if (s != AMX_OK) return s;
    }
    // launches past the final one exited early
    time_drop(c, rec0, [K](const amx_fv_ctx::Rec& r) { return r.cls == CLS_MAG && r.k > K; });
    c->last_k = K;
  }
  ++c->tstep;
  if (k_out) *k_out = K;
  return AMX_OK;
}

}  // namespace

extern "C" amx_status amx_fv_workspace_bytes(const amx_fv_config* cfg, size_t* bytes) {
  amx_status s = validate(cfg);
  if (s != AMX_OK) return s;
  if (!bytes) return fail(AMX_ERR_INVALID, "null output%s");
  *bytes = sizeof(float) * (size_t)cells(cfg) * ws_fields(cfg, 0) + sizeof(Ctl) + sizeof(Phys);
  if (cfg->riemann == AMX_HYBRID_HLLC || cfg->riemann == AMX_AM_HLLC) *bytes += (size_t)cells(cfg);
  if (cfg->limiter == AMX_VAN_ALBADA_PP) *bytes += sizeof(float) * (size_t)cells(cfg) * (2 * cfg->dim + 1);
  return AMX_OK;
}

extern "C" amx_status amx_fv_create(const amx_fv_config* cfg, amx_fv_ctx** out) {
  return make_ctx(cfg, 1, 0, out);
}

extern "C" amx_status amx_fv_destroy(amx_fv_ctx* c) {
  free_ctx(c);
  return AMX_OK;
}

extern "C" amx_status amx_fv_set_timing(amx_fv_ctx* c, int enable) {
  if (!c) return fail(AMX_ERR_INVALID, "null context%s");
  HIPCHK(hipDeviceSynchronize());
  time_drop(c, 0, [](const amx_fv_ctx::Rec&) { return true; });
  c->timing = enable > 0 ? enable : 0;
  c->tstep = 0;
  c->ms[0] = c->ms[1] = 0.0;
  c->launches[0] = c->launches[1] = 0;
  return AMX_OK;
}

extern "C" amx_status amx_fv_get_timing(amx_fv_ctx* c, amx_fv_timing* out) {
  if (!c || !out) return fail(AMX_ERR_INVALID, "null argument%s");
  HIPCHK(hipDeviceSynchronize());
  time_resolve(c);
  out->stage_ms = c->ms[CLS_STAGE];
  out->stage_launches = c->launches[CLS_STAGE];
  out->mag_ms = c->ms[CLS_MAG];
  out->mag_launches = c->launches[CLS_MAG];
  return AMX_OK;
}

extern "C" amx_status amx_fv_set_step_hook(amx_fv_ctx* c, amx_step_hook hook, void* user) {
  if (!c) return fail(AMX_ERR_INVALID, "null context%s");
  if (c->own_state) return fail(AMX_ERR_INVALID, "pools have no step hook%s");
  c->hook = hook;
  c->hook_user = user;
  return AMX_OK;
}

extern "C" amx_status amx_fv_set_physics(amx_fv_ctx* c, float gamma, float C_cfl) {
  if (!c) return fail(AMX_ERR_INVALID, "null context%s");
  if (c->own_state) return fail(AMX_ERR_INVALID, "pool slots take gamma at amx_fv_pool_load%s");
  if (!(gamma > 1.0f) || !(C_cfl > 0.0f)) return fail(AMX_ERR_INVALID, "need gamma > 1 and C_cfl > 0%s");
  HIPCHK(hipDeviceSynchronize());  // work in flight still reads the old values
  c->cfg.gamma = gamma;
  c->cfg.C_cfl = C_cfl;
  c->hph[0] = make_phys(gamma, c->cfg.grid_spacing);
  HIPCHK(hipMemcpy(c->dph, c->hph, sizeof(Phys), hipMemcpyHostToDevice));
  return AMX_OK;
}

extern "C" amx_status amx_fv_cfl_dt(amx_fv_ctx* c, const float* state, float* d_dt, void* stream) {
  if (!c || !state || !d_dt) return fail(AMX_ERR_INVALID, "null argument%s");
  if (c->own_state) return fail(AMX_ERR_INVALID, "a pool context is stepped with amx_fv_pool_*%s");
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_zero_cfl, dim3(1), dim3(1), 0, st, c->ctl);
  HIPCHK(hipGetLastError());
  HIPCHK(launch_cfl(c, state, 0, 1, st));
  hipLaunchKernelGGL(k_cfl_out, dim3(1), dim3(1), 0, st, c->ctl, c->cfg.C_cfl, c->dx, c->cfg.dim, d_dt);
  HIPCHK(hipGetLastError());
  return AMX_OK;
}

extern "C" amx_status amx_fv_evolve_state(amx_fv_ctx* c, float* state, const float* d_dt, int* iterations,
                                          void* stream) {
  if (!c || !state || !d_dt) return fail(AMX_ERR_INVALID, "null argument%s");
  if (c->own_state) return fail(AMX_ERR_INVALID, "a pool context is stepped with amx_fv_pool_*%s");
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_set_dt, dim3(1), dim3(1), 0, st, c->ctl, d_dt);
  HIPCHK(hipGetLastError());
  int K = 0;
  amx_status s = do_step(c, state, st, 0, &K);
  if (s != AMX_OK) return s;
  if (iterations) *iterations = K;
  return AMX_OK;
}

// ---- Allocation-free step (amx_fv_evolve_state_noalloc). The workspace,
// the control block and the physics constants of one step live in `scratch`,
// a file-scope device array each GPU holds from the library's load until the
// process ends; the host addresses it through the pointer hipGetSymbolAddress
// returns on that device (the array's name is not a device address in host
// code). Calls on one device share the array: calls that overlap on a device
// race. Each call re-initialises what it reads before writing it (control
// block, constants; the stage and iterate buffers are written before read).
namespace {
constexpr size_t NOALLOC_SCRATCH_BYTES = AMX_NOALLOC_SCRATCH_BYTES;
}  // namespace

static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int NOALLOC_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch[NOALLOC_MAX_DEVICES];

// Device address of this device's `scratch`, looked up once per device.
amx_status scratch_base(unsigned char** out) {
  int dev = -1;
  HIPCHK(hipGetDevice(&dev));
  if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return fail(AMX_ERR_INVALID, "device index past the scratch table%s");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return fail(AMX_ERR_HIP, "hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
    g_scratch[dev] = (unsigned char*)p;
  }
  *out = g_scratch[dev];
  return AMX_OK;
}

struct ScratchLayout {
  size_t ctl, phys, ws, amg, ppf, total;
};

ScratchLayout scratch_layout(const amx_fv_config* cfg) {
  auto up = [](size_t x) { return (x + 255) & ~(size_t)255; };
  const size_t n = (size_t)cells(cfg);
  ScratchLayout L;
  size_t o = 0;
  L.ctl = o;
  o = up(o + sizeof(Ctl));
  L.phys = o;
  o = up(o + sizeof(Phys));
  L.ws = o;
  o = up(o + sizeof(float) * n * (size_t)ws_fields(cfg, 0));
  L.amg = o;
  if (cfg->riemann == AMX_HYBRID_HLLC || cfg->riemann == AMX_AM_HLLC) o = up(o + n);
  L.ppf = o;
  if (cfg->limiter == AMX_VAN_ALBADA_PP) o = up(o + sizeof(float) * n * (size_t)(2 * cfg->dim + 1));
  L.total = o;
  return L;
}

}  // namespace

// Control block, constants and step size of the allocation-free step: slot 0
// live and running, B0 in the state's rows, fixed point and CFL slots cleared.
__global__ void k_noalloc_begin(Ctl* c, Phys* ph, Phys v, const float* dtp) {
  *ph = v;
  c->t = 0.f;
  c->steps = 0;
  c->live = 1;
  c->t_end = 0.f;
  c->held = 0;
  c->stall = 0;
  c->iters_sum = 0;
  c->iters_max = 0;
  c->iters_bmax = 0;
  c->dt = *dtp;
  c->run = 1;
  c->bset = -1;
  c->kprev = -1;
  c->regen_ev = 0;
  for (int i = 0; i < 4; ++i) c->cfl[i] = 0u;
  zero_cflp(c);
  fixed_point_reset(c);
}

extern "C" amx_status amx_fv_evolve_state_noalloc(const amx_fv_config* cfg, float* state, const float* d_dt,
                                                  int* iterations, void* stream) {
  amx_status s = validate(cfg);
  if (s != AMX_OK) return s;
  if (!state || !d_dt) return fail(AMX_ERR_INVALID, "null argument%s");
  const ScratchLayout L = scratch_layout(cfg);
  if (L.total > NOALLOC_SCRATCH_BYTES) {
    char msg[160];
    snprintf(msg, sizeof(msg), "this grid needs %zu bytes of scratch, the allocation-free step has %zu",
             L.total, NOALLOC_SCRATCH_BYTES);
    return fail(AMX_ERR_INVALID, "%s", msg);
  }
  unsigned char* base = nullptr;
  s = scratch_base(&base);
  if (s != AMX_OK) return s;
  // host-side context on the stack: geometry only, its containers stay empty
  amx_fv_ctx c{};
  init_geometry(&c, cfg, 1, 0);
  c.mem = (float*)(base + L.ws);
  c.ctl = (Ctl*)(base + L.ctl);
  c.dph = (Phys*)(base + L.phys);
  c.amg = (cfg->riemann == AMX_HYBRID_HLLC || cfg->riemann == AMX_AM_HLLC) ? base + L.amg : nullptr;
  c.ppf = cfg->limiter == AMX_VAN_ALBADA_PP ? (float*)(base + L.ppf) : nullptr;
  Ctl hctl;
  c.hctl = &hctl;
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_noalloc_begin, dim3(1), dim3(1), 0, st, c.ctl, c.dph,
                     make_phys(cfg->gamma, cfg->grid_spacing), d_dt);
  HIPCHK(hipGetLastError());
  const Bufs b = bufs(&c, state);
  int klast = 0, K = 0;
  s = step_phase1(&c, b, 1, st, 0, &klast);
  if (s != AMX_OK) return s;
  if (cfg->mhd) {
    // the fixed point's residuals, read into host stack memory
    HIPCHK(hipMemcpyAsync(c.hctl, c.ctl, sizeof(Ctl), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    if (!fixed_point_settled(&c, 1, &K)) {
      s = fixed_point_more(&c, b, 1, st, &klast);
      if (s != AMX_OK) return s;
      fixed_point_settled(&c, 1, &K);
    }
    s = step_phase3(&c, b, 1, st, 0, 0);
    if (s != AMX_OK) return s;
  }
  if (iterations) *iterations = K;
  return AMX_OK;
}

namespace {

// Layout of a single-simulation FV state for the snapshot diagnostics.
amx_diag_desc fv_diag_desc(const amx_fv_ctx* c, double box_size) {
  amx_diag_desc d;
  d.num_vars = c->NS;
  d.dim = c->cfg.dim;
  d.nx = c->cfg.nx;
  d.ny = c->cfg.ny;
  d.nz = c->cfg.dim == 3 ? c->cfg.nz : 1;
  d.rho_row = 0;
  d.vel_row = 1;
  d.p_row = c->cfg.dim + 1;
  d.mag_row = c->cfg.mhd ? c->cfg.dim + 2 : 0;
  d.div_mode = c->cfg.mhd ? AMX_DIV_CENTRAL : AMX_DIV_NONE;
  d.gamma = c->cfg.gamma;
  d.grid_spacing = c->cfg.grid_spacing;
  d.box_size = box_size;
  return d;
}

// The final B of slot 0 from its iterate set into the state's B rows; gate:
// nothing when a step of the queued batch stalled.
hipError_t copy_b_home(amx_fv_ctx* c, const Bufs& bl, hipStream_t st, int gate) {
  BsetArgs ba;
  for (int j = 0; j < 4; ++j)
    for (int k = 0; k < 3; ++k) ba.PB[j][k] = bl.P[j][k];
  for (int k = 0; k < 3; ++k) ba.dst[k] = bl.B[k];
  ba.n = c->ncell;
  ba.gate = gate;
  long blocks = (c->ncell + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  hipLaunchKernelGGL(k_copy_bset, dim3(blocks), dim3(256), 0, st, c->ctl, ba);
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_bset_clear, dim3(1), dim3(1), 0, st, c->ctl, gate);
  return hipGetLastError();
}

// The step loop of _time_integration (time_integration.py:629-801); with a
// snapshot driver, rows are taken as in time_integration.py:539-610, 736-745.
amx_status fv_loop(amx_fv_ctx* c, float* state, const amx_fv_run* run, snap::Driver* snp, amx_fv_stats* stats,
                   hipStream_t st) {
  hipLaunchKernelGGL(k_slot_reset, dim3(1), dim3(1), 0, st, c->ctl, 0, run->t_end, 1, 0);
  HIPCHK(hipGetLastError());
  // Without snapshot rows or a step hook nothing reads the state's B rows
  // between steps: the final B of each step stays in its iterate set (the
  // next step's B0) and is copied into the state once, after the loop.
  struct KeepB {
    amx_fv_ctx* c;
    ~KeepB() { c->keep_b = 0; }
  } keep_guard{c};
  c->keep_b = c->cfg.mhd && !snp && !c->hook ? 1 : 0;
  const Bufs bl = bufs(c, state);
  if (!run->fixed_timestep) HIPCHK(launch_cfl(c, state, 0, 1, st));
  StepParams sp;
  sp.fixed = run->fixed_timestep;
  sp.num_timesteps = run->num_timesteps;
  sp.exact_end = run->exact_end_time;
  sp.C = c->cfg.C_cfl;
  sp.dx = c->dx;
  sp.dim = c->cfg.dim;
  // the end-time clamp is off with snapshot time points (time_integration.py:677)
  if (snp && snp->spec->use_timepoints) sp.exact_end = 0;
  long steps = 0, iters = 0;
  int kmax = 0;
  float t = 0.f;
  // Hydro steps without snapshot rows or a hook need nothing from the host
  // between steps: they are queued in batches and t is read once per batch
  // (k_step_begin clears ctl.run past the end, so surplus steps do nothing).
  const bool batch = !c->cfg.mhd && !snp && !c->hook;
  // MHD steps of such loops are queued in batches too, each with its
  // speculative fixed-point launches and its second half behind them; the
  // device decides whether each step's launches settled the fixed point. A
  // step that needs more launches stalls the batch (the steps behind it do
  // nothing); the host reads the control block once per batch, finishes a
  // stalled step with the further launches and goes on. Event timing keeps
  // the step-by-step path, which sorts out early-exited launches per step.
  const bool mbatch = c->cfg.mhd && c->keep_b && c->timing == 0;
  constexpr long BATCH = 16;
  bool b_home = false;  // the final B is in the state's rows and hctl is current
  for (;;) {
    if (run->fixed_timestep) {
      if (steps >= run->num_timesteps) break;
    } else if (!(t < run->t_end)) {
      break;
    }
    if (run->max_steps >= 0 && steps >= run->max_steps) break;
    if (batch) {
      long n = BATCH;
      if (run->max_steps >= 0 && run->max_steps - steps < n) n = run->max_steps - steps;
      if (run->fixed_timestep && run->num_timesteps - steps < n) n = run->num_timesteps - steps;
      hipLaunchKernelGGL(k_step_begin, dim3(1), dim3(1), 0, st, c->ctl, sp);
      HIPCHK(hipGetLastError());
      for (long i = 0; i < n; ++i) {
        amx_status s = do_step(c, state, st, !run->fixed_timestep, nullptr);
        if (s != AMX_OK) return s;
        // this step's end and the next one's beginning
        if (i + 1 < n)
          hipLaunchKernelGGL(k_step_end_begin, dim3(1), dim3(1), 0, st, c->ctl, 0, sp);
        else
          hipLaunchKernelGGL(k_step_end, dim3(1), dim3(1), 0, st, c->ctl, 0);
        HIPCHK(hipGetLastError());
      }
      HIPCHK(hipMemcpyAsync(c->hctl, c->ctl, offsetof(Ctl, cflp), hipMemcpyDeviceToHost, st));
      HIPCHK(hipStreamSynchronize(st));
      time_resolve(c);
      t = c->hctl[0].t;
      steps = c->hctl[0].steps;
      continue;
    }
    if (mbatch) {
      // steps left when the loop has a step bound; a tail of up to half a
      // batch joins the batch before it (one readback fewer)
      long left = -1;
      if (run->max_steps >= 0) left = run->max_steps - steps;
      if (run->fixed_timestep && (left < 0 || run->num_timesteps - steps < left)) left = run->num_timesteps - steps;
      const long n = left >= 0 && left <= BATCH + BATCH / 2 ? left : BATCH;
      int klast = 0;
      StepParams spi = sp;
      spi.batch_first = 1;
      hipLaunchKernelGGL(k_step_begin, dim3(1), dim3(1), 0, st, c->ctl, spi);
      HIPCHK(hipGetLastError());
      spi.batch_first = 0;
      for (long i = 0; i < n; ++i) {
        amx_status s = step_phase1(c, bl, 1, st, !run->fixed_timestep, &klast);
        if (s != AMX_OK) return s;
        s = step_phase3(c, bl, 1, st, !run->fixed_timestep, 1, c->regen_q);
        if (s != AMX_OK) return s;
        // this step's end and the next one's beginning
        if (i + 1 < n)
          hipLaunchKernelGGL(k_step_end_begin, dim3(1), dim3(1), 0, st, c->ctl, c->keep_b, spi);
        else
          hipLaunchKernelGGL(k_step_end, dim3(1), dim3(1), 0, st, c->ctl, c->keep_b);
        HIPCHK(hipGetLastError());
      }
      // the loop's last batch: B goes home behind it, unless a step stalled
      const bool last = n == left;
      if (last) HIPCHK(copy_b_home(c, bl, st, 1));
      HIPCHK(hipMemcpyAsync(c->hctl, c->ctl, sizeof(Ctl), hipMemcpyDeviceToHost, st));
      HIPCHK(hipStreamSynchronize(st));
      b_home = last && !c->hctl[0].stall;
      // regeneration launches in the next batch only if a step of this one
      // ended on an unstored first iterate (otherwise such a step stalls)
      c->regen_q = c->hctl[0].regen_ev > 0;
      if (c->hctl[0].stall) {
        // the stalled step's launches up to klast have run (every launch
        // queued behind it exited at once): the further pairs, then its
        // second half and its end
        int K = 0;
        hipLaunchKernelGGL(k_resume, dim3(1), dim3(1), 0, st, c->ctl);
        HIPCHK(hipGetLastError());
        c->hctl[0].run = 1;
        amx_status s = fixed_point_more(c, bl, 1, st, &klast);
        if (s != AMX_OK) return s;
        fixed_point_settled(c, 1, &K);
        s = step_phase3(c, bl, 1, st, !run->fixed_timestep, 0);
        if (s != AMX_OK) return s;
        hipLaunchKernelGGL(k_step_end, dim3(1), dim3(1), 0, st, c->ctl, c->keep_b);
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(c->hctl, c->ctl, offsetof(Ctl, cfl), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
      }
      t = c->hctl[0].t;
      steps = c->hctl[0].steps;
      iters = c->hctl[0].iters_sum;
      kmax = c->hctl[0].iters_max;
      // the next batch queues launches for the largest count of this one
      if (c->hctl[0].iters_bmax > 0) c->last_k = c->hctl[0].iters_bmax;
      continue;
    }
    if (snp) {
      amx_status s = snp->begin(t, state, st);
      if (s != AMX_OK) return s;
      sp.clamp_tp = snp->clamp();
      if (sp.clamp_tp) sp.tp = snp->clamp_target();
    }
    hipLaunchKernelGGL(k_step_begin, dim3(1), dim3(1), 0, st, c->ctl, sp);
    HIPCHK(hipGetLastError());
    if (c->hook && c->hook(c->hook_user, state, &c->ctl[0].dt, steps, st) != 0)
      return fail(AMX_ERR_INVALID, "the step hook stopped the run%s");
    int K = 0;
    amx_status s = do_step(c, state, st, !run->fixed_timestep, &K);
    if (s != AMX_OK) return s;
    hipLaunchKernelGGL(k_step_end, dim3(1), dim3(1), 0, st, c->ctl, c->keep_b);
    HIPCHK(hipGetLastError());
    if (!c->cfg.mhd) {
      // hydro steps have no fixed-point readback: read dt here
      HIPCHK(hipMemcpyAsync(c->hctl, c->ctl, offsetof(Ctl, cflp), hipMemcpyDeviceToHost, st));
      HIPCHK(hipStreamSynchronize(st));
      time_resolve(c);
      t = c->hctl[0].t;
    } else {
      // hctl holds dt and t from the fixed-point readback (before k_step_end)
      t = c->hctl[0].t + c->hctl[0].dt;
    }
    iters += K;
    if (K > kmax) kmax = K;
    ++steps;
    if (snp) {
      s = snp->end(t, state, st);
      if (s != AMX_OK) return s;
    }
  }
  if (c->keep_b && !b_home) {
    HIPCHK(copy_b_home(c, bl, st, 0));
  }
  if (!b_home) {
    HIPCHK(hipMemcpyAsync(c->hctl, c->ctl, offsetof(Ctl, cflp), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
  }
  time_resolve(c);
  if (snp) snp->finish(steps);
  if (stats) {
    stats->steps = steps;
    stats->t = c->hctl[0].t;
    stats->fixed_point_iterations = iters;
    stats->max_fixed_point_iterations = kmax;
  }
  return AMX_OK;
}

}  // namespace

extern "C" amx_status amx_fv_time_integration(amx_fv_ctx* c, float* state, const amx_fv_run* run,
                                              amx_fv_stats* stats, void* stream) {
  if (!c || !state || !run) return fail(AMX_ERR_INVALID, "null argument%s");
  if (c->own_state) return fail(AMX_ERR_INVALID, "a pool context is stepped with amx_fv_pool_*%s");
  if (run->fixed_timestep && run->num_timesteps <= 0) return fail(AMX_ERR_INVALID, "num_timesteps must be positive%s");
  return fv_loop(c, state, run, nullptr, stats, (hipStream_t)stream);
}

extern "C" amx_status amx_fv_time_integration_snapshots(amx_fv_ctx* c, float* state, const amx_fv_run* run,
                                                        const amx_snapshot_spec* spec, amx_snapshot_out* out,
                                                        amx_fv_stats* stats, void* stream) {
  if (!c || !state || !run || !spec) return fail(AMX_ERR_INVALID, "null argument%s");
  if (c->own_state) return fail(AMX_ERR_INVALID, "a pool context is stepped with amx_fv_pool_*%s");
  if (run->fixed_timestep && run->num_timesteps <= 0) return fail(AMX_ERR_INVALID, "num_timesteps must be positive%s");
  snap::Driver d;
  d.spec = spec;
  d.out = out;
  d.desc = fv_diag_desc(c, spec->box_size);
  d.t_end = run->t_end;
  d.work = nullptr;
  amx_status s = snap::validate(spec, out, d.desc);
  if (s != AMX_OK) return s;
  if (!spec->callback) {
    if (!c->snapwork && hipMalloc(&c->snapwork, snap::work_bytes()) != hipSuccess) {
      c->snapwork = nullptr;
      return fail(AMX_ERR_NOMEM, "snapshot workspace allocation failed%s");
    }
    d.work = c->snapwork;
  }
  return fv_loop(c, state, run, &d, stats, (hipStream_t)stream);
}

// ============================================================================
// Reverse mode of one hydro SSP-RK2 step
// ============================================================================

namespace {

template <int DIM, int AXI>
void launch_adj_iface(const amx_fv_ctx* c, dim3 grid, dim3 blk, hipStream_t st, const float* Win, const float* NU,
                      float* con, adj::AdjGeom g, Phys ph, bool duals) {
  if (c->cfg.riemann == AMX_HLLC && duals)
    hipLaunchKernelGGL((adj::k_adj_iface<DIM, AXI, AMX_HLLC, true>), grid, blk, 0, st, Win, NU, con, c->ctl, g, ph);
  else if (c->cfg.riemann == AMX_HLLC)
    hipLaunchKernelGGL((adj::k_adj_iface<DIM, AXI, AMX_HLLC, false>), grid, blk, 0, st, Win, NU, con, c->ctl, g, ph);
  else if (duals)
    hipLaunchKernelGGL((adj::k_adj_iface<DIM, AXI, AMX_HLL, true>), grid, blk, 0, st, Win, NU, con, c->ctl, g, ph);
  else
    hipLaunchKernelGGL((adj::k_adj_iface<DIM, AXI, AMX_HLL, false>), grid, blk, 0, st, Win, NU, con, c->ctl, g, ph);
}

template <int DIM>
amx_status adjoint_step_t(amx_fv_ctx* c, const float* w0, const float* lam_out, float* lam_in, hipStream_t st) {
  constexpr int NV = DIM + 2;
  const long n = c->ncell;
  adj::AdjGeom g;
  if (DIM == 3) {
    g.n[0] = c->cfg.nx;
    g.n[1] = c->cfg.ny;
    g.n[2] = c->cfg.nz;
    g.st[2] = 1;
    g.st[1] = c->cfg.nz;
    g.st[0] = (long)c->cfg.ny * c->cfg.nz;
  } else {
    g.n[0] = c->cfg.nx;
    g.n[1] = c->cfg.ny;
    g.n[2] = 1;
    g.st[1] = 1;
    g.st[0] = c->cfg.ny;
    g.st[2] = n;
  }
  g.nc = n;
  float* G1 = c->mem;                  // W1 = E(W0) (the context's stage buffer)
  float* G2 = c->adj;                  // W2 = E(W1)
  float* L2 = G2 + (long)NV * n;       // lam of W2
  float* NU = L2 + (long)NV * n;
  float* L1 = NU + (long)NV * n;       // lam of W1
  float* CON = L1 + (long)NV * n;      // DIM x 4 x NV interface contributions
  // recompute the two stage states with the forward kernel (dt in ctl[0])
  StageArgs a1 = stage_base(c, 1.0f);
  for (int v = 0; v < NV; ++v) {
    a1.src[v] = w0 + (long)v * n;
    a1.out[v] = G1 + (long)v * n;
  }
  HIPCHK((launch_stage<SRC_PLAIN, false, SRC_PLAIN>(c, a1, 1, st)));
  StageArgs a2 = stage_base(c, 1.0f);
  for (int v = 0; v < NV; ++v) {
    a2.src[v] = G1 + (long)v * n;
    a2.out[v] = G2 + (long)v * n;
  }
  HIPCHK((launch_stage<SRC_PLAIN, false, SRC_PLAIN>(c, a2, 1, st)));
  const Phys ph = c->hph[0];  // the interface kernels read dt from ctl[0]
  // AMX_ADJ_DUAL=1: the flux Jacobians by dual numbers (a cross-check of the reverse sweep)
  const bool duals = env_int("AMX_ADJ_DUAL", 0) != 0;
  long blocks = (n + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  const dim3 grid(blocks), blk(256);
  hipLaunchKernelGGL(adj::k_adj_heun<DIM>, grid, blk, 0, st, w0, G2, lam_out, lam_in, L2, g, ph);
  HIPCHK(hipGetLastError());
  // two stages in reverse: (W1 -> W2, lam = L2) then (W0 -> W1, lam = L1)
  for (int stage = 1; stage >= 0; --stage) {
    const float* Win = stage == 1 ? G1 : w0;
    const float* Wout = stage == 1 ? G2 : G1;
    const float* lam = stage == 1 ? L2 : L1;
    float* dst = stage == 1 ? L1 : lam_in;
    hipLaunchKernelGGL(adj::k_adj_nu<DIM>, grid, blk, 0, st, Wout, lam, NU, g, ph);
    HIPCHK(hipGetLastError());
    for (int ax = 0; ax < DIM; ++ax) {
      float* con = CON + (long)ax * 4 * NV * n;
      if (ax == 0)
        launch_adj_iface<DIM, 0>(c, grid, blk, st, Win, NU, con, g, ph, duals);
      else if (ax == 1)
        launch_adj_iface<DIM, 1>(c, grid, blk, st, Win, NU, con, g, ph, duals);
      else
        launch_adj_iface<DIM, 2>(c, grid, blk, st, Win, NU, con, g, ph, duals);
      HIPCHK(hipGetLastError());
    }
    hipLaunchKernelGGL(adj::k_adj_gather<DIM>, grid, blk, 0, st, Win, NU, CON, dst, stage == 0 ? 1 : 0, g, ph);
    HIPCHK(hipGetLastError());
  }
  return AMX_OK;
}

}  // namespace

// Fields of the MHD reverse mode's workspace before the iterates: state copy
// (NS), W_B (NV), lam of W_B (NV), lam of W_A (NV), two 6-row cotangents, X
// and Y (3 + 3), the accumulated cotangents of B0, v0, rho (7).
long adjm_fixed_fields(const amx_fv_ctx* c) { return (long)c->NS + 3L * c->NV + 12 + 6 + 7; }

amx_status adjm_reserve(amx_fv_ctx* c, int iters, hipStream_t st) {
  if (c->adjm && iters <= c->adjm_iters) return AMX_OK;
  int cap = c->adjm_iters > 0 ? c->adjm_iters : 8;
  while (cap < iters) cap *= 2;
  const size_t bytes = sizeof(float) * (size_t)(adjm_fixed_fields(c) + 6L * cap) * c->ncell;
  float* m = nullptr;
  if (hipMalloc((void**)&m, bytes) != hipSuccess) return fail(AMX_ERR_NOMEM, "reverse-mode workspace allocation failed%s");
  if (c->adjm) {
    HIPCHK(hipStreamSynchronize(st));
    HIPCHK(hipMemcpy(m, c->adjm, sizeof(float) * (size_t)(adjm_fixed_fields(c) + 6L * c->adjm_iters) * c->ncell,
                     hipMemcpyDeviceToDevice));
    (void)hipFree(c->adjm);
  }
  c->adjm = m;
  c->adjm_iters = cap;
  return AMX_OK;
}

// Reverse mode of one MHD step S_A(dt/2) S_B(dt) S_A(dt/2)
// (evolve_state.py:496-600): the forward pass is recomputed from a copy of
// the state with the forward kernels (so the stage states, the fixed-point
// iterates and their count are the forward's own), keeping every iterate;
// then the second gas half, the magnetic sub-step and the first gas half
// are reversed (adj::k_adjm_*, adjoint_step_t at dt/2).
template <int DIM>
amx_status adjoint_step_mhd(amx_fv_ctx* c, const float* state, const float* lam_out, float* lam_in,
                            hipStream_t st) {
  constexpr int NV = DIM + 2;
  const long n = c->ncell;
  amx_status r = adjm_reserve(c, 8, st);
  if (r != AMX_OK) return r;
  HIPCHK(hipMemcpyAsync(c->adjm, state, sizeof(float) * (size_t)c->NS * n, hipMemcpyDeviceToDevice, st));
  Bufs b = bufs(c, c->adjm);
  // ---- forward: S_A(dt/2) into S's gas rows (as step_phase1 does)
  StageArgs a1 = stage_base(c, 0.5f);
  for (int v = 0; v < NV; ++v) {
    a1.src[v] = b.gas[v];
    a1.out[v] = b.G1 + (long)v * n;
  }
  HIPCHK((launch_stage<SRC_PLAIN, false, SRC_PLAIN>(c, a1, 1, st)));
  StageArgs a2 = stage_base(c, 0.5f);
  for (int v = 0; v < NV; ++v) {
    a2.src[v] = b.G1 + (long)v * n;
    a2.w0[v] = b.gas[v];
    a2.out[v] = b.gas[v];
  }
  HIPCHK((launch_stage<SRC_PLAIN, true, SRC_PLAIN>(c, a2, 1, st)));
  // ---- forward: the fixed point pair by pair, every iterate kept (launch k's
  // output at iterate slot k), until the residuals settle it
  c->skip_mask = 0u;
  int F = -1;
  for (int L = 0; F < 0; L += 2) {
    r = adjm_reserve(c, L + 2, st);  // may move the workspace: the buffers are taken again
    if (r != AMX_OK) return r;
    b = bufs(c, c->adjm);
    float* IT = c->adjm + adjm_fixed_fields(c) * n;
    HIPCHK(launch_mag_pair(c, b, L, 1, st, 1, 0));
    for (int k = L; k <= L + 1; ++k)
      for (int f = 0; f < 6; ++f)
        HIPCHK(hipMemcpyAsync(IT + (long)(6 * k + f) * n, b.P[iterate_set(-1, k)][f], sizeof(float) * n,
                              hipMemcpyDeviceToDevice, st));
    HIPCHK(hipMemcpyAsync(c->hctl, c->ctl, sizeof(Ctl), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    F = slot_final(c, c->hctl[0]);
  }
  float* S2 = c->adjm;
  const Bufs bb = bufs(c, S2);
  float* WB = S2 + (long)c->NS * n;
  float* LWB = WB + (long)NV * n;
  float* LWA = LWB + (long)NV * n;
  float* LB0 = LWA + (long)NV * n;  // two 6-row cotangents
  float* LB1 = LB0 + 6L * n;
  float* XY = LB1 + 6L * n;
  float* ACC = XY + 6L * n;
  float* IT = ACC + 7L * n;
  adj::AdjGeom g;
  if (DIM == 3) {
    g.n[0] = c->cfg.nx;
    g.n[1] = c->cfg.ny;
    g.n[2] = c->cfg.nz;
    g.st[2] = 1;
    g.st[1] = c->cfg.nz;
    g.st[0] = (long)c->cfg.ny * c->cfg.nz;
  } else {
    g.n[0] = c->cfg.nx;
    g.n[1] = c->cfg.ny;
    g.n[2] = 1;
    g.st[1] = 1;
    g.st[0] = c->cfg.ny;
    g.st[2] = n;
  }
  g.nc = n;
  long blocks = (n + 255) / 256;
  if (blocks > 8192) blocks = 8192;
  const dim3 grid(blocks), blk(256);
  hipLaunchKernelGGL(adj::k_adjm_wb<DIM>, grid, blk, 0, st, bb.gas[0], IT + 6L * F * n, WB, g);
  HIPCHK(hipGetLastError());
  // ---- reverse: second gas half (dt/2) from lam_out's gas rows
  hipLaunchKernelGGL(adj::k_scale_dt, dim3(1), dim3(1), 0, st, c->ctl, 0.5f);
  HIPCHK(hipGetLastError());
  amx_status s = adjoint_step_t<DIM>(c, WB, lam_out, LWB, st);
  if (s != AMX_OK) return s;
  hipLaunchKernelGGL(adj::k_scale_dt, dim3(1), dim3(1), 0, st, c->ctl, 2.0f);
  HIPCHK(hipGetLastError());
  // ---- reverse: the magnetic sub-step, launch F down to 0
  hipLaunchKernelGGL(adj::k_adjm_init<DIM>, grid, blk, 0, st, lam_out + (long)NV * n, LWB, LB0, ACC, g);
  HIPCHK(hipGetLastError());
  adj::MagAdjArgs m;
  memset(&m, 0, sizeof(m));
  for (int k = 0; k < 3; ++k) {
    m.B0[k] = bb.B[k];
    m.v0[k] = k < DIM ? bb.gas[1 + k] : nullptr;
  }
  m.rho = bb.gas[0];
  m.X = XY;
  m.Y = XY + 3L * n;
  m.acc = ACC;
  m.ctl = c->ctl;
  m.inv_dx = c->hph[0].inv_dx;
  m.euler = c->cfg.magnetic_integrator == AMX_IMPLICIT_EULER;
  m.g = g;
  float* lcur = LB0;
  float* lnext = LB1;
  for (int j = F; j >= 0; --j) {
    m.first = j == 0;
    for (int k = 0; k < 3; ++k) {
      m.Bj[k] = j == 0 ? bb.B[k] : IT + (long)(6 * (j - 1) + k) * n;
      m.vj[k] = j == 0 ? m.v0[k] : IT + (long)(6 * (j - 1) + 3 + k) * n;
    }
    m.lB = lcur;
    m.lBn = lnext;
    hipLaunchKernelGGL(adj::k_adjm_xy<DIM>, grid, blk, 0, st, m);
    HIPCHK(hipGetLastError());
    hipLaunchKernelGGL(adj::k_adjm_back<DIM>, grid, blk, 0, st, m);
    HIPCHK(hipGetLastError());
    std::swap(lcur, lnext);
  }
  hipLaunchKernelGGL(adj::k_adjm_final<DIM>, grid, blk, 0, st, LWB, lcur, ACC, LWA, lam_in + (long)NV * n, g);
  HIPCHK(hipGetLastError());
  // ---- reverse: first gas half (dt/2) at the step's input state
  hipLaunchKernelGGL(adj::k_scale_dt, dim3(1), dim3(1), 0, st, c->ctl, 0.5f);
  HIPCHK(hipGetLastError());
  s = adjoint_step_t<DIM>(c, state, LWA, lam_in, st);
  if (s != AMX_OK) return s;
  hipLaunchKernelGGL(adj::k_scale_dt, dim3(1), dim3(1), 0, st, c->ctl, 2.0f);
  HIPCHK(hipGetLastError());
  return AMX_OK;
}

extern "C" amx_status amx_fv_adjoint_step(amx_fv_ctx* c, const float* state, const float* d_dt, const float* lam_out,
                                          float* lam_in, void* stream) {
  if (!c || !state || !d_dt || !lam_out || !lam_in) return fail(AMX_ERR_INVALID, "null argument%s");
  if (c->own_state) return fail(AMX_ERR_INVALID, "a pool context has no reverse mode%s");
  if ((c->cfg.riemann != AMX_HLL && c->cfg.riemann != AMX_HLLC) || c->cfg.limiter != AMX_MINMOD)
    return fail(AMX_ERR_UNSUPPORTED, "reverse mode covers the HLL / HLLC fluxes with the minmod limiter%s");
  if (c->bcx) return fail(AMX_ERR_UNSUPPORTED, "reverse mode covers periodic boundaries only%s");
  if (lam_in == lam_out || lam_in == state) return fail(AMX_ERR_INVALID, "lam_in must not alias its inputs%s");
  const int NV = c->NV, dim = c->cfg.dim;
  if (!c->adj) {
    const size_t fields = (size_t)4 * NV + (size_t)dim * 4 * NV;
    if (hipMalloc((void**)&c->adj, sizeof(float) * fields * c->ncell) != hipSuccess) {
      c->adj = nullptr;
      return fail(AMX_ERR_NOMEM, "reverse-mode workspace allocation failed%s");
    }
  }
  hipStream_t st = (hipStream_t)stream;
  hipLaunchKernelGGL(k_set_dt, dim3(1), dim3(1), 0, st, c->ctl, d_dt);
  HIPCHK(hipGetLastError());
  if (c->cfg.mhd)
    return dim == 3 ? adjoint_step_mhd<3>(c, state, lam_out, lam_in, st) : adjoint_step_mhd<2>(c, state, lam_out, lam_in, st);
  return dim == 3 ? adjoint_step_t<3>(c, state, lam_out, lam_in, st) : adjoint_step_t<2>(c, state, lam_out, lam_in, st);
}

// ============================================================================
// Pools: continuous batching of independent simulations of one grid
// ============================================================================

namespace {

// Runs `enqueue` on `st`: directly, or -- with graphs on -- by replaying the
// graph cached under `slot` (captured from `enqueue` on first use). The
// enqueued work is kernels only (no memset nodes).
template <class F>
amx_status replay_or_capture(amx_fv_ctx* c, hipStream_t st, hipGraphExec_t* slot, bool have, F&& enqueue) {
  if (!c->use_graphs) return enqueue();
  if (!have) {
    HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    const amx_status s = enqueue();
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(st, &g);
    if (s != AMX_OK) {
      if (g) (void)hipGraphDestroy(g);
      return s;
    }
    HIPCHK(e);
    hipGraphExec_t x = nullptr;
    const hipError_t ei = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    HIPCHK(ei);
    *slot = x;
  }
  HIPCHK(hipGraphLaunch(*slot, st));
  return AMX_OK;
}

}  // namespace

extern "C" amx_status amx_fv_pool_slot_bytes(const amx_fv_config* cfg, size_t* bytes) {
  amx_status s = validate(cfg);
  if (s != AMX_OK) return s;
  if (!bytes) return fail(AMX_ERR_INVALID, "null output%s");
  *bytes = sizeof(float) * (size_t)cells(cfg) * (ws_fields(cfg, 1) + state_fields(cfg)) + sizeof(Ctl) + sizeof(Phys);
  if (cfg->riemann == AMX_HYBRID_HLLC || cfg->riemann == AMX_AM_HLLC) *bytes += (size_t)cells(cfg);
  if (cfg->limiter == AMX_VAN_ALBADA_PP) *bytes += sizeof(float) * (size_t)cells(cfg) * (2 * cfg->dim + 1);
  return AMX_OK;
}

extern "C" amx_status amx_fv_pool_create(const amx_fv_config* cfg, const amx_fv_run* run, int capacity,
                                         amx_fv_pool** out) {
  if (!out || !run) return fail(AMX_ERR_INVALID, "null argument%s");
  if (run->fixed_timestep && run->num_timesteps <= 0) return fail(AMX_ERR_INVALID, "num_timesteps must be positive%s");
  amx_fv_ctx* c = nullptr;
  amx_status s = make_ctx(cfg, capacity, 1, &c);
  if (s != AMX_OK) return s;
  c->sp.fixed = run->fixed_timestep;
  c->sp.num_timesteps = run->num_timesteps;
  c->sp.exact_end = run->exact_end_time;
  c->sp.C = cfg->C_cfl;
  c->sp.dx = c->dx;
  c->sp.dim = cfg->dim;
  c->use_graphs = env_int("AMX_POOL_GRAPHS", 1) != 0;
  c->batch = 1;
  // pool slots keep B in their iterate sets between steps; amx_fv_pool_read
  // writes it into the slot's state rows
  c->keep_b = c->cfg.mhd ? 1 : 0;
  amx_fv_pool* p = new (std::nothrow) amx_fv_pool();
  if (!p) {
    free_ctx(c);
    return fail(AMX_ERR_NOMEM, "host allocation failed%s");
  }
  p->c = c;
  *out = p;
  return AMX_OK;
}

extern "C" amx_status amx_fv_pool_destroy(amx_fv_pool* p) {
  if (!p) return AMX_OK;
  free_ctx(p->c);
  delete p;
  return AMX_OK;
}

static amx_status pool_slot_check(amx_fv_pool* p, int slot) {
  if (!p || !p->c) return fail(AMX_ERR_INVALID, "null pool%s");
  if (slot < 0 || slot >= p->c->cap) return fail(AMX_ERR_INVALID, "slot out of range%s");
  if (p->c->stepping) return fail(AMX_ERR_INVALID, "a step is in flight: call amx_fv_pool_step_finish first%s");
  return AMX_OK;
}

extern "C" amx_status amx_fv_pool_load(amx_fv_pool* p, int slot, const float* state, float gamma, float t_end,
                                       void* stream) {
  amx_status s = pool_slot_check(p, slot);
  if (s != AMX_OK) return s;
  if (!state) return fail(AMX_ERR_INVALID, "null state%s");
  if (!(gamma > 1.0f)) return fail(AMX_ERR_INVALID, "gamma must exceed 1%s");
  if (!(t_end >= 0.0f)) return fail(AMX_ERR_INVALID, "t_end must be >= 0%s");
  amx_fv_ctx* c = p->c;
  hipStream_t st = (hipStream_t)stream;
  float* dst = c->mem + (long)slot * c->sstride;
  HIPCHK(hipMemcpyAsync(dst, state, sizeof(float) * (size_t)c->NS * c->ncell, hipMemcpyDefault, st));
  HIPCHK(hipStreamSynchronize(st));  // the pinned parameter mirror is rewritten below
  c->hph[slot] = make_phys(gamma, c->cfg.grid_spacing);
  HIPCHK(hipMemcpyAsync(c->dph + slot, c->hph + slot, sizeof(Phys), hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_slot_reset, dim3(1), dim3(1), 0, st, c->ctl, slot, t_end, 1, 1);
  HIPCHK(hipGetLastError());
  if (!c->sp.fixed) HIPCHK(launch_cfl(c, c->mem, slot, 1, st));
  c->live[slot] = 1;
  c->fresh[slot] = 1;
  return AMX_OK;
}

extern "C" amx_status amx_fv_pool_release(amx_fv_pool* p, int slot, void* stream) {
  amx_status s = pool_slot_check(p, slot);
  if (s != AMX_OK) return s;
  hipLaunchKernelGGL(k_slot_reset, dim3(1), dim3(1), 0, (hipStream_t)stream, p->c->ctl, slot, 0.f, 0, 1);
  HIPCHK(hipGetLastError());
  p->c->live[slot] = 0;
  p->c->fresh[slot] = 0;
  return AMX_OK;
}

extern "C" amx_status amx_fv_pool_read(amx_fv_pool* p, int slot, float* dst, void* stream) {
  amx_status s = pool_slot_check(p, slot);
  if (s != AMX_OK) return s;
  if (!dst) return fail(AMX_ERR_INVALID, "null destination%s");
  amx_fv_ctx* c = p->c;
  hipStream_t st = (hipStream_t)stream;
  if (c->keep_b) {
    // the slot's B lives in its iterate set between steps: into its state rows first
    const Bufs b = bufs(c, nullptr);
    const long so = (long)slot * c->sstride;
    BsetArgs ba;
    for (int j = 0; j < 4; ++j)
      for (int k = 0; k < 3; ++k) ba.PB[j][k] = b.P[j][k] + so;
    for (int k = 0; k < 3; ++k) ba.dst[k] = c->mem + so + (long)(c->NV + k) * c->ncell;
    ba.n = c->ncell;
    ba.gate = 0;
    long blocks = (c->ncell + 255) / 256;
    if (blocks > 8192) blocks = 8192;
    hipLaunchKernelGGL(k_copy_bset, dim3(blocks), dim3(256), 0, st, c->ctl + slot, ba);
    HIPCHK(hipGetLastError());
  }
  HIPCHK(hipMemcpyAsync(dst, c->mem + (long)slot * c->sstride, sizeof(float) * (size_t)c->NS * c->ncell,
                        hipMemcpyDefault, st));
  return AMX_OK;
}

extern "C" amx_status amx_fv_pool_step_launch(amx_fv_pool* p, void* stream) {
  if (!p || !p->c) return fail(AMX_ERR_INVALID, "null pool%s");
  amx_fv_ctx* c = p->c;
  if (c->stepping) return fail(AMX_ERR_INVALID, "a step is already in flight%s");
  int span = 0;
  for (int s = 0; s < c->cap; ++s)
    if (c->live[s]) span = s + 1;
  c->span = span;
  if (span == 0) return AMX_OK;
  hipStream_t st = (hipStream_t)stream;
  const Bufs b = bufs(c, nullptr);
  int klast = 0;
  if (c->cfg.mhd) {
    // the speculative fixed-point batch of step_phase1: pairs L = 0, 2, ... <= upto
    const int maxit = c->cfg.max_fixed_point_iterations;
    int upto = c->last_k < 1 ? 1 : c->last_k;
    if (upto > maxit) upto = maxit;
    klast = (upto / 2) * 2 + 1;
    c->skip_mask = skip_mask_for(upto);  // graph replays do not run step_phase1 on the host
  } else {
    c->skip_mask = 0u;
  }
  const int lk = c->cfg.mhd ? c->last_k : 0;  // fixes the speculative pairs launched
  if (c->batch > 1) {
    // c->batch whole steps queued at once, each with its second half behind
    // its speculative launches; the device decides per slot whether they
    // settled the fixed point (a slot whose step stalls does nothing until
    // step_finish has completed that step)
    const int nb = c->batch;
    const auto bkey = std::make_tuple(span, lk, nb);
    auto bit = c->g_batch.find(bkey);
    hipGraphExec_t bx = bit == c->g_batch.end() ? nullptr : bit->second;
    amx_status bs = replay_or_capture(c, st, &bx, bx != nullptr, [&]() -> amx_status {
      StepParams spi = c->sp;
      spi.batch_first = 1;
      hipLaunchKernelGGL(k_step_begin, dim3(span), dim3(1), 0, st, c->ctl, spi);
      HIPCHK(hipGetLastError());
      spi.batch_first = 0;
      for (int i = 0; i < nb; ++i) {
        int kl = 0;
        amx_status r = step_phase1(c, b, span, st, !c->sp.fixed, &kl);
        if (r != AMX_OK) return r;
        r = step_phase3(c, b, span, st, !c->sp.fixed, 1);
        if (r != AMX_OK) return r;
        // this step's end and the next one's beginning
        if (i + 1 < nb)
          hipLaunchKernelGGL(k_step_end_begin, dim3(span), dim3(1), 0, st, c->ctl, c->keep_b, spi);
        else
          hipLaunchKernelGGL(k_step_end, dim3(span), dim3(1), 0, st, c->ctl, c->keep_b);
        HIPCHK(hipGetLastError());
      }
      return AMX_OK;
    });
    if (bs != AMX_OK) return bs;
    if (bx) c->g_batch[bkey] = bx;
    c->klast = klast;
    HIPCHK(hipMemcpyAsync(c->hctl, c->ctl, sizeof(Ctl) * span, hipMemcpyDeviceToHost, st));
    HIPCHK(hipEventRecord(c->ev_status, st));
    c->stepping = 1;
    c->batching = 1;
    return AMX_OK;
  }
  const auto key = std::make_pair(span, lk);
  auto it = c->g_launch.find(key);
  hipGraphExec_t x = it == c->g_launch.end() ? nullptr : it->second;
  amx_status s = replay_or_capture(c, st, &x, x != nullptr, [&]() -> amx_status {
    hipLaunchKernelGGL(k_step_begin, dim3(span), dim3(1), 0, st, c->ctl, c->sp);
    HIPCHK(hipGetLastError());
    int kl = 0;
    return step_phase1(c, b, span, st, !c->sp.fixed, &kl);
  });
  if (s != AMX_OK) return s;
  if (x) c->g_launch[key] = x;
  c->klast = klast;  // the launched batch, read back by step_finish
  HIPCHK(hipMemcpyAsync(c->hctl, c->ctl, sizeof(Ctl) * span, hipMemcpyDeviceToHost, st));
  HIPCHK(hipEventRecord(c->ev_status, st));
  c->stepping = 1;
  return AMX_OK;
}

extern "C" amx_status amx_fv_pool_step_finish(amx_fv_pool* p, void* stream) {
  if (!p || !p->c) return fail(AMX_ERR_INVALID, "null pool%s");
  amx_fv_ctx* c = p->c;
  if (!c->stepping) return AMX_OK;
  c->stepping = 0;
  const int span = c->span;
  hipStream_t st = (hipStream_t)stream;
  HIPCHK(hipEventSynchronize(c->ev_status));
  const Bufs b = bufs(c, nullptr);
  int klast = c->klast;
  int kmax = 0;
  if (c->batching) {
    c->batching = 0;
    int stalled = 0, bmax = 0;
    for (int i = 0; i < span; ++i) {
      stalled |= c->hctl[i].stall;
      bmax = c->hctl[i].iters_bmax > bmax ? c->hctl[i].iters_bmax : bmax;
    }
    if (stalled) {
      // complete the stalled slots' steps (the others sit out: run = 0)
      hipLaunchKernelGGL(k_resume, dim3(span), dim3(1), 0, st, c->ctl);
      HIPCHK(hipGetLastError());
      for (int i = 0; i < span; ++i) c->hctl[i].run = c->hctl[i].stall;
      amx_status s = fixed_point_more(c, b, span, st, &klast);
      if (s != AMX_OK) return s;
      int km = 0;
      fixed_point_settled(c, span, &km);
      bmax = km > bmax ? km : bmax;
      s = step_phase3(c, b, span, st, !c->sp.fixed, 0);
      if (s != AMX_OK) return s;
      hipLaunchKernelGGL(k_step_end, dim3(span), dim3(1), 0, st, c->ctl, c->keep_b);
      HIPCHK(hipGetLastError());
    }
    if (c->cfg.mhd && bmax > 0) c->last_k = bmax;
    HIPCHK(hipMemcpyAsync(c->hctl, c->ctl, sizeof(Ctl) * span, hipMemcpyDeviceToHost, st));
    HIPCHK(hipEventRecord(c->ev_status, st));
    for (int i = 0; i < span; ++i) c->fresh[i] = 0;
    return AMX_OK;
  }
  if (c->cfg.mhd) {
    amx_status s = fixed_point_more(c, b, span, st, &klast);
    if (s != AMX_OK) return s;
    fixed_point_settled(c, span, &kmax);
    c->last_k = kmax;

  }
  const auto fkey = std::make_pair(span, c->skip_mask);
  auto it = c->g_finish.find(fkey);
  hipGraphExec_t x = it == c->g_finish.end() ? nullptr : it->second;
  amx_status s = replay_or_capture(c, st, &x, x != nullptr, [&]() -> amx_status {
    amx_status r = step_phase3(c, b, span, st, !c->sp.fixed, 0);
    if (r != AMX_OK) return r;
    hipLaunchKernelGGL(k_step_end, dim3(span), dim3(1), 0, st, c->ctl, c->keep_b);
    HIPCHK(hipGetLastError());
    return AMX_OK;
  });
  if (s != AMX_OK) return s;
  if (x) c->g_finish[fkey] = x;
  HIPCHK(hipMemcpyAsync(c->hctl, c->ctl, sizeof(Ctl) * span, hipMemcpyDeviceToHost, st));
  HIPCHK(hipEventRecord(c->ev_status, st));
  for (int i = 0; i < span; ++i) c->fresh[i] = 0;
  return AMX_OK;
}

extern "C" amx_status amx_fv_pool_set_graphs(amx_fv_pool* p, int enable) {
  if (!p || !p->c) return fail(AMX_ERR_INVALID, "null pool%s");
  if (p->c->stepping) return fail(AMX_ERR_INVALID, "a step is in flight%s");
  p->c->use_graphs = enable ? 1 : 0;
  return AMX_OK;
}

extern "C" amx_status amx_fv_pool_set_batch(amx_fv_pool* p, int steps) {
  if (!p || !p->c) return fail(AMX_ERR_INVALID, "null pool%s");
  if (p->c->stepping) return fail(AMX_ERR_INVALID, "a step is in flight%s");
  if (steps < 1 || steps > 256) return fail(AMX_ERR_INVALID, "steps per launch must be in 1..256%s");
  p->c->batch = steps;
  return AMX_OK;
}

extern "C" amx_status amx_fv_pool_status(amx_fv_pool* p, int slot, amx_fv_slot_status* out) {
  if (!p || !p->c || !out) return fail(AMX_ERR_INVALID, "null argument%s");
  amx_fv_ctx* c = p->c;
  if (slot < 0 || slot >= c->cap) return fail(AMX_ERR_INVALID, "slot out of range%s");
  if (c->stepping) return fail(AMX_ERR_INVALID, "a step is in flight: call amx_fv_pool_step_finish first%s");
  HIPCHK(hipEventSynchronize(c->ev_status));
  memset(out, 0, sizeof(*out));
  out->live = c->live[slot];
  if (!out->live) return AMX_OK;
  if (c->fresh[slot] || slot >= c->span) {
    // loaded after the last step: nothing stepped yet
    return AMX_OK;
  }
  const Ctl& h = c->hctl[slot];
  out->steps = h.steps;
  out->t = h.t;
  out->t_end = h.t_end;
  out->fixed_point_iterations = h.iters_sum;
  out->finished = c->sp.fixed ? (h.steps >= c->sp.num_timesteps) : !(h.t < h.t_end);
  return AMX_OK;
}

extern "C" amx_status amx_copy_f32(float* dst, const float* src, size_t n, void* stream) {
  if ((!dst || !src) && n) return fail(AMX_ERR_INVALID, "null argument%s");
  if (((uintptr_t)dst | (uintptr_t)src) & 15) return fail(AMX_ERR_INVALID, "pointers must be 16-byte aligned%s");
  hipStream_t st = (hipStream_t)stream;
  const long n4 = (long)(n / 4);
  if (n4) {
    long blocks = (n4 + 255) / 256;
    if (blocks > 256 * 32) blocks = 256 * 32;
    hipLaunchKernelGGL(k_copy, dim3(blocks), dim3(256), 0, st, (float4*)dst, (const float4*)src, n4);
    HIPCHK(hipGetLastError());
  }
  if (n % 4) {
    hipLaunchKernelGGL(k_copy_tail, dim3(1), dim3(4), 0, st, dst, src, 4 * n4, (long)n);
    HIPCHK(hipGetLastError());
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

amx_status amx_fail(amx_status s, const char* fmt, const char* what) { return fail(s, fmt, what); }
