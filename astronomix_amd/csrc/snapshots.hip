// Snapshot rows of time_integration (time_integration.py:404-537): the time
// point, a copy of the state and the conserved-quantity diagnostics
// (_fluid_equations/total_quantities.py:29-176) plus max |div B|.
//
// One streaming pass over the state: each thread walks cells with a grid
// stride and accumulates the per-cell terms in float64; waves and blocks
// reduce in a fixed order into per-block partials, and one block reduces the
// partials. The result does not depend on scheduling (deterministic), and the
// max |div B| is exact. The per-cell terms are evaluated in float32 in the
// reference's operation order, without contraction into FMAs.

#include "snapshots.h"

#include <stdint.h>

namespace amx {
namespace snap {

constexpr int NQ = 5;  // sums: mass, total energy, internal energy, kinetic energy, radial momentum
constexpr int TPB = 256;
constexpr int NWV = TPB / 64;
constexpr int MAXBLK = 2048;

enum { Q_MASS = 0, Q_TOT, Q_INT, Q_KIN, Q_RAD };

struct DiagArgs {
  const float* rho;
  const float* v[3];
  const float* p;
  const float* b[3];
  int dim, nx, ny, nz, div_mode;
  unsigned ncell;
  int want_energy, want_rad, want_div;
  float gm1;   // gamma - 1 in float32
  float vol;   // float32(dx^dim)
  float den;   // float32(2 dx) (central) or float32(dx) (faces)
  float ctr;   // float32(box_size / 2)
  double dx;
  double* part;    // [gridDim.x][NQ]
  unsigned* pmax;  // [gridDim.x]
};

struct FinalArgs {
  int nblk, index;
  float t;
  float mass_scale;  // float32(box_size^dim)
  float* time_points;
  float* out[NQ];    // total_mass, total_energy, internal_energy, kinetic_energy, radial_momentum
  float* divb;
};

#pragma clang fp contract(off)

// cell-centre offset from the box centre along one axis: float32 linspace
// value (i + 1/2) dx minus float32(box_size / 2) (simulation_helper_data.py:94-131)
__device__ __forceinline__ float centre_offset(int i, double dx, float ctr) {
  return (float)(((double)i + 0.5) * dx) - ctr;
}

// sixth-order interface difference (finite_difference_int6, _differencing.py:9-30):
// f(o) is the face value at offset o along the axis; index i holds face i + 1/2
__device__ __forceinline__ float int6(float fm3, float fm2, float fm1, float f0, float fp1, float fp2) {
  const float c1 = 75.0f / 64.0f, c2 = -25.0f / 384.0f, c3 = 3.0f / 640.0f;
  return (c1 * (f0 - fm1) + c2 * (fp1 - fm2)) + c3 * (fp2 - fm3);
}

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Block reduction of NQ sums and one max; results valid in thread 0.
__device__ __forceinline__ void block_reduce(double (&s)[NQ], unsigned& m) {
  __shared__ double ss[NWV][NQ];
  __shared__ unsigned sm[NWV];
#pragma unroll
  for (int q = 0; q < NQ; ++q) s[q] = wave_sum(s[q]);
  m = wave_max_u(m);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) {
#pragma unroll
    for (int q = 0; q < NQ; ++q) ss[w][q] = s[q];
    sm[w] = m;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < NQ; ++q) {
      double a = ss[0][q];
      for (int i = 1; i < NWV; ++i) a += ss[i][q];
      s[q] = a;
    }
    unsigned mm = sm[0];
    for (int i = 1; i < NWV; ++i) mm = sm[i] > mm ? sm[i] : mm;
    m = mm;
  }
}

__global__ __launch_bounds__(TPB) void k_diag(const DiagArgs a) {
  double s[NQ] = {0.0, 0.0, 0.0, 0.0, 0.0};
  unsigned m = 0u;
  const unsigned plane = (unsigned)a.ny * (unsigned)a.nz;
  const unsigned stride = gridDim.x * TPB;
  for (unsigned c = blockIdx.x * TPB + threadIdx.x; c < a.ncell; c += stride) {
    const int k = (int)(c % (unsigned)a.nz);
    const unsigned r = c / (unsigned)a.nz;
    const int j = (int)(r % (unsigned)a.ny);
    const int i = (int)(r / (unsigned)a.ny);
    const float rho = a.rho[c];
    s[Q_MASS] += (double)rho;
    if (a.want_energy || a.want_rad) {
      const float vx = a.v[0][c], vy = a.v[1][c];
      const float vz = a.dim == 3 ? a.v[2][c] : 0.f;
      if (a.want_energy) {
        // get_absolute_velocity (_equations.py:165-197): +1e-20 in 3D only
        float usq = vx * vx + vy * vy;
        if (a.dim == 3) usq = (usq + vz * vz) + 1e-20f;
        const float u = sqrtf(usq);
        const float kin = (0.5f * rho) * (u * u);
        const float inte = a.p[c] / a.gm1;
        const float tot = inte + kin;
        s[Q_TOT] += (double)(tot * a.vol);
        s[Q_INT] += (double)(inte * a.vol);
        s[Q_KIN] += (double)(kin * a.vol);
      }
      if (a.want_rad) {
        // geometric_centers come from jnp.meshgrid(x, y[, z]) with its default
        // 'xy' indexing (simulation_helper_data.py:120-127): at array index
        // (i, j[, k]) component 0 is the coordinate of j and component 1 that of i
        const float o0 = centre_offset(j, a.dx, a.ctr);
        const float o1 = centre_offset(i, a.dx, a.ctr);
        float ur;
        if (a.dim == 3) {
          const float o2 = centre_offset(k, a.dx, a.ctr);
          const float nrm = sqrtf((o0 * o0 + o1 * o1) + o2 * o2);
          ur = (vx * (o0 / nrm) + vy * (o1 / nrm)) + vz * (o2 / nrm);
        } else {
          const float nrm = sqrtf(o0 * o0 + o1 * o1);
          ur = vx * (o0 / nrm) + vy * (o1 / nrm);
        }
        s[Q_RAD] += (double)((rho * ur) * a.vol);
      }
    }
    if (a.want_div) {
      float div;
      if (a.div_mode == AMX_DIV_CENTRAL) {
        // divergence3D (_vector_maths.py:33-51), periodic rolls
        const int ip = i + 1 == a.nx ? 0 : i + 1, im = i == 0 ? a.nx - 1 : i - 1;
        const int jp = j + 1 == a.ny ? 0 : j + 1, jm = j == 0 ? a.ny - 1 : j - 1;
        const int kp = k + 1 == a.nz ? 0 : k + 1, km = k == 0 ? a.nz - 1 : k - 1;
        const unsigned base = c - (unsigned)i * plane;
        const unsigned bj = c - (unsigned)j * (unsigned)a.nz;
        const unsigned bk = c - (unsigned)k;
        const float dfx = (a.b[0][base + (unsigned)ip * plane] - a.b[0][base + (unsigned)im * plane]) / a.den;
        const float dfy = (a.b[1][bj + (unsigned)jp * a.nz] - a.b[1][bj + (unsigned)jm * a.nz]) / a.den;
        const float dfz = (a.b[2][bk + kp] - a.b[2][bk + km]) / a.den;
        div = (dfx + dfy) + dfz;
      } else {
        // _interface_field_divergence (_differencing.py:33-52)
        float f[6];
        const unsigned base = c - (unsigned)i * plane;
#pragma unroll
        for (int o = 0; o < 6; ++o) f[o] = a.b[0][base + (unsigned)wrapi(i + o - 3, a.nx) * plane];
        const float ddx = int6(f[0], f[1], f[2], f[3], f[4], f[5]) / a.den;
        const unsigned bj = c - (unsigned)j * (unsigned)a.nz;
#pragma unroll
        for (int o = 0; o < 6; ++o) f[o] = a.b[1][bj + (unsigned)wrapi(j + o - 3, a.ny) * (unsigned)a.nz];
        const float ddy = int6(f[0], f[1], f[2], f[3], f[4], f[5]) / a.den;
        const unsigned bk = c - (unsigned)k;
#pragma unroll
        for (int o = 0; o < 6; ++o) f[o] = a.b[2][bk + (unsigned)wrapi(k + o - 3, a.nz)];
        const float ddz = int6(f[0], f[1], f[2], f[3], f[4], f[5]) / a.den;
        div = (ddx + ddy) + ddz;
      }
      const unsigned bits = canon_bits(fabsf(div));
      m = bits > m ? bits : m;
    }
  }
  block_reduce(s, m);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < NQ; ++q) a.part[(size_t)blockIdx.x * NQ + q] = s[q];
    a.pmax[blockIdx.x] = m;
  }
}

__global__ __launch_bounds__(TPB) void k_diag_final(const double* part, const unsigned* pmax, const FinalArgs f) {
  double s[NQ] = {0.0, 0.0, 0.0, 0.0, 0.0};
  unsigned m = 0u;
  for (int b = threadIdx.x; b < f.nblk; b += TPB) {
#pragma unroll
    for (int q = 0; q < NQ; ++q) s[q] += part[(size_t)b * NQ + q];
    m = pmax[b] > m ? pmax[b] : m;
  }
  block_reduce(s, m);
  if (threadIdx.x != 0) return;
  if (f.time_points) f.time_points[f.index] = f.t;
  // jnp.sum(rho) * box_size**dim (total_quantities.py:176)
  if (f.out[Q_MASS]) f.out[Q_MASS][f.index] = (float)s[Q_MASS] * f.mass_scale;
  for (int q = 1; q < NQ; ++q)
    if (f.out[q]) f.out[q][f.index] = (float)s[q];
  if (f.divb) f.divb[f.index] = __uint_as_float(m);
}

#pragma clang fp contract(on)

size_t work_bytes() { return (size_t)MAXBLK * (NQ * sizeof(double) + sizeof(unsigned)); }

namespace {

amx_status check_desc(const amx_diag_desc& d) {
  if (d.dim != 2 && d.dim != 3) return amx_fail(AMX_ERR_INVALID, "diagnostics: dim must be 2 or 3%s");
  if (d.nx < 1 || d.ny < 1 || d.nz < 1 || (d.dim == 2 && d.nz != 1))
    return amx_fail(AMX_ERR_INVALID, "diagnostics: bad grid sizes%s");
  const double n = (double)d.nx * d.ny * d.nz;
  if (n >= (double)(1u << 30)) return amx_fail(AMX_ERR_INVALID, "diagnostics: grids are limited to < 2^30 cells%s");
  const int nv = d.num_vars;
  if (d.rho_row < 0 || d.rho_row >= nv || d.p_row < 0 || d.p_row >= nv || d.vel_row < 0 || d.vel_row + d.dim > nv)
    return amx_fail(AMX_ERR_INVALID, "diagnostics: row index out of range%s");
  if (d.div_mode != AMX_DIV_NONE && d.div_mode != AMX_DIV_CENTRAL && d.div_mode != AMX_DIV_FACES)
    return amx_fail(AMX_ERR_INVALID, "diagnostics: unknown div_mode%s");
  if (d.div_mode != AMX_DIV_NONE && (d.mag_row < 0 || d.mag_row + 3 > nv))
    return amx_fail(AMX_ERR_INVALID, "diagnostics: magnetic rows out of range%s");
  if (d.div_mode == AMX_DIV_FACES && d.dim != 3)
    return amx_fail(AMX_ERR_INVALID, "diagnostics: face-field divergence is 3D%s");
  if (!(d.grid_spacing > 0.0)) return amx_fail(AMX_ERR_INVALID, "diagnostics: grid_spacing must be positive%s");
  return AMX_OK;
}

amx_status check_out(const amx_snapshot_out* o, const amx_diag_desc& d) {
  if (!o) return amx_fail(AMX_ERR_INVALID, "null snapshot output%s");
  if (o->magnetic_divergence) {
    if (d.div_mode == AMX_DIV_NONE)
      return amx_fail(AMX_ERR_INVALID, "magnetic divergence requested for a state without a magnetic field%s");
    // divergence3D rolls the third component along axis 2 (_vector_maths.py:47-48)
    if (d.div_mode == AMX_DIV_CENTRAL && d.dim != 3)
      return amx_fail(AMX_ERR_UNSUPPORTED, "the cell-centred magnetic divergence (divergence3D) needs a 3D grid%s");
  }
  return AMX_OK;
}

}  // namespace

amx_status validate(const amx_snapshot_spec* spec, const amx_snapshot_out* out, const amx_diag_desc& d) {
  if (!spec) return amx_fail(AMX_ERR_INVALID, "null snapshot spec%s");
  if (spec->num_snapshots < 1) return amx_fail(AMX_ERR_INVALID, "num_snapshots must be positive%s");
  if (spec->use_timepoints && (!spec->timepoints || spec->num_timepoints < 1))
    return amx_fail(AMX_ERR_INVALID, "use_timepoints needs at least one snapshot time point%s");
  if (spec->callback) return AMX_OK;
  amx_status s = check_desc(d);
  if (s != AMX_OK) return s;
  return check_out(out, d);
}

amx_status record(const amx_diag_desc& d, const float* state, float t, const amx_snapshot_out& o, int index,
                  void* work, hipStream_t st) {
  const size_t ncell = (size_t)d.nx * d.ny * d.nz;
  if (o.states)
    AMX_HIPCHK(hipMemcpyAsync(o.states + (size_t)index * d.num_vars * ncell, state,
                              sizeof(float) * d.num_vars * ncell, hipMemcpyDeviceToDevice, st));
  const bool want_energy = o.total_energy || o.internal_energy || o.kinetic_energy;
  const bool want_rad = o.radial_momentum != nullptr;
  const bool want_div = o.magnetic_divergence != nullptr;
  const bool want_any = want_energy || want_rad || want_div || o.total_mass;
  double* part = (double*)work;
  unsigned* pmax = work ? (unsigned*)(part + (size_t)MAXBLK * NQ) : nullptr;
  int nblk = 0;
  if (want_any) {
    if (!work) return amx_fail(AMX_ERR_INVALID, "null diagnostics workspace%s");
    DiagArgs a;
    a.rho = state + (size_t)d.rho_row * ncell;
    for (int q = 0; q < 3; ++q) a.v[q] = q < d.dim ? state + (size_t)(d.vel_row + q) * ncell : nullptr;
    a.p = state + (size_t)d.p_row * ncell;
    for (int q = 0; q < 3; ++q) a.b[q] = d.div_mode != AMX_DIV_NONE ? state + (size_t)(d.mag_row + q) * ncell : nullptr;
    a.dim = d.dim;
    a.nx = d.nx;
    a.ny = d.ny;
    a.nz = d.nz;
    a.div_mode = d.div_mode;
    a.ncell = (unsigned)ncell;
    a.want_energy = want_energy;
    a.want_rad = want_rad;
    a.want_div = want_div;
    a.gm1 = d.gamma - 1.0f;
    a.vol = (float)(d.dim == 3 ? d.grid_spacing * d.grid_spacing * d.grid_spacing : d.grid_spacing * d.grid_spacing);
    a.den = (float)(d.div_mode == AMX_DIV_CENTRAL ? 2.0 * d.grid_spacing : d.grid_spacing);
    a.ctr = (float)(d.box_size / 2.0);
    a.dx = d.grid_spacing;
    a.part = part;
    a.pmax = pmax;
    const size_t want_blocks = (ncell + (size_t)TPB * 4 - 1) / ((size_t)TPB * 4);
    nblk = (int)(want_blocks < (size_t)MAXBLK ? want_blocks : (size_t)MAXBLK);
    if (nblk < 1) nblk = 1;
    hipLaunchKernelGGL(k_diag, dim3(nblk), dim3(TPB), 0, st, a);
    AMX_HIPCHK(hipGetLastError());
  }
  FinalArgs f;
  f.nblk = nblk;
  f.index = index;
  f.t = t;
  f.mass_scale = (float)(d.dim == 3 ? d.box_size * d.box_size * d.box_size : d.box_size * d.box_size);
  f.time_points = o.time_points;
  f.out[Q_MASS] = o.total_mass;
  f.out[Q_TOT] = o.total_energy;
  f.out[Q_INT] = o.internal_energy;
  f.out[Q_KIN] = o.kinetic_energy;
  f.out[Q_RAD] = o.radial_momentum;
  f.divb = o.magnetic_divergence;
  hipLaunchKernelGGL(k_diag_final, dim3(1), dim3(TPB), 0, st, (const double*)part, (const unsigned*)pmax, f);
  AMX_HIPCHK(hipGetLastError());
  return AMX_OK;
}

}  // namespace snap
}  // namespace amx

extern "C" size_t amx_snapshot_workspace_bytes(void) { return amx::snap::work_bytes(); }

extern "C" amx_status amx_snapshot_record(const amx_diag_desc* d, const float* state, float t,
                                          const amx_snapshot_out* out, int index, void* work, void* stream) {
  if (!d || !state || !out) return amx_fail(AMX_ERR_INVALID, "null argument%s");
  if (index < 0) return amx_fail(AMX_ERR_INVALID, "index must be >= 0%s");
  amx_status s = amx::snap::check_desc(*d);
  if (s != AMX_OK) return s;
  s = amx::snap::check_out(out, *d);
  if (s != AMX_OK) return s;
  return amx::snap::record(*d, state, t, *out, index, work, (hipStream_t)stream);
}
