// Turbulent forcing of the velocity between steps
// (_physics_modules/_turbulent_forcing/_turbulent_forcing.py:14-104).
//
// amx_forcing_spectrum writes the solenoidal forcing spectrum of one step:
// complex normal noise scaled by sqrt(P(k)), P(k) = k^6 exp(-8 k / k_p),
// k_p = 4 pi / L, the DC mode removed and the compressive part projected out
// (_turbulent_forcing.py:14-75). The noise comes from this library's own
// counter-based generator (Philox4x64-10 of Salmon et al., SC'11, keyed by
// the seed, counter = (element, step)) -- the reference draws from JAX's
// threefry keys, which are not reproducible here. The inverse FFT is the
// caller's (rocFFT). amx_forcing_apply then solves the reference's quadratic
// for the amplitude that injects energy_injection_rate * dt of kinetic
// energy and adds amp * w to the velocity (_turbulent_forcing.py:77-104).

#include <math.h>

#include "common.h"

namespace amx {
namespace forcing {

constexpr int TPB = 256;
constexpr int MAXBLK = 1024;

// Philox4x64-10 block: 10 rounds of (hi, lo) = mulhilo(M, x), with the Weyl
// key schedule (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy
// as 1, 2, 3", SC 2011).
__device__ __forceinline__ void philox4x64(unsigned long long c[4], unsigned long long k0, unsigned long long k1) {
  const unsigned long long M0 = 0xD2E7470EE14C6C93ull, M1 = 0xCA5A826395121157ull;
  const unsigned long long W0 = 0x9E3779B97F4A7C15ull, W1 = 0xBB67AE8584CAA73Bull;
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const unsigned long long hi0 = __umul64hi(M0, c[0]), lo0 = M0 * c[0];
    const unsigned long long hi1 = __umul64hi(M1, c[2]), lo1 = M1 * c[2];
    const unsigned long long n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
    c[0] = n0;
    c[1] = lo1;
    c[2] = n2;
    c[3] = lo0;
    k0 += W0;
    k1 += W1;
  }
}

struct SpecArgs {
  float2* cw;  // (3, nx, ny, nz) complex
  int nx, ny, nz;
  unsigned ncell;
  float twopi_over_l;  // 2 pi / L in float32
  float kp;            // 4 pi / L in float32
  unsigned long long seed, step;
};

__device__ __forceinline__ int fft_index(int i, int n) { return i < (n + 1) / 2 ? i : i - n; }

// One normal pair per element from one Philox block (Box-Muller in float64):
// u1 = (a >> 11 + 1) 2^-53 in (0, 1], u2 = (b >> 11) 2^-53 in [0, 1).
__device__ __forceinline__ float2 normal_pair(unsigned long long e, unsigned long long step, unsigned long long seed) {
  unsigned long long c[4] = {e, step, 0ull, 0ull};
  philox4x64(c, seed, 0ull);
  const double u1 = (double)((c[0] >> 11) + 1ull) * 0x1.0p-53;
  const double u2 = (double)(c[1] >> 11) * 0x1.0p-53;
  const double r = sqrt(-2.0 * log(u1));
  const double th = 6.283185307179586 * u2;
  return make_float2((float)(r * cos(th)), (float)(r * sin(th)));
}

// one thread per wavevector: the three components' noise, scaling, projection
__global__ __launch_bounds__(TPB) void k_spectrum(const SpecArgs a) {
  for (unsigned c = blockIdx.x * TPB + threadIdx.x; c < a.ncell; c += gridDim.x * TPB) {
    const int l = (int)(c % (unsigned)a.nz);
    const unsigned r = c / (unsigned)a.nz;
    const int j = (int)(r % (unsigned)a.ny);
    const int i = (int)(r / (unsigned)a.ny);
    // 2 pi fftfreq(n, d = L / n): (index, index - n) / L
    const float kx = a.twopi_over_l * (float)fft_index(i, a.nx);
    const float ky = a.twopi_over_l * (float)fft_index(j, a.ny);
    const float kz = a.twopi_over_l * (float)fft_index(l, a.nz);
    const float k2 = (kx * kx + ky * ky) + kz * kz;
    const float kk = sqrtf(k2);
    const float pk = (((kk * kk) * (kk * kk)) * (kk * kk)) * expf((-8.f * kk) / a.kp);
    const float sp = sqrtf(pk);
    float2 w[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      const float2 z = normal_pair((unsigned long long)d * a.ncell + c, a.step, a.seed);
      w[d] = make_float2(sp * z.x, sp * z.y);
    }
    if (c == 0) {
      w[0] = w[1] = w[2] = make_float2(0.f, 0.f);  // DC mode
    } else {
      // remove the compressive part: w -= k (k . w) / k^2
      const float ik2 = 1.f / k2;
      const float dre = ((kx * w[0].x + ky * w[1].x) + kz * w[2].x) * ik2;
      const float dim = ((kx * w[0].y + ky * w[1].y) + kz * w[2].y) * ik2;
      w[0] = make_float2(w[0].x - kx * dre, w[0].y - kx * dim);
      w[1] = make_float2(w[1].x - ky * dre, w[1].y - ky * dim);
      w[2] = make_float2(w[2].x - kz * dre, w[2].y - kz * dim);
    }
#pragma unroll
    for (int d = 0; d < 3; ++d) a.cw[(size_t)d * a.ncell + c] = w[d];
  }
}

struct ApplyArgs {
  float* v[3];          // velocity rows of the state
  const float* rho;
  const float* w[3];    // real forcing field
  unsigned ncell;
  double* part;         // [gridDim.x][2]: sum rho |w|^2 / 2, sum rho v . w
  float* amp;           // device scalar
  const float* dt;      // device scalar
  float edot, dv;
};

template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__device__ __forceinline__ void block_sum2(double& a, double& b) {
  __shared__ double s[2][TPB / 64];
  a = wave_sum(a);
  b = wave_sum(b);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) {
    s[0][w] = a;
    s[1][w] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    a = s[0][0];
    b = s[1][0];
    for (int i = 1; i < TPB / 64; ++i) {
      a += s[0][i];
      b += s[1][i];
    }
  }
}

__global__ __launch_bounds__(TPB) void k_sums(const ApplyArgs a) {
  double sa = 0.0, sb = 0.0;
  for (unsigned c = blockIdx.x * TPB + threadIdx.x; c < a.ncell; c += gridDim.x * TPB) {
    const float r = a.rho[c];
    const float wx = a.w[0][c], wy = a.w[1][c], wz = a.w[2][c];
    sa += (double)(r * ((wx * wx + wy * wy) + wz * wz));
    sb += (double)((r * a.v[0][c] * wx + r * a.v[1][c] * wy) + r * a.v[2][c] * wz);
  }
  block_sum2(sa, sb);
  if (threadIdx.x == 0) {
    a.part[2 * blockIdx.x] = sa;
    a.part[2 * blockIdx.x + 1] = sb;
  }
}

// a amp^2 + b amp + c = 0 with a = sum rho |w|^2 / 2, b = sum rho v . w,
// c = -Edot dt / dV; amp = (-b + sqrt(b^2 - 4 a c)) / (2 a), or 0 when the
// discriminant is negative or |a| <= 1e-10 (_turbulent_forcing.py:90-104)
__global__ __launch_bounds__(TPB) void k_amp(const ApplyArgs a, int nblk) {
  double sa = 0.0, sb = 0.0;
  for (int i = threadIdx.x; i < nblk; i += TPB) {
    sa += a.part[2 * i];
    sb += a.part[2 * i + 1];
  }
  block_sum2(sa, sb);
  if (threadIdx.x != 0) return;
  const float ta = 0.5f * (float)sa, tb = (float)sb;
  const float tc = -((a.edot * *a.dt) / a.dv);
  const float disc = tb * tb - (4.f * ta) * tc;
  *a.amp = (disc >= 0.f && fabsf(ta) > 1e-10f) ? (-tb + sqrtf(disc)) / (2.f * ta) : 0.f;
}

__global__ __launch_bounds__(TPB) void k_apply(const ApplyArgs a) {
  const float amp = *a.amp;
  for (unsigned c = blockIdx.x * TPB + threadIdx.x; c < a.ncell; c += gridDim.x * TPB) {
#pragma unroll
    for (int d = 0; d < 3; ++d) a.v[d][c] = a.v[d][c] + amp * a.w[d][c];
  }
}

int nblocks(size_t n) {
  const size_t want = (n + (size_t)TPB * 4 - 1) / ((size_t)TPB * 4);
  return (int)(want < 1 ? 1 : (want > (size_t)MAXBLK ? (size_t)MAXBLK : want));
}

}  // namespace forcing
}  // namespace amx

using namespace amx::forcing;

extern "C" size_t amx_forcing_workspace_bytes(void) { return (size_t)MAXBLK * 2 * sizeof(double) + 256; }

extern "C" amx_status amx_forcing_spectrum(float* cw, int nx, int ny, int nz, double box_size, unsigned long long seed,
                                           unsigned long long step, void* stream) {
  if (!cw) return amx_fail(AMX_ERR_INVALID, "null argument%s");
  if (nx < 2 || ny < 2 || nz < 2 || !(box_size > 0.0)) return amx_fail(AMX_ERR_INVALID, "bad grid%s");
  const size_t n = (size_t)nx * ny * nz;
  if (n >= (size_t)1 << 30) return amx_fail(AMX_ERR_INVALID, "grids are limited to < 2^30 cells%s");
  SpecArgs a;
  a.cw = (float2*)cw;
  a.nx = nx;
  a.ny = ny;
  a.nz = nz;
  a.ncell = (unsigned)n;
  a.twopi_over_l = (float)(2.0 * M_PI / box_size);
  a.kp = (float)(4.0 * M_PI / box_size);
  a.seed = seed;
  a.step = step;
  hipLaunchKernelGGL(k_spectrum, dim3(nblocks(n)), dim3(TPB), 0, (hipStream_t)stream, a);
  AMX_HIPCHK(hipGetLastError());
  return AMX_OK;
}

extern "C" amx_status amx_forcing_apply(float* state, int num_vars, int nx, int ny, int nz, int rho_row, int vel_row,
                                        const float* w, const float* d_dt, float energy_injection_rate,
                                        double grid_spacing, void* work, float* d_amp, void* stream) {
  if (!state || !w || !d_dt || !work) return amx_fail(AMX_ERR_INVALID, "null argument%s");
  if (nx < 1 || ny < 1 || nz < 1 || rho_row < 0 || rho_row >= num_vars || vel_row < 0 || vel_row + 3 > num_vars)
    return amx_fail(AMX_ERR_INVALID, "bad layout%s");
  const size_t n = (size_t)nx * ny * nz;
  if (n >= (size_t)1 << 30) return amx_fail(AMX_ERR_INVALID, "grids are limited to < 2^30 cells%s");
  hipStream_t st = (hipStream_t)stream;
  ApplyArgs a;
  for (int d = 0; d < 3; ++d) {
    a.v[d] = state + (size_t)(vel_row + d) * n;
    a.w[d] = w + (size_t)d * n;
  }
  a.rho = state + (size_t)rho_row * n;
  a.ncell = (unsigned)n;
  a.part = (double*)work;
  a.amp = d_amp ? d_amp : (float*)((char*)work + (size_t)MAXBLK * 2 * sizeof(double));
  a.dt = d_dt;
  a.edot = energy_injection_rate;
  a.dv = (float)(grid_spacing * grid_spacing * grid_spacing);
  const int nb = nblocks(n);
  hipLaunchKernelGGL(k_sums, dim3(nb), dim3(TPB), 0, st, a);
  AMX_HIPCHK(hipGetLastError());
  hipLaunchKernelGGL(k_amp, dim3(1), dim3(TPB), 0, st, a, nb);
  AMX_HIPCHK(hipGetLastError());
  hipLaunchKernelGGL(k_apply, dim3(nb), dim3(TPB), 0, st, a);
  AMX_HIPCHK(hipGetLastError());
  return AMX_OK;
}
