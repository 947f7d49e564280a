// Shared helpers for the libmemi HIP kernels (gfx950 / CDNA4 only).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <cstdarg>
#include <cmath>
#include <algorithm>

#include "../../include/memi.h"

namespace memi {

// last error text, per host thread
void set_error(const char* fmt, ...);

#define MEMI_CHECK_ARG(cond, ...)                 \
  do {                                            \
    if (!(cond)) {                                \
      ::memi::set_error(__VA_ARGS__);             \
      return MEMI_EINVAL;                         \
    }                                             \
  } while (0)

#define MEMI_CHECK_LAUNCH(what)                                              \
  do {                                                                       \
    hipError_t e_ = hipGetLastError();                                       \
    if (e_ != hipSuccess) {                                                  \
      ::memi::set_error("%s: %s", what, hipGetErrorString(e_));              \
      return MEMI_ELAUNCH;                                                   \
    }                                                                        \
  } while (0)

constexpr int kThreads = 256;

inline int grid_for(int64_t work, int per_block = kThreads, int cap = 65536) {
  int64_t g = (work + per_block - 1) / per_block;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return static_cast<int>(g);
}

// Signed FFT frequency of index i on an n-point grid, matching the ordering of
// ifftshift(arange(-n/2, n/2)) for even n (Nyquist is -n/2) and of
// ifftshift(arange(-(n-1)/2, (n-1)/2 + 1)) for odd n.
__host__ __device__ inline int fft_freq(int i, int n) {
  return (i < (n + 1) / 2) ? i : i - n;
}

// Radial frequency Q = sqrt(fx^2 + fy^2) / (N/2) of create_grid
// (getDistanceCTF_local_Conj9combinedS2.py:139-154) after ifftshift.
__device__ inline double radial_q(int row, int col, int N) {
  double fy = fft_freq(row, N);
  double fx = fft_freq(col, N);
  return sqrt(fx * fx + fy * fy) / (0.5 * N);
}

// Half-spectrum weight of column kx of an rfft2 layout: self-conjugate columns
// (kx = 0 and, for even N, kx = N/2) count once, the others stand for two columns.
__device__ inline float half_weight(int kx, int N) {
  return (kx == 0 || (N % 2 == 0 && kx == N / 2)) ? 1.0f : 2.0f;
}

// CTF of ctemh_cryoFrank.py:24-44 at radial frequency k (1/A) for defocus df (A).
// Evaluated in float64: the phase wr reaches several hundred radians.
__device__ inline double ctf_value(double k, double df, double cs_mm, double ekv,
                                   double ampc, double gauss_env) {
  const double cs = cs_mm * 1.0e7;
  const double wav = 12.3986 / sqrt((2.0 * 511.0 + ekv) * ekv);
  const double w1 = M_PI * cs * wav * wav * wav;
  const double w2 = M_PI * wav * df;
  const double k2 = k * k;
  double wi = 1.0;
  if (isfinite(gauss_env)) {
    const double sigm = gauss_env / sqrt(2.0 * log(2.0));
    wi = exp(-k2 / (2.0 * sigm * sigm));
  }
  const double wr = (0.5 * w1 * k2 - w2) * k2;
  double s, c;
  sincos(wr, &s, &c);
  return (s - ampc * c) * wi;
}

// Wave sum of a double by DPP moves (plain VALU operations, no LDS traffic): an inclusive
// scan within each 16-lane row (row_shr 1, 2, 4, 8; lanes without a source read 0), then
// row_bcast:15 adds row 0's total into row 1 and row 2's into row 3, and row_bcast:31 adds
// rows 0 + 1 into rows 2 and 3; lane 63 then holds the wave's total, read back as a uniform
// value.  Both 32-bit halves of the double move with the same control.  Call it with all 64
// lanes of the wave active (inactive lanes would feed stale values into the scan).
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double dpp_move(double v) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_update_dpp(0, (int)b, CTRL, ROW_MASK, 0xf, true);
  const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, ROW_MASK, 0xf, true);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

__device__ __forceinline__ double wave_total(double v) {
  v += dpp_move<0x111, 0xf>(v);   // row_shr:1
  v += dpp_move<0x112, 0xf>(v);   // row_shr:2
  v += dpp_move<0x114, 0xf>(v);   // row_shr:4
  v += dpp_move<0x118, 0xf>(v);   // row_shr:8
  v += dpp_move<0x142, 0xa>(v);   // row_bcast:15 into rows 1 and 3
  v += dpp_move<0x143, 0xc>(v);   // row_bcast:31 into rows 2 and 3
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)b, 63);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), 63);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// Block-wide sum (whole waves, up to 512 threads) of a double; result valid in all threads.
// The wave part by DPP (wave_total): a shuffle-down tree goes through LDS permutes, six
// dependent round trips per sum.
__device__ inline double block_sum(double v, double* sh) {
  v = wave_total(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  __syncthreads();
  if (lane == 0) sh[wid] = v;
  __syncthreads();
  double t = 0.0;
  const int nw = (blockDim.x + 63) >> 6;
  for (int w = 0; w < nw; ++w) t += sh[w];
  __syncthreads();
  return t;
}

}  // namespace memi
