// Image and mask ingest on the GPU: the cubic-spline sub-pixel shift of RELION particle stacks
// and the volume-mask projection of a PD.
//
// Reference: getDistanceCTF_local_Conj9combinedS2.py:259-264 (RELION images:
// scipy.ndimage.shift(img, (sh[1] - 0.5, sh[0] - 0.5), order=3, mode='wrap')).
//
// scipy's shift with mode='wrap' (the container's scipy, the version the oracle is pinned on)
// computes, per axis of length L:
//   * cubic B-spline coefficients with the whole-sample mirror extension
//     (x[-k] = x[k], x[L-1+k] = x[L-1-k]): the tridiagonal system (c[k-1] + 4c[k] + c[k+1])/6 =
//     x[k] with c[-1] = c[1], c[L] = c[L-2];
//   * output(i) = sum over the four taps around y = i - shift of c * beta3, where y outside
//     [0, L-1] is first wrapped with period L-1 (y < 0 -> into (0, L-1], y > L-1 -> into
//     [0, L-1)), and taps outside [0, L-1] take the mirrored coefficient.
// Both are restated here exactly (float64), on the separable tensor-product spline.
#include "common.h"

namespace memi {

// One cubic B-spline coefficient solve per line, in place on float64 data.  LU factors of the
// L x L system (times 6), diagonal 4 and off-diagonals 1 inside; the boundary rows are
// [4 2] / [2 4] for the whole-sample mirror extension (up0 = 2) and [5 1] / [1 5] for the
// half-sample reflection (up0 = 1): m[k] multiplies the previous row out, ib[k] = 1 / pivot k.
__device__ inline void mirror_spline_solve(double* x, int64_t stride, int L,
                                           const double* __restrict__ m,
                                           const double* __restrict__ ib, double up0 = 2.0) {
  // forward: d'[k] = 6 x[k] - m[k] d'[k-1]
  double prev = 6.0 * x[0];
  x[0] = prev;
  for (int k = 1; k < L; ++k) {
    prev = 6.0 * x[(int64_t)k * stride] - m[k] * prev;
    x[(int64_t)k * stride] = prev;
  }
  // backward: c[L-1] = d'[L-1] ib[L-1]; c[k] = (d'[k] - up_k c[k+1]) ib[k]
  double nxt = x[(int64_t)(L - 1) * stride] * ib[L - 1];
  x[(int64_t)(L - 1) * stride] = nxt;
  for (int k = L - 2; k >= 0; --k) {
    const double up = k == 0 ? up0 : 1.0;
    nxt = (x[(int64_t)k * stride] - up * nxt) * ib[k];
    x[(int64_t)k * stride] = nxt;
  }
}

// rows: coef[i][r][:] <- spline coefficients of in[i][r][:] (float32 -> float64)
__global__ void k_spline_rows(const float* __restrict__ in, int64_t n, int L,
                              const double* __restrict__ m, const double* __restrict__ ib,
                              double* __restrict__ coef) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= n * L) return;
  const float* src = in + t * L;
  double* dst = coef + t * L;
  for (int k = 0; k < L; ++k) dst[k] = (double)src[k];
  mirror_spline_solve(dst, 1, L, m, ib);
}

// columns, in place: coef[i][:][c]
__global__ void k_spline_cols(double* __restrict__ coef, int64_t n, int L,
                              const double* __restrict__ m, const double* __restrict__ ib) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= n * L) return;
  const int64_t img = t / L;
  const int c = (int)(t % L);
  mirror_spline_solve(coef + img * (int64_t)L * L + c, L, L, m, ib);
}

__device__ inline double bspline3(double t) {
  t = fabs(t);
  if (t < 1.0) return 2.0 / 3.0 - t * t + 0.5 * t * t * t;
  if (t < 2.0) {
    const double u = 2.0 - t;
    return u * u * u / 6.0;
  }
  return 0.0;
}

// coordinate wrap of scipy's mode='wrap' (period L-1, the first and last samples overlap)
__device__ inline double wrap_coord(double y, int L) {
  const double P = (double)(L - 1);
  if (y < 0.0) {
    const double r = fmod(-y, P);
    return r != 0.0 ? P - r : P;
  }
  if (y > P) return fmod(y, P);
  return y;
}

__device__ inline int mirror_index(int k, int L) {
  if (k < 0) return -k;
  if (k > L - 1) return 2 * (L - 1) - k;
  return k;
}

// out[i][x][y] = shifted image i at (y, x): the transposed layout of a SPIDER stack, so the
// loaders' transpose gives back the shifted image itself
__global__ void k_shift_interp(const double* __restrict__ coef, int64_t n, int L,
                               const double* __restrict__ sh0, const double* __restrict__ sh1,
                               int transpose_out, float* __restrict__ out) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= n * L * L) return;
  const int64_t img = t / ((int64_t)L * L);
  const int rem = (int)(t % ((int64_t)L * L));
  const int r = rem / L, c = rem % L;
  const double y = wrap_coord((double)r - sh0[img], L);
  const double x = wrap_coord((double)c - sh1[img], L);
  const int fy = (int)floor(y), fx = (int)floor(x);
  const double* C = coef + img * (int64_t)L * L;
  double s = 0.0;
  for (int a = fy - 1; a <= fy + 2; ++a) {
    const double wy = bspline3(y - a);
    if (wy == 0.0) continue;
    const int ra = mirror_index(a, L);
    double row = 0.0;
    for (int b = fx - 1; b <= fx + 2; ++b) {
      const double wx = bspline3(x - b);
      if (wx == 0.0) continue;
      row += wx * C[(int64_t)ra * L + mirror_index(b, L)];
    }
    s += wy * row;
  }
  const int64_t o = transpose_out ? ((int64_t)c * L + r) : ((int64_t)r * L + c);
  out[img * (int64_t)L * L + o] = (float)s;
}

// mask = proj > 1 (CC/projectMask.py:139), as 1.0 / 0.0
__global__ void k_mask_threshold(const double* __restrict__ proj, int64_t n,
                                 float* __restrict__ mask) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x)
    mask[t] = proj[t] > 1.0 ? 1.0f : 0.0f;
}

// ---- volume-mask projection (CC/projectMask.py:118-140) ----
// scipy's affine_transform(order=3, mode='nearest') pads the volume by NPAD samples per side
// with its edge values, takes the spline coefficients of the padded array with the half-sample
// reflection, and interpolates at the input coordinate + NPAD with taps clamped to the padded
// array (coordinates themselves are not clamped).
constexpr int MASK_NPAD = 12;

// padded[a][b][c] = vol[clamp(a - NPAD)][clamp(b - NPAD)][clamp(c - NPAD)] (float64)
__global__ void k_mask_pad(const float* __restrict__ vol, int L, double* __restrict__ padded) {
  const int Lp = L + 2 * MASK_NPAD;
  const int64_t total = (int64_t)Lp * Lp * Lp;
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < total;
       t += (int64_t)gridDim.x * blockDim.x) {
    const int c = (int)(t % Lp), b = (int)((t / Lp) % Lp), a = (int)(t / ((int64_t)Lp * Lp));
    const int ia = min(max(a - MASK_NPAD, 0), L - 1), ib = min(max(b - MASK_NPAD, 0), L - 1),
              ic = min(max(c - MASK_NPAD, 0), L - 1);
    padded[t] = (double)vol[((int64_t)ia * L + ib) * L + ic];
  }
}

// spline solve along one axis of the Lp^3 array: stride 1 (axis 2), Lp (axis 1), Lp^2 (axis 0)
__global__ void k_mask_spline_axis(double* __restrict__ a, int Lp, int axis,
                                   const double* __restrict__ m, const double* __restrict__ ib) {
  const int64_t lines = (int64_t)Lp * Lp;
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= lines) return;
  const int64_t u = t / Lp, v = t % Lp;
  int64_t base, stride;
  if (axis == 2) { base = (u * Lp + v) * Lp; stride = 1; }
  else if (axis == 1) { base = u * Lp * Lp + v; stride = Lp; }
  else { base = u * Lp + v; stride = (int64_t)Lp * Lp; }
  mirror_spline_solve(a + base, stride, Lp, m, ib, 1.0);
}

// proj[j][i] = sum_k f(T (i, j, k) + off) over the output grid of the swapped volume
// (CC/projectMask.py:127-139: rho = affine_transform(vol, T, off), msk = sum(rho, axis 2).T)
__global__ void k_mask_project(const double* __restrict__ coef, int L, const double* __restrict__ T,
                               double* __restrict__ proj) {
  const int Lp = L + 2 * MASK_NPAD;
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= (int64_t)L * L) return;
  const int i = (int)(t / L), j = (int)(t % L);
  double s = 0.0;
  for (int k = 0; k < L; ++k) {
    double c[3];
#pragma unroll
    for (int r = 0; r < 3; ++r)
      c[r] = T[3 * r + 0] * i + T[3 * r + 1] * j + T[3 * r + 2] * k + T[9 + r] + MASK_NPAD;
    int f[3];
    double w[3][4];
#pragma unroll
    for (int r = 0; r < 3; ++r) {
      f[r] = (int)floor(c[r]);
#pragma unroll
      for (int q = 0; q < 4; ++q) w[r][q] = bspline3(c[r] - (f[r] - 1 + q));
    }
    double v = 0.0;
    for (int qa = 0; qa < 4; ++qa) {
      const int a = min(max(f[0] - 1 + qa, 0), Lp - 1);
      double va = 0.0;
      for (int qb = 0; qb < 4; ++qb) {
        const int b = min(max(f[1] - 1 + qb, 0), Lp - 1);
        const double* line = coef + ((int64_t)a * Lp + b) * Lp;
        double vb = 0.0;
#pragma unroll
        for (int qc = 0; qc < 4; ++qc) vb += w[2][qc] * line[min(max(f[2] - 1 + qc, 0), Lp - 1)];
        va += w[1][qb] * vb;
      }
      v += w[0][qa] * va;
    }
    s += v;
  }
  proj[(int64_t)j * L + i] = s;
}

// S^2 tessellation (S2tessellation.py:61-65, classS2): the index of the nearest bin centre of
// every particle direction, by squared Euclidean distance in float64 (the same ordering as the
// distances of the reference's 1-NN search); equal distances go to the lower bin index.  One
// thread per particle; the centres stream through LDS in tiles of NB_TILE.
constexpr int NB_TILE = 1024;
__global__ __launch_bounds__(256) void k_nearest_bins(const double* __restrict__ pts, int64_t n,
                                                      const double* __restrict__ ctr, int m,
                                                      int32_t* __restrict__ idx) {
  __shared__ double c[3 * NB_TILE];
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const bool live = i < n;
  const double px = live ? pts[3 * i] : 0.0, py = live ? pts[3 * i + 1] : 0.0,
               pz = live ? pts[3 * i + 2] : 0.0;
  double best = INFINITY;
  int arg = 0;
  for (int t0 = 0; t0 < m; t0 += NB_TILE) {
    const int nt = min(NB_TILE, m - t0);
    __syncthreads();
    for (int e = threadIdx.x; e < 3 * nt; e += blockDim.x) c[e] = ctr[3 * (int64_t)t0 + e];
    __syncthreads();
    for (int j = 0; j < nt; ++j) {
      const double dx = px - c[3 * j], dy = py - c[3 * j + 1], dz = pz - c[3 * j + 2];
      const double d = fma(dx, dx, fma(dy, dy, dz * dz));
      if (d < best) {
        best = d;
        arg = t0 + j;
      }
    }
  }
  if (live) idx[i] = arg;
}

}  // namespace memi

using namespace memi;

extern "C" {

int memi_shift_images(const float* in, int64_t n, int N, const double* shift0,
                      const double* shift1, const double* lu, int transpose_out, float* out,
                      double* ws, int64_t ws_doubles, void* stream) {
  MEMI_CHECK_ARG(in && shift0 && shift1 && lu && out && ws, "memi_shift_images: null pointer");
  MEMI_CHECK_ARG(N >= 3 && N <= 4096 && n >= 0, "memi_shift_images: bad N=%d / n", N);
  MEMI_CHECK_ARG(ws_doubles >= n * (int64_t)N * N,
                 "memi_shift_images: workspace %lld < %lld doubles", (long long)ws_doubles,
                 (long long)(n * (int64_t)N * N));
  MEMI_CHECK_ARG((const void*)in != (const void*)out, "memi_shift_images: in and out alias");
  if (n == 0) return MEMI_OK;
  hipStream_t s = (hipStream_t)stream;
  const int64_t lines = n * N;
  hipLaunchKernelGGL(k_spline_rows, dim3(grid_for(lines, 256, 1 << 30)), dim3(256), 0, s, in, n,
                     N, lu, lu + N, ws);
  MEMI_CHECK_LAUNCH("k_spline_rows");
  hipLaunchKernelGGL(k_spline_cols, dim3(grid_for(lines, 256, 1 << 30)), dim3(256), 0, s, ws, n,
                     N, lu, lu + N);
  MEMI_CHECK_LAUNCH("k_spline_cols");
  hipLaunchKernelGGL(k_shift_interp, dim3(grid_for(lines * N, 256, 1 << 30)), dim3(256), 0, s,
                     ws, n, N, shift0, shift1, transpose_out, out);
  MEMI_CHECK_LAUNCH("k_shift_interp");
  return MEMI_OK;
}

int memi_project_mask(const float* vol, int L, const double* T, const double* lu,
                      double* proj, float* mask, double* ws, int64_t ws_doubles, void* stream) {
  MEMI_CHECK_ARG(vol && T && lu && proj && mask && ws, "memi_project_mask: null pointer");
  MEMI_CHECK_ARG(L >= 2 && L <= 1024, "memi_project_mask: bad volume side %d", L);
  const int64_t Lp = L + 2 * MASK_NPAD;
  MEMI_CHECK_ARG(ws_doubles >= Lp * Lp * Lp, "memi_project_mask: workspace %lld < %lld doubles",
                 (long long)ws_doubles, (long long)(Lp * Lp * Lp));
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_mask_pad, dim3(grid_for(Lp * Lp * Lp)), dim3(256), 0, s, vol, L, ws);
  MEMI_CHECK_LAUNCH("k_mask_pad");
  for (int axis = 0; axis < 3; ++axis) {
    hipLaunchKernelGGL(k_mask_spline_axis, dim3(grid_for(Lp * Lp, 256, 1 << 30)), dim3(256), 0, s,
                       ws, (int)Lp, axis, lu, lu + Lp);
    MEMI_CHECK_LAUNCH("k_mask_spline_axis");
  }
  hipLaunchKernelGGL(k_mask_project, dim3(grid_for((int64_t)L * L, 64, 1 << 30)), dim3(64), 0, s,
                     ws, L, T, proj);
  MEMI_CHECK_LAUNCH("k_mask_project");
  hipLaunchKernelGGL(k_mask_threshold, dim3(grid_for((int64_t)L * L)), dim3(256), 0, s, proj,
                     (int64_t)L * L, mask);
  MEMI_CHECK_LAUNCH("k_mask_threshold");
  return MEMI_OK;
}

int memi_nearest_bins(const double* pts, int64_t n, const double* centres, int m, int32_t* idx,
                      void* stream) {
  MEMI_CHECK_ARG(pts && centres && idx, "memi_nearest_bins: null pointer");
  MEMI_CHECK_ARG(n >= 0 && m >= 1, "memi_nearest_bins: bad sizes n=%lld m=%d", (long long)n, m);
  if (n == 0) return MEMI_OK;
  MEMI_CHECK_ARG((n + 255) / 256 < 0x7fffffffLL, "memi_nearest_bins: too many points");
  hipLaunchKernelGGL(k_nearest_bins, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     (hipStream_t)stream, pts, n, centres, m, idx);
  MEMI_CHECK_LAUNCH("k_nearest_bins");
  return MEMI_OK;
}

}  // extern "C"
