// Stage 3 NLSA kernels.  Reference: NLSA.py:23-158, get_wiener.py:9-21, svdRF.py:9-35,
// L2_distance.py:8-41, psiAnalysisParS2.py:89-90.
#include "common.h"

namespace memi {

// ---------------------------------------------------------------- delay embedding -------
// ConD[i][j] = sum_{c<C} D[ord[i+c]][ord[j+c]] (NLSA.py:31-34 after the psi sort of
// psiAnalysisParS2.py:89-90).  A 32x32 output tile needs the (32+C-1)^2 gathered block,
// staged once in LDS; the diagonal sums then run in increasing c like the reference loop.
constexpr int CT = 32;
constexpr int CMAX = 97;  // largest delay order the LDS block holds (32 + 96 = 128 rows)

template <typename T>
__global__ __launch_bounds__(256) void k_condist(const T* __restrict__ D,
                                                 const int64_t* __restrict__ ord, int n, int C,
                                                 double* __restrict__ ConD) {
  extern __shared__ double blk[];  // (CT+C-1)^2
  const int m = n - C;
  const int S = CT + C - 1;
  const int i0 = blockIdx.y * CT, j0 = blockIdx.x * CT;
  for (int e = threadIdx.x; e < S * S; e += blockDim.x) {
    const int r = e / S, c = e % S;
    const int gi = i0 + r, gj = j0 + c;
    double v = 0.0;
    if (gi < n && gj < n) v = (double)D[ord[gi] * (int64_t)n + ord[gj]];
    blk[e] = v;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < CT * CT; e += blockDim.x) {
    const int r = e / CT, c = e % CT;
    const int i = i0 + r, j = j0 + c;
    if (i >= m || j >= m) continue;
    double s = 0.0;
    for (int k = 0; k < C; ++k) s += blk[(r + k) * S + (c + k)];
    ConD[(int64_t)i * m + j] = s;
  }
}

// ------------------------------------------------------------ supervector spectra -------
// out[ii][c][k] = sum_{i<m} mu_psi[i][c] * spec[j][k] C[j][k] / W_i[k], j = C-1-ii+i,
// W_i[k] = sum_{t=i+1}^{i+C} C[t][k]^2 + 1/snr.  Thread = (frequency k, group of G delays);
// the G spectra a thread needs slide by one image per step, so they live in a register ring.
constexpr int SG = 5;
constexpr int ELLMAX = 8;

__global__ __launch_bounds__(256) void k_supervectors(const float2* __restrict__ spec,
                                                      const float* __restrict__ ctf, int n,
                                                      int64_t Kh, int C,
                                                      const double* __restrict__ mu_psi, int ell,
                                                      double snr, float2* __restrict__ out) {
  const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const int g0 = blockIdx.y * SG;
  if (k >= Kh) return;
  const int m = n - C;
  // W_0 = sum_{t=1}^{C} C_t^2
  double W = 1.0 / snr;
  for (int t = 1; t <= C; ++t) {
    const double c = ctf[(int64_t)t * Kh + k];
    W += c * c;
  }
  float2 ring[SG];
  const int jmax0 = C - 1 - g0;
#pragma unroll
  for (int t = 0; t < SG; ++t) {
    // ring[SG-1-t] holds G_{jmax - t}; preload all but the newest (t = 0)
    const int j = jmax0 - t;
    float2 g = make_float2(0.f, 0.f);
    if (t > 0 && j >= 0 && g0 + t < C) {
      const float2 x = spec[(int64_t)j * Kh + k];
      const float c = ctf[(int64_t)j * Kh + k];
      g = make_float2(x.x * c, x.y * c);
    }
    ring[SG - 1 - t] = g;
  }
  float acc[SG][ELLMAX][2];
#pragma unroll
  for (int t = 0; t < SG; ++t)
#pragma unroll
    for (int c = 0; c < ELLMAX; ++c) acc[t][c][0] = acc[t][c][1] = 0.f;
  for (int i = 0; i < m; ++i) {
    const int jn = jmax0 + i;  // newest image needed at this step
    {
      const float2 x = spec[(int64_t)jn * Kh + k];
      const float c = ctf[(int64_t)jn * Kh + k];
      ring[SG - 1] = make_float2(x.x * c, x.y * c);
    }
    const double invW = 1.0 / W;
    float u[ELLMAX];
#pragma unroll
    for (int c = 0; c < ELLMAX; ++c) u[c] = (c < ell) ? (float)(mu_psi[(int64_t)i * ell + c] * invW) : 0.f;
#pragma unroll
    for (int t = 0; t < SG; ++t) {
      const float2 g = ring[SG - 1 - t];
#pragma unroll
      for (int c = 0; c < ELLMAX; ++c) {
        acc[t][c][0] += u[c] * g.x;
        acc[t][c][1] += u[c] * g.y;
      }
    }
    // slide: drop the oldest, W_{i+1} = W_i - C_{i+1}^2 + C_{i+1+C}^2
#pragma unroll
    for (int t = 0; t < SG - 1; ++t) ring[t] = ring[t + 1];
    if (i + 1 < m) {
      const double a = ctf[(int64_t)(i + 1) * Kh + k];
      const double b = ctf[(int64_t)(i + 1 + C) * Kh + k];
      W += b * b - a * a;
    }
  }
#pragma unroll
  for (int t = 0; t < SG; ++t) {
    const int ii = g0 + t;
    if (ii >= C) break;
    for (int c = 0; c < ell; ++c)
      out[((int64_t)ii * ell + c) * Kh + k] = make_float2(acc[t][c][0], acc[t][c][1]);
  }
}

// ------------------------------------------------------------- tall-skinny algebra -------
constexpr int TGMAX = 8;

__global__ __launch_bounds__(256) void k_tall_gram(const double* __restrict__ A, int64_t M, int c,
                                                   double* __restrict__ G) {
  __shared__ double red[TGMAX * TGMAX][4];
  double s[TGMAX * (TGMAX + 1) / 2];
#pragma unroll
  for (int q = 0; q < TGMAX * (TGMAX + 1) / 2; ++q) s[q] = 0.0;
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < M;
       r += (int64_t)gridDim.x * blockDim.x) {
    double a[TGMAX];
#pragma unroll
    for (int x = 0; x < TGMAX; ++x) a[x] = (x < c) ? A[r * c + x] : 0.0;
    int q = 0;
#pragma unroll
    for (int x = 0; x < TGMAX; ++x)
#pragma unroll
      for (int y = x; y < TGMAX; ++y) s[q++] += a[x] * a[y];
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  int q = 0;
#pragma unroll
  for (int x = 0; x < TGMAX; ++x)
#pragma unroll
    for (int y = x; y < TGMAX; ++y) {
      double v = s[q++];
      for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
      if (lane == 0) red[x * TGMAX + y][wid] = v;
    }
  __syncthreads();
  for (int e = threadIdx.x; e < c * c; e += blockDim.x) {
    const int x = e / c, y = e % c;
    const int xx = min(x, y), yy = max(x, y);
    const double v = red[xx * TGMAX + yy][0] + red[xx * TGMAX + yy][1] +
                     red[xx * TGMAX + yy][2] + red[xx * TGMAX + yy][3];
    atomicAdd(&G[e], v);
  }
}

struct SmallMat {
  double v[TGMAX * TGMAX];
};

__global__ void k_tall_apply(const double* __restrict__ A, int64_t M, int c, SmallMat B, int d,
                             double* __restrict__ Y) {
  for (int64_t r = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; r < M;
       r += (int64_t)gridDim.x * blockDim.x) {
    double a[TGMAX];
    for (int x = 0; x < c; ++x) a[x] = A[r * c + x];
    for (int y = 0; y < d; ++y) {
      double s = 0.0;
      for (int x = 0; x < c; ++x) s += a[x] * B.v[x * d + y];
      Y[r * d + y] = s;
    }
  }
}

// column means of U2 [npix][q] (accumulated into mean[q]), then centre in place
__global__ void k_col_sum(const double* __restrict__ U, int64_t npix, int q,
                          double* __restrict__ colsum) {
  const int col = blockIdx.y * blockDim.x + threadIdx.x;
  if (col >= q) return;
  double s = 0.0;
  for (int64_t p = blockIdx.x; p < npix; p += gridDim.x) s += U[p * q + col];
  atomicAdd(&colsum[col], s);
}

__global__ void k_col_center(double* __restrict__ U, int64_t npix, int q,
                             const double* __restrict__ colsum) {
  const int64_t total = npix * q;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x)
    U[e] -= colsum[e % q] / (double)npix;
}

// G[a][b] += sum_p U[p][a] U[p][b] over this block's pixel range; 16x16 tiles of G
__global__ __launch_bounds__(256) void k_gram_cols(const double* __restrict__ U, int64_t npix,
                                                   int q, int64_t pchunk, double* __restrict__ G) {
  __shared__ double ta[64][17], tb[64][17];
  const int a0 = blockIdx.x * 16, b0 = blockIdx.y * 16;
  const int ta_i = threadIdx.x & 15, tb_i = threadIdx.x >> 4;
  const int64_t p0 = blockIdx.z * pchunk, p1 = min(npix, p0 + pchunk);
  double s = 0.0;
  for (int64_t pb = p0; pb < p1; pb += 64) {
    for (int e = threadIdx.x; e < 64 * 16; e += blockDim.x) {
      const int r = e / 16, cc = e % 16;
      const int64_t p = pb + r;
      ta[r][cc] = (p < p1 && a0 + cc < q) ? U[p * q + a0 + cc] : 0.0;
      tb[r][cc] = (p < p1 && b0 + cc < q) ? U[p * q + b0 + cc] : 0.0;
    }
    __syncthreads();
#pragma unroll 8
    for (int r = 0; r < 64; ++r) s += ta[r][ta_i] * tb[r][tb_i];
    __syncthreads();
  }
  const int a = a0 + ta_i, b = b0 + tb_i;
  if (a < q && b < q) atomicAdd(&G[a * q + b], s);
}

// IMGT[p][j] = inv_sd[j] * sum_q Ubar[p][q] H[q][j]   (float64, 64x64 tiles, 4x4 per thread)
__global__ __launch_bounds__(256) void k_imgt(const double* __restrict__ U,
                                              const double* __restrict__ H,
                                              const double* __restrict__ inv_sd, int64_t npix,
                                              int q, int ncols, double* __restrict__ out) {
  __shared__ double us[16][65], hs[16][65];
  const int64_t p0 = (int64_t)blockIdx.y * 64;
  const int j0 = blockIdx.x * 64;
  const int tr = threadIdx.x >> 4, tc = threadIdx.x & 15;
  double acc[4][4] = {};
  for (int k0 = 0; k0 < q; k0 += 16) {
    for (int e = threadIdx.x; e < 16 * 64; e += blockDim.x) {
      const int kk = e / 64, x = e % 64;
      const int64_t p = p0 + x;
      us[kk][x] = (p < npix && k0 + kk < q) ? U[p * q + k0 + kk] : 0.0;
      hs[kk][x] = (j0 + x < ncols && k0 + kk < q) ? H[(int64_t)(k0 + kk) * ncols + j0 + x] : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < 16; ++kk) {
      double a[4], b[4];
#pragma unroll
      for (int x = 0; x < 4; ++x) {
        a[x] = us[kk][tr * 4 + x];
        b[x] = hs[kk][tc * 4 + x];
      }
#pragma unroll
      for (int x = 0; x < 4; ++x)
#pragma unroll
        for (int y = 0; y < 4; ++y) acc[x][y] += a[x] * b[y];
    }
    __syncthreads();
  }
#pragma unroll
  for (int x = 0; x < 4; ++x) {
    const int64_t p = p0 + tr * 4 + x;
    if (p >= npix) continue;
#pragma unroll
    for (int y = 0; y < 4; ++y) {
      const int j = j0 + tc * 4 + y;
      if (j < ncols) out[p * ncols + j] = acc[x][y] * inv_sd[j];
    }
  }
}

__global__ void k_zero_d(double* p, int64_t n) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * blockDim.x)
    p[e] = 0.0;
}

}  // namespace memi

using namespace memi;

extern "C" {

int memi_nlsa_condist(const void* D, int d_is_f64, const int64_t* ord, int n, int con_order,
                      double* ConD, void* stream) {
  MEMI_CHECK_ARG(D && ord && ConD, "memi_nlsa_condist: null pointer");
  MEMI_CHECK_ARG(con_order >= 1 && con_order <= CMAX && n > con_order,
                 "memi_nlsa_condist: need 1 <= ConOrder <= %d and n > ConOrder (n=%d, C=%d)",
                 CMAX, n, con_order);
  const int m = n - con_order;
  const int S = CT + con_order - 1;
  const size_t shm = (size_t)S * S * sizeof(double);
  dim3 g((m + CT - 1) / CT, (m + CT - 1) / CT);
  if (d_is_f64)
    hipLaunchKernelGGL(k_condist<double>, g, dim3(256), shm, (hipStream_t)stream,
                       (const double*)D, ord, n, con_order, ConD);
  else
    hipLaunchKernelGGL(k_condist<float>, g, dim3(256), shm, (hipStream_t)stream,
                       (const float*)D, ord, n, con_order, ConD);
  MEMI_CHECK_LAUNCH("k_condist");
  return MEMI_OK;
}

int memi_nlsa_supervectors(const void* spec, const float* ctf_half, int n, int N, int con_order,
                           const double* mu_psi, int ell, double snr, void* out, void* stream) {
  MEMI_CHECK_ARG(spec && ctf_half && mu_psi && out, "memi_nlsa_supervectors: null pointer");
  MEMI_CHECK_ARG(ell >= 1 && ell <= ELLMAX, "memi_nlsa_supervectors: ell=%d must be 1..%d", ell,
                 ELLMAX);
  MEMI_CHECK_ARG(con_order >= 1 && n > con_order && snr > 0,
                 "memi_nlsa_supervectors: bad n=%d ConOrder=%d", n, con_order);
  const int64_t Kh = (int64_t)N * (N / 2 + 1);
  dim3 g((unsigned)((Kh + 255) / 256), (unsigned)((con_order + SG - 1) / SG));
  hipLaunchKernelGGL(k_supervectors, g, dim3(256), 0, (hipStream_t)stream, (const float2*)spec,
                     ctf_half, n, Kh, con_order, mu_psi, ell, snr, (float2*)out);
  MEMI_CHECK_LAUNCH("k_supervectors");
  return MEMI_OK;
}

int memi_tall_gram(const double* A, int64_t M, int c, double* G, void* stream) {
  MEMI_CHECK_ARG(A && G && M > 0 && c >= 1 && c <= TGMAX, "memi_tall_gram: bad arguments");
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_zero_d, dim3(1), dim3(64), 0, s, G, (int64_t)c * c);
  MEMI_CHECK_LAUNCH("k_zero_d");
  hipLaunchKernelGGL(k_tall_gram, dim3(grid_for(M, 256, 2048)), dim3(256), 0, s, A, M, c, G);
  MEMI_CHECK_LAUNCH("k_tall_gram");
  return MEMI_OK;
}

int memi_tall_apply(const double* A, int64_t M, int c, const double* B_host, int d, double* Y,
                    void* stream) {
  MEMI_CHECK_ARG(A && B_host && Y && M > 0 && c >= 1 && c <= TGMAX && d >= 1 && d <= TGMAX,
                 "memi_tall_apply: bad arguments");
  SmallMat B{};
  for (int e = 0; e < c * d; ++e) B.v[e] = B_host[e];
  hipLaunchKernelGGL(k_tall_apply, dim3(grid_for(M)), dim3(256), 0, (hipStream_t)stream, A, M, c,
                     B, d, Y);
  MEMI_CHECK_LAUNCH("k_tall_apply");
  return MEMI_OK;
}

int memi_center_gram(double* U2, int64_t npix, int q_dim, double* G, void* stream) {
  MEMI_CHECK_ARG(U2 && G && npix > 0 && q_dim >= 1 && q_dim <= 4096,
                 "memi_center_gram: bad arguments");
  hipStream_t s = (hipStream_t)stream;
  // column sums are accumulated in G's first q_dim entries before G is cleared for the Gram
  hipLaunchKernelGGL(k_zero_d, dim3(grid_for((int64_t)q_dim * q_dim)), dim3(256), 0, s, G,
                     (int64_t)q_dim * q_dim);
  MEMI_CHECK_LAUNCH("k_zero_d");
  dim3 gs(256, (unsigned)((q_dim + 255) / 256));
  hipLaunchKernelGGL(k_col_sum, gs, dim3(256), 0, s, U2, npix, q_dim, G);
  MEMI_CHECK_LAUNCH("k_col_sum");
  hipLaunchKernelGGL(k_col_center, dim3(grid_for(npix * q_dim)), dim3(256), 0, s, U2, npix,
                     q_dim, G);
  MEMI_CHECK_LAUNCH("k_col_center");
  hipLaunchKernelGGL(k_zero_d, dim3(1), dim3(256), 0, s, G, (int64_t)q_dim);
  MEMI_CHECK_LAUNCH("k_zero_d");
  const int64_t pchunk = 1024;
  dim3 gg((unsigned)((q_dim + 15) / 16), (unsigned)((q_dim + 15) / 16),
          (unsigned)((npix + pchunk - 1) / pchunk));
  hipLaunchKernelGGL(k_gram_cols, gg, dim3(256), 0, s, U2, npix, q_dim, pchunk, G);
  MEMI_CHECK_LAUNCH("k_gram_cols");
  return MEMI_OK;
}

int memi_nlsa_imgt(const double* Ubar, const double* H, const double* inv_sd, int64_t npix,
                   int q_dim, int ncols, double* IMGT, void* stream) {
  MEMI_CHECK_ARG(Ubar && H && inv_sd && IMGT && npix > 0 && q_dim > 0 && ncols > 0,
                 "memi_nlsa_imgt: bad arguments");
  dim3 g((unsigned)((ncols + 63) / 64), (unsigned)((npix + 63) / 64));
  hipLaunchKernelGGL(k_imgt, g, dim3(256), 0, (hipStream_t)stream, Ubar, H, inv_sd, npix, q_dim,
                     ncols, IMGT);
  MEMI_CHECK_LAUNCH("k_imgt");
  return MEMI_OK;
}

}  // extern "C"
