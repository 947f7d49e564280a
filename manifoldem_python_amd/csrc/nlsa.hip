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
#pragma unroll
    for (int c = 0; c < ELLMAX; ++c)       // constant indices keep acc in registers
      if (c < ell) out[((int64_t)ii * ell + c) * Kh + k] = make_float2(acc[t][c][0], acc[t][c][1]);
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

// ---------------------------------------------------------- reconstruction distances ----
// T[a][j] = sum_b Gq[a][b] H[b][j]   (Gq [q][q], H [q][ncols])
__global__ void k_gq_h(const double* __restrict__ Gq, const double* __restrict__ H, int q,
                       int ncols, double* __restrict__ T) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int a = blockIdx.y;
  if (j >= ncols) return;
  double s = 0.0;
  for (int b = 0; b < q; ++b) s = fma(Gq[a * q + b], H[(int64_t)b * ncols + j], s);
  T[(int64_t)a * ncols + j] = s;
}

// Gram_jj = sum_a H[a][j] T[a][j] in increasing a with fma, the order k_recon_dist's tiles use,
// so its diagonal entries equal these bit for bit; inv_sd_j = 1 / sqrt(Gram_jj / npix).
__global__ void k_recon_diag(const double* __restrict__ H, const double* __restrict__ T, int q,
                             int ncols, double npix, double* __restrict__ dg,
                             double* __restrict__ inv_sd) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= ncols) return;
  double s = 0.0;
  for (int a = 0; a < q; ++a) s = fma(H[(int64_t)a * ncols + j], T[(int64_t)a * ncols + j], s);
  dg[j] = s;
  inv_sd[j] = 1.0 / sqrt(s / npix);
}

// 64x64 tiles of Gram = H^T T with the L2_distance epilogue of NLSA.py:141-146:
// C_ij = Gram_ij (inv_i inv_j), d = (C_ii + C_jj) - 2 C_ij, d < 1e-8 -> 0, out = sqrt(d)^2.
__global__ __launch_bounds__(256) void k_recon_dist(const double* __restrict__ H,
                                                    const double* __restrict__ T,
                                                    const double* __restrict__ dg,
                                                    const double* __restrict__ inv_sd, int q,
                                                    int ncols, float* __restrict__ out) {
  __shared__ double hs[16][65], ts[16][65];
  const int i0 = blockIdx.y * 64, j0 = blockIdx.x * 64;
  const int tr = threadIdx.x >> 4, tc = threadIdx.x & 15;
  double acc[4][4] = {};
  for (int k0 = 0; k0 < q; k0 += 16) {
    for (int e = threadIdx.x; e < 16 * 64; e += blockDim.x) {
      const int kk = e / 64, x = e % 64;
      const bool kin = k0 + kk < q;
      hs[kk][x] = (kin && i0 + x < ncols) ? H[(int64_t)(k0 + kk) * ncols + i0 + x] : 0.0;
      ts[kk][x] = (kin && j0 + x < ncols) ? T[(int64_t)(k0 + kk) * ncols + j0 + x] : 0.0;
    }
    __syncthreads();
    const int kend = min(16, q - k0);
    for (int kk = 0; kk < kend; ++kk) {
      double a[4], b[4];
#pragma unroll
      for (int x = 0; x < 4; ++x) {
        a[x] = hs[kk][tr * 4 + x];
        b[x] = ts[kk][tc * 4 + x];
      }
#pragma unroll
      for (int x = 0; x < 4; ++x)
#pragma unroll
        for (int y = 0; y < 4; ++y) acc[x][y] = fma(a[x], b[y], acc[x][y]);
    }
    __syncthreads();
  }
#pragma unroll
  for (int x = 0; x < 4; ++x) {
    const int i = i0 + tr * 4 + x;
    if (i >= ncols) continue;
    const double ii = inv_sd[i];
    const double aai = dg[i] * (ii * ii);
#pragma unroll
    for (int y = 0; y < 4; ++y) {
      const int j = j0 + tc * 4 + y;
      if (j >= ncols) continue;
      const double ij = inv_sd[j];
      const double c = acc[x][y] * (ii * ij);
      double d = (aai + dg[j] * (ij * ij)) - 2.0 * c;
      if (d < 1e-8) d = 0.0;
      const double r = sqrt(d);
      out[(int64_t)i * ncols + j] = (float)(r * r);
    }
  }
}

// ------------------------------------------------------------- 1-D open-manifold fit ----
// tau update of fit_1D_open_manifold_3D.py:64-160: for a point x and curve (a, b),
// R(tau) = sum_j (x_j - b_j - a_j cos(j pi tau))^2 (R_p.py:41-49) is stationary in (0, 1)
// where beta = cos(pi tau) is a root of the quintic of solve_d_R_d_tau_p_3D.py:35-60; tau is
// the candidate arccos(beta)/pi (real beta in [-1, 1]) or 0 or 1 of least R, first in order on
// ties.  The reference takes the roots from np.roots (a LAPACK eigensolve of the companion
// matrix) and keeps those with zero imaginary part.  Here the real roots come from monotone
// segments: the roots of p^(k) bracket those of p^(k-1), so isolating from the linear fourth
// derivative up gives every real root of p in [-1.01, 1.01] by safeguarded Newton steps.
// Where LAPACK's real/complex decision or the arg min is within rounding of a tie (a stationary
// value of p within 1e-10 of its scale: a near-double root; p(+-1) that small: a root at the
// edge; two candidates with R equal to 1e-12 of its scale; a zero leading or trailing
// coefficient, which np.roots treats specially) the point is flagged for the host to re-solve.
constexpr int FIT_BLOCK = 256;
constexpr double FIT_LO = -1.01, FIT_HI = 1.01;
constexpr double FIT_EPS = 2.220446049250313e-16;   // double precision unit roundoff x 2

struct FitCurve {
  double a[3], b[3];
};

__device__ inline double peval(const double* e, int d, double x) {
  double r = e[d];
  for (int k = d - 1; k >= 0; --k) r = fma(r, x, e[k]);
  return r;
}

// root of e (degree d) in the bracket [u, v] where e(u), e(v) have opposite signs and e is
// monotone; de is its derivative (degree d - 1)
__device__ double monotone_root(const double* e, int d, const double* de, double u, double v,
                                double pu) {
  double x = 0.5 * (u + v);
  for (int it = 0; it < 200; ++it) {
    const double px = peval(e, d, x);
    if (px == 0.0) return x;
    if ((px < 0.0) == (pu < 0.0)) {
      u = x;
      pu = px;
    } else {
      v = x;
    }
    if (v - u <= 4.0 * FIT_EPS * fmax(fabs(u), fabs(v)) || v - u <= 1e-300) break;
    const double dx = peval(de, d - 1, x);
    if (dx != 0.0) {
      const double step = px / dx;
      // a Newton step of a few ulps: x is the root to rounding (further steps only bounce
      // between the bracket's ends while bisection halves it to the last bit)
      if (fabs(step) <= 16.0 * FIT_EPS * fabs(x)) {
        const double xn = x - step;
        return (xn > u && xn < v) ? xn : x;
      }
    }
    double xn = dx != 0.0 ? x - px / dx : u;
    if (!(xn > u && xn < v)) xn = 0.5 * (u + v);
    if (xn == x) break;
    x = xn;
  }
  return x;
}

// real roots of e (degree d) in [FIT_LO, FIT_HI], ascending, given the ascending roots cr[0..nc)
// of its derivative strictly inside the interval
__device__ int segment_roots(const double* e, int d, const double* de, const double* cr, int nc,
                             double* out) {
  int n = 0;
  double u = FIT_LO, pu = peval(e, d, u);
  if (pu == 0.0) out[n++] = u;
  for (int s = 0; s <= nc; ++s) {
    const double v = s < nc ? cr[s] : FIT_HI;
    const double pv = peval(e, d, v);
    if (pv == 0.0) {
      if (n == 0 || out[n - 1] != v) out[n++] = v;
    } else if (pu != 0.0 && ((pu < 0.0) != (pv < 0.0))) {
      out[n++] = monotone_root(e, d, de, u, v, pu);
    }
    u = v;
    pu = pv;
  }
  return n;
}

__device__ inline int interior(const double* r, int n, double* out) {
  int m = 0;
  for (int i = 0; i < n; ++i)
    if (r[i] > FIT_LO && r[i] < FIT_HI) out[m++] = r[i];
  return m;
}

__device__ inline double fit_residual(const double x[3], const FitCurve& c, double tau) {
  double r = 0.0;
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    const double e = x[j] - c.b[j] - c.a[j] * cos(tau * (double)(j + 1) * M_PI);
    r += e * e;
  }
  return r;
}

__global__ __launch_bounds__(FIT_BLOCK) void k_fit1d_tau(const double* __restrict__ X, int n,
                                                         FitCurve c, double* __restrict__ tau,
                                                         int32_t* __restrict__ flag,
                                                         double* __restrict__ part) {
  __shared__ double red[13][FIT_BLOCK];
  const int i = blockIdx.x * FIT_BLOCK + threadIdx.x;
  double acc[13];
#pragma unroll
  for (int k = 0; k < 13; ++k) acc[k] = 0.0;
  if (i < n) {
    const double x[3] = {X[i * 3], X[i * 3 + 1], X[i * 3 + 2]};
    // np.roots order c0..c5 (solve_d_R_d_tau_p_3D.py:35-60), stored ascending: e[k] = c_{5-k}
    double e[6];
    e[5] = 48.0 * c.a[2] * c.a[2];
    e[4] = 0.0;
    e[3] = 8.0 * c.a[1] * c.a[1] - 48.0 * c.a[2] * c.a[2];
    e[2] = -12.0 * c.a[2] * (x[2] - c.b[2]);
    e[1] = c.a[0] * c.a[0] - 4.0 * c.a[1] * c.a[1] + 9.0 * c.a[2] * c.a[2] -
           4.0 * c.a[1] * (x[1] - c.b[1]);
    e[0] = -c.a[0] * (x[0] - c.b[0]) + 3.0 * c.a[2] * (x[2] - c.b[2]);
    int amb = (e[5] == 0.0 || e[0] == 0.0) ? 1 : 0;
    double t_best = 0.0;
    if (!amb) {
      // derivatives, each ascending
      double d1[5], d2[4], d3[3], d4[2];
      for (int k = 0; k < 5; ++k) d1[k] = (k + 1) * e[k + 1];
      for (int k = 0; k < 4; ++k) d2[k] = (k + 1) * d1[k + 1];
      for (int k = 0; k < 3; ++k) d3[k] = (k + 1) * d2[k + 1];
      for (int k = 0; k < 2; ++k) d4[k] = (k + 1) * d3[k + 1];
      double r4[1], r3[3], r2[4], r1[5], r0[6], tmp[6];
      int n4 = 0;
      {
        const double z = -d4[0] / d4[1];
        if (z > FIT_LO && z < FIT_HI) r4[n4++] = z;
      }
      int n3 = interior(tmp, segment_roots(d3, 2, d4, r4, n4, tmp), r3);
      int n2 = interior(tmp, segment_roots(d2, 3, d3, r3, n3, tmp), r2);
      int n1 = interior(tmp, segment_roots(d1, 4, d2, r2, n2, tmp), r1);
      const int n0 = segment_roots(e, 5, d1, r1, n1, r0);
      const double scale = fabs(e[0]) + fabs(e[1]) + fabs(e[2]) + fabs(e[3]) + fabs(e[5]);
      const double tol = 1e-10 * scale;
      for (int k = 0; k < n1; ++k)
        if (fabs(peval(e, 5, r1[k])) <= tol) amb = 1;
      if (fabs(peval(e, 5, 1.0)) <= tol || fabs(peval(e, 5, -1.0)) <= tol) amb = 1;
      if (scale > 1e8 * fabs(e[5])) amb = 1;
      // candidates: real roots in [-1, 1], then 0, then 1
      double cand[8], rc[8];
      int nc = 0;
      for (int k = 0; k < n0; ++k)
        if (fabs(r0[k]) <= 1.0) cand[nc++] = acos(r0[k]) / M_PI;
      cand[nc++] = 0.0;
      cand[nc++] = 1.0;
      int best = 0;
      for (int k = 0; k < nc; ++k) {
        rc[k] = fit_residual(x, c, cand[k]);
        if (rc[k] < rc[best]) best = k;
      }
      double rs = 0.0;
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        const double m = fabs(x[j] - c.b[j]) + fabs(c.a[j]);
        rs += m * m;
      }
      for (int k = 0; k < nc; ++k)
        if (k != best && fabs(cand[k] - cand[best]) > 1e-12 && rc[k] - rc[best] <= 1e-12 * rs)
          amb = 1;
      t_best = cand[best];
    }
    tau[i] = t_best;
    flag[i] = amb;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      // tau . (pi j), the operand order of fit_1D_open_manifold_3D.py:85
      const double cj = cos(t_best * (M_PI * (double)(j + 1)));
      acc[j] = cj * cj;
      acc[3 + j] = cj;
      acc[6 + j] = x[j] * cj;
      acc[9 + j] = x[j];
    }
    acc[12] = amb;
  }
#pragma unroll
  for (int k = 0; k < 13; ++k) red[k][threadIdx.x] = acc[k];
  __syncthreads();
  for (int w = FIT_BLOCK / 2; w > 0; w >>= 1) {
    if (threadIdx.x < w)
#pragma unroll
      for (int k = 0; k < 13; ++k) red[k][threadIdx.x] += red[k][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x < 13) part[blockIdx.x * 13 + threadIdx.x] = red[threadIdx.x][0];
}

// sums[k] = sum over blocks of part[b][k], in block order
__global__ void k_fit1d_sums(const double* __restrict__ part, int nb, double* __restrict__ sums) {
  const int k = threadIdx.x;
  if (k >= 13) return;
  double s = 0.0;
  for (int b = 0; b < nb; ++b) s += part[b * 13 + k];
  sums[k] = s;
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

int memi_nlsa_recon_dist(const double* Gq, const double* H, int q_dim, int ncols, int64_t npix,
                         double* inv_sd, float* Drecon2, double* ws, void* stream) {
  MEMI_CHECK_ARG(Gq && H && inv_sd && Drecon2 && ws && q_dim > 0 && q_dim <= 4096 &&
                     ncols > 0 && npix > 0,
                 "memi_nlsa_recon_dist: bad arguments");
  hipStream_t s = (hipStream_t)stream;
  double* T = ws;
  double* dg = ws + (int64_t)q_dim * ncols;
  hipLaunchKernelGGL(k_gq_h, dim3((ncols + 255) / 256, q_dim), dim3(256), 0, s, Gq, H, q_dim,
                     ncols, T);
  MEMI_CHECK_LAUNCH("k_gq_h");
  hipLaunchKernelGGL(k_recon_diag, dim3((ncols + 255) / 256), dim3(256), 0, s, H, T, q_dim, ncols,
                     (double)npix, dg, inv_sd);
  MEMI_CHECK_LAUNCH("k_recon_diag");
  const unsigned nt = (unsigned)((ncols + 63) / 64);
  hipLaunchKernelGGL(k_recon_dist, dim3(nt, nt), dim3(256), 0, s, H, T, dg, inv_sd, q_dim, ncols,
                     Drecon2);
  MEMI_CHECK_LAUNCH("k_recon_dist");
  return MEMI_OK;
}

int memi_fit1d_tau(const double* x, int n, const double* ab_host, double* tau, int32_t* flag,
                   double* sums, double* ws, void* stream) {
  MEMI_CHECK_ARG(x && ab_host && tau && flag && sums && ws && n > 0,
                 "memi_fit1d_tau: bad arguments");
  FitCurve c;
  for (int j = 0; j < 3; ++j) {
    c.a[j] = ab_host[j];
    c.b[j] = ab_host[3 + j];
  }
  const int nb = (n + FIT_BLOCK - 1) / FIT_BLOCK;
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_fit1d_tau, dim3(nb), dim3(FIT_BLOCK), 0, s, x, n, c, tau, flag, ws);
  MEMI_CHECK_LAUNCH("k_fit1d_tau");
  hipLaunchKernelGGL(k_fit1d_sums, dim3(1), dim3(64), 0, s, ws, nb, sums);
  MEMI_CHECK_LAUNCH("k_fit1d_sums");
  return MEMI_OK;
}

}  // extern "C"
