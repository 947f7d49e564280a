// Stage 1 image preparation kernels (the elementwise work between batched FFTs).
// Reference: getDistanceCTF_local_Conj9combinedS2.py:244-367 and its leaves
// annularMask.py:19-32, rotatefill.py:21-28, ctemh_cryoFrank.py:24-44.
#include "common.h"
#include "fft_lds.h"

#include <mutex>

namespace memi {

// annularMask.op(0, N/2, N, N): r^2 < (N/2)^2 about row N/2-1, column N/2.
__device__ inline bool in_disk(int r, int c, int N) {
  const double xd = r - 0.5 * N + 1.0;
  const double yd = c - 0.5 * N;
  return xd * xd + yd * yd < 0.25 * (double)N * (double)N;
}

// One block per image: background statistics then a tiled transpose into the output.
__global__ __launch_bounds__(256) void k_load_normalize(const float* __restrict__ stack,
                                                        const int64_t* __restrict__ src,
                                                        const int32_t* __restrict__ conj, int N,
                                                        float* __restrict__ out) {
  __shared__ double red[8];
  __shared__ float tile[32][33];
  const int64_t img = blockIdx.x;
  const float* raw = stack + src[img] * (int64_t)N * N;
  const bool cj = conj != nullptr && conj[img] != 0;
  const int NN = N * N;
  // raw[a][b] is t[b][a] (t = raw.T); with flipud, t'[r][c] = raw[c][N-1-r].
  double s = 0.0;
  for (int p = threadIdx.x; p < NN; p += blockDim.x) {
    const int a = p / N, b = p % N;
    const int r = cj ? N - 1 - b : b, c = a;
    if (!in_disk(r, c, N)) s += (double)raw[p];
  }
  const double mean = block_sum(s, red) / NN;
  double v = 0.0;
  for (int p = threadIdx.x; p < NN; p += blockDim.x) {
    const int a = p / N, b = p % N;
    const int r = cj ? N - 1 - b : b, c = a;
    const double bg = in_disk(r, c, N) ? 0.0 : (double)raw[p];
    v += (bg - mean) * (bg - mean);
  }
  const double sd = sqrt(block_sum(v, red) / NN);
  float* o = out + img * (int64_t)NN;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  for (int a0 = 0; a0 < N; a0 += 32) {
    for (int b0 = 0; b0 < N; b0 += 32) {
      for (int y = ty; y < 32; y += 8) {
        const int a = a0 + y, b = b0 + tx;
        tile[y][tx] = (a < N && b < N) ? raw[(int64_t)a * N + b] : 0.f;
      }
      __syncthreads();
      // output row r from raw column b, output column c = raw row a
      for (int y = ty; y < 32; y += 8) {
        const int b = b0 + y, a = a0 + tx;
        if (a < N && b < N) {
          const int r = cj ? N - 1 - b : b;
          o[(int64_t)r * N + a] = (float)(((double)tile[tx][y] - mean) / sd);
        }
      }
      __syncthreads();
    }
  }
}

// The same with the raw image staged once in LDS (row stride N + 1, so the transposed reads
// are conflict-free): background moments in one pass (float64 sum and sum of squares; the
// variance of the zero-filled background array is E[bg^2] - mean^2), then the transposed,
// normalised write.  One HBM read and one write per image.
template <int NMAX>
__global__ __launch_bounds__(512) void k_load_normalize_lds(const float* __restrict__ stack,
                                                            const int64_t* __restrict__ src,
                                                            const int32_t* __restrict__ conj,
                                                            int N, float* __restrict__ out) {
  __shared__ float t[NMAX * (NMAX + 1)];
  __shared__ double red[16];
  const int64_t img = blockIdx.x;
  const float* raw = stack + src[img] * (int64_t)N * N;
  const bool cj = conj != nullptr && conj[img] != 0;
  const int S = N + 1;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
  double s = 0.0, s2 = 0.0;
  for (int a = wv; a < N; a += nw)
    for (int b = lane; b < N; b += 64) {
      const float v = raw[a * N + b];
      t[a * S + b] = v;
      // raw[a][b] lands at output (r, c) = (cj ? N-1-b : b, a)
      const int r = cj ? N - 1 - b : b;
      const int xd = 2 * r - N + 2, yd = 2 * a - N;      // annularMask.op(0, N/2, N, N)
      if (xd * xd + yd * yd >= N * N) {
        s += (double)v;
        s2 += (double)v * v;
      }
    }
  const double NN = (double)N * N;
  const double mean = block_sum(s, red) / NN;
  const double var = block_sum(s2, red + 8) / NN - mean * mean;
  const double sd = sqrt(var);
  float* o = out + img * (int64_t)N * N;
  for (int r = wv; r < N; r += nw) {
    const int b = cj ? N - 1 - r : r;
    for (int c = lane; c < N; c += 64)
      o[r * N + c] = (float)(((double)t[c * S + b] - mean) / sd);
  }
}

// cubic B-spline transfer function (4 + 2 cos w) / 6 at FFT index i of n
__device__ inline double bspline_hat(int i, int n) {
  return (4.0 + 2.0 * cospi(2.0 * (double)i / (double)n)) / 6.0;
}

// thread per half-spectrum bin and SPEC_IMG images (factor evaluated once per thread)
constexpr int SPEC_IMG = 4;
__global__ void k_spectrum_scale(float2* __restrict__ spec, int64_t n, int N, int kind,
                                 double Qc, int order) {
  const int Nh = N / 2 + 1;
  const int64_t Kh = (int64_t)N * Nh;
  const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (k >= Kh) return;
  const int col = (int)(k % Nh);
  const int row = (int)(k / Nh);
  double f = 1.0;
  if (kind & (MEMI_FILTER_BUTTER | MEMI_FILTER_GAUSS)) {
    const double q = radial_q(row, col, N) / Qc;
    if (kind & MEMI_FILTER_BUTTER) f *= sqrt(1.0 / (1.0 + pow(q, 2.0 * order)));
    if (kind & MEMI_FILTER_GAUSS) f *= exp(-(log(2.0) / 2.0) * q * q);
  }
  if (kind & MEMI_FILTER_BSPLINE) f /= bspline_hat(row, N) * bspline_hat(col, N);
  const float ff = (float)f;
  // all loads of the chunk are issued before any store (SPEC_IMG images per thread)
  const int64_t i0 = (int64_t)blockIdx.y * SPEC_IMG;
  float2 v[SPEC_IMG];
#pragma unroll
  for (int u = 0; u < SPEC_IMG; ++u)
    if (i0 + u < n) v[u] = spec[(i0 + u) * Kh + k];
#pragma unroll
  for (int u = 0; u < SPEC_IMG; ++u)
    if (i0 + u < n) spec[(i0 + u) * Kh + k] = make_float2(v[u].x * ff, v[u].y * ff);
}

__global__ void k_mask_multiply(float* __restrict__ img, int64_t total, int N, int disk,
                                const float* __restrict__ msk2) {
  const int NN = N * N;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int p = (int)(e % NN);
    float m = 1.f;
    if (disk && !in_disk(p / N, p % N, N)) m = 0.f;
    if (msk2) m *= msk2[p];
    img[e] *= m;
  }
}

// rotatefill: output (r, c) samples the periodic cubic spline at
//   src_r = h + cos*(r-h) + sin*(c-h),  src_c = h - sin*(r-h) + cos*(c-h),  h = (N-1)/2,
// which is scipy.ndimage.rotate's affine map rot_matrix = [[cos, sin], [-sin, cos]] about
// the centre of the 3N tile, restricted to the central copy.
__device__ inline void bspline_w(float t, float w[4]) {
  const float u = 1.f - t;
  const float t2 = t * t, t3 = t2 * t;
  w[0] = u * u * u * (1.f / 6.f);
  w[1] = 2.f / 3.f - t2 + 0.5f * t3;
  w[2] = 2.f / 3.f - u * u + 0.5f * u * u * u;
  w[3] = t3 * (1.f / 6.f);
}

__global__ __launch_bounds__(256) void k_spline_rotate(const float* __restrict__ coef,
                                                       const double* __restrict__ angle_deg,
                                                       int N, float* __restrict__ out) {
  const int NN = N * N;
  const int64_t img = blockIdx.x;
  double sn, cs;
  sincospi(angle_deg[img] / 180.0, &sn, &cs);
  const double h = 0.5 * (N - 1);
  const float* cf = coef + img * NN;
  float* o = out + img * NN;
  for (int p = threadIdx.x; p < NN; p += blockDim.x) {
    const int r = p / N, c = p % N;
    const double dr = r - h, dc = c - h;
    const double sr = h + cs * dr + sn * dc;
    const double sc = h - sn * dr + cs * dc;
    const double fr = floor(sr), fc = floor(sc);
    float wr[4], wc[4];
    bspline_w((float)(sr - fr), wr);
    bspline_w((float)(sc - fc), wc);
    int ir = ((int)fr - 1) % N;
    if (ir < 0) ir += N;
    int ic0 = ((int)fc - 1) % N;
    if (ic0 < 0) ic0 += N;
    int ics[4];
#pragma unroll
    for (int b = 0; b < 4; ++b) ics[b] = (ic0 + b) % N;
    float acc = 0.f;
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      const float* rowp = cf + ir * N;
      float racc = 0.f;
#pragma unroll
      for (int b = 0; b < 4; ++b) racc += wc[b] * rowp[ics[b]];
      acc += wr[a] * racc;
      ir = (ir + 1 == N) ? 0 : ir + 1;
    }
    o[p] = acc;
  }
}

// ---- fused alignment: disk mask, then two cubic-spline rotations, image resident in LDS ----
//
// The periodic cubic B-spline coefficients c of a line s (c = s / ((4 + 2 cos w) / 6) in
// frequency) come from the exact recursive filter with pole z = sqrt(3) - 2:
//   c+[0] = sum_j z^j s[-j mod N] / (1 - z^N),       c+[k] = s[k] + z c+[k-1],
//   c-[N-1] = -z / (1 - z^N) sum_j z^j c+[(N-1+j) mod N],  c-[k] = z (c-[k+1] - c+[k]),
//   c = 6 c-,
// the periodic sums truncated after 24 terms (z^24 < 1e-13).  Rows then columns.
constexpr double BSPLINE_Z = -0.2679491924311227;   // sqrt(3) - 2
constexpr int BSPLINE_NT = 24;

// Each line is cut into nseg segments handled by one thread each (N lines x nseg <= block):
// every segment starts from its own truncated periodic sum, so the recursions are short and
// independent.  Phases: causal starts (read only), causal recursion, anticausal starts (read
// the finished c+), anticausal recursion; the block synchronises between phases.  The start
// sums run in two pieces (before and after the line's wrap) so no term needs a modulo.
template <bool ROWS>
__device__ inline void bspline_lines(float* buf, int N, int S, double zN) {
  const int es = ROWS ? 1 : S;
  const int nt = min(N, BSPLINE_NT);
  const double z = BSPLINE_Z;
  const int nseg = max(1, min((int)blockDim.x / N, (N + 7) / 8));
  const int Ls = (N + nseg - 1) / nseg;
  const int t = threadIdx.x;
  const int l = t % N, sg = t / N;              // consecutive threads: consecutive lines
  const int a = sg * Ls, e = min(a + Ls, N);
  const bool act = t < N * nseg && a < N;
  float* x = buf + (ROWS ? l * S : l);
  double c0 = 0.0;
  if (act) {
    // sum_{j < nt} z^j x[a - j mod N]: j <= a directly, then from the line's end
    double zj = 1.0;
    const int j1 = min(a + 1, nt);
    for (int j = 0; j < j1; ++j) {
      c0 += zj * (double)x[(a - j) * es];
      zj *= z;
    }
    for (int j = j1; j < nt; ++j) {
      c0 += zj * (double)x[(N + a - j) * es];
      zj *= z;
    }
    c0 /= 1.0 - zN;
  }
  __syncthreads();
  if (act) {
    double prev = c0;
    x[a * es] = (float)prev;
    for (int k = a + 1; k < e; ++k) {
      prev = (double)x[k * es] + z * prev;
      x[k * es] = (float)prev;
    }
  }
  __syncthreads();
  double cm = 0.0;
  if (act) {
    // sum_{j < nt} z^j c+[e - 1 + j mod N]: up to the line's end, then from its start
    double zj = 1.0;
    const int j1 = min(N - e + 1, nt);
    for (int j = 0; j < j1; ++j) {
      cm += zj * (double)x[(e - 1 + j) * es];
      zj *= z;
    }
    for (int j = j1; j < nt; ++j) {
      cm += zj * (double)x[(e - 1 + j - N) * es];
      zj *= z;
    }
    cm *= -z / (1.0 - zN);
  }
  __syncthreads();
  if (act) {
    double nxt = cm;
    x[(e - 1) * es] = (float)(6.0 * nxt);
    for (int k = e - 2; k >= a; --k) {
      nxt = z * (nxt - (double)x[k * es]);
      x[k * es] = (float)(6.0 * nxt);
    }
  }
}

// The image is stored with row stride S = N + 3 and three extra rows and columns that repeat
// rows / columns 0..2 (the periodic continuation), so the 4x4 spline taps never wrap.
__device__ inline void refresh_pad(float* buf, int N, int S) {
  for (int q = threadIdx.x; q < N * 3; q += blockDim.x) {
    const int r = q / 3, c = q % 3;
    buf[r * S + N + c] = buf[r * S + c];
  }
  __syncthreads();
  for (int q = threadIdx.x; q < 3 * S; q += blockDim.x) {
    const int r = q / S, c = q % S;
    buf[(N + r) * S + c] = buf[r * S + c];
  }
  __syncthreads();
}

__device__ inline void bspline_prefilter(float* buf, int N, int S, double zN) {
  bspline_lines<true>(buf, N, S, zN);
  __syncthreads();
  bspline_lines<false>(buf, N, S, zN);
  __syncthreads();
  refresh_pad(buf, N, S);
}

// Sample the coefficient image `src` (LDS, row stride S, padded as above) rotated by
// angle_deg into dst (row stride ds).  Source rows/columns are clamped to [-N+2, 2N-3] before
// the taps, which changes nothing for finite angles (|src - h| <= sqrt(2) h) and, with one
// wrap of the first tap into [0, N), keeps all four taps inside the padded image.
__device__ inline void spline_sample(const float* src, int N, int S, double angle_deg, float* dst,
                                     int ds) {
  double sn, cs;
  sincospi(angle_deg / 180.0, &sn, &cs);
  const double h = 0.5 * (N - 1);
  const double lo = 2 - N, hi = 2 * N - 3;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
  for (int r = wv; r < N; r += nw) {
    const double dr = r - h;
    const double br = h + cs * dr, bc = h - sn * dr;
    for (int c = lane; c < N; c += 64) {
      const double dc = c - h;
      const double sr = br + sn * dc;
      const double sc = bc + cs * dc;
      const double fr = floor(sr), fc = floor(sc);
      float wr[4], wc[4];
      bspline_w((float)(sr - fr), wr);
      bspline_w((float)(sc - fc), wc);
      int ir = (int)fmin(fmax(fr, lo), hi) - 1;
      int ic = (int)fmin(fmax(fc, lo), hi) - 1;
      ir += ir < 0 ? N : 0;
      ir -= ir >= N ? N : 0;
      ic += ic < 0 ? N : 0;
      ic -= ic >= N ? N : 0;
      const float* p = src + ir * S + ic;
      float acc = 0.f;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const float* rp = p + q * S;
        acc += wr[q] * (wc[0] * rp[0] + wc[1] * rp[1] + wc[2] * rp[2] + wc[3] * rp[3]);
      }
      dst[r * ds + c] = acc;
    }
  }
}

// spline_sample into the same LDS image: every thread keeps its outputs in registers (a
// 512-thread block gives each thread at most ceil(NMAX/8) rows x ceil(NMAX/64) columns), the
// block synchronises, then writes them back -- one image of LDS per rotation instead of two.
template <int NMAX>
__device__ inline void spline_sample_inplace(float* buf, int N, int S, double angle_deg) {
  constexpr int NW = 8, RI = (NMAX + NW - 1) / NW, CI = (NMAX + 63) / 64;
  double sn, cs;
  sincospi(angle_deg / 180.0, &sn, &cs);
  const double h = 0.5 * (N - 1);
  const double lo = 2 - N, hi = 2 * N - 3;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  float v[RI][CI];
#pragma unroll
  for (int i = 0; i < RI; ++i) {
    const int r = wv + i * NW;
    const double dr = r - h;
    const double br = h + cs * dr, bc = h - sn * dr;
#pragma unroll
    for (int q = 0; q < CI; ++q) {
      const int c = lane + q * 64;
      float acc = 0.f;
      if (r < N && c < N) {
        const double dc = c - h;
        const double sr = br + sn * dc;
        const double sc = bc + cs * dc;
        const double fr = floor(sr), fc = floor(sc);
        float wr[4], wc[4];
        bspline_w((float)(sr - fr), wr);
        bspline_w((float)(sc - fc), wc);
        int ir = (int)fmin(fmax(fr, lo), hi) - 1;
        int ic = (int)fmin(fmax(fc, lo), hi) - 1;
        ir += ir < 0 ? N : 0;
        ir -= ir >= N ? N : 0;
        ic += ic < 0 ? N : 0;
        ic -= ic >= N ? N : 0;
        const float* p = buf + ir * S + ic;
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          const float* rp = p + u * S;
          acc += wr[u] * (wc[0] * rp[0] + wc[1] * rp[1] + wc[2] * rp[2] + wc[3] * rp[3]);
        }
      }
      v[i][q] = acc;
    }
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < RI; ++i)
#pragma unroll
    for (int q = 0; q < CI; ++q) {
      const int r = wv + i * NW, c = lane + q * 64;
      if (r < N && c < N) buf[r * S + c] = v[i][q];
    }
  __syncthreads();
}

// One block of 512 threads per image, the image (row stride N + 3, padded) in LDS.
// NBUF = 2: mask, rotate by ang1 in place, rotate by ang2 into out.  NBUF = 1: (mask if disk,)
// rotate by ang1 into out; the host runs it twice.  in and out may be the same buffer (each
// block reads its whole image before writing).
template <int NMAX, int NBUF>
__global__ __launch_bounds__(512) void k_align_images(const float* in, const double* ang1,
                                                      const double* ang2, int N, int disk,
                                                      float* out) {
  __shared__ float lds[(NMAX + 3) * (NMAX + 3)];
  const int S = N + 3;
  const int64_t img = blockIdx.x;
  const float* src = in + img * N * N;
  float* A = lds;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
  for (int r = wv; r < N; r += nw)
    for (int c = lane; c < N; c += 64) {
      float v = src[r * N + c];
      if (disk) {
        const int xd = 2 * r - N + 2, yd = 2 * c - N;     // annularMask.op(0, N/2, N, N)
        if (xd * xd + yd * yd >= N * N) v = 0.f;
      }
      A[r * S + c] = v;
    }
  __syncthreads();
  const double zN = pow(BSPLINE_Z, (double)N);
  bspline_prefilter(A, N, S, zN);
  float* o = out + img * N * N;
  if constexpr (NBUF == 2) {
    spline_sample_inplace<NMAX>(A, N, S, ang1[img]);
    bspline_prefilter(A, N, S, zN);
    spline_sample(A, N, S, ang2[img], o, N);
  } else {
    spline_sample(A, N, S, ang1[img], o, N);
  }
}

__device__ inline uint16_t bf16_rne(float f) {
  uint32_t u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}
__device__ inline float bf16_to_f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

// thread per half-spectrum bin k (grid.y: image chunks).  The CTF phase
// wr = (w1/2 k^2 - w2) k^2 = a_k - c_k df is formed in float64, reduced to [-pi, pi] in float64,
// and only then passed to the float32 sincos (|C| error ~1e-7).
constexpr int CTF_IMG = 4;
__global__ void k_ctf_pack(const float2* __restrict__ fy, const double* __restrict__ df,
                           int64_t n, int N, memi_ctf_params cp, int prec, int64_t Khp,
                           int64_t row0, void* __restrict__ op_hi, void* __restrict__ op_lo,
                           float* __restrict__ ctf_half) {
  const int Nh = N / 2 + 1;
  const int64_t Kh = (int64_t)N * Nh;
  const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (k >= Khp) return;
  const bool live = k < Kh;
  const int row = live ? (int)(k / Nh) : 0, col = live ? (int)(k % Nh) : 0;
  const double kk = live ? radial_q(row, col, N) / (2.0 * cp.pix_size) : 0.0;
  const double cs = cp.cs_mm * 1.0e7;
  const double wav = 12.3986 / sqrt((2.0 * 511.0 + cp.ekv) * cp.ekv);
  const double k2 = kk * kk;
  const double a_k = 0.5 * (M_PI * cs * wav * wav * wav) * k2 * k2;
  const double c_k = M_PI * wav * k2;
  double env = 1.0;
  if (isfinite(cp.gauss_env)) {
    const double sigm = cp.gauss_env / sqrt(2.0 * log(2.0));
    env = exp(-k2 / (2.0 * sigm * sigm));
  }
  const double w = live ? half_weight(col, N) : 0.0;
  const double sw = sqrt(w);
  const double inv2pi = 0.15915494309189535, twopi = 6.283185307179586;
  // all loads of the chunk are issued before any store (CTF_IMG images per thread)
  const int64_t i0 = (int64_t)blockIdx.y * CTF_IMG;
  double dfv[CTF_IMG];
  float2 Xv[CTF_IMG];
#pragma unroll
  for (int u = 0; u < CTF_IMG; ++u) {
    const bool ok = i0 + u < n;
    dfv[u] = ok ? df[i0 + u] : 0.0;
    Xv[u] = (ok && live) ? fy[(i0 + u) * Kh + k] : make_float2(0.f, 0.f);
  }
#pragma unroll
  for (int u = 0; u < CTF_IMG; ++u) {
    const int64_t i = i0 + u;
    if (i >= n) break;
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    if (live) {
      const double wr = a_k - c_k * dfv[u];
      const double red = wr - twopi * rint(wr * inv2pi);
      float sf, cf;
      sincosf((float)red, &sf, &cf);
      const double C = ((double)sf - cp.ampc * (double)cf) * env;
      ctf_half[i * Kh + k] = (float)C;
      const float2 X = Xv[u];
      v[0] = w * ((double)X.x * X.x + (double)X.y * X.y);
      v[1] = C * C;
      v[2] = sw * C * X.x;
      v[3] = sw * C * X.y;
    }
    const int64_t base = (row0 + i) * 4 * Khp + k;
    if (prec == MEMI_PREC_FP32) {
      float* o = (float*)op_hi;
#pragma unroll
      for (int sg = 0; sg < 4; ++sg) o[base + sg * Khp] = (float)v[sg];
    } else {
      uint16_t* oh = (uint16_t*)op_hi;
      uint16_t* ol = (uint16_t*)op_lo;
#pragma unroll
      for (int sg = 0; sg < 4; ++sg) {
        const float f = (float)v[sg];
        const uint16_t h = bf16_rne(f);
        oh[base + sg * Khp] = h;
        ol[base + sg * Khp] = bf16_rne((float)((double)f - (double)bf16_to_f(h)));
      }
    }
  }
}

__global__ void k_phase_flip(const float2* __restrict__ fy, const float* __restrict__ ctf,
                             int64_t total, float2* __restrict__ out) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const float c = ctf[e];
    const float s = (c > 0.f) ? 1.f : ((c < 0.f) ? -1.f : 0.f);
    float2 v = fy[e];
    v.x *= s;
    v.y *= s;
    out[e] = v;
  }
}

// thread per half-spectrum bin: loop over images (coalesced across bins)
__global__ void k_wiener_average(const float2* __restrict__ spec, const float* __restrict__ ctf,
                                 int64_t n, int64_t Kh, double snr, float2* __restrict__ out) {
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < Kh;
       k += (int64_t)gridDim.x * blockDim.x) {
    double nr = 0.0, ni = 0.0, den = 0.0;
    for (int64_t i = 0; i < n; ++i) {
      const double c = ctf[i * Kh + k];
      const float2 x = spec[i * Kh + k];
      nr += x.x * c;
      ni += x.y * c;
      den += c * c;
    }
    const double w = -(den + 1.0 / snr);
    out[k] = make_float2((float)(nr / w), (float)(ni / w));
  }
}

__global__ void k_image_moments(const float* __restrict__ img, int64_t n, int64_t npix,
                                double* __restrict__ sum, double* __restrict__ sumsq) {
  for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < npix;
       p += (int64_t)gridDim.x * blockDim.x) {
    double s = 0.0, s2 = 0.0;
    for (int64_t i = 0; i < n; ++i) {
      const double v = img[i * npix + p];
      s += v;
      s2 += v * v;
    }
    sum[p] = s;
    sumsq[p] = s2;
  }
}

// ---- fused image preparation (a1-a6 in one kernel) ------------------------------------
// Per half-spectrum bin k (row-major [ky][kx]): {Butterworth factor / N^2, a_k, c_k,
// envelope, w_k, sqrt(w_k)} with the CTF phase wr = a_k - c_k df (ctemh_cryoFrank.py:24-44).
constexpr int PT = 6;

__global__ void k_prep_tables(int N, int kind, double Qc, int order, memi_ctf_params cp,
                              double* __restrict__ tab) {
  const int Nh = N / 2 + 1;
  const int64_t Kh = (int64_t)N * Nh;
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < Kh;
       k += (int64_t)gridDim.x * blockDim.x) {
    const int row = (int)(k / Nh), col = (int)(k % Nh);
    double f = 1.0;
    if (kind & (MEMI_FILTER_BUTTER | MEMI_FILTER_GAUSS)) {
      const double q = radial_q(row, col, N) / Qc;
      if (kind & MEMI_FILTER_BUTTER) f *= sqrt(1.0 / (1.0 + pow(q, 2.0 * order)));
      if (kind & MEMI_FILTER_GAUSS) f *= exp(-(log(2.0) / 2.0) * q * q);
    }
    const double kk = radial_q(row, col, N) / (2.0 * cp.pix_size);
    const double cs = cp.cs_mm * 1.0e7;
    const double wav = 12.3986 / sqrt((2.0 * 511.0 + cp.ekv) * cp.ekv);
    const double k2 = kk * kk;
    double env = 1.0;
    if (isfinite(cp.gauss_env)) {
      const double sigm = cp.gauss_env / sqrt(2.0 * log(2.0));
      env = exp(-k2 / (2.0 * sigm * sigm));
    }
    const double w = half_weight(col, N);
    double* t = tab + k * PT;
    t[0] = f / ((double)N * N);
    t[1] = 0.5 * (M_PI * cs * wav * wav * wav) * k2 * k2;
    t[2] = M_PI * wav * k2;
    t[3] = env;
    t[4] = w;
    t[5] = sqrt(w);
  }
}

template <int NMAX>
__global__ __launch_bounds__(512) void k_prep_images(
    const float* __restrict__ stack, const int64_t* __restrict__ src,
    const int32_t* __restrict__ conj, const double* __restrict__ rot1,
    const double* __restrict__ rot2, const double* __restrict__ df, int N,
    const double* __restrict__ tab, double ampc, int prec, int64_t Khp, int64_t row0,
    void* __restrict__ op_hi, void* __restrict__ op_lo, float2* __restrict__ fy,
    float* __restrict__ ctf_half, float* __restrict__ img_out) {
  // one image buffer: the padded real image ((N+3)^2 floats) is the largest of its layouts
  // (half spectrum 2 N (N/2+1) floats, row-pair lines 2 ceil(N/2) N floats)
  constexpr int BUF = (NMAX + 3) * (NMAX + 3);
  __shared__ float lds[BUF];
  __shared__ cpx tw[NMAX];
  __shared__ int pos[NMAX], rev[NMAX];              // frequency <-> position after the passes
  __shared__ double red[16];
  float* A = lds;
  const int S = N + 3;
  const int Nh = N / 2 + 1;
  const int64_t Kh = (int64_t)N * Nh;
  const int64_t img = blockIdx.x;
  int rad[8];
  const int nrad = radix_plan(N, rad);
  fft_twiddles(tw, N);
  fft_digit_tables(N, rad, nrad, pos, rev);
  // a1: load (conjugates flipped, SPIDER transpose) and background statistics
  const float* raw = stack + src[img] * (int64_t)N * N;
  const bool cj = conj != nullptr && conj[img] != 0;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
  double s1 = 0.0, s2 = 0.0;
  // every load of this thread in flight at once, then the LDS stores and the statistics
  constexpr int PER = (NMAX * NMAX + 511) / 512;
  const int NNi = N * N;
  float vv[PER];
#pragma unroll
  for (int u = 0; u < PER; ++u) {
    const int p = threadIdx.x + u * 512;
    vv[u] = p < NNi ? raw[p] : 0.f;
  }
  const float inv_n = 1.0f / N;
#pragma unroll
  for (int u = 0; u < PER; ++u) {
    const int p = threadIdx.x + u * 512;
    if (p < NNi) {
      const int a = fdivi(p, N, inv_n), b = p - a * N;
      const int r = cj ? N - 1 - b : b;           // raw[a][b] lands at (r, a)
      A[r * S + a] = vv[u];
      const int xd = 2 * r - N + 2, yd = 2 * a - N;
      if (xd * xd + yd * yd >= N * N) {
        s1 += (double)vv[u];
        s2 += (double)vv[u] * vv[u];
      }
    }
  }
  const double NN = (double)N * N;
  const double mean = block_sum(s1, red) / NN;
  const double sd = sqrt(block_sum(s2, red + 8) / NN - mean * mean);
  for (int r = wv; r < N; r += nw)
    for (int c = lane; c < N; c += 64) A[r * S + c] = (float)(((double)A[r * S + c] - mean) / sd);
  __syncthreads();
  // a2: Butterworth in frequency (the table folds in the 1/N^2 of the inverse)
  cpx* H = rfft2_ip<NMAX>(A, N, S, tw, rad, nrad, pos);
  const float inv_nh = 1.0f / Nh;
  for (int q = threadIdx.x; q < Kh; q += blockDim.x) {   // H row p holds ky = rev[p]
    const int py = fdivi(q, Nh, inv_nh), kx = q - py * Nh;
    const float f = (float)tab[(int64_t)(rev[py] * Nh + kx) * PT];
    H[q] = make_float2(H[q].x * f, H[q].y * f);
  }
  __syncthreads();
  irfft2_ip<NMAX>(A, N, S, tw, rad, nrad, pos);
  // a4: disk mask, then the two spline rotations, each in place
  for (int r = wv; r < N; r += nw)
    for (int c = lane; c < N; c += 64) {
      const int xd = 2 * r - N + 2, yd = 2 * c - N;
      if (xd * xd + yd * yd >= N * N) A[r * S + c] = 0.f;
    }
  __syncthreads();
  const double zN = pow(BSPLINE_Z, (double)N);
  bspline_prefilter(A, N, S, zN);
  spline_sample_inplace<NMAX>(A, N, S, rot1[img]);
  bspline_prefilter(A, N, S, zN);
  spline_sample_inplace<NMAX>(A, N, S, rot2[img]);
  if (img_out) {
    float* o = img_out + img * (int64_t)N * N;
    for (int r = wv; r < N; r += nw)
      for (int c = lane; c < N; c += 64) o[r * N + c] = A[r * S + c];
  }
  // a5/a6: spectrum of the aligned image, CTF, operand rows (as k_ctf_pack)
  cpx* F = rfft2_ip<NMAX>(A, N, S, tw, rad, nrad, pos);
  const double dfi = df[img];
  const double inv2pi = 0.15915494309189535, twopi = 6.283185307179586;
  const int64_t base = (row0 + img) * 4 * Khp;
  for (int64_t q = threadIdx.x; q < Khp; q += blockDim.x) {
    double v[4] = {0.0, 0.0, 0.0, 0.0};
    if (q < Kh) {
      const double* t = tab + q * PT;
      const double wr = t[1] - t[2] * dfi;
      const double rd = wr - twopi * rint(wr * inv2pi);
      float sf, cf;
      sincosf((float)rd, &sf, &cf);
      const double C = ((double)sf - ampc * (double)cf) * t[3];
      const int ky = fdivi((int)q, Nh, inv_nh), kx = (int)q - ky * Nh;
      const cpx X = F[pos[ky] * Nh + kx];
      ctf_half[img * Kh + q] = (float)C;
      fy[img * Kh + q] = X;
      v[0] = t[4] * ((double)X.x * X.x + (double)X.y * X.y);
      v[1] = C * C;
      v[2] = t[5] * C * X.x;
      v[3] = t[5] * C * X.y;
    }
    if (prec == MEMI_PREC_FP32) {
      float* o = (float*)op_hi;
#pragma unroll
      for (int sg = 0; sg < 4; ++sg) o[base + sg * Khp + q] = (float)v[sg];
    } else {
      uint16_t* oh = (uint16_t*)op_hi;
      uint16_t* ol = (uint16_t*)op_lo;
#pragma unroll
      for (int sg = 0; sg < 4; ++sg) {
        const float f = (float)v[sg];
        const uint16_t h = bf16_rne(f);
        oh[base + sg * Khp + q] = h;
        ol[base + sg * Khp + q] = bf16_rne((float)((double)f - (double)bf16_to_f(h)));
      }
    }
  }
}

// Device scratch of memi_prep_images_noalloc: the per-bin tables (PT doubles per half-spectrum
// bin) of the largest supported image side.  One copy per GPU, held until the process ends;
// host code reaches it only through the address hipGetSymbolAddress returns for the current
// device (the name itself is a host-side stand-in).
constexpr int64_t PREP_SCRATCH_DOUBLES =
    (int64_t)PT * MEMI_PREP_MAX_N * (MEMI_PREP_MAX_N / 2 + 1);
static __device__ __align__(256) double scratch[PREP_SCRATCH_DOUBLES];

constexpr int SCRATCH_MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static double* g_scratch_ptr[SCRATCH_MAX_DEVICES] = {};

// The current device's address of `scratch`, looked up on the first call on that device.
static int scratch_for_current_device(double** out) {
  int dev = -1;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= SCRATCH_MAX_DEVICES) {
    set_error("memi_prep_images_noalloc: no usable current device (%d)", dev);
    return MEMI_ELAUNCH;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[dev] == nullptr) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || p == nullptr) {
      set_error("memi_prep_images_noalloc: hipGetSymbolAddress(scratch) failed on device %d: %s",
                dev, hipGetErrorString(e));
      return MEMI_ELAUNCH;
    }
    g_scratch_ptr[dev] = (double*)p;
  }
  *out = g_scratch_ptr[dev];
  return MEMI_OK;
}

}  // namespace memi

using namespace memi;

extern "C" {

int memi_load_normalize(const float* stack, int64_t n_stack, const int64_t* src,
                        const int32_t* conj, int64_t n, int N, float* out, void* stream) {
  MEMI_CHECK_ARG(stack && src && out, "memi_load_normalize: null pointer");
  MEMI_CHECK_ARG(n >= 0 && N > 0 && n_stack >= 0, "memi_load_normalize: bad sizes");
  if (n == 0) return MEMI_OK;
  hipStream_t s = (hipStream_t)stream;
  if (N <= 64)
    hipLaunchKernelGGL(k_load_normalize_lds<64>, dim3((unsigned)n), dim3(512), 0, s, stack, src,
                       conj, N, out);
  else if (N <= 96)
    hipLaunchKernelGGL(k_load_normalize_lds<96>, dim3((unsigned)n), dim3(512), 0, s, stack, src,
                       conj, N, out);
  else if (N <= 141)
    hipLaunchKernelGGL(k_load_normalize_lds<141>, dim3((unsigned)n), dim3(512), 0, s, stack,
                       src, conj, N, out);
  else if (N <= 200)
    hipLaunchKernelGGL(k_load_normalize_lds<200>, dim3((unsigned)n), dim3(512), 0, s, stack,
                       src, conj, N, out);
  else
    hipLaunchKernelGGL(k_load_normalize, dim3((unsigned)n), dim3(256), 0, s, stack, src, conj, N,
                       out);
  MEMI_CHECK_LAUNCH("k_load_normalize");
  return MEMI_OK;
}

int memi_spectrum_scale(void* spec, int64_t n, int N, int kind, double Qc, int order,
                        void* stream) {
  MEMI_CHECK_ARG(spec && n >= 0 && N > 0, "memi_spectrum_scale: bad arguments");
  MEMI_CHECK_ARG((kind & ~(MEMI_FILTER_BUTTER | MEMI_FILTER_GAUSS | MEMI_FILTER_BSPLINE)) == 0,
                 "memi_spectrum_scale: unknown filter kind %d", kind);
  MEMI_CHECK_ARG(!(kind & (MEMI_FILTER_BUTTER | MEMI_FILTER_GAUSS)) || Qc > 0,
                 "memi_spectrum_scale: Qc must be positive");
  const int64_t Kh = (int64_t)N * (N / 2 + 1);
  if (n == 0) return MEMI_OK;
  MEMI_CHECK_ARG((n + SPEC_IMG - 1) / SPEC_IMG < 65536, "memi_spectrum_scale: too many images "
                 "(at most %d per call)", 65535 * SPEC_IMG);
  dim3 g((unsigned)((Kh + 255) / 256), (unsigned)((n + SPEC_IMG - 1) / SPEC_IMG));
  hipLaunchKernelGGL(k_spectrum_scale, g, dim3(256), 0, (hipStream_t)stream, (float2*)spec, n, N,
                     kind, Qc, order);
  MEMI_CHECK_LAUNCH("k_spectrum_scale");
  return MEMI_OK;
}

int memi_mask_multiply(float* img, int64_t n, int N, int disk, const float* msk2, void* stream) {
  MEMI_CHECK_ARG(img && n >= 0 && N > 0, "memi_mask_multiply: bad arguments");
  const int64_t total = n * (int64_t)N * N;
  if (total == 0 || (!disk && !msk2)) return MEMI_OK;
  hipLaunchKernelGGL(k_mask_multiply, dim3(grid_for(total)), dim3(256), 0, (hipStream_t)stream,
                     img, total, N, disk, msk2);
  MEMI_CHECK_LAUNCH("k_mask_multiply");
  return MEMI_OK;
}

int memi_spline_rotate(const float* coef, const double* angle_deg, int64_t n, int N, float* out,
                       void* stream) {
  MEMI_CHECK_ARG(coef && angle_deg && out && n >= 0 && N > 0, "memi_spline_rotate: bad args");
  MEMI_CHECK_ARG(coef != out, "memi_spline_rotate: in-place rotation is not supported");
  if (n == 0) return MEMI_OK;
  MEMI_CHECK_ARG(n < ((int64_t)1 << 31), "memi_spline_rotate: too many images");
  hipLaunchKernelGGL(k_spline_rotate, dim3((unsigned)n), dim3(256), 0, (hipStream_t)stream, coef,
                     angle_deg, N, out);
  MEMI_CHECK_LAUNCH("k_spline_rotate");
  return MEMI_OK;
}

int memi_align_images(const float* in, const double* rot1_deg, const double* rot2_deg, int64_t n,
                      int N, int disk, float* out, void* stream) {
  MEMI_CHECK_ARG(in && rot1_deg && rot2_deg && out && n >= 0, "memi_align_images: bad arguments");
  MEMI_CHECK_ARG(N >= 8 && N <= MEMI_ALIGN_MAX_N,
                 "memi_align_images: N=%d outside [8, %d]; use memi_spline_rotate", N,
                 MEMI_ALIGN_MAX_N);
  if (n == 0) return MEMI_OK;
  MEMI_CHECK_ARG(n < ((int64_t)1 << 31), "memi_align_images: too many images");
  hipStream_t s = (hipStream_t)stream;
  const dim3 g((unsigned)n), b(512);
  if (N <= 64) {
    hipLaunchKernelGGL((k_align_images<64, 2>), g, b, 0, s, in, rot1_deg, rot2_deg, N, disk, out);
  } else if (N <= 96) {          // 39 KB of LDS: four blocks per CU
    hipLaunchKernelGGL((k_align_images<96, 2>), g, b, 0, s, in, rot1_deg, rot2_deg, N, disk, out);
  } else if (N <= 140) {
    hipLaunchKernelGGL((k_align_images<140, 2>), g, b, 0, s, in, rot1_deg, rot2_deg, N, disk, out);
  } else {
    hipLaunchKernelGGL((k_align_images<MEMI_ALIGN_MAX_N, 1>), g, b, 0, s, in, rot1_deg, rot2_deg,
                       N, disk, out);
    MEMI_CHECK_LAUNCH("k_align_images");
    hipLaunchKernelGGL((k_align_images<MEMI_ALIGN_MAX_N, 1>), g, b, 0, s, out, rot2_deg, rot2_deg,
                       N, 0, out);
  }
  MEMI_CHECK_LAUNCH("k_align_images");
  return MEMI_OK;
}

int memi_prep_images(const float* stack, int64_t n_stack, const int64_t* src,
                     const int32_t* conj, const double* rot1_deg, const double* rot2_deg,
                     const double* df, int64_t n, int N, int filter_kind, double Qc,
                     int order, const memi_ctf_params* ctf, int precision, int64_t Khp,
                     int64_t row0, void* op_hi, void* op_lo, void* fy, float* ctf_half,
                     float* img_out, double* ws, int64_t ws_doubles, void* stream) {
  MEMI_CHECK_ARG(stack && src && rot1_deg && rot2_deg && df && ctf && op_hi && fy && ctf_half &&
                 ws && n >= 0 && n_stack >= 0, "memi_prep_images: null pointer or bad size");
  MEMI_CHECK_ARG(precision == MEMI_PREC_FP32 || (precision == MEMI_PREC_SPLIT_BF16 && op_lo),
                 "memi_prep_images: precision %d needs its operand buffers", precision);
  int rad[8];
  MEMI_CHECK_ARG(N >= 8 && N <= MEMI_PREP_MAX_N && radix_plan(N, rad) > 0,
                 "memi_prep_images: N=%d unsupported (8 <= N <= %d, factors 2, 3, 5, 7)", N,
                 MEMI_PREP_MAX_N);
  const int64_t Kh = (int64_t)N * (N / 2 + 1);
  MEMI_CHECK_ARG(Khp >= Kh && Khp % 64 == 0, "memi_prep_images: bad Khp %lld", (long long)Khp);
  MEMI_CHECK_ARG(ws_doubles >= Kh * PT, "memi_prep_images: workspace %lld < %lld doubles",
                 (long long)ws_doubles, (long long)(Kh * PT));
  MEMI_CHECK_ARG((filter_kind & ~(MEMI_FILTER_BUTTER | MEMI_FILTER_GAUSS)) == 0 &&
                     (!filter_kind || Qc > 0),
                 "memi_prep_images: bad filter kind %d / Qc", filter_kind);
  MEMI_CHECK_ARG(ctf->pix_size > 0, "memi_prep_images: pixel size must be positive");
  if (n == 0) return MEMI_OK;
  MEMI_CHECK_ARG(n < ((int64_t)1 << 31), "memi_prep_images: too many images");
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_prep_tables, dim3(grid_for(Kh)), dim3(256), 0, s, N, filter_kind, Qc,
                     order, *ctf, ws);
  MEMI_CHECK_LAUNCH("k_prep_tables");
  const dim3 g((unsigned)n), b(512);
#define MEMI_PREP(NM)                                                                          \
  hipLaunchKernelGGL(k_prep_images<NM>, g, b, 0, s, stack, src, conj, rot1_deg, rot2_deg, df, N, \
                     ws, ctf->ampc, precision, Khp, row0, op_hi, op_lo, (float2*)fy, ctf_half,  \
                     img_out)
  if (N <= 64) MEMI_PREP(64);
  else if (N <= 96) MEMI_PREP(96);
  else if (N <= 128) MEMI_PREP(128);
  else MEMI_PREP(160);
#undef MEMI_PREP
  MEMI_CHECK_LAUNCH("k_prep_images");
  return MEMI_OK;
}

int memi_prep_images_noalloc(const float* stack, int64_t n_stack, const int64_t* src,
                             const int32_t* conj, const double* rot1_deg, const double* rot2_deg,
                             const double* df, int64_t n, int N, int filter_kind, double Qc,
                             int order, const memi_ctf_params* ctf, int precision, int64_t Khp,
                             int64_t row0, void* op_hi, void* op_lo, void* fy, float* ctf_half,
                             float* img_out, void* stream) {
  const int64_t need = (int64_t)PT * N * (N / 2 + 1);
  MEMI_CHECK_ARG(N >= 8 && N <= MEMI_PREP_MAX_N && need <= PREP_SCRATCH_DOUBLES,
                 "memi_prep_images_noalloc: N=%d needs %lld doubles of tables, beyond the %lld of "
                 "the static scratch (at most N = %d)", N, (long long)need,
                 (long long)PREP_SCRATCH_DOUBLES, MEMI_PREP_MAX_N);
  double* ws = nullptr;
  const int st = scratch_for_current_device(&ws);
  if (st != MEMI_OK) return st;
  // the tables are rewritten in full by every call (k_prep_tables), so nothing is cleared
  return memi_prep_images(stack, n_stack, src, conj, rot1_deg, rot2_deg, df, n, N, filter_kind,
                          Qc, order, ctf, precision, Khp, row0, op_hi, op_lo, fy, ctf_half,
                          img_out, ws, PREP_SCRATCH_DOUBLES, stream);
}

int memi_ctf_pack(const void* fy, const double* df, int64_t n, int N, const memi_ctf_params* ctf,
                  int precision, int64_t Khp, int64_t row0, void* op_hi, void* op_lo,
                  float* ctf_half, void* stream) {
  MEMI_CHECK_ARG(fy && df && ctf && op_hi && ctf_half, "memi_ctf_pack: null pointer");
  MEMI_CHECK_ARG(precision == MEMI_PREC_FP32 || (precision == MEMI_PREC_SPLIT_BF16 && op_lo),
                 "memi_ctf_pack: precision %d needs its operand buffers", precision);
  const int64_t Kh = (int64_t)N * (N / 2 + 1);
  MEMI_CHECK_ARG(Khp >= Kh && Khp % 64 == 0, "memi_ctf_pack: Khp=%lld must be >= Kh=%lld and a "
                 "multiple of 64", (long long)Khp, (long long)Kh);
  MEMI_CHECK_ARG(ctf->pix_size > 0, "memi_ctf_pack: pixel size must be positive");
  if (n == 0) return MEMI_OK;
  MEMI_CHECK_ARG((n + CTF_IMG - 1) / CTF_IMG < 65536, "memi_ctf_pack: too many images (at most "
                 "%d per call)", 65535 * CTF_IMG);
  dim3 g((unsigned)((Khp + 255) / 256), (unsigned)((n + CTF_IMG - 1) / CTF_IMG));
  hipLaunchKernelGGL(k_ctf_pack, g, dim3(256), 0, (hipStream_t)stream, (const float2*)fy, df, n, N,
                     *ctf, precision, Khp, row0, op_hi, op_lo, ctf_half);
  MEMI_CHECK_LAUNCH("k_ctf_pack");
  return MEMI_OK;
}

int memi_phase_flip(const void* fy, const float* ctf_half, int64_t n, int N, void* out,
                    void* stream) {
  MEMI_CHECK_ARG(fy && ctf_half && out, "memi_phase_flip: null pointer");
  const int64_t total = n * (int64_t)N * (N / 2 + 1);
  if (total == 0) return MEMI_OK;
  hipLaunchKernelGGL(k_phase_flip, dim3(grid_for(total)), dim3(256), 0, (hipStream_t)stream,
                     (const float2*)fy, ctf_half, total, (float2*)out);
  MEMI_CHECK_LAUNCH("k_phase_flip");
  return MEMI_OK;
}

int memi_wiener_average(const void* spec, const float* ctf_half, int64_t n, int N, double snr,
                        void* out, void* stream) {
  MEMI_CHECK_ARG(spec && ctf_half && out && snr > 0, "memi_wiener_average: bad arguments");
  const int64_t Kh = (int64_t)N * (N / 2 + 1);
  hipLaunchKernelGGL(k_wiener_average, dim3(grid_for(Kh, 64)), dim3(64), 0, (hipStream_t)stream,
                     (const float2*)spec, ctf_half, n, Kh, snr, (float2*)out);
  MEMI_CHECK_LAUNCH("k_wiener_average");
  return MEMI_OK;
}

int memi_image_moments(const float* img, int64_t n, int64_t npix, double* sum, double* sumsq,
                       void* stream) {
  MEMI_CHECK_ARG(img && sum && sumsq && npix > 0, "memi_image_moments: bad arguments");
  hipLaunchKernelGGL(k_image_moments, dim3(grid_for(npix, 64)), dim3(64), 0,
                     (hipStream_t)stream, img, n, npix, sum, sumsq);
  MEMI_CHECK_LAUNCH("k_image_moments");
  return MEMI_OK;
}

}  // extern "C"
