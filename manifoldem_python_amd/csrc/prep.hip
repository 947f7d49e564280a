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

// 6 x bspline_w: {u^3, 4 - 6 t^2 + 3 t^3, 4 - 6 u^2 + 3 u^3, t^3}, u = 1 - t
__device__ __forceinline__ void bspline_w6(float t, float w[4]) {
  const float u = 1.f - t;
  const float t2 = t * t, u2 = u * u;
  w[0] = u2 * u;
  w[1] = fmaf(t2, fmaf(3.f, t, -6.f), 4.f);
  w[2] = fmaf(u2, fmaf(3.f, u, -6.f), 4.f);
  w[3] = t2 * t;
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
// the periodic sums truncated after BSPLINE_NT terms.  Rows then columns.
constexpr double BSPLINE_Z = -0.2679491924311227;   // sqrt(3) - 2
constexpr int BSPLINE_NT = 16;   // |z|^16 / (1 - |z|) < 1e-9: below float32 rounding

// Cubic B-spline prefilter (the interpolation coefficients of scipy's spline_filter with
// periodic lines) of every row (ROWS) or column of the image, in place, in float32: the
// causal recursion c+[k] = x[k] + z c+[k-1] and the anticausal c[k] = z (c[k+1] - c+[k]),
// times 6.  Each line is cut into nseg segments run by consecutive threads; a segment's first
// value comes from the truncated periodic sum over the BSPLINE_NT preceding samples (causal)
// or following coefficients (anticausal), exact to float rounding since |z|^16 < 1e-9.
template <bool ROWS>
__device__ __forceinline__ void bspline_lines(float* buf, int N, int S, float zN) {
  const int es = ROWS ? 1 : S;
  const int nt = min(N, BSPLINE_NT);
  const float z = (float)BSPLINE_Z;
  const int nseg = max(1, min((int)blockDim.x / N, (N + 7) / 8));
  const int Ls = (N + nseg - 1) / nseg;
  const int t = threadIdx.x;
  const int l = t % N, sg = t / N;              // consecutive threads: consecutive lines
  const int a = sg * Ls, e = min(a + Ls, N);
  const bool act = t < N * nseg && a < N;
  float* x = buf + (ROWS ? l * S : l);
  float c0 = 0.f;
  if (act) {
    // sum_{j < nt} z^j x[a - j mod N]: j <= a directly, then from the line's end
    float zj = 1.f;
    const int j1 = min(a + 1, nt);
    for (int j = 0; j < j1; ++j) {
      c0 = fmaf(zj, x[(a - j) * es], c0);
      zj *= z;
    }
    for (int j = j1; j < nt; ++j) {
      c0 = fmaf(zj, x[(N + a - j) * es], c0);
      zj *= z;
    }
    c0 /= 1.f - zN;
  }
  __syncthreads();
  if (act) {
    float prev = c0;
    x[a * es] = prev;
    for (int k = a + 1; k < e; ++k) {
      prev = fmaf(z, prev, x[k * es]);
      x[k * es] = prev;
    }
  }
  __syncthreads();
  float cm = 0.f;
  if (act) {
    // sum_{j < nt} z^j c+[e - 1 + j mod N]: up to the line's end, then from its start
    float zj = 1.f;
    const int j1 = min(N - e + 1, nt);
    for (int j = 0; j < j1; ++j) {
      cm = fmaf(zj, x[(e - 1 + j) * es], cm);
      zj *= z;
    }
    for (int j = j1; j < nt; ++j) {
      cm = fmaf(zj, x[(e - 1 + j - N) * es], cm);
      zj *= z;
    }
    cm *= -z / (1.f - zN);
  }
  __syncthreads();
  if (act) {
    float nxt = cm;
    x[(e - 1) * es] = 6.f * nxt;
    for (int k = e - 2; k >= a; --k) {
      nxt = z * (nxt - x[k * es]);
      x[k * es] = 6.f * nxt;
    }
  }
}

// The image is stored with row stride S = N + 3 and three extra rows and columns that repeat
// rows / columns 0..2 (the periodic continuation), so the 4x4 spline taps never wrap.
__device__ __forceinline__ void refresh_pad(float* buf, int N, int S) {
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

__device__ __forceinline__ void bspline_prefilter(float* buf, int N, int S, float zN) {
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

// spline_sample into the same LDS image: thread t evaluates pixels p = t + NT k (row p / N,
// column p % N) into registers, the block synchronises, then writes them back -- one image of
// LDS per rotation instead of two.  Source coordinates in float64 (as the reference's), the
// tap indices clamped as in spline_sample.  Inlined into its kernel, so the image is addressed
// as LDS and, for a compile-time side NF, p / N is a multiply-high.
// x = hi + lo with hi on the 2^-12 grid (at most 13 significant bits for |x| <= 1, 21 for
// |x| < 2^9) and lo = x - hi rounded to float32
__device__ __forceinline__ void split12(double x, float* hi, float* lo) {
  const double h = rint(x * 4096.0) * (1.0 / 4096.0);
  *hi = (float)h;
  *lo = (float)(x - h);
}

// floor and fractional part of hi + lo, hi exact: hi - floor(hi) is exact, adding lo rounds once
__device__ __forceinline__ float floor_frac(float hi, float lo, int* fl) {
  float f0 = floorf(hi);
  float t = (hi - f0) + lo;
  if (t < 0.f) {
    t += 1.f;
    f0 -= 1.f;
  } else if (t >= 1.f) {
    t -= 1.f;
    f0 += 1.f;
  }
  *fl = (int)f0;
  return t;
}

template <int NMAX, int NF, int NT>
__device__ __forceinline__ void spline_sample_inplace(float* buf, int N, int S,
                                                      double angle_deg) {
  constexpr int PIX = (NMAX * NMAX + NT - 1) / NT;
  static_assert(NMAX < 256, "pixel indices must stay below 2^8 for the exact products");
  double sn, cs;
  sincospi(angle_deg / 180.0, &sn, &cs);
  const double h = 0.5 * (N - 1);
  // src_r = ar + cs r + sn c and src_c = ac - sn r + cs c, evaluated in float32 from split
  // coefficients: with r, c < 2^8 the hi products and their sums are exact, the lo terms sum
  // to less than 2^-4 with rounding below 2^-28, so the fractional parts carry the float64
  // coordinates' values to float32 rounding
  float cs_h, cs_l, sn_h, sn_l, ar_h, ar_l, ac_h, ac_l;
  split12(cs, &cs_h, &cs_l);
  split12(sn, &sn_h, &sn_l);
  split12(h - cs * h - sn * h, &ar_h, &ar_l);
  split12(h + sn * h - cs * h, &ac_h, &ac_l);
  const int NN = N * N;
  const float inv_n = 1.0f / N;
  float v[PIX];
#pragma unroll
  for (int k = 0; k < PIX; ++k) {
    const int p = threadIdx.x + k * NT;
    float acc = 0.f;
    if (p < NN) {
      const int r = qdiv<NF>(p, N, inv_n), c = p - __umul24(r, N);
      const float rf = (float)r, cf = (float)c;
      const float sr_h = fmaf(sn_h, cf, fmaf(cs_h, rf, ar_h));
      const float sr_l = fmaf(sn_l, cf, fmaf(cs_l, rf, ar_l));
      const float sc_h = fmaf(cs_h, cf, fmaf(-sn_h, rf, ac_h));
      const float sc_l = fmaf(cs_l, cf, fmaf(-sn_l, rf, ac_l));
      int ir, ic;
      float wr[4], wc[4];
      bspline_w6(floor_frac(sr_h, sr_l, &ir), wr);
      bspline_w6(floor_frac(sc_h, sc_l, &ic), wc);
      // |src - h| <= sqrt(2) h for |sin|, |cos| <= 1 (a NaN angle converts to 0), so the first
      // tap lies in [-N, 2N) and one wrap brings it into [0, N); the three after it come from
      // the padding
      ir -= 1;
      ic -= 1;
      ir += ir < 0 ? N : 0;
      ir -= ir >= N ? N : 0;
      ic += ic < 0 ? N : 0;
      ic -= ic >= N ? N : 0;
      const float* q = buf + __umul24(ir, S) + ic;
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const float* rp = q + u * S;
        acc += wr[u] * (wc[0] * rp[0] + wc[1] * rp[1] + wc[2] * rp[2] + wc[3] * rp[3]);
      }
      acc *= 1.f / 36.f;          // the two sets of weights were 6x the B-spline's
    }
    v[k] = acc;
    // one pixel at a time: without this the scheduler hoists every pixel's taps at once
    __builtin_amdgcn_sched_barrier(0);
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < PIX; ++k) {
    const int p = threadIdx.x + k * NT;
    if (p < NN) {
      const int r = qdiv<NF>(p, N, inv_n), c = p - __umul24(r, N);
      buf[__umul24(r, S) + c] = v[k];
    }
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
  const float zN = (float)pow(BSPLINE_Z, (double)N);
  bspline_prefilter(A, N, S, zN);
  float* o = out + img * N * N;
  if constexpr (NBUF == 2) {
    spline_sample_inplace<NMAX, 0, 512>(A, N, S, ang1[img]);
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
// sin and cos of a CTF phase x (float64, |x| < 2^40): q = rint(x / (pi/2)) and
// r = x - q pi/2 in float64 (|r| <= pi/4), then the Taylor series of sin r (to r^9) and cos r
// (to r^10) in float32, whose truncation errors at |r| = pi/4 are 1.7e-9 and 1.1e-10, and
// the quadrant q mod 4 picks the signs and the swap.
__device__ inline void sincos_phase(double x, float* s, float* c) {
  const double q = rint(x * 0.63661977236758134);
  const float r = (float)fma(-q, 1.5707963267948966, x);
  const float r2 = r * r;
  const float sr =
      r * fmaf(r2, fmaf(r2, fmaf(r2, fmaf(r2, 2.7557319e-6f, -1.9841270e-4f), 8.3333333e-3f),
                        -1.6666667e-1f), 1.f);
  const float cr = fmaf(
      r2, fmaf(r2, fmaf(r2, fmaf(r2, fmaf(r2, -2.7557319e-7f, 2.4801587e-5f), -1.3888889e-3f),
                        4.1666667e-2f), -0.5f), 1.f);
  const int k = (int)((long long)q & 3);
  *s = k == 0 ? sr : k == 1 ? cr : k == 2 ? -sr : -cr;
  *c = k == 0 ? cr : k == 1 ? -sr : k == 2 ? -cr : sr;
}

// Operand entries of one half-spectrum bin from its spectrum value X, CTF value C and the
// bin's half-spectrum weight w (1 or 2): {w |X|^2, C^2, sqrt(w) C Re X, sqrt(w) C Im X}.
__device__ inline void operand_bin(float2 X, float C, float w, float v[4]) {
  const float swc = (w == 1.f ? 1.f : 1.41421356f) * C;
  v[0] = w * fmaf(X.x, X.x, X.y * X.y);
  v[1] = C * C;
  v[2] = swc * X.x;
  v[3] = swc * X.y;
}

// The four segment entries of one operand row at element offset base (segment stride Khp):
// float32, or the bf16 pair hi = rne(v), lo = rne(v - hi) (v - hi is exact in float32).
__device__ inline void store_operand(int prec, void* op_hi, void* op_lo, int64_t base,
                                     int64_t Khp, const float v[4]) {
  if (prec == MEMI_PREC_FP32) {
    float* o = (float*)op_hi;
#pragma unroll
    for (int sg = 0; sg < 4; ++sg) o[base + sg * Khp] = v[sg];
  } else {
    uint16_t* oh = (uint16_t*)op_hi;
    uint16_t* ol = (uint16_t*)op_lo;
#pragma unroll
    for (int sg = 0; sg < 4; ++sg) {
      const uint16_t h = bf16_rne(v[sg]);
      oh[base + sg * Khp] = h;
      ol[base + sg * Khp] = bf16_rne(v[sg] - bf16_to_f(h));
    }
  }
}

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
  const float w = live ? half_weight(col, N) : 0.f;
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
    float v[4] = {0.f, 0.f, 0.f, 0.f};
    if (live) {
      float sf, cf;
      sincos_phase(a_k - c_k * dfv[u], &sf, &cf);
      const float C = (sf - (float)cp.ampc * cf) * (float)env;
      ctf_half[i * Kh + k] = C;
      operand_bin(Xv[u], C, w, v);
    }
    store_operand(prec, op_hi, op_lo, (row0 + i) * 4 * Khp + k, Khp, v);
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
// Tables over the half-spectrum bins k (row-major [ky][kx]), as arrays: a_k and c_k (float64,
// the CTF phase wr = a_k - c_k df of ctemh_cryoFrank.py:24-44), then the Butterworth factor
// / N^2 and the envelope (float32) -- 3 Kh doubles of the PT Kh the workspace holds.
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
    // structure of arrays, so that the kernels' loads of one field are contiguous
    tab[k] = 0.5 * (M_PI * cs * wav * wav * wav) * k2 * k2;
    tab[Kh + k] = M_PI * wav * k2;
    float* tf = (float*)(tab + 2 * Kh);
    tf[k] = (float)(f / ((double)N * N));
    tf[Kh + k] = (float)env;
  }
}

// NF > 0: a kernel for image side NF only (every size, count and divisor a constant);
// NF = 0: any side up to NMAX, given at run time.  NT threads per image: 256 where four
// images fit in one CU's LDS (sides up to 96: four blocks of four waves interleave their
// barriers and serial spline recursions), else 512.  The 256-thread and the 128 kernels are
// held to 128 registers: 4 waves per SIMD, i.e. four (two) images per CU, as their LDS allows.
constexpr int prep_threads(int nmax) { return nmax <= 96 ? 256 : 512; }
template <int NMAX, int NF, int NT = prep_threads(NMAX)>
__global__ __launch_bounds__(NT, (NMAX == 128 || NT == 256) ? 4 : 1) void k_prep_images(
    const float* __restrict__ stack, const int64_t* __restrict__ src,
    const int32_t* __restrict__ conj, const double* __restrict__ rot1,
    const double* __restrict__ rot2, const double* __restrict__ df, int N_arg,
    const double* __restrict__ tab, double ampc, int prec, int64_t Khp, int64_t row0,
    void* __restrict__ op_hi, void* __restrict__ op_lo, float2* __restrict__ fy,
    float* __restrict__ ctf_half, float* __restrict__ img_out) {
  // one image buffer: the padded real image ((N+3)^2 floats) is the largest of its layouts
  // (half spectrum 2 N (N/2+1) floats, row-pair lines 2 ceil(N/2) N floats)
  static_assert(NF == 0 || NF == NMAX, "a fixed-size kernel is sized for its own side");
  const int N = NF > 0 ? NF : N_arg;
  constexpr int BUF = (NMAX + 3) * (NMAX + 3);
  __shared__ float lds[BUF];
  __shared__ cpx tw[NMAX];
  __shared__ int pos[NMAX], rev[NMAX];              // frequency <-> position after the passes
  __shared__ double red[16];
  float* A = lds;
  const int S = N + 3;
  const int Nh = N / 2 + 1;
  const int64_t Kh = (int64_t)N * Nh;
  const float* tabf = (const float*)(tab + 2 * Kh);    // filter and envelope (k_prep_tables)
  const int64_t img = blockIdx.x;
  int rad[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  int nrad = 0;
  if constexpr (NF > 0) {
    fft_digit_tables_c<NF>(pos, rev);
  } else {
    nrad = radix_plan(N, rad);
    fft_digit_tables(N, rad, nrad, pos, rev);
  }
  fft_twiddles(tw, N);
  // a1: load (conjugates flipped, SPIDER transpose) and background statistics
  const float* raw = stack + src[img] * (int64_t)N * N;
  const bool cj = conj != nullptr && conj[img] != 0;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
  double s1 = 0.0, s2 = 0.0;
  // every load of this thread in flight at once, then the LDS stores and the statistics
  constexpr int PER = (NMAX * NMAX + NT - 1) / NT;
  const int NNi = N * N;
  float vv[PER];
#pragma unroll
  for (int u = 0; u < PER; ++u) {
    const int p = threadIdx.x + u * NT;
    vv[u] = p < NNi ? raw[p] : 0.f;
  }
  const float inv_n = 1.0f / N;
#pragma unroll
  for (int u = 0; u < PER; ++u) {
    const int p = threadIdx.x + u * NT;
    if (p < NNi) {
      const int a = qdiv<NF>(p, N, inv_n), b = p - a * N;
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
  const double inv_sd = 1.0 / sd;
  for (int r = wv; r < N; r += nw)
    for (int c = lane; c < N; c += 64)
      A[r * S + c] = (float)(((double)A[r * S + c] - mean) * inv_sd);
  __syncthreads();
  // a2: Butterworth in frequency (the table folds in the 1/N^2 of the inverse)
  cpx* H = rfft2_ip<NMAX, NF, NT>(A, N, S, tw, rad, nrad, pos);
  const float inv_nh = 1.0f / Nh;
  for (int q = threadIdx.x; q < Kh; q += blockDim.x) {   // H row p holds ky = rev[p]
    const int py = qdiv<NF>(q, Nh, inv_nh), kx = q - py * Nh;
    const float f = tabf[rev[py] * Nh + kx];
    H[q] = make_float2(H[q].x * f, H[q].y * f);
  }
  __syncthreads();
  irfft2_ip<NMAX, NF, NT>(A, N, S, tw, rad, nrad, pos);
  // a4: disk mask, then the two spline rotations, each in place
  for (int r = wv; r < N; r += nw)
    for (int c = lane; c < N; c += 64) {
      const int xd = 2 * r - N + 2, yd = 2 * c - N;
      if (xd * xd + yd * yd >= N * N) A[r * S + c] = 0.f;
    }
  __syncthreads();
  const float zN = (float)pow(BSPLINE_Z, (double)N);
  bspline_prefilter(A, N, S, zN);
  // the rotations run with the side as a run-time value: with every pixel's coordinates
  // constant-folded the compiler computes them all ahead and spills
  int n_rt = N;
  asm volatile("" : "+s"(n_rt));
  spline_sample_inplace<NMAX, 0, NT>(A, n_rt, n_rt + 3, rot1[img]);
  bspline_prefilter(A, N, S, zN);
  spline_sample_inplace<NMAX, 0, NT>(A, n_rt, n_rt + 3, rot2[img]);
  if (img_out) {
    float* o = img_out + img * (int64_t)N * N;
    for (int r = wv; r < N; r += nw)
      for (int c = lane; c < N; c += 64) o[r * N + c] = A[r * S + c];
  }
  // a5/a6: spectrum of the aligned image, CTF, operand rows (as k_ctf_pack)
  cpx* F = rfft2_ip<NMAX, NF, NT>(A, N, S, tw, rad, nrad, pos);
  const double dfi = df[img];
  const float ampf = (float)ampc;
  const int64_t base = (row0 + img) * 4 * Khp;
  for (int q = threadIdx.x; q < Khp; q += blockDim.x) {
    float v[4] = {0.f, 0.f, 0.f, 0.f};
    if (q < Kh) {
      float sf, cf;
      sincos_phase(tab[q] - tab[Kh + q] * dfi, &sf, &cf);
      const float C = (sf - ampf * cf) * tabf[Kh + q];
      const int ky = qdiv<NF>(q, Nh, inv_nh), kx = q - ky * Nh;
      const cpx X = F[pos[ky] * Nh + kx];
      ctf_half[img * Kh + q] = C;
      fy[img * Kh + q] = X;
      operand_bin(X, C, half_weight(kx, N), v);
    }
    store_operand(prec, op_hi, op_lo, base + q, Khp, v);
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
  const dim3 g((unsigned)n);
#define MEMI_PREP(NM, NF)                                                                      \
  hipLaunchKernelGGL((k_prep_images<NM, NF>), g, dim3(prep_threads(NM)), 0, s, stack, src, conj, \
                     rot1_deg, rot2_deg, df, N, ws, ctf->ampc, precision, Khp, row0, op_hi,     \
                     op_lo, (float2*)fy, ctf_half, img_out)
  // the power-of-two-rich sides of the standard workloads get their own kernels (a fixed
  // 70 = 2 x 5 x 7 kernel needs more registers than the generic one); any other side runs on
  // the kernel of the next larger size class
  switch (N) {
    case 64: MEMI_PREP(64, 64); break;
    case 96: MEMI_PREP(96, 96); break;
    case 128: MEMI_PREP(128, 128); break;
    case 160: MEMI_PREP(160, 160); break;
    default:
      if (N <= 64) MEMI_PREP(64, 0);
      else if (N <= 96) MEMI_PREP(96, 0);
      else if (N <= 128) MEMI_PREP(128, 0);
      else MEMI_PREP(160, 0);
  }
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
