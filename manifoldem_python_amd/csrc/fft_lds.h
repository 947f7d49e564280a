// Mixed-radix Stockham FFTs on images resident in LDS (used by the fused image preparation in
// prep.hip).  Radices 8, 4, 2, 3, 5, 7; real rows are transformed two at a time as one complex
// line (x_a + i x_b), columns as complex lines of the half spectrum.  Twiddles come from a table
// of the N-th roots of unity computed in float64 (fft_twiddles).
#pragma once
#include "common.h"

namespace memi {

typedef float2 cpx;

__device__ inline cpx cmul(cpx a, cpx b) {
  return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ inline cpx cadd(cpx a, cpx b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ inline cpx csub(cpx a, cpx b) { return make_float2(a.x - b.x, a.y - b.y); }
// multiply by +i (inv) or -i (forward)
__device__ inline cpx rot90(cpx a, bool inv) {
  return inv ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
}

// in-register DFTs, sign -1 (forward) or +1 (inv)
template <int R>
struct Dft;
template <>
struct Dft<2> {
  __device__ static void run(cpx* v, bool) {
    const cpx a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  }
};
template <>
struct Dft<4> {
  __device__ static void run(cpx* v, bool inv) {
    const cpx s02 = cadd(v[0], v[2]), d02 = csub(v[0], v[2]);
    const cpx s13 = cadd(v[1], v[3]), d13 = rot90(csub(v[1], v[3]), inv);
    v[0] = cadd(s02, s13);
    v[2] = csub(s02, s13);
    v[1] = cadd(d02, d13);
    v[3] = csub(d02, d13);
  }
};
template <>
struct Dft<8> {
  __device__ static void run(cpx* v, bool inv) {
    cpx e[4] = {v[0], v[2], v[4], v[6]}, o[4] = {v[1], v[3], v[5], v[7]};
    Dft<4>::run(e, inv);
    Dft<4>::run(o, inv);
    const float h = 0.70710678118654752f;
    // o[k] *= w8^k, w8 = exp(-+ i pi / 4)
    o[1] = inv ? make_float2(h * (o[1].x - o[1].y), h * (o[1].x + o[1].y))
               : make_float2(h * (o[1].x + o[1].y), h * (o[1].y - o[1].x));
    o[2] = rot90(o[2], inv);
    o[3] = inv ? make_float2(-h * (o[3].x + o[3].y), h * (o[3].x - o[3].y))
               : make_float2(h * (o[3].y - o[3].x), -h * (o[3].x + o[3].y));
    for (int k = 0; k < 4; ++k) {
      v[k] = cadd(e[k], o[k]);
      v[k + 4] = csub(e[k], o[k]);
    }
  }
};
// odd radices by the direct sum: cos / sin (2 pi t / R), t = 0 .. R-1
template <int R>
struct OddRoots;
template <>
struct OddRoots<3> {
  static constexpr float c[3] = {1.f, -0.5f, -0.5f};
  static constexpr float s[3] = {0.f, 0.86602540378443865f, -0.86602540378443865f};
};
template <>
struct OddRoots<5> {
  static constexpr float c[5] = {1.f, 0.30901699437494742f, -0.80901699437494742f,
                                 -0.80901699437494742f, 0.30901699437494742f};
  static constexpr float s[5] = {0.f, 0.95105651629515357f, 0.58778525229247313f,
                                 -0.58778525229247313f, -0.95105651629515357f};
};
template <>
struct OddRoots<7> {
  static constexpr float c[7] = {1.f, 0.62348980185873353f, -0.22252093395631440f,
                                 -0.90096886790241913f, -0.90096886790241913f,
                                 -0.22252093395631440f, 0.62348980185873353f};
  static constexpr float s[7] = {0.f, 0.78183148246802981f, 0.97492791218182361f,
                                 0.43388373911755812f, -0.43388373911755812f,
                                 -0.97492791218182361f, -0.78183148246802981f};
};
template <int R>
struct DftOdd {
  __device__ static void run(cpx* v, bool inv) {
    cpx y[R];
#pragma unroll
    for (int s = 0; s < R; ++s) {
      float re = 0.f, im = 0.f;
#pragma unroll
      for (int q = 0; q < R; ++q) {
        const int t = (q * s) % R;
        const float c = OddRoots<R>::c[t], d = inv ? OddRoots<R>::s[t] : -OddRoots<R>::s[t];
        re += v[q].x * c - v[q].y * d;
        im += v[q].x * d + v[q].y * c;
      }
      y[s] = make_float2(re, im);
    }
#pragma unroll
    for (int s = 0; s < R; ++s) v[s] = y[s];
  }
};
template <>
struct Dft<3> : DftOdd<3> {};
template <>
struct Dft<5> : DftOdd<5> {};
template <>
struct Dft<7> : DftOdd<7> {};

// q / d for 0 <= q < 2^22, d >= 1, from a float reciprocal and one correction each way
__device__ inline int fdivi(int q, int d, float inv_d) {
  int t = (int)((float)q * inv_d);
  t += (t + 1) * d <= q;
  t -= t * d > q;
  return t;
}

// One Stockham pass of radix R over L lines of length N (line l at base + l*ls, element i at
// + i*es) from src to dst; Ns = product of the radices of the earlier passes.
// y[(j/Ns) Ns R + j%Ns + r Ns] = DFT_R( x[j + q N/R] w_N^{q (j%Ns) N/(Ns R)} )_r
template <int R>
__device__ inline void stockham_pass(const cpx* src, cpx* dst, int N, int L, int ls, int es,
                                     int Ns, const cpx* tw, bool inv) {
  const int m = N / R;
  const int total = L * m;
  const int tstep = N / (Ns * R);
  const float inv_m = 1.0f / m, inv_ns = 1.0f / Ns;
  for (int q = threadIdx.x; q < total; q += blockDim.x) {
    const int line = fdivi(q, m, inv_m), j = q - line * m;
    const cpx* s = src + line * ls;
    cpx* d = dst + line * ls;
    const int jg = fdivi(j, Ns, inv_ns);
    const int k = j - jg * Ns;
    cpx v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) v[r] = s[(j + r * m) * es];
#pragma unroll
    for (int r = 1; r < R; ++r) {
      cpx w = tw[r * k * tstep];
      if (inv) w.y = -w.y;
      v[r] = cmul(v[r], w);
    }
    Dft<R>::run(v, inv);
    const int idx = jg * Ns * R + k;
#pragma unroll
    for (int r = 0; r < R; ++r) d[(idx + r * Ns) * es] = v[r];
  }
}

// radix plan of N: up to 8 passes from {8, 4, 2, 3, 5, 7}; returns the count (0: unsupported)
__host__ __device__ inline int radix_plan(int N, int* rad) {
  int n = N, c = 0;
  const int order[6] = {8, 4, 2, 3, 5, 7};
  for (int o = 0; o < 6; ++o)
    while (n % order[o] == 0 && c < 8) {
      rad[c++] = order[o];
      n /= order[o];
    }
  return n == 1 ? c : 0;
}

// the passes of one batch of line transforms; returns the buffer holding the result
__device__ inline cpx* fft_lines(cpx* a, cpx* b, int N, int L, int ls, int es, const cpx* tw,
                                 const int* rad, int nrad, bool inv) {
  int Ns = 1;
  for (int p = 0; p < nrad; ++p) {
    switch (rad[p]) {
      case 8: stockham_pass<8>(a, b, N, L, ls, es, Ns, tw, inv); break;
      case 4: stockham_pass<4>(a, b, N, L, ls, es, Ns, tw, inv); break;
      case 2: stockham_pass<2>(a, b, N, L, ls, es, Ns, tw, inv); break;
      case 3: stockham_pass<3>(a, b, N, L, ls, es, Ns, tw, inv); break;
      case 5: stockham_pass<5>(a, b, N, L, ls, es, Ns, tw, inv); break;
      default: stockham_pass<7>(a, b, N, L, ls, es, Ns, tw, inv); break;
    }
    __syncthreads();
    Ns *= rad[p];
    cpx* t = a;
    a = b;
    b = t;
  }
  return a;
}

// 2-D real-to-half-spectrum FFT of the real image in `re` (row stride S floats).  `other` is
// a second buffer of the same size.  The half spectrum [N][Nh] (complex, row stride Nh) is
// returned; it lives in `re` or `other`.
__device__ inline cpx* rfft2_lds(float* re, float* other, int N, int S, const cpx* tw,
                                 const int* rad, int nrad) {
  const int Nh = N / 2 + 1, P = (N + 1) / 2;       // row pairs
  cpx* X = (cpx*)other;
  const float inv_n = 1.0f / N, inv_nh = 1.0f / Nh;
  for (int q = threadIdx.x; q < P * N; q += blockDim.x) {
    const int l = fdivi(q, N, inv_n), c = q - l * N;
    const float a = re[(2 * l) * S + c];
    const float b = 2 * l + 1 < N ? re[(2 * l + 1) * S + c] : 0.f;
    X[l * N + c] = make_float2(a, b);
  }
  __syncthreads();
  cpx* Z = fft_lines(X, (cpx*)re, N, P, N, 1, tw, rad, nrad, false);
  cpx* H = Z == X ? (cpx*)re : X;
  // unpack: X_a[k] = (Z[k] + conj Z[-k]) / 2, X_b[k] = (Z[k] - conj Z[-k]) / (2i)
  for (int q = threadIdx.x; q < P * Nh; q += blockDim.x) {
    const int l = fdivi(q, Nh, inv_nh), k = q - l * Nh;
    const cpx z = Z[l * N + k], zc = Z[l * N + (k == 0 ? 0 : N - k)];
    H[(2 * l) * Nh + k] = make_float2(0.5f * (z.x + zc.x), 0.5f * (z.y - zc.y));
    if (2 * l + 1 < N)
      H[(2 * l + 1) * Nh + k] = make_float2(0.5f * (z.y + zc.y), 0.5f * (zc.x - z.x));
  }
  __syncthreads();
  return fft_lines(H, Z, N, Nh, 1, Nh, tw, rad, nrad, false);
}

// 2-D inverse (unnormalised) of the half spectrum H (row stride Nh) into the real image
// `dst_re` (row stride S); H and `spare` are the two buffers, dst_re is one of them.
__device__ inline void irfft2_lds(cpx* H, cpx* spare, float* dst_re, int N, int S,
                                  const cpx* tw, const int* rad, int nrad) {
  const int Nh = N / 2 + 1, P = (N + 1) / 2;
  const float inv_n = 1.0f / N;
  cpx* G = fft_lines(H, spare, N, Nh, 1, Nh, tw, rad, nrad, true);
  cpx* Zb = G == H ? spare : H;
  // row pairs: Z[k] = G_a[k] + i G_b[k] with G_r[k] = conj G_r[N-k] for k > N/2
  for (int q = threadIdx.x; q < P * N; q += blockDim.x) {
    const int l = fdivi(q, N, inv_n), k = q - l * N;
    const bool lowk = k < Nh;
    const int kk = lowk ? k : N - k;
    cpx ga = G[(2 * l) * Nh + kk];
    cpx gb = 2 * l + 1 < N ? G[(2 * l + 1) * Nh + kk] : make_float2(0.f, 0.f);
    if (!lowk) {
      ga.y = -ga.y;
      gb.y = -gb.y;
    }
    Zb[l * N + k] = make_float2(ga.x - gb.y, ga.y + gb.x);
  }
  __syncthreads();
  cpx* Z = fft_lines(Zb, G, N, P, N, 1, tw, rad, nrad, true);
  // Z is in Zb or G; write the real rows into dst_re, which must not be Z's buffer
  float* tmp = (float*)(Z == Zb ? G : Zb);
  for (int q = threadIdx.x; q < P * N; q += blockDim.x) {
    const int l = fdivi(q, N, inv_n), c = q - l * N;
    const cpx z = Z[l * N + c];
    tmp[(2 * l) * S + c] = z.x;
    if (2 * l + 1 < N) tmp[(2 * l + 1) * S + c] = z.y;
  }
  __syncthreads();
  if (tmp != dst_re) {
    for (int q = threadIdx.x; q < N * S; q += blockDim.x) dst_re[q] = tmp[q];
    __syncthreads();
  }
}

// tw[t] = exp(-2 pi i t / N), t < N, rounded from float64
__device__ inline void fft_twiddles(cpx* tw, int N) {
  for (int t = threadIdx.x; t < N; t += blockDim.x) {
    double sn, cs;
    sincospi(2.0 * t / N, &sn, &cs);
    tw[t] = make_float2((float)cs, -(float)sn);
  }
}

}  // namespace memi
