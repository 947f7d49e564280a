// Mixed-radix FFTs on images resident in LDS (used by the fused image preparation in prep.hip),
// in place in one buffer per image.  Radices 8, 4, 2, 3, 5, 7; real rows are transformed two at
// a time as one complex line (x_a + i x_b), columns as complex lines of the half spectrum.
// Twiddles come from a table of the N-th roots of unity computed in float64 (fft_twiddles).
// Blocks of NT threads (a template parameter of the 2-D transforms).
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
  t += (int)__umul24(t + 1, d) <= q;   // 24-bit products: full-rate multiplies
  t -= (int)__umul24(t, d) > q;
  return t;
}

// q / d for the compile-time image side NF > 0 (d then folds to a constant after inlining,
// and the division becomes a multiply-high) or by fdivi for a run-time side (NF = 0)
template <int NF>
__device__ inline int qdiv(int q, int d, float inv_d) {
  if constexpr (NF > 0)
    return q / d;
  else
    return fdivi(q, d, inv_d);
}

// Every transform below works in ONE LDS buffer.  The line transforms are in-place
// mixed-radix FFTs: decimation in frequency forward (each butterfly writes its outputs over its
// inputs, twiddles after the butterfly), which leaves the frequencies in digit-reversed order,
// and decimation in time for the inverse, which takes that order back to natural order -- so a
// forward / inverse pair needs no reordering pass at all.  pos[k] (natural k -> position) and
// rev[p] (position -> k) translate where a consumer needs natural frequencies.  Only the
// packing of two real rows into one complex line and its separation stage values in registers
// (all reads, a barrier, all writes).


// items per thread of a step over `total` work items, for the largest image of the template
__host__ __device__ constexpr int fft_items(int total, int nt) { return (total + nt - 1) / nt; }

// the same plan at compile time (fixed-size kernels)
struct RadixPlan {
  int n;
  int r[8];
};
constexpr RadixPlan radix_plan_c(int N) {
  RadixPlan P{0, {0, 0, 0, 0, 0, 0, 0, 0}};
  int n = N;
  const int order[6] = {8, 4, 2, 3, 5, 7};
  for (int o = 0; o < 6; ++o)
    while (n % order[o] == 0 && P.n < 8) {
      P.r[P.n++] = order[o];
      n /= order[o];
    }
  if (n != 1) P.n = 0;
  return P;
}
// sub-transform length at pass I of N's plan: N / (r_0 ... r_{I-1})
constexpr int plan_span(int N, int I) {
  const RadixPlan P = radix_plan_c(N);
  int s = N;
  for (int p = 0; p < I; ++p) s /= P.r[p];
  return s;
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

// position of natural frequency k after the forward passes: with k = d1 + R1 d2 + R1 R2 d3 + ..,
// p = d1 N/R1 + d2 N/(R1 R2) + ...; fills pos[] and rev[] (LDS, N entries each)
__device__ inline void fft_digit_tables(int N, const int* rad, int nrad, int* pos, int* rev) {
  for (int k = threadIdx.x; k < N; k += blockDim.x) {
    int rem = k, span = N, p = 0;
    for (int i = 0; i < nrad; ++i) {
      span /= rad[i];
      p += (rem % rad[i]) * span;
      rem /= rad[i];
    }
    pos[k] = p;
    rev[p] = k;
  }
}

// fft_digit_tables for a compile-time size (no run-time plan array)
template <int NF>
__device__ __forceinline__ void fft_digit_tables_c(int* pos, int* rev) {
  constexpr RadixPlan P = radix_plan_c(NF);
  for (int k = threadIdx.x; k < NF; k += blockDim.x) {
    int rem = k, span = NF, p = 0;
#pragma unroll
    for (int i = 0; i < P.n; ++i) {
      span /= P.r[i];
      p += (rem % P.r[i]) * span;
      rem /= P.r[i];
    }
    pos[k] = p;
    rev[p] = k;
  }
}

// one in-place pass of radix R over L lines of length N (line l at l*ls, element i at i*es):
// sub-transforms of length `span` at multiples of span; butterfly (block, n1) takes
// idx_r = block*span + n1 + r*span/R.  Forward (DIF): DFT_R, then y_s *= w_span^(n1 s).
// Inverse (DIT): x_s *= conj w_span^(n1 s), then the conjugate DFT_R.
// col: lines side by side (ls == 1): consecutive threads take consecutive lines.
template <int R, int NF>
__device__ __forceinline__ void fft_pass_ip(cpx* buf, int N, int L, int ls, int es, int span,
                                   const cpx* tw, bool inv) {
  const int sub = span / R, per = N / R;
  const int total = L * per;
  const int tstep = N / span;
  const float inv_sub = 1.0f / sub, inv_per = 1.0f / per, inv_l = 1.0f / L;
  const bool col = ls == 1;
  for (int q = threadIdx.x; q < total; q += blockDim.x) {
    int line, t;
    if (col) {
      t = qdiv<NF>(q, L, inv_l);
      line = q - t * L;
    } else {
      line = qdiv<NF>(q, per, inv_per);
      t = q - line * per;
    }
    const int blk = qdiv<NF>(t, sub, inv_sub), n1 = t - blk * sub;
    cpx* x = buf + line * ls + (blk * span + n1) * es;
    const int st = sub * es;
    cpx v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) v[r] = x[r * st];
    if (!inv) {
      Dft<R>::run(v, false);
#pragma unroll
      for (int r = 1; r < R; ++r) v[r] = cmul(v[r], tw[n1 * r * tstep]);
    } else {
#pragma unroll
      for (int r = 1; r < R; ++r) {
        cpx w = tw[n1 * r * tstep];
        w.y = -w.y;
        v[r] = cmul(v[r], w);
      }
      Dft<R>::run(v, true);
    }
#pragma unroll
    for (int r = 0; r < R; ++r) x[r * st] = v[r];
  }
  __syncthreads();
}

// all passes of one batch of line transforms, in place: forward in plan order (spans N,
// N/R1, ...), inverse in reverse order
__device__ __forceinline__ void fft_lines_ip(cpx* buf, int N, int L, int ls, int es, const cpx* tw,
                                    const int* rad, int nrad, bool inv) {
  int spans[8];
  int span = N;
  for (int p = 0; p < nrad; ++p) {
    spans[p] = span;
    span /= rad[p];
  }
  for (int i = 0; i < nrad; ++i) {
    const int p = inv ? nrad - 1 - i : i;
    switch (rad[p]) {
      case 8: fft_pass_ip<8, 0>(buf, N, L, ls, es, spans[p], tw, inv); break;
      case 4: fft_pass_ip<4, 0>(buf, N, L, ls, es, spans[p], tw, inv); break;
      case 2: fft_pass_ip<2, 0>(buf, N, L, ls, es, spans[p], tw, inv); break;
      case 3: fft_pass_ip<3, 0>(buf, N, L, ls, es, spans[p], tw, inv); break;
      case 5: fft_pass_ip<5, 0>(buf, N, L, ls, es, spans[p], tw, inv); break;
      default: fft_pass_ip<7, 0>(buf, N, L, ls, es, spans[p], tw, inv); break;
    }
  }
}

// the passes of a compile-time size NF, each with its radix and span as constants
template <int NF, int I, bool INV>
__device__ __forceinline__ void fixed_passes(cpx* buf, int L, int ls, int es, const cpx* tw) {
  constexpr RadixPlan P = radix_plan_c(NF);
  if constexpr (I < P.n) {
    constexpr int p = INV ? P.n - 1 - I : I;
    constexpr int span = plan_span(NF, p);
    fft_pass_ip<P.r[p], NF>(buf, NF, L, ls, es, span, tw, INV);
    fixed_passes<NF, I + 1, INV>(buf, L, ls, es, tw);
  }
}

template <int NF>
__device__ __forceinline__ void fft_lines(cpx* buf, int N, int L, int ls, int es, const cpx* tw,
                                 const int* rad, int nrad, bool inv) {
  if constexpr (NF > 0) {
    static_assert(radix_plan_c(NF).n > 0, "image side without a radix plan");
    if (inv)
      fixed_passes<NF, 0, true>(buf, L, ls, es, tw);
    else
      fixed_passes<NF, 0, false>(buf, L, ls, es, tw);
  } else {
    fft_lines_ip(buf, N, L, ls, es, tw, rad, nrad, inv);
  }
}

// 2-D real-to-half-spectrum FFT in place: the real image (row stride S floats) in `buf`
// becomes the half spectrum, kx natural along each row (row stride Nh), rows at ky's
// digit-reversed position: spectrum (ky, kx) is at [pos[ky] * Nh + kx].
template <int NMAX, int NF, int NT>
__device__ inline cpx* rfft2_ip(float* buf, int N, int S, const cpx* tw, const int* rad,
                                int nrad, const int* pos) {
  constexpr int PK = fft_items(((NMAX + 1) / 2) * NMAX, NT);
  constexpr int UK = fft_items(((NMAX + 1) / 2) * (NMAX / 2 + 1), NT);
  const int Nh = N / 2 + 1, P = (N + 1) / 2;       // row pairs
  cpx* X = (cpx*)buf;
  const float inv_n = 1.0f / N, inv_nh = 1.0f / Nh;
  {
    cpx st[PK];
#pragma unroll
    for (int b = 0; b < PK; ++b) {
      const int q = threadIdx.x + b * NT;
      if (q < P * N) {
        const int l = qdiv<NF>(q, N, inv_n), c = q - l * N;
        st[b] = make_float2(buf[(2 * l) * S + c], 2 * l + 1 < N ? buf[(2 * l + 1) * S + c] : 0.f);
      }
    }
    __syncthreads();
#pragma unroll
    for (int b = 0; b < PK; ++b) {
      const int q = threadIdx.x + b * NT;
      if (q < P * N) X[q] = st[b];                 // line l at l * N
    }
    __syncthreads();
  }
  fft_lines<NF>(X, N, P, N, 1, tw, rad, nrad, false);
  // separate the row pairs: X_a[k] = (Z[k] + conj Z[-k]) / 2, X_b[k] = (Z[k] - conj Z[-k]) / 2i,
  // Z[k] read at its digit-reversed position
  {
    cpx ha[UK], hb[UK];
#pragma unroll
    for (int b = 0; b < UK; ++b) {
      const int q = threadIdx.x + b * NT;
      if (q < P * Nh) {
        const int l = qdiv<NF>(q, Nh, inv_nh), k = q - l * Nh;
        const cpx z = X[l * N + pos[k]], zc = X[l * N + pos[k == 0 ? 0 : N - k]];
        ha[b] = make_float2(0.5f * (z.x + zc.x), 0.5f * (z.y - zc.y));
        hb[b] = make_float2(0.5f * (z.y + zc.y), 0.5f * (zc.x - z.x));
      }
    }
    __syncthreads();
#pragma unroll
    for (int b = 0; b < UK; ++b) {
      const int q = threadIdx.x + b * NT;
      if (q < P * Nh) {
        const int l = qdiv<NF>(q, Nh, inv_nh), k = q - l * Nh;
        X[(2 * l) * Nh + k] = ha[b];
        if (2 * l + 1 < N) X[(2 * l + 1) * Nh + k] = hb[b];
      }
    }
    __syncthreads();
  }
  fft_lines<NF>(X, N, Nh, 1, Nh, tw, rad, nrad, false);
  return X;
}

// 2-D inverse (unnormalised) of a half spectrum laid out as rfft2_ip leaves it, into the real
// image (row stride S) in the same buffer.
template <int NMAX, int NF, int NT>
__device__ inline void irfft2_ip(float* buf, int N, int S, const cpx* tw, const int* rad,
                                 int nrad, const int* pos) {
  constexpr int PK = fft_items(((NMAX + 1) / 2) * NMAX, NT);
  const int Nh = N / 2 + 1, P = (N + 1) / 2;
  cpx* G = (cpx*)buf;
  const float inv_n = 1.0f / N;
  fft_lines<NF>(G, N, Nh, 1, Nh, tw, rad, nrad, true);     // ky digit-reversed -> y natural
  // row pairs Z[k] = G_a[k] + i G_b[k] (G_r[k] = conj G_r[N-k] for k > N/2), stored at the
  // digit-reversed position the row inverse expects
  {
    cpx st[PK];
#pragma unroll
    for (int b = 0; b < PK; ++b) {
      const int q = threadIdx.x + b * NT;
      if (q < P * N) {
        const int l = qdiv<NF>(q, N, inv_n), k = q - l * N;
        const bool lowk = k < Nh;
        const int kk = lowk ? k : N - k;
        cpx ga = G[(2 * l) * Nh + kk];
        cpx gb = 2 * l + 1 < N ? G[(2 * l + 1) * Nh + kk] : make_float2(0.f, 0.f);
        if (!lowk) {
          ga.y = -ga.y;
          gb.y = -gb.y;
        }
        st[b] = make_float2(ga.x - gb.y, ga.y + gb.x);
      }
    }
    __syncthreads();
#pragma unroll
    for (int b = 0; b < PK; ++b) {
      const int q = threadIdx.x + b * NT;
      if (q < P * N) {
        const int l = qdiv<NF>(q, N, inv_n), k = q - l * N;
        G[l * N + pos[k]] = st[b];
      }
    }
    __syncthreads();
  }
  fft_lines<NF>(G, N, P, N, 1, tw, rad, nrad, true);
  // the real rows, row stride S
  {
    cpx st[PK];
#pragma unroll
    for (int b = 0; b < PK; ++b) {
      const int q = threadIdx.x + b * NT;
      if (q < P * N) st[b] = G[q];
    }
    __syncthreads();
#pragma unroll
    for (int b = 0; b < PK; ++b) {
      const int q = threadIdx.x + b * NT;
      if (q < P * N) {
        const int l = qdiv<NF>(q, N, inv_n), c = q - l * N;
        buf[(2 * l) * S + c] = st[b].x;
        if (2 * l + 1 < N) buf[(2 * l + 1) * S + c] = st[b].y;
      }
    }
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
