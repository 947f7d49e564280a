// Lanczos eigensolver steps for the normalised diffusion kernels of a PD batch (float64).
//
// Reference: sembeddingonFly.py:17-39 calls ARPACK eigsh(l, k=nEigs+1, which='LM'), an
// implicitly restarted Lanczos method.  Here every PD of a batch runs Lanczos with full
// reorthogonalisation in lockstep: one step is a grouped matvec w = L v_j over all PDs
// (k_lanczos_matvec, many workgroups per PD; it also leaves each row's w_i v_j[i]), then two
// orthogonalisation passes: the three-term recurrence w' = w - alpha v_j - beta_{j-1} v_{j-1}
// (alpha = v_j . w) removes the components Lanczos puts there, and one classical Gram-Schmidt
// pass against the whole basis removes what rounding left (components of order eps), so w' is
// orthogonalised twice as in CGS2, in one pass over the basis instead of two.  The step
// records alpha_j, beta_j and appends v_{j+1} = w'' / beta_j.  One workgroup per PD does both
// passes, or, for batches of few large PDs, two grid-wide kernels after the matvec.  The host diagonalises the small tridiagonal T; the Ritz
// vectors V S of every PD come from one batched kernel (k_ritz_vectors).
//
// Layout: L float32 or float64, PD p's n*n block at plan d_off; V float64 column-major per PD, PD p at
// element p_vo = m * row0(p) (row0 = packed row offset), column c of length n at p_vo + c*n;
// W float64 [n_rows]; ab float64 [npd][m][2] = {alpha_j, beta_j}; active int32 [npd].
#include "common.h"

namespace memi {

// Every block of the last kernel of a step arrives here; the last to arrive advances the step
// index jp[0] and re-arms the arrival counter jp[1] (all blocks read jp[0] before arriving).
__device__ inline void step_arrive(int32_t* jp, int j) {
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    const int prev = atomicAdd(&jp[1], 1);
    if (prev == (int)(gridDim.x * gridDim.y) - 1) {
      jp[1] = 0;
      jp[0] = j + 1;
    }
  }
}

// w = L v_j: each wave takes `rpw` consecutive rows of the plan's row list (the row index is
// uniform over the wave, so the row pair, the PD descriptor and the active flag are scalar
// loads, and one descriptor serves the run of rows of one PD).  Four rows of one PD go
// together (16 loads of L per lane in flight, each load of v_j shared by the four), others one
// at a time with eight loads in flight; each row is summed by DPP.  rpw is 1 up to 8192 rows
// (a wave per row fills the GPU) and 4 beyond: a wave per row spent most of its time on the
// descriptor loads and the reduction of a short row, and long runs of rows left the waves
// holding the largest PDs' rows as a tail.
template <typename T>
__device__ __forceinline__ void matvec_row(const T* __restrict__ Lr, const double* __restrict__ v,
                                           int n, int lane, double* __restrict__ out,
                                           double* __restrict__ aout, double vi) {
  double s[8];
#pragma unroll
  for (int u = 0; u < 8; ++u) s[u] = 0.0;
  int c = lane;
  for (; c + 7 * 64 < n; c += 8 * 64) {
#pragma unroll
    for (int u = 0; u < 8; ++u) s[u] += (double)Lr[c + 64 * u] * v[c + 64 * u];
  }
#pragma unroll
  for (int u = 0; u < 7; ++u)
    if (c + 64 * u < n) s[u] += (double)Lr[c + 64 * u] * v[c + 64 * u];
  const double t = wave_total(((s[0] + s[1]) + (s[2] + s[3])) + ((s[4] + s[5]) + (s[6] + s[7])));
  if (lane == 0) {
    *out = t;
    if (aout) *aout = t * vi;
  }
}

template <typename T>
__device__ __forceinline__ void matvec_rows4(const T* __restrict__ L0, int64_t stride,
                                             const double* __restrict__ v, int n, int lane,
                                             double* __restrict__ out, double* __restrict__ aout,
                                             const double* __restrict__ vrow) {
  double s[4][4];
#pragma unroll
  for (int r = 0; r < 4; ++r)
#pragma unroll
    for (int u = 0; u < 4; ++u) s[r][u] = 0.0;
  int c = lane;
  for (; c + 3 * 64 < n; c += 4 * 64) {
    double x[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) x[u] = v[c + 64 * u];
#pragma unroll
    for (int r = 0; r < 4; ++r)
#pragma unroll
      for (int u = 0; u < 4; ++u) s[r][u] += (double)L0[r * stride + c + 64 * u] * x[u];
  }
#pragma unroll
  for (int u = 0; u < 3; ++u)
    if (c + 64 * u < n) {
      const double x = v[c + 64 * u];
#pragma unroll
      for (int r = 0; r < 4; ++r) s[r][u] += (double)L0[r * stride + c + 64 * u] * x;
    }
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const double t = wave_total((s[r][0] + s[r][1]) + (s[r][2] + s[r][3]));
    if (lane == 0) {
      out[r] = t;
      if (aout) aout[r] = t * vrow[r];
    }
  }
}

template <typename T>
__global__ __launch_bounds__(256) void k_lanczos_matvec(const int32_t* __restrict__ rows,
                                                        int64_t n_rows,
                                                        const memi_pd_desc* __restrict__ pds,
                                                        const T* __restrict__ L,
                                                        const double* __restrict__ V, int m,
                                                        const int32_t* __restrict__ jp,
                                                        const int32_t* __restrict__ active,
                                                        double* __restrict__ W,
                                                        double* __restrict__ A, int rpw,
                                                        const int32_t* __restrict__ skip) {
  const int j = *jp;
  const int lane = threadIdx.x & 63;
  const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int64_t g0 = ((int64_t)blockIdx.x * 4 + wid) * rpw;
  const int64_t g1 = min(g0 + (int64_t)rpw, n_rows);
  for (int64_t g = g0; g < g1;) {                     // uniform over the wave
    const int pd = rows[2 * g], i = rows[2 * g + 1];
    const memi_pd_desc P = pds[pd];
    const int n = P.n;
    const bool live = active[pd] && j < min(n, m) && !(skip && skip[pd]);
    // rows g..g+3 are rows i..i+3 of the same PD when the run and the PD both hold them
    const double* vj = V + (int64_t)m * P.row0 + (int64_t)j * n;
    if (g + 3 < g1 && i + 3 < n) {
      if (live)
        matvec_rows4<T>(L + P.d_off + (int64_t)i * n, n, vj, n, lane, W + P.row0 + i,
                        A ? A + P.row0 + i : nullptr, vj + i);
      g += 4;
    } else {
      if (live)
        matvec_row<T>(L + P.d_off + (int64_t)i * n, vj, n, lane, W + P.row0 + i,
                      A ? A + P.row0 + i : nullptr, live ? vj[i] : 0.0);
      g += 1;
    }
  }
}

// ---- sparse form of the kernel (sembeddingonFly.py passes eigsh a scipy sparse matrix) ----
// Row i of PD p keeps its nonzero entries at the dense row's offset d_off + i*n: columns in
// increasing order (col, int32) and values (val), cnt[row0 + i] of them.  A kNN kernel with
// k << n has about 2k nonzeros per row, so the Lanczos matvec reads 12 B per nonzero instead
// of 8 B per entry of the dense row.  A batch may mix the two forms: with a selection sel
// (int32 [npd]) only PDs with sel[pd] != 0 get sparse rows and the sparse matvec, the others
// keep the dense matvec (which skips the selected PDs).

// one wave per row: ballot compaction of the row's nonzeros, 64 columns at a time
__global__ __launch_bounds__(256) void k_kernel_ell(const int32_t* __restrict__ rows,
                                                    int64_t n_rows,
                                                    const memi_pd_desc* __restrict__ pds,
                                                    const double* __restrict__ L,
                                                    const int32_t* __restrict__ sel,
                                                    int32_t* __restrict__ col,
                                                    double* __restrict__ val,
                                                    int32_t* __restrict__ cnt) {
  const int lane = threadIdx.x & 63;
  const int64_t g = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (g >= n_rows) return;                          // uniform over the wave
  const int pd = rows[2 * g], i = rows[2 * g + 1];
  if (sel && !sel[pd]) return;
  const memi_pd_desc P = pds[pd];
  const int n = P.n;
  const int64_t off = P.d_off + (int64_t)i * n;
  const double* Lr = L + off;
  int base = 0;
  for (int c0 = 0; c0 < n; c0 += 64) {
    const int c = c0 + lane;
    const double x = c < n ? Lr[c] : 0.0;
    const uint64_t nz = __ballot(x != 0.0);
    if (x != 0.0) {
      const int pos = base + __popcll(nz & ((1ull << lane) - 1ull));
      col[off + pos] = c;
      val[off + pos] = x;
    }
    base += __popcll(nz);
  }
  if (lane == 0) cnt[P.row0 + i] = base;
}

// w = L v_j from the sparse rows: each wave takes `rpw` consecutive rows of the plan's row
// list, as the dense matvec does; four rows of one PD go together (their nonzeros strided over
// the lanes, two per row and lane in flight, so eight index loads and eight gathers of v_j per
// lane), others one at a time with four in flight.  Each row's w_i v_j[i] is left for the
// orthogonalisation.
__device__ __forceinline__ void ell_row(const int32_t* __restrict__ col,
                                        const double* __restrict__ val, int64_t off, int c,
                                        const double* __restrict__ v, int lane,
                                        double* __restrict__ out, double* __restrict__ aout,
                                        double vi) {
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
  int t = lane;
  for (; t + 3 * 64 < c; t += 4 * 64) {
    s0 += val[off + t] * v[col[off + t]];
    s1 += val[off + t + 64] * v[col[off + t + 64]];
    s2 += val[off + t + 128] * v[col[off + t + 128]];
    s3 += val[off + t + 192] * v[col[off + t + 192]];
  }
  for (; t < c; t += 64) s0 += val[off + t] * v[col[off + t]];
  const double w = wave_total((s0 + s1) + (s2 + s3));
  if (lane == 0) {
    *out = w;
    if (aout) *aout = w * vi;
  }
}

__device__ __forceinline__ void ell_rows4(const int32_t* __restrict__ col,
                                          const double* __restrict__ val, int64_t off0,
                                          int64_t stride, const int32_t* __restrict__ cnt4,
                                          const double* __restrict__ v, int lane,
                                          double* __restrict__ out, double* __restrict__ aout,
                                          const double* __restrict__ vrow) {
  int c[4];
  int cmax = 0;
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    c[r] = cnt4[r];
    cmax = max(cmax, c[r]);
  }
  double s[4][2];
#pragma unroll
  for (int r = 0; r < 4; ++r) s[r][0] = s[r][1] = 0.0;
  for (int t = lane; t < cmax; t += 2 * 64) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int64_t o = off0 + r * stride + t;
      if (t < c[r]) s[r][0] += val[o] * v[col[o]];
      if (t + 64 < c[r]) s[r][1] += val[o + 64] * v[col[o + 64]];
    }
  }
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    const double w = wave_total(s[r][0] + s[r][1]);
    if (lane == 0) {
      out[r] = w;
      if (aout) aout[r] = w * vrow[r];
    }
  }
}

__global__ __launch_bounds__(256) void k_lanczos_matvec_ell(const int32_t* __restrict__ rows,
                                                            int64_t n_rows,
                                                            const memi_pd_desc* __restrict__ pds,
                                                            const int32_t* __restrict__ col,
                                                            const double* __restrict__ val,
                                                            const int32_t* __restrict__ cnt,
                                                            const double* __restrict__ V, int m,
                                                            const int32_t* __restrict__ jp,
                                                            const int32_t* __restrict__ active,
                                                            double* __restrict__ W,
                                                            double* __restrict__ A,
                                                            const int32_t* __restrict__ sel,
                                                            int rpw) {
  const int j = *jp;
  const int lane = threadIdx.x & 63;
  const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int64_t g0 = ((int64_t)blockIdx.x * 4 + wid) * rpw;
  const int64_t g1 = min(g0 + (int64_t)rpw, n_rows);
  for (int64_t g = g0; g < g1;) {                     // uniform over the wave
    const int pd = rows[2 * g], i = rows[2 * g + 1];
    const memi_pd_desc P = pds[pd];
    const int n = P.n;
    const bool live = active[pd] && !(sel && !sel[pd]) && j < min(n, m);
    const double* v = V + (int64_t)m * P.row0 + (int64_t)j * n;
    const int64_t off = P.d_off + (int64_t)i * n;
    if (g + 3 < g1 && i + 3 < n) {
      if (live)
        ell_rows4(col, val, off, n, cnt + P.row0 + i, v, lane, W + P.row0 + i,
                  A ? A + P.row0 + i : nullptr, v + i);
      g += 4;
    } else {
      if (live)
        ell_row(col, val, off, cnt[P.row0 + i], v, lane, W + P.row0 + i,
                A ? A + P.row0 + i : nullptr, v[i]);
      g += 1;
    }
  }
}

// one block per PD: the three-term recurrence, one CGS pass against columns 0..j, then
// normalise into column j+1
__device__ inline void lanczos_orth_body(const memi_pd_desc* __restrict__ pds,
                                         double* __restrict__ V, int m, int j,
                                         const int32_t* __restrict__ active,
                                         double* __restrict__ W, double* __restrict__ ab,
                                         double* h, double* red) {
  const int pd = blockIdx.x;
  if (!active[pd]) return;
  const memi_pd_desc P = pds[pd];
  const int n = P.n;
  if (j >= min(n, m)) return;
  double* Vp = V + (int64_t)m * P.row0;
  double* w = W + P.row0;
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  // pass 1: w' = w - alpha v_j - beta_{j-1} v_{j-1}, alpha = v_j . w
  const double* vj = Vp + (int64_t)j * n;
  const double* vjm = j > 0 ? vj - n : vj;
  const double bp = j > 0 ? ab[((int64_t)pd * m + j - 1) * 2 + 1] : 0.0;
  double sa = 0.0;
  for (int r = threadIdx.x; r < n; r += blockDim.x) sa += vj[r] * w[r];
  double alpha = block_sum(sa, red);
  for (int r = threadIdx.x; r < n; r += blockDim.x) w[r] = (w[r] - alpha * vj[r]) - bp * vjm[r];
  __syncthreads();
  // pass 2: h = V[:, 0..j]^T w' (one wave per column), w'' = w' - V h
  for (int c = wid; c <= j; c += 4) {
    const double* vc = Vp + (int64_t)c * n;
    double s = 0.0;
    for (int r = lane; r < n; r += 64) s += vc[r] * w[r];
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    if (lane == 0) h[c] = s;
  }
  __syncthreads();
  alpha += h[j];
  for (int r = threadIdx.x; r < n; r += blockDim.x) {
    double acc = 0.0;
    for (int c = 0; c <= j; ++c) acc += Vp[(int64_t)c * n + r] * h[c];
    w[r] -= acc;
  }
  __syncthreads();
  double ss = 0.0;
  for (int r = threadIdx.x; r < n; r += blockDim.x) ss += w[r] * w[r];
  const double beta = sqrt(block_sum(ss, red));
  if (threadIdx.x == 0) {
    ab[((int64_t)pd * m + j) * 2 + 0] = alpha;
    ab[((int64_t)pd * m + j) * 2 + 1] = beta;
  }
  if (j + 1 < m) {
    const double inv = beta > 0.0 ? 1.0 / beta : 0.0;
    double* vn = Vp + (int64_t)(j + 1) * n;
    for (int r = threadIdx.x; r < n; r += blockDim.x) vn[r] = w[r] * inv;
  }
}

__global__ __launch_bounds__(256) void k_lanczos_orth(const memi_pd_desc* __restrict__ pds,
                                                      double* __restrict__ V, int m,
                                                      int32_t* __restrict__ jp,
                                                      const int32_t* __restrict__ active,
                                                      double* __restrict__ W,
                                                      double* __restrict__ ab) {
  extern __shared__ double h[];  // m coefficients
  __shared__ double red[8];
  const int j = jp[0];
  lanczos_orth_body(pds, V, m, j, active, W, ab, h, red);
  step_arrive(jp, j);
}


// Grid-parallel orthogonalisation for large PDs (one workgroup per PD leaves most CUs idle
// when a batch has few PDs), two kernels after the matvec:
//   k_orth_dot (one workgroup per (column, PD)) or k_orth_dot_pairs (many PDs): alpha = the sum
//     of the rows' w_i v_j[i] that the matvec left in A (summed identically by every block), and
//     with w' = w - alpha v_j - beta_{j-1} v_{j-1} formed on the fly, h = V^T w' and w'.w';
//   k_orth_update_rows: w'' = w' - V h, and since V is orthonormal and h is of order eps
//     relative to w', beta^2 = w'.w' - ||h||^2 is known before w'' is, so each row of
//     v_{j+1} = w'' / beta is written where it is computed; alpha_j = alpha + h_j.
// H: float64 [npd][m + 2]: h in slots 0..j, w'.w' in slot m, alpha in slot m + 1.
constexpr int H_STRIDE_EXTRA = 2;

// w'_r of the three-term recurrence
__device__ __forceinline__ double w_local(const double* __restrict__ w,
                                          const double* __restrict__ vj,
                                          const double* __restrict__ vjm, double alpha,
                                          double bp, int r) {
  return (w[r] - alpha * vj[r]) - bp * vjm[r];
}

__global__ __launch_bounds__(256) void k_orth_dot(const memi_pd_desc* __restrict__ pds,
                                                  const double* __restrict__ V, int m,
                                                  const int32_t* __restrict__ jp,
                                                  const int32_t* __restrict__ active,
                                                  const double* __restrict__ W,
                                                  const double* __restrict__ A,
                                                  const double* __restrict__ ab,
                                                  double* __restrict__ H) {
  __shared__ double red[8];
  const int j = *jp;
  const int pd = blockIdx.y;
  if (!active[pd]) return;
  const memi_pd_desc P = pds[pd];
  const int n = P.n;
  if (j >= min(n, m)) return;
  const double* w = W + P.row0;
  const double* Vp = V + (int64_t)m * P.row0;
  const double* vj = Vp + (int64_t)j * n;
  const double* vjm = j > 0 ? vj - n : vj;
  const double bp = j > 0 ? ab[((int64_t)pd * m + j - 1) * 2 + 1] : 0.0;
  double sa = 0.0;
  for (int r = threadIdx.x; r < n; r += blockDim.x) sa += A[P.row0 + r];
  const double alpha = block_sum(sa, red);
  const int64_t hs = m + H_STRIDE_EXTRA;
  // a fixed grid of blocks walks the live columns and, as entry j + 1, w' itself (to slot m)
  for (int c = blockIdx.x; c < j + 2; c += gridDim.x) {   // uniform over the block
    const bool norm = c == j + 1;
    const double* vc = Vp + (int64_t)(norm ? j : c) * n;
    double s = 0.0;
    for (int r = threadIdx.x; r < n; r += blockDim.x) {
      const double wr = w_local(w, vj, vjm, alpha, bp, r);
      s += (norm ? wr : vc[r]) * wr;
    }
    s = block_sum(s, red);
    if (threadIdx.x == 0) H[(int64_t)pd * hs + (norm ? m : c)] = s;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) H[(int64_t)pd * hs + m + 1] = alpha;
}

// Many-PD form of the dot pass: a fixed grid whose waves walk the npd x (j + 2) live
// (PD, column) pairs, one wave per pair (grids sized for the largest j leave most blocks empty
// when there are hundreds of PDs).  The pair index is uniform over the wave (scalar loads of
// the PD descriptor and flag); each wave sums alpha itself, in one fixed order, and the wave of
// column 0 stores it.
__global__ __launch_bounds__(256) void k_orth_dot_pairs(const memi_pd_desc* __restrict__ pds,
                                                        int64_t npd,
                                                        const double* __restrict__ V, int m,
                                                        const int32_t* __restrict__ jp,
                                                        const int32_t* __restrict__ active,
                                                        const double* __restrict__ W,
                                                        const double* __restrict__ A,
                                                        const double* __restrict__ ab,
                                                        double* __restrict__ H) {
  const int j = *jp;
  const int lane = threadIdx.x & 63;
  const int per = j + 2;
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
  const int64_t total = npd * (int64_t)per;
  const int wid = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int64_t hs = m + H_STRIDE_EXTRA;
  for (int64_t q = (int64_t)blockIdx.x * (blockDim.x >> 6) + wid; q < total; q += nwaves) {
    const int pd = (int)(q / per), c = (int)(q % per);
    if (!active[pd]) continue;                    // uniform over the wave
    const memi_pd_desc P = pds[pd];
    const int n = P.n;
    if (j >= min(n, m)) continue;
    const double* w = W + P.row0;
    const double* Vp = V + (int64_t)m * P.row0;
    const double* vj = Vp + (int64_t)j * n;
    const double* vjm = j > 0 ? vj - n : vj;
    const double bp = j > 0 ? ab[((int64_t)pd * m + j - 1) * 2 + 1] : 0.0;
    double sa = 0.0;
    for (int r = lane; r < n; r += 64) sa += A[P.row0 + r];
    const double alpha = wave_total(sa);
    const bool norm = c == j + 1;
    const double* vc = Vp + (int64_t)(norm ? j : c) * n;
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    int r = lane;
    for (; r + 3 * 64 < n; r += 4 * 64) {
      const double w0 = w_local(w, vj, vjm, alpha, bp, r);
      const double w1 = w_local(w, vj, vjm, alpha, bp, r + 64);
      const double w2 = w_local(w, vj, vjm, alpha, bp, r + 128);
      const double w3 = w_local(w, vj, vjm, alpha, bp, r + 192);
      s0 += (norm ? w0 : vc[r]) * w0;
      s1 += (norm ? w1 : vc[r + 64]) * w1;
      s2 += (norm ? w2 : vc[r + 128]) * w2;
      s3 += (norm ? w3 : vc[r + 192]) * w3;
    }
    for (; r < n; r += 64) {
      const double w0 = w_local(w, vj, vjm, alpha, bp, r);
      s0 += (norm ? w0 : vc[r]) * w0;
    }
    const double s = wave_total((s0 + s1) + (s2 + s3));
    if (lane == 0) {
      H[(int64_t)pd * hs + (norm ? m : c)] = s;
      if (c == 0) H[(int64_t)pd * hs + m + 1] = alpha;
    }
  }
}

// Row-list form of the finishing update: 64 consecutive rows of the plan per block (rows of two
// PDs may share a block), 4 column groups each: w''_r = w'_r - sum_c V[c][r] h[c] and
// v_{j+1}[r] = w''_r / beta with beta^2 = w'.w' - ||h||^2; the PD's first row records
// alpha_j = alpha + h[j] and beta_j, and the grid advances j.
__global__ __launch_bounds__(256) void k_orth_update_rows(const int32_t* __restrict__ rows,
                                                          int64_t n_rows,
                                                          const memi_pd_desc* __restrict__ pds,
                                                          double* __restrict__ V, int m,
                                                          int32_t* __restrict__ jp,
                                                          const int32_t* __restrict__ active,
                                                          const double* __restrict__ W,
                                                          const double* __restrict__ H,
                                                          int arrive, double* __restrict__ ab) {
  __shared__ double part[4][64];
  __shared__ double qpart[4][64];
  const int j = jp[0];
  const int64_t g = (int64_t)blockIdx.x * 64 + (threadIdx.x & 63);
  const int cg = threadIdx.x >> 6;
  const int64_t hs = m + H_STRIDE_EXTRA;
  double acc = 0.0, q = 0.0;
  bool live = false;
  int pd = 0, i = 0;
  memi_pd_desc P{};
  if (g < n_rows) {
    pd = rows[2 * g];
    i = rows[2 * g + 1];
    P = pds[pd];
    if (active[pd] && j < min(P.n, m)) {
      live = true;
      const double* Vp = V + (int64_t)m * P.row0 + i;
      const double* h = H + (int64_t)pd * hs;
      for (int c = cg; c <= j; c += 4) {
        const double hc = h[c];
        acc += Vp[(int64_t)c * P.n] * hc;
        q += hc * hc;                  // ||h||^2 in the same column groups, for beta
      }
    }
  }
  part[cg][threadIdx.x & 63] = acc;
  qpart[cg][threadIdx.x & 63] = q;
  __syncthreads();
  if (cg == 0 && live) {
    const int t = threadIdx.x;
    const double* h = H + (int64_t)pd * hs;
    const double alpha = h[m + 1];
    const double* Vp = V + (int64_t)m * P.row0;
    const double* vj = Vp + (int64_t)j * P.n;
    const double* vjm = j > 0 ? vj - P.n : vj;
    const double bp = j > 0 ? ab[((int64_t)pd * m + j - 1) * 2 + 1] : 0.0;
    const double w = w_local(W + P.row0, vj, vjm, alpha, bp, i) -
                     ((part[0][t] + part[1][t]) + (part[2][t] + part[3][t]));
    // every row of the PD sums the same column-group partials in the same order: one beta
    const double qq = (qpart[0][t] + qpart[1][t]) + (qpart[2][t] + qpart[3][t]);
    const double beta = sqrt(fmax(h[m] - qq, 0.0));
    if (j + 1 < m) {
      const double inv = beta > 0.0 ? 1.0 / beta : 0.0;
      V[(int64_t)m * P.row0 + (int64_t)(j + 1) * P.n + i] = w * inv;
    }
    if (i == 0) {
      ab[((int64_t)pd * m + j) * 2 + 0] = alpha + h[j];
      ab[((int64_t)pd * m + j) * 2 + 1] = beta;
    }
  }
  if (arrive) step_arrive(jp, j);
}

// Advances the step index after a finishing update launched without arrivals (large grids,
// where thousands of blocks contending for one counter cost more than this launch).
__global__ void k_step_advance(int32_t* __restrict__ jp) {
  if (threadIdx.x == 0) jp[0] = jp[0] + 1;
}

// ---- convergence check: Ritz pairs of the tridiagonal T_p on the GPU -----------------------
// One workgroup per active PD with J = min(steps, n, m) Lanczos steps.
//  1. candidates: the k smallest and k largest eigenvalues of T (all J when J <= 2k), each by
//     Sturm-count multisection on the Gershgorin interval;
//  2. the k of largest |theta| (first of equal magnitudes in candidate order), sorted by
//     descending theta -- the pairs eigsh(which='LM') returns;
//  3. their eigenvectors by the twisted factorisation of T - theta I (forward and backward
//     pivots d+, d-, twist r = argmin |gamma_r|, then the two recurrences from z_r = 1),
//     normalised, then modified Gram-Schmidt in descending-theta order;
//  4. residuals |beta_{J-1} s_{J-1}|; the PD is done when all are <= tol * max|theta|, or
//     J = n (the whole space), or J = m.  res = the largest residual (0 when J = n), written at
//     every check; done PDs get active = 0, steps = J; theta [npd][16] and S (row-major J x k at
//     pd*m*16) stay as the final Ritz pairs.
// piv: float64 scratch [npd][16][2][m].
// Sturm count (eigenvalues below x) from the LDS copy q2[i] = {alpha_i, beta_{i-1}^2},
// division free: the sign changes of p_0 = 1, p_i = (alpha_i - x) p_{i-1} - beta_{i-1}^2 p_{i-2}
// (a zero p_i counts as a change, as a -pivmin pivot does in the ratio form), rescaled by exact
// powers of two when |p| leaves [2^-600, 2^600].  Four steps per load batch and range check: the
// loads go out together instead of one LDS round trip per step, and four steps move |p| by far
// less than the 2^400 of headroom on either side.
__device__ inline void sturm_step(double2 q, double x, double& p, double& pm, int& c) {
  double pn = (q.x - x) * p - q.y * pm;
  if (pn == 0.0) pn = p < 0.0 ? 1e-300 : -1e-300;
  c += (pn < 0.0) != (p < 0.0);
  pm = p;
  p = pn;
}

__device__ inline void sturm_rescale(double& p, double& pm) {
  const double a = fabs(p);
  if (a > 0x1p600 || a < 0x1p-600) {
    int e;
    frexp(a, &e);
    pm = ldexp(pm, -e);
    p = ldexp(p, -e);
  }
}

__device__ inline int sturm_count(const double2* q2, int J, double x) {
  double pm = 1.0, p = q2[0].x - x;
  if (p == 0.0) p = -1e-300;
  int c = p < 0.0;
  int i = 1;
  for (; i + 3 < J; i += 4) {
    const double2 u0 = q2[i], u1 = q2[i + 1], u2 = q2[i + 2], u3 = q2[i + 3];
    sturm_step(u0, x, p, pm, c);
    sturm_step(u1, x, p, pm, c);
    sturm_step(u2, x, p, pm, c);
    sturm_step(u3, x, p, pm, c);
    sturm_rescale(p, pm);
  }
  for (; i < J; ++i) sturm_step(q2[i], x, p, pm, c);
  return c;
}

// MAXM: the largest m of the instantiation.  MAXM <= 512 keeps alpha, beta^2 and the twisted
// factorisation's pivots in LDS (pivot (i, side, pair) at (2 i + side) * 16 + pair, so the
// pairs' threads hit consecutive banks); larger MAXM keeps the pivots in `piv` (global).
// RITZ_NT threads: all of them multisect the candidate eigenvalues (a group of G lanes of one
// wave per eigenvalue evaluates G Sturm counts per round, so each round narrows the bracket
// G + 1 fold instead of 2); the selection, vectors and Gram-Schmidt use the first wave.
constexpr int RITZ_NT = 1024;

template <int MAXM>
__global__ __launch_bounds__(RITZ_NT) void k_lanczos_ritz(const memi_pd_desc* __restrict__ pds,
                                                          const double* __restrict__ ab, int m,
                                                          int steps_done,
                                                          const int32_t* __restrict__ ks,
                                                          double tol,
                                                          int32_t* __restrict__ active,
                                                          double* __restrict__ theta,
                                                          double* __restrict__ S,
                                                          int32_t* __restrict__ steps,
                                                          double* __restrict__ res_out,
                                                          double* __restrict__ piv,
                                                          double* __restrict__ summary) {
  constexpr int KM = MEMI_RITZ_KMAX;
  constexpr bool PLDS = MAXM <= 512;
  __shared__ double2 q2[MAXM];                       // {alpha_i, beta_{i-1}^2}
  __shared__ double bt[MAXM];                         // beta_i
  __shared__ double pv_lds[PLDS ? 2 * MAXM * KM : 1];
  __shared__ double cand[2 * KM];
  __shared__ double sel[KM];
  __shared__ double resv[KM];
  __shared__ double gersh[3];
  __shared__ int done_sh;
  const int pd = blockIdx.x;
  if (!active[pd]) {                                  // uniform over the block
    if (summary && threadIdx.x == 0) {
      summary[pd] = 0.0;
      summary[gridDim.x + pd] = 0.0;
      summary[2 * gridDim.x + pd] = 0.0;
    }
    return;
  }
  const int t = threadIdx.x;
  const int n = pds[pd].n;
  const int J = min(steps_done, min(n, m));
  const int k = min(ks[pd], J);
  const double* a = ab + (int64_t)pd * m * 2;      // a[2i] = alpha_i, a[2i+1] = beta_i
  for (int i = t; i < J; i += RITZ_NT) {
    const double bp = i > 0 ? a[2 * (i - 1) + 1] : 0.0;
    q2[i] = make_double2(a[2 * i], bp * bp);
    bt[i] = a[2 * i + 1];
  }
  // Gershgorin interval and pivot floor (first wave)
  if (t < 64) {
    double lo = INFINITY, hi = -INFINITY, bmax = 0.0;
    for (int i = t; i < J; i += 64) {
      const double bl = i > 0 ? fabs(a[2 * (i - 1) + 1]) : 0.0;
      const double br = i + 1 < J ? fabs(a[2 * i + 1]) : 0.0;
      lo = fmin(lo, a[2 * i] - bl - br);
      hi = fmax(hi, a[2 * i] + bl + br);
      bmax = fmax(bmax, br);
    }
    for (int off = 32; off > 0; off >>= 1) {
      lo = fmin(lo, __shfl_xor(lo, off, 64));
      hi = fmax(hi, __shfl_xor(hi, off, 64));
      bmax = fmax(bmax, __shfl_xor(bmax, off, 64));
    }
    if (t == 0) {
      gersh[0] = lo;
      gersh[1] = hi;
      gersh[2] = bmax;
    }
  }
  __syncthreads();
  double lo = gersh[0], hi = gersh[1];
  const double bmax = gersh[2];
  const double span = fmax(hi - lo, 1e-300);
  lo -= 1e-14 * span;
  hi += 1e-14 * span;
  const double pivmin = fmax(1e-290, 1e-300 * bmax * bmax);
  const double atol = 1e-15 * span;    // eigenvalues to 1e-15 of the spectrum's width
  // 1. candidate eigenvalues by multisection: G lanes (a power of two, within one wave) per
  //    candidate; lane g counts the eigenvalues below X(g + 1), X(j) = l + (h - l) j / (G + 1),
  //    and the bracket becomes [X(f), X(f + 1)] around the first point with more than idx below
  const int nc = J <= 2 * k ? J : 2 * k;
  // 8 lanes per candidate: 2k <= 32 candidates keep up to 4 waves busy; more lanes per candidate
  // (fewer rounds, 9-fold narrowing each) cost more issue slots than they save in latency
  // (measured at J = 32...96: 8 lanes fastest, 32 lanes 15-20% slower)
  int G = 8;
  while (G > 1 && G * nc > RITZ_NT) G >>= 1;
  const int ci = t / G, g = t % G;
  if (ci < nc) {                                      // whole groups, so ballots stay in-group
    const int idx = J <= 2 * k ? ci : (ci < k ? ci : J - 2 * k + ci);
    const int base = (t & 63) & ~(G - 1);
    const uint64_t gmask = G == 64 ? ~0ull : ((1ull << G) - 1ull);
    double l = lo, h = hi;
    for (int it = 0; it < 64 && h - l > fmax(2.2e-16 * fmax(fabs(l), fabs(h)), atol); ++it) {
      const double w = h - l;
      const double x = l + (w * (double)(g + 1)) / (double)(G + 1);
      const bool above = sturm_count(q2, J, x) > idx;
      const uint64_t bits = (__ballot(above) >> base) & gmask;
      const int f = bits ? __ffsll((unsigned long long)bits) - 1 : G;
      const double nl = f == 0 ? l : l + (w * (double)f) / (double)(G + 1);
      const double nh = f == G ? h : l + (w * (double)(f + 1)) / (double)(G + 1);
      if (nl <= l && nh >= h) break;                  // no progress: the bracket is at ulps
      l = nl;
      h = nh;
    }
    if (g == 0) cand[ci] = 0.5 * (l + h);
  }
  __syncthreads();
  // 2. selection: the k largest |theta| (of equal magnitudes, the lower candidate index first),
  //    placed by descending theta.  Thread c ranks candidate c against all others (a serial
  //    scan on one thread waits on one LDS read after another)
  __shared__ int taken[2 * KM];
  double tc = 0.0;
  if (t < nc) {
    tc = cand[t];
    const double ac = fabs(tc);
    int rank = 0;
    for (int c = 0; c < nc; ++c) {
      const double a2 = fabs(cand[c]);
      rank += (a2 > ac) || (a2 == ac && c < t);
    }
    taken[t] = rank < k;
  }
  __syncthreads();
  if (t < nc && taken[t]) {
    int pos = 0;
    for (int c = 0; c < nc; ++c)
      pos += taken[c] && (cand[c] > tc || (cand[c] == tc && c < t));
    sel[pos] = tc;
  }
  __syncthreads();
  // 3. twisted factorisation per selected pair, two adjacent threads of the first wave per
  //    pair (the forward pivots d+ on one, the backward d- on the other; then each takes half
  //    of the twist search and one side of the vector's recurrences): pivots, twist, the
  //    vector's norm and last component (the vector itself is written only once the PD is done)
  double* pvb = PLDS ? pv_lds : piv + (int64_t)pd * 2 * m * KM;
  const int e = t >> 1, side = t & 1;
  const bool tw = t < 2 * k;
  auto DP = [&](int i) -> double& { return pvb[(2 * i) * KM + e]; };
  auto DM = [&](int i) -> double& { return pvb[(2 * i + 1) * KM + e]; };
  int r = 0;
  double inv = 0.0, th = 0.0;
  if (tw) {
    th = sel[e];
    if (side == 0) {
      double q = q2[0].x - th;
      DP(0) = q;
      for (int i = 1; i < J; ++i) {
        if (fabs(q) < pivmin) q = -pivmin;
        q = q2[i].x - th - q2[i].y / q;
        DP(i) = q;
      }
    } else {
      double q = q2[J - 1].x - th;
      DM(J - 1) = q;
      for (int i = J - 2; i >= 0; --i) {
        if (fabs(q) < pivmin) q = -pivmin;
        q = q2[i].x - th - q2[i + 1].y / q;
        DM(i) = q;
      }
    }
  }
  __syncthreads();
  if (tw) {
    // twist r = argmin |d+_i + d-_i - (alpha_i - theta)|, first index on ties: the lower half
    // on side 0, the upper on side 1
    const int hmid = (J + 1) / 2;
    double gb = INFINITY;
    int rb = side ? hmid : 0;
    for (int i = side ? hmid : 0; i < (side ? J : hmid); ++i) {
      const double g = fabs(DP(i) + DM(i) - (q2[i].x - th));
      if (g < gb) {
        gb = g;
        rb = i;
      }
    }
    const double go = __shfl_xor(gb, 1, 64);
    const int ro = __shfl_xor(rb, 1, 64);
    r = side == 0 ? (go < gb ? ro : rb) : (gb < go ? rb : ro);
    // norm and last component from z_r = 1: side 0 the rows below r, side 1 the rows after
    double zz = 1.0, part = 0.0, zl = r == J - 1 ? 1.0 : 0.0;
    if (side == 0) {
      for (int i = r - 1; i >= 0; --i) {
        const double d = fabs(DP(i)) < pivmin ? -pivmin : DP(i);
        zz = -(bt[i] / d) * zz;
        part += zz * zz;
      }
    } else {
      for (int i = r; i + 1 < J; ++i) {
        const double d = fabs(DM(i + 1)) < pivmin ? -pivmin : DM(i + 1);
        zz = -(bt[i] / d) * zz;
        part += zz * zz;
        if (i + 1 == J - 1) zl = zz;
      }
    }
    const double po = __shfl_xor(part, 1, 64);
    const double zo = __shfl_xor(zl, 1, 64);
    inv = 1.0 / sqrt(1.0 + (side == 0 ? part + po : po + part));
    // 4. residual |beta_{J-1} s_{J-1}|
    if (side == 0) resv[e] = fabs(bt[J - 1] * zo * inv);
  }
  __syncthreads();
  if (t == 0) {
    double mx = 0.0, rmax = 0.0;
    for (int q = 0; q < k; ++q) {
      mx = fmax(mx, fabs(sel[q]));
      rmax = fmax(rmax, resv[q]);
    }
    const bool done = J >= n || J >= m || rmax <= tol * mx;
    for (int q = 0; q < KM; ++q) theta[(int64_t)pd * KM + q] = q < k ? sel[q] : 0.0;
    done_sh = done;
    res_out[pd] = J >= n ? 0.0 : rmax;               // every check: the caller paces the next
    if (done) {
      active[pd] = 0;
      steps[pd] = J;
    }
    if (summary) {
      summary[pd] = done ? 0.0 : 1.0;
      summary[gridDim.x + pd] = J >= n ? 0.0 : rmax;
      summary[2 * gridDim.x + pd] = mx;
    }
  }
  __syncthreads();
  if (!done_sh) return;                               // uniform over the block
  // 5. the converged vectors: the recurrences again, normalised, into S (row-major J x k)
  double* Sp = S + (int64_t)pd * m * KM;
  if (tw) {
    double zz = 1.0;
    if (side == 0) {
      Sp[(int64_t)r * k + e] = inv;
      for (int i = r - 1; i >= 0; --i) {
        const double d = fabs(DP(i)) < pivmin ? -pivmin : DP(i);
        zz = -(bt[i] / d) * zz;
        Sp[(int64_t)i * k + e] = zz * inv;
      }
    } else {
      for (int i = r; i + 1 < J; ++i) {
        const double d = fabs(DM(i + 1)) < pivmin ? -pivmin : DM(i + 1);
        zz = -(bt[i] / d) * zz;
        Sp[(int64_t)(i + 1) * k + e] = zz * inv;
      }
    }
  }
  __syncthreads();
  // modified Gram-Schmidt, descending theta, only within clusters (relative gap < 1e-6):
  // vectors of well-separated pairs are orthogonal to ~1e-16 / gap already
  double mx = 0.0;
  for (int q = 0; q < k; ++q) mx = fmax(mx, fabs(sel[q]));
  for (int c = 1; c < k; ++c) {
    for (int b = 0; b < c; ++b) {
      if (fabs(sel[c] - sel[b]) > 1e-6 * mx) continue;   // uniform over the block
      if (t < 64) {                                     // first wave only
        double d = 0.0;
        for (int i = t; i < J; i += 64) d += Sp[(int64_t)i * k + c] * Sp[(int64_t)i * k + b];
        for (int off = 32; off > 0; off >>= 1) d += __shfl_xor(d, off, 64);
        for (int i = t; i < J; i += 64) Sp[(int64_t)i * k + c] -= d * Sp[(int64_t)i * k + b];
      }
      __syncthreads();
      if (t < 64) {
        double nn = 0.0;
        for (int i = t; i < J; i += 64) nn += Sp[(int64_t)i * k + c] * Sp[(int64_t)i * k + c];
        for (int off = 32; off > 0; off >>= 1) nn += __shfl_xor(nn, off, 64);
        const double iv = 1.0 / sqrt(nn);
        for (int i = t; i < J; i += 64) Sp[(int64_t)i * k + c] *= iv;
      }
      __syncthreads();
    }
  }
}

// Ritz vectors Y_p = V_p S_p of every PD in one launch, one thread per row: V_p the n x J
// Lanczos basis (column-major at m*row0), S_p the J x k eigenvectors of T_p (row-major at
// meta.s_off), Y [n_rows][kmax] (columns k..kmax-1 zero).  meta: int64 [npd][3] = {J, k, s_off}.
// 64-thread blocks (a single PD of 2000 rows still spreads over 32 CUs); the S of the block's
// first PD is staged in LDS (read by the whole wave at one address), a second PD's rows in the
// same block read theirs from global memory.  Every output is summed over j in order.
constexpr int RV_NT = 64;
constexpr int RV_SMAX = 2048;

template <bool LDS_S>
__device__ __forceinline__ void ritz_row(const double* __restrict__ v, int n, int J, int k,
                                         const double* __restrict__ Sp, double* acc) {
  int j = 0;
  for (; j + 8 <= J; j += 8) {
    double x[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) x[u] = v[(int64_t)(j + u) * n];
#pragma unroll
    for (int u = 0; u < 8; ++u)
#pragma unroll
      for (int e = 0; e < MEMI_RITZ_KMAX; ++e)
        if (e < k) acc[e] += x[u] * Sp[(int64_t)(j + u) * k + e];
  }
  for (; j < J; ++j) {
    const double x = v[(int64_t)j * n];
#pragma unroll
    for (int e = 0; e < MEMI_RITZ_KMAX; ++e)
      if (e < k) acc[e] += x * Sp[(int64_t)j * k + e];
  }
}

__global__ __launch_bounds__(RV_NT) void k_ritz_vectors(const int32_t* __restrict__ rows,
                                                        int64_t n_rows,
                                                        const memi_pd_desc* __restrict__ pds,
                                                        const double* __restrict__ V, int m,
                                                        const double* __restrict__ S,
                                                        const int64_t* __restrict__ meta,
                                                        int kmax, double* __restrict__ Y) {
  __shared__ double s_sh[RV_SMAX];
  const int64_t g0 = (int64_t)blockIdx.x * RV_NT;          // < n_rows by the grid size
  const int pd0 = rows[2 * g0];
  const int64_t sz0 = meta[3 * pd0] * meta[3 * pd0 + 1];
  const bool staged = sz0 <= RV_SMAX;
  if (staged) {
    const double* src = S + meta[3 * pd0 + 2];
    for (int q = threadIdx.x; q < sz0; q += RV_NT) s_sh[q] = src[q];
  }
  __syncthreads();
  const int64_t g = g0 + threadIdx.x;
  if (g >= n_rows) return;
  const int pd = rows[2 * g], i = rows[2 * g + 1];
  const memi_pd_desc P = pds[pd];
  const int n = P.n;
  const int J = (int)meta[3 * pd], k = (int)meta[3 * pd + 1];
  const double* v = V + (int64_t)m * P.row0 + i;
  double acc[MEMI_RITZ_KMAX];
#pragma unroll
  for (int e = 0; e < MEMI_RITZ_KMAX; ++e) acc[e] = 0.0;
  if (staged && pd == pd0)
    ritz_row<true>(v, n, J, k, s_sh, acc);
  else
    ritz_row<false>(v, n, J, k, S + meta[3 * pd + 2], acc);
#pragma unroll
  for (int e = 0; e < MEMI_RITZ_KMAX; ++e)
    if (e < kmax) Y[g * kmax + e] = e < k ? acc[e] : 0.0;
}

// The start of a Lanczos run in one launch: V[idx[i]] = v0[i] (column 0 of every PD's basis),
// ab = 0, j = {0, 0}, active = 1.
__global__ void k_lanczos_init(double* __restrict__ V, const double* __restrict__ v0,
                               const int64_t* __restrict__ idx, int64_t n_rows,
                               double* __restrict__ ab, int64_t n_ab, int32_t* __restrict__ j,
                               int32_t* __restrict__ active, int64_t npd) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  const int64_t t0 = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  for (int64_t t = t0; t < n_rows; t += stride) V[idx[t]] = v0[t];
  for (int64_t t = t0; t < n_ab; t += stride) ab[t] = 0.0;
  for (int64_t t = t0; t < npd; t += stride) active[t] = 1;
  if (t0 < 2) j[t0] = 0;
}

}  // namespace memi

using namespace memi;

extern "C" {

int memi_lanczos_init(double* V, const double* v0, const int64_t* idx, int64_t n_rows,
                      double* ab, int64_t n_ab, int32_t* j, int32_t* active, int64_t npd,
                      void* stream) {
  MEMI_CHECK_ARG(V && v0 && idx && ab && j && active && n_rows >= 0 && n_ab >= 0 && npd >= 0,
                 "memi_lanczos_init: bad arguments");
  const int64_t work = std::max(std::max(n_rows, n_ab), std::max(npd, (int64_t)2));
  hipLaunchKernelGGL(k_lanczos_init, dim3(grid_for(work)), dim3(256), 0, (hipStream_t)stream, V,
                     v0, idx, n_rows, ab, n_ab, j, active, npd);
  MEMI_CHECK_LAUNCH("k_lanczos_init");
  return MEMI_OK;
}

// L: the dense kernel (every PD when ell_col is null, else the PDs with sel[pd] == 0);
// ell_*: the sparse rows (every PD when sel is null, else the PDs with sel[pd] != 0)
static int lanczos_steps_impl(const void* plan_dev, const memi_plan_info* info, const void* L,
                              int l_is_f64, const int32_t* ell_col, const double* ell_val,
                              const int32_t* ell_cnt, const int32_t* sel, double* V, int m,
                              int32_t* j_dev, int nsteps, const int32_t* active, double* W,
                              double* H, double* ab, int split, void* stream) {
  MEMI_CHECK_ARG(plan_dev && info && (L || ell_col) && V && j_dev && active && W && ab &&
                     (!split || H) && (!sel || (ell_col && L && l_is_f64)),
                 "memi_lanczos_steps: null pointer");
  MEMI_CHECK_ARG(m >= 2 && m <= 4096 && nsteps >= 0, "memi_lanczos_steps: bad m=%d nsteps=%d",
                 m, nsteps);
  MEMI_CHECK_ARG(!split || (info->n_pd <= 65535 && info->max_n <= 65535 * 64),
                 "memi_lanczos_steps: batch too large for the split orthogonalisation");
  if (info->n_pd == 0 || nsteps == 0) return MEMI_OK;
  hipStream_t s = (hipStream_t)stream;
  const char* plan = (const char*)plan_dev;
  const memi_pd_desc* pds = (const memi_pd_desc*)(plan + info->off_pds);
  const int32_t* rows = (const int32_t*)(plan + info->off_rows);
  const unsigned npd = (unsigned)info->n_pd;
  const int32_t* jp = j_dev;
  // rows per wave: one for batches up to 8192 rows (a wave per row fills the GPU), else one
  // group of four (waves of equal row counts but unequal row lengths must stay short, or the
  // waves holding the largest PDs' rows leave a long tail)
  const int64_t rpw = info->n_rows > 8192 ? 4 : 1;
  const int64_t n_waves = (info->n_rows + rpw - 1) / rpw;
  const dim3 gm((unsigned)((n_waves + 3) / 4));
  // the rows' w_i v_j[i] for the grid-parallel orthogonalisation: the second half of W
  double* A = split ? W + info->n_rows : nullptr;
  for (int st = 0; st < nsteps; ++st) {
    if (ell_col) {
      hipLaunchKernelGGL(k_lanczos_matvec_ell, gm, dim3(256), 0, s, rows, info->n_rows, pds,
                         ell_col, ell_val, ell_cnt, V, m, jp, active, W, A, sel, (int)rpw);
      MEMI_CHECK_LAUNCH("k_lanczos_matvec_ell");
    }
    if (!ell_col || sel) {
      if (l_is_f64)
        hipLaunchKernelGGL(k_lanczos_matvec<double>, gm, dim3(256), 0, s, rows, info->n_rows,
                           pds, (const double*)L, V, m, jp, active, W, A, (int)rpw, sel);
      else
        hipLaunchKernelGGL(k_lanczos_matvec<float>, gm, dim3(256), 0, s, rows, info->n_rows,
                           pds, (const float*)L, V, m, jp, active, W, A, (int)rpw, sel);
      MEMI_CHECK_LAUNCH("k_lanczos_matvec");
    }
    if (!split) {
      hipLaunchKernelGGL(k_lanczos_orth, dim3(npd), dim3(256), (size_t)m * sizeof(double), s,
                         pds, V, m, j_dev, active, W, ab);
      MEMI_CHECK_LAUNCH("k_lanczos_orth");
    } else {
      if (npd < 64) {       // few PDs: a 256-thread block per (column, PD)
        hipLaunchKernelGGL(k_orth_dot, dim3((unsigned)min(m + 1, 64), npd), dim3(256), 0, s,
                           pds, V, m, jp, active, W, A, ab, H);
      } else {              // many PDs: waves walk the live (PD, column) pairs; 2048
                            // blocks fill the CUs (1024 left the short, latency-bound waves
                            // half the occupancy: 3 % slower on a 500-PD batch)
        hipLaunchKernelGGL(k_orth_dot_pairs, dim3(2048), dim3(256), 0, s, pds, info->n_pd, V, m,
                           jp, active, W, A, ab, H);
      }
      MEMI_CHECK_LAUNCH("k_orth_dot");
      const unsigned ub = (unsigned)((info->n_rows + 63) / 64);
      const int arrive = ub <= 256;   // else one thread advances the step after the update
      hipLaunchKernelGGL(k_orth_update_rows, dim3(ub), dim3(256), 0, s, rows, info->n_rows, pds,
                         V, m, j_dev, active, W, H, arrive, ab);
      MEMI_CHECK_LAUNCH("k_orth_update_rows");
      if (!arrive) {
        hipLaunchKernelGGL(k_step_advance, dim3(1), dim3(64), 0, s, j_dev);
        MEMI_CHECK_LAUNCH("k_step_advance");
      }
    }
  }
  return MEMI_OK;
}

int memi_lanczos_steps(const void* plan_dev, const memi_plan_info* info, const void* L,
                       int l_is_f64, double* V, int m, int32_t* j_dev, int nsteps,
                       const int32_t* active, double* W, double* H, double* ab, int split,
                       void* stream) {
  MEMI_CHECK_ARG(L, "memi_lanczos_steps: null pointer");
  return lanczos_steps_impl(plan_dev, info, L, l_is_f64, nullptr, nullptr, nullptr, nullptr, V,
                            m, j_dev, nsteps, active, W, H, ab, split, stream);
}

int memi_lanczos_steps_sparse(const void* plan_dev, const memi_plan_info* info,
                              const int32_t* ell_col, const double* ell_val,
                              const int32_t* ell_cnt, const int32_t* sparse_pd, const double* L,
                              double* V, int m, int32_t* j_dev, int nsteps,
                              const int32_t* active, double* W, double* H, double* ab,
                              int split, void* stream) {
  MEMI_CHECK_ARG(ell_col && ell_val && ell_cnt && (!sparse_pd || L),
                 "memi_lanczos_steps_sparse: null pointer (a selection needs the dense L)");
  return lanczos_steps_impl(plan_dev, info, L, 1, ell_col, ell_val, ell_cnt, sparse_pd, V, m,
                            j_dev, nsteps, active, W, H, ab, split, stream);
}

int memi_kernel_ell(const void* plan_dev, const memi_plan_info* info, const double* L,
                    const int32_t* sparse_pd, int32_t* ell_col, double* ell_val,
                    int32_t* ell_cnt, void* stream) {
  MEMI_CHECK_ARG(plan_dev && info && L && ell_col && ell_val && ell_cnt,
                 "memi_kernel_ell: null pointer");
  if (info->n_rows == 0) return MEMI_OK;
  const char* plan = (const char*)plan_dev;
  const memi_pd_desc* pds = (const memi_pd_desc*)(plan + info->off_pds);
  const int32_t* rows = (const int32_t*)(plan + info->off_rows);
  hipLaunchKernelGGL(k_kernel_ell, dim3((unsigned)((info->n_rows + 3) / 4)), dim3(256), 0,
                     (hipStream_t)stream, rows, info->n_rows, pds, L, sparse_pd, ell_col, ell_val,
                     ell_cnt);
  MEMI_CHECK_LAUNCH("k_kernel_ell");
  return MEMI_OK;
}

int memi_ritz_vectors(const void* plan_dev, const memi_plan_info* info, const double* V, int m,
                      const double* S, const int64_t* meta, int kmax, double* Y, void* stream) {
  MEMI_CHECK_ARG(plan_dev && info && V && S && meta && Y, "memi_ritz_vectors: null pointer");
  MEMI_CHECK_ARG(m >= 1 && kmax >= 1 && kmax <= MEMI_RITZ_KMAX,
                 "memi_ritz_vectors: bad m=%d kmax=%d (at most %d)", m, kmax, MEMI_RITZ_KMAX);
  if (info->n_rows == 0) return MEMI_OK;
  hipStream_t s = (hipStream_t)stream;
  const char* plan = (const char*)plan_dev;
  const memi_pd_desc* pds = (const memi_pd_desc*)(plan + info->off_pds);
  const int32_t* rows = (const int32_t*)(plan + info->off_rows);
  hipLaunchKernelGGL(k_ritz_vectors, dim3((unsigned)((info->n_rows + RV_NT - 1) / RV_NT)),
                     dim3(RV_NT), 0,
                     s, rows, info->n_rows, pds, V, m, S, meta, kmax, Y);
  MEMI_CHECK_LAUNCH("k_ritz_vectors");
  return MEMI_OK;
}

int memi_lanczos_ritz(const void* plan_dev, const memi_plan_info* info, const double* ab, int m,
                      int steps_done, const int32_t* ks, double tol, int32_t* active,
                      double* theta, double* S, int32_t* steps, double* res, double* piv,
                      double* summary, void* stream) {
  MEMI_CHECK_ARG(plan_dev && info && ab && ks && active && theta && S && steps && res && piv,
                 "memi_lanczos_ritz: null pointer");
  MEMI_CHECK_ARG(m >= 2 && m <= 4096 && steps_done >= 1 && steps_done <= m && tol >= 0,
                 "memi_lanczos_ritz: bad m=%d steps=%d", m, steps_done);
  if (info->n_pd == 0) return MEMI_OK;
  const char* plan = (const char*)plan_dev;
  const memi_pd_desc* pds = (const memi_pd_desc*)(plan + info->off_pds);
  if (m <= 512)
    hipLaunchKernelGGL(k_lanczos_ritz<512>, dim3((unsigned)info->n_pd), dim3(RITZ_NT), 0,
                       (hipStream_t)stream, pds, ab, m, steps_done, ks, tol, active, theta, S,
                       steps, res, piv, summary);
  else
    hipLaunchKernelGGL(k_lanczos_ritz<4096>, dim3((unsigned)info->n_pd), dim3(RITZ_NT), 0,
                       (hipStream_t)stream, pds, ab, m, steps_done, ks, tol, active, theta, S,
                       steps, res, piv, summary);
  MEMI_CHECK_LAUNCH("k_lanczos_ritz");
  return MEMI_OK;
}

}  // extern "C"
