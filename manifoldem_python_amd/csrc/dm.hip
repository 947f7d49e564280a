// Stage 2 diffusion-map kernels: kNN selection, Ferguson bandwidth sweep, Markov
// normalisation.  Reference: DMembeddingII.py:43-140, fergusonE.py:25-43,
// slaplacianonFly.py:42-81.
//
// The reference builds a COO list from the kNN lists (zero entries first, then the
// symmetrised non-zero entries) and then works on that list.  Here the list is implicit:
// entry (i, j) of a PD belongs to the kernel
//   * as a zero entry (value 0, weight exp(0) = 1) when j is on i's kNN list and
//     D_ij < 1e-6 (always true for i == j: the self entry is reset to 0,
//     DMembeddingII.py:53,115);
//   * as a non-zero entry with value D_ij when i != j, D_ij >= 1e-6 and j is on i's list or
//     i is on j's list (construct_matrix0, DMembeddingII.py:73-78, gives d^2 either way);
//   * not at all otherwise.
// M[i][j] = [j on i's kNN list] is the bitmask written by k_knn_select.
#include "common.h"

namespace memi {

constexpr int TB = 128;  // tile side of the full-matrix tiles of the plan
// blocks per 128 x 128 tile in the Ferguson and normalisation passes, each a quarter of the
// tile (one block per tile left most of each CU idle for batches of few tiles)
// blocks per 128 x 128 tile: the normalisation kernels (rows split MARKOV_RQ ways; 8 keeps
// eight waves per SIMD on a one-PD batch of 2000 rows, whose float64 chains are latency-bound
// at fewer), the moment sums (entries split FM_RQ ways; 80 KB of LDS per block allow two per
// CU) and the statistics (FS_RQ ways; their counters are spread over FER_SLOTS slots per PD)
constexpr int MARKOV_RQ = 8;
constexpr int FM_RQ = 4;
constexpr int FS_RQ = 4;
constexpr int FM_GP = 8;  // grid parameters per PD of the Ferguson moments (k_fer_grid)

__device__ inline uint32_t float_key(float v) {
  const uint32_t u = __float_as_uint(v);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}

__device__ inline bool mbit(const uint32_t* __restrict__ mask, int64_t off, int W, int i, int j) {
  return (mask[off + (int64_t)i * W + (j >> 5)] >> (j & 31)) & 1u;
}

// Value of entry (i, j) in the kernel, or -1 when it is not an entry.  dij = D_ij = D_ji
// (D is exactly symmetric), mij = [j on i's list], mji = [i on j's list].
__device__ inline double entry_from(float dij, bool mij, bool mji, bool diag) {
  if (diag) return 0.0;
  if (!((double)dij >= 1e-6)) return mij ? 0.0 : -1.0;
  return (mij || mji) ? (double)dij : -1.0;
}

__device__ inline double entry_value(const float* __restrict__ D, const uint32_t* __restrict__ mask,
                                     const memi_pd_desc& P, int W, bool full, int i, int j) {
  const float dij = D[P.d_off + (int64_t)i * P.n + j];
  const bool mij = full || mbit(mask, P.mask_off, W, i, j);
  const bool mji = full || mbit(mask, P.mask_off, W, j, i);
  return entry_from(dij, mij, mji, i == j);
}

// The transposed list bits of one 128 x 128 tile (ti, tj) in LDS: bit il of mt[jl][il / 32] is
// [i on j's list] for i = ti TB + il, j = tj TB + jl.  Read row by row from the mask (4 words
// per row of the tile) once per block; read per entry it was a gather with a row stride of W
// words across the lanes, 64 cache lines per wave-wide load.  Call before any thread leaves.
typedef uint32_t TileMaskT[TB][TB / 32];
__device__ inline void load_tile_mask_t(TileMaskT& mt, const uint32_t* __restrict__ mask,
                                        const memi_pd_desc& P, int W, bool full, int ti, int tj) {
  if (!full)
    for (int q = threadIdx.x; q < TB * (TB / 32); q += blockDim.x) {
      const int jl = q / (TB / 32), w = q % (TB / 32);
      const int j = tj * TB + jl, word = ti * (TB / 32) + w;
      mt[jl][w] = (j < P.n && word < W) ? mask[P.mask_off + (int64_t)j * W + word] : 0u;
    }
  __syncthreads();
}

__device__ inline bool tile_mji(const TileMaskT& mt, int il, int jl) {
  return (mt[jl][il >> 5] >> (il & 31)) & 1u;
}

// entry_value with [i on j's list] from the block's LDS copy; il, jl: i, j within the tile
__device__ inline double entry_value_t(const float* __restrict__ D,
                                       const uint32_t* __restrict__ mask, const memi_pd_desc& P,
                                       int W, bool full, int i, int j, const TileMaskT& mt,
                                       int il, int jl) {
  const float dij = D[P.d_off + (int64_t)i * P.n + j];
  const bool mij = full || mbit(mask, P.mask_off, W, i, j);
  const bool mji = full || tile_mji(mt, il, jl);
  return entry_from(dij, mij, mji, i == j);
}

// inclusive block scan (256 threads) of an int; returns exclusive prefix, total in *tot
__device__ inline int block_excl_scan(int v, int* sh, int* tot) {
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  int x = v;
  for (int off = 1; off < 64; off <<= 1) {
    const int y = __shfl_up(x, off, 64);
    if (lane >= off) x += y;
  }
  if (lane == 63) sh[wid] = x;
  __syncthreads();
  int base = 0, all = 0;
  for (int w = 0; w < 4; ++w) {
    if (w < wid) base += sh[w];
    all += sh[w];
  }
  __syncthreads();
  *tot = all;
  return base + x - v;
}

// ----------------------------------------------------------------------------- kNN ------
// One block per row.  Dynamic LDS: keys[max_n] (u32) + member flags[max_n] (u8) and, for the
// optional sorted lists, pairs[pow2 >= k] (u64).
__global__ __launch_bounds__(256) void k_knn_select(const int32_t* __restrict__ rows,
                                                    const memi_pd_desc* __restrict__ pds,
                                                    const float* __restrict__ D,
                                                    uint32_t* __restrict__ mask,
                                                    int32_t* __restrict__ idx_out,
                                                    float* __restrict__ val_out, int max_n,
                                                    int sort_cap) {
  extern __shared__ __attribute__((aligned(16))) char dyn[];
  uint32_t* keys = (uint32_t*)dyn;
  unsigned char* mem = (unsigned char*)(keys + max_n);
  unsigned long long* pairs =
      (unsigned long long*)(dyn + (((size_t)max_n * 5 + 15) & ~(size_t)15));
  __shared__ int hist[256];
  __shared__ int scan_sh[8];
  __shared__ int sel_bin, sel_rem;

  const int p = rows[2 * blockIdx.x], i = rows[2 * blockIdx.x + 1];
  const memi_pd_desc P = pds[p];
  const int n = P.n;
  const int k = min(P.k, n);
  const int W = (n + 31) / 32;
  const float* Drow = D + P.d_off + (int64_t)i * n;
  for (int j = threadIdx.x; j < n; j += blockDim.x)
    keys[j] = (j == i) ? float_key(-INFINITY) : float_key(Drow[j]);
  __syncthreads();

  uint32_t T = 0xFFFFFFFFu;
  int need = n;
  if (k < n) {
    uint32_t prefix = 0, pm = 0;
    int rem = k;
    for (int shift = 24; shift >= 0; shift -= 8) {
      hist[threadIdx.x] = 0;
      __syncthreads();
      for (int j = threadIdx.x; j < n; j += blockDim.x) {
        const uint32_t key = keys[j];
        if ((key & pm) == prefix) atomicAdd(&hist[(key >> shift) & 255], 1);
      }
      __syncthreads();
      // the bin holding the rem-th smallest key: thread b owns bin b, an exclusive scan gives
      // the keys in bins below it, and the one bin with excl < rem <= excl + count is selected
      // (a serial walk over the 256 bins on one thread waited on one LDS read after another)
      {
        const int hb = hist[threadIdx.x];
        int tot_b;
        const int excl = block_excl_scan(hb, scan_sh, &tot_b);
        if (excl < rem && rem <= excl + hb) {
          sel_bin = threadIdx.x;
          sel_rem = rem - excl;
        }
      }
      __syncthreads();
      prefix |= (uint32_t)sel_bin << shift;
      pm |= 0xFFu << shift;
      rem = sel_rem;
      __syncthreads();
    }
    T = prefix;
    need = rem;  // elements equal to T to keep, lowest column first
  }
  // member flags: key < T, or key == T with tie rank < need (index order)
  const int chunk = (n + blockDim.x - 1) / blockDim.x;
  const int j0 = min(n, (int)threadIdx.x * chunk), j1 = min(n, j0 + chunk);
  int eq = 0;
  for (int j = j0; j < j1; ++j) eq += (keys[j] == T);
  int tot;
  int rank = block_excl_scan(eq, scan_sh, &tot);
  for (int j = j0; j < j1; ++j) {
    const uint32_t key = keys[j];
    unsigned char m = 0;
    if (k >= n || key < T) m = 1;
    else if (key == T) m = (rank++ < need) ? 1 : 0;
    mem[j] = m;
  }
  __syncthreads();
  uint32_t* mrow = mask + P.mask_off + (int64_t)i * W;
  for (int w = threadIdx.x; w < W; w += blockDim.x) {
    uint32_t bits = 0;
    for (int b = 0; b < 32; ++b) {
      const int j = w * 32 + b;
      if (j < n && mem[j]) bits |= 1u << b;
    }
    mrow[w] = bits;
  }
  if (idx_out == nullptr && val_out == nullptr) return;
  // sorted lists: compact members, bitonic sort on (key, index)
  int cnt = 0;
  for (int j = j0; j < j1; ++j) cnt += mem[j];
  int pos = block_excl_scan(cnt, scan_sh, &tot);
  for (int j = j0; j < j1; ++j)
    if (mem[j]) pairs[pos++] = ((unsigned long long)keys[j] << 32) | (unsigned)j;
  int L = 1;
  while (L < tot) L <<= 1;
  for (int e = tot + threadIdx.x; e < L; e += blockDim.x) pairs[e] = ~0ULL;
  __syncthreads();
  for (int size = 2; size <= L; size <<= 1) {
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int e = threadIdx.x; e < L; e += blockDim.x) {
        const int o = e ^ stride;
        if (o > e) {
          const bool up = ((e & size) == 0);
          const unsigned long long a = pairs[e], b = pairs[o];
          if ((a > b) == up) {
            pairs[e] = b;
            pairs[o] = a;
          }
        }
      }
      __syncthreads();
    }
  }
  for (int r = threadIdx.x; r < k; r += blockDim.x) {
    const int j = (int)(pairs[r] & 0xFFFFFFFFu);
    if (idx_out) idx_out[P.list_off + (int64_t)r * n + i] = j;
    if (val_out) val_out[P.list_off + (int64_t)r * n + i] = (r == 0) ? 0.f : Drow[j];
  }
}

// ------------------------------------------------------------------------ Ferguson ------
// thr pass: per PD, count of entries and sum exp(-v / (2 max eps))  -> acc[pd][2]
__global__ __launch_bounds__(256) void k_ferguson_thr(const int32_t* __restrict__ ftiles,
                                                      const memi_pd_desc* __restrict__ pds,
                                                      const float* __restrict__ D,
                                                      const uint32_t* __restrict__ mask,
                                                      const double* __restrict__ logEps,
                                                      int n_eps, double* __restrict__ acc) {
  __shared__ double red[8];
  __shared__ double mx_sh[4];
  // fergusonE.py:27: d = D2 / (2 max(eps))
  double mx = -INFINITY;
  for (int e = threadIdx.x; e < n_eps; e += blockDim.x) mx = fmax(mx, logEps[e]);
  for (int off = 32; off > 0; off >>= 1) mx = fmax(mx, __shfl_down(mx, off, 64));
  if ((threadIdx.x & 63) == 0) mx_sh[threadIdx.x >> 6] = mx;
  __syncthreads();
  mx = fmax(fmax(mx_sh[0], mx_sh[1]), fmax(mx_sh[2], mx_sh[3]));
  const double cmin = 1.0 / (2.0 * exp(mx));
  const int32_t* f = ftiles + 4 * (int64_t)blockIdx.x;
  const memi_pd_desc P = pds[f[0]];
  const int W = (P.n + 31) / 32;
  const bool full = P.k >= P.n;
  double s = 0.0, c = 0.0;
  for (int e = threadIdx.x; e < TB * TB; e += blockDim.x) {
    const int i = f[1] * TB + e / TB, j = f[2] * TB + e % TB;
    if (i >= P.n || j >= P.n) continue;
    const double v = entry_value(D, mask, P, W, full, i, j);
    if (v < 0) continue;
    s += exp(-v * cmin);
    c += 1.0;
  }
  s = block_sum(s, red);
  c = block_sum(c, red);
  if (threadIdx.x == 0) {
    atomicAdd(&acc[2 * f[0]], c);
    atomicAdd(&acc[2 * f[0] + 1], s);
  }
}

// sweep: block = (256 eps values) x (one full tile); entries staged through LDS
__global__ __launch_bounds__(256) void k_ferguson_sweep(const int32_t* __restrict__ ftiles,
                                                        const memi_pd_desc* __restrict__ pds,
                                                        const float* __restrict__ D,
                                                        const uint32_t* __restrict__ mask,
                                                        const double* __restrict__ logEps,
                                                        int n_eps,
                                                        const double* __restrict__ thr_acc,
                                                        int thr_stride,
                                                        const double* __restrict__ gate,
                                                        double* __restrict__ out) {
  constexpr int SUB = 2048;
  __shared__ float vals[SUB];
  __shared__ int scan_sh[8];
  const int32_t* f = ftiles + 4 * (int64_t)blockIdx.y;
  const int pd = f[0];
  // gate (the moments path's grid parameters): only PDs marked for the direct sweep
  if (gate != nullptr && gate[FM_GP * (int64_t)pd + 7] == 0.0) return;
  const memi_pd_desc P = pds[pd];
  const int W = (P.n + 31) / 32;
  const bool full = P.k >= P.n;
  const int e_idx = blockIdx.x * blockDim.x + threadIdx.x;
  const bool active = e_idx < n_eps;
  // fergusonE.py:25-31: thr = max(-log(0.01 * ss / len), 10)
  const double thr = fmax(-log(0.01 * thr_acc[(int64_t)thr_stride * pd + 1] /
                              thr_acc[(int64_t)thr_stride * pd]), 10.0);
  const double c = active ? 1.0 / (2.0 * exp(logEps[e_idx])) : 0.0;
  const double L2E = 1.4426950408889634;
  double S = 0.0, SX = 0.0, CN = 0.0;
  for (int base = 0; base < TB * TB; base += SUB) {
    // stage this thread's 8 entries, compact the kernel entries
    float v8[SUB / 256];
    int cnt = 0;
#pragma unroll
    for (int q = 0; q < SUB / 256; ++q) {
      const int e = base + threadIdx.x * (SUB / 256) + q;
      const int i = f[1] * TB + e / TB, j = f[2] * TB + e % TB;
      double v = -1.0;
      if (i < P.n && j < P.n) v = entry_value(D, mask, P, W, full, i, j);
      v8[q] = (float)v;
      cnt += (v >= 0);
    }
    int tot;
    int pos = block_excl_scan(cnt, scan_sh, &tot);
#pragma unroll
    for (int q = 0; q < SUB / 256; ++q)
      if (v8[q] >= 0.f) vals[pos++] = v8[q];
    __syncthreads();
    if (active) {
      for (int t = 0; t < tot; ++t) {
        const double x = (double)vals[t] * c;
        if (x < thr) {
          // exp(-x) = 2^-(y) with y = x log2(e) split into integer and fraction in float64
          const double y = x * L2E;
          const double fl = floor(y);
          const float r = exp2f(-(float)(y - fl));
          const double ex = ldexp((double)r, -(int)fl);
          S += ex;
          SX += x * ex;
          CN += 1.0;
        }
      }
    }
    __syncthreads();
  }
  if (active) {
    double* o = out + ((int64_t)pd * n_eps + e_idx) * 3;
    atomicAdd(o + 0, S);
    atomicAdd(o + 1, SX);
    atomicAdd(o + 2, CN);
  }
}

// ---------------------------------------------------- Ferguson sweep by bucket moments ----
// For a log-uniform eps grid (logEps[e] = le0 + e*dle) the threshold cuts of all eps,
// v < 2 thr eps_e, are log-uniform too.  Positive entry values go into log-uniform buckets
// whose edges E_k = exp(base + k w), base = log(2 thr) + le0, w = dle / s, contain every cut
// (the cut of eps_e is E_{e s}), so bucket k is wholly inside (k < e s) or outside the sum of
// eps_e.  Per bucket, with centre c_k and half-width h_k, the power sums M_m = sum u^m of
// u = (v - c_k) / h_k (m < FM_NM) give, for x = v / (2 eps) and t = h_k / (2 eps),
//   sum exp(-x)   = exp(-c_k / (2 eps)) sum_m (-t)^m M_m / m!,
//   sum x exp(-x) = exp(-c_k / (2 eps)) sum_m (-t)^m (c_k/(2 eps) M_m + t M_{m+1}) / m!,
// truncated at m = 16: the terms fall as t^m / m!, so t <= thr * (e^{w/2} - 1) <= FM_TMAX = 1
// leaves < 3e-15 relative.  s is the largest of 1..8 that keeps the PD's value range within
// FM_BMAX buckets.  Zero entries (v = 0, x = 0) count 1 each.  Results equal the direct
// sweep to ~1e-14 relative.  A PD whose values span more than FM_BMAX buckets of w = dle (a
// fine grid over many decades) or whose w gives t > FM_TMAX (a coarse grid) is marked in its
// grid parameters and summed by the direct sweep instead, in the same launch sequence.
constexpr int FM_BMAX = 544;  // buckets per PD: ln(3.4e38 / 1e-6) / 0.2 + 2 < 544
constexpr int FM_NM = 18;     // power sums M_0 .. M_17
constexpr double FM_TMAX = 1.0;

__device__ inline unsigned long long dbits(double v) { return (unsigned long long)__double_as_longlong(v); }

// per PD: #entries, the threshold sum, #zero entries, min and max positive value, summed into
// FER_SLOTS slots of 5 per PD (the block index modulo FER_SLOTS) at the start of the PD's moment table
// (zeroed, and zeroed again by k_fer_grid once it has summed the slots): one set of five
// counters for the hundreds of tiles of one large PD serialised their atomics (~60 us at
// n = 2000).  The minimum is kept as the maximum of the complemented bit pattern, so that a
// zeroed slot is neutral for both.
constexpr int FER_SLOTS = 16;
static_assert(FER_SLOTS * 5 <= FM_BMAX * FM_NM && FER_SLOTS <= 64 &&
                  (FER_SLOTS & (FER_SLOTS - 1)) == 0, "slots: a power of two, in a wave");

__global__ __launch_bounds__(256) void k_fer_stats(const int32_t* __restrict__ ftiles,
                                                   const memi_pd_desc* __restrict__ pds,
                                                   const float* __restrict__ D,
                                                   const uint32_t* __restrict__ mask,
                                                   const double* __restrict__ logEps, int n_eps,
                                                   double* __restrict__ slots) {
  __shared__ double red[8];
  __shared__ double mx_sh[4];
  __shared__ unsigned long long mm_sh[8];
  double mx = -INFINITY;
  for (int e = threadIdx.x; e < n_eps; e += blockDim.x) mx = fmax(mx, logEps[e]);
  for (int off = 32; off > 0; off >>= 1) mx = fmax(mx, __shfl_down(mx, off, 64));
  if ((threadIdx.x & 63) == 0) mx_sh[threadIdx.x >> 6] = mx;
  __syncthreads();
  mx = fmax(fmax(mx_sh[0], mx_sh[1]), fmax(mx_sh[2], mx_sh[3]));
  const double cmin = 1.0 / (2.0 * exp(mx));
  const int32_t* f = ftiles + 4 * (int64_t)blockIdx.x;
  const memi_pd_desc P = pds[f[0]];
  const int W = (P.n + 31) / 32;
  const bool full = P.k >= P.n;
  __shared__ TileMaskT mt;
  load_tile_mask_t(mt, mask, P, W, full, f[1], f[2]);
  double sum = 0.0, cnt = 0.0, zer = 0.0;
  unsigned long long vmin = dbits(INFINITY), vmax = 0ull;
  // blockIdx.y: which gridDim.y-th of the tile's entries
  const int part = TB * TB / gridDim.y;
  const int e0 = blockIdx.y * part, e1 = e0 + part;
  // EU entries per thread per round, their loads issued together: with one block per tile a
  // one-PD batch has a block per CU, and a load-then-exp loop one entry at a time waited a
  // memory latency per entry
  constexpr int EU = 8;
  for (int eb = e0; eb < e1; eb += EU * 256) {
    double v[EU];
#pragma unroll
    for (int u = 0; u < EU; ++u) {
      const int e = eb + u * 256 + threadIdx.x;
      const int il = e / TB, jl = e % TB;
      const int i = f[1] * TB + il, j = f[2] * TB + jl;
      v[u] = (e < e1 && i < P.n && j < P.n)
                 ? entry_value_t(D, mask, P, W, full, i, j, mt, il, jl) : -1.0;
    }
#pragma unroll
    for (int u = 0; u < EU; ++u) {
      if (v[u] < 0) continue;
      sum += exp(-v[u] * cmin);
      cnt += 1.0;
      if (v[u] == 0.0) {
        zer += 1.0;
      } else {
        vmin = min(vmin, dbits(v[u]));
        vmax = max(vmax, dbits(v[u]));
      }
    }
  }
  sum = block_sum(sum, red);
  cnt = block_sum(cnt, red);
  zer = block_sum(zer, red);
  for (int off = 32; off > 0; off >>= 1) {
    vmin = min(vmin, (unsigned long long)__shfl_down((long long)vmin, off, 64));
    vmax = max(vmax, (unsigned long long)__shfl_down((long long)vmax, off, 64));
  }
  if ((threadIdx.x & 63) == 0) {
    mm_sh[threadIdx.x >> 6] = vmin;
    mm_sh[4 + (threadIdx.x >> 6)] = vmax;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < 4; ++w) {
      vmin = min(vmin, mm_sh[w]);
      vmax = max(vmax, mm_sh[4 + w]);
    }
    double* sl = slots + (int64_t)f[0] * FM_BMAX * FM_NM +
                 ((blockIdx.x * gridDim.y + blockIdx.y) % FER_SLOTS) * 5;
    atomicAdd(sl + 0, cnt);
    atomicAdd(sl + 1, sum);
    atomicAdd(sl + 2, zer);
    atomicMax((unsigned long long*)(sl + 3), ~vmin);
    atomicMax((unsigned long long*)(sl + 4), vmax);
  }
}

__device__ inline int fm_bucket(double v, double base, double w) {
  int k = (int)floor((log(v) - base) / w);
  if (v < exp(base + k * w)) --k;
  else if (v >= exp(base + (k + 1) * w)) ++k;
  return k;
}

// per PD (one thread each): thr and the bucket grid {thr, w, base, b0, B, s, zeros, -}
__global__ __launch_bounds__(256) void k_fer_grid(int64_t npd, double* __restrict__ st,
                                                  double* __restrict__ slots, double le0,
                                                  double dle, double* __restrict__ gp) {
  // one wave per PD: lane q < FER_SLOTS reads (and clears, for the moment sums) slot q of
  // k_fer_stats; the sums run in one fixed order (a shuffle tree), lane 0 does the rest
  const int64_t p = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (p >= npd) return;                               // uniform over the wave
  double* sl = slots + p * FM_BMAX * FM_NM + 5 * lane;
  double cnt = 0.0, sum = 0.0, zer = 0.0;
  unsigned long long cmin = 0ull, vmx = 0ull;
  if (lane < FER_SLOTS) {
    cnt = sl[0];
    sum = sl[1];
    zer = sl[2];
    cmin = dbits(sl[3]);
    vmx = dbits(sl[4]);
    for (int q = 0; q < 5; ++q) sl[q] = 0.0;
  }
  for (int off = FER_SLOTS / 2; off > 0; off >>= 1) {
    cnt += __shfl_xor(cnt, off, 64);
    sum += __shfl_xor(sum, off, 64);
    zer += __shfl_xor(zer, off, 64);
    cmin = max(cmin, (unsigned long long)__shfl_xor((long long)cmin, off, 64));
    vmx = max(vmx, (unsigned long long)__shfl_xor((long long)vmx, off, 64));
  }
  if (lane != 0) return;
  double* sp = st + 6 * p;
  const double vmin = __longlong_as_double(cmin ? (long long)~cmin : 0x7FF0000000000000LL);
  const double vmax = __longlong_as_double((long long)vmx);
  sp[0] = cnt;                                            // entries
  sp[1] = sum;                                            // sum exp(-v / (2 max eps))
  sp[2] = zer;                                            // zero entries
  sp[3] = vmin;                                           // min positive v; +inf: none
  sp[4] = vmax;                                           // max v
  sp[5] = 0.0;
  double* g = gp + FM_GP * p;
  // fergusonE.py:25-31
  const double thr = fmax(-log(0.01 * sum / cnt), 10.0);
  int B = 0, s = 1, b0 = 0;
  double w = dle, base = log(2.0 * thr) + le0;
  bool direct = false;
  if (vmax > 0.0 && vmin <= vmax) {
    const double range = fmax(log(vmax) - log(vmin), 1e-300);
    s = (int)fmin(8.0, fmax(1.0, floor((FM_BMAX - 32) * dle / range)));
    w = dle / s;
    // buckets needed, in float64 before any int conversion (the +2 covers the edge buckets)
    const double need = range / w + 2.0;
    direct = !(need <= (double)FM_BMAX) || !(thr * (exp(0.5 * w) - 1.0) <= FM_TMAX);
    if (!direct) {
      b0 = fm_bucket(vmin, base, w);
      B = fm_bucket(vmax, base, w) - b0 + 1;
      B = min(max(B, 0), FM_BMAX);    // bounds the tables for any input (NaN, inf)
    }
  }
  g[0] = thr;
  g[1] = w;
  g[2] = base;
  g[3] = b0;
  g[4] = B;
  g[5] = s;
  g[6] = zer;
  g[7] = direct ? 1.0 : 0.0;          // 1: summed by the direct sweep
}

// power sums of every positive entry into its bucket (LDS accumulation, then one atomic
// per nonzero sum); M: float64 [npd][FM_BMAX][FM_NM], zeroed by the caller
__global__ __launch_bounds__(256) void k_fer_moments(const int32_t* __restrict__ ftiles,
                                                     const memi_pd_desc* __restrict__ pds,
                                                     const float* __restrict__ D,
                                                     const uint32_t* __restrict__ mask,
                                                     const double* __restrict__ gp,
                                                     double* __restrict__ M) {
  __shared__ double ms[FM_BMAX * FM_NM];
  const int32_t* f = ftiles + 4 * (int64_t)blockIdx.x;
  const int pd = f[0];
  const memi_pd_desc P = pds[pd];
  const double* g = gp + FM_GP * (int64_t)pd;
  const double w = g[1], base = g[2];
  const int b0 = (int)g[3], B = (int)g[4];
  if (B <= 0) return;                                 // uniform over the block
  for (int q = threadIdx.x; q < B * FM_NM; q += blockDim.x) ms[q] = 0.0;
  const int W = (P.n + 31) / 32;
  const bool full = P.k >= P.n;
  __shared__ TileMaskT mt;
  load_tile_mask_t(mt, mask, P, W, full, f[1], f[2]);   // and the barrier after the zeroing
  // blockIdx.y: which FM_RQ-th of the tile's entries
  const int e1 = (blockIdx.y + 1) * (TB * TB / FM_RQ);
  for (int e = blockIdx.y * (TB * TB / FM_RQ) + threadIdx.x; e < e1; e += blockDim.x) {
    const int il = e / TB, jl = e % TB;
    const int i = f[1] * TB + il, j = f[2] * TB + jl;
    if (i >= P.n || j >= P.n) continue;
    const double v = entry_value_t(D, mask, P, W, full, i, j, mt, il, jl);
    if (v <= 0.0) continue;
    const int k = min(max(fm_bucket(v, base, w) - b0, 0), B - 1);
    const double lo = exp(base + (b0 + k) * w), hi = exp(base + (b0 + k + 1) * w);
    const double u = (v - 0.5 * (lo + hi)) / (0.5 * (hi - lo));
    double pw = 1.0;
    for (int m = 0; m < FM_NM; ++m) {
      atomicAdd(&ms[k * FM_NM + m], pw);
      pw *= u;
    }
  }
  __syncthreads();
  double* Mp = M + (int64_t)pd * FM_BMAX * FM_NM;
  for (int q = threadIdx.x; q < B * FM_NM; q += blockDim.x)
    if (ms[q] != 0.0) atomicAdd(&Mp[q], ms[q]);
}

__global__ void k_fer_thr_copy(int64_t npd, const double* __restrict__ st,
                               double* __restrict__ dst) {
  for (int64_t p = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; p < npd;
       p += (int64_t)gridDim.x * blockDim.x) {
    dst[2 * p] = st[6 * p];
    dst[2 * p + 1] = st[6 * p + 1];
  }
}

// one thread per eps: the series over the buckets below the eps's cut
__global__ __launch_bounds__(256) void k_fer_sweep_moments(const double* __restrict__ logEps,
                                                           int n_eps,
                                                           const double* __restrict__ gp,
                                                           const double* __restrict__ M,
                                                           double* __restrict__ out) {
  __shared__ double cen[FM_BMAX], hw[FM_BMAX];
  const int pd = blockIdx.y;
  const double* g = gp + FM_GP * (int64_t)pd;
  const double w = g[1], base = g[2];
  const int b0 = (int)g[3], B = (int)g[4], s = (int)g[5];
  const double zeros = g[6];
  for (int k = threadIdx.x; k < B; k += blockDim.x) {
    const double lo = exp(base + (b0 + k) * w), hi = exp(base + (b0 + k + 1) * w);
    cen[k] = 0.5 * (lo + hi);
    hw[k] = 0.5 * (hi - lo);
  }
  __syncthreads();
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n_eps) return;
  if (g[7] != 0.0) {                  // direct-sweep PD: cleared for its accumulation
    double* o = out + ((int64_t)pd * n_eps + e) * 3;
    o[0] = o[1] = o[2] = 0.0;
    return;
  }
  constexpr double invf[FM_NM - 1] = {1.0, 1.0, 0.5, 1.0 / 6, 1.0 / 24, 1.0 / 120, 1.0 / 720,
                                      1.0 / 5040, 1.0 / 40320, 1.0 / 362880, 1.0 / 3628800,
                                      1.0 / 39916800, 1.0 / 479001600, 1.0 / 6227020800.0,
                                      1.0 / 87178291200.0, 1.0 / 1307674368000.0,
                                      1.0 / 20922789888000.0};
  const double c = 1.0 / (2.0 * exp(logEps[e]));
  const int jmax = min(B, max(0, e * s - b0));
  const double* Mp = M + (int64_t)pd * FM_BMAX * FM_NM;
  double S = zeros, SX = 0.0, CN = zeros;
  for (int k = 0; k < jmax; ++k) {
    const double xb = cen[k] * c;
    if (xb > 800.0) break;                            // later buckets underflow too
    const double t = -hw[k] * c;
    const double* m = Mp + (int64_t)k * FM_NM;
    double s0 = m[FM_NM - 2] * invf[FM_NM - 2], s1 = m[FM_NM - 1] * invf[FM_NM - 2];
    for (int q = FM_NM - 3; q >= 0; --q) {
      s0 = s0 * t + m[q] * invf[q];
      s1 = s1 * t + m[q + 1] * invf[q];
    }
    const double ex = exp(-xb);
    S += ex * s0;
    SX += ex * (xb * s0 - t * s1);
    CN += m[0];
  }
  double* o = out + ((int64_t)pd * n_eps + e) * 3;
  o[0] = S;
  o[1] = SX;
  o[2] = CN;
}

// ------------------------------------------------------------------ normalisation ------
__device__ inline double kernel_weight(double v, double inv_s2) {
  return v < 0 ? 0.0 : (v == 0.0 ? 1.0 : exp(-v * inv_s2));
}

// pass 0: deg1[j] += sum_i W_ij ; pass 1: deg2[j] += sum_i W_ij / deg1[i]
template <int PASS>
__global__ __launch_bounds__(256) void k_markov_colsum(const int32_t* __restrict__ ftiles,
                                                       const memi_pd_desc* __restrict__ pds,
                                                       const float* __restrict__ D,
                                                       const uint32_t* __restrict__ mask,
                                                       const double* __restrict__ sigma,
                                                       double* __restrict__ deg,
                                                       int64_t n_rows) {
  const int32_t* f = ftiles + 4 * (int64_t)blockIdx.x;
  const int pd = f[0];
  const memi_pd_desc P = pds[pd];
  const int W = (P.n + 31) / 32;
  const bool full = P.k >= P.n;
  const double inv_s2 = 1.0 / (sigma[pd] * sigma[pd]);
  const int cl = threadIdx.x & (TB - 1), half = threadIdx.x >> 7;
  const int j = f[2] * TB + cl;
  __shared__ TileMaskT mt;
  load_tile_mask_t(mt, mask, P, W, full, f[1], f[2]);
  if (j >= P.n) return;
  double s = 0.0;
  // blockIdx.y: which MARKOV_RQ-th of the tile's rows (several blocks per tile keep the CUs
  // busy when a batch has few tiles)
  const int r1 = (blockIdx.y + 1) * (TB / MARKOV_RQ);
  for (int r = blockIdx.y * (TB / MARKOV_RQ) + half; r < r1; r += 2) {
    const int i = f[1] * TB + r;
    if (i >= P.n) break;
    const double w =
        kernel_weight(entry_value_t(D, mask, P, W, full, i, j, mt, r, cl), inv_s2);
    s += (PASS == 0) ? w : w / deg[P.row0 + i];
  }
  atomicAdd(&deg[(PASS == 0 ? 0 : n_rows) + P.row0 + j], s);
}

template <typename LT>
__global__ __launch_bounds__(256) void k_markov_write(const int32_t* __restrict__ ftiles,
                                                      const memi_pd_desc* __restrict__ pds,
                                                      const float* __restrict__ D,
                                                      const uint32_t* __restrict__ mask,
                                                      const double* __restrict__ sigma,
                                                      const double* __restrict__ deg,
                                                      int64_t n_rows, LT* __restrict__ L) {
  const int32_t* f = ftiles + 4 * (int64_t)blockIdx.x;
  const int pd = f[0];
  const memi_pd_desc P = pds[pd];
  const int W = (P.n + 31) / 32;
  const bool full = P.k >= P.n;
  const double inv_s2 = 1.0 / (sigma[pd] * sigma[pd]);
  const int cl = threadIdx.x & (TB - 1), half = threadIdx.x >> 7;
  const int j = f[2] * TB + cl;
  __shared__ TileMaskT mt;
  load_tile_mask_t(mt, mask, P, W, full, f[1], f[2]);
  if (j >= P.n) return;
  const double dj = deg[P.row0 + j];
  const double dpj = sqrt(deg[n_rows + P.row0 + j] / dj);
  const int r1 = (blockIdx.y + 1) * (TB / MARKOV_RQ);
  for (int r = blockIdx.y * (TB / MARKOV_RQ) + half; r < r1; r += 2) {
    const int i = f[1] * TB + r;
    if (i >= P.n) break;
    const double di = deg[P.row0 + i];
    const double dpi = sqrt(deg[n_rows + P.row0 + i] / di);
    const float dij = D[P.d_off + (int64_t)i * P.n + j];
    const bool mij = full || mbit(mask, P.mask_off, W, i, j);
    const bool mji = full || tile_mji(mt, r, cl);
    const double wij = kernel_weight(entry_from(dij, mij, mji, i == j), inv_s2);
    const double wji = kernel_weight(entry_from(dij, mji, mij, i == j), inv_s2);
    // slaplacianonFly.py:63-78: W/(d_i d_j), then /(d'_i d'_j), then |l + l^T| / 2
    const double lij = wij / (di * dj) / (dpi * dpj);
    const double lji = wji / (dj * di) / (dpj * dpi);
    // LT = double keeps the float64 kernel the eigensolve is specified on; LT = float halves
    // the bytes of every Lanczos matvec at ~6e-8 relative rounding of each entry
    L[P.d_off + (int64_t)i * P.n + j] = (LT)(fabs(lij + lji) / 2.0);
  }
}

// ------------------------------------------------------------------ tanh fit (a13) ----
// Levenberg-Marquardt fit of f(x) = a3 + a2 tanh(a0 x + a1) to y (fergusonE.py:45-57), one
// block per PD, float64.  Scaled LM in the manner of MINPACK lmdif (diagonal scaling by the
// running max column norm of J, ftol = xtol = 1.49012e-8), analytic Jacobian.  Returns popt,
// J^T J at popt (for pcov = inv(J^T J) ss_res / (m - 4)), ss_res and the iteration count.
__device__ inline void tanh_terms(double x, const double a[4], double& f, double J[4]) {
  const double u = a[0] * x + a[1];
  const double t = tanh(u);
  const double s2 = 1.0 - t * t;
  f = a[3] + a[2] * t;
  J[0] = a[2] * s2 * x;
  J[1] = a[2] * s2;
  J[2] = t;
  J[3] = 1.0;
}

// 256 threads per PD evaluate the curve (~6 points each).  Each wave sums its 15 values by
// DPP, lane 0 stores them, and after one barrier every thread adds the wave sums in wave order,
// so all threads hold the same totals and run the LM bookkeeping (scaling, the damped 4x4
// Cholesky solve, accept / reject) themselves, in registers: no serial section on one thread
// and no state in LDS.  The wave sums alternate between two LDS buffers, so a buffer is
// rewritten only after the barrier of the next evaluation has passed, and one barrier per
// evaluation suffices.
constexpr int TANH_NT = 256;
constexpr int TANH_NW = TANH_NT / 64;

struct LmState {
  double a[4], at[4], diag[4], d[4];
  double jtj[10], jtr[4];          // at a
  double cost, lam;
  int it, done;
};

// cost, J^T J and J^T r at the point pt, summed over the curve into v (every thread gets the
// same totals); sh: this evaluation's buffer of wave sums
__device__ __forceinline__ void tanh_eval(const double* __restrict__ x,
                                          const double* __restrict__ yp, int m,
                                          const double pt[4], double v[15],
                                          double (*sh)[15]) {
#pragma unroll
  for (int q = 0; q < 15; ++q) v[q] = 0.0;
  for (int i = threadIdx.x; i < m; i += TANH_NT) {
    double f, J[4];
    tanh_terms(x[i], pt, f, J);
    const double r = yp[i] - f;
    v[14] += r * r;
    int k = 0;
#pragma unroll
    for (int p = 0; p < 4; ++p)
#pragma unroll
      for (int q = p; q < 4; ++q) v[k++] += J[p] * J[q];
#pragma unroll
    for (int p = 0; p < 4; ++p) v[10 + p] += J[p] * r;
  }
  const int wid = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < 15; ++q) v[q] = wave_total(v[q]);
  if ((threadIdx.x & 63) == 0)
#pragma unroll
    for (int q = 0; q < 15; ++q) sh[wid][q] = v[q];
  __syncthreads();
#pragma unroll
  for (int q = 0; q < 15; ++q) {
    double t = sh[0][q];
#pragma unroll
    for (int w = 1; w < TANH_NW; ++w) t += sh[w][q];
    v[q] = t;
  }
}

// scaling update and the damped step from the current point into st.at (more damping while
// the system is not positive definite); st.done when the fit must stop
__device__ inline void lm_propose(LmState& st, int max_iter) {
  const int dg[4] = {0, 4, 7, 9};  // packed (p, p)
  for (;;) {
    ++st.it;
#pragma unroll
    for (int p = 0; p < 4; ++p) st.diag[p] = fmax(st.diag[p], sqrt(st.jtj[dg[p]]));
    double A[4][4];
    {
      int k = 0;
#pragma unroll
      for (int p = 0; p < 4; ++p)
#pragma unroll
        for (int q = p; q < 4; ++q) {
          A[p][q] = st.jtj[k];
          A[q][p] = st.jtj[k];
          ++k;
        }
    }
#pragma unroll
    for (int p = 0; p < 4; ++p) A[p][p] += st.lam * st.diag[p] * st.diag[p];
    double Lc[4][4];
    bool ok = true;
#pragma unroll
    for (int p = 0; p < 4; ++p) {
#pragma unroll
      for (int q = 0; q <= p; ++q) {
        double sm = A[p][q];
#pragma unroll
        for (int r = 0; r < q; ++r) sm -= Lc[p][r] * Lc[q][r];
        if (p == q) {
          ok = ok && sm > 0.0;
          Lc[p][p] = sqrt(fmax(sm, 1e-300));
        } else {
          Lc[p][q] = sm / Lc[q][q];
        }
      }
    }
    if (!ok) {                     // not positive definite: more damping, same point
      st.lam *= 10.0;
      if (st.lam > 1e32 || st.it >= max_iter) {
        st.done = 1;
        return;
      }
      continue;
    }
    double z[4];
#pragma unroll
    for (int p = 0; p < 4; ++p) {
      double sm = st.jtr[p];
#pragma unroll
      for (int r = 0; r < p; ++r) sm -= Lc[p][r] * z[r];
      z[p] = sm / Lc[p][p];
    }
#pragma unroll
    for (int p = 3; p >= 0; --p) {
      double sm = z[p];
#pragma unroll
      for (int r = p + 1; r < 4; ++r) sm -= Lc[r][p] * st.d[r];
      st.d[p] = sm / Lc[p][p];
    }
#pragma unroll
    for (int p = 0; p < 4; ++p) st.at[p] = st.a[p] + st.d[p];
    return;
  }
}

__global__ __launch_bounds__(TANH_NT) void k_tanh_fit(const double* __restrict__ x,
                                                      const double* __restrict__ y, int m,
                                                      const double* __restrict__ p0,
                                                      const int32_t* __restrict__ active,
                                                      int max_iter, double* __restrict__ out) {
  __shared__ double sh[2][TANH_NW][15];
  const int pd = blockIdx.x;
  if (active && !active[pd]) return;
  const double* yp = y + (int64_t)pd * m;
  LmState st;
  for (int q = 0; q < 4; ++q) {
    st.a[q] = p0[pd * 4 + q];
    st.diag[q] = 0.0;
    st.d[q] = 0.0;
  }
  st.lam = 1e-3;
  st.it = 0;
  st.done = 0;
  int buf = 0;
  {
    double v[15];
    tanh_eval(x, yp, m, st.a, v, sh[buf]);
    buf ^= 1;
    for (int k = 0; k < 10; ++k) st.jtj[k] = v[k];
    for (int p = 0; p < 4; ++p) st.jtr[p] = v[10 + p];
    st.cost = v[14];
    if (max_iter <= 0) st.done = 1;
    else lm_propose(st, max_iter);
  }
  while (!st.done) {                                  // the same in every thread
    // the trial point's J^T J and J^T r come with its cost: an accepted step then needs no
    // second pass (they are the values a re-evaluation at the new point would give)
    double v[15];
    tanh_eval(x, yp, m, st.at, v, sh[buf]);
    buf ^= 1;
    const double cost_t = v[14];
    double pred = 0.0;
    for (int p = 0; p < 4; ++p)
      pred += st.d[p] * (st.jtr[p] + st.lam * st.diag[p] * st.diag[p] * st.d[p]);
    if (cost_t < st.cost) {
      const double actred = (st.cost - cost_t) / fmax(st.cost, 1e-300);
      double xn = 0.0, dn = 0.0;
      for (int p = 0; p < 4; ++p) {
        xn += (st.diag[p] * st.at[p]) * (st.diag[p] * st.at[p]);
        dn += (st.diag[p] * st.d[p]) * (st.diag[p] * st.d[p]);
      }
      for (int p = 0; p < 4; ++p) st.a[p] = st.at[p];
      st.cost = cost_t;
      for (int k = 0; k < 10; ++k) st.jtj[k] = v[k];
      for (int p = 0; p < 4; ++p) st.jtr[p] = v[10 + p];
      st.lam = fmax(st.lam * 0.1, 1e-15);
      const double prered = pred / fmax(cost_t + pred, 1e-300);
      if ((actred <= 1.49012e-8 && prered <= 1.49012e-8) || sqrt(dn) <= 1.49012e-8 * sqrt(xn))
        st.done = 1;
    } else {
      st.lam *= 10.0;
      if (st.lam > 1e32) st.done = 1;
    }
    if (!st.done && st.it >= max_iter) st.done = 1;
    if (!st.done) lm_propose(st, max_iter);
  }
  if (threadIdx.x == 0) {
    double* o = out + (int64_t)pd * 16;
    for (int p = 0; p < 4; ++p) o[p] = st.a[p];
    for (int k = 0; k < 10; ++k) o[4 + k] = st.jtj[k];
    o[14] = st.cost;
    o[15] = st.it;
  }
}

// ------------------------------------------------- list forms (fergusonE / slaplacian) ----
// fergusonE.op(D, logEps, a0) on a flat list: D2 = D*D; thr pass then the eps sweep.
__global__ __launch_bounds__(256) void k_ferguson_values_thr(const double* __restrict__ v,
                                                             int64_t nv,
                                                             const double* __restrict__ logEps,
                                                             int n_eps, double* __restrict__ acc) {
  __shared__ double red[8];
  double mx = -INFINITY;
  for (int e = threadIdx.x; e < n_eps; e += blockDim.x) mx = fmax(mx, logEps[e]);
  __shared__ double mx_sh[4];
  for (int off = 32; off > 0; off >>= 1) mx = fmax(mx, __shfl_down(mx, off, 64));
  if ((threadIdx.x & 63) == 0) mx_sh[threadIdx.x >> 6] = mx;
  __syncthreads();
  mx = fmax(fmax(mx_sh[0], mx_sh[1]), fmax(mx_sh[2], mx_sh[3]));
  const double cmin = 1.0 / (2.0 * exp(mx));
  double s = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nv;
       i += (int64_t)gridDim.x * blockDim.x)
    s += exp(-(v[i] * v[i]) * cmin);
  s = block_sum(s, red);
  if (threadIdx.x == 0) atomicAdd(acc, s);
}

__global__ __launch_bounds__(256) void k_ferguson_values(const double* __restrict__ v, int64_t nv,
                                                         const double* __restrict__ logEps,
                                                         int n_eps, const double* __restrict__ ss,
                                                         double* __restrict__ out) {
  constexpr int CH = 2048;
  __shared__ double vals[CH];
  const int e_idx = blockIdx.x * blockDim.x + threadIdx.x;
  const bool active = e_idx < n_eps;
  const double thr = fmax(-log(0.01 * ss[0] / (double)nv), 10.0);
  const double c = active ? 1.0 / (2.0 * exp(logEps[e_idx])) : 0.0;
  const int64_t b0 = (int64_t)blockIdx.y * CH;
  const int cnt = (int)min((int64_t)CH, nv - b0);
  for (int t = threadIdx.x; t < cnt; t += blockDim.x) {
    const double d = v[b0 + t];
    vals[t] = d * d;
  }
  __syncthreads();
  if (!active) return;
  double S = 0.0, SX = 0.0, CN = 0.0;
  for (int t = 0; t < cnt; ++t) {
    const double x = vals[t] * c;
    if (x < thr) {
      const double ex = exp(-x);
      S += ex;
      SX += x * ex;
      CN += 1.0;
    }
  }
  atomicAdd(out + e_idx * 3 + 0, S);
  atomicAdd(out + e_idx * 3 + 1, SX);
  atomicAdd(out + e_idx * 3 + 2, CN);
}

// slaplacianonFly.op on COO triplets: W = exp(-val/sigma^2); d = column sums (^alpha);
// W' = W/(d_r d_c); d' = sqrt(column sums of W'); out = W'/(d'_r d'_c).
template <int PASS>
__global__ void k_coo_norm(const int64_t* __restrict__ row, const int64_t* __restrict__ col,
                           const double* __restrict__ val, int64_t nnz, double inv_s2,
                           double alpha, double* __restrict__ deg, int64_t nS,
                           double* __restrict__ out) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < nnz;
       e += (int64_t)gridDim.x * blockDim.x) {
    const double w = exp(-val[e] * inv_s2);
    const int64_t r = row[e], c = col[e];
    if (PASS == 0) {
      atomicAdd(&deg[c], w);
    } else {
      double dr = deg[r], dc = deg[c];
      if (alpha != 1.0) {
        dr = pow(dr, alpha);
        dc = pow(dc, alpha);
      }
      const double w1 = w / (dr * dc);
      if (PASS == 1) {
        atomicAdd(&deg[nS + c], w1);
      } else {
        out[e] = w1 / (sqrt(deg[nS + r]) * sqrt(deg[nS + c]));
      }
    }
  }
}

__global__ void k_zero_f64(double* p, int64_t n) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * blockDim.x)
    p[e] = 0.0;
}

// Trimming rounds (manifoldTrimmingAuto.py:63-75): PD p of the destination plan is the block of
// source PD src_off[p] / src_n[p] restricted to rows and columns sel[row0_p .. row0_p + n_p).
// One block per destination row; sel is ascending, so a row's reads are near-contiguous.
__global__ __launch_bounds__(256) void k_gather_blocks(const int32_t* __restrict__ rows,
                                                       const memi_pd_desc* __restrict__ pds,
                                                       const float* __restrict__ src,
                                                       const int64_t* __restrict__ src_off,
                                                       const int32_t* __restrict__ src_n,
                                                       const int32_t* __restrict__ sel,
                                                       float* __restrict__ dst) {
  const int p = rows[2 * blockIdx.x], i = rows[2 * blockIdx.x + 1];
  const memi_pd_desc P = pds[p];
  const int32_t* s = sel + P.row0;
  const float* srow = src + src_off[p] + (int64_t)s[i] * src_n[p];
  float* drow = dst + P.d_off + (int64_t)i * P.n;
  for (int j = threadIdx.x; j < P.n; j += blockDim.x) drow[j] = srow[s[j]];
}

// The restart test of fergusonE.py:51-57 on one fit of memi_tanh_fit's output: pcov =
// inv(J^T J) ss_res / (m - 4) as scipy's curve_fit forms it, value = sum_i sqrt|pcov_ii|
// (+inf when J^T J is exactly singular or anything is not finite).  The diagonal of the inverse
// comes from an LU factorisation with partial pivoting (first largest |pivot|, as LAPACK's
// getrf) and one solve per unit vector.
__device__ inline double tanh_fit_resnorm(const double* o, int m) {
  double A[4][4];
  int k = 4;
  for (int i = 0; i < 4; ++i)
    for (int j = i; j < 4; ++j) {
      A[i][j] = o[k];
      A[j][i] = o[k];
      ++k;
    }
  int perm[4] = {0, 1, 2, 3};
  for (int c = 0; c < 4; ++c) {
    int pr = c;
    for (int r = c + 1; r < 4; ++r)
      if (fabs(A[r][c]) > fabs(A[pr][c])) pr = r;
    if (A[pr][c] == 0.0) return INFINITY;
    if (pr != c) {
      for (int j = 0; j < 4; ++j) {
        const double t = A[c][j];
        A[c][j] = A[pr][j];
        A[pr][j] = t;
      }
      const int t = perm[c];
      perm[c] = perm[pr];
      perm[pr] = t;
    }
    for (int r = c + 1; r < 4; ++r) {
      const double l = A[r][c] / A[c][c];
      A[r][c] = l;
      for (int j = c + 1; j < 4; ++j) A[r][j] -= l * A[c][j];
    }
  }
  const double scale = o[14] / (double)(m - 4);
  double sum = 0.0;
  for (int e = 0; e < 4; ++e) {
    // solve L U x = P e_e; only x_e is needed, but the back substitution needs the rest
    double y[4];
    for (int r = 0; r < 4; ++r) {
      double v = perm[r] == e ? 1.0 : 0.0;
      for (int j = 0; j < r; ++j) v -= A[r][j] * y[j];
      y[r] = v;
    }
    double x[4];
    for (int r = 3; r >= 0; --r) {
      double v = y[r];
      for (int j = r + 1; j < 4; ++j) v -= A[r][j] * x[j];
      x[r] = v / A[r][r];
    }
    sum += sqrt(fabs(x[e] * scale));
  }
  return isfinite(sum) ? sum : INFINITY;
}

// One thread per PD: of its R fits, the first whose restart value is <= max_res; ok = 1 + its
// index (0: none), sigma = tune sqrt(2 exp(-a1/a0)) (DMembeddingII.py:170) from its popt.
__global__ void k_tanh_select(const double* __restrict__ fit, int npd, int R, int m, double tune,
                              double max_res, int32_t* __restrict__ ok,
                              double* __restrict__ sigma, double* __restrict__ res) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= npd) return;
  int pick = 0;
  for (int r = 0; r < R; ++r) {
    const double* o = fit + ((int64_t)p * R + r) * 16;
    const double v = tanh_fit_resnorm(o, m);
    if (res) res[(int64_t)p * R + r] = v;
    if (!pick && v <= max_res) pick = r + 1;
  }
  ok[p] = pick;
  if (pick) {
    const double* o = fit + ((int64_t)p * R + pick - 1) * 16;
    sigma[p] = tune * sqrt(2.0 * exp(-o[1] / o[0]));
  }
}

}  // namespace memi

using namespace memi;

namespace {
struct PlanView {
  const memi_pd_desc* pds;
  const int32_t* ftiles;
  const int32_t* rows;
};
PlanView view(const void* plan, const memi_plan_info* info) {
  const char* b = (const char*)plan;
  return PlanView{(const memi_pd_desc*)(b + info->off_pds),
                  (const int32_t*)(b + info->off_ftiles), (const int32_t*)(b + info->off_rows)};
}
}  // namespace

extern "C" {

int memi_knn_select(const void* plan_dev, const memi_plan_info* info, const float* D,
                    uint32_t* mask, int32_t* idx_out, float* val_out, void* stream) {
  MEMI_CHECK_ARG(plan_dev && info && D && mask, "memi_knn_select: null pointer");
  if (info->n_rows == 0) return MEMI_OK;
  const int max_n = info->max_n;
  const bool lists = idx_out || val_out;
  int L = 1;
  while (L < max_n) L <<= 1;
  size_t shm = (((size_t)max_n * 5 + 15) & ~(size_t)15) + (lists ? (size_t)L * 8 : 0);
  MEMI_CHECK_ARG(shm <= 160 * 1024 - 2048, "memi_knn_select: n=%d too large for %s", max_n,
                 lists ? "sorted lists (n <= 8192)" : "the row buffer");
  PlanView v = view(plan_dev, info);
  hipLaunchKernelGGL(k_knn_select, dim3((unsigned)info->n_rows), dim3(256), shm,
                     (hipStream_t)stream, v.rows, v.pds, D, mask, idx_out, val_out, max_n, L);
  MEMI_CHECK_LAUNCH("k_knn_select");
  return MEMI_OK;
}

int memi_ferguson_sweep(const void* plan_dev, const memi_plan_info* info, const float* D,
                        const uint32_t* mask, const double* logEps, int n_eps, double* out,
                        void* stream) {
  // out holds npd*n_eps*3 results followed by 2*npd threshold sums (count, sum exp)
  MEMI_CHECK_ARG(plan_dev && info && D && mask && logEps && out && n_eps > 0,
                 "memi_ferguson_sweep: bad arguments");
  MEMI_CHECK_ARG(info->n_ftiles < 65536, "memi_ferguson_sweep: too many tiles (%lld)",
                 (long long)info->n_ftiles);
  if (info->n_pd == 0) return MEMI_OK;
  hipStream_t s = (hipStream_t)stream;
  PlanView v = view(plan_dev, info);
  // scratch for the threshold sums lives after the outputs: out[npd*n_eps*3 + 2*npd]
  double* thr_acc = out + info->n_pd * (int64_t)n_eps * 3;
  const int64_t nz = info->n_pd * ((int64_t)n_eps * 3 + 2);
  hipLaunchKernelGGL(k_zero_f64, dim3(grid_for(nz)), dim3(256), 0, s, out, nz);
  MEMI_CHECK_LAUNCH("k_zero_f64");
  hipLaunchKernelGGL(k_ferguson_thr, dim3((unsigned)info->n_ftiles), dim3(256), 0, s, v.ftiles,
                     v.pds, D, mask, logEps, n_eps, thr_acc);
  MEMI_CHECK_LAUNCH("k_ferguson_thr");
  dim3 grid((unsigned)((n_eps + 255) / 256), (unsigned)info->n_ftiles);
  hipLaunchKernelGGL(k_ferguson_sweep, grid, dim3(256), 0, s, v.ftiles, v.pds, D, mask, logEps,
                     n_eps, thr_acc, 2, (const double*)nullptr, out);
  MEMI_CHECK_LAUNCH("k_ferguson_sweep");
  return MEMI_OK;
}

int memi_ferguson_moments(const void* plan_dev, const memi_plan_info* info, const float* D,
                          const uint32_t* mask, const double* logEps, int n_eps, double le0,
                          double dle, double* out, double* ws, int64_t ws_doubles,
                          void* stream) {
  MEMI_CHECK_ARG(plan_dev && info && D && mask && logEps && out && ws && n_eps > 0,
                 "memi_ferguson_moments: bad arguments");
  MEMI_CHECK_ARG(dle > 0.0 && std::isfinite(le0), "memi_ferguson_moments: the eps grid must be "
                 "log-uniform with a positive step (dle = %g)", dle);
  const int64_t npd = info->n_pd;
  const int64_t need = npd * (6 + FM_GP + (int64_t)FM_BMAX * FM_NM);
  MEMI_CHECK_ARG(ws_doubles >= need, "memi_ferguson_moments: workspace %lld < %lld doubles",
                 (long long)ws_doubles, (long long)need);
  MEMI_CHECK_ARG(npd < 65536 && n_eps <= (1 << 24) && info->n_ftiles < 65536,
                 "memi_ferguson_moments: too many PDs/eps/tiles");
  if (npd == 0) return MEMI_OK;
  hipStream_t s = (hipStream_t)stream;
  PlanView v = view(plan_dev, info);
  double* st = ws;
  double* gp = st + 6 * npd;
  double* M = gp + FM_GP * npd;
  const int64_t nm = npd * (int64_t)FM_BMAX * FM_NM;
  hipLaunchKernelGGL(k_zero_f64, dim3(grid_for(nm)), dim3(256), 0, s, M, nm);
  MEMI_CHECK_LAUNCH("k_zero_f64");
  hipLaunchKernelGGL(k_fer_stats, dim3((unsigned)info->n_ftiles, FS_RQ), dim3(256), 0, s,
                     v.ftiles, v.pds, D, mask, logEps, n_eps, M);
  MEMI_CHECK_LAUNCH("k_fer_stats");
  hipLaunchKernelGGL(k_fer_grid, dim3((unsigned)((npd + 3) / 4)), dim3(256), 0, s, npd, st, M,
                     le0, dle, gp);
  MEMI_CHECK_LAUNCH("k_fer_grid");
  hipLaunchKernelGGL(k_fer_moments, dim3((unsigned)info->n_ftiles, FM_RQ), dim3(256), 0, s,
                     v.ftiles,
                     v.pds, D, mask, gp, M);
  MEMI_CHECK_LAUNCH("k_fer_moments");
  hipLaunchKernelGGL(k_fer_sweep_moments, dim3((unsigned)((n_eps + 255) / 256), (unsigned)npd),
                     dim3(256), 0, s, logEps, n_eps, gp, M, out);
  MEMI_CHECK_LAUNCH("k_fer_sweep_moments");
  // PDs whose value range or grid step the moments cannot carry (marked by k_fer_grid): the
  // direct sweep, its blocks of every other PD exiting at once
  hipLaunchKernelGGL(k_ferguson_sweep, dim3((unsigned)((n_eps + 255) / 256),
                                            (unsigned)info->n_ftiles),
                     dim3(256), 0, s, v.ftiles, v.pds, D, mask, logEps, n_eps, st, 6,
                     (const double*)gp, out);
  MEMI_CHECK_LAUNCH("k_ferguson_sweep");
  // the threshold sums after the results, as memi_ferguson_sweep leaves them
  hipLaunchKernelGGL(k_fer_thr_copy, dim3(grid_for(npd)), dim3(256), 0, s, npd, st,
                     out + npd * (int64_t)n_eps * 3);
  MEMI_CHECK_LAUNCH("k_fer_thr_copy");
  return MEMI_OK;
}

int memi_markov_normalize(const void* plan_dev, const memi_plan_info* info, const float* D,
                          const uint32_t* mask, const double* sigma, void* L, int l_is_f64,
                          double* deg, void* stream) {
  MEMI_CHECK_ARG(plan_dev && info && D && mask && sigma && L && deg,
                 "memi_markov_normalize: null pointer");
  if (info->n_pd == 0) return MEMI_OK;
  hipStream_t s = (hipStream_t)stream;
  PlanView v = view(plan_dev, info);
  const int64_t nr = info->n_rows;
  hipLaunchKernelGGL(k_zero_f64, dim3(grid_for(2 * nr)), dim3(256), 0, s, deg, 2 * nr);
  MEMI_CHECK_LAUNCH("k_zero_f64");
  const dim3 g((unsigned)info->n_ftiles, MARKOV_RQ);
  hipLaunchKernelGGL(k_markov_colsum<0>, g, dim3(256), 0, s, v.ftiles, v.pds, D, mask, sigma,
                     deg, nr);
  MEMI_CHECK_LAUNCH("k_markov_colsum<0>");
  hipLaunchKernelGGL(k_markov_colsum<1>, g, dim3(256), 0, s, v.ftiles, v.pds, D, mask, sigma,
                     deg, nr);
  MEMI_CHECK_LAUNCH("k_markov_colsum<1>");
  if (l_is_f64)
    hipLaunchKernelGGL(k_markov_write<double>, g, dim3(256), 0, s, v.ftiles, v.pds, D, mask,
                       sigma, deg, nr, (double*)L);
  else
    hipLaunchKernelGGL(k_markov_write<float>, g, dim3(256), 0, s, v.ftiles, v.pds, D, mask,
                       sigma, deg, nr, (float*)L);
  MEMI_CHECK_LAUNCH("k_markov_write");
  return MEMI_OK;
}

int memi_ferguson_values(const double* vals, int64_t nv, const double* logEps, int n_eps,
                         double* out, void* stream) {
  MEMI_CHECK_ARG(vals && logEps && out && nv > 0 && n_eps > 0,
                 "memi_ferguson_values: bad arguments");
  hipStream_t s = (hipStream_t)stream;
  const int64_t nz = (int64_t)n_eps * 3 + 1;
  hipLaunchKernelGGL(k_zero_f64, dim3(grid_for(nz)), dim3(256), 0, s, out, nz);
  MEMI_CHECK_LAUNCH("k_zero_f64");
  double* ss = out + (int64_t)n_eps * 3;
  hipLaunchKernelGGL(k_ferguson_values_thr, dim3(grid_for(nv, 256, 1024)), dim3(256), 0, s, vals,
                     nv, logEps, n_eps, ss);
  MEMI_CHECK_LAUNCH("k_ferguson_values_thr");
  dim3 g((unsigned)((n_eps + 255) / 256), (unsigned)((nv + 2047) / 2048));
  MEMI_CHECK_ARG(g.y < 65536u * 32, "memi_ferguson_values: too many values");
  hipLaunchKernelGGL(k_ferguson_values, g, dim3(256), 0, s, vals, nv, logEps, n_eps, ss, out);
  MEMI_CHECK_LAUNCH("k_ferguson_values");
  return MEMI_OK;
}

int memi_markov_coo(const int64_t* row, const int64_t* col, const double* val, int64_t nnz,
                    int64_t nS, double sigma, double alpha, double* out, double* deg,
                    void* stream) {
  MEMI_CHECK_ARG(row && col && val && out && deg && nnz >= 0 && nS > 0 && sigma > 0,
                 "memi_markov_coo: bad arguments");
  if (nnz == 0) return MEMI_OK;
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_zero_f64, dim3(grid_for(2 * nS)), dim3(256), 0, s, deg, 2 * nS);
  MEMI_CHECK_LAUNCH("k_zero_f64");
  const double inv_s2 = 1.0 / (sigma * sigma);
  const dim3 g(grid_for(nnz));
  hipLaunchKernelGGL(k_coo_norm<0>, g, dim3(256), 0, s, row, col, val, nnz, inv_s2, alpha, deg,
                     nS, out);
  MEMI_CHECK_LAUNCH("k_coo_norm<0>");
  hipLaunchKernelGGL(k_coo_norm<1>, g, dim3(256), 0, s, row, col, val, nnz, inv_s2, alpha, deg,
                     nS, out);
  MEMI_CHECK_LAUNCH("k_coo_norm<1>");
  hipLaunchKernelGGL(k_coo_norm<2>, g, dim3(256), 0, s, row, col, val, nnz, inv_s2, alpha, deg,
                     nS, out);
  MEMI_CHECK_LAUNCH("k_coo_norm<2>");
  return MEMI_OK;
}

int memi_gather_blocks(const void* plan_dev, const memi_plan_info* info, const float* src,
                       const int64_t* src_off, const int32_t* src_n, const int32_t* sel,
                       float* dst, void* stream) {
  MEMI_CHECK_ARG(plan_dev && info && src && src_off && src_n && sel && dst,
                 "memi_gather_blocks: null pointer");
  if (info->n_rows == 0) return MEMI_OK;
  MEMI_CHECK_ARG(info->n_rows < 0x7fffffffLL, "memi_gather_blocks: %lld rows",
                 (long long)info->n_rows);
  PlanView v = view(plan_dev, info);
  hipLaunchKernelGGL(k_gather_blocks, dim3((unsigned)info->n_rows), dim3(256), 0,
                     (hipStream_t)stream, v.rows, v.pds, src, src_off, src_n, sel, dst);
  MEMI_CHECK_LAUNCH("k_gather_blocks");
  return MEMI_OK;
}

int memi_tanh_fit(const double* x, const double* y, int m, int npd, const double* p0,
                  const int32_t* active, int max_iter, double* out, void* stream) {
  MEMI_CHECK_ARG(x && y && p0 && out && m >= 5 && npd >= 0 && max_iter > 0,
                 "memi_tanh_fit: bad arguments");
  if (npd == 0) return MEMI_OK;
  hipLaunchKernelGGL(k_tanh_fit, dim3((unsigned)npd), dim3(TANH_NT), 0, (hipStream_t)stream, x, y, m,
                     p0, active, max_iter, out);
  MEMI_CHECK_LAUNCH("k_tanh_fit");
  return MEMI_OK;
}

int memi_tanh_select(const double* fit, int npd, int R, int m, double tune, double max_res,
                     int32_t* ok, double* sigma, double* res, void* stream) {
  MEMI_CHECK_ARG(fit && ok && sigma && npd >= 0 && R >= 1 && m >= 5,
                 "memi_tanh_select: bad arguments");
  if (npd == 0) return MEMI_OK;
  hipLaunchKernelGGL(k_tanh_select, dim3((unsigned)((npd + 63) / 64)), dim3(64), 0,
                     (hipStream_t)stream, fit, npd, R, m, tune, max_res, ok, sigma, res);
  MEMI_CHECK_LAUNCH("k_tanh_select");
  return MEMI_OK;
}

}  // extern "C"
