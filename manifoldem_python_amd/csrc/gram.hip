// CTF-weighted distance Gram on the MFMA units of gfx950.
//
// Reference: getDistanceCTF_local_Conj9combinedS2.py:391-397
//   CTFfy = CTF.conj()*fy;  D = |CTF|^2 . (|fy|^2)^T;  D = D + D^T - 2 Re(CTFfy . CTFfy^H)
// i.e. D_ij = sum_k |C_jk X_ik - C_ik X_jk|^2 (the dead direct form at :106-122).
//
// Formulation.  With the packed operand row of image i
//     L_i = [ A_i | B_i | Yr_i | Yi_i ],  A = w|X|^2, B = C^2, Y = sqrt(w) C X
// (half spectrum, column weights w folded in), the four segment products
//     S1 = A B^T,  S2 = B A^T,  R1 = Yr Yr^T,  R2 = Yi Yi^T
// give P = S1 + S2 and R = R1 + R2, and D = P - 2R.  Each (tile, segment, chunk) is an
// independent work item: an MFMA GEMM over one K range into its own float32 partial tile,
// so every accumulator chain is short and the float64 finalize sees P and R separately,
// which also yields the cancellation ratio kappa = (P + 2|R|)/D per entry.
//
// Tiles: 128x128 outputs per 256-thread workgroup (4 waves, 2x2, 64x64 each as 4x4 blocks of
// 16x16 MFMA outputs).  K steps of 64 bytes per operand row are staged through LDS, double
// buffered, 16-byte chunks XOR-swizzled per row.  Split-BF16 issues hi*hi, hi*lo and lo*hi
// v_mfma_f32_16x16x32_bf16 per block; FP32 issues four v_mfma_f32_16x16x4_f32.  Both
// operands of an MFMA read the same k slots of their lanes, so the sum over a block's k range
// is exact whatever order the hardware uses inside an instruction.
#include "common.h"

namespace memi {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float floatx4_t __attribute__((ext_vector_type(4)));

constexpr int BM = 128;
constexpr int BN = 128;
constexpr int ROW_STEP = 64;           // bytes of one operand row per K step
constexpr int LDS_ROW = 64;            // LDS row stride (bytes); 16-byte chunks swizzled
constexpr int LDS_TILE = BM * LDS_ROW; // 10240 bytes per operand tile

// LDS placement of 16-byte chunk c of tile row r: chunk c ^ swz(r), with swz's two bits
// row bit 2 and row bits 1 ^ 3 -- conflict-free for the 16-row fragment reads and the 4-lane
// row stores (measured 2-5% faster than rows padded to 80 bytes)
__device__ inline int swz(int r) { return ((r >> 2) & 1) | ((((r >> 1) ^ (r >> 3)) & 1) << 1); }

// segment of the right operand that pairs with left segment s: A<->B, Yr, Yi
__device__ inline int seg_pair(int s) { return s < 2 ? (s ^ 1) : s; }

template <int PREC>
struct GramCfg;
template <>
struct GramCfg<MEMI_PREC_FP32> {
  static constexpr int ESIZE = 4;
  static constexpr int NARR = 2;  // A, B
};
template <>
struct GramCfg<MEMI_PREC_SPLIT_BF16> {
  static constexpr int ESIZE = 2;
  static constexpr int NARR = 4;  // A hi, A lo, B hi, B lo
};

// workgroups per CU every variant is built for (depth-2 split-BF16 fits in 253 VGPRs)
template <int PREC, int DEPTH>
struct GramOcc {
  static constexpr int value = 2;
};

template <int PREC, int DEPTH>
__global__ __launch_bounds__(256, (GramOcc<PREC, DEPTH>::value)) void k_gram_tiles(
    const char* __restrict__ op_hi, const char* __restrict__ op_lo, int64_t Khp, int spc,
    int steps, const int32_t* __restrict__ items, int64_t n_items,
    const memi_pd_desc* __restrict__ pds, float* __restrict__ ws) {
  using Cfg = GramCfg<PREC>;
  constexpr int ES = Cfg::ESIZE;
  constexpr int NA = Cfg::NARR;
  __shared__ __attribute__((aligned(16))) char lds[2 * NA * LDS_TILE];

  // XCD-aware order: workgroup w runs on XCD w % 8; give each XCD a contiguous range of
  // items (consecutive items share an operand panel).  gridDim.x is a multiple of 8.
  const int64_t q = gridDim.x / 8;
  const int64_t it = (int64_t)(blockIdx.x % 8) * q + blockIdx.x / 8;
  if (it >= n_items) return;
  const int32_t* itm = items + it * 6;
  const int pd = itm[0], ti = itm[1], tj = itm[2], seg = itm[3], chunk = itm[4], slot = itm[5];
  const memi_pd_desc P = pds[pd];
  const int mi = min(BM, P.n - ti * BM);
  const int mj = min(BN, P.n - tj * BN);
  const int64_t row_bytes = 4 * Khp * ES;
  // chunk c covers K steps [c*spc, min((c+1)*spc, steps)); the host guarantees none is empty
  const int64_t a_k0 = (int64_t)seg * Khp * ES + (int64_t)chunk * spc * ROW_STEP;
  const int64_t b_k0 = (int64_t)seg_pair(seg) * Khp * ES + (int64_t)chunk * spc * ROW_STEP;
  const int nks = min(spc, steps - chunk * spc);

  const int t = threadIdx.x;
  const int lane = t & 63, wv = t >> 6;
  const int wr = wv >> 1, wc = wv & 1;

  // global -> LDS copy assignment: 2 rows x 16 bytes per thread per operand tile
  const int ld_row0 = t >> 2, ld_part = t & 3;
  int64_t oA0, oA1, oB0, oB1;
  {
    const int r0 = ld_row0, r1 = ld_row0 + 64;
    oA0 = (P.row0 + (int64_t)ti * BM + min(r0, mi - 1)) * row_bytes + a_k0 + ld_part * 16;
    oA1 = (P.row0 + (int64_t)ti * BM + min(r1, mi - 1)) * row_bytes + a_k0 + ld_part * 16;
    oB0 = (P.row0 + (int64_t)tj * BN + min(r0, mj - 1)) * row_bytes + b_k0 + ld_part * 16;
    oB1 = (P.row0 + (int64_t)tj * BN + min(r1, mj - 1)) * row_bytes + b_k0 + ld_part * 16;
  }
  const int lds_st = ld_row0 * LDS_ROW + ((ld_part ^ swz(ld_row0)) * 16);
  constexpr int HALF = 64 * LDS_ROW;

  // Two register staging sets (X, Y) give a prefetch distance of two K steps: the loads for
  // step k+2 are issued before step k's MFMAs and land in LDS only after step k+1's.
  uint4 X0h, X1h, X2h, X3h, X0l, X1l, X2l, X3l;
  uint4 Y0h, Y1h, Y2h, Y3h, Y0l, Y1l, Y2l, Y3l;
#define MEMI_GLOAD(R, ks)                                                     \
  do {                                                                        \
    const int64_t ko_ = (int64_t)(ks) * ROW_STEP;                             \
    R##0h = *(const uint4*)(op_hi + oA0 + ko_);                               \
    R##1h = *(const uint4*)(op_hi + oA1 + ko_);                               \
    R##2h = *(const uint4*)(op_hi + oB0 + ko_);                               \
    R##3h = *(const uint4*)(op_hi + oB1 + ko_);                               \
    if (NA == 4) {                                                            \
      R##0l = *(const uint4*)(op_lo + oA0 + ko_);                             \
      R##1l = *(const uint4*)(op_lo + oA1 + ko_);                             \
      R##2l = *(const uint4*)(op_lo + oB0 + ko_);                             \
      R##3l = *(const uint4*)(op_lo + oB1 + ko_);                             \
    }                                                                         \
  } while (0)
  // LDS tile order: A hi, B hi (FP32: A, B) then A lo, B lo
#define MEMI_LSTORE(R, buf)                                                   \
  do {                                                                        \
    char* b_ = lds + (buf) * NA * LDS_TILE + lds_st;                          \
    *(uint4*)(b_ + 0 * LDS_TILE) = R##0h;                                     \
    *(uint4*)(b_ + 0 * LDS_TILE + HALF) = R##1h;                              \
    *(uint4*)(b_ + 1 * LDS_TILE) = R##2h;                                     \
    *(uint4*)(b_ + 1 * LDS_TILE + HALF) = R##3h;                              \
    if (NA == 4) {                                                            \
      *(uint4*)(b_ + 2 * LDS_TILE) = R##0l;                                   \
      *(uint4*)(b_ + 2 * LDS_TILE + HALF) = R##1l;                            \
      *(uint4*)(b_ + 3 * LDS_TILE) = R##2l;                                   \
      *(uint4*)(b_ + 3 * LDS_TILE + HALF) = R##3l;                            \
    }                                                                         \
  } while (0)

  floatx4_t acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = floatx4_t{0.f, 0.f, 0.f, 0.f};

  const int frag_off = (lane & 15) * LDS_ROW + (((lane >> 4) ^ swz(lane & 15)) * 16);

#define MEMI_COMPUTE(cur, G)                                                                 \
  do {                                                                                      \
    const char* base = lds + (cur) * NA * LDS_TILE;                                         \
    if constexpr (PREC == MEMI_PREC_SPLIT_BF16) {                                           \
      /* B hi / lo fragments stay resident; A hi / lo of one 16-row block at a time keeps    \
         the live fragments at 40 VGPRs.  Each accumulator still gets hi*hi, hi*lo, lo*hi. */ \
      bf16x8_t fb[4], fc[4];                                                                \
      _Pragma("unroll") for (int j = 0; j < 4; ++j) {                                       \
        fb[j] = *(const bf16x8_t*)(base + 1 * LDS_TILE + (wc * 64 + j * 16) * LDS_ROW + frag_off); \
        fc[j] = *(const bf16x8_t*)(base + 3 * LDS_TILE + (wc * 64 + j * 16) * LDS_ROW + frag_off); \
      }                                                                                     \
      _Pragma("unroll") for (int i = 0; i < 4; ++i) {                                       \
        const bf16x8_t ah = *(const bf16x8_t*)(base + 0 * LDS_TILE + (wr * 64 + i * 16) * LDS_ROW + frag_off); \
        const bf16x8_t al = *(const bf16x8_t*)(base + 2 * LDS_TILE + (wr * 64 + i * 16) * LDS_ROW + frag_off); \
        _Pragma("unroll") for (int j = 0; j < 4; ++j)                                       \
          if (!(G) || (i < nbi && j < nbj)) {                                               \
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, fb[j], acc[i][j], 0, 0, 0); \
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, fc[j], acc[i][j], 0, 0, 0); \
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, fb[j], acc[i][j], 0, 0, 0); \
          }                                                                                 \
      }                                                                                     \
    } else {                                                                                \
      floatx4_t a[4], b[4];                                                                 \
      _Pragma("unroll") for (int i = 0; i < 4; ++i) {                                       \
        a[i] = *(const floatx4_t*)(base + 0 * LDS_TILE + (wr * 64 + i * 16) * LDS_ROW + frag_off); \
        b[i] = *(const floatx4_t*)(base + 1 * LDS_TILE + (wc * 64 + i * 16) * LDS_ROW + frag_off); \
      }                                                                                     \
      _Pragma("unroll") for (int e = 0; e < 4; ++e) _Pragma("unroll") for (int i = 0; i < 4; ++i) \
        _Pragma("unroll") for (int j = 0; j < 4; ++j)                                       \
          if (!(G) || (i < nbi && j < nbj))                                                 \
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[i][e], b[j][e], acc[i][j], 0, 0, 0); \
    }                                                                                       \
  } while (0)

  // The K loop, G = true for edge tiles: MFMA blocks outside the tile's rows/columns are
  // skipped (nbi x nbj of this wave's 4x4 blocks hold data), full tiles run unguarded.
#define MEMI_KLOOP(G)\
  do {                                                                                      \
  if constexpr (DEPTH == 1) {                                                               \
    /* prefetch distance one K step: loads for ks+1 are issued before step ks's MFMAs */    \
    MEMI_GLOAD(X, 0);                                                                       \
    MEMI_LSTORE(X, 0);                                                                      \
    __syncthreads();                                                                        \
    for (int ks = 0; ks < nks; ++ks) {                                                      \
      const int cur = ks & 1;                                                               \
      if (ks + 1 < nks) MEMI_GLOAD(X, ks + 1);                                              \
      MEMI_COMPUTE(cur, G);                                                                 \
      if (ks + 1 < nks) MEMI_LSTORE(X, cur ^ 1);                                            \
      __syncthreads();                                                                      \
    }                                                                                       \
  } else {                                                                                  \
    /* prologue: step 0 into LDS buffer 0, step 1 in flight in Y */                         \
    MEMI_GLOAD(X, 0);                                                                       \
    if (nks > 1) MEMI_GLOAD(Y, 1);                                                          \
    MEMI_LSTORE(X, 0);                                                                      \
    __syncthreads();                                                                        \
    /* steady state, unrolled by two so the register sets alternate by name */              \
    int ks = 0;                                                                             \
    while (ks < nks) {                                                                      \
      /* even step: compute buffer 0; X is free -> fetch ks+2; Y (step ks+1) -> buffer 1 */ \
      if (ks + 2 < nks) MEMI_GLOAD(X, ks + 2);                                              \
      MEMI_COMPUTE(0, G);                                                                   \
      if (ks + 1 < nks) MEMI_LSTORE(Y, 1);                                                  \
      __syncthreads();                                                                      \
      ++ks;                                                                                 \
      if (ks >= nks) break;                                                                 \
      /* odd step: compute buffer 1; Y is free -> fetch ks+2; X (step ks+1) -> buffer 0 */  \
      if (ks + 2 < nks) MEMI_GLOAD(Y, ks + 2);                                              \
      MEMI_COMPUTE(1, G);                                                                   \
      if (ks + 1 < nks) MEMI_LSTORE(X, 0);                                                  \
      __syncthreads();                                                                      \
      ++ks;                                                                                 \
    }                                                                                       \
  }                                                                                         \
  } while (0)
  const int nbi = max(0, min(4, (mi - wr * 64 + 15) / 16));
  const int nbj = max(0, min(4, (mj - wc * 64 + 15) / 16));
  if (mi == BM && mj == BN) {
    MEMI_KLOOP(false);
  } else {
    MEMI_KLOOP(true);
  }
#undef MEMI_KLOOP
#undef MEMI_GLOAD
#undef MEMI_LSTORE
#undef MEMI_COMPUTE

  // 16x16 output block: lane holds rows 4*(lane/16) + r, column lane % 16
  float* out = ws + (int64_t)slot * BM * BN;
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = wr * 64 + i * 16 + (lane >> 4) * 4 + r;
        const int col = wc * 64 + j * 16 + (lane & 15);
        out[row * BN + col] = acc[i][j][r];
      }
}

constexpr int FIN_SPLIT = 8;   // workgroups per upper tile in the finalize

// Reduce the segment partials of one upper tile in float64, write D and its mirror,
// and queue entries with a large cancellation ratio for direct recomputation.
__global__ __launch_bounds__(256) void k_gram_finalize(const int32_t* __restrict__ utiles,
                                                       const memi_pd_desc* __restrict__ pds,
                                                       const float* __restrict__ ws, int chunks,
                                                       double kappa, float* __restrict__ D,
                                                       int32_t* __restrict__ flags,
                                                       int64_t flag_cap,
                                                       int32_t* __restrict__ flag_count) {
  const int32_t* u = utiles + (int64_t)blockIdx.x * 4;
  const int pd = u[0], ti = u[1], tj = u[2];
  const int64_t slot0 = u[3];
  const memi_pd_desc P = pds[pd];
  const int n = P.n;
  const int64_t tile = (int64_t)BM * BN;
  // blockIdx.y: which FIN_SPLIT-th of the tile's entries this block reduces
  const int e0 = blockIdx.y * (BM * BN / FIN_SPLIT), e1 = e0 + BM * BN / FIN_SPLIT;
  for (int e = e0 + threadIdx.x; e < e1; e += blockDim.x) {
    const int r = e / BN, c = e % BN;
    const int i = ti * BM + r, j = tj * BN + c;
    if (i >= n || j >= n) continue;
    if (ti == tj && c < r) continue;
    double d = 0.0;
    if (i != j) {
      double Pp = 0.0, R = 0.0;
      for (int ch = 0; ch < chunks; ++ch) {
        Pp += (double)ws[(slot0 + 0 * chunks + ch) * tile + e];
        Pp += (double)ws[(slot0 + 1 * chunks + ch) * tile + e];
        R += (double)ws[(slot0 + 2 * chunks + ch) * tile + e];
        R += (double)ws[(slot0 + 3 * chunks + ch) * tile + e];
      }
      d = Pp - 2.0 * R;
      if (Pp + 2.0 * fabs(R) > kappa * d) {
        const int32_t slot = atomicAdd(flag_count, 1);
        if (slot < flag_cap) {
          flags[3 * (int64_t)slot + 0] = pd;
          flags[3 * (int64_t)slot + 1] = i;
          flags[3 * (int64_t)slot + 2] = j;
        }
      }
    }
    D[P.d_off + (int64_t)i * n + j] = (float)d;
    if (i != j) D[P.d_off + (int64_t)j * n + i] = (float)d;
  }
}

// Direct form sum_k w_k |C_jk X_ik - C_ik X_jk|^2 for queued entries, one wave per entry.
__global__ __launch_bounds__(256) void k_gram_recompute(const int32_t* __restrict__ flags,
                                                        const int32_t* __restrict__ flag_count,
                                                        int64_t flag_cap,
                                                        const memi_pd_desc* __restrict__ pds,
                                                        const float2* __restrict__ fy,
                                                        const float* __restrict__ ctf, int N,
                                                        float* __restrict__ D) {
  const int Nh = N / 2 + 1;
  const int64_t Kh = (int64_t)N * Nh;
  const int64_t cnt = min((int64_t)*flag_count, flag_cap);
  const int lane = threadIdx.x & 63;
  const int64_t wave = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  const int64_t nwaves = ((int64_t)gridDim.x * blockDim.x) >> 6;
  for (int64_t e = wave; e < cnt; e += nwaves) {
    const int pd = flags[3 * e], i = flags[3 * e + 1], j = flags[3 * e + 2];
    const memi_pd_desc P = pds[pd];
    const int64_t ri = (P.row0 + i) * Kh, rj = (P.row0 + j) * Kh;
    double s = 0.0;
    for (int64_t k = lane; k < Kh; k += 64) {
      const float2 xi = fy[ri + k], xj = fy[rj + k];
      const float ci = ctf[ri + k], cj = ctf[rj + k];
      const float dr = cj * xi.x - ci * xj.x;
      const float di = cj * xi.y - ci * xj.y;
      s += (double)half_weight((int)(k % Nh), N) * ((double)dr * dr + (double)di * di);
    }
    for (int off = 32; off > 0; off >>= 1) s += __shfl_down(s, off, 64);
    if (lane == 0) {
      D[P.d_off + (int64_t)i * P.n + j] = (float)s;
      D[P.d_off + (int64_t)j * P.n + i] = (float)s;
    }
  }
}

__global__ void k_zero_i32(int32_t* p, int64_t n) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < n;
       e += (int64_t)gridDim.x * blockDim.x)
    p[e] = 0;
}

}  // namespace memi

using namespace memi;

extern "C" {

int memi_plan(const memi_pd_desc* pds_in, int npd, int chunks, void* plan_host,
              int64_t plan_cap, memi_plan_info* info) {
  MEMI_CHECK_ARG(info != nullptr && (npd == 0 || pds_in != nullptr), "memi_plan: null pointer");
  MEMI_CHECK_ARG(npd >= 0 && chunks >= 1 && chunks <= 64, "memi_plan: bad npd/chunks");
  memi_plan_info I{};
  I.n_pd = npd;
  I.chunks = chunks;
  int64_t d = 0, mw = 0, le = 0, rows = 0, items = 0, ut = 0, ft = 0, fc = 0;
  int max_n = 0;
  for (int p = 0; p < npd; ++p) {
    const int64_t n = pds_in[p].n;
    MEMI_CHECK_ARG(n >= 1 && n <= 65536 && pds_in[p].k >= 1,
                   "memi_plan: PD %d has n=%lld k=%d", p, (long long)n, pds_in[p].k);
    MEMI_CHECK_ARG(pds_in[p].row0 == rows,
                   "memi_plan: PD %d starts at row %lld; PD rows must be packed contiguously "
                   "in order (expected %lld)", p, (long long)pds_in[p].row0, (long long)rows);
    const int64_t T = (n + BM - 1) / BM;
    d += n * n;
    mw += n * ((n + 31) / 32);
    le += (int64_t)std::min<int64_t>(pds_in[p].k, n) * n;
    rows += n;
    ut += T * (T + 1) / 2;
    ft += T * T;
    fc += n * (n - 1) / 2;
    if (n > max_n) max_n = (int)n;
  }
  items = ut * 4 * chunks;
  I.n_items = items;
  I.n_utiles = ut;
  I.n_ftiles = ft;
  I.n_rows = rows;
  I.d_elems = d;
  I.mask_words = mw;
  I.list_elems = le;
  I.ws_floats = ut * 4 * chunks * (int64_t)BM * BN;
  I.flag_cap = std::max<int64_t>(fc, 1);
  I.max_n = max_n;
  int64_t off = 0;
  I.off_pds = off;
  off += (int64_t)npd * sizeof(memi_pd_desc);
  off = (off + 15) & ~15LL;
  I.off_items = off;
  off += items * 6 * 4;
  I.off_utiles = off;
  off += ut * 4 * 4;
  I.off_ftiles = off;
  off += ft * 4 * 4;
  I.off_rows = off;
  off += rows * 2 * 4;
  I.plan_bytes = (off + 15) & ~15LL;
  *info = I;
  if (plan_host == nullptr) return MEMI_OK;
  MEMI_CHECK_ARG(plan_cap >= I.plan_bytes, "memi_plan: plan buffer too small (%lld < %lld)",
                 (long long)plan_cap, (long long)I.plan_bytes);
  char* base = (char*)plan_host;
  memi_pd_desc* pds = (memi_pd_desc*)(base + I.off_pds);
  int32_t* it = (int32_t*)(base + I.off_items);
  int32_t* uts = (int32_t*)(base + I.off_utiles);
  int32_t* fts = (int32_t*)(base + I.off_ftiles);
  int32_t* rw = (int32_t*)(base + I.off_rows);
  d = mw = le = 0;
  int64_t uidx = 0, nit = 0, nft = 0, nr = 0;
  for (int p = 0; p < npd; ++p) {
    memi_pd_desc P = pds_in[p];
    const int n = P.n;
    P.d_off = d;
    P.mask_off = mw;
    P.list_off = le;
    pds[p] = P;
    d += (int64_t)n * n;
    mw += (int64_t)n * ((n + 31) / 32);
    le += (int64_t)std::min(P.k, n) * n;
    const int T = (n + BM - 1) / BM;
    const int64_t u0 = uidx;
    for (int a = 0; a < T; ++a)
      for (int b = a; b < T; ++b) {
        int32_t* u = uts + 4 * uidx;
        u[0] = p; u[1] = a; u[2] = b; u[3] = (int32_t)(uidx * 4 * chunks);
        ++uidx;
      }
    // items: segment-major, then chunk, then row panel, so neighbours share operand panels
    for (int s = 0; s < 4; ++s)
      for (int c = 0; c < chunks; ++c) {
        int64_t uu = u0;
        for (int a = 0; a < T; ++a)
          for (int b = a; b < T; ++b) {
            int32_t* x = it + 6 * nit++;
            x[0] = p; x[1] = a; x[2] = b; x[3] = s; x[4] = c;
            x[5] = (int32_t)(uu * 4 * chunks + s * chunks + c);
            ++uu;
          }
      }
    for (int a = 0; a < T; ++a)
      for (int b = 0; b < T; ++b) {
        int32_t* f = fts + 4 * nft++;
        f[0] = p; f[1] = a; f[2] = b; f[3] = 0;
      }
    for (int i = 0; i < n; ++i) {
      rw[2 * nr] = p;
      rw[2 * nr + 1] = i;
      ++nr;
    }
  }
  return MEMI_OK;
}

int memi_ctf_distance(const void* plan_dev, const memi_plan_info* info, const void* op_hi,
                      const void* op_lo, int64_t Khp, int precision, double kappa,
                      const void* fy, const float* ctf_half, int N, float* D, float* ws,
                      int32_t* flags, int32_t* flag_count, int variant, void* ev_start,
                      void* ev_stop, void* stream) {
  MEMI_CHECK_ARG(plan_dev && info && op_hi && fy && ctf_half && D && ws && flags && flag_count,
                 "memi_ctf_distance: null pointer");
  MEMI_CHECK_ARG(precision == MEMI_PREC_FP32 || (precision == MEMI_PREC_SPLIT_BF16 && op_lo),
                 "memi_ctf_distance: bad precision %d", precision);
  MEMI_CHECK_ARG(Khp > 0 && Khp % 64 == 0 && Khp >= (int64_t)N * (N / 2 + 1),
                 "memi_ctf_distance: bad Khp %lld", (long long)Khp);
  const int chunks = info->chunks;
  const int kstep = ROW_STEP / (precision == MEMI_PREC_FP32 ? 4 : 2);
  // K steps per segment, split into chunks of spc steps (the last one may be shorter)
  const int64_t steps = Khp / kstep;
  const int64_t spc = (steps + chunks - 1) / chunks;
  MEMI_CHECK_ARG(steps < (1LL << 30) && (int64_t)(chunks - 1) * spc < steps,
                 "memi_ctf_distance: %d chunks leave an empty chunk of the %lld K steps",
                 chunks, (long long)steps);
  MEMI_CHECK_ARG(kappa > 0, "memi_ctf_distance: kappa must be positive");
  if (info->n_pd == 0) return MEMI_OK;
  hipStream_t s = (hipStream_t)stream;
  const char* plan = (const char*)plan_dev;
  const memi_pd_desc* pds = (const memi_pd_desc*)(plan + info->off_pds);
  const int32_t* items = (const int32_t*)(plan + info->off_items);
  const int32_t* ut = (const int32_t*)(plan + info->off_utiles);
  const int64_t grid = ((info->n_items + 7) / 8) * 8;
  MEMI_CHECK_ARG(grid < (int64_t)1 << 31, "memi_ctf_distance: too many work items");
  hipLaunchKernelGGL(k_zero_i32, dim3(1), dim3(64), 0, s, flag_count, (int64_t)1);
  MEMI_CHECK_LAUNCH("k_zero_i32");
  if (ev_start && hipEventRecord((hipEvent_t)ev_start, s) != hipSuccess) {
    set_error("memi_ctf_distance: hipEventRecord failed");
    return MEMI_ELAUNCH;
  }
  MEMI_CHECK_ARG(variant >= 0 && variant <= 2, "memi_ctf_distance: unknown variant %d", variant);
  const int v = variant == 0 ? 1 : variant;
  const char* lo = precision == MEMI_PREC_FP32 ? nullptr : (const char*)op_lo;
#define MEMI_LAUNCH_GRAM(P, V)                                                                \
  hipLaunchKernelGGL((k_gram_tiles<P, V>), dim3((unsigned)grid), dim3(256), 0, s,              \
                     (const char*)op_hi, lo, Khp, (int)spc, (int)steps, items, info->n_items, \
                     pds, ws)
  if (precision == MEMI_PREC_FP32) {
    if (v == 1) MEMI_LAUNCH_GRAM(MEMI_PREC_FP32, 1); else MEMI_LAUNCH_GRAM(MEMI_PREC_FP32, 2);
  } else {
    if (v == 1) MEMI_LAUNCH_GRAM(MEMI_PREC_SPLIT_BF16, 1);
    else MEMI_LAUNCH_GRAM(MEMI_PREC_SPLIT_BF16, 2);
  }
#undef MEMI_LAUNCH_GRAM
  MEMI_CHECK_LAUNCH("k_gram_tiles");
  if (ev_stop && hipEventRecord((hipEvent_t)ev_stop, s) != hipSuccess) {
    set_error("memi_ctf_distance: hipEventRecord failed");
    return MEMI_ELAUNCH;
  }
  hipLaunchKernelGGL(k_gram_finalize, dim3((unsigned)info->n_utiles, FIN_SPLIT), dim3(256), 0, s,
                     ut, pds, ws, chunks, kappa, D, flags, info->flag_cap, flag_count);
  MEMI_CHECK_LAUNCH("k_gram_finalize");
  hipLaunchKernelGGL(k_gram_recompute, dim3(1024), dim3(256), 0, s, flags, flag_count,
                     info->flag_cap, pds, (const float2*)fy, ctf_half, N, D);
  MEMI_CHECK_LAUNCH("k_gram_recompute");
  return MEMI_OK;
}

}  // extern "C"
