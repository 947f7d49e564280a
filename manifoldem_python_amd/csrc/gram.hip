// CTF-weighted distance Gram on the MFMA units of gfx950.
//
// Reference: getDistanceCTF_local_Conj9combinedS2.py:391-397
//   CTFfy = CTF.conj()*fy;  D = |CTF|^2 . (|fy|^2)^T;  D = D + D^T - 2 Re(CTFfy . CTFfy^H)
// i.e. D_ij = sum_k |C_jk X_ik - C_ik X_jk|^2 (the dead direct form at :106-122).
//
// Formulation.  With the packed operand row of image i
//     L_i = [ A_i | B_i | Yr_i | Yi_i ],  A = w|X|^2, B = C^2, Y = sqrt(w) C X
// (half spectrum, column weights w folded in), the two halves of the contraction
//     P = [A B] . [B A]^T      (segments 0, 1 against 1, 0)
//     R = [Yr Yi] . [Yr Yi]^T  (segments 2, 3 against themselves)
// give D = P - 2R.  A work item is one (tile, half, K chunk): an MFMA GEMM over its K range
// into its own float32 partial tile, so every accumulator chain is short and the float64
// finalize sees P and R separately, which also yields the cancellation ratio
// kappa = (P + 2|R|)/D per entry.
//
// Tiles: BT x BT outputs per workgroup (BT = 128: 4 waves of 64 x 64, two workgroups per CU;
// BT = 256: 8 waves of 128 x 64, one workgroup per CU, 3/4 of the operand bytes staged per
// MFMA).  K steps of 64 bytes per operand row are staged through LDS, double buffered,
// 16-byte chunks XOR-swizzled per row.  Split-BF16 issues hi*hi, hi*lo and lo*hi
// v_mfma_f32_16x16x32_bf16 per block; FP32 issues four v_mfma_f32_16x16x4_f32.  Both operands
// of an MFMA read the same k slots of their lanes, so the sum over a block's k range is exact
// whatever order the hardware uses inside an instruction.  On a diagonal tile the blocks
// wholly below the diagonal are skipped (the finalize reads the upper triangle only).
#include "common.h"

namespace memi {

typedef __bf16 bf16x8_t __attribute__((ext_vector_type(8)));
typedef float floatx4_t __attribute__((ext_vector_type(4)));
typedef unsigned int u32x4_t __attribute__((ext_vector_type(4)));   // 16 staged bytes

constexpr int TB_FIN = 2048;           // tile entries per finalize workgroup
constexpr int ROW_STEP = 64;           // bytes of one operand row per K step
constexpr int LDS_ROW = 64;            // LDS row stride (bytes); 16-byte chunks swizzled

// LDS placement of 16-byte chunk c of tile row r: chunk c ^ swz(r), with swz's two bits
// row bit 2 and row bits 1 ^ 3 -- conflict-free for the 16-row fragment reads (measured 2-5%
// faster than rows padded to 80 bytes).  Rows r and r + 64q get the same swizzle.
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
  static constexpr int NARR = 4;  // A hi, B hi, A lo, B lo
};

// Workgroup shapes: WAVES_M x WAVES_N waves, each computing WM x WN outputs as
// (WM/16) x (WN/16) blocks of 16x16 MFMA outputs; the workgroup tile is BT x BT.
//   Shape128: 2 x 2 waves of 64 x 64 (256 threads, two workgroups per CU)
//   Shape256: 2 x 4 waves of 128 x 64 (512 threads, one workgroup per CU, two waves per SIMD)
template <int WM_, int WN_, int WAVES_M_, int WAVES_N_, int OCC_>
struct GramShape {
  static constexpr int WM = WM_, WN = WN_, WAVES_M = WAVES_M_, WAVES_N = WAVES_N_;
  static constexpr int NT = 64 * WAVES_M * WAVES_N;       // threads
  static constexpr int BT = WAVES_M * WM;                 // tile side
  static constexpr int NBI = WM / 16, NBJ = WN / 16;      // MFMA blocks per wave
  static constexpr int RPT = BT / (NT / 4);               // staged rows per thread per array
  static constexpr int OCC = OCC_;                        // workgroups per CU
  static_assert(WAVES_N * WN == BT, "square workgroup tile");
};
using Shape128 = GramShape<64, 64, 2, 2, 2>;
using Shape256 = GramShape<128, 64, 2, 4, 1>;

// register staging of one K step: NA arrays x RPT rows of 16 bytes per thread
template <int NA, int RPT>
struct GramStage {
  u32x4_t v[NA][RPT];
};

template <int NA, int RPT>
__device__ __forceinline__ void gram_gload(GramStage<NA, RPT>& R, const char* gAh,
                                           const char* gBh, const char* gAl, const char* gBl,
                                           const uint32_t (&oA)[RPT], const uint32_t (&oB)[RPT],
                                           uint32_t ao, uint32_t bo) {
#pragma unroll
  for (int q = 0; q < RPT; ++q) {
    R.v[0][q] = *(const u32x4_t*)(gAh + (oA[q] + ao));
    R.v[1][q] = *(const u32x4_t*)(gBh + (oB[q] + bo));
    if constexpr (NA == 4) {
      R.v[2][q] = *(const u32x4_t*)(gAl + (oA[q] + ao));
      R.v[3][q] = *(const u32x4_t*)(gBl + (oB[q] + bo));
    }
  }
}

// LDS array order: A hi, B hi (FP32: A, B) then A lo, B lo; row q of a thread is
// ld_row0 + q * RSTRIDE
template <int NA, int RPT, int TILE, int RSTRIDE>
__device__ __forceinline__ void gram_lstore(const GramStage<NA, RPT>& R, char* b) {
#pragma unroll
  for (int a = 0; a < NA; ++a)
#pragma unroll
    for (int q = 0; q < RPT; ++q)
      *(u32x4_t*)(b + a * TILE + q * RSTRIDE * LDS_ROW) = R.v[a][q];
}

// Direct-to-LDS staging of one K step (global_load_lds_dwordx4): each instruction of a wave
// fills 16 rows x 64 bytes of LDS, lane l landing at byte 16 l, so lane l loads row l / 4,
// chunk (l % 4) ^ swz(l / 4) of its 16-row group: the swizzled layout the fragment reads
// expect.  The wave's groups are wvu + q * NW (q < RPT); the LDS destination is uniform.
template <int NA, class S, int TILE>
__device__ __forceinline__ void gram_dload(char* buf, const char* gAh, const char* gBh,
                                           const char* gAl, const char* gBl,
                                           const uint32_t (&oA)[S::RPT],
                                           const uint32_t (&oB)[S::RPT], uint32_t ao,
                                           uint32_t bo, int wvu) {
  constexpr int NW = S::NT / 64;
  typedef __attribute__((address_space(3))) void lds_void;
#pragma unroll
  for (int q = 0; q < S::RPT; ++q) {
    char* d = buf + (wvu + q * NW) * 16 * LDS_ROW;
    __builtin_amdgcn_global_load_lds((const void*)(gAh + (oA[q] + ao)), (lds_void*)(d), 16, 0,
                                     0);
    __builtin_amdgcn_global_load_lds((const void*)(gBh + (oB[q] + bo)),
                                     (lds_void*)(d + 1 * TILE), 16, 0, 0);
    if constexpr (NA == 4) {
      __builtin_amdgcn_global_load_lds((const void*)(gAl + (oA[q] + ao)),
                                       (lds_void*)(d + 2 * TILE), 16, 0, 0);
      __builtin_amdgcn_global_load_lds((const void*)(gBl + (oB[q] + bo)),
                                       (lds_void*)(d + 3 * TILE), 16, 0, 0);
    }
  }
}

// MFMAs of one K step for this wave's WM x WN outputs.  G: guarded blocks (inside the PD:
// nbi x nbj of the wave's blocks hold data; on a diagonal tile, not wholly below the
// diagonal); full off-diagonal tiles run unguarded.  Split-BF16 keeps the B hi / lo fragments
// of JB column blocks resident and streams A hi / lo one 16-row block at a time; each
// accumulator gets hi*hi, hi*lo, lo*hi.
template <int PREC, class S, int TILE, bool G>
__device__ __forceinline__ void gram_compute(floatx4_t (&acc)[S::NBI][S::NBJ],
                                             const char* base, int frag_off, int wr, int wc,
                                             int nbi, int nbj, bool diag) {
  constexpr int NBI = S::NBI, NBJ = S::NBJ, WM = S::WM, WN = S::WN;
  if constexpr (PREC == MEMI_PREC_SPLIT_BF16) {
    constexpr int JB = NBJ > 4 ? 4 : NBJ;    // column blocks per pass (bounds live fragments)
#pragma unroll
    for (int j0 = 0; j0 < NBJ; j0 += JB) {
      bf16x8_t fb[JB], fc[JB];
#pragma unroll
      for (int j = 0; j < JB; ++j) {
        const int c = (wc * WN + (j0 + j) * 16) * LDS_ROW + frag_off;
        fb[j] = *(const bf16x8_t*)(base + 1 * TILE + c);
        fc[j] = *(const bf16x8_t*)(base + 3 * TILE + c);
      }
#pragma unroll
      for (int i = 0; i < NBI; ++i) {
        const int r = (wr * WM + i * 16) * LDS_ROW + frag_off;
        const bf16x8_t ah = *(const bf16x8_t*)(base + 0 * TILE + r);
        const bf16x8_t al = *(const bf16x8_t*)(base + 2 * TILE + r);
#pragma unroll
        for (int j = 0; j < JB; ++j) {
          const int jj = j0 + j;
          if (!G || (i < nbi && jj < nbj && !(diag && wr * NBI + i > wc * NBJ + jj))) {
            acc[i][jj] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, fb[j], acc[i][jj], 0, 0, 0);
            acc[i][jj] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, fc[j], acc[i][jj], 0, 0, 0);
            acc[i][jj] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, fb[j], acc[i][jj], 0, 0, 0);
          }
        }
      }
    }
  } else {
    floatx4_t a[NBI], b[NBJ];
#pragma unroll
    for (int i = 0; i < NBI; ++i)
      a[i] = *(const floatx4_t*)(base + 0 * TILE + (wr * WM + i * 16) * LDS_ROW + frag_off);
#pragma unroll
    for (int j = 0; j < NBJ; ++j)
      b[j] = *(const floatx4_t*)(base + 1 * TILE + (wc * WN + j * 16) * LDS_ROW + frag_off);
#pragma unroll
    for (int e = 0; e < 4; ++e)
#pragma unroll
      for (int i = 0; i < NBI; ++i)
#pragma unroll
        for (int j = 0; j < NBJ; ++j)
          if (!G || (i < nbi && j < nbj && !(diag && wr * NBI + i > wc * NBJ + j)))
            acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(a[i][e], b[j][e], acc[i][j], 0, 0, 0);
  }
}

// The K loop over nks steps of 64 bytes per row: prefetch distance DEPTH (one or two register
// stages), two LDS buffers, one barrier per step.
template <int PREC, class S, int DEPTH, bool G>
__device__ __forceinline__ void gram_kloop(floatx4_t (&acc)[S::NBI][S::NBJ], char* lds,
                                           char* lds_b, int nks,
                                           int s0, int steps, int half, uint32_t seg_bytes,
                                           const char* gAh, const char* gBh, const char* gAl,
                                           const char* gBl, const uint32_t (&oA)[S::RPT],
                                           const uint32_t (&oB)[S::RPT], int lds_st,
                                           int frag_off, int wr, int wc, int nbi, int nbj,
                                           bool diag, int wvu) {
  using Cfg = GramCfg<PREC>;
  constexpr int NA = Cfg::NARR;
  constexpr int RPT = S::RPT;
  constexpr int TILE = S::BT * LDS_ROW;
  constexpr int RS = S::NT / 4;
  // byte offsets of K step ks (within the chunk) in the left and right operand rows
#define MEMI_GLOAD(R, ks)                                                                 \
  do {                                                                                    \
    const int gs_ = s0 + (ks);                                                            \
    const int sec_ = gs_ >= steps;                                                        \
    const int sl_ = 2 * half + sec_;                                                      \
    const uint32_t kk_ = (uint32_t)(gs_ - sec_ * steps) * ROW_STEP;                       \
    gram_gload<NA, RPT>(R, gAh, gBh, gAl, gBl, oA, oB, sl_ * seg_bytes + kk_,             \
                        seg_pair(sl_) * seg_bytes + kk_);                                 \
  } while (0)
  // The steady-state loops prefetch on every step with no condition around the loads: a load
  // that only some paths issue makes the compiler's wait-count merge at the join assume the
  // fewer outstanding loads, so the LDS store of the older stage would wait for the prefetch
  // just issued and the prefetch distance would collapse to one step's MFMAs.
#define MEMI_DLOAD(B, ks)                                                                 \
  do {                                                                                    \
    const int gs_ = s0 + (ks);                                                            \
    const int sec_ = gs_ >= steps;                                                        \
    const int sl_ = 2 * half + sec_;                                                      \
    const uint32_t kk_ = (uint32_t)(gs_ - sec_ * steps) * ROW_STEP;                       \
    gram_dload<NA, S, TILE>(B, gAh, gBh, gAl, gBl, oA, oB, sl_ * seg_bytes + kk_,         \
                            seg_pair(sl_) * seg_bytes + kk_, wvu);                        \
  } while (0)
  if constexpr (DEPTH == 3) {
    // direct-to-LDS, two buffers: the loads of step ks + 1 fly during step ks's MFMAs; each
    // wave waits for its own loads, then the barrier publishes the buffer
    char* const b0 = lds;
    char* const b1 = lds_b;
    MEMI_DLOAD(b0, 0);
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    int ks = 0;
    for (; ks + 2 < nks; ks += 2) {
      MEMI_DLOAD(b1, ks + 1);
      __builtin_amdgcn_sched_barrier(0);     // the loads go out before the MFMAs
      gram_compute<PREC, S, TILE, G>(acc, b0, frag_off, wr, wc, nbi, nbj, diag);
      __builtin_amdgcn_s_waitcnt(0);
      __syncthreads();
      MEMI_DLOAD(b0, ks + 2);
      __builtin_amdgcn_sched_barrier(0);
      gram_compute<PREC, S, TILE, G>(acc, b1, frag_off, wr, wc, nbi, nbj, diag);
      __builtin_amdgcn_s_waitcnt(0);
      __syncthreads();
    }
    if (ks + 1 < nks) {
      MEMI_DLOAD(b1, ks + 1);
      gram_compute<PREC, S, TILE, G>(acc, b0, frag_off, wr, wc, nbi, nbj, diag);
      __builtin_amdgcn_s_waitcnt(0);
      __syncthreads();
      gram_compute<PREC, S, TILE, G>(acc, b1, frag_off, wr, wc, nbi, nbj, diag);
    } else {
      gram_compute<PREC, S, TILE, G>(acc, b0, frag_off, wr, wc, nbi, nbj, diag);
    }
    return;
  }
#undef MEMI_DLOAD
  GramStage<NA, RPT> X;
  if constexpr (DEPTH == 1) {
    MEMI_GLOAD(X, 0);
    gram_lstore<NA, RPT, TILE, RS>(X, lds + lds_st);
    __syncthreads();
    int ks = 0;
    for (; ks + 1 < nks; ++ks) {
      const int cur = ks & 1;
      MEMI_GLOAD(X, ks + 1);
      __builtin_amdgcn_sched_barrier(0);     // issue the prefetch before this step's MFMAs
      gram_compute<PREC, S, TILE, G>(acc, lds + cur * NA * TILE, frag_off, wr, wc, nbi, nbj,
                                     diag);
      gram_lstore<NA, RPT, TILE, RS>(X, lds + (cur ^ 1) * NA * TILE + lds_st);
      __syncthreads();
    }
    gram_compute<PREC, S, TILE, G>(acc, lds + (ks & 1) * NA * TILE, frag_off, wr, wc, nbi, nbj,
                                   diag);
  } else {
    // prologue: step 0 into LDS buffer 0, step 1 in flight in Y; steady state unrolled by two
    // so the register stages alternate by name; at every loop top LDS buffer 0 holds step ks
    // and Y holds step ks + 1
    GramStage<NA, RPT> Y;
    MEMI_GLOAD(X, 0);
    if (nks > 1) MEMI_GLOAD(Y, 1);
    gram_lstore<NA, RPT, TILE, RS>(X, lds + lds_st);
    __syncthreads();
    int ks = 0;
    for (; ks + 3 < nks; ks += 2) {
      MEMI_GLOAD(X, ks + 2);
      __builtin_amdgcn_sched_barrier(0);     // issue the prefetch before this step's MFMAs
      gram_compute<PREC, S, TILE, G>(acc, lds, frag_off, wr, wc, nbi, nbj, diag);
      gram_lstore<NA, RPT, TILE, RS>(Y, lds + NA * TILE + lds_st);
      __syncthreads();
      MEMI_GLOAD(Y, ks + 3);
      __builtin_amdgcn_sched_barrier(0);
      gram_compute<PREC, S, TILE, G>(acc, lds + NA * TILE, frag_off, wr, wc, nbi, nbj, diag);
      gram_lstore<NA, RPT, TILE, RS>(X, lds + lds_st);
      __syncthreads();
    }
    // the last one to three steps
    while (ks < nks) {
      if (ks + 2 < nks) MEMI_GLOAD(X, ks + 2);
      gram_compute<PREC, S, TILE, G>(acc, lds, frag_off, wr, wc, nbi, nbj, diag);
      if (ks + 1 < nks) gram_lstore<NA, RPT, TILE, RS>(Y, lds + NA * TILE + lds_st);
      __syncthreads();
      ++ks;
      if (ks >= nks) break;
      if (ks + 2 < nks) MEMI_GLOAD(Y, ks + 2);
      gram_compute<PREC, S, TILE, G>(acc, lds + NA * TILE, frag_off, wr, wc, nbi, nbj, diag);
      if (ks + 1 < nks) gram_lstore<NA, RPT, TILE, RS>(X, lds + lds_st);
      __syncthreads();
      ++ks;
    }
  }
#undef MEMI_GLOAD
}

template <int PREC, class S, int DEPTH>
__global__ __launch_bounds__(S::NT, S::OCC) void k_gram_tiles(
    const char* __restrict__ op_hi, const char* __restrict__ op_lo, int64_t Khp, int spc,
    int steps, const int32_t* __restrict__ items, int64_t n_items,
    const memi_pd_desc* __restrict__ pds, float* __restrict__ ws) {
  using Cfg = GramCfg<PREC>;
  constexpr int ES = Cfg::ESIZE;
  constexpr int NA = Cfg::NARR;
  constexpr int BT = S::BT;
  constexpr int NBI = S::NBI, NBJ = S::NBJ;
  constexpr int RPT = S::RPT;
  constexpr int TILE = BT * LDS_ROW;         // LDS bytes of one array's K step
  static_assert(DEPTH >= 1 && DEPTH <= 3, "prefetch depth 1 or 2, or 3 = direct to LDS");
  // DEPTH 3 keeps its two buffers in separate arrays, so that the compiler can tell the
  // direct-to-LDS loads into one from the fragment reads of the other
  __shared__ __attribute__((aligned(16))) char lds[(DEPTH == 3 ? 1 : 2) * NA * TILE];
  __shared__ __attribute__((aligned(16))) char lds_b[DEPTH == 3 ? NA * TILE : 16];

  // XCD-aware order: workgroup w runs on XCD w % 8; give each XCD a contiguous range of
  // items (consecutive items share operand panels and K ranges).  gridDim.x % 8 == 0.
  const int64_t q8 = gridDim.x / 8;
  const int64_t it = (int64_t)(blockIdx.x % 8) * q8 + blockIdx.x / 8;
  if (it >= n_items) return;
  const int32_t* itm = items + it * 6;
  const int pd = itm[0], ti = itm[1], tj = itm[2], half = itm[3], chunk = itm[4];
  const int64_t slot = itm[5];
  const memi_pd_desc P = pds[pd];
  const int mi = min(BT, P.n - ti * BT);
  const int mj = min(BT, P.n - tj * BT);
  const int64_t row_bytes = 4 * Khp * ES;
  // chunk c covers steps [c*spc, min((c+1)*spc, 2*steps)) of the half's two segments; the
  // host guarantees none is empty
  const int s0 = chunk * spc;
  const int nks = min(spc, 2 * steps - s0);

  const int t = threadIdx.x;
  const int lane = t & 63, wv = t >> 6;
  const int wr = wv / S::WAVES_N, wc = wv % S::WAVES_N;

  // global -> LDS copy assignment: RPT rows x 16 bytes per thread per array, four lanes per
  // 64-byte row (coalesced).  Tile bases are uniform 64-bit pointers; per-thread offsets
  // within a tile fit 32 bits.
  const int ld_row0 = t >> 2, ld_part = t & 3;
  const int64_t tA = (P.row0 + (int64_t)ti * BT) * row_bytes;
  const int64_t tB = (P.row0 + (int64_t)tj * BT) * row_bytes;
  const char* const gAh = op_hi + tA;
  const char* const gBh = op_hi + tB;
  const char* const gAl = NA == 4 ? op_lo + tA : nullptr;
  const char* const gBl = NA == 4 ? op_lo + tB : nullptr;
  uint32_t oA[RPT], oB[RPT];
  const int wvu = __builtin_amdgcn_readfirstlane(wv);
  if constexpr (DEPTH == 3) {
    // row (wvu + q NW) * 16 + lane / 4, swizzled chunk, for the direct-to-LDS loads
    static_assert(S::BT == 16 * (S::NT / 64) * RPT, "16-row groups cover the tile");
    const int ch = (lane & 3) ^ swz(lane >> 2);
#pragma unroll
    for (int q = 0; q < RPT; ++q) {
      const int r = (wvu + q * (S::NT / 64)) * 16 + (lane >> 2);
      oA[q] = (uint32_t)(min(r, mi - 1) * row_bytes + ch * 16);
      oB[q] = (uint32_t)(min(r, mj - 1) * row_bytes + ch * 16);
    }
  } else {
#pragma unroll
    for (int q = 0; q < RPT; ++q) {
      const int r = ld_row0 + (S::NT / 4) * q;
      oA[q] = (uint32_t)(min(r, mi - 1) * row_bytes + ld_part * 16);
      oB[q] = (uint32_t)(min(r, mj - 1) * row_bytes + ld_part * 16);
    }
  }
  const int lds_st = ld_row0 * LDS_ROW + ((ld_part ^ swz(ld_row0)) * 16);
  const int frag_off = (lane & 15) * LDS_ROW + (((lane >> 4) ^ swz(lane & 15)) * 16);
  const int nbi = max(0, min(NBI, (mi - wr * S::WM + 15) / 16));
  const int nbj = max(0, min(NBJ, (mj - wc * S::WN + 15) / 16));
  const bool diag = ti == tj;

  floatx4_t acc[NBI][NBJ];
#pragma unroll
  for (int i = 0; i < NBI; ++i)
#pragma unroll
    for (int j = 0; j < NBJ; ++j) acc[i][j] = floatx4_t{0.f, 0.f, 0.f, 0.f};

  const uint32_t seg_bytes = (uint32_t)(Khp * ES);
  if (mi == BT && mj == BT && !diag)
    gram_kloop<PREC, S, DEPTH, false>(acc, lds, lds_b, nks, s0, steps, half, seg_bytes, gAh,
                                      gBh, gAl, gBl, oA, oB, lds_st, frag_off, wr, wc, nbi, nbj,
                                      diag, wvu);
  else
    gram_kloop<PREC, S, DEPTH, true>(acc, lds, lds_b, nks, s0, steps, half, seg_bytes, gAh, gBh,
                                     gAl, gBl, oA, oB, lds_st, frag_off, wr, wc, nbi, nbj, diag,
                                     wvu);

  // 16x16 output block: lane holds rows 4*(lane/16) + r, column lane % 16
  float* out = ws + slot * BT * BT;
#pragma unroll
  for (int i = 0; i < NBI; ++i)
#pragma unroll
    for (int j = 0; j < NBJ; ++j) {
      if (diag && wr * NBI + i > wc * NBJ + j) continue;   // never read by the finalize
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = wr * S::WM + i * 16 + (lane >> 4) * 4 + r;
        const int col = wc * S::WN + j * 16 + (lane & 15);
        out[row * BT + col] = acc[i][j][r];
      }
    }
}

// Reduce the P and R partials of one upper tile in float64, write D and its mirror, and queue
// entries with a large cancellation ratio for direct recomputation.  blockIdx.y: which
// TB_FIN-entry part of the BT x BT tile this block reduces.
__global__ __launch_bounds__(256) void k_gram_finalize(const int32_t* __restrict__ utiles,
                                                       const memi_pd_desc* __restrict__ pds,
                                                       const float* __restrict__ ws, int chunks,
                                                       int BT, double kappa,
                                                       float* __restrict__ D,
                                                       int32_t* __restrict__ flags,
                                                       int64_t flag_cap,
                                                       int32_t* __restrict__ flag_count) {
  const int32_t* u = utiles + (int64_t)blockIdx.x * 4;
  const int pd = u[0], ti = u[1], tj = u[2];
  const int64_t slot0 = u[3];
  const memi_pd_desc P = pds[pd];
  const int n = P.n;
  const int64_t tile = (int64_t)BT * BT;
  const int e0 = blockIdx.y * TB_FIN, e1 = e0 + TB_FIN;
  for (int e = e0 + threadIdx.x; e < e1; e += blockDim.x) {
    const int r = e / BT, c = e % BT;
    const int i = ti * BT + r, j = tj * BT + c;
    if (i >= n || j >= n) continue;
    if (ti == tj && c < r) continue;
    double d = 0.0;
    if (i != j) {
      double Pp = 0.0, R = 0.0;
      for (int ch = 0; ch < chunks; ++ch) {
        Pp += (double)ws[(slot0 + ch) * tile + e];
        R += (double)ws[(slot0 + chunks + ch) * tile + e];
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

int memi_plan(const memi_pd_desc* pds_in, int npd, int chunks, int gram_tile,
              void* plan_host, int64_t plan_cap, memi_plan_info* info) {
  MEMI_CHECK_ARG(info != nullptr && (npd == 0 || pds_in != nullptr), "memi_plan: null pointer");
  MEMI_CHECK_ARG(npd >= 0 && chunks >= 1 && chunks <= 64, "memi_plan: bad npd/chunks");
  MEMI_CHECK_ARG(gram_tile == 128 || gram_tile == 256,
                 "memi_plan: gram_tile must be 128 or 256 (got %d)", gram_tile);
  constexpr int FT = 128;             // side of the full-matrix tiles of the Stage-2 kernels
  const int BT = gram_tile;
  memi_plan_info I{};
  I.n_pd = npd;
  I.chunks = chunks;
  I.gram_tile = BT;
  int64_t d = 0, mw = 0, le = 0, rows = 0, items = 0, ut = 0, ft = 0, fc = 0;
  int max_n = 0;
  for (int p = 0; p < npd; ++p) {
    const int64_t n = pds_in[p].n;
    MEMI_CHECK_ARG(n >= 1 && n <= 65536 && pds_in[p].k >= 1,
                   "memi_plan: PD %d has n=%lld k=%d", p, (long long)n, pds_in[p].k);
    MEMI_CHECK_ARG(pds_in[p].row0 == rows,
                   "memi_plan: PD %d starts at row %lld; PD rows must be packed contiguously "
                   "in order (expected %lld)", p, (long long)pds_in[p].row0, (long long)rows);
    const int64_t T = (n + BT - 1) / BT;
    const int64_t TF = (n + FT - 1) / FT;
    d += n * n;
    mw += n * ((n + 31) / 32);
    le += (int64_t)std::min<int64_t>(pds_in[p].k, n) * n;
    rows += n;
    ut += T * (T + 1) / 2;
    ft += TF * TF;
    fc += n * (n - 1) / 2;
    if (n > max_n) max_n = (int)n;
  }
  items = ut * 2 * chunks;
  MEMI_CHECK_ARG(items < (1LL << 31), "memi_plan: too many Gram work items");
  I.n_items = items;
  I.n_utiles = ut;
  I.n_ftiles = ft;
  I.n_rows = rows;
  I.d_elems = d;
  I.mask_words = mw;
  I.list_elems = le;
  I.ws_floats = ut * 2 * chunks * (int64_t)BT * BT;
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
    const int T = (n + BT - 1) / BT;
    const int64_t u0 = uidx;
    for (int a = 0; a < T; ++a)
      for (int b = a; b < T; ++b) {
        int32_t* u = uts + 4 * uidx;
        u[0] = p; u[1] = a; u[2] = b; u[3] = (int32_t)(uidx * 2 * chunks);
        ++uidx;
      }
    // items: half-major, then chunk, then row panel, so neighbours share operand panels and
    // K ranges; slot = utile * 2 chunks + half * chunks + chunk
    for (int h = 0; h < 2; ++h)
      for (int c = 0; c < chunks; ++c) {
        int64_t uu = u0;
        for (int a = 0; a < T; ++a)
          for (int b = a; b < T; ++b) {
            int32_t* x = it + 6 * nit++;
            x[0] = p; x[1] = a; x[2] = b; x[3] = h; x[4] = c;
            x[5] = (int32_t)(uu * 2 * chunks + h * chunks + c);
            ++uu;
          }
      }
    const int TF = (n + FT - 1) / FT;
    for (int a = 0; a < TF; ++a)
      for (int b = 0; b < TF; ++b) {
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
  const int BT = info->gram_tile;
  MEMI_CHECK_ARG(BT == 128 || BT == 256, "memi_ctf_distance: plan has Gram tile %d", BT);
  MEMI_CHECK_ARG(variant >= 0 && variant <= 3 && !(BT == 256 && variant == 2),
                 "memi_ctf_distance: variant %d is not available for Gram tile %d", variant, BT);
  const int chunks = info->chunks;
  const int kstep = ROW_STEP / (precision == MEMI_PREC_FP32 ? 4 : 2);
  // K steps per segment; a half spans two segments, split into chunks of spc steps (the last
  // one may be shorter)
  const int64_t steps = Khp / kstep;
  const int64_t spc = (2 * steps + chunks - 1) / chunks;
  MEMI_CHECK_ARG(steps < (1LL << 29) && (int64_t)(chunks - 1) * spc < 2 * steps,
                 "memi_ctf_distance: %d chunks leave an empty chunk of the %lld K steps",
                 chunks, (long long)(2 * steps));
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
  // default schedule: prefetch two K steps for 128-wide tiles, one for 256-wide
  const int depth = variant == 0 ? (BT == 128 ? 2 : 1) : variant;
  const char* lo = precision == MEMI_PREC_FP32 ? nullptr : (const char*)op_lo;
#define MEMI_LAUNCH_GRAM(P, SH, V)                                                            \
  hipLaunchKernelGGL((k_gram_tiles<P, SH, V>), dim3((unsigned)grid), dim3(SH::NT), 0, s,       \
                     (const char*)op_hi, lo, Khp, (int)spc, (int)steps, items, info->n_items, \
                     pds, ws)
  if (precision == MEMI_PREC_FP32) {
    if (BT == 256 && depth == 3) MEMI_LAUNCH_GRAM(MEMI_PREC_FP32, Shape256, 3);
    else if (BT == 256) MEMI_LAUNCH_GRAM(MEMI_PREC_FP32, Shape256, 1);
    else if (depth == 1) MEMI_LAUNCH_GRAM(MEMI_PREC_FP32, Shape128, 1);
    else if (depth == 3) MEMI_LAUNCH_GRAM(MEMI_PREC_FP32, Shape128, 3);
    else MEMI_LAUNCH_GRAM(MEMI_PREC_FP32, Shape128, 2);
  } else {
    if (BT == 256 && depth == 3) MEMI_LAUNCH_GRAM(MEMI_PREC_SPLIT_BF16, Shape256, 3);
    else if (BT == 256) MEMI_LAUNCH_GRAM(MEMI_PREC_SPLIT_BF16, Shape256, 1);
    else if (depth == 1) MEMI_LAUNCH_GRAM(MEMI_PREC_SPLIT_BF16, Shape128, 1);
    else if (depth == 3) MEMI_LAUNCH_GRAM(MEMI_PREC_SPLIT_BF16, Shape128, 3);
    else MEMI_LAUNCH_GRAM(MEMI_PREC_SPLIT_BF16, Shape128, 2);
  }
#undef MEMI_LAUNCH_GRAM
  MEMI_CHECK_LAUNCH("k_gram_tiles");
  if (ev_stop && hipEventRecord((hipEvent_t)ev_stop, s) != hipSuccess) {
    set_error("memi_ctf_distance: hipEventRecord failed");
    return MEMI_ELAUNCH;
  }
  hipLaunchKernelGGL(k_gram_finalize, dim3((unsigned)info->n_utiles, BT * BT / TB_FIN),
                     dim3(256), 0, s, ut, pds, ws, chunks, BT, kappa, D, flags, info->flag_cap,
                     flag_count);
  MEMI_CHECK_LAUNCH("k_gram_finalize");
  hipLaunchKernelGGL(k_gram_recompute, dim3(1024), dim3(256), 0, s, flags, flag_count,
                     info->flag_cap, pds, (const float2*)fy, ctf_half, N, D);
  MEMI_CHECK_LAUNCH("k_gram_recompute");
  return MEMI_OK;
}

}  // extern "C"
