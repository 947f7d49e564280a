/*
 * libmemi - MI355X (gfx950) kernels for ManifoldEM's per-projection-direction hot path.
 *
 * C ABI: plain pointers and sizes; every array argument is DEVICE memory unless its name ends
 * in _host.  Every entry point enqueues work on `stream` (a hipStream_t passed as void*;
 * NULL = the default stream), returns 0 on success or a negative MEMI_E* code, and never
 * throws.  memi_last_error() gives the text of the last failure on the calling host thread.
 * Entry points keep no mutable global state: two calls on different streams do not interact
 * beyond the buffers they are given.
 *
 * Each declaration names the reference function (file:line in evanseitz/ManifoldEM_Python,
 * modules/) whose arithmetic it replaces.  The Python drop-in modules in
 * manifoldem_python_amd/ keep the reference call signatures and call these.
 *
 * Layout conventions
 *   image   : float32 [n][N][N], row-major, the reference's in-memory orientation
 *             (the SPIDER memmap transposed, getDistanceCTF_local_Conj9combinedS2.py:256-258)
 *   spectrum: complex64 [n][N][Nh] (interleaved re, im), Nh = N/2 + 1, the rfft2 layout;
 *             the reference's full fft2 spectra are Hermitian (real images, even CTF), so the
 *             half spectrum with column weights (1 for kx = 0 and kx = N/2, 2 otherwise)
 *             reproduces every full-spectrum sum.
 *   Kh      : N * Nh half-spectrum length; Khp = Kh rounded up to a multiple of 64.
 *   operand : the packed distance operand, one row per image, four segments of Khp each:
 *             [ w|X|^2 | C^2 | sqrt(w) Re(CX) | sqrt(w) Im(CX) ] (zero padded);
 *             float32 [rows][4*Khp] (precision FP32) or a bf16 hi/lo pair (SPLIT_BF16).
 *   PD batch: a memi_pd_desc array; PD p owns operand/image rows row0 .. row0+n-1 and the
 *             n*n block of every per-PD matrix starting at element d_off.
 */
#ifndef MEMI_H
#define MEMI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MEMI_OK 0
#define MEMI_EINVAL (-1)      /* bad argument (shape, size, null pointer) */
#define MEMI_ELAUNCH (-2)     /* a kernel launch or HIP runtime call failed */
#define MEMI_ENOTSUP (-3)     /* unsupported option */

#define MEMI_PREC_FP32 0       /* plain FP32 MFMA Gram (v_mfma_f32_16x16x4_f32) */
#define MEMI_PREC_SPLIT_BF16 1 /* split-BF16: hi*hi + hi*lo + lo*hi, FP32 accumulate */

#define MEMI_FILTER_BUTTER 1   /* sqrt(1 / (1 + (Q/Qc)^(2*order)))          */
#define MEMI_FILTER_GAUSS 2    /* exp(-(log 2 / 2) (Q/Qc)^2)                  */
#define MEMI_FILTER_BSPLINE 4  /* divide by the cubic B-spline transfer fn    */

typedef struct memi_pd_desc {
  int64_t row0;      /* in : first image / operand row of the PD                         */
  int64_t d_off;     /* out: element offset of the PD's n*n block in per-PD matrices     */
  int64_t mask_off;  /* out: word offset of the PD's kNN bitmask (n rows of ceil(n/32))   */
  int64_t list_off;  /* out: element offset of the PD's k*n kNN lists                    */
  int32_t n;         /* in : particles in the PD                                         */
  int32_t k;         /* in : nearest neighbours kept (k >= n: full kernel)               */
} memi_pd_desc;

/* Sizes and table offsets of a batch plan (memi_plan). */
typedef struct memi_plan_info {
  int64_t n_pd;
  int64_t n_items;     /* Gram work items: (PD, tile i, tile j, half, chunk)             */
  int64_t n_utiles;    /* gram_tile^2 tiles of the upper triangles                      */
  int64_t n_ftiles;    /* 128x128 tiles of the full matrices                            */
  int64_t n_rows;      /* sum of n                                                      */
  int64_t off_pds, off_items, off_utiles, off_ftiles, off_rows;  /* byte offsets         */
  int64_t plan_bytes;  /* size of the plan buffer                                       */
  int64_t d_elems;     /* sum n^2: per-PD matrices (D, L)                               */
  int64_t mask_words;  /* sum n*ceil(n/32): kNN bitmasks                                */
  int64_t list_elems;  /* sum min(k,n)*n: kNN lists                                     */
  int64_t ws_floats;   /* Gram partial-sum workspace (float32)                          */
  int64_t flag_cap;    /* sum n(n-1)/2: capacity of the recompute list (int32 triplets) */
  int32_t chunks;      /* K chunks per half (P or R: two operand segments)               */
  int32_t max_n;
  int32_t gram_tile;   /* side of the Gram's output tiles (128 or 256)                  */
  int32_t reserved;
} memi_plan_info;

typedef struct memi_ctf_params {
  double pix_size;   /* A per pixel           (p.pix_size)    */
  double cs_mm;      /* spherical aberration  (p.Cs)          */
  double ekv;        /* voltage kV            (p.EkV)         */
  double ampc;       /* amplitude contrast    (p.AmpContrast) */
  double gauss_env;  /* envelope half width; +inf = none (p.gaussEnv) */
} memi_ctf_params;

int memi_version(void);
const char* memi_last_error(void);

/* ---------------- Stage 1: image preparation (getDistanceCTF_local_Conj9combinedS2) ------ */

/* Load + normalise, getDistanceCTF_local_Conj9combinedS2.py:246-283 (annularMask.py:19-32):
 * out[i] = t.T (flipud'ed when conj[i] != 0) of stack[src[i]], normalised by the mean and
 * ddof-0 std of t*(1-msk) taken over all N*N pixels, msk the disk r < N/2 centred on
 * (N/2-1, N/2).  stack: float32 [n_stack][N][N] in SPIDER order. */
int memi_load_normalize(const float* stack, int64_t n_stack, const int64_t* src,
                        const int32_t* conj, int64_t n, int N, float* out, void* stream);

/* Sub-pixel shift of RELION particle images, getDistanceCTF_local_Conj9combinedS2.py:259-264:
 * out[i] = scipy.ndimage.shift(in[i], (shift0[i], shift1[i]), order=3, mode='wrap') (the
 * reference passes (sh[1] - 0.5, sh[0] - 0.5)): mirror-boundary cubic B-spline coefficients,
 * coordinates outside [0, N-1] wrapped with period N-1, mirrored taps; float64 arithmetic.
 * in: float32 [n][N][N] (mrcfile's image layout); out: float32 [n][N][N], transposed when
 * transpose_out != 0 (the SPIDER stack layout the loaders above read); in and out must not
 * alias.  lu: float64 [2N] LU factors of the mirror spline system, {m[0..N-1], 1/pivot[0..N-1]}
 * (diagonal 4, off-diagonals 1, first row [4 2], last row [2 4]); ws: float64 scratch of
 * n N^2 doubles. */
int memi_shift_images(const float* in, int64_t n, int N, const double* shift0,
                      const double* shift1, const double* lu, int transpose_out, float* out,
                      double* ws, int64_t ws_doubles, void* stream);

/* Volume-mask projection of one PD, CC/projectMask.py:118-140 (getDistanceCTF...:304-310):
 * proj[j][i] = sum_k f(T (i, j, k) + t) with f scipy's affine_transform interpolant of the
 * (already axis-swapped, CC/projectMask.py:120) volume vol float32 [L][L][L] -- order-3 spline,
 * mode 'nearest': the volume padded by 12 edge samples per side, half-sample-reflection spline
 * coefficients of the padded array, taps clamped to it -- and mask = proj > 1 as 1.0 / 0.0.
 * T: float64 [12] = {the 3x3 matrix row-major, the offset t}; lu: float64 [2(L+24)] LU factors
 * of the reflection spline system of length L + 24 (first / last rows [5 1] / [1 5]);
 * proj: float64 [L][L]; mask: float32 [L][L]; ws: float64 scratch of (L+24)^3 doubles. */
int memi_project_mask(const float* vol, int L, const double* T, const double* lu, double* proj,
                      float* mask, double* ws, int64_t ws_doubles, void* stream);

/* S2tessellation.py:61-65 (classS2): for each particle direction pts[i] (float64 [n][3]), the
 * index of the nearest bin centre (float64 [m][3]) by Euclidean distance, float64 arithmetic,
 * ties to the lower index; idx int32 [n].  Replaces the reference's sklearn ball-tree 1-NN. */
int memi_nearest_bins(const double* pts, int64_t n, const double* centres, int m, int32_t* idx,
                      void* stream);

/* Spectral multiply on rfft2 spectra, getDistanceCTF_local_Conj9combinedS2.py:285-293
 * (create_grid/create_filter :139-167) and the cubic-spline prefilter of rotatefill.py:24
 * (scipy ndimage.rotate order 3): kind is an OR of MEMI_FILTER_* flags. */
int memi_spectrum_scale(void* spec, int64_t n, int N, int kind, double Qc, int order,
                        void* stream);

/* img[i] *= disk mask (annularMask.op(0, N/2, N, N)) when disk != 0, and *= msk2 [N][N]
 * when msk2 != NULL; getDistanceCTF_local_Conj9combinedS2.py:326 and :344. */
int memi_mask_multiply(float* img, int64_t n, int N, int disk, const float* msk2,
                       void* stream);

/* In-plane rotation, rotatefill.py:21-28: order-3 B-spline interpolation of the
 * periodically extended coefficient image coef[i] (prefiltered with MEMI_FILTER_BSPLINE),
 * rotated by angle_deg[i] about ((N-1)/2, (N-1)/2) with scipy.ndimage.rotate's axes.
 * Equals the reference's 3x3-tile rotation up to the tile's ~0.27^(0.79 N) boundary term. */
int memi_spline_rotate(const float* coef, const double* angle_deg, int64_t n, int N,
                       float* out, void* stream);

/* Image preparation in one kernel, getDistanceCTF_local_Conj9combinedS2.py:244-349: for each
 * of the n images (stack index src[i], conjugate flag conj[i] or conj = NULL), load +
 * normalise (a1), Butterworth / Gaussian filter (a2, filter_kind as memi_spectrum_scale, no
 * B-spline bit), disk mask and the rotations by rot1_deg[i] then rot2_deg[i] (a4), the
 * half spectrum fy = rfft2 (complex64 [n][N][N/2+1], a6), the CTF (float32 [n][N(N/2+1)],
 * a5) and the packed operand rows row0 + i (as memi_ctf_pack).  img_out: float32 [n][N][N]
 * aligned images, or NULL.  The image stays in LDS from load to operand (mixed-radix FFTs in
 * LDS); 8 <= N <= MEMI_PREP_MAX_N with prime factors 2, 3, 5, 7.  ws: float64 scratch of
 * 6 N(N/2+1) doubles (per-bin tables). */
#define MEMI_PREP_MAX_N 160
int memi_prep_images(const float* stack, int64_t n_stack, const int64_t* src,
                     const int32_t* conj, const double* rot1_deg, const double* rot2_deg,
                     const double* df, int64_t n, int N, int filter_kind, double Qc,
                     int order, const memi_ctf_params* ctf, int precision, int64_t Khp,
                     int64_t row0, void* op_hi, void* op_lo, void* fy, float* ctf_half,
                     float* img_out, double* ws, int64_t ws_doubles, void* stream);

/* memi_prep_images without a workspace argument: the per-bin tables live in a static device
 * array of 6 * 160 * 81 doubles (622 KB, enough for N <= MEMI_PREP_MAX_N) that each GPU holds
 * until the process ends.  No counterpart in the reference: it is for callers that cannot
 * allocate device memory (code under HIP graph capture, serving loops).  It calls no
 * allocation routine; its first call on a device looks up that device's copy of the array.
 * Every call on one GPU shares that array (each call rewrites it in full before use), so calls
 * that overlap on one GPU (other streams, other host threads) race.  N beyond the array's
 * size is rejected with MEMI_EINVAL.  Other arguments and results as memi_prep_images. */
int memi_prep_images_noalloc(const float* stack, int64_t n_stack, const int64_t* src,
                             const int32_t* conj, const double* rot1_deg, const double* rot2_deg,
                             const double* df, int64_t n, int N, int filter_kind, double Qc,
                             int order, const memi_ctf_params* ctf, int precision, int64_t Khp,
                             int64_t row0, void* op_hi, void* op_lo, void* fy, float* ctf_half,
                             float* img_out, void* stream);

/* Fused in-plane alignment, getDistanceCTF_local_Conj9combinedS2.py:315-330 with
 * rotatefill.py:21-28 and annularMask.py:19-32: optionally the disk mask (disk != 0), then
 * the periodic cubic-spline rotation by rot1_deg[i] followed by the one by rot2_deg[i], each
 * image resident in LDS (exact recursive B-spline prefilter in place of the FFT division of
 * memi_spectrum_scale + memi_spline_rotate).  in, out: float32 [n][N][N] (may be the same
 * buffer); 8 <= N <= MEMI_ALIGN_MAX_N. */
#define MEMI_ALIGN_MAX_N 196
int memi_align_images(const float* in, const double* rot1_deg, const double* rot2_deg, int64_t n,
                      int N, int disk, float* out, void* stream);

/* CTF + packed distance operand, ctemh_cryoFrank.py:24-44 and
 * getDistanceCTF_local_Conj9combinedS2.py:336-344,392-396: from fy (rfft2 of the aligned
 * images) and per-image defocus df (A), writes ctf_half float32 [n][Kh] (C on the half
 * spectrum) and operand rows (op_hi only for FP32; op_hi + op_lo bf16 for SPLIT_BF16),
 * starting at operand row `row0`. */
int memi_ctf_pack(const void* fy, const double* df, int64_t n, int N,
                  const memi_ctf_params* ctf, int precision, int64_t Khp, int64_t row0,
                  void* op_hi, void* op_lo, float* ctf_half, void* stream);

/* Phase flip, getDistanceCTF_local_Conj9combinedS2.py:346: out = sign(C) * fy. */
int memi_phase_flip(const void* fy, const float* ctf_half, int64_t n, int N, void* out,
                    void* stream);

/* Wiener average numerator, getDistanceCTF_local_Conj9combinedS2.py:352-366,422-430:
 * out[k] = sum_i spec_i[k] * C_i[k] / -(sum_i C_i[k]^2 + 1/snr) over the half spectrum. */
int memi_wiener_average(const void* spec, const float* ctf_half, int64_t n, int N, double snr,
                        void* out, void* stream);

/* Per-pixel image moments over the image axis (getDistanceCTF...:348,367,400):
 * sum[p] = sum_i img[i][p], sumsq[p] = sum_i img[i][p]^2 (float64). */
int memi_image_moments(const float* img, int64_t n, int64_t npix, double* sum, double* sumsq,
                       void* stream);

/* ---------------- Stage 1: distance Gram (getDistanceCTF...:391-397) --------------------- */

/* Batch plan (host only).  From pds_in (row0, n, k per PD; rows packed contiguously, i.e.
 * row0 of PD p = sum of n over PDs before p) compute the table sizes into *info; when
 * plan_host is non-NULL (capacity plan_cap bytes >= info->plan_bytes) also write the tables:
 * PD descriptors with their offsets filled in, Gram work items over the gram_tile x gram_tile
 * upper-triangle tiles of every PD x 2 halves (P, R) x `chunks`, those upper tiles, the
 * 128 x 128 full tiles of the Stage-2 kernels and the row list.  gram_tile: 128 or 256.  The
 * caller copies the plan to device memory once and passes it to every batched entry point
 * below (plan_dev). */
int memi_plan(const memi_pd_desc* pds_in_host, int npd, int chunks, int gram_tile,
              void* plan_host, int64_t plan_cap, memi_plan_info* info);

/* CTF-weighted distances D_ij = sum_k |C_j X_ik - C_i X_jk|^2 for every PD of the plan,
 * getDistanceCTF_local_Conj9combinedS2.py:391-397 (dead direct form :106-122).  MFMA Gram of
 * the packed operand into partial tiles of its two halves (ws, info->ws_floats floats),
 * P = [A B].[B A]^T (the |X|^2 C^2 terms) and R = [Yr Yi].[Yr Yi]^T (the cross term), float64
 * reduction D = P - 2R, exact zero diagonal, and float32 direct-form recomputation of entries
 * whose cancellation ratio (P + 2|R|)/D exceeds `kappa` (flags: 3*info->flag_cap int32;
 * *flag_count: int32 on the device, zeroed here, gets the number flagged).  D: float32
 * [info->d_elems].  fy/ctf_half: the spectra and CTFs that produced the operand (rows indexed
 * like the operand).
 * Each half's K range (two segments of Khp elements, 64-byte steps) is split into
 * info->chunks chunks of ceil(2 steps / chunks) steps, the last one shorter; a chunk count
 * that leaves an empty chunk is rejected.
 * variant: Gram kernel schedule, 0 = default; for 128-wide tiles (two workgroups per CU)
 * 1 = prefetch one K step through registers, 2 = prefetch two (the default), 3 = load each K
 * step straight into LDS (global_load_lds) during the previous step's MFMAs; 256-wide tiles
 * (one workgroup per CU, 128 x 64 per wave) take 1 (the default) or 3.  Every schedule gives
 * the same D bit for bit.
 * ev_start / ev_stop: optional hipEvent_t recorded on `stream` right before and after the
 * MFMA Gram kernel (NULL: not recorded), for timing that kernel alone. */
int memi_ctf_distance(const void* plan_dev, const memi_plan_info* info, const void* op_hi,
                      const void* op_lo, int64_t Khp, int precision, double kappa,
                      const void* fy, const float* ctf_half, int N, float* D, float* ws,
                      int32_t* flags, int32_t* flag_count, int variant, void* ev_start,
                      void* ev_stop, void* stream);

/* ---------------- Stage 2: diffusion map (DMembeddingII / fergusonE / slaplacianonFly) ---- */

/* kNN selection, DMembeddingII.py:43-57: per row i of every PD, D_ii := -inf and the k
 * smallest entries of row i are kept (ties broken by lower column index).  Writes the
 * membership bitmask (bit j of row i at word mask_off + i*ceil(n/32) + j/32).  When idx_out /
 * val_out are non-NULL, also the sorted lists in the reference layout of yInd1 / yVal1
 * (k x n, column i = point i, first entry self with value 0) at list_off. */
int memi_knn_select(const void* plan_dev, const memi_plan_info* info, const float* D,
                    uint32_t* mask, int32_t* idx_out, float* val_out, void* stream);

/* Ferguson bandwidth sweep, fergusonE.py:25-43 with the graph of DMembeddingII.py:97-140:
 * for each PD and each logEps[e], sums over the kernel's entries x = v/(2 eps), v = 0 on the
 * zero entries of the kNN lists (D < 1e-6, self included) and v = D_ij on the symmetrised
 * kNN entries, of exp(-x) where x < thr = max(-log(0.01 * mean exp(-v/(2 max eps))), 10).
 * out: float64 [npd][n_eps][3] = {sum exp(-x), sum x exp(-x), #terms} followed by
 * [npd][2] = {#entries, sum exp(-v/(2 max eps))} (the threshold sums); zeroed here. */
int memi_ferguson_sweep(const void* plan_dev, const memi_plan_info* info, const float* D,
                        const uint32_t* mask, const double* logEps, int n_eps, double* out,
                        void* stream);

/* The same sums for a log-uniform eps grid (logEps[e] = le0 + e*dle, dle > 0) by bucket
 * moments: positive entry values go into log-uniform buckets whose edges contain every
 * eps's threshold cut, each bucket keeps the power sums of its values' offsets from its
 * centre, and each eps sums a 17-term series per bucket below its cut (~1e-14 relative to
 * the direct sums; cost O(n_eps * buckets) per PD instead of O(n_eps * entries)).  A PD whose
 * values span more than 544 buckets at the grid's step, or whose step is too coarse for the
 * 17-term series (thr (e^{dle/2} - 1) > 1), is summed by the direct sweep of
 * memi_ferguson_sweep instead, within the same call.
 * out as memi_ferguson_sweep; ws: float64 scratch of npd * (6 + 8 + 544 * 18) doubles. */
int memi_ferguson_moments(const void* plan_dev, const memi_plan_info* info, const float* D,
                          const uint32_t* mask, const double* logEps, int n_eps, double le0,
                          double dle, double* out, double* ws, int64_t ws_doubles,
                          void* stream);

/* Batched tanh fit of fergusonE.py:45-57: for each PD (block), Levenberg-Marquardt from p0
 * (float64 [npd][4]) of f = a3 + a2 tanh(a0 x + a1) to y (float64 [npd][m]) at x [m].
 * out float64 [npd][16] = {popt[4], J^T J at popt (upper triangle, 10), ss_res, iterations};
 * the caller forms pcov = inv(J^T J) ss_res/(m-4) and applies the reference's restart rule.
 * active: optional int32 [npd] (0 = skip). */
int memi_tanh_fit(const double* x, const double* y, int m, int npd, const double* p0,
                  const int32_t* active, int max_iter, double* out, void* stream);

/* The restart test of fergusonE.py:51-57 and the bandwidth of DMembeddingII.py:170 on the
 * device: fit is memi_tanh_fit's out for R starts per PD ([npd][R][16]); for each PD the first
 * start whose sum_i sqrt|diag(inv(J^T J))_i ss_res/(m-4)| <= max_res (100 in the reference; a
 * singular J^T J or a non-finite value fails) gives ok[p] = 1 + its index and
 * sigma[p] = tune sqrt(2 exp(-a1/a0)); ok[p] = 0 (sigma untouched) when none passes.
 * res: optional float64 [npd][R], every start's value. */
int memi_tanh_select(const double* fit, int npd, int R, int m, double tune, double max_res,
                     int32_t* ok, double* sigma, double* res, void* stream);

/* Markov normalisation, slaplacianonFly.py:42-81 with alpha = 1: W = exp(-v/sigma^2) on the
 * kernel's entries, d = column sums, W/(d_i d_j), d' = sqrt(column sums), L = |L + L^T| / 2.
 * Sums in float64.  sigma: float64 [npd]; L: [info->d_elems] float64 (l_is_f64 = 1, the
 * float64 kernel of the reference) or float32 (l_is_f64 = 0: each entry rounded once, half the
 * bytes for the Lanczos matvecs); deg: float64 scratch [2 * n_rows]. */
int memi_markov_normalize(const void* plan_dev, const memi_plan_info* info, const float* D,
                          const uint32_t* mask, const double* sigma, void* L, int l_is_f64,
                          double* deg, void* stream);

/* fergusonE.py:25-43 on a flat value list (the reference's D argument; squared here):
 * out float64 [n_eps][3] = {sum exp(-x), sum x exp(-x), #terms} for x = D^2/(2 eps) < thr,
 * followed by 1 slot for the threshold sum; zeroed here. */
int memi_ferguson_values(const double* vals, int64_t nv, const double* logEps, int n_eps,
                         double* out, void* stream);

/* Sub-matrices for trimming rounds (manifoldTrimmingAuto.py:63-65, D1 = D[posPath1][:, posPath1]
 * for every PD of a batch at once): block p of dst (the plan's n_p x n_p block at d_off) gets
 * src[src_off[p] + sel[row0_p + a] * src_n[p] + sel[row0_p + b]] at (a, b).  src_off int64
 * [npd], src_n int32 [npd] and sel int32 [n_rows] are device arrays; the caller guarantees
 * 0 <= sel < src_n[p] and that each source block lies inside src. */
int memi_gather_blocks(const void* plan_dev, const memi_plan_info* info, const float* src,
                       const int64_t* src_off, const int32_t* src_n, const int32_t* sel,
                       float* dst, void* stream);

/* slaplacianonFly.py:42-78 on COO triplets (rows, cols, squared distances): out[e] =
 * W_e / (d_r d_c) / (d'_r d'_c) with W = exp(-val/sigma^2), d = column sums (^alpha),
 * d' = sqrt(column sums of W/(d d)); duplicates behave as the reference's csc sums.
 * deg: float64 scratch [2*nS].  The caller forms |l + l^T| / 2. */
int memi_markov_coo(const int64_t* row, const int64_t* col, const double* val, int64_t nnz,
                    int64_t nS, double sigma, double alpha, double* out, double* deg,
                    void* stream);

/* The start of a Lanczos run (sembeddingonFly.py:17-39 starts ARPACK from a start vector) in
 * one launch: V[idx[i]] = v0[i] for i < n_rows (column 0 of every PD's basis), ab[0..n_ab) = 0,
 * j[0] = j[1] = 0, active[0..npd) = 1.  idx must lie inside V. */
int memi_lanczos_init(double* V, const double* v0, const int64_t* idx, int64_t n_rows,
                      double* ab, int64_t n_ab, int32_t* j, int32_t* active, int64_t npd,
                      void* stream);

/* nsteps Lanczos steps (full reorthogonalisation: the three-term recurrence, then one classical
 * Gram-Schmidt pass against the whole basis) on every active PD's L, for the
 * largest-magnitude eigenpairs that sembeddingonFly.py:27 asks ARPACK eigsh for.
 * L: float32 (l_is_f64 = 0, as memi_markov_normalize writes it) or float64 (l_is_f64 = 1, a
 * caller's own float64 kernel); products and sums in float64.  V: float64 bases, column-major per PD at m*row0 (column 0 = unit start vector, set by the
 * caller); W: float64 [n_rows] scratch ([2 * n_rows] with split = 1); ab: float64 [npd][m][2]
 * gets {alpha_j, beta_j};
 * active: int32 [npd] (0 = skip).  The step index j lives in device memory: j_dev is int32
 * [2] = {j, arrival counter}, both zero before the first step.  Step j reads column j, writes
 * column j+1, and its last kernel advances j_dev[0] (the last of its workgroups to finish);
 * steps with j >= min(n, m) do nothing -- so the launches of a block of steps can be captured
 * once in a HIP graph and replayed.
 * split = 0: one workgroup per PD orthogonalises (H unused, may be NULL); split = 1: the
 * orthogonalisation is spread over the grid (for few large PDs), H: float64 [npd*(m+2)]. */
int memi_lanczos_steps(const void* plan_dev, const memi_plan_info* info, const void* L,
                       int l_is_f64, double* V, int m, int32_t* j_dev, int nsteps,
                       const int32_t* active, double* W, double* H, double* ab, int split,
                       void* stream);

/* The sparse form of L that sembeddingonFly.py:27 hands eigsh (a scipy sparse matrix): row i of
 * PD p keeps its nonzero entries at the dense row's offset d_off + i*n, columns in increasing
 * order in ell_col (int32 [d_elems]) and values in ell_val (float64 [d_elems]), ell_cnt[row0 + i]
 * of them (int32 [n_rows]).  L float64 as memi_markov_normalize writes it.  sparse_pd (int32
 * [npd], may be null = every PD): only the rows of PDs with sparse_pd[p] != 0 are written. */
int memi_kernel_ell(const void* plan_dev, const memi_plan_info* info, const double* L,
                    const int32_t* sparse_pd, int32_t* ell_col, double* ell_val,
                    int32_t* ell_cnt, void* stream);

/* memi_lanczos_steps with the matvec over memi_kernel_ell's sparse rows (the same steps and
 * outputs; the products of each row are summed over its nonzeros only).  For kNN kernels with
 * k much smaller than n.  sparse_pd null: every PD is sparse (L may be null).  Otherwise the
 * PDs with sparse_pd[p] != 0 (memi_kernel_ell called with the same selection) use the sparse
 * rows and the others the dense float64 L (required). */
int memi_lanczos_steps_sparse(const void* plan_dev, const memi_plan_info* info,
                              const int32_t* ell_col, const double* ell_val,
                              const int32_t* ell_cnt, const int32_t* sparse_pd, const double* L,
                              double* V, int m, int32_t* j_dev, int nsteps,
                              const int32_t* active, double* W, double* H, double* ab,
                              int split, void* stream);

/* Convergence check of the batched Lanczos run, the Ritz pairs eigsh(which='LM') returns
 * (sembeddingonFly.py:27-35), entirely on the GPU.  For every active PD with
 * J = min(steps_done, n, m): the k = ks[pd] (<= MEMI_RITZ_KMAX) eigenvalues of largest
 * magnitude of the tridiagonal T (ab as written by memi_lanczos_steps) by Sturm-count
 * multisection, sorted descending, their eigenvectors by twisted factorisation + modified
 * Gram-Schmidt, and the residuals |beta_{J-1} s_{J-1}|; res[pd] = the largest residual (0 when
 * J = n) at every check.  A PD is done when every residual is <= tol * max|theta|, or J = n,
 * or J = m: then active[pd] = 0 and steps[pd] = J.  theta: float64 [npd][16] (every check);
 * S: float64 [npd][m][16] (PD p's J x k row-major
 * block at p*m*16, ready for memi_ritz_vectors); piv: float64 scratch [npd][16][2][m].
 * summary (may be null): float64 [3][npd] = {still active (1 or 0), res, max |theta|} of every
 * PD after this check (0, 0, 0 for PDs inactive on entry) -- what the caller reads back to
 * decide and pace the next check, in one copy. */
int memi_lanczos_ritz(const void* plan_dev, const memi_plan_info* info, const double* ab, int m,
                      int steps_done, const int32_t* ks, double tol, int32_t* active,
                      double* theta, double* S, int32_t* steps, double* res, double* piv,
                      double* summary, void* stream);

/* Ritz vectors of every PD after Lanczos, the eigenvectors sembeddingonFly.py:27-35 takes
 * from eigsh: Y_p = V_p S_p with V_p the basis of memi_lanczos_step (first J columns) and S_p
 * the J x k eigenvectors of the tridiagonal T_p, row-major.  meta: int64 on the device,
 * [npd][3] = {J, k, s_off} (the caller checks J <= min(n, m), k <= kmax, s_off + J*k within S);
 * S: float64 [sum J*k]; Y: float64 [n_rows][kmax], kmax <= MEMI_RITZ_KMAX, columns from k on
 * zero. */
#define MEMI_RITZ_KMAX 16
int memi_ritz_vectors(const void* plan_dev, const memi_plan_info* info, const double* V, int m,
                      const double* S, const int64_t* meta, int kmax, double* Y, void* stream);

/* ---------------- Stage 3: NLSA (NLSA.py / get_wiener.py / svdRF.py / L2_distance.py) ---- */

/* Reordered distances and delay embedding, psiAnalysisParS2.py:89-90 and NLSA.py:31-34:
 * ConD[i][j] = sum_{c < con_order} D[ord[i+c]][ord[j+c]], summed in increasing c (float64),
 * D [n][n] float32 (d_is_f64 = 0) or float64 (d_is_f64 = 1), ConD float64 [m][m],
 * m = n - con_order, 1 <= con_order <= 97. */
int memi_nlsa_condist(const void* D, int d_is_f64, const int64_t* ord, int n, int con_order,
                      double* ConD, void* stream);

/* Wiener-filtered supervector spectra, NLSA.py:50-86 and get_wiener.py:9-21 by linearity:
 * out[ii][c][k] = sum_{i<m} mu_psi[i][c] * G[s+i][k] / W_i[k], s = con_order-1-ii,
 * W_i = sum_{j=i+1}^{i+con_order} C_j^2 + 1/snr, G_j = spec_j * C_j (images and CTFs already
 * in psi-sorted order, half spectrum).  out: complex64 [con_order][ell][Kh]. */
int memi_nlsa_supervectors(const void* spec, const float* ctf_half, int n, int N,
                           int con_order, const double* mu_psi, int ell, double snr, void* out,
                           void* stream);

/* Tall-skinny products for svdRF.py:9-35: G = A^T A (A float64 [M][c], G [c][c]). */
int memi_tall_gram(const double* A, int64_t M, int c, double* G, void* stream);
/* Y = A * B (A float64 [M][c], B [c][d] given on the host, Y [M][d]). */
int memi_tall_apply(const double* A, int64_t M, int c, const double* B_host, int d, double* Y,
                    void* stream);

/* Rank-2 reconstruction, NLSA.py:96-138 in factored form: IMGT[p][j] =
 * (sum_q Ubar[p][q] H[q][j]) * inv_sd[j], Ubar float64 [npix][q_dim], H float64
 * [q_dim][ncols], inv_sd [ncols], IMGT float64 [npix][ncols]. */
int memi_nlsa_imgt(const double* Ubar, const double* H, const double* inv_sd, int64_t npix,
                   int q_dim, int ncols, double* IMGT, void* stream);

/* Distances between the z-scored reconstructed frames, NLSA.py:112-146 with
 * L2_distance.py:8-41, from the factored Gram: T = Gq H, Gram = H^T T (Gq float64
 * [q_dim][q_dim] = Ubar^T Ubar, H float64 [q_dim][ncols]), inv_sd[j] = 1/sqrt(Gram_jj / npix),
 * C_ij = Gram_ij inv_sd_i inv_sd_j, d_ij = C_ii + C_jj - 2 C_ij with d < 1e-8 set to 0,
 * Drecon2[i][j] = (float)(sqrt(d_ij)^2) (float32 [ncols][ncols]).
 * ws: float64 scratch of (q_dim + 1) * ncols.  Replaces the host products and L2_distance. */
int memi_nlsa_recon_dist(const double* Gq, const double* H, int q_dim, int ncols, int64_t npix,
                         double* inv_sd, float* Drecon2, double* ws, void* stream);

/* Column-centre U2 (float64 [npix][q_dim]) in place and return its Gram Ubar^T Ubar
 * (float64 [q_dim][q_dim]); NLSA.py:112-138 + L2_distance.py:26-41 in factored form. */
int memi_center_gram(double* U2, int64_t npix, int q_dim, double* G, void* stream);

/* One tau update of the 1-D open-manifold fit, fit_1D_open_manifold_3D.py:64-160 with
 * solve_d_R_d_tau_p_3D.py:35-60 and R_p.py:41-49 (replaces the per-point np.roots loop):
 * for every point x_i (float64 [n][3]) and the curve ab_host = {a1, a2, a3, b1, b2, b3}
 * (host), tau[i] = the candidate of least R_p among arccos(real roots in [-1, 1] of the
 * quintic)/pi, 0 and 1.  flag[i] (int32) = 1 where the reference's LAPACK root classification
 * or arg min is within rounding of a tie (near-double root, root at +-1, equal R, zero leading
 * or trailing coefficient): the caller re-solves those points with np.roots.
 * sums (float64 [13], device): sum cos^2(j pi tau) [3], sum cos(j pi tau) [3],
 * sum x_j cos(j pi tau) [3], sum x_j [3], number flagged [1].
 * ws: float64 scratch of 13 * ceil(n / 256). */
int memi_fit1d_tau(const double* x, int n, const double* ab_host, double* tau, int32_t* flag,
                   double* sums, double* ws, void* stream);

/* ---------------- Differentiation of the fused distance -> kernel op --------------------- */

/* Forward: for each eps_e, S[e][i] = sum_j exp(-D_ij / (2 eps_e)) and
 * T[e][i] = sum_j exp(-D_ij/(2 eps_e)) * D_ij (float64), D float32 [n][n]
 * (the kernel of fergusonE.py:40). */
int memi_kernel_rowsums(const float* D, int n, const double* eps, int n_eps, double* S,
                        double* T, void* stream);

/* Backward helper: Wt[i][j] = sum_e g[e] * (-1/(2 eps_e)) * (K_ij/S[e][i] + K_ji/S[e][j]),
 * K recomputed from D (float32 [n][n]); Wt float32 [n][n]. */
int memi_kernel_grad_weights(const float* D, int n, const double* eps, int n_eps,
                             const double* S, const double* g, float* Wt, void* stream);

/* ---------------- Optical flow of NLSA movies (CC stage) --------------------------------- */
/* Movies are float32 [nmov][F][H][W] (frame f of a movie is the reference's
 * np.resize(Mov, (F, dim, dim))[f]).  Arithmetic follows numpy/scipy: float32 where numpy
 * computes in float32, float64 accumulation rounded once where scipy.ndimage filters. */

/* OpticalFlowMovie.py:202-341 (anisodiff3): niter - 1 Perona-Malik updates of every movie
 * (option 1: exp(-(d/kappa)^2), option 2: 1/(1 + (d/kappa)^2); step = voxel spacing (z, y, x)).
 * The result lands in out; tmp (same size) is needed when niter > 2.  in, out, tmp distinct. */
int memi_flow_anisodiff(const float* in, float* out, float* tmp, int64_t nmov, int F, int H,
                        int W, int niter, double kappa, double gamma, double step0, double step1,
                        double step2, int option, void* stream);

/* OpticalFlowMovie.py:447-457: out float64 [nmov][ceil(F/bs)][hw], the mean of each block of bs
 * consecutive frames (the last block may be shorter). */
int memi_flow_block_average(const float* in, int64_t nmov, int F, int64_t hw, int bs,
                            double* out, void* stream);

/* hornschunck_simple.py:35-38 (lowpassfilt = scipy gaussian_filter, truncate 4, mode
 * 'reflect') on count images [H][W]: float64 -> float64, float64 -> float32 (the input is
 * rounded to float32 first, as the reference's astype) or float32 -> float32.  tmp holds one
 * pass in the output type (count*H*W elements). */
int memi_gaussian_filter(const void* in, int in_f64, void* out, int out_f64, void* tmp,
                         int64_t count, int H, int W, double sigma, void* stream);

/* hornschunck_simple.py:24-31 (computeDerivatives) for nP frame pairs: pairs int32 [nP][2]
 * holds the frame indices (im1, im2) into frames float32 [*][H][W]; fx, fy, ft float32
 * [nP][H][W]. */
int memi_hs_derivatives(const float* frames, const int32_t* pairs, int64_t nP, int H, int W,
                        float* fx, float* fy, float* ft, void* stream);

/* hornschunck_simple.py:70-82: niter Horn-Schunck iterations in place on U, V [nP][H][W]
 * (float64 when is_f64, else float32 -- the reference's dtype follows the initial flow's);
 * tmpU, tmpV are buffers of the same size. */
int memi_hs_iterate(const float* fx, const float* fy, const float* ft, int64_t nP, int H, int W,
                    double alpha, int niter, void* U, void* V, int is_f64, void* tmpU,
                    void* tmpV, void* stream);

/* OpticalFlowMovie.py:519-534 and getOrientMag :21-24: vx, vy float64 [nmov][hw] = the sum of
 * U, V [nmov][A][hw] over A frame pairs (from 0.0 when from_zero, as the reference's
 * VxM = zeros + Vx), orient = (atan2(vy, vx) % 2) * pi, mag = sqrt(vx^2 + vy^2), orient_rev the
 * orientation of (-vx, -vy).  orient, mag and orient_rev may be null. */
int memi_flow_sum(const void* U, const void* V, int is_f64, int64_t nmov, int A, int64_t hw,
                  int from_zero, double* vx, double* vy, double* orient, double* mag,
                  double* orient_rev, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* MEMI_H */
