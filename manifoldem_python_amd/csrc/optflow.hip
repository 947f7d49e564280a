// Optical flow of NLSA movies (the CC stage): 3-D Perona-Malik diffusion of each movie, block
// averaging of its frames, Gaussian low-pass, Horn-Schunck flow between consecutive averaged
// frames, and the sum of the flows with its orientation and magnitude.
//
// Reference: CC/hornschunck_simple.py:9-84, CC/OpticalFlowMovie.py:21-24, 202-341, 385-555.
//
// Every kernel reproduces the reference's arithmetic: float32 where numpy computes in float32
// (the diffusion, the derivative images, the denominator alpha^2 + fx^2 + fy^2), float64
// accumulation rounded once to the output type where scipy.ndimage filters (it accumulates each
// output in double), and the reference's operation order with floating-point contraction off.
// Boundaries are scipy's mode 'reflect' (half-sample mirror), which at distance 1 is a clamp.
//
// Data layout: a batch of movies is (movie, frame, row, column) row-major; each frame is the
// reference's Mov[f] = np.resize(row, (dim, dim)).
#include "common.h"

namespace memi {

// --- 3-D anisotropic diffusion (OpticalFlowMovie.py:270-336) -----------------------------------

// g(delta) * delta of one face: option 1 exp(-(d/kappa)^2)/step, option 2 1/(1+(d/kappa)^2)/step.
// numpy evaluates the float32 exp; here it is the float64 exp rounded to float32.
template <int OPTION>
__device__ inline float pm_flux(float d, float kappa, float step) {
#pragma clang fp contract(off)
  const float q = d / kappa;
  const float q2 = q * q;
  float g;
  if (OPTION == 1) g = (float)exp(-(double)q2);
  else g = 1.0f / (1.0f + q2);
  g = g / step;
  return g * d;
}

// One update of every voxel: out = s + gamma * ((UD + NS) + EW), where UD = D(z) - D(z-1) with
// D(z) = g(s[z+1]-s[z]) (s[z+1]-s[z]) for z < F-1 and 0 on the last frame (no flux through the
// stack's faces); NS along rows, EW along columns alike.
template <int OPTION>
__global__ void k_anisodiff_step(const float* __restrict__ in, float* __restrict__ out,
                                 int64_t nvox, int F, int H, int W, float kappa, float gamma,
                                 float s0, float s1, float s2) {
#pragma clang fp contract(off)
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= nvox) return;
  const int x = (int)(t % W);
  int64_t r = t / W;
  const int y = (int)(r % H);
  const int z = (int)((r / H) % F);
  const int64_t plane = (int64_t)H * W;
  const float c = in[t];
  const float Dz = z < F - 1 ? pm_flux<OPTION>(in[t + plane] - c, kappa, s0) : 0.0f;
  const float UD = z > 0 ? Dz - pm_flux<OPTION>(c - in[t - plane], kappa, s0) : Dz;
  const float Sy = y < H - 1 ? pm_flux<OPTION>(in[t + W] - c, kappa, s1) : 0.0f;
  const float NS = y > 0 ? Sy - pm_flux<OPTION>(c - in[t - W], kappa, s1) : Sy;
  const float Ex = x < W - 1 ? pm_flux<OPTION>(in[t + 1] - c, kappa, s2) : 0.0f;
  const float EW = x > 0 ? Ex - pm_flux<OPTION>(c - in[t - 1], kappa, s2) : Ex;
  out[t] = c + gamma * ((UD + NS) + EW);
}

// --- block averages of frames (OpticalFlowMovie.py:447-457) -------------------------------------

// out[m][b][p] = (sum of in[m][f][p] over the block's frames, in frame order, float64) / count:
// the float64 np.mean of float32 frames.
__global__ void k_block_average(const float* __restrict__ in, int64_t nmov, int F, int64_t hw,
                                int bs, int A, double* __restrict__ out) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= nmov * A * hw) return;
  const int64_t p = t % hw;
  const int64_t mb = t / hw;
  const int b = (int)(mb % A);
  const int64_t m = mb / A;
  const int f0 = b * bs;
  const int f1 = min(f0 + bs, F);
  const float* src = in + (m * F + f0) * hw + p;
  double s = (double)src[0];
  for (int f = 1; f < f1 - f0; ++f) s += (double)src[(int64_t)f * hw];
  out[t] = s / (double)(f1 - f0);
}

// --- Gaussian low-pass (hornschunck_simple.py:35-38, scipy gaussian_filter) ---------------------

constexpr int kGaussMaxR = 32;
struct GaussTaps {
  double w[kGaussMaxR + 1];  // w[d] = weight at distance d (normalised, symmetric)
  int r;
};

__device__ inline int reflect_index(int i, int n) {
  const int m = 2 * n;
  i %= m;
  if (i < 0) i += m;
  return i < n ? i : m - 1 - i;
}

// One separable pass along AXIS (0: down the rows, 1: along a row).  The input is first rounded
// to the output type (the reference casts before it filters; the second pass reads the first's
// output), then out = w0 c + sum_{d = r..1} (in[-d] + in[+d]) w_d in float64, rounded once.
template <typename TI, typename TO, int AXIS>
__global__ void k_gauss_pass(const TI* __restrict__ in, TO* __restrict__ out, int64_t count,
                             int H, int W, GaussTaps g) {
#pragma clang fp contract(off)
  const int64_t hw = (int64_t)H * W;
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= count * hw) return;
  const int x = (int)(t % W);
  const int y = (int)((t / W) % H);
  const TI* img = in + (t / hw) * hw;
  auto ld = [&](int yy, int xx) { return (double)(TO)img[(int64_t)yy * W + xx]; };
  double acc = ld(y, x) * g.w[0];
  for (int d = g.r; d >= 1; --d) {
    double a, b;
    if (AXIS == 0) {
      a = ld(reflect_index(y - d, H), x);
      b = ld(reflect_index(y + d, H), x);
    } else {
      a = ld(y, reflect_index(x - d, W));
      b = ld(y, reflect_index(x + d, W));
    }
    acc += (a + b) * g.w[d];
  }
  out[t] = (TO)acc;
}

// --- Horn-Schunck derivatives (hornschunck_simple.py:24-31) ------------------------------------

// scipy's convolve with a 2 x 2 kernel k: out[i][j] = k11 a[i][j] + k10 a[i][j+1]
// + k01 a[i+1][j] + k00 a[i+1][j+1] (float64, rounded to float32), the last row and column
// mirrored onto themselves.  fx, fy, ft are float32 sums of the two frames' convolutions.
__global__ void k_hs_derivatives(const float* __restrict__ frames,
                                 const int32_t* __restrict__ pairs, int64_t nP, int H, int W,
                                 float* __restrict__ fx, float* __restrict__ fy,
                                 float* __restrict__ ft) {
#pragma clang fp contract(off)
  const int64_t hw = (int64_t)H * W;
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= nP * hw) return;
  const int x = (int)(t % W);
  const int y = (int)((t / W) % H);
  const int64_t p = t / hw;
  const int x1 = min(x + 1, W - 1);
  const int y1 = min(y + 1, H - 1);
  double v[2][4];
  for (int f = 0; f < 2; ++f) {
    const float* im = frames + (int64_t)pairs[2 * p + f] * hw;
    v[f][0] = im[(int64_t)y * W + x];
    v[f][1] = im[(int64_t)y * W + x1];
    v[f][2] = im[(int64_t)y1 * W + x];
    v[f][3] = im[(int64_t)y1 * W + x1];
  }
  float cx[2], cy[2], ct[2];
  for (int f = 0; f < 2; ++f) {
    double s = 0.0;
    s += 0.25 * v[f][0]; s += -0.25 * v[f][1]; s += 0.25 * v[f][2]; s += -0.25 * v[f][3];
    cx[f] = (float)s;
    s = 0.0;
    s += 0.25 * v[f][0]; s += 0.25 * v[f][1]; s += -0.25 * v[f][2]; s += -0.25 * v[f][3];
    cy[f] = (float)s;
    const double kt = f == 0 ? 0.25 : -0.25;
    s = 0.0;
    s += kt * v[f][0]; s += kt * v[f][1]; s += kt * v[f][2]; s += kt * v[f][3];
    ct[f] = (float)s;
  }
  fx[t] = cx[0] + cx[1];
  fy[t] = cy[0] + cy[1];
  ft[t] = ct[0] + ct[1];
}

// --- Horn-Schunck iterations (hornschunck_simple.py:70-82) -------------------------------------

// `iters` (<= R) iterations per launch on a TS x TS output tile, temporally blocked in LDS: the
// tile plus an R-pixel halo of U, V, fx, fy, ft is loaded once; sub-iteration s updates the
// cells within iters - s of the tile (their 3 x 3 neighbours were updated by sub-iteration
// s - 1), so after `iters` sub-iterations the tile holds the exact iterate.  A neighbour outside
// the image is the nearest edge cell (mode 'reflect' at distance 1).
//
//   uAvg = sum of the 8 neighbours x HSKERN (scipy's raster order, float64, rounded to T)
//   der  = ((fx uAvg + fy vAvg) + ft) / den,   den = (float)((alpha^2 + fx^2) + fy^2)
//   U = uAvg - fx der,  V = vAvg - fy der       (in T: float64 for a float64 start, float32 for
//                                                a float32 start, as numpy promotes)
template <typename T, int TS, int R, int NT>
__global__ __launch_bounds__(NT) void k_hs_iterate(
    const float* __restrict__ gfx, const float* __restrict__ gfy, const float* __restrict__ gft,
    int H, int W, int tiles_x, int ntiles, float alpha2, int iters,
    const T* __restrict__ Uin, const T* __restrict__ Vin, T* __restrict__ Uout,
    T* __restrict__ Vout) {
#pragma clang fp contract(off)
  constexpr int E = TS + 2 * R;
  constexpr int NC = E * E;
  constexpr int PER = (NC + NT - 1) / NT;
  __shared__ T sU[NC];
  __shared__ T sV[NC];
  __shared__ float sfx[NC];
  __shared__ float sfy[NC];
  __shared__ float sft[NC];
  const int64_t p = blockIdx.x / ntiles;
  const int tile = (int)(blockIdx.x % ntiles);
  const int ty0 = (tile / tiles_x) * TS;
  const int tx0 = (tile % tiles_x) * TS;
  const int64_t base = p * (int64_t)H * W;

  // per owned cell: LDS index, image coordinates, Chebyshev distance to the tile (-1: outside
  // the image)
  int dist[PER];
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    const int c = threadIdx.x + k * NT;
    const int ly = c / E, lx = c % E;
    const int gy = ty0 - R + ly, gx = tx0 - R + lx;
    int d = -1;
    if (c < NC && gy >= 0 && gy < H && gx >= 0 && gx < W) {
      const int dy = max(max(ty0 - gy, gy - (ty0 + TS - 1)), 0);
      const int dx = max(max(tx0 - gx, gx - (tx0 + TS - 1)), 0);
      d = max(dy, dx);
    }
    dist[k] = d;
    if (d >= 0 && d <= iters) {
      const int64_t g = base + (int64_t)gy * W + gx;
      sU[c] = Uin[g];
      sV[c] = Vin[g];
      sfx[c] = gfx[g];
      sfy[c] = gfy[g];
      sft[c] = gft[g];
    }
  }
  __syncthreads();

  constexpr double w1 = 1.0 / 12, w2 = 1.0 / 6;
  for (int s = 1; s <= iters; ++s) {
    T nu[PER], nv[PER];
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const int c = threadIdx.x + k * NT;
      if (dist[k] >= 0 && dist[k] <= iters - s) {
        const int ly = c / E, lx = c % E;
        const int gy = ty0 - R + ly, gx = tx0 - R + lx;
        const int up = gy > 0 ? -E : 0, dn = gy < H - 1 ? E : 0;
        const int lf = gx > 0 ? -1 : 0, rt = gx < W - 1 ? 1 : 0;
        double au = 0.0, av = 0.0;
        au += (double)sU[c + up + lf] * w1; av += (double)sV[c + up + lf] * w1;
        au += (double)sU[c + up] * w2;      av += (double)sV[c + up] * w2;
        au += (double)sU[c + up + rt] * w1; av += (double)sV[c + up + rt] * w1;
        au += (double)sU[c + lf] * w2;      av += (double)sV[c + lf] * w2;
        au += (double)sU[c + rt] * w2;      av += (double)sV[c + rt] * w2;
        au += (double)sU[c + dn + lf] * w1; av += (double)sV[c + dn + lf] * w1;
        au += (double)sU[c + dn] * w2;      av += (double)sV[c + dn] * w2;
        au += (double)sU[c + dn + rt] * w1; av += (double)sV[c + dn + rt] * w1;
        const T ua = (T)au, va = (T)av;
        const float fx = sfx[c], fy = sfy[c];
        const float den = (alpha2 + fx * fx) + fy * fy;
        const T der = (((T)fx * ua + (T)fy * va) + (T)sft[c]) / (T)den;
        nu[k] = ua - (T)fx * der;
        nv[k] = va - (T)fy * der;
      }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const int c = threadIdx.x + k * NT;
      if (dist[k] >= 0 && dist[k] <= iters - s) {
        sU[c] = nu[k];
        sV[c] = nv[k];
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int k = 0; k < PER; ++k) {
    if (dist[k] == 0) {
      const int c = threadIdx.x + k * NT;
      const int gy = ty0 - R + c / E, gx = tx0 - R + c % E;
      const int64_t g = base + (int64_t)gy * W + gx;
      Uout[g] = sU[c];
      Vout[g] = sV[c];
    }
  }
}

// --- flow sum, orientation, magnitude (OpticalFlowMovie.py:21-24, 519-555) ----------------------

// numpy's remainder x % 2.0 (result in [0, 2), a zero remainder is +0)
__device__ inline double py_mod2(double a) {
  double m = fmod(a, 2.0);
  if (m != 0.0) {
    if (m < 0.0) m += 2.0;
  } else {
    m = 0.0;
  }
  return m;
}

// VxM = 0 + U[0] + U[1] + ... over a movie's frame pairs (float64, in order; from_zero = 0 starts
// from U[0] itself, keeping a -0.0), orient = (atan2(VyM, VxM) % 2) pi, mag = sqrt(Vx^2 + Vy^2),
// orient_rev = the orientation of (-VxM, -VyM).
template <typename T>
__global__ void k_flow_sum(const T* __restrict__ U, const T* __restrict__ V, int64_t nmov, int A,
                           int64_t hw, int from_zero, double* __restrict__ vx,
                           double* __restrict__ vy, double* __restrict__ orient,
                           double* __restrict__ mag, double* __restrict__ orient_rev) {
#pragma clang fp contract(off)
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= nmov * hw) return;
  const int64_t m = t / hw, p = t % hw;
  double sx, sy;
  int a0;
  if (from_zero) {
    sx = 0.0; sy = 0.0; a0 = 0;
  } else {
    sx = (double)U[m * A * hw + p]; sy = (double)V[m * A * hw + p]; a0 = 1;
  }
  for (int a = a0; a < A; ++a) {
    sx = sx + (double)U[(m * A + a) * hw + p];
    sy = sy + (double)V[(m * A + a) * hw + p];
  }
  vx[t] = sx;
  vy[t] = sy;
  if (orient) orient[t] = py_mod2(atan2(sy, sx)) * M_PI;
  if (mag) mag[t] = sqrt(sx * sx + sy * sy);
  if (orient_rev) orient_rev[t] = py_mod2(atan2(-sy, -sx)) * M_PI;
}

constexpr int kHsTile = 32;
constexpr int kHsHalo = 4;
constexpr int kHsThreads = 256;

}  // namespace memi

using namespace memi;

extern "C" {

int memi_flow_anisodiff(const float* in, float* out, float* tmp, int64_t nmov, int F, int H,
                        int W, int niter, double kappa, double gamma, double step0, double step1,
                        double step2, int option, void* stream) {
  MEMI_CHECK_ARG(in && out, "memi_flow_anisodiff: null pointer");
  MEMI_CHECK_ARG(nmov >= 0 && F >= 1 && H >= 1 && W >= 1,
                 "memi_flow_anisodiff: bad sizes nmov=%lld F=%d H=%d W=%d", (long long)nmov, F,
                 H, W);
  MEMI_CHECK_ARG(option == 1 || option == 2, "memi_flow_anisodiff: option must be 1 or 2");
  MEMI_CHECK_ARG(in != out, "memi_flow_anisodiff: in and out must not alias");
  const int steps = niter - 1;  // the reference's loop runs over range(1, niter)
  MEMI_CHECK_ARG(steps < 2 || (tmp && tmp != in && tmp != out),
                 "memi_flow_anisodiff: tmp buffer needed for niter > 2");
  const int64_t nvox = nmov * F * (int64_t)H * W;
  hipStream_t s = (hipStream_t)stream;
  if (nvox == 0) return MEMI_OK;
  if (steps <= 0) {
    if (hipMemcpyAsync(out, in, nvox * sizeof(float), hipMemcpyDeviceToDevice, s) != hipSuccess) {
      set_error("memi_flow_anisodiff: copy failed");
      return MEMI_ELAUNCH;
    }
    return MEMI_OK;
  }
  // the last step writes `out`
  float* buf[2];
  buf[(steps - 1) % 2] = out;
  buf[steps % 2] = tmp;
  const dim3 grid((unsigned)((nvox + 255) / 256));
  for (int k = 0; k < steps; ++k) {
    const float* src = k == 0 ? in : buf[(k - 1) % 2];
    float* dst = buf[k % 2];
    if (option == 1)
      hipLaunchKernelGGL(k_anisodiff_step<1>, grid, dim3(256), 0, s, src, dst, nvox, F, H, W,
                         (float)kappa, (float)gamma, (float)step0, (float)step1, (float)step2);
    else
      hipLaunchKernelGGL(k_anisodiff_step<2>, grid, dim3(256), 0, s, src, dst, nvox, F, H, W,
                         (float)kappa, (float)gamma, (float)step0, (float)step1, (float)step2);
    MEMI_CHECK_LAUNCH("k_anisodiff_step");
  }
  return MEMI_OK;
}

int memi_flow_block_average(const float* in, int64_t nmov, int F, int64_t hw, int bs,
                            double* out, void* stream) {
  MEMI_CHECK_ARG(in && out, "memi_flow_block_average: null pointer");
  MEMI_CHECK_ARG(nmov >= 0 && F >= 1 && hw >= 1 && bs >= 1,
                 "memi_flow_block_average: bad sizes nmov=%lld F=%d hw=%lld bs=%d",
                 (long long)nmov, F, (long long)hw, bs);
  const int A = (F + bs - 1) / bs;
  const int64_t n = nmov * A * hw;
  if (n == 0) return MEMI_OK;
  hipLaunchKernelGGL(k_block_average, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     (hipStream_t)stream, in, nmov, F, hw, bs, A, out);
  MEMI_CHECK_LAUNCH("k_block_average");
  return MEMI_OK;
}

int memi_gaussian_filter(const void* in, int in_f64, void* out, int out_f64, void* tmp,
                         int64_t count, int H, int W, double sigma, void* stream) {
  MEMI_CHECK_ARG(in && out && tmp, "memi_gaussian_filter: null pointer");
  MEMI_CHECK_ARG(count >= 0 && H >= 1 && W >= 1, "memi_gaussian_filter: bad sizes");
  MEMI_CHECK_ARG(in_f64 || !out_f64,
                 "memi_gaussian_filter: float32 input with float64 output is not a reference "
                 "case");
  MEMI_CHECK_ARG(tmp != in && tmp != out, "memi_gaussian_filter: tmp must not alias");
  MEMI_CHECK_ARG(sigma > 0.0, "memi_gaussian_filter: sigma must be positive");
  // scipy: radius = int(truncate * sigma + 0.5) with truncate 4, weights exp(-x^2 / (2 sigma^2))
  // normalised to sum 1
  const int r = (int)(4.0 * sigma + 0.5);
  MEMI_CHECK_ARG(r <= kGaussMaxR, "memi_gaussian_filter: sigma %.3f needs radius %d > %d", sigma,
                 r, kGaussMaxR);
  GaussTaps g{};
  g.r = r;
  double raw[2 * kGaussMaxR + 1];
  const int n_taps = 2 * r + 1;
  const double c = -0.5 / (sigma * sigma);
  for (int i = -r; i <= r; ++i) raw[i + r] = std::exp(c * (double)(i * i));
  // numpy's sum of up to 128 values: a plain loop below 8, else 8 strided partial sums combined
  // pairwise, then the tail
  double sum = 0.0;
  if (n_taps < 8) {
    for (int i = 0; i < n_taps; ++i) sum += raw[i];
  } else {
    double acc[8];
    for (int j = 0; j < 8; ++j) acc[j] = raw[j];
    int i = 8;
    for (; i < n_taps - n_taps % 8; i += 8)
      for (int j = 0; j < 8; ++j) acc[j] += raw[i + j];
    sum = ((acc[0] + acc[1]) + (acc[2] + acc[3])) + ((acc[4] + acc[5]) + (acc[6] + acc[7]));
    for (; i < n_taps; ++i) sum += raw[i];
  }
  for (int d = 0; d <= r; ++d) g.w[d] = raw[d + r] / sum;
  const int64_t n = count * (int64_t)H * W;
  if (n == 0) return MEMI_OK;
  const dim3 grid((unsigned)((n + 255) / 256));
  hipStream_t s = (hipStream_t)stream;
  if (in_f64 && out_f64) {
    hipLaunchKernelGGL((k_gauss_pass<double, double, 0>), grid, dim3(256), 0, s,
                       (const double*)in, (double*)tmp, count, H, W, g);
    hipLaunchKernelGGL((k_gauss_pass<double, double, 1>), grid, dim3(256), 0, s,
                       (const double*)tmp, (double*)out, count, H, W, g);
  } else if (in_f64) {
    hipLaunchKernelGGL((k_gauss_pass<double, float, 0>), grid, dim3(256), 0, s,
                       (const double*)in, (float*)tmp, count, H, W, g);
    hipLaunchKernelGGL((k_gauss_pass<float, float, 1>), grid, dim3(256), 0, s,
                       (const float*)tmp, (float*)out, count, H, W, g);
  } else {
    hipLaunchKernelGGL((k_gauss_pass<float, float, 0>), grid, dim3(256), 0, s,
                       (const float*)in, (float*)tmp, count, H, W, g);
    hipLaunchKernelGGL((k_gauss_pass<float, float, 1>), grid, dim3(256), 0, s,
                       (const float*)tmp, (float*)out, count, H, W, g);
  }
  MEMI_CHECK_LAUNCH("k_gauss_pass");
  return MEMI_OK;
}

int memi_hs_derivatives(const float* frames, const int32_t* pairs, int64_t nP, int H, int W,
                        float* fx, float* fy, float* ft, void* stream) {
  MEMI_CHECK_ARG(frames && pairs && fx && fy && ft, "memi_hs_derivatives: null pointer");
  MEMI_CHECK_ARG(nP >= 0 && H >= 1 && W >= 1, "memi_hs_derivatives: bad sizes");
  const int64_t n = nP * (int64_t)H * W;
  if (n == 0) return MEMI_OK;
  hipLaunchKernelGGL(k_hs_derivatives, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                     (hipStream_t)stream, frames, pairs, nP, H, W, fx, fy, ft);
  MEMI_CHECK_LAUNCH("k_hs_derivatives");
  return MEMI_OK;
}

int memi_hs_iterate(const float* fx, const float* fy, const float* ft, int64_t nP, int H, int W,
                    double alpha, int niter, void* U, void* V, int is_f64, void* tmpU,
                    void* tmpV, void* stream) {
  MEMI_CHECK_ARG(fx && fy && ft && U && V && tmpU && tmpV, "memi_hs_iterate: null pointer");
  MEMI_CHECK_ARG(nP >= 0 && H >= 1 && W >= 1 && niter >= 0, "memi_hs_iterate: bad sizes");
  MEMI_CHECK_ARG(tmpU != U && tmpV != V && tmpU != tmpV && U != V,
                 "memi_hs_iterate: U, V, tmpU, tmpV must be distinct");
  const int tiles_x = (W + kHsTile - 1) / kHsTile;
  const int tiles_y = (H + kHsTile - 1) / kHsTile;
  const int ntiles = tiles_x * tiles_y;
  MEMI_CHECK_ARG(nP * ntiles < 0x7fffffffLL, "memi_hs_iterate: too many tiles");
  if (nP == 0 || niter == 0) return MEMI_OK;
  const int launches = (niter + kHsHalo - 1) / kHsHalo;
  const size_t bytes = (size_t)nP * H * W * (is_f64 ? 8 : 4);
  hipStream_t s = (hipStream_t)stream;
  // ping-pong so that the last launch writes U, V
  void* cur_u = U;
  void* cur_v = V;
  void* nxt_u = tmpU;
  void* nxt_v = tmpV;
  if (launches % 2) {
    if (hipMemcpyAsync(tmpU, U, bytes, hipMemcpyDeviceToDevice, s) != hipSuccess ||
        hipMemcpyAsync(tmpV, V, bytes, hipMemcpyDeviceToDevice, s) != hipSuccess) {
      set_error("memi_hs_iterate: copy failed");
      return MEMI_ELAUNCH;
    }
    std::swap(cur_u, nxt_u);
    std::swap(cur_v, nxt_v);
  }
  // numpy adds the Python float alpha**2 to a float32 array as float32
  const float alpha2 = (float)(alpha * alpha);
  const dim3 grid((unsigned)(nP * ntiles));
  for (int l = 0, done = 0; l < launches; ++l) {
    const int it = std::min(kHsHalo, niter - done);
    if (is_f64)
      hipLaunchKernelGGL((k_hs_iterate<double, kHsTile, kHsHalo, kHsThreads>), grid,
                         dim3(kHsThreads), 0, s, fx, fy, ft, H, W, tiles_x, ntiles, alpha2, it,
                         (const double*)cur_u, (const double*)cur_v, (double*)nxt_u,
                         (double*)nxt_v);
    else
      hipLaunchKernelGGL((k_hs_iterate<float, kHsTile, kHsHalo, kHsThreads>), grid,
                         dim3(kHsThreads), 0, s, fx, fy, ft, H, W, tiles_x, ntiles, alpha2, it,
                         (const float*)cur_u, (const float*)cur_v, (float*)nxt_u,
                         (float*)nxt_v);
    MEMI_CHECK_LAUNCH("k_hs_iterate");
    std::swap(cur_u, nxt_u);
    std::swap(cur_v, nxt_v);
    done += it;
  }
  return MEMI_OK;
}

int memi_flow_sum(const void* U, const void* V, int is_f64, int64_t nmov, int A, int64_t hw,
                  int from_zero, double* vx, double* vy, double* orient, double* mag,
                  double* orient_rev, void* stream) {
  MEMI_CHECK_ARG(U && V && vx && vy, "memi_flow_sum: null pointer");
  MEMI_CHECK_ARG(nmov >= 0 && A >= 1 && hw >= 1, "memi_flow_sum: bad sizes");
  const int64_t n = nmov * hw;
  if (n == 0) return MEMI_OK;
  const dim3 grid((unsigned)((n + 255) / 256));
  if (is_f64)
    hipLaunchKernelGGL(k_flow_sum<double>, grid, dim3(256), 0, (hipStream_t)stream,
                       (const double*)U, (const double*)V, nmov, A, hw, from_zero, vx, vy,
                       orient, mag, orient_rev);
  else
    hipLaunchKernelGGL(k_flow_sum<float>, grid, dim3(256), 0, (hipStream_t)stream,
                       (const float*)U, (const float*)V, nmov, A, hw, from_zero, vx, vy, orient,
                       mag, orient_rev);
  MEMI_CHECK_LAUNCH("k_flow_sum");
  return MEMI_OK;
}

}  // extern "C"
