// Fused distance -> Gaussian kernel -> log-sum op and its backward helpers.
// The kernel is the one the Ferguson sweep sums, K_ij = exp(-D_ij / (2 eps))
// (fergusonE.py:36-43); the loss per eps is L = sum_i log sum_j K_ij.
#include "common.h"

namespace memi {

constexpr int EG = 8;  // eps values per block

// S[e][i] = sum_j K^e_ij, T[e][i] = sum_j K^e_ij D_ij; block = (row i, group of EG eps)
__global__ __launch_bounds__(256) void k_kernel_rowsums(const float* __restrict__ D, int n,
                                                        const double* __restrict__ eps,
                                                        int n_eps, double* __restrict__ S,
                                                        double* __restrict__ T) {
  __shared__ double red[8];
  const int i = blockIdx.x;
  const int e0 = blockIdx.y * EG;
  double c[EG], s[EG], t[EG];
#pragma unroll
  for (int q = 0; q < EG; ++q) {
    c[q] = (e0 + q < n_eps) ? 1.0 / (2.0 * eps[e0 + q]) : 0.0;
    s[q] = t[q] = 0.0;
  }
  const float* row = D + (int64_t)i * n;
  for (int j = threadIdx.x; j < n; j += blockDim.x) {
    const double d = row[j];
#pragma unroll
    for (int q = 0; q < EG; ++q) {
      const double k = exp(-d * c[q]);
      s[q] += k;
      t[q] += k * d;
    }
  }
#pragma unroll
  for (int q = 0; q < EG; ++q) {
    const double a = block_sum(s[q], red);
    const double b = block_sum(t[q], red);
    if (threadIdx.x == 0 && e0 + q < n_eps) {
      S[(int64_t)(e0 + q) * n + i] = a;
      T[(int64_t)(e0 + q) * n + i] = b;
    }
  }
}

// Wt[i][j] = sum_e g[e] (-1/(2 eps_e)) K^e_ij (1/S[e][i] + 1/S[e][j])   (K symmetric), and
// Wt[i][i] = 0: the i = j term of the gradient, w_k (X_ik C_ik^2 - C_ik C_ik X_ik) Wt_ii, is
// identically zero, while Wt_ii itself (K_ii = 1 at every eps, so the smallest eps dominates)
// can exceed the off-diagonal weights by 1e6 or more and would only feed rounding error of
// that size into the two float32 products Wt B and Wt Y, which cancel it in exact arithmetic.
__global__ void k_kernel_grad_weights(const float* __restrict__ D, int n,
                                      const double* __restrict__ eps, int n_eps,
                                      const double* __restrict__ S, const double* __restrict__ g,
                                      float* __restrict__ Wt) {
  const int64_t total = (int64_t)n * n;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < total;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int i = (int)(e / n), j = (int)(e % n);
    const double d = D[e];
    double w = 0.0;
    for (int q = 0; q < n_eps; ++q) {
      const double c = 1.0 / (2.0 * eps[q]);
      const double k = exp(-d * c);
      w += g[q] * (-c) * k * (1.0 / S[(int64_t)q * n + i] + 1.0 / S[(int64_t)q * n + j]);
    }
    Wt[e] = i == j ? 0.f : (float)w;
  }
}

}  // namespace memi

using namespace memi;

extern "C" {

int memi_kernel_rowsums(const float* D, int n, const double* eps, int n_eps, double* S, double* T,
                        void* stream) {
  MEMI_CHECK_ARG(D && eps && S && T && n > 0 && n_eps > 0, "memi_kernel_rowsums: bad arguments");
  dim3 g((unsigned)n, (unsigned)((n_eps + EG - 1) / EG));
  hipLaunchKernelGGL(k_kernel_rowsums, g, dim3(256), 0, (hipStream_t)stream, D, n, eps, n_eps, S,
                     T);
  MEMI_CHECK_LAUNCH("k_kernel_rowsums");
  return MEMI_OK;
}

int memi_kernel_grad_weights(const float* D, int n, const double* eps, int n_eps, const double* S,
                             const double* g, float* Wt, void* stream) {
  MEMI_CHECK_ARG(D && eps && S && g && Wt && n > 0 && n_eps > 0,
                 "memi_kernel_grad_weights: bad arguments");
  hipLaunchKernelGGL(k_kernel_grad_weights, dim3(grid_for((int64_t)n * n)), dim3(256), 0,
                     (hipStream_t)stream, D, n, eps, n_eps, S, g, Wt);
  MEMI_CHECK_LAUNCH("k_kernel_grad_weights");
  return MEMI_OK;
}

}  // extern "C"
