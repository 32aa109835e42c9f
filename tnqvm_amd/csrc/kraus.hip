// Kernels of the locally purified MPS (ExaTnPmpsVisitor.cpp).  A purified site carries a
// Kraus leg k between its physical index p and its right bond: A[l][p][k][r] (kap = 1 is
// the plain MPS site).
//   k_swap_mid ...... [l][p][k][m] <-> [l][k][p][m], so that a 2-qubit gate sees the left
//                     site's Kraus leg fused into its left bond and the right site's into its
//                     right bond (the merge of ExaTnPmpsVisitor.cpp:337-434)
//   k_kraus_expand .. the channel's Kraus operators on the physical index, one row per
//                     (old Kraus index, operator) (ExaTnPmpsVisitor.cpp:887-987: merge of Q
//                     with its conjugate, the Choi operator, then a split back)
//   k_kraus_gather .. compressed Kraus vectors (rows of the Jacobi result) back into the site
#include <hip/hip_runtime.h>

#include <algorithm>

#include "device_util.h"
#include "kernels.h"

namespace hipmps {

template <typename T>
__global__ __launch_bounds__(256) void k_swap_mid(const T* __restrict__ src, T* __restrict__ dst, int L, int A,
                                                  int B, int m_max, const int* m_dyn) {
  const int M = m_dyn ? *m_dyn : m_max;
  const size_t total = (size_t)L * A * B * M;
  for (size_t e = (size_t)blockIdx.x * 256 + threadIdx.x; e < total; e += (size_t)gridDim.x * 256) {
    const int m = (int)(e % M);
    size_t t = e / M;
    const int b = (int)(t % B);
    t /= B;
    const int a = (int)(t % A);
    const size_t l = t / A;
    dst[((l * B + b) * A + a) * M + m] = src[e];
  }
}

hipError_t launch_swap_mid(const void* src, void* dst, int L, int A, int B, int m_max, const int* m_dyn, int fp32,
                           hipStream_t s) {
  const size_t total = (size_t)L * A * B * m_max;
  if (total == 0) return hipSuccess;
  const unsigned blocks = (unsigned)std::min<size_t>((total + 255) / 256, 65536);
  if (fp32)
    hipLaunchKernelGGL(k_swap_mid<float2>, dim3(blocks), dim3(256), 0, s, (const float2*)src, (float2*)dst, L, A,
                       B, m_max, m_dyn);
  else
    hipLaunchKernelGGL(k_swap_mid<double2>, dim3(blocks), dim3(256), 0, s, (const double2*)src, (double2*)dst, L,
                       A, B, m_max, m_dyn);
  return hipGetLastError();
}

template <typename T>
__global__ __launch_bounds__(256) void k_kraus_expand(const T* __restrict__ site, int cl, int kap, int cr,
                                                      const double2* __restrict__ K, int n_ops,
                                                      double2* __restrict__ X, int rows_pad) {
  const size_t c = (size_t)2 * cl * cr;
  const size_t total = (size_t)rows_pad * c;
  for (size_t e = (size_t)blockIdx.x * 256 + threadIdx.x; e < total; e += (size_t)gridDim.x * 256) {
    const size_t row = e / c, col = e % c;
    double2 v = make_double2(0.0, 0.0);
    if (row < (size_t)kap * n_ops) {
      const int k = (int)(row / n_ops), o = (int)(row % n_ops);
      const int r = (int)(col % cr);
      const size_t lp = col / cr;
      const int p = (int)(lp & 1);
      const size_t l = lp >> 1;
      const double2 a0 = ld<T>(site, ((l * 2 + 0) * kap + k) * cr + r);
      const double2 a1 = ld<T>(site, ((l * 2 + 1) * kap + k) * cr + r);
      v = cadd(cmul(K[o * 4 + 2 * p], a0), cmul(K[o * 4 + 2 * p + 1], a1));
    }
    X[e] = v;
  }
}

hipError_t launch_kraus_expand(const void* site, int fp32, int cl, int kap, int cr, const double2* d_kraus,
                               int n_ops, double2* X, int rows_pad, hipStream_t s) {
  const size_t total = (size_t)rows_pad * 2 * cl * cr;
  const unsigned blocks = (unsigned)std::min<size_t>((total + 255) / 256, 65536);
  if (fp32)
    hipLaunchKernelGGL(k_kraus_expand<float2>, dim3(blocks), dim3(256), 0, s, (const float2*)site, cl, kap, cr,
                       d_kraus, n_ops, X, rows_pad);
  else
    hipLaunchKernelGGL(k_kraus_expand<double2>, dim3(blocks), dim3(256), 0, s, (const double2*)site, cl, kap, cr,
                       d_kraus, n_ops, X, rows_pad);
  return hipGetLastError();
}

template <typename T>
__global__ __launch_bounds__(256) void k_kraus_gather(const double2* __restrict__ Y, int ldy,
                                                      const int* __restrict__ perm, const int* Kp, int k_max,
                                                      T* __restrict__ site, int cl, int cr) {
  const int K = *Kp;
  const size_t c = (size_t)2 * cl * cr;
  const size_t total = (size_t)K * c;
  for (size_t e = (size_t)blockIdx.x * 256 + threadIdx.x; e < total; e += (size_t)gridDim.x * 256) {
    const int j = (int)(e / c);
    const size_t col = e % c;
    const int r = (int)(col % cr);
    const size_t lp = col / cr;
    st<T>(site, (lp * K + j) * cr + r, Y[(size_t)perm[j] * ldy + col]);
  }
}

// Adjoint layout, for a channel whose rows (kap n_ops) outnumber its columns c = 2 cl cr: the
// Jacobi then runs on the c rows of X^H, each followed by a row of the identity, so that its
// rotations W (W X^H has orthogonal rows) accumulate there.  Y[j][i] = conj(X[i][j]) for i <
// rows, Y[j][rows + t] = delta_jt, pad rows j >= c zero; X as k_kraus_expand forms it.
template <typename T>
__global__ __launch_bounds__(256) void k_kraus_expand_adj(const T* __restrict__ site, int cl, int kap, int cr,
                                                          const double2* __restrict__ K, int n_ops,
                                                          double2* __restrict__ Y, int ldy, int rows_pad) {
  const size_t c = (size_t)2 * cl * cr, rows = (size_t)kap * n_ops;
  const size_t total = (size_t)rows_pad * ldy;
  for (size_t e = (size_t)blockIdx.x * 256 + threadIdx.x; e < total; e += (size_t)gridDim.x * 256) {
    const size_t j = e / ldy, i = e % ldy;
    double2 v = make_double2(0.0, 0.0);
    if (j < c) {
      if (i < rows) {
        const int k = (int)(i / n_ops), o = (int)(i % n_ops);
        const int r = (int)(j % cr);
        const size_t lp = j / cr;
        const int p = (int)(lp & 1);
        const size_t l = lp >> 1;
        const double2 a0 = ld<T>(site, ((l * 2 + 0) * kap + k) * cr + r);
        const double2 a1 = ld<T>(site, ((l * 2 + 1) * kap + k) * cr + r);
        v = cconj(cadd(cmul(K[o * 4 + 2 * p], a0), cmul(K[o * 4 + 2 * p + 1], a1)));
      } else if (i - rows == j) {
        v = make_double2(1.0, 0.0);
      }
    }
    Y[e] = v;
  }
}

hipError_t launch_kraus_expand_adj(const void* site, int fp32, int cl, int kap, int cr, const double2* d_kraus,
                                   int n_ops, double2* Y, int ldy, int rows_pad, hipStream_t s) {
  const size_t total = (size_t)rows_pad * ldy;
  const unsigned blocks = (unsigned)std::min<size_t>((total + 255) / 256, 65536);
  if (fp32)
    hipLaunchKernelGGL(k_kraus_expand_adj<float2>, dim3(blocks), dim3(256), 0, s, (const float2*)site, cl, kap, cr,
                       d_kraus, n_ops, Y, ldy, rows_pad);
  else
    hipLaunchKernelGGL(k_kraus_expand_adj<double2>, dim3(blocks), dim3(256), 0, s, (const double2*)site, cl, kap,
                       cr, d_kraus, n_ops, Y, ldy, rows_pad);
  return hipGetLastError();
}

// Kraus vectors from the adjoint layout: X' = Sigma W (X'^H X' = W^H Sigma^2 W = X^H X), row j
// = sigma_perm[j] times the identity part of row perm[j], sigma the norm of its data part.
template <typename T>
__global__ __launch_bounds__(256) void k_kraus_gather_adj(const double2* __restrict__ Y, int ldy, int off,
                                                          const double* __restrict__ rownorm,
                                                          const int* __restrict__ perm, const int* Kp,
                                                          T* __restrict__ site, int cl, int cr) {
  const int K = *Kp;
  const size_t c = (size_t)2 * cl * cr;
  const size_t total = (size_t)K * c;
  for (size_t e = (size_t)blockIdx.x * 256 + threadIdx.x; e < total; e += (size_t)gridDim.x * 256) {
    const int j = (int)(e / c);
    const size_t col = e % c;
    const int r = (int)(col % cr);
    const size_t lp = col / cr;
    const int pj = perm[j];
    st<T>(site, (lp * K + j) * cr + r, cscale(Y[(size_t)pj * ldy + off + col], rownorm[pj]));
  }
}

hipError_t launch_kraus_gather_adj(const double2* Y, int ldy, int off, const double* rownorm, const int* perm,
                                   const int* K, int k_max, void* site, int fp32, int cl, int cr, hipStream_t s) {
  const size_t total = (size_t)k_max * 2 * cl * cr;
  if (total == 0) return hipSuccess;
  const unsigned blocks = (unsigned)std::min<size_t>((total + 255) / 256, 65536);
  if (fp32)
    hipLaunchKernelGGL(k_kraus_gather_adj<float2>, dim3(blocks), dim3(256), 0, s, Y, ldy, off, rownorm, perm, K,
                       (float2*)site, cl, cr);
  else
    hipLaunchKernelGGL(k_kraus_gather_adj<double2>, dim3(blocks), dim3(256), 0, s, Y, ldy, off, rownorm, perm, K,
                       (double2*)site, cl, cr);
  return hipGetLastError();
}

hipError_t launch_kraus_gather(const double2* Y, int ldy, const int* perm, const int* K, int k_max, void* site,
                               int fp32, int cl, int cr, hipStream_t s) {
  const size_t total = (size_t)k_max * 2 * cl * cr;
  if (total == 0) return hipSuccess;
  const unsigned blocks = (unsigned)std::min<size_t>((total + 255) / 256, 65536);
  if (fp32)
    hipLaunchKernelGGL(k_kraus_gather<float2>, dim3(blocks), dim3(256), 0, s, Y, ldy, perm, K, k_max,
                       (float2*)site, cl, cr);
  else
    hipLaunchKernelGGL(k_kraus_gather<double2>, dim3(blocks), dim3(256), 0, s, Y, ldy, perm, K, k_max,
                       (double2*)site, cl, cr);
  return hipGetLastError();
}

}  // namespace hipmps
