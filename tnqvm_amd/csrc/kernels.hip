// kernels.hip -- gfx950 kernels of libhipmps.
//
// Kernel map (reference function each one replaces, tnqvm/visitors/exatn-mps/):
//   k_apply_1q ......... 1-qubit gate into a site      ExaTnMpsVisitor.cpp:1231-1273
//   k_theta_mfma ....... merge + 2-qubit gate          ExaTnMpsVisitor.cpp:1401-1541
//                        (+ stabilizeTensorBody, :141-161)
//   k_jacobi_prep,
//   k_jacobi_round ..... full SVD (one-sided block Jacobi, FP64)  :1555-1627
//   k_rownorm, k_select  partial norms, cutoff, max-bond clamp    :2366-2536
//   k_gather, k_gemm ... sqrt(S) split into the two new sites     :1623
//   k_gemm, k_chain,
//   k_trace_prod ....... environments / amplitudes / slices       :593-643, :2588-2675
#include "kernels.h"
#include "device_util.h"

#include <algorithm>
#include <cfloat>
#include <cmath>

namespace hipmps {

// ------------------------------------------------------------------ 1-qubit gate
// new[l][b][r] = sum_i M[b][i] A[l][i][r]  (pattern Result(a,b,c)=Q(a,i,c)*G(i,b),
// ExaTnMpsVisitor.cpp:1246, with G(i,b) = M[b][i] from the row-major flattening)
template <typename T>
__global__ __launch_bounds__(256) void k_apply_1q(T* A, int cl, int cr, double2 m0, double2 m1,
                                                  double2 m2, double2 m3) {
  size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x;
  size_t total = (size_t)cl * cr;
  if (idx >= total) return;
  size_t l = idx / cr, r = idx % cr;
  size_t i0 = (l * 2 + 0) * cr + r, i1 = (l * 2 + 1) * cr + r;
  double2 a0 = ld(A, i0), a1 = ld(A, i1);
  st(A, i0, cfma(m1, a1, cmul(m0, a0)));
  st(A, i1, cfma(m3, a1, cmul(m2, a0)));
}

hipError_t launch_apply_1q(void* site, int fp32, int cl, int cr, const double2 m[4], hipStream_t s) {
  size_t total = (size_t)cl * cr;
  dim3 grid((unsigned)((total + 255) / 256));
  if (fp32)
    hipLaunchKernelGGL(k_apply_1q<float2>, grid, dim3(256), 0, s, (float2*)site, cl, cr, m[0], m[1], m[2], m[3]);
  else
    hipLaunchKernelGGL(k_apply_1q<double2>, grid, dim3(256), 0, s, (double2*)site, cl, cr, m[0], m[1], m[2], m[3]);
  return hipGetLastError();
}

// ------------------------------------------------------------------ theta = G (Ai Aj)
// Workgroup tile: 16 values of l (32 rows (l,a)) x 32 values of r (64 columns (b,r)).
constexpr int TH_L = 16, TH_R = 32;

// ------------------------------------------------------------------ theta on MFMA
// theta'[(l,a)][(b,r)] = sum_{a'b'} G[(a,b),(a',b')] sum_m Ai[l][a'][m] Aj[m][b'][r] as a complex
// GEMM on the matrix cores: complex64 sites use v_mfma_f32_16x16x4f32 (FP32 products and FP32
// accumulation), complex128 sites v_mfma_f64_16x16x4f64; each complex product is four real
// MFMAs.  Workgroup tile: 16 l x 32 r; wave w: 8 l x 16 r, i.e. one 16-row MFMA
// tile (rows = 8 l x 2 a) against two 16-column tiles (b = 0, 1).  The MFMA rows are ordered
// so that the a = 0 and a = 1 rows of one l land in the same lane, as do b = 0 and b = 1, so
// the 4x4 gate mixes the (a, b) block of each (l, r) in registers:
//   f32 D[i][j]: lane j + 16 (i / 4), register i % 4  -> row i = 4g + v: l = 2g + (v & 1), a = v >> 1
//   f64 D[i][j]: lane j + 16 (i % 4), register i / 4  -> row i = 4v + g: l = 2g + (v & 1), a = v >> 1
template <typename T>
struct MfmaTraits;
template <>
struct MfmaTraits<float2> {
  typedef float scalar;
  typedef float acc __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ acc mma(float a, float b, acc c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  // (l offset, a) of MFMA row i
  static __device__ __forceinline__ int row_l(int i) { return 2 * (i >> 2) + (i & 1); }
  static __device__ __forceinline__ int row_a(int i) { return (i >> 1) & 1; }
  // (lane group g, register v) holding D[i][.]
  static __device__ __forceinline__ int out_row(int g, int v) { return 4 * g + v; }
};
template <>
struct MfmaTraits<double2> {
  typedef double scalar;
  typedef double acc __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ acc mma(double a, double b, acc c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row_l(int i) { return 2 * (i & 3) + ((i >> 2) & 1); }
  static __device__ __forceinline__ int row_a(int i) { return i >> 3; }
  static __device__ __forceinline__ int out_row(int g, int v) { return 4 * v + g; }
};

template <typename T>
__global__ __launch_bounds__(256) void k_theta_mfma(const ThetaTask* tasks, int ntasks) {
  typedef MfmaTraits<T> M;
  typedef typename M::scalar S;
  typedef typename M::acc A4;
  constexpr int KC = 16;  // contraction chunk staged in LDS
  __shared__ T sAi[TH_L][2][KC + 1];
  __shared__ T sAj[KC][2][TH_R + 1];
  const int t = find_task(tasks, ntasks, (int)blockIdx.x, [](const ThetaTask& x) { return x.tile_begin; });
  const ThetaTask& task = tasks[t];
  const int tile = blockIdx.x - task.tile_begin;
  if (tile >= task.tiles_l * task.tiles_r) return;
  const int l0 = (tile / task.tiles_r) * TH_L, r0 = (tile % task.tiles_r) * TH_R;
  const int cl = task.cl, cm = task.cm, cr = task.cr;
  const T* Ai = (const T*)task.Ai;
  const T* Aj = (const T*)task.Aj;
  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  const int l16 = lane & 15, k4 = lane >> 4;
  const int wl = (wave >> 1) * 8, wr = (wave & 1) * 16;  // this wave's l and r offsets in the tile
  const int my_l = wl + M::row_l(l16), my_a = M::row_a(l16);  // A operand row of this lane
  A4 re0 = {0, 0, 0, 0}, im0 = {0, 0, 0, 0}, re1 = {0, 0, 0, 0}, im1 = {0, 0, 0, 0};
  for (int m0 = 0; m0 < cm; m0 += KC) {
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const int e = tid + 256 * q;  // 512 = 16 l x 2 a x 16 k
      const int l = e >> 5, a = (e >> 4) & 1, k = e & 15;
      T v = T{0, 0};
      if (l0 + l < cl && m0 + k < cm) v = Ai[((size_t)(l0 + l) * 2 + a) * cm + m0 + k];
      sAi[l][a][k] = v;
    }
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int e = tid + 256 * q;  // 1024 = 16 k x 2 b x 32 r
      const int k = e >> 6, b = (e >> 5) & 1, rr = e & 31;
      T v = T{0, 0};
      if (m0 + k < cm && r0 + rr < cr) v = Aj[((size_t)(m0 + k) * 2 + b) * cr + r0 + rr];
      sAj[k][b][rr] = v;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < KC / 4; ++kk) {
      const T x = sAi[my_l][my_a][kk * 4 + k4];            // A[i][k], i = l16
      const T y0 = sAj[kk * 4 + k4][0][wr + l16];          // B[k][j], b = 0
      const T y1 = sAj[kk * 4 + k4][1][wr + l16];          // b = 1
      const S xr = x.x, xi = x.y, nxi = -x.y;
      re0 = M::mma(xr, y0.x, re0);
      re0 = M::mma(nxi, y0.y, re0);
      im0 = M::mma(xr, y0.y, im0);
      im0 = M::mma(xi, y0.x, im0);
      re1 = M::mma(xr, y1.x, re1);
      re1 = M::mma(nxi, y1.y, re1);
      im1 = M::mma(xr, y1.y, im1);
      im1 = M::mma(xi, y1.x, im1);
    }
    __syncthreads();
  }
  // registers v and v + 2 of this lane hold a = 0, 1 of one l; re0/im0 b = 0, re1/im1 b = 1
  const int r = r0 + wr + l16;
  const size_t ldt = (size_t)2 * cr;
#pragma unroll
  for (int v = 0; v < 2; ++v) {
    const int l = l0 + wl + M::row_l(M::out_row(k4, v));
    if (l >= cl || r >= cr) continue;
    double2 x[4];  // (a', b') = (0,0), (0,1), (1,0), (1,1)
    x[0] = make_double2(re0[v], im0[v]);
    x[1] = make_double2(re1[v], im1[v]);
    x[2] = make_double2(re0[v + 2], im0[v + 2]);
    x[3] = make_double2(re1[v + 2], im1[v + 2]);
#pragma unroll
    for (int ab = 0; ab < 4; ++ab) {
      double2 o = make_double2(0, 0);
#pragma unroll
      for (int k = 0; k < 4; ++k) o = cfma(task.G[ab * 4 + k], x[k], o);
      // stabilizeTensorBody: |x| < 1e-100 -> 0 (ExaTnMpsVisitor.cpp:141-161)
      if (cabs2(o) < 1e-200) o = make_double2(0, 0);
      const int a = ab >> 1, b = ab & 1;
      task.theta[((size_t)l * 2 + a) * ldt + (size_t)b * cr + r] = o;
    }
  }
}

hipError_t launch_theta(const ThetaTask* d_tasks, int ntasks, int total_tiles, int fp32, hipStream_t s) {
  if (total_tiles <= 0) return hipSuccess;
  if (fp32)
    hipLaunchKernelGGL(k_theta_mfma<float2>, dim3(total_tiles), dim3(256), 0, s, d_tasks, ntasks);
  else
    hipLaunchKernelGGL(k_theta_mfma<double2>, dim3(total_tiles), dim3(256), 0, s, d_tasks, ntasks);
  return hipGetLastError();
}

// ------------------------------------------------------------------ Jacobi: prepare Y
// Y = X with X = theta (m <= n) or theta^H (m > n); rows r..r_pad-1 are zero.
__global__ __launch_bounds__(256) void k_jacobi_prep(const SplitTask* sp, const JacTask* jt) {
  __shared__ double2 tile[32][33];
  const SplitTask& S = sp[blockIdx.y];
  const JacTask& J = jt[blockIdx.y];
  const int tiles_c = (J.c_tot + 31) / 32;
  const int tiles_r = J.r_pad / 32;
  const int bt = blockIdx.x;
  if (bt >= tiles_c * tiles_r) return;
  const int i0 = (bt / tiles_c) * 32, c0 = (bt % tiles_c) * 32;
  const int tx = threadIdx.x % 32, ty = threadIdx.x / 32;  // 32 x 8
  if (!S.transposed) {
    for (int y = ty; y < 32; y += 8) {
      int i = i0 + y, c = c0 + tx;
      if (c < J.c_tot) {
        double2 v = make_double2(0, 0);
        if (i < J.r && c < J.c) v = S.theta[(size_t)i * S.n + c];
        J.Y[(size_t)i * J.ldy + c] = v;
      }
    }
  } else {
    // Y[i][c] = conj(theta[c][i]); read theta rows coalesced along i
    for (int y = ty; y < 32; y += 8) {
      int c = c0 + y, i = i0 + tx;
      double2 v = make_double2(0, 0);
      if (c < J.c && i < J.r) v = cconj(S.theta[(size_t)c * S.n + i]);
      tile[y][tx] = v;
    }
    __syncthreads();
    for (int y = ty; y < 32; y += 8) {
      int i = i0 + y, c = c0 + tx;
      if (c < J.c_tot) J.Y[(size_t)i * J.ldy + c] = tile[tx][y];
    }
  }
}

hipError_t launch_jacobi_prep(const SplitTask* d_split, const JacTask* d_jac, int ntasks, int max_rows,
                              int max_cols, hipStream_t s) {  // max_cols: largest c + r
  int tr = (max_rows + 31) / 32, tc = (max_cols + 31) / 32;
  hipLaunchKernelGGL(k_jacobi_prep, dim3(tr * tc, ntasks), dim3(256), 0, s, d_split, d_jac);
  return hipGetLastError();
}

// ------------------------------------------------------------------ Jacobi: one round
// One workgroup per block pair (the rows of two JB-row blocks).  Round-robin ("circle")
// ordering: slot 0 holds block 0, slot s>0 holds block 1 + (s-1+round) mod (nb-1); pair
// i couples slots i and nb-1-i, so every unordered block pair meets once per sweep.  The
// same ordering, with 2*JB single rows in place of blocks, drives the inner sweep.
//
// For the 2*JB rows:  Gram  Gam[i][j] = sum_c conj(Y[i][c]) Y[j][c]
//                     Gam ~ W Lambda W^H  by one cyclic two-sided Jacobi sweep (FP64)
//                     Y' = W^T Y          (rows become mutually orthogonal)
__device__ __forceinline__ int circle_slot(int s, int round, int nb) {
  return s == 0 ? 0 : 1 + (s - 1 + round) % (nb - 1);
}

// 1/sqrt(a) and 1/d: the hardware approximations refined by one Newton step, which
// brings them to the last bit or two of FP64 (all that a rotation needs: only
// c^2 + |s|^2 = 1 has to hold to working precision, not the angle itself).
__device__ __forceinline__ double nr_rsq(double a) {
  const double y = __builtin_amdgcn_rsq(a);
  const double hy = 0.5 * a * y;
  return fma(y, fma(-hy, y, 0.5), y);
}
__device__ __forceinline__ double nr_rcp(double d) {
  const double x = __builtin_amdgcn_rcp(d);
  return fma(x, fma(-d, x, 1.0), x);
}

// Rotation that annihilates gamma in the Hermitian 2x2 [[alpha, gamma], [conj(gamma), beta]]
// (Golub & Van Loan Alg. 8.4.1 with a phase): h = beta - alpha,
// tau = 2 sign(h) gamma / (|h| + sqrt(h^2 + 4|gamma|^2))  (|tau| = t, the smaller root),
// c = 1/sqrt(1 + |tau|^2), se = c tau;  J = [[c, se], [-conj(se), c]].
// Returns the identity when |gamma| <= tol sqrt(alpha beta) or a diagonal entry is 0.
__device__ __forceinline__ void jacobi_rotation(double alpha, double beta, double2 gamma, double tol, double floor2,
                                                double& c, double2& se) {
  const double g2 = cabs2(gamma);
  if (!(alpha > floor2) || !(beta > floor2) || !(g2 > tol * tol * alpha * beta)) {
    c = 1.0;
    se = make_double2(0.0, 0.0);
    return;
  }
  const double h = beta - alpha;
  const double a = fma(h, h, 4.0 * g2);
  const double d = fabs(h) + a * nr_rsq(a);
  const double f = (h >= 0.0 ? 2.0 : -2.0) * nr_rcp(d);
  const double2 tau = make_double2(f * gamma.x, f * gamma.y);
  c = nr_rsq(1.0 + cabs2(tau));
  se = make_double2(c * tau.x, c * tau.y);
}

// Thread -> 2x2 block of rotation pairs (bk, bl) of the inner sweep, packed bk << 4 | bl.  The 16
// blocks that form the next step's rotations -- (0, 1), (j + 1, j - 1) for 1 <= j <= 14, (15, 14)
// -- go to threads 0..15 (wave 0), so that the other three waves skip that code; the remaining
// 240 blocks follow in row order.
struct InnerBlockMap {
  unsigned char v[256];
};
constexpr bool inner_owner(int bk, int bl) {
  return (bk == 0 && bl == 1) || (bk >= 2 && bl == bk - 2) || (bk == 15 && bl == 14);
}
constexpr InnerBlockMap make_inner_block_map() {
  InnerBlockMap m{};
  m.v[0] = (unsigned char)(0 << 4 | 1);
  for (int j = 1; j <= 14; ++j) m.v[j] = (unsigned char)((j + 1) << 4 | (j - 1));
  m.v[15] = (unsigned char)(15 << 4 | 14);
  int n = 16;
  for (int bk = 0; bk < 16; ++bk)
    for (int bl = 0; bl < 16; ++bl)
      if (!inner_owner(bk, bl)) m.v[n++] = (unsigned char)(bk << 4 | bl);
  return m;
}
__constant__ InnerBlockMap c_inner_blocks = make_inner_block_map();

// Design notes (gfx950):
//  * Gram and apply stream their MFMA operands straight from memory into registers; no
//    LDS staging and no barriers in either loop.  The Gram is Hermitian, so only tiles
//    (0,0), (0,1), (1,1) are formed; each wave covers every fourth 4-column unit and the
//    four partial Grams are summed once in LDS.
//  * In the apply each wave owns whole 32-row x 16-column tiles, so the in-place update
//    never overwrites a column another wave still reads.
//  * Inner sweep: thread (k, l) owns the 2x2 block of rotation pairs k and l, updates it
//    two-sided in one pass (B' = Jk^H B Jl, G double-buffered in LDS), updates two W
//    entries pairs, and -- when the next step's pair (p', q') has p' in pair k and q' in
//    pair l -- also forms that pair's rotation from values it already holds.  One barrier
//    per step; every LDS read of a step is issued before any arithmetic.
//    In the circle ordering the pair owner is static: next pair 0 <- (0,1), next pair j
//    <- (j+1, j-1) for 1 <= j <= 14, next pair 15 <- (15,14); and each row index moves
//    to x+1 (31 -> 1) from one step to the next, except row 0.
// BLK is the block size JB (a template parameter: the kernel's register allocation on
// gfx950 comes out at 150 VGPRs, three waves per SIMD, in this form).
template <int BLK, bool TILED>
__global__ __launch_bounds__(256, 2) void k_jacobi_round(const JacTask* tasks, int ntasks, int round,
                                                         int inner_sweeps) {
  static_assert(BLK == 16 && JB == BLK && JP == 2 * BLK, "256 threads map onto 16 x 16 rotation-pair blocks");
  __shared__ double2 sG[JP][JP + 1];
  __shared__ double2 sG2[JP][JP + 1];
  __shared__ double2 sW[JP][JP + 1];
  __shared__ double s_c[2][JP / 2 + 1];  // rotations of the current / next step;
  __shared__ double2 s_se[2][JP / 2 + 1];  // entry JP/2 takes the writes of non-owners
  __shared__ double s_red[4];  // per-wave maxima of the convergence test

  const int t = find_task(tasks, ntasks, (int)blockIdx.x, [](const JacTask& x) { return x.pair_begin; });
  const JacTask& T = tasks[t];
  const int pi = blockIdx.x - T.pair_begin;
  const int nb = T.nb;
  if (pi >= nb / 2 || round >= nb - 1) return;
  const int ba = circle_slot(pi, round, nb), bb = circle_slot(nb - 1 - pi, round, nb);
  const int C = T.c, CT = T.c_tot;
  const size_t ld = T.ldy;
  const double tol = T.tol;
  const double floor2 = T.fro2 ? JAC_FLOOR2 * *T.fro2 : 0.0;
  double2* Y = TILED ? T.Yt : T.Y;
  const int ntc = T.ntc;
  const int tid = threadIdx.x;
  const int wave = tid >> 6, lane = tid & 63;
  const int l16 = lane & 15, k4 = lane >> 4;
  // element (row, col) of the pair's rows; row-major: row * ld + col
  auto at = [&](int row, int col) -> size_t {
    return TILED ? jac_tile_offset(row, col, ntc) : (size_t)row * ld + col;
  };
  const int rowa = ba * JB + l16, rowb = bb * JB + l16;  // pair rows l16 and 16 + l16

  // ---- Gram tiles (0,0), (0,1), (1,1) over this wave's 4-column units u = wave, wave + 4, ... ----
  // FP64 MFMA, or in the early sweeps of large SVDs FP32 MFMA at twice the rate: the rotations
  // they give are still applied in FP64 and are unitary, so the rows only lose the accuracy of
  // the rotation angles, which the FP64 sweeps that end every SVD make up for
  const int nu = (C + 3) >> 2;
  constexpr int PD = 2;
  auto ld_unit = [&](int u, double2& v0, double2& v1) {
    const int col = 4 * u + k4;
    if (u < nu && col < C) {
      v0 = Y[at(rowa, col)];
      v1 = Y[at(rowb, col)];
    } else {
      v0 = make_double2(0, 0);
      v1 = make_double2(0, 0);
    }
  };
  // Complex products on real MFMA with three real products each instead of four:
  //  * diagonal tiles (Hermitian): Re = Ar Ar^T + Ai Ai^T, and Im = X - X^T with X = Ar Ai^T;
  //  * off-diagonal tile conj(A) B^T (Gauss): K1 = (Ar - Ai) Br^T, K2 = Ar (Bi - Br)^T,
  //    K3 = -Ai (Br + Bi)^T;  Re = K1 - K3, Im = K1 + K2.
  // Each partial sum of the four waves goes to LDS as (Re, X) for the diagonal tiles and
  // (Re, Im) for the off-diagonal one; Im of the diagonal tiles is formed after the sum.
  if (!T.gram32) {
    mfma_d4 a00r = {0, 0, 0, 0}, a00x = {0, 0, 0, 0}, a11r = {0, 0, 0, 0}, a11x = {0, 0, 0, 0};
    mfma_d4 k1 = {0, 0, 0, 0}, k2 = {0, 0, 0, 0}, k3 = {0, 0, 0, 0};
    double2 x0[PD], x1[PD];
#pragma unroll
    for (int d = 0; d < PD; ++d) ld_unit(wave + 4 * d, x0[d], x1[d]);
    for (int u0 = wave; u0 < nu; u0 += 4 * PD) {
#pragma unroll
      for (int d = 0; d < PD; ++d) {
        const double2 v0 = x0[d], v1 = x1[d];
        ld_unit(u0 + 4 * (d + PD), x0[d], x1[d]);
        a00r = mfma64(v0.x, v0.x, a00r);
        a11r = mfma64(v1.x, v1.x, a11r);
        a00x = mfma64(v0.x, v0.y, a00x);
        a11x = mfma64(v1.x, v1.y, a11x);
        k1 = mfma64(v0.x - v0.y, v1.x, k1);
        k2 = mfma64(v0.x, v1.y - v1.x, k2);
        k3 = mfma64(-v0.y, v1.x + v1.y, k3);
        a00r = mfma64(v0.y, v0.y, a00r);
        a11r = mfma64(v1.y, v1.y, a11r);
      }
    }
    // f64 D[i][j]: lane j + 16 (i % 4), register i / 4
    for (int w = 0; w < 4; ++w) {
      if (wave == w) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int i = 4 * v + k4, j = l16;
          double2 g00 = make_double2(a00r[v], a00x[v]);
          double2 g01 = make_double2(k1[v] - k3[v], k1[v] + k2[v]);
          double2 g11 = make_double2(a11r[v], a11x[v]);
          if (w) {
            g00 = cadd(g00, sG[i][j]);
            g01 = cadd(g01, sG[i][JB + j]);
            g11 = cadd(g11, sG[JB + i][JB + j]);
          }
          sG[i][j] = g00;
          sG[i][JB + j] = g01;
          sG[JB + i][JB + j] = g11;
        }
      }
      __syncthreads();
    }
  } else {
    typedef float mfma_f4 __attribute__((ext_vector_type(4)));
    mfma_f4 b00r = {0, 0, 0, 0}, b00x = {0, 0, 0, 0}, b11r = {0, 0, 0, 0}, b11x = {0, 0, 0, 0};
    mfma_f4 h1 = {0, 0, 0, 0}, h2 = {0, 0, 0, 0}, h3 = {0, 0, 0, 0};
    double2 x0[PD], x1[PD];
#pragma unroll
    for (int d = 0; d < PD; ++d) ld_unit(wave + 4 * d, x0[d], x1[d]);
    for (int u0 = wave; u0 < nu; u0 += 4 * PD) {
#pragma unroll
      for (int d = 0; d < PD; ++d) {
        const float p0 = (float)x0[d].x, q0 = (float)x0[d].y, p1 = (float)x1[d].x, q1 = (float)x1[d].y;
        ld_unit(u0 + 4 * (d + PD), x0[d], x1[d]);
        b00r = __builtin_amdgcn_mfma_f32_16x16x4f32(p0, p0, b00r, 0, 0, 0);
        b11r = __builtin_amdgcn_mfma_f32_16x16x4f32(p1, p1, b11r, 0, 0, 0);
        b00x = __builtin_amdgcn_mfma_f32_16x16x4f32(p0, q0, b00x, 0, 0, 0);
        b11x = __builtin_amdgcn_mfma_f32_16x16x4f32(p1, q1, b11x, 0, 0, 0);
        h1 = __builtin_amdgcn_mfma_f32_16x16x4f32(p0 - q0, p1, h1, 0, 0, 0);
        h2 = __builtin_amdgcn_mfma_f32_16x16x4f32(p0, q1 - p1, h2, 0, 0, 0);
        h3 = __builtin_amdgcn_mfma_f32_16x16x4f32(-q0, p1 + q1, h3, 0, 0, 0);
        b00r = __builtin_amdgcn_mfma_f32_16x16x4f32(q0, q0, b00r, 0, 0, 0);
        b11r = __builtin_amdgcn_mfma_f32_16x16x4f32(q1, q1, b11r, 0, 0, 0);
      }
    }
    // f32 D[i][j]: lane j + 16 (i / 4), register i % 4; the four waves' partial tiles are
    // summed in FP64
    for (int w = 0; w < 4; ++w) {
      if (wave == w) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int i = 4 * k4 + v, j = l16;
          double2 g00 = make_double2((double)b00r[v], (double)b00x[v]);
          double2 g01 = make_double2((double)h1[v] - (double)h3[v], (double)h1[v] + (double)h2[v]);
          double2 g11 = make_double2((double)b11r[v], (double)b11x[v]);
          if (w) {
            g00 = cadd(g00, sG[i][j]);
            g01 = cadd(g01, sG[i][JB + j]);
            g11 = cadd(g11, sG[JB + i][JB + j]);
          }
          sG[i][j] = g00;
          sG[i][JB + j] = g01;
          sG[JB + i][JB + j] = g11;
        }
      }
      __syncthreads();
    }
  }
  {
    // Im of the diagonal tiles from X (X[i][j] - X[j][i]), and the lower off-diagonal tile
    const int i = tid >> 4, j = tid & 15;
    const double x00 = sG[i][j].y - sG[j][i].y, x11 = sG[JB + i][JB + j].y - sG[JB + j][JB + i].y;
    const double2 g10 = cconj(sG[j][JB + i]);
    __syncthreads();
    sG[i][j].y = x00;
    sG[JB + i][JB + j].y = x11;
    sG[JB + i][j] = g10;
  }
  __syncthreads();
  // ---- convergence test of this block pair: max |G_ij|^2 / (G_ii G_jj) over i != j ----
  {
    double m = 0.0;
    for (int e = tid; e < JP * JP; e += 256) {
      int i = e / JP, j = e % JP;
      if (i == j) continue;
      double aii = sG[i][i].x, ajj = sG[j][j].x;
      if (aii > floor2 && ajj > floor2) m = fmax(m, cabs2(sG[i][j]) / (aii * ajj));
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
    if (lane == 0) s_red[wave] = m;
    __syncthreads();
  }
  if (tid == 0) atomicAdd(T.visited, 1);
  if (fmax(fmax(s_red[0], s_red[1]), fmax(s_red[2], s_red[3])) <= tol * tol) return;

  // ---- apply prefetch: this wave's first column tile, all 32 rows ----
  const int nt = (CT + 15) >> 4;
  double2 bq[8];
  auto ld_tile = [&](int ctile, double2* b) {
    const int col = ctile * 16 + l16;
    const bool ok = ctile < nt && col < CT;
#pragma unroll
    for (int kk = 0; kk < 8; ++kk) {
      const int row = kk * 4 + k4;
      b[kk] = ok ? Y[at(row < JB ? ba * JB + row : bb * JB + row - JB, col)] : make_double2(0, 0);
    }
  };
  ld_tile(wave, bq);

  for (int e = tid; e < JP * JP; e += 256) sW[e / JP][e % JP] = make_double2((e / JP) == (e % JP) ? 1.0 : 0.0, 0.0);
  const double itol = tol * 0.25;
  if (tid < JP / 2) {
    const int p = circle_slot(tid, 0, JP), q = circle_slot(JP - 1 - tid, 0, JP);
    double c;
    double2 se;
    jacobi_rotation(sG[p][p].x, sG[q][q].x, sG[p][q], itol, floor2, c, se);
    s_c[0][tid] = c;
    s_se[0][tid] = se;
  }
  __syncthreads();
  // ---- inner sweep (one barrier per step) ----
  // G is held by slot: at every step, pair k is rows k and JP-1-k of the buffer.  Each step
  // writes its rotated 2x2 blocks where the next step reads them (a row moves from slot x to
  // x-1, slot 1 to JP-1, slot 0 stays), so every G address is fixed per thread; after JP-1
  // steps the slots are back in row order.  W stays in row order (single buffer).
  const int blk = c_inner_blocks.v[tid];
  const int bk = blk >> 4, bl = blk & 15;  // G block of rotation pairs bk, bl
  const int wi = tid >> 3, wl = tid & 7;   // W row wi, rotation pairs wl and wl + 8
  auto nxt = [](int x) { return x == JP - 1 ? 1 : x + 1; };  // row index one step later
  auto prv = [](int x) { return x == 0 ? 0 : (x == 1 ? JP - 1 : x - 1); };  // slot one step later
  const int pk = bk, qk = JP - 1 - bk, pl = bl, ql = JP - 1 - bl;
  const int dpk = prv(pk), dqk = prv(qk), dpl = prv(pl), dql = prv(ql);
  int pw0 = wl, qw0 = JP - 1 - wl, pw1 = wl + 8, qw1 = JP - 1 - (wl + 8);
  // next-step pair owned by this thread (jn), and which row of pair bk (rra) and of pair
  // bl (rrb) its members are
  int jn = -1, rra = 0, rrb = 0;
  if (bk == 0 && bl == 1) { jn = 0; rra = 0; rrb = 0; }
  else if (bk >= 2 && bl == bk - 2) { jn = bk - 1; rra = 0; rrb = 1; }
  else if (bk == 15 && bl == 14) { jn = 15; rra = 1; rrb = 1; }
  const int nsteps = inner_sweeps * (JP - 1);
  double2 (*gA)[JP + 1] = sG;
  double2 (*gB)[JP + 1] = sG2;
  const int slot = jn >= 0 ? jn : JP / 2;  // non-owners write a dummy slot
  for (int st = 0; st < nsteps; ++st) {
    const int buf = st & 1;
    // ---- all LDS reads of the step ----
    const double ck = s_c[buf][bk], cl = s_c[buf][bl];
    const double2 sek = s_se[buf][bk], sel = s_se[buf][bl];
    const double2 b00 = gA[pk][pl], b01 = gA[pk][ql], b10 = gA[qk][pl], b11 = gA[qk][ql];
    const double c2 = s_c[buf][wl], c3 = s_c[buf][wl + 8];
    const double2 se2 = s_se[buf][wl], se3 = s_se[buf][wl + 8];
    const double2 w0p = sW[wi][pw0], w0q = sW[wi][qw0], w1p = sW[wi][pw1], w1q = sW[wi][qw1];
    // ---- block B' = Jk^H B Jl,  J = [[c, se], [-conj(se), c]] ----
    const double2 nsel = make_double2(-sel.x, sel.y);
    const double2 m00 = cfma(b01, nsel, cscale(b00, cl)), m01 = cfma(b00, sel, cscale(b01, cl));
    const double2 m10 = cfma(b11, nsel, cscale(b10, cl)), m11 = cfma(b10, sel, cscale(b11, cl));
    const double2 nsek = make_double2(-sek.x, -sek.y), csek = cconj(sek);
    const double2 n00 = cfma(nsek, m10, cscale(m00, ck)), n01 = cfma(nsek, m11, cscale(m01, ck));
    const double2 n10 = cfma(csek, m00, cscale(m10, ck)), n11 = cfma(csek, m01, cscale(m11, ck));
    // ---- W' = W J on this thread's two pairs (identity rotations leave W unchanged) ----
    const double2 v0p = cfma(w0q, make_double2(-se2.x, se2.y), cscale(w0p, c2));
    const double2 v0q = cfma(w0p, se2, cscale(w0q, c2));
    const double2 v1p = cfma(w1q, make_double2(-se3.x, se3.y), cscale(w1p, c3));
    const double2 v1q = cfma(w1p, se3, cscale(w1q, c3));
    // ---- the next step's rotation: wave 0 (threads 0..15 own one each) ----
    double cn = 1.0;
    double2 sen = make_double2(0.0, 0.0);
    if (wave == 0) {
      const double ak = gA[pk][pk].x, bk2 = gA[qk][qk].x, al2 = gA[pl][pl].x, bl2 = gA[ql][ql].x;
      const double2 gk = gA[pk][qk], gl = gA[pl][ql];
      const double rk = 2.0 * ck * (gk.x * sek.x + gk.y * sek.y);
      const double ek = cabs2(sek), c2k = ck * ck;
      const double alpha = rra ? (ek * ak + c2k * bk2 + rk) : (c2k * ak + ek * bk2 - rk);
      const double rl = 2.0 * cl * (gl.x * sel.x + gl.y * sel.y);
      const double el = cabs2(sel), c2l = cl * cl;
      const double beta = rrb ? (el * al2 + c2l * bl2 + rl) : (c2l * al2 + el * bl2 - rl);
      const double2 gam = rra ? (rrb ? n11 : n10) : (rrb ? n01 : n00);
      jacobi_rotation(alpha, beta, gam, itol, floor2, cn, sen);
    }
    // ---- all LDS writes of the step ----
    gB[dpk][dpl] = n00;
    gB[dpk][dql] = n01;
    gB[dqk][dpl] = n10;
    gB[dqk][dql] = n11;
    sW[wi][pw0] = v0p;
    sW[wi][qw0] = v0q;
    sW[wi][pw1] = v1p;
    sW[wi][qw1] = v1q;
    if (wave == 0) {
      s_c[buf ^ 1][slot] = cn;
      s_se[buf ^ 1][slot] = sen;
    }
    __syncthreads();
    double2 (*tmp)[JP + 1] = gA;
    gA = gB;
    gB = tmp;
    if (wl) pw0 = nxt(pw0);
    qw0 = nxt(qw0);
    pw1 = nxt(pw1);
    qw1 = nxt(qw1);
  }

  // ---- apply Y' = W^T Y; wave owns column tiles wave, wave + 4, ... (all 32 rows) ----
  for (int ctile = wave; ctile < nt; ctile += 4) {
    // (A + iB)(C + iD) with three real products (Gauss): K1 = (A + B) C, K2 = A (D - C),
    // K3 = B (C + D); Re = K1 - K3, Im = K1 + K2
    mfma_d4 e0 = {0, 0, 0, 0}, f0 = {0, 0, 0, 0}, g0 = {0, 0, 0, 0};
    mfma_d4 e1 = {0, 0, 0, 0}, f1 = {0, 0, 0, 0}, g1 = {0, 0, 0, 0};
    const int ncol = (ctile + 4) * 16 + l16;
    const bool nok = ctile + 4 < nt && ncol < CT;
#pragma unroll
    for (int kk = 0; kk < 8; ++kk) {
      const double2 b = bq[kk];
      const double2 w0 = sW[kk * 4 + k4][l16];       // A[i][k] = W[k][i], rows 0..15
      const double2 w1 = sW[kk * 4 + k4][16 + l16];  // rows 16..31
      const int row = kk * 4 + k4;
      bq[kk] = nok ? Y[at(row < JB ? ba * JB + row : bb * JB + row - JB, ncol)] : make_double2(0, 0);
      const double dmc = b.y - b.x, cpd = b.x + b.y;
      e0 = mfma64(w0.x + w0.y, b.x, e0);
      e1 = mfma64(w1.x + w1.y, b.x, e1);
      f0 = mfma64(w0.x, dmc, f0);
      f1 = mfma64(w1.x, dmc, f1);
      g0 = mfma64(w0.y, cpd, g0);
      g1 = mfma64(w1.y, cpd, g1);
    }
    const int col = ctile * 16 + l16;
    if (col < CT) {
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        Y[at(ba * JB + 4 * v + k4, col)] = make_double2(e0[v] - g0[v], e0[v] + f0[v]);
        Y[at(bb * JB + 4 * v + k4, col)] = make_double2(e1[v] - g1[v], e1[v] + f1[v]);
      }
    }
  }
  if (tid == 0) atomicAdd(T.counter, 1);
}

__global__ __launch_bounds__(256) void k_fro2(const SplitTask* sp, const JacTask* jt) {
  __shared__ double red[256];
  const SplitTask& S = sp[blockIdx.x];
  const size_t n = (size_t)S.m * S.n;
  double a = 0.0;
  for (size_t e = threadIdx.x; e < n; e += 256) a += cabs2(S.theta[e]);
  red[threadIdx.x] = a;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0 && jt[blockIdx.x].fro2) *jt[blockIdx.x].fro2 = red[0];
}

hipError_t launch_fro2(const SplitTask* d_split, const JacTask* d_jac, int ntasks, hipStream_t s) {
  if (ntasks <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fro2, dim3(ntasks), dim3(256), 0, s, d_split, d_jac);
  return hipGetLastError();
}

hipError_t launch_jacobi_round(const JacTask* d_tasks, int ntasks, int total_pairs, int round,
                               int inner_sweeps, hipStream_t s, bool tiled) {
  if (total_pairs <= 0) return hipSuccess;
  if (tiled)
    hipLaunchKernelGGL((k_jacobi_round<JB, true>), dim3(total_pairs), dim3(256), 0, s, d_tasks, ntasks, round,
                       inner_sweeps);
  else
    hipLaunchKernelGGL((k_jacobi_round<JB, false>), dim3(total_pairs), dim3(256), 0, s, d_tasks, ntasks, round,
                       inner_sweeps);
  return hipGetLastError();
}

// One 16 x 16 tile per workgroup (blockIdx.x = 16-row block * max_ntc + tile column, blockIdx.y =
// task); thread (i, j) = (tid / 16, tid % 16) moves element (16 b + i, 16 t + j).
template <bool TO>
__global__ __launch_bounds__(256) void k_tiles(const JacTask* tasks, int max_ntc) {
  const JacTask& T = tasks[blockIdx.y];
  const int b = blockIdx.x / max_ntc, t = blockIdx.x % max_ntc;
  if (b >= T.r_pad / 16 || t >= T.ntc) return;
  const int row = 16 * b + (threadIdx.x >> 4), col = 16 * t + (threadIdx.x & 15);
  const size_t o = jac_tile_offset(row, col, T.ntc);
  if (TO) T.Yt[o] = col < T.c_tot ? T.Y[(size_t)row * T.ldy + col] : make_double2(0, 0);
  else if (col < T.c_tot) T.Y[(size_t)row * T.ldy + col] = T.Yt[o];
}

hipError_t launch_to_tiles(const JacTask* d_tasks, int ntasks, int max_blocks, int max_ntc, hipStream_t s) {
  if (ntasks <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_tiles<true>, dim3(max_blocks * max_ntc, ntasks), dim3(256), 0, s, d_tasks, max_ntc);
  return hipGetLastError();
}

hipError_t launch_from_tiles(const JacTask* d_tasks, int ntasks, int max_blocks, int max_ntc, hipStream_t s) {
  if (ntasks <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_tiles<false>, dim3(max_blocks * max_ntc, ntasks), dim3(256), 0, s, d_tasks, max_ntc);
  return hipGetLastError();
}

// ------------------------------------------------------------------ row norms
__global__ __launch_bounds__(256) void k_rownorm(const SplitTask* sp) {
  const SplitTask& S = sp[blockIdx.y];
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  const int i = blockIdx.x * 4 + wave;
  if (i >= S.r) return;
  double a = 0.0;
  const double2* row = S.Y + (size_t)i * S.ldy;
  for (int c = lane; c < S.c; c += 64) a += cabs2(row[c]);
  for (int off = 32; off > 0; off >>= 1) a += __shfl_down(a, off, 64);
  if (lane == 0) S.rownorm[i] = sqrt(a);
}

hipError_t launch_rownorm(const SplitTask* d_split, int ntasks, int max_r, hipStream_t s) {
  hipLaunchKernelGGL(k_rownorm, dim3((max_r + 3) / 4, ntasks), dim3(256), 0, s, d_split);
  return hipGetLastError();
}

// ------------------------------------------------------------------ sort + truncation rule
// Singular values sorted descending (bitonic in LDS, ties by row index).  Truncation
// follows truncateSvdTensors (ExaTnMpsVisitor.cpp:2424-2445): with the sqrt(S) split
// both partial norms of bond index k equal sigma_k (squared definition) or
// sqrt(sigma_k) (2-norm definition); K = i+1 for the first i whose norm is below the
// cutoff (the sub-threshold index itself is kept), else K0; then K = min(K, max_bond).
constexpr int SEL_MAX = 4096;
__global__ __launch_bounds__(256) void k_select(const SplitTask* sp) {
  __shared__ double key[SEL_MAX];
  __shared__ int idx[SEL_MAX];
  __shared__ double red[256];
  __shared__ int s_first;
  const SplitTask& S = sp[blockIdx.x];
  const int r = S.r;
  int P = 1;
  while (P < r) P <<= 1;
  const int tid = threadIdx.x;
  __shared__ int s_bad;
  if (tid == 0) {
    s_first = r;
    s_bad = 0;
  }
  __syncthreads();
  // a NaN or Inf singular value (non-finite input or overflow) is sorted as 0 and flagged:
  // the keys of real rows then stay >= 0 > -1 (the padding), so the first r indices of the
  // sorted order are always rows of Y
  for (int i = tid; i < P; i += 256) {
    double v = -1.0;
    if (i < r) {
      v = S.rownorm[i];
      if (!(v >= 0.0 && v <= DBL_MAX)) {
        v = 0.0;
        s_bad = 1;
      }
    }
    key[i] = v;
    idx[i] = i;
  }
  __syncthreads();
  for (int k = 2; k <= P; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = tid; i < P; i += 256) {
        int ixj = i ^ j;
        if (ixj > i) {
          bool desc = (i & k) == 0;
          double a = key[i], b = key[ixj];
          int ia = idx[i], ib = idx[ixj];
          // "a before b" in descending order with index tie-break
          bool a_first = (a > b) || (a == b && ia < ib);
          bool swap = desc ? !a_first : a_first;
          if (swap) {
            key[i] = b; key[ixj] = a;
            idx[i] = ib; idx[ixj] = ia;
          }
        }
      }
      __syncthreads();
    }
  }
  // first sub-threshold index
  // partial norm of index i: across a sqrt(S) split each factor carries sqrt(sigma_i), so
  // sigma_i (squared norm) or sqrt(sigma_i) (2-norm); a Kraus vector sigma_i v_i carries
  // sigma_i^2 or sigma_i
  for (int i = tid; i < r; i += 256) {
    const double sg = key[i];
    double pn = S.keep_rule ? (S.pn_l2 ? sg : sg * sg) : (S.pn_l2 ? sqrt(sg) : sg);
    if (pn < S.cutoff) atomicMin(&s_first, i);
  }
  __syncthreads();
  int K = s_first < r ? (S.keep_rule ? s_first : s_first + 1) : r;
  if (K > S.max_bond) K = S.max_bond;
  if (K < 1) K = 1;
  // kept / discarded weights (FP64)
  double kept = 0.0, disc = 0.0;
  for (int i = tid; i < r; i += 256) {
    double w = key[i] * key[i];
    if (i < K) kept += w; else disc += w;
  }
  red[tid] = kept;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) { if (tid < w) red[tid] += red[tid + w]; __syncthreads(); }
  kept = red[0];
  __syncthreads();
  red[tid] = disc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) { if (tid < w) red[tid] += red[tid + w]; __syncthreads(); }
  disc = red[0];
  __syncthreads();
  // sum of all singular values: the squared Frobenius norm of either sqrt(S) factor
  // before truncation, for the post-SVD norm validation (ExaTnMpsVisitor.cpp:1634-1661)
  double ssum = 0.0;
  for (int i = tid; i < r; i += 256) ssum += key[i];
  red[tid] = ssum;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) { if (tid < w) red[tid] += red[tid + w]; __syncthreads(); }
  ssum = red[0];
  double f = 1.0;
  if (S.renorm && kept > 0.0) f = sqrt((kept + disc) / kept);
  // scales: left factor U (s)^aL, right factor (s)^aR V^H with s = f * sigma.
  // Row k of Y (sigma_k xh_k^H) gives the factor on the X-row side (times sigma^-1);
  // the other factor is a unit vector recovered by a GEMM and scaled to (s)^a.
  const double aL = S.absorb == 0 ? 0.5 : 0.0, aR = 1.0 - aL;
  const double a_y = S.transposed ? aL : aR;
  const double a_w = S.transposed ? aR : aL;
  for (int k = tid; k < r; k += 256) {
    S.sig[k] = key[k];
    S.perm[k] = idx[k];
    if (k < K) {
      double sg = key[k];
      double gy = 0.0, gw = 0.0, is = 0.0;
      if (sg > 1e-150) {
        double sf = sg * f;
        gy = pow(sf, a_y) / sg;
        gw = pow(sf, a_w);
        is = 1.0 / sg;
      }
      S.scale_y[k] = gy;
      S.scale_w[k] = gw;
      S.inv_sig[k] = is;
    }
  }
  if (tid == 0) {
    *S.K = K;
    S.wstats[0] = disc;
    S.wstats[1] = kept;
    S.wstats[2] = kept + disc;
    S.wstats[3] = f;
    S.wstats[4] = ssum;
    S.wstats[5] = s_bad ? 1.0 : 0.0;
  }
}

hipError_t launch_select(const SplitTask* d_split, int ntasks, hipStream_t s) {
  hipLaunchKernelGGL(k_select, dim3(ntasks), dim3(256), 0, s, d_split);
  return hipGetLastError();
}

// ------------------------------------------------------------------ factor from Y rows
// not transposed (X = theta):  Aj[k][j] = Y[perm k][j] * scale_y[k]         (K x n)
// transposed (X = theta^H):    Ai[i][k] = conj(Y[perm k][i]) * scale_y[k]   (m x K)
template <typename T>
__global__ __launch_bounds__(256) void k_gather(const SplitTask* sp) {
  __shared__ double2 tile[32][33];
  const SplitTask& S = sp[blockIdx.y];
  const int K = *S.K;
  const int tx = threadIdx.x % 32, ty = threadIdx.x / 32;
  const int bt = blockIdx.x;
  if (!S.transposed) {
    const int len = S.n;
    const int tiles_c = (len + 31) / 32;
    const int k0 = (bt / tiles_c) * 32, c0 = (bt % tiles_c) * 32;
    if (k0 >= K) return;
    T* out = (T*)S.Aj;
    for (int y = ty; y < 32; y += 8) {
      int k = k0 + y, c = c0 + tx;
      if (k < K && c < len) {
        double2 v = S.Y[(size_t)S.perm[k] * S.ldy + c];
        st(out, (size_t)k * len + c, cscale(v, S.scale_y[k]));
      }
    }
  } else {
    const int len = S.m;
    const int tiles_i = (len + 31) / 32;
    const int k0 = (bt / tiles_i) * 32, i0 = (bt % tiles_i) * 32;
    if (k0 >= K) return;
    T* out = (T*)S.Ai;
    for (int y = ty; y < 32; y += 8) {
      int k = k0 + y, i = i0 + tx;
      double2 v = make_double2(0, 0);
      if (k < K && i < len) v = cscale(cconj(S.Y[(size_t)S.perm[k] * S.ldy + i]), S.scale_y[k]);
      tile[y][tx] = v;
    }
    __syncthreads();
    for (int y = ty; y < 32; y += 8) {
      int i = i0 + y, k = k0 + tx;
      if (i < len && k < K) st(out, (size_t)i * K + k, tile[tx][y]);
    }
  }
}

hipError_t launch_gather(const SplitTask* d_split, int ntasks, int max_k, int max_len, int fp32,
                         hipStream_t s) {
  int tiles = ((max_k + 31) / 32) * ((max_len + 31) / 32);
  if (fp32)
    hipLaunchKernelGGL(k_gather<float2>, dim3(tiles, ntasks), dim3(256), 0, s, d_split);
  else
    hipLaunchKernelGGL(k_gather<double2>, dim3(tiles, ntasks), dim3(256), 0, s, d_split);
  return hipGetLastError();
}

// ------------------------------------------------------------------ normalise recovered factor
// not transposed: columns k of Ai (m x K) scaled to norm scale_w[k];
// transposed:     rows k of Aj (K x n) scaled to norm scale_w[k].
// A vector that came out exactly zero stays zero.
template <typename T>
__global__ __launch_bounds__(256) void k_normalize(const SplitTask* sp) {
  __shared__ double part[8][33];
  __shared__ double fac[32];
  const SplitTask& S = sp[blockIdx.y];
  const int K = *S.K;
  const int k0 = blockIdx.x * 32;
  if (k0 >= K) return;
  const int tx = threadIdx.x % 32, ty = threadIdx.x / 32;
  if (!S.transposed) {
    T* A = (T*)S.Ai;
    const int k = k0 + tx;
    double a = 0.0;
    if (k < K)
      for (int i = ty; i < S.m; i += 8) a += cabs2(ld(A, (size_t)i * K + k));
    part[ty][tx] = a;
    __syncthreads();
    if (ty == 0) {
      double t = 0.0;
      for (int q = 0; q < 8; ++q) t += part[q][tx];
      fac[tx] = (k < K && t > 0.0) ? S.scale_w[k] / sqrt(t) : 0.0;
    }
    __syncthreads();
    if (k < K)
      for (int i = ty; i < S.m; i += 8) {
        size_t o = (size_t)i * K + k;
        st(A, o, cscale(ld(A, o), fac[tx]));
      }
  } else {
    T* A = (T*)S.Aj;
    // 8 rows per block pass, 32 threads per row
    for (int kr = ty; kr < 32; kr += 8) {
      const int k = k0 + kr;
      double a = 0.0;
      if (k < K)
        for (int j = tx; j < S.n; j += 32) a += cabs2(ld(A, (size_t)k * S.n + j));
      for (int off = 16; off > 0; off >>= 1) a += __shfl_xor(a, off, 32);
      const double fk = (k < K && a > 0.0) ? S.scale_w[k] / sqrt(a) : 0.0;
      if (k < K)
        for (int j = tx; j < S.n; j += 32) {
          size_t o = (size_t)k * S.n + j;
          st(A, o, cscale(ld(A, o), fk));
        }
    }
  }
}

hipError_t launch_normalize(const SplitTask* d_split, int ntasks, int max_k, int fp32, hipStream_t s) {
  dim3 grid((max_k + 31) / 32, ntasks);
  if (fp32)
    hipLaunchKernelGGL(k_normalize<float2>, grid, dim3(256), 0, s, d_split);
  else
    hipLaunchKernelGGL(k_normalize<double2>, grid, dim3(256), 0, s, d_split);
  return hipGetLastError();
}

// ------------------------------------------------------------------ generic batched GEMM
// C[i][j] = rowscale_i * colscale_j * sum_k a(i,k) b(k,j)
//   a(i,k): N  A[gA(i)*lda + k]   C conj(A[gA(i)*lda + k])
//           H  conj(A[k*lda + i]) T A[k*lda + i]
//   b(k,j): N  B[k*ldb + j]       T B[gB(j)*ldb + k]   H conj(B[gB(j)*ldb + k])
// optional diagonal operator diag(kd0, kd1) on a(i,k), selected by the physical index
// (k / sign_div) & 1 (Z, projectors, ... inserted on the contracted physical leg).
constexpr int GM = 32, GN = 32, GK = 16;

template <typename TA, typename TB, typename TC, int OPA, int OPB>
__global__ __launch_bounds__(256) void k_gemm(const GemmTask* tasks, int ntasks) {
  __shared__ double2 sA[GM][GK + 1];
  __shared__ double2 sB[GK][GN + 1];
  const int t = find_task(tasks, ntasks, (int)blockIdx.x, [](const GemmTask& x) { return x.tile_begin; });
  const GemmTask& T = tasks[t];
  const int tile = blockIdx.x - T.tile_begin;
  if (tile >= T.tiles_m * T.tiles_n) return;
  int M = T.M, N = T.N;
  if (T.Mdyn) M = min(M, *T.Mdyn);
  if (T.Ndyn) N = min(N, *T.Ndyn);
  const int i0 = (tile / T.tiles_n) * GM, j0 = (tile % T.tiles_n) * GN;
  if (i0 >= M || j0 >= N) return;
  const int Kd = T.K;
  const int ldc = T.ldc > 0 ? T.ldc : N;
  const TA* A = (const TA*)T.A;
  const TB* B = (const TB*)T.B;
  const int tid = threadIdx.x;
  const int ty = tid / 16, tx = tid % 16;
  double2 acc[2][2];
#pragma unroll
  for (int x = 0; x < 2; ++x)
#pragma unroll
    for (int y = 0; y < 2; ++y) acc[x][y] = make_double2(0, 0);

  for (int k0 = 0; k0 < Kd; k0 += GK) {
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      int e = tid + 256 * q;  // 512 = 32 x 16
      int i, k;
      if (OPA == 0 || OPA == 1) { i = e / GK; k = e % GK; }
      else { k = e / GM; i = e % GM; }
      double2 v = make_double2(0, 0);
      int gi = i0 + i, gk = k0 + k;
      if (gi < M && gk < Kd) {
        if (OPA == 0 || OPA == 1) {
          int sr = T.gatherA ? T.gatherA[gi] : gi;
          v = ld(A, (size_t)sr * T.lda + gk);
        } else {
          v = ld(A, (size_t)gk * T.lda + gi);
        }
        if (OPA == 1 || OPA == 2) v = cconj(v);
        if (T.sign_div > 0) v = cscale(v, ((gk / T.sign_div) & 1) ? T.kd1 : T.kd0);
      }
      sA[i][k] = v;
    }
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      int e = tid + 256 * q;  // 512 = 16 x 32
      int k, j;
      if (OPB == 0) { k = e / GN; j = e % GN; }
      else { j = e / GK; k = e % GK; }
      double2 v = make_double2(0, 0);
      int gj = j0 + j, gk = k0 + k;
      if (gj < N && gk < Kd) {
        if (OPB == 0) {
          v = ld(B, (size_t)gk * T.ldb + gj);
        } else {
          int sr = T.gatherB ? T.gatherB[gj] : gj;
          v = ld(B, (size_t)sr * T.ldb + gk);
          if (OPB == 2) v = cconj(v);
        }
      }
      sB[k][j] = v;
    }
    __syncthreads();
#pragma unroll 4
    for (int k = 0; k < GK; ++k) {
      double2 a0 = sA[ty * 2][k], a1 = sA[ty * 2 + 1][k];
      double2 b0 = sB[k][tx * 2], b1 = sB[k][tx * 2 + 1];
      acc[0][0] = cfma(a0, b0, acc[0][0]);
      acc[0][1] = cfma(a0, b1, acc[0][1]);
      acc[1][0] = cfma(a1, b0, acc[1][0]);
      acc[1][1] = cfma(a1, b1, acc[1][1]);
    }
    __syncthreads();
  }
  TC* Cp = (TC*)T.C;
#pragma unroll
  for (int x = 0; x < 2; ++x) {
    int i = i0 + ty * 2 + x;
    if (i >= M) continue;
    double rs = T.rowscale ? T.rowscale[i] : 1.0;
#pragma unroll
    for (int y = 0; y < 2; ++y) {
      int j = j0 + tx * 2 + y;
      if (j >= N) continue;
      double s = rs * (T.colscale ? T.colscale[j] : 1.0);
      double2 v = cscale(acc[x][y], s);
      size_t o = (size_t)i * ldc + j;
      if (T.accumulate) v = cadd(v, ld(Cp, o));
      st(Cp, o, v);
    }
  }
}

template <int OPA, int OPB>
static hipError_t gemm_types(const GemmTask* d, int nt, int tiles, int ta, int tb, int tc, hipStream_t s) {
  dim3 g(tiles), b(256);
#define HIPMPS_GEMM_CASE(A_, B_, C_, a, b_, c)                                                   \
  if (ta == a && tb == b_ && tc == c) {                                                          \
    hipLaunchKernelGGL((k_gemm<A_, B_, C_, OPA, OPB>), g, b, 0, s, d, nt);                       \
    return hipGetLastError();                                                                    \
  }
  HIPMPS_GEMM_CASE(double2, double2, double2, 0, 0, 0)
  HIPMPS_GEMM_CASE(double2, double2, float2, 0, 0, 1)
  HIPMPS_GEMM_CASE(float2, double2, double2, 1, 0, 0)
  HIPMPS_GEMM_CASE(double2, float2, double2, 0, 1, 0)
  HIPMPS_GEMM_CASE(float2, float2, double2, 1, 1, 0)
#undef HIPMPS_GEMM_CASE
  return hipErrorInvalidValue;
}

hipError_t launch_gemm(const GemmTask* d_tasks, int ntasks, int total_tiles, int opa, int opb, int ta, int tb,
                       int tc, hipStream_t s) {
  if (total_tiles <= 0) return hipSuccess;
#define HIPMPS_OP_CASE(a, b)                                                                     \
  if (opa == a && opb == b) return gemm_types<a, b>(d_tasks, ntasks, total_tiles, ta, tb, tc, s);
  HIPMPS_OP_CASE(0, 0)
  HIPMPS_OP_CASE(0, 2)
  HIPMPS_OP_CASE(0, 3)
  HIPMPS_OP_CASE(1, 3)
  HIPMPS_OP_CASE(2, 0)
#undef HIPMPS_OP_CASE
  return hipErrorInvalidValue;
}

// ------------------------------------------------------------------ chain step (amplitudes / slices)
// amplitude: vout[row][r] = sum_l vin[row][l] A[l][bits[row][site]][r]
// slice:     vout[o][r]   = sum_l vin[o % rows_in][l] A[l][p][r], p = open ? o / rows_in : bit
template <typename T>
__global__ __launch_bounds__(256) void k_chain(const ChainTask* tasks, int ntasks) {
  const int t = find_task(tasks, ntasks, (int)blockIdx.x, [](const ChainTask& x) { return x.block_begin; });
  const ChainTask& C = tasks[t];
  const int b = blockIdx.x - C.block_begin;
  if (b >= C.blocks) return;
  const int rblocks = (C.cr + 255) / 256;
  const int row = b / rblocks, r = (b % rblocks) * 256 + threadIdx.x;
  if (row >= C.rows_out || r >= C.cr) return;
  int src, p;
  if (C.mode == 0) {
    src = row;
    p = C.bits[(size_t)row * C.n + C.site];
  } else {
    src = row % C.rows_in;
    p = C.slice_bit < 0 ? row / C.rows_in : C.slice_bit;
  }
  const T* A = (const T*)C.A;
  const double2* v = C.vin + (size_t)src * C.cl;
  double2 acc = make_double2(0, 0);
  for (int l = 0; l < C.cl; ++l) acc = cfma(v[l], ld(A, ((size_t)l * 2 + p) * C.cr + r), acc);
  C.vout[(size_t)row * C.cr + r] = acc;
}

hipError_t launch_chain(const ChainTask* d_tasks, int ntasks, int total_blocks, int fp32, hipStream_t s) {
  if (total_blocks <= 0) return hipSuccess;
  if (fp32)
    hipLaunchKernelGGL(k_chain<float2>, dim3(total_blocks), dim3(256), 0, s, d_tasks, ntasks);
  else
    hipLaunchKernelGGL(k_chain<double2>, dim3(total_blocks), dim3(256), 0, s, d_tasks, ntasks);
  return hipGetLastError();
}

// ------------------------------------------------------------------ sum_{ij} a_ij b_ij
__global__ __launch_bounds__(256) void k_trace_prod(const double2* const* a, const double2* const* b,
                                                    const int* dims, double2* out) {
  __shared__ double2 red[256];
  const int t = blockIdx.x;
  const size_t n = (size_t)dims[t] * dims[t];
  double2 acc = make_double2(0, 0);
  for (size_t e = threadIdx.x; e < n; e += 256) acc = cfma(a[t][e], b[t][e], acc);
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if (threadIdx.x < w) red[threadIdx.x] = cadd(red[threadIdx.x], red[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) out[t] = red[0];
}

hipError_t launch_trace_prod(const double2* const* d_a, const double2* const* d_b, const int* d_dims,
                             double2* d_out, int ntasks, hipStream_t s) {
  if (ntasks <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_trace_prod, dim3(ntasks), dim3(256), 0, s, d_a, d_b, d_dims, d_out);
  return hipGetLastError();
}

// ------------------------------------------------------------------ misc
__global__ void k_check_finite(const double* p, size_t n, int* flag) {
  size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n && !isfinite(p[i])) atomicOr(flag, 1);
}
__global__ void k_check_finite_f(const float* p, size_t n, int* flag) {
  size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n && !isfinite(p[i])) atomicOr(flag, 1);
}
hipError_t launch_check_finite(const void* p, size_t n_complex, int fp32, int* d_flag, hipStream_t s) {
  size_t n = 2 * n_complex;
  if (n == 0) return hipSuccess;
  dim3 g((unsigned)((n + 255) / 256));
  if (fp32) hipLaunchKernelGGL(k_check_finite_f, g, dim3(256), 0, s, (const float*)p, n, d_flag);
  else hipLaunchKernelGGL(k_check_finite, g, dim3(256), 0, s, (const double*)p, n, d_flag);
  return hipGetLastError();
}

__global__ void k_fill_identity1(double2* p) { p[0] = make_double2(1.0, 0.0); }
hipError_t launch_fill_identity1(double2* p, hipStream_t s) {
  hipLaunchKernelGGL(k_fill_identity1, dim3(1), dim3(1), 0, s, p);
  return hipGetLastError();
}

__global__ __launch_bounds__(256) void k_fill_c(double2* p, size_t n, double2 v) {
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i < n) p[i] = v;
}
hipError_t launch_fill_c(double2* p, size_t n, double2 v, hipStream_t s) {
  if (n == 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill_c, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, p, n, v);
  return hipGetLastError();
}

// density-matrix readout: label A holds, in base-4 digit t, the ket (2s) and bra (1s) component
// of the t-th measured site; element A goes to out[x + 2^k x'] with x, x' the components placed
// at the bits P.pos[t]
__global__ __launch_bounds__(256) void k_rdm_scatter(const double2* __restrict__ E, double2* __restrict__ out,
                                                     size_t count, RdmPerm P) {
  for (size_t a = (size_t)blockIdx.x * 256 + threadIdx.x; a < count; a += (size_t)gridDim.x * 256) {
    size_t x = 0, xb = 0;
    for (int t = 0; t < P.k; ++t) {
      const size_t d = (a >> (2 * t)) & 3;
      x |= (d >> 1) << P.pos[t];
      xb |= (d & 1) << P.pos[t];
    }
    out[x + (xb << P.k)] = E[a];
  }
}

hipError_t launch_rdm_scatter(const double2* E, double2* out, const RdmPerm& P, hipStream_t s) {
  const size_t count = (size_t)1 << (2 * P.k);
  const unsigned blocks = (unsigned)std::min<size_t>((count + 255) / 256, 65536);
  hipLaunchKernelGGL(k_rdm_scatter, dim3(blocks), dim3(256), 0, s, E, out, count, P);
  return hipGetLastError();
}

}  // namespace hipmps
