// kernels.h -- device task descriptors and launchers of libhipmps (internal).
//
// All kernels work on batches described by task arrays that live in device memory,
// so one launch serves every circuit / bond of a wave.  Site tensors are stored
// A[l][p][r] row-major (r fastest) as complex128 (double2) or complex64 (float2).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace hipmps {

// Block size of the one-sided block-Jacobi SVD: a block pair is 2*JB rows.
constexpr int JB = 16;
constexpr int JP = 2 * JB;
constexpr int WS_N = 6;  // doubles of split statistics per SVD (SplitTask::wstats)

// theta' = G . (A_i x A_j): merge of two neighbouring sites and the 2-qubit gate
// (ExaTnMpsVisitor.cpp:1401-1524), written as the (2cl) x (2cr) matrix with
// rows (l, a) and columns (b, r).
struct ThetaTask {
  const void* Ai;    // cl x 2 x cm
  const void* Aj;    // cm x 2 x cr
  double2* theta;    // 2cl x 2cr, ld = 2cr
  int cl, cm, cr;
  int tile_begin;    // first block of this task in the grid
  int tiles_l, tiles_r;
  int pad[2];
  double2 G[16];     // G[(a,b)*4 + (a',b')], a = bit of the left site
};

// One Jacobi problem: orthogonalise the rows of Y (r_pad x c_tot, row-major, ld = ldy)
// with respect to their first c columns (the data X); every rotation is applied to all
// c_tot columns (c_tot = c: the rotation product is not carried).
struct JacTask {
  double2* Y;
  int ldy, r, r_pad, c;
  int c_tot;
  int nb;            // r_pad / JB (even)
  int pair_begin;    // first block of this task in the grid
  int* counter;      // rotated block pairs in the current sweep
  int* visited;      // block pairs visited in the current sweep
  double* fro2;      // ||X||_F^2 (k_fro2), or null: rows with squared norm below JAC_FLOOR2 * fro2 are
                     // numerically zero (rounding residue) and neither rotate nor count against
                     // convergence
  double tol;        // relative off-diagonal below which a block pair counts as orthogonal
  int gram32;        // this sweep forms the Gram with FP32 MFMA (rotations and updates stay FP64)
  int pad;
  double2* Yt;       // the same rows in tiles (jac_tile_offset), or null: the sweeps of the main path
                     // run on Yt (launch_to_tiles before, launch_from_tiles after)
  int ntc;           // 16-column tiles per 16-row block of Yt: ceil(c_tot / 16)
  int pad2;
};

// Tiled row layout of the Jacobi sweeps: 16-row blocks, each a sequence of 16 x 16 tiles, each
// tile 4 x 4 sub-tiles of 4 rows x 4 columns (256 B, row-major) in row-major order.  A Gram load
// of 16 rows x 4 columns is then four 256 B segments (not sixteen 64 B ones 8 KB apart), and an
// apply load of 4 rows x 16 columns one 1 KB segment.
__host__ __device__ inline size_t jac_tile_offset(int row, int col, int ntc) {
  const int i = row & 15;
  return ((size_t)((row >> 4) * ntc + (col >> 4)) * 16 + (i >> 2) * 4 + ((col & 15) >> 2)) * 16 + (i & 3) * 4 +
         (col & 3);
}

// Everything needed after convergence: singular values, truncation, factors.
// Split after convergence.  Row k of the final Y is sigma_k xh_k^H (xh_k the k-th right
// singular vector of X); the other side's singular vector is recovered as
// normalize(X xh_k), whose norm is fixed to 1 so that noise in the direction of a
// numerically-zero singular value cannot be amplified.
struct SplitTask {
  const double2* theta;  // m x n (ld n)
  const double2* Y;      // r_pad x c (ld ldy): final rows sigma_k xh_k^H
  int m, n, r, c, ldy;
  int transposed;        // 0: X = theta (m <= n); 1: X = theta^H (m > n)
  double* rownorm;       // r
  double* sig;           // r (sorted, descending)
  int* perm;             // r
  double* scale_y;       // K: factor gathered from row k of Y (sigma^-1 (f sigma)^a)
  double* scale_w;       // K: target norm (f sigma)^a of the recovered factor
  double* inv_sig;       // K: 1 / sigma_k (0 for sigma_k = 0), GEMM operand scale
  int* K;                // out: new bond dimension
  double* wstats;        // out: [discarded, kept, total, renorm factor, sum of all sigma,
                         //       1 if a singular value was NaN / Inf]
  void* Ai;              // out site i: cl x 2 x K
  void* Aj;              // out site j: K x 2 x cr
  int cl, cr;
  double cutoff;
  int max_bond;
  int pn_l2;
  int absorb;            // 0 sqrt split, 1 all to the right factor
  int renorm;
  int site_fp32;         // site storage is complex64
  int keep_rule;         // 0: keep i + 1 for the first sub-threshold index i (bond);
                         // 1: keep i (Kraus leg, ExaTnPmpsVisitor.cpp:1060-1072)
};

// Generic batched complex GEMM: C (M x N) = rowscale . op(A) op(B) . colscale
// op codes: 0 N, 1 C (conjugate), 2 H (conjugate transpose), 3 T (transpose).
struct GemmTask {
  const void* A;
  const void* B;
  void* C;
  int M, N, K;
  int lda, ldb, ldc;
  const int* Mdyn;        // if set, M = min(M, *Mdyn)
  const int* Ndyn;        // if set, N = min(N, *Ndyn)
  const int* gatherA;     // N/C: storage row of A for output row i
  const int* gatherB;     // T/H: storage row of B for output column j
  const double* rowscale; // per output row
  const double* colscale; // per output column
  int sign_div;           // >0: a(i,k) *= ((k / sign_div) & 1) ? kd1 : kd0
  int tile_begin, tiles_m, tiles_n;
  int accumulate;         // C += result
  double kd0, kd1;        // diagonal operator on the physical index (Z: 1, -1)
};

// Row-oriented blocked Householder LQ of M (rows x cols, rows <= cols, ld), 16-row panels.
// After the factorization row k of M holds L[k][0..k] and, above the diagonal, the tail of
// the unit reflector row w_k (support columns k..cols-1; head w_k[k] in W0[k]); T holds per
// panel the 16 x 16 upper-triangular factor with H_k0 ... H_k0+15 = I - W^H T W.
// Tgt (rows_t x cols, ld_t) is the matrix the inverse transform is applied to.
struct LqTask {
  double2* M;
  double2* W0;
  double2* T;
  double2* Tgt;
  int ld, rows, cols;
  int ld_t, rows_t;
  int pad[3];
};

// Copy tasks around the LQ: src (ld_src), dst (ld_dst), n columns, rows_dst rows of dst.
struct TriTask {
  const double2* src;
  double2* dst;
  int ld_src, n, rows_dst, ld_dst;
};

// zero the rows of the n x n factor tasks[t].dst whose norm is below 1e-14 of its Frobenius norm,
// where that moves no singular value by more than 1e-3 sigma_cut[t] (<= 0: skip); n <= 4096
hipError_t launch_rank_cut(const TriTask* d_tasks, const double* d_sigma_cut, int ntasks, hipStream_t s);

// Chain step for amplitudes / slices: rows of partial contractions v[row][l].
struct ChainTask {
  const double2* vin;   // rows_in x cl
  double2* vout;        // rows_out x cr
  const void* A;        // cl x 2 x cr
  int cl, cr, rows_in, rows_out;
  const int8_t* bits;   // amplitude mode: bits[row * n + site]; slice mode: unused
  int n, site;
  int slice_bit;        // slice mode: 0/1 fixed, -1 open (rows double)
  int mode;             // 0 amplitude, 1 slice
  int block_begin, blocks;
};

// ---- launchers (stream-ordered, return hipError_t) --------------------------------
hipError_t launch_apply_1q(void* site, int fp32, int cl, int cr, const double2 m[4], hipStream_t s);
hipError_t launch_theta(const ThetaTask* d_tasks, int ntasks, int total_tiles, int fp32,
                        hipStream_t s);
hipError_t launch_jacobi_prep(const SplitTask* d_split, const JacTask* d_jac, int ntasks,
                              int max_rows, int max_cols, hipStream_t s);
// (1e-15)^2: a row whose norm is below 1e-15 ||X||_F lies under FP64 resolution of any result
// computed from X, so the relative convergence test has no meaning for it
constexpr double JAC_FLOOR2 = 1e-30;
// *jt[t].fro2 = ||theta'_t||_F^2 (= ||X||_F^2 = ||R2||_F^2), one workgroup per task, fixed order
hipError_t launch_fro2(const SplitTask* d_split, const JacTask* d_jac, int ntasks, hipStream_t s);
hipError_t launch_jacobi_round(const JacTask* d_tasks, int ntasks, int total_pairs, int round,
                               int inner_sweeps, hipStream_t s, bool tiled = false);
// Y (row-major, ldy) <-> Yt (tiled, ntc) for every task: rows [0, r_pad), columns [0, c_tot);
// to_tiles writes zeros into the tile columns past c_tot.  max_blocks = largest r_pad / 16.
hipError_t launch_to_tiles(const JacTask* d_tasks, int ntasks, int max_blocks, int max_ntc, hipStream_t s);
hipError_t launch_from_tiles(const JacTask* d_tasks, int ntasks, int max_blocks, int max_ntc, hipStream_t s);
// Householder LQ of every task's M (panel by panel, one launch per panel).
hipError_t launch_lq(const LqTask* d_tasks, int ntasks, int max_rows, hipStream_t s);
// The same LQ with look-ahead on two streams: on sa, the next panel's rows take the current block
// reflector and the panel is factored with its rows in registers; on sb, the trailing rows take the
// block reflector meanwhile (events ea, eb order the two).  Rows of at most LQ_REG_MAX_COLS columns
// (hipErrorNotSupported otherwise); on return every later launch on sa follows all of it.
constexpr int LQ_REG_MAX_COLS = 512;
hipError_t launch_lq_lookahead(const LqTask* d_tasks, int ntasks, int max_rows, int max_cols, hipStream_t sa,
                               hipStream_t sb, hipEvent_t ea, hipEvent_t eb);
// Tgt <- Tgt Q~ (Q~ = H_{rows-1} ... H_0 restricted to the first rows of M's row space, applied
// panel by panel from the last), i.e. [Y | 0] Q~ for Tgt = [Y | 0].
hipError_t launch_lq_apply_q(const LqTask* d_tasks, int ntasks, int max_rows, int max_rows_t, hipStream_t s);
// dst (rows_dst x n, ld ld_dst) = conjugate transpose of the lower triangle of src (n x n, ld
// ld_src); rows n .. rows_dst - 1 zero.
hipError_t launch_lower_to_upper_h(const TriTask* d_tasks, int ntasks, int max_n, int max_rows_dst, hipStream_t s);
// dst[i][j] = src[i][j] for j < n, 0 for n <= j < ld_dst, i < rows_dst.
hipError_t launch_widen(const TriTask* d_tasks, int ntasks, int max_cols, int max_rows_dst, hipStream_t s);
hipError_t launch_rownorm(const SplitTask* d_split, int ntasks, int max_r, hipStream_t s);
hipError_t launch_select(const SplitTask* d_split, int ntasks, hipStream_t s);
hipError_t launch_gather(const SplitTask* d_split, int ntasks, int max_k, int max_len, int fp32,
                         hipStream_t s);
// rescale the GEMM-recovered factor: columns of Ai (not transposed) or rows of Aj
// (transposed) to norm scale_w[k]
hipError_t launch_normalize(const SplitTask* d_split, int ntasks, int max_k, int fp32, hipStream_t s);
// element types: 0 = complex128, 1 = complex64
hipError_t launch_gemm(const GemmTask* d_tasks, int ntasks, int total_tiles, int opa, int opb,
                       int ta, int tb, int tc, hipStream_t s);
hipError_t launch_chain(const ChainTask* d_tasks, int ntasks, int total_blocks, int fp32,
                        hipStream_t s);
hipError_t launch_trace_prod(const double2* const* d_a, const double2* const* d_b,
                             const int* d_dims, double2* d_out, int ntasks, hipStream_t s);
hipError_t launch_check_finite(const void* p, size_t n_complex, int fp32, int* d_flag,
                               hipStream_t s);
hipError_t launch_fill_identity1(double2* p, hipStream_t s);
// p[i] = v for i < n
hipError_t launch_fill_c(double2* p, size_t n, double2 v, hipStream_t s);
// reduced density matrix of k <= 15 qubits: E[A] (4^k labels, base-4 digit t = 2 b_t + b'_t)
// scattered to out[x + 2^k x'], x = sum_t b_t 2^pos[t], x' = sum_t b'_t 2^pos[t]
struct RdmPerm {
  int k;
  int pos[15];
};
hipError_t launch_rdm_scatter(const double2* E, double2* out, const RdmPerm& P, hipStream_t s);

// ---- purified sites (Kraus leg), kraus.hip ------------------------------------------------
// dst[l][b][a][m] = src[l][a][b][m] for m < M (M = *m_dyn when m_dyn is not null, at most
// m_max); element type complex128 (fp32 = 0) or complex64.
hipError_t launch_swap_mid(const void* src, void* dst, int L, int A, int B, int m_max, const int* m_dyn,
                           int fp32, hipStream_t s);
// X[k * M + o][(l * 2 + p) * cr + r] = sum_p' K_o[p][p'] A[l][p'][k][r] for the site A
// (cl x 2 x kap x cr), rows kap * M .. rows_pad - 1 zero; K: M row-major 2 x 2 matrices.
hipError_t launch_kraus_expand(const void* site, int fp32, int cl, int kap, int cr, const double2* d_kraus,
                               int n_ops, double2* X, int rows_pad, hipStream_t s);
// site[l][p][j][r] = Y[perm[j]][(l * 2 + p) * cr + r] for j < *K (at most k_max).
// the adjoint layout for channels with more rows than columns (rows = kap n_ops, c = 2 cl cr):
// Y (rows_pad x ldy, ldy = rows + c) = [X^H | I_c], pad rows zero; after the Jacobi, Kraus row j =
// rownorm[perm[j]] * Y[perm[j]][off + col] (off = rows)
hipError_t launch_kraus_expand_adj(const void* site, int fp32, int cl, int kap, int cr, const double2* d_kraus,
                                   int n_ops, double2* Y, int ldy, int rows_pad, hipStream_t s);
hipError_t launch_kraus_gather_adj(const double2* Y, int ldy, int off, const double* rownorm, const int* perm,
                                   const int* K, int k_max, void* site, int fp32, int cl, int cr, hipStream_t s);
hipError_t launch_kraus_gather(const double2* Y, int ldy, const int* perm, const int* K, int k_max, void* site,
                               int fp32, int cl, int cr, hipStream_t s);

}  // namespace hipmps
