// mps_oracle.cpp -- CPU FP64 restatement of the exatn-mps circuit simulator.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP library: only
// tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
// The product path (libhipmps) never links, calls or falls back to it.
//
// What it restates (file:line citations are relative to the tnqvm source tree):
//   * gate matrices ........................ tnqvm/base/Gates.hpp:131-339
//   * gate flattening / operand order ...... tnqvm/visitors/exatn-mps/ExatnUtils.cpp:36-133
//                                            (row-major 4x4, first operand = MSB)
//   * dense state vector ................... tnqvm/utils/GateMatrixAlgebra.hpp:63-123, 213-219
//                                            (qubit k <-> bit k of the index)
//   * sampling ............................. GateMatrixAlgebra.hpp:125-156 + utils/RandomEngine.hpp
//   * sequential sampling (n >= 20) ........ ExaTnMpsVisitor.cpp:2211-2364 (getMeasureSample)
//   * MPS initial state |0..0>, bonds 1 .... ExaTnMpsVisitor.cpp:280-326
//   * 1q gate into a site .................. ExaTnMpsVisitor.cpp:1185-1292
//   * 2q gate: merge, gate, full SVD with
//     sqrt(S) on both factors, stabilise ... ExaTnMpsVisitor.cpp:1387-1627, 141-161
//   * truncation by partial norms, keep the
//     first sub-threshold index, clamp ..... ExaTnMpsVisitor.cpp:2366-2536
//   * exp-val-z parity (unnormalised) ...... ExaTnMpsVisitor.cpp:616-643
//   * bit-string amplitude / slice ......... ExaTnMpsVisitor.cpp:2588-2675, 794-820
//
// Parity pinning (see tests/test_oracle.py): the dense state vector is checked against
// closed forms from the reference tests; the MPS restatement against the dense
// state vector; the Jacobi SVD against LAPACK zgesdd spectra written to
// tests/golden by numpy; the truncation rule against its special cases.
//
// Site layout: A_k[l][p][r] row-major (r fastest), dims (chi_k, 2, chi_{k+1}),
// chi_0 = chi_n = 1.  Complex numbers cross the C ABI as interleaved (re, im) doubles.

#include <algorithm>
#include <array>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <random>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

using cd = std::complex<double>;

namespace {

// ------------------------------------------------------------------------------------
// Gate matrices: Gates.hpp:131-339.  Out = row-major 2x2 or 4x4.
// ------------------------------------------------------------------------------------
int gate_matrix(const std::string& g, const double* p, int np, cd* M) {
  const cd I(0.0, 1.0);
  auto need = [&](int k) { return np >= k; };
  auto set2 = [&](cd a, cd b, cd c, cd d) { M[0] = a; M[1] = b; M[2] = c; M[3] = d; return 1; };
  auto diag4 = [&](cd a, cd b, cd c, cd d) {
    for (int i = 0; i < 16; ++i) M[i] = 0.0;
    M[0] = a; M[5] = b; M[10] = c; M[15] = d;
    return 2;
  };
  if (g == "I") return set2(1, 0, 0, 1);
  if (g == "H") return set2(M_SQRT1_2, M_SQRT1_2, M_SQRT1_2, -M_SQRT1_2);  // Gates.hpp:142-148
  if (g == "X") return set2(0, 1, 1, 0);
  if (g == "Y") return set2(0, -I, I, 0);
  if (g == "Z") return set2(1, 0, 0, -1);
  if (g == "S") return set2(1, 0, 0, I);    // not in Gates.hpp; standard phase gate
  if (g == "Sdg") return set2(1, 0, 0, -I);
  if (g == "T") return set2(1, 0, 0, std::exp(I * M_PI_4));   // Gates.hpp:206-213
  if (g == "Tdg") return set2(1, 0, 0, std::exp(-I * M_PI_4));
  if (g == "Rx" && need(1)) {  // Gates.hpp:177-184
    double c = std::cos(0.5 * p[0]), s = std::sin(0.5 * p[0]);
    return set2(c, -I * s, -I * s, c);
  }
  if (g == "Ry" && need(1)) {  // Gates.hpp:186-194
    double c = std::cos(0.5 * p[0]), s = std::sin(0.5 * p[0]);
    return set2(c, -s, s, c);
  }
  if (g == "Rz" && need(1)) {  // Gates.hpp:196-204
    return set2(std::exp(-0.5 * I * p[0]), 0, 0, std::exp(0.5 * I * p[0]));
  }
  if (g == "U" && need(3)) {  // Gates.hpp:224-235
    double t = p[0], ph = p[1], la = p[2];
    return set2(std::cos(t / 2.0), -std::exp(I * la) * std::sin(t / 2.0),
                std::exp(I * ph) * std::sin(t / 2.0), std::exp(I * (ph + la)) * std::cos(t / 2.0));
  }
  if (g == "CNOT" || g == "CX") {  // Gates.hpp:237-246
    diag4(1, 1, 0, 0);
    M[11] = 1; M[14] = 1;
    return 2;
  }
  if (g == "CZ") return diag4(1, 1, 1, -1);
  if (g == "CY") {
    diag4(1, 1, 0, 0);
    M[11] = -I; M[14] = I;
    return 2;
  }
  if (g == "CH") {
    diag4(1, 1, M_SQRT1_2, -M_SQRT1_2);
    M[11] = M_SQRT1_2; M[14] = M_SQRT1_2;
    return 2;
  }
  if (g == "CRZ" && need(1)) return diag4(1, 1, std::exp(-0.5 * I * p[0]), std::exp(0.5 * I * p[0]));
  if (g == "CPhase" && need(1)) return diag4(1, 1, 1, std::exp(I * p[0]));
  if (g == "Swap") {
    for (int i = 0; i < 16; ++i) M[i] = 0.0;
    M[0] = 1; M[6] = 1; M[9] = 1; M[15] = 1;
    return 2;
  }
  if (g == "iSwap") {
    for (int i = 0; i < 16; ++i) M[i] = 0.0;
    M[0] = 1; M[6] = I; M[9] = I; M[15] = 1;
    return 2;
  }
  if (g == "fSim" && need(2)) {  // Gates.hpp:325-334
    for (int i = 0; i < 16; ++i) M[i] = 0.0;
    M[0] = 1;
    M[5] = std::cos(p[0]); M[6] = -I * std::sin(p[0]);
    M[9] = -I * std::sin(p[0]); M[10] = std::cos(p[0]);
    M[15] = std::exp(-I * p[1]);
    return 2;
  }
  return 0;
}

void load(const double* in, cd* out, size_t n) {
  for (size_t i = 0; i < n; ++i) out[i] = cd(in[2 * i], in[2 * i + 1]);
}
void store(const cd* in, double* out, size_t n) {
  for (size_t i = 0; i < n; ++i) { out[2 * i] = in[i].real(); out[2 * i + 1] = in[i].imag(); }
}

// ------------------------------------------------------------------------------------
// One-sided (Hestenes) Jacobi SVD, FP64, cyclic ordering with a parallel-friendly
// round-robin schedule.  Input A (m x n, row-major).  Produces the factors of the
// sqrt(S)-split used by the exatn-mps visitor: L = U sqrt(S) (m x K0), R = sqrt(S) Vh
// (K0 x n), K0 = min(m, n), singular values sorted descending.
// Rotation (Golub & Van Loan, Alg. 8.4.1 with a phase alignment for the complex case):
//   for columns a_p, a_q:  alpha=|a_p|^2, beta=|a_q|^2, gamma = a_p^H a_q = |gamma| e^{i phi}
//   zeta = (beta-alpha)/(2|gamma|), t = sign(zeta)/(|zeta|+sqrt(1+zeta^2)), c = 1/sqrt(1+t^2), s = c t
//   [a_p, a_q] <- [a_p, a_q] [[c, s e^{i phi}], [-s e^{-i phi}, c]]
// ------------------------------------------------------------------------------------
struct JacobiResult {
  std::vector<double> s;  // descending
  int sweeps = 0;
};

// Column Jacobi on X (rows x cols, column-major storage colmaj[j*rows + i]), accumulating
// V (cols x cols, column-major).  Returns sweeps used.
int column_jacobi(int rows, int cols, std::vector<cd>& X, std::vector<cd>& V, int max_sweeps,
                  double tol, int nthreads) {
  V.assign((size_t)cols * cols, 0.0);
  for (int j = 0; j < cols; ++j) V[(size_t)j * cols + j] = 1.0;
  // round-robin (tournament) ordering over an even number of slots
  int slots = cols + (cols & 1);
  std::vector<int> ring(slots);
  std::iota(ring.begin(), ring.end(), 0);
  int sweep = 0;
  for (; sweep < max_sweeps; ++sweep) {
    int rotated = 0;
    for (int round = 0; round < slots - 1; ++round) {
      // pairs (ring[i], ring[slots-1-i])
#pragma omp parallel for num_threads(nthreads) reduction(+ : rotated) schedule(dynamic, 1)
      for (int i = 0; i < slots / 2; ++i) {
        int p = ring[i], q = ring[slots - 1 - i];
        if (p >= cols || q >= cols) continue;
        if (p > q) std::swap(p, q);
        cd* ap = &X[(size_t)p * rows];
        cd* aq = &X[(size_t)q * rows];
        double alpha = 0, beta = 0;
        cd gamma = 0;
        for (int r = 0; r < rows; ++r) {
          alpha += std::norm(ap[r]);
          beta += std::norm(aq[r]);
          gamma += std::conj(ap[r]) * aq[r];
        }
        double g = std::abs(gamma);
        if (g == 0.0 || g <= tol * std::sqrt(alpha * beta)) continue;
        rotated++;
        double zeta = (beta - alpha) / (2.0 * g);
        double t = (zeta >= 0 ? 1.0 : -1.0) / (std::fabs(zeta) + std::sqrt(1.0 + zeta * zeta));
        double c = 1.0 / std::sqrt(1.0 + t * t), s = c * t;
        cd e = gamma / g;  // e^{i phi}
        cd se = s * e, sec = s * std::conj(e);
        for (int r = 0; r < rows; ++r) {
          cd x = ap[r], y = aq[r];
          ap[r] = c * x - sec * y;
          aq[r] = se * x + c * y;
        }
        cd* vp = &V[(size_t)p * cols];
        cd* vq = &V[(size_t)q * cols];
        for (int r = 0; r < cols; ++r) {
          cd x = vp[r], y = vq[r];
          vp[r] = c * x - sec * y;
          vq[r] = se * x + c * y;
        }
      }
      // rotate ring keeping slot 0 fixed
      int last = ring[slots - 1];
      for (int i = slots - 1; i > 1; --i) ring[i] = ring[i - 1];
      ring[1] = last;
    }
    if (rotated == 0) { ++sweep; break; }
  }
  return sweep;
}

// sqrt(S)-split SVD of A (m x n row-major).  L: m x K0 row-major, R: K0 x n row-major.
JacobiResult svd_split(int m, int n, const cd* A, std::vector<cd>& L, std::vector<cd>& R,
                       int nthreads, double tol = 1e-15, int max_sweeps = 60) {
  JacobiResult res;
  const int K0 = std::min(m, n);
  L.assign((size_t)m * K0, 0.0);
  R.assign((size_t)K0 * n, 0.0);
  std::vector<cd> X, V;
  if (m >= n) {
    // columns of A (m x n): column-major copy
    X.resize((size_t)m * n);
    for (int i = 0; i < m; ++i)
      for (int j = 0; j < n; ++j) X[(size_t)j * m + i] = A[(size_t)i * n + j];
    res.sweeps = column_jacobi(m, n, X, V, max_sweeps, tol, nthreads);
    // A V = W, w_k = s_k u_k ; L[:,k] = w_k / sqrt(s_k) ; R[k,:] = sqrt(s_k) V[:,k]^H
    std::vector<double> s(n);
    for (int j = 0; j < n; ++j) {
      double a = 0;
      for (int i = 0; i < m; ++i) a += std::norm(X[(size_t)j * m + i]);
      s[j] = std::sqrt(a);
    }
    std::vector<int> ord(n);
    std::iota(ord.begin(), ord.end(), 0);
    std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return s[a] > s[b]; });
    res.s.resize(K0);
    for (int k = 0; k < K0; ++k) {
      int j = ord[k];
      double sk = s[j];
      res.s[k] = sk;
      double rs = std::sqrt(sk);
      double inv = sk > 0 ? 1.0 / rs : 0.0;
      for (int i = 0; i < m; ++i) L[(size_t)i * K0 + k] = X[(size_t)j * m + i] * inv;
      for (int c = 0; c < n; ++c) R[(size_t)k * n + c] = rs * std::conj(V[(size_t)j * n + c]);
    }
  } else {
    // work on A^H (n x m): columns are conj rows of A
    X.resize((size_t)n * m);
    for (int i = 0; i < m; ++i)
      for (int j = 0; j < n; ++j) X[(size_t)i * n + j] = std::conj(A[(size_t)i * n + j]);
    res.sweeps = column_jacobi(n, m, X, V, max_sweeps, tol, nthreads);
    // A^H V' = W', w'_k = s_k u'_k ; A = V' S U'^H
    // L[:,k] = sqrt(s_k) V'[:,k] ; R[k,:] = sqrt(s_k) u'_k^H = conj(w'_k)^T / sqrt(s_k)
    std::vector<double> s(m);
    for (int j = 0; j < m; ++j) {
      double a = 0;
      for (int i = 0; i < n; ++i) a += std::norm(X[(size_t)j * n + i]);
      s[j] = std::sqrt(a);
    }
    std::vector<int> ord(m);
    std::iota(ord.begin(), ord.end(), 0);
    std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return s[a] > s[b]; });
    res.s.resize(K0);
    for (int k = 0; k < K0; ++k) {
      int j = ord[k];
      double sk = s[j];
      res.s[k] = sk;
      double rs = std::sqrt(sk);
      double inv = sk > 0 ? 1.0 / rs : 0.0;
      for (int i = 0; i < m; ++i) L[(size_t)i * K0 + k] = rs * V[(size_t)j * m + i];
      for (int c = 0; c < n; ++c) R[(size_t)k * n + c] = std::conj(X[(size_t)j * n + c]) * inv;
    }
  }
  return res;
}

// ------------------------------------------------------------------------------------
// MPS restatement
// ------------------------------------------------------------------------------------
struct Mps {
  int n = 0;
  std::vector<int> chi;              // n+1 bonds
  std::vector<std::vector<cd>> A;    // A[k] : chi[k] x 2 x chi[k+1]
  double cutoff = 0;                 // svd-cutoff, default DBL_MIN (ExaTnMpsVisitor.cpp:257-263)
  int max_bond = 0x7ffffffe;         // max-bond-dim, default INT_MAX-1 (:265-271)
  int pn_l2 = 0;                     // 0: partial norms squared (comment :2421-2423), 1: 2-norms
  int nthreads = 1;
  double discarded = 0.0;            // sum of discarded squared singular values (diagnostic)
  int max_chi = 1;
  long long n2q = 0;
};

// |0..0>, all bonds 1, body {1,0}: ExaTnMpsVisitor.cpp:46, 280-326
void mps_init(Mps& s, int n) {
  s.n = n;
  s.chi.assign(n + 1, 1);
  s.A.assign(n, std::vector<cd>(2, 0.0));
  for (int k = 0; k < n; ++k) s.A[k][0] = 1.0;
}

// 1q: new[l][b][r] = sum_i M[b][i] A[l][i][r]   (pattern :1246 with G(i,b) = M[b][i])
void mps_apply_1q(Mps& s, int q, const cd* M) {
  auto& A = s.A[q];
  int cl = s.chi[q], cr = s.chi[q + 1];
  for (int l = 0; l < cl; ++l)
    for (int r = 0; r < cr; ++r) {
      cd a0 = A[((size_t)l * 2 + 0) * cr + r], a1 = A[((size_t)l * 2 + 1) * cr + r];
      A[((size_t)l * 2 + 0) * cr + r] = M[0] * a0 + M[1] * a1;
      A[((size_t)l * 2 + 1) * cr + r] = M[2] * a0 + M[3] * a1;
    }
}

// 2q on adjacent sites (:1387-1627) followed by truncation (:2366-2536)
int mps_apply_2q(Mps& s, int q0, int q1, const cd* M) {
  if (std::abs(q0 - q1) != 1) return -1;  // :1397 neighbours only
  const int i = std::min(q0, q1), j = i + 1;
  const int cl = s.chi[i], cm = s.chi[i + 1], cr = s.chi[j + 1];
  // local gate G[(a,b),(a',b')] with a = left-site bit, b = right-site bit.
  // The matrix row/col index has the FIRST operand as MSB (Gates.hpp:237-246; :1463-1513).
  cd G[16];
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b)
      for (int a2 = 0; a2 < 2; ++a2)
        for (int b2 = 0; b2 < 2; ++b2) {
          int row = (q0 == i) ? (2 * a + b) : (2 * b + a);
          int col = (q0 == i) ? (2 * a2 + b2) : (2 * b2 + a2);
          G[(a * 2 + b) * 4 + (a2 * 2 + b2)] = M[row * 4 + col];
        }
  // merge: theta[l][a][b][r] = sum_m A_i[l][a][m] A_j[m][b][r]   (:1401-1440)
  const int m = 2 * cl, n = 2 * cr;
  std::vector<cd> th((size_t)m * n, 0.0);
  const auto& Ai = s.A[i];
  const auto& Aj = s.A[j];
  for (int l = 0; l < cl; ++l)
    for (int a = 0; a < 2; ++a)
      for (int k = 0; k < cm; ++k) {
        cd x = Ai[((size_t)l * 2 + a) * cm + k];
        if (x == 0.0) continue;
        const cd* rowj = &Aj[(size_t)k * 2 * cr];
        cd* out = &th[((size_t)l * 2 + a) * n];
        for (int c = 0; c < n; ++c) out[c] += x * rowj[c];
      }
  // gate (:1442-1524): theta'[l][a][b][r] = sum G[(a,b),(a',b')] theta[l][a'][b'][r]
  std::vector<cd> tp((size_t)m * n);
  for (int l = 0; l < cl; ++l)
    for (int r = 0; r < cr; ++r) {
      cd v[4];
      for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b) v[a * 2 + b] = th[((size_t)l * 2 + a) * n + b * cr + r];
      for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b) {
          cd acc = 0;
          for (int k = 0; k < 4; ++k) acc += G[(a * 2 + b) * 4 + k] * v[k];
          // stabilizeTensorBody (:141-161): |x| < 1e-100 -> 0
          if (std::abs(acc) < 1e-100) acc = 0.0;
          tp[((size_t)l * 2 + a) * n + b * cr + r] = acc;
        }
    }
  // full SVD, sqrt(S) into both factors; new bond = min(volQ1, volQ2) = min(m, n) (:1594)
  std::vector<cd> L, R;
  JacobiResult jr = svd_split(m, n, tp.data(), L, R, s.nthreads);
  const int K0 = std::min(m, n);
  // truncation: partial norms of each factor along the shared bond (:2424-2429)
  std::vector<double> ln(K0, 0.0), rn(K0, 0.0);
  for (int r = 0; r < m; ++r)
    for (int k = 0; k < K0; ++k) ln[k] += std::norm(L[(size_t)r * K0 + k]);
  for (int k = 0; k < K0; ++k)
    for (int c = 0; c < n; ++c) rn[k] += std::norm(R[(size_t)k * n + c]);
  if (s.pn_l2)
    for (int k = 0; k < K0; ++k) { ln[k] = std::sqrt(ln[k]); rn[k] = std::sqrt(rn[k]); }
  // findCutoffDim (:2434-2443): i+1 for the first i with both below eps
  int K = K0;
  for (int k = 0; k < K0; ++k)
    if (ln[k] < s.cutoff && rn[k] < s.cutoff) { K = k + 1; break; }
  K = std::min(K, s.max_bond);  // :2445
  for (int k = K; k < K0; ++k) s.discarded += jr.s[k] * jr.s[k];
  // slice (:2481-2486); no renormalisation afterwards
  std::vector<cd> Ni((size_t)cl * 2 * K), Nj((size_t)K * 2 * cr);
  for (int r = 0; r < m; ++r)
    for (int k = 0; k < K; ++k) Ni[(size_t)r * K + k] = L[(size_t)r * K0 + k];
  for (int k = 0; k < K; ++k)
    for (int c = 0; c < n; ++c) Nj[(size_t)k * n + c] = R[(size_t)k * n + c];
  s.A[i].swap(Ni);
  s.A[j].swap(Nj);
  s.chi[i + 1] = K;
  s.max_chi = std::max(s.max_chi, K);
  s.n2q++;
  return jr.sweeps;
}

// Left-to-right environment contraction with a per-site diagonal operator d_k (2 entries).
// Returns <psi| (prod_k diag(d_k)) |psi>.
cd mps_diag_expect(const Mps& s, const std::vector<std::array<double, 2>>& ops) {
  std::vector<cd> E(1, 1.0);  // chi_0 x chi_0
  for (int k = 0; k < s.n; ++k) {
    int cl = s.chi[k], cr = s.chi[k + 1];
    const auto& A = s.A[k];
    // X[l][p][r'] = sum_{l'} E[l][l'] A[l'][p][r']
    std::vector<cd> X((size_t)cl * 2 * cr, 0.0);
    for (int l = 0; l < cl; ++l)
      for (int l2 = 0; l2 < cl; ++l2) {
        cd e = E[(size_t)l * cl + l2];
        if (e == 0.0) continue;
        const cd* a = &A[(size_t)l2 * 2 * cr];
        cd* x = &X[(size_t)l * 2 * cr];
        for (int c = 0; c < 2 * cr; ++c) x[c] += e * a[c];
      }
    // E'[r][r'] = sum_{l,p} conj(A[l][p][r]) d_p X[l][p][r']
    std::vector<cd> En((size_t)cr * cr, 0.0);
    for (int l = 0; l < cl; ++l)
      for (int p = 0; p < 2; ++p) {
        double d = ops[k][p];
        if (d == 0.0) continue;
        for (int r = 0; r < cr; ++r) {
          cd a = std::conj(A[((size_t)l * 2 + p) * cr + r]) * d;
          if (a == 0.0) continue;
          const cd* x = &X[((size_t)l * 2 + p) * cr];
          cd* e = &En[(size_t)r * cr];
          for (int r2 = 0; r2 < cr; ++r2) e[r2] += a * x[r2];
        }
      }
    E.swap(En);
  }
  return E[0];
}

// Open-leg contraction of the MPS with a product bra (:2588-2675).  bits[k] in {0,1,-1}.
// Output: 2^{#open} amplitudes, first open qubit fastest.
std::vector<cd> mps_slice(const Mps& s, const int* bits) {
  // T[(x), r] with x enumerating open bits seen so far
  std::vector<cd> T(1, 1.0);
  size_t nx = 1;
  for (int k = 0; k < s.n; ++k) {
    int cl = s.chi[k], cr = s.chi[k + 1];
    const auto& A = s.A[k];
    int b = bits[k];
    size_t nx2 = (b < 0) ? nx * 2 : nx;
    std::vector<cd> Tn(nx2 * cr, 0.0);
    for (int pp = 0; pp < 2; ++pp) {
      if (b >= 0 && pp != b) continue;
      size_t off = (b < 0) ? pp * nx : 0;
      for (size_t x = 0; x < nx; ++x)
        for (int l = 0; l < cl; ++l) {
          cd t = T[x * cl + l];
          if (t == 0.0) continue;
          const cd* a = &A[((size_t)l * 2 + pp) * cr];
          cd* o = &Tn[(off + x) * cr];
          for (int r = 0; r < cr; ++r) o[r] += t * a[r];
        }
    }
    T.swap(Tn);
    nx = nx2;
  }
  return T;  // chi_n = 1
}

}  // namespace

// ======================================================================================
// C ABI
// ======================================================================================
extern "C" {

int orc_gate_matrix(const char* name, const double* params, int nparams, double* out) {
  cd M[16];
  int k = gate_matrix(name, params, nparams, M);
  if (k == 1) store(M, out, 4);
  if (k == 2) store(M, out, 16);
  return k;  // 0 unknown, 1 one-qubit, 2 two-qubit
}

// ---------------- dense state vector (GateMatrixAlgebra.hpp) ----------------
void orc_sv_init(int n, double* psi) {
  size_t N = (size_t)1 << n;
  std::memset(psi, 0, sizeof(double) * 2 * N);
  psi[0] = 1.0;
}

void orc_sv_apply_1q(int n, double* psi_, int q, const double* m_) {
  cd M[4];
  load(m_, M, 4);
  cd* psi = reinterpret_cast<cd*>(psi_);
  size_t N = (size_t)1 << n, k = (size_t)1 << q;
  for (size_t g = 0; g < N; g += 2 * k)
    for (size_t i = g; i < g + k; ++i) {
      cd a = psi[i], b = psi[i + k];
      psi[i] = M[0] * a + M[1] * b;
      psi[i + k] = M[2] * a + M[3] * b;
    }
}

// first operand q0 is the MSB of the 4x4 matrix index
void orc_sv_apply_2q(int n, double* psi_, int q0, int q1, const double* m_) {
  cd M[16];
  load(m_, M, 16);
  cd* psi = reinterpret_cast<cd*>(psi_);
  size_t N = (size_t)1 << n;
  size_t b0 = (size_t)1 << q0, b1 = (size_t)1 << q1;
  for (size_t x = 0; x < N; ++x) {
    if (x & (b0 | b1)) continue;
    size_t idx[4] = {x, x | b1, x | b0, x | b0 | b1};  // (a0 a1) = 00, 01, 10, 11
    cd v[4];
    for (int t = 0; t < 4; ++t) v[t] = psi[idx[t]];
    for (int r = 0; r < 4; ++r) {
      cd acc = 0;
      for (int c = 0; c < 4; ++c) acc += M[r * 4 + c] * v[c];
      psi[idx[r]] = acc;
    }
  }
}

// exp-val-z = sum_x (-1)^{popcount(x & mask)} |psi_x|^2, unnormalised (:616-643)
double orc_sv_expval_z_parity(int n, const double* psi_, const int* qs, int k) {
  const cd* psi = reinterpret_cast<const cd*>(psi_);
  size_t N = (size_t)1 << n;
  uint64_t mask = 0;
  for (int i = 0; i < k; ++i) mask |= (uint64_t)1 << qs[i];
  double acc = 0;
  for (size_t x = 0; x < N; ++x) acc += (__builtin_popcountll(x & mask) & 1 ? -1.0 : 1.0) * std::norm(psi[x]);
  return acc;
}

double orc_sv_norm(int n, const double* psi_) {
  const cd* psi = reinterpret_cast<const cd*>(psi_);
  size_t N = (size_t)1 << n;
  double acc = 0;
  for (size_t x = 0; x < N; ++x) acc += std::norm(psi[x]);
  return acc;
}

// GenerateSamples (GateMatrixAlgebra.hpp:125-156) with randomEngine (RandomEngine.hpp):
// mt19937_64 seeded with `seed`, uniform_real_distribution(0,1), sort, CDF walk.
// out: shots * k chars, char i <-> meas[i].
int orc_sv_sample(int n, const double* psi_, uint64_t seed, int shots, const int* meas, int k,
                  char* out) {
  const cd* psi = reinterpret_cast<const cd*>(psi_);
  std::mt19937_64 eng;
  eng.seed(seed);
  std::vector<double> rs;
  rs.reserve(shots);
  for (int i = 0; i < shots; ++i) rs.push_back(std::uniform_real_distribution<double>(0.0, 1.0)(eng));
  std::sort(rs.begin(), rs.end());
  size_t N = (size_t)1 << n;
  double csum = 0;
  int m = 0;
  for (size_t x = 0; x < N && m < shots; ++x) {
    csum += std::norm(psi[x]);
    while (m < shots && rs[m] < csum) {
      for (int i = 0; i < k; ++i) out[(size_t)m * k + i] = (x >> meas[i]) & 1 ? '1' : '0';
      ++m;
    }
  }
  return m;
}

// ---------------- SVD (sqrt(S) split) ----------------
// A: m x n interleaved complex row-major; s: min(m,n); L: m x K0; R: K0 x n.
int orc_svd_split(int m, int n, const double* a, double* s, double* L, double* R, int nthreads) {
  std::vector<cd> A((size_t)m * n);
  load(a, A.data(), A.size());
  std::vector<cd> Lv, Rv;
  JacobiResult jr = svd_split(m, n, A.data(), Lv, Rv, nthreads);
  int K0 = std::min(m, n);
  for (int k = 0; k < K0; ++k) s[k] = jr.s[k];
  if (L) store(Lv.data(), L, Lv.size());
  if (R) store(Rv.data(), R, Rv.size());
  return jr.sweeps;
}

// ---------------- MPS ----------------
void* orc_mps_create(int n, double cutoff, int max_bond, int pn_l2, int nthreads) {
  Mps* s = new Mps();
  mps_init(*s, n);
  s->cutoff = cutoff;
  s->max_bond = max_bond > 0 ? max_bond : 0x7ffffffe;
  s->pn_l2 = pn_l2;
  s->nthreads = nthreads > 0 ? nthreads : 1;
  return s;
}
void orc_mps_destroy(void* h) { delete static_cast<Mps*>(h); }

int orc_mps_apply_1q(void* h, int q, const double* m) {
  cd M[4];
  load(m, M, 4);
  mps_apply_1q(*static_cast<Mps*>(h), q, M);
  return 0;
}
int orc_mps_apply_2q(void* h, int q0, int q1, const double* m) {
  cd M[16];
  load(m, M, 16);
  return mps_apply_2q(*static_cast<Mps*>(h), q0, q1, M);
}
void orc_mps_bonds(void* h, int* out) {
  Mps* s = static_cast<Mps*>(h);
  for (int k = 0; k <= s->n; ++k) out[k] = s->chi[k];
}
double orc_mps_discarded(void* h) { return static_cast<Mps*>(h)->discarded; }

// site tensor access (for uploading a known state to the GPU in tests)
size_t orc_mps_site_size(void* h, int k) {
  Mps* s = static_cast<Mps*>(h);
  return s->A[k].size();
}
void orc_mps_get_site(void* h, int k, double* out) {
  Mps* s = static_cast<Mps*>(h);
  store(s->A[k].data(), out, s->A[k].size());
}
int orc_mps_set_site(void* h, int k, int chil, int chir, const double* in) {
  Mps* s = static_cast<Mps*>(h);
  if (s->chi[k] != chil || s->chi[k + 1] != chir) {
    // caller must keep neighbouring sites consistent
    s->chi[k] = chil;
    s->chi[k + 1] = chir;
  }
  s->A[k].assign((size_t)chil * 2 * chir, 0.0);
  load(in, s->A[k].data(), s->A[k].size());
  return 0;
}

// full contraction to a 2^n vector (qubit k <-> bit k), n <= 26
int orc_mps_state_vector(void* h, double* out) {
  Mps* s = static_cast<Mps*>(h);
  if (s->n > 26) return -1;
  std::vector<int> bits(s->n, -1);
  auto v = mps_slice(*s, bits.data());
  store(v.data(), out, v.size());
  return 0;
}

double orc_mps_expval_z_parity(void* h, const int* qs, int k) {
  Mps* s = static_cast<Mps*>(h);
  std::vector<std::array<double, 2>> ops(s->n, {1.0, 1.0});
  for (int i = 0; i < k; ++i) ops[qs[i]][1] *= -1.0;
  return mps_diag_expect(*s, ops).real();
}
double orc_mps_norm(void* h) {
  Mps* s = static_cast<Mps*>(h);
  std::vector<std::array<double, 2>> ops(s->n, {1.0, 1.0});
  return mps_diag_expect(*s, ops).real();
}

// getMeasureSample (ExaTnMpsVisitor.cpp:2211-2364), `shots` times in a row (the n >= 20 path of
// finalize, :651-670), draws from one mt19937_64 seeded with `seed` through
// uniform_real_distribution(0, 1) as randomEngine::randProb (RandomEngine.hpp:11-14), one draw
// per measured qubit per shot in Measure order.  Per qubit: the diagonal of its reduced density
// matrix with the earlier outcomes applied as collapse projectors scaled by 1/p of their outcome
// (:2229-2276) -- the first qubit's probabilities are therefore those of the unnormalised state;
// values below 1e-12 in magnitude count as 0 (:2329-2337); outcome 0 when the draw is <= p0
// (:2342-2345).  out: shots * k chars, char i <-> meas[i].
int orc_mps_measure_samples(void* h, uint64_t seed, int shots, const int* meas, int k, char* out) {
  Mps* s = static_cast<Mps*>(h);
  std::mt19937_64 eng;
  eng.seed(seed);
  constexpr double PROB_EPS = 1e-12;
  for (int t = 0; t < shots; ++t) {
    std::vector<std::array<double, 2>> ops(s->n, {1.0, 1.0});
    for (int i = 0; i < k; ++i) {
      const int q = meas[i];
      const std::array<double, 2> keep = ops[q];
      double prob[2];
      for (int b = 0; b < 2; ++b) {
        ops[q] = {b == 0 ? keep[0] : 0.0, b == 1 ? keep[1] : 0.0};
        const double v = mps_diag_expect(*s, ops).real();
        prob[b] = std::fabs(v) < PROB_EPS ? 0.0 : v;
      }
      const double r = std::uniform_real_distribution<double>(0.0, 1.0)(eng);
      const int bit = r <= prob[0] ? 0 : 1;
      const double p = prob[bit];
      // collapse onto the outcome, the ket scaled by 1/p (the bra is not projected: for the
      // diagonal entries used here one projector gives the same value)
      ops[q] = {bit == 0 ? keep[0] / p : 0.0, bit == 1 ? keep[1] / p : 0.0};
      out[(size_t)t * k + i] = bit ? '1' : '0';
    }
  }
  return shots;
}

// amplitudes of k full bit strings (bits: k*n int8, 0/1)
void orc_mps_amplitudes(void* h, const int8_t* bits, int k, double* out) {
  Mps* s = static_cast<Mps*>(h);
  std::vector<int> b(s->n);
  for (int t = 0; t < k; ++t) {
    for (int i = 0; i < s->n; ++i) b[i] = bits[(size_t)t * s->n + i];
    auto v = mps_slice(*s, b.data());
    out[2 * t] = v[0].real();
    out[2 * t + 1] = v[0].imag();
  }
}

// wave-function slice for bits in {0,1,-1}; returns 2^{#open} values (first open fastest)
int orc_mps_slice(void* h, const int* bits, double* out) {
  Mps* s = static_cast<Mps*>(h);
  auto v = mps_slice(*s, bits);
  store(v.data(), out, v.size());
  return (int)v.size();
}

}  // extern "C"
