// Exercises the CPU oracle (oracle/mps_oracle.cpp, compiled into this binary) under AddressSanitizer
// and UndefinedBehaviorSanitizer: dense state vector and sampling, the Jacobi SVD split, the MPS with
// and without truncation, amplitudes, slices and the sequential sampler.  Exit status 0 when every
// check holds and the sanitizers reported nothing.
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstdio>
#include <random>
#include <vector>

extern "C" {
int orc_gate_matrix(const char* name, const double* params, int nparams, double* out);
void orc_sv_init(int n, double* psi);
void orc_sv_apply_1q(int n, double* psi, int q, const double* m);
void orc_sv_apply_2q(int n, double* psi, int q0, int q1, const double* m);
double orc_sv_norm(int n, const double* psi);
int orc_sv_sample(int n, const double* psi, uint64_t seed, int shots, const int* meas, int k, char* out);
int orc_svd_split(int m, int n, const double* a, double* s, double* L, double* R, int nthreads);
void* orc_mps_create(int n, double cutoff, int max_bond, int pn_l2, int nthreads);
void orc_mps_destroy(void* h);
int orc_mps_apply_1q(void* h, int q, const double* m);
int orc_mps_apply_2q(void* h, int q0, int q1, const double* m);
void orc_mps_bonds(void* h, int* out);
int orc_mps_state_vector(void* h, double* out);
double orc_mps_norm(void* h);
void orc_mps_amplitudes(void* h, const int8_t* bits, int k, double* out);
int orc_mps_slice(void* h, const int* bits, double* out);
int orc_mps_measure_samples(void* h, uint64_t seed, int shots, const int* meas, int k, char* out);
}

static int fails = 0;
#define CHECK(c) do { if (!(c)) { std::printf("FAIL %s:%d %s\n", __FILE__, __LINE__, #c); ++fails; } } while (0)

int main() {
  const int n = 8;
  std::mt19937_64 rng(3);
  std::uniform_real_distribution<double> ang(0.0, 6.28);
  std::vector<double> psi(2 << n);
  orc_sv_init(n, psi.data());
  void* exact = orc_mps_create(n, 0.0, 0, 0, 1);
  void* cut = orc_mps_create(n, 1e-3, 4, 0, 2);
  double m[32];
  for (int layer = 0; layer < 6; ++layer) {
    for (int q = 0; q < n; ++q) {
      const double p[3] = {ang(rng), ang(rng), ang(rng)};
      CHECK(orc_gate_matrix("U", p, 3, m) == 1);
      orc_sv_apply_1q(n, psi.data(), q, m);
      orc_mps_apply_1q(exact, q, m);
      orc_mps_apply_1q(cut, q, m);
    }
    for (int q = layer % 2; q + 1 < n; q += 2) {
      CHECK(orc_gate_matrix("CNOT", nullptr, 0, m) == 2);
      const int a = layer % 3 ? q : q + 1, b = layer % 3 ? q + 1 : q;
      orc_sv_apply_2q(n, psi.data(), a, b, m);
      orc_mps_apply_2q(exact, a, b, m);
      orc_mps_apply_2q(cut, a, b, m);
    }
  }
  std::vector<double> sv(2 << n);
  CHECK(orc_mps_state_vector(exact, sv.data()) == 0);
  double err = 0.0;
  for (size_t i = 0; i < sv.size(); ++i) err = std::fmax(err, std::fabs(sv[i] - psi[i]));
  CHECK(err < 1e-12);
  CHECK(std::fabs(orc_sv_norm(n, psi.data()) - 1.0) < 1e-12);
  std::vector<int> bonds(n + 1);
  orc_mps_bonds(cut, bonds.data());
  for (int b : bonds) CHECK(b <= 4);
  CHECK(orc_mps_norm(cut) < 1.0);
  const int meas[4] = {6, 1, 3, 0};
  std::vector<char> a(400 * 4), b(400 * 4);
  CHECK(orc_sv_sample(n, psi.data(), 11, 400, meas, 4, a.data()) == 400);
  CHECK(orc_mps_measure_samples(exact, 11, 400, meas, 4, b.data()) == 400);
  std::vector<int8_t> bits(5 * n);
  for (auto& x : bits) x = (int8_t)(rng() & 1);
  std::vector<double> amp(10);
  orc_mps_amplitudes(exact, bits.data(), 5, amp.data());
  std::vector<int> sl(n, -1);
  sl[2] = 1;
  sl[5] = 0;
  std::vector<double> slice(2 << (n - 2));
  CHECK(orc_mps_slice(exact, sl.data(), slice.data()) == (1 << (n - 2)));
  // SVD split of a rank-deficient matrix: reconstruction
  const int r = 12, c = 9;
  std::vector<double> A(2 * r * c, 0.0), S(std::min(r, c)), L(2 * r * std::min(r, c)), R(2 * std::min(r, c) * c);
  for (int i = 0; i < r; ++i)
    for (int j = 0; j < c; ++j) A[2 * (i * c + j)] = (i % 3 + 1) * (j + 1);  // rank 1
  orc_svd_split(r, c, A.data(), S.data(), L.data(), R.data(), 1);
  double rec = 0.0;
  for (int i = 0; i < r; ++i)
    for (int j = 0; j < c; ++j) {
      std::complex<double> v = 0.0;
      for (int k = 0; k < std::min(r, c); ++k)
        v += std::complex<double>(L[2 * (i * std::min(r, c) + k)], L[2 * (i * std::min(r, c) + k) + 1]) *
             std::complex<double>(R[2 * (k * c + j)], R[2 * (k * c + j) + 1]);
      rec = std::fmax(rec, std::abs(v - std::complex<double>(A[2 * (i * c + j)], 0.0)));
    }
  CHECK(rec < 1e-10);
  orc_mps_destroy(exact);
  orc_mps_destroy(cut);
  std::printf("oracle driver: %d failures\n", fails);
  return fails ? 1 : 0;
}
