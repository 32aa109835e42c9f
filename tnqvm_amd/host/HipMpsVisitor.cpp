// HipMpsVisitor.cpp -- "hip-mps" visitor over libhipmps.
#include "HipMpsVisitor.hpp"

#include <cfloat>
#include <climits>
#include <cmath>
#include <map>
#include <mutex>

#include "RandomEngine.hpp"

namespace tnqvm {

namespace {
// The reference contracts the full state vector below this size (finalize,
// ExaTnMpsVisitor.cpp:55, :588) and samples shot by shot above it (:651-670).
constexpr int MAX_QUBITS_FOR_STATE_VEC = 20;

const double SWAP_M[32] = {1, 0, 0, 0, 0, 0, 0, 0,  //
                           0, 0, 0, 0, 1, 0, 0, 0,  //
                           0, 0, 1, 0, 0, 0, 0, 0,  //
                           0, 0, 0, 0, 0, 0, 1, 0};
}  // namespace

hipmps_engine HipMpsVisitor::engine(int device) {
  static std::mutex mu;
  static std::map<int, hipmps_engine> engines;
  std::lock_guard<std::mutex> lk(mu);
  auto it = engines.find(device);
  if (it != engines.end()) return it->second;
  hipmps_engine_opts o{};
  o.device = device;
  hipmps_engine e = nullptr;
  int rc = hipmps_engine_create(&o, &e);
  if (rc != 0) xacc::error(std::string("hip-mps: engine creation failed: ") + hipmps_last_error());
  engines[device] = e;
  return e;
}

HipMpsVisitor::~HipMpsVisitor() {
  if (m_state) hipmps_state_destroy(m_state);
}

void HipMpsVisitor::check(int rc, const char* what) const {
  if (rc != 0) xacc::error(std::string("hip-mps: ") + what + " failed (" + std::to_string(rc) + "): " + hipmps_last_error());
}

hipmps_trunc_opts HipMpsVisitor::trunc_options(double default_cutoff) const {
  hipmps_trunc_opts t{};
  t.svd_cutoff = default_cutoff;
  t.max_bond_dim = 0;  // INT_MAX-1, ExaTnMpsVisitor.cpp:265-271
  if (options.keyExists<double>("svd-cutoff")) t.svd_cutoff = options.get<double>("svd-cutoff");
  if (options.keyExists<int>("max-bond-dim")) t.max_bond_dim = options.get<int>("max-bond-dim");
  if (options.stringExists("precision")) {
    const auto p = options.getString("precision");
    if (p == "mixed") t.precision = HIPMPS_MIXED;
    else if (p == "fp64") t.precision = HIPMPS_FP64;
    else xacc::error("hip-mps: precision must be 'fp64' or 'mixed'");
  }
  if (options.keyExists<bool>("renormalize")) t.renormalize = options.get<bool>("renormalize") ? 1 : 0;
  if (options.stringExists("partial-norm")) t.partial_norm = options.getString("partial-norm") == "l2" ? HIPMPS_PN_L2 : HIPMPS_PN_SQUARED;
  if (options.stringExists("absorb")) t.absorb = options.getString("absorb") == "right" ? HIPMPS_ABSORB_RIGHT : HIPMPS_ABSORB_SQRT;
  return t;
}

void HipMpsVisitor::initialize(std::shared_ptr<AcceleratorBuffer> in_buffer, int nbShots) {
  buffer = std::move(in_buffer);
  m_n = buffer->size();
  m_shots = nbShots;
  m_measured.clear();
  hipmps_trunc_opts t = trunc_options(DBL_MIN);  // ExaTnMpsVisitor.cpp:257-263
  const int device = options.keyExists<int>("device") ? options.get<int>("device") : 0;
  if (m_state) {
    hipmps_state_destroy(m_state);
    m_state = nullptr;
  }
  check(hipmps_state_create(run_engine(device), m_n, &t, &m_state), "state creation");
}

// The engine is shared by every run on the device; its Jacobi sweep cap is set from this run's
// "jacobi-max-sweeps" option (default 40).
hipmps_engine HipMpsVisitor::run_engine(int device) const {
  hipmps_engine e = engine(device);
  hipmps_tuning tu{};
  check(hipmps_engine_get_tuning(e, &tu), "tuning");
  tu.jacobi_max_sweeps = options.keyExists<int>("jacobi-max-sweeps") ? options.get<int>("jacobi-max-sweeps") : 40;
  check(hipmps_engine_set_tuning(e, &tu), "jacobi-max-sweeps");
  return e;
}

void HipMpsVisitor::apply(xacc::Instruction& g) {
  std::vector<double> p;
  for (auto& x : g.getParameters()) p.push_back(x.as<double>());
  double m[32];
  const int k = hipmps_gate_matrix(g.name().c_str(), p.empty() ? nullptr : p.data(), (int)p.size(), m);
  if (k < 0) xacc::error("hip-mps: no matrix for gate " + g.name());
  apply_matrix(g.bits(), m);
  after_gate(g);
}

// Non-neighbouring 2-qubit gates are routed with the meet-in-the-middle swap scheme of
// NearestNeighborTransform (tnqvm/visitors/exatn-mps/NearestNeighborTransform.hpp:77-121):
// TNQVM applies XACC's nnizer only to the visitor names it knows (TNQVM.cpp:119-124).
void HipMpsVisitor::apply_matrix(const std::vector<size_t>& bits, const double* m) {
  if (bits.size() == 1) {
    check(hipmps_apply_1q(m_state, (int)bits[0], m), "1-qubit gate");
    return;
  }
  const int q0 = (int)bits[0], q1 = (int)bits[1];
  if (std::abs(q0 - q1) == 1) {
    check(hipmps_apply_2q(m_state, q0, q1, m), "2-qubit gate");
    return;
  }
  const int lower0 = std::min(q0, q1), upper0 = std::max(q0, q1);
  int lo = lower0, up = upper0;
  for (;;) {
    check(hipmps_apply_2q(m_state, lo, lo + 1, SWAP_M), "routing swap");
    ++lo;
    if (up - lo <= 1) break;
    check(hipmps_apply_2q(m_state, up, up - 1, SWAP_M), "routing swap");
    --up;
    if (up - lo <= 1) break;
  }
  check(hipmps_apply_2q(m_state, q0 < q1 ? lo : up, q0 < q1 ? up : lo, m), "routed 2-qubit gate");
  for (int i = lo; i > lower0; --i) check(hipmps_apply_2q(m_state, i, i - 1, SWAP_M), "routing swap");
  for (int i = up; i < upper0; ++i) check(hipmps_apply_2q(m_state, i, i + 1, SWAP_M), "routing swap");
}

// GenerateSamples (tnqvm/utils/GateMatrixAlgebra.hpp:125-156): sorted uniforms walked
// along the cumulative distribution of |psi_x|^2; char i <-> i-th measured qubit.
std::vector<std::string> HipMpsVisitor::sample_state_vector(int shots) {
  std::vector<double> sv(2 * ((size_t)1 << m_n));
  check(hipmps_state_vector(m_state, sv.data(), (size_t)1 << m_n), "state vector");
  const auto rs = randomEngine::get_instance().sortedRandProbs(shots);
  std::vector<std::string> out;
  out.reserve(shots);
  double csum = 0.0;
  size_t m = 0;
  const size_t N = (size_t)1 << m_n;
  for (size_t x = 0; x < N && m < (size_t)shots; ++x) {
    csum += sv[2 * x] * sv[2 * x] + sv[2 * x + 1] * sv[2 * x + 1];
    while (m < (size_t)shots && rs[m] < csum) {
      std::string s(m_measured.size(), '0');
      for (size_t i = 0; i < m_measured.size(); ++i)
        if ((x >> m_measured[i]) & 1) s[i] = '1';
      out.push_back(s);
      ++m;
    }
  }
  return out;
}

// getMeasureSample (ExaTnMpsVisitor.cpp:2211-2364): qubit by qubit in Measure order, the
// diagonal of the qubit's reduced density matrix with the earlier outcomes applied as collapse
// projectors scaled by 1/p of their outcome (:2229-2276), so the first qubit sees the
// probabilities of the unnormalised state and every later one its conditional probabilities;
// values below 1e-12 in magnitude count as 0 (:2329-2337); outcome 0 when the draw is <= p0
// (:2342-2345).  One draw per measured qubit per shot, shots in a row.
std::vector<std::string> HipMpsVisitor::sample_sequential(int shots) {
  std::vector<std::string> out;
  const int k = (int)m_measured.size();
  bool increasing = true;
  for (int i = 1; i < k; ++i) increasing = increasing && m_measured[i] > m_measured[i - 1];
  if (increasing && k > 0) {  // every shot's marginals advanced together on the device
    std::vector<double> draws((size_t)shots * k);
    for (auto& r : draws) r = randomEngine::get_instance().randProb();
    std::vector<int> q(m_measured.begin(), m_measured.end());
    std::vector<char> bits((size_t)shots * k + 1);
    check(hipmps_sample_sequential(m_state, q.data(), k, shots, draws.data(), bits.data()), "sampling");
    for (int s = 0; s < shots; ++s) out.emplace_back(bits.data() + (size_t)s * k, (size_t)k);
    return out;
  }
  std::vector<double> diag(2 * m_n);
  for (int s = 0; s < shots; ++s) {
    std::fill(diag.begin(), diag.end(), 1.0);
    std::string bits;
    for (int q : m_measured) {
      const double k0 = diag[2 * q], k1 = diag[2 * q + 1];
      double prob[2];
      for (int b = 0; b < 2; ++b) {
        diag[2 * q] = b == 0 ? k0 : 0.0;
        diag[2 * q + 1] = b == 1 ? k1 : 0.0;
        check(hipmps_expval_diag(m_state, diag.data(), &prob[b]), "marginal");
        if (std::fabs(prob[b]) < 1e-12) prob[b] = 0.0;
      }
      const int b = randomEngine::get_instance().randProb() <= prob[0] ? 0 : 1;
      bits.push_back(b ? '1' : '0');
      // collapse onto outcome b, scaled by 1/p of that outcome
      const double sc = prob[b] > 0.0 ? 1.0 / prob[b] : 1.0;
      diag[2 * q] = b == 0 ? k0 * sc : 0.0;
      diag[2 * q + 1] = b == 1 ? k1 * sc : 0.0;
    }
    out.push_back(bits);
  }
  return out;
}

void HipMpsVisitor::finalize() {
  int rc = hipmps_flush(m_state);
  if (rc != 0 && rc != HIPMPS_E_CONVERGENCE) check(rc, "flush");
  double norm = 0.0;
  check(hipmps_norm(m_state, &norm), "norm");
  buffer->addExtraInfo("norm", norm);
  hipmps_stats_t st{};
  hipmps_stats(m_state, &st);
  // an SVD that hit the Jacobi sweep cap: results are still written, flagged on the buffer
  buffer->addExtraInfo("svd-converged", st.svds_nonconverged == 0 ? 1 : 0);
  buffer->addExtraInfo("svd-nonconverged-count", st.svds_nonconverged);
  if (st.svds_nonconverged > 0)
    std::cerr << "[WARNING] " << st.svds_nonconverged
              << " SVD(s) reached the Jacobi sweep cap without converging; results may be inaccurate\n";
  buffer->addExtraInfo("discarded-weight", st.discarded_weight);
  buffer->addExtraInfo("truncation-fidelity", st.fidelity_est);
  buffer->addExtraInfo("max-bond-dim-reached", st.max_bond);
  buffer->addExtraInfo("norm-validation-failures", st.norm_warnings);
  if (st.norm_warnings > 0)
    std::cerr << "[ERROR] Tensor norm validation failed after " << st.norm_warnings
              << " SVD(s): a factor's Frobenius norm fell below 1e-3\n";

  if (options.keyExists<bool>("exp-val-z-per-qubit") && options.get<bool>("exp-val-z-per-qubit")) {
    std::vector<double> z(m_n);
    check(hipmps_expval_z_all(m_state, z.data()), "per-qubit <Z>");
    buffer->addExtraInfo("exp-val-z-per-qubit", z);
  }
  if (options.keyExists<std::vector<int>>("zz-pairs")) {
    const auto p = options.get<std::vector<int>>("zz-pairs");
    std::vector<double> zz(p.size() / 2);
    if (!zz.empty()) check(hipmps_expval_zz(m_state, p.data(), (int)zz.size(), zz.data()), "<ZZ>");
    buffer->addExtraInfo("exp-val-zz", zz);
  }
  // amplitude or wave-function slice (ExaTnMpsVisitor.cpp:776-822)
  if (options.keyExists<std::vector<int>>("bitstring")) {
    const auto bits = options.get<std::vector<int>>("bitstring");
    if ((int)bits.size() != m_n) xacc::error("Bitstring size must match the number of qubits.");
    int open = 0;
    for (int b : bits) open += b < 0;
    std::vector<double> v(2 * ((size_t)1 << open));
    check(hipmps_wavefunction_slice(m_state, bits.data(), v.data(), (size_t)1 << open), "wave-function slice");
    if (open == 0) {
      buffer->addExtraInfo("amplitude-real", v[0]);
      buffer->addExtraInfo("amplitude-imag", v[1]);
    } else {
      double nrm = 0.0;
      for (size_t i = 0; i < v.size(); i += 2) nrm += v[i] * v[i] + v[i + 1] * v[i + 1];
      const double scale = nrm > 1e-12 ? 1.0 / std::sqrt(nrm) : 1.0;
      std::vector<double> re, im;
      for (size_t i = 0; i < v.size(); i += 2) {
        re.push_back(v[i] * scale);
        im.push_back(v[i + 1] * scale);
      }
      buffer->addExtraInfo("amplitude-real-vec", re);
      buffer->addExtraInfo("amplitude-imag-vec", im);
    }
  }
  if (!m_measured.empty()) {
    if (m_shots < 1) {
      // exp-val-z: parity over the measured qubits, not divided by the norm (:616-643)
      double e = 0.0;
      check(hipmps_expval_z_parity(m_state, m_measured.data(), (int)m_measured.size(), &e), "exp-val-z");
      buffer->addExtraInfo("exp-val-z", e);
    } else {
      const auto samples = m_n < MAX_QUBITS_FOR_STATE_VEC ? sample_state_vector(m_shots) : sample_sequential(m_shots);
      for (auto& s : samples) buffer->appendMeasurement(s);
    }
  }
}

// Exact <Z..Z> of the observable sub-circuit applied after the current state (the
// reference estimates it from 100000 shots, ExaTnMpsVisitor.cpp:1087-1117).  Basis
// changes made of 1-qubit gates are applied and undone in place; anything else runs on
// a copy of the state.
const double HipMpsVisitor::getExpectationValueZ(std::shared_ptr<CompositeInstruction> fn) {
  std::vector<std::shared_ptr<xacc::Instruction>> gates;
  std::vector<int> meas;
  bool only1q = true;
  xacc::InstructionIterator it(fn);
  while (it.hasNext()) {
    auto inst = it.next();
    if (!inst->isEnabled()) continue;
    if (inst->name() == "Measure") {
      meas.push_back((int)inst->bits()[0]);
      continue;
    }
    if (inst->name() == "I") continue;
    if (inst->bits().size() != 1) only1q = false;
    gates.push_back(inst);
  }
  if (meas.empty()) return 0.0;
  auto matrix = [](xacc::Instruction& g, double* m) {
    std::vector<double> p;
    for (auto& x : g.getParameters()) p.push_back(x.as<double>());
    if (hipmps_gate_matrix(g.name().c_str(), p.empty() ? nullptr : p.data(), (int)p.size(), m) < 0)
      xacc::error("hip-mps: no matrix for gate " + g.name());
  };
  double e = 0.0;
  if (only1q) {
    for (auto& g : gates) apply(*g);
    check(hipmps_expval_z_parity(m_state, meas.data(), (int)meas.size(), &e), "exp-val-z");
    for (auto r = gates.rbegin(); r != gates.rend(); ++r) {
      double m[32], mh[8];
      matrix(**r, m);
      // conjugate transpose of the 2x2
      for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) {
          mh[2 * (i * 2 + j)] = m[2 * (j * 2 + i)];
          mh[2 * (i * 2 + j) + 1] = -m[2 * (j * 2 + i) + 1];
        }
      check(hipmps_apply_1q(m_state, (int)(*r)->bits()[0], mh), "basis change undo");
    }
    return e;
  }
  hipmps_state saved = m_state;
  hipmps_state copy = nullptr;
  check(hipmps_state_clone(saved, &copy), "state clone");
  m_state = copy;
  try {
    for (auto& g : gates) apply(*g);
    check(hipmps_expval_z_parity(m_state, meas.data(), (int)meas.size(), &e), "exp-val-z");
  } catch (...) {
    m_state = saved;
    hipmps_state_destroy(copy);
    throw;
  }
  m_state = saved;
  hipmps_state_destroy(copy);
  return e;
}

const std::vector<std::complex<double>> HipMpsVisitor::getState() {
  if (!m_state || m_n > 26) return {};
  std::vector<double> sv(2 * ((size_t)1 << m_n));
  check(hipmps_state_vector(m_state, sv.data(), (size_t)1 << m_n), "state vector");
  std::vector<std::complex<double>> out((size_t)1 << m_n);
  for (size_t i = 0; i < out.size(); ++i) out[i] = {sv[2 * i], sv[2 * i + 1]};
  return out;
}

}  // namespace tnqvm
