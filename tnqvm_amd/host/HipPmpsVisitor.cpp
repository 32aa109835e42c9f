// HipPmpsVisitor.cpp -- "hip-pmps" visitor over libhipmps.
#include "HipPmpsVisitor.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>

#include "RandomEngine.hpp"

namespace tnqvm {

namespace {
constexpr int MAX_QUBITS_FOR_MEASURE = 15;  // ExaTnPmpsVisitor.cpp:658
constexpr int DEFAULT_NOISY_SHOTS = 1024;   // :490
constexpr double PMPS_CUTOFF = 1e-9;        // truncateSvdTensors' default eps, :1255
}  // namespace

void HipPmpsVisitor::initialize(std::shared_ptr<AcceleratorBuffer> in_buffer, int nbShots) {
  buffer = std::move(in_buffer);
  m_n = buffer->size();
  m_shots = nbShots < 1 ? DEFAULT_NOISY_SHOTS : nbShots;
  m_measured.clear();
  m_noise = options.pointerLikeExists<xacc::NoiseModel>("noise-model")
                ? options.getPointerLike<xacc::NoiseModel>("noise-model")
                : nullptr;
  hipmps_trunc_opts t = trunc_options(PMPS_CUTOFF);
  const int device = options.keyExists<int>("device") ? options.get<int>("device") : 0;
  if (m_state) {
    hipmps_state_destroy(m_state);
    m_state = nullptr;
  }
  check(hipmps_state_create(engine(device), m_n, &t, &m_state), "state creation");
}

// applySingleQubitGate / applyTwoQubitGate tails (ExaTnPmpsVisitor.cpp:798-808, :836-846):
// every single-qubit channel the model gives for the gate, in its order
void HipPmpsVisitor::after_gate(xacc::Instruction& g) {
  if (!m_noise) return;
  auto* gate = dynamic_cast<xacc::quantum::Gate*>(&g);
  if (!gate) return;
  static bool warned = false;
  for (const auto& ch : m_noise->getNoiseChannels(*gate)) {
    if (ch.noise_qubits.size() != 1) {
      if (!warned) {
        std::fprintf(stderr, "Multi-qubit channels are not supported. Ignoring.\n");
        warned = true;
      }
      continue;
    }
    std::vector<double> m;
    for (const auto& k : ch.mats) {
      if (k.size() != 2 || k[0].size() != 2 || k[1].size() != 2) xacc::error("hip-pmps: Kraus operators must be 2 x 2");
      for (const auto& row : k)
        for (const auto& v : row) {
          m.push_back(v.real());
          m.push_back(v.imag());
        }
    }
    check(hipmps_apply_kraus(m_state, (int)ch.noise_qubits[0], (int)ch.mats.size(), m.data()), "noise channel");
  }
}

// finalize (ExaTnPmpsVisitor.cpp:657-782): the outcome distribution of the measured qubits
// (the diagonal of their reduced density matrix), then the shots drawn from it with the
// model's readout errors
void HipPmpsVisitor::finalize() {
  int rc = hipmps_flush(m_state);
  if (rc != 0 && rc != HIPMPS_E_CONVERGENCE) check(rc, "flush");
  double tr = 0.0;
  check(hipmps_norm(m_state, &tr), "trace");
  buffer->addExtraInfo("norm", tr);
  std::vector<int> kd(m_n);
  check(hipmps_kraus_dims(m_state, kd.data()), "Kraus dims");
  buffer->addExtraInfo("kraus-dims", kd);
  if (m_measured.empty()) return;
  const int k = (int)m_measured.size();
  if (k > MAX_QUBITS_FOR_MEASURE)
    xacc::error("Measure more than " + std::to_string(MAX_QUBITS_FOR_MEASURE) + " is not supported.");
  std::vector<double> probs((size_t)1 << k), diag(2 * m_n);
  double ez = 0.0;
  for (size_t x = 0; x < probs.size(); ++x) {
    std::fill(diag.begin(), diag.end(), 1.0);
    int parity = 0;
    for (int i = 0; i < k; ++i) {
      const int b = (int)((x >> i) & 1);
      parity ^= b;
      const int q = m_measured[i];
      // several Measure on one qubit: the outcome bits must agree
      diag[2 * q + (1 - b)] = 0.0;
    }
    check(hipmps_expval_diag(m_state, diag.data(), &probs[x]), "outcome probability");
    ez += parity ? -probs[x] : probs[x];
  }
  buffer->addExtraInfo("density-matrix-diagonal", probs);
  std::vector<int> sorted(m_measured);
  std::sort(sorted.begin(), sorted.end());
  const bool distinct = std::adjacent_find(sorted.begin(), sorted.end()) == sorted.end();
  if (k <= 8 && distinct) {  // the reduced density matrix, flattened ket index fastest (:684-757)
    const size_t dim = (size_t)1 << k;
    std::vector<double> rho(2 * dim * dim);
    check(hipmps_reduced_density_matrix(m_state, m_measured.data(), k, rho.data(), dim * dim), "density matrix");
    std::vector<std::pair<double, double>> flat(dim * dim);
    for (size_t t = 0; t < flat.size(); ++t) flat[t] = {rho[2 * t], rho[2 * t + 1]};
    buffer->addExtraInfo("density_matrix", flat);
  }
  buffer->addExtraInfo("exp-val-z", ez);
  double tot = 0.0;
  for (double p : probs) tot += std::max(p, 0.0);
  auto& rng = randomEngine::get_instance();
  for (int s = 0; s < m_shots; ++s) {
    const double u = rng.randProb() * tot;
    double acc = 0.0;
    size_t x = 0;
    for (; x + 1 < probs.size(); ++x) {
      acc += std::max(probs[x], 0.0);
      if (u < acc) break;
    }
    std::string bits;
    for (int i = 0; i < k; ++i) {
      int b = (int)((x >> i) & 1);
      if (m_noise) {
        const auto ro = m_noise->readoutError((size_t)m_measured[i]);  // P(0 | 1), P(1 | 0)
        if (rng.randProb() < (b ? ro.first : ro.second)) b ^= 1;
      }
      bits.push_back(b ? '1' : '0');
    }
    buffer->appendMeasurement(bits);
  }
}

}  // namespace tnqvm
