// HipPmpsVisitor.cpp -- "hip-pmps" visitor over libhipmps.
#include "HipPmpsVisitor.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <iostream>

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
  check(hipmps_state_create(run_engine(device), m_n, &t, &m_state), "state creation");
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

// finalize (ExaTnPmpsVisitor.cpp:657-782): the density matrix the reference stores under
// "density_matrix" (all qubits when n <= 15, :664-683; else the measured qubits, in qubit order,
// :701-761), the outcome distribution of the measured qubits from its diagonal, then the shots
// drawn from it with the model's readout errors
void HipPmpsVisitor::finalize() {
  int rc = hipmps_flush(m_state);
  if (rc != 0 && rc != HIPMPS_E_CONVERGENCE) check(rc, "flush");
  double tr = 0.0;
  check(hipmps_norm(m_state, &tr), "trace");
  buffer->addExtraInfo("norm", tr);
  hipmps_stats_t st{};
  hipmps_stats(m_state, &st);
  buffer->addExtraInfo("svd-converged", st.svds_nonconverged == 0 ? 1 : 0);
  buffer->addExtraInfo("svd-nonconverged-count", st.svds_nonconverged);
  if (st.svds_nonconverged > 0)
    std::cerr << "[WARNING] " << st.svds_nonconverged
              << " SVD(s) reached the Jacobi sweep cap without converging; results may be inaccurate\n";
  std::vector<int> kd(m_n);
  check(hipmps_kraus_dims(m_state, kd.data()), "Kraus dims");
  buffer->addExtraInfo("kraus-dims", kd);
  if (m_measured.empty()) return;
  const int k = (int)m_measured.size();
  if (k > MAX_QUBITS_FOR_MEASURE)
    xacc::error("Measure more than " + std::to_string(MAX_QUBITS_FOR_MEASURE) + " is not supported.");
  std::vector<int> dmq;
  if (m_n <= MAX_QUBITS_FOR_MEASURE) {
    for (int q = 0; q < m_n; ++q) dmq.push_back(q);
  } else {
    dmq = m_measured;
    std::sort(dmq.begin(), dmq.end());
    dmq.erase(std::unique(dmq.begin(), dmq.end()), dmq.end());
  }
  const size_t dim = (size_t)1 << dmq.size();
  // flattened ket index fastest, as the reference's; written straight into the result pairs
  std::vector<std::pair<double, double>> flat(dim * dim);
  static_assert(sizeof(std::pair<double, double>) == 2 * sizeof(double), "pair layout");
  check(hipmps_reduced_density_matrix(m_state, dmq.data(), (int)dmq.size(), reinterpret_cast<double*>(flat.data()),
                                      dim * dim),
        "density matrix");
  // outcome y of the measured qubits (bit i = qubit m_measured[i]); several Measure on one
  // qubit read the same bit
  std::vector<int> slot(k);
  for (int i = 0; i < k; ++i)
    slot[i] = (int)(std::lower_bound(dmq.begin(), dmq.end(), m_measured[i]) - dmq.begin());
  std::vector<double> probs((size_t)1 << k, 0.0);
  for (size_t x = 0; x < dim; ++x) {
    size_t y = 0;
    for (int i = 0; i < k; ++i) y |= ((x >> slot[i]) & 1) << i;
    probs[y] += flat[x + dim * x].first;
  }
  buffer->addExtraInfo("density_matrix", flat);
  flat.clear();
  flat.shrink_to_fit();
  buffer->addExtraInfo("density-matrix-diagonal", probs);
  double ez = 0.0;
  for (size_t y = 0; y < probs.size(); ++y) ez += (__builtin_popcountll(y) & 1) ? -probs[y] : probs[y];
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
