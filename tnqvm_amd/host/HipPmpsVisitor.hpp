// HipPmpsVisitor.hpp -- the "hip-pmps" TNQVM visitor: noisy circuits on a locally purified
// MPS whose arithmetic runs in libhipmps (hipmps_apply_kraus and the shared gate path).
//
// Drop-in for ExaTnPmpsVisitor (tnqvm/visitors/exatn-mpo/ExaTnPmpsVisitor.hpp:64-127):
//   noise-model   xacc::NoiseModel (pointer-like option, as the reference reads it,
//                 ExaTnPmpsVisitor.cpp:470-474): after every gate, its single-qubit Kraus
//                 channels (multi-qubit channels are ignored, as there, :787-797); readout
//                 errors on the sampled bits (:158-170)
//   shots         default 1024 for a noisy run (:490)
//   svd-cutoff    default 1e-9, no bond cap (truncateSvdTensors' eps, :1255); the other
//                 hip-mps truncation and precision keys apply too
// Results: measurements (char i <-> i-th Measure) sampled from the outcome distribution of
//   the measured qubits (at most 15, :763-766), that distribution as
//   "density-matrix-diagonal", "exp-val-z" (Z parity of the measured qubits), "norm"
//   (trace of rho), "kraus-dims".
#pragma once

#include "HipMpsVisitor.hpp"

namespace tnqvm {

class HipPmpsVisitor : public HipMpsVisitor {
 public:
  const std::string name() const override { return "hip-pmps"; }
  const std::string description() const override { return "Purified-MPS (noisy) visitor on MI355X (HIP)"; }
  std::shared_ptr<TNQVMVisitor> clone() override { return std::make_shared<HipPmpsVisitor>(); }
  bool supportVqeMode() const override { return false; }

  void initialize(std::shared_ptr<AcceleratorBuffer> buffer, int nbShots) override;
  void finalize() override;
  // the reference applies the identity, and so the noise that comes with it (:1128-1131)
  void visit(xacc::quantum::Identity& g) override { after_gate(g); }

 protected:
  void after_gate(xacc::Instruction& g) override;

 private:
  xacc::NoiseModel* m_noise = nullptr;
};

}  // namespace tnqvm
