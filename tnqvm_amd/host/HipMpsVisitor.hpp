// HipMpsVisitor.hpp -- the "hip-mps" TNQVM visitor: an MPS circuit simulator whose
// arithmetic runs in libhipmps (HIP kernels on one MI355X, include/hipmps.h).
//
// Drop-in for ExatnMpsVisitor (tnqvm/visitors/exatn-mps/ExaTnMpsVisitor.hpp:53-161):
// same service interface (TNQVMVisitor), same option keys with the same meaning, same
// result keys on the AcceleratorBuffer.
//
// Options (HeterogeneousMap)                                  reference
//   svd-cutoff      double  default DBL_MIN                    ExaTnMpsVisitor.cpp:257-263
//   max-bond-dim    int     default INT_MAX-1                  :265-271
//   bitstring       vector<int> 0/1/-1 -> amplitude / slice    :776-822 (honoured always)
//   precision       string  "fp64" (default) | "mixed"         new
//   renormalize     bool    default false (reference: none)    new
//   partial-norm    string  "squared" (default) | "l2"         :2421-2429 ambiguity
//   absorb          string  "sqrt" (default) | "right"         :1623 (SVDLR)
//   zz-pairs        vector<int> flattened pairs -> exp-val-zz  new
//   device          int     HIP device ordinal                 new
// Results: norm, exp-val-z (parity over measured qubits, unnormalised, :616-643),
//   measurements (shots > 0), amplitude-real/-imag or amplitude-real-vec/-imag-vec,
//   plus exp-val-z-per-qubit, exp-val-zz, discarded-weight, truncation-fidelity,
//   max-bond-dim-reached.
#pragma once

#include <vector>

#include "TNQVMVisitor.hpp"
#include "hipmps.h"

namespace tnqvm {

class HipMpsVisitor : public TNQVMVisitor {
 public:
  HipMpsVisitor() = default;
  ~HipMpsVisitor() override;

  const std::string name() const override { return "hip-mps"; }
  const std::string description() const override { return "MPS visitor on MI355X (HIP)"; }
  std::shared_ptr<TNQVMVisitor> clone() override { return std::make_shared<HipMpsVisitor>(); }

  void initialize(std::shared_ptr<AcceleratorBuffer> buffer, int nbShots) override;
  void finalize() override;
  const double getExpectationValueZ(std::shared_ptr<CompositeInstruction> function) override;
  const std::vector<std::complex<double>> getState() override;
  bool supportVqeMode() const override { return true; }

  // one-qubit gates
  void visit(xacc::quantum::Identity&) override {}  // skipped, :870-880
  void visit(xacc::quantum::Hadamard& g) override { apply(g); }
  void visit(xacc::quantum::X& g) override { apply(g); }
  void visit(xacc::quantum::Y& g) override { apply(g); }
  void visit(xacc::quantum::Z& g) override { apply(g); }
  void visit(xacc::quantum::S& g) override { apply(g); }
  void visit(xacc::quantum::Sdg& g) override { apply(g); }
  void visit(xacc::quantum::T& g) override { apply(g); }
  void visit(xacc::quantum::Tdg& g) override { apply(g); }
  void visit(xacc::quantum::Rx& g) override { apply(g); }
  void visit(xacc::quantum::Ry& g) override { apply(g); }
  void visit(xacc::quantum::Rz& g) override { apply(g); }
  void visit(xacc::quantum::U& g) override { apply(g); }
  // two-qubit gates
  void visit(xacc::quantum::CNOT& g) override { apply(g); }
  void visit(xacc::quantum::CY& g) override { apply(g); }
  void visit(xacc::quantum::CZ& g) override { apply(g); }
  void visit(xacc::quantum::CH& g) override { apply(g); }
  void visit(xacc::quantum::CRZ& g) override { apply(g); }
  void visit(xacc::quantum::CPhase& g) override { apply(g); }
  void visit(xacc::quantum::Swap& g) override { apply(g); }
  void visit(xacc::quantum::iSwap& g) override { apply(g); }
  void visit(xacc::quantum::fSim& g) override { apply(g); }
  void visit(xacc::quantum::Measure& g) override { m_measured.push_back((int)g.bits()[0]); }

  // the engine shared by all hip-mps visitors of this process (one per device)
  static hipmps_engine engine(int device);

 protected:
  void apply(xacc::Instruction& g);
  // called after every gate has been queued on the state (noise channels of the purified-MPS
  // visitor); nothing here
  virtual void after_gate(xacc::Instruction&) {}
  void apply_matrix(const std::vector<size_t>& bits, const double* m);
  void check(int rc, const char* what) const;
  std::vector<std::string> sample_state_vector(int shots);
  std::vector<std::string> sample_sequential(int shots);
  // truncation options from the shared option keys, with the given default cutoff
  hipmps_trunc_opts trunc_options(double default_cutoff) const;
  // the device's shared engine with this run's Jacobi sweep cap ("jacobi-max-sweeps")
  hipmps_engine run_engine(int device) const;

  hipmps_state m_state = nullptr;
  int m_n = 0;
  int m_shots = -1;
  std::vector<int> m_measured;
};

}  // namespace tnqvm
