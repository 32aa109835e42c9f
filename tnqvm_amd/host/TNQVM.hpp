// TNQVM.hpp -- the "tnqvm" accelerator driver in front of the hip-mps visitor.
//
// Mirrors tnqvm/TNQVM.hpp:46-125 and tnqvm/TNQVM.cpp:47-200: options are parsed by
// updateConfiguration (tnqvm-visitor / backend, shots, seed, vqe-mode, verbose) and
// forwarded to the visitor; execute() walks the IR into the visitor between
// initialize() and finalize(); the multi-kernel form runs each kernel into a child
// buffer, or, in VQE mode with a visitor that supports it, simulates the shared ansatz
// once and evaluates every observable sub-circuit on that state.
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "TNQVMVisitor.hpp"

namespace tnqvm {

class TNQVM : public xacc::Accelerator {
 public:
  void initialize(const HeterogeneousMap& params = {}) override {
    options.clear();
    updateConfiguration(params);
  }
  void updateConfiguration(const HeterogeneousMap& config) override;
  void execute(std::shared_ptr<AcceleratorBuffer> buffer, const std::shared_ptr<CompositeInstruction> kernel) override;
  void execute(std::shared_ptr<AcceleratorBuffer> buffer,
               const std::vector<std::shared_ptr<CompositeInstruction>> functions) override;
  const std::vector<std::complex<double>> getAcceleratorState(std::shared_ptr<CompositeInstruction> program);
  const std::string name() const override { return "tnqvm"; }
  const std::string description() const override { return "TNQVM accelerator (hip-mps build)"; }
  const std::string& getVisitorName() const { return backendName; }
  HeterogeneousMap getExecutionInfo() const override {
    auto r = visitor ? visitor->getExecutionInfo() : HeterogeneousMap{};
    if (visitor) r.insert("visitor", visitor->name());
    return r;
  }

  // This build registers "hip-mps" only; it is also the fallback for unknown names
  // (the reference falls back to "itensor-mps", TNQVM.cpp:45).
  static const std::string DEFAULT_VISITOR_BACKEND;

 protected:
  std::shared_ptr<TNQVMVisitor> visitor;

 private:
  bool vqeMode = true;
  std::string backendName = DEFAULT_VISITOR_BACKEND;
  int nbShots = -1;
  HeterogeneousMap options;
};

}  // namespace tnqvm
