// TNQVMVisitor.hpp -- the TNQVM visitor plug-in interface.
//
// Same members and meaning as tnqvm/visitors/TNQVMVisitor.hpp:58-82 (initialize,
// getExpectationValueZ, getState, finalize, setOptions, setKernelName, supportVqeMode,
// getExecutionInfo), declared over the XACC API subset in xacc_lite.hpp.
#pragma once

#include <complex>
#include <memory>
#include <string>
#include <vector>

#include "xacc_lite.hpp"

namespace tnqvm {

using xacc::AcceleratorBuffer;
using xacc::CompositeInstruction;
using xacc::HeterogeneousMap;

class TNQVMVisitor : public xacc::quantum::AllGateVisitor {
 public:
  virtual ~TNQVMVisitor() = default;
  virtual const std::string name() const = 0;
  virtual const std::string description() const = 0;
  virtual std::shared_ptr<TNQVMVisitor> clone() = 0;
  virtual void initialize(std::shared_ptr<AcceleratorBuffer> buffer, int nbShots = 1) = 0;
  virtual const double getExpectationValueZ(std::shared_ptr<CompositeInstruction> function) = 0;
  virtual const std::vector<std::complex<double>> getState() { return {}; }
  virtual void finalize() = 0;
  void setOptions(const HeterogeneousMap& in_options) { options = in_options; }
  virtual void setKernelName(const std::string&) {}
  virtual bool supportVqeMode() const { return false; }
  HeterogeneousMap getExecutionInfo() const { return executionInfo; }

 protected:
  std::shared_ptr<AcceleratorBuffer> buffer;
  HeterogeneousMap options;
  HeterogeneousMap executionInfo;
};

}  // namespace tnqvm
