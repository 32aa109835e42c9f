// TNQVM.cpp -- accelerator driver, the lnn-transform IR pass and service registration.
#include "TNQVM.hpp"

#include "HipMpsVisitor.hpp"
#include "HipPmpsVisitor.hpp"
#include "RandomEngine.hpp"

namespace tnqvm {

const std::string TNQVM::DEFAULT_VISITOR_BACKEND = "hip-mps";

namespace {
int shotCountOption(const HeterogeneousMap& o) { return o.keyExists<int>("shots") ? o.get<int>("shots") : -1; }

bool isMpsVisitor(const std::string& n) {
  // visitors that need nearest-neighbour gates (TNQVM.cpp:119-124); hip-mps routes itself,
  // hip-pmps takes the routed circuit like exatn-pmps, so that routing swaps carry noise
  return n == "itensor-mps" || n == "exatn-mps" || n == "exatn-pmps" || n == "hip-pmps";
}

void walk(const std::shared_ptr<CompositeInstruction>& k, const std::shared_ptr<TNQVMVisitor>& v) {
  xacc::InstructionIterator it(k);
  while (it.hasNext()) {
    auto inst = it.next();
    if (inst->isEnabled()) inst->accept(v);
  }
}
}  // namespace

// TNQVM.hpp:61-125
void TNQVM::updateConfiguration(const HeterogeneousMap& config) {
  if (config.keyExists<bool>("verbose") && config.get<bool>("verbose")) xacc::verbose = true;
  if (config.keyExists<bool>("vqe-mode")) vqeMode = config.get<bool>("vqe-mode");
  if (config.stringExists("tnqvm-visitor") || config.stringExists("backend")) {
    const auto requested = config.stringExists("tnqvm-visitor") ? config.getString("tnqvm-visitor") : config.getString("backend");
    if (xacc::hasService<TNQVMVisitor>(requested)) {
      backendName = requested;
    } else {
      backendName = DEFAULT_VISITOR_BACKEND;
      xacc::warning("The requested TNQVM visitor backend '" + requested +
                    "' cannot be found in the service registry. The default visitor backend of type '" +
                    DEFAULT_VISITOR_BACKEND + "' will be used.");
    }
  }
  if (config.keyExists<int>("shots")) {
    nbShots = config.get<int>("shots");
    if (nbShots < 1) xacc::error("Invalid 'shots' parameter.");
  }
  if (config.keyExists<int>("seed")) randomEngine::get_instance().setSeed((size_t)config.get<int>("seed"));
  options.merge(config);
}

// TNQVM.cpp:106-139
void TNQVM::execute(std::shared_ptr<AcceleratorBuffer> buffer, const std::shared_ptr<CompositeInstruction> kernel) {
  visitor = xacc::getService<TNQVMVisitor>(getVisitorName());
  visitor->setOptions(options);
  visitor->initialize(buffer, shotCountOption(options));
  visitor->setKernelName(kernel->name());
  if (isMpsVisitor(visitor->name()) && xacc::hasService<xacc::IRTransformation>("lnn-transform")) {
    auto opt = xacc::getService<xacc::IRTransformation>("lnn-transform");
    HeterogeneousMap o;
    o.insert("max-distance", 1);
    opt->apply(kernel, nullptr, o);
  }
  xacc::info("Number of instructions: " + std::to_string(kernel->nInstructions()));
  walk(kernel, visitor);
  visitor->finalize();
}

// TNQVM.cpp:47-104.  In VQE mode the kernels are split into their longest common
// instruction prefix (the ansatz) and the per-kernel remainders (basis change +
// measurements), which is what XACC's ObservedAnsatz decomposition produces for
// observable-generated kernels.
void TNQVM::execute(std::shared_ptr<AcceleratorBuffer> buffer,
                    const std::vector<std::shared_ptr<CompositeInstruction>> functions) {
  visitor = xacc::getService<TNQVMVisitor>(getVisitorName())->clone();
  if (vqeMode && functions.size() > 1 && visitor->supportVqeMode()) {
    std::vector<std::vector<std::shared_ptr<xacc::Instruction>>> flat;
    for (auto& f : functions) {
      std::vector<std::shared_ptr<xacc::Instruction>> v;
      xacc::InstructionIterator it(f);
      while (it.hasNext()) {
        auto i = it.next();
        if (i->isEnabled()) v.push_back(i);
      }
      flat.push_back(v);
    }
    auto same = [](const std::shared_ptr<xacc::Instruction>& a, const std::shared_ptr<xacc::Instruction>& b) {
      if (a->name() != b->name() || a->bits() != b->bits()) return false;
      auto pa = a->getParameters(), pb = b->getParameters();
      if (pa.size() != pb.size()) return false;
      for (size_t i = 0; i < pa.size(); ++i)
        if (pa[i].as<double>() != pb[i].as<double>()) return false;
      return true;
    };
    size_t common = flat[0].size();
    for (auto& v : flat) {
      size_t c = 0;
      while (c < common && c < v.size() && same(v[c], flat[0][c]) && v[c]->name() != "Measure") ++c;
      common = c;
    }
    visitor->setOptions(options);
    visitor->initialize(buffer, shotCountOption(options));
    for (size_t i = 0; i < common; ++i) flat[0][i]->accept(visitor);
    for (size_t k = 0; k < functions.size(); ++k) {
      auto obs = std::make_shared<CompositeInstruction>(functions[k]->name());
      for (size_t i = common; i < flat[k].size(); ++i) obs->addInstruction(flat[k][i]);
      auto tmp = std::make_shared<AcceleratorBuffer>(functions[k]->name(), buffer->size());
      tmp->addExtraInfo("exp-val-z", visitor->getExpectationValueZ(obs));
      buffer->appendChild(functions[k]->name(), tmp);
    }
    visitor->finalize();
    return;
  }
  for (auto& f : functions) {
    auto tmp = std::make_shared<AcceleratorBuffer>(f->name(), buffer->size());
    execute(tmp, f);
    buffer->appendChild(f->name(), tmp);
  }
}

// TNQVM.cpp:203-244
const std::vector<std::complex<double>> TNQVM::getAcceleratorState(std::shared_ptr<CompositeInstruction> program) {
  visitor = xacc::getService<TNQVMVisitor>(getVisitorName());
  size_t maxBit = 0;
  xacc::InstructionIterator it1(program);
  while (it1.hasNext()) {
    auto i = it1.next();
    for (auto b : i->bits()) maxBit = std::max(maxBit, b);
  }
  if (maxBit == 0) maxBit++;
  auto buffer = std::make_shared<AcceleratorBuffer>("q", (int)maxBit + 1);
  visitor->setOptions(options);
  visitor->initialize(buffer, shotCountOption(options));
  walk(program, visitor);
  visitor->finalize();
  return visitor->getState();
}

// "lnn-transform": 2-qubit gates farther apart than max-distance become swaps toward
// each other (meet in the middle), the gate, and swaps back
// (tnqvm/visitors/exatn-mps/NearestNeighborTransform.hpp:43-135).
class NearestNeighborTransform : public xacc::IRTransformation {
 public:
  void apply(std::shared_ptr<CompositeInstruction> program, const std::shared_ptr<xacc::Accelerator>,
             const HeterogeneousMap& opts = {}) override {
    int maxDistance = opts.keyExists<int>("max-distance") ? opts.get<int>("max-distance") : 1;
    std::vector<std::shared_ptr<xacc::Instruction>> flat, out;
    xacc::InstructionIterator it(program);
    while (it.hasNext()) {
      auto i = it.next();
      if (i->isEnabled()) flat.push_back(i->clone());
    }
    auto far = [&](long a, long b) { return std::labs(a - b) > maxDistance; };
    for (auto& inst : flat) {
      if (inst->bits().size() != 2 || !far((long)inst->bits()[0], (long)inst->bits()[1])) {
        out.push_back(inst);
        continue;
      }
      const size_t lower0 = std::min(inst->bits()[0], inst->bits()[1]);
      const size_t upper0 = std::max(inst->bits()[0], inst->bits()[1]);
      size_t lo = lower0, up = upper0;
      for (;;) {
        out.push_back(xacc::createGate("Swap", {lo, lo + 1}));
        ++lo;
        if (!far((long)lo, (long)up)) break;
        out.push_back(xacc::createGate("Swap", {up, up - 1}));
        --up;
        if (!far((long)lo, (long)up)) break;
      }
      const bool ordered = inst->bits()[0] < inst->bits()[1];
      inst->setBits(ordered ? std::vector<size_t>{lo, up} : std::vector<size_t>{up, lo});
      out.push_back(inst);
      for (size_t i = lo; i > lower0; --i) out.push_back(xacc::createGate("Swap", {i, i - 1}));
      for (size_t i = up; i < upper0; ++i) out.push_back(xacc::createGate("Swap", {i, i + 1}));
    }
    program->clear();
    program->addInstructions(out);
  }
  const std::string name() const override { return "lnn-transform"; }
};

namespace {
// service registration, as the bundle activators do (TNQVMActivator.cpp:9-22,
// ExaTnMpsActivator.cpp:14-19)
struct Activator {
  Activator() {
    xacc::registerService<xacc::Accelerator>("tnqvm", std::make_shared<TNQVM>());
    xacc::registerService<TNQVMVisitor>("hip-mps", std::make_shared<HipMpsVisitor>());
    xacc::registerService<TNQVMVisitor>("hip-pmps", std::make_shared<HipPmpsVisitor>());
    xacc::registerService<xacc::IRTransformation>("lnn-transform", std::make_shared<NearestNeighborTransform>());
  }
};
Activator activator;
}  // namespace

}  // namespace tnqvm
