// xacc_lite.hpp -- the subset of the XACC C++ API that the TNQVM driver and its
// visitors use, so that the hip-mps visitor compiles and runs where XACC itself is not
// installed.  Class and function names follow XACC's public API as TNQVM uses it
// (tnqvm/TNQVM.{hpp,cpp}, tnqvm/visitors/TNQVMVisitor.hpp, the exatn-mps tests); the
// implementation is this project's own and covers only what those call sites need:
//   HeterogeneousMap, AcceleratorBuffer (extra info, measurements, children),
//   Instruction / Gate classes, CompositeInstruction, InstructionIterator,
//   AllGateVisitor, Accelerator, IRTransformation, a name -> service registry,
//   getAccelerator / qalloc / getIRProvider("quantum") / getCompiler("xasm").
// Building the visitor against real XACC replaces this header by "xacc.hpp".
#pragma once

#include <algorithm>
#include <any>
#include <cmath>
#include <complex>
#include <functional>
#include <iostream>
#include <map>
#include <memory>
#include <numeric>
#include <sstream>
#include <stdexcept>
#include <string>
#include <typeindex>
#include <utility>
#include <vector>

namespace xacc {

// ------------------------------------------------------------------ logging / errors
inline bool verbose = false;
inline void info(const std::string& msg) {
  if (verbose) std::cout << "[xacc] " << msg << "\n";
}
inline void warning(const std::string& msg) { std::cerr << "[xacc warning] " << msg << "\n"; }
inline void error(const std::string& msg) { throw std::runtime_error(msg); }

// ------------------------------------------------------------------ HeterogeneousMap
class HeterogeneousMap {
 public:
  HeterogeneousMap() = default;
  HeterogeneousMap(std::initializer_list<std::pair<std::string, std::any>> l) {
    for (auto& kv : l) insert(kv.first, kv.second);
  }
  template <typename T>
  void insert(const std::string& key, const T& v) {
    m_[key] = normalize(std::any(v));
  }
  void insert(const std::string& key, const char* v) { m_[key] = std::string(v); }
  template <typename T>
  bool keyExists(const std::string& key) const {
    auto it = m_.find(key);
    return it != m_.end() && it->second.type() == typeid(T);
  }
  bool stringExists(const std::string& key) const { return keyExists<std::string>(key); }
  template <typename T>
  T get(const std::string& key) const {
    auto it = m_.find(key);
    if (it == m_.end()) error("HeterogeneousMap: missing key " + key);
    return std::any_cast<T>(it->second);
  }
  std::string getString(const std::string& key) const { return get<std::string>(key); }
  // options that carry an object by pointer (XACC's pointerLikeExists / getPointerLike):
  // a raw T* or a std::shared_ptr<T>
  template <typename T>
  bool pointerLikeExists(const std::string& key) const {
    return keyExists<T*>(key) || keyExists<std::shared_ptr<T>>(key);
  }
  template <typename T>
  T* getPointerLike(const std::string& key) const {
    if (keyExists<T*>(key)) return get<T*>(key);
    if (keyExists<std::shared_ptr<T>>(key)) return get<std::shared_ptr<T>>(key).get();
    error("HeterogeneousMap: no pointer-like value for key " + key);
    return nullptr;
  }
  void merge(const HeterogeneousMap& o) {
    for (auto& kv : o.m_) m_[kv.first] = kv.second;
  }
  void clear() { m_.clear(); }
  size_t size() const { return m_.size(); }

 private:
  static std::any normalize(const std::any& a) {
    if (a.type() == typeid(const char*)) return std::string(std::any_cast<const char*>(a));
    return a;
  }
  std::map<std::string, std::any> m_;
};

// ------------------------------------------------------------------ AcceleratorBuffer
class ExtraInfo {
 public:
  ExtraInfo() = default;
  explicit ExtraInfo(std::any v) : v_(std::move(v)) {}
  template <typename T>
  T as() const {
    return std::any_cast<T>(v_);
  }
  bool empty() const { return !v_.has_value(); }

 private:
  std::any v_;
};

class AcceleratorBuffer {
 public:
  AcceleratorBuffer(const std::string& name, int n) : name_(name), n_(n) {}
  explicit AcceleratorBuffer(int n) : name_("q"), n_(n) {}
  int size() const { return n_; }
  const std::string& name() const { return name_; }
  template <typename T>
  void addExtraInfo(const std::string& key, const T& v) {
    info_[key] = ExtraInfo(std::any(v));
  }
  bool hasExtraInfoKey(const std::string& key) const { return info_.count(key) > 0; }
  ExtraInfo operator[](const std::string& key) const {
    auto it = info_.find(key);
    if (it == info_.end()) error("AcceleratorBuffer: no extra info " + key);
    return it->second;
  }
  void appendMeasurement(const std::string& bits) { counts_[bits]++; }
  std::map<std::string, int> getMeasurementCounts() const { return counts_; }
  double computeMeasurementProbability(const std::string& bits) const {
    int tot = 0;
    for (auto& kv : counts_) tot += kv.second;
    auto it = counts_.find(bits);
    return (tot == 0 || it == counts_.end()) ? 0.0 : (double)it->second / tot;
  }
  // <Z...Z> of the measured bits: exp-val-z if present, else from the counts (parity)
  double getExpectationValueZ() const {
    auto it = info_.find("exp-val-z");
    if (it != info_.end()) return it->second.as<double>();
    int tot = 0;
    double acc = 0;
    for (auto& kv : counts_) {
      int ones = (int)std::count(kv.first.begin(), kv.first.end(), '1');
      acc += (ones % 2 ? -1.0 : 1.0) * kv.second;
      tot += kv.second;
    }
    return tot ? acc / tot : 0.0;
  }
  void appendChild(const std::string& name, std::shared_ptr<AcceleratorBuffer> b) {
    children_.emplace_back(name, std::move(b));
  }
  std::vector<std::shared_ptr<AcceleratorBuffer>> getChildren() const {
    std::vector<std::shared_ptr<AcceleratorBuffer>> out;
    for (auto& c : children_) out.push_back(c.second);
    return out;
  }
  void print() const {
    std::cout << "AcceleratorBuffer " << name_ << " (" << n_ << " qubits)\n";
    for (auto& kv : counts_) std::cout << "  " << kv.first << ": " << kv.second << "\n";
  }

 private:
  std::string name_;
  int n_;
  std::map<std::string, ExtraInfo> info_;
  std::map<std::string, int> counts_;
  std::vector<std::pair<std::string, std::shared_ptr<AcceleratorBuffer>>> children_;
};

// ------------------------------------------------------------------ instructions
class InstructionParameter {
 public:
  InstructionParameter(double v) : d_(v), isd_(true) {}
  InstructionParameter(int v) : d_(v), isd_(true) {}
  InstructionParameter(const std::string& s) : s_(s), isd_(false) {}
  template <typename T>
  T as() const;
  bool isVariable() const { return !isd_; }
  const std::string& variable() const { return s_; }

 private:
  double d_ = 0;
  std::string s_;
  bool isd_ = true;
};
template <>
inline double InstructionParameter::as<double>() const {
  if (!isd_) error("unbound instruction parameter " + s_);
  return d_;
}
template <>
inline int InstructionParameter::as<int>() const { return (int)as<double>(); }

namespace quantum {
class AllGateVisitor;
}

class Instruction : public std::enable_shared_from_this<Instruction> {
 public:
  virtual ~Instruction() = default;
  virtual const std::string name() const = 0;
  virtual int nRequiredBits() const = 0;
  virtual void accept(std::shared_ptr<quantum::AllGateVisitor> v) = 0;
  virtual void accept(quantum::AllGateVisitor* v) = 0;
  virtual std::shared_ptr<Instruction> clone() = 0;
  virtual bool isComposite() const { return false; }
  bool isEnabled() const { return enabled_; }
  void disable() { enabled_ = false; }
  void enable() { enabled_ = true; }
  const std::vector<size_t>& bits() const { return bits_; }
  void setBits(const std::vector<size_t>& b) { bits_ = b; }
  const std::vector<InstructionParameter>& getParameters() const { return params_; }
  InstructionParameter getParameter(int i) const { return params_.at(i); }
  void setParameter(int i, const InstructionParameter& p) { params_.at(i) = p; }
  void setParameters(const std::vector<InstructionParameter>& p) { params_ = p; }
  std::string toString() const {
    std::ostringstream o;
    o << name();
    for (auto& p : params_) o << (p.isVariable() ? " " + p.variable() : " " + std::to_string(p.as<double>()));
    for (auto b : bits_) o << " q" << b;
    return o.str();
  }

 protected:
  std::vector<size_t> bits_;
  std::vector<InstructionParameter> params_;
  bool enabled_ = true;
};

class CompositeInstruction : public Instruction {
 public:
  explicit CompositeInstruction(const std::string& n) : name_(n) {}
  const std::string name() const override { return name_; }
  int nRequiredBits() const override { return 0; }
  void accept(std::shared_ptr<quantum::AllGateVisitor>) override {}
  void accept(quantum::AllGateVisitor*) override {}
  std::shared_ptr<Instruction> clone() override {
    auto c = std::make_shared<CompositeInstruction>(name_);
    for (auto& i : insts_) c->addInstruction(i->clone());
    c->vars_ = vars_;
    return c;
  }
  bool isComposite() const override { return true; }
  void addInstruction(std::shared_ptr<Instruction> i) { insts_.push_back(std::move(i)); }
  void addInstructions(const std::vector<std::shared_ptr<Instruction>>& v) {
    for (auto& i : v) addInstruction(i);
  }
  std::vector<std::shared_ptr<Instruction>> getInstructions() const { return insts_; }
  std::shared_ptr<Instruction> getInstruction(size_t i) const { return insts_.at(i); }
  size_t nInstructions() const { return insts_.size(); }
  void clear() { insts_.clear(); }
  void addVariable(const std::string& v) { vars_.push_back(v); }
  const std::vector<std::string>& getVariables() const { return vars_; }
  std::string toString() const {
    std::string s;
    for (auto& i : insts_) s += i->toString() + "\n";
    return s;
  }
  // bind the declared variables (in order) and return a new composite
  std::shared_ptr<CompositeInstruction> operator()(const std::vector<double>& values);

 private:
  std::string name_;
  std::vector<std::shared_ptr<Instruction>> insts_;
  std::vector<std::string> vars_;
};

// depth-first iteration over the leaves of a composite (composites are skipped)
class InstructionIterator {
 public:
  explicit InstructionIterator(std::shared_ptr<CompositeInstruction> root) { flatten(root); }
  bool hasNext() const { return pos_ < list_.size(); }
  std::shared_ptr<Instruction> next() { return list_.at(pos_++); }

 private:
  void flatten(const std::shared_ptr<CompositeInstruction>& c) {
    for (auto& i : c->getInstructions()) {
      if (i->isComposite()) flatten(std::static_pointer_cast<CompositeInstruction>(i));
      else list_.push_back(i);
    }
  }
  std::vector<std::shared_ptr<Instruction>> list_;
  size_t pos_ = 0;
};

namespace quantum {

class Gate;
#define XACC_LITE_GATES(GATE_)                                                                      \
  GATE_(Identity, "I", 1, 0) GATE_(Hadamard, "H", 1, 0) GATE_(X, "X", 1, 0) GATE_(Y, "Y", 1, 0)       \
  GATE_(Z, "Z", 1, 0) GATE_(S, "S", 1, 0) GATE_(Sdg, "Sdg", 1, 0) GATE_(T, "T", 1, 0)                  \
  GATE_(Tdg, "Tdg", 1, 0) GATE_(Rx, "Rx", 1, 1) GATE_(Ry, "Ry", 1, 1) GATE_(Rz, "Rz", 1, 1)            \
  GATE_(U, "U", 1, 3) GATE_(Measure, "Measure", 1, 0) GATE_(CNOT, "CNOT", 2, 0) GATE_(CY, "CY", 2, 0)  \
  GATE_(CZ, "CZ", 2, 0) GATE_(CH, "CH", 2, 0) GATE_(CRZ, "CRZ", 2, 1) GATE_(CPhase, "CPhase", 2, 1)    \
  GATE_(Swap, "Swap", 2, 0) GATE_(iSwap, "iSwap", 2, 0) GATE_(fSim, "fSim", 2, 2)

#define XACC_LITE_FWD(cls, nm, nb, np) class cls;
XACC_LITE_GATES(XACC_LITE_FWD)
#undef XACC_LITE_FWD

// Default visit() of every gate is a no-op, as XACC's AllGateVisitor.
class AllGateVisitor {
 public:
  virtual ~AllGateVisitor() = default;
#define XACC_LITE_VISIT(cls, nm, nb, np) \
  virtual void visit(cls&) {}
  XACC_LITE_GATES(XACC_LITE_VISIT)
#undef XACC_LITE_VISIT
};

class Gate : public Instruction {
 public:
  int nRequiredBits() const override { return nbits_; }

 protected:
  int nbits_ = 1;
};

#define XACC_LITE_CLASS(cls, nm, nb, np)                                                         \
  class cls : public Gate {                                                                    \
   public:                                                                                     \
    cls() {                                                                                    \
      nbits_ = nb;                                                                             \
      params_.assign(np, InstructionParameter(0.0));                                           \
    }                                                                                          \
    cls(const std::vector<size_t>& b) : cls() { bits_ = b; }                                   \
    const std::string name() const override { return nm; }                                     \
    void accept(std::shared_ptr<AllGateVisitor> v) override { v->visit(*this); }               \
    void accept(AllGateVisitor* v) override { v->visit(*this); }                               \
    std::shared_ptr<Instruction> clone() override { return std::make_shared<cls>(*this); }     \
  };
XACC_LITE_GATES(XACC_LITE_CLASS)
#undef XACC_LITE_CLASS

}  // namespace quantum

inline std::shared_ptr<Instruction> createGate(const std::string& name, const std::vector<size_t>& bits,
                                               const std::vector<InstructionParameter>& params = {}) {
  std::shared_ptr<Instruction> g;
  std::string n = name == "CX" ? "CNOT" : (name == "U3" ? "U" : name);
#define XACC_LITE_MAKE(cls, nm, nb, np) \
  if (n == nm) g = std::make_shared<quantum::cls>();
  XACC_LITE_GATES(XACC_LITE_MAKE)
#undef XACC_LITE_MAKE
  if (!g) error("unknown gate " + name);
  if ((int)bits.size() != g->nRequiredBits()) error("gate " + name + ": wrong number of qubits");
  g->setBits(bits);
  if (!params.empty()) {
    if (params.size() != g->getParameters().size()) error("gate " + name + ": wrong number of parameters");
    g->setParameters(params);
  }
  return g;
}

inline std::shared_ptr<CompositeInstruction> CompositeInstruction::operator()(const std::vector<double>& values) {
  if (values.size() != vars_.size()) error("composite " + name_ + ": expected " + std::to_string(vars_.size()) + " values");
  std::map<std::string, double> env;
  for (size_t i = 0; i < vars_.size(); ++i) env[vars_[i]] = values[i];
  auto out = std::make_shared<CompositeInstruction>(name_);
  for (auto& i : insts_) {
    auto c = i->clone();
    auto ps = c->getParameters();
    for (auto& p : ps) {
      if (!p.isVariable()) continue;
      // "t", "-t", "2*t", "t/2"
      std::string e = p.variable();
      double sign = 1.0, mul = 1.0, div = 1.0;
      if (!e.empty() && e[0] == '-') { sign = -1.0; e = e.substr(1); }
      auto star = e.find('*');
      if (star != std::string::npos) { mul = std::stod(e.substr(0, star)); e = e.substr(star + 1); }
      auto slash = e.find('/');
      if (slash != std::string::npos) { div = std::stod(e.substr(slash + 1)); e = e.substr(0, slash); }
      auto it = env.find(e);
      if (it == env.end()) error("unbound variable " + e);
      p = InstructionParameter(sign * mul * it->second / div);
    }
    c->setParameters(ps);
    out->addInstruction(c);
  }
  return out;
}

// ------------------------------------------------------------------ noise models
// The interface the reference's purified-MPS visitor reads its noise from
// (ExaTnPmpsVisitor.cpp:470-487, :775-829): per gate, Kraus channels on some qubits; per
// qubit, the readout errors (P(measure 0 | prepared 1), P(measure 1 | prepared 0)).
struct NoiseChannelKraus {
  std::vector<size_t> noise_qubits;
  std::vector<std::vector<std::vector<std::complex<double>>>> mats;  // Kraus operators, row-major
};

class NoiseModel {
 public:
  virtual ~NoiseModel() = default;
  virtual std::vector<NoiseChannelKraus> getNoiseChannels(quantum::Gate& gate) = 0;
  virtual std::pair<double, double> readoutError(size_t qubit) = 0;
};

// ------------------------------------------------------------------ services
class Accelerator {
 public:
  virtual ~Accelerator() = default;
  virtual const std::string name() const = 0;
  virtual const std::string description() const { return ""; }
  virtual void initialize(const HeterogeneousMap& params = {}) = 0;
  virtual void updateConfiguration(const HeterogeneousMap& config) = 0;
  virtual void execute(std::shared_ptr<AcceleratorBuffer> buffer, const std::shared_ptr<CompositeInstruction> kernel) = 0;
  virtual void execute(std::shared_ptr<AcceleratorBuffer> buffer,
                       const std::vector<std::shared_ptr<CompositeInstruction>> kernels) = 0;
  virtual HeterogeneousMap getExecutionInfo() const { return {}; }
};

class IRTransformation {
 public:
  virtual ~IRTransformation() = default;
  virtual void apply(std::shared_ptr<CompositeInstruction> program, const std::shared_ptr<Accelerator> acc,
                     const HeterogeneousMap& options = {}) = 0;
  virtual const std::string name() const = 0;
};

struct Registry {
  std::map<std::type_index, std::map<std::string, std::shared_ptr<void>>> services;
  static Registry& get() {
    static Registry r;
    return r;
  }
};

template <typename T>
void registerService(const std::string& name, std::shared_ptr<T> svc) {
  Registry::get().services[std::type_index(typeid(T))][name] = std::static_pointer_cast<void>(svc);
}
template <typename T>
std::shared_ptr<T> getService(const std::string& name) {
  auto& m = Registry::get().services[std::type_index(typeid(T))];
  auto it = m.find(name);
  if (it == m.end()) error("no service named " + name);
  return std::static_pointer_cast<T>(it->second);
}
template <typename T>
bool hasService(const std::string& name) {
  auto& m = Registry::get().services[std::type_index(typeid(T))];
  return m.count(name) > 0;
}
template <typename T>
std::vector<std::shared_ptr<T>> getServices() {
  std::vector<std::shared_ptr<T>> out;
  for (auto& kv : Registry::get().services[std::type_index(typeid(T))]) out.push_back(std::static_pointer_cast<T>(kv.second));
  return out;
}

std::shared_ptr<Accelerator> getAccelerator(const std::string& name, const HeterogeneousMap& params = {});

inline std::shared_ptr<AcceleratorBuffer> qalloc(int n) { return std::make_shared<AcceleratorBuffer>("q", n); }

struct IRProvider {
  std::shared_ptr<Instruction> createInstruction(const std::string& name, const std::vector<size_t>& bits,
                                                 const std::vector<InstructionParameter>& params = {}) {
    return createGate(name, bits, params);
  }
  std::shared_ptr<CompositeInstruction> createComposite(const std::string& name) {
    return std::make_shared<CompositeInstruction>(name);
  }
};
inline std::shared_ptr<IRProvider> getIRProvider(const std::string&) { return std::make_shared<IRProvider>(); }

// Minimal XASM front end: "__qpu__ void f(qbit q, double t, ...) { G(q[i], ..., expr); ... }"
// Gate arguments are qubits q[k] then numbers (with pi, +-*/) or declared variables.
struct IR {
  std::vector<std::shared_ptr<CompositeInstruction>> composites;
  std::shared_ptr<CompositeInstruction> getComposite(const std::string& n) const {
    for (auto& c : composites)
      if (c->name() == n) return c;
    error("no composite " + n);
    return nullptr;
  }
  std::vector<std::shared_ptr<CompositeInstruction>> getComposites() const { return composites; }
};

struct XasmCompiler {
  std::shared_ptr<IR> compile(const std::string& src, std::shared_ptr<Accelerator> = nullptr);
};
inline std::shared_ptr<XasmCompiler> getCompiler(const std::string&) { return std::make_shared<XasmCompiler>(); }

inline std::vector<double> linspace(double a, double b, size_t n) {
  std::vector<double> v(n);
  for (size_t i = 0; i < n; ++i) v[i] = a + (b - a) * (double)i / (double)(n - 1);
  return v;
}
namespace constants {
constexpr double pi = 3.141592653589793238462643383279502884;
}

inline void Initialize() {}
inline void Finalize() {}

}  // namespace xacc
