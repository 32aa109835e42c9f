// visitor_tests.cpp -- the hip-mps visitor through the TNQVM accelerator API.
//
// Cases follow the exatn-mps gtest suite (tnqvm/visitors/exatn-mps/tests/*.cpp) and the
// closed-form checks of the other visitors' tests, run against "hip-mps":
//   ./visitor_tests --cpu    IR / routing / option cases (no GPU)
//   ./visitor_tests --gpu    circuit cases on the GPU
// Exit status 0 when every selected case passes.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <map>
#include <random>
#include <string>
#include <vector>

#include "TNQVM.hpp"

using namespace xacc;

// The CPU oracle (oracle/mps_oracle.cpp, the checker) for the sampling parity cases
extern "C" {
int orc_gate_matrix(const char* name, const double* params, int nparams, double* out);
void orc_sv_init(int n, double* psi);
void orc_sv_apply_1q(int n, double* psi, int q, const double* m);
void orc_sv_apply_2q(int n, double* psi, int q0, int q1, const double* m);
int orc_sv_sample(int n, const double* psi, uint64_t seed, int shots, const int* meas, int k, char* out);
void* orc_mps_create(int n, double cutoff, int max_bond, int pn_l2, int nthreads);
void orc_mps_destroy(void* h);
int orc_mps_apply_1q(void* h, int q, const double* m);
int orc_mps_apply_2q(void* h, int q0, int q1, const double* m);
int orc_mps_measure_samples(void* h, uint64_t seed, int shots, const int* meas, int k, char* out);
}

namespace {
int g_fail = 0, g_checks = 0;
std::string g_case;

#define EXPECT_NEAR(a, b, tol)                                                                  \
  do {                                                                                          \
    ++g_checks;                                                                                 \
    double _a = (a), _b = (b);                                                                  \
    if (!(std::fabs(_a - _b) <= (tol))) {                                                       \
      ++g_fail;                                                                                 \
      std::printf("  FAIL %s:%d [%s] %s = %.12g vs %.12g (tol %g)\n", __FILE__, __LINE__,        \
                  g_case.c_str(), #a, _a, _b, (double)(tol));                                    \
    }                                                                                           \
  } while (0)
#define EXPECT_TRUE(c)                                                                          \
  do {                                                                                          \
    ++g_checks;                                                                                 \
    if (!(c)) {                                                                                 \
      ++g_fail;                                                                                 \
      std::printf("  FAIL %s:%d [%s] %s\n", __FILE__, __LINE__, g_case.c_str(), #c);           \
    }                                                                                           \
  } while (0)
#define EXPECT_EQ(a, b) EXPECT_TRUE((a) == (b))

struct Case {
  const char* name;
  bool gpu;
  std::function<void()> fn;
};
std::vector<Case>& cases() {
  static std::vector<Case> c;
  return c;
}
struct Reg {
  Reg(const char* n, bool gpu, std::function<void()> f) { cases().push_back({n, gpu, std::move(f)}); }
};
#define CASE(name, gpu) \
  static void name();   \
  static Reg reg_##name(#name, gpu, name); \
  static void name()

std::shared_ptr<CompositeInstruction> compile1(const std::string& src) {
  return getCompiler("xasm")->compile(src)->getComposites()[0];
}
std::shared_ptr<Accelerator> mps(std::initializer_list<std::pair<std::string, std::any>> extra = {}) {
  HeterogeneousMap m{{"tnqvm-visitor", std::string("hip-mps")}};
  for (auto& kv : extra) {
    if (kv.second.type() == typeid(int)) m.insert(kv.first, std::any_cast<int>(kv.second));
    else if (kv.second.type() == typeid(double)) m.insert(kv.first, std::any_cast<double>(kv.second));
    else if (kv.second.type() == typeid(bool)) m.insert(kv.first, std::any_cast<bool>(kv.second));
    else if (kv.second.type() == typeid(std::string)) m.insert(kv.first, std::any_cast<std::string>(kv.second));
    else if (kv.second.type() == typeid(const char*)) m.insert(kv.first, std::string(std::any_cast<const char*>(kv.second)));
    else if (kv.second.type() == typeid(std::vector<int>)) m.insert(kv.first, std::any_cast<std::vector<int>>(kv.second));
  }
  return getAccelerator("tnqvm", m);
}
int count_swaps(const std::shared_ptr<CompositeInstruction>& f) {
  int n = 0;
  for (auto& i : f->getInstructions()) n += i->name() == "Swap";
  return n;
}

// ------------------------------------------------------------------ CPU cases
CASE(xasm_parse, false) {
  auto f = compile1(R"(__qpu__ void t(qbit q, double th) {
      H(q[0]);
      for (int i = 0; i < 3; i++) { CNOT(q[i], q[i + 1]); }
      U(q[1], th, -pi/2, pi/2);
      Rx(q[2], 2*th);
      Measure(q[0]);
  })");
  EXPECT_EQ(f->nInstructions(), (size_t)7);
  auto g = (*f)({0.25});
  EXPECT_NEAR(g->getInstruction(4)->getParameter(0).as<double>(), 0.25, 0);
  EXPECT_NEAR(g->getInstruction(4)->getParameter(1).as<double>(), -constants::pi / 2, 1e-15);
  EXPECT_NEAR(g->getInstruction(5)->getParameter(0).as<double>(), 0.5, 0);
  EXPECT_EQ(g->getInstruction(3)->bits()[1], (size_t)3);
}

// NearestNeighborTransformTester.cpp:24-64
CASE(lnn_transform_simple, false) {
  auto f = compile1(R"(__qpu__ void test1(qbit q) {
        X(q[0]);
        CNOT(q[1], q[3]);
        H(q[7]);
        CNOT(q[6], q[7]);
        CNOT(q[7], q[3]);
        X(q[6]);
        Swap(q[4], q[9]);
        S(q[5]);
    })");
  auto opt = getService<IRTransformation>("lnn-transform");
  const auto before = f->nInstructions();
  const int sw0 = count_swaps(f);
  opt->apply(f, nullptr);
  const int added = count_swaps(f) - sw0;
  EXPECT_EQ(f->nInstructions(), before + added);
  for (auto& i : f->getInstructions())
    if (i->bits().size() == 2) EXPECT_EQ(std::labs((long)i->bits()[0] - (long)i->bits()[1]), 1L);
}

// NearestNeighborTransformTester.cpp:67-100
CASE(lnn_transform_max_distance, false) {
  auto f = compile1(R"(__qpu__ void test2(qbit q) {
        X(q[0]);
        CNOT(q[1], q[9]);
        H(q[7]);
        CNOT(q[3], q[1]);
    })");
  auto opt = getService<IRTransformation>("lnn-transform");
  const auto before = f->nInstructions();
  const int sw0 = count_swaps(f);
  HeterogeneousMap o;
  o.insert("max-distance", 2);
  opt->apply(f, nullptr, o);
  const int added = count_swaps(f) - sw0;
  EXPECT_EQ(f->nInstructions(), before + added);
  auto last = f->getInstruction(f->nInstructions() - 1);
  EXPECT_EQ(last->name(), std::string("CNOT"));
  EXPECT_EQ(last->bits()[0], (size_t)3);
  EXPECT_EQ(last->bits()[1], (size_t)1);
  for (auto& i : f->getInstructions())
    if (i->bits().size() == 2) EXPECT_TRUE(std::labs((long)i->bits()[0] - (long)i->bits()[1]) <= 2);
}

CASE(unknown_visitor_falls_back, false) {
  auto acc = getAccelerator("tnqvm", {{"tnqvm-visitor", std::string("no-such-visitor")}});
  EXPECT_EQ(std::static_pointer_cast<tnqvm::TNQVM>(acc)->getVisitorName(), std::string("hip-mps"));
  bool threw = false;
  try {
    getAccelerator("tnqvm", {{"shots", 0}});
  } catch (const std::exception&) {
    threw = true;
  }
  EXPECT_TRUE(threw);  // TNQVM.hpp:103-108
}

// ------------------------------------------------------------------ GPU cases
// MpsGateTester.cpp:7-88
CASE(gate_simple, true) {
  const char* names[4] = {"1000", "0100", "0010", "0001"};
  for (int k = 0; k < 4; ++k) {
    auto f = compile1("__qpu__ void t(qbit q) { H(q[" + std::to_string(k) +
                      "]); Measure(q[0]); Measure(q[1]); Measure(q[2]); Measure(q[3]); }");
    auto acc = mps({{"shots", 10000}});
    auto q = qalloc(4);
    acc->execute(q, f);
    EXPECT_NEAR(q->computeMeasurementProbability("0000"), 0.5, 0.05);
    EXPECT_NEAR(q->computeMeasurementProbability(names[k]), 0.5, 0.05);
  }
}

// MpsGateTester.cpp:90-218
CASE(gate_two_qubit, true) {
  struct T { const char* src; int n; const char* a; double pa; const char* b; double pb; };
  std::vector<T> ts = {
      {"H(q[0]); CNOT(q[0], q[1]); Measure(q[0]); Measure(q[1]);", 4, "00", 0.5, "11", 0.5},
      {"H(q[1]); CNOT(q[1], q[0]); Measure(q[0]); Measure(q[1]);", 4, "00", 0.5, "11", 0.5},
      {"H(q[2]); CNOT(q[2], q[3]); Measure(q[2]); Measure(q[3]);", 4, "00", 0.5, "11", 0.5},
      {"H(q[3]); CNOT(q[3], q[2]); Measure(q[2]); Measure(q[3]);", 4, "00", 0.5, "11", 0.5},
      {"X(q[0]); CNOT(q[0], q[1]); CNOT(q[1], q[2]); CNOT(q[2], q[3]); Measure(q[0]); Measure(q[1]); Measure(q[2]); Measure(q[3]);",
       4, "1111", 1.0, "0000", 0.0},
  };
  for (auto& t : ts) {
    g_case = t.src;
    auto acc = mps({{"shots", 10000}});
    auto q = qalloc(t.n);
    acc->execute(q, compile1(std::string("__qpu__ void f(qbit q) {") + t.src + "}"));
    EXPECT_NEAR(q->computeMeasurementProbability(t.a), t.pa, 0.05);
    EXPECT_NEAR(q->computeMeasurementProbability(t.b), t.pb, 0.05);
  }
}

// MpsGateTester.cpp:220-240 (non-neighbouring CNOT, routed by the visitor)
CASE(gate_distance, true) {
  auto acc = mps({{"shots", 10000}});
  auto q = qalloc(4);
  acc->execute(q, compile1("__qpu__ void f(qbit q) { X(q[0]); CNOT(q[0], q[3]); Measure(q[0]); Measure(q[1]); Measure(q[2]); Measure(q[3]); }"));
  EXPECT_NEAR(q->computeMeasurementProbability("1001"), 1.0, 0.01);
  // exact amplitude of the routed result
  auto acc2 = mps({{"bitstring", std::vector<int>{1, 0, 0, 1}}});
  auto q2 = qalloc(4);
  acc2->execute(q2, compile1("__qpu__ void f(qbit q) { X(q[0]); CNOT(q[0], q[3]); }"));
  EXPECT_NEAR((*q2)["amplitude-real"].as<double>(), 1.0, 1e-12);
}

// MpsGateTester.cpp:242-310: 2- and 1-qubit registers
CASE(gate_small_registers, true) {
  {
    auto acc = mps({{"shots", 10000}});
    auto q = qalloc(2);
    acc->execute(q, compile1("__qpu__ void f(qbit q) { X(q[0]); CNOT(q[0], q[1]); Measure(q[0]); Measure(q[1]); }"));
    EXPECT_NEAR(q->computeMeasurementProbability("11"), 1.0, 0.01);
  }
  {
    auto acc = mps({{"shots", 10000}});
    auto q = qalloc(2);
    acc->execute(q, compile1("__qpu__ void f(qbit q) { X(q[1]); CNOT(q[1], q[0]); Measure(q[0]); Measure(q[1]); }"));
    EXPECT_NEAR(q->computeMeasurementProbability("11"), 1.0, 0.01);
  }
  {
    auto acc = mps({{"shots", 10000}});
    auto q = qalloc(1);
    acc->execute(q, compile1("__qpu__ void f(qbit q) { H(q[0]); Z(q[0]); H(q[0]); Measure(q[0]); }"));
    EXPECT_NEAR(q->computeMeasurementProbability("1"), 1.0, 0.01);
  }
}

// MpsGateTester.cpp:312-357
CASE(gate_iswap_fsim, true) {
  {
    auto acc = mps({{"shots", 100}});
    auto q = qalloc(3);
    acc->execute(q, compile1("__qpu__ void f(qbit q) { X(q[0]); iSwap(q[0], q[1]); Measure(q[0]); Measure(q[1]); Measure(q[2]); }"));
    EXPECT_NEAR(q->computeMeasurementProbability("010"), 1.0, 0.01);
  }
  auto prog = compile1("__qpu__ void f(qbit q, double x, double y) { X(q[0]); fSim(q[0], q[1], x, y); Measure(q[0]); Measure(q[1]); }");
  for (double a : linspace(-constants::pi, constants::pi, 10)) {
    auto acc = mps({{"shots", 1000}});
    auto q = qalloc(2);
    acc->execute(q, (*prog)({a, 0.0}));
    EXPECT_NEAR(q->computeMeasurementProbability("01"), std::sin(a) * std::sin(a), 0.1);
    EXPECT_NEAR(q->computeMeasurementProbability("10"), 1.0 - std::sin(a) * std::sin(a), 0.1);
    // exact: <Z0 Z1> = -1 for every angle (one excitation)
    auto acc2 = mps();
    auto q2 = qalloc(2);
    acc2->execute(q2, (*prog)({a, 0.3}));
    EXPECT_NEAR(q2->getExpectationValueZ(), -1.0, 1e-12);
  }
}

// MpsGateTester.cpp:359-407 (table over all angles, exact exp-val-z instead of shots)
CASE(deuteron, true) {
  const std::vector<double> expected{0.0,      -0.324699, -0.614213, -0.837166, -0.9694,  -0.996584, -0.915773,
                                     -0.735724, -0.475947, -0.164595, 0.164595,  0.475947, 0.735724,  0.915773,
                                     0.996584,  0.9694,    0.837166,  0.614213,  0.324699, 0.0};
  auto prog = compile1(R"(__qpu__ void ansatz(qbit q, double t) {
      X(q[0]); Ry(q[1], t); CX(q[1], q[0]); H(q[0]); H(q[1]); Measure(q[0]); Measure(q[1]); })");
  auto angles = linspace(-constants::pi, constants::pi, 20);
  for (size_t i = 0; i < angles.size(); ++i) {
    auto acc = mps();
    auto q = qalloc(2);
    acc->execute(q, (*prog)({angles[i]}));
    EXPECT_NEAR(q->getExpectationValueZ(), expected[i], 1e-6);
  }
}

// MpsGateTester.cpp:409-499
CASE(grover, true) {
  auto acc = mps({{"shots", 10000}});
  auto q = qalloc(3);
  acc->execute(q, compile1(R"(__qpu__ void g(qbit q) {
      H(q[0]); H(q[1]); H(q[2]); X(q[0]); H(q[2]); H(q[2]); CNOT(q[1], q[2]); Tdg(q[2]); CNOT(q[0], q[2]);
      T(q[2]); CNOT(q[1], q[2]); Tdg(q[2]); CNOT(q[0], q[2]); T(q[2]); H(q[2]); T(q[1]); CNOT(q[0], q[1]);
      T(q[0]); Tdg(q[1]); CNOT(q[0], q[1]); X(q[0]); H(q[2]); H(q[0]); H(q[1]); H(q[2]); X(q[0]); X(q[1]);
      X(q[2]); H(q[2]); H(q[2]); CNOT(q[1], q[2]); Tdg(q[2]); CNOT(q[0], q[2]); T(q[2]); CNOT(q[1], q[2]);
      Tdg(q[2]); CNOT(q[0], q[2]); T(q[2]); H(q[2]); T(q[1]); CNOT(q[0], q[1]); T(q[0]); Tdg(q[1]);
      CNOT(q[0], q[1]); H(q[2]); X(q[0]); X(q[1]); X(q[2]); H(q[0]); H(q[1]); H(q[2]);
      Measure(q[2]); Measure(q[1]); Measure(q[0]); })"));
  EXPECT_TRUE(q->computeMeasurementProbability("110") > 0.5);
}

// MpsMeasurementTester.cpp:7-35: 35-qubit GHZ sampled through the MPS (n >= 20 path)
CASE(measure_ghz35, true) {
  auto acc = mps({{"shots", 64}});
  auto q = qalloc(35);
  acc->execute(q, compile1(R"(__qpu__ void t(qbit q) {
      H(q[0]);
      for (int i = 0; i < 34; i++) { CNOT(q[i], q[i + 1]); }
      Measure(q[5]); Measure(q[3]); Measure(q[7]); Measure(q[34]); })"));
  EXPECT_NEAR(q->computeMeasurementProbability("0000") + q->computeMeasurementProbability("1111"), 1.0, 1e-12);
}

// ---- sampling parity with the oracle under the same seed ------------------------------------
// A seeded line circuit of U(theta, phi, lambda) layers and CNOT brickwork as XASM
std::string line_circuit(int n, int depth, uint64_t seed, const std::vector<int>& meas) {
  std::mt19937_64 rng(seed);
  std::uniform_real_distribution<double> ang(0.0, 2.0 * constants::pi);
  std::string src = "__qpu__ void c(qbit q) {\n";
  char buf[160];
  for (int l = 0; l < depth; ++l) {
    for (int q = 0; q < n; ++q) {
      std::snprintf(buf, sizeof(buf), "  U(q[%d], %.17g, %.17g, %.17g);\n", q, ang(rng), ang(rng), ang(rng));
      src += buf;
    }
    for (int q = l % 2; q + 1 < n; q += 2) {
      std::snprintf(buf, sizeof(buf), "  CNOT(q[%d], q[%d]);\n", q, q + 1);
      src += buf;
    }
  }
  for (int q : meas) {
    std::snprintf(buf, sizeof(buf), "  Measure(q[%d]);\n", q);
    src += buf;
  }
  return src + "}";
}

// gates of a composite through the oracle's matrices: apply1(q, m), apply2(q0, q1, m)
template <class F1, class F2>
void run_on_oracle(const std::shared_ptr<CompositeInstruction>& f, F1 apply1, F2 apply2) {
  for (auto& g : f->getInstructions()) {
    if (g->name() == "Measure" || g->name() == "I") continue;
    std::vector<double> p;
    for (auto& x : g->getParameters()) p.push_back(x.as<double>());
    double m[32];
    const int nq = orc_gate_matrix(g->name().c_str(), p.data(), (int)p.size(), m);
    if (nq == 1) apply1((int)g->bits()[0], m);
    else apply2((int)g->bits()[0], (int)g->bits()[1], m);
  }
}

std::map<std::string, int> count(const std::vector<char>& raw, int shots, int k) {
  std::map<std::string, int> c;
  for (int t = 0; t < shots; ++t) c[std::string(raw.data() + (size_t)t * k, (size_t)k)]++;
  return c;
}

// n < 20: GenerateSamples over the state vector (GateMatrixAlgebra.hpp:125-156)
CASE(sample_parity_state_vector, true) {
  const int n = 10, shots = 3000;
  const std::vector<int> meas = {7, 2, 9, 0, 4};
  auto f = compile1(line_circuit(n, 6, 11, meas));
  auto acc = mps({{"shots", shots}, {"seed", 2024}});
  auto q = qalloc(n);
  acc->execute(q, f);
  std::vector<double> psi(2 * ((size_t)1 << n));
  orc_sv_init(n, psi.data());
  run_on_oracle(f, [&](int a, const double* m) { orc_sv_apply_1q(n, psi.data(), a, m); },
                [&](int a, int b, const double* m) { orc_sv_apply_2q(n, psi.data(), a, b, m); });
  std::vector<char> raw((size_t)shots * meas.size());
  EXPECT_EQ(orc_sv_sample(n, psi.data(), 2024, shots, meas.data(), (int)meas.size(), raw.data()), shots);
  EXPECT_TRUE(q->getMeasurementCounts() == count(raw, shots, (int)meas.size()));
}

// n >= 20: getMeasureSample (ExaTnMpsVisitor.cpp:2211-2364), measured qubits in increasing order
// (every shot advanced together on the device) and in a scrambled order (shot by shot), untruncated
// and with max-bond-dim 8 (norm below 1)
CASE(sample_parity_sequential, true) {
  const int n = 22, shots = 150;
  const std::vector<std::vector<int>> orders = {{0, 3, 4, 9, 15, 21}, {15, 3, 21, 0}};
  for (int cap : {0, 8}) {
    for (const auto& meas : orders) {
      auto f = compile1(line_circuit(n, 5, 12, meas));
      auto acc = cap ? mps({{"shots", shots}, {"seed", 77}, {"max-bond-dim", cap}}) : mps({{"shots", shots}, {"seed", 77}});
      auto q = qalloc(n);
      acc->execute(q, f);
      void* o = orc_mps_create(n, 2.2250738585072014e-308, cap, 0, 1);
      run_on_oracle(f, [&](int a, const double* m) { orc_mps_apply_1q(o, a, m); },
                    [&](int a, int b, const double* m) { orc_mps_apply_2q(o, a, b, m); });
      std::vector<char> raw((size_t)shots * meas.size());
      orc_mps_measure_samples(o, 77, shots, meas.data(), (int)meas.size(), raw.data());
      orc_mps_destroy(o);
      EXPECT_TRUE(q->getMeasurementCounts() == count(raw, shots, (int)meas.size()));
    }
  }
}

// MpsMeasurementTester.cpp:37-66
CASE(measure_seed_determinism, true) {
  std::vector<std::map<std::string, int>> results;
  for (int i = 0; i < 20; ++i) {
    auto acc = mps({{"shots", 8192}, {"seed", 123}});
    auto q = qalloc(4);
    acc->execute(q, compile1(R"(__qpu__ void t(qbit q) {
        H(q[0]); for (int i = 0; i < 3; i++) { CNOT(q[i], q[i + 1]); }
        Measure(q[0]); Measure(q[1]); Measure(q[2]); Measure(q[3]); })"));
    results.push_back(q->getMeasurementCounts());
  }
  for (auto& r : results) EXPECT_TRUE(r == results[0]);
}

// NumericalTesterCheckNorm.cpp:15-62, with a seeded layer circuit instead of "rcs"
CASE(norm_check, true) {
  std::mt19937_64 rng(7);
  std::uniform_int_distribution<int> pick(0, 7);
  std::uniform_real_distribution<double> ang(-constants::pi, constants::pi);
  auto prov = getIRProvider("quantum");
  auto c = prov->createComposite("rcs");
  const int n = 10;
  for (int layer = 0; layer < 15; ++layer) {
    for (size_t q = 0; q < (size_t)n; ++q) {
      const char* g1[8] = {"H", "X", "Y", "Z", "T", "Rx", "Ry", "Rz"};
      int k = pick(rng);
      if (k < 5) c->addInstruction(prov->createInstruction(g1[k], {q}));
      else c->addInstruction(prov->createInstruction(g1[k], {q}, {InstructionParameter(ang(rng))}));
    }
    for (size_t q = 0; q + 1 < (size_t)n; ++q) c->addInstruction(prov->createInstruction("CNOT", {q, q + 1}));
  }
  for (double cut : {0.0, 1e-16}) {
    auto acc = cut == 0.0 ? mps({{"shots", 1}}) : mps({{"shots", 1}, {"svd-cutoff", 1e-16}});
    auto q = qalloc(n);
    acc->execute(q, c);
    EXPECT_NEAR((*q)["norm"].as<double>(), 1.0, 1e-10);
  }
}

// Failure detection: an SVD that reaches the Jacobi sweep cap is flagged on the buffer
// ("svd-converged" 0, with the count) instead of passing silently; the same circuit under the
// default cap converges.  The reference has no such guard beyond its post-SVD norm check
// (ExaTnMpsVisitor.cpp:1634-1661).
CASE(svd_sweep_cap_flagged, true) {
  std::mt19937_64 rng(11);
  std::uniform_real_distribution<double> ang(-constants::pi, constants::pi);
  auto prov = getIRProvider("quantum");
  auto c = prov->createComposite("cap");
  const int n = 8;
  for (int layer = 0; layer < 8; ++layer) {
    for (size_t q = 0; q < (size_t)n; ++q)
      c->addInstruction(prov->createInstruction("U", {q}, {InstructionParameter(ang(rng)), InstructionParameter(ang(rng)),
                                                           InstructionParameter(ang(rng))}));
    for (size_t q = layer % 2; q + 1 < (size_t)n; q += 2) c->addInstruction(prov->createInstruction("CNOT", {q, q + 1}));
  }
  auto capped = qalloc(n);
  mps({{"jacobi-max-sweeps", 1}})->execute(capped, c);
  EXPECT_EQ((*capped)["svd-converged"].as<int>(), 0);
  EXPECT_TRUE((*capped)["svd-nonconverged-count"].as<int>() > 0);
  auto plain = qalloc(n);
  mps()->execute(plain, c);
  EXPECT_EQ((*plain)["svd-converged"].as<int>(), 1);
  EXPECT_EQ((*plain)["svd-nonconverged-count"].as<int>(), 0);
}

// ITensorMPSVisitorTester.cpp:310-392 closed forms, exact
CASE(rotation_kats, true) {
  auto rx = compile1("__qpu__ void r(qbit q, double t) { Rx(q[0], t); Measure(q[0]); }");
  auto u3 = compile1("__qpu__ void r(qbit q, double t) { U(q[0], t, -pi/2, pi/2); Measure(q[0]); }");
  auto cp = compile1("__qpu__ void r(qbit q, double t) { X(q[0]); H(q[1]); CPhase(q[0], q[1], t); H(q[1]); Measure(q[1]); }");
  for (double a : linspace(-constants::pi, constants::pi, 20)) {
    const double want = 1.0 - 2.0 * std::sin(a / 2) * std::sin(a / 2);
    for (auto& p : {rx, u3}) {
      auto q = qalloc(1);
      mps()->execute(q, (*p)({a}));
      EXPECT_NEAR(q->getExpectationValueZ(), want, 1e-12);
    }
    auto q = qalloc(2);
    mps()->execute(q, (*cp)({a}));
    EXPECT_NEAR(q->getExpectationValueZ(), want, 1e-12);
  }
}

// amplitudes (ITensorMPSVisitorTester.cpp:522-551; bitstring_amplitude example)
CASE(amplitudes, true) {
  auto q = qalloc(3);
  mps({{"bitstring", std::vector<int>{1, 1, 1}}})->execute(q, compile1("__qpu__ void g(qbit q) { H(q[0]); Z(q[0]); CX(q[0], q[1]); CX(q[1], q[2]); }"));
  EXPECT_NEAR((*q)["amplitude-real"].as<double>(), -1.0 / std::sqrt(2.0), 1e-12);
  EXPECT_NEAR((*q)["amplitude-imag"].as<double>(), 0.0, 1e-12);
  // slice over all qubits of an n-qubit GHZ: 1/sqrt2 at 0 and 2^n - 1 (ExaTnGenTester.cpp:188-218)
  auto q2 = qalloc(5);
  mps({{"bitstring", std::vector<int>{-1, -1, -1, -1, -1}}})->execute(
      q2, compile1("__qpu__ void g(qbit q) { H(q[0]); for (int i = 0; i < 4; i++) { CX(q[i], q[i + 1]); } }"));
  auto re = (*q2)["amplitude-real-vec"].as<std::vector<double>>();
  EXPECT_EQ(re.size(), (size_t)32);
  EXPECT_NEAR(re[0], 1.0 / std::sqrt(2.0), 1e-12);
  EXPECT_NEAR(re[31], 1.0 / std::sqrt(2.0), 1e-12);
}

// VQE mode: the shared ansatz once, observables on that state == each kernel alone
CASE(vqe_mode, true) {
  auto prov = getIRProvider("quantum");
  auto build = [&](const std::string& basis) {
    auto c = prov->createComposite("k_" + basis);
    c->addInstruction(prov->createInstruction("X", {0}));
    c->addInstruction(prov->createInstruction("Ry", {1}, {InstructionParameter(0.59)}));
    c->addInstruction(prov->createInstruction("CNOT", {1, 0}));
    c->addInstruction(prov->createInstruction("Rz", {2}, {InstructionParameter(0.3)}));
    c->addInstruction(prov->createInstruction("CNOT", {1, 2}));
    if (basis == "XX") {
      c->addInstruction(prov->createInstruction("H", {0}));
      c->addInstruction(prov->createInstruction("H", {1}));
    }
    if (basis == "YY") {
      c->addInstruction(prov->createInstruction("Rx", {0}, {InstructionParameter(constants::pi / 2)}));
      c->addInstruction(prov->createInstruction("Rx", {1}, {InstructionParameter(constants::pi / 2)}));
    }
    if (basis == "CZ") c->addInstruction(prov->createInstruction("CNOT", {0, 2}));
    c->addInstruction(prov->createInstruction("Measure", {0}));
    c->addInstruction(prov->createInstruction("Measure", {1}));
    return c;
  };
  std::vector<std::shared_ptr<CompositeInstruction>> ks = {build("ZZ"), build("XX"), build("YY"), build("CZ")};
  auto qv = qalloc(3);
  mps({{"vqe-mode", true}})->execute(qv, ks);
  auto qn = qalloc(3);
  mps({{"vqe-mode", false}})->execute(qn, ks);
  auto cv = qv->getChildren(), cn = qn->getChildren();
  EXPECT_EQ(cv.size(), (size_t)4);
  for (size_t i = 0; i < cv.size() && i < cn.size(); ++i)
    EXPECT_NEAR(cv[i]->getExpectationValueZ(), cn[i]->getExpectationValueZ(), 1e-12);
}

CASE(accelerator_state, true) {
  auto acc = std::static_pointer_cast<tnqvm::TNQVM>(mps());
  auto sv = acc->getAcceleratorState(compile1("__qpu__ void b(qbit q) { H(q[0]); CNOT(q[0], q[1]); }"));
  EXPECT_EQ(sv.size(), (size_t)4);
  EXPECT_NEAR(std::abs(sv[0]), 1.0 / std::sqrt(2.0), 1e-14);
  EXPECT_NEAR(std::abs(sv[3]), 1.0 / std::sqrt(2.0), 1e-14);
}

// ------------------------------------------------------------------ hip-pmps (noisy)
// ExaTnPmpsTester.cpp cases.  The reference builds its model from the ibmqx2 backend JSON
// through XACC; here a model with the same structure (amplitude damping and dephasing after
// every non-virtual gate, readout errors) is given directly as the noise-model option.
class DampingModel : public xacc::NoiseModel {
 public:
  DampingModel(double gamma, double lam, double ro01, double ro10) : g_(gamma), l_(lam), ro_{ro01, ro10} {}
  std::vector<xacc::NoiseChannelKraus> getNoiseChannels(xacc::quantum::Gate& gate) override {
    std::vector<xacc::NoiseChannelKraus> out;
    if (gate.name() == "Rz" || gate.name() == "Measure") return out;
    using C = std::complex<double>;
    for (size_t q : gate.bits()) {
      xacc::NoiseChannelKraus ad, pd;
      ad.noise_qubits = {q};
      ad.mats = {{{C(1), C(0)}, {C(0), C(std::sqrt(1 - g_))}}, {{C(0), C(std::sqrt(g_))}, {C(0), C(0)}}};
      pd.noise_qubits = {q};
      pd.mats = {{{C(1), C(0)}, {C(0), C(std::sqrt(1 - l_))}}, {{C(0), C(0)}, {C(0), C(std::sqrt(l_))}}};
      out.push_back(ad);
      out.push_back(pd);
    }
    return out;
  }
  std::pair<double, double> readoutError(size_t) override { return ro_; }

 private:
  double g_, l_;
  std::pair<double, double> ro_;
};

std::shared_ptr<Accelerator> pmps(xacc::NoiseModel* model, int shots = 1024) {
  HeterogeneousMap m{{"tnqvm-visitor", std::string("hip-pmps")}};
  if (model) m.insert("noise-model", model);
  m.insert("shots", shots);
  m.insert("seed", 7);
  return getAccelerator("tnqvm", m);
}

// ExaTnPmpsTester.cpp:20-42: 100 X gates with damping leave a visible "1" population
CASE(pmps_check_simple, true) {
  DampingModel model(0.002, 0.001, 0.05, 0.02);
  auto f = compile1(R"(__qpu__ void test1(qbit q) {
    for (int i = 0; i < 100; i++) { X(q[0]); }
    Measure(q[0]);
  })");
  auto buffer = std::make_shared<AcceleratorBuffer>(1);
  pmps(&model)->execute(buffer, f);
  EXPECT_TRUE(buffer->computeMeasurementProbability("1") > 0.01);
  // X^100 = I; damping after each X: P(1) = sum over the alternation, exactly from rho
  auto diag = (*buffer)["density-matrix-diagonal"].as<std::vector<double>>();
  EXPECT_TRUE(diag[1] > 0.01 && diag[1] < 0.5);
  EXPECT_NEAR(diag[0] + diag[1], 1.0, 1e-12);
}

// :44-67 and :69-89: |111> and a Bell pair, most likely but not certain under noise
CASE(pmps_check_multiple_qubits_and_bell, true) {
  DampingModel model(0.01, 0.005, 0.03, 0.01);
  auto f = compile1(R"(__qpu__ void test2(qbit q) {
    X(q[0]); X(q[1]); X(q[2]);
    Measure(q[0]); Measure(q[1]); Measure(q[2]);
  })");
  auto b3 = std::make_shared<AcceleratorBuffer>(3);
  pmps(&model)->execute(b3, f);
  const double p111 = b3->computeMeasurementProbability("111");
  EXPECT_TRUE(p111 > 0.8 && p111 < 0.99);
  auto g = compile1(R"(__qpu__ void testBell(qbit q) {
    X(q[0]); CX(q[0], q[1]);
    Measure(q[0]); Measure(q[1]);
  })");
  auto b2 = std::make_shared<AcceleratorBuffer>(2);
  pmps(&model)->execute(b2, g);
  const double p11 = b2->computeMeasurementProbability("11");
  EXPECT_TRUE(p11 > 0.8 && p11 < 0.99);
  // without readout errors and noise, the distribution is exact
  auto b2c = std::make_shared<AcceleratorBuffer>(2);
  pmps(nullptr)->execute(b2c, g);
  auto d = (*b2c)["density-matrix-diagonal"].as<std::vector<double>>();
  EXPECT_NEAR(d[3], 1.0, 1e-12);
  // and the density matrix is |11><11| (flattened, ket index fastest)
  auto rho = (*b2c)["density_matrix"].as<std::vector<std::pair<double, double>>>();
  EXPECT_EQ(rho.size(), (size_t)16);
  EXPECT_NEAR(rho[3 + 4 * 3].first, 1.0, 1e-12);
  // a Bell pair: <00|rho|11> = 1/2
  auto h = compile1(R"(__qpu__ void bell2(qbit q) { H(q[0]); CX(q[0], q[1]); Measure(q[0]); Measure(q[1]); })");
  auto bb = std::make_shared<AcceleratorBuffer>(2);
  pmps(nullptr)->execute(bb, h);
  auto rb = (*bb)["density_matrix"].as<std::vector<std::pair<double, double>>>();
  EXPECT_NEAR(rb[0 + 4 * 3].first, 0.5, 1e-12);
  EXPECT_NEAR(rb[3 + 4 * 0].first, 0.5, 1e-12);
  EXPECT_NEAR(rb[1 + 4 * 1].first, 0.0, 1e-12);
}

// :91-115: a 100-qubit GHZ chain, 6 of its qubits measured
CASE(pmps_check_many_qubits, true) {
  auto f = compile1(R"(__qpu__ void testGHz(qbit q) {
    H(q[0]);
    for (int i = 0; i < 99; i++) { CX(q[i], q[i + 1]); }
    Measure(q[0]); Measure(q[6]); Measure(q[8]); Measure(q[36]); Measure(q[72]); Measure(q[98]);
  })");
  auto buffer = std::make_shared<AcceleratorBuffer>(100);
  pmps(nullptr)->execute(buffer, f);
  EXPECT_NEAR(buffer->computeMeasurementProbability("111111"), 0.5, 0.1);
  EXPECT_NEAR(buffer->computeMeasurementProbability("000000"), 0.5, 0.1);
  auto d = (*buffer)["density-matrix-diagonal"].as<std::vector<double>>();
  EXPECT_NEAR(d[0], 0.5, 1e-10);
  EXPECT_NEAR(d[63], 0.5, 1e-10);
  // n > 15: the density matrix of the measured qubits; a proper subset of GHZ qubits has no
  // coherence left, diag(1/2, 0, ..., 0, 1/2)
  auto rho = (*buffer)["density_matrix"].as<std::vector<std::pair<double, double>>>();
  EXPECT_EQ(rho.size(), (size_t)4096);
  EXPECT_NEAR(rho[0].first, 0.5, 1e-10);
  EXPECT_NEAR(rho[63 + 64 * 63].first, 0.5, 1e-10);
  EXPECT_NEAR(rho[0 + 64 * 63].first, 0.0, 1e-10);
}

// a routed gate: lnn-transform swaps are visited, and carry noise like any other gate
CASE(pmps_routed_gate_noise, true) {
  DampingModel model(0.05, 0.0, 0.0, 0.0);
  auto f = compile1(R"(__qpu__ void r(qbit q) { X(q[0]); CX(q[0], q[3]); Measure(q[3]); })");
  auto b = std::make_shared<AcceleratorBuffer>(4);
  pmps(&model)->execute(b, f);
  auto d = (*b)["density-matrix-diagonal"].as<std::vector<double>>();
  EXPECT_TRUE(d[1] < 0.95 && d[1] > 0.5);
  auto kd = (*b)["kraus-dims"].as<std::vector<int>>();
  EXPECT_TRUE(kd[3] > 1);
  // n <= 15: the full density matrix of all 4 qubits (ExaTnPmpsVisitor.cpp:664-683); its
  // diagonal summed over qubits 0..2 is the distribution of the measured qubit 3
  auto rho = (*b)["density_matrix"].as<std::vector<std::pair<double, double>>>();
  EXPECT_EQ(rho.size(), (size_t)256);
  double p1 = 0.0, tr = 0.0;
  for (size_t x = 0; x < 16; ++x) {
    tr += rho[x + 16 * x].first;
    if (x & 8) p1 += rho[x + 16 * x].first;
  }
  EXPECT_NEAR(p1, d[1], 1e-12);
  EXPECT_NEAR(tr, (*b)["norm"].as<double>(), 1e-10);
}

}  // namespace

int main(int argc, char** argv) {
  bool want_gpu = false, want_cpu = false;
  const char* filter = nullptr;
  for (int i = 1; i < argc; ++i) {
    if (!strcmp(argv[i], "--gpu")) want_gpu = true;
    else if (!strcmp(argv[i], "--cpu")) want_cpu = true;
    else filter = argv[i];
  }
  if (!want_gpu && !want_cpu) want_cpu = true;
  int ran = 0;
  for (auto& c : cases()) {
    if ((c.gpu && !want_gpu) || (!c.gpu && !want_cpu)) continue;
    if (filter && !strstr(c.name, filter)) continue;
    g_case = c.name;
    const int before = g_fail;
    std::printf("[case] %s\n", c.name);
    std::fflush(stdout);
    try {
      c.fn();
    } catch (const std::exception& e) {
      ++g_fail;
      std::printf("  FAIL [%s] exception: %s\n", c.name, e.what());
    }
    std::printf("[%s] %s\n", g_fail == before ? "ok" : "FAILED", c.name);
    std::fflush(stdout);
    ++ran;
  }
  std::printf("%d cases, %d checks, %d failures\n", ran, g_checks, g_fail);
  return g_fail ? 1 : 0;
}
