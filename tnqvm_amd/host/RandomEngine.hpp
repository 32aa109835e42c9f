// RandomEngine.hpp -- process-wide seeded random source of the TNQVM driver.
//
// Same interface and draws as tnqvm/utils/RandomEngine.hpp:8-45: a std::mt19937_64
// seeded by the "seed" option (random_device otherwise), uniform_real_distribution(0, 1)
// per draw, and sortedRandProbs for inverse-CDF sampling.
#pragma once

#include <algorithm>
#include <mutex>
#include <random>
#include <vector>

namespace tnqvm {

class randomEngine {
 public:
  randomEngine(const randomEngine&) = delete;
  randomEngine& operator=(const randomEngine&) = delete;
  static randomEngine& get_instance() {
    static randomEngine inst;
    return inst;
  }
  void setSeed(size_t seed) {
    std::lock_guard<std::mutex> lk(m_);
    eng_.seed(seed);
  }
  double randProb() {
    std::lock_guard<std::mutex> lk(m_);
    return std::uniform_real_distribution<double>(0.0, 1.0)(eng_);
  }
  std::vector<double> sortedRandProbs(uint64_t n) {
    std::vector<double> r;
    r.reserve(n);
    std::lock_guard<std::mutex> lk(m_);
    for (uint64_t i = 0; i < n; ++i) r.push_back(std::uniform_real_distribution<double>(0.0, 1.0)(eng_));
    std::sort(r.begin(), r.end());
    return r;
  }

 private:
  randomEngine() {
    std::random_device rd;
    eng_.seed(rd());
  }
  std::mt19937_64 eng_;
  std::mutex m_;
};

}  // namespace tnqvm
