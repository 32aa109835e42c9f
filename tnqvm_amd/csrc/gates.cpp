// gates.cpp -- gate matrices of the hip-mps visitor (host side).
//
// Same matrices and conventions as tnqvm/base/Gates.hpp:131-339 and the flattening of
// GateTensorConstructor::getGateTensor (tnqvm/visitors/exatn-mps/ExatnUtils.cpp:36-133):
// row-major, and for 2-qubit gates the first operand is the most significant bit of the
// row/column index.  S and Sdg, which Gates.hpp lacks (the exatn-mps visitor does not
// override them), get their standard matrices here instead of being dropped.
#include <cmath>
#include <complex>
#include <cstring>
#include <string>

#include "hipmps.h"

namespace {
using cd = std::complex<double>;

int fill(const std::string& g, const double* p, int np, cd* M) {
  const cd I(0.0, 1.0);
  auto two = [&](cd a, cd b, cd c, cd d) {
    M[0] = a; M[1] = b; M[2] = c; M[3] = d;
    return 1;
  };
  auto four_diag = [&](cd a, cd b, cd c, cd d) {
    for (int i = 0; i < 16; ++i) M[i] = 0.0;
    M[0] = a; M[5] = b; M[10] = c; M[15] = d;
    return 2;
  };
  const double h = std::sqrt(0.5);
  if (g == "I") return two(1, 0, 0, 1);
  if (g == "H") return two(h, h, h, -h);
  if (g == "X") return two(0, 1, 1, 0);
  if (g == "Y") return two(0, -I, I, 0);
  if (g == "Z") return two(1, 0, 0, -1);
  if (g == "S") return two(1, 0, 0, I);
  if (g == "Sdg") return two(1, 0, 0, -I);
  if (g == "T") return two(1, 0, 0, std::exp(I * (M_PI / 4)));
  if (g == "Tdg") return two(1, 0, 0, std::exp(-I * (M_PI / 4)));
  if (g == "Rx" && np >= 1) {
    const double c = std::cos(p[0] / 2), s = std::sin(p[0] / 2);
    return two(c, -I * s, -I * s, c);
  }
  if (g == "Ry" && np >= 1) {
    const double c = std::cos(p[0] / 2), s = std::sin(p[0] / 2);
    return two(c, -s, s, c);
  }
  if (g == "Rz" && np >= 1) return two(std::exp(-I * (p[0] / 2)), 0, 0, std::exp(I * (p[0] / 2)));
  if ((g == "U" || g == "U3") && np >= 3) {
    const double c = std::cos(p[0] / 2), s = std::sin(p[0] / 2);
    return two(c, -std::exp(I * p[2]) * s, std::exp(I * p[1]) * s, std::exp(I * (p[1] + p[2])) * c);
  }
  if (g == "CNOT" || g == "CX") {
    four_diag(1, 1, 0, 0);
    M[11] = 1;
    M[14] = 1;
    return 2;
  }
  if (g == "CZ") return four_diag(1, 1, 1, -1);
  if (g == "CY") {
    four_diag(1, 1, 0, 0);
    M[11] = -I;
    M[14] = I;
    return 2;
  }
  if (g == "CH") {
    four_diag(1, 1, h, -h);
    M[11] = h;
    M[14] = h;
    return 2;
  }
  if (g == "CRZ" && np >= 1) return four_diag(1, 1, std::exp(-I * (p[0] / 2)), std::exp(I * (p[0] / 2)));
  if (g == "CPhase" && np >= 1) return four_diag(1, 1, 1, std::exp(I * p[0]));
  if (g == "Swap") {
    for (int i = 0; i < 16; ++i) M[i] = 0.0;
    M[0] = 1; M[6] = 1; M[9] = 1; M[15] = 1;
    return 2;
  }
  if (g == "iSwap") {
    for (int i = 0; i < 16; ++i) M[i] = 0.0;
    M[0] = 1; M[6] = I; M[9] = I; M[15] = 1;
    return 2;
  }
  if (g == "fSim" && np >= 2) {
    for (int i = 0; i < 16; ++i) M[i] = 0.0;
    M[0] = 1;
    M[5] = std::cos(p[0]);
    M[6] = -I * std::sin(p[0]);
    M[9] = -I * std::sin(p[0]);
    M[10] = std::cos(p[0]);
    M[15] = std::exp(-I * p[1]);
    return 2;
  }
  return 0;
}
}  // namespace

extern "C" int hipmps_gate_matrix(const char* name, const double* params, int n_params, double* out) {
  if (!name || !out || (n_params > 0 && !params)) return HIPMPS_E_ARG;
  cd M[16];
  const int k = fill(name, params, n_params, M);
  if (k == 0) return HIPMPS_E_ARG;
  const int cnt = k == 1 ? 4 : 16;
  for (int i = 0; i < cnt; ++i) {
    out[2 * i] = M[i].real();
    out[2 * i + 1] = M[i].imag();
  }
  return k;
}
