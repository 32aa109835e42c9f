// device_util.h -- device helpers shared by the libhipmps kernel files (internal).
#pragma once
#include <hip/hip_runtime.h>

namespace hipmps {

// ------------------------------------------------------------------ complex helpers
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cconj(double2 a) { return make_double2(a.x, -a.y); }
__device__ __forceinline__ double2 cscale(double2 a, double s) { return make_double2(a.x * s, a.y * s); }
// c + a*b
__device__ __forceinline__ double2 cfma(double2 a, double2 b, double2 c) {
  c.x = fma(a.x, b.x, c.x);
  c.x = fma(-a.y, b.y, c.x);
  c.y = fma(a.x, b.y, c.y);
  c.y = fma(a.y, b.x, c.y);
  return c;
}
// c + conj(a)*b
__device__ __forceinline__ double2 cfmac(double2 a, double2 b, double2 c) {
  c.x = fma(a.x, b.x, c.x);
  c.x = fma(a.y, b.y, c.x);
  c.y = fma(a.x, b.y, c.y);
  c.y = fma(-a.y, b.x, c.y);
  return c;
}
__device__ __forceinline__ double cabs2(double2 a) { return a.x * a.x + a.y * a.y; }

template <typename T>
__device__ __forceinline__ double2 ld(const T* p, size_t i);
template <>
__device__ __forceinline__ double2 ld<double2>(const double2* p, size_t i) { return p[i]; }
template <>
__device__ __forceinline__ double2 ld<float2>(const float2* p, size_t i) {
  float2 v = p[i];
  return make_double2((double)v.x, (double)v.y);
}
template <typename T>
__device__ __forceinline__ void st(T* p, size_t i, double2 v);
template <>
__device__ __forceinline__ void st<double2>(double2* p, size_t i, double2 v) { p[i] = v; }
template <>
__device__ __forceinline__ void st<float2>(float2* p, size_t i, double2 v) {
  p[i] = make_float2((float)v.x, (float)v.y);
}

// Largest t with begin(t) <= b, for task arrays sorted by their first block.
template <typename Task, typename F>
__device__ __forceinline__ int find_task(const Task* tasks, int ntasks, int b, F begin) {
  int lo = 0, hi = ntasks - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (begin(tasks[mid]) <= b) lo = mid; else hi = mid - 1;
  }
  return lo;
}


// FP64 MFMA v_mfma_f64_16x16x4f64 operand layout on gfx950 (measured):
//   A[i][k] in lane i + 16k, B[k][j] in lane j + 16k, D[i][j] in lane j + 16*(i%4),
//   accumulator register i/4.
typedef double mfma_d4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ mfma_d4 mfma64(double a, double b, mfma_d4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

}  // namespace hipmps
