// engine.hip -- host side of libhipmps: engine, MPS state, wave scheduler, batched
// gate application, measurement sweeps and the continuous-batching serving loop.
//
// Reference mapping (tnqvm/visitors/exatn-mps/ExaTnMpsVisitor.cpp):
//   state create / reset ....... initialize (:173-345): |0..0>, all bonds 1
//   apply_1q ................... applyGate (:1185-1292)
//   apply_2q + wave execution .. applyTwoQubitGate (:1387-1730) + truncateSvdTensors
//                                (:2366-2536)
//   norm / expval_z_parity ..... finalize (:593-643)
//   amplitudes / slice ......... computeWaveFuncSlice (:2588-2675) and the slice
//                                normalisation (:794-820, done by the caller)
//
// Execution model.  Every 2-qubit gate becomes a task of a "wave": a set of gates on
// pairwise disjoint site pairs, possibly from many circuits.  One wave runs as a
// fixed sequence of batched launches on the engine stream:
//   k_theta_mfma -> k_jacobi_prep -> (k_jacobi_round x rounds) x sweeps -> k_rownorm ->
//   k_select -> k_gather + k_gemm
// so the device sees all independent SVDs of the wave at once.  1-qubit gates are
// fused into the next 2-qubit gate on the same qubit (U_2q . (U_a x U_b)); whatever is
// left over is applied by k_apply_1q when the state is flushed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <atomic>
#include <condition_variable>
#include <cfloat>
#include <chrono>
#include <climits>
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "hipmps.h"
#include "kernels.h"

using cd = std::complex<double>;
using namespace hipmps;

// Static device scratch of hipmps_amplitudes_noalloc: the chain vectors, the bit strings and the
// chain descriptors of one call.  Every GPU the library runs on holds its own copy from module
// load to process end.  Host code never passes this name to a kernel or a copy: it is a host-side
// stand-in, and the device address of each GPU's copy comes from hipGetSymbolAddress, taken once
// per device when an engine is created on it (engine_scratch_setup).
static __device__ __align__(256) char scratch[HIPMPS_NOALLOC_SCRATCH_BYTES];

// Test hook (HIPMPS_TEST_STAGE_DELAY_US): one wave that sleeps for `ticks` of the constant-rate
// wall clock, queued on a wave's stream ahead of its descriptor upload, so that the upload runs
// late, as it does when the stream waits behind other work on a shared hardware queue.
__global__ void k_stage_delay(long long ticks) {
  const long long t0 = wall_clock64();
  while (wall_clock64() - t0 < ticks) __builtin_amdgcn_s_sleep(8);
}

namespace {

thread_local std::string t_err;

int fail(int code, const std::string& msg) {
  t_err = msg;
  return code;
}

#define HIP_TRY(expr)                                                                          \
  do {                                                                                         \
    hipError_t _e = (expr);                                                                    \
    if (_e != hipSuccess) return fail(HIPMPS_E_DEVICE, std::string(#expr) + ": " + hipGetErrorString(_e)); \
  } while (0)

#define RC_TRY(expr)          \
  do {                        \
    int _rc = (expr);         \
    if (_rc != 0) return _rc; \
  } while (0)

constexpr size_t ALIGN = 256;
inline size_t align_up(size_t x) { return (x + ALIGN - 1) / ALIGN * ALIGN; }

struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  int ensure(size_t need) {
    if (need <= bytes) return 0;
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
    size_t nb = std::max(need, (size_t)1 << 20);
    nb = std::max(nb, need + need / 4);
    if (hipMalloc(&p, nb) != hipSuccess) {
      p = nullptr;
      (void)hipGetLastError();
      return fail(HIPMPS_E_NOMEM, "device workspace allocation of " + std::to_string(nb) + " bytes failed");
    }
    bytes = nb;
    return 0;
  }
  void release() {
    if (p) (void)hipFree(p);
    p = nullptr;
    bytes = 0;
  }
};

struct HostBuf {
  void* p = nullptr;
  size_t bytes = 0;
  int ensure(size_t need) {
    if (need <= bytes) return 0;
    if (p) (void)hipHostFree(p);
    p = nullptr;
    size_t nb = std::max(need, (size_t)1 << 16);
    if (hipHostMalloc(&p, nb, hipHostMallocDefault) != hipSuccess) {
      p = nullptr;
      bytes = 0;
      (void)hipGetLastError();
      return fail(HIPMPS_E_NOMEM, "pinned host allocation failed");
    }
    bytes = nb;
    return 0;
  }
  void release() {
    if (p) (void)hipHostFree(p);
    p = nullptr;
    bytes = 0;
  }
};

// 2x2 / 4x4 helpers (row-major)
void mat_mul4(const cd* A, const cd* B, cd* C) {
  for (int i = 0; i < 4; ++i)
    for (int j = 0; j < 4; ++j) {
      cd s = 0;
      for (int k = 0; k < 4; ++k) s += A[i * 4 + k] * B[k * 4 + j];
      C[i * 4 + j] = s;
    }
}
void mat_mul2(const cd* A, const cd* B, cd* C) {
  cd t[4];
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) t[i * 2 + j] = A[i * 2] * B[j] + A[i * 2 + 1] * B[2 + j];
  for (int i = 0; i < 4; ++i) C[i] = t[i];
}
void kron2(const cd* A, const cd* B, cd* K) {  // A (x) B, A on the most significant bit
  for (int a = 0; a < 2; ++a)
    for (int b = 0; b < 2; ++b)
      for (int a2 = 0; a2 < 2; ++a2)
        for (int b2 = 0; b2 < 2; ++b2) K[(a * 2 + b) * 4 + (a2 * 2 + b2)] = A[a * 2 + a2] * B[b * 2 + b2];
}

}  // namespace

// ======================================================================================
// engine / state
// ======================================================================================
struct Site {
  void* p = nullptr;
  size_t cap = 0;  // elements
};

struct hipmps_state_s;

struct PendingCircuit;

struct hipmps_engine_s {
  int device = 0;
  hipStream_t stream = nullptr;
  // streams for the Jacobi rounds: the wave's SVDs are dealt to `jacobi_lanes` lists whose
  // rounds run on their own streams (lane 0 on `stream`), so that the launch tail of one list
  // overlaps the work of the others
  static constexpr int MAX_LANES = 4;
  hipStream_t lane_stream[MAX_LANES] = {nullptr, nullptr, nullptr, nullptr};
  hipEvent_t lane_ev[MAX_LANES + 1] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  int jacobi_lanes = 2;
  int lane_min_pairs = 1536;  // two generations of the 768 resident Jacobi workgroups
  // mixed path: the first sweeps of SVDs with min(m, n) >= 256 form the Gram with FP32 MFMA;
  // such a sweep never counts as the converged one, so every SVD ends with FP64 sweeps.  The
  // FP64 path never does: rotation angles good to ~1e-7 mix large rows into small ones, which
  // costs one-sided Jacobi its relative accuracy on graded spectra (1e-12 absolute at 2048)
  int gram32_sweeps = 5;
  size_t budget = 0;
  size_t used = 0;  // bytes of site storage held by states
  int max_sweeps = 40;
  double tol = 1e-13;       // FP64 path: relative orthogonality of converged rows
  double tol_mixed = 1e-8;  // mixed path: factors are stored in complex64
  int inner_sweeps = 1;
  int debug_sweeps = 0;  // HIPMPS_DEBUG_SWEEPS: per-sweep rotated / visited block pairs on stderr
  int lq_min_rank = 64;  // Jacobi on the double-LQ factor when min(m, n) >= this (0: never)
  // LQ with look-ahead (next panel factored in LDS on one stream while the trailing rows update
  // on another) for rows of at most LQ_REG_MAX_COLS columns; HIPMPS_LQ_LOOKAHEAD=0 turns it off
  int lq_lookahead = 1;
  hipEvent_t lq_ev[2 * MAX_LANES] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  // completion of a wave's descriptor upload: the host edits the staged Jacobi descriptors
  // after enqueuing that copy, and the copy may still wait behind earlier work on the stream
  hipEvent_t stage_ev = nullptr;
  long long stage_delay_ticks = 0;  // test hook: the upload held back this long (wall-clock ticks)
  DevBuf ws;     // per-wave arena
  DevBuf desc;   // task descriptors
  DevBuf small;  // counters / K / stats
  DevBuf env;    // environments for measurements
  HostBuf hdesc;
  HostBuf hsmall;
  std::mutex mu;
  bool timing = false;
  hipmps_profile_t prof{};
  hipEvent_t ev[7] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  // serving
  uint64_t next_ticket = 1;
  std::deque<PendingCircuit*> queue;
  std::vector<PendingCircuit*> running;
  std::map<uint64_t, PendingCircuit*> by_ticket;
  // background lane of the serving loop: circuits with max-bond-dim >= big_min_chi run on an
  // engine of their own (stream, arena, lanes), whose waves a worker thread runs back to back
  // while the steps advance the other circuits; the worker alone touches that engine and its
  // circuits, and hands them back through `done`.  Off by default: the large circuit then
  // finishes sooner, but its launches slow every wave of the others (see big_lane_min_chi)
  int big_min_chi = 0;
  // density-matrix sweep: the per-site intermediate is held a block of (p, kappa) at a time when
  // it exceeds max(this many elements, the largest environment)
  int rdm_block_elems = 1 << 24;
  // device address of this device's copy of `scratch` (hipGetSymbolAddress at creation)
  void* scratch = nullptr;
  hipmps_engine_s* big = nullptr;
  std::thread big_worker;
  std::mutex big_mu;
  std::condition_variable big_cv;
  std::deque<PendingCircuit*> big_inbox;
  bool big_stop = false;
  int big_rc = 0;
  std::string big_err;
  // wave groups of batches and serving: the circuits are split over wave_groups engines (this
  // one and grp), whose waves advance concurrently, grp's on a worker thread, so that one
  // group's latency-bound phases (the LQ panels) overlap the other's Jacobi rounds.  A circuit
  // stays on the engine it started on; each engine alone touches its circuits.
  int wave_groups = 2;
  hipmps_engine_s* grp = nullptr;
  std::thread grp_worker;
  std::mutex grp_mu;
  std::condition_variable grp_cv;
  std::vector<PendingCircuit*>* grp_job = nullptr;  // posted by the step, cleared by the worker
  bool grp_stop = false;
  int grp_rc = 0;
  std::string grp_err;
};

struct hipmps_state_s {
  hipmps_engine e = nullptr;
  int n = 0;
  hipmps_trunc_opts t{};
  int fp32 = 0;
  size_t es = 16;
  std::vector<int> chi;
  std::vector<int> kap;  // Kraus leg of each site (1 everywhere for a pure state)
  std::vector<Site> sites;
  std::vector<std::array<cd, 4>> pend;
  std::vector<char> has_pend;
  struct Q2 {
    int i;
    cd G[16];
  };
  std::vector<Q2> wave;
  std::vector<char> busy;
  // noise channels queued after the open wave of gates, on distinct sites; they run as one
  // wave of compressions after the gates
  struct QK {
    int q, n_ops;
    cd K[4 * HIPMPS_MAX_KRAUS_OPS];
  };
  std::vector<QK> cwave;
  std::vector<char> cbusy;
  double discarded = 0.0;
  double fid = 1.0;
  int maxb = 1;
  int n2q = 0, n1q = 0;
  int sweeps_max = 0;
  long long sweeps_total = 0;
  size_t bytes = 0;
  int status = 0;
  int norm_warnings = 0;  // SVDs whose factor norm fell below 1e-3 (post-SVD validation)
  int nonconverged = 0;   // SVDs that reached the Jacobi sweep cap without converging
  size_t peak = 0;        // most site bytes held at once (old and new buffers during a growth)
  size_t wave_new = 0;    // scratch of grow_sites
  int kraus_ops = 0;      // noise channels applied
  std::vector<void*> deferred;
};

struct PendingCircuit {
  uint64_t ticket = 0;
  // finished: the lane that runs the circuit has written its results; done: published once
  // that lane no longer touches it (the caller may then poll and free it)
  bool finished = false;
  std::atomic<bool> done{false};
  bool big = false;  // runs on the background lane
  int grp = 0;       // wave group (engine) it runs on
  hipmps_circuit c{};
  std::vector<hipmps_gate> gates;
  std::vector<int> zz;
  hipmps_result* res = nullptr;
  hipmps_state_s* st = nullptr;
  size_t cursor = 0;
  size_t footprint = 0;
  std::chrono::steady_clock::time_point t0;
};

namespace {

// hipGetSymbolAddress of `scratch`, once per device (the device current), kept per device
constexpr int SCRATCH_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[SCRATCH_MAX_DEVICES] = {};

int engine_scratch_setup(hipmps_engine e) {
  if (e->device >= SCRATCH_MAX_DEVICES) return 0;  // the no-alloc entry point refuses such a device
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[e->device]) {
    void* p = nullptr;
    const hipError_t rc = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (rc != hipSuccess || p == nullptr) {
      (void)hipGetLastError();
      return fail(HIPMPS_E_DEVICE, std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(rc));
    }
    g_scratch_ptr[e->device] = p;
  }
  e->scratch = g_scratch_ptr[e->device];
  return 0;
}

int set_device(hipmps_engine e) {
  HIP_TRY(hipSetDevice(e->device));
  return 0;
}

// bond capacity bound: min(2^b, 2^(n-b), chi_max)
size_t bond_bound(int n, int b, int maxb) {
  int e = std::min(b, n - b);
  size_t v = e >= 30 ? ((size_t)1 << 30) : ((size_t)1 << e);
  return std::min(v, (size_t)maxb);
}

size_t projected_bytes(int n, int maxb, int fp32) {
  size_t es = fp32 ? 8 : 16, tot = 0;
  for (int k = 0; k < n; ++k) tot += bond_bound(n, k, maxb) * 2 * bond_bound(n, k + 1, maxb) * es;
  return tot;
}

// largest bond a circuit can reach: min(max-bond-dim, 2^floor(n/2))
int projected_max_bond(const hipmps_circuit& c) {
  const int half = c.n_qubits / 2;
  const long long full = half >= 30 ? (1LL << 30) : (1LL << half);
  const long long cap = c.trunc.max_bond_dim > 0 ? c.trunc.max_bond_dim : (1LL << 30);
  return (int)std::min(full, cap);
}

// HBM a circuit may hold while it runs, for admission: its sites at the largest bonds they can
// reach, twice (a growing site holds its old buffer until theta has read it), plus its share of
// the wave arena (theta', the Jacobi matrix and the LQ factors of up to n/2 gates of bond
// chi: about 6 (2 chi)^2 complex128 each, the tiled copy of the Jacobi matrix included)
size_t circuit_footprint(const hipmps_circuit& c) {
  const int maxb = projected_max_bond(c);
  const size_t sites = projected_bytes(c.n_qubits, maxb, c.trunc.precision == HIPMPS_MIXED);
  const size_t gates = (size_t)std::max(1, c.n_qubits / 2);
  return 2 * sites + gates * 6 * (size_t)(2 * maxb) * (size_t)(2 * maxb) * 16;
}

// capacity a pure site grows to: doubled, but never past the largest the site can ever hold
// (bonds bounded by min(2^k, 2^(n-k), max-bond-dim)), so that a state never holds more than
// its projected footprint
size_t grown_capacity(const hipmps_state_s* s, int k, size_t need, size_t cap) {
  size_t ncap = std::max(need, cap * 2);
  if (s->kap[k] == 1) {
    const int maxb = s->t.max_bond_dim > 0 ? s->t.max_bond_dim : INT_MAX - 1;
    const size_t bound = bond_bound(s->n, k, maxb) * 2 * bond_bound(s->n, k + 1, maxb);
    ncap = std::max(need, std::min(ncap, bound));
  }
  return ncap;
}

size_t site_elems(const hipmps_state_s* s, int k) { return (size_t)s->chi[k] * 2 * s->kap[k] * s->chi[k + 1]; }

bool is_purified(const hipmps_state_s* s) {
  for (int k : s->kap)
    if (k > 1) return true;
  return false;
}

int max_kap(const hipmps_state_s* s) {
  int m = 1;
  for (int k : s->kap) m = std::max(m, k);
  return m;
}

int ensure_site(hipmps_state_s* s, int k, size_t elems, bool keep_old_for_read, void** old_out) {
  Site& S = s->sites[k];
  if (old_out) *old_out = S.p;
  if (elems <= S.cap) return 0;
  size_t ncap = std::max(elems, S.cap * 2);
  size_t nbytes = ncap * s->es;
  hipmps_engine e = s->e;
  if (e->used + nbytes > e->budget)
    return fail(HIPMPS_E_NOMEM, "HBM budget exceeded growing site " + std::to_string(k));
  void* p = nullptr;
  if (hipMalloc(&p, nbytes) != hipSuccess) {
    (void)hipGetLastError();
    return fail(HIPMPS_E_NOMEM, "device allocation for site " + std::to_string(k) + " failed");
  }
  if (S.p) {
    if (keep_old_for_read) s->deferred.push_back(S.p);
    else (void)hipFree(S.p);
    e->used -= S.cap * s->es;
    s->bytes -= S.cap * s->es;
  }
  S.p = p;
  S.cap = ncap;
  e->used += nbytes;
  s->bytes += nbytes;
  s->peak = std::max(s->peak, s->bytes);
  return 0;
}

void release_deferred(hipmps_state_s* s) {
  for (void* p : s->deferred) (void)hipFree(p);
  s->deferred.clear();
}

// Site growth of a whole wave, in two phases: every new buffer is reserved (budget) and
// allocated first; only when all of them exist are they committed, the old buffers parked in
// `deferred` for theta to read.  On failure every new buffer is freed again and no state of the
// wave is touched, so the wave can be retried (or its states run apart) with their sites intact.
struct SiteGrowth {
  hipmps_state_s* s;
  int k;
  size_t ncap;  // elements
  void* p;
};

int grow_sites(hipmps_engine E, std::vector<SiteGrowth>& g) {
  size_t add = 0;
  for (auto& x : g) add += x.ncap * x.s->es;
  if (E->used + add > E->budget)
    return fail(HIPMPS_E_NOMEM, "HBM budget exceeded growing " + std::to_string(g.size()) + " site(s) by " +
                                    std::to_string(add) + " bytes");
  for (size_t a = 0; a < g.size(); ++a) {
    if (hipMalloc(&g[a].p, g[a].ncap * g[a].s->es) != hipSuccess) {
      (void)hipGetLastError();
      for (size_t b = 0; b < a; ++b) (void)hipFree(g[b].p);
      for (auto& x : g) x.p = nullptr;
      return fail(HIPMPS_E_NOMEM, "device allocation for site " + std::to_string(g[a].k) + " failed");
    }
  }
  for (auto& x : g) x.s->wave_new += x.ncap * x.s->es;
  for (auto& x : g) {
    if (x.s->wave_new) {  // held during the wave: the old buffers and the new ones
      x.s->peak = std::max(x.s->peak, x.s->bytes + x.s->wave_new);
      x.s->wave_new = 0;
    }
  }
  for (auto& x : g) {
    Site& S = x.s->sites[x.k];
    if (S.p) {
      x.s->deferred.push_back(S.p);
      E->used -= S.cap * x.s->es;
      x.s->bytes -= S.cap * x.s->es;
    }
    S.p = x.p;
    S.cap = x.ncap;
    E->used += x.ncap * x.s->es;
    x.s->bytes += x.ncap * x.s->es;
  }
  return 0;
}

int upload_site_host(hipmps_state_s* s, int k, const cd* data, size_t n) {
  hipmps_engine e = s->e;
  RC_TRY(e->hsmall.ensure(n * 16));
  if (s->fp32) {
    float* h = (float*)e->hsmall.p;
    for (size_t i = 0; i < n; ++i) {
      h[2 * i] = (float)data[i].real();
      h[2 * i + 1] = (float)data[i].imag();
    }
    HIP_TRY(hipMemcpyAsync(s->sites[k].p, h, n * 8, hipMemcpyHostToDevice, e->stream));
  } else {
    memcpy(e->hsmall.p, data, n * 16);
    HIP_TRY(hipMemcpyAsync(s->sites[k].p, e->hsmall.p, n * 16, hipMemcpyHostToDevice, e->stream));
  }
  HIP_TRY(hipStreamSynchronize(e->stream));
  return 0;
}

int state_init_product(hipmps_state_s* s) {
  // |0..0>: every site 1 x 2 x 1 with body {1, 0} (ExaTnMpsVisitor.cpp:46, :323)
  s->chi.assign(s->n + 1, 1);
  s->kap.assign(s->n, 1);
  const cd body[2] = {1.0, 0.0};
  for (int k = 0; k < s->n; ++k) {
    RC_TRY(ensure_site(s, k, 2, false, nullptr));
    RC_TRY(upload_site_host(s, k, body, 2));
  }
  s->pend.assign(s->n, {cd(1), cd(0), cd(0), cd(1)});
  s->has_pend.assign(s->n, 0);
  s->busy.assign(s->n, 0);
  s->cbusy.assign(s->n, 0);
  s->wave.clear();
  s->cwave.clear();
  s->discarded = 0.0;
  s->fid = 1.0;
  s->maxb = 1;
  s->n2q = s->n1q = 0;
  s->sweeps_max = 0;
  s->sweeps_total = 0;
  s->status = 0;
  s->norm_warnings = 0;
  s->nonconverged = 0;
  s->peak = s->bytes;
  s->kraus_ops = 0;
  return 0;
}

// ------------------------------------------------------------------------------------
// wave execution (many states, many bonds)
// ------------------------------------------------------------------------------------
struct WaveItem {
  hipmps_state_s* s;
  int i;          // left site
  const cd* G;    // 16 entries, local order (a = left-site bit)
  // derived
  int cl, cm, cr, m, n, r, c, r_pad, kub, transposed, fp32;
  void* old_i;
  void* old_j;
  size_t off_theta, off_y, off_vec;
  int lq;                            // Jacobi on the double-LQ factor R2 of X
  size_t off_z, off_j, off_yt, off_lq;
  int ki, kj;                        // Kraus legs of the two sites (fused into cl and cr)
  size_t off_pin, off_pout;          // left site with its Kraus leg moved next to the left bond
  size_t off_jt;                     // the Jacobi matrix in tiles during the sweeps
};

int run_wave(hipmps_engine E, std::vector<WaveItem>& items) {
  if (items.empty()) return 0;
  const hipStream_t st = E->stream;
  // order: fp64 before fp32, non-transposed before transposed (contiguous launch groups)
  for (auto& it : items) {
    hipmps_state_s* s = it.s;
    it.ki = s->kap[it.i];
    it.kj = s->kap[it.i + 1];
    it.cl = s->chi[it.i] * it.ki;
    it.cm = s->chi[it.i + 1];
    it.cr = it.kj * s->chi[it.i + 2];
    it.m = 2 * it.cl;
    it.n = 2 * it.cr;
    it.r = std::min(it.m, it.n);
    it.c = std::max(it.m, it.n);
    it.transposed = it.m > it.n;
    it.r_pad = (it.r + JP - 1) / JP * JP;
    int maxb = s->t.max_bond_dim > 0 ? s->t.max_bond_dim : INT_MAX - 1;
    it.kub = std::min(it.r, maxb);
    it.fp32 = s->fp32;
    it.lq = E->lq_min_rank > 0 && it.r >= E->lq_min_rank;
  }
  std::stable_sort(items.begin(), items.end(), [](const WaveItem& a, const WaveItem& b) {
    if (a.fp32 != b.fp32) return a.fp32 < b.fp32;
    return a.transposed < b.transposed;
  });
  const int N = (int)items.size();
  for (auto& it : items)
    if (it.r > 4096) return fail(HIPMPS_E_UNSUPPORTED, "bond dimension above 2048 is not supported");
  // arena
  size_t off = 0;
  for (auto& it : items) {
    it.off_theta = off;
    off = align_up(off + (size_t)it.m * it.n * 16);
    it.off_y = off;
    off = align_up(off + (size_t)it.r_pad * it.c * 16);
    it.off_vec = off;
    off = align_up(off + (size_t)it.r_pad * (8 * 5 + 4));
    it.off_jt = off;  // r_pad x (columns of the Jacobi matrix rounded up to 16)
    off = align_up(off + (size_t)it.r_pad * (((it.lq ? it.r : it.c) + 15) / 16 * 16) * 16);
    if (it.lq) {
      const size_t panels = (it.r + 15) / 16;  // LQ panels of 16 rows, T factor 16 x 16 each
      it.off_z = off;  // L^H, then the second LQ (r x r)
      off = align_up(off + (size_t)it.r * it.r * 16);
      it.off_j = off;  // R2, the Jacobi matrix (r_pad x r)
      off = align_up(off + (size_t)it.r_pad * it.r * 16);
      it.off_yt = off;  // [Y_J | 0] Q~ (r_pad x c)
      off = align_up(off + (size_t)it.r_pad * it.c * 16);
      it.off_lq = off;  // W0 and T of both factorizations
      off = align_up(off + 2 * ((size_t)it.r + panels * 256) * 16);
    }
    if (it.ki > 1) {
      const size_t es = it.fp32 ? 8 : 16;
      it.off_pin = off;  // old left site as [l][k][p][m]
      off = align_up(off + (size_t)it.cl * 2 * it.cm * es);
      it.off_pout = off;  // new left site as [l][k][p][K]
      off = align_up(off + (size_t)it.cl * 2 * it.kub * es);
    }
  }
  RC_TRY(E->ws.ensure(off));
  char* ws = (char*)E->ws.p;
  // small: counts (2N ints), K (N ints), wstats (WS_N doubles per item), ||X||_F^2 (N doubles)
  size_t small_bytes = align_up(3 * N * sizeof(int)) + align_up(WS_N * N * sizeof(double)) + N * sizeof(double);
  RC_TRY(E->small.ensure(small_bytes));
  int* d_counts = (int*)E->small.p;
  int* d_K = d_counts + 2 * N;
  double* d_w = (double*)((char*)E->small.p + align_up(3 * N * sizeof(int)));
  double* d_fro2 = (double*)((char*)d_w + align_up(WS_N * N * sizeof(double)));
  // descriptors
  size_t desc_bytes = align_up(N * sizeof(ThetaTask)) + align_up(N * sizeof(SplitTask)) +
                      align_up(N * sizeof(JacTask)) + align_up(N * sizeof(GemmTask)) +
                      align_up(2 * N * sizeof(LqTask)) + align_up(3 * N * sizeof(TriTask)) +
                      align_up(N * sizeof(double));
  RC_TRY(E->desc.ensure(desc_bytes));
  // host staging: descriptors | counts (2N ints) | active Jacobi list / results
  const size_t stage_bytes = desc_bytes + align_up(2 * N * sizeof(int)) + N * sizeof(JacTask) + WS_N * N * sizeof(double) + 4096;
  RC_TRY(E->hdesc.ensure(stage_bytes));
  // site capacity, after every engine buffer exists (new sites are written after theta has
  // read the old ones); all-or-nothing across the wave
  {
    std::vector<SiteGrowth> grow;
    for (auto& it : items) {
      const size_t need[2] = {(size_t)it.cl * 2 * it.kub, (size_t)it.kub * 2 * it.cr};
      for (int h = 0; h < 2; ++h) {
        Site& S = it.s->sites[it.i + h];
        (h ? it.old_j : it.old_i) = S.p;
        if (need[h] > S.cap) grow.push_back({it.s, it.i + h, grown_capacity(it.s, it.i + h, need[h], S.cap), nullptr});
      }
    }
    if (!grow.empty()) RC_TRY(grow_sites(E, grow));
  }
  char* hd = (char*)E->hdesc.p;
  ThetaTask* hth = (ThetaTask*)hd;
  SplitTask* hsp = (SplitTask*)(hd + align_up(N * sizeof(ThetaTask)));
  JacTask* hjt = (JacTask*)((char*)hsp + align_up(N * sizeof(SplitTask)));
  GemmTask* hgm = (GemmTask*)((char*)hjt + align_up(N * sizeof(JacTask)));
  LqTask* hlq = (LqTask*)((char*)hgm + align_up(N * sizeof(GemmTask)));
  TriTask* htr = (TriTask*)((char*)hlq + align_up(2 * N * sizeof(LqTask)));
  double* hcut = (double*)((char*)htr + align_up(3 * N * sizeof(TriTask)));  // rank cut, per LQ item
  char* dd = (char*)E->desc.p;
  ThetaTask* dth = (ThetaTask*)dd;
  SplitTask* dsp = (SplitTask*)(dd + align_up(N * sizeof(ThetaTask)));
  JacTask* djt = (JacTask*)((char*)dsp + align_up(N * sizeof(SplitTask)));
  GemmTask* dgm = (GemmTask*)((char*)djt + align_up(N * sizeof(JacTask)));
  LqTask* dlq = (LqTask*)((char*)dgm + align_up(N * sizeof(GemmTask)));
  TriTask* dtr = (TriTask*)((char*)dlq + align_up(2 * N * sizeof(LqTask)));
  const double* dcut = (const double*)((char*)dtr + align_up(3 * N * sizeof(TriTask)));

  // group boundaries
  int n64 = 0;
  while (n64 < N && !items[n64].fp32) ++n64;
  int max_r = 0, max_c = 0, max_rpad = 0, max_k = 0;
  int tile_acc[2] = {0, 0};
  for (int q = 0; q < N; ++q) {
    auto& it = items[q];
    const int grp = it.fp32;
    ThetaTask& T = hth[q];
    memset(&T, 0, sizeof(T));
    T.Ai = it.ki > 1 ? (void*)(ws + it.off_pin) : it.old_i;
    T.Aj = it.old_j;
    T.theta = (double2*)(ws + it.off_theta);
    T.cl = it.cl;
    T.cm = it.cm;
    T.cr = it.cr;
    T.tiles_l = (it.cl + 15) / 16;
    T.tiles_r = (it.cr + 31) / 32;
    T.tile_begin = tile_acc[grp];
    tile_acc[grp] += T.tiles_l * T.tiles_r;
    for (int k = 0; k < 16; ++k) T.G[k] = make_double2(it.G[k].real(), it.G[k].imag());

    SplitTask& S = hsp[q];
    memset(&S, 0, sizeof(S));
    S.theta = T.theta;
    S.Y = (double2*)(ws + (it.lq ? it.off_yt : it.off_y));
    S.m = it.m;
    S.n = it.n;
    S.r = it.r;
    S.c = it.c;
    S.ldy = it.c;
    S.transposed = it.transposed;
    char* v = ws + it.off_vec;
    S.rownorm = (double*)v;
    S.sig = (double*)(v + 8 * it.r_pad);
    S.scale_y = (double*)(v + 16 * it.r_pad);
    S.scale_w = (double*)(v + 24 * it.r_pad);
    S.inv_sig = (double*)(v + 32 * it.r_pad);
    S.perm = (int*)(v + 40 * it.r_pad);
    S.K = d_K + q;
    S.wstats = d_w + WS_N * q;
    S.Ai = it.ki > 1 ? (void*)(ws + it.off_pout) : it.s->sites[it.i].p;
    S.Aj = it.s->sites[it.i + 1].p;
    S.cl = it.cl;
    S.cr = it.cr;
    const hipmps_trunc_opts& to = it.s->t;
    S.cutoff = to.svd_cutoff;
    S.max_bond = to.max_bond_dim > 0 ? to.max_bond_dim : INT_MAX - 1;
    S.pn_l2 = to.partial_norm == HIPMPS_PN_L2;
    S.absorb = to.absorb == HIPMPS_ABSORB_RIGHT ? 1 : 0;
    S.renorm = to.renormalize ? 1 : 0;
    S.site_fp32 = it.fp32;

    JacTask& J = hjt[q];
    J.Y = (double2*)(ws + it.off_y);  // X; for LQ items replaced by R2 after the factorizations
    J.ldy = S.ldy;
    J.r = it.r;
    J.r_pad = it.r_pad;
    J.c = it.c;
    J.c_tot = it.c;
    J.nb = it.r_pad / JB;
    J.counter = d_counts + 2 * q;
    J.visited = d_counts + 2 * q + 1;
    J.fro2 = d_fro2 + q;
    J.tol = it.fp32 ? E->tol_mixed : E->tol;
    J.gram32 = 0;
    J.pad = 0;

    max_r = std::max(max_r, it.r);
    max_rpad = std::max(max_rpad, it.r_pad);
    max_c = std::max(max_c, it.c);
    max_k = std::max(max_k, it.kub);
  }
  // double-LQ preconditioning: X = L Q~, L^H = L_Z Q_Z, Jacobi on R2 = L_Z^H
  // LQ tasks of items whose rows fit the look-ahead LQ (c <= LQ_REG_MAX_COLS, so r too) come first
  // (nlq_a of them): the method is chosen per item, so that an item's result does not depend on
  // the other items of its wave
  int nlq = 0, nlq_a = 0, lq_max_r = 0, lq_max_rpad = 0, lq_max_c = 0;
  bool any_cut = false;
  std::vector<int> lq_order;
  for (int pass = 0; pass < 2; ++pass)
    for (int q = 0; q < N; ++q)
      if (items[q].lq && ((items[q].c <= LQ_REG_MAX_COLS) == (pass == 0))) lq_order.push_back(q);
  for (int q : lq_order) {
    auto& it = items[q];
    if (it.c <= LQ_REG_MAX_COLS) ++nlq_a;
    const size_t panels = (it.r + 15) / 16;
    double2* Y = (double2*)(ws + it.off_y);
    double2* Z = (double2*)(ws + it.off_z);
    double2* Jm = (double2*)(ws + it.off_j);
    double2* Yt = (double2*)(ws + it.off_yt);
    double2* lqb = (double2*)(ws + it.off_lq);
    LqTask& A = hlq[nlq];
    memset(&A, 0, sizeof(A));
    A.M = Y; A.ld = it.c; A.rows = it.r; A.cols = it.c;
    A.W0 = lqb; A.T = lqb + it.r;
    A.Tgt = Yt; A.ld_t = it.c; A.rows_t = it.r_pad;
    LqTask& B = hlq[N + nlq];
    memset(&B, 0, sizeof(B));
    B.M = Z; B.ld = it.r; B.rows = it.r; B.cols = it.r;
    B.W0 = lqb + it.r + panels * 256; B.T = B.W0 + it.r;
    htr[nlq] = TriTask{Y, Z, it.c, it.r, it.r, it.r};               // Z = L^H
    htr[N + nlq] = TriTask{Z, Jm, it.r, it.r, it.r_pad, it.r};      // J = L_Z^H, pad rows zero
    htr[2 * N + nlq] = TriTask{Jm, Yt, it.r, it.r, it.r_pad, it.c};  // Yt = [J | 0]
    // rank cut, for gates on purified sites, whose theta' has 2 chi kap rows but rank about chi
    // (a pure site's theta' is full rank in practice, and the cut would only read it): the
    // singular value below which the truncation drops an index (partial norm sigma, or
    // sqrt(sigma) with the 2-norm definition), far above the cut's perturbation
    if (it.ki > 1 || it.kj > 1) {
      const double cutoff = it.s->t.svd_cutoff;
      hcut[nlq] = it.s->t.partial_norm == HIPMPS_PN_L2 ? cutoff * cutoff : cutoff;
      any_cut = true;
    } else {
      hcut[nlq] = 0.0;
    }
    ++nlq;
    lq_max_r = std::max(lq_max_r, it.r);
    lq_max_rpad = std::max(lq_max_rpad, it.r_pad);
    lq_max_c = std::max(lq_max_c, it.c);
  }
  // GEMM recovering the second factor, one per item; tile_begin relative to its
  // (precision, transposed) range
  struct Range {
    int b, e, tiles;
  };
  std::vector<Range> granges;
  for (int q = 0; q < N;) {
    int e = q;
    while (e < N && items[e].fp32 == items[q].fp32 && items[e].transposed == items[q].transposed) ++e;
    int tiles = 0;
    for (int k = q; k < e; ++k) {
      auto& it = items[k];
      SplitTask& S = hsp[k];
      GemmTask& G = hgm[k];
      memset(&G, 0, sizeof(G));
      if (!it.transposed) {
        // Ai (m x K) = theta (m x n) . Ysel^H (n x K) . diag(1/sigma)
        G.A = S.theta; G.lda = it.n;
        G.B = S.Y; G.ldb = S.ldy; G.gatherB = S.perm;
        G.C = S.Ai; G.ldc = -1;
        G.M = it.m; G.N = it.kub; G.K = it.n;
        G.Ndyn = S.K;
        G.colscale = S.inv_sig;
      } else {
        // Aj (K x n) = diag(1/sigma) . Ysel (K x m) . theta (m x n)
        G.A = S.Y; G.lda = S.ldy; G.gatherA = S.perm;
        G.B = S.theta; G.ldb = it.n;
        G.C = S.Aj; G.ldc = it.n;
        G.M = it.kub; G.N = it.n; G.K = it.m;
        G.Mdyn = S.K;
        G.rowscale = S.inv_sig;
      }
      G.tiles_m = (G.M + 31) / 32;
      G.tiles_n = (G.N + 31) / 32;
      G.tile_begin = tiles;
      tiles += G.tiles_m * G.tiles_n;
    }
    granges.push_back({q, e, tiles});
    q = e;
  }
  if (E->stage_delay_ticks > 0) {
    hipLaunchKernelGGL(k_stage_delay, dim3(1), dim3(64), 0, st, E->stage_delay_ticks);
    HIP_TRY(hipGetLastError());
  }
  HIP_TRY(hipMemcpyAsync(dd, hd, desc_bytes, hipMemcpyHostToDevice, st));
  HIP_TRY(hipEventRecord(E->stage_ev, st));
  // the copy reads the staging buffer when the stream reaches it; the host edits the Jacobi
  // descriptors there (LQ items switch to R2) only once it has been read
  bool staged = false;
  auto wait_staged = [&]() -> int {
    if (!staged) HIP_TRY(hipEventSynchronize(E->stage_ev));
    staged = true;
    return 0;
  };

  if (E->timing) HIP_TRY(hipEventRecord(E->ev[0], st));
  // purified left sites: Kraus leg next to the left bond, [l][p][k][m] -> [l][k][p][m]
  for (auto& it : items)
    if (it.ki > 1)
      HIP_TRY(launch_swap_mid(it.old_i, ws + it.off_pin, it.s->chi[it.i], 2, it.ki, it.cm, nullptr, it.fp32, st));
  // ---- theta ----
  if (n64 > 0) HIP_TRY(launch_theta(dth, n64, tile_acc[0], 0, st));
  if (N - n64 > 0) HIP_TRY(launch_theta(dth + n64, N - n64, tile_acc[1], 1, st));
  HIP_TRY(launch_jacobi_prep(dsp, djt, N, max_rpad, max_c, st));
  HIP_TRY(launch_fro2(dsp, djt, N, st));
  if (E->timing) HIP_TRY(hipEventRecord(E->ev[6], st));
  // LQ lanes: with many items, contiguous slices of the LQ task arrays run their panel
  // sequences on separate streams, so that one slice's launch tails overlap another's work
  const int lq_lanes = nlq >= 64 ? std::max(1, std::min(E->jacobi_lanes, nlq)) : 1;
  auto lq_slice = [&](int l, int& b, int& n) {
    b = (int)((long long)nlq * l / lq_lanes);
    n = (int)((long long)nlq * (l + 1) / lq_lanes) - b;
  };
  auto fork_lanes = [&](int lanes) -> int {
    if (lanes < 2) return 0;
    HIP_TRY(hipEventRecord(E->lane_ev[0], st));
    for (int l = 1; l < lanes; ++l) HIP_TRY(hipStreamWaitEvent(E->lane_stream[l], E->lane_ev[0], 0));
    return 0;
  };
  auto join_lanes = [&](int lanes) -> int {
    for (int l = 1; l < lanes; ++l) {
      HIP_TRY(hipEventRecord(E->lane_ev[l], E->lane_stream[l]));
      HIP_TRY(hipStreamWaitEvent(st, E->lane_ev[l], 0));
    }
    return 0;
  };
  // one LQ of a lane's slice: with look-ahead on the lane's stream and a partner stream when the
  // rows fit (lanes 0, 1 pair with streams 2, 3), else panel by panel on the lane's stream
  const bool lookahead = E->lq_lookahead && 2 * lq_lanes <= hipmps_engine_s::MAX_LANES && E->lq_ev[0] != nullptr;
  // the LQ of tasks [b, b + n) of one array (second = 0: X, 1: L^H): look-ahead for the tasks of
  // items in the first group (columns <= LQ_REG_MAX_COLS), panel by panel for the others
  auto run_lq = [&](const LqTask* tasks, int b, int n, int second, int l) -> int {
    const hipStream_t ls = E->lane_stream[l];
    const int a_end = std::min(b + n, nlq_a), b_beg = std::max(b, nlq_a);
    const bool la = lookahead && E->lane_stream[l + lq_lanes] != nullptr;
    auto extent = [&](int from, int to, int& rows, int& cols) {
      rows = cols = 0;
      for (int k = from; k < to; ++k) {
        const auto& it = items[lq_order[k]];
        rows = std::max(rows, it.r);
        cols = std::max(cols, second ? it.r : it.c);
      }
    };
    int rows, cols;
    if (a_end > b) {
      extent(b, a_end, rows, cols);
      if (la) HIP_TRY(launch_lq_lookahead(tasks + b, a_end - b, rows, cols, ls, E->lane_stream[l + lq_lanes],
                                         E->lq_ev[2 * l], E->lq_ev[2 * l + 1]));
      else HIP_TRY(launch_lq(tasks + b, a_end - b, rows, ls));
    }
    if (b + n > b_beg) {
      extent(b_beg, b + n, rows, cols);
      HIP_TRY(launch_lq(tasks + b_beg, b + n - b_beg, rows, ls));
    }
    return 0;
  };
  if (nlq > 0) {
    RC_TRY(fork_lanes(lq_lanes));
    for (int l = 0; l < lq_lanes; ++l) {
      int b, n;
      lq_slice(l, b, n);
      const hipStream_t ls = E->lane_stream[l];
      RC_TRY(run_lq(dlq, b, n, 0, l));
      HIP_TRY(launch_lower_to_upper_h(dtr + b, n, lq_max_r, lq_max_r, ls));
      RC_TRY(run_lq(dlq + N, b, n, 1, l));
      HIP_TRY(launch_lower_to_upper_h(dtr + N + b, n, lq_max_r, lq_max_rpad, ls));
    }
    RC_TRY(join_lanes(lq_lanes));
    if (any_cut) HIP_TRY(launch_rank_cut(dtr + N, dcut, nlq, st));
    RC_TRY(wait_staged());
    for (int q = 0; q < N; ++q) {
      auto& it = items[q];
      if (!it.lq) continue;
      JacTask& J = hjt[q];
      J.Y = (double2*)(ws + it.off_j);
      J.ldy = it.r;
      J.c = it.r;
      J.c_tot = it.r;
    }
  }
  if (E->timing) HIP_TRY(hipEventRecord(E->ev[1], st));

  // ---- Jacobi sweeps, on a tiled copy of every Jacobi matrix ----
  int jt_blocks = 0, jt_ntc = 0;
  RC_TRY(wait_staged());
  for (int q = 0; q < N; ++q) {
    JacTask& J = hjt[q];
    J.Yt = (double2*)(ws + items[q].off_jt);
    J.ntc = (J.c_tot + 15) / 16;
    jt_blocks = std::max(jt_blocks, J.r_pad / 16);
    jt_ntc = std::max(jt_ntc, J.ntc);
  }
  HIP_TRY(hipMemcpyAsync(djt, hjt, N * sizeof(JacTask), hipMemcpyHostToDevice, st));
  HIP_TRY(launch_to_tiles(djt, N, jt_blocks, jt_ntc, st));
  std::vector<int> active(N);
  for (int q = 0; q < N; ++q) active[q] = q;
  std::vector<int> sweeps_used(N, E->max_sweeps);
  std::vector<char> converged(N, 0);
  // pairs each item rotated in the previous sweep: within a lane, the busiest items' block pairs
  // come first in the launch, so that the workgroups starting last (in the launch's tail) are
  // mostly the short ones of nearly converged items, which only form their Gram
  std::vector<long long> prev_rot(N, LLONG_MAX);
  int* hcounts = (int*)((char*)E->hdesc.p + desc_bytes);
  for (int sweep = 0; sweep < E->max_sweeps && !active.empty(); ++sweep) {
    // active task list with pair prefix; with several lanes, the items are dealt to lists
    // of about equal work (largest first), each a prefix-indexed launch on its own stream
    constexpr int ML = hipmps_engine_s::MAX_LANES;
    int nact = (int)active.size();
    JacTask* hja = (JacTask*)((char*)E->hdesc.p + desc_bytes + align_up(2 * N * sizeof(int)));
    // lanes only pay when one launch would already fill the chip twice over (768 resident
    // workgroups: 256 CUs x 3 at the kernel's register count); below that they add launches
    int total_pairs = 0;
    for (int q : active) total_pairs += hjt[q].nb / 2;
    const int lanes = total_pairs >= E->lane_min_pairs ? std::max(1, std::min(E->jacobi_lanes, nact)) : 1;
    std::vector<int> order(active);
    if (lanes > 1)
      std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return hjt[a].nb > hjt[b].nb; });
    std::vector<int> lane_items[ML];
    long long lane_load[ML] = {0, 0, 0, 0};
    for (int q : order) {
      int l = 0;
      for (int k = 1; k < lanes; ++k)
        if (lane_load[k] < lane_load[l]) l = k;
      lane_items[l].push_back(q);
      lane_load[l] += (long long)(hjt[q].nb / 2) * hjt[q].c * (hjt[q].nb - 1);
    }
    for (int l = 0; l < lanes; ++l)
      std::stable_sort(lane_items[l].begin(), lane_items[l].end(),
                       [&](int a, int b) { return prev_rot[a] > prev_rot[b]; });
    int lane_begin[ML] = {0, 0, 0, 0}, lane_n[ML] = {0, 0, 0, 0}, lane_pairs[ML] = {0, 0, 0, 0},
        lane_rounds[ML] = {0, 0, 0, 0};
    {
      int a = 0;
      for (int l = 0; l < lanes; ++l) {
        lane_begin[l] = a;
        for (int q : lane_items[l]) {
          hja[a] = hjt[q];
          hja[a].gram32 = (sweep < E->gram32_sweeps && items[q].fp32 && hjt[q].r >= 256) ? 1 : 0;
          hja[a].pair_begin = lane_pairs[l];
          lane_pairs[l] += hja[a].nb / 2;
          lane_rounds[l] = std::max(lane_rounds[l], hja[a].nb - 1);
          ++a;
        }
        lane_n[l] = a - lane_begin[l];
      }
    }
    // reuse the JacTask region of the device descriptor block for the active list
    HIP_TRY(hipMemcpyAsync(djt, hja, nact * sizeof(JacTask), hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemsetAsync(d_counts, 0, 2 * N * sizeof(int), st));
    if (E->timing) HIP_TRY(hipEventRecord(E->ev[4], st));
    int max_rounds = 0;
    for (int l = 0; l < lanes; ++l) max_rounds = std::max(max_rounds, lane_rounds[l]);
    if (lanes > 1) {
      HIP_TRY(hipEventRecord(E->lane_ev[0], st));
      for (int l = 1; l < lanes; ++l) HIP_TRY(hipStreamWaitEvent(E->lane_stream[l], E->lane_ev[0], 0));
    }
    for (int round = 0; round < max_rounds; ++round)
      for (int l = 0; l < lanes; ++l)
        if (lane_n[l] > 0 && round < lane_rounds[l])
          HIP_TRY(launch_jacobi_round(djt + lane_begin[l], lane_n[l], lane_pairs[l], round, E->inner_sweeps,
                                      E->lane_stream[l], true));
    for (int l = 1; l < lanes; ++l) {
      HIP_TRY(hipEventRecord(E->lane_ev[l], E->lane_stream[l]));
      HIP_TRY(hipStreamWaitEvent(st, E->lane_ev[l], 0));
    }
    if (E->timing) HIP_TRY(hipEventRecord(E->ev[5], st));
    for (int l = 0; l < lanes; ++l) E->prof.jacobi_launches += lane_rounds[l];
    HIP_TRY(hipMemcpyAsync(hcounts, d_counts, 2 * N * sizeof(int), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    if (E->timing) {
      float ms = 0;
      (void)hipEventElapsedTime(&ms, E->ev[4], E->ev[5]);
      E->prof.ms_jacobi_rounds += ms;
    }
    if (E->debug_sweeps) {
      long long R = 0, V = 0;
      for (int a = 0; a < nact; ++a) R += hcounts[2 * active[a]], V += hcounts[2 * active[a] + 1];
      fprintf(stderr, "[hipmps sweep] wave %lld sweep %d active %d rotated %lld visited %lld\n", E->prof.waves, sweep,
              nact, R, V);
      if (sweep >= 25)
        for (int a = 0; a < nact; ++a) {
          const int q = active[a];
          fprintf(stderr, "[hipmps sweep]   item %d: m %d n %d r %d lq %d fp32 %d rotated %d visited %d\n", q,
                  items[q].m, items[q].n, items[q].r, items[q].lq, items[q].fp32, hcounts[2 * q], hcounts[2 * q + 1]);
        }
    }
    std::vector<int> next;
    for (int a = 0; a < nact; ++a) {
      int q = active[a];
      long long rot = hcounts[2 * q], vis = hcounts[2 * q + 1];
      prev_rot[q] = rot;
      E->prof.jacobi_pairs_rotated += rot;
      E->prof.jacobi_pairs_visited += vis;
      const double c = items[q].lq ? items[q].r : items[q].c;  // columns the Jacobi works on
      E->prof.jacobi_flops += 8.0 * JP * JP * c * (double)(vis + rot);
      const bool g32 = sweep < E->gram32_sweeps && items[q].fp32 && hjt[q].r >= 256;
      if (rot == 0 && !g32) {
        sweeps_used[q] = sweep + 1;
        converged[q] = 1;
      } else {
        next.push_back(q);
      }
    }
    active.swap(next);
  }
  // rows back from the tiles (the full task list: the device copy held the last active list)
  HIP_TRY(hipMemcpyAsync(djt, hjt, N * sizeof(JacTask), hipMemcpyHostToDevice, st));
  HIP_TRY(launch_from_tiles(djt, N, jt_blocks, jt_ntc, st));
  if (nlq > 0) {  // rows of the Jacobi result back to the column space of X
    RC_TRY(fork_lanes(lq_lanes));
    for (int l = 0; l < lq_lanes; ++l) {
      int b, n;
      lq_slice(l, b, n);
      const hipStream_t ls = E->lane_stream[l];
      HIP_TRY(launch_widen(dtr + 2 * N + b, n, lq_max_c, lq_max_rpad, ls));
      HIP_TRY(launch_lq_apply_q(dlq + b, n, lq_max_r, lq_max_rpad, ls));
    }
    RC_TRY(join_lanes(lq_lanes));
  }
  if (E->timing) HIP_TRY(hipEventRecord(E->ev[2], st));

  // ---- singular values, truncation, factors ----
  HIP_TRY(launch_rownorm(dsp, N, max_r, st));
  HIP_TRY(launch_select(dsp, N, st));
  {
    int max_len64 = 0, max_len32 = 0, mk64 = 0, mk32 = 0;
    for (int q = 0; q < N; ++q) {
      int len = std::max(items[q].m, items[q].n);
      if (items[q].fp32) { max_len32 = std::max(max_len32, len); mk32 = std::max(mk32, items[q].kub); }
      else { max_len64 = std::max(max_len64, len); mk64 = std::max(mk64, items[q].kub); }
    }
    if (n64 > 0) HIP_TRY(launch_gather(dsp, n64, mk64, max_len64, 0, st));
    if (N - n64 > 0) HIP_TRY(launch_gather(dsp + n64, N - n64, mk32, max_len32, 1, st));
    for (auto& R : granges) {
      const auto& it = items[R.b];
      HIP_TRY(launch_gemm(dgm + R.b, R.e - R.b, R.tiles, 0, it.transposed ? 0 : 2, 0, 0, it.fp32 ? 1 : 0, st));
    }
    if (n64 > 0) HIP_TRY(launch_normalize(dsp, n64, mk64, 0, st));
    if (N - n64 > 0) HIP_TRY(launch_normalize(dsp + n64, N - n64, mk32, 1, st));
  }
  // purified left sites back to [l][p][k][K]
  for (int q = 0; q < N; ++q) {
    auto& it = items[q];
    if (it.ki > 1)
      HIP_TRY(launch_swap_mid(ws + it.off_pout, it.s->sites[it.i].p, it.s->chi[it.i], it.ki, 2, it.kub, d_K + q,
                              it.fp32, st));
  }
  // results
  int* hK = hcounts;  // reuse staging
  double* hw = (double*)((char*)E->hdesc.p + desc_bytes + align_up(2 * N * sizeof(int)));
  HIP_TRY(hipMemcpyAsync(hK, d_K, N * sizeof(int), hipMemcpyDeviceToHost, st));
  HIP_TRY(hipMemcpyAsync(hw, d_w, WS_N * N * sizeof(double), hipMemcpyDeviceToHost, st));
  if (E->timing) HIP_TRY(hipEventRecord(E->ev[3], st));
  HIP_TRY(hipStreamSynchronize(st));
  if (E->timing) {
    float a = 0, b = 0, c = 0, d = 0;
    (void)hipEventElapsedTime(&a, E->ev[0], E->ev[6]);
    (void)hipEventElapsedTime(&d, E->ev[6], E->ev[1]);
    (void)hipEventElapsedTime(&b, E->ev[1], E->ev[2]);
    (void)hipEventElapsedTime(&c, E->ev[2], E->ev[3]);
    E->prof.ms_theta += a;
    E->prof.ms_lq += d;
    E->prof.ms_jacobi += b;
    E->prof.ms_split += c;
  }
  E->prof.waves += 1;
  E->prof.wave_tasks += N;
  for (int q = 0; q < N; ++q) {
    int lg = 0;
    while (lg < 13 && (2 << lg) <= items[q].r) ++lg;
    E->prof.applies_by_rank[lg] += 1;
  }
  for (int q = 0; q < N; ++q) {
    E->prof.jacobi_sweeps += sweeps_used[q];
    if (items[q].r >= 256) {
      E->prof.svds_large += 1;
      E->prof.sweeps_large += sweeps_used[q];
    }
  }
  for (int q = 0; q < N; ++q) {
    auto& it = items[q];
    hipmps_state_s* s = it.s;
    int K = hK[q];
    s->chi[it.i + 1] = K;
    s->maxb = std::max(s->maxb, K);
    const double* w = hw + WS_N * q;
    double disc = w[0], kept = w[1];
    if (w[5] != 0.0 || !std::isfinite(kept + disc)) s->status = HIPMPS_E_NUMERIC;
    s->discarded += disc;
    if (kept + disc > 0) s->fid *= kept / (kept + disc);
    // post-SVD norm validation (ExaTnMpsVisitor.cpp:1634-1661): Frobenius norms of the two
    // factors before truncation, sqrt(sum sigma) each for the sqrt(S) split; the reference
    // prints an error when either is below 1e-3 and carries on (its assert is debug-only)
    const double nf_i = s->t.absorb == HIPMPS_ABSORB_RIGHT ? std::sqrt((double)it.r) : std::sqrt(w[4]);
    const double nf_j = s->t.absorb == HIPMPS_ABSORB_RIGHT ? std::sqrt(w[2]) : std::sqrt(w[4]);
    if (std::fabs(nf_i) < 1e-3 || std::fabs(nf_j) < 1e-3) s->norm_warnings++;
    s->n2q++;
    s->sweeps_max = std::max(s->sweeps_max, sweeps_used[q]);
    s->sweeps_total += sweeps_used[q];
    if (!converged[q]) {
      if (s->status == 0) s->status = HIPMPS_E_CONVERGENCE;  // NaN / Inf is the graver report
      s->nonconverged++;
    }
  }
  for (auto& it : items) release_deferred(it.s);
  return 0;
}

// ------------------------------------------------------------------------------------
// per-state gate queue
// ------------------------------------------------------------------------------------
int run_kraus_waves(hipmps_engine E, const std::vector<hipmps_state_s*>& states);

// the open wave of gates, then the channels queued after it
int state_run_wave(hipmps_state_s* s) {
  if (!s->wave.empty()) {
    std::vector<WaveItem> items;
    items.reserve(s->wave.size());
    for (auto& q : s->wave) {
      WaveItem it{};
      it.s = s;
      it.i = q.i;
      it.G = q.G;
      items.push_back(it);
    }
    // growth is all or nothing: when the wave cannot run (HIPMPS_E_NOMEM), its 2-qubit gates
    // are dropped with the error and the sites stay those of the last completed wave
    const int rc = run_wave(s->e, items);
    s->wave.clear();
    std::fill(s->busy.begin(), s->busy.end(), 0);
    if (rc != 0) return rc;
  }
  return run_kraus_waves(s->e, {s});
}

// Queue a 2-qubit gate; returns 1 if the gate conflicts with the open wave (caller runs
// the wave and retries), 0 when queued.
int queue_2q(hipmps_state_s* s, int q0, int q1, const cd* M, bool* conflict) {
  const int i = std::min(q0, q1);
  *conflict = false;
  // a gate on a site with a queued channel comes after that channel: next wave
  if (s->busy[i] || s->busy[i + 1] || s->cbusy[i] || s->cbusy[i + 1]) {
    *conflict = true;
    return 0;
  }
  // fuse pending 1-qubit gates: G' = M . (P_q0 (x) P_q1), q0 on the most significant bit
  cd P0[4] = {1, 0, 0, 1}, P1[4] = {1, 0, 0, 1};
  if (s->has_pend[q0]) std::copy(s->pend[q0].begin(), s->pend[q0].end(), P0);
  if (s->has_pend[q1]) std::copy(s->pend[q1].begin(), s->pend[q1].end(), P1);
  cd Kp[16], Gf[16];
  kron2(P0, P1, Kp);
  mat_mul4(M, Kp, Gf);
  s->has_pend[q0] = s->has_pend[q1] = 0;
  s->pend[q0] = {cd(1), cd(0), cd(0), cd(1)};
  s->pend[q1] = {cd(1), cd(0), cd(0), cd(1)};
  hipmps_state_s::Q2 g;
  g.i = i;
  if (q0 == i) {
    std::copy(Gf, Gf + 16, g.G);
  } else {  // first operand is the right site: swap the two bits of row and column index
    for (int a = 0; a < 2; ++a)
      for (int b = 0; b < 2; ++b)
        for (int a2 = 0; a2 < 2; ++a2)
          for (int b2 = 0; b2 < 2; ++b2) g.G[(a * 2 + b) * 4 + (a2 * 2 + b2)] = Gf[(b * 2 + a) * 4 + (b2 * 2 + a2)];
  }
  s->wave.push_back(g);
  s->busy[i] = s->busy[i + 1] = 1;
  return 0;
}

void queue_1q(hipmps_state_s* s, int q, const cd* M) {
  cd P[4];
  std::copy(s->pend[q].begin(), s->pend[q].end(), P);
  cd R[4];
  mat_mul2(M, P, R);  // later gate on the left
  for (int k = 0; k < 4; ++k) s->pend[q][k] = R[k];
  s->has_pend[q] = 1;
  s->n1q++;
}

int flush_pending_1q(hipmps_state_s* s) {
  for (int q = 0; q < s->n; ++q) {
    if (!s->has_pend[q]) continue;
    double2 m[4];
    for (int k = 0; k < 4; ++k) m[k] = make_double2(s->pend[q][k].real(), s->pend[q][k].imag());
    HIP_TRY(launch_apply_1q(s->sites[q].p, s->fp32, s->chi[q], s->kap[q] * s->chi[q + 1], m, s->e->stream));
    s->has_pend[q] = 0;
    s->pend[q] = {cd(1), cd(0), cd(0), cd(1)};
  }
  return 0;
}

int state_flush(hipmps_state_s* s) {
  RC_TRY(set_device(s->e));
  RC_TRY(state_run_wave(s));
  RC_TRY(flush_pending_1q(s));
  HIP_TRY(hipStreamSynchronize(s->e->stream));
  return 0;
}

// Noise channel rho -> sum_o K_o rho K_o^H on qubit q (ExaTnPmpsVisitor.cpp:887-1130).  The
// reference merges the site with its conjugate over the Kraus leg, applies the channel's Choi
// operator and splits the result back; the same site comes out here without forming that
// (2 chiL chiR)^2 operator: the operators open one row per (old Kraus index, operator), and a
// one-sided FP64 Jacobi on those rows gives orthogonal Kraus vectors sigma_j v_j (the
// reference's U sqrt(S) of the merged operator, up to a unitary on the Kraus leg, which
// leaves rho unchanged).  Kept: the rows up to the first whose partial norm is below 1e-9 (that
// row excluded, ExaTnPmpsVisitor.cpp:1060-1072), at most min(2 chiL chiR, 2 kap) of them, the
// reference's new Kraus extent (:985-996).
// Channels are queued (one per site) after the open wave of gates and compressed together:
// one expansion per channel, then one Jacobi launch per round for all of them, one selection
// and one readback of the new extents.  A 1-qubit gate pending on the site when its channel
// is queued is folded into the operators (K_o P); a gate on a site with a queued channel opens
// the next wave.
constexpr double KRAUS_EPS = 1e-9;

int run_kraus_waves(hipmps_engine E, const std::vector<hipmps_state_s*>& states) {
  std::vector<std::pair<hipmps_state_s*, const hipmps_state_s::QK*>> list;
  for (auto* st_ : states)
    for (const auto& k : st_->cwave) list.push_back({st_, &k});
  if (list.empty()) return 0;
  const hipStream_t st = E->stream;
  const int C = (int)list.size();
  // adj: more rows (kap n_ops) than columns (2 cl cr); the Jacobi then runs on the rows of
  // [X^H | I] (jr = c rows, ld = rows + c), a sweep of few rounds instead of many
  struct Ch {
    hipmps_state_s* s;
    const hipmps_state_s::QK* k;
    int q, cl, cr, kap, n_ops, rows, r_pad, jr, jc, ld;
    bool adj;
    size_t c, off_x, off_v, off_k;
  };
  std::vector<Ch> ch(C);
  size_t off = 0, kbytes = 0;
  int max_rows = 1;
  for (int i = 0; i < C; ++i) {
    Ch& h = ch[i];
    h.s = list[i].first;
    h.k = list[i].second;
    h.q = h.k->q;
    h.n_ops = h.k->n_ops;
    h.cl = h.s->chi[h.q];
    h.cr = h.s->chi[h.q + 1];
    h.kap = h.s->kap[h.q];
    h.rows = h.kap * h.n_ops;
    if (h.rows > 4096) return fail(HIPMPS_E_UNSUPPORTED, "Kraus leg above 4096 rows");
    h.c = (size_t)2 * h.cl * h.cr;
    h.adj = (size_t)h.rows > h.c;
    h.jr = h.adj ? (int)h.c : h.rows;
    h.jc = h.adj ? h.rows : (int)h.c;
    h.ld = h.adj ? h.rows + (int)h.c : (int)h.c;
    h.r_pad = (h.jr + JP - 1) / JP * JP;
    h.off_x = off;
    off = align_up(off + (size_t)h.r_pad * h.ld * 16);
    h.off_v = off;
    off = align_up(off + (size_t)h.r_pad * (8 * 5 + 4));
    h.off_k = off;
    off = align_up(off + (size_t)h.n_ops * 4 * 16);
    kbytes += (size_t)h.n_ops * 4 * sizeof(double2);
    max_rows = std::max(max_rows, h.jr);
  }
  RC_TRY(E->ws.ensure(off));
  char* ws = (char*)E->ws.p;
  const size_t cnt_bytes = align_up((size_t)3 * C * sizeof(int));
  RC_TRY(E->small.ensure(cnt_bytes + (size_t)C * WS_N * sizeof(double)));
  int* d_counts = (int*)E->small.p;  // rotated, visited per channel
  int* d_K = d_counts + 2 * C;       // kept Kraus vectors per channel
  double* d_w = (double*)((char*)E->small.p + cnt_bytes);
  const size_t sp_bytes = align_up((size_t)C * sizeof(SplitTask)), jt_bytes = align_up((size_t)C * sizeof(JacTask));
  const size_t desc_bytes = sp_bytes + jt_bytes;
  RC_TRY(E->desc.ensure(desc_bytes));
  const size_t h_jac = desc_bytes, h_k = h_jac + jt_bytes, h_cnt = h_k + align_up(kbytes),
               h_w = h_cnt + cnt_bytes;
  RC_TRY(E->hdesc.ensure(h_w + (size_t)C * WS_N * sizeof(double)));
  char* hd = (char*)E->hdesc.p;
  SplitTask* hs = (SplitTask*)hd;
  JacTask* hj = (JacTask*)(hd + sp_bytes);
  JacTask* hja = (JacTask*)(hd + h_jac);  // active list of a sweep
  double2* hk = (double2*)(hd + h_k);
  int* hcount = (int*)(hd + h_cnt);
  double* hw = (double*)(hd + h_w);
  memset(hs, 0, desc_bytes);
  size_t ko = 0;
  for (int i = 0; i < C; ++i) {
    const Ch& h = ch[i];
    hipmps_state_s* s = h.s;
    for (int t = 0; t < 4 * h.n_ops; ++t) hk[ko + t] = make_double2(h.k->K[t].real(), h.k->K[t].imag());
    char* v = ws + h.off_v;
    double2* X = (double2*)(ws + h.off_x);
    SplitTask& S = hs[i];
    S.Y = X;
    S.r = h.jr;
    S.c = h.jc;
    S.ldy = h.ld;
    S.m = h.jr;
    S.n = h.jc;
    S.rownorm = (double*)v;
    S.sig = (double*)(v + 8 * h.r_pad);
    S.scale_y = (double*)(v + 16 * h.r_pad);
    S.scale_w = (double*)(v + 24 * h.r_pad);
    S.inv_sig = (double*)(v + 32 * h.r_pad);
    S.perm = (int*)(v + 40 * h.r_pad);
    S.K = d_K + i;
    S.wstats = d_w + (size_t)i * WS_N;
    S.cutoff = KRAUS_EPS;
    S.max_bond = (int)std::min<size_t>(h.c, (size_t)2 * h.kap);
    S.pn_l2 = s->t.partial_norm == HIPMPS_PN_L2;
    S.site_fp32 = s->fp32;
    S.keep_rule = 1;
    JacTask& J = hj[i];
    J.Y = X;
    J.ldy = h.ld;
    J.r = h.jr;
    J.r_pad = h.r_pad;
    J.c = h.jc;
    J.c_tot = h.ld;
    J.nb = h.r_pad / JB;
    J.counter = d_counts + 2 * i;
    J.visited = d_counts + 2 * i + 1;
    J.fro2 = nullptr;
    J.tol = s->fp32 ? E->tol_mixed : E->tol;
    J.gram32 = 0;
    J.pad = 0;
    J.Yt = nullptr;  // row-major sweeps
    J.ntc = 0;
    J.pad2 = 0;
    ko += 4 * (size_t)h.n_ops;
  }
  char* dd = (char*)E->desc.p;
  const SplitTask* dsp = (const SplitTask*)dd;
  JacTask* djt = (JacTask*)(dd + sp_bytes);
  HIP_TRY(hipMemcpyAsync(dd, hd, sp_bytes, hipMemcpyHostToDevice, st));
  ko = 0;
  for (int i = 0; i < C; ++i) {
    const Ch& h = ch[i];
    hipmps_state_s* s = h.s;
    HIP_TRY(hipMemcpyAsync(ws + h.off_k, hk + ko, (size_t)h.n_ops * 4 * sizeof(double2), hipMemcpyHostToDevice, st));
    if (h.adj)
      HIP_TRY(launch_kraus_expand_adj(s->sites[h.q].p, s->fp32, h.cl, h.kap, h.cr, (const double2*)(ws + h.off_k),
                                      h.n_ops, (double2*)(ws + h.off_x), h.ld, h.r_pad, st));
    else
      HIP_TRY(launch_kraus_expand(s->sites[h.q].p, s->fp32, h.cl, h.kap, h.cr, (const double2*)(ws + h.off_k),
                                  h.n_ops, (double2*)(ws + h.off_x), h.r_pad, st));
    ko += 4 * (size_t)h.n_ops;
  }
  // sweeps over the channels not yet converged, one launch per round for all of them
  std::vector<int> active(C);
  for (int i = 0; i < C; ++i) active[i] = i;
  for (int sweep = 0; sweep < E->max_sweeps && !active.empty(); ++sweep) {
    int pairs = 0, rounds = 0;
    for (size_t a = 0; a < active.size(); ++a) {
      hja[a] = hj[active[a]];
      hja[a].pair_begin = pairs;
      pairs += hja[a].nb / 2;
      rounds = std::max(rounds, hja[a].nb - 1);
    }
    HIP_TRY(hipMemcpyAsync(djt, hja, active.size() * sizeof(JacTask), hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemsetAsync(d_counts, 0, 2 * (size_t)C * sizeof(int), st));
    for (int round = 0; round < rounds; ++round)
      HIP_TRY(launch_jacobi_round(djt, (int)active.size(), pairs, round, E->inner_sweeps, st));
    HIP_TRY(hipMemcpyAsync(hcount, d_counts, 2 * (size_t)C * sizeof(int), hipMemcpyDeviceToHost, st));
    HIP_TRY(hipStreamSynchronize(st));
    std::vector<int> next;
    for (int i : active)
      if (hcount[2 * i] != 0) next.push_back(i);
    active.swap(next);
  }
  for (int i : active) {
    if (ch[i].s->status == 0) ch[i].s->status = HIPMPS_E_CONVERGENCE;
    ch[i].s->nonconverged++;
  }
  HIP_TRY(launch_rownorm(dsp, C, max_rows, st));
  HIP_TRY(launch_select(dsp, C, st));
  HIP_TRY(hipMemcpyAsync(hcount + 2 * C, d_K, (size_t)C * sizeof(int), hipMemcpyDeviceToHost, st));
  HIP_TRY(hipMemcpyAsync(hw, d_w, (size_t)C * WS_N * sizeof(double), hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  // new Kraus extents: every site that must grow is reserved before any is written (all or
  // nothing, as in run_wave), so a failure leaves every state and its queued channels as they were
  {
    std::vector<SiteGrowth> grow;
    for (int i = 0; i < C; ++i) {
      const Ch& h = ch[i];
      const int Kn = hcount[2 * C + i];
      if (Kn < 1 || Kn > h.jr) return fail(HIPMPS_E_NUMERIC, "Kraus compression returned no rows");
      const size_t need = (size_t)h.cl * 2 * Kn * h.cr;
      const Site& S = h.s->sites[h.q];
      if (need > S.cap) grow.push_back({h.s, h.q, std::max(need, S.cap * 2), nullptr});
    }
    if (!grow.empty()) RC_TRY(grow_sites(E, grow));
  }
  for (int i = 0; i < C; ++i) {
    const Ch& h = ch[i];
    hipmps_state_s* s = h.s;
    const int Kn = hcount[2 * C + i];
    const double* w = hw + (size_t)i * WS_N;
    if (w[5] != 0.0 || !std::isfinite(w[0] + w[1])) s->status = HIPMPS_E_NUMERIC;
    const int* perm = (const int*)(ws + h.off_v + 40 * (size_t)h.r_pad);
    if (h.adj)
      HIP_TRY(launch_kraus_gather_adj((double2*)(ws + h.off_x), h.ld, h.rows, (const double*)(ws + h.off_v), perm,
                                      d_K + i, Kn, s->sites[h.q].p, s->fp32, h.cl, h.cr, st));
    else
      HIP_TRY(launch_kraus_gather((double2*)(ws + h.off_x), h.ld, perm, d_K + i, Kn, s->sites[h.q].p, s->fp32,
                                  h.cl, h.cr, st));
    s->kap[h.q] = Kn;
    s->kraus_ops++;
  }
  HIP_TRY(hipStreamSynchronize(st));
  for (auto* st_ : states) {
    st_->cwave.clear();
    std::fill(st_->cbusy.begin(), st_->cbusy.end(), 0);
    release_deferred(st_);
  }
  return 0;
}

bool all_finite(const double* v, size_t n) {
  for (size_t i = 0; i < n; ++i)
    if (!std::isfinite(v[i])) return false;
  return true;
}

bool gate_finite(const hipmps_gate& g) { return all_finite(g.m, g.nq == 2 ? 32 : 8); }

int apply_gate_list(hipmps_state_s* s, int n, const hipmps_gate* gates) {
  for (int k = 0; k < n; ++k) {
    const hipmps_gate& g = gates[k];
    if (!gate_finite(g)) return fail(HIPMPS_E_NUMERIC, "gate matrix holds NaN or Inf");
    cd M[16];
    if (g.nq == 1) {
      if (g.q0 < 0 || g.q0 >= s->n) return fail(HIPMPS_E_ARG, "qubit index out of range");
      for (int t = 0; t < 4; ++t) M[t] = cd(g.m[2 * t], g.m[2 * t + 1]);
      queue_1q(s, g.q0, M);
    } else if (g.nq == 2) {
      if (g.q0 < 0 || g.q0 >= s->n || g.q1 < 0 || g.q1 >= s->n || g.q0 == g.q1)
        return fail(HIPMPS_E_ARG, "qubit index out of range");
      if (std::abs(g.q0 - g.q1) != 1)
        return fail(HIPMPS_E_NONADJACENT, "2-qubit gate on non-neighbouring qubits " + std::to_string(g.q0) +
                                              "," + std::to_string(g.q1));
      for (int t = 0; t < 16; ++t) M[t] = cd(g.m[2 * t], g.m[2 * t + 1]);
      bool conflict = false;
      RC_TRY(queue_2q(s, g.q0, g.q1, M, &conflict));
      if (conflict) {
        RC_TRY(state_run_wave(s));
        RC_TRY(queue_2q(s, g.q0, g.q1, M, &conflict));
      }
    } else {
      return fail(HIPMPS_E_ARG, "gate must act on 1 or 2 qubits");
    }
  }
  return 0;
}

// ------------------------------------------------------------------------------------
// measurement: environments
// ------------------------------------------------------------------------------------
struct EnvJob {
  const void* A;
  int cl, cr;
  const double2* Ein;
  double2* X;
  double2* Eout;
  int z;              // 1: apply diag(d0, d1) on the physical index of this site
  double d0 = 1.0, d1 = -1.0;
  int kap = 1;        // Kraus leg of the site, traced with the physical index
  const void* Abra = nullptr;  // bra-side site when it differs from A (projected copies)
};

// left step: E'(cr x cr) = A^H D (E A), D = diag(+-1) on the physical index when z; the two
// GEMMs of one step (tile_begin 0)
void env_tasks_left(const EnvJob& j, GemmTask& a, GemmTask& b) {
  memset(&a, 0, sizeof(a));
  const int d = 2 * j.kap;  // physical index and Kraus leg, (p, k) with k fastest
  a.A = j.Ein; a.lda = j.cl;
  a.B = j.A; a.ldb = d * j.cr;
  a.C = j.X; a.ldc = d * j.cr;
  a.M = j.cl; a.N = d * j.cr; a.K = j.cl;
  a.tiles_m = (a.M + 31) / 32; a.tiles_n = (a.N + 31) / 32;
  memset(&b, 0, sizeof(b));
  b.A = j.Abra ? j.Abra : j.A; b.lda = j.cr;
  b.B = j.X; b.ldb = j.cr;
  b.C = j.Eout; b.ldc = j.cr;
  b.M = j.cr; b.N = j.cr; b.K = d * j.cl;
  b.sign_div = j.z ? j.kap : 0;
  b.kd0 = j.d0;
  b.kd1 = j.d1;
  b.tiles_m = (b.M + 31) / 32; b.tiles_n = (b.N + 31) / 32;
}

// right step: E'(cl x cl) = conj(A) (A R^T)^T with the physical sign when z
void env_tasks_right(const EnvJob& j, GemmTask& a, GemmTask& b) {
  memset(&a, 0, sizeof(a));
  const int d = 2 * j.kap;
  a.A = j.A; a.lda = j.cr;                 // (d cl) x cr
  a.B = j.Ein; a.ldb = j.cr;               // R^T
  a.C = j.X; a.ldc = j.cr;
  a.M = d * j.cl; a.N = j.cr; a.K = j.cr;
  a.tiles_m = (a.M + 31) / 32; a.tiles_n = (a.N + 31) / 32;
  memset(&b, 0, sizeof(b));
  b.A = j.A; b.lda = d * j.cr;             // conj, cl x (d cr)
  b.B = j.X; b.ldb = d * j.cr;             // X viewed cl x (d cr), transposed
  b.C = j.Eout; b.ldc = j.cl;
  b.M = j.cl; b.N = j.cl; b.K = d * j.cr;
  b.sign_div = j.z ? j.kap * j.cr : 0;
  b.kd0 = j.d0;
  b.kd1 = j.d1;
  b.tiles_m = (b.M + 31) / 32; b.tiles_n = (b.N + 31) / 32;
}

// independent steps, one launch per GEMM for all of them
int env_steps(hipmps_engine E, std::vector<EnvJob>& jobs, bool left, int fp32) {
  if (jobs.empty()) return 0;
  const int N = (int)jobs.size();
  RC_TRY(E->hdesc.ensure(2 * N * sizeof(GemmTask)));
  RC_TRY(E->desc.ensure(2 * N * sizeof(GemmTask)));
  GemmTask* h = (GemmTask*)E->hdesc.p;
  int t1 = 0, t2 = 0;
  for (int q = 0; q < N; ++q) {
    GemmTask& a = h[q];
    GemmTask& b = h[N + q];
    if (left) env_tasks_left(jobs[q], a, b);
    else env_tasks_right(jobs[q], a, b);
    a.tile_begin = t1; t1 += a.tiles_m * a.tiles_n;
    b.tile_begin = t2; t2 += b.tiles_m * b.tiles_n;
  }
  GemmTask* d = (GemmTask*)E->desc.p;
  HIP_TRY(hipMemcpyAsync(d, h, 2 * N * sizeof(GemmTask), hipMemcpyHostToDevice, E->stream));
  if (left) {
    HIP_TRY(launch_gemm(d, N, t1, 0, 0, 0, fp32, 0, E->stream));
    HIP_TRY(launch_gemm(d + N, N, t2, 2, 0, fp32, 0, 0, E->stream));
  } else {
    HIP_TRY(launch_gemm(d, N, t1, 0, 3, fp32, 0, 0, E->stream));
    HIP_TRY(launch_gemm(d + N, N, t2, 1, 3, fp32, 0, 0, E->stream));
  }
  HIP_TRY(hipStreamSynchronize(E->stream));  // staging reuse
  return 0;
}

int env_left(hipmps_engine E, std::vector<EnvJob>& jobs, int fp32) { return env_steps(E, jobs, true, fp32); }
int env_right(hipmps_engine E, std::vector<EnvJob>& jobs, int fp32) { return env_steps(E, jobs, false, fp32); }

// a chain of steps, each reading the environment the one before wrote (a sweep over the
// sites): every descriptor is staged at once and the steps launch back to back in stream
// order (the X scratch they share is rewritten only after the step before has read it), with
// one synchronisation at the end instead of one per site
int env_chain(hipmps_engine E, const std::vector<EnvJob>& jobs, bool left, int fp32) {
  if (jobs.empty()) return 0;
  const int N = (int)jobs.size();
  RC_TRY(E->hdesc.ensure(2 * N * sizeof(GemmTask)));
  RC_TRY(E->desc.ensure(2 * N * sizeof(GemmTask)));
  GemmTask* h = (GemmTask*)E->hdesc.p;
  for (int q = 0; q < N; ++q) {
    if (left) env_tasks_left(jobs[q], h[2 * q], h[2 * q + 1]);
    else env_tasks_right(jobs[q], h[2 * q], h[2 * q + 1]);
  }
  GemmTask* d = (GemmTask*)E->desc.p;
  HIP_TRY(hipMemcpyAsync(d, h, 2 * N * sizeof(GemmTask), hipMemcpyHostToDevice, E->stream));
  for (int q = 0; q < N; ++q) {
    const GemmTask& a = h[2 * q];
    const GemmTask& b = h[2 * q + 1];
    if (left) {
      HIP_TRY(launch_gemm(d + 2 * q, 1, a.tiles_m * a.tiles_n, 0, 0, 0, fp32, 0, E->stream));
      HIP_TRY(launch_gemm(d + 2 * q + 1, 1, b.tiles_m * b.tiles_n, 2, 0, fp32, 0, 0, E->stream));
    } else {
      HIP_TRY(launch_gemm(d + 2 * q, 1, a.tiles_m * a.tiles_n, 0, 3, fp32, 0, 0, E->stream));
      HIP_TRY(launch_gemm(d + 2 * q + 1, 1, b.tiles_m * b.tiles_n, 1, 3, fp32, 0, 0, E->stream));
    }
  }
  HIP_TRY(hipStreamSynchronize(E->stream));  // staging reuse
  return 0;
}

// environment storage: all left (L_0..L_n) and right (R_0..R_n) environments + scratch
struct EnvLayout {
  std::vector<size_t> offL, offR;
  size_t offX = 0, offT = 0, total = 0;
};

EnvLayout env_layout(const hipmps_state_s* s, size_t extra_mats, size_t extra_elems) {
  EnvLayout L;
  size_t off = 0;
  L.offL.resize(s->n + 1);
  L.offR.resize(s->n + 1);
  int cmax = 1;
  for (int k = 0; k <= s->n; ++k) cmax = std::max(cmax, s->chi[k]);
  for (int k = 0; k <= s->n; ++k) {
    L.offL[k] = off;
    off = align_up(off + (size_t)s->chi[k] * s->chi[k] * 16);
  }
  for (int k = 0; k <= s->n; ++k) {
    L.offR[k] = off;
    off = align_up(off + (size_t)s->chi[k] * s->chi[k] * 16);
  }
  L.offX = off;
  off = align_up(off + align_up((size_t)cmax * 2 * max_kap(s) * cmax * 16) * std::max<size_t>(1, extra_mats));
  L.offT = off;
  off = align_up(off + std::max<size_t>(extra_elems, 1) * 16);
  L.total = off;
  return L;
}

// diag: empty (identity everywhere) or 2n values diag(d_k0, d_k1) per site
int compute_left_envs_diag(hipmps_state_s* s, const EnvLayout& L, const std::vector<double>& diag) {
  hipmps_engine E = s->e;
  char* base = (char*)E->env.p;
  HIP_TRY(launch_fill_identity1((double2*)(base + L.offL[0]), E->stream));
  std::vector<EnvJob> jobs(s->n);
  for (int k = 0; k < s->n; ++k) {
    jobs[k] = {s->sites[k].p, s->chi[k], s->chi[k + 1], (const double2*)(base + L.offL[k]),
               (double2*)(base + L.offX), (double2*)(base + L.offL[k + 1]), 0};
    jobs[k].kap = s->kap[k];
    if (!diag.empty() && !(diag[2 * k] == 1.0 && diag[2 * k + 1] == 1.0)) {
      jobs[k].z = 1;
      jobs[k].d0 = diag[2 * k];
      jobs[k].d1 = diag[2 * k + 1];
    }
  }
  return env_chain(E, jobs, true, s->fp32);
}

int compute_left_envs(hipmps_state_s* s, const EnvLayout& L, const std::vector<int>& zmask) {
  std::vector<double> diag;
  if (!zmask.empty()) {
    diag.assign(2 * s->n, 1.0);
    for (int k = 0; k < s->n; ++k)
      if (zmask[k]) diag[2 * k + 1] = -1.0;
  }
  return compute_left_envs_diag(s, L, diag);
}

int compute_right_envs(hipmps_state_s* s, const EnvLayout& L) {
  hipmps_engine E = s->e;
  char* base = (char*)E->env.p;
  HIP_TRY(launch_fill_identity1((double2*)(base + L.offR[s->n]), E->stream));
  std::vector<EnvJob> jobs;
  for (int k = s->n - 1; k >= 0; --k) {
    EnvJob j{s->sites[k].p, s->chi[k], s->chi[k + 1], (const double2*)(base + L.offR[k + 1]),
             (double2*)(base + L.offX), (double2*)(base + L.offR[k]), 0};
    j.kap = s->kap[k];
    jobs.push_back(j);
  }
  return env_chain(E, jobs, false, s->fp32);
}

int read_scalar(hipmps_engine E, const void* dptr, double2* out) {
  HIP_TRY(hipMemcpyAsync(out, dptr, sizeof(double2), hipMemcpyDeviceToHost, E->stream));
  HIP_TRY(hipStreamSynchronize(E->stream));
  return 0;
}

// sum over pairs M_t o R_t (elementwise), results to host
int trace_products(hipmps_engine E, const std::vector<const double2*>& A, const std::vector<const double2*>& B,
                   const std::vector<int>& dims, char* scratch, std::vector<double2>& out) {
  const int N = (int)A.size();
  if (N == 0) return 0;
  size_t bytes = align_up(N * sizeof(void*)) * 2 + align_up(N * sizeof(int)) + N * sizeof(double2);
  RC_TRY(E->hdesc.ensure(bytes));
  char* h = (char*)E->hdesc.p;
  const double2** ha = (const double2**)h;
  const double2** hb = (const double2**)(h + align_up(N * sizeof(void*)));
  int* hdim = (int*)(h + 2 * align_up(N * sizeof(void*)));
  for (int q = 0; q < N; ++q) {
    ha[q] = A[q];
    hb[q] = B[q];
    hdim[q] = dims[q];
  }
  size_t head = 2 * align_up(N * sizeof(void*)) + align_up(N * sizeof(int));
  HIP_TRY(hipMemcpyAsync(scratch, h, head, hipMemcpyHostToDevice, E->stream));
  double2* dout = (double2*)(scratch + head);
  HIP_TRY(launch_trace_prod((const double2* const*)scratch, (const double2* const*)(scratch + align_up(N * sizeof(void*))),
                            (const int*)(scratch + 2 * align_up(N * sizeof(void*))), dout, N, E->stream));
  out.resize(N);
  HIP_TRY(hipMemcpyAsync(out.data(), dout, N * sizeof(double2), hipMemcpyDeviceToHost, E->stream));
  HIP_TRY(hipStreamSynchronize(E->stream));
  return 0;
}

// <Z_i> for every site; with norm, also <psi|psi> (the closed left environment L_n against the
// identity R_n, one more pair of the same trace products)
int measure_z_all(hipmps_state_s* s, double* out, double* norm = nullptr) {
  hipmps_engine E = s->e;
  const int n = s->n;
  int cmax = 1;
  for (int k = 0; k <= n; ++k) cmax = std::max(cmax, s->chi[k]);
  // M_i buffers (n of them) + pointer tables
  size_t mbytes = 0;
  std::vector<size_t> offM(n);
  for (int i = 0; i < n; ++i) {
    offM[i] = mbytes;
    mbytes = align_up(mbytes + (size_t)s->chi[i + 1] * s->chi[i + 1] * 16);
  }
  EnvLayout L = env_layout(s, (size_t)n, mbytes / 16 + 4096);
  RC_TRY(E->env.ensure(L.total));
  char* base = (char*)E->env.p;
  RC_TRY(compute_left_envs(s, L, {}));
  RC_TRY(compute_right_envs(s, L));
  // M_i = env update of L_i with Z at i (batched); X scratch per job
  std::vector<EnvJob> jobs;
  size_t xstride = align_up((size_t)cmax * 2 * max_kap(s) * cmax * 16);
  for (int i = 0; i < n; ++i) {
    jobs.push_back({s->sites[i].p, s->chi[i], s->chi[i + 1], (const double2*)(base + L.offL[i]),
                    (double2*)(base + L.offX + i * xstride), (double2*)(base + L.offT + offM[i]), 1});
    jobs.back().kap = s->kap[i];
  }
  RC_TRY(env_left(E, jobs, s->fp32));
  std::vector<const double2*> A, B;
  std::vector<int> dims;
  for (int i = 0; i < n; ++i) {
    A.push_back((const double2*)(base + L.offT + offM[i]));
    B.push_back((const double2*)(base + L.offR[i + 1]));
    dims.push_back(s->chi[i + 1]);
  }
  if (norm) {
    A.push_back((const double2*)(base + L.offL[n]));
    B.push_back((const double2*)(base + L.offR[n]));
    dims.push_back(s->chi[n]);
  }
  std::vector<double2> res;
  // pointer tables go right after the M buffers
  RC_TRY(trace_products(E, A, B, dims, base + L.offT + align_up(mbytes), res));
  for (int i = 0; i < n; ++i) out[i] = res[i].x;
  if (norm) *norm = res[n].x;
  return 0;
}

int measure_zz(hipmps_state_s* s, const int* pairs, int k, double* out) {
  hipmps_engine E = s->e;
  const int n = s->n;
  int cmax = 1;
  for (int b = 0; b <= n; ++b) cmax = std::max(cmax, s->chi[b]);
  size_t mat = align_up((size_t)cmax * cmax * 16);
  size_t xmat = align_up((size_t)cmax * 2 * max_kap(s) * cmax * 16);
  EnvLayout L = env_layout(s, 1, 0);
  size_t per_pair = 2 * mat + xmat;
  size_t total = L.total + per_pair * (size_t)k + 4096 + (size_t)k * 64;
  RC_TRY(E->env.ensure(total));
  char* base = (char*)E->env.p;
  RC_TRY(compute_left_envs(s, L, {}));
  RC_TRY(compute_right_envs(s, L));
  char* pb = base + L.total;
  struct P { int i, j, cur; int pp; };
  std::vector<P> ps(k);
  std::vector<double> norm_val(1, 0.0);
  double2 nrm;
  RC_TRY(read_scalar(E, base + L.offL[n], &nrm));
  for (int t = 0; t < k; ++t) {
    int a = pairs[2 * t], b = pairs[2 * t + 1];
    if (a < 0 || a >= n || b < 0 || b >= n) return fail(HIPMPS_E_ARG, "zz pair out of range");
    if (a > b) std::swap(a, b);
    ps[t] = {a, b, a, 0};
  }
  // advance all pairs in lock-step
  bool any = true;
  while (any) {
    any = false;
    std::vector<EnvJob> jobs;
    std::vector<int> who;
    for (int t = 0; t < k; ++t) {
      P& p = ps[t];
      if (p.i == p.j || p.cur > p.j) continue;
      any = true;
      char* buf = pb + per_pair * t;
      const double2* ein = (p.cur == p.i) ? (const double2*)(base + L.offL[p.i]) : (const double2*)(buf + mat * p.pp);
      double2* eout = (double2*)(buf + mat * (1 - p.pp));
      int z = (p.cur == p.i || p.cur == p.j) ? 1 : 0;
      jobs.push_back({s->sites[p.cur].p, s->chi[p.cur], s->chi[p.cur + 1], ein, (double2*)(buf + 2 * mat), eout, z});
      jobs.back().kap = s->kap[p.cur];
      who.push_back(t);
    }
    if (!any) break;
    RC_TRY(env_left(E, jobs, s->fp32));
    for (int t : who) {
      ps[t].cur++;
      ps[t].pp = 1 - ps[t].pp;
    }
  }
  std::vector<const double2*> A, B;
  std::vector<int> dims, idx;
  for (int t = 0; t < k; ++t) {
    P& p = ps[t];
    if (p.i == p.j) {
      out[t] = nrm.x;  // Z_i Z_i = I
      continue;
    }
    A.push_back((const double2*)(pb + per_pair * t + mat * p.pp));
    B.push_back((const double2*)(base + L.offR[p.j + 1]));
    dims.push_back(s->chi[p.j + 1]);
    idx.push_back(t);
  }
  std::vector<double2> res;
  RC_TRY(trace_products(E, A, B, dims, pb + per_pair * k, res));
  for (size_t q = 0; q < idx.size(); ++q) out[idx[q]] = res[q].x;
  return 0;
}

int measure_parity(hipmps_state_s* s, const std::vector<int>& zmask, double* out) {
  hipmps_engine E = s->e;
  EnvLayout L = env_layout(s, 1, 0);
  RC_TRY(E->env.ensure(L.total));
  RC_TRY(compute_left_envs(s, L, zmask));
  double2 v;
  RC_TRY(read_scalar(E, (char*)E->env.p + L.offL[s->n], &v));
  *out = v.x;
  return 0;
}

int measure_diag(hipmps_state_s* s, const std::vector<double>& diag, double* out) {
  hipmps_engine E = s->e;
  EnvLayout L = env_layout(s, 1, 0);
  RC_TRY(E->env.ensure(L.total));
  RC_TRY(compute_left_envs_diag(s, L, diag));
  double2 v;
  RC_TRY(read_scalar(E, (char*)E->env.p + L.offL[s->n], &v));
  *out = v.x;
  return 0;
}

// getMeasureSample (ExaTnMpsVisitor.cpp:2211-2364) for many shots at once, the measured qubits in
// increasing site order.  Every shot carries its own left environment E_t across the chain
// (the traced sites contract ket with bra, a measured site keeps only the shot's outcome, scaled
// by 1/p as the reference's collapse tensors are); at a measured site one launch forms both
// outcome branches A_b^H E_t A_b of every shot, one trace against the shared right environment
// gives their probabilities, and the host applies the shot's draw (outcome 0 when the draw is <=
// p0, values below 1e-12 counted as 0).  The first measured qubit therefore sees the
// probabilities of the unnormalised state, every later one its conditional probabilities, as in
// the reference.  Shots run in blocks whose environments fit a bounded workspace.
int sample_measure(hipmps_state_s* s, const int* qubits, int k, int shots, const double* draws, char* out) {
  hipmps_engine E = s->e;
  const int n = s->n;
  int cmax = 1;
  for (int b = 0; b <= n; ++b) cmax = std::max(cmax, s->chi[b]);
  const EnvLayout L = env_layout(s, 1, 0);
  const size_t mat = align_up((size_t)cmax * cmax * 16);
  const size_t xs = align_up((size_t)cmax * 2 * max_kap(s) * cmax * 16);
  const size_t per_shot = 3 * mat + 2 * xs;  // E and two candidates, two X scratches
  const size_t budget = std::max<size_t>((size_t)256 << 20, 8 * per_shot);
  const int B = (int)std::max<size_t>(1, std::min<size_t>((size_t)shots, budget / per_shot));
  const size_t tables = 2 * align_up(2 * (size_t)B * sizeof(void*)) + align_up(2 * (size_t)B * sizeof(int)) +
                        2 * (size_t)B * sizeof(double2) + 256;
  RC_TRY(E->env.ensure(L.total + (size_t)B * per_shot + tables));
  char* base = (char*)E->env.p;
  RC_TRY(compute_right_envs(s, L));
  char* sb = base + L.total;
  char* tb = sb + (size_t)B * per_shot;
  auto mat_at = [&](int t, int i) { return (double2*)(sb + (size_t)t * per_shot + (size_t)i * mat); };
  auto x_at = [&](int t, int b) { return (double2*)(sb + (size_t)t * per_shot + 3 * mat + (size_t)b * xs); };
  constexpr double PROB_EPS = 1e-12;
  for (int s0 = 0; s0 < shots; s0 += B) {
    const int nb = std::min(B, shots - s0);
    std::vector<int> cur(nb, 0);
    std::vector<double> f(nb, 1.0);  // product of 1/p of the outcomes so far
    {  // E_t = [1] (bond 0 has dimension 1); the stream is idle here (every step above ended synchronised)
      const std::vector<double2> ones(nb, make_double2(1.0, 0.0));
      HIP_TRY(hipMemcpy2D(mat_at(0, 0), per_shot, ones.data(), sizeof(double2), sizeof(double2), nb,
                          hipMemcpyHostToDevice));
    }
    int mi = 0;
    for (int site = 0; site < n && mi < k; ++site) {
      const bool meas = qubits[mi] == site;
      std::vector<EnvJob> jobs;
      jobs.reserve((size_t)nb * 2);
      for (int t = 0; t < nb; ++t) {
        for (int b = 0; b < (meas ? 2 : 1); ++b) {
          EnvJob j{s->sites[site].p, s->chi[site], s->chi[site + 1], mat_at(t, cur[t]), x_at(t, b),
                   mat_at(t, (cur[t] + 1 + b) % 3), meas ? 1 : 0};
          j.kap = s->kap[site];
          if (meas) {
            j.d0 = b == 0 ? 1.0 : 0.0;
            j.d1 = b == 1 ? 1.0 : 0.0;
          }
          jobs.push_back(j);
        }
      }
      RC_TRY(env_steps(E, jobs, true, s->fp32));
      if (!meas) {
        for (int t = 0; t < nb; ++t) cur[t] = (cur[t] + 1) % 3;
        continue;
      }
      std::vector<const double2*> A, R;
      std::vector<int> dims;
      for (int t = 0; t < nb; ++t)
        for (int b = 0; b < 2; ++b) {
          A.push_back(mat_at(t, (cur[t] + 1 + b) % 3));
          R.push_back((const double2*)(base + L.offR[site + 1]));
          dims.push_back(s->chi[site + 1]);
        }
      std::vector<double2> pr;
      RC_TRY(trace_products(E, A, R, dims, tb, pr));
      for (int t = 0; t < nb; ++t) {
        double prob[2];
        for (int b = 0; b < 2; ++b) {
          prob[b] = pr[2 * t + b].x * f[t];
          if (std::fabs(prob[b]) < PROB_EPS) prob[b] = 0.0;
        }
        const int bit = draws[(size_t)(s0 + t) * k + mi] <= prob[0] ? 0 : 1;
        out[(size_t)(s0 + t) * k + mi] = bit ? '1' : '0';
        if (prob[bit] > 0.0) f[t] /= prob[bit];
        cur[t] = (cur[t] + 1 + bit) % 3;
      }
      ++mi;
    }
  }
  return 0;
}

// Reduced density matrix of `qubits` (the trace over the other qubits that the reference's
// finalize contracts, ExaTnPmpsVisitor.cpp:657-761; at most 15 qubits, :661): out[x + 2^k x'] =
// <x| rho |x'>, x = sum_i b_i 2^i over qubits[i] (the ket index fastest, as the reference's
// flattened tensor).  Pure and purified states alike.
//
// Two half-sweeps meet at a bond m: sites 0..m-1 from the left, sites n-1..m from the right
// (on copies with the bonds swapped, so that both halves run the same step).  A half-sweep
// keeps the environments of all its labels A (ket and bra components of the measured sites
// passed so far, 4 per site) as one tensor E[l'][A][l] (bra bond, label, ket bond), so that a
// site costs two GEMMs over all labels at once:
//   ket side  X[s][l'][A][r] = sum_l E[l'][A][l] A[l][s][r]            one task per s = (p, kappa)
//   bra side  E'[r'][A][r]   = sum_(s, l') conj(B[s][l'][r']) X[s][l'][A][r]
// with B the site copied to s-major order.  An unmeasured site is traced (ket and bra are the
// same site); at a measured one the ket and bra take the site's copies with |0><b| applied
// (component b moved to slot 0, the other slot zero), and the four (b, b') give the labels
// A' = (2b + b') |A| + A.  X holds a block of s at a time when all of them would not fit; the
// bra side then accumulates over the blocks.  The halves join in one GEMM over both bonds,
// rho[A_L][A_R] = sum_(l', l) E_L[l'][A_L][l] E_R[l'][A_R][l], and a scatter puts the labels in
// the output order.  The split m minimises the device workspace: about 4^(k/2) chi^2 per half
// instead of 4^k chi^2 for a single sweep.
constexpr int RDM_MAX_QUBITS = 15;

struct RdmSite {
  const void* ket[2];  // [l][s][r]
  const void* bra[2];  // [s][l][r]
  int cl, d, cr;
  bool meas;
};

struct RdmPlan {
  int m = -1;
  size_t es = 0, xs = 0, el = 0, er = 0, nl = 1, nr = 1, copies = 0, xcap = 0, total = 0;
};

constexpr size_t RDM_INT_MAX = ((size_t)1 << 31) - 1;
// k_gemm runs one 256-thread block per tile and a dispatch holds at most 2^32 work-items
constexpr size_t RDM_TILES_MAX = ((size_t)1 << 24) - 1;

// block of s that X holds for a site with `per` = cl |A| cr elements and `tiles` ket-side GEMM
// tiles per s
inline int rdm_block(int d, size_t per, size_t xcap, size_t tiles) {
  size_t w = (size_t)d * per <= xcap ? (size_t)d : std::max<size_t>(1, std::min<size_t>(d, xcap / per));
  w = std::min(w, std::max<size_t>(1, RDM_TILES_MAX / std::max<size_t>(1, tiles)));
  return (int)w;
}

int rdm_half(hipmps_engine E, const std::vector<RdmSite>& v, int fp32, size_t es_site, char* ea, char* eb, char* xs,
             char* dst, size_t xcap) {
  const hipStream_t st = E->stream;
  HIP_TRY(launch_fill_identity1((double2*)(v.empty() ? dst : ea), st));
  size_t nA = 1;
  for (size_t t = 0; t < v.size(); ++t) {
    const RdmSite& S = v[t];
    const int cl = S.cl, d = S.d, cr = S.cr;
    char* out = (t + 1 == v.size()) ? dst : eb;
    const size_t per = (size_t)cl * nA * cr;
    const int w = rdm_block(d, per, xcap, ((cl * nA + 31) / 32) * (((size_t)cr + 31) / 32));
    const int nket = S.meas ? 2 : 1, nbra = nket;
    const int nch = (d + w - 1) / w;
    const size_t ndesc = (size_t)nket * (d + (size_t)nch * nbra);
    RC_TRY(E->hdesc.ensure(ndesc * sizeof(GemmTask)));
    RC_TRY(E->desc.ensure(ndesc * sizeof(GemmTask)));
    GemmTask* h = (GemmTask*)E->hdesc.p;
    memset(h, 0, ndesc * sizeof(GemmTask));
    struct Launch {
      size_t off;
      int ntask, tiles, bra;
    };
    std::vector<Launch> launches;
    size_t o = 0;
    for (int b = 0; b < nket; ++b)
      for (int c = 0; c < nch; ++c) {
        const int s0 = c * w, ws = std::min(w, d - s0);
        int t1 = 0;
        for (int sl = 0; sl < ws; ++sl) {
          GemmTask& g = h[o + sl];
          g.A = ea;
          g.lda = cl;
          g.M = (int)(cl * nA);
          g.K = cl;
          g.B = (const char*)S.ket[b] + (size_t)(s0 + sl) * cr * es_site;
          g.ldb = d * cr;
          g.N = cr;
          g.C = (double2*)xs + (size_t)sl * per;
          g.ldc = cr;
          g.tiles_m = (g.M + 31) / 32;
          g.tiles_n = (g.N + 31) / 32;
          g.tile_begin = t1;
          t1 += g.tiles_m * g.tiles_n;
        }
        launches.push_back({o, ws, t1, 0});
        o += ws;
        int t2 = 0;
        for (int bp = 0; bp < nbra; ++bp) {
          GemmTask& g = h[o + bp];
          const int slot = S.meas ? 2 * b + bp : 0;
          g.A = (const char*)S.bra[bp] + (size_t)s0 * cl * cr * es_site;
          g.lda = cr;
          g.M = cr;
          g.K = ws * cl;
          g.B = xs;
          g.ldb = (int)(nA * cr);
          g.N = (int)(nA * cr);
          g.C = (double2*)out + (size_t)slot * nA * cr;
          g.ldc = (int)((S.meas ? 4 : 1) * nA * cr);
          g.accumulate = s0 > 0;
          g.tiles_m = (g.M + 31) / 32;
          g.tiles_n = (g.N + 31) / 32;
          g.tile_begin = t2;
          t2 += g.tiles_m * g.tiles_n;
        }
        launches.push_back({o, nbra, t2, 1});
        o += nbra;
      }
    GemmTask* dd = (GemmTask*)E->desc.p;
    HIP_TRY(hipMemcpyAsync(dd, h, ndesc * sizeof(GemmTask), hipMemcpyHostToDevice, st));
    // stream order: each block's bra side reads X before the next block's ket side rewrites it
    for (const auto& L : launches)
      HIP_TRY(launch_gemm(dd + L.off, L.ntask, L.tiles, L.bra ? 2 : 0, 0, L.bra ? fp32 : 0, L.bra ? 0 : fp32, 0, st));
    HIP_TRY(hipStreamSynchronize(st));  // staging reuse
    nA *= S.meas ? 4 : 1;
    if (out == eb) std::swap(ea, eb);
  }
  return 0;
}

int measure_rdm(hipmps_state_s* s, const std::vector<int>& qubits, double* out) {
  hipmps_engine E = s->e;
  const hipStream_t st = E->stream;
  const int n = s->n, k = (int)qubits.size();
  if (k < 1 || k > RDM_MAX_QUBITS) return fail(HIPMPS_E_UNSUPPORTED, "reduced density matrix of 1 to 15 qubits");
  std::vector<int> pos(n, -1);
  for (int i = 0; i < k; ++i) {
    const int q = qubits[i];
    if (q < 0 || q >= n || pos[q] >= 0) return fail(HIPMPS_E_ARG, "qubits out of range or repeated");
    pos[q] = i;
  }
  const size_t nout = (size_t)1 << (2 * k);
  std::vector<size_t> sb(n);
  for (int q = 0; q < n; ++q) sb[q] = align_up(site_elems(s, q) * s->es);
  // the workspace of a split at bond m; false when an int argument would overflow
  auto plan = [&](int m, RdmPlan& P) -> bool {
    P = RdmPlan{};
    P.m = m;
    size_t max_e = 1;
    auto pass_e = [&](int q0, int q1, int step, size_t& labels) {
      size_t nA = 1;
      for (int q = q0; q != q1; q += step) {
        const size_t cl = step > 0 ? s->chi[q] : s->chi[q + 1];
        max_e = std::max(max_e, nA * cl * cl);
        nA *= pos[q] >= 0 ? 4 : 1;
      }
      labels = nA;
    };
    pass_e(0, m, 1, P.nl);
    pass_e(n - 1, m - 1, -1, P.nr);
    const size_t cm = s->chi[m];
    P.el = P.nl * cm * cm;
    P.er = P.nr * cm * cm;
    max_e = std::max(max_e, std::max(P.el, P.er));
    P.es = max_e;
    P.xcap = std::max(max_e, (size_t)E->rdm_block_elems);
    bool ok = cm * cm <= RDM_INT_MAX && ((P.nl + 31) / 32) * ((P.nr + 31) / 32) <= RDM_TILES_MAX;
    auto pass_x = [&](int q0, int q1, int step) {
      size_t nA = 1;
      for (int q = q0; q != q1; q += step) {
        const size_t cl = step > 0 ? s->chi[q] : s->chi[q + 1], cr = step > 0 ? s->chi[q + 1] : s->chi[q];
        const int d = 2 * s->kap[q];
        const size_t per = cl * nA * cr;
        const size_t tiles = ((cl * nA + 31) / 32) * ((cr + 31) / 32);
        const int w = rdm_block(d, per, P.xcap, tiles);
        P.xs = std::max(P.xs, (size_t)w * per);
        ok = ok && cl * nA <= RDM_INT_MAX && tiles * w <= RDM_TILES_MAX && 4 * nA * cr <= RDM_INT_MAX &&
             ((cr + 31) / 32) * ((nA * cr + 31) / 32) * 2 <= RDM_TILES_MAX && (size_t)w * cl <= RDM_INT_MAX;
        nA *= pos[q] >= 0 ? 4 : 1;
      }
    };
    pass_x(0, m, 1);
    pass_x(n - 1, m - 1, -1);
    for (int q = 0; q < n; ++q) P.copies += sb[q] * (q < m ? (pos[q] >= 0 ? 4 : 1) : (pos[q] >= 0 ? 6 : 2));
    P.total = 2 * align_up(P.es * 16) + align_up(P.xs * 16) + align_up(P.el * 16) + align_up(P.er * 16) +
              2 * align_up(nout * 16) + P.copies + 4096;
    return ok;
  };
  RdmPlan best;
  for (int m = 0; m <= n; ++m) {
    RdmPlan P;
    if (plan(m, P) && (best.m < 0 || P.total < best.total)) best = P;
  }
  if (best.m < 0) return fail(HIPMPS_E_UNSUPPORTED, "reduced density matrix too large for these bonds");
  const int m = best.m;
  // large sweeps get a workspace of their own, released at the end; small ones use the arena
  const bool own = best.total > ((size_t)1 << 30);
  DevBuf work;
  if (own) {
    E->env.release();
    if (hipMalloc(&work.p, best.total) != hipSuccess) {
      work.p = nullptr;
      (void)hipGetLastError();
      return fail(HIPMPS_E_NOMEM, "reduced density matrix: workspace of " + std::to_string(best.total) + " bytes");
    }
    work.bytes = best.total;
  } else {
    RC_TRY(E->env.ensure(best.total));
  }
  struct Release {
    DevBuf& b;
    ~Release() { b.release(); }
  } release_work{work};
  char* cur = (char*)(own ? work.p : E->env.p);
  auto take = [&](size_t bytes) {
    char* p = cur;
    cur += align_up(bytes);
    return p;
  };
  char* p0 = take(best.es * 16);
  char* p1 = take(best.es * 16);
  char* xs = take(best.xs * 16);
  char* el = take(best.el * 16);
  char* er = take(best.er * 16);
  char* cbuf = take(nout * 16);
  char* obuf = take(nout * 16);
  // the site copies below take exactly best.copies bytes
  if ((size_t)(cur - (char*)(own ? work.p : E->env.p)) + best.copies > best.total)
    return fail(HIPMPS_E_UNSUPPORTED, "reduced density matrix: workspace layout exceeds its plan");
  std::vector<RdmSite> left, right;
  for (int q = 0; q < n; ++q) {
    const bool rev = q >= m, meas = pos[q] >= 0;
    const int cl = s->chi[q], cr = s->chi[q + 1], d = 2 * s->kap[q];
    const size_t bytes = site_elems(s, q) * s->es;
    const void* ver[2] = {s->sites[q].p, nullptr};
    if (meas)
      for (int b = 0; b < 2; ++b) {
        char* p = take(bytes);
        HIP_TRY(hipMemcpyAsync(p, s->sites[q].p, bytes, hipMemcpyDeviceToDevice, st));
        const double2 mb[4] = {make_double2(b == 0 ? 1.0 : 0.0, 0), make_double2(b == 1 ? 1.0 : 0.0, 0),
                               make_double2(0, 0), make_double2(0, 0)};
        HIP_TRY(launch_apply_1q(p, s->fp32, cl, s->kap[q] * cr, mb, st));
        ver[b] = p;
      }
    RdmSite S{};
    S.d = d;
    S.meas = meas;
    for (int v = 0; v < (meas ? 2 : 1); ++v) {
      if (!rev) {
        char* bra = take(bytes);  // [l][s][r] -> [s][l][r]
        HIP_TRY(launch_swap_mid(ver[v], bra, 1, cl, d, cr, nullptr, s->fp32, st));
        S.ket[v] = ver[v];
        S.bra[v] = bra;
        S.cl = cl;
        S.cr = cr;
      } else {
        char* bra = take(bytes);  // [l][s][r] -> [s][r][l]: the s-major copy with the bonds swapped
        HIP_TRY(launch_swap_mid(ver[v], bra, 1, cl, d * cr, 1, nullptr, s->fp32, st));
        char* ket = take(bytes);  // [s][r][l] -> [r][s][l]
        HIP_TRY(launch_swap_mid(bra, ket, 1, d, cr, cl, nullptr, s->fp32, st));
        S.ket[v] = ket;
        S.bra[v] = bra;
        S.cl = cr;
        S.cr = cl;
      }
    }
    (rev ? right : left).push_back(S);
  }
  std::reverse(right.begin(), right.end());
  RC_TRY(rdm_half(E, left, s->fp32, s->es, p0, p1, xs, el, best.xcap));
  RC_TRY(rdm_half(E, right, s->fp32, s->es, p0, p1, xs, er, best.xcap));
  // join: both environments label-major, [A][l'][l], then one GEMM over (l', l)
  const int cm = s->chi[m];
  HIP_TRY(launch_swap_mid(el, p0, 1, cm, (int)best.nl, cm, nullptr, 0, st));
  HIP_TRY(launch_swap_mid(er, p1, 1, cm, (int)best.nr, cm, nullptr, 0, st));
  RC_TRY(E->hdesc.ensure(sizeof(GemmTask)));
  RC_TRY(E->desc.ensure(sizeof(GemmTask)));
  GemmTask* g = (GemmTask*)E->hdesc.p;
  memset(g, 0, sizeof(GemmTask));
  g->A = p0;
  g->lda = cm * cm;
  g->M = (int)best.nl;
  g->K = cm * cm;
  g->B = p1;
  g->ldb = cm * cm;
  g->N = (int)best.nr;
  g->C = cbuf;
  g->ldc = (int)best.nr;
  g->tiles_m = (g->M + 31) / 32;
  g->tiles_n = (g->N + 31) / 32;
  HIP_TRY(hipMemcpyAsync(E->desc.p, g, sizeof(GemmTask), hipMemcpyHostToDevice, st));
  HIP_TRY(launch_gemm((GemmTask*)E->desc.p, 1, g->tiles_m * g->tiles_n, 0, 3, 0, 0, 0, st));
  // joined label A_L |A_R| + A_R: the right half's measured sites (in its sweep order, from
  // site n-1) in the low base-4 digits, then the left half's
  RdmPerm P{};
  P.k = k;
  int t = 0;
  for (int q = n - 1; q >= m; --q)
    if (pos[q] >= 0) P.pos[t++] = pos[q];
  for (int q = 0; q < m; ++q)
    if (pos[q] >= 0) P.pos[t++] = pos[q];
  HIP_TRY(launch_rdm_scatter((const double2*)cbuf, (double2*)obuf, P, st));
  HIP_TRY(hipMemcpyAsync(out, obuf, nout * 16, hipMemcpyDeviceToHost, st));
  HIP_TRY(hipStreamSynchronize(st));
  if (!own && best.total > ((size_t)256 << 20)) E->env.release();
  return 0;
}

// amplitude / slice chain
int chain_contract(hipmps_state_s* s, int rows0, const int8_t* d_bits, const int* slice_bits, std::vector<double2>& out) {
  if (is_purified(s))
    return fail(HIPMPS_E_UNSUPPORTED, "amplitudes are not defined for a purified (mixed) state");
  hipmps_engine E = s->e;
  const int n = s->n;
  // rows after each site
  std::vector<long long> rows(n + 1);
  rows[0] = rows0;
  for (int k = 0; k < n; ++k) rows[k + 1] = rows[k] * ((slice_bits && slice_bits[k] < 0) ? 2 : 1);
  size_t maxv = 1;
  for (int k = 0; k <= n; ++k) maxv = std::max(maxv, (size_t)rows[k] * s->chi[k]);
  if (maxv > ((size_t)1 << 31)) return fail(HIPMPS_E_UNSUPPORTED, "slice too large");
  size_t vb = align_up(maxv * 16);
  RC_TRY(E->env.ensure(2 * vb + 4096));
  char* base = (char*)E->env.p;
  double2* v0 = (double2*)base;
  double2* v1 = (double2*)(base + vb);
  // v0 = ones (rows0 x 1)
  {
    std::vector<double2> ones(rows0, make_double2(1.0, 0.0));
    HIP_TRY(hipMemcpyAsync(v0, ones.data(), rows0 * sizeof(double2), hipMemcpyHostToDevice, E->stream));
    HIP_TRY(hipStreamSynchronize(E->stream));
  }
  RC_TRY(E->hdesc.ensure(sizeof(ChainTask)));
  RC_TRY(E->desc.ensure(sizeof(ChainTask)));
  for (int k = 0; k < n; ++k) {
    ChainTask* h = (ChainTask*)E->hdesc.p;
    memset(h, 0, sizeof(*h));
    h->vin = v0;
    h->vout = v1;
    h->A = s->sites[k].p;
    h->cl = s->chi[k];
    h->cr = s->chi[k + 1];
    h->rows_in = (int)rows[k];
    h->rows_out = (int)rows[k + 1];
    h->bits = d_bits;
    h->n = n;
    h->site = k;
    h->mode = slice_bits ? 1 : 0;
    h->slice_bit = slice_bits ? slice_bits[k] : 0;
    h->block_begin = 0;
    h->blocks = (int)rows[k + 1] * ((h->cr + 255) / 256);
    HIP_TRY(hipMemcpyAsync(E->desc.p, h, sizeof(ChainTask), hipMemcpyHostToDevice, E->stream));
    HIP_TRY(launch_chain((const ChainTask*)E->desc.p, 1, h->blocks, s->fp32, E->stream));
    HIP_TRY(hipStreamSynchronize(E->stream));
    std::swap(v0, v1);
  }
  out.resize(rows[n]);
  HIP_TRY(hipMemcpyAsync(out.data(), v0, rows[n] * sizeof(double2), hipMemcpyDeviceToHost, E->stream));
  HIP_TRY(hipStreamSynchronize(E->stream));
  return 0;
}

}  // namespace

// ======================================================================================
// C ABI
// ======================================================================================
extern "C" {

const char* hipmps_last_error(void) { return t_err.c_str(); }
const char* hipmps_version(void) { return "hipmps 0.1 (gfx950)"; }

int hipmps_engine_create(const hipmps_engine_opts* opts, hipmps_engine* out) {
  if (!out) return fail(HIPMPS_E_ARG, "null output");
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
    (void)hipGetLastError();
    return fail(HIPMPS_E_DEVICE, "no HIP device available");
  }
  auto* e = new hipmps_engine_s();
  e->device = opts ? opts->device : 0;
  if (e->device < 0 || e->device >= ndev) {
    delete e;
    return fail(HIPMPS_E_ARG, "device ordinal out of range");
  }
  if (hipSetDevice(e->device) != hipSuccess || hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess) {
    delete e;
    return fail(HIPMPS_E_DEVICE, "stream creation failed");
  }
  size_t freeb = 0, totalb = 0;
  (void)hipMemGetInfo(&freeb, &totalb);
  e->budget = (opts && opts->hbm_budget_bytes) ? opts->hbm_budget_bytes : (size_t)(0.9 * (double)totalb);
  if (opts && opts->jacobi_max_sweeps > 0) e->max_sweeps = opts->jacobi_max_sweeps;
  if (opts && opts->jacobi_tol > 0) e->tol = opts->jacobi_tol;
  if (opts && opts->jacobi_tol_mixed > 0) e->tol_mixed = opts->jacobi_tol_mixed;
  for (auto& ev : e->ev) (void)hipEventCreate(&ev);
  if (const char* v = getenv("HIPMPS_GRAM32_SWEEPS")) e->gram32_sweeps = std::max(0, atoi(v));
  if (const char* v = getenv("HIPMPS_DEBUG_SWEEPS")) e->debug_sweeps = atoi(v);
  if (const char* v = getenv("HIPMPS_JACOBI_LANES"))
    e->jacobi_lanes = std::max(1, std::min(hipmps_engine_s::MAX_LANES, atoi(v)));
  e->lane_stream[0] = e->stream;
  bool ok = true;
  for (int l = 1; l < hipmps_engine_s::MAX_LANES; ++l)
    ok = ok && hipStreamCreateWithFlags(&e->lane_stream[l], hipStreamNonBlocking) == hipSuccess;
  for (auto& ev : e->lane_ev) ok = ok && hipEventCreateWithFlags(&ev, hipEventDisableTiming) == hipSuccess;
  if (hipEventCreateWithFlags(&e->stage_ev, hipEventDisableTiming) != hipSuccess) {
    e->stage_ev = nullptr;
    (void)hipGetLastError();
    (void)hipmps_engine_destroy(e);
    return fail(HIPMPS_E_DEVICE, "event creation failed");
  }
  bool lq_ok = true;
  for (auto& ev : e->lq_ev) lq_ok = lq_ok && hipEventCreateWithFlags(&ev, hipEventDisableTiming) == hipSuccess;
  if (!lq_ok) {
    (void)hipGetLastError();
    e->lq_lookahead = 0;
  }
  if (const char* v = getenv("HIPMPS_LQ_LOOKAHEAD")) e->lq_lookahead = atoi(v) != 0;
  if (const char* v = getenv("HIPMPS_TEST_STAGE_DELAY_US")) {
    int khz = 0;  // wall-clock rate
    if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, e->device) != hipSuccess || khz <= 0) {
      (void)hipGetLastError();
      khz = 100000;
    }
    e->stage_delay_ticks = (long long)std::max(0, std::min(100000, atoi(v))) * khz / 1000;
  }
  if (const char* v = getenv("HIPMPS_WAVE_GROUPS")) e->wave_groups = std::max(1, std::min(2, atoi(v)));
  if (!ok) {
    (void)hipGetLastError();
    e->jacobi_lanes = 1;
  }
  if (const int rc = engine_scratch_setup(e)) {
    const std::string msg = t_err;
    (void)hipmps_engine_destroy(e);
    return fail(rc, msg);
  }
  *out = e;
  return 0;
}

int hipmps_engine_destroy(hipmps_engine e) {
  if (!e) return 0;
  if (e->big) {  // stop the background lane before its circuits are freed
    {
      std::lock_guard<std::mutex> lk(e->big_mu);
      e->big_stop = true;
    }
    e->big_cv.notify_all();
    if (e->big_worker.joinable()) e->big_worker.join();
  }
  if (e->grp) {  // the group worker is idle between steps; stop it before the circuits are freed
    {
      std::lock_guard<std::mutex> lk(e->grp_mu);
      e->grp_stop = true;
    }
    e->grp_cv.notify_all();
    if (e->grp_worker.joinable()) e->grp_worker.join();
  }
  (void)hipSetDevice(e->device);
  (void)hipStreamSynchronize(e->stream);
  // every submitted circuit that was not polled yet is owned through by_ticket
  for (auto& kv : e->by_ticket) {
    if (kv.second->st) hipmps_state_destroy(kv.second->st);
    delete kv.second;
  }
  e->by_ticket.clear();
  e->running.clear();
  e->queue.clear();
  e->ws.release();
  e->desc.release();
  e->small.release();
  e->env.release();
  e->hdesc.release();
  e->hsmall.release();
  for (auto& ev : e->ev)
    if (ev) (void)hipEventDestroy(ev);
  for (auto& ev : e->lane_ev)
    if (ev) (void)hipEventDestroy(ev);
  for (auto& ev : e->lq_ev)
    if (ev) (void)hipEventDestroy(ev);
  if (e->stage_ev) (void)hipEventDestroy(e->stage_ev);
  for (int l = 1; l < hipmps_engine_s::MAX_LANES; ++l)
    if (e->lane_stream[l]) {
      (void)hipStreamSynchronize(e->lane_stream[l]);
      (void)hipStreamDestroy(e->lane_stream[l]);
    }
  (void)hipStreamDestroy(e->stream);
  if (e->big) hipmps_engine_destroy(e->big);
  if (e->grp) hipmps_engine_destroy(e->grp);
  delete e;
  return 0;
}

int hipmps_engine_set_lq_precondition(hipmps_engine e, int min_rank) {
  if (!e) return fail(HIPMPS_E_ARG, "null engine");
  if (min_rank < 0) return fail(HIPMPS_E_ARG, "min_rank must be >= 0");
  e->lq_min_rank = min_rank;
  if (e->grp) e->grp->lq_min_rank = min_rank;
  return 0;
}

int hipmps_engine_get_tuning(hipmps_engine e, hipmps_tuning* out) {
  if (!e || !out) return fail(HIPMPS_E_ARG, "null argument");
  out->jacobi_lanes = e->jacobi_lanes;
  out->lane_min_pairs = e->lane_min_pairs;
  out->gram32_sweeps = e->gram32_sweeps;
  out->lq_min_rank = e->lq_min_rank;
  out->big_lane_min_chi = e->big_min_chi;
  out->rdm_block_elems = e->rdm_block_elems;
  out->jacobi_max_sweeps = e->max_sweeps;
  out->wave_groups = e->wave_groups;
  return 0;
}

int hipmps_engine_set_tuning(hipmps_engine e, const hipmps_tuning* t) {
  if (!e || !t) return fail(HIPMPS_E_ARG, "null argument");
  if (t->jacobi_lanes < 1 || t->jacobi_lanes > hipmps_engine_s::MAX_LANES || t->lane_min_pairs < 0 ||
      t->gram32_sweeps < 0 || t->lq_min_rank < 0 || t->big_lane_min_chi < 0 || t->rdm_block_elems < 1 ||
      t->jacobi_max_sweeps < 1 || t->jacobi_max_sweeps > 1000 || t->wave_groups < 1 || t->wave_groups > 2)
    return fail(HIPMPS_E_ARG, "tuning values out of range");
  std::lock_guard<std::mutex> lock(e->mu);
  int made = 1;  // lanes whose stream exists
  while (made < hipmps_engine_s::MAX_LANES && e->lane_stream[made]) ++made;
  e->jacobi_lanes = std::min(t->jacobi_lanes, made);
  e->lane_min_pairs = t->lane_min_pairs;
  e->gram32_sweeps = t->gram32_sweeps;
  e->lq_min_rank = t->lq_min_rank;
  e->big_min_chi = t->big_lane_min_chi;
  e->rdm_block_elems = t->rdm_block_elems;
  e->max_sweeps = t->jacobi_max_sweeps;
  e->wave_groups = t->wave_groups;
  if (e->grp) {  // the group engine follows, with no lane or groups of its own
    hipmps_tuning g = *t;
    g.big_lane_min_chi = 0;
    g.wave_groups = 1;
    const int rc = hipmps_engine_set_tuning(e->grp, &g);
    if (rc) return rc;
  }
  return 0;
}

int hipmps_engine_set_timing(hipmps_engine e, int enabled) {
  if (!e) return fail(HIPMPS_E_ARG, "null engine");
  e->timing = enabled != 0;
  if (e->grp) e->grp->timing = e->timing;
  return 0;
}

int hipmps_engine_profile(hipmps_engine e, hipmps_profile_t* out, int reset) {
  if (!e) return fail(HIPMPS_E_ARG, "null engine");
  if (out) {
    *out = e->prof;
    if (e->grp) {  // both wave groups; their phase times add up although the groups overlap in time
      const hipmps_profile_t& g = e->grp->prof;
      out->ms_theta += g.ms_theta;
      out->ms_jacobi += g.ms_jacobi;
      out->ms_split += g.ms_split;
      out->ms_other += g.ms_other;
      out->ms_jacobi_rounds += g.ms_jacobi_rounds;
      out->jacobi_launches += g.jacobi_launches;
      out->jacobi_pairs_rotated += g.jacobi_pairs_rotated;
      out->jacobi_pairs_visited += g.jacobi_pairs_visited;
      out->jacobi_flops += g.jacobi_flops;
      out->waves += g.waves;
      out->wave_tasks += g.wave_tasks;
      out->jacobi_sweeps += g.jacobi_sweeps;
      out->svds_large += g.svds_large;
      out->sweeps_large += g.sweeps_large;
      out->ms_lq += g.ms_lq;
      for (int k = 0; k < 14; ++k) out->applies_by_rank[k] += g.applies_by_rank[k];
    }
  }
  if (reset) {
    e->prof = hipmps_profile_t{};
    if (e->grp) e->grp->prof = hipmps_profile_t{};
  }
  return 0;
}

int hipmps_state_create(hipmps_engine e, int n_qubits, const hipmps_trunc_opts* t, hipmps_state* out) {
  if (!e || !out || n_qubits < 1) return fail(HIPMPS_E_ARG, "bad arguments to hipmps_state_create");
  *out = nullptr;
  RC_TRY(set_device(e));
  auto* s = new hipmps_state_s();
  s->e = e;
  s->n = n_qubits;
  if (t) s->t = *t;
  else {
    s->t.svd_cutoff = DBL_MIN;
    s->t.max_bond_dim = 0;
  }
  s->fp32 = s->t.precision == HIPMPS_MIXED;
  s->es = s->fp32 ? 8 : 16;
  s->sites.resize(n_qubits);
  int rc = state_init_product(s);
  if (rc) {
    hipmps_state_destroy(s);
    return rc;
  }
  *out = s;
  return 0;
}

int hipmps_state_destroy(hipmps_state s) {
  if (!s) return 0;
  (void)hipSetDevice(s->e->device);
  (void)hipStreamSynchronize(s->e->stream);
  release_deferred(s);
  for (auto& S : s->sites) {
    if (S.p) (void)hipFree(S.p);
    s->e->used -= S.cap * s->es;
  }
  delete s;
  return 0;
}

int hipmps_state_reset(hipmps_state s) {
  if (!s) return fail(HIPMPS_E_ARG, "null state");
  RC_TRY(set_device(s->e));
  return state_init_product(s);
}

int hipmps_apply_1q(hipmps_state s, int q, const double m[8]) {
  if (!s || !m) return fail(HIPMPS_E_ARG, "null argument");
  hipmps_gate g{};
  g.nq = 1;
  g.q0 = q;
  memcpy(g.m, m, 8 * sizeof(double));
  RC_TRY(set_device(s->e));
  return apply_gate_list(s, 1, &g);
}

int hipmps_apply_2q(hipmps_state s, int q0, int q1, const double m[32]) {
  if (!s || !m) return fail(HIPMPS_E_ARG, "null argument");
  hipmps_gate g{};
  g.nq = 2;
  g.q0 = q0;
  g.q1 = q1;
  memcpy(g.m, m, 32 * sizeof(double));
  RC_TRY(set_device(s->e));
  return apply_gate_list(s, 1, &g);
}

int hipmps_apply_gates(hipmps_state s, int n, const hipmps_gate* gates) {
  if (!s || (n > 0 && !gates)) return fail(HIPMPS_E_ARG, "null argument");
  RC_TRY(set_device(s->e));
  return apply_gate_list(s, n, gates);
}

int hipmps_flush(hipmps_state s) {
  if (!s) return fail(HIPMPS_E_ARG, "null state");
  RC_TRY(state_flush(s));
  if (s->status == HIPMPS_E_CONVERGENCE)
    return fail(s->status, std::to_string(s->nonconverged) + " SVD(s) reached the Jacobi sweep cap unconverged");
  if (s->status != 0) return fail(s->status, "a tensor of the state held NaN or Inf");
  return 0;
}

int hipmps_apply_kraus(hipmps_state s, int q, int n_ops, const double* mats) {
  if (!s || !mats || q < 0 || q >= s->n || n_ops < 1 || n_ops > HIPMPS_MAX_KRAUS_OPS)
    return fail(HIPMPS_E_ARG, "bad arguments");
  if (!all_finite(mats, (size_t)8 * n_ops)) return fail(HIPMPS_E_NUMERIC, "Kraus operator holds NaN or Inf");
  RC_TRY(set_device(s->e));
  if (s->kap[q] * n_ops > 4096) {
    RC_TRY(state_flush(s));  // the Kraus extent of q after the channels already queued
    if (s->kap[q] * n_ops > 4096) return fail(HIPMPS_E_UNSUPPORTED, "Kraus leg above 4096 rows");
  }
  if (s->cbusy[q]) RC_TRY(state_run_wave(s));  // one queued channel per site
  hipmps_state_s::QK ch;
  ch.q = q;
  ch.n_ops = n_ops;
  for (int t = 0; t < 4 * n_ops; ++t) ch.K[t] = cd(mats[2 * t], mats[2 * t + 1]);
  if (s->has_pend[q]) {  // 1-qubit gates still pending on q act before the channel: K_o P
    for (int o = 0; o < n_ops; ++o) mat_mul2(ch.K + 4 * o, s->pend[q].data(), ch.K + 4 * o);
    s->has_pend[q] = 0;
    s->pend[q] = {cd(1), cd(0), cd(0), cd(1)};
  }
  s->cwave.push_back(ch);
  s->cbusy[q] = 1;
  return 0;
}

int hipmps_queue_gates(hipmps_state s, int n, const hipmps_gate* gates, int* n_queued) {
  if (!s || !n_queued || (n > 0 && !gates)) return fail(HIPMPS_E_ARG, "null argument");
  *n_queued = 0;
  RC_TRY(set_device(s->e));
  for (int k = 0; k < n; ++k) {
    const hipmps_gate& g = gates[k];
    if (g.nq == 2 && g.q0 >= 0 && g.q0 < s->n && g.q1 >= 0 && g.q1 < s->n && std::abs(g.q0 - g.q1) == 1) {
      const int i = std::min(g.q0, g.q1);
      if (s->busy[i] || s->busy[i + 1] || s->cbusy[i] || s->cbusy[i + 1]) return 0;  // after the open waves
    }
    RC_TRY(apply_gate_list(s, 1, &g));
    ++*n_queued;
  }
  return 0;
}

int hipmps_queue_kraus(hipmps_state s, int q, int n_ops, const double* mats, int* queued) {
  if (!s || !queued || q < 0 || q >= s->n) return fail(HIPMPS_E_ARG, "bad arguments");
  *queued = 0;
  if (s->cbusy[q]) return 0;  // one queued channel per site
  RC_TRY(hipmps_apply_kraus(s, q, n_ops, mats));
  *queued = 1;
  return 0;
}

int hipmps_run_waves(hipmps_state* states, int n) {
  if (n < 0 || (n > 0 && !states)) return fail(HIPMPS_E_ARG, "bad arguments");
  if (n == 0) return 0;
  hipmps_engine E = states[0] ? states[0]->e : nullptr;
  for (int k = 0; k < n; ++k) {
    if (!states[k]) return fail(HIPMPS_E_ARG, "null state");
    if (states[k]->e != E) return fail(HIPMPS_E_ARG, "states of different engines");
    for (int j = 0; j < k; ++j)
      if (states[j] == states[k]) return fail(HIPMPS_E_ARG, "a state is listed twice");
  }
  RC_TRY(set_device(E));
  std::vector<WaveItem> items;
  std::vector<hipmps_state_s*> list(states, states + n);
  for (auto* st : list)
    for (auto& q : st->wave) {
      WaveItem it{};
      it.s = st;
      it.i = q.i;
      it.G = q.G;
      items.push_back(it);
    }
  int rc = items.empty() ? 0 : run_wave(E, items);
  for (auto* st : list) {
    st->wave.clear();
    std::fill(st->busy.begin(), st->busy.end(), 0);
  }
  if (rc != 0) return rc;
  return run_kraus_waves(E, list);
}

int hipmps_kraus_dims(hipmps_state s, int* out) {
  if (!s || !out) return fail(HIPMPS_E_ARG, "null argument");
  RC_TRY(state_flush(s));
  for (int k = 0; k < s->n; ++k) out[k] = s->kap[k];
  return 0;
}

int hipmps_bond_dims(hipmps_state s, int* out) {
  if (!s || !out) return fail(HIPMPS_E_ARG, "null argument");
  RC_TRY(state_flush(s));
  for (int k = 0; k <= s->n; ++k) out[k] = s->chi[k];
  return 0;
}

int hipmps_norm(hipmps_state s, double* out) {
  if (!s || !out) return fail(HIPMPS_E_ARG, "null argument");
  RC_TRY(state_flush(s));
  return measure_parity(s, {}, out);
}

int hipmps_stats(hipmps_state s, hipmps_stats_t* out) {
  if (!s || !out) return fail(HIPMPS_E_ARG, "null argument");
  RC_TRY(state_flush(s));
  memset(out, 0, sizeof(*out));
  RC_TRY(measure_parity(s, {}, &out->norm));
  out->discarded_weight = s->discarded;
  out->fidelity_est = s->fid;
  out->max_bond = s->maxb;
  out->n_2q = s->n2q;
  out->n_1q = s->n1q;
  out->jacobi_sweeps_max = s->sweeps_max;
  out->jacobi_sweeps_total = s->sweeps_total;
  out->device_bytes = s->bytes;
  out->norm_warnings = s->norm_warnings;
  out->svds_nonconverged = s->nonconverged;
  out->device_bytes_peak = s->peak;
  return s->status;
}

int hipmps_expval_z_parity(hipmps_state s, const int* qubits, int k, double* out) {
  if (!s || !out || (k > 0 && !qubits)) return fail(HIPMPS_E_ARG, "null argument");
  RC_TRY(state_flush(s));
  std::vector<int> zmask(s->n, 0);
  for (int t = 0; t < k; ++t) {
    if (qubits[t] < 0 || qubits[t] >= s->n) return fail(HIPMPS_E_ARG, "qubit out of range");
    zmask[qubits[t]] ^= 1;  // parity over the listed bits (:618-629)
  }
  return measure_parity(s, zmask, out);
}

int hipmps_expval_z_all(hipmps_state s, double* out) {
  if (!s || !out) return fail(HIPMPS_E_ARG, "null argument");
  RC_TRY(state_flush(s));
  return measure_z_all(s, out);
}

int hipmps_expval_zz(hipmps_state s, const int* pairs, int k, double* out) {
  if (!s || !out || (k > 0 && !pairs)) return fail(HIPMPS_E_ARG, "null argument");
  RC_TRY(state_flush(s));
  if (k == 0) return 0;
  return measure_zz(s, pairs, k, out);
}

int hipmps_reduced_density_matrix(hipmps_state s, const int* qubits, int k, double* out, size_t cap) {
  if (!s || !qubits || !out || k < 1) return fail(HIPMPS_E_ARG, "bad arguments");
  if (k > RDM_MAX_QUBITS) return fail(HIPMPS_E_UNSUPPORTED, "reduced density matrix of 1 to 15 qubits");
  if (cap < ((size_t)1 << (2 * k))) return fail(HIPMPS_E_ARG, "output buffer too small");
  RC_TRY(state_flush(s));
  return measure_rdm(s, std::vector<int>(qubits, qubits + k), out);
}

int hipmps_expval_diag(hipmps_state s, const double* diag, double* out) {
  if (!s || !diag || !out) return fail(HIPMPS_E_ARG, "null argument");
  RC_TRY(state_flush(s));
  std::vector<double> d(diag, diag + 2 * s->n);
  return measure_diag(s, d, out);
}

int hipmps_sample_sequential(hipmps_state s, const int* qubits, int k, int shots, const double* draws, char* out) {
  if (!s || !qubits || k < 1 || shots < 0 || (shots > 0 && (!draws || !out))) return fail(HIPMPS_E_ARG, "bad arguments");
  for (int i = 0; i < k; ++i) {
    if (qubits[i] < 0 || qubits[i] >= s->n) return fail(HIPMPS_E_ARG, "measured qubit out of range");
    if (i > 0 && qubits[i] <= qubits[i - 1])
      return fail(HIPMPS_E_UNSUPPORTED, "measured qubits must be in increasing order");
  }
  if (shots == 0) return 0;
  RC_TRY(state_flush(s));
  return sample_measure(s, qubits, k, shots, draws, out);
}

int hipmps_state_clone(hipmps_state src, hipmps_state* out) {
  if (!src || !out) return fail(HIPMPS_E_ARG, "null argument");
  *out = nullptr;
  RC_TRY(state_flush(src));
  hipmps_state dst = nullptr;
  RC_TRY(hipmps_state_create(src->e, src->n, &src->t, &dst));
  for (int k = 0; k < src->n; ++k) {
    size_t n = site_elems(src, k);
    int rc = ensure_site(dst, k, n, false, nullptr);
    if (rc) {
      hipmps_state_destroy(dst);
      return rc;
    }
    if (hipMemcpyAsync(dst->sites[k].p, src->sites[k].p, n * src->es, hipMemcpyDeviceToDevice, src->e->stream) !=
        hipSuccess) {
      hipmps_state_destroy(dst);
      return fail(HIPMPS_E_DEVICE, "site copy failed");
    }
  }
  HIP_TRY(hipStreamSynchronize(src->e->stream));
  dst->chi = src->chi;
  dst->kap = src->kap;
  dst->discarded = src->discarded;
  dst->fid = src->fid;
  dst->maxb = src->maxb;
  dst->n2q = src->n2q;
  dst->n1q = src->n1q;
  dst->sweeps_max = src->sweeps_max;
  dst->sweeps_total = src->sweeps_total;
  dst->status = src->status;
  dst->norm_warnings = src->norm_warnings;
  dst->nonconverged = src->nonconverged;
  *out = dst;
  return 0;
}

int hipmps_amplitudes(hipmps_state s, const int8_t* bits, int k, double* out) {
  if (!s || !bits || !out || k <= 0) return fail(HIPMPS_E_ARG, "bad arguments");
  RC_TRY(state_flush(s));
  for (size_t t = 0; t < (size_t)k * s->n; ++t)
    if (bits[t] != 0 && bits[t] != 1) return fail(HIPMPS_E_ARG, "bits must be 0 or 1");
  hipmps_engine E = s->e;
  // bits live after the chain buffers; upload into a dedicated small buffer
  DevBuf db;
  RC_TRY(db.ensure((size_t)k * s->n));
  HIP_TRY(hipMemcpy(db.p, bits, (size_t)k * s->n, hipMemcpyHostToDevice));
  std::vector<double2> v;
  int rc = chain_contract(s, k, (const int8_t*)db.p, nullptr, v);
  db.release();
  (void)E;
  if (rc) return rc;
  for (int t = 0; t < k; ++t) {
    out[2 * t] = v[t].x;
    out[2 * t + 1] = v[t].y;
  }
  return 0;
}

// hipmps_amplitudes with every intermediate in this device's copy of `scratch`: no device or
// host allocation, the same kernel on the same inputs (results equal bit for bit).  Layout in
// the scratch: n chain descriptors | k x n bits | two k x max(chi) chain vectors.
int hipmps_amplitudes_noalloc(hipmps_state s, const int8_t* bits, int k, double* out) {
  if (!s || !bits || !out || k <= 0) return fail(HIPMPS_E_ARG, "bad arguments");
  hipmps_engine E = s->e;
  if (!E->scratch) return fail(HIPMPS_E_UNSUPPORTED, "no static scratch on this device");
  if (is_purified(s))
    return fail(HIPMPS_E_UNSUPPORTED, "amplitudes are not defined for a purified (mixed) state");
  bool queued = !s->wave.empty() || !s->cwave.empty();
  for (int q = 0; q < s->n; ++q) queued = queued || s->has_pend[q];
  if (queued)
    return fail(HIPMPS_E_ARG, "the state has queued gates; hipmps_flush it first (a flush may grow site tensors)");
  const int n = s->n;
  for (size_t t = 0; t < (size_t)k * n; ++t)
    if (bits[t] != 0 && bits[t] != 1) return fail(HIPMPS_E_ARG, "bits must be 0 or 1");
  int maxchi = 1;
  for (int b = 0; b <= n; ++b) maxchi = std::max(maxchi, s->chi[b]);
  const size_t desc_b = align_up((size_t)n * sizeof(ChainTask));
  const size_t bits_b = align_up((size_t)k * n);
  const size_t vec_b = align_up((size_t)k * maxchi * sizeof(double2));
  const size_t need = desc_b + bits_b + 2 * vec_b;
  if (need > HIPMPS_NOALLOC_SCRATCH_BYTES)
    return fail(HIPMPS_E_NOMEM, "hipmps_amplitudes_noalloc needs " + std::to_string(need) + " bytes for " +
                                    std::to_string(k) + " bit strings at bond " + std::to_string(maxchi) +
                                    "; the static scratch holds " + std::to_string(HIPMPS_NOALLOC_SCRATCH_BYTES));
  RC_TRY(set_device(E));
  char* base = (char*)E->scratch;
  ChainTask* d_desc = (ChainTask*)base;
  int8_t* d_bits = (int8_t*)(base + desc_b);
  double2* v0 = (double2*)(base + desc_b + bits_b);
  double2* v1 = (double2*)(base + desc_b + bits_b + vec_b);
  // descriptors of all sites, a block at a time through a stack buffer (synchronous copies)
  constexpr int BLK = 32;
  ChainTask h[BLK];
  for (int k0 = 0; k0 < n; k0 += BLK) {
    const int m = std::min(BLK, n - k0);
    for (int j = 0; j < m; ++j) {
      const int site = k0 + j;
      ChainTask& c = h[j];
      memset(&c, 0, sizeof(c));
      c.vin = (site & 1) ? v1 : v0;
      c.vout = (site & 1) ? v0 : v1;
      c.A = s->sites[site].p;
      c.cl = s->chi[site];
      c.cr = s->chi[site + 1];
      c.rows_in = k;
      c.rows_out = k;
      c.bits = d_bits;
      c.n = n;
      c.site = site;
      c.mode = 0;
      c.block_begin = 0;
      c.blocks = k * ((c.cr + 255) / 256);
    }
    HIP_TRY(hipMemcpy(d_desc + k0, h, (size_t)m * sizeof(ChainTask), hipMemcpyHostToDevice));
  }
  HIP_TRY(hipMemcpy(d_bits, bits, (size_t)k * n, hipMemcpyHostToDevice));
  HIP_TRY(launch_fill_c(v0, (size_t)k, make_double2(1.0, 0.0), E->stream));  // rows of ones (chi_0 = 1)
  for (int site = 0; site < n; ++site) {
    const int blocks = k * ((s->chi[site + 1] + 255) / 256);
    HIP_TRY(launch_chain(d_desc + site, 1, blocks, s->fp32, E->stream));
  }
  const double2* res = (n & 1) ? v1 : v0;
  HIP_TRY(hipMemcpyAsync(out, res, (size_t)k * sizeof(double2), hipMemcpyDeviceToHost, E->stream));
  HIP_TRY(hipStreamSynchronize(E->stream));
  return 0;
}

int hipmps_wavefunction_slice(hipmps_state s, const int* bits, double* out, size_t cap) {
  if (!s || !bits || !out) return fail(HIPMPS_E_ARG, "bad arguments");
  RC_TRY(state_flush(s));
  int open = 0;
  for (int k = 0; k < s->n; ++k) {
    if (bits[k] < -1 || bits[k] > 1) return fail(HIPMPS_E_ARG, "bits must be 0, 1 or -1");
    open += bits[k] < 0;
  }
  if (open > 30) return fail(HIPMPS_E_UNSUPPORTED, "too many open legs");
  if (cap < ((size_t)1 << open)) return fail(HIPMPS_E_ARG, "output buffer too small");
  std::vector<double2> v;
  RC_TRY(chain_contract(s, 1, nullptr, bits, v));
  for (size_t t = 0; t < v.size(); ++t) {
    out[2 * t] = v[t].x;
    out[2 * t + 1] = v[t].y;
  }
  return 0;
}

int hipmps_state_vector(hipmps_state s, double* out, size_t cap) {
  if (!s || !out) return fail(HIPMPS_E_ARG, "bad arguments");
  if (s->n > 26) return fail(HIPMPS_E_UNSUPPORTED, "state vector limited to 26 qubits");
  std::vector<int> bits(s->n, -1);
  return hipmps_wavefunction_slice(s, bits.data(), out, cap);
}

int hipmps_get_site(hipmps_state s, int k, double* out, size_t cap) {
  if (!s || !out || k < 0 || k >= s->n) return fail(HIPMPS_E_ARG, "bad arguments");
  RC_TRY(state_flush(s));
  size_t n = site_elems(s, k);
  if (cap < n) return fail(HIPMPS_E_ARG, "output buffer too small");
  hipmps_engine E = s->e;
  if (s->fp32) {
    std::vector<float> h(2 * n);
    HIP_TRY(hipMemcpy(h.data(), s->sites[k].p, n * 8, hipMemcpyDeviceToHost));
    for (size_t i = 0; i < 2 * n; ++i) out[i] = h[i];
  } else {
    HIP_TRY(hipMemcpy(out, s->sites[k].p, n * 16, hipMemcpyDeviceToHost));
  }
  (void)E;
  return 0;
}

int hipmps_set_site(hipmps_state s, int k, int chi_l, int chi_r, const double* in) {
  if (!s || !in || k < 0 || k >= s->n || chi_l < 1 || chi_r < 1) return fail(HIPMPS_E_ARG, "bad arguments");
  if ((k == 0 && chi_l != 1) || (k == s->n - 1 && chi_r != 1)) return fail(HIPMPS_E_ARG, "boundary bonds must be 1");
  size_t n = (size_t)chi_l * 2 * chi_r;
  if (!all_finite(in, 2 * n)) return fail(HIPMPS_E_NUMERIC, "site tensor holds NaN or Inf");
  RC_TRY(state_flush(s));
  RC_TRY(ensure_site(s, k, n, false, nullptr));
  s->chi[k] = chi_l;
  s->chi[k + 1] = chi_r;
  s->kap[k] = 1;
  s->maxb = std::max(s->maxb, std::max(chi_l, chi_r));
  return upload_site_host(s, k, reinterpret_cast<const cd*>(in), n);
}

// ---------------------------------------------------------------------------------------
// batches and serving
// ---------------------------------------------------------------------------------------
}  // extern "C"

namespace {

int finish_circuit(PendingCircuit* pc) {
  hipmps_state_s* s = pc->st;
  hipmps_result* r = pc->res;
  RC_TRY(state_flush(s));
  if (r) {
    if (r->expval_z) RC_TRY(measure_z_all(s, r->expval_z, &r->norm));
    else RC_TRY(measure_parity(s, {}, &r->norm));
    if (r->expval_zz && !pc->zz.empty()) RC_TRY(measure_zz(s, pc->zz.data(), (int)pc->zz.size() / 2, r->expval_zz));
    r->discarded_weight = s->discarded;
    r->fidelity_est = s->fid;
    r->max_bond = s->maxb;
    r->status = s->status;
    r->latency_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - pc->t0).count();
  }
  pc->finished = true;
  return 0;
}

// A circuit that fails on its own (a site it cannot grow within the budget, a gate the host
// refuses) ends with that status; the circuits beside it go on.
void fail_circuit(PendingCircuit* pc, int rc) {
  if (pc->res) {
    pc->res->status = rc;
    pc->res->latency_ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - pc->t0).count();
  }
  pc->finished = true;
}

std::vector<WaveItem> wave_items(hipmps_state_s* s) {
  std::vector<WaveItem> v;
  for (auto& q : s->wave) {
    WaveItem it{};
    it.s = s;
    it.i = q.i;
    it.G = q.G;
    v.push_back(it);
  }
  return v;
}

// Feed each running circuit up to its next wave conflict, then run all waves together.
// Memory: a wave grows its sites all or nothing (run_wave), so when the joint wave does not fit,
// the circuits' waves run one by one; a circuit whose own wave does not fit waits (its wave stays
// queued) while others progress and free memory, and fails with HIPMPS_E_NOMEM only when no
// circuit could progress, the largest first.
int advance_all(hipmps_engine E, std::vector<PendingCircuit*>& run) {
  std::vector<WaveItem> items;
  std::vector<PendingCircuit*> live;
  for (auto* pc : run) {
    if (pc->finished) continue;
    hipmps_state_s* s = pc->st;
    int rc = 0;
    while (pc->cursor < pc->gates.size()) {
      const hipmps_gate& g = pc->gates[pc->cursor];
      if (g.nq == 2) {
        const int i = std::min(g.q0, g.q1);
        if (i >= 0 && i + 1 < s->n && (s->busy[i] || s->busy[i + 1])) break;  // next wave
      }
      rc = apply_gate_list(s, 1, &g);
      if (rc != 0) break;
      pc->cursor++;
    }
    if (rc != 0) {
      fail_circuit(pc, rc);
      continue;
    }
    live.push_back(pc);
    for (auto& it : wave_items(s)) items.push_back(it);
  }
  std::vector<char> ran(live.size(), 1);
  int rc = run_wave(E, items);
  if (rc == HIPMPS_E_NOMEM) {
    bool progress = false;
    for (size_t k = 0; k < live.size(); ++k) {
      std::vector<WaveItem> one = wave_items(live[k]->st);
      const int r = run_wave(E, one);
      if (r == HIPMPS_E_NOMEM) ran[k] = 0;
      else if (r != 0) return r;
      else progress = true;
    }
    if (!progress) {  // nothing can move: the largest stalled circuit gives way
      size_t worst = 0;
      for (size_t k = 1; k < live.size(); ++k)
        if (live[k]->st->bytes > live[worst]->st->bytes) worst = k;
      fail_circuit(live[worst], HIPMPS_E_NOMEM);
    }
  } else if (rc != 0) {
    return rc;
  }
  for (size_t k = 0; k < live.size(); ++k) {
    PendingCircuit* pc = live[k];
    if (pc->finished || !ran[k]) continue;
    hipmps_state_s* s = pc->st;
    s->wave.clear();
    std::fill(s->busy.begin(), s->busy.end(), 0);
    if (pc->cursor >= pc->gates.size()) {
      const int r = finish_circuit(pc);
      if (r == HIPMPS_E_NOMEM) fail_circuit(pc, r);  // its measurement workspace does not fit
      else if (r != 0) return r;
    }
  }
  return 0;
}

// Host-side checks of a submitted circuit, so that a bad request is refused at the door and
// cannot stall the circuits it would run beside.
int validate_circuit(const hipmps_circuit& c) {
  if (c.n_qubits < 1) return fail(HIPMPS_E_ARG, "a circuit needs at least one qubit");
  if (c.n_gates < 0 || (c.n_gates > 0 && !c.gates)) return fail(HIPMPS_E_ARG, "bad gate list");
  if (c.n_zz < 0 || (c.n_zz > 0 && !c.zz_pairs)) return fail(HIPMPS_E_ARG, "bad <ZZ> pair list");
  if (c.trunc.precision != HIPMPS_FP64 && c.trunc.precision != HIPMPS_MIXED)
    return fail(HIPMPS_E_ARG, "unknown precision");
  const int n = c.n_qubits;
  for (int k = 0; k < c.n_gates; ++k) {
    const hipmps_gate& g = c.gates[k];
    const std::string at = "gate " + std::to_string(k) + ": ";
    if (g.nq == 1) {
      if (g.q0 < 0 || g.q0 >= n) return fail(HIPMPS_E_ARG, at + "qubit index out of range");
    } else if (g.nq == 2) {
      if (g.q0 < 0 || g.q0 >= n || g.q1 < 0 || g.q1 >= n || g.q0 == g.q1)
        return fail(HIPMPS_E_ARG, at + "qubit index out of range");
      if (std::abs(g.q0 - g.q1) != 1) return fail(HIPMPS_E_NONADJACENT, at + "2-qubit gate on non-neighbouring qubits");
    } else {
      return fail(HIPMPS_E_ARG, at + "a gate acts on 1 or 2 qubits");
    }
    if (!gate_finite(g)) return fail(HIPMPS_E_NUMERIC, at + "matrix holds NaN or Inf");
  }
  for (int k = 0; k < 2 * c.n_zz; ++k)
    if (c.zz_pairs[k] < 0 || c.zz_pairs[k] >= n) return fail(HIPMPS_E_ARG, "<ZZ> pair qubit out of range");
  return 0;
}

int start_circuit(hipmps_engine E, PendingCircuit* pc) {
  hipmps_state st = nullptr;
  RC_TRY(hipmps_state_create(E, pc->c.n_qubits, &pc->c.trunc, &st));
  pc->st = st;
  return 0;
}

// The background lane's worker: take newly admitted circuits, advance everything it owns by
// one wave, free the finished ones and publish them (done), and repeat while it has work.
void big_lane_loop(hipmps_engine E) {
  hipmps_engine B = E->big;
  (void)hipSetDevice(B->device);
  std::vector<PendingCircuit*> run;
  for (;;) {
    {
      std::unique_lock<std::mutex> lk(E->big_mu);
      E->big_cv.wait(lk, [&] { return E->big_stop || !E->big_inbox.empty() || !run.empty(); });
      if (E->big_stop) return;  // circuits still owned are freed by hipmps_engine_destroy
      while (!E->big_inbox.empty()) {
        run.push_back(E->big_inbox.front());
        E->big_inbox.pop_front();
      }
    }
    int rc = 0;
    for (auto* pc : run)
      if (!pc->st && rc == 0) rc = start_circuit(B, pc);
    if (rc == 0) rc = advance_all(B, run);
    std::string err = rc ? hipmps_last_error() : "";
    std::vector<PendingCircuit*> keep;
    for (auto* pc : run) {
      if (rc != 0 && !pc->finished) {
        if (pc->res) pc->res->status = rc;
        pc->finished = true;
      }
      if (pc->finished) {
        if (pc->st) hipmps_state_destroy(pc->st);
        pc->st = nullptr;
        pc->done.store(true, std::memory_order_release);
      } else {
        keep.push_back(pc);
      }
    }
    run.swap(keep);
    {
      std::lock_guard<std::mutex> lk(E->big_mu);
      if (rc != 0 && E->big_rc == 0) {
        E->big_rc = rc;
        E->big_err = err;
      }
    }
    E->big_cv.notify_all();
  }
}

int ensure_big_lane(hipmps_engine E) {
  if (E->big) return 0;
  hipmps_engine_opts o{};
  o.device = E->device;
  o.hbm_budget_bytes = E->budget;
  o.jacobi_max_sweeps = E->max_sweeps;
  o.jacobi_tol = E->tol;
  o.jacobi_tol_mixed = E->tol_mixed;
  hipmps_engine B = nullptr;
  RC_TRY(hipmps_engine_create(&o, &B));
  hipmps_tuning t{};
  (void)hipmps_engine_get_tuning(E, &t);
  t.big_lane_min_chi = 0;
  t.wave_groups = 1;  // the worker advances the lane's circuits with advance_all, in one group
  (void)hipmps_engine_set_tuning(B, &t);
  E->big = B;
  E->big_stop = false;
  E->big_worker = std::thread(big_lane_loop, E);
  return 0;
}

// The second wave group's worker: run advance_all on the group engine for every posted job.
void group_loop(hipmps_engine E) {
  hipmps_engine B = E->grp;
  (void)hipSetDevice(B->device);
  for (;;) {
    std::vector<PendingCircuit*>* job = nullptr;
    {
      std::unique_lock<std::mutex> lk(E->grp_mu);
      E->grp_cv.wait(lk, [&] { return E->grp_stop || E->grp_job != nullptr; });
      if (E->grp_stop) return;
      job = E->grp_job;
    }
    const int rc = advance_all(B, *job);
    const std::string err = rc ? hipmps_last_error() : "";
    {
      std::lock_guard<std::mutex> lk(E->grp_mu);
      E->grp_rc = rc;
      E->grp_err = err;
      E->grp_job = nullptr;
    }
    E->grp_cv.notify_all();
  }
}

int ensure_group_engine(hipmps_engine E) {
  if (E->grp) return 0;
  hipmps_engine_opts o{};
  o.device = E->device;
  o.hbm_budget_bytes = E->budget;  // admission bounds the projected footprint of both groups together
  o.jacobi_max_sweeps = E->max_sweeps;
  o.jacobi_tol = E->tol;
  o.jacobi_tol_mixed = E->tol_mixed;
  hipmps_engine B = nullptr;
  RC_TRY(hipmps_engine_create(&o, &B));
  hipmps_tuning t{};
  (void)hipmps_engine_get_tuning(E, &t);
  t.big_lane_min_chi = 0;
  t.wave_groups = 1;
  (void)hipmps_engine_set_tuning(B, &t);
  B->timing = E->timing;
  E->grp = B;
  E->grp_stop = false;
  try {
    E->grp_worker = std::thread(group_loop, E);
  } catch (const std::exception& ex) {
    E->grp = nullptr;
    (void)hipmps_engine_destroy(B);
    return fail(HIPMPS_E_DEVICE, std::string("wave group worker: ") + ex.what());
  }
  return 0;
}

// Start a circuit on the wave group holding less projected footprint among the running ones.
constexpr int GROUP_MIN_BOND = 128;
int start_grouped(hipmps_engine E, PendingCircuit* pc, const std::vector<PendingCircuit*>& running) {
  pc->grp = 0;
  if (E->wave_groups >= 2 && ensure_group_engine(E) != 0) {
    // no second engine on this device (streams or memory): run every circuit in one group,
    // which the tuning then reports
    (void)hipGetLastError();
    E->wave_groups = 1;
  }
  // only circuits whose bond can reach GROUP_MIN_BOND go to the second group: their SVDs have
  // LQ panel chains long enough for the other group's Jacobi rounds to fill; the waves of small
  // circuits are launch-bound, and splitting them only adds a join to every step
  if (E->wave_groups >= 2 && projected_max_bond(pc->c) >= GROUP_MIN_BOND) {
    size_t load[2] = {0, 0};
    for (auto* r : running)
      if (!r->big && !r->finished) load[r->grp] += r->footprint;
    pc->grp = load[1] < load[0] ? 1 : 0;
  }
  return start_circuit(pc->grp ? E->grp : E, pc);
}

// advance_all over the wave groups: group 1 on the worker, group 0 on the calling thread, joined
// before returning (each group's circuits are touched by its own thread only).
int advance_groups(hipmps_engine E, std::vector<PendingCircuit*>& run) {
  if (!E->grp) return advance_all(E, run);
  std::vector<PendingCircuit*> g0, g1;
  for (auto* pc : run) (pc->grp ? g1 : g0).push_back(pc);
  if (g1.empty()) return advance_all(E, g0);
  {
    std::lock_guard<std::mutex> lk(E->grp_mu);
    E->grp_rc = 0;
    E->grp_job = &g1;
  }
  E->grp_cv.notify_all();
  const int rc0 = g0.empty() ? 0 : advance_all(E, g0);
  const std::string err0 = rc0 ? hipmps_last_error() : "";
  int rc1 = 0;
  std::string err1;
  {
    std::unique_lock<std::mutex> lk(E->grp_mu);
    E->grp_cv.wait(lk, [&] { return E->grp_job == nullptr; });
    rc1 = E->grp_rc;
    err1 = E->grp_err;
  }
  if (rc0) return fail(rc0, err0);
  if (rc1) return fail(rc1, err1);
  return 0;
}

}  // namespace

extern "C" {

int hipmps_circuit_footprint(const hipmps_circuit* c, size_t* out) {
  if (!c || !out) return fail(HIPMPS_E_ARG, "null argument");
  RC_TRY(validate_circuit(*c));
  *out = circuit_footprint(*c);
  return 0;
}

int hipmps_run_batch(hipmps_engine e, int n, const hipmps_circuit* circuits, hipmps_result* results) {
  if (!e || n < 0 || (n > 0 && (!circuits || !results))) return fail(HIPMPS_E_ARG, "bad arguments");
  std::lock_guard<std::mutex> lock(e->mu);
  RC_TRY(set_device(e));
  std::vector<PendingCircuit*> all, queue, running;
  auto t0 = std::chrono::steady_clock::now();
  for (int k = 0; k < n; ++k) {
    auto* pc = new PendingCircuit();
    pc->res = &results[k];
    pc->t0 = t0;
    all.push_back(pc);
    const int v = validate_circuit(circuits[k]);
    if (v != 0) {  // refused alone; the others run
      fail_circuit(pc, v);
      continue;
    }
    pc->c = circuits[k];
    pc->gates.assign(circuits[k].gates, circuits[k].gates + circuits[k].n_gates);
    if (circuits[k].n_zz > 0) pc->zz.assign(circuits[k].zz_pairs, circuits[k].zz_pairs + 2 * circuits[k].n_zz);
    pc->footprint = circuit_footprint(circuits[k]);
    queue.push_back(pc);
  }
  // admission by projected footprint, in submission order (a circuit larger than the budget
  // runs when it is alone); finished circuits free their sites at once
  int rc = 0;
  size_t next = 0, inflight = 0;
  while (rc == 0 && (next < queue.size() || !running.empty())) {
    while (next < queue.size() && (running.empty() || inflight + queue[next]->footprint <= e->budget)) {
      PendingCircuit* pc = queue[next++];
      const int r = start_grouped(e, pc, running);
      if (r == HIPMPS_E_NOMEM) {
        fail_circuit(pc, r);
        continue;
      }
      if (r != 0) {
        rc = r;
        break;
      }
      running.push_back(pc);
      inflight += pc->footprint;
    }
    if (rc != 0) break;
    rc = advance_groups(e, running);
    std::vector<PendingCircuit*> keep;
    for (auto* pc : running) {
      if (pc->finished) {
        hipmps_state_destroy(pc->st);
        pc->st = nullptr;
        inflight -= pc->footprint;
      } else {
        keep.push_back(pc);
      }
    }
    running.swap(keep);
  }
  for (auto* pc : all) {
    if (pc->st) hipmps_state_destroy(pc->st);
    delete pc;
  }
  return rc;
}

int hipmps_serve_submit(hipmps_engine e, const hipmps_circuit* c, hipmps_result* result, uint64_t* ticket) {
  if (!e || !c || !result || !ticket) return fail(HIPMPS_E_ARG, "bad arguments");
  std::lock_guard<std::mutex> lock(e->mu);
  RC_TRY(validate_circuit(*c));
  auto* pc = new PendingCircuit();
  pc->c = *c;
  pc->gates.assign(c->gates, c->gates + c->n_gates);
  if (c->n_zz > 0) pc->zz.assign(c->zz_pairs, c->zz_pairs + 2 * c->n_zz);
  pc->res = result;
  pc->footprint = circuit_footprint(*c);
  if (pc->footprint > e->budget) {
    delete pc;
    return fail(HIPMPS_E_NOMEM, "circuit's projected footprint exceeds the HBM budget");
  }
  pc->t0 = std::chrono::steady_clock::now();
  pc->ticket = e->next_ticket++;
  *ticket = pc->ticket;
  e->queue.push_back(pc);
  e->by_ticket[pc->ticket] = pc;
  return 0;
}

int hipmps_serve_step(hipmps_engine e, int* n_done) {
  if (!e) return fail(HIPMPS_E_ARG, "null engine");
  std::lock_guard<std::mutex> lock(e->mu);
  RC_TRY(set_device(e));
  // admission by projected HBM footprint; circuits that can reach big_min_chi go to the
  // background lane
  size_t inflight = 0;
  for (auto* pc : e->running) inflight += pc->footprint;
  while (!e->queue.empty()) {
    PendingCircuit* pc = e->queue.front();
    if (!e->running.empty() && inflight + pc->footprint > e->budget) break;
    e->queue.pop_front();
    pc->big = e->big_min_chi > 0 && projected_max_bond(pc->c) >= e->big_min_chi;
    if (pc->big) {
      RC_TRY(ensure_big_lane(e));
      {
        std::lock_guard<std::mutex> lk(e->big_mu);
        e->big_inbox.push_back(pc);
      }
      e->big_cv.notify_all();
    } else {
      const int r = start_grouped(e, pc, e->running);
      if (r == HIPMPS_E_NOMEM) fail_circuit(pc, r);
      else RC_TRY(r);
    }
    e->running.push_back(pc);
    inflight += pc->footprint;
  }
  std::vector<PendingCircuit*> small;
  bool any_big = false;
  for (auto* pc : e->running) {
    if (pc->big) any_big = true;
    else small.push_back(pc);
  }
  if (!small.empty()) {
    RC_TRY(advance_groups(e, small));
  } else if (any_big) {  // only background circuits in flight: wait for some progress
    std::unique_lock<std::mutex> lk(e->big_mu);
    e->big_cv.wait_for(lk, std::chrono::milliseconds(20));
  }
  int done = 0;
  std::vector<PendingCircuit*> keep;
  for (auto* pc : e->running) {
    if (!pc->big && pc->finished) {
      hipmps_state_destroy(pc->st);
      pc->st = nullptr;
      pc->done.store(true, std::memory_order_release);
      ++done;
    } else if (pc->big && pc->done.load(std::memory_order_acquire)) {
      ++done;
    } else {
      keep.push_back(pc);
    }
  }
  e->running.swap(keep);
  if (n_done) *n_done = done;
  std::lock_guard<std::mutex> lk(e->big_mu);
  if (e->big_rc != 0) {
    const int rc = e->big_rc;
    e->big_rc = 0;
    return fail(rc, "background lane: " + e->big_err);
  }
  return 0;
}

int hipmps_serve_poll(hipmps_engine e, uint64_t ticket, int* done) {
  if (!e || !done) return fail(HIPMPS_E_ARG, "bad arguments");
  std::lock_guard<std::mutex> lock(e->mu);
  auto it = e->by_ticket.find(ticket);
  if (it == e->by_ticket.end()) return fail(HIPMPS_E_ARG, "unknown ticket");
  const bool d = it->second->done.load(std::memory_order_acquire);
  *done = d ? 1 : 0;
  if (d) {
    // a background-lane circuit is published by its worker and may still be listed as running
    // until the next step: drop it from that list before it is freed
    auto r = std::find(e->running.begin(), e->running.end(), it->second);
    if (r != e->running.end()) e->running.erase(r);
    delete it->second;
    e->by_ticket.erase(it);
  }
  return 0;
}

int hipmps_serve_pending(hipmps_engine e, int* queued, int* running) {
  if (!e) return fail(HIPMPS_E_ARG, "null engine");
  std::lock_guard<std::mutex> lock(e->mu);
  if (queued) *queued = (int)e->queue.size();
  if (running) *running = (int)e->running.size();
  return 0;
}

}  // extern "C"
