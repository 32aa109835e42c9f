/*
 * hipmps.h -- C ABI of libhipmps, the MI355X (gfx950) matrix-product-state circuit
 * simulator behind the "hip-mps" TNQVM visitor.
 *
 * Every entry point replaces one piece of the exatn-mps visitor's ExaTN-backed
 * arithmetic (citations relative to the tnqvm source tree):
 *
 *   hipmps_state_create ..... ExatnMpsVisitor::initialize, |0..0> MPS with bonds 1
 *                             (tnqvm/visitors/exatn-mps/ExaTnMpsVisitor.cpp:173-345)
 *   hipmps_apply_1q ......... ExatnMpsVisitor::applyGate, 1-qubit path (:1185-1292)
 *   hipmps_apply_2q ......... ExatnMpsVisitor::applyTwoQubitGate + truncateSvdTensors
 *                             (:1387-1730, :2366-2536)
 *   hipmps_flush ............ the exatn::sync() calls that close a gate (:1289, :1730)
 *   hipmps_expval_z_parity .. finalize's exp-val-z (:616-643)
 *   hipmps_norm ............. finalize's "norm" (:604-612)
 *   hipmps_amplitudes ....... computeWaveFuncSlice with a full bit string (:2588-2675)
 *   hipmps_wavefunction_slice computeWaveFuncSlice with open legs (-1) (:2588-2675, :794-820)
 *   hipmps_state_vector ..... finalize's full network evaluation for n < 20 (:593-597)
 *   hipmps_expval_z_all,
 *   hipmps_expval_zz ........ no counterpart (per-qubit <Z_i> / <Z_i Z_j> by environments;
 *                             the ITensor visitor's environment method is the model,
 *                             tnqvm/visitors/itensor/mps/ITensorMPSVisitor.cpp:173-250)
 *   hipmps_gate_matrix ...... GateTensorConstructor::getGateTensor + GetGateMatrix<>
 *                             (tnqvm/visitors/exatn-mps/ExatnUtils.cpp:36-133,
 *                             tnqvm/base/Gates.hpp:131-339)
 *   hipmps_batch_* .......... no counterpart: many circuits advanced together
 *   hipmps_serve_* .......... no counterpart: continuous-batching request loop
 *
 * Conventions
 *   - complex numbers are interleaved (re, im) doubles;
 *   - gate matrices are row-major; for a 2-qubit gate the FIRST qubit argument is the
 *     most significant bit of the 4x4 row/column index (tnqvm/base/Gates.hpp:237-246);
 *   - qubit k is bit k of a state-vector index (ExaTnMpsVisitor.cpp:617-633);
 *   - every call returns 0 on success or a negative HIPMPS_E_* code; the message of the
 *     last failure on the calling thread is hipmps_last_error().  Nothing aborts.
 *
 * Threading: an engine may be used from several threads for batch/serve calls; a state
 * handle must be used by one thread at a time.
 */
#ifndef HIPMPS_H_
#define HIPMPS_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HIPMPS_OK 0
#define HIPMPS_E_ARG -1          /* bad argument (null, range, shape) */
#define HIPMPS_E_NONADJACENT -2  /* 2-qubit gate on non-neighbouring qubits */
#define HIPMPS_E_NOMEM -3        /* device allocation / admission failed */
#define HIPMPS_E_CONVERGENCE -4  /* Jacobi SVD did not converge within its sweep cap */
#define HIPMPS_E_DEVICE -5       /* HIP runtime error */
#define HIPMPS_E_NUMERIC -6      /* NaN / Inf met in a tensor */
#define HIPMPS_E_UNSUPPORTED -7  /* request outside what the build supports */

/* precision */
#define HIPMPS_FP64 0   /* complex128 everywhere ("plain") */
#define HIPMPS_MIXED 1  /* complex64 site tensors and contractions, FP64 SVD and norms */

/* partial-norm definition used by the truncation rule (ExaTnMpsVisitor.cpp:2421-2429) */
#define HIPMPS_PN_SQUARED 0  /* sum |x|^2 (the reference's comment) */
#define HIPMPS_PN_L2 1       /* sqrt(sum |x|^2) */

/* where singular values go after the split */
#define HIPMPS_ABSORB_SQRT 0  /* sqrt(S) into both factors (decomposeTensorSVDLR, :1623) */
#define HIPMPS_ABSORB_RIGHT 1 /* S into the right factor */

typedef struct hipmps_engine_s* hipmps_engine;
typedef struct hipmps_state_s* hipmps_state;

typedef struct {
  int device;               /* HIP device ordinal */
  size_t hbm_budget_bytes;  /* admission budget; 0 = 90% of the device's memory */
  int jacobi_max_sweeps;    /* 0 = default (40) */
  double jacobi_tol;        /* 0 = default: relative off-diagonal 1e-13 (fp64 states) */
  double jacobi_tol_mixed;  /* 0 = default: 1e-8 (mixed states, factors kept in complex64) */
} hipmps_engine_opts;

typedef struct {
  double svd_cutoff;  /* "svd-cutoff"; default DBL_MIN (:257-263) */
  int max_bond_dim;   /* "max-bond-dim"; <= 0 means INT_MAX-1 (:265-271) */
  int precision;      /* HIPMPS_FP64 | HIPMPS_MIXED */
  int renormalize;    /* 0 = reference behaviour (no renormalisation) */
  int partial_norm;   /* HIPMPS_PN_* */
  int absorb;         /* HIPMPS_ABSORB_* */
} hipmps_trunc_opts;

typedef struct {
  int nq;        /* 1 or 2 */
  int q0, q1;    /* q1 ignored for nq == 1 */
  int pad;
  double m[32];  /* row-major 2x2 (8 doubles) or 4x4 (32 doubles), interleaved */
} hipmps_gate;

typedef struct {
  double norm;             /* <psi|psi> after the last flush */
  double discarded_weight; /* sum of squared discarded singular values */
  double fidelity_est;     /* prod_g (1 - eps_g), eps_g = discarded / kept+discarded */
  int max_bond;            /* largest bond dimension reached */
  int n_2q;                /* 2-qubit gates applied */
  int n_1q;                /* 1-qubit gates applied (fused or not) */
  int jacobi_sweeps_max;   /* worst sweep count of one SVD */
  long long jacobi_sweeps_total;
  size_t device_bytes;     /* device memory held by the state */
  int norm_warnings;       /* SVDs after which a sqrt(S) factor had Frobenius norm < 1e-3: the
                              post-SVD validation of ExaTnMpsVisitor.cpp:1634-1661, which
                              prints an error there and continues (its assert is debug-only) */
  int svds_nonconverged;   /* SVDs that reached the Jacobi sweep cap (jacobi_max_sweeps) without
                              converging; the state's status is then HIPMPS_E_CONVERGENCE */
  size_t device_bytes_peak; /* most site bytes held at once since creation / reset (a growing site
                              holds its old and its new buffer during the wave) */
} hipmps_stats_t;

/* ---- engine ---------------------------------------------------------------------- */
int hipmps_engine_create(const hipmps_engine_opts* opts, hipmps_engine* out);
int hipmps_engine_destroy(hipmps_engine e);
const char* hipmps_last_error(void);
const char* hipmps_version(void);

/* gate matrix by name (I H X Y Z S Sdg T Tdg Rx Ry Rz U CNOT CX CZ CY CH CRZ CPhase Swap
 * iSwap fSim); returns 1 or 2 (qubits) or a negative code for an unknown gate */
int hipmps_gate_matrix(const char* name, const double* params, int n_params, double* out);

/* ---- one circuit -------------------------------------------------------------------- */
int hipmps_state_create(hipmps_engine e, int n_qubits, const hipmps_trunc_opts* t,
                        hipmps_state* out);
int hipmps_state_destroy(hipmps_state s);
int hipmps_state_reset(hipmps_state s); /* back to |0..0> */
int hipmps_apply_1q(hipmps_state s, int q, const double m[8]);
int hipmps_apply_2q(hipmps_state s, int q0, int q1, const double m[32]); /* |q0-q1| == 1 */
int hipmps_apply_gates(hipmps_state s, int n, const hipmps_gate* gates);
/* Gates queue into waves (2-qubit gates on disjoint sites) that run when a gate conflicts with
 * the open wave or at a flush.  A wave whose site growth does not fit the HBM budget returns
 * HIPMPS_E_NOMEM from the call that ran it; its 2-qubit gates are dropped and every site is left
 * as after the last completed wave.  flush returns HIPMPS_E_CONVERGENCE (results still written)
 * when an SVD reached the Jacobi sweep cap, HIPMPS_E_NUMERIC when NaN/Inf appeared. */
int hipmps_flush(hipmps_state s);

int hipmps_bond_dims(hipmps_state s, int* out_n_plus_1);
int hipmps_stats(hipmps_state s, hipmps_stats_t* out);
int hipmps_norm(hipmps_state s, double* out);
int hipmps_expval_z_parity(hipmps_state s, const int* qubits, int k, double* out);
int hipmps_expval_z_all(hipmps_state s, double* out_n);
int hipmps_expval_zz(hipmps_state s, const int* pairs, int k, double* out);
int hipmps_amplitudes(hipmps_state s, const int8_t* bits, int k, double* out_re_im);
int hipmps_wavefunction_slice(hipmps_state s, const int* bits, double* out_re_im, size_t cap);
/* hipmps_amplitudes for callers that cannot allocate memory (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  No counterpart in the reference.  It allocates
 * nothing: every intermediate lives in a static device array of HIPMPS_NOALLOC_SCRATCH_BYTES that
 * each GPU holds from library load to process end, and results equal hipmps_amplitudes' bit for
 * bit.  Limits and rules:
 *   - it needs n x 72 + k x n + 2 x k x max_bond x 16 bytes (each part rounded up to 256); a
 *     larger request fails with HIPMPS_E_NOMEM (16 MiB holds 2000 strings at bond 256);
 *   - the state must have no queued gates (hipmps_flush it first; a flush may grow site tensors),
 *     else HIPMPS_E_ARG; purified states are refused as in hipmps_amplitudes;
 *   - all calls on one GPU share its array: calls that overlap on one GPU (two threads, two
 *     engines on the same device) race and must be serialised by the caller. */
#define HIPMPS_NOALLOC_SCRATCH_BYTES ((size_t)16 << 20)
int hipmps_amplitudes_noalloc(hipmps_state s, const int8_t* bits, int k, double* out_re_im);
/* <psi| (x)_k diag(d[2k], d[2k+1]) |psi>: Z strings, projectors (sequential sampling,
 * getMeasureSample :2211-2364), occupation numbers */
int hipmps_expval_diag(hipmps_state s, const double* diag_2n, double* out);
/* getMeasureSample (ExaTnMpsVisitor.cpp:2211-2364, the reference's sampling for n >= 20) for `shots`
 * shots at once, every shot's marginals advanced together on the device.  qubits: the k measured
 * qubits in Measure order, which must be increasing (HIPMPS_E_UNSUPPORTED otherwise); draws: shots * k
 * uniforms in [0, 1) in the order the reference takes them (shot by shot, one per measured qubit);
 * out: shots * k chars '0' / '1', char i <-> qubits[i].  The first measured qubit is drawn against
 * the probabilities of the unnormalised state, later ones against their conditionals; values below
 * 1e-12 count as 0; outcome 0 when the draw is <= p0. */
int hipmps_sample_sequential(hipmps_state s, const int* qubits, int k, int shots, const double* draws,
                             char* out);
/* deep copy of a state on the same engine (observable circuits in VQE mode) */
int hipmps_state_clone(hipmps_state src, hipmps_state* out);
int hipmps_state_vector(hipmps_state s, double* out_re_im, size_t cap); /* n <= 26 */
/* copy one site tensor (chi_l x 2 x kap x chi_r, complex128; kap = 1 unless noise channels
 * were applied) to / from the host; set_site writes a site with kap = 1 */
int hipmps_get_site(hipmps_state s, int k, double* out, size_t cap);
int hipmps_set_site(hipmps_state s, int k, int chi_l, int chi_r, const double* in);

/* ---- noisy simulation: locally purified MPS (exatn-pmps) ------------------------------
 * A state becomes a purified MPS the first time a channel is applied: every site then has a
 * Kraus leg kap_k (A[l][p][k][r]) and the state represents rho = sum_k |psi_k><psi_k|.
 * Gates, <Z>, <ZZ>, Z parities and diagonal operators (hipmps_expval_diag: outcome
 * probabilities) apply to rho; amplitudes, slices and the state vector are refused
 * (HIPMPS_E_UNSUPPORTED).
 *
 * hipmps_apply_kraus replaces ExaTnPmpsVisitor::applyKrausOp / applyLocalKrausOp
 * (ExaTnPmpsVisitor.cpp:887-1130): rho -> sum_o K_o rho K_o^H on qubit q, with n_ops 2 x 2
 * complex Kraus operators (row-major, 8 doubles each, 1 <= n_ops <= HIPMPS_MAX_KRAUS_OPS).
 * The Kraus leg is compressed as the reference does: vectors with partial norm below 1e-9
 * dropped, at most min(2 chi_l chi_r, 2 kap) kept.  Gates and channels apply in call order. */
/* The channel is queued like a gate: channels on distinct sites run together after the open
 * wave of gates, and an error of the compression is returned by the call that runs it. */
#define HIPMPS_MAX_KRAUS_OPS 16
int hipmps_apply_kraus(hipmps_state s, int q, int n_ops, const double* mats);
int hipmps_kraus_dims(hipmps_state s, int* out_n); /* kap_k per site */
/* Drivers of many states at once (no counterpart in the reference, whose visitor runs one
 * circuit).  hipmps_queue_gates queues gates up to the first 2-qubit gate that would have to
 * wait for the open waves (it shares a site with a queued gate or channel) and reports how many
 * it queued; hipmps_queue_kraus queues a channel unless one is already queued on the site
 * (*queued = 0).  hipmps_run_waves then runs the open waves of all the given states (one
 * engine, each listed once) together: their gates as one wave, then their channels as one, after
 * which the waiting operation can be queued.  Results equal those of the states run alone. */
int hipmps_queue_gates(hipmps_state s, int n, const hipmps_gate* gates, int* n_queued);
int hipmps_queue_kraus(hipmps_state s, int q, int n_ops, const double* mats, int* queued);
int hipmps_run_waves(hipmps_state* states, int n);
/* Reduced density matrix of k qubits (1 <= k <= 15, the reference's limit), the other qubits
 * traced out; with k = n the full density matrix.  Replaces the "density_matrix" contraction of
 * ExaTnPmpsVisitor::finalize (ExaTnPmpsVisitor.cpp:657-761).  HIPMPS_E_NOMEM when its device
 * workspace (about 3 x 4^k x 16 bytes at the end of the sweep) does not fit.
 * out[2 (x + 2^k x')] = Re <x|rho|x'>, + 1 the imaginary part, x = sum_i b_i 2^i over
 * qubits[i]: the ket index runs fastest, as in the reference's flattened tensor.  cap >= 4^k. */
int hipmps_reduced_density_matrix(hipmps_state s, const int* qubits, int k, double* out, size_t cap);

/* ---- many circuits advanced together (no counterpart in the reference) ------------- */
typedef struct {
  int n_qubits;
  int n_gates;
  const hipmps_gate* gates;
  hipmps_trunc_opts trunc;
  int n_zz;           /* number of <Z_i Z_j> pairs to report */
  const int* zz_pairs;
} hipmps_circuit;

typedef struct {
  double* expval_z;   /* caller buffer of n_qubits doubles, or NULL */
  double* expval_zz;  /* caller buffer of n_zz doubles, or NULL */
  double norm;
  double discarded_weight;
  double fidelity_est;
  int max_bond;
  int status;
  double latency_ms;  /* admission to completion */
} hipmps_result;

/* HBM a circuit may hold while it runs, the admission unit of hipmps_run_batch and serving: its
 * sites at the largest bonds they can reach (min(2^k, 2^(n-k), max-bond-dim)), twice (a growing
 * site holds its old buffer during the wave), plus its share of the wave arena.  Fails with the
 * codes of the host checks below for a circuit they refuse. */
int hipmps_circuit_footprint(const hipmps_circuit* c, size_t* out);
/* Runs every circuit to completion, admitting them in order while the sum of their footprints
 * fits hbm_budget_bytes (a circuit above the budget runs alone).  A circuit that fails alone --
 * refused by the host checks (qubit range, adjacency, NaN/Inf: HIPMPS_E_ARG / _NONADJACENT /
 * _NUMERIC), or unable to grow its sites within hbm_budget_bytes even after the others finished
 * (HIPMPS_E_NOMEM) -- reports that code in results[k].status, and the other circuits' results
 * are those they have alone.
 * results[k].status is HIPMPS_E_CONVERGENCE when an SVD of the circuit hit the sweep cap.  The
 * call itself returns non-zero only for an engine-level failure (bad arguments, device error). */
int hipmps_run_batch(hipmps_engine e, int n, const hipmps_circuit* circuits,
                     hipmps_result* results);

/* ---- serving: continuous batching (no counterpart in the reference) ------------------ */
/* Queues a circuit; refused at once (nothing queued) when it fails the host checks of
 * hipmps_run_batch or when its projected footprint exceeds the HBM budget (HIPMPS_E_NOMEM).  A
 * queued circuit that fails later fails alone, with its code in result->status. */
int hipmps_serve_submit(hipmps_engine e, const hipmps_circuit* c, hipmps_result* result,
                        uint64_t* ticket);
/* advance every in-flight circuit by one wave (the two wave groups concurrently, see
 * hipmps_tuning.wave_groups); circuits on the background lane (see
 * hipmps_tuning.big_lane_min_chi) advance on a worker thread of the engine meanwhile */
int hipmps_serve_step(hipmps_engine e, int* n_done);
int hipmps_serve_poll(hipmps_engine e, uint64_t ticket, int* done);
int hipmps_serve_pending(hipmps_engine e, int* queued, int* running);

/* ---- measurement hooks for the bench ------------------------------------------------ */
typedef struct {
  double ms_theta, ms_jacobi, ms_split, ms_other;
  double ms_jacobi_rounds;  /* HIP-event time of the back-to-back k_jacobi_round launches */
  long long jacobi_launches;
  long long jacobi_pairs_rotated;
  long long jacobi_pairs_visited;
  double jacobi_flops; /* flops executed by the Jacobi block-rotation kernel */
  long long waves;      /* batched waves run */
  long long wave_tasks; /* 2-qubit gates applied in those waves */
  long long jacobi_sweeps; /* sweeps summed over all SVDs */
  long long svds_large;    /* SVDs with min(m, n) >= 256 */
  long long sweeps_large;  /* sweeps summed over those */
  double ms_lq;            /* HIP-event time of the double-LQ factorisations (between theta and
                              the Jacobi sweeps); ms_jacobi holds the sweeps and the application
                              of the LQ reflectors back to the rows */
  long long applies_by_rank[14]; /* 2-qubit applies by floor(log2(min(m, n))) of theta' (m x n):
                                    index 9 counts 512 <= min(m, n) < 1024, i.e. theta' 512 x 512
                                    when chi_max = 256 */
} hipmps_profile_t;
/* with two wave groups (hipmps_tuning.wave_groups) every field is the sum over both groups'
 * engines, whose phase times overlap in time; the background lane's engine is not included */
int hipmps_engine_profile(hipmps_engine e, hipmps_profile_t* out, int reset);
int hipmps_engine_set_timing(hipmps_engine e, int enabled);
/* Preconditioning of the SVD: for min(m, n) >= min_rank the one-sided Jacobi runs on R2 of the
 * double Householder LQ (X = L Q, L^H = L_Z Q_Z, R2 = L_Z^H) instead of on X, which it
 * converges on in fewer sweeps; 0 turns it off.  Default 64.  No counterpart in the reference,
 * whose SVD is ExaTN's (ExaTnMpsVisitor.cpp:1623); the singular values and the truncation are
 * the same either way. */
int hipmps_engine_set_lq_precondition(hipmps_engine e, int min_rank);
/* Launch and precision tunables of the batched SVD (no counterpart in the reference).  The
 * defaults are the choice of tnqvm_amd/autotune.py on MI355X; get_tuning reports the values
 * in force (after HIPMPS_JACOBI_LANES / HIPMPS_GRAM32_SWEEPS from the environment). */
typedef struct {
  int jacobi_lanes;    /* streams that the Jacobi rounds and LQ panels of large waves are dealt to, 1..4 */
  int lane_min_pairs;  /* block pairs per sweep from which more than one lane is used */
  int gram32_sweeps;   /* mixed states: leading sweeps of SVDs with min(m, n) >= 256 whose Gram uses
                          FP32 MFMA (rotations and updates stay FP64); 0 = none */
  int lq_min_rank;     /* double-LQ preconditioning for min(m, n) >= this; 0 = off */
  int big_lane_min_chi; /* serving: circuits whose bond can reach this (min(max-bond-dim, 2^(n/2)))
                           run on a background lane of their own, concurrently with the steps; 0 = off
                           (default).  The large circuits then finish sooner and the others later */
  int rdm_block_elems;  /* density matrices: the per-site intermediate of the sweep is held a block of
                           (p, kappa) at a time above max(this, largest environment) complex
                           elements; default 2^24 */
  int jacobi_max_sweeps; /* sweep cap of every Jacobi SVD (1..1000), after which the SVD counts as
                            non-converged (HIPMPS_E_CONVERGENCE, hipmps_stats_t.svds_nonconverged);
                            default: hipmps_engine_opts.jacobi_max_sweeps */
  int wave_groups;      /* batches and serving: circuits whose bond can reach 128 are balanced over
                           this many groups (1 or 2, default 2; smaller circuits run in the first),
                           each with an engine of its own (streams, arena); the groups' waves advance
                           concurrently, the second on a worker thread, so that one group's LQ
                           panels overlap the other's Jacobi rounds.  A circuit's results do not
                           depend on its group.  hipmps_engine_profile sums both groups */
} hipmps_tuning;
int hipmps_engine_get_tuning(hipmps_engine e, hipmps_tuning* out);
int hipmps_engine_set_tuning(hipmps_engine e, const hipmps_tuning* t);

#ifdef __cplusplus
}
#endif
#endif /* HIPMPS_H_ */
