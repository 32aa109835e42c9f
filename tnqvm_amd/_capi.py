"""ctypes binding of libhipmps (include/hipmps.h).

The library is loaded from this package directory only (built in-tree by
``python -m tnqvm_amd.build``).  There is no fallback: if the shared object is
missing or cannot be loaded, every call raises.
"""
from __future__ import annotations

import ctypes
import os

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libhipmps.so")

HIPMPS_FP64 = 0
HIPMPS_MIXED = 1
HIPMPS_PN_SQUARED = 0
HIPMPS_PN_L2 = 1
HIPMPS_ABSORB_SQRT = 0
HIPMPS_ABSORB_RIGHT = 1

ERRORS = {
    -1: "HIPMPS_E_ARG", -2: "HIPMPS_E_NONADJACENT", -3: "HIPMPS_E_NOMEM", -4: "HIPMPS_E_CONVERGENCE",
    -5: "HIPMPS_E_DEVICE", -6: "HIPMPS_E_NUMERIC", -7: "HIPMPS_E_UNSUPPORTED",
}
HIPMPS_E_ARG, HIPMPS_E_NONADJACENT, HIPMPS_E_NOMEM, HIPMPS_E_CONVERGENCE = -1, -2, -3, -4
HIPMPS_E_DEVICE, HIPMPS_E_NUMERIC, HIPMPS_E_UNSUPPORTED = -5, -6, -7

# every symbol declared in include/hipmps.h
EXPORTS = [
    "hipmps_engine_create", "hipmps_engine_destroy", "hipmps_last_error", "hipmps_version",
    "hipmps_gate_matrix", "hipmps_state_create", "hipmps_state_destroy", "hipmps_state_reset",
    "hipmps_apply_1q", "hipmps_apply_2q", "hipmps_apply_gates", "hipmps_flush", "hipmps_bond_dims",
    "hipmps_stats", "hipmps_norm", "hipmps_expval_z_parity", "hipmps_expval_z_all", "hipmps_expval_zz",
    "hipmps_amplitudes", "hipmps_wavefunction_slice", "hipmps_state_vector", "hipmps_get_site",
    "hipmps_set_site", "hipmps_run_batch", "hipmps_serve_submit", "hipmps_serve_step",
    "hipmps_serve_poll", "hipmps_serve_pending", "hipmps_engine_profile", "hipmps_engine_set_timing", "hipmps_engine_set_lq_precondition",
    "hipmps_expval_diag", "hipmps_state_clone", "hipmps_apply_kraus", "hipmps_kraus_dims",
    "hipmps_reduced_density_matrix", "hipmps_engine_get_tuning", "hipmps_engine_set_tuning",
    "hipmps_queue_gates", "hipmps_queue_kraus", "hipmps_run_waves", "hipmps_circuit_footprint",
    "hipmps_amplitudes_noalloc", "hipmps_sample_sequential",
]

HIPMPS_NOALLOC_SCRATCH_BYTES = 16 << 20


class HipMpsError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class EngineOpts(ctypes.Structure):
    _fields_ = [("device", ctypes.c_int), ("hbm_budget_bytes", ctypes.c_size_t),
                ("jacobi_max_sweeps", ctypes.c_int), ("jacobi_tol", ctypes.c_double),
                ("jacobi_tol_mixed", ctypes.c_double)]


class TruncOpts(ctypes.Structure):
    _fields_ = [("svd_cutoff", ctypes.c_double), ("max_bond_dim", ctypes.c_int), ("precision", ctypes.c_int),
                ("renormalize", ctypes.c_int), ("partial_norm", ctypes.c_int), ("absorb", ctypes.c_int)]


class Gate(ctypes.Structure):
    _fields_ = [("nq", ctypes.c_int), ("q0", ctypes.c_int), ("q1", ctypes.c_int), ("pad", ctypes.c_int),
                ("m", ctypes.c_double * 32)]


class Tuning(ctypes.Structure):
    _fields_ = [("jacobi_lanes", ctypes.c_int), ("lane_min_pairs", ctypes.c_int), ("gram32_sweeps", ctypes.c_int),
                ("lq_min_rank", ctypes.c_int), ("big_lane_min_chi", ctypes.c_int),
                ("rdm_block_elems", ctypes.c_int), ("jacobi_max_sweeps", ctypes.c_int), ("wave_groups", ctypes.c_int)]


class Stats(ctypes.Structure):
    _fields_ = [("norm", ctypes.c_double), ("discarded_weight", ctypes.c_double),
                ("fidelity_est", ctypes.c_double), ("max_bond", ctypes.c_int), ("n_2q", ctypes.c_int),
                ("n_1q", ctypes.c_int), ("jacobi_sweeps_max", ctypes.c_int),
                ("jacobi_sweeps_total", ctypes.c_longlong), ("device_bytes", ctypes.c_size_t),
                ("norm_warnings", ctypes.c_int), ("svds_nonconverged", ctypes.c_int),
                ("device_bytes_peak", ctypes.c_size_t)]


class Circuit(ctypes.Structure):
    _fields_ = [("n_qubits", ctypes.c_int), ("n_gates", ctypes.c_int), ("gates", ctypes.POINTER(Gate)),
                ("trunc", TruncOpts), ("n_zz", ctypes.c_int), ("zz_pairs", ctypes.POINTER(ctypes.c_int))]


class Result(ctypes.Structure):
    _fields_ = [("expval_z", ctypes.POINTER(ctypes.c_double)), ("expval_zz", ctypes.POINTER(ctypes.c_double)),
                ("norm", ctypes.c_double), ("discarded_weight", ctypes.c_double),
                ("fidelity_est", ctypes.c_double), ("max_bond", ctypes.c_int), ("status", ctypes.c_int),
                ("latency_ms", ctypes.c_double)]


class Profile(ctypes.Structure):
    _fields_ = [("ms_theta", ctypes.c_double), ("ms_jacobi", ctypes.c_double), ("ms_split", ctypes.c_double),
                ("ms_other", ctypes.c_double), ("ms_jacobi_rounds", ctypes.c_double),
                ("jacobi_launches", ctypes.c_longlong), ("jacobi_pairs_rotated", ctypes.c_longlong),
                ("jacobi_pairs_visited", ctypes.c_longlong), ("jacobi_flops", ctypes.c_double),
                ("waves", ctypes.c_longlong), ("wave_tasks", ctypes.c_longlong),
                ("jacobi_sweeps", ctypes.c_longlong), ("svds_large", ctypes.c_longlong),
                ("sweeps_large", ctypes.c_longlong), ("ms_lq", ctypes.c_double),
                ("applies_by_rank", ctypes.c_longlong * 14)]


_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise HipMpsError(-7, f"{LIB_PATH} is not built (python -m tnqvm_amd.build)")
    L = ctypes.CDLL(LIB_PATH)
    vp, ip, dp = ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_double)
    pint = ctypes.POINTER(ctypes.c_int)
    sig = {
        "hipmps_engine_create": ([ctypes.POINTER(EngineOpts), ctypes.POINTER(vp)], ip),
        "hipmps_engine_destroy": ([vp], ip),
        "hipmps_last_error": ([], ctypes.c_char_p),
        "hipmps_version": ([], ctypes.c_char_p),
        "hipmps_gate_matrix": ([ctypes.c_char_p, dp, ip, dp], ip),
        "hipmps_state_create": ([vp, ip, ctypes.POINTER(TruncOpts), ctypes.POINTER(vp)], ip),
        "hipmps_state_destroy": ([vp], ip),
        "hipmps_state_reset": ([vp], ip),
        "hipmps_apply_1q": ([vp, ip, dp], ip),
        "hipmps_apply_2q": ([vp, ip, ip, dp], ip),
        "hipmps_apply_gates": ([vp, ip, ctypes.POINTER(Gate)], ip),
        "hipmps_flush": ([vp], ip),
        "hipmps_bond_dims": ([vp, pint], ip),
        "hipmps_stats": ([vp, ctypes.POINTER(Stats)], ip),
        "hipmps_norm": ([vp, dp], ip),
        "hipmps_expval_z_parity": ([vp, pint, ip, dp], ip),
        "hipmps_expval_z_all": ([vp, dp], ip),
        "hipmps_expval_zz": ([vp, pint, ip, dp], ip),
        "hipmps_amplitudes": ([vp, ctypes.POINTER(ctypes.c_int8), ip, dp], ip),
        "hipmps_amplitudes_noalloc": ([vp, ctypes.POINTER(ctypes.c_int8), ip, dp], ip),
        "hipmps_sample_sequential": ([vp, pint, ip, ip, dp, ctypes.c_char_p], ip),
        "hipmps_wavefunction_slice": ([vp, pint, dp, ctypes.c_size_t], ip),
        "hipmps_state_vector": ([vp, dp, ctypes.c_size_t], ip),
        "hipmps_get_site": ([vp, ip, dp, ctypes.c_size_t], ip),
        "hipmps_set_site": ([vp, ip, ip, ip, dp], ip),
        "hipmps_run_batch": ([vp, ip, ctypes.POINTER(Circuit), ctypes.POINTER(Result)], ip),
        "hipmps_serve_submit": ([vp, ctypes.POINTER(Circuit), ctypes.POINTER(Result),
                                 ctypes.POINTER(ctypes.c_uint64)], ip),
        "hipmps_serve_step": ([vp, pint], ip),
        "hipmps_serve_poll": ([vp, ctypes.c_uint64, pint], ip),
        "hipmps_serve_pending": ([vp, pint, pint], ip),
        "hipmps_engine_profile": ([vp, ctypes.POINTER(Profile), ip], ip),
        "hipmps_engine_set_timing": ([vp, ip], ip),
        "hipmps_engine_set_lq_precondition": ([vp, ip], ip),
        "hipmps_expval_diag": ([vp, dp, dp], ip),
        "hipmps_state_clone": ([vp, ctypes.POINTER(vp)], ip),
        "hipmps_apply_kraus": ([vp, ip, ip, dp], ip),
        "hipmps_kraus_dims": ([vp, pint], ip),
        "hipmps_reduced_density_matrix": ([vp, pint, ip, dp, ctypes.c_size_t], ip),
        "hipmps_engine_get_tuning": ([vp, ctypes.POINTER(Tuning)], ip),
        "hipmps_engine_set_tuning": ([vp, ctypes.POINTER(Tuning)], ip),
        "hipmps_queue_gates": ([vp, ip, ctypes.POINTER(Gate), pint], ip),
        "hipmps_queue_kraus": ([vp, ip, ip, dp, pint], ip),
        "hipmps_run_waves": ([ctypes.POINTER(vp), ip], ip),
    }
    for name, (args, res) in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = res
    _lib = L
    return L


def check(rc: int) -> int:
    if rc < 0:
        msg = lib().hipmps_last_error()
        raise HipMpsError(rc, msg.decode() if msg else "")
    return rc
