"""ctypes front end of the CPU FP64 oracle (``liboracle.so`` built from mps_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of bench.py, always as the checker.  The product library
(tnqvm_amd / libhipmps) never imports this module.

The restated reference functions are listed with file:line citations at the top of
mps_oracle.cpp.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def build(force: bool = False) -> str:
    """Compile liboracle.so with the committed Makefile."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-B" if force else "-s", "-C", _HERE], check=True,
                       stdout=subprocess.DEVNULL)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        L.orc_gate_matrix.argtypes = [ctypes.c_char_p, _dp, ctypes.c_int, _dp]
        L.orc_sv_init.argtypes = [ctypes.c_int, _dp]
        L.orc_sv_apply_1q.argtypes = [ctypes.c_int, _dp, ctypes.c_int, _dp]
        L.orc_sv_apply_2q.argtypes = [ctypes.c_int, _dp, ctypes.c_int, ctypes.c_int, _dp]
        L.orc_sv_expval_z_parity.argtypes = [ctypes.c_int, _dp, _ip, ctypes.c_int]
        L.orc_sv_expval_z_parity.restype = ctypes.c_double
        L.orc_sv_norm.argtypes = [ctypes.c_int, _dp]
        L.orc_sv_norm.restype = ctypes.c_double
        L.orc_sv_sample.argtypes = [ctypes.c_int, _dp, ctypes.c_uint64, ctypes.c_int, _ip,
                                    ctypes.c_int, ctypes.c_char_p]
        L.orc_svd_split.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp, _dp, _dp, ctypes.c_int]
        L.orc_mps_create.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int,
                                     ctypes.c_int]
        L.orc_mps_create.restype = ctypes.c_void_p
        L.orc_mps_destroy.argtypes = [ctypes.c_void_p]
        L.orc_mps_apply_1q.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp]
        L.orc_mps_apply_2q.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _dp]
        L.orc_mps_bonds.argtypes = [ctypes.c_void_p, _ip]
        L.orc_mps_discarded.argtypes = [ctypes.c_void_p]
        L.orc_mps_discarded.restype = ctypes.c_double
        L.orc_mps_site_size.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_mps_site_size.restype = ctypes.c_size_t
        L.orc_mps_get_site.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp]
        L.orc_mps_set_site.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp]
        L.orc_mps_state_vector.argtypes = [ctypes.c_void_p, _dp]
        L.orc_mps_expval_z_parity.argtypes = [ctypes.c_void_p, _ip, ctypes.c_int]
        L.orc_mps_expval_z_parity.restype = ctypes.c_double
        L.orc_mps_norm.argtypes = [ctypes.c_void_p]
        L.orc_mps_norm.restype = ctypes.c_double
        L.orc_mps_amplitudes.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int8),
                                         ctypes.c_int, _dp]
        L.orc_mps_slice.argtypes = [ctypes.c_void_p, _ip, _dp]
        L.orc_mps_measure_samples.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int, _ip, ctypes.c_int,
                                              ctypes.c_char_p]
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


def gate_matrix(name: str, params=()) -> np.ndarray:
    """Row-major gate matrix as restated from Gates.hpp (4x4: first operand = MSB)."""
    p = np.ascontiguousarray(np.asarray(params, dtype=np.float64).reshape(-1))
    if p.size == 0:
        p = np.zeros(1)
    out = np.zeros(32, dtype=np.float64)
    k = lib().orc_gate_matrix(name.encode(), _d(p), len(params), _d(out))
    if k == 0:
        raise ValueError(f"unknown gate {name!r} / params {params!r}")
    dim = 2 if k == 1 else 4
    return out[: 2 * dim * dim].view(np.complex128).reshape(dim, dim).copy()


def _as_c128(m) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(m, dtype=np.complex128))


def gate_of(g) -> np.ndarray:
    if "matrix" in g:
        return _as_c128(g["matrix"])
    return gate_matrix(g["name"], g.get("params", ()))


class StateVector:
    """Dense FP64 state vector (GateMatrixAlgebra.hpp conventions, qubit k = bit k)."""

    def __init__(self, n: int):
        if n > 26:
            raise ValueError("dense oracle limited to 26 qubits")
        self.n = n
        self.psi = np.zeros(1 << n, dtype=np.complex128)
        lib().orc_sv_init(n, _d(self.psi.view(np.float64)))

    def apply(self, g):
        m = gate_of(g)
        q = g["qubits"]
        if len(q) == 1:
            lib().orc_sv_apply_1q(self.n, _d(self.psi.view(np.float64)), q[0], _d(m.view(np.float64)))
        else:
            lib().orc_sv_apply_2q(self.n, _d(self.psi.view(np.float64)), q[0], q[1],
                                  _d(m.view(np.float64)))

    def run(self, gates):
        for g in gates:
            if g["name"] in ("Measure", "I"):
                continue
            self.apply(g)
        return self

    def expval_z_parity(self, qubits) -> float:
        q = np.ascontiguousarray(np.asarray(qubits, dtype=np.int32))
        return lib().orc_sv_expval_z_parity(self.n, _d(self.psi.view(np.float64)), _i(q), len(q))

    def norm(self) -> float:
        return lib().orc_sv_norm(self.n, _d(self.psi.view(np.float64)))

    def amplitudes(self, bits: np.ndarray) -> np.ndarray:
        bits = np.asarray(bits)
        idx = (bits.astype(np.int64) << np.arange(self.n, dtype=np.int64)).sum(axis=1)
        return self.psi[idx]

    def sample(self, seed: int, shots: int, meas) -> list:
        m = np.ascontiguousarray(np.asarray(meas, dtype=np.int32))
        buf = ctypes.create_string_buffer(shots * len(m))
        got = lib().orc_sv_sample(self.n, _d(self.psi.view(np.float64)), seed, shots, _i(m), len(m), buf)
        raw = buf.raw
        return [raw[i * len(m):(i + 1) * len(m)].decode() for i in range(got)]


def svd_split(a: np.ndarray, nthreads: int = 1):
    """sqrt(S)-split Jacobi SVD: returns (s, L, R) with a = L @ R, s descending."""
    a = _as_c128(a)
    m, n = a.shape
    k0 = min(m, n)
    s = np.zeros(k0)
    L = np.zeros((m, k0), dtype=np.complex128)
    R = np.zeros((k0, n), dtype=np.complex128)
    lib().orc_svd_split(m, n, _d(a.view(np.float64)), _d(s), _d(L.view(np.float64)),
                        _d(R.view(np.float64)), nthreads)
    return s, L, R


class Mps:
    """FP64 MPS restatement of the exatn-mps visitor (see mps_oracle.cpp header)."""

    def __init__(self, n: int, cutoff: float = np.finfo(np.float64).tiny, max_bond: int = 0,
                 pn_l2: bool = False, nthreads: int = 1):
        self.n = n
        self.h = lib().orc_mps_create(n, cutoff, max_bond, int(pn_l2), nthreads)

    def __del__(self):
        h = getattr(self, "h", None)
        if h:
            lib().orc_mps_destroy(h)
            self.h = None

    def apply(self, g):
        m = gate_of(g)
        q = g["qubits"]
        if len(q) == 1:
            lib().orc_mps_apply_1q(self.h, q[0], _d(m.view(np.float64)))
            return 0
        rc = lib().orc_mps_apply_2q(self.h, q[0], q[1], _d(m.view(np.float64)))
        if rc < 0:
            raise ValueError(f"oracle MPS: non-adjacent 2q gate {q}")
        return rc

    def run(self, gates):
        for g in gates:
            if g["name"] in ("Measure", "I"):
                continue
            self.apply(g)
        return self

    def bonds(self) -> np.ndarray:
        out = np.zeros(self.n + 1, dtype=np.int32)
        lib().orc_mps_bonds(self.h, _i(out))
        return out

    def discarded(self) -> float:
        return lib().orc_mps_discarded(self.h)

    def site(self, k: int) -> np.ndarray:
        b = self.bonds()
        out = np.zeros(b[k] * 2 * b[k + 1], dtype=np.complex128)
        lib().orc_mps_get_site(self.h, k, _d(out.view(np.float64)))
        return out.reshape(b[k], 2, b[k + 1])

    def set_site(self, k: int, a: np.ndarray):
        a = np.ascontiguousarray(np.asarray(a, dtype=np.complex128))
        lib().orc_mps_set_site(self.h, k, a.shape[0], a.shape[2], _d(a.view(np.float64)))

    def state_vector(self) -> np.ndarray:
        out = np.zeros(1 << self.n, dtype=np.complex128)
        if lib().orc_mps_state_vector(self.h, _d(out.view(np.float64))) != 0:
            raise ValueError("state vector limited to 26 qubits")
        return out

    def expval_z_parity(self, qubits) -> float:
        q = np.ascontiguousarray(np.asarray(qubits, dtype=np.int32))
        return lib().orc_mps_expval_z_parity(self.h, _i(q), len(q))

    def norm(self) -> float:
        return lib().orc_mps_norm(self.h)

    def amplitudes(self, bits: np.ndarray) -> np.ndarray:
        bits = np.ascontiguousarray(np.asarray(bits, dtype=np.int8))
        k = bits.shape[0]
        out = np.zeros(k, dtype=np.complex128)
        lib().orc_mps_amplitudes(self.h, bits.ctypes.data_as(ctypes.POINTER(ctypes.c_int8)), k,
                                 _d(out.view(np.float64)))
        return out

    def slice(self, bits) -> np.ndarray:
        b = np.ascontiguousarray(np.asarray(bits, dtype=np.int32))
        nopen = int((b < 0).sum())
        out = np.zeros(1 << nopen, dtype=np.complex128)
        lib().orc_mps_slice(self.h, _i(b), _d(out.view(np.float64)))
        return out

    def measure_samples(self, seed: int, shots: int, meas) -> list:
        """getMeasureSample `shots` times (the reference's n >= 20 sampling), mt19937_64(seed)."""
        m = np.ascontiguousarray(np.asarray(meas, dtype=np.int32))
        buf = ctypes.create_string_buffer(shots * len(m))
        lib().orc_mps_measure_samples(self.h, seed, shots, _i(m), len(m), buf)
        raw = buf.raw
        return [raw[i * len(m):(i + 1) * len(m)].decode() for i in range(shots)]
