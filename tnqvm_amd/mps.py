"""Python front end of libhipmps: engine, MPS state, batches and serving.

This module is a thin layer over the C ABI (include/hipmps.h); every numeric
operation runs in the HIP library.  Gate dicts follow tnqvm_amd.circuits.
"""
from __future__ import annotations

import ctypes
import sys
import weakref

import numpy as np

from . import _capi as C
from ._capi import check, lib

DBL_MIN = sys.float_info.min

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def gate_matrix(name: str, params=()) -> np.ndarray:
    """Row-major gate matrix from the library's gate table (2-qubit: first operand = MSB)."""
    p = np.ascontiguousarray(np.asarray(params, dtype=np.float64).reshape(-1))
    buf = np.zeros(32)
    k = check(lib().hipmps_gate_matrix(name.encode(), p.ctypes.data_as(_dp) if p.size else None, int(p.size),
                                       buf.ctypes.data_as(_dp)))
    d = 2 if k == 1 else 4
    return buf[: 2 * d * d].view(np.complex128).reshape(d, d).copy()


def matrix_of(g) -> np.ndarray:
    if "matrix" in g:
        return np.ascontiguousarray(np.asarray(g["matrix"], dtype=np.complex128))
    return gate_matrix(g["name"], g.get("params", ()))


_GATE_DTYPE = np.dtype([("nq", np.int32), ("q0", np.int32), ("q1", np.int32), ("pad", np.int32),
                        ("m", np.float64, 32)])  # the layout of hipmps_gate


def to_c_gates(gates):
    """Gate dicts -> ctypes array of hipmps_gate (Measure / I are skipped)."""
    gl = [g for g in gates if g["name"] not in ("Measure", "I")]
    n = len(gl)
    buf = np.zeros(max(1, n), dtype=_GATE_DTYPE)
    if n:
        mats = [matrix_of(g).reshape(-1) for g in gl]
        nq = np.fromiter((len(g["qubits"]) for g in gl), np.int32, n)
        buf["nq"][:n] = nq
        buf["q0"][:n] = np.fromiter((g["qubits"][0] for g in gl), np.int32, n)
        buf["q1"][:n] = np.fromiter((g["qubits"][1] if len(g["qubits"]) > 1 else -1 for g in gl), np.int32, n)
        for width, nqv in ((4, 1), (16, 2)):
            idx = np.flatnonzero(nq == nqv)
            if len(idx):
                buf["m"][idx, : 2 * width] = np.stack([mats[k] for k in idx]).view(np.float64)
    assert _GATE_DTYPE.itemsize == ctypes.sizeof(C.Gate)
    arr = (C.Gate * len(buf)).from_buffer(buf)
    return arr, n


def trunc_opts(cutoff=DBL_MIN, max_bond=0, precision="fp64", renormalize=False, partial_norm="squared",
               absorb="sqrt") -> C.TruncOpts:
    t = C.TruncOpts()
    t.svd_cutoff = float(cutoff)
    t.max_bond_dim = int(max_bond)
    t.precision = C.HIPMPS_MIXED if precision == "mixed" else C.HIPMPS_FP64
    t.renormalize = int(bool(renormalize))
    t.partial_norm = C.HIPMPS_PN_L2 if partial_norm == "l2" else C.HIPMPS_PN_SQUARED
    t.absorb = C.HIPMPS_ABSORB_RIGHT if absorb == "right" else C.HIPMPS_ABSORB_SQRT
    return t


class Engine:
    def __init__(self, device: int = 0, hbm_budget_bytes: int = 0, jacobi_max_sweeps: int = 0,
                 jacobi_tol: float = 0.0, jacobi_tol_mixed: float = 0.0):
        o = C.EngineOpts(device, hbm_budget_bytes, jacobi_max_sweeps, jacobi_tol, jacobi_tol_mixed)
        h = ctypes.c_void_p()
        check(lib().hipmps_engine_create(ctypes.byref(o), ctypes.byref(h)))
        self.h = h
        self._states = weakref.WeakSet()

    def close(self):
        if getattr(self, "h", None):
            # states hold device memory owned through the engine: release them first
            for s in list(self._states):
                s.close()
            lib().hipmps_engine_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_timing(self, on: bool):
        check(lib().hipmps_engine_set_timing(self.h, int(on)))

    def set_lq_precondition(self, min_rank: int):
        """Jacobi on the double-LQ factor for min(m, n) >= min_rank (0: never)."""
        check(lib().hipmps_engine_set_lq_precondition(self.h, int(min_rank)))

    def tuning(self) -> dict:
        t = C.Tuning()
        check(lib().hipmps_engine_get_tuning(self.h, ctypes.byref(t)))
        return {k: getattr(t, k) for k, _ in C.Tuning._fields_}

    def set_tuning(self, **kw):
        t = C.Tuning()
        check(lib().hipmps_engine_get_tuning(self.h, ctypes.byref(t)))
        for k, v in kw.items():
            setattr(t, k, int(v))
        check(lib().hipmps_engine_set_tuning(self.h, ctypes.byref(t)))
        return self

    def profile(self, reset: bool = False) -> dict:
        p = C.Profile()
        check(lib().hipmps_engine_profile(self.h, ctypes.byref(p), int(reset)))
        out = {k: getattr(p, k) for k, _ in C.Profile._fields_}
        out["applies_by_rank"] = list(out["applies_by_rank"])
        return out

    # -- batches -------------------------------------------------------------------
    def _circuit(self, n, gates, trunc, zz):
        garr, ng = to_c_gates(gates)
        c = C.Circuit()
        c.n_qubits = n
        c.n_gates = ng
        c.gates = ctypes.cast(garr, ctypes.POINTER(C.Gate))
        c.trunc = trunc
        zarr = None
        if zz:
            zarr = np.ascontiguousarray(np.asarray(zz, dtype=np.int32).reshape(-1))
            c.n_zz = len(zz)
            c.zz_pairs = zarr.ctypes.data_as(_ip)
        return c, (garr, zarr)

    def footprint(self, n, gates, trunc=None) -> int:
        """Admission footprint of a circuit in bytes (hipmps_circuit_footprint)."""
        c, hold = self._circuit(n, gates, trunc or trunc_opts(), None)
        out = ctypes.c_size_t()
        check(lib().hipmps_circuit_footprint(ctypes.byref(c), ctypes.byref(out)))
        del hold
        return out.value

    def run_batch(self, jobs):
        """jobs: list of dicts {n, gates, trunc (TruncOpts), zz (list of pairs)}.
        Returns list of dicts with expval_z, expval_zz, norm, discarded_weight, ..."""
        keep = []
        circs = (C.Circuit * len(jobs))()
        res = (C.Result * len(jobs))()
        outs = []
        for k, j in enumerate(jobs):
            c, hold = self._circuit(j["n"], j["gates"], j.get("trunc", trunc_opts()), j.get("zz"))
            keep.append(hold)
            circs[k] = c
            ez = np.zeros(j["n"])
            ezz = np.zeros(max(1, len(j.get("zz") or [])))
            outs.append((ez, ezz))
            res[k].expval_z = ez.ctypes.data_as(_dp)
            res[k].expval_zz = ezz.ctypes.data_as(_dp)
        check(lib().hipmps_run_batch(self.h, len(jobs), circs, res))
        out = []
        for k, j in enumerate(jobs):
            r = res[k]
            out.append({"expval_z": outs[k][0], "expval_zz": outs[k][1][: len(j.get("zz") or [])],
                        "norm": r.norm, "discarded_weight": r.discarded_weight, "fidelity_est": r.fidelity_est,
                        "max_bond": r.max_bond, "status": r.status, "latency_ms": r.latency_ms})
        return out


class Server:
    """Continuous-batching request loop on top of hipmps_serve_*."""

    def __init__(self, engine: Engine):
        self.e = engine
        self.live = {}

    def submit(self, n, gates, trunc=None, zz=None) -> int:
        c, hold = self.e._circuit(n, gates, trunc or trunc_opts(), zz)
        res = C.Result()
        ez = np.zeros(n)
        ezz = np.zeros(max(1, len(zz or [])))
        res.expval_z = ez.ctypes.data_as(_dp)
        res.expval_zz = ezz.ctypes.data_as(_dp)
        t = ctypes.c_uint64()
        check(lib().hipmps_serve_submit(self.e.h, ctypes.byref(c), ctypes.byref(res), ctypes.byref(t)))
        self.live[t.value] = (c, hold, res, ez, ezz, len(zz or []))
        return t.value

    def step(self) -> int:
        d = ctypes.c_int()
        check(lib().hipmps_serve_step(self.e.h, ctypes.byref(d)))
        return d.value

    def pending(self):
        q, r = ctypes.c_int(), ctypes.c_int()
        check(lib().hipmps_serve_pending(self.e.h, ctypes.byref(q), ctypes.byref(r)))
        return q.value, r.value

    def poll(self, ticket: int):
        d = ctypes.c_int()
        check(lib().hipmps_serve_poll(self.e.h, ticket, ctypes.byref(d)))
        if not d.value:
            return None
        c, hold, res, ez, ezz, nzz = self.live.pop(ticket)
        return {"expval_z": ez, "expval_zz": ezz[:nzz], "norm": res.norm, "discarded_weight": res.discarded_weight,
                "fidelity_est": res.fidelity_est, "max_bond": res.max_bond, "status": res.status,
                "latency_ms": res.latency_ms}


class MpsState:
    """One MPS circuit state resident in HBM (hipmps_state)."""

    def __init__(self, engine: Engine, n: int, **trunc):
        self.e = engine
        self.n = n
        self.trunc = trunc_opts(**trunc)
        h = ctypes.c_void_p()
        check(lib().hipmps_state_create(engine.h, n, ctypes.byref(self.trunc), ctypes.byref(h)))
        self.h = h
        engine._states.add(self)

    def close(self):
        if getattr(self, "h", None):
            if getattr(self.e, "h", None):
                lib().hipmps_state_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def reset(self):
        check(lib().hipmps_state_reset(self.h))

    def apply(self, g):
        m = matrix_of(g)
        q = g["qubits"]
        flat = np.ascontiguousarray(m.reshape(-1)).view(np.float64)
        if len(q) == 1:
            check(lib().hipmps_apply_1q(self.h, int(q[0]), flat.ctypes.data_as(_dp)))
        else:
            check(lib().hipmps_apply_2q(self.h, int(q[0]), int(q[1]), flat.ctypes.data_as(_dp)))

    def run(self, gates):
        arr, ng = to_c_gates(gates)
        check(lib().hipmps_apply_gates(self.h, ng, arr))
        return self

    def flush(self):
        return check(lib().hipmps_flush(self.h))

    def bonds(self) -> np.ndarray:
        out = np.zeros(self.n + 1, dtype=np.int32)
        check(lib().hipmps_bond_dims(self.h, out.ctypes.data_as(_ip)))
        return out

    def stats(self) -> dict:
        s = C.Stats()
        lib().hipmps_stats(self.h, ctypes.byref(s))
        return {k: getattr(s, k) for k, _ in C.Stats._fields_}

    def norm(self) -> float:
        v = ctypes.c_double()
        check(lib().hipmps_norm(self.h, ctypes.byref(v)))
        return v.value

    def expval_z_parity(self, qubits) -> float:
        q = np.ascontiguousarray(np.asarray(qubits, dtype=np.int32))
        v = ctypes.c_double()
        check(lib().hipmps_expval_z_parity(self.h, q.ctypes.data_as(_ip), len(q), ctypes.byref(v)))
        return v.value

    def expval_z_all(self) -> np.ndarray:
        out = np.zeros(self.n)
        check(lib().hipmps_expval_z_all(self.h, out.ctypes.data_as(_dp)))
        return out

    def expval_zz(self, pairs) -> np.ndarray:
        p = np.ascontiguousarray(np.asarray(pairs, dtype=np.int32).reshape(-1))
        out = np.zeros(len(p) // 2)
        check(lib().hipmps_expval_zz(self.h, p.ctypes.data_as(_ip), len(p) // 2, out.ctypes.data_as(_dp)))
        return out

    def amplitudes(self, bits) -> np.ndarray:
        b = np.ascontiguousarray(np.asarray(bits, dtype=np.int8))
        out = np.zeros(b.shape[0], dtype=np.complex128)
        check(lib().hipmps_amplitudes(self.h, b.ctypes.data_as(ctypes.POINTER(ctypes.c_int8)), b.shape[0],
                                      out.view(np.float64).ctypes.data_as(_dp)))
        return out

    def amplitudes_noalloc(self, bits) -> np.ndarray:
        """amplitudes() through hipmps_amplitudes_noalloc (static device scratch, no allocation);
        the state must be flushed first."""
        b = np.ascontiguousarray(np.asarray(bits, dtype=np.int8))
        out = np.zeros(b.shape[0], dtype=np.complex128)
        check(lib().hipmps_amplitudes_noalloc(self.h, b.ctypes.data_as(ctypes.POINTER(ctypes.c_int8)), b.shape[0],
                                              out.view(np.float64).ctypes.data_as(_dp)))
        return out

    def sample_sequential(self, qubits, draws) -> list:
        """getMeasureSample for len(draws) // len(qubits) shots with the given uniforms (shot-major,
        one per measured qubit); qubits in increasing order.  Returns the strings."""
        q = np.ascontiguousarray(np.asarray(qubits, dtype=np.int32))
        d = np.ascontiguousarray(np.asarray(draws, dtype=np.float64).reshape(-1))
        k = len(q)
        shots = d.size // k
        buf = ctypes.create_string_buffer(max(1, shots * k))
        check(lib().hipmps_sample_sequential(self.h, q.ctypes.data_as(_ip), k, shots, d.ctypes.data_as(_dp), buf))
        raw = buf.raw
        return [raw[i * k:(i + 1) * k].decode() for i in range(shots)]

    def slice(self, bits) -> np.ndarray:
        b = np.ascontiguousarray(np.asarray(bits, dtype=np.int32))
        nopen = int((b < 0).sum())
        out = np.zeros(1 << nopen, dtype=np.complex128)
        check(lib().hipmps_wavefunction_slice(self.h, b.ctypes.data_as(_ip), out.view(np.float64).ctypes.data_as(_dp),
                                              out.size))
        return out

    def state_vector(self) -> np.ndarray:
        out = np.zeros(1 << self.n, dtype=np.complex128)
        check(lib().hipmps_state_vector(self.h, out.view(np.float64).ctypes.data_as(_dp), out.size))
        return out

    def get_site(self, k: int) -> np.ndarray:
        """Site k as (chi_l, 2, chi_r), or (chi_l, 2, kap, chi_r) once it has a Kraus leg."""
        b = self.bonds()
        kap = int(self.kraus_dims()[k])
        out = np.zeros(b[k] * 2 * kap * b[k + 1], dtype=np.complex128)
        check(lib().hipmps_get_site(self.h, k, out.view(np.float64).ctypes.data_as(_dp), out.size))
        return out.reshape(b[k], 2, b[k + 1]) if kap == 1 else out.reshape(b[k], 2, kap, b[k + 1])

    def apply_kraus(self, q: int, ops):
        """Noise channel rho -> sum_o K_o rho K_o^H on qubit q (2 x 2 Kraus operators)."""
        k = np.ascontiguousarray(np.asarray(ops, dtype=np.complex128).reshape(-1, 2, 2))
        check(lib().hipmps_apply_kraus(self.h, int(q), len(k), k.view(np.float64).ctypes.data_as(_dp)))
        return self

    def queue_gates(self, arr, start: int, count: int) -> int:
        """Queue gates arr[start:start + count] (a to_c_gates array) up to the first 2-qubit gate
        that must wait for the open waves; returns how many were queued."""
        nq = ctypes.c_int(0)
        ptr = ctypes.cast(ctypes.addressof(arr) + start * ctypes.sizeof(C.Gate), ctypes.POINTER(C.Gate))
        check(lib().hipmps_queue_gates(self.h, count, ptr, ctypes.byref(nq)))
        return nq.value

    def queue_kraus(self, q: int, mats: np.ndarray) -> bool:
        """Queue a channel (mats: float64 view of n_ops 2 x 2 complex operators) unless one is
        already queued on q."""
        ok = ctypes.c_int(0)
        check(lib().hipmps_queue_kraus(self.h, int(q), mats.size // 8, mats.ctypes.data_as(_dp), ctypes.byref(ok)))
        return bool(ok.value)

    def clone(self) -> "MpsState":
        """Deep copy on the same engine (hipmps_state_clone)."""
        h = ctypes.c_void_p()
        check(lib().hipmps_state_clone(self.h, ctypes.byref(h)))
        c = MpsState.__new__(MpsState)
        c.e, c.n, c.trunc, c.h = self.e, self.n, self.trunc, h
        self.e._states.add(c)
        return c

    def expval_diag(self, diag) -> float:
        """<(x)_k diag(d[2k], d[2k+1])>, a (2n,) array of per-site diagonals."""
        d = np.ascontiguousarray(np.asarray(diag, dtype=np.float64).reshape(-1))
        v = ctypes.c_double()
        check(lib().hipmps_expval_diag(self.h, d.ctypes.data_as(_dp), ctypes.byref(v)))
        return v.value

    def outcome_probabilities(self, qubits) -> np.ndarray:
        """P(outcome) over `qubits` (index sum_i b_i 2^i, b_i the bit of qubits[i]): one
        projector string per outcome through the environment contraction, for pure and
        purified states alike."""
        k = len(qubits)
        if k > 16:
            raise ValueError("at most 16 measured qubits")
        out = np.zeros(1 << k)
        for x in range(1 << k):
            d = np.ones(2 * self.n)
            for i, q in enumerate(qubits):
                b = (x >> i) & 1
                d[2 * q + (1 - b)] = 0.0
            out[x] = self.expval_diag(d)
        return out

    def reduced_density_matrix(self, qubits) -> np.ndarray:
        """rho[x, x'] = <x|rho_red|x'> over `qubits` (x = sum_i b_i 2^i), others traced out."""
        q = np.ascontiguousarray(np.asarray(qubits, dtype=np.int32))
        dim = 1 << len(q)
        flat = np.zeros(dim * dim, dtype=np.complex128)
        check(lib().hipmps_reduced_density_matrix(self.h, q.ctypes.data_as(_ip), len(q),
                                                  flat.view(np.float64).ctypes.data_as(_dp), flat.size))
        return flat.reshape(dim, dim).T  # flat[x + dim x'] -> rho[x, x']

    def kraus_dims(self) -> np.ndarray:
        out = np.zeros(self.n, dtype=np.int32)
        check(lib().hipmps_kraus_dims(self.h, out.ctypes.data_as(_ip)))
        return out

    def set_site(self, k: int, a: np.ndarray):
        a = np.ascontiguousarray(np.asarray(a, dtype=np.complex128))
        check(lib().hipmps_set_site(self.h, k, a.shape[0], a.shape[2], a.view(np.float64).ctypes.data_as(_dp)))


def run_waves(states):
    """Run the open waves of several states (one engine) together: gates, then channels."""
    arr = (ctypes.c_void_p * len(states))(*[st.h for st in states])
    check(lib().hipmps_run_waves(arr, len(states)))


def run_noisy_batch(engine: Engine, circuits, noise, precision: str = "fp64", max_bond: int = 0):
    """Many noisy circuits advanced together: each is fed up to the next operation that must wait
    for its open waves, then the open waves of all of them run as one gate wave and one channel
    wave.  circuits: [(n, gates)]; noise as in run_noisy.  Returns per circuit the same fields as
    run_noisy's expval_z, bonds, kraus_dims and trace (equal to running each alone)."""
    prepared = []
    cache = {}
    for n, gates in circuits:
        gates = [g for g in gates if g["name"] not in ("Measure", "I")]
        arr, _ = to_c_gates(gates)
        ops = []
        for k, g in enumerate(gates):
            ops.append((k, None, None))
            if noise is not None:
                for q, kops in noise.channels(g):
                    key = id(kops)
                    if key not in cache:
                        m = np.ascontiguousarray(np.asarray(kops, dtype=np.complex128).reshape(-1, 2, 2))
                        cache[key] = (kops, m.view(np.float64).reshape(-1))
                    ops.append((None, q, cache[key][1]))
        prepared.append((MpsState(engine, n, cutoff=1e-9, max_bond=max_bond, precision=precision), arr, ops))
    try:
        cursor = [0] * len(prepared)
        while True:
            live = []
            for i, (st, arr, ops) in enumerate(prepared):
                c = cursor[i]
                while c < len(ops):
                    k, q, m = ops[c]
                    if k is not None:
                        if st.queue_gates(arr, k, 1) == 0:
                            break
                    elif not st.queue_kraus(q, m):
                        break
                    c += 1
                cursor[i] = c
                live.append(st)
            run_waves(live)
            if all(cursor[i] == len(p[2]) for i, p in enumerate(prepared)):
                break
        return [{"expval_z": st.expval_z_all(), "bonds": st.bonds(), "kraus_dims": st.kraus_dims(),
                 "trace": st.norm()} for st, _, _ in prepared]
    finally:
        for st, _, _ in prepared:
            st.close()


def run_noisy(engine: Engine, n: int, gates, noise=None, measured=(), shots: int = 1024, seed=None,
              precision: str = "fp64", max_bond: int = 0):
    """A noisy circuit the way the reference's purified-MPS visitor runs one
    (ExaTnPmpsVisitor.cpp): each gate followed by the single-qubit channels the noise model
    gives for it (:775-829), bonds cut at 1e-9 with no cap (truncateSvdTensors' default,
    :1255), then the outcome distribution of the measured qubits and `shots` strings sampled
    from it with the model's readout errors (finalize, :657-782; the default of 1024 shots
    for a noisy run, :490).  String character i is the outcome of measured[i].

    noise: None or an object with channels(gate) -> [(qubit, Kraus ops)] and
    readout_error(qubit) -> (P(0 | 1), P(1 | 0)) (tnqvm_amd.noise.NoiseModel).
    max_bond: 0 (the reference: no cap) or a cap on the bonds, as the visitor's max-bond-dim."""
    s = MpsState(engine, n, cutoff=1e-9, max_bond=max_bond, precision=precision)
    try:
        for g in gates:
            if g["name"] in ("Measure", "I"):
                continue
            s.run([g])
            if noise is not None:
                for q, ops in noise.channels(g):
                    s.apply_kraus(q, ops)
        measured = list(measured)
        probs = s.outcome_probabilities(measured) if measured else np.ones(1)
        rng = np.random.default_rng(seed)
        p = np.clip(probs, 0.0, None)
        picks = rng.choice(len(p), size=shots, p=p / p.sum()) if measured else np.zeros(0, dtype=int)
        counts = {}
        for x in picks:
            bits = []
            for i, q in enumerate(measured):
                b = (int(x) >> i) & 1
                if noise is not None:
                    p01, p10 = noise.readout_error(q)  # P(read 0 | 1), P(read 1 | 0)
                    if rng.random() < (p01 if b else p10):
                        b ^= 1
                bits.append("1" if b else "0")
            key = "".join(bits)
            counts[key] = counts.get(key, 0) + 1
        out = {"probabilities": probs, "counts": counts, "expval_z": s.expval_z_all(),
               "bonds": s.bonds(), "kraus_dims": s.kraus_dims(), "trace": s.norm()}
        if measured:
            # what the reference stores as "density_matrix": all qubits when n <= 15 (:664-683),
            # else the distinct measured qubits in qubit order (:701-761); rho[x, x']
            dmq = list(range(n)) if n <= 15 else sorted(set(measured))
            out["density_matrix"] = s.reduced_density_matrix(dmq)
            out["density_matrix_qubits"] = dmq
        return out
    finally:
        s.close()
