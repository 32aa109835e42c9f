"""Noise channels and a noise model for the purified-MPS (noisy) path.

The reference's noisy visitor (tnqvm/visitors/exatn-mpo/ExaTnPmpsVisitor.cpp) takes its
channels from an XACC NoiseModel (`getNoiseChannels(gate)`, `readoutError(qubit)`,
ExaTnPmpsVisitor.cpp:470-487, :784-807), usually XACC's IBM model built from a backend
JSON.  XACC is not part of the reference tree, so its model is not restated here line by
line; `NoiseModel.from_backend_json` builds the textbook device model from the same JSON
fields (gate_error, gate_length, T1, T2, prob_meas0_prep1, prob_meas1_prep0): thermal
relaxation for the gate's duration, then depolarizing noise for the rest of the gate error
(parity against XACC's model unpinned).  Only single-qubit channels are produced, as the
reference ignores multi-qubit channels (:787-797).
"""
from __future__ import annotations

import json

import numpy as np

_I = np.eye(2, dtype=np.complex128)
_X = np.array([[0, 1], [1, 0]], dtype=np.complex128)
_Y = np.array([[0, -1j], [1j, 0]], dtype=np.complex128)
_Z = np.array([[1, 0], [0, -1]], dtype=np.complex128)


def amplitude_damping(gamma: float):
    return [np.array([[1, 0], [0, np.sqrt(1 - gamma)]], dtype=np.complex128),
            np.array([[0, np.sqrt(gamma)], [0, 0]], dtype=np.complex128)]


def phase_damping(lam: float):
    return [np.array([[1, 0], [0, np.sqrt(1 - lam)]], dtype=np.complex128),
            np.array([[0, 0], [0, np.sqrt(lam)]], dtype=np.complex128)]


def depolarizing(p: float):
    """rho -> (1 - p) rho + p I/2 (Tr rho)."""
    return [np.sqrt(1 - 3 * p / 4) * _I, np.sqrt(p / 4) * _X, np.sqrt(p / 4) * _Y, np.sqrt(p / 4) * _Z]


def bit_flip(p: float):
    return [np.sqrt(1 - p) * _I, np.sqrt(p) * _X]


def phase_flip(p: float):
    return [np.sqrt(1 - p) * _I, np.sqrt(p) * _Z]


def compose(a, b):
    """Kraus operators of channel b after channel a."""
    return [kb @ ka for ka in a for kb in b]


def thermal_relaxation(t: float, t1: float, t2: float):
    """Amplitude damping gamma = 1 - exp(-t/T1) followed by pure dephasing with
    1/T_phi = 1/T2 - 1/(2 T1) (T2 <= 2 T1)."""
    gamma = 1.0 - np.exp(-t / t1) if t1 > 0 else 0.0
    rate_phi = (1.0 / t2 - 0.5 / t1) if (t2 > 0 and t1 > 0) else 0.0
    lam = 1.0 - np.exp(-2.0 * t * max(rate_phi, 0.0))
    return compose(amplitude_damping(gamma), phase_damping(lam))


def random_channel(rng, m: int):
    """m Kraus operators of a random CPTP map (an isometry 2 -> 2m, QR of a Gaussian)."""
    g = rng.standard_normal((2 * m, 2)) + 1j * rng.standard_normal((2 * m, 2))
    q, r = np.linalg.qr(g)
    q = q * (np.diag(r) / np.abs(np.diag(r)))
    return [q[2 * k:2 * k + 2, :] for k in range(m)]


_VIRTUAL = {"Rz", "U1", "CPhase", "T", "Tdg", "S", "Sdg", "Z", "I", "Measure"}


class NoiseModel:
    """Per-gate single-qubit channels and readout errors.

    channels(gate) -> [(qubit, [K_o ...]), ...] applied after the gate;
    readout_error(q) -> (P(measure 0 | prepared 1), P(measure 1 | prepared 0))."""

    def __init__(self, one_qubit=None, two_qubit=None, readout=None):
        self.one_qubit = one_qubit or {}  # qubit -> Kraus list after a 1-qubit gate
        self.two_qubit = two_qubit or {}  # (a, b) sorted -> {qubit: Kraus list}
        self.readout = readout or {}

    def channels(self, gate):
        q = tuple(gate["qubits"])
        if gate["name"] in _VIRTUAL:
            return []
        if len(q) == 1:
            ops = self.one_qubit.get(q[0])
            return [(q[0], ops)] if ops else []
        per = self.two_qubit.get(tuple(sorted(q)), {})
        return [(k, per[k]) for k in q if per.get(k)]

    def readout_error(self, q):
        return self.readout.get(q, (0.0, 0.0))

    @classmethod
    def from_backend_json(cls, text: str):
        d = json.loads(text)
        qubit_props = []
        for props in d["qubits"]:
            p = {x["name"]: x["value"] for x in props}
            qubit_props.append(p)

        def us(v):  # properties in microseconds, gate lengths in nanoseconds
            return 1e3 * v

        one, two = {}, {}
        for g in d["gates"]:
            p = {x["name"]: x["value"] for x in g["parameters"]}
            err, length = p.get("gate_error", 0.0), p.get("gate_length", 0.0)
            qs = g["qubits"]
            if g["gate"] in ("u2", "sx") and len(qs) == 1:
                one[qs[0]] = _gate_channel(qubit_props[qs[0]], length, err, us)
            elif g["gate"] == "cx" and len(qs) == 2:
                key = tuple(sorted(qs))
                per = {}
                for q in qs:
                    # the 2-qubit depolarizing part is not representable by single-qubit
                    # channels; it is split evenly between the two qubits
                    per[q] = _gate_channel(qubit_props[q], length, err / 2.0, us)
                two.setdefault(key, per)
        readout = {}
        for q, p in enumerate(qubit_props):
            readout[q] = (p.get("prob_meas0_prep1", p.get("readout_error", 0.0)),
                          p.get("prob_meas1_prep0", p.get("readout_error", 0.0)))
        return cls(one, two, readout)


def _gate_channel(props, length_ns, gate_error, us):
    t1, t2 = us(props.get("T1", 0.0)), us(props.get("T2", 0.0))
    relax = thermal_relaxation(length_ns, t1, min(t2, 2 * t1)) if t1 > 0 else [_I.copy()]
    # average infidelity of the relaxation part, 1 - F_avg = (2 - (|tr K|^2 sum)/2) / 3
    f_pro = sum(abs(np.trace(k)) ** 2 for k in relax) / 4.0
    relax_err = (1.0 - (2.0 * f_pro + 1.0) / 3.0)
    rest = max(0.0, gate_error - relax_err)
    p_dep = min(1.0, 2.0 * rest)  # depolarizing(p) has average infidelity p / 2
    return compose(relax, depolarizing(p_dep)) if p_dep > 0 else relax
