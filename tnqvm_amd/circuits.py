"""Seeded synthetic circuit generators for the hip-mps workloads.

A circuit is a list of gate dicts ``{"name": str, "qubits": tuple, "params": tuple}``
or, for explicit unitaries, ``{"name": "unitary", "qubits": tuple, "matrix": ndarray}``.
Two-qubit matrices are row-major 4x4 with the FIRST listed qubit as the most
significant bit of the row/column index (tnqvm/base/Gates.hpp:237-246).

Workloads (all seeded; the seed is recorded by the callers):
  * ``u3_cnot_line``   12-qubit U3 + alternating CNOT line (plain small case)
  * ``haar_brickwork`` Haar-random 1q/2q brickwork (40 qubits, depth 40 at full size)
  * ``qaoa_maxcut``    QAOA on a random 3-regular graph routed by a swap network
  * ``grid_fsim``      6x6 grid fSim circuits, snake-mapped onto a line
"""
from __future__ import annotations

import numpy as np

TWO_PI = 2.0 * np.pi


def haar_unitary(rng: np.random.Generator, d: int) -> np.ndarray:
    """Haar-random d x d unitary: QR of a complex Gaussian with the phase fix
    Q diag(r_ii/|r_ii|) (Mezzadri, Notices AMS 54 (2007) 592)."""
    z = (rng.standard_normal((d, d)) + 1j * rng.standard_normal((d, d))) / np.sqrt(2.0)
    q, r = np.linalg.qr(z)
    ph = np.diag(r) / np.abs(np.diag(r))
    return q * ph[None, :]


def u3_cnot_line(n: int = 12, depth: int = 20, seed: int = 0x7E5A0001):
    """Each layer: U(theta, phi, lambda ~ U[0, 2pi)) on every qubit, then CNOT on
    pairs (0,1),(2,3).. on even layers and (1,2),(3,4).. on odd layers."""
    rng = np.random.default_rng(seed)
    gates = []
    for layer in range(depth):
        for q in range(n):
            th, ph, la = rng.uniform(0.0, TWO_PI, size=3)
            gates.append({"name": "U", "qubits": (q,), "params": (float(th), float(ph), float(la))})
        start = layer % 2
        for q in range(start, n - 1, 2):
            gates.append({"name": "CNOT", "qubits": (q, q + 1), "params": ()})
    return gates


def haar_brickwork(n: int = 40, depth: int = 40, seed: int = 0x7E5A0002):
    """Each layer: a Haar 2x2 on every qubit, then Haar 4x4 gates on the even
    (layer even) or odd (layer odd) nearest-neighbour pairs."""
    rng = np.random.default_rng(seed)
    gates = []
    for layer in range(depth):
        for q in range(n):
            gates.append({"name": "unitary", "qubits": (q,), "matrix": haar_unitary(rng, 2)})
        start = layer % 2
        for q in range(start, n - 1, 2):
            gates.append({"name": "unitary", "qubits": (q, q + 1), "matrix": haar_unitary(rng, 4)})
    return gates


def brickwork_layers(gates, n: int):
    """Split a brickwork circuit (haar_brickwork / u3_cnot_line) into its layers: layer l
    holds n one-qubit gates followed by the 2-qubit gates on pairs starting at l % 2."""
    out, pos, layer = [], 0, 0
    while pos < len(gates):
        size = n + len(range(layer % 2, n - 1, 2))
        out.append(gates[pos:pos + size])
        pos += size
        layer += 1
    return out


def random_bitstrings(n: int, k: int, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed)
    return rng.integers(0, 2, size=(k, n), dtype=np.int8)


def count_2q(gates) -> int:
    return sum(1 for g in gates if len(g["qubits"]) == 2)


def layers_2q(gates):
    """Split a gate list into consecutive groups of gates acting on disjoint qubits
    (a greedy as-soon-as-possible layering that preserves the order on each qubit)."""
    out = []
    cur, busy = [], set()
    for g in gates:
        qs = set(g["qubits"])
        if qs & busy:
            out.append(cur)
            cur, busy = [], set()
        cur.append(g)
        busy |= qs
    if cur:
        out.append(cur)
    return out
