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
  * ``mixed_batch``    many brickwork circuits of uneven width, depth and chi_max
  * ``read_qflex``     qFlex circuit files (Google Sycamore random circuits)
  * ``to_reference_gates``  explicit unitaries as U3 / CNOT (KAK: 3 CNOT + 8 U3 per 2q gate)
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


def route_nearest_neighbour(gates):
    """Replace every 2-qubit gate on qubits at distance > 1 by swaps that move the two
    qubits towards each other from both ends, the gate on the met pair, and the inverse
    swaps (NearestNeighborTransform.hpp:77-121, the same scheme as the hip-mps visitor)."""
    out = []
    for g in gates:
        q = g["qubits"]
        if len(q) != 2 or abs(q[0] - q[1]) <= 1:
            out.append(g)
            continue
        lower0, upper0 = min(q), max(q)
        lo, up = lower0, upper0
        while True:
            out.append({"name": "Swap", "qubits": (lo, lo + 1), "params": ()})
            lo += 1
            if up - lo <= 1:
                break
            out.append({"name": "Swap", "qubits": (up, up - 1), "params": ()})
            up -= 1
            if up - lo <= 1:
                break
        moved = dict(g)
        moved["qubits"] = (lo, up) if q[0] < q[1] else (up, lo)
        out.append(moved)
        for i in range(lo, lower0, -1):
            out.append({"name": "Swap", "qubits": (i, i - 1), "params": ()})
        for i in range(up, upper0):
            out.append({"name": "Swap", "qubits": (i, i + 1), "params": ()})
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


# ------------------------------------------------------------------ QAOA MaxCut
def random_regular_graph(n: int, d: int, rng: np.random.Generator, max_tries: int = 1000):
    """Seeded d-regular simple graph by the pairing (configuration) model with rejection."""
    if (n * d) % 2:
        raise ValueError("n * d must be even")
    for _ in range(max_tries):
        stubs = np.repeat(np.arange(n), d)
        rng.shuffle(stubs)
        pairs = stubs.reshape(-1, 2)
        if np.any(pairs[:, 0] == pairs[:, 1]):
            continue
        edges = {tuple(sorted(map(int, p))) for p in pairs}
        if len(edges) == len(pairs):
            return sorted(edges)
    raise RuntimeError("no simple regular graph found")


def zz_phase(gamma: float) -> np.ndarray:
    """exp(-i gamma Z x Z) = CNOT . (I x Rz(2 gamma)) . CNOT."""
    return np.diag(np.exp(-1j * gamma * np.array([1, -1, -1, 1]))).astype(np.complex128)


SWAP = np.array([[1, 0, 0, 0], [0, 0, 1, 0], [0, 1, 0, 0], [0, 0, 0, 1]], dtype=np.complex128)


def qaoa_maxcut(n: int = 60, p: int = 4, seed: int = 0x7E5A0003):
    """QAOA MaxCut on a seeded random 3-regular graph, gamma, beta ~ U[0, pi].
    Each cost layer runs an odd-even transposition swap network (n layers of alternating
    neighbour swaps): every logical pair becomes adjacent exactly once, and an edge's ZZ
    phase is fused with the swap that passes it (one 4x4 gate).  Mixer: Rx(2 beta).
    Returns (gates, edges, final_position) with final_position[logical] = physical site."""
    rng = np.random.default_rng(seed)
    edges = random_regular_graph(n, 3, rng)
    eset = set(edges)
    gammas = rng.uniform(0.0, np.pi, size=p)
    betas = rng.uniform(0.0, np.pi, size=p)
    gates = [{"name": "H", "qubits": (q,), "params": ()} for q in range(n)]
    phys = list(range(n))  # phys[site] = logical qubit on that site
    for layer in range(p):
        zz = zz_phase(gammas[layer])
        fused = SWAP @ zz
        for t in range(n):
            for s0 in range(t % 2, n - 1, 2):
                a, b = phys[s0], phys[s0 + 1]
                m = fused if tuple(sorted((a, b))) in eset else SWAP
                gates.append({"name": "unitary", "qubits": (s0, s0 + 1), "matrix": m})
                phys[s0], phys[s0 + 1] = b, a
        for q in range(n):
            gates.append({"name": "Rx", "qubits": (q,), "params": (float(2.0 * betas[layer]),)})
    pos = [0] * n
    for site, logical in enumerate(phys):
        pos[logical] = site
    return gates, edges, pos, (gammas, betas)


def qaoa_logical(n: int, edges, gammas, betas):
    """The same QAOA circuit on logical qubits without routing (for dense checks)."""
    gates = [{"name": "H", "qubits": (q,), "params": ()} for q in range(n)]
    for g, b in zip(gammas, betas):
        for (i, j) in edges:
            gates.append({"name": "unitary", "qubits": (i, j), "matrix": zz_phase(g)})
        for q in range(n):
            gates.append({"name": "Rx", "qubits": (q,), "params": (float(2.0 * b),)})
    return gates


# ------------------------------------------------------------------ 2D grid fSim
def snake_site(r: int, c: int, cols: int) -> int:
    return r * cols + (c if r % 2 == 0 else cols - 1 - c)


def grid_couplers(rows: int, cols: int):
    """Coupler sets A, B (horizontal, even / odd column) and C, D (vertical, even / odd row)."""
    A = [((r, c), (r, c + 1)) for r in range(rows) for c in range(0, cols - 1, 2)]
    B = [((r, c), (r, c + 1)) for r in range(rows) for c in range(1, cols - 1, 2)]
    Cv = [((r, c), (r + 1, c)) for r in range(0, rows - 1, 2) for c in range(cols)]
    D = [((r, c), (r + 1, c)) for r in range(1, rows - 1, 2) for c in range(cols)]
    return {"A": A, "B": B, "C": Cv, "D": D}


def grid_fsim(rows: int = 6, cols: int = 6, cycles: int = 8, seed: int = 0x7E5A0005,
              theta: float = np.pi / 2, phi: float = np.pi / 6):
    """Random grid circuit: each cycle applies sqrt(X) = Rx(pi/2), sqrt(Y) = Ry(pi/2) or
    sqrt(W) = Rz(-pi/4) Rx(pi/2) Rz(pi/4) (never the same on a qubit twice in a row) to every
    qubit, then fSim(theta, phi) on one coupler pattern (ABCDCDAB order).  Qubits are placed
    on the line by the snake map site = cols*r + (c if r even else cols-1-c); vertical
    couplers become non-neighbouring 2-qubit gates (route_nearest_neighbour makes the
    circuit nearest-neighbour)."""
    rng = np.random.default_rng(seed)
    pats = grid_couplers(rows, cols)
    order = "ABCDCDAB"
    n = rows * cols
    last = [-1] * n
    gates = []
    for cyc in range(cycles):
        for q in range(n):
            k = int(rng.integers(0, 3))
            while k == last[q]:
                k = int(rng.integers(0, 3))
            last[q] = k
            if k == 0:
                gates.append({"name": "Rx", "qubits": (q,), "params": (np.pi / 2,)})
            elif k == 1:
                gates.append({"name": "Ry", "qubits": (q,), "params": (np.pi / 2,)})
            else:
                gates.append({"name": "Rz", "qubits": (q,), "params": (-np.pi / 4,)})
                gates.append({"name": "Rx", "qubits": (q,), "params": (np.pi / 2,)})
                gates.append({"name": "Rz", "qubits": (q,), "params": (np.pi / 4,)})
        for (a, b) in pats[order[cyc % len(order)]]:
            sa, sb = snake_site(*a, cols), snake_site(*b, cols)
            gates.append({"name": "fSim", "qubits": (sa, sb), "params": (float(theta), float(phi))})
    return gates


# ------------------------------------------------------------------ batch of uneven circuits
def mixed_batch(count: int = 256, seed: int = 0x7E5A0004, n_range=(16, 48), depth_range=(10, 60),
                chis=(32, 64, 128, 256)):
    """Brickwork circuits with n ~ U{n_range}, depth ~ U{depth_range}, chi_max from `chis`."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        n = int(rng.integers(n_range[0], n_range[1] + 1))
        d = int(rng.integers(depth_range[0], depth_range[1] + 1))
        chi = int(rng.choice(chis))
        out.append({"n": n, "depth": d, "chi": chi, "gates": haar_brickwork(n, d, seed=seed * 1000 + k)})
    return out


# ------------------------------------------------------------------ qFlex circuit files
# Sycamore lattice index -> position on the line (the converter's table,
# examples/sycamore/parse_qflex.py:9-61).
SYCAMORE_53 = dict(zip(
    [5, 6, 14, 15, 16, 17, 24, 25, 26, 27, 28, 32, 33, 34, 35, 36, 37, 38, 39, 41, 42, 43, 44, 45, 46, 47, 48,
     49, 50, 51, 52, 53, 54, 55, 56, 57, 58, 61, 62, 63, 64, 65, 66, 67, 72, 73, 74, 75, 76, 83, 84, 85, 94],
    range(53)))
# qFlex gives angles as fractions of pi; the reference converter multiplies by this
# truncated value (parse_qflex.py:75, :115-116), and the same constant is used here so that
# the gate parameters are bit-identical to the XASM it produced.
QFLEX_PI = 3.14159265359
_HALF_PI = 1.57079632679
_QUARTER_PI = 0.78539816339


def _qflex_line_gates(line: str, qubit_map, conjugate: bool):
    """Gates of one qFlex line "<cycle> <gate> <qubit> [<qubit>]" (parse_qflex.py:65-124)."""
    parts = line.replace(", ", ",").split()
    if len(parts) == 3:
        _, name, q = parts
        q = qubit_map[int(q)]
        sign = -1.0 if conjugate else 1.0
        if name == "hz_1_2":
            # Cirq PhasedXPowGate(0.25, 0.5) as Rz(-pi/4) Rx(pi/2) Rz(pi/4) in time order; the
            # converter's conjugate form only negates the Rx angle (parse_qflex.py:83-98)
            return [{"name": "Rz", "qubits": (q,), "params": (-_QUARTER_PI,)},
                    {"name": "Rx", "qubits": (q,), "params": (sign * _HALF_PI,)},
                    {"name": "Rz", "qubits": (q,), "params": (_QUARTER_PI,)}]
        if name == "x_1_2":
            return [{"name": "Rx", "qubits": (q,), "params": (sign * _HALF_PI,)}]
        if name == "y_1_2":
            return [{"name": "Ry", "qubits": (q,), "params": (sign * _HALF_PI,)}]
        if name.lower().startswith("rz("):
            return [{"name": "Rz", "qubits": (q,), "params": (sign * (float(name[3:-1]) * QFLEX_PI),)}]
        raise ValueError(f"unknown qFlex 1-qubit gate {name!r}")
    if len(parts) == 4:
        _, name, q0, q1 = parts
        if not name.startswith("fsim("):
            raise ValueError(f"unknown qFlex 2-qubit gate {name!r}")
        th, ph = (float(x) * QFLEX_PI for x in name[5:-1].split(","))
        if conjugate:
            th, ph = -th, -ph
        return [{"name": "fSim", "qubits": (qubit_map[int(q0)], qubit_map[int(q1)]), "params": (th, ph)}]
    raise ValueError(f"malformed qFlex line {line!r}")


def read_qflex(text: str, qubit_map=None, conjugate_last_layer: int = 0):
    """Gate list of a qFlex circuit file (first line: qubit count; then one gate per line,
    "<cycle> <gate> <qubits>").  Qubits are placed on the line by `qubit_map` (default:
    the 53-qubit Sycamore map).  With conjugate_last_layer = L > 0, the gates of layer L-1
    (cycles 4(L-1)+1 .. 4L) are appended again, reversed and conjugated, as the reference
    converter does for its debug circuits (parse_qflex.py:126-195)."""
    qubit_map = SYCAMORE_53 if qubit_map is None else qubit_map
    lines = [ln for ln in text.splitlines() if ln.strip()]
    gates, last = [], []
    for ln in lines[1:]:
        cycle = int(ln.split()[0])
        gates.extend(_qflex_line_gates(ln, qubit_map, False))
        if conjugate_last_layer and cycle > 0 and (cycle - 1) // 4 == conjugate_last_layer - 1:
            last.append(ln)
    for ln in reversed(last):
        gates.extend(_qflex_line_gates(ln, qubit_map, True))
    return gates


def restrict(gates, keep: int):
    """The gates that act only on qubits < keep (a sub-circuit small enough for dense checks)."""
    return [g for g in gates if max(g["qubits"]) < keep]


# ------------------------------------------------------------------ KAK emission (3 CNOT + 8 U3)
# Magic basis: B^H (A x C) B is real orthogonal for A, C in SU(2), and XX, YY, ZZ are
# diagonal in it (Kraus & Cirac, PRA 63, 062309 (2001); Vatan & Williams, PRA 69,
# 032315 (2004)).
_MAGIC = np.array([[1, 0, 0, 1j], [0, 1j, 1, 0], [0, 1j, -1, 0], [1, 0, 0, -1j]]) / np.sqrt(2.0)
_PX = np.array([[0, 1], [1, 0]], dtype=np.complex128)
_PY = np.array([[0, -1j], [1j, 0]], dtype=np.complex128)
_PZ = np.diag([1.0, -1.0]).astype(np.complex128)


def _rz(t):
    return np.diag([np.exp(-0.5j * t), np.exp(0.5j * t)])


def _ry(t):
    c, s = np.cos(0.5 * t), np.sin(0.5 * t)
    return np.array([[c, -s], [s, c]], dtype=np.complex128)


def u3_params(v: np.ndarray):
    """(theta, phi, lambda) with v = e^{i g} U3(theta, phi, lambda), U3 as in Gates.hpp:224-235:
    [[cos t/2, -e^{i l} sin t/2], [e^{i p} sin t/2, e^{i(p+l)} cos t/2]]."""
    a00, a01, a10, a11 = v[0, 0], v[0, 1], v[1, 0], v[1, 1]
    theta = 2.0 * np.arctan2(abs(a10), abs(a00))
    g = np.angle(a00) if abs(a00) > 1e-12 else np.angle(a10)
    phi = np.angle(a10) - g if abs(a10) > 1e-12 else 0.0
    lam = np.angle(-a01) - g if abs(a01) > 1e-12 else np.angle(a11) - g - phi
    return float(theta), float(phi), float(lam)


def _split_local(k: np.ndarray):
    """A1, A2 with k = A1 x A2 (rank-1 factor of the realigned 4x4)."""
    r = k.reshape(2, 2, 2, 2).transpose(0, 2, 1, 3).reshape(4, 4)
    u, s, vh = np.linalg.svd(r)
    return np.sqrt(s[0]) * u[:, 0].reshape(2, 2), np.sqrt(s[0]) * vh[0].reshape(2, 2)


def kak(u: np.ndarray):
    """u = phase (A1 x A2) exp(i(a XX + b YY + c ZZ)) (B1 x B2); returns ((A1, A2), (a, b, c), (B1, B2))."""
    u = np.asarray(u, dtype=np.complex128)
    u = u / np.linalg.det(u) ** 0.25
    um = _MAGIC.conj().T @ u @ _MAGIC
    m2 = um.T @ um  # symmetric unitary; Re and Im commute and share a real eigenbasis
    rng = np.random.default_rng(0x7E5A)
    for _ in range(32):
        _, p = np.linalg.eigh(m2.real + rng.standard_normal() * m2.imag)
        d = p.T @ m2 @ p
        if np.abs(d - np.diag(np.diag(d))).max() < 1e-9:
            break
    else:
        raise RuntimeError("kak: no common real eigenbasis")
    if np.linalg.det(p) < 0:
        p[:, 0] *= -1.0
    th = 0.5 * np.angle(np.diag(p.T @ m2 @ p))
    if np.prod(np.exp(1j * th)).real < 0:
        th[0] += np.pi
    o1 = um @ p @ np.diag(np.exp(-1j * th))
    if np.abs(o1.imag).max() > 1e-7:
        raise RuntimeError("kak: left factor is not real")
    k1 = _MAGIC @ o1.real @ _MAGIC.conj().T
    k2 = _MAGIC @ p.T @ _MAGIC.conj().T
    basis = np.stack([np.diag(_MAGIC.conj().T @ np.kron(m, m) @ _MAGIC).real for m in (_PX, _PY, _PZ)]
                     + [np.ones(4)], axis=1)
    a, b, c, _ = np.linalg.solve(basis, th)
    return _split_local(k1), (float(a), float(b), float(c)), _split_local(k2)


def kak_emit(u: np.ndarray, q0: int, q1: int):
    """Gates (U3 and CNOT, in time order) equal to the 4x4 u on (q0, q1), q0 = most significant
    bit of u's index, up to a global phase.  exp(i(aXX + bYY + cZZ)) is emitted as
    (I x Rz(pi/2)), CNOT(q1,q0), (Rz(pi/2-2c) x Ry(pi/2-2a)), CNOT(q0,q1), (I x Ry(2b-pi/2)),
    CNOT(q1,q0), (Rz(-pi/2) x I) (Vatan & Williams, Fig. 6), and the one-qubit factors are merged
    into the surrounding U3s: 3 CNOT + 8 U3."""
    (a1, a2), (a, b, c), (b1, b2) = kak(u)
    seq = [(q0, b1), (q1, _rz(np.pi / 2) @ b2), ("CNOT", q1, q0),
           (q0, _rz(np.pi / 2 - 2 * c)), (q1, _ry(np.pi / 2 - 2 * a)), ("CNOT", q0, q1),
           (q0, np.eye(2)), (q1, _ry(2 * b - np.pi / 2)), ("CNOT", q1, q0),
           (q0, a1 @ _rz(-np.pi / 2)), (q1, a2)]
    out = []
    for it in seq:
        if it[0] == "CNOT":
            out.append({"name": "CNOT", "qubits": (it[1], it[2]), "params": ()})
        else:
            out.append({"name": "U", "qubits": (it[0],), "params": u3_params(it[1])})
    return out


def to_reference_gates(gates):
    """Explicit unitaries rewritten in the reference gate set: 1q -> U3, 2q -> 3 CNOT + 8 U3.
    Each gate is reproduced up to a global phase."""
    out = []
    for g in gates:
        if g["name"] != "unitary":
            out.append(g)
        elif len(g["qubits"]) == 1:
            out.append({"name": "U", "qubits": g["qubits"], "params": u3_params(np.asarray(g["matrix"]))})
        else:
            out.extend(kak_emit(g["matrix"], *g["qubits"]))
    return out
