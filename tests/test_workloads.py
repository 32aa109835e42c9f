"""CPU checks of the seeded workload generators (QAOA swap network, 2D grid with snake
map and routing, mixed batch) against the dense FP64 oracle.

The routed / swap-network forms are compared with the plain logical circuits on the
dense state vector, so a wrong swap, permutation or coupler pattern fails here before
any GPU run.  The oracle MPS restatement is run on the routed forms without
truncation and must reproduce the dense state to 1e-12.
"""
import itertools

import numpy as np
import pytest

from oracle import oracle as O
from tnqvm_amd import circuits as C


def permute_state(psi, n, pos):
    """psi over sites -> psi over logical qubits, logical i sits on site pos[i]
    (state-vector index bit k = qubit k)."""
    t = psi.reshape([2] * n)  # axis a <-> bit n-1-a
    axes = [n - 1 - pos[n - 1 - a] for a in range(n)]
    return np.transpose(t, axes).reshape(-1)


def test_regular_graph_is_simple_3_regular_and_seeded():
    for n in (8, 12, 60):
        e1 = C.random_regular_graph(n, 3, np.random.default_rng(5))
        e2 = C.random_regular_graph(n, 3, np.random.default_rng(5))
        assert e1 == e2
        assert len(e1) == 3 * n // 2
        deg = np.zeros(n, int)
        for i, j in e1:
            assert i < j
            deg[i] += 1
            deg[j] += 1
        assert (deg == 3).all()
    with pytest.raises(ValueError):
        C.random_regular_graph(7, 3, np.random.default_rng(0))


def test_qaoa_swap_network_meets_every_pair_once():
    n = 9
    phys = list(range(n))
    met = set()
    for t in range(n):
        for s0 in range(t % 2, n - 1, 2):
            pair = tuple(sorted((phys[s0], phys[s0 + 1])))
            assert pair not in met
            met.add(pair)
            phys[s0], phys[s0 + 1] = phys[s0 + 1], phys[s0]
    assert met == set(itertools.combinations(range(n), 2))
    assert phys == list(range(n))[::-1]


@pytest.mark.parametrize("n,p", [(8, 1), (10, 2), (6, 3)])
def test_qaoa_swap_network_equals_logical_circuit(n, p):
    gates, edges, pos, (gam, bet) = C.qaoa_maxcut(n, p, seed=11 + n)
    assert all(len(g["qubits"]) == 1 or abs(g["qubits"][0] - g["qubits"][1]) == 1 for g in gates)
    assert C.count_2q(gates) == p * n * (n - 1) // 2
    routed = O.StateVector(n).run(gates)
    logical = O.StateVector(n).run(C.qaoa_logical(n, edges, gam, bet))
    psi = permute_state(routed.psi, n, pos)
    assert np.abs(psi - logical.psi).max() < 1e-12
    e_routed = sum(routed.expval_z_parity([pos[i], pos[j]]) for i, j in edges)
    e_logic = sum(logical.expval_z_parity([i, j]) for i, j in edges)
    assert abs(e_routed - e_logic) < 1e-12
    assert abs(e_logic) <= len(edges) + 1e-12


def test_qaoa_p1_closed_form():
    """p=1 QAOA on a triangle-free graph (Wang, Hadfield, Jiang, Rieffel, PRA 97, 022304
    (2018)): with U_C = exp(-i gamma C), C = sum (1 - Z_u Z_v)/2 and U_B = exp(-i beta X),
    <Z_u Z_v> = -1/2 sin(4 beta) sin(gamma) (cos^du(gamma) + cos^dv(gamma)), du = deg(u) - 1.
    The generator's exp(-i g Z Z) is U_C with gamma = -2g up to a global phase, and its
    mixer Rx(2b) is U_B with beta = b, so on a 3-regular graph
    <Z_u Z_v> = sin(4b) sin(2g) cos^2(2g)."""
    n = 8
    # the cube graph Q3: 3-regular and bipartite (no triangles)
    edges = sorted({tuple(sorted((v, v ^ (1 << k)))) for v in range(8) for k in range(3)})
    for g, b in ((0.37, 1.21), (2.5, 0.3)):
        sv = O.StateVector(n).run(C.qaoa_logical(n, edges, [g], [b]))
        ref = np.sin(4 * b) * np.sin(2 * g) * np.cos(2 * g) ** 2
        for (u, v) in edges:
            assert abs(sv.expval_z_parity([u, v]) - ref) < 1e-12


def test_snake_map_and_couplers():
    rows, cols = 6, 6
    sites = {C.snake_site(r, c, cols) for r in range(rows) for c in range(cols)}
    assert sites == set(range(36))
    pats = C.grid_couplers(rows, cols)
    allc = set(pats["A"]) | set(pats["B"]) | set(pats["C"]) | set(pats["D"])
    assert len(allc) == rows * (cols - 1) + (rows - 1) * cols
    for ((r0, c0), (r1, c1)) in allc:
        assert abs(r0 - r1) + abs(c0 - c1) == 1
    dist = sorted({abs(C.snake_site(*a, cols) - C.snake_site(*b, cols)) for a, b in allc})
    assert dist[0] == 1 and dist[-1] == 11
    # every pattern acts on disjoint qubits
    for k in "ABCD":
        qs = [q for a, b in pats[k] for q in (a, b)]
        assert len(qs) == len(set(qs))


def test_sqrt_w_is_the_qflex_gate():
    """sqrt(W) = Rz(-pi/4) Rx(pi/2) Rz(pi/4) squares to W = (X + Y)/sqrt(2) up to phase."""
    rz = lambda t: O.gate_matrix("Rz", (t,))
    # applied in time order Rz(-pi/4), Rx(pi/2), Rz(pi/4)
    m = rz(np.pi / 4) @ O.gate_matrix("Rx", (np.pi / 2,)) @ rz(-np.pi / 4)
    w = np.array([[0, 1 - 1j], [1 + 1j, 0]]) / np.sqrt(2)
    m2 = m @ m
    ph = m2[0, 1] / w[0, 1]
    assert abs(abs(ph) - 1) < 1e-14
    assert np.abs(m2 - ph * w).max() < 1e-14


@pytest.mark.parametrize("rows,cols,cycles", [(3, 3, 8), (4, 3, 6), (2, 5, 8)])
def test_grid_routing_equals_unrouted(rows, cols, cycles):
    n = rows * cols
    gates = C.grid_fsim(rows, cols, cycles, seed=3 + n)
    routed = C.route_nearest_neighbour(gates)
    assert all(len(g["qubits"]) == 1 or abs(g["qubits"][0] - g["qubits"][1]) == 1 for g in routed)
    a = O.StateVector(n).run(gates)
    b = O.StateVector(n).run(routed)
    assert np.abs(a.psi - b.psi).max() < 1e-12
    m = O.Mps(n, cutoff=0.0).run(routed)
    assert np.abs(m.state_vector() - a.psi).max() < 1e-12


def test_route_keeps_operand_order():
    rng = np.random.default_rng(4)
    n = 7
    gates = []
    for _ in range(30):
        q0, q1 = rng.choice(n, size=2, replace=False)
        gates.append({"name": "unitary", "qubits": (int(q0), int(q1)), "matrix": C.haar_unitary(rng, 4)})
    a = O.StateVector(n).run(gates)
    b = O.StateVector(n).run(C.route_nearest_neighbour(gates))
    assert np.abs(a.psi - b.psi).max() < 1e-12


def test_mixed_batch_ranges_and_seed():
    b1 = C.mixed_batch(40, seed=9)
    b2 = C.mixed_batch(40, seed=9)
    for x, y in zip(b1, b2):
        assert (x["n"], x["depth"], x["chi"]) == (y["n"], y["depth"], y["chi"])
        assert 16 <= x["n"] <= 48 and 10 <= x["depth"] <= 60 and x["chi"] in (32, 64, 128, 256)
        assert C.count_2q(x["gates"]) == sum(len(range(l % 2, x["n"] - 1, 2)) for l in range(x["depth"]))


# ------------------------------------------------------------------ qFlex ingest
import os
import re

SYC = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "sycamore")


def _read_xasm(path):
    """Gate list of the reference converter's XASM output (fixture)."""
    out = []
    pat = re.compile(r"^(\w+)\(q\[(\d+)\](?:, q\[(\d+)\])?((?:, [-+0-9.eE]+)*)\);")
    for line in open(path):
        m = pat.match(line.strip())
        if not m:
            continue
        qs = (int(m.group(2)),) if m.group(3) is None else (int(m.group(2)), int(m.group(3)))
        ps = tuple(float(x) for x in m.group(4).split(",")[1:]) if m.group(4) else ()
        out.append({"name": m.group(1), "qubits": qs, "params": ps})
    return out


@pytest.mark.parametrize("conj,fixture", [(0, "sycamore_53_10_0.xasm"), (10, "sycamore_53_10_0_Conjugate.xasm")])
def test_qflex_reader_matches_reference_converter(conj, fixture):
    """The reference's own converter output (examples/sycamore/resources) is the golden
    file: names, qubits and parameters must agree bit for bit."""
    text = open(os.path.join(SYC, "sycamore_53_10_0.txt")).read()
    mine = C.read_qflex(text, conjugate_last_layer=conj)
    ref = _read_xasm(os.path.join(SYC, fixture))
    assert len(mine) == len(ref)
    for a, b in zip(mine, ref):
        assert a["name"] == b["name"] and tuple(a["qubits"]) == b["qubits"], (a, b)
        assert tuple(a["params"]) == b["params"], (a, b)


def test_qflex_subcircuit_routed_equals_dense():
    text = open(os.path.join(SYC, "sycamore_53_10_0.txt")).read()
    gates = C.restrict(C.read_qflex(text), 12)
    assert C.count_2q(gates) > 20
    a = O.StateVector(12).run(gates)
    b = O.StateVector(12).run(C.route_nearest_neighbour(gates))
    assert np.abs(a.psi - b.psi).max() < 1e-12


# ------------------------------------------------------------------ KAK emission
def _matrix_of_2q(gates, n=2):
    """4x4 matrix (first qubit = MSB) of a gate list on qubits 0, 1 through the dense oracle."""
    cols = []
    for k in range(4):
        s = O.StateVector(n)
        s.psi[:] = 0
        s.psi[(k >> 1) | ((k & 1) << 1)] = 1
        s.run(gates)
        cols.append([s.psi[(j >> 1) | ((j & 1) << 1)] for j in range(4)])
    return np.array(cols).T


def _equal_up_to_phase(a, b, tol):
    k = np.argmax(np.abs(b))
    ph = a.flat[k] / b.flat[k]
    return abs(abs(ph) - 1) < tol and np.abs(a - ph * b).max() < tol


def test_kak_emission_reconstructs_haar_and_named_gates():
    rng = np.random.default_rng(17)
    mats = [C.haar_unitary(rng, 4) for _ in range(60)]
    mats += [O.gate_matrix(n, p) for n, p in [("CNOT", ()), ("CZ", ()), ("Swap", ()), ("iSwap", ()),
                                              ("fSim", (0.3, 0.7)), ("CPhase", (0.4,))]]
    mats += [np.eye(4), np.kron(C.haar_unitary(rng, 2), C.haar_unitary(rng, 2))]
    for u in mats:
        g = C.kak_emit(u, 0, 1)
        assert sum(x["name"] == "CNOT" for x in g) == 3 and sum(x["name"] == "U" for x in g) == 8
        assert _equal_up_to_phase(_matrix_of_2q(g), u, 1e-12)


def test_kak_coordinates_of_known_gates():
    """CNOT is locally exp(i pi/4 XX) (one coordinate pi/4, two zero); SWAP has all three pi/4."""
    _, abc, _ = C.kak(O.gate_matrix("CNOT", ()))
    assert sorted(np.round(np.abs(abc), 12)) == [0, 0, round(np.pi / 4, 12)]
    _, abc, _ = C.kak(O.gate_matrix("Swap", ()))
    assert np.allclose(np.abs(abc), np.pi / 4, atol=1e-12)


def test_reference_gate_brickwork_equals_unitary_brickwork():
    n = 8
    gates = C.haar_brickwork(n, 6, seed=3)
    ref = C.to_reference_gates(gates)
    assert all(g["name"] in ("U", "CNOT") for g in ref)
    a = O.StateVector(n).run(gates)
    b = O.StateVector(n).run(ref)
    assert abs(abs(np.vdot(a.psi, b.psi)) - 1) < 1e-12
    for q in range(n):
        assert abs(a.expval_z_parity([q]) - b.expval_z_parity([q])) < 1e-12


def test_bench_refuses_several_gpus():
    """bench.py measures one GPU: asked for more, it exits with code 2 and says why, before it
    touches a device."""
    import os
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 2
    assert "one GPU only" in r.stderr and r.stdout == ""
