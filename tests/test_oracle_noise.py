"""Pinning of the noisy-simulation oracle (oracle/noise.py), CPU only.

DensityMatrix against closed forms and the pure-state oracle; PurifiedMps (the restated
ExaTnPmpsVisitor algorithm) against DensityMatrix, exactly where the reference's 2 kap bound
on the Kraus extent cannot bite and with the bound lifted where it can; the bound itself on a
case where it must drop weight.
"""
import numpy as np
import pytest

from oracle import oracle as O
from oracle.noise import DensityMatrix, PurifiedMps
from tnqvm_amd import circuits as C
from tnqvm_amd import mps
from tnqvm_amd import noise as N


def test_amplitude_damping_closed_form():
    for g in (0.0, 0.1, 0.37, 1.0):
        dm = DensityMatrix(1).gate(mps.gate_matrix("X"), (0,))
        for k in range(1, 4):
            dm.channel(0, N.amplitude_damping(g))
            assert abs(dm.probabilities([0])[1] - (1 - g) ** k) < 1e-14
            assert abs(dm.trace() - 1) < 1e-14


def test_dephasing_closed_form():
    for lam in (0.0, 0.2, 0.9):
        dm = DensityMatrix(1).gate(mps.gate_matrix("H"), (0,))
        dm.channel(0, N.phase_damping(lam))
        assert abs(dm.rho[0, 1] - 0.5 * np.sqrt(1 - lam)) < 1e-14
        assert abs(dm.rho[0, 0] - 0.5) < 1e-14


def test_depolarizing_bell_closed_form():
    for p in (0.0, 0.1, 0.5, 1.0):
        dm = DensityMatrix(2).gate(mps.gate_matrix("H"), (0,)).gate(mps.gate_matrix("CNOT"), (0, 1))
        dm.channel(0, N.depolarizing(p))
        assert abs(dm.zz(0, 1) - (1 - p)) < 1e-14
        assert abs(dm.z(0)) < 1e-14 and abs(dm.z(1)) < 1e-14


def test_random_channels_trace_preserving():
    rng = np.random.default_rng(4)
    for _ in range(5):
        ops = N.random_channel(rng, 3)
        assert np.abs(sum(k.conj().T @ k for k in ops) - np.eye(2)).max() < 1e-12
        dm = DensityMatrix(3)
        for g in C.u3_cnot_line(3, 2, seed=int(rng.integers(1 << 30))):
            dm.gate(mps.matrix_of(g), g["qubits"])
        dm.channel(1, ops)
        assert abs(dm.trace() - 1) < 1e-13
        assert np.min(np.linalg.eigvalsh(dm.rho)) > -1e-13


def test_density_matrix_matches_state_vector_oracle():
    n = 5
    gates = C.u3_cnot_line(n, 5, seed=17)
    dm = DensityMatrix(n)
    for g in gates:
        dm.gate(mps.matrix_of(g), g["qubits"])
    sv = O.StateVector(n).run(gates)
    assert np.abs(dm.rho - np.outer(sv.psi, sv.psi.conj())).max() < 1e-13
    for q in range(n):
        assert abs(dm.z(q) - sv.expval_z_parity([q])) < 1e-13


def _noisy(n, depth, seed, channels):
    gates = C.u3_cnot_line(n, depth, seed=seed)
    ops = []
    for g in gates:
        ops.append(("g", g))
        for q in g["qubits"]:
            ops.append(("c", q, channels[q % len(channels)]))
    return ops


def _run(sim, ops):
    for op in ops:
        if op[0] == "g":
            sim.gate(mps.matrix_of(op[1]), op[1]["qubits"])
        else:
            sim.channel(op[1], op[2])
    return sim


# with the 2-norm convention the 1e-9 cut only drops singular values below 1e-18 (sigma) or
# 1e-9 (Kraus weight sigma^2 = 1e-18): numerical noise survives and the extents grow, so that
# case runs on a smaller circuit
@pytest.mark.parametrize("pn,n,depth", [("squared", 5, 5), ("l2", 3, 2)])
def test_purified_mps_matches_density_matrix(pn, n, depth):
    ops = _noisy(n, depth, 23, [N.amplitude_damping(0.07), N.phase_damping(0.13), N.bit_flip(0.05)])
    dm = _run(DensityMatrix(n), ops)
    pm = _run(PurifiedMps(n, partial_norm=pn), ops)
    assert np.abs(pm.density_matrix() - dm.rho).max() < 1e-8


def test_purified_mps_expval_z_matches_its_density_matrix():
    """The environment route to <Z_q> (any n) against the dense route on the same state,
    with Kraus legs of several extents and a sign pattern no product state has."""
    n = 5
    ops = _noisy(n, 4, 37, [N.amplitude_damping(0.2), N.depolarizing(0.1), N.random_channel(np.random.default_rng(3), 3)])
    pm = _run(PurifiedMps(n), ops)
    assert max(pm.kraus_dims()) > 1
    dm = pm.as_density()
    ez = pm.expval_z()
    for q in range(n):
        assert abs(ez[q] - dm.z(q)) < 1e-12
    assert np.ptp([dm.z(q) for q in range(n)]) > 0.05


def test_purified_mps_uncapped_matches_depolarizing():
    n = 4
    ops = _noisy(n, 4, 29, [N.depolarizing(0.08), N.random_channel(np.random.default_rng(2), 4)])
    dm = _run(DensityMatrix(n), ops)
    pm = _run(PurifiedMps(n, kraus_cap=False), ops)
    assert np.abs(pm.density_matrix() - dm.rho).max() < 1e-8


def test_kraus_cap_drops_weight_on_a_bell_pair():
    # one half of a Bell pair: the site vector spans (p, r) = 4 dims with kap = 1, so a
    # depolarizing channel gives rank 4 > 2 kap: the capped split keeps 2 of 4 components
    p = 0.3
    pm = PurifiedMps(2).gate(mps.gate_matrix("H"), (0,)).gate(mps.gate_matrix("CNOT"), (0, 1))
    pm.channel(0, N.depolarizing(p))
    assert pm.kraus_dims()[0] == 2
    rho = pm.density_matrix()
    # the depolarized Bell pair has eigenvalues 1 - 3p/4, p/4, p/4, p/4; two are kept
    assert abs(np.real(np.trace(rho)) - (1 - 3 * p / 4 + p / 4)) < 1e-12
    full = PurifiedMps(2, kraus_cap=False).gate(mps.gate_matrix("H"), (0,)).gate(mps.gate_matrix("CNOT"), (0, 1))
    full.channel(0, N.depolarizing(p))
    assert abs(np.real(np.trace(full.density_matrix())) - 1) < 1e-12


def test_reduced_density_matrix():
    n = 4
    ops = _noisy(n, 3, 7, [N.amplitude_damping(0.2), N.phase_damping(0.1)])
    dm = _run(DensityMatrix(n), ops)
    assert np.abs(dm.reduced(list(range(n))) - dm.rho).max() < 1e-15  # all qubits, in order: rho
    for qs in ([1], [3, 0], [2, 0, 1]):
        r = dm.reduced(qs)
        assert abs(np.trace(r) - dm.trace()) < 1e-13
        assert np.abs(np.real(np.diag(r)) - dm.probabilities(qs)).max() < 1e-14
        assert np.abs(r - r.conj().T).max() < 1e-14
    # a product state: the reduced matrix of one qubit is its own density matrix
    p = DensityMatrix(3).gate(mps.gate_matrix("H"), (1,)).gate(mps.gate_matrix("X"), (2,))
    assert np.abs(p.reduced([1]) - 0.5 * np.ones((2, 2))).max() < 1e-15
    assert np.abs(p.reduced([2]) - np.diag([0.0, 1.0])).max() < 1e-15
