"""Noisy circuits on the purified MPS (the exatn-pmps path) on the GPU.

* Channels of Kraus rank 2 (amplitude damping, phase damping, bit flip) after every gate:
  <Z_i>, <Z_i Z_j>, outcome probabilities and the trace against the dense density matrix
  (FP64 1e-8: the 1e-9 cut of Kraus vectors and bonds is the only approximation; mixed
  1e-5), and the Kraus extents against the restated reference algorithm.
* Depolarizing noise where the reference's 2 kap bound on the Kraus extent drops weight:
  against the restated algorithm (which applies the same bound).
* Noise-model runs in the shape of the reference's own tests (ExaTnPmpsTester.cpp), with the
  ibmqx2 backend JSON fixture from the reference's test resources.
* A purified state refuses amplitudes; clone keeps the Kraus legs.
* Reduced and full density matrices up to the reference's 15-qubit limit, against the dense
  density matrix, the state vector and the GHZ closed form.
"""
import os
import sys

import numpy as np
import pytest

from oracle import oracle as O
from oracle.noise import DensityMatrix, PurifiedMps
from tnqvm_amd import _capi as capi
from tnqvm_amd import circuits as C
from tnqvm_amd import mps
from tnqvm_amd import noise as N

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def eng():
    e = mps.Engine()
    yield e
    e.close()


def noisy_ops(n, depth, seed, channels):
    ops = []
    for g in C.u3_cnot_line(n, depth, seed=seed):
        ops.append(("g", g))
        for q in g["qubits"]:
            ops.append(("c", q, channels[q % len(channels)]))
    return ops


def run_oracle(sim, ops):
    for op in ops:
        if op[0] == "g":
            sim.gate(mps.matrix_of(op[1]), op[1]["qubits"])
        else:
            sim.channel(op[1], op[2])
    return sim


def run_device(eng, n, ops, precision):
    s = mps.MpsState(eng, n, cutoff=1e-9, precision=precision)
    for op in ops:
        if op[0] == "g":
            s.run([op[1]])
        else:
            s.apply_kraus(op[1], op[2])
    return s


RANK2 = [N.amplitude_damping(0.07), N.phase_damping(0.13), N.bit_flip(0.05)]


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-8), ("mixed", 1e-5)])
@pytest.mark.parametrize("n,depth", [(4, 4), (6, 5)])
def test_noisy_vs_density_matrix(eng, n, depth, precision, tol):
    say("noisy vs dense", n, depth, precision)
    ops = noisy_ops(n, depth, 31 + n, RANK2)
    dm = run_oracle(DensityMatrix(n), ops)
    s = run_device(eng, n, ops, precision)
    ez = s.expval_z_all()
    for q in range(n):
        assert abs(ez[q] - dm.z(q)) < tol
    pairs = [(q, q + 1) for q in range(n - 1)] + [(0, n - 1)]
    zz = s.expval_zz(pairs)
    for (a, b), v in zip(pairs, zz):
        assert abs(v - dm.zz(a, b)) < tol
    meas = [0, n // 2, n - 1]
    assert np.abs(s.outcome_probabilities(meas) - dm.probabilities(meas)).max() < tol
    assert abs(s.norm() - dm.trace()) < tol
    if precision == "fp64":
        pm = run_oracle(PurifiedMps(n), ops)
        assert list(s.kraus_dims()) == pm.kraus_dims()
    s.close()


def test_kraus_cap_matches_restated_reference(eng):
    say("kraus cap")
    n = 5
    # entangle first, then depolarize every qubit: each site then has rank-4 channel output
    # against a bound of 2 kap, so the bound drops weight; then more gates and channels
    ops = [("g", g) for g in C.u3_cnot_line(n, 3, seed=41)]
    ops += [("c", q, N.depolarizing(0.2)) for q in range(n)]
    for g in C.u3_cnot_line(n, 2, seed=42):
        ops.append(("g", g))
        ops += [("c", q, N.random_channel(np.random.default_rng(5), 3)) for q in g["qubits"]]
    pm = run_oracle(PurifiedMps(n), ops)
    full = run_oracle(DensityMatrix(n), ops)
    s = run_device(eng, n, ops, "fp64")
    ez = s.expval_z_all()
    ref = pm.as_density()
    for q in range(n):
        assert abs(ez[q] - ref.z(q)) < 1e-8
    assert abs(s.norm() - ref.trace()) < 1e-8
    assert list(s.kraus_dims()) == pm.kraus_dims()
    assert ref.trace() < full.trace() - 1e-6  # the bound did drop weight in this circuit
    s.close()


def test_bell_depolarized(eng):
    say("bell depolarized")
    p = 0.3
    s = mps.MpsState(eng, 2, cutoff=1e-9)
    s.run([{"name": "H", "qubits": (0,)}, {"name": "CNOT", "qubits": (0, 1)}])
    s.apply_kraus(0, N.depolarizing(p))
    # the reference's bound keeps 2 of the 4 eigen-components: weight 1 - 3p/4 + p/4
    assert abs(s.norm() - (1 - p / 2)) < 1e-12
    assert list(s.kraus_dims()) == [2, 1]
    s.close()


def _model():
    with open(os.path.join(HERE, "golden", "noise", "ibmqx2.json")) as f:
        return N.NoiseModel.from_backend_json(f.read())


def test_reference_scenarios(eng):
    model = _model()
    say("x100 amplitude damping")
    r = mps.run_noisy(eng, 1, [{"name": "X", "qubits": (0,)}] * 100, model, measured=[0], shots=1024, seed=1)
    assert r["probabilities"][1] > 0.01 and r["probabilities"][0] > 0.5
    say("x on three qubits")
    r = mps.run_noisy(eng, 3, [{"name": "X", "qubits": (q,)} for q in range(3)], model, measured=[0, 1, 2],
                      shots=1024, seed=2)
    p111 = r["counts"].get("111", 0) / 1024
    assert 0.8 < p111 < 0.99
    say("bell")
    r = mps.run_noisy(eng, 2, [{"name": "X", "qubits": (0,)}, {"name": "CNOT", "qubits": (0, 1)}], model,
                      measured=[0, 1], shots=1024, seed=3)
    p11 = r["counts"].get("11", 0) / 1024
    assert 0.8 < p11 < 0.99
    rho = r["density_matrix"]  # n <= 15: the full matrix, whose diagonal is the distribution
    assert rho.shape == (4, 4) and np.abs(np.real(np.diag(rho)) - r["probabilities"]).max() < 1e-12
    say("ghz 100")
    gates = [{"name": "H", "qubits": (0,)}] + [{"name": "CNOT", "qubits": (i, i + 1)} for i in range(99)]
    r = mps.run_noisy(eng, 100, gates, None, measured=[0, 6, 8, 36, 72, 98], shots=1024, seed=4)
    pr = r["probabilities"]
    assert abs(pr[0] - 0.5) < 1e-10 and abs(pr[-1] - 0.5) < 1e-10
    assert abs(r["counts"].get("111111", 0) / 1024 - 0.5) < 0.1
    # n > 15: the measured qubits' matrix; of a proper subset of GHZ qubits it has no coherence
    # (the traced qubits tell the two branches apart): diag(1/2, 0, ..., 0, 1/2)
    rho = r["density_matrix"]
    assert r["density_matrix_qubits"] == [0, 6, 8, 36, 72, 98] and rho.shape == (64, 64)
    assert abs(rho[0, 0] - 0.5) < 1e-10 and abs(rho[63, 63] - 0.5) < 1e-10
    assert abs(np.abs(rho).sum() - 1.0) < 1e-9


def test_purified_state_limits(eng):
    say("purified limits")
    s = mps.MpsState(eng, 3, cutoff=1e-9)
    s.run([{"name": "H", "qubits": (0,)}, {"name": "CNOT", "qubits": (0, 1)}])
    s.apply_kraus(1, N.amplitude_damping(0.2))
    with pytest.raises(capi.HipMpsError) as ei:
        s.amplitudes(np.zeros((1, 3), dtype=np.int8))
    assert ei.value.code == capi.HIPMPS_E_UNSUPPORTED
    c = s.clone()
    assert list(c.kraus_dims()) == list(s.kraus_dims())
    assert np.abs(c.expval_z_all() - s.expval_z_all()).max() < 1e-14
    c.close()
    s.close()


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-8), ("mixed", 1e-5)])
def test_reduced_density_matrix(eng, precision, tol):
    say("reduced density matrix", precision)
    n = 5
    ops = noisy_ops(n, 4, 53, RANK2)
    dm = run_oracle(DensityMatrix(n), ops)
    s = run_device(eng, n, ops, precision)
    for qs in ([2], [0, 4], [3, 1, 0]):
        assert np.abs(s.reduced_density_matrix(qs) - dm.reduced(qs)).max() < tol
    s.close()
    # a pure state (no channel): the same contraction with every Kraus leg of extent 1
    gates = C.u3_cnot_line(6, 4, seed=3)
    p = mps.MpsState(eng, 6, precision="fp64").run(gates)
    ref = DensityMatrix(6)
    for g in gates:
        ref.gate(mps.matrix_of(g), g["qubits"])
    assert np.abs(p.reduced_density_matrix([5, 2, 0]) - ref.reduced([5, 2, 0])).max() < 1e-10
    p.close()


def _sv_reduced(psi, n, qs):
    """rho_red[x, x'] of a dense state (qubit q = bit q), x = sum_i b_i 2^i over qs."""
    t = psi.reshape((2,) * n)
    keep = [n - 1 - q for q in qs]
    rest = [a for a in range(n) if a not in keep]
    m = np.transpose(t, list(reversed(keep)) + rest).reshape(1 << len(qs), -1)
    return m @ m.conj().T


@pytest.mark.parametrize("precision,tol", [("fp64", 1e-8), ("mixed", 1e-5)])
def test_full_density_matrix(eng, precision, tol):
    """k = n: the full density matrix that the reference stores for n <= 15
    (ExaTnPmpsVisitor.cpp:664-683), in qubit order and permuted."""
    say("full density matrix", precision)
    n = 6
    ops = noisy_ops(n, 4, 61, RANK2)
    dm = run_oracle(DensityMatrix(n), ops)
    s = run_device(eng, n, ops, precision)
    for qs in (list(range(n)), [4, 1, 5, 0, 3, 2]):
        assert np.abs(s.reduced_density_matrix(qs) - dm.reduced(qs)).max() < tol
    s.close()


def test_density_matrix_many_qubits(eng):
    """Up to the reference's 15-qubit limit: 10 of 12 qubits of a pure state against the state
    vector; 12 of a 20-qubit GHZ state (4^12 labels), a proper subset whose matrix is
    diag(1/2, 0, ..., 0, 1/2); all 13 qubits of a 13-qubit GHZ state, which keeps the coherences."""
    say("density matrix 10 of 12")
    n = 12
    gates = C.u3_cnot_line(n, 5, seed=71)
    s = mps.MpsState(eng, n, precision="fp64").run(gates)
    sv = O.StateVector(n).run(gates)
    qs = [11, 0, 5, 3, 7, 1, 9, 2, 10, 6]
    assert np.abs(s.reduced_density_matrix(qs) - _sv_reduced(sv.psi, n, qs)).max() < 1e-10
    s.close()
    say("density matrix 12 of GHZ 20")
    n = 20
    g = mps.MpsState(eng, n, precision="fp64")
    g.run([{"name": "H", "qubits": (0,)}] + [{"name": "CNOT", "qubits": (i, i + 1)} for i in range(n - 1)])
    qs = [0, 2, 3, 5, 8, 9, 11, 13, 14, 16, 18, 19]
    rho = g.reduced_density_matrix(qs)
    d = 1 << len(qs)
    assert rho.shape == (d, d)
    assert abs(rho[0, 0] - 0.5) < 1e-10 and abs(rho[d - 1, d - 1] - 0.5) < 1e-10
    assert abs(np.abs(rho).sum() - 1.0) < 1e-9
    g.close()
    say("density matrix of GHZ 13")
    n = 13
    g = mps.MpsState(eng, n, precision="fp64")
    g.run([{"name": "H", "qubits": (0,)}] + [{"name": "CNOT", "qubits": (i, i + 1)} for i in range(n - 1)])
    rho = g.reduced_density_matrix(list(range(n)))
    d = 1 << n
    for a, b in ((0, 0), (0, d - 1), (d - 1, 0), (d - 1, d - 1)):
        assert abs(rho[a, b] - 0.5) < 1e-10
    assert abs(np.abs(rho).sum() - 2.0) < 1e-9
    g.close()
    # beyond 15 qubits: refused before any work (a one-element buffer suffices to show it)
    big = mps.MpsState(eng, 16)
    q16 = np.arange(16, dtype=np.int32)
    buf = np.zeros(2)
    rc = capi.lib().hipmps_reduced_density_matrix(big.h, q16.ctypes.data_as(capi.ctypes.POINTER(capi.ctypes.c_int)),
                                                  16, buf.ctypes.data_as(capi.ctypes.POINTER(capi.ctypes.c_double)), 1)
    assert rc == capi.HIPMPS_E_UNSUPPORTED
    big.close()


@pytest.mark.parametrize("block", [1 << 24, 64, 1])
def test_density_matrix_wide_kraus(eng, block):
    """Amplitude damping after every gate: wide Kraus legs.  With a small rdm_block_elems the
    sweep holds its per-site intermediate a block of (p, kappa) at a time and accumulates the
    bra side over the blocks; every block size gives the dense density matrix (rank-2 channels:
    the 1e-9 cut is the only approximation)."""
    say("density matrix, wide Kraus legs, block", block)
    n = 8
    ops = noisy_ops(n, 4, 81, [N.amplitude_damping(0.05)])
    dm = run_oracle(DensityMatrix(n), ops)
    s = run_device(eng, n, ops, "fp64")
    assert max(s.kraus_dims()) >= 16
    base = eng.tuning()
    try:
        eng.set_tuning(rdm_block_elems=block)
        assert np.abs(s.reduced_density_matrix(list(range(n))) - dm.reduced(list(range(n)))).max() < 1e-8
        qs = [7, 2, 5]
        assert np.abs(s.reduced_density_matrix(qs) - dm.reduced(qs)).max() < 1e-8
    finally:
        eng.set_tuning(**base)
    s.close()


def test_channels_queued_equal_one_at_a_time(eng):
    """Channels queue after the open wave of gates and run together; flushing after every
    channel instead (kraus_dims) runs them one at a time.  Both orders apply the same
    operations to the same tensors (queued channels only pass gates on other sites), so the
    results agree to rounding, pending 1-qubit gates folded into the operators included."""
    say("channels queued vs one at a time")
    n = 6
    ops = []
    for g in C.u3_cnot_line(n, 4, seed=91):
        ops.append(("g", g))
        for q in g["qubits"]:
            ops.append(("c", q, [N.amplitude_damping(0.07), N.depolarizing(0.1), N.phase_damping(0.2)][q % 3]))
    a = run_device(eng, n, ops, "fp64")
    b = mps.MpsState(eng, n, cutoff=1e-9)
    for op in ops:
        if op[0] == "g":
            b.run([op[1]])
        else:
            b.apply_kraus(op[1], op[2])
            b.kraus_dims()
    assert list(a.kraus_dims()) == list(b.kraus_dims())
    assert np.abs(a.expval_z_all() - b.expval_z_all()).max() < 1e-12
    assert abs(a.norm() - b.norm()) < 1e-12
    a.close()
    b.close()


def test_reset_drops_queued_channels(eng):
    """Channels still queued when a state is reset are dropped with it: |0..0>, Kraus legs 1."""
    say("reset with queued channels")
    s = mps.MpsState(eng, 4, cutoff=1e-9)
    s.run([{"name": "H", "qubits": (0,)}, {"name": "CNOT", "qubits": (0, 1)}])
    s.apply_kraus(0, N.amplitude_damping(0.3))
    s.apply_kraus(1, N.depolarizing(0.2))
    s.reset()
    assert list(s.kraus_dims()) == [1, 1, 1, 1]
    assert np.abs(s.expval_z_all() - 1.0).max() < 1e-14
    s.run([{"name": "X", "qubits": (2,)}])
    s.apply_kraus(2, N.amplitude_damping(0.25))
    assert abs(s.expval_z_all()[2] - (-1 + 2 * 0.25)) < 1e-12
    s.close()


class _EveryGate:
    """A channel list per gate: the given channels on each qubit the gate touches."""

    def __init__(self, chans):
        self.chans = chans

    def channels(self, g):
        return [(q, self.chans[q % len(self.chans)]) for q in g["qubits"]]


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_noisy_batch_equals_alone(eng, precision):
    """Noisy circuits of different sizes advanced together (hipmps_queue_gates / queue_kraus /
    run_waves) give what each gives alone: the joint waves only share launches."""
    say("noisy batch vs alone", precision)
    noise = _EveryGate([N.amplitude_damping(0.07), N.depolarizing(0.1), N.phase_damping(0.2)])
    circuits = [(n, C.u3_cnot_line(n, d, seed=100 + n + d)) for n, d in ((5, 3), (8, 4), (3, 6), (12, 2), (6, 5))]
    got = mps.run_noisy_batch(eng, circuits, noise, precision=precision)
    for (n, gates), r in zip(circuits, got):
        s = mps.MpsState(eng, n, cutoff=1e-9, precision=precision)
        for g in gates:
            s.run([g])
            for q, k in noise.channels(g):
                s.apply_kraus(q, k)
        assert list(r["kraus_dims"]) == list(s.kraus_dims())
        assert list(r["bonds"]) == list(s.bonds())
        assert np.abs(r["expval_z"] - s.expval_z_all()).max() < 1e-12
        s.close()
    # against the restated reference algorithm (same 2 kap bound) for one of them
    n, gates = circuits[2]
    ref = PurifiedMps(n)
    for g in gates:
        ref.gate(mps.matrix_of(g), g["qubits"])
        for q, k in noise.channels(g):
            ref.channel(q, k)
    if precision == "fp64":
        assert list(got[2]["kraus_dims"]) == ref.kraus_dims()
    assert np.abs(got[2]["expval_z"] - ref.expval_z()).max() < (1e-8 if precision == "fp64" else 1e-5)
    # argument checks: a state listed twice, states of two engines
    a = mps.MpsState(eng, 2)
    with pytest.raises(capi.HipMpsError):
        mps.run_waves([a, a])
    other = mps.Engine()
    b = mps.MpsState(other, 2)
    with pytest.raises(capi.HipMpsError):
        mps.run_waves([a, b])
    b.close()
    other.close()
    a.close()
