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
"""
import os
import sys

import numpy as np
import pytest

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
    say("ghz 100")
    gates = [{"name": "H", "qubits": (0,)}] + [{"name": "CNOT", "qubits": (i, i + 1)} for i in range(99)]
    r = mps.run_noisy(eng, 100, gates, None, measured=[0, 6, 8, 36, 72, 98], shots=1024, seed=4)
    pr = r["probabilities"]
    assert abs(pr[0] - 0.5) < 1e-10 and abs(pr[-1] - 0.5) < 1e-10
    assert abs(r["counts"].get("111111", 0) / 1024 - 0.5) < 0.1


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
