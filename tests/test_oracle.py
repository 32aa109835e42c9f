"""Pins the CPU FP64 oracle (oracle/mps_oracle.cpp) before it is trusted as the checker.

Each check is chosen so that a plausible slip (transposed gate, wrong operand order,
wrong bit convention, dropped term, lost singular value) fails it:
  * gate matrices: unitarity + identities between gates (U3 = Rx, CNOT = H CZ H, ...)
  * dense state vector: closed-form results the tnqvm tests assert
  * Jacobi SVD: LAPACK zgesdd spectra (tests/golden/svd_spectra.npz)
  * MPS restatement: dense state vector on the same circuits, both operand orders
  * truncation: special cases (cutoff 0, max-bond-dim k = Eckart-Young rank k,
    keep-the-first-sub-threshold-index rule)
  * sampling: determinism under a seed and the outcome support
"""
import json
import os

import numpy as np
import pytest

from mt64 import Mt64 as _Mt64
from oracle import oracle as O
from tnqvm_amd import circuits as C

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def g(name, *qubits, params=()):
    return {"name": name, "qubits": tuple(qubits), "params": tuple(params)}


# ---------------------------------------------------------------- gate matrices
@pytest.mark.parametrize("name,params", [
    ("H", ()), ("X", ()), ("Y", ()), ("Z", ()), ("T", ()), ("Tdg", ()), ("S", ()), ("Sdg", ()),
    ("Rx", (0.3,)), ("Ry", (1.1,)), ("Rz", (-2.0,)), ("U", (0.4, 1.3, -0.7)),
    ("CNOT", ()), ("CZ", ()), ("CY", ()), ("CH", ()), ("CRZ", (0.9,)), ("CPhase", (0.5,)),
    ("Swap", ()), ("iSwap", ()), ("fSim", (0.7, 0.2))])
def test_gates_unitary(name, params):
    m = O.gate_matrix(name, params)
    assert np.allclose(m.conj().T @ m, np.eye(m.shape[0]), atol=1e-15)


def test_gate_identities():
    rng = np.random.default_rng(3)
    for th in rng.uniform(-np.pi, np.pi, 8):
        # U3(theta, -pi/2, pi/2) = Rx(theta) (ITensorMPSVisitorTester.cpp:366-392)
        assert np.allclose(O.gate_matrix("U", (th, -np.pi / 2, np.pi / 2)), O.gate_matrix("Rx", (th,)),
                           atol=1e-15)
        # Rz(theta) = e^{-i theta/2} diag(1, e^{i theta})
        assert np.allclose(O.gate_matrix("Rz", (th,)),
                           np.exp(-0.5j * th) * np.diag([1, np.exp(1j * th)]), atol=1e-15)
    H = O.gate_matrix("H")
    I2 = np.eye(2)
    # CNOT (control = first operand = MSB) = (I x H) CZ (I x H)
    assert np.allclose(np.kron(I2, H) @ O.gate_matrix("CZ") @ np.kron(I2, H), O.gate_matrix("CNOT"),
                       atol=1e-15)
    assert np.allclose(O.gate_matrix("T") @ O.gate_matrix("T"), O.gate_matrix("S"), atol=1e-15)
    # fSim(pi/2, 0) = iSwap^dagger on the middle block
    assert np.allclose(O.gate_matrix("fSim", (np.pi / 2, 0.0)), O.gate_matrix("iSwap").conj(), atol=1e-15)


# ---------------------------------------------------------------- dense state vector
def test_sv_bit_convention():
    # X on q0 of 3 qubits -> index 1 (qubit k <-> bit k)
    sv = O.StateVector(3).run([g("X", 0)])
    assert abs(sv.psi[1] - 1) < 1e-15
    # CNOT(q1, q0) with q1=1 flips q0: |q1 q0> = |10> -> |11> = index 3
    sv = O.StateVector(2).run([g("X", 1), g("CNOT", 1, 0)])
    assert abs(sv.psi[3] - 1) < 1e-15
    sv = O.StateVector(2).run([g("X", 0), g("CNOT", 1, 0)])  # control 0 -> no flip
    assert abs(sv.psi[1] - 1) < 1e-15


def test_sv_kats():
    # Bell pair, both orders (MpsGateTester.cpp:90-126)
    for ops in ([g("H", 0), g("CNOT", 0, 1)], [g("H", 1), g("CNOT", 1, 0)]):
        sv = O.StateVector(2).run(ops)
        assert np.allclose(np.abs(sv.psi) ** 2, [0.5, 0, 0, 0.5], atol=1e-15)
    # iSwap: X(q0); iSwap(q0,q1) -> "010" with prob 1 (MpsGateTester.cpp:312-330)
    sv = O.StateVector(3).run([g("X", 0), g("iSwap", 0, 1)])
    assert abs(abs(sv.psi[2]) ** 2 - 1) < 1e-15
    # fSim: X(q0); fSim(q0,q1,theta,0): P(q0=0,q1=1) = sin^2 theta (MpsGateTester.cpp:332-357)
    for th in np.linspace(-np.pi, np.pi, 10):
        sv = O.StateVector(2).run([g("X", 0), g("fSim", 0, 1, params=(th, 0.0))])
        p = np.abs(sv.psi) ** 2
        assert abs(p[2] - np.sin(th) ** 2) < 1e-14 and abs(p[1] - np.cos(th) ** 2) < 1e-14
    # Rx / CPhase rotations: <Z> = cos(theta) (ITensorMPSVisitorTester.cpp:310-364)
    for th in np.linspace(-np.pi, np.pi, 20):
        assert abs(O.StateVector(1).run([g("Rx", 0, params=(th,))]).expval_z_parity([0]) - np.cos(th)) < 1e-14
        sv = O.StateVector(2).run([g("X", 0), g("H", 1), g("CPhase", 0, 1, params=(th,)), g("H", 1)])
        assert abs(sv.expval_z_parity([1]) - np.cos(th)) < 1e-14


def test_sv_deuteron_table():
    kat = json.load(open(os.path.join(GOLD, "kat.json")))
    for t, ref in zip(kat["deuteron_angles"], kat["deuteron_x0x1"]):
        sv = O.StateVector(2).run([g("X", 0), g("Ry", 1, params=(t,)), g("CNOT", 1, 0), g("H", 0), g("H", 1)])
        # the table carries 6 significant digits
        assert abs(sv.expval_z_parity([0, 1]) - ref) < 1e-6


def test_sv_sampling_determinism():
    sv = O.StateVector(4).run([g("H", 0), g("CNOT", 0, 1), g("CNOT", 1, 2), g("CNOT", 2, 3)])
    a = sv.sample(123, 8192, [0, 1, 2, 3])
    b = sv.sample(123, 8192, [0, 1, 2, 3])
    assert a == b and len(a) == 8192
    assert set(a) == {"0000", "1111"}
    frac = a.count("0000") / len(a)
    assert abs(frac - 0.5) < 0.03
    # char i <-> i-th measured qubit (GateMatrixAlgebra.hpp:129-136)
    sv = O.StateVector(3).run([g("X", 2)])
    assert set(sv.sample(1, 10, [2, 0])) == {"10"}


def test_sv_sampling_inverse_cdf():
    """GenerateSamples (GateMatrixAlgebra.hpp:125-156) is inverse-CDF sampling with sorted draws:
    shot j is the first basis index x whose cumulative probability exceeds the j-th smallest of
    the shots' mt19937_64 uniforms.  Pinned on a product state Ry(t_q)|0>, whose probabilities
    are products of cos^2 and sin^2 of t_q / 2, with the draws from an independent MT19937-64."""
    n, shots, seed = 4, 300, 2024
    th = [0.4, 1.3, 2.2, 2.8]
    sv = O.StateVector(n).run([{"name": "Ry", "qubits": (q,), "params": (t,)} for q, t in enumerate(th)])
    meas = [2, 0, 3]
    got = sv.sample(seed, shots, meas)
    p = np.array([np.prod([np.sin(th[q] / 2) ** 2 if (x >> q) & 1 else np.cos(th[q] / 2) ** 2 for q in range(n)])
                  for x in range(1 << n)])
    cdf = np.cumsum(p)
    gen = _Mt64(seed)
    draws = sorted(gen.uniform() for _ in range(shots))
    want = []
    for r in draws:
        x = int(np.searchsorted(cdf, r, side="right"))  # first x with cdf[x] > r
        if x < (1 << n):
            want.append("".join("1" if (x >> q) & 1 else "0" for q in meas))
    assert got == want


# ---------------------------------------------------------------- SVD vs LAPACK
def test_svd_golden_spectra():
    z = np.load(os.path.join(GOLD, "svd_spectra.npz"))
    keys = [k[2:] for k in z.files if k.startswith("A_")]
    assert len(keys) >= 10
    for k in keys:
        a, sref = z["A_" + k], z["S_" + k]
        s, L, R = O.svd_split(a)
        assert np.allclose(s, sref, atol=1e-13 * sref[0], rtol=0), k
        assert np.abs(L @ R - a).max() < 1e-13 * sref[0], k
        # sqrt(S) split: |L[:,k]|^2 = |R[k,:]|^2 = s_k
        assert np.allclose((np.abs(L) ** 2).sum(0), s, atol=1e-13 * sref[0])
        assert np.allclose((np.abs(R) ** 2).sum(1), s, atol=1e-13 * sref[0])


# ---------------------------------------------------------------- MPS vs dense
def random_adjacent_circuit(n, n_gates, seed):
    rng = np.random.default_rng(seed)
    gates = []
    for _ in range(n_gates):
        if n > 1 and rng.random() < 0.5:
            q = int(rng.integers(0, n - 1))
            pair = (q, q + 1) if rng.random() < 0.5 else (q + 1, q)
            gates.append({"name": "unitary", "qubits": pair, "matrix": C.haar_unitary(rng, 4)})
        else:
            gates.append({"name": "unitary", "qubits": (int(rng.integers(0, n)),),
                          "matrix": C.haar_unitary(rng, 2)})
    return gates


@pytest.mark.parametrize("n", [1, 2, 3, 5, 8])
def test_mps_matches_dense(n):
    gates = random_adjacent_circuit(n, 60, seed=n)
    sv = O.StateVector(n).run(gates)
    m = O.Mps(n).run(gates)
    assert np.abs(m.state_vector() - sv.psi).max() < 1e-12
    assert abs(m.norm() - 1.0) < 1e-12
    for q in range(n):
        assert abs(m.expval_z_parity([q]) - sv.expval_z_parity([q])) < 1e-12
    if n >= 3:
        assert abs(m.expval_z_parity([0, n - 1]) - sv.expval_z_parity([0, n - 1])) < 1e-12


def test_mps_u3_cnot_line_full():
    gates = C.u3_cnot_line()
    sv = O.StateVector(12).run(gates)
    m = O.Mps(12).run(gates)
    assert list(m.bonds()) == [1, 2, 4, 8, 16, 32, 64, 32, 16, 8, 4, 2, 1]
    for q in range(12):
        assert abs(m.expval_z_parity([q]) - sv.expval_z_parity([q])) < 1e-10
    bits = C.random_bitstrings(12, 16, 5)
    assert np.abs(m.amplitudes(bits) - sv.amplitudes(bits)).max() < 1e-10


def test_mps_slice():
    gates = random_adjacent_circuit(6, 40, seed=11)
    sv = O.StateVector(6).run(gates)
    m = O.Mps(6).run(gates)
    bits = [1, -1, 0, -1, -1, 1]
    sl = m.slice(bits)
    # open legs in increasing qubit order, first open fastest
    ref = []
    for x in range(8):
        b = [1, x & 1, 0, (x >> 1) & 1, (x >> 2) & 1, 1]
        ref.append(sv.psi[sum(bb << i for i, bb in enumerate(b))])
    assert np.abs(sl - np.array(ref)).max() < 1e-13


# ---------------------------------------------------------------- truncation rule
def test_truncation_keep_first_subthreshold():
    # CNOT on |00>: theta' has singular values [1, 0] exactly.  With the default
    # cutoff (DBL_MIN) the rule returns i+1 for the first zero -> bond 2 (:2434-2443).
    m = O.Mps(3).run([g("CNOT", 0, 1)])
    assert list(m.bonds()) == [1, 2, 1, 1]
    # a cutoff above 0 but below 1 gives the same bond (keep index of the first zero)
    m = O.Mps(3, cutoff=0.5).run([g("CNOT", 0, 1)])
    assert list(m.bonds()) == [1, 2, 1, 1]
    # cutoff above 1 cuts at the first index (keeps 1)
    m = O.Mps(3, cutoff=2.0).run([g("CNOT", 0, 1)])
    assert list(m.bonds()) == [1, 1, 1, 1]


def test_truncation_max_bond_is_eckart_young():
    # 4 qubits; build a state with a full-rank middle bond, then a final gate on (1,2)
    # with max-bond-dim 2: the result must equal the rank-2 SVD truncation of theta'.
    rng = np.random.default_rng(7)
    pre = random_adjacent_circuit(4, 30, seed=3)
    last = {"name": "unitary", "qubits": (1, 2), "matrix": C.haar_unitary(rng, 4)}
    m_full = O.Mps(4).run(pre)
    m_cut = O.Mps(4, max_bond=4).run(pre)     # no cut needed before the last gate
    assert np.abs(m_full.state_vector() - m_cut.state_vector()).max() < 1e-12
    # left-canonicalise is not available in the restatement, so compare in the dense
    # picture: psi' = (I x G x I) psi, and the cut state has Schmidt rank <= 2 across (1|2)
    m_cut2 = O.Mps(4, max_bond=2).run(pre + [last])
    assert m_cut2.bonds()[2] == 2
    psi = m_cut2.state_vector().reshape(2, 2, 2, 2)  # [q3][q2][q1][q0]
    mat = psi.reshape(4, 4)                          # rows (q3,q2), cols (q1,q0)
    sv = np.linalg.svd(mat, compute_uv=False)
    assert sv[2] < 1e-13 and sv[3] < 1e-13
    # cutoff 0 never truncates
    m0 = O.Mps(4, cutoff=0.0).run(pre + [last])
    ref = O.StateVector(4).run(pre + [last])
    assert np.abs(m0.state_vector() - ref.psi).max() < 1e-12


def test_truncation_discards_smallest():
    # single 2q gate on a 2-qubit register with max-bond 1: the kept product state is
    # the top Schmidt component, norm = s_0^2 (no renormalisation).
    rng = np.random.default_rng(9)
    G = C.haar_unitary(rng, 4)
    gate = {"name": "unitary", "qubits": (0, 1), "matrix": G}
    ref = O.StateVector(2).run([gate]).psi.reshape(2, 2)  # [q1][q0]
    s = np.linalg.svd(ref, compute_uv=False)
    m = O.Mps(2, max_bond=1).run([gate])
    assert abs(m.norm() - s[0] ** 2) < 1e-13
    assert abs(m.discarded() - s[1] ** 2) < 1e-13


def _rand_adjacent(n, n_gates, seed):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n_gates):
        if rng.random() < 0.5:
            q = int(rng.integers(0, n - 1))
            pair = (q, q + 1) if rng.random() < 0.5 else (q + 1, q)
            out.append({"name": "unitary", "qubits": pair, "matrix": C.haar_unitary(rng, 4)})
        else:
            out.append({"name": "unitary", "qubits": (int(rng.integers(0, n)),), "matrix": C.haar_unitary(rng, 2)})
    return out


@pytest.mark.parametrize("cutoff,max_bond,pn_l2", [(0.0, 0, False), (1e-3, 0, False), (1e-8, 4, False),
                                                   (1e-2, 0, True)])
def test_lapack_mps_matches_oracle(cutoff, max_bond, pn_l2):
    """The zgesdd restatement (the LAPACK-class CPU baseline) against the Jacobi oracle on the same
    truncated circuits: bonds, state and discarded weight; without truncation, also the dense vector."""
    from oracle.lapack_mps import LapackMps
    for seed in range(3):
        n = 7
        gates = _rand_adjacent(n, 50, seed=40 + seed)
        ref = O.Mps(n, cutoff=cutoff, max_bond=max_bond, pn_l2=pn_l2).run(gates)
        lap = LapackMps(n, cutoff=cutoff, max_bond=max_bond, pn_l2=pn_l2).run(gates, O.gate_of)
        assert lap.bonds() == list(ref.bonds())
        assert np.abs(lap.state_vector() - ref.state_vector()).max() < 1e-10
        assert abs(lap.discarded - ref.discarded()) < 1e-10
        if cutoff == 0.0 and max_bond == 0:
            assert np.abs(lap.state_vector() - O.StateVector(n).run(gates).psi).max() < 1e-12


def test_mt64_standard_value():
    # [rand.predef]: the 10000th consecutive output of a default-constructed mt19937_64
    g = _Mt64()
    for _ in range(9999):
        g.next()
    assert g.next() == 9981545732273789042


def test_measure_samples_product_state_closed_form():
    """getMeasureSample on a product state Ry(t_q)|0> whose first site is scaled by c (norm c^2):
    the first measured qubit is drawn against its unnormalised probability c^2 cos^2(t/2), every
    later one against its conditional cos^2(t/2); one mt19937_64 draw per qubit per shot."""
    n, shots, seed, c = 5, 200, 99, 0.8
    th = [0.3, 1.1, 2.0, 2.9, 0.7]
    m = O.Mps(n).run([{"name": "Ry", "qubits": (q,), "params": (t,)} for q, t in enumerate(th)])
    a0 = m.site(0) * c
    m.set_site(0, a0)
    meas = [3, 0, 4, 1]
    got = m.measure_samples(seed, shots, meas)
    g = _Mt64(seed)
    p0 = {q: np.cos(th[q] / 2) ** 2 for q in range(n)}
    want = []
    for _ in range(shots):
        s = ""
        for i, q in enumerate(meas):
            p = p0[q] * (c * c if i == 0 else 1.0)  # the first measured qubit sees the norm c^2
            s += "0" if g.uniform() <= p else "1"
        want.append(s)
    assert got == want


def test_measure_samples_ghz_and_dense_statistics():
    """GHZ: the first draw decides every bit; a random 8-qubit state: the counts of 4000 shots agree
    with the dense distribution of the measured qubits (chi-square below a loose bound)."""
    n = 6
    m = O.Mps(n).run([{"name": "H", "qubits": (0,)}] + [{"name": "CNOT", "qubits": (i, i + 1)} for i in range(n - 1)])
    got = m.measure_samples(5, 64, [2, 0, 5])
    g = _Mt64(5)
    want = []
    for _ in range(64):
        r = [g.uniform() for _ in range(3)]
        want.append("000" if r[0] <= 0.5 else "111")
    assert got == want
    gates = _rand_adjacent(8, 60, seed=77)
    sv = O.StateVector(8).run(gates)
    mm = O.Mps(8).run(gates)
    meas = [1, 4, 6]
    got = mm.measure_samples(11, 4000, meas)
    p = np.abs(sv.psi) ** 2
    probs = {}
    for x in range(256):
        key = "".join("1" if (x >> q) & 1 else "0" for q in meas)
        probs[key] = probs.get(key, 0.0) + p[x]
    chi2 = sum((got.count(k) - 4000 * v) ** 2 / (4000 * v) for k, v in probs.items() if v > 1e-9)
    assert chi2 < 40.0  # 7 degrees of freedom
