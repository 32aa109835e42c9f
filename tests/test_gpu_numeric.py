"""Numerical guards of the 2-qubit path on the GPU.

* The post-SVD norm validation of the reference (ExaTnMpsVisitor.cpp:1634-1661): after each
  split, an error is reported when the Frobenius norm of either sqrt(S) factor is below
  1e-3.  Here it is counted in hipmps_stats_t.norm_warnings; the expected count comes from
  the factor norms of the oracle's own sqrt(S) split of the same theta.
* NaN / Inf: gate matrices and site tensors holding them are refused on the host
  (HIPMPS_E_NUMERIC) before anything reaches the device; an overflow on the device
  (finite inputs whose Gram matrix exceeds the FP64 range) ends in HIPMPS_E_NUMERIC.
"""
import sys

import numpy as np
import pytest

from oracle import oracle as O
from tnqvm_amd import _capi as capi
from tnqvm_amd import circuits as C
from tnqvm_amd import mps

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def eng():
    e = mps.Engine()
    yield e
    e.close()


def theta_of(a, b, g):
    """theta[(l, i), (j, r)] = sum G[(i, j), (i', j')] a[l, i', m] b[m, j', r]; the gate's first
    operand (the left site here) is the most significant bit of its row index."""
    ab = np.einsum("lim,mjr->lijr", a, b)
    t = np.einsum("ijkl,aklr->aijr", g.reshape(2, 2, 2, 2), ab)
    return t.reshape(a.shape[0] * 2, 2 * b.shape[2])


def expected_warnings(n, planted, gates):
    """Replays the gates on the oracle MPS; before each 2-qubit gate splits the same theta
    with the oracle's sqrt(S) SVD and counts the splits whose factor norms fall below 1e-3."""
    m = O.Mps(n)
    for k, a in planted.items():
        m.set_site(k, a)
    count = 0
    for g in gates:
        if len(g["qubits"]) == 2:
            k = min(g["qubits"])
            _, L, R = O.svd_split(theta_of(m.site(k), m.site(k + 1), mps.matrix_of(g)))
            if min(np.linalg.norm(L), np.linalg.norm(R)) < 1e-3:
                count += 1
        m.apply(g)
    return count


@pytest.mark.parametrize("scale", [1e-10, 1e-4])
def test_norm_validation_count(eng, scale):
    say("norm validation", scale)
    rng = np.random.default_rng(11)
    n = 5
    planted = {0: np.array([scale, 0.0], dtype=np.complex128).reshape(1, 2, 1)}
    gates = [{"name": "unitary", "qubits": (q, q + 1), "matrix": C.haar_unitary(rng, 4)}
             for q in (0, 1, 2, 3, 0, 1)]
    s = mps.MpsState(eng, n, precision="fp64")
    for k, a in planted.items():
        s.set_site(k, a)
    s.run(gates)
    got = s.stats()["norm_warnings"]
    want = expected_warnings(n, planted, gates)
    assert got == want
    if scale == 1e-10:
        assert want >= 1  # the planted site is tiny enough to trip the check
    s.close()


def test_nonfinite_gate_refused(eng):
    say("nonfinite gate")
    s = mps.MpsState(eng, 3, precision="fp64")
    bad = np.eye(4, dtype=np.complex128)
    bad[1, 2] = np.nan
    with pytest.raises(capi.HipMpsError) as ei:
        s.run([{"name": "unitary", "qubits": (0, 1), "matrix": bad}])
    assert ei.value.code == capi.HIPMPS_E_NUMERIC
    inf1 = np.array([[1, np.inf], [0, 1]], dtype=np.complex128)
    with pytest.raises(capi.HipMpsError) as ei:
        s.run([{"name": "unitary", "qubits": (2,), "matrix": inf1}])
    assert ei.value.code == capi.HIPMPS_E_NUMERIC
    # the state is untouched and usable
    s.run([{"name": "X", "qubits": (1,)}])
    assert np.allclose(s.expval_z_all(), [1.0, -1.0, 1.0], atol=1e-12)
    with pytest.raises(capi.HipMpsError) as ei:
        s.set_site(0, np.array([np.nan, 0.0], dtype=np.complex128).reshape(1, 2, 1))
    assert ei.value.code == capi.HIPMPS_E_NUMERIC
    s.close()
    srv = mps.Server(eng)
    with pytest.raises(capi.HipMpsError) as ei:
        srv.submit(3, [{"name": "unitary", "qubits": (0, 1), "matrix": bad}])
    assert ei.value.code == capi.HIPMPS_E_NUMERIC


def test_overflow_reported(eng):
    say("overflow")
    s = mps.MpsState(eng, 3, precision="fp64")
    s.set_site(0, np.array([1e300, 1e300], dtype=np.complex128).reshape(1, 2, 1))
    s.set_site(1, np.array([1e300, 1e300], dtype=np.complex128).reshape(1, 2, 1))
    rng = np.random.default_rng(3)
    s.run([{"name": "unitary", "qubits": (0, 1), "matrix": C.haar_unitary(rng, 4)}])
    with pytest.raises(capi.HipMpsError) as ei:
        s.flush()
    assert ei.value.code == capi.HIPMPS_E_NUMERIC
    s.close()


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_rank_deficient_theta_converges(eng, precision):
    """The 60-qubit QAOA swap network produces exactly rank-deficient theta' (here 50 x 56 of rank
    28) whose remaining rows are rounding residue (norms 1e-16 ... 1e-93).  A purely relative
    orthogonality test keeps rotating such rows against the big ones for ever; rows below
    1e-15 ||X||_F count as zero, and every SVD converges."""
    from tnqvm_amd import circuits as C
    gates, _, _, _ = C.qaoa_maxcut(60, 4, seed=0x7E5A0003)
    s = mps.MpsState(eng, 60, cutoff=1e-12, max_bond=512, precision=precision)
    for k in range(0, 660, 30):
        s.run(gates[k:k + 30])
        s.flush()
    st = s.stats()
    assert st["svds_nonconverged"] == 0
    assert st["max_bond"] == 512
    assert np.all(np.isfinite(s.expval_z_all()))
    s.close()
