"""hipmps_amplitudes_noalloc: amplitudes with every intermediate in the library's static device
scratch (no allocation), against hipmps_amplitudes on the same state -- equal bit for bit, on
two calls in a row, the second with fewer bit strings -- and against the dense state vector.
Requests beyond the scratch and states with queued gates are refused on the host."""
import sys

import numpy as np
import pytest

from oracle import oracle as O
from tnqvm_amd import _capi
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


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_noalloc_equals_amplitudes_line(eng, precision):
    n = 12
    gates = C.u3_cnot_line(n=n, depth=12, seed=0x7E5A0201)
    ref = O.StateVector(n).run(gates)
    s = mps.MpsState(eng, n, precision=precision).run(gates)
    s.flush()
    tol = 1e-10 if precision == "fp64" else 1e-5
    for k, seed in ((300, 11), (5, 12)):
        say("line", precision, "k", k)
        bits = C.random_bitstrings(n, k, seed)
        want = s.amplitudes(bits)
        got = s.amplitudes_noalloc(bits)
        assert np.array_equal(got, want)
        assert np.abs(got - ref.amplitudes(bits)).max() < tol
    s.close()


def test_noalloc_two_states_in_a_row(eng):
    """A truncated 17-qubit brickwork state (odd n, bonds up to 64) then a 6-qubit one: the second
    call's smaller chain reads none of what the first left in the scratch."""
    a = mps.MpsState(eng, 17, cutoff=1e-10, max_bond=64, precision="mixed").run(
        C.haar_brickwork(17, 14, seed=0x7E5A0202))
    a.flush()
    assert a.bonds().max() == 64
    b = mps.MpsState(eng, 6).run(C.haar_brickwork(6, 5, seed=0x7E5A0203))
    b.flush()
    for st, k, seed in ((a, 1024, 21), (b, 3, 22), (a, 40, 23)):
        say("brickwork n", st.n, "k", k)
        bits = C.random_bitstrings(st.n, k, seed)
        assert np.array_equal(st.amplitudes_noalloc(bits), st.amplitudes(bits))
    a.close()
    b.close()


def test_noalloc_refusals(eng):
    n = 20
    s = mps.MpsState(eng, n, cutoff=1e-10, max_bond=256).run(C.haar_brickwork(n, 16, seed=0x7E5A0204))
    s.flush()
    chi = int(s.bonds().max())
    # largest k that fits, then one more string than the scratch holds
    def need(k):
        up = lambda x: (x + 255) // 256 * 256
        return up(n * 72) + up(k * n) + 2 * up(k * chi * 16)
    k_max = max(k for k in range(1, 1 << 16) if need(k) <= _capi.HIPMPS_NOALLOC_SCRATCH_BYTES) if chi > 1 else 0
    assert k_max > 0
    bits = C.random_bitstrings(n, k_max + 1, 31)
    with pytest.raises(_capi.HipMpsError) as ei:
        s.amplitudes_noalloc(bits)
    assert ei.value.code == _capi.HIPMPS_E_NOMEM
    say("refusals: largest k", k_max, "at bond", chi)
    ok = s.amplitudes_noalloc(bits[:k_max])
    assert np.array_equal(ok, s.amplitudes(bits[:k_max]))
    # a queued gate: refused (a flush could allocate)
    s.run([{"name": "H", "qubits": (3,)}])
    with pytest.raises(_capi.HipMpsError) as ei:
        s.amplitudes_noalloc(bits[:4])
    assert ei.value.code == _capi.HIPMPS_E_ARG
    s.flush()
    assert np.array_equal(s.amplitudes_noalloc(bits[:4]), s.amplitudes(bits[:4]))
    s.close()
