"""GPU parity of the workload families on hip-mps, through the C ABI, against the oracle.

* QAOA MaxCut (swap network, fused ZZ+swap gates): energy sum_edges <Z_i Z_j> on the
  final qubit positions against the dense logical circuit, FP64 to 1e-10, mixed to 1e-5;
  with truncation against the oracle MPS on the same routed gate list.
* 2D grid fSim circuits, snake map + nearest-neighbour routing: <Z_i> and amplitudes
  against the dense state vector of the unrouted circuit.
* A mixed batch of uneven brickwork circuits with per-circuit chi_max, through
  hipmps_run_batch, against the oracle MPS run circuit by circuit.
"""
import sys

import numpy as np
import pytest

from oracle import oracle as O
from tnqvm_amd import circuits as C
from tnqvm_amd import mps

pytestmark = pytest.mark.gpu
TOL64 = 1e-10
TOLMIX = 1e-5


def say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def eng():
    e = mps.Engine()
    yield e
    e.close()


@pytest.mark.parametrize("precision,tol", [("fp64", TOL64), ("mixed", TOLMIX)])
def test_qaoa_energy_vs_dense(eng, precision, tol):
    n, p = 12, 2
    say("qaoa dense", n, p, precision)
    gates, edges, pos, (gam, bet) = C.qaoa_maxcut(n, p, seed=0x7E5A0003)
    ref = O.StateVector(n).run(C.qaoa_logical(n, edges, gam, bet))
    s = mps.MpsState(eng, n, cutoff=0.0, precision=precision).run(gates)
    pairs = [(pos[i], pos[j]) for i, j in edges]
    zz = s.expval_zz(pairs)
    for (i, j), v in zip(edges, zz):
        assert abs(v - ref.expval_z_parity([i, j])) < tol
    e_ref = sum(ref.expval_z_parity([i, j]) for i, j in edges)
    assert abs(zz.sum() - e_ref) < tol * len(edges)
    ez = s.expval_z_all()
    for q in range(n):
        assert abs(ez[pos[q]] - ref.expval_z_parity([q])) < tol


def test_qaoa_cutoff_vs_dense(eng):
    """Cutoff-only truncation (1e-12, no chi cap) keeps the state to 1e-10."""
    n, p = 14, 2
    say("qaoa cutoff", n, p)
    gates, edges, pos, (gam, bet) = C.qaoa_maxcut(n, p, seed=0x7E5A0013)
    ref = O.StateVector(n).run(C.qaoa_logical(n, edges, gam, bet))
    s = mps.MpsState(eng, n, cutoff=1e-12).run(gates)
    zz = s.expval_zz([(pos[i], pos[j]) for i, j in edges])
    for (i, j), v in zip(edges, zz):
        assert abs(v - ref.expval_z_parity([i, j])) < TOL64


def test_qaoa_chi_capped_invariants(eng):
    """With a chi cap the cut can fall inside an exactly degenerate Schmidt multiplet (the
    spectra of independent edges across a cut multiply), where the kept subspace depends on
    the SVD's gauge, in the reference as in any SVD.  Checked here: the same bond dimensions
    as the oracle MPS and the energy bound |sum <ZZ>| <= |E|."""
    n, p, chi = 16, 1, 16
    say("qaoa capped", n, p, chi)
    gates, edges, pos, _ = C.qaoa_maxcut(n, p, seed=0x7E5A0013)
    m = O.Mps(n, cutoff=1e-12, max_bond=chi).run(gates)
    s = mps.MpsState(eng, n, cutoff=1e-12, max_bond=chi).run(gates)
    assert list(s.bonds()) == list(m.bonds())
    assert max(s.bonds()) == chi
    zz = s.expval_zz([(pos[i], pos[j]) for i, j in edges])
    assert np.all(np.abs(zz) <= s.norm() + 1e-12)
    assert abs(zz.sum()) <= len(edges) * s.norm() + 1e-12


@pytest.mark.parametrize("precision,tol", [("fp64", TOL64), ("mixed", TOLMIX)])
@pytest.mark.parametrize("rows,cols,cycles", [(3, 4, 8), (4, 4, 6)])
def test_grid_fsim_vs_dense(eng, rows, cols, cycles, precision, tol):
    n = rows * cols
    say("grid", rows, cols, cycles, precision)
    gates = C.grid_fsim(rows, cols, cycles, seed=0x7E5A0005)
    ref = O.StateVector(n).run(gates)
    s = mps.MpsState(eng, n, cutoff=0.0, precision=precision).run(C.route_nearest_neighbour(gates))
    ez = s.expval_z_all()
    for q in range(n):
        assert abs(ez[q] - ref.expval_z_parity([q])) < tol
    bits = C.random_bitstrings(n, 16, seed=5)
    amp = s.amplitudes(bits)
    assert np.abs(amp - ref.amplitudes(bits)).max() < tol
    if precision == "fp64":
        assert np.abs(s.state_vector() - ref.psi).max() < TOL64


def test_mixed_batch_vs_oracle_mps(eng):
    say("mixed batch")
    circs = C.mixed_batch(8, seed=0x7E5A0004, n_range=(6, 12), depth_range=(4, 16), chis=(4, 8, 16, 32))
    jobs = []
    for c in circs:
        jobs.append({"n": c["n"], "gates": c["gates"], "trunc": mps.trunc_opts(1e-10, c["chi"]),
                     "zz": [(q, q + 1) for q in range(c["n"] - 1)]})
    res = eng.run_batch(jobs)
    for c, j, r in zip(circs, jobs, res):
        m = O.Mps(c["n"], cutoff=1e-10, max_bond=c["chi"]).run(c["gates"])
        for q in range(c["n"]):
            assert abs(r["expval_z"][q] - m.expval_z_parity([q])) < TOL64
        for (a, b), v in zip(j["zz"], r["expval_zz"]):
            assert abs(v - m.expval_z_parity([a, b])) < TOL64
        assert abs(r["discarded_weight"] - m.discarded()) < 1e-10
        assert r["max_bond"] == max(m.bonds())


@pytest.mark.parametrize("precision,tol", [("fp64", TOL64), ("mixed", TOLMIX)])
def test_sycamore_subcircuit_vs_dense(eng, precision, tol):
    """qFlex ingest end to end: the reference's Sycamore depth-10 circuit file, restricted
    to the gates on the first 14 line qubits, routed, against the dense state vector."""
    import os
    say("sycamore", precision)
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "sycamore", "sycamore_53_10_0.txt")
    gates = C.restrict(C.read_qflex(open(path).read()), 14)
    ref = O.StateVector(14).run(gates)
    s = mps.MpsState(eng, 14, cutoff=0.0, precision=precision).run(C.route_nearest_neighbour(gates))
    ez = s.expval_z_all()
    for q in range(14):
        assert abs(ez[q] - ref.expval_z_parity([q])) < tol
    bits = C.random_bitstrings(14, 16, seed=11)
    assert np.abs(s.amplitudes(bits) - ref.amplitudes(bits)).max() < tol


def test_brickwork_in_reference_gate_set(eng):
    """A Haar brickwork emitted as U3 + CNOT (KAK) gives the same <Z> and |amplitudes| on
    hip-mps as the dense run of the original 4x4 unitaries."""
    n = 12
    say("kak brickwork", n)
    gates = C.haar_brickwork(n, 10, seed=21)
    ref = O.StateVector(n).run(gates)
    s = mps.MpsState(eng, n, cutoff=0.0).run(C.to_reference_gates(gates))
    ez = s.expval_z_all()
    for q in range(n):
        assert abs(ez[q] - ref.expval_z_parity([q])) < TOL64
    bits = C.random_bitstrings(n, 16, seed=4)
    assert np.abs(np.abs(s.amplitudes(bits)) - np.abs(ref.amplitudes(bits))).max() < TOL64
