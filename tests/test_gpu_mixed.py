"""GPU parity of the mixed-precision path (complex64 sites and theta, FP64 Jacobi SVD, norms
and truncation weights) on truncating workloads, against the FP64 oracle.

The truncation rule is the exatn-mps one (tnqvm/visitors/exatn-mps/ExaTnMpsVisitor.cpp:
2366-2536): partial norms, first sub-threshold index kept, max-bond-dim clamp, no
renormalisation.  A heavily truncated state can be ill-conditioned: rounding only its gate
matrices to complex64 can move the FP64 result itself.  Each case first measures that
sensitivity on the oracle (FP64 MPS with the gates rounded to complex64) and asserts it is
below 1e-6, so that the 1e-5 bar on the mixed path tests its arithmetic and not the
conditioning of the circuit.

Cases:
* Haar brickwork (the 40-qubit workload's generator) at n = 16-20, chi_max 32-64,
  cutoff 1e-10: bonds equal to the oracle's, <Z_i> and <Z_i Z_i+1> within 1e-5;
* a batch of uneven brickwork circuits through hipmps_run_batch in mixed precision;
* a staggered serving mix in mixed precision with a 3x4 grid fSim circuit capped at
  chi 8: each result equal to its alone run (1e-13) and to the oracle (1e-5);
* QAOA MaxCut p = 4 (the 60-qubit workload's generator) at n = 16: energy and every
  <Z_i Z_j> against the dense state vector, FP64 to 1e-10 and mixed to 1e-5.
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
SENS_MAX = 1e-6


def say(*a):
    print("[gpu-case]", *a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def eng():
    e = mps.Engine()
    yield e
    e.close()


def rounded_c64(gates):
    out = []
    for g in gates:
        h = dict(g)
        m = np.asarray(O.gate_of(g))
        h["matrix"] = m.astype(np.complex64).astype(np.complex128)
        h.pop("params", None)
        out.append(h)
    return out


def oracle_obs(n, gates, cutoff, chi, pairs):
    m = O.Mps(n, cutoff=cutoff, max_bond=chi).run(gates)
    z = np.array([m.expval_z_parity([q]) for q in range(n)])
    zz = np.array([m.expval_z_parity(list(p)) for p in pairs])
    return m, z, zz


def sensitivity(n, gates, cutoff, chi, pairs, z, zz):
    """Largest move of the FP64 oracle's <Z> / <ZZ> when only the gates are rounded to complex64."""
    _, z2, zz2 = oracle_obs(n, rounded_c64(gates), cutoff, chi, pairs)
    return max(np.abs(z - z2).max(), np.abs(zz - zz2).max() if len(zz) else 0.0)


@pytest.mark.parametrize("n,depth,chi,seed", [(16, 10, 32, 0x7E5A0300), (18, 12, 32, 0x7E5A0302),
                                              (20, 12, 64, 0x7E5A0301)])
def test_brickwork_mixed_vs_oracle_mps(eng, n, depth, chi, seed):
    say("brickwork mixed", n, depth, chi, hex(seed))
    gates = C.haar_brickwork(n, depth, seed=seed)
    pairs = [(q, q + 1) for q in range(n - 1)]
    m, z, zz = oracle_obs(n, gates, 1e-10, chi, pairs)
    sens = sensitivity(n, gates, 1e-10, chi, pairs, z, zz)
    assert sens < SENS_MAX, sens
    assert max(m.bonds()) == chi  # the cap is active: the state truncates
    s = mps.MpsState(eng, n, cutoff=1e-10, max_bond=chi, precision="mixed").run(gates)
    assert list(s.bonds()) == list(m.bonds())
    ez = s.expval_z_all()
    ezz = s.expval_zz(pairs)
    err = max(np.abs(ez - z).max(), np.abs(ezz - zz).max())
    say("  sensitivity %.2e, mixed error %.2e" % (sens, err))
    assert err < TOLMIX
    st = s.stats()
    assert abs(st["discarded_weight"] - m.discarded()) < TOLMIX * max(1.0, m.discarded())
    s.close()


@pytest.mark.parametrize("precision,tol", [("fp64", TOL64), ("mixed", TOLMIX)])
def test_uneven_batch_vs_oracle_mps(eng, precision, tol):
    """hipmps_run_batch on uneven brickwork circuits (per-circuit qubits, depth and chi_max), each
    result against the oracle MPS of that circuit alone."""
    say("uneven batch", precision)
    circs = C.mixed_batch(8, seed=0x7E5A0310, n_range=(8, 16), depth_range=(6, 14), chis=(8, 16, 32))
    jobs = [{"n": c["n"], "gates": c["gates"], "trunc": mps.trunc_opts(1e-10, c["chi"], precision=precision),
             "zz": [(q, q + 1) for q in range(c["n"] - 1)]} for c in circs]
    res = eng.run_batch(jobs)
    worst = 0.0
    for c, j, r in zip(circs, jobs, res):
        assert r["status"] == 0
        m, z, zz = oracle_obs(c["n"], c["gates"], 1e-10, c["chi"], j["zz"])
        if precision == "mixed":
            assert sensitivity(c["n"], c["gates"], 1e-10, c["chi"], j["zz"], z, zz) < SENS_MAX
        err = max(np.abs(r["expval_z"] - z).max(), np.abs(r["expval_zz"] - zz).max())
        worst = max(worst, err)
        assert err < tol
        assert r["max_bond"] == max(m.bonds())
        assert abs(r["discarded_weight"] - m.discarded()) < tol * max(1.0, m.discarded())
    say("  worst error %.2e" % worst)


def test_serving_mixed_alone_equals_mix(eng):
    """Staggered arrivals of brickwork circuits and a 3x4 grid fSim circuit (snake map, routed,
    capped at chi 8 so that it truncates), mixed precision: every result equals its alone run
    to 1e-13 and the FP64 oracle to 1e-5."""
    say("serving mixed")
    circs = C.mixed_batch(4, seed=0x7E5A0320, n_range=(8, 14), depth_range=(6, 12), chis=(8, 16))
    jobs = [(c["n"], c["gates"], c["chi"]) for c in circs]
    grid = C.route_nearest_neighbour(C.grid_fsim(3, 4, 8, seed=0x7E5A0321))
    jobs.insert(2, (12, grid, 8))
    srv = mps.Server(eng)

    def submit(n, gates, chi):
        return srv.submit(n, gates, mps.trunc_opts(1e-10, chi, precision="mixed"),
                          zz=[(q, q + 1) for q in range(n - 1)])

    alone = []
    for n, gates, chi in jobs:
        t = submit(n, gates, chi)
        r = None
        while r is None:
            srv.step()
            r = srv.poll(t)
        alone.append(r)
    tickets = []
    for job in jobs:
        tickets.append(submit(*job))
        srv.step()
    results = {}
    for _ in range(100000):
        if len(results) == len(tickets):
            break
        srv.step()
        for t in tickets:
            if t not in results:
                r = srv.poll(t)
                if r is not None:
                    results[t] = r
    assert len(results) == len(tickets)
    for (n, gates, chi), t, a in zip(jobs, tickets, alone):
        r = results[t]
        assert r["status"] == 0
        pairs = [(q, q + 1) for q in range(n - 1)]
        m, z, zz = oracle_obs(n, gates, 1e-10, chi, pairs)
        assert sensitivity(n, gates, 1e-10, chi, pairs, z, zz) < SENS_MAX
        assert np.abs(r["expval_z"] - a["expval_z"]).max() < 1e-13
        assert np.abs(r["expval_zz"] - a["expval_zz"]).max() < 1e-13
        assert max(np.abs(r["expval_z"] - z).max(), np.abs(r["expval_zz"] - zz).max()) < TOLMIX
        assert r["max_bond"] == max(m.bonds())
    grid_res = results[tickets[2]]
    assert grid_res["max_bond"] == 8


@pytest.mark.parametrize("precision,tol", [("fp64", TOL64), ("mixed", TOLMIX)])
def test_qaoa_p4_vs_dense(eng, precision, tol):
    """QAOA MaxCut at p = 4 (swap network, fused ZZ + swap), cutoff 1e-12 as in the 60-qubit
    workload: energy and each edge's <Z_i Z_j> against the dense logical circuit."""
    n, p = 16, 4
    say("qaoa p4", n, precision)
    gates, edges, pos, (gam, bet) = C.qaoa_maxcut(n, p, seed=0x7E5A0330)
    ref = O.StateVector(n).run(C.qaoa_logical(n, edges, gam, bet))
    s = mps.MpsState(eng, n, cutoff=1e-12, precision=precision).run(gates)
    zz = s.expval_zz([(pos[i], pos[j]) for i, j in edges])
    want = np.array([ref.expval_z_parity([i, j]) for i, j in edges])
    say("  max bond", int(s.bonds().max()), "energy", float(zz.sum()), "error %.2e" % np.abs(zz - want).max())
    assert np.abs(zz - want).max() < tol
    assert abs(zz.sum() - want.sum()) < tol * len(edges)
    assert -len(edges) - 1e-9 <= zz.sum() <= 1e-9 + len(edges)
