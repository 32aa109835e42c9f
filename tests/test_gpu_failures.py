"""Failure isolation through the C ABI: a circuit that fails -- refused by the host checks,
unable to grow its sites within the HBM budget, or stopped by the Jacobi sweep cap -- reports
its own status, and the circuits beside it give the results they give alone.

The failing conditions are forced from the host (a small hbm_budget_bytes, a sweep cap of one,
a gate on non-neighbouring qubits); every kernel still exits normally.  The reference runs one
circuit at a time and has no such isolation; its only guard is the post-SVD norm check
(tnqvm/visitors/exatn-mps/ExaTnMpsVisitor.cpp:1634-1661).
"""
import sys

import numpy as np
import pytest

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


def _jobs(circs, precision="fp64"):
    return [{"n": c["n"], "gates": c["gates"], "trunc": mps.trunc_opts(1e-10, c["chi"], precision=precision),
             "zz": [(q, q + 1) for q in range(c["n"] - 1)]} for c in circs]


def _alone(eng, job):
    return eng.run_batch([job])[0]


def _peak_bytes(eng, job):
    """Most site bytes the circuit holds at once in a run alone (old and new buffers of a wave)."""
    s = mps.MpsState(eng, job["n"], cutoff=1e-10, max_bond=job["trunc"].max_bond_dim).run(job["gates"])
    b = s.stats()["device_bytes_peak"]
    s.close()
    return b


def test_batch_budget_one_circuit_cannot_grow(eng):
    """A budget the big circuit can never fit: it runs last and alone (its footprint is above
    the budget) and fails with HIPMPS_E_NOMEM; every small circuit of the batch gives its alone
    result."""
    small = C.mixed_batch(4, seed=0x7E5A0101, n_range=(6, 8), depth_range=(6, 10), chis=(4, 8))
    big = {"n": 16, "gates": C.haar_brickwork(16, 12, seed=0x7E5A0102), "chi": 128}
    jobs = _jobs(small[:2] + [big] + small[2:])
    want = [_alone(eng, j) for j in jobs[:2] + jobs[3:]]
    fps = [eng.footprint(j["n"], j["gates"], j["trunc"]) for j in jobs]
    budget = sum(fps[:2] + fps[3:]) + (1 << 16)
    assert fps[2] > budget and _peak_bytes(eng, jobs[2]) > budget
    say("budget", budget, "footprints", fps)
    e2 = mps.Engine(hbm_budget_bytes=budget)
    try:
        res = e2.run_batch(jobs)
    finally:
        e2.close()
    assert res[2]["status"] == _capi.HIPMPS_E_NOMEM
    for r, w in zip(res[:2] + res[3:], want):
        assert r["status"] == 0
        assert np.abs(r["expval_z"] - w["expval_z"]).max() < 1e-12
        assert np.abs(r["expval_zz"] - w["expval_zz"]).max() < 1e-12
        assert r["max_bond"] == w["max_bond"]


def test_batch_budget_waits_for_memory(eng):
    """Two circuits whose footprints fit one at a time but not together: the second is admitted
    when the first has finished and freed its sites; both complete with their alone results."""
    a = {"n": 14, "gates": C.haar_brickwork(14, 10, seed=0x7E5A0103), "chi": 64}
    b = {"n": 14, "gates": C.haar_brickwork(14, 12, seed=0x7E5A0104), "chi": 64}
    jobs = _jobs([a, b])
    want = [_alone(eng, j) for j in jobs]
    fps = [eng.footprint(j["n"], j["gates"], j["trunc"]) for j in jobs]
    budget = max(fps) + (1 << 16)
    assert sum(fps) > budget
    assert max(_peak_bytes(eng, j) for j in jobs) <= budget
    say("budget", budget, "footprints", fps)
    e2 = mps.Engine(hbm_budget_bytes=budget)
    try:
        res = e2.run_batch(jobs)
    finally:
        e2.close()
    for r, w in zip(res, want):
        assert r["status"] == 0
        assert np.abs(r["expval_z"] - w["expval_z"]).max() < 1e-12
        assert r["max_bond"] == w["max_bond"]
    assert res[1]["latency_ms"] > res[0]["latency_ms"]


def test_state_growth_failure_keeps_state(eng):
    """One state whose next wave cannot grow within the budget: the call fails with
    HIPMPS_E_NOMEM, the wave's 2-qubit gates are dropped and the sites are those of the last
    completed wave (same amplitudes as before the call); a second try fails the same way."""
    n = 12
    layers = C.brickwork_layers(C.haar_brickwork(n, 8, seed=0x7E5A0105), n)
    head = [g for l in layers[:3] for g in l]
    tail = [g for l in layers[3:] for g in l if len(g["qubits"]) == 2]  # no 1-qubit gates left pending
    s = mps.MpsState(eng, n, cutoff=0.0).run(head)
    head_peak = s.stats()["device_bytes_peak"]
    s.run(tail)
    full_bytes = s.stats()["device_bytes"]
    s.close()
    assert full_bytes > 2 * head_peak
    # room for the head's sites, not for the tail's
    e2 = mps.Engine(hbm_budget_bytes=head_peak + 256)
    try:
        s = mps.MpsState(e2, n, cutoff=0.0).run(head)
        s.flush()
        bits = C.random_bitstrings(n, 8, seed=4)
        before = s.amplitudes(bits)
        for _ in range(2):
            with pytest.raises(_capi.HipMpsError) as ei:
                s.run(tail)
                s.flush()
            assert ei.value.code == _capi.HIPMPS_E_NOMEM
            assert np.abs(s.amplitudes(bits) - before).max() == 0.0
        s.close()
    finally:
        e2.close()


def test_bad_circuits_refused_alone(eng):
    """serve_submit refuses a circuit with a gate on non-neighbouring qubits or a qubit out of
    range (nothing queued); in a batch such a circuit reports the code in its status and the
    good circuits beside it run as alone."""
    good = C.mixed_batch(3, seed=0x7E5A0106, n_range=(6, 8), depth_range=(4, 8), chis=(8,))
    jobs = _jobs(good)
    want = [_alone(eng, j) for j in jobs]
    cnot = mps.gate_matrix("CNOT")
    nonadj = {"n": 6, "gates": [{"name": "unitary", "qubits": (0, 2), "matrix": cnot}], "trunc": mps.trunc_opts()}
    outside = {"n": 6, "gates": [{"name": "unitary", "qubits": (5, 6), "matrix": cnot}], "trunc": mps.trunc_opts()}
    srv = mps.Server(eng)
    for bad, code in ((nonadj, _capi.HIPMPS_E_NONADJACENT), (outside, _capi.HIPMPS_E_ARG)):
        with pytest.raises(_capi.HipMpsError) as ei:
            srv.submit(bad["n"], bad["gates"], bad["trunc"])
        assert ei.value.code == code
    assert srv.pending() == (0, 0)
    res = eng.run_batch([jobs[0], nonadj, jobs[1], outside, jobs[2]])
    assert res[1]["status"] == _capi.HIPMPS_E_NONADJACENT
    assert res[3]["status"] == _capi.HIPMPS_E_ARG
    for r, w in zip([res[0], res[2], res[4]], want):
        assert r["status"] == 0
        assert np.abs(r["expval_z"] - w["expval_z"]).max() < 1e-12


def test_sweep_cap_surfaces_convergence(eng):
    """A Jacobi sweep cap of one: every SVD that needs a second sweep is reported
    (HIPMPS_E_CONVERGENCE from flush, svds_nonconverged in the stats, the batch result's
    status), and results are still written."""
    e2 = mps.Engine(jacobi_max_sweeps=1)
    try:
        n = 10
        gates = C.haar_brickwork(n, 6, seed=0x7E5A0107)
        s = mps.MpsState(e2, n, cutoff=1e-10)
        s.run(gates)
        with pytest.raises(_capi.HipMpsError) as ei:
            s.flush()
        assert ei.value.code == _capi.HIPMPS_E_CONVERGENCE
        assert s.stats()["svds_nonconverged"] > 0
        s.close()
        res = e2.run_batch(_jobs([{"n": n, "gates": gates, "chi": 32}]))
        assert res[0]["status"] == _capi.HIPMPS_E_CONVERGENCE
        assert np.all(np.isfinite(res[0]["expval_z"]))
    finally:
        e2.close()
    # the default cap converges on the same circuit
    s = mps.MpsState(eng, 10, cutoff=1e-10).run(gates)
    s.flush()
    assert s.stats()["svds_nonconverged"] == 0
    s.close()


def test_serve_admission(eng):
    """Serving admission by projected footprint: hipmps_serve_submit refuses a circuit whose
    footprint exceeds the engine's HBM budget (HIPMPS_E_NOMEM, nothing queued); two circuits that
    fit only one at a time are both accepted, the second waits in the queue until the first has
    finished, and both give their alone results."""
    a = {"n": 14, "gates": C.haar_brickwork(14, 10, seed=0x7E5A0108), "chi": 64}
    b = {"n": 14, "gates": C.haar_brickwork(14, 12, seed=0x7E5A0109), "chi": 64}
    big = {"n": 16, "gates": C.haar_brickwork(16, 12, seed=0x7E5A010A), "chi": 128}
    jobs = _jobs([a, b, big])
    want = [_alone(eng, j) for j in jobs[:2]]
    fps = [eng.footprint(j["n"], j["gates"], j["trunc"]) for j in jobs]
    budget = max(fps[:2]) + (1 << 16)
    assert sum(fps[:2]) > budget and fps[2] > budget
    say("serve budget", budget, "footprints", fps)
    e2 = mps.Engine(hbm_budget_bytes=budget)
    try:
        srv = mps.Server(e2)
        with pytest.raises(_capi.HipMpsError) as ei:
            srv.submit(jobs[2]["n"], jobs[2]["gates"], jobs[2]["trunc"], zz=jobs[2]["zz"])
        assert ei.value.code == _capi.HIPMPS_E_NOMEM
        assert srv.pending() == (0, 0)
        tickets = [srv.submit(j["n"], j["gates"], j["trunc"], zz=j["zz"]) for j in jobs[:2]]
        srv.step()
        assert srv.pending() == (1, 1)  # the second waits for the first's memory
        results = {}
        for _ in range(100000):
            if len(results) == 2:
                break
            srv.step()
            for t in tickets:
                if t not in results:
                    r = srv.poll(t)
                    if r is not None:
                        results[t] = r
                        if len(results) == 1:
                            assert t == tickets[0]
        assert len(results) == 2
        for t, w in zip(tickets, want):
            r = results[t]
            assert r["status"] == 0
            assert np.abs(r["expval_z"] - w["expval_z"]).max() < 1e-12
            assert np.abs(r["expval_zz"] - w["expval_zz"]).max() < 1e-12
            assert r["max_bond"] == w["max_bond"]
    finally:
        e2.close()


def test_descriptor_upload_held_back(eng, monkeypatch):
    """A wave's descriptor upload held back on the device (HIPMPS_TEST_STAGE_DELAY_US queues a
    2 ms sleeping wave ahead of it, as a stream waiting behind other work on a shared hardware
    queue would): the host edits the staged Jacobi descriptors of LQ-preconditioned items only
    after the copy has read them, so every batch result equals the undelayed engine's, bit for
    bit, and a state run alone on the delayed engine agrees to 1e-12."""
    circs = [{"n": 12, "gates": C.haar_brickwork(12, 8, seed=0x7E5A0B00 + k), "chi": 32} for k in range(3)]
    circs.append({"n": 14, "gates": C.haar_brickwork(14, 10, seed=0x7E5A0108), "chi": 64})
    jobs = _jobs(circs, "mixed") + _jobs(circs, "fp64")
    want = eng.run_batch(jobs)
    monkeypatch.setenv("HIPMPS_TEST_STAGE_DELAY_US", "2000")
    e2 = mps.Engine()
    monkeypatch.delenv("HIPMPS_TEST_STAGE_DELAY_US")
    try:
        say("held-back upload", len(jobs), "circuits")
        got = e2.run_batch(jobs)
        s = mps.MpsState(e2, 14, cutoff=1e-10, max_bond=64).run(circs[3]["gates"])
        z_state = s.expval_z_all()
        s.close()
    finally:
        e2.close()
    for g, w in zip(got, want):
        assert g["status"] == 0 and w["status"] == 0
        assert np.array_equal(g["expval_z"], w["expval_z"])
        assert np.array_equal(g["expval_zz"], w["expval_zz"])
        assert g["max_bond"] == w["max_bond"]
    assert np.abs(z_state - want[len(circs) + 3]["expval_z"]).max() < 1e-12
