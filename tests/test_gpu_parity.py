"""GPU parity of libhipmps against the CPU FP64 oracle, through the C ABI.

Bars: FP64 path <Z> / amplitudes within 1e-10 of the dense state vector (no
truncation) and of the oracle MPS restatement (with truncation); mixed path <Z>
within 1e-5.  Singular values against LAPACK spectra in tests/golden.
"""
import os
import sys

import numpy as np
import pytest

from oracle import oracle as O
from tnqvm_amd import circuits as Cc
from tnqvm_amd import mps

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
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


def g(name, *qubits, params=()):
    return {"name": name, "qubits": tuple(qubits), "params": tuple(params)}


def rand_adjacent(n, n_gates, seed):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n_gates):
        if n > 1 and rng.random() < 0.5:
            q = int(rng.integers(0, n - 1))
            pair = (q, q + 1) if rng.random() < 0.5 else (q + 1, q)
            out.append({"name": "unitary", "qubits": pair, "matrix": Cc.haar_unitary(rng, 4)})
        else:
            out.append({"name": "unitary", "qubits": (int(rng.integers(0, n)),), "matrix": Cc.haar_unitary(rng, 2)})
    return out


def test_kats(eng):
    say("kats")
    # Bell, both operand orders
    for ops in ([g("H", 0), g("CNOT", 0, 1)], [g("H", 1), g("CNOT", 1, 0)]):
        s = mps.MpsState(eng, 4).run(ops)
        p = np.abs(s.state_vector()) ** 2
        assert abs(p[0] - 0.5) < 1e-14 and abs(p[3] - 0.5) < 1e-14
    # iSwap: X(q0); iSwap(q0,q1) -> |010>
    s = mps.MpsState(eng, 3).run([g("X", 0), g("iSwap", 0, 1)])
    assert abs(abs(s.state_vector()[2]) ** 2 - 1) < 1e-14
    # X chain with CNOT ladder -> 1111
    s = mps.MpsState(eng, 4).run([g("X", 0), g("CNOT", 0, 1), g("CNOT", 1, 2), g("CNOT", 2, 3)])
    assert abs(s.amplitudes([[1, 1, 1, 1]])[0] - 1) < 1e-14
    # Rx: <Z> = cos theta
    for th in np.linspace(-np.pi, np.pi, 7):
        s = mps.MpsState(eng, 1).run([g("Rx", 0, params=(th,))])
        assert abs(s.expval_z_parity([0]) - np.cos(th)) < 1e-14


def test_deuteron_table(eng):
    say("deuteron")
    import json
    kat = json.load(open(os.path.join(GOLD, "kat.json")))
    for t, ref in zip(kat["deuteron_angles"], kat["deuteron_x0x1"]):
        s = mps.MpsState(eng, 2).run([g("X", 0), g("Ry", 1, params=(t,)), g("CNOT", 1, 0), g("H", 0), g("H", 1)])
        assert abs(s.expval_z_parity([0, 1]) - ref) < 1e-6


@pytest.mark.parametrize("n", [1, 2, 3, 5, 8])
def test_random_adjacent_vs_dense(eng, n):
    say("random_adjacent", n)
    gates = rand_adjacent(n, 60, seed=100 + n)
    ref = O.StateVector(n).run(gates)
    s = mps.MpsState(eng, n).run(gates)
    assert np.abs(s.state_vector() - ref.psi).max() < 1e-12
    ez = s.expval_z_all()
    for q in range(n):
        assert abs(ez[q] - ref.expval_z_parity([q])) < 1e-12
    assert abs(s.norm() - 1.0) < 1e-12


def test_u3_cnot_line_fp64(eng):
    say("u3_cnot_line fp64")
    gates = Cc.u3_cnot_line()
    ref = O.StateVector(12).run(gates)
    s = mps.MpsState(eng, 12).run(gates)
    assert list(s.bonds()) == [1, 2, 4, 8, 16, 32, 64, 32, 16, 8, 4, 2, 1]
    ez = s.expval_z_all()
    for q in range(12):
        assert abs(ez[q] - ref.expval_z_parity([q])) < TOL64
        assert abs(s.expval_z_parity([q]) - ref.expval_z_parity([q])) < TOL64
    bits = Cc.random_bitstrings(12, 16, 5)
    assert np.abs(s.amplitudes(bits) - ref.amplitudes(bits)).max() < TOL64
    assert abs(s.norm() - ref.norm()) < TOL64
    pairs = [(q, q + 1) for q in range(11)] + [(0, 11), (3, 7)]
    zz = s.expval_zz(pairs)
    for (a, b), v in zip(pairs, zz):
        assert abs(v - ref.expval_z_parity([a, b])) < TOL64


def test_u3_cnot_line_mixed(eng):
    say("u3_cnot_line mixed")
    gates = Cc.u3_cnot_line()
    ref = O.StateVector(12).run(gates)
    s = mps.MpsState(eng, 12, precision="mixed").run(gates)
    ez = s.expval_z_all()
    for q in range(12):
        assert abs(ez[q] - ref.expval_z_parity([q])) < TOLMIX


def test_slice(eng):
    say("slice")
    gates = rand_adjacent(6, 40, seed=11)
    ref = O.StateVector(6).run(gates)
    s = mps.MpsState(eng, 6).run(gates)
    bits = [1, -1, 0, -1, -1, 1]
    sl = s.slice(bits)
    exp = []
    for x in range(8):
        b = [1, x & 1, 0, (x >> 1) & 1, (x >> 2) & 1, 1]
        exp.append(ref.psi[sum(bb << i for i, bb in enumerate(b))])
    assert np.abs(sl - np.array(exp)).max() < 1e-13


@pytest.mark.parametrize("cut,maxb", [(0.5, 0), (1e-3, 4), (1e-8, 3), (0.0, 2)])
def test_truncation_vs_oracle_mps(eng, cut, maxb):
    say("truncation", cut, maxb)
    gates = rand_adjacent(7, 80, seed=7)
    m = O.Mps(7, cutoff=cut, max_bond=maxb).run(gates)
    s = mps.MpsState(eng, 7, cutoff=cut, max_bond=maxb).run(gates)
    assert list(s.bonds()) == list(m.bonds())
    assert np.abs(s.state_vector() - m.state_vector()).max() < TOL64
    assert abs(s.stats()["discarded_weight"] - m.discarded()) < 1e-10


def test_svd_golden_spectra(eng):
    """theta' = golden matrix through a 4-site MPS with an identity gate on (1,2);
    the new left site's bond-column norms^2 are the singular values (sqrt split)."""
    z = np.load(os.path.join(GOLD, "svd_spectra.npz"))
    for key in ["gauss_16x16", "gauss_32x32", "gauss_64x64", "graded_64x64", "rank3_16x16", "gauss_96x128",
                "gauss_128x96"]:
        say("svd", key)
        a, sref = z["A_" + key], z["S_" + key]
        m, n = a.shape
        cl, cr = m // 2, n // 2
        s = mps.MpsState(eng, 4)
        s.set_site(0, np.ones((1, 2, cl)) / np.sqrt(2 * cl))
        s.set_site(1, a.reshape(cl, 2, n))          # chi_M = n
        s.set_site(2, np.eye(n).reshape(n, 2, cr))  # theta = A1 A2 = a
        s.set_site(3, np.ones((cr, 2, 1)) / np.sqrt(2 * cr))
        s.apply({"name": "unitary", "qubits": (1, 2), "matrix": np.eye(4)})
        s.flush()
        A1 = s.get_site(1).reshape(m, -1)
        A2 = s.get_site(2).reshape(-1, n)
        sig = (np.abs(A1) ** 2).sum(0)
        kk = min(len(sig), len(sref))
        assert np.allclose(sig[:kk], sref[:kk], atol=1e-12 * sref[0]), key
        assert np.abs(A1 @ A2 - a).max() < 1e-12 * sref[0], key


def test_batch_matches_single(eng):
    say("batch")
    jobs = []
    for k, n in enumerate([5, 8, 12]):
        gates = rand_adjacent(n, 50, seed=200 + k)
        jobs.append({"n": n, "gates": gates, "zz": [(0, 1), (1, n - 1)]})
    res = eng.run_batch(jobs)
    for j, r in zip(jobs, res):
        ref = O.StateVector(j["n"]).run(j["gates"])
        for q in range(j["n"]):
            assert abs(r["expval_z"][q] - ref.expval_z_parity([q])) < TOL64
        for (a, b), v in zip(j["zz"], r["expval_zz"]):
            assert abs(v - ref.expval_z_parity([a, b])) < TOL64
        assert abs(r["norm"] - 1.0) < 1e-10


def test_serving_alone_equals_mixed(eng):
    say("serving")
    srv = mps.Server(eng)
    jobs = [(n, rand_adjacent(n, 40 + 5 * k, seed=300 + k)) for k, n in enumerate([4, 9, 6, 11])]
    alone = []
    for n, gates in jobs:
        t = srv.submit(n, gates)
        r = None
        while r is None:
            srv.step()
            r = srv.poll(t)
        alone.append(r)
    # resubmit all at once, staggered arrival
    tickets = []
    for k, (n, gates) in enumerate(jobs):
        tickets.append(srv.submit(n, gates))
        srv.step()
    results = {}
    while len(results) < len(tickets):
        srv.step()
        for t in tickets:
            if t not in results:
                r = srv.poll(t)
                if r is not None:
                    results[t] = r
    for (n, gates), t, a in zip(jobs, tickets, alone):
        ref = O.StateVector(n).run(gates)
        ez = results[t]["expval_z"]
        for q in range(n):
            assert abs(ez[q] - ref.expval_z_parity([q])) < TOL64
        # the same circuit alone and in the mix gives the same numbers
        assert np.abs(ez - a["expval_z"]).max() < 1e-13


def test_serving_background_lane(eng):
    """Circuits whose bond can reach big_lane_min_chi run on the background lane, concurrently
    with the steps of the others; every result equals the circuit run alone and the oracle."""
    say("serving background lane")
    base = eng.tuning()
    try:
        eng.set_tuning(big_lane_min_chi=16)
        srv = mps.Server(eng)
        # n = 10 or 11 can reach 32 (big, cap 32); n = 6 reaches 8 (small); n = 9 capped at 8 (small)
        jobs = [(10, rand_adjacent(10, 60, seed=501), 32), (6, rand_adjacent(6, 30, seed=502), 0),
                (11, rand_adjacent(11, 70, seed=503), 32), (9, rand_adjacent(9, 40, seed=504), 8),
                (10, rand_adjacent(10, 50, seed=505), 0)]
        tickets = []
        for n, gates, cap in jobs:
            tickets.append(srv.submit(n, gates, mps.trunc_opts(cutoff=1e-12, max_bond=cap)))
            srv.step()
        results = {}
        for _ in range(2000):
            if len(results) == len(tickets):
                break
            srv.step()
            for t in tickets:
                if t not in results:
                    r = srv.poll(t)
                    if r is not None:
                        results[t] = r
        assert len(results) == len(tickets)
        assert srv.pending() == (0, 0)
        for (n, gates, cap), t in zip(jobs, tickets):
            alone = mps.MpsState(eng, n, cutoff=1e-12, max_bond=cap).run(gates)
            ez_alone = alone.expval_z_all()
            alone.close()
            assert np.abs(results[t]["expval_z"] - ez_alone).max() < 1e-12
            if cap == 0:
                ref = O.StateVector(n).run(gates)
                for q in range(n):
                    assert abs(results[t]["expval_z"][q] - ref.expval_z_parity([q])) < TOL64
    finally:
        eng.set_tuning(**base)


@pytest.mark.parametrize("n,depth,chi", [(12, 14, 32), (14, 12, 64)])
def test_lq_preconditioning_same_results(n, depth, chi):
    """The Jacobi on R2 of the double Householder LQ gives the same truncated MPS as the
    Jacobi on X itself, and both match the oracle MPS."""
    say("lq precondition", n, depth, chi)
    gates = Cc.haar_brickwork(n, depth, seed=31 + n)
    m = O.Mps(n, cutoff=1e-10, max_bond=chi).run(gates)
    res = {}
    for rank in (0, 16):
        e = mps.Engine()
        e.set_lq_precondition(rank)
        s = mps.MpsState(e, n, cutoff=1e-10, max_bond=chi).run(gates)
        res[rank] = (list(s.bonds()), s.expval_z_all(), s.stats()["discarded_weight"])
        s.close()
        e.close()
    assert res[0][0] == res[16][0] == list(m.bonds())
    assert np.abs(res[0][1] - res[16][1]).max() < TOL64
    for q in range(n):
        assert abs(res[16][1][q] - m.expval_z_parity([q])) < TOL64
    assert abs(res[16][2] - m.discarded()) < 1e-10


def _graded(m, n, seed, decades=12.0):
    """m x n matrix with known singular values logspace(0, -decades, min(m, n))."""
    rng = np.random.default_rng(seed)
    k = min(m, n)
    u = np.linalg.qr(rng.standard_normal((m, k)) + 1j * rng.standard_normal((m, k)))[0]
    v = np.linalg.qr(rng.standard_normal((n, k)) + 1j * rng.standard_normal((n, k)))[0]
    s = np.logspace(0.0, -decades, k)
    return (u * s) @ v.conj().T, s


@pytest.mark.parametrize("m,n", [(1024, 1024), (512, 1024), (1024, 512), (2048, 2048)])
def test_svd_large_graded(eng, m, n):
    """Bond dimension 512 (theta' up to 1024 x 1024) and 1024 (2048 x 2048, the chi_max of the
    6x6 grid circuits) through the LQ-preconditioned Jacobi, with singular values spread over
    12 decades (known exactly by construction)."""
    say("svd large", m, n)
    a, sref = _graded(m, n, seed=m + 3 * n)
    cl, cr = m // 2, n // 2
    s = mps.MpsState(eng, 4)
    s.set_site(0, np.ones((1, 2, cl)) / np.sqrt(2 * cl))
    s.set_site(1, a.reshape(cl, 2, n))
    s.set_site(2, np.eye(n).reshape(n, 2, cr))
    s.set_site(3, np.ones((cr, 2, 1)) / np.sqrt(2 * cr))
    s.apply({"name": "unitary", "qubits": (1, 2), "matrix": np.eye(4)})
    s.flush()
    A1 = s.get_site(1).reshape(m, -1)
    A2 = s.get_site(2).reshape(-1, n)
    sig = (np.abs(A1) ** 2).sum(0)
    assert len(sig) == min(m, n)
    assert np.abs(sig - sref).max() < 1e-12
    assert np.abs(A1 @ A2 - a).max() < 1e-12
    s.close()


def test_empty_and_one_qubit_circuits(eng):
    """Edge cases of the request path: a circuit with no gates, a one-qubit register, both alone,
    in a batch beside a regular circuit, and through the serving loop."""
    say("empty and one-qubit circuits")
    s = mps.MpsState(eng, 5)
    assert np.abs(s.expval_z_all() - 1.0).max() < 1e-15
    assert abs(s.norm() - 1.0) < 1e-15
    assert abs(s.amplitudes(np.zeros((1, 5), dtype=np.int8))[0] - 1.0) < 1e-15
    s.close()
    one = [{"name": "H", "qubits": (0,)}, {"name": "Rz", "qubits": (0,), "params": (0.3,)},
           {"name": "Ry", "qubits": (0,), "params": (1.1,)}]
    ref1 = O.StateVector(1).run(one)
    s1 = mps.MpsState(eng, 1).run(one)
    assert abs(s1.expval_z_all()[0] - ref1.expval_z_parity([0])) < 1e-14
    assert np.abs(s1.state_vector() - ref1.psi).max() < 1e-14
    s1.close()
    line = Cc.u3_cnot_line(4, 3, seed=9)
    ref4 = O.StateVector(4).run(line)
    jobs = [{"n": 3, "gates": [], "trunc": mps.trunc_opts(), "zz": [(0, 1)]},
            {"n": 1, "gates": one, "trunc": mps.trunc_opts(), "zz": []},
            {"n": 4, "gates": line, "trunc": mps.trunc_opts(), "zz": [(1, 2)]}]
    res = eng.run_batch(jobs)
    assert np.abs(res[0]["expval_z"] - 1.0).max() < 1e-15 and abs(res[0]["expval_zz"][0] - 1.0) < 1e-15
    assert abs(res[1]["expval_z"][0] - ref1.expval_z_parity([0])) < 1e-14
    for q in range(4):
        assert abs(res[2]["expval_z"][q] - ref4.expval_z_parity([q])) < 1e-12
    assert abs(res[2]["expval_zz"][0] - ref4.expval_z_parity([1, 2])) < 1e-12
    srv = mps.Server(eng)
    t0, t1 = srv.submit(3, []), srv.submit(1, one)
    got = {}
    for _ in range(10):
        srv.step()
        for t in (t0, t1):
            if t not in got:
                r = srv.poll(t)
                if r is not None:
                    got[t] = r
        if len(got) == 2:
            break
    assert len(got) == 2
    assert np.abs(got[t0]["expval_z"] - 1.0).max() < 1e-15
    assert abs(got[t1]["expval_z"][0] - ref1.expval_z_parity([0])) < 1e-14


def test_tuning_roundtrip_and_results(eng):
    """Lanes and wave groups only change the schedule (identical results, bit for bit); FP32-Gram
    sweeps only the mixed path's rounding; bad values are refused."""
    say("tuning")
    base = eng.tuning()
    with pytest.raises(Exception):
        eng.set_tuning(jacobi_lanes=0)
    with pytest.raises(Exception):
        eng.set_tuning(wave_groups=3)
    circs = [Cc.haar_brickwork(12, 8, seed=0x7E5A0B00 + k) for k in range(6)]

    def run(precision):
        jobs = [{"n": 12, "gates": c, "trunc": mps.trunc_opts(1e-10, 32, precision=precision), "zz": []}
                for c in circs]
        return np.array([r["expval_z"] for r in eng.run_batch(jobs)])

    try:
        eng.set_tuning(jacobi_lanes=1, lane_min_pairs=0, gram32_sweeps=0)
        a64, amx = run("fp64"), run("mixed")
        eng.set_tuning(jacobi_lanes=2, lane_min_pairs=0, gram32_sweeps=0)
        assert eng.tuning()["jacobi_lanes"] == 2
        b64 = run("fp64")
        assert np.array_equal(a64, b64)
        eng.set_tuning(jacobi_lanes=2, lane_min_pairs=0, gram32_sweeps=5)
        c64, cmx = run("fp64"), run("mixed")
        assert np.array_equal(a64, c64)  # FP64 states never take the FP32 Gram
        assert np.abs(cmx - amx).max() < 1e-6
        # wave groups: circuits whose bond can reach 128 are balanced over the two groups, the
        # small ones stay in the first
        heavy = [Cc.haar_brickwork(16, 10, seed=0x7E5A0B10 + k) for k in range(4)]

        def run_mix(precision):
            jobs = [{"n": 12, "gates": c, "trunc": mps.trunc_opts(1e-10, 32, precision=precision), "zz": []}
                    for c in circs]
            jobs += [{"n": 16, "gates": c, "trunc": mps.trunc_opts(1e-10, 128, precision=precision), "zz": []}
                     for c in heavy]
            return np.array([r["expval_z"] for r in eng.run_batch(jobs)], dtype=object)

        ref = {}
        for groups in (1, 2):
            eng.set_tuning(wave_groups=groups)
            assert eng.tuning()["wave_groups"] == groups
            assert np.array_equal(run("fp64"), c64) and np.array_equal(run("mixed"), cmx)
            for p in ("fp64", "mixed"):
                z = run_mix(p)
                if groups == 1:
                    ref[p] = z
                else:
                    assert all(np.array_equal(a, b) for a, b in zip(z, ref[p]))
    finally:
        eng.set_tuning(**base)


def test_autotune_small_shape(eng):
    """The autotuner times each candidate tuning on the same circuits and leaves the fastest set
    on the engine (here: wave groups 1 and 2 at a small shape)."""
    from tnqvm_amd import autotune
    say("autotune")
    base = eng.tuning()
    try:
        out = autotune.tune(eng, n=12, chi=16, batch=4, warmup=4, steps=2, lanes=(2,), gram32=(5,), lq=(64,),
                            groups=(1, 2))
        assert [r["wave_groups"] for r in out["candidates"]] == [1, 2]
        assert all(r["applies_per_s"] > 0 for r in out["candidates"])
        assert eng.tuning()["wave_groups"] == out["choice"]["wave_groups"]
    finally:
        eng.set_tuning(**base)
