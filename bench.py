#!/usr/bin/env python3
"""hip-mps headline benchmark: 2-qubit-gate applies/s at chi = 256, with circuit
latency p50/p99, batched mixed precision against the plain per-circuit FP64 path.

Headline workload (synthetic, seeded): 8 circuits of the 40-qubit brickwork of Haar-random
gates (a Haar 2x2 on every qubit, then Haar 4x4 gates on alternating even/odd neighbour
pairs), depth 40, chi_max = 256, svd-cutoff 1e-10.  A "step" is one brickwork layer
advanced for every circuit of the batch.  The first layers (at least 12; 40 - 1 - warmup -
steps when that is more) run the full brickwork and bring bonds 8..32 to chi = 256.  The
--warmup and --steps layers that follow apply the brickwork's Haar gates only on the pairs
(i, i+1) with 8 <= i <= 30, whose theta' is 512 x 512 once those bonds are saturated: every
timed apply is theta 512 x 512 -> SVD -> clamp to 256, the unit of the headline metric.
The engine counts the timed applies by theta' size; `value` counts only the 512 x 512 ones
over the whole timed span.  The remaining layers (at least one) run untimed.

Printed: ONE JSON line (rank 0, N = 1).  Fields beyond the driver contract:
  plain_fp64_value   the same layers, one circuit at a time, complex128 everywhere
  speedup_vs_plain   value / plain_fp64_value
  plain_fp64_gate_at_a_time_value
                     plain FP64, one gate per call with a synchronisation after each
  uneven_batch       256 brickwork circuits of uneven size (qubits U{16..48}, depth
                     U{10..60}, chi_max from {32,64,128,256}) submitted together, against
                     each circuit run alone in plain FP64; circuits/s, applies/s, the ratio,
                     and the <Z> difference between the two paths per circuit
  serving            Poisson arrivals (40/s, above what either path sustains) of those 256
                     brickwork circuits and two 6x6 grid fSim circuits (chi_max 1024) through the
                     continuous-batching loop, circuits whose bond can reach 512 on the background
                     lane; against the plain path one circuit at a time: sustained circuits/s and
                     p50/p99 latency for both, latency per class, the grid circuit's time alone
  serving_moderate   the same loop below capacity: 96 of the brickwork circuits and one grid
                     circuit at 5 arrivals/s
  latency_ms_p50/p99 the batched latencies (arrival to result) of serving_moderate
  qaoa_60            the 60-qubit QAOA MaxCut circuit (3-regular graph, p = 4, odd-even swap
                     network, chi_max 512, cutoff 1e-12), mixed and FP64: applies/s, energy,
                     their difference, discarded weight
  roofline           the SVD (double LQ + Jacobi rounds) of the timed applies, FP64 bound:
                     nominal 704 chi^3 flops per apply over the wall time of the timed steps
                     (two wave groups overlap); the Jacobi rounds' executed flops and HBM
                     traffic beside it
  noisy              the purified-MPS (noisy) path: a 40-qubit u3/CNOT circuit of depth 4 with the
                     ibmqx2 backend's single-qubit channels after every gate (qubit q takes the
                     channels of backend qubit q mod 5), no bond cap, cut 1e-9: wall time per
                     circuit with its <Z> read back; batch_64, 64 copies advanced together
  cpu_baseline       the exatn-mps step with a LAPACK SVD (numpy zgesdd) on the host cores at
                     chi = 256, all cores and one thread, CPU model recorded; beside it the
                     Jacobi oracle, the per-circuit oracle check of 16 uneven circuits (dense or
                     oracle MPS, error per result) and, under "noisy", the restated
                     purified-MPS algorithm on the same noisy circuit
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from tnqvm_amd import circuits as C  # noqa: E402
from tnqvm_amd import mps  # noqa: E402

METRIC = "2q-gate applies/s @χ=256 + p50/p99 circuit latency, mixed/batched vs plain FP64"
N_QUBITS = 40
CHI = 256
CUTOFF = 1e-10
PEAK_FP64_TFLOPS = 78.6  # MI355X dense FP64 (vector and MFMA)


def log(*a):
    print("[bench]", *a, file=sys.stderr, flush=True)


DEPTH = 40
SAT_MIN = 12          # full brickwork layers that saturate bonds 8..32 of the 40-qubit chain at 256
PAIR_LO, PAIR_HI = 8, 30  # pairs (i, i+1) whose theta' is (2 chi_i) x (2 chi_(i+2)) = 512 x 512 then
RANK_512 = 9          # engine histogram bucket of min(m, n) in [512, 1024)
NOMINAL_SVD = 704.0   # 4 (14 m n^2 + 8 n^3) at m = n = 2 chi, over chi^3
NOMINAL_APPLY = 736.0  # theta (32 chi^3) + nominal SVD; the gate term (128 chi^2) is below 1%


def headline_plan(warmup, steps):
    """Layers of the headline circuits: (saturation, warm, timed, tail) counts."""
    sat = max(SAT_MIN, DEPTH - 1 - warmup - steps)
    depth = max(DEPTH, sat + warmup + steps + 1)
    return sat, warmup, steps, depth - sat - warmup - steps


def interior(layer):
    """The gates of one brickwork layer on the pairs PAIR_LO..PAIR_HI (and the 1-qubit gates on
    their qubits, which fuse into them)."""
    keep2 = [g for g in layer if len(g["qubits"]) == 2 and PAIR_LO <= min(g["qubits"]) <= PAIR_HI]
    qs = {q for g in keep2 for q in g["qubits"]}
    return [g for g in layer if len(g["qubits"]) == 1 and g["qubits"][0] in qs] + keep2


def headline_circuit(seed, warmup, steps):
    sat, warm, timed, tail = headline_plan(warmup, steps)
    layers = C.brickwork_layers(C.haar_brickwork(N_QUBITS, sat + warm + timed + tail, seed=seed), N_QUBITS)
    out = [g for l in layers[:sat] for g in l]
    out += [g for l in layers[sat:sat + warm + timed] for g in interior(l)]
    out += [g for l in layers[sat + warm + timed:] for g in l]
    return out


def run_layers(eng, circuits, precision, warmup, steps):
    """Advance circuits together through the serving loop (one layer per step); time the
    `steps` interior layers after the saturation and warm layers."""
    sat, warm, timed, _ = headline_plan(warmup, steps)
    srv = mps.Server(eng)
    tr = mps.trunc_opts(cutoff=CUTOFF, max_bond=CHI, precision=precision)
    tickets = [srv.submit(N_QUBITS, c, tr) for c in circuits]
    for _ in range(sat + warm):
        srv.step()
    eng.profile(reset=True)
    step_ms = []
    t0 = time.perf_counter()
    for _ in range(timed):
        ts = time.perf_counter()
        srv.step()
        step_ms.append(1e3 * (time.perf_counter() - ts))
    total = time.perf_counter() - t0
    prof = eng.profile(reset=True)
    results = []
    for t in tickets:  # the tail layers and the measurements, untimed
        r = srv.poll(t)
        while r is None:
            srv.step()
            r = srv.poll(t)
        results.append(r)
    hist = prof["applies_by_rank"]
    applies = hist[RANK_512]
    return applies, sum(hist) - applies, total, step_ms, prof, results


def uneven_batch(count, seed):
    """Brickwork circuits of uneven size: qubits ~ U{16..48}, depth ~ U{10..60}, chi_max drawn
    from {32, 64, 128, 256}, Haar 2x2 and 4x4 gates, svd-cutoff 1e-10."""
    return [{"n": c["n"], "gates": c["gates"], "chi": c["chi"], "kind": "brickwork"}
            for c in C.mixed_batch(count, seed=seed)]


def serving_stream(brick, n_grid, grid_cycles, seed):
    """The serving workload: the given brickwork circuits mixed with 6x6 grid circuits
    (fSim(pi/2, pi/6) on the ABCDCDAB coupler patterns, sqrt(X)/sqrt(Y)/sqrt(W) single-qubit
    gates, snake map, routed to nearest neighbours) at chi_max 1024, in a seeded random order.
    Each job keeps its index in `brick` ("src", -1 for grid circuits)."""
    rng = np.random.default_rng(seed)
    jobs = [dict(j, src=k) for k, j in enumerate(brick)]
    for k in range(n_grid):
        g = C.route_nearest_neighbour(C.grid_fsim(6, 6, grid_cycles, seed=seed + 2 + k))
        jobs.append({"n": 36, "gates": g, "chi": 1024, "kind": "grid", "src": -1, "gid": k})
    return [jobs[i] for i in rng.permutation(len(jobs))]


def fifo_latency(arrivals, service):
    """Latency (ms) and span (s) of a single server that runs jobs one at a time in arrival
    order with the given service times (s): finish_k = max(arrival_k, finish_(k-1)) + service_k."""
    t, lat = 0.0, np.zeros(len(arrivals))
    for k, (a, s) in enumerate(zip(arrivals, service)):
        t = max(t, a) + s
        lat[k] = 1e3 * (t - a)
    return lat, t - arrivals[0]


def serve_batched(eng, jobs, arrivals, precision):
    """Continuous batching: every circuit whose arrival time (s, wall clock) has passed is
    submitted at the next step boundary and advances with all others in flight. Returns
    per-circuit latency (arrival -> result observed, ms), the span from the first arrival to
    the last result (s), and the results."""
    srv = mps.Server(eng)
    n = len(jobs)
    lat, res = np.zeros(n), [None] * n
    live, nxt, done = {}, 0, 0
    t0 = time.perf_counter()
    while done < n:
        now = time.perf_counter() - t0
        while nxt < n and arrivals[nxt] <= now:
            j = jobs[nxt]
            tk = srv.submit(j["n"], j["gates"], mps.trunc_opts(cutoff=CUTOFF, max_bond=j["chi"], precision=precision))
            live[tk] = nxt
            nxt += 1
        if not live:
            time.sleep(max(0.0, arrivals[nxt] - now))
            continue
        srv.step()
        now = time.perf_counter() - t0
        for tk in list(live):
            r = srv.poll(tk)
            if r is not None:
                k = live.pop(tk)
                lat[k] = 1e3 * (now - arrivals[k])
                res[k] = r
                done += 1
    return lat, time.perf_counter() - t0 - arrivals[0], res


def serve_plain(eng, jobs, arrivals):
    """The plain per-circuit FP64 path on the same arrivals: one circuit at a time in arrival
    order (complex128 everywhere, the bonds of one layer batched), <Z> of every qubit read
    back before the next circuit starts. Returns latency (ms), span (s), <Z> per circuit and
    each circuit's own run time (s)."""
    n = len(jobs)
    lat, ez, own = np.zeros(n), [None] * n, np.zeros(n)
    t0 = time.perf_counter()
    for k, j in enumerate(jobs):
        wait = arrivals[k] - (time.perf_counter() - t0)
        if wait > 0:
            time.sleep(wait)
        ts = time.perf_counter()
        s = mps.MpsState(eng, j["n"], cutoff=CUTOFF, max_bond=j["chi"], precision="fp64")
        ez[k] = s.run(j["gates"]).expval_z_all()
        s.close()
        own[k] = time.perf_counter() - ts
        lat[k] = 1e3 * (time.perf_counter() - t0 - arrivals[k])
    return lat, time.perf_counter() - t0 - arrivals[0], ez, own


def fp64_sensitivity(eng, job, ez_ref):
    """How far the plain FP64 <Z> moves when only the gate matrices are rounded to complex64
    (a relative perturbation of about 6e-8, the size of one complex64 rounding): the
    conditioning of the truncated result itself, for comparison with the mixed path's error."""
    gates = []
    for g in job["gates"]:
        m = mps.matrix_of(g).astype(np.complex64).astype(np.complex128)
        gates.append({"name": "unitary", "qubits": g["qubits"], "matrix": m})
    s = mps.MpsState(eng, job["n"], cutoff=CUTOFF, max_bond=job["chi"], precision="fp64")
    ez = s.run(gates).expval_z_all()
    s.close()
    return float(np.abs(ez - ez_ref).max())


def gate_at_a_time(eng, seed, warmup, steps, max_applies=40):
    """Plain FP64, one 2-qubit gate per wave (the reference visitor's gate-by-gate model):
    the circuit is brought through its saturation and warm layers normally, then 2-qubit
    gates of the timed (interior) layers are applied and flushed one at a time; a bounded
    sample of `max_applies` gates is timed."""
    sat, warm, timed, tail = headline_plan(warmup, steps)
    s = mps.MpsState(eng, N_QUBITS, cutoff=CUTOFF, max_bond=CHI, precision="fp64")
    layers = C.brickwork_layers(C.haar_brickwork(N_QUBITS, sat + warm + timed + tail, seed=seed), N_QUBITS)
    per_layer = layers[:sat] + [interior(l) for l in layers[sat:]]
    for l in range(sat + warm):
        s.run(per_layer[l])
    s.flush()
    done, t0 = 0, time.perf_counter()
    for l in range(sat + warm, sat + warm + timed):
        for g in per_layer[l]:
            s.run([g])
            if len(g["qubits"]) == 2:
                s.flush()
                done += 1
                if done >= max_applies:
                    break
        if done >= max_applies:
            break
    el = time.perf_counter() - t0
    s.close()
    return done, el


def cpu_info():
    model = None
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"nproc": os.cpu_count(), "model": model}


def cpu_baseline_lapack(threads, seconds=8.0):
    """The LAPACK-class CPU comparator (oracle/lapack_mps.py: the exatn-mps step with numpy's zgesdd
    for the SVD) at chi = 256, Haar gates on the two middle bonds of an 18-site chain whose bonds are
    min(2^k, 2^(18-k), 256), with the BLAS/LAPACK thread pool limited to `threads`."""
    from threadpoolctl import threadpool_limits

    from oracle.lapack_mps import LapackMps
    n = 18
    bonds = [min(2 ** min(k, n - k), CHI) for k in range(n + 1)]
    rng = np.random.default_rng(0xC0FFEE)
    m = LapackMps(n, cutoff=CUTOFF, max_bond=CHI)
    for k in range(n):
        a = (rng.standard_normal((bonds[k], 2, bonds[k + 1])) + 1j * rng.standard_normal((bonds[k], 2, bonds[k + 1])))
        m.set_site(k, a / np.sqrt(2 * bonds[k + 1]))
    applies = 0
    with threadpool_limits(limits=threads):
        t0 = time.perf_counter()
        while True:
            q = 8 + (applies % 2)
            m.apply_2q(q, q + 1, C.haar_unitary(rng, 4))
            applies += 1
            el = time.perf_counter() - t0
            if (el >= seconds and applies >= 2) or applies >= 400:
                break
    return {"value": applies / el, "unit": "2q-gate applies/s @chi=256", "cores": threads, "kind": "port",
            "sample": f"{applies} Haar 2q applies on the middle bonds (theta 512x512, chi_max 256) of an 18-site "
                      f"FP64 MPS, {el:.1f} s, oracle/lapack_mps.py (numpy zgesdd, {threads} BLAS thread(s))"}


def oracle_check_sample(sample, threads, budget_s=25.0):
    """Per-result error of the uneven batch against the FP64 oracle (oracle/mps_oracle.cpp, the same
    truncation), on the seeded sample chosen in the uneven block, cheapest first, within a time budget."""
    from oracle import oracle as O
    out, t0 = [], time.perf_counter()
    for rec in sample:
        if time.perf_counter() - t0 > budget_s:
            break
        job = rec["job"]
        m = O.Mps(job["n"], cutoff=CUTOFF, max_bond=job["chi"], nthreads=threads).run(job["gates"])
        z = np.array([m.expval_z_parity([q]) for q in range(job["n"])])
        out.append({"qubits": job["n"], "chi_max": job["chi"], "two_qubit_gates": C.count_2q(job["gates"]),
                    "fidelity_est": rec["fidelity_est"], "bond_equal": int(max(m.bonds())) == rec["max_bond"],
                    "z_error_mixed": float(np.abs(rec["ez_mixed"] - z).max()),
                    "z_error_plain_fp64": float(np.abs(rec["ez_fp64"] - z).max()) if rec["ez_fp64"] is not None else None,
                    "fp64_sensitivity": rec["fp64_sensitivity"]})
    return out, time.perf_counter() - t0


def cpu_baseline(seconds=12.0):
    """FP64 oracle (CPU restatement of the exatn-mps algorithm) at chi = 256: Haar gates on the
    two middle bonds of an 18-site chain whose bonds are min(2^k, 2^(18-k), 256)."""
    from oracle import oracle as O
    threads = min(os.cpu_count() or 1, 16)
    n = 18
    bonds = [min(2 ** min(k, n - k), CHI) for k in range(n + 1)]
    rng = np.random.default_rng(0xC0FFEE)
    m = O.Mps(n, cutoff=CUTOFF, max_bond=CHI, nthreads=threads)
    for k in range(n):
        a = (rng.standard_normal((bonds[k], 2, bonds[k + 1])) + 1j * rng.standard_normal((bonds[k], 2, bonds[k + 1])))
        m.set_site(k, a / np.sqrt(2 * bonds[k + 1]))
    applies, t0 = 0, time.perf_counter()
    while True:
        q = 8 + (applies % 2)
        m.apply({"name": "unitary", "qubits": (q, q + 1), "matrix": C.haar_unitary(rng, 4)})
        applies += 1
        el = time.perf_counter() - t0
        if (el >= seconds and applies >= 2) or applies >= 64:
            break
    return {"value": applies / el, "unit": "2q-gate applies/s @chi=256", "cores": threads, "kind": "port",
            "sample": f"{applies} Haar 2q applies on the middle bonds (theta 512x512, chi_max 256) of an "
                      f"18-site FP64 MPS, {el:.1f} s, oracle/mps_oracle.cpp with {threads} OpenMP threads"}


def qaoa_60(eng, precision):
    """60-qubit QAOA MaxCut on a seeded random 3-regular graph (90 edges), p = 4, routed by the
    swap network (7080 2-qubit applies), chi_max 512, svd-cutoff 1e-12, one circuit alone: wall
    time of the gates, the energy sum_edges <Z_i Z_j> unnormalised (the reference's exp-val
    semantics: no renormalisation after truncation) and divided by <psi|psi>."""
    gates, edges, pos, _ = C.qaoa_maxcut(60, 4, seed=0x7E5A0003)
    pairs = [(pos[i], pos[j]) for i, j in edges]
    eng.profile(reset=True)
    t0 = time.perf_counter()
    s = mps.MpsState(eng, 60, cutoff=1e-12, max_bond=512, precision=precision).run(gates)
    s.flush()
    dt = time.perf_counter() - t0
    zz = s.expval_zz(pairs)
    nrm = s.norm()
    st = s.stats()
    prof = eng.profile(reset=True)
    s.close()
    applies = C.count_2q(gates)
    e = float(zz.sum())
    return {"precision": precision, "seconds": round(dt, 3), "applies": applies,
            "applies_per_s": round(applies / dt, 2), "applies_theta_1024x1024": prof["applies_by_rank"][10],
            "energy_unnormalized": e, "norm": nrm, "energy": e / nrm if nrm > 0 else None,
            "fidelity_est": st["fidelity_est"], "discarded_weight": st["discarded_weight"],
            "max_bond": st["max_bond"], "jacobi_sweeps_max": st["jacobi_sweeps_max"]}


def oracle_cost(job):
    """Rough FP64 oracle cost of a brickwork circuit: 2-qubit gates x (its largest bond)^3."""
    return C.count_2q(job["gates"]) * float(min(job["chi"], 2 ** (job["n"] // 2))) ** 3


NOISY_N, NOISY_DEPTH = 40, 4


def noisy_workload():
    """The noisy circuit and its channels: [(gate, [(qubit, Kraus ops)])]."""
    from tnqvm_amd import noise as N
    with open(os.path.join(ROOT, "tests", "golden", "noise", "ibmqx2.json")) as f:
        model = N.NoiseModel.from_backend_json(f.read())
    ops = []
    for g in C.u3_cnot_line(NOISY_N, NOISY_DEPTH, seed=5):
        ops.append((g, [(q, model.one_qubit[q % 5]) for q in g["qubits"] if model.one_qubit.get(q % 5)]))
    return ops


def noisy_gpu(eng, ops, reps=5):
    """Median wall time of the noisy circuit on the device (channels queued after each wave of
    gates), <Z> of every qubit read back inside the timed region."""
    times, ez, kd, bonds = [], None, None, None
    for _ in range(reps + 1):  # the first run warms the arena
        s = mps.MpsState(eng, NOISY_N, cutoff=1e-9)
        t0 = time.perf_counter()
        for g, chans in ops:
            s.run([g])
            for q, k in chans:
                s.apply_kraus(q, k)
        ez = s.expval_z_all()
        times.append(time.perf_counter() - t0)
        kd, bonds = int(max(s.kraus_dims())), int(max(s.bonds()))
        s.close()
    return float(np.median(times[1:])), ez, kd, bonds


def noisy_batch(eng, ops, count=64):
    """The same noisy circuit `count` times, advanced together (mps.run_noisy_batch: the open
    waves of all of them run as one gate wave and one channel wave); wall time and circuits/s."""
    chans = {id(g): c for g, c in ops}

    class Model:
        def channels(self, g):
            return chans[id(g)]

    gates = [g for g, _ in ops]
    mps.run_noisy_batch(eng, [(NOISY_N, gates)] * 2, Model())  # warm the arena at batch shapes
    t0 = time.perf_counter()
    res = mps.run_noisy_batch(eng, [(NOISY_N, gates)] * count, Model())
    dt = time.perf_counter() - t0
    spread = max(float(np.abs(r["expval_z"] - res[0]["expval_z"]).max()) for r in res)
    return dt, spread


def cpu_baseline_noisy(ops, ez_gpu, seconds=3.0):
    """The restated purified-MPS algorithm (oracle/noise.py, numpy) on the same noisy circuit."""
    from threadpoolctl import threadpool_limits

    from oracle.noise import PurifiedMps
    runs = 0
    with threadpool_limits(limits=1):  # one core: the BLAS under numpy's SVDs single-threaded
        t0 = time.perf_counter()
        while True:
            pm = PurifiedMps(NOISY_N)
            for g, chans in ops:
                pm.gate(mps.matrix_of(g), g["qubits"])
                for q, k in chans:
                    pm.channel(q, k)
            runs += 1
            el = time.perf_counter() - t0
            if el >= seconds or runs >= 20:
                break
        ez = pm.expval_z()
    return {"value": runs / el, "unit": "noisy circuits/s", "cores": 1, "kind": "port",
            "sample": f"{runs} runs of the noisy circuit, {el:.1f} s, oracle/noise.py PurifiedMps (numpy, one BLAS thread)",
            "max_abs_dz_vs_gpu": float(np.abs(ez - ez_gpu).max())}


def measure_kernels(precision: str, amps: int = 1024):
    """Workload of the per-kernel counter runs of the memory-bound kernels: one 40-qubit Haar
    brickwork state brought to saturation (12 full layers, bonds 8..32 at chi 256), then a Haar
    1-qubit gate on every qubit and a flush (k_apply_1q on every site), `amps` bit-string
    amplitudes (k_chain on every site), <Z_i> of every qubit (environment GEMMs) and the norm."""
    eng = mps.Engine()
    rng = np.random.default_rng(0x7E5A0040)
    s = mps.MpsState(eng, N_QUBITS, cutoff=CUTOFF, max_bond=CHI, precision=precision)
    s.run(C.haar_brickwork(N_QUBITS, SAT_MIN, seed=0x7E5A0041))
    s.flush()
    one = [{"name": "unitary", "qubits": (q,), "matrix": C.haar_unitary(rng, 2)} for q in range(N_QUBITS)]
    t0 = time.perf_counter()
    s.run(one)
    s.flush()
    t1 = time.perf_counter()
    bits = C.random_bitstrings(N_QUBITS, amps, seed=0x7E5A0042)
    amp = s.amplitudes(bits)
    t2 = time.perf_counter()
    ez = s.expval_z_all()
    nrm = s.norm()
    t3 = time.perf_counter()
    out = {"precision": precision, "bonds": [int(b) for b in s.bonds()], "amplitudes": amps,
           "ms_1q_all_sites": round(1e3 * (t1 - t0), 3), "ms_amplitudes": round(1e3 * (t2 - t1), 3),
           "ms_expval_z_all_and_norm": round(1e3 * (t3 - t2), 3), "norm": nrm,
           "checksum": float(np.abs(amp).sum() + np.abs(ez).sum())}
    s.close()
    eng.close()
    print(json.dumps(out), flush=True)


def calibrate_copy(nbytes: int, reps: int = 4):
    """Plain device-to-device copies of a known size, for scaling rocprofv3 FETCH_SIZE /
    WRITE_SIZE readings (run under the profiler: bench.py --calibrate-copy BYTES)."""
    import ctypes
    hip = ctypes.CDLL("libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(a), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(ctypes.byref(b), ctypes.c_size_t(nbytes)) == 0
    assert hip.hipMemset(a, 1, ctypes.c_size_t(nbytes)) == 0
    for _ in range(reps):
        assert hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3) == 0  # device to device
    assert hip.hipDeviceSynchronize() == 0
    hip.hipFree(a)
    hip.hipFree(b)
    print(json.dumps({"calibrate_copy_bytes": nbytes, "reps": reps}), flush=True)


def serving_block(eng, args, brick, n_grid, rate, plain_own, plain_grid):
    """Poisson arrivals at `rate` of the brickwork circuits `brick` (indices into the uneven batch
    kept as "src") and `n_grid` 6x6 grid circuits through the continuous-batching loop, against
    the plain FP64 path one circuit at a time (FIFO over each circuit's measured plain run time;
    `plain_own` holds those of the uneven batch, `plain_grid` caches the grid circuits')."""
    jobs = serving_stream(brick, n_grid, args.serve_grid_cycles, 0x7E5A0010)
    arrivals = np.cumsum(np.random.default_rng(0x5E7).exponential(1.0 / rate, size=len(jobs)))
    arrivals -= arrivals[0]
    log(f"serving: {len(brick)} brickwork + {n_grid} 6x6 grid circuits, Poisson arrivals at {rate}/s")
    base = eng.tuning()
    eng.set_tuning(big_lane_min_chi=args.serve_big_lane)
    lat, span, _ = serve_batched(eng, jobs, arrivals, args.precision)
    eng.set_tuning(**base)
    grid_lat = [float(lat[k]) for k, j in enumerate(jobs) if j["kind"] == "grid"]
    brick_lat = np.array([lat[k] for k, j in enumerate(jobs) if j["kind"] != "grid"])
    log(f"  batched {args.precision}: {len(jobs) / span:.2f} circuits/s, p50 {np.percentile(lat, 50):.0f} ms, "
        f"p99 {np.percentile(lat, 99):.0f} ms, grid {grid_lat}")
    out = {"workload": f"Poisson arrivals of {len(brick)} of the uneven brickwork circuits and {n_grid} 6x6 grid "
                       f"fSim circuit(s) ({args.serve_grid_cycles} cycles, snake map, routed, chi_max 1024), "
                       "svd-cutoff 1e-10",
           "circuits": len(jobs), "brickwork": len(brick), "grid_6x6": n_grid,
           "arrival_rate_per_s": rate, "precision": args.precision, "big_lane_min_chi": args.serve_big_lane,
           "circuits_per_s": round(len(jobs) / span, 3),
           "latency_ms_p50": round(float(np.percentile(lat, 50)), 2),
           "latency_ms_p99": round(float(np.percentile(lat, 99)), 2),
           "grid_latency_ms": [round(x, 1) for x in grid_lat],
           "brickwork_latency_ms_p50": round(float(np.percentile(brick_lat, 50)), 2),
           "brickwork_latency_ms_p99": round(float(np.percentile(brick_lat, 99)), 2)}
    if n_grid and "grid_alone_ms" not in plain_grid:  # one grid circuit alone on the same path
        g0 = next(j for j in jobs if j["kind"] == "grid")
        t0 = time.perf_counter()
        st_ = mps.MpsState(eng, g0["n"], cutoff=CUTOFF, max_bond=g0["chi"], precision=args.precision).run(g0["gates"])
        st_.expval_z_all()
        st_.close()
        plain_grid["grid_alone_ms"] = round(1e3 * (time.perf_counter() - t0), 1)
    if n_grid:
        out["grid_alone_ms"] = plain_grid["grid_alone_ms"]
    if not args.no_plain:
        # the plain path serves one circuit at a time, so its latencies follow from each
        # circuit's measured plain run time through the FIFO recurrence
        service = np.zeros(len(jobs))
        for k, j in enumerate(jobs):
            key = ("grid", j["gid"]) if j["src"] < 0 else None
            if j["src"] >= 0 and plain_own is not None:
                service[k] = plain_own[j["src"]]
            elif key in plain_grid:
                service[k] = plain_grid[key]
            else:
                _, _, _, own = serve_plain(eng, [j], np.zeros(1))
                service[k] = own[0]
                if key is not None:
                    plain_grid[key] = own[0]
        plat, pspan = fifo_latency(arrivals, service)
        log(f"  plain fp64: {len(jobs) / pspan:.2f} circuits/s, p50 {np.percentile(plat, 50):.0f} ms, "
            f"p99 {np.percentile(plat, 99):.0f} ms")
        out["plain_fp64"] = {"circuits_per_s": round(len(jobs) / pspan, 3),
                             "latency_ms_p50": round(float(np.percentile(plat, 50)), 2),
                             "latency_ms_p99": round(float(np.percentile(plat, 99)), 2),
                             "method": "FIFO recurrence over the measured plain FP64 run time of each circuit (one "
                                       "circuit at a time)"}
        out["speedup_vs_plain"] = round(pspan / span, 3)
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20, help="timed interior layers (every apply chi = 256)")
    ap.add_argument("--warmup", type=int, default=7,
                    help="untimed interior layers after the (at least 12) saturation layers")
    ap.add_argument("--batch", type=int, default=8, help="circuits advanced together")
    ap.add_argument("--precision", default="mixed", choices=["mixed", "fp64"])
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-uneven", action="store_true")
    ap.add_argument("--no-noisy", action="store_true")
    ap.add_argument("--no-qaoa60", action="store_true")
    ap.add_argument("--oracle-sample", type=int, default=16,
                    help="uneven-batch circuits (the cheapest) checked against the FP64 oracle")
    ap.add_argument("--uneven-circuits", type=int, default=256, help="circuits of the uneven batch")
    ap.add_argument("--serve-circuits", type=int, default=256, help="brickwork circuits in the serving stream")
    ap.add_argument("--serve-grid", type=int, default=2, help="6x6 grid circuits in the serving stream")
    ap.add_argument("--serve-grid-cycles", type=int, default=4)
    ap.add_argument("--serve-rate", type=float, default=40.0,
                    help="Poisson arrival rate, circuits/s (above both paths' capacity: sustained throughput)")
    ap.add_argument("--serve-rate-moderate", type=float, default=5.0,
                    help="arrival rate of the second serving block (96 brickwork + 1 grid circuit)")
    ap.add_argument("--serve-big-lane", type=int, default=512,
                    help="serving: circuits whose bond can reach this run on the background lane (0 = off)")
    ap.add_argument("--lane-min-pairs", type=int, default=-1,
                    help="block pairs from which a wave's Jacobi rounds use several streams (default: engine default)")
    ap.add_argument("--lq-min-rank", type=int, default=-1,
                    help="Jacobi on the double-LQ factor for min(m,n) >= this (0 off; default: engine default)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="BYTES",
                    help="only run plain device copies of BYTES (profiler counter calibration)")
    ap.add_argument("--noisy-only", action="store_true",
                    help="only the noisy block (for profiling the purified-MPS path)")
    ap.add_argument("--measure-kernels", action="store_true",
                    help="only the workload of the memory-bound kernels (1q gates, amplitudes, <Z>) on a "
                         "saturated chi=256 state, for counter runs")
    args = ap.parse_args()
    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)
        return
    if args.measure_kernels:
        measure_kernels(args.precision)
        return
    if args.noisy_only:
        eng = mps.Engine()
        nt, _, kd, bonds = noisy_gpu(eng, noisy_workload())
        eng.close()
        print(json.dumps({"noisy_ms_per_circuit": round(nt * 1e3, 2), "max_kraus": kd, "max_bond": bonds}), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); independent circuits shard "
              "across GPUs with no communication, so the per-GPU figure at N=1 is the measurement.",
              file=sys.stderr)
        sys.exit(2)
    sat, warm, timed, tail = headline_plan(args.warmup, args.steps)
    depth = sat + warm + timed + tail
    eng = mps.Engine()
    if args.lq_min_rank >= 0:
        eng.set_lq_precondition(args.lq_min_rank)
    if args.lane_min_pairs >= 0:
        eng.set_tuning(lane_min_pairs=args.lane_min_pairs)
    eng.set_timing(True)

    log(f"batched {args.precision}: {args.batch} x 40-qubit depth-{depth} Haar brickwork, chi_max {CHI}; "
        f"{sat} saturation + {warm} warm layers, {timed} timed interior layers")
    seeds = [0x7E5A0002 + k for k in range(args.batch)]
    circuits = [headline_circuit(sd, args.warmup, args.steps) for sd in seeds]
    applies, other, total, step_ms, prof, res = run_layers(eng, circuits, args.precision, args.warmup, args.steps)
    value = applies / total
    svd_ms = prof["ms_lq"] + prof["ms_jacobi"]
    log(f"  {applies} applies at theta 512x512 (+{other} other) in {total:.3f} s -> {value:.1f} applies/s; "
        f"theta {prof['ms_theta']:.0f} ms, lq {prof['ms_lq']:.0f} ms, jacobi {prof['ms_jacobi']:.0f} ms "
        f"(rounds {prof['ms_jacobi_rounds']:.0f} ms, {prof['jacobi_launches']} launches), split {prof['ms_split']:.0f} ms")

    # roofline of the SVD (double LQ + Jacobi rounds + reflectors back), the dominant phase:
    # nominal 704 chi^3 flops per chi = 256 apply (the Golub-Reinsch count for a 512 x 512
    # complex SVD, fixed by convention so that fewer executed flops count as faster) over the
    # wall time of the timed steps.  The circuits run in two wave groups whose phases overlap
    # (one group's LQ beside the other's Jacobi rounds), so the HIP-event phase times, summed
    # over the groups, exceed the wall time; each serving step returns only after both groups'
    # streams are synchronized, so the host clock around the steps brackets all of their work.
    svd_nominal = applies * NOMINAL_SVD * CHI ** 3
    svd_tflops = svd_nominal / max(1e3 * total, 1e-9) / 1e9
    waves = max(prof["waves"], 1)
    groups = eng.tuning()["wave_groups"]
    # executed flops of k_jacobi_round (counted on the device) over the span of its launches,
    # per group (HIP events on each group's engine stream), summed over the groups
    jac_exec_tflops = prof["jacobi_flops"] / max(prof["ms_jacobi_rounds"], 1e-9) / 1e9
    avg_launch_us = 1e3 * prof["ms_jacobi_rounds"] / max(prof["jacobi_launches"], 1)
    # HBM bytes per k_jacobi_round launch: per-block-pair FETCH/WRITE from the committed rocprofv3
    # counter runs times the block pairs this run launched per dispatch; against the algorithmic
    # bytes of the SVDs (read and write the 512 x 512 complex128 matrix once: 2 m n 16)
    traffic, traffic_ratio = None, None
    tpath = os.path.join(ROOT, "profiles", "k_jacobi_round_traffic.json")
    if os.path.exists(tpath):
        tj = json.load(open(tpath))
        per_pair = tj["fetch_bytes_per_pair"] + tj["write_bytes_per_pair"]
        traffic = round(per_pair * prof["jacobi_pairs_visited"] / max(prof["jacobi_launches"], 1))
        traffic_ratio = round(per_pair * prof["jacobi_pairs_visited"] / max(1.0, applies * 2 * 512 * 512 * 16), 2)
    roofline = {"bound": "mfma", "achieved": round(svd_tflops, 3), "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                "frac": round(svd_tflops / PEAK_FP64_TFLOPS, 4), "traffic": traffic,
                "kernel": "SVD of theta' 512x512: double LQ (k_lq_factor_lds/k_lq_next/k_lq_rest with look-ahead, "
                          "k_lq_panel0/k_lq_update), k_jacobi_round, k_lq_apply_q",
                "flops_per_unit": NOMINAL_SVD * CHI ** 3, "unit_of_work": "one chi=256 apply's SVD (nominal 704 chi^3)",
                "wave_groups": groups, "svd_event_ms": round(svd_ms, 2),
                "avg_wave_ms": round(svd_ms / waves, 3), "svds_per_wave": round(applies / waves, 1),
                "timing": "host clock around the timed serving steps (each returns after both wave groups' "
                          "streams are synchronized); svd_event_ms: HIP events around the LQ and Jacobi phases "
                          "of each wave on its group's engine stream, summed over the groups, which overlap",
                "jacobi_round": {"executed_tflops": round(jac_exec_tflops, 3),
                                 "executed_over_nominal": round(prof["jacobi_flops"] / max(svd_nominal, 1.0), 3),
                                 "avg_launch_us": round(avg_launch_us, 2), "launches": prof["jacobi_launches"],
                                 "timing": "span of each sweep's launches on its group's streams (HIP events), "
                                           "summed over the groups, per launch",
                                 "traffic_bytes_per_launch": traffic, "traffic_over_algorithmic": traffic_ratio},
                "precision_note": "mixed path: the Gram of the first 5 sweeps of large SVDs on FP32 MFMA; rotations, "
                                  "row updates and the final sweeps on FP64 MFMA; peak is the FP64 one"}
    nominal_flops = NOMINAL_APPLY * CHI ** 3
    out = {
        "metric": METRIC, "value": round(value, 2), "unit": "2q-gate applies/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(1e3 * total / max(timed, 1), 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "c64+f64" if args.precision == "mixed" else "c128",
        "data": "synthetic",
        "config": {"workload": f"{args.batch} x 40-qubit Haar brickwork, depth {depth}, chi_max {CHI}, svd-cutoff "
                               f"{CUTOFF:g}; layers 0-{sat - 1} full (bonds 8-32, 25 of 39, saturate at 256), then "
                               f"{warm} warm + {timed} timed layers of the brickwork's gates on pairs "
                               f"{PAIR_LO}-{PAIR_HI} only (every timed apply theta 512x512), {tail} full tail layer(s)",
                   "batch": args.batch, "precision": args.precision,
                   "precision_detail": "complex64 sites and theta (FP32 MFMA), FP64 Jacobi SVD, norms and weights"
                   if args.precision == "mixed" else "complex128 everywhere"},
        "timed_applies": {"theta_512x512": applies, "other": other},
        "nominal_tflops": round(value * nominal_flops / 1e12, 3),
        "nominal_frac_fp64": round(value * nominal_flops / 1e12 / PEAK_FP64_TFLOPS, 4),
        "fidelity_est": float(np.mean([r["fidelity_est"] for r in res])),
        "roofline": roofline,
        "phase_ms": {k: round(prof[k], 2) for k in ("ms_theta", "ms_lq", "ms_jacobi", "ms_split")},
        "jacobi_sweeps_per_svd_chi256": round(prof["sweeps_large"] / max(1, prof["svds_large"]), 2),
    }
    if not args.no_plain:
        log("plain fp64, one circuit at a time (bonds of a layer batched)")
        pa, po, pt, _, pprof, _ = run_layers(eng, circuits[:1], "fp64", args.warmup, args.steps)
        out["plain_fp64_value"] = round(pa / pt, 2)
        out["speedup_vs_plain"] = round(value / (pa / pt), 3)
        log(f"  {pa} applies (+{po} other) in {pt:.3f} s -> {pa / pt:.1f} applies/s")
        log("plain fp64, one circuit, one 2-qubit gate at a time")
        ga, gt = gate_at_a_time(eng, seeds[0], args.warmup, args.steps)
        out["plain_fp64_gate_at_a_time_value"] = round(ga / gt, 2)
        out["speedup_vs_gate_at_a_time"] = round(value / (ga / gt), 3)
        log(f"  {ga} applies in {gt:.3f} s -> {ga / gt:.1f} applies/s")
    plain_own = None  # plain FP64 run time (s) of each uneven-batch circuit
    oracle_sample = []  # uneven-batch results to check against the FP64 oracle in the cpu_baseline leg
    uneven = uneven_batch(args.uneven_circuits, 0x7E5A0004)
    if not args.no_uneven:
        jobs = uneven
        applies_u = sum(C.count_2q(j["gates"]) for j in jobs)
        log(f"uneven batch: {len(jobs)} brickwork circuits ({applies_u} 2q gates) submitted together, "
            f"{args.precision}")
        _, bspan, bres = serve_batched(eng, jobs, np.zeros(len(jobs)), args.precision)
        log(f"  {len(jobs) / bspan:.2f} circuits/s, {applies_u / bspan:.1f} applies/s")
        ub = {"workload": "Haar brickwork, qubits U{16..48}, depth U{10..60}, chi_max from {32,64,128,256}, "
                          "svd-cutoff 1e-10, all submitted at once",
              "circuits": len(jobs), "applies": applies_u, "precision": args.precision, "span_s": round(bspan, 3),
              "circuits_per_s": round(len(jobs) / bspan, 3), "applies_per_s": round(applies_u / bspan, 2)}
        if not args.no_plain:
            log("  plain fp64, each circuit alone")
            _, _, pez, plain_own = serve_plain(eng, jobs, np.zeros(len(jobs)))
            ptot = float(plain_own.sum())
            log(f"  plain {len(jobs) / ptot:.2f} circuits/s -> x{ptot / bspan:.2f}")
            dz = np.array([np.abs(r["expval_z"] - z).max() for r, z in zip(bres, pez)])
            fid = np.array([r["fidelity_est"] for r in bres])
            worst = int(np.argmax(dz))
            ub["plain_fp64"] = {"span_s": round(ptot, 3), "circuits_per_s": round(len(jobs) / ptot, 3),
                                "applies_per_s": round(applies_u / ptot, 2)}
            ub["speedup_vs_plain"] = round(ptot / bspan, 3)
            ub["z_error_vs_plain_fp64"] = {
                "max": float(dz.max()), "median": float(np.median(dz)),
                "circuits_within_1e-5": int((dz <= 1e-5).sum()),
                "worst_circuit": {"qubits": jobs[worst]["n"], "chi_max": jobs[worst]["chi"],
                                  "fidelity_est": float(fid[worst]),
                                  "fp64_sensitivity": fp64_sensitivity(eng, jobs[worst], pez[worst])}}
        # the cheapest circuits for the FP64 oracle, checked in the cpu_baseline leg against the oracle
        order = sorted(range(len(jobs)), key=lambda k: oracle_cost(jobs[k]))[:args.oracle_sample]
        for k in order:
            oracle_sample.append({"job": jobs[k], "ez_mixed": bres[k]["expval_z"], "max_bond": bres[k]["max_bond"],
                                  "fidelity_est": bres[k]["fidelity_est"],
                                  "ez_fp64": pez[k] if not args.no_plain else None,
                                  "fp64_sensitivity": fp64_sensitivity(eng, jobs[k], pez[k]) if not args.no_plain
                                  else None})
        out["uneven_batch"] = ub
    if not args.no_serving:
        plain_grid = {}  # plain FP64 run time of each grid circuit, by seed
        serving = serving_block(eng, args, uneven[:args.serve_circuits], args.serve_grid, args.serve_rate,
                                plain_own, plain_grid)
        out["serving"] = serving
        # the same loop below capacity: 96 of the brickwork circuits and one grid circuit at a rate
        # the batched path sustains, so that latency is service time rather than queueing
        moderate = serving_block(eng, args, uneven[:96], min(1, args.serve_grid), args.serve_rate_moderate,
                                 plain_own, plain_grid)
        out["serving_moderate"] = moderate
        out["latency_ms_p50"] = moderate["latency_ms_p50"]
        out["latency_ms_p99"] = moderate["latency_ms_p99"]
    if not args.no_qaoa60:
        log("qaoa60: 60-qubit QAOA MaxCut p=4, swap network, chi_max 512, cutoff 1e-12")
        q = {}
        for prec in ("mixed", "fp64"):
            q[prec] = qaoa_60(eng, prec)
            log(f"  {prec}: {q[prec]['seconds']:.1f} s, {q[prec]['applies_per_s']:.1f} applies/s, energy "
                f"{q[prec]['energy']} (unnormalised {q[prec]['energy_unnormalized']:.3e}), fidelity est "
                f"{q[prec]['fidelity_est']:.2e}")
        em, ef = q["mixed"]["energy"], q["fp64"]["energy"]
        out["qaoa_60"] = {"workload": "60-qubit QAOA MaxCut, seeded random 3-regular graph (90 edges), p=4, gamma, "
                                      "beta ~ U[0, pi], odd-even swap network (7080 2q applies), chi_max 512, "
                                      "svd-cutoff 1e-12, one circuit alone",
                          "mixed": q["mixed"], "fp64": q["fp64"],
                          "energy_diff_mixed_vs_fp64": abs(em - ef) if em is not None and ef is not None else None,
                          "speedup_mixed_vs_fp64": round(q["fp64"]["seconds"] / q["mixed"]["seconds"], 3)}
    noisy_ops, noisy_ez = None, None
    if not args.no_noisy:
        noisy_ops = noisy_workload()
        nt, noisy_ez, kd, bonds = noisy_gpu(eng, noisy_ops)
        nch = sum(len(c) for _, c in noisy_ops)
        log(f"noisy: {len(noisy_ops)} gates, {nch} channels, {nt * 1e3:.1f} ms per circuit (Kraus legs up to {kd}, "
            f"bonds up to {bonds})")
        out["noisy"] = {"workload": f"{NOISY_N}-qubit u3/CNOT line, depth {NOISY_DEPTH}, ibmqx2 single-qubit "
                                    "channels after every gate (backend qubit q mod 5), cut 1e-9, no bond cap",
                        "gates": len(noisy_ops), "channels": nch, "ms_per_circuit": round(nt * 1e3, 2),
                        "circuits_per_s": round(1.0 / nt, 3), "max_kraus": kd, "max_bond": bonds}
        bt, spread = noisy_batch(eng, noisy_ops)
        log(f"noisy batch: 64 circuits together, {bt:.2f} s, {64 / bt:.1f} circuits/s")
        out["noisy"]["batch_64"] = {"seconds": round(bt, 3), "circuits_per_s": round(64 / bt, 2),
                                    "max_abs_dz_across_copies": spread}
    eng.close()
    if not args.no_cpu_baseline:
        threads = min(os.cpu_count() or 1, 16)
        log(f"cpu baseline: LAPACK zgesdd step on {threads} cores, then 1 core; Jacobi oracle")
        out["cpu_baseline"] = cpu_baseline_lapack(threads)
        out["cpu_baseline"]["cpu"] = cpu_info()
        out["cpu_baseline"]["lapack_1_thread"] = cpu_baseline_lapack(1, seconds=6.0)
        out["cpu_baseline"]["jacobi_oracle"] = cpu_baseline(seconds=6.0)
        if oracle_sample:
            log(f"cpu baseline: uneven-batch results against the FP64 oracle ({len(oracle_sample)} cheapest)")
            checks, el = oracle_check_sample(oracle_sample, threads, budget_s=20.0)
            zm = [c["z_error_mixed"] for c in checks]
            out["cpu_baseline"]["uneven_oracle_check"] = {
                "circuits": len(checks), "seconds": round(el, 2), "oracle": "oracle/mps_oracle.cpp FP64, same truncation",
                "z_error_mixed_max": max(zm) if zm else None,
                "bonds_equal": all(c["bond_equal"] for c in checks), "per_circuit": checks}
        if noisy_ops is not None:
            log("cpu baseline, noisy (purified-MPS oracle)")
            out["cpu_baseline"]["noisy"] = cpu_baseline_noisy(noisy_ops, noisy_ez)
            cpu_rate = out["cpu_baseline"]["noisy"]["value"]
            out["noisy"]["speedup_vs_cpu"] = round(out["noisy"]["circuits_per_s"] / cpu_rate, 2)
            out["noisy"]["batch_64"]["speedup_vs_cpu"] = round(out["noisy"]["batch_64"]["circuits_per_s"] / cpu_rate, 2)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
