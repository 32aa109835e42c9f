"""Autotuner for the batched SVD's launch and precision tunables (hipmps_tuning).

For a workload shape (qubits, bond cap chi, circuits per wave) it runs Haar brickwork circuits
through the continuous-batching loop: warm-up layers grow the bonds to chi, then timed layers
run at chi.  Every candidate tuning runs the same circuits, and the fastest is set on the
engine.  The candidates cover:
  jacobi_lanes   1, 2, 4 streams for the Jacobi rounds and LQ panels of large waves
  gram32_sweeps  0, 3, 5, 7 leading sweeps with an FP32-MFMA Gram (mixed states only)
  lq_min_rank    0 (no preconditioning) or 64 (double LQ for min(m, n) >= 64)
  wave_groups    1 or 2 engines whose waves advance concurrently

    python -m tnqvm_amd.autotune [--qubits 40] [--chi 256] [--batch 8] [--warmup 12] [--steps 6]
prints one JSON object: the table of candidates (applies/s, Jacobi sweeps per SVD) and the
choice.  The engine's defaults are the choice on MI355X at the headline shape.
"""
from __future__ import annotations

import argparse
import itertools
import json
import sys
import time

from . import circuits as C
from . import mps


def time_tuning(eng, circuits, n, chi, warmup, steps, precision="mixed"):
    srv = mps.Server(eng)
    tr = mps.trunc_opts(cutoff=1e-10, max_bond=chi, precision=precision)
    tickets = [srv.submit(n, c, tr) for c in circuits]
    for _ in range(warmup):
        srv.step()
    eng.profile(reset=True)
    t0 = time.perf_counter()
    for _ in range(steps):
        srv.step()
    el = time.perf_counter() - t0
    prof = eng.profile(reset=True)
    for t in tickets:  # drain
        while srv.poll(t) is None:
            srv.step()
    applies = sum(len(range(l % 2, n - 1, 2)) for l in range(warmup, warmup + steps)) * len(circuits)
    sweeps = prof["sweeps_large"] / max(1, prof["svds_large"])
    return applies / el, sweeps


def tune(eng, n=40, chi=256, batch=8, warmup=12, steps=6, lanes=(1, 2, 4), gram32=(0, 3, 5, 7), lq=(0, 64),
         groups=(1, 2), precision="mixed", log=None):
    depth = warmup + steps
    circuits = [C.haar_brickwork(n, depth, seed=0x7E5A0A00 + k) for k in range(batch)]
    base = eng.tuning()
    rows = []
    for jl, g, q, wg in itertools.product(lanes, gram32, lq, groups):
        eng.set_tuning(jacobi_lanes=jl, gram32_sweeps=g, lq_min_rank=q, wave_groups=wg)
        rate, sweeps = time_tuning(eng, circuits, n, chi, warmup, steps, precision)
        rows.append({"jacobi_lanes": jl, "gram32_sweeps": g, "lq_min_rank": q, "wave_groups": wg,
                     "applies_per_s": round(rate, 2), "sweeps_per_svd": round(sweeps, 2)})
        if log:
            log(rows[-1])
    best = max(rows, key=lambda r: r["applies_per_s"])
    eng.set_tuning(jacobi_lanes=best["jacobi_lanes"], gram32_sweeps=best["gram32_sweeps"],
                   lq_min_rank=best["lq_min_rank"], wave_groups=best["wave_groups"],
                   lane_min_pairs=base["lane_min_pairs"])
    return {"workload": {"qubits": n, "chi": chi, "batch": batch, "warmup_layers": warmup, "timed_layers": steps,
                         "precision": precision},
            "candidates": rows, "choice": best}


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--qubits", type=int, default=40)
    ap.add_argument("--chi", type=int, default=256)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--warmup", type=int, default=12)
    ap.add_argument("--steps", type=int, default=6)
    ap.add_argument("--precision", default="mixed", choices=["mixed", "fp64"])
    a = ap.parse_args(argv)
    eng = mps.Engine()
    eng.set_timing(True)
    out = tune(eng, a.qubits, a.chi, a.batch, a.warmup, a.steps, precision=a.precision,
               log=lambda r: print("[autotune]", json.dumps(r), file=sys.stderr, flush=True))
    eng.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
