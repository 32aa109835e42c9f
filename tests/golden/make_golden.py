"""Writes the golden fixtures under tests/golden/.

* svd_spectra.npz -- seeded complex matrices and their singular values from LAPACK
  (numpy.linalg.svd -> zgesdd).  Shapes cover square, tall, wide and graded spectra
  (singular values from 1 down to 1e-14) so that the Jacobi SVDs of the oracle and of
  the HIP library are pinned by an independent library routine.
* kat.json -- known-answer values taken from the tnqvm tests:
    deuteron ansatz <X0 X1> table, tnqvm/visitors/exatn-mps/tests/MpsGateTester.cpp:375-396
    (angles = linspace(-pi, pi, 20), :398)
  plus the closed forms the reference tests assert (Rx / U3 / CPhase: cos(theta),
  ITensorMPSVisitorTester.cpp:310-392; GHZ amplitude 1/sqrt(2); conjugate-circuit
  amplitude -1/sqrt(2), ITensorMPSVisitorTester.cpp:522-551).

Run:  python tests/golden/make_golden.py
"""
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    rng = np.random.default_rng(0x5EED)
    mats, svals, names = {}, {}, []
    shapes = [(2, 2), (4, 4), (8, 3), (3, 8), (16, 16), (32, 32), (64, 64), (96, 128), (128, 96)]
    for (m, n) in shapes:
        a = (rng.standard_normal((m, n)) + 1j * rng.standard_normal((m, n))) / np.sqrt(2)
        key = f"gauss_{m}x{n}"
        mats[key] = a
        svals[key] = np.linalg.svd(a, compute_uv=False)
        names.append(key)
    # graded spectra: U diag(s) V^H with s = logspace(0, -14)
    for (m, n) in [(16, 16), (64, 64), (64, 40)]:
        k = min(m, n)
        u, _ = np.linalg.qr(rng.standard_normal((m, m)) + 1j * rng.standard_normal((m, m)))
        v, _ = np.linalg.qr(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
        s = np.logspace(0, -14, k)
        a = (u[:, :k] * s[None, :]) @ v[:, :k].conj().T
        key = f"graded_{m}x{n}"
        mats[key] = a
        svals[key] = np.linalg.svd(a, compute_uv=False)
        names.append(key)
    # exactly rank-deficient: rank 3 in 16x16
    b = (rng.standard_normal((16, 3)) + 1j * rng.standard_normal((16, 3))) @ \
        (rng.standard_normal((3, 16)) + 1j * rng.standard_normal((3, 16)))
    mats["rank3_16x16"] = b
    svals["rank3_16x16"] = np.linalg.svd(b, compute_uv=False)
    names.append("rank3_16x16")

    out = {}
    for k in names:
        out["A_" + k] = mats[k]
        out["S_" + k] = svals[k]
    np.savez_compressed(os.path.join(HERE, "svd_spectra.npz"), **out)

    kat = {
        "source": {
            "deuteron": "tnqvm/visitors/exatn-mps/tests/MpsGateTester.cpp:375-396",
            "rotation_cos": "tnqvm/visitors/itensor/mps/tests/ITensorMPSVisitorTester.cpp:310-392",
            "conj_amplitude": "tnqvm/visitors/itensor/mps/tests/ITensorMPSVisitorTester.cpp:522-551",
        },
        "deuteron_angles": list(np.linspace(-np.pi, np.pi, 20)),
        "deuteron_x0x1": [0.0, -0.324699, -0.614213, -0.837166, -0.9694, -0.996584, -0.915773,
                          -0.735724, -0.475947, -0.164595, 0.164595, 0.475947, 0.735724,
                          0.915773, 0.996584, 0.9694, 0.837166, 0.614213, 0.324699, 0.0],
        "deuteron_tol": 0.05,
        "conj_amplitude_ghz_xxx": -1.0 / np.sqrt(2.0),
    }
    with open(os.path.join(HERE, "kat.json"), "w") as f:
        json.dump(kat, f, indent=1)


if __name__ == "__main__":
    main()
