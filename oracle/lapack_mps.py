"""LAPACK-class FP64 restatement of the exatn-mps two-qubit step (numpy, zgesdd).

TEST INFRASTRUCTURE ONLY: used by tests/ (pinned against the C++ oracle and the dense state
vector) and by the ``cpu_baseline`` leg of bench.py, where it is the CPU comparator with the SVD
a LAPACK routine -- as ExaTN's split is (tnqvm/visitors/exatn-mps/ExaTnMpsVisitor.cpp:1623).
The product library never imports it.

The step follows ExaTnMpsVisitor.cpp and is the same as oracle/mps_oracle.cpp's mps_apply_2q
with the Jacobi replaced by numpy.linalg.svd (LAPACK zgesdd):
  merge of the two sites                                   :1387-1440
  gate contraction, first operand = MSB, |x| < 1e-100 -> 0  :1442-1524, :141-161
  full SVD, sqrt(S) into both factors, new bond min(m, n)   :1555-1627
  partial norms of each factor along the bond (squared, or
  2-norms), first index below the cutoff kept (i + 1),
  max-bond-dim clamp, no renormalisation                    :2366-2536
"""
from __future__ import annotations

import numpy as np

DBL_MIN = np.finfo(np.float64).tiny


class LapackMps:
    """MPS of n qubits, sites A[k] of shape (chi_k, 2, chi_k+1), |0..0> with bonds 1."""

    def __init__(self, n: int, cutoff: float = DBL_MIN, max_bond: int = 0, pn_l2: bool = False):
        self.n = n
        self.cutoff = cutoff
        self.max_bond = max_bond if max_bond > 0 else 0x7FFFFFFE
        self.pn_l2 = pn_l2
        self.discarded = 0.0
        self.A = []
        for _ in range(n):
            a = np.zeros((1, 2, 1), dtype=np.complex128)
            a[0, 0, 0] = 1.0
            self.A.append(a)

    def bonds(self):
        return [1] + [a.shape[2] for a in self.A]

    def set_site(self, k: int, a):
        self.A[k] = np.ascontiguousarray(np.asarray(a, dtype=np.complex128))

    def apply_1q(self, q: int, m):
        # new[l][b][r] = sum_i M[b][i] A[l][i][r]  (:1246 with G(i, b) = M[b][i])
        self.A[q] = np.einsum("bi,lir->lbr", np.asarray(m, dtype=np.complex128), self.A[q])

    def apply_2q(self, q0: int, q1: int, m):
        if abs(q0 - q1) != 1:
            raise ValueError("neighbouring qubits only (ExaTnMpsVisitor.cpp:1397)")
        i = min(q0, q1)
        g = np.asarray(m, dtype=np.complex128).reshape(2, 2, 2, 2)  # [row a, row b, col a', col b']
        if q0 != i:  # the first operand is the most significant bit: reorder to (left, right) site bits
            g = g.transpose(1, 0, 3, 2)
        ai, aj = self.A[i], self.A[i + 1]
        cl, cr = ai.shape[0], aj.shape[2]
        theta = np.tensordot(ai, aj, axes=([2], [0]))  # l, a', b', r
        tp = np.einsum("abxy,lxyr->labr", g, theta)
        tp[np.abs(tp) < 1e-100] = 0.0
        mat = tp.reshape(2 * cl, 2 * cr)
        u, s, vh = np.linalg.svd(mat, full_matrices=False)  # LAPACK zgesdd
        rs = np.sqrt(s)
        left = u * rs[None, :]
        right = rs[:, None] * vh
        ln = np.sum(np.abs(left) ** 2, axis=0)
        rn = np.sum(np.abs(right) ** 2, axis=1)
        if self.pn_l2:
            ln, rn = np.sqrt(ln), np.sqrt(rn)
        k0 = s.size
        below = np.nonzero((ln < self.cutoff) & (rn < self.cutoff))[0]
        k = int(below[0]) + 1 if below.size else k0
        k = min(k, self.max_bond)
        self.discarded += float(np.sum(s[k:] ** 2))
        self.A[i] = np.ascontiguousarray(left[:, :k].reshape(cl, 2, k))
        self.A[i + 1] = np.ascontiguousarray(right[:k, :].reshape(k, 2, cr))
        return k

    def run(self, gates, gate_matrix):
        """gates: dicts with "qubits" and "matrix" or "name"/"params"; gate_matrix(g) gives the matrix."""
        for g in gates:
            if g["name"] in ("Measure", "I"):
                continue
            q = g["qubits"]
            if len(q) == 1:
                self.apply_1q(q[0], gate_matrix(g))
            else:
                self.apply_2q(q[0], q[1], gate_matrix(g))
        return self

    def state_vector(self):
        """psi[x], qubit k = bit k of x (small n only)."""
        v = self.A[0].reshape(2, -1)  # [p0, r]
        for k in range(1, self.n):
            a = self.A[k]
            v = np.tensordot(v, a, axes=([v.ndim - 1], [0]))  # [..., p_k, r]
            v = v.reshape(-1, a.shape[2])
        psi = v.reshape([2] * self.n)  # index order p0, p1, ..., p_{n-1}
        return psi.transpose(list(range(self.n))[::-1]).reshape(-1)
