"""Noisy-simulation oracle: a dense density matrix, and a restatement of the reference's
locally purified MPS (tnqvm/visitors/exatn-mpo/ExaTnPmpsVisitor.cpp), both in numpy.

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker, and by the cpu_baseline leg of
bench.py (PurifiedMps timed on the host as the noisy path's CPU baseline, and its <Z> compared
with the device's).  The product library (tnqvm_amd / libhipmps) never imports this module.

Conventions, as in the reference: qubit k is bit k of a basis index (LSB = qubit 0); the first
operand of a 2-qubit gate is the most significant bit of its matrix index (Gates.hpp:237-246).

* DensityMatrix: rho (2^n x 2^n) under rho -> U rho U^H and rho -> sum_o K_o rho K_o^H.
  Pinned by closed forms in tests/test_oracle_noise.py (amplitude damping of |1>, dephasing
  of |+>, depolarizing of a Bell pair, trace preservation, agreement with the pure-state
  oracle for unitary circuits).
* PurifiedMps.expval_z: Tr(rho Z_q) through environments, pinned against the dense route
  (as_density().z) in tests/test_oracle_noise.py.
* PurifiedMps: sites Q_k[l][p][k][r] (bond, physical, Kraus, bond), following the
  reference's steps:
    - initial |0..0>, all legs of extent 1 (buildInitialNetwork, :501-531);
    - 1-qubit gate on the physical leg (contractSingleQubitGateTensor, :240-290);
    - 2-qubit gate: merge the two sites, contract the gate, split back with sqrt(S) on
      both sides (contractTwoQubitGateTensor, :337-434), then truncateSvdTensors with
      eps = 1e-9: keep i + 1 for the first index i whose partial norms on both sides are
      below eps (:1255-1290);
    - channel (applyLocalKrausOp, :898-1127): merge Q with its conjugate over the Kraus
      leg, D[(l p r), (l' p' r')]; apply the channel on (p, p'); split D' = U S U^H back
      into Q' = U sqrt(S) with the new Kraus extent min(2 chiL chiR, 2 kap) (:985-996);
      drop Kraus vectors from the first whose partial norm is below 1e-9 (that one
      excluded, :1060-1072).
  The reference builds the channel from the Choi matrix of XACC's noise utilities (an
  external library, not in the reference tree); D' = sum_o K_o D K_o^H is what that operator
  does on (p, p').  Its split bond may be wider than the rank (zero-padded singular values);
  the economy SVD here gives the same state with at most one zero column fewer.
  Partial norms: "squared" = sum of |x|^2 over a slice, "l2" its square root (the
  convention of ExaTN's computePartialNormsSync is not fixed by the reference tree).
  Pinned against DensityMatrix wherever the 2 kap cap cannot bite.
"""
from __future__ import annotations

import numpy as np

EPS = 1e-9


def _axis(n, q):
    return n - 1 - q  # numpy's first axis is the most significant bit


class DensityMatrix:
    def __init__(self, n: int):
        self.n = n
        self.rho = np.zeros((1 << n, 1 << n), dtype=np.complex128)
        self.rho[0, 0] = 1.0

    def _apply_left(self, t, m, qubits):
        """t (2,)*2n; m acts on the ket axes of `qubits` (first operand = MSB)."""
        n, k = self.n, len(qubits)
        mt = np.asarray(m, dtype=np.complex128).reshape((2,) * (2 * k))
        axes = [_axis(n, q) for q in qubits]
        out = np.tensordot(mt, t, axes=(list(range(k, 2 * k)), axes))
        return np.moveaxis(out, list(range(k)), axes)

    def _sandwich(self, a, b, qubits):
        """a rho b^H restricted to `qubits`."""
        n = self.n
        t = self.rho.reshape((2,) * (2 * n))
        t = self._apply_left(t, a, qubits)
        # bra side: (rho b^H)^T = conj(b) rho^T on the bra axes
        k = len(qubits)
        bt = np.conj(np.asarray(b, dtype=np.complex128)).reshape((2,) * (2 * k))
        axes = [n + _axis(n, q) for q in qubits]
        t = np.tensordot(bt, t, axes=(list(range(k, 2 * k)), axes))
        t = np.moveaxis(t, list(range(k)), axes)
        return t.reshape(1 << n, 1 << n)

    def gate(self, m, qubits):
        self.rho = self._sandwich(m, m, tuple(qubits))
        return self

    def channel(self, q, ops):
        acc = np.zeros_like(self.rho)
        for k in ops:
            acc += self._sandwich(k, k, (q,))
        self.rho = acc
        return self

    def z(self, q):
        d = np.real(np.diag(self.rho))
        bits = (np.arange(1 << self.n) >> q) & 1
        return float(np.sum(d * (1 - 2 * bits)))

    def zz(self, a, b):
        d = np.real(np.diag(self.rho))
        x = np.arange(1 << self.n)
        return float(np.sum(d * (1 - 2 * ((x >> a) & 1)) * (1 - 2 * ((x >> b) & 1))))

    def probabilities(self, qubits):
        """P(outcome) over `qubits`, outcome index sum_i b_i 2^i (b_i the bit of qubits[i])."""
        d = np.real(np.diag(self.rho))
        x = np.arange(1 << self.n)
        idx = np.zeros_like(x)
        for i, q in enumerate(qubits):
            idx |= ((x >> q) & 1) << i
        out = np.zeros(1 << len(qubits))
        np.add.at(out, idx, d)
        return out

    def trace(self):
        return float(np.real(np.trace(self.rho)))

    def reduced(self, qubits):
        """rho_red[x, x'] over `qubits` (x = sum_i b_i 2^i), the other qubits traced out."""
        n, k = self.n, len(qubits)
        t = self.rho.reshape((2,) * (2 * n))
        keep = [_axis(n, q) for q in qubits]
        rest = [a for a in range(n) if a not in keep]
        # ket axes ordered qubits[k-1] .. qubits[0] (most significant first), bra the same
        order = list(reversed(keep)) + rest + [n + a for a in reversed(keep)] + [n + a for a in rest]
        t = np.transpose(t, order).reshape(1 << k, 1 << (n - k), 1 << k, 1 << (n - k))
        return np.einsum("arbr->ab", t)


class PurifiedMps:
    def __init__(self, n: int, partial_norm: str = "squared", kraus_cap: bool = True):
        """kraus_cap=False lifts the reference's 2 kap bound on the new Kraus extent (the
        channel is then exact up to the 1e-9 cut), which is how the tests pin the channel
        step against DensityMatrix where the cap would bite."""
        self.n = n
        self.l2 = partial_norm == "l2"
        self.kraus_cap = kraus_cap
        q0 = np.zeros((1, 2, 1, 1), dtype=np.complex128)
        q0[0, 0, 0, 0] = 1.0
        self.Q = [q0.copy() for _ in range(n)]

    def bonds(self):
        return [1] + [q.shape[3] for q in self.Q]

    def kraus_dims(self):
        return [q.shape[2] for q in self.Q]

    def gate(self, m, qubits):
        m = np.asarray(m, dtype=np.complex128)
        if len(qubits) == 1:
            q = qubits[0]
            self.Q[q] = np.einsum("ap,lpkr->lakr", m, self.Q[q])
            return self
        q0, q1 = qubits
        assert abs(q0 - q1) == 1
        i = min(q0, q1)
        g = m.reshape(2, 2, 2, 2)
        if q0 > q1:  # first operand on the right site: swap the bit order of rows and columns
            g = g.transpose(1, 0, 3, 2)
        a, b = self.Q[i], self.Q[i + 1]
        th = np.einsum("lpkm,mqjr->lpkqjr", a, b)
        th = np.einsum("abpq,lpkqjr->lakbjr", g, th)
        cl, ki, kj, cr = a.shape[0], a.shape[2], b.shape[2], b.shape[3]
        mat = th.reshape(cl * 2 * ki, 2 * kj * cr)
        u, s, vh = np.linalg.svd(mat, full_matrices=False)
        pn = np.sqrt(s) if self.l2 else s  # both sides carry sqrt(s)
        K = len(s)
        for t in range(len(s)):
            if pn[t] < EPS:
                K = t + 1
                break
        rs = np.sqrt(s[:K])
        self.Q[i] = (u[:, :K] * rs).reshape(cl, 2, ki, K)
        self.Q[i + 1] = (rs[:, None] * vh[:K]).reshape(K, 2, kj, cr)
        return self

    def channel(self, q, ops):
        Q = self.Q[q]
        cl, _, kap, cr = Q.shape
        # D[(l p r), (l' p' r')] = sum_k Q[l p k r] conj(Q[l' p' k r'])
        v = Q.transpose(0, 1, 3, 2).reshape(cl * 2 * cr, kap)
        D = (v @ v.conj().T).reshape(cl, 2, cr, cl, 2, cr)
        Dn = np.zeros_like(D)
        for k in ops:
            k = np.asarray(k, dtype=np.complex128)
            Dn += np.einsum("ap,lprmqs,bq->larmbs", k, D, k.conj())
        M = Dn.reshape(cl * 2 * cr, cl * 2 * cr)
        M = 0.5 * (M + M.conj().T)
        lam, U = np.linalg.eigh(M)
        order = np.argsort(lam)[::-1]
        lam = np.clip(lam[order], 0.0, None)
        U = U[:, order]
        knew = min(cl * 2 * cr, 2 * kap) if self.kraus_cap else cl * 2 * cr
        lam, U = lam[:knew], U[:, :knew]
        sig = np.sqrt(lam)
        pn = sig if self.l2 else lam
        K = knew
        for t in range(knew):
            if pn[t] < EPS:
                K = t
                break
        K = max(K, 1)
        Qn = (U[:, :K] * sig[:K]).reshape(cl, 2, cr, K).transpose(0, 1, 3, 2)
        self.Q[q] = np.ascontiguousarray(Qn)
        return self

    def expval_z(self):
        """Tr(rho Z_q) for every qubit through left and right environments (any n), the Kraus
        leg traced with the physical one."""
        left = [np.ones((1, 1), dtype=np.complex128)]
        for Q in self.Q:
            left.append(np.einsum("lm,lpkr,mpks->rs", left[-1], Q, Q.conj()))
        right = [np.ones((1, 1), dtype=np.complex128)]
        for Q in reversed(self.Q):
            right.append(np.einsum("rs,lpkr,mpks->lm", right[-1], Q, Q.conj()))
        right = right[::-1]  # right[q]: the sites q.. n-1 contracted, at bond q
        z = np.array([1.0, -1.0])
        return np.array([np.einsum("lm,lpkr,p,mpks,rs->", left[q], Q, z, Q.conj(), right[q + 1]).real
                         for q, Q in enumerate(self.Q)])

    def density_matrix(self):
        """rho = sum over Kraus legs of |psi><psi| (small n only)."""
        R = np.ones((1, 1, 1, 1), dtype=np.complex128)
        for Q in self.Q:
            d = R.shape[0]
            Rn = np.einsum("xylm,lpkr,mqks->pxqyrs", R, Q, Q.conj())
            R = Rn.reshape(2 * d, 2 * d, Rn.shape[4], Rn.shape[5])
        return R[:, :, 0, 0]

    def as_density(self) -> DensityMatrix:
        dm = DensityMatrix(self.n)
        dm.rho = self.density_matrix()
        return dm
