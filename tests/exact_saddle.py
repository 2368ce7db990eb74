"""Checker-side exact solution of the reference saddle system
[[I3 (x) A, C^T],[C, 0]] [x; lambda] = [r; s]   (linsolve.ts:771-790)
for a given FP64 A, C and right-hand side: LAPACK Cholesky of
A_hat = A + mu 11^T (mu = trace(A) / V^2) through the same Schur reduction the
device uses, then iterative refinement with the oracle's double-double
residual (rco_saddle_residual_dd) until the correction is at the rounding
level. The fixed point is the exact solution of that system rounded to FP64
(kappa(A_hat) ~ 1e9 at V = 8192: an FP64 solve alone is ~2e-8 away).

TEST INFRASTRUCTURE ONLY (uses the oracle and scipy)."""
from __future__ import annotations

import numpy as np
import scipy.linalg as sla

from oracle import rc_oracle as rco


class ExactSaddle:
    def __init__(self, A, Cm):
        A = np.asarray(A, dtype=np.float64)
        self.A = A
        self.V = V = A.shape[0]
        self.C = np.ascontiguousarray(np.asarray(Cm, dtype=np.float64).reshape(-1, 3 * V))
        self.k = k = self.C.shape[0]
        mu = np.trace(A) / float(V) ** 2
        self.cf = sla.cho_factor(A + mu, lower=True)
        Cb = [self.C[:, c * V:(c + 1) * V] for c in range(3)]
        self.Cb = Cb
        self.Y = [sla.cho_solve(self.cf, Cb[c].T) for c in range(3)]
        CY = sum(Cb[c] @ self.Y[c] for c in range(3))
        CN = np.stack([Cb[c].sum(1) for c in range(3)], 1)
        M = np.zeros((k + 3, k + 3))
        M[:k, :k] = CY
        M[:k, k:] = -CN / V
        M[k:, :k] = CN.T
        self.M = M

    def solve_fp64(self, rt, rb):
        """The Schur-reduced solve in FP64 (one pass, no refinement)."""
        V, k = self.V, self.k
        Ar = [sla.cho_solve(self.cf, rt[c * V:(c + 1) * V]) for c in range(3)]
        rhs = np.concatenate([sum(self.Cb[c] @ Ar[c] for c in range(3)) - rb,
                              [rt[c * V:(c + 1) * V].sum() for c in range(3)]])
        sol = np.linalg.solve(self.M, rhs)
        lam, t = sol[:k], sol[k:]
        x = np.concatenate([Ar[c] - self.Y[c] @ lam + t[c] / V for c in range(3)])
        return x, lam

    def solve(self, r, s=None, max_steps=6):
        """Exact (refined) solution: x (3V coordinate blocks), lambda (k),
        and the relative size of each correction."""
        s = np.zeros(self.k) if s is None else np.asarray(s, dtype=np.float64)
        r = np.asarray(r, dtype=np.float64).reshape(3 * self.V)
        x, lam = self.solve_fp64(r, s)
        hist = []
        for _ in range(max_steps):
            rt, rb = rco.saddle_residual_dd(self.A, self.C, r, s, x, lam)
            dx, dl = self.solve_fp64(rt, rb)
            x, lam = x + dx, lam + dl
            hist.append(float(np.linalg.norm(dx) / np.linalg.norm(x)))
            if hist[-1] < 1e-16:
                break
        return x, lam, hist


def blocks(aos):
    """(V, 3) AoS -> coordinate-block vector [x.., y.., z..]."""
    return np.ascontiguousarray(np.asarray(aos, dtype=np.float64).T).reshape(-1)


def aos(blk, V):
    return np.ascontiguousarray(np.asarray(blk).reshape(3, V).T)
