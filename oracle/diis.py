"""Host restatement of the reference DIIS (oracle, test infrastructure only).

Follows distr_cc/diis.py:105-177 at one rank: error vector x - x_prev
(:114), B-matrix rows from dot products (:118-132), eigh with a 1e-14
cutoff or a plain solve (:157-167), x_new = sum_i c_i x_i (:169-176).

Pin: the coefficients minimise |sum_i c_i e_i| subject to sum_i c_i = 1,
which tests/test_oracle_diis.py checks against numpy.linalg.lstsq on the
constrained least-squares problem (a different algorithm).
"""
import numpy as np


class DIIS:
    def __init__(self, space=6):
        self.space = int(space)
        self.xs, self.es = [None] * self.space, [None] * self.space
        self._head, self._H, self._xprev = 0, None, None

    def update(self, x):
        x = np.asarray(x, dtype=np.float64).ravel()
        if self._xprev is None:
            self._xprev = x.copy()
            return x.copy()
        err = x - self._xprev
        self.xs[self._head], self.es[self._head] = x.copy(), err
        nd = sum(v is not None for v in self.xs)
        if self._H is None:
            self._H = np.zeros((self.space + 1, self.space + 1))
            self._H[0, 1:] = self._H[1:, 0] = 1
        for i in range(nd):
            val = float(np.dot(err, self.es[i]))
            self._H[self._head + 1, i + 1] = val
            self._H[i + 1, self._head + 1] = val
        self._head = (self._head + 1) % self.space
        c = self.coefficients(nd)
        xnew = sum(ci * self.xs[i] for i, ci in enumerate(c[1:]))
        self._xprev = xnew.copy()
        return xnew

    def coefficients(self, nd):
        h = self._H[:nd + 1, :nd + 1]
        g = np.zeros(nd + 1)
        g[0] = 1
        w, v = np.linalg.eigh(h)
        if np.any(abs(w) < 1e-14):
            idx = abs(w) > 1e-14
            return np.dot(v[:, idx] * (1 / w[idx]), np.dot(v[:, idx].T, g))
        return np.linalg.solve(h, g)
