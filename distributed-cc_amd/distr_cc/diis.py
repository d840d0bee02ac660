"""Pulay DIIS on device vectors (reference distr_cc/diis.py:25-177).

The history vectors, error vectors and the extrapolation live in HBM; only
the (space+1)^2 B matrix and its solve are on the host, as in the reference
(diis.py:118-167).  Error vectors are x - x_prev (diis.py:114); the B-matrix
row of the newest error is built from dot products (diis.py:122-132); the
coefficients come from eigh with a 1e-14 cutoff or a plain solve
(diis.py:157-167).  Every reduction is a deterministic FP64 device dot.
"""
import ctypes

import numpy as np

from distr_cc import _lib as L
from distr_cc.device import DeviceArray


class DeviceDIIS:
    def __init__(self, size, space=6, stream=None):
        if int(space) <= 0:
            raise ValueError("DIIS space must be a positive integer")
        self.size, self.space, self.stream = int(size), int(space), stream
        self.xs = [None] * self.space
        self.es = [None] * self.space
        self._head, self._H, self._xprev = 0, None, None
        self.last_delta_norm_sq = None
        self._partials = DeviceArray((int(L._lib.dcc_dot_partials()),))
        self._out = DeviceArray((1,))
        self._tmp = DeviceArray((self.size,))

    # device primitives
    def _h(self):
        return self.stream.handle if self.stream is not None else None

    def _axpby(self, dst, src, alpha, beta):
        lay = L.Strided()
        lay.ndim = 1
        lay.shape[0] = self.size
        lay.dst_stride[0] = 1
        lay.src_stride[0] = 1
        L.call("dcc_tensor_axpby_f64", dst.vp, src.vp, ctypes.byref(lay), float(alpha), float(beta), self._h())

    def _dot(self, x, y):
        L.call("dcc_dot_f64", x.vp, y.vp, self.size, self._partials.vp, self._out.vp, self._h())
        return float(self._out.download(stream=self.stream)[0])

    def _copy(self, src):
        dst = DeviceArray((self.size,))
        self._axpby(dst, src, 1.0, 0.0)
        return dst

    def update(self, x):
        """x: DeviceArray of ``size`` doubles; returns the extrapolated vector (a new
        DeviceArray) and keeps a copy of it as the next x_prev (diis.py:105-145)."""
        if self._xprev is None:
            self._xprev = self._copy(x)
            self.last_delta_norm_sq = None
            return self._copy(x)
        h = self._head
        if self.xs[h] is None:
            self.xs[h] = DeviceArray((self.size,))
            self.es[h] = DeviceArray((self.size,))
        self._axpby(self.xs[h], x, 1.0, 0.0)
        self._axpby(self.es[h], x, 1.0, 0.0)
        self._axpby(self.es[h], self._xprev, -1.0, 1.0)
        nd = sum(v is not None for v in self.xs)
        if self._H is None:
            self._H = np.zeros((self.space + 1, self.space + 1))
            self._H[0, 1:] = self._H[1:, 0] = 1
        for i in range(nd):
            val = self._dot(self.es[h], self.es[i])
            self._H[h + 1, i + 1] = val
            self._H[i + 1, h + 1] = val
        self._head = (h + 1) % self.space
        xnew = self.extrapolate(nd)
        self._axpby(self._tmp, xnew, 1.0, 0.0)
        self._axpby(self._tmp, self._xprev, -1.0, 1.0)
        self.last_delta_norm_sq = max(self._dot(self._tmp, self._tmp), 0.0)
        self._axpby(self._xprev, xnew, 1.0, 0.0)
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

    def extrapolate(self, nd):
        c = self.coefficients(nd)
        xnew = DeviceArray((self.size,))
        for i, ci in enumerate(c[1:]):
            self._axpby(xnew, self.xs[i], float(ci), 0.0 if i == 0 else 1.0)
        return xnew
