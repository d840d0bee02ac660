"""Brute-force determinant-space coupled cluster for tiny closed-shell systems.

Test infrastructure only: nothing in the product imports this module.

It evaluates the similarity-transformed Hamiltonian exactly,

    Hbar|0> = exp(-T) H exp(T) |0>,   T = T1 + T2 + T3,

with the spin-free closed-shell cluster operator

    T1 = sum t1[i,a] E_ai
    T2 = 1/2 sum t2[i,j,a,b] E_ai E_bj
    T3 = 1/6 sum t3[i,j,k,a,b,c] E_ai E_bj E_ck

(E_pq = sum_sigma a+_{p sigma} a_{q sigma}) in the full determinant space, and
returns the coefficients of Hbar|0> on all determinants.  The spin-adapted
residuals rho1, rho2, rho3 of any CCSDT implementation are defined by

    Hbar|0> = E |0> + sum rho1 E_ai|0> + 1/2 sum rho2 E_ai E_bj|0>
            + 1/6 sum rho3 E_ai E_bj E_ck|0> + (quadruples and higher),

so ``excitation_image(rho1, rho2, rho3)`` must equal ``hbar_vector`` on every
determinant of excitation rank 1..3.  This fixes the residual equations up to
the null space of the closed-shell triples (the part of rho3 that is totally
symmetric in abc for fixed ijk), independently of how they are factorised.

The Hamiltonian is H = sum h[p,q] E_pq + 1/2 sum g[p,q,r,s] (E_pr E_qs -
delta_qr E_ps) with physicist integrals g[p,q,r,s] = <pq|rs>: creators p, q,
annihilators s, r (the convention of distr_cc/rccsdt.py:216-221, where
eris.pppp[:o,:o,o:,o:] is <ij|ab>).  h and g need not be Hermitian; g only
needs g[p,q,r,s] = g[q,p,s,r], which the T1-dressed integrals keep.
"""
import itertools

import numpy as np
import scipy.sparse as sps


class DetSpace:
    """Determinants of ``nocc`` alpha and ``nocc`` beta electrons in ``nmo`` orbitals."""

    def __init__(self, nmo, nocc):
        self.nmo, self.nocc = nmo, nocc
        self.strings = [sum(1 << p for p in c) for c in itertools.combinations(range(nmo), nocc)]
        self.index = {s: n for n, s in enumerate(self.strings)}
        self.ns = len(self.strings)
        self.dim = self.ns * self.ns
        occ_mask = (1 << nocc) - 1
        exc = np.array([bin(s & ~occ_mask).count("1") for s in self.strings])
        # excitation rank of determinant (alpha string A, beta string B) at A * ns + B
        self.rank = (exc[:, None] + exc[None, :]).ravel()
        self.ref = self.index[occ_mask] * self.ns + self.index[occ_mask]
        self._e = {}
        self._eye = sps.identity(self.ns, format="csr")

    def _string_op(self, p, q):
        """a+_p a_q on the strings of one spin, as an ns x ns matrix."""
        rows, cols, vals = [], [], []
        for n, s in enumerate(self.strings):
            if not (s >> q) & 1:
                continue
            sign = -1 if bin(s & ((1 << q) - 1)).count("1") % 2 else 1
            t = s & ~(1 << q)
            if (t >> p) & 1:
                continue
            if bin(t & ((1 << p) - 1)).count("1") % 2:
                sign = -sign
            t |= 1 << p
            rows.append(self.index[t])
            cols.append(n)
            vals.append(sign)
        return sps.csr_matrix((vals, (rows, cols)), shape=(self.ns, self.ns), dtype=np.float64)

    def E(self, p, q):
        """Spin-free excitation operator E_pq (alpha and beta parts commute with the
        other spin's string: a pair of operators of one spin passes the other
        spin's electrons with sign +1)."""
        key = (p, q)
        if key not in self._e:
            e = self._string_op(p, q)
            self._e[key] = (sps.kron(e, self._eye) + sps.kron(self._eye, e)).tocsr()
        return self._e[key]

    def hamiltonian(self, h, g):
        n = self.nmo
        H = sps.csr_matrix((self.dim, self.dim))
        for p in range(n):
            for q in range(n):
                if h[p, q] != 0.0:
                    H = H + h[p, q] * self.E(p, q)
        # 1/2 sum_pqrs g[p,q,r,s] (E_pr E_qs - delta_qr E_ps)
        for p in range(n):
            for r in range(n):
                acc = sps.csr_matrix((self.dim, self.dim))
                for q in range(n):
                    for s in range(n):
                        if g[p, q, r, s] != 0.0:
                            acc = acc + g[p, q, r, s] * self.E(q, s)
                H = H + 0.5 * (self.E(p, r) @ acc)
                for s in range(n):
                    # delta_qr term: q == r
                    if g[p, r, r, s] != 0.0:
                        H = H - 0.5 * g[p, r, r, s] * self.E(p, s)
        return H.tocsr()

    def cluster_operator(self, t1=None, t2=None, t3=None):
        o, n = self.nocc, self.nmo
        v = n - o
        T = sps.csr_matrix((self.dim, self.dim))
        if t1 is not None:
            for i in range(o):
                for a in range(v):
                    if t1[i, a] != 0.0:
                        T = T + t1[i, a] * self.E(o + a, i)
        if t2 is not None:
            for i in range(o):
                for a in range(v):
                    acc = sps.csr_matrix((self.dim, self.dim))
                    for j in range(o):
                        for b in range(v):
                            if t2[i, j, a, b] != 0.0:
                                acc = acc + t2[i, j, a, b] * self.E(o + b, j)
                    T = T + 0.5 * (self.E(o + a, i) @ acc)
        if t3 is not None:
            for i in range(o):
                for a in range(v):
                    for j in range(o):
                        for b in range(v):
                            acc = sps.csr_matrix((self.dim, self.dim))
                            for k in range(o):
                                for c in range(v):
                                    if t3[i, j, k, a, b, c] != 0.0:
                                        acc = acc + t3[i, j, k, a, b, c] * self.E(o + c, k)
                            T = T + (1.0 / 6.0) * (self.E(o + a, i) @ (self.E(o + b, j) @ acc))
        return T.tocsr()

    def reference(self):
        x = np.zeros(self.dim)
        x[self.ref] = 1.0
        return x

    @staticmethod
    def _exp_apply(T, x, sign=1.0, max_terms=40):
        out = x.copy()
        term = x.copy()
        for n in range(1, max_terms):
            term = sign * (T @ term) / n
            if not np.any(term):
                return out
            out = out + term
        raise RuntimeError("cluster operator is not nilpotent on this space")

    def hbar_vector(self, h, g, t1=None, t2=None, t3=None, H=None):
        """exp(-T) H exp(T) |0> as a determinant-space vector."""
        if H is None:
            H = self.hamiltonian(h, g)
        T = self.cluster_operator(t1, t2, t3)
        psi = self._exp_apply(T, self.reference())
        return self._exp_apply(T, H @ psi, sign=-1.0)

    def excitation_image(self, rho0=0.0, rho1=None, rho2=None, rho3=None):
        """E|0> + sum rho1 E_ai|0> + 1/2 sum rho2 E_ai E_bj|0> + 1/6 sum rho3 E_ai E_bj E_ck|0>."""
        o = self.nocc
        v = self.nmo - o
        x0 = self.reference()
        out = rho0 * x0
        single = {}
        for i in range(o):
            for a in range(v):
                single[i, a] = self.E(o + a, i) @ x0
        if rho1 is not None:
            for (i, a), y in single.items():
                out = out + rho1[i, a] * y
        if rho2 is not None or rho3 is not None:
            double = {}
            for (j, b), y in single.items():
                for i in range(o):
                    for a in range(v):
                        double[i, j, a, b] = self.E(o + a, i) @ y
            if rho2 is not None:
                for (i, j, a, b), y in double.items():
                    out = out + 0.5 * rho2[i, j, a, b] * y
            if rho3 is not None:
                for (j, k, b, c), y in double.items():
                    for i in range(o):
                        for a in range(v):
                            coef = rho3[i, j, k, a, b, c]
                            if coef != 0.0:
                                out = out + (coef / 6.0) * (self.E(o + a, i) @ y)
        return out

    def rank_mask(self, rank):
        return self.rank == rank


def random_hamiltonian(nmo, rng, scale_g=0.1, scale_h=1.0, nocc=None, gap=0.0):
    """A real Hermitian one-electron matrix and two-electron integrals with the
    full eightfold symmetry of real orbitals, in physicist order <pq|rs>.

    ``gap`` lowers the first ``nocc`` diagonal levels further (a molecule-like
    spectrum with a HOMO-LUMO gap, for iterative solves)."""
    h = scale_h * rng.standard_normal((nmo, nmo))
    h = 0.5 * (h + h.T)
    h += np.diag(np.arange(nmo, dtype=float) * 1.5 - nmo)  # occupied levels below virtual ones
    if nocc:
        h[np.arange(nocc), np.arange(nocc)] -= gap
    x = rng.standard_normal((nmo, nmo, nmo, nmo))
    # chemist (pq|rs) with (pq|rs) = (qp|rs) = (pq|sr) = (rs|pq)
    c = x + x.transpose(1, 0, 2, 3)
    c = c + c.transpose(0, 1, 3, 2)
    c = c + c.transpose(2, 3, 0, 1)
    c *= scale_g / 8.0
    g = c.transpose(0, 2, 1, 3)  # <pq|rs> = (pr|qs)
    return h, np.ascontiguousarray(g)


def fock(h, g, nocc):
    """f[p,q] = h[p,q] + sum_k (2 <pk|qk> - <pk|kq>) over doubly occupied k."""
    o = slice(0, nocc)
    return h + 2.0 * np.einsum("pkqk->pq", g[:, o, :, o]) - np.einsum("pkkq->pq", g[:, o, o, :])


def symmetrize_t3(x):
    """Sum of x over the six simultaneous permutations of (i a), (j b), (k c)."""
    out = np.zeros_like(x)
    for p in itertools.permutations(range(3)):
        out += x.transpose(p + tuple(3 + q for q in p))
    return out
