"""Seeded synthetic RCCSDT inputs of the shapes the triples path runs at.

Generator (host NumPy, FP64):
  1. ``rng = np.random.default_rng(seed)``.
  2. occupied energies ~ U[-1.5,-0.3] (sorted), virtual energies ~ U[0.1,3.0]
     (sorted); ``fock = diag(mo_energy)``.
  3. ``B ~ N(0, 0.05^2)`` of shape (naux, nmo, nmo), symmetrized in (p,q);
     chemist ``(pq|rs) = sum_Q B[Q,p,q] B[Q,r,s]``; physicist
     ``pppp[p,q,r,s] = <pq|rs> = (pr|qs)``.
  4. ``t1 = 0``; ``t2[i,j,a,b] = (ia|jb) / (e_i + e_j - e_a - e_b)``.
  5. each canonical block n = lin(i<=j<=k) is N(0, sigma_t3^2) noise drawn from
     its own stream ``default_rng(SeedSequence([seed, 7919, n]))``, symmetrized
     for tied occupied indices so that the simultaneous-permutation symmetry
     holds, then divided by D_ijk^abc (blocks are generated in parallel threads).

The ERI tensor is only materialized when ``nmo`` is small enough; large shapes
build the slices the triples path needs directly from ``B``.  Integral slices are
one GEMM over the auxiliary index.
"""
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np


def _threads():
    return max(1, min(16, len(os.sched_getaffinity(0))))


def _parallel_transpose_0213(X):
    """np.ascontiguousarray(X.transpose(0, 2, 1, 3)) with the copy split over threads along axis 0."""
    out = np.empty((X.shape[0], X.shape[2], X.shape[1], X.shape[3]), dtype=X.dtype)

    def part(p):
        out[p] = X[p].transpose(1, 0, 2)

    if X.shape[0] * X.shape[1] * X.shape[2] * X.shape[3] < (1 << 22):
        for p in range(X.shape[0]):
            part(p)
    else:
        with ThreadPoolExecutor(_threads()) as ex:
            list(ex.map(part, range(X.shape[0])))
    return out


class SyntheticSystem:
    """Holds the seeded inputs; slices are built lazily from the B factor."""

    def __init__(self, nocc, nvir, naux=300, seed=20261202, sigma_b=0.05, sigma_t3=1e-4,
                 dtype=np.float64):
        self.nocc, self.nvir, self.naux, self.seed = int(nocc), int(nvir), int(naux), int(seed)
        self.nmo = self.nocc + self.nvir
        rng = np.random.default_rng(self.seed)
        e_occ = np.sort(rng.uniform(-1.5, -0.3, self.nocc))
        e_vir = np.sort(rng.uniform(0.1, 3.0, self.nvir))
        self.mo_energy = np.concatenate([e_occ, e_vir])
        self.fock = np.diag(self.mo_energy)
        B = rng.normal(0.0, sigma_b, (self.naux, self.nmo, self.nmo))
        self.B = 0.5 * (B + B.transpose(0, 2, 1))
        self.sigma_t3 = sigma_t3
        self.dtype = dtype

    # physicist <pq|rs> = (pr|qs) = sum_Q B[Q,p,r] B[Q,q,s]
    def eri_phys(self, ps, qs, rs, ss):
        """Slice <p q | r s> for index slices (ps, qs, rs, ss) of the MO range: one GEMM
        (p r) x Q x (q s), then the (p r q s) -> (p q r s) reordering."""
        Bpr = self.B[:, ps, rs]
        Bqs = self.B[:, qs, ss]
        nq, np_, nr = Bpr.shape
        _, nq2, ns = Bqs.shape
        X = Bpr.reshape(nq, np_ * nr).T @ Bqs.reshape(nq, nq2 * ns)
        return _parallel_transpose_0213(X.reshape(np_, nr, nq2, ns))

    def pppp(self):
        """Full physicist ERI tensor (nmo^4); only for small systems."""
        s = slice(0, self.nmo)
        return self.eri_phys(s, s, s, s)

    def t2(self):
        o = self.nocc
        so, sv = slice(0, o), slice(o, self.nmo)
        eia = self.mo_energy[:o, None] - self.mo_energy[None, o:]
        d = eia[:, None, :, None] + eia[None, :, None, :]
        return self.eri_phys(so, so, sv, sv) / d

    def t3_tri(self):
        """Canonical-block T3, shape (nocc3, v, v, v), lexicographic i<=j<=k."""
        o, v = self.nocc, self.nvir
        eia = self.mo_energy[:o, None] - self.mo_energy[None, o:]
        triples = [(i, j, k) for i in range(o) for j in range(i, o) for k in range(j, o)]
        out = np.empty((len(triples), v, v, v), dtype=self.dtype)

        def block(n):
            # drawn straight into the output block and divided slab by slab, so that no block-sized
            # temporaries are allocated (first-touch page faults dominate at nvir = 160)
            i, j, k = triples[n]
            rng = np.random.default_rng(np.random.SeedSequence([self.seed, 7919, n]))
            x = out[n] if out.dtype == np.float64 else np.empty((v, v, v))
            rng.standard_normal(out=x)
            x *= self.sigma_t3
            if i == j == k:
                x[...] = (x + x.transpose(0, 2, 1) + x.transpose(1, 0, 2) + x.transpose(1, 2, 0)
                          + x.transpose(2, 0, 1) + x.transpose(2, 1, 0)) / 6.0
            elif i == j:
                x[...] = 0.5 * (x + x.transpose(1, 0, 2))
            elif j == k:
                x[...] = 0.5 * (x + x.transpose(0, 2, 1))
            for a in range(v):
                x[a] /= (eia[i, a] + eia[j, :])[:, None] + eia[k, None, :]
            if x is not out[n]:
                out[n] = x

        if len(triples) * v ** 3 < (1 << 22):
            for n in range(len(triples)):
                block(n)
        else:
            with ThreadPoolExecutor(_threads()) as ex:
                list(ex.map(block, range(len(triples))))
        return out

    def bare_intermediates(self):
        """Leading (T1 = T2 = 0) values of the triples intermediates: ERI slices.

        Shapes follow distr_cc/rccsdt.py:380,445,625: F_oo (o,o), F_vv (v,v),
        W_oooo (o,o,o,o), W_ovvo (o,v,v,o), W_ovov (o,v,o,v), W_vooo (v,o,o,o),
        W_vvvo (v,v,v,o), W_vvvv (v,v,v,v).
        """
        o = self.nocc
        so, sv = slice(0, o), slice(o, self.nmo)
        f = self.fock
        return {
            'F_oo': np.ascontiguousarray(f[so, so]),
            'F_vv': np.ascontiguousarray(f[sv, sv]),
            'W_oooo': self.eri_phys(so, so, so, so),
            'W_ovvo': self.eri_phys(so, sv, sv, so),
            'W_ovov': self.eri_phys(so, sv, so, sv),
            'W_vooo': self.eri_phys(sv, so, so, so),
            'W_vvvo': self.eri_phys(sv, sv, sv, so),
            'W_vvvv': self.eri_phys(sv, sv, sv, sv),
        }
