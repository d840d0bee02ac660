"""Minimal molecular front end: Gaussian basis sets, one- and two-electron integrals, RHF.

It stands in for the two PySCF objects the reference's drivers are built on
(``gto.Mole`` and ``scf.RHF``, distr_cc/rccsdt.py:1162-1213 and the tests'
set-up, tests/test_rccsdt.py:46-63): ``Molecule`` gives the AO integrals,
``RHF`` the canonical orbitals, and an ``RHF`` object has the fields
``RCCSDT(mf, comm)`` reads (mo_coeff, mo_occ, mo_energy, e_tot, get_hcore(),
_eri, energy_nuc()).  This is the input side of the triples path, not the
path itself; integrals are evaluated once per molecule on the host.

Integrals use the McMurchie-Davidson scheme over Cartesian primitives
(Hermite expansion coefficients E, Hermite Coulomb integrals R with the Boys
function F_n(T) = Gamma(n+1/2) P(n+1/2, T) / (2 T^(n+1/2))), vectorised over the primitive pairs
and quartets of each shell quartet.  Shells of angular momentum
l >= 2 are transformed to real solid harmonics (spherical basis, as PySCF's
default) when ``cart=False``.  Lengths are converted from Angstrom with the
Bohr radius 0.52917721092 Angstrom.
"""
import itertools
import math

import numpy as np
from scipy.special import gamma, gammainc

BOHR = 0.52917721092
Z_OF = {"H": 1, "He": 2, "Li": 3, "Be": 4, "B": 5, "C": 6, "N": 7, "O": 8, "F": 9, "Ne": 10}
SYMBOL = {v: k for k, v in Z_OF.items()}

# Basis sets as {element: [(l, [(exponent, coefficient), ...]), ...]} (EMSL basis set exchange data;
# SP shells are split into an s and a p shell with shared exponents).
BASIS = {
    "6-31g": {
        "H": [(0, [(18.7311370, 0.03349460), (2.8253937, 0.23472695), (0.6401217, 0.81375733)]),
              (0, [(0.1612778, 1.0)])],
        "O": [(0, [(5484.6717000, 0.0018311), (825.2349500, 0.0139501), (188.0469600, 0.0684451),
                   (52.9645000, 0.2327143), (16.8975700, 0.4701930), (5.7996353, 0.3585209)]),
              (0, [(15.5396160, -0.1107775), (3.5999336, -0.1480263), (1.0137618, 1.1307670)]),
              (1, [(15.5396160, 0.0708743), (3.5999336, 0.3397528), (1.0137618, 0.7271586)]),
              (0, [(0.2700058, 1.0)]),
              (1, [(0.2700058, 1.0)])],
    },
    "cc-pvdz": {
        "H": [(0, [(13.0100000, 0.0196850), (1.9620000, 0.1379770), (0.4446000, 0.4781480),
                   (0.1220000, 0.5012400)]),
              (0, [(0.1220000, 1.0)]),
              (1, [(0.7270000, 1.0)])],
        "N": [(0, [(9046.0000000, 0.0007000), (1357.0000000, 0.0053890), (309.3000000, 0.0274060),
                   (87.7300000, 0.1032070), (28.5600000, 0.2787230), (10.2100000, 0.4485400),
                   (3.8380000, 0.2782380), (0.7466000, 0.0154400), (0.2248000, -0.0028640)]),
              (0, [(9046.0000000, -0.0001530), (1357.0000000, -0.0012080), (309.3000000, -0.0059920),
                   (87.7300000, -0.0245440), (28.5600000, -0.0674590), (10.2100000, -0.1580780),
                   (3.8380000, -0.1218310), (0.7466000, 0.5490030), (0.2248000, 0.5788150)]),
              (0, [(0.2248000, 1.0)]),
              (1, [(13.5500000, 0.0399190), (2.9170000, 0.2171690), (0.7973000, 0.5103190),
                   (0.2185000, 0.4622140)]),
              (1, [(0.2185000, 1.0)]),
              (2, [(0.8170000, 1.0)])],
        "O": [(0, [(11720.0000000, 0.0007100), (1759.0000000, 0.0054700), (400.8000000, 0.0278370),
                   (113.7000000, 0.1048000), (37.0300000, 0.2830620), (13.2700000, 0.4487190),
                   (5.0250000, 0.2709520), (1.0130000, 0.0154580), (0.3023000, -0.0025850)]),
              (0, [(11720.0000000, -0.0001600), (1759.0000000, -0.0012630), (400.8000000, -0.0062670),
                   (113.7000000, -0.0257160), (37.0300000, -0.0709240), (13.2700000, -0.1654110),
                   (5.0250000, -0.1169550), (1.0130000, 0.5573680), (0.3023000, 0.5727590)]),
              (0, [(0.3023000, 1.0)]),
              (1, [(17.7000000, 0.0430180), (3.8540000, 0.2289130), (1.0460000, 0.5087280),
                   (0.2753000, 0.4605310)]),
              (1, [(0.2753000, 1.0)]),
              (2, [(1.1850000, 1.0)])],
    },
}
BASIS["631g"] = BASIS["6-31g"]
BASIS["ccpvdz"] = BASIS["cc-pvdz"]


def cart_components(l):
    """Cartesian exponents (lx, ly, lz) of shell l in the usual order (xx, xy, xz, yy, yz, zz)."""
    return [(lx, ly, l - lx - ly) for lx in range(l, -1, -1) for ly in range(l - lx, -1, -1)]


def _dfact(n):
    return 1.0 if n <= 0 else float(np.prod(np.arange(n, 0, -2, dtype=np.float64)))


def primitive_norm(alpha, lmn):
    """Normalisation of x^l y^m z^n exp(-alpha r^2)."""
    l, m, n = lmn
    L = l + m + n
    return math.sqrt((2 * alpha / math.pi) ** 1.5 * (4 * alpha) ** L
                     / (_dfact(2 * l - 1) * _dfact(2 * m - 1) * _dfact(2 * n - 1)))


class Shell:
    def __init__(self, l, exps, coefs, center, atom):
        self.l = int(l)
        self.exps = np.asarray(exps, dtype=np.float64)
        self.center = np.asarray(center, dtype=np.float64)
        self.atom = atom
        # normalise the contraction for the axis-aligned component (x^l) as PySCF does
        lmn = (self.l, 0, 0)
        c = np.asarray(coefs, dtype=np.float64) * np.array([primitive_norm(a, lmn) for a in self.exps])
        s = 0.0
        for ci, ai in zip(c, self.exps):
            for cj, aj in zip(c, self.exps):
                s += ci * cj * _overlap_1d_total(ai, aj, lmn)
        self.coefs = c / math.sqrt(s)


def _overlap_1d_total(a, b, lmn):
    """<x^l y^m z^n e^{-a r^2} | same with exponent b> for two functions on one centre."""
    p = a + b
    out = (math.pi / p) ** 1.5
    for l in lmn:
        out *= _dfact(2 * l - 1) / (2 * p) ** l
    return out


def hermite_e(i, j, t, Qx, a, b):
    """E^{ij}_t Hermite expansion coefficient (McMurchie-Davidson recursion); Qx = A_x - B_x."""
    p = a + b
    q = a * b / p
    if t < 0 or t > i + j:
        return 0.0
    if i == j == t == 0:
        return math.exp(-q * Qx * Qx)
    if j == 0:
        return (1 / (2 * p) * hermite_e(i - 1, j, t - 1, Qx, a, b) - q * Qx / a * hermite_e(i - 1, j, t, Qx, a, b)
                + (t + 1) * hermite_e(i - 1, j, t + 1, Qx, a, b))
    return (1 / (2 * p) * hermite_e(i, j - 1, t - 1, Qx, a, b) + q * Qx / b * hermite_e(i, j - 1, t, Qx, a, b)
            + (t + 1) * hermite_e(i, j - 1, t + 1, Qx, a, b))


def boys(n, T):
    """F_n(T) = int_0^1 t^{2n} exp(-T t^2) dt: incomplete-gamma form, series for tiny T."""
    T = np.asarray(T, dtype=np.float64)
    small = T < 1e-7
    Ts = np.where(small, 1.0, T)
    big = gamma(n + 0.5) * gammainc(n + 0.5, Ts) / (2.0 * Ts ** (n + 0.5))
    series = 1.0 / (2 * n + 1) - T / (2 * n + 3) + T * T / (2 * (2 * n + 5))
    return np.where(small, series, big)


def _e_table(la, lb, a, b, Qx):
    """E^{ij}_t for i <= la, j <= lb over arrays of exponents a, b (broadcast); dict (i,j,t) -> array."""
    p = a + b
    q = a * b / p
    memo = {}

    def E(i, j, t):
        if t < 0 or t > i + j or i < 0 or j < 0:
            return 0.0
        key = (i, j, t)
        if key in memo:
            return memo[key]
        if i == j == t == 0:
            val = np.exp(-q * Qx * Qx)
        elif j == 0:
            val = E(i - 1, j, t - 1) / (2 * p) - q * Qx / a * E(i - 1, j, t) + (t + 1) * E(i - 1, j, t + 1)
        else:
            val = E(i, j - 1, t - 1) / (2 * p) + q * Qx / b * E(i, j - 1, t) + (t + 1) * E(i, j - 1, t + 1)
        memo[key] = val
        return val

    return E


def _r_table(L, alpha, PQ):
    """Hermite Coulomb integrals R_{tuv}(alpha, PQ) for t+u+v <= L over arrays; dict (t,u,v) -> array."""
    X, Y, Z = PQ[..., 0], PQ[..., 1], PQ[..., 2]
    T = alpha * (X * X + Y * Y + Z * Z)
    F = [boys(n, T) * (-2 * alpha) ** n for n in range(L + 1)]
    memo = {}

    def R(t, u, v, n):
        if t < 0 or u < 0 or v < 0:
            return 0.0
        key = (t, u, v, n)
        if key in memo:
            return memo[key]
        if t == u == v == 0:
            val = F[n]
        elif t > 0:
            val = (t - 1) * R(t - 2, u, v, n + 1) + X * R(t - 1, u, v, n + 1)
        elif u > 0:
            val = (u - 1) * R(t, u - 2, v, n + 1) + Y * R(t, u - 1, v, n + 1)
        else:
            val = (v - 1) * R(t, u, v - 2, n + 1) + Z * R(t, u, v - 1, n + 1)
        memo[key] = val
        return val

    return {(t, u, v): R(t, u, v, 0) for t in range(L + 1) for u in range(L + 1 - t) for v in range(L + 1 - t - u)}


def _hermite_index(L):
    return [(t, u, v) for t in range(L + 1) for u in range(L + 1 - t) for v in range(L + 1 - t - u)]


class _ShellPair:
    """Primitive-pair data of two shells: exponents, centres, weights and E[comp pair, tuv, prim]."""

    def __init__(self, sa, sb):
        a = np.repeat(sa.exps, len(sb.exps))
        b = np.tile(sb.exps, len(sa.exps))
        self.p = a + b
        self.P = (a[:, None] * sa.center + b[:, None] * sb.center) / self.p[:, None]
        self.K = np.repeat(sa.coefs, len(sb.coefs)) * np.tile(sb.coefs, len(sa.coefs))
        self.L = sa.l + sb.l
        self.tuv = _hermite_index(self.L)
        Ex = [_e_table(sa.l, sb.l, a, b, sa.center[x] - sb.center[x]) for x in range(3)]
        ca, cb = cart_components(sa.l), cart_components(sb.l)
        self.ncomp = (len(ca), len(cb))
        E = np.zeros((len(ca) * len(cb), len(self.tuv), len(a)))
        for ia, A in enumerate(ca):
            for ib, B in enumerate(cb):
                for it, (t, u, v) in enumerate(self.tuv):
                    val = Ex[0](A[0], B[0], t) * Ex[1](A[1], B[1], u) * Ex[2](A[2], B[2], v)
                    if not np.isscalar(val) or val != 0.0:
                        E[ia * len(cb) + ib, it] = val
        self.E = E


def hermite_r(tmax, umax, vmax, p, PC):
    """R_{tuv}(p, PC) for all t<=tmax, u<=umax, v<=vmax (returns dict keyed by (t,u,v))."""
    X, Y, Z = PC
    T = p * (X * X + Y * Y + Z * Z)
    nmax = tmax + umax + vmax
    F = [float(boys(n, T)) for n in range(nmax + 1)]
    cache = {}

    def R(t, u, v, n):
        key = (t, u, v, n)
        if key in cache:
            return cache[key]
        if t < 0 or u < 0 or v < 0:
            val = 0.0
        elif t == u == v == 0:
            val = (-2 * p) ** n * F[n]
        elif t > 0:
            val = (t - 1) * R(t - 2, u, v, n + 1) + X * R(t - 1, u, v, n + 1)
        elif u > 0:
            val = (u - 1) * R(t, u - 2, v, n + 1) + Y * R(t, u - 1, v, n + 1)
        else:
            val = (v - 1) * R(t, u, v - 2, n + 1) + Z * R(t, u, v - 1, n + 1)
        cache[key] = val
        return val

    return {(t, u, v): R(t, u, v, 0) for t in range(tmax + 1) for u in range(umax + 1) for v in range(vmax + 1)}


def _cart2sph(l):
    """Real solid-harmonic rows over the Cartesian components of shell l (orthonormal combos for
    normalised Cartesians), ordered m = -l..l as PySCF."""
    if l < 2:
        if l == 1:
            return np.eye(3)   # px, py, pz
        return np.eye(1)
    if l == 2:
        # components: xx, xy, xz, yy, yz, zz (each normalised as x^l in the contraction)
        s3 = math.sqrt(3.0)
        return np.array([
            [0, 1, 0, 0, 0, 0],                 # dxy
            [0, 0, 0, 0, 1, 0],                 # dyz
            [-0.5, 0, 0, -0.5, 0, 1],           # dz2
            [0, 0, 1, 0, 0, 0],                 # dxz
            [s3 / 2, 0, 0, -s3 / 2, 0, 0],      # dx2-y2
        ])
    raise NotImplementedError("spherical transformation implemented for l <= 2")


def chemcore(mol):
    """Number of core orbitals to freeze (pyscf.data.elements.chemcore semantics for H..Ne: none for
    H and He, the 1s orbital for Li..Ne), as the reference's examples pass to RCCSDT(frozen=...)."""
    return int(sum(0 if z <= 2 else 1 for z in mol.charges))


class Molecule:
    """Atoms (symbol, (x, y, z) in Angstrom) and a basis set name."""

    def __init__(self, atoms, basis="6-31g", charge=0, cart=False):
        self.atoms = [(SYMBOL.get(a, a) if isinstance(a, int) else a, np.asarray(xyz, dtype=np.float64) / BOHR)
                      for a, xyz in atoms]
        self.basis = basis.lower()
        self.cart = cart
        self.charges = np.array([Z_OF[a] for a, _ in self.atoms], dtype=np.float64)
        self.coords = np.array([xyz for _, xyz in self.atoms])
        self.nelectron = int(self.charges.sum()) - int(charge)
        self.shells = []
        for ia, (sym, xyz) in enumerate(self.atoms):
            for l, prims in BASIS[self.basis][sym]:
                ex, co = zip(*prims)
                self.shells.append(Shell(l, ex, co, xyz, ia))
        # AO functions: Cartesian components per shell (normalised as x^l: relative norms of the other
        # components are applied in the Cartesian -> AO transformation)
        self.cart_funcs = []
        for si, sh in enumerate(self.shells):
            for lmn in cart_components(sh.l):
                self.cart_funcs.append((si, lmn))
        self.ncart = len(self.cart_funcs)
        blocks, scale = [], []
        for sh in self.shells:
            comps = cart_components(sh.l)
            # component normalisation relative to x^l
            rel = [primitive_norm(1.0, lmn) / primitive_norm(1.0, (sh.l, 0, 0)) for lmn in comps]
            scale += rel
            blocks.append(np.eye(len(comps)) if (cart or sh.l < 2) else _cart2sph(sh.l))
        self.c2ao = np.zeros((self.ncart, sum(b.shape[0] for b in blocks)))
        r = c = 0
        for b in blocks:
            self.c2ao[r:r + b.shape[1], c:c + b.shape[0]] = b.T
            r += b.shape[1]
            c += b.shape[0]
        self.c2ao = np.diag(scale) @ self.c2ao
        self.nao = self.c2ao.shape[1]

    def energy_nuc(self):
        e = 0.0
        for i in range(len(self.atoms)):
            for j in range(i):
                e += self.charges[i] * self.charges[j] / np.linalg.norm(self.coords[i] - self.coords[j])
        return e

    # -- primitive lists per Cartesian function
    def _prims(self):
        out = []
        for si, lmn in self.cart_funcs:
            sh = self.shells[si]
            out.append((sh.center, lmn, sh.exps, sh.coefs))
        return out

    def _to_ao(self, mat_cart):
        return self.c2ao.T @ mat_cart @ self.c2ao

    def int1e(self):
        """(S, T, V) in the AO basis."""
        prims = self._prims()
        n = self.ncart
        S, T, V = np.zeros((n, n)), np.zeros((n, n)), np.zeros((n, n))
        for i in range(n):
            A, la, ea, ca = prims[i]
            for j in range(i + 1):
                B, lb, eb, cb = prims[j]
                s = t = v = 0.0
                for a, cA in zip(ea, ca):
                    for b, cB in zip(eb, cb):
                        w = cA * cB
                        s += w * _ovlp(a, la, A, b, lb, B)
                        t += w * _kin(a, la, A, b, lb, B)
                        v += w * _nuc(a, la, A, b, lb, B, self.coords, self.charges)
                S[i, j] = S[j, i] = s
                T[i, j] = T[j, i] = t
                V[i, j] = V[j, i] = v
        return self._to_ao(S), self._to_ao(T), self._to_ao(V)

    def int2e(self):
        """Chemist (pq|rs) over AO functions, full (nao,)*4 (shell quartets, vectorised)."""
        shells = self.shells
        offs = np.cumsum([0] + [len(cart_components(sh.l)) for sh in shells])
        pairs = {}
        for i in range(len(shells)):
            for j in range(i + 1):
                pairs[i, j] = _ShellPair(shells[i], shells[j])
        n = self.ncart
        eri = np.zeros((n, n, n, n))
        keys = sorted(pairs)
        for ki, (i, j) in enumerate(keys):
            A = pairs[i, j]
            for (k, l) in keys[: ki + 1]:
                B = pairs[k, l]
                Ltot = A.L + B.L
                pq = A.p[:, None] * B.p[None, :]
                spq = A.p[:, None] + B.p[None, :]
                alpha = pq / spq
                PQ = A.P[:, None, :] - B.P[None, :, :]
                R = _r_table(Ltot, alpha, PQ)
                pref = 2 * math.pi ** 2.5 / (pq * np.sqrt(spq)) * A.K[:, None] * B.K[None, :]
                Rm = np.empty((len(A.tuv), len(B.tuv)) + pref.shape)
                for it, (t, u, v) in enumerate(A.tuv):
                    for jt, (tt, uu, vv) in enumerate(B.tuv):
                        sgn = -1.0 if (tt + uu + vv) % 2 else 1.0
                        Rm[it, jt] = sgn * R[t + tt, u + uu, v + vv] * pref
                blk = np.einsum("aiN,ijNM,bjM->ab", A.E, Rm, B.E, optimize=True)
                na, nb = A.ncomp
                nc, nd = B.ncomp
                blk = blk.reshape(na, nb, nc, nd)
                si, sj, sk, sl = (slice(offs[x], offs[x + 1]) for x in (i, j, k, l))
                eri[si, sj, sk, sl] = blk
                eri[sj, si, sk, sl] = blk.transpose(1, 0, 2, 3)
                eri[si, sj, sl, sk] = blk.transpose(0, 1, 3, 2)
                eri[sj, si, sl, sk] = blk.transpose(1, 0, 3, 2)
                eri[sk, sl, si, sj] = blk.transpose(2, 3, 0, 1)
                eri[sl, sk, si, sj] = blk.transpose(3, 2, 0, 1)
                eri[sk, sl, sj, si] = blk.transpose(2, 3, 1, 0)
                eri[sl, sk, sj, si] = blk.transpose(3, 2, 1, 0)
        C = self.c2ao
        return np.einsum("pqrs,pP,qQ,rR,sS->PQRS", eri, C, C, C, C, optimize=True)

    def int2e_reference(self):
        """Chemist (pq|rs), primitive by primitive (slow; kept as a cross-check of int2e)."""
        prims = self._prims()
        n = self.ncart
        # primitive pair data
        pairs = {}
        for i in range(n):
            A, la, ea, ca = prims[i]
            for j in range(i + 1):
                B, lb, eb, cb = prims[j]
                lst = []
                for a, cA in zip(ea, ca):
                    for b, cB in zip(eb, cb):
                        p = a + b
                        P = (a * A + b * B) / p
                        E = [[hermite_e(la[x], lb[x], t, A[x] - B[x], a, b) for t in range(la[x] + lb[x] + 1)]
                             for x in range(3)]
                        lst.append((p, P, cA * cB, E, [la[x] + lb[x] for x in range(3)]))
                pairs[i, j] = lst
        eri = np.zeros((n, n, n, n))
        keys = sorted(pairs)
        for ki, (i, j) in enumerate(keys):
            for (k, l) in keys[: ki + 1]:
                val = 0.0
                for p, P, w1, E1, L1 in pairs[i, j]:
                    for q, Q, w2, E2, L2 in pairs[k, l]:
                        alpha = p * q / (p + q)
                        R = hermite_r(L1[0] + L2[0], L1[1] + L2[1], L1[2] + L2[2], alpha, P - Q)
                        acc = 0.0
                        for t, u, v in itertools.product(range(L1[0] + 1), range(L1[1] + 1), range(L1[2] + 1)):
                            e1 = E1[0][t] * E1[1][u] * E1[2][v]
                            if e1 == 0.0:
                                continue
                            for tau, nu, phi in itertools.product(range(L2[0] + 1), range(L2[1] + 1),
                                                                  range(L2[2] + 1)):
                                e2 = E2[0][tau] * E2[1][nu] * E2[2][phi]
                                if e2 == 0.0:
                                    continue
                                sgn = -1.0 if (tau + nu + phi) % 2 else 1.0
                                acc += e1 * e2 * sgn * R[t + tau, u + nu, v + phi]
                        val += w1 * w2 * 2 * math.pi ** 2.5 / (p * q * math.sqrt(p + q)) * acc
                for (a_, b_) in ((i, j), (j, i)):
                    for (c_, d_) in ((k, l), (l, k)):
                        eri[a_, b_, c_, d_] = val
                        eri[c_, d_, a_, b_] = val
        C = self.c2ao
        return np.einsum("pqrs,pP,qQ,rR,sS->PQRS", eri, C, C, C, C, optimize=True)


def _ovlp(a, la, A, b, lb, B):
    p = a + b
    out = (math.pi / p) ** 1.5
    for x in range(3):
        out *= hermite_e(la[x], lb[x], 0, A[x] - B[x], a, b)
    return out


def _kin(a, la, A, b, lb, B):
    l, m, n = lb
    t0 = b * (2 * (l + m + n) + 3) * _ovlp(a, la, A, b, lb, B)
    t1 = -2 * b * b * (_ovlp(a, la, A, b, (l + 2, m, n), B) + _ovlp(a, la, A, b, (l, m + 2, n), B)
                       + _ovlp(a, la, A, b, (l, m, n + 2), B))
    t2 = -0.5 * (l * (l - 1) * _ovlp(a, la, A, b, (l - 2, m, n), B)
                 + m * (m - 1) * _ovlp(a, la, A, b, (l, m - 2, n), B)
                 + n * (n - 1) * _ovlp(a, la, A, b, (l, m, n - 2), B))
    return t0 + t1 + t2


def _nuc(a, la, A, b, lb, B, coords, charges):
    p = a + b
    P = (a * A + b * B) / p
    out = 0.0
    Ex = [hermite_e(la[0], lb[0], t, A[0] - B[0], a, b) for t in range(la[0] + lb[0] + 1)]
    Ey = [hermite_e(la[1], lb[1], u, A[1] - B[1], a, b) for u in range(la[1] + lb[1] + 1)]
    Ez = [hermite_e(la[2], lb[2], v, A[2] - B[2], a, b) for v in range(la[2] + lb[2] + 1)]
    for Cc, Zc in zip(coords, charges):
        R = hermite_r(len(Ex) - 1, len(Ey) - 1, len(Ez) - 1, p, P - Cc)
        acc = 0.0
        for t, ex in enumerate(Ex):
            for u, ey in enumerate(Ey):
                for v, ez in enumerate(Ez):
                    acc += ex * ey * ez * R[t, u, v]
        out += -Zc * 2 * math.pi / p * acc
    return out


class RHF:
    """Closed-shell Roothaan-Hall SCF with DIIS; exposes the fields RCCSDT reads."""

    def __init__(self, mol, conv_tol=1e-12, max_cycle=200):
        self.mol = mol
        self.conv_tol = conv_tol
        self.max_cycle = max_cycle
        S, T, V = mol.int1e()
        self.S, self.hcore = S, T + V
        self._eri = mol.int2e()
        self.e_tot = None

    def get_hcore(self, *args):
        return self.hcore

    def get_ovlp(self, *args):
        return self.S

    def energy_nuc(self):
        return self.mol.energy_nuc()

    def kernel(self):
        nocc = self.mol.nelectron // 2
        S, h, eri = self.S, self.hcore, self._eri
        w, U = np.linalg.eigh(S)
        X = U @ np.diag(w ** -0.5) @ U.T
        e, C = np.linalg.eigh(X.T @ h @ X)
        C = X @ C
        hist_F, hist_e = [], []
        e_old = 0.0
        for it in range(self.max_cycle):
            D = 2 * C[:, :nocc] @ C[:, :nocc].T
            F = h + np.einsum("pqrs,rs->pq", eri, D) - 0.5 * np.einsum("prqs,rs->pq", eri, D)
            e_elec = 0.5 * np.sum(D * (h + F))
            err = F @ D @ S - S @ D @ F
            hist_F.append(F)
            hist_e.append(X.T @ err @ X)
            hist_F, hist_e = hist_F[-8:], hist_e[-8:]
            if len(hist_F) > 1:
                nd = len(hist_F)
                B = -np.ones((nd + 1, nd + 1))
                B[nd, nd] = 0
                for i in range(nd):
                    for j in range(nd):
                        B[i, j] = np.sum(hist_e[i] * hist_e[j])
                rhs = np.zeros(nd + 1)
                rhs[nd] = -1
                c = np.linalg.solve(B, rhs)[:nd]
                F = sum(ci * Fi for ci, Fi in zip(c, hist_F))
            e, Cp = np.linalg.eigh(X.T @ F @ X)
            C = X @ Cp
            if abs(e_elec - e_old) < self.conv_tol and np.abs(err).max() < 1e-9:
                break
            e_old = e_elec
        D = 2 * C[:, :nocc] @ C[:, :nocc].T
        F = h + np.einsum("pqrs,rs->pq", eri, D) - 0.5 * np.einsum("prqs,rs->pq", eri, D)
        e, Cp = np.linalg.eigh(X.T @ F @ X)
        self.mo_coeff = X @ Cp
        self.mo_energy = e
        self.mo_occ = np.array([2.0] * nocc + [0.0] * (len(e) - nocc))
        Cocc = self.mo_coeff[:, :nocc]
        D = 2 * Cocc @ Cocc.T
        F = h + np.einsum("pqrs,rs->pq", eri, D) - 0.5 * np.einsum("prqs,rs->pq", eri, D)
        self.e_tot = 0.5 * np.sum(D * (h + F)) + self.energy_nuc()
        self.converged = True
        return self.e_tot
