"""Small closed-shell test systems for the full-iteration tests (CPU and GPU)."""
import numpy as np

from oracle.fci_cc import fock, random_hamiltonian, symmetrize_t3
from oracle.t3_layout import canonical_triples


def canonical(h, g, nocc):
    """Rotate a random Hamiltonian to the basis where its closed-shell Fock matrix is diagonal."""
    F = fock(h, g, nocc)
    for _ in range(200):
        _, C = np.linalg.eigh(F)
        h2 = C.T @ h @ C
        g2 = np.einsum("pqrs,pP,qQ,rR,sS->PQRS", g, C, C, C, C, optimize=True)
        F2 = fock(h2, g2, nocc)
        if np.abs(F2 - np.diag(np.diag(F2))).max() < 1e-13:
            return h2, g2, np.diag(F2).copy()
        h, g, F = h2, g2, F2
    raise RuntimeError("no SCF convergence")


def molecule_like(nocc, nvir, seed):
    """(fock, pppp, mo_energy, h) of a weakly correlated random closed-shell system."""
    rng = np.random.default_rng(seed)
    h, g = random_hamiltonian(nocc + nvir, rng, scale_g=0.3, scale_h=0.2, nocc=nocc, gap=3.0)
    h, g, e = canonical(h, g, nocc)
    return fock(h, g, nocc), g, e, h


def random_amplitudes(nocc, nvir, seed, s1=0.05, s2=0.05, s3=0.02):
    rng = np.random.default_rng(seed)
    t1 = s1 * rng.standard_normal((nocc, nvir))
    t2 = s2 * rng.standard_normal((nocc, nocc, nvir, nvir))
    t2 = t2 + t2.transpose(1, 0, 3, 2)
    T3 = symmetrize_t3(s3 * rng.standard_normal((nocc,) * 3 + (nvir,) * 3))
    t3 = np.array([T3[i, j, k] for (i, j, k) in canonical_triples(nocc)])
    return t1, t2, t3


class DuckMF:
    """The fields of a mean-field object that RCCSDT reads."""

    def __init__(self, nocc, nvir, mo_energy):
        self.mo_coeff = np.eye(nocc + nvir)
        self.mo_occ = np.array([2.0] * nocc + [0.0] * nvir)
        self.mo_energy = mo_energy
        self.e_tot = 0.0


def with_fov(F, nocc, scale, seed):
    """F with a symmetric occupied-virtual block added (a non-canonical Fock matrix, so t1 != 0)."""
    rng = np.random.default_rng(seed)
    F = F.copy()
    X = scale * rng.standard_normal((nocc, F.shape[0] - nocc))
    F[:nocc, nocc:] += X
    F[nocc:, :nocc] += X.T
    return F


def oracle_kernel(nocc, F, g, e, cycles, level_shift=0.0, nvir_diis=None, diis=True, with_t3=True, space=6):
    """``cycles`` iterations of the reference kernel loop (distr_cc/rccsdt.py:1069-1148) on the oracle:
    MP2 start (unshifted denominators), oracle update_amps, then DIIS (oracle.diis) over the unified
    vector [t1, t2 over pairs i <= j, t3[:, :nd, :nd, :nd]] (run_diis_t3, rccsdt.py:1026-1067) or, with
    with_t3=False, over [t1, t2 pairs] only (the standard DIIS of do_diis_max_t = False, T3 not
    extrapolated).  Returns (t1, t2, t3, energy after every cycle)."""
    from oracle import rccsdt_lower as lo, rccsdt_tri as tri
    from oracle.diis import DIIS
    o = nocc
    v = F.shape[0] - o
    _, t1, t2 = tri.init_amps_rhf(o, F, e, g)
    n3 = o * (o + 1) * (o + 2) // 6
    t3 = np.zeros((n3, v, v, v))
    nd = v if nvir_diis is None else min(int(nvir_diis), v)
    d = DIIS(space) if diis else None
    pairs = [(i, j) for i in range(o) for j in range(i, o)]
    energies = []
    for _ in range(cycles):
        t1, t2, t3, _res = lo.update_amps(o, F, g, e, t1, t2, t3, level_shift=level_shift, nvir_diis=nd)
        if d is not None:
            parts = [t1.ravel(), np.concatenate([t2[i, j].ravel() for i, j in pairs])]
            if with_t3:
                parts.append(t3[:, :nd, :nd, :nd].ravel())
            x = d.update(np.concatenate(parts))
            t1 = x[:o * v].reshape(o, v).copy()
            off = o * v
            t2 = t2.copy()
            for i, j in pairs:
                blk = x[off:off + v * v].reshape(v, v)
                off += v * v
                t2[i, j] = blk
                t2[j, i] = blk.T
            if with_t3:
                t3 = t3.copy()
                t3[:, :nd, :nd, :nd] = x[off:].reshape(n3, nd, nd, nd)
        energies.append(tri.energy_rhf(o, F, g, t1, t2))
    return t1, t2, t3, energies
