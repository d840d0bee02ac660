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
