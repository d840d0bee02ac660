"""The lower-order RCCSDT oracle and the hot-path restatement against a brute-force
evaluation of exp(-T) H exp(T)|0> in the determinant space.

A wrong sign, index or factor anywhere in the T1 dressing, the singles and
doubles residual, the T2-dressed triples intermediates, T3 -> r1/r2, T3 -> W
or the triples residual itself changes some determinant coefficient of the
image below, because the brute force contracts nothing by hand.
"""
import numpy as np
import pytest
import scipy.sparse.linalg as spl

from oracle import rccsdt_lower as lo
from oracle import rccsdt_tri as tri
from oracle.fci_cc import DetSpace, fock, random_hamiltonian, symmetrize_t3
from oracle.t3_layout import canonical_triples, full_t3_from_tri


def _system(nocc, nvir, seed):
    rng = np.random.default_rng(seed)
    nmo = nocc + nvir
    h, g = random_hamiltonian(nmo, rng)
    t1 = 0.1 * rng.standard_normal((nocc, nvir))
    t2 = 0.1 * rng.standard_normal((nocc, nocc, nvir, nvir))
    t2 = t2 + t2.transpose(1, 0, 3, 2)
    T3 = symmetrize_t3(0.05 * rng.standard_normal((nocc,) * 3 + (nvir,) * 3))
    t3_local = np.array([T3[i, j, k] for (i, j, k) in canonical_triples(nocc)])
    return h, g, t1, t2, T3, t3_local


def _hf_energy(h, g, nocc):
    o = slice(0, nocc)
    return 2 * np.trace(h[o, o]) + 2 * np.einsum("ijij", g[o, o, o, o]) - np.einsum("ijji", g[o, o, o, o])


def test_t1_dressing_is_the_similarity_transform():
    nocc, nvir = 2, 3
    h, g, t1, *_ = _system(nocc, nvir, 11)
    ds = DetSpace(nocc + nvir, nocc)
    t1_fock, t1_eris = lo.update_t1_fock_eris(nocc, fock(h, g, nocc), g, t1)
    h_dressed = lo.core_hamiltonian(nocc, t1_fock, t1_eris)
    T1 = ds.cluster_operator(t1).tocsc()
    H = ds.hamiltonian(h, g)
    ref = (spl.expm(-T1) @ H @ spl.expm(T1)).toarray()
    got = ds.hamiltonian(h_dressed, t1_eris).toarray()
    assert np.abs(ref - got).max() < 1e-12 * np.abs(ref).max()


@pytest.mark.parametrize("nocc,nvir,seed", [(3, 3, 5), (2, 4, 6), (3, 4, 7), (1, 4, 8)])
def test_residuals_match_brute_force(nocc, nvir, seed):
    h, g, t1, t2, T3, t3_local = _system(nocc, nvir, seed)
    ds = DetSpace(nocc + nvir, nocc)
    F = fock(h, g, nocc)
    hbar = ds.hbar_vector(h, g, t1, t2, T3)

    t1_fock, t1_eris = lo.update_t1_fock_eris(nocc, F, g, t1)
    imds = lo.intermediates_t1t2(nocc, t1_fock, t1_eris, t2)
    r1, r2 = lo.compute_r1r2(nocc, t1_fock, t1_eris, t2, imds)
    tri.r1r2_add_t3_tri(nocc, nvir, t1_fock, t1_eris, t3_local, r1, r2)
    imds3 = lo.intermediates_t3(nocc, t1_fock, t1_eris, t2)
    tri.intermediates_t3_add_t3_tri(nocc, nvir, t1_eris, t3_local, imds3["W_vooo"], imds3["W_vvvo"])
    r3 = full_t3_from_tri(tri.compute_r3_tri(nocc, nvir, imds3, t2, t3_local), nocc)
    energy = _hf_energy(h, g, nocc) + tri.energy_rhf(nocc, F, g, t1, t2)

    image = ds.excitation_image(energy, r1, r2 + r2.transpose(1, 0, 3, 2), r3)
    for rank in range(4):
        m = ds.rank_mask(rank)
        if not m.any():
            continue
        scale = np.abs(hbar[m]).max()
        assert np.abs(hbar[m] - image[m]).max() < 1e-12 * max(scale, 1.0), rank


def test_projection_removes_only_the_null_space():
    """P3_full changes r3 only by a part whose determinant image is zero."""
    nocc, nvir = 3, 3
    *_, T3, t3_local = _system(nocc, nvir, 12)
    ds = DetSpace(nocc + nvir, nocc)
    projected = full_t3_from_tri(tri.project_p3_full(t3_local.copy()), nocc)
    m = ds.rank_mask(3)
    a = ds.excitation_image(0.0, None, None, T3)[m]
    b = ds.excitation_image(0.0, None, None, projected)[m]
    assert np.abs(a - b).max() < 1e-13


def _solve(nocc, nvir, h, g, mo_energy, iters=200, tol=1e-11):
    F = fock(h, g, nocc)
    _, t1, t2 = tri.init_amps_rhf(nocc, F, mo_energy, g)
    t3 = np.zeros((nocc * (nocc + 1) * (nocc + 2) // 6, nvir, nvir, nvir))
    for _ in range(iters):
        t1, t2, t3, res = lo.update_amps(nocc, F, g, mo_energy, t1, t2, t3)
        if max(res) < tol:
            break
    return F, t1, t2, t3, res


def _canonical(h, g, nocc):
    """Rotate a random Hamiltonian to its RHF-like canonical basis (Fock diagonal)."""
    F = fock(h, g, nocc)
    for _ in range(200):   # Roothaan iterations in the orbital basis
        e, C = np.linalg.eigh(F)
        h2 = C.T @ h @ C
        g2 = np.einsum("pqrs,pP,qQ,rR,sS->PQRS", g, C, C, C, C, optimize=True)
        F2 = fock(h2, g2, nocc)
        if np.abs(F2 - np.diag(np.diag(F2))).max() < 1e-13:
            return h2, g2, np.diag(F2).copy()
        h, g, F = h2, g2, F2
    raise RuntimeError("no SCF convergence")


def test_two_electron_ccsdt_is_exact():
    """For two electrons RCCSDT is exact (tests/test_rccsdt.py:221-230 uses He)."""
    nocc, nvir = 1, 4
    rng = np.random.default_rng(21)
    h, g = random_hamiltonian(nocc + nvir, rng, scale_g=0.3, scale_h=0.2, nocc=nocc, gap=3.0)
    h, g, mo_energy = _canonical(h, g, nocc)
    F, t1, t2, t3, res = _solve(nocc, nvir, h, g, mo_energy)
    assert max(res) < 1e-10
    e_cc = tri.energy_rhf(nocc, F, g, t1, t2)
    ds = DetSpace(nocc + nvir, nocc)
    H = ds.hamiltonian(h, g).toarray()
    w, U = np.linalg.eigh(H)
    root = np.argmax(np.abs(U[ds.ref]))
    assert abs((w[root] - _hf_energy(h, g, nocc)) - e_cc) < 1e-10


def test_converged_amplitudes_solve_ccsdt():
    """At the fixed point of update_amps the brute-force residual vanishes."""
    nocc, nvir = 2, 3
    rng = np.random.default_rng(22)
    h, g = random_hamiltonian(nocc + nvir, rng, scale_g=0.3, scale_h=0.2, nocc=nocc, gap=3.0)
    h, g, mo_energy = _canonical(h, g, nocc)
    F, t1, t2, t3, res = _solve(nocc, nvir, h, g, mo_energy)
    assert max(res) < 1e-10
    ds = DetSpace(nocc + nvir, nocc)
    hbar = ds.hbar_vector(h, g, t1, t2, full_t3_from_tri(t3, nocc))
    for rank in (1, 2, 3):
        assert np.abs(hbar[ds.rank_mask(rank)]).max() < 1e-9
    assert abs(hbar[ds.ref] - _hf_energy(h, g, nocc) - tri.energy_rhf(nocc, F, g, t1, t2)) < 1e-12
