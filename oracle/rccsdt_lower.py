"""Lower-order RCCSDT pieces around the triples update (oracle, test infrastructure only).

The reference takes these from PySCF 2.13.1 (``pyscf.cc.rccsdt``), called at
distr_cc/rccsdt.py:909 (update_t1_fock_eris), :912 (intermediates_t1t2),
:915 (compute_r1r2), :924 (r1r2_divide_e_) and :935 (intermediates_t3).
PySCF is not part of the reference tree, so this module restates the
published closed-shell CCSDT theory instead of PySCF's code:

  * T1 similarity transformation of the Hamiltonian (creators dressed by
    1 - t1^T, annihilators by 1 + t1^T);
  * spin-adapted residuals rho defined by
        exp(-T) H exp(T)|0> = E|0> + sum rho1 E_ai|0> + 1/2 sum rho2 E_ai E_bj|0>
                              + 1/6 sum rho3 E_ai E_bj E_ck|0> + ...
    whose Fock diagonal part is -(e_i + ... - e_a - ...) t, the
    normalisation the Jacobi update t += rho / D of rccsdt.py:917-971 needs;
  * r2 returned as half of rho2 (the reference adds r2.transpose(1,0,3,2),
    rccsdt.py:919).

Integrals are physicist <pq|rs> (creators p, q; annihilators r, s), the
convention of rccsdt.py:216-221 and of every hot-path function
(r1r2_add_t3_tri_, intermediates_t3_add_t3_tri, compute_r3_tri).

Pins (tests/test_oracle_lower.py): a brute-force determinant-space
evaluation of exp(-T) H exp(T)|0> (oracle/fci_cc.py) on random closed-shell
Hamiltonians and random t1, t2, t3 must reproduce rho1, rho2 and rho3 (up
to the closed-shell triples null space) built from these functions together
with the hot-path restatement in oracle/rccsdt_tri.py; for two electrons
the converged RCCSDT energy must equal the exact (FCI) correlation energy.
The split of rho2 between r2 and its transpose, and the level-shift
placement, are not fixed by that pin (parity unpinned at function level).
"""
import numpy as np

from oracle import rccsdt_tri as tri

_E = np.einsum


def _es(sub, *ops):
    return np.einsum(sub, *ops, optimize=True)


def core_hamiltonian(nocc, fock, pppp):
    """h = f - sum_k (2 <pk|qk> - <pk|kq>) over the doubly occupied orbitals."""
    o = slice(0, nocc)
    return fock - 2.0 * _es("pkqk->pq", pppp[:, o, :, o]) + _es("pkkq->pq", pppp[:, o, o, :])


def dressing_matrices(nocc, nmo, t1):
    """Creator (L) and annihilator (R) transforms of exp(-T1) ... exp(T1)."""
    th = np.zeros((nmo, nmo))
    th[nocc:, :nocc] = t1.T
    return np.eye(nmo) - th, np.eye(nmo) + th


def update_t1_fock_eris(nocc, fock, pppp, t1):
    """T1-dressed Fock matrix and integrals (call site rccsdt.py:909)."""
    nmo = fock.shape[0]
    L, R = dressing_matrices(nocc, nmo, t1)
    h = core_hamiltonian(nocc, fock, pppp)
    t1_eris = _es("Pp,Qq,pqrs,rR,sS->PQRS", L, L, pppp, R, R)
    o = slice(0, nocc)
    t1_fock = L @ h @ R + 2.0 * _es("pkqk->pq", t1_eris[:, o, :, o]) - _es("pkkq->pq", t1_eris[:, o, o, :])
    return t1_fock, t1_eris


def _blocks(nocc, t1_eris):
    nmo = t1_eris.shape[0]
    O, V = slice(0, nocc), slice(nocc, nmo)
    return O, V


def intermediates_t1t2(nocc, t1_fock, t1_eris, t2):
    """Intermediates of the singles/doubles residual (call site rccsdt.py:912)."""
    O, V = _blocks(nocc, t1_eris)
    g = t1_eris
    oovv = g[O, O, V, V]
    imds = {}
    imds["F_oo"] = t1_fock[O, O] + 2.0 * _es("klcd,ilcd->ki", oovv, t2) - _es("kldc,ilcd->ki", oovv, t2)
    imds["F_vv"] = t1_fock[V, V] - 2.0 * _es("klcd,klad->ac", oovv, t2) + _es("kldc,klad->ac", oovv, t2)
    imds["W_oooo"] = g[O, O, O, O] + _es("klcd,ijcd->klij", oovv, t2)
    imds["W_voov"] = (g[V, O, O, V] - 0.5 * _es("lkdc,ilda->akic", oovv, t2)
                      - 0.5 * _es("lkcd,ilad->akic", oovv, t2) + _es("lkdc,ilad->akic", oovv, t2))
    imds["W_vovo"] = g[V, O, V, O] - 0.5 * _es("lkcd,ilda->akci", oovv, t2)
    return imds


def compute_r1r2(nocc, t1_fock, t1_eris, t2, imds):
    """Singles residual and half of the doubles residual without T3 (call site rccsdt.py:915)."""
    O, V = _blocks(nocc, t1_eris)
    g, f = t1_eris, t1_fock
    fov = f[O, V]
    r1 = f[V, O].T.copy()
    r1 += 2.0 * _es("kc,kica->ia", fov, t2) - _es("kc,ikca->ia", fov, t2)
    r1 += 2.0 * _es("akcd,ikcd->ia", g[V, O, V, V], t2) - _es("akdc,ikcd->ia", g[V, O, V, V], t2)
    r1 += -2.0 * _es("lkci,klac->ia", g[O, O, V, O], t2) + _es("klci,klac->ia", g[O, O, V, O], t2)
    Wvoov, Wvovo = imds["W_voov"], imds["W_vovo"]
    r2 = 0.5 * g[V, V, O, O].transpose(2, 3, 0, 1)
    r2 = r2 + 0.5 * _es("klij,klab->ijab", imds["W_oooo"], t2)
    r2 += 0.5 * _es("abcd,ijcd->ijab", g[V, V, V, V], t2)
    r2 += _es("ac,ijcb->ijab", imds["F_vv"], t2) - _es("ki,kjab->ijab", imds["F_oo"], t2)
    r2 += 2.0 * _es("akic,kjcb->ijab", Wvoov, t2) - _es("akci,kjcb->ijab", Wvovo, t2)
    r2 -= _es("akic,kjbc->ijab", Wvoov, t2) + _es("bkci,kjac->ijab", Wvovo, t2)
    return r1, r2


def r1r2_divide_e(nocc, r1, r2, mo_energy, level_shift=0.0):
    """r1 /= e_i - e_a - shift, r2 /= e_i + e_j - e_a - e_b - 2 shift (call site rccsdt.py:924)."""
    eia = mo_energy[:nocc, None] - mo_energy[None, nocc:] - level_shift
    r1 /= eia
    r2 /= eia[:, None, :, None] + eia[None, :, None, :]
    return r1, r2


def intermediates_t3(nocc, t1_fock, t1_eris, t2):
    """T2-dressed intermediates of the triples residual (call site rccsdt.py:935).

    Returns F_oo, F_vv, W_oooo, W_ovvo, W_ovov, W_vooo, W_vvvo, W_vvvv in the
    index order compute_r3_tri reads them (rccsdt.py:388-426, 540-596,
    632-723).  W_ovvo is the spin-summed ring vertex 2<la|di> - <la|id>,
    W_ovov the exchange one <lb|id>, each dressed by T2 in full (the T2 T3
    terms of the triples equation carry no 1/2).
    """
    O, V = _blocks(nocc, t1_eris)
    g, f = t1_eris, t1_fock
    oovv = g[O, O, V, V]
    d = {}
    d["F_oo"] = f[O, O] + 2.0 * _es("lmde,imde->li", oovv, t2) - _es("lmed,imde->li", oovv, t2)
    d["F_vv"] = f[V, V] - 2.0 * _es("lmde,lmae->ad", oovv, t2) + _es("lmed,lmae->ad", oovv, t2)
    d["W_oooo"] = g[O, O, O, O] + _es("lmde,ijde->lmij", oovv, t2)
    d["W_vvvv"] = g[V, V, V, V] + _es("lmde,lmab->abde", oovv, t2)
    h_ovvo = (g[O, V, V, O] - _es("mled,imea->ladi", oovv, t2) - _es("mlde,imae->ladi", oovv, t2)
              + 2.0 * _es("mled,imae->ladi", oovv, t2))
    h_ovov = g[O, V, O, V] - _es("mlde,imeb->lbid", oovv, t2)
    d["W_ovvo"] = 2.0 * h_ovvo - h_ovov.transpose(0, 1, 3, 2)
    d["W_ovov"] = h_ovov
    gvovv = g[V, O, V, V]
    d["W_vvvo"] = (g[V, V, V, O]
                   + 2.0 * _es("amde,jmbe->abdj", gvovv, t2) - _es("amde,jmeb->abdj", gvovv, t2)
                   - _es("amed,jmbe->abdj", gvovv, t2) - _es("bmed,jmea->abdj", gvovv, t2)
                   + _es("mndj,mnab->abdj", g[O, O, V, O], t2)
                   - _es("md,mjab->abdj", f[O, V], t2))
    gooov = g[O, O, O, V]
    d["W_vooo"] = (g[V, O, O, O]
                   + _es("alef,ijef->alij", g[V, O, V, V], t2)
                   + 2.0 * _es("lmje,imae->alij", gooov, t2) - _es("lmje,imea->alij", gooov, t2)
                   - _es("lmej,imae->alij", g[O, O, V, O], t2) - _es("lmei,jmea->alij", g[O, O, V, O], t2))
    return {k: np.ascontiguousarray(x) for k, x in d.items()}


def update_amps(nocc, fock, pppp, mo_energy, t1, t2, t3_local, level_shift=0.0, nvir_diis=None):
    """One Jacobi iteration in the order of update_amps_rccsdt_tri_ (rccsdt.py:890-977).

    Returns new (t1, t2, t3_local) and [|r1|, |r2|, |r3|]; the inputs are not
    modified.
    """
    nvir = fock.shape[0] - nocc
    t1_fock, t1_eris = update_t1_fock_eris(nocc, fock, pppp, t1)               # :909
    imds = intermediates_t1t2(nocc, t1_fock, t1_eris, t2)                       # :912
    r1, r2 = compute_r1r2(nocc, t1_fock, t1_eris, t2, imds)                     # :915
    tri.r1r2_add_t3_tri(nocc, nvir, t1_fock, t1_eris, t3_local, r1, r2)         # :916
    r2 = r2 + r2.transpose(1, 0, 3, 2)                                          # :919
    r1r2_divide_e(nocc, r1, r2, mo_energy, level_shift)                         # :924
    res = [np.linalg.norm(r1), np.linalg.norm(r2)]
    t1 = t1 + r1
    t2 = t2 + r2
    imds3 = intermediates_t3(nocc, t1_fock, t1_eris, t2)                        # :935
    tri.intermediates_t3_add_t3_tri(nocc, nvir, t1_eris, t3_local, imds3["W_vooo"], imds3["W_vvvo"])
    r3 = tri.compute_r3_tri(nocc, nvir, imds3, t2, t3_local)                    # :941
    t3_new, _, norm_sq, _ = tri.t3_update(nocc, r3, t3_local, mo_energy, level_shift, nvir_diis)
    res.append(np.sqrt(norm_sq))
    return t1, t2, t3_new, res
