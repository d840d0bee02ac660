"""Block-loop restatement of the RCCSDT triples update (oracle, test infrastructure only).

Each function follows the cited lines of ``distr_cc/rccsdt.py`` einsum by
einsum, at one rank (every canonical block is local, lexicographic order).
All arithmetic is FP64.  ``einsum(..., alpha, beta)`` means
``out = beta*out + alpha*einsum(...)`` as in PySCF's ``_einsum`` wrapper.

Pins (tests/test_oracle_*.py):
  * the independent full-tensor restatement in ``rccsdt_dense`` (different
    derivation: dense einsums over the ordered-triple tensor) must agree to
    round-off for every function here;
  * invariance of ``compute_r3_tri`` to ``batch_size`` (the reference's own
    invariance test, tests/test_rccsdt.py:85-101);
  * the r3 of a dense T3 that has the simultaneous-permutation symmetry must
    itself have it;
  * ``P3_full`` projection idempotence; denominators by closed form.
"""
import numpy as np

from oracle.t3_layout import canonical_triples, nocc3, take_t3
from oracle.spin_sum import t3_spin_sum, t3_transpose_add, t3_triple_sym


def einsum(subscripts, a, b, out, alpha=1.0, beta=1.0):
    """out <- beta*out + alpha*einsum(subscripts, a, b) (PySCF ``_einsum`` semantics)."""
    res = np.einsum(subscripts, a, b, optimize=True)
    if beta == 0.0:
        out[...] = alpha * res
    else:
        out *= beta
        out += alpha * res
    return out


def distinct_perms(i, j, k):
    """distr_cc/rccsdt.py:297-304 (distinct ordered permutations of a canonical triple)."""
    if i < j < k:
        return [(i, j, k), (i, k, j), (j, i, k), (j, k, i), (k, i, j), (k, j, i)]
    if i == j < k:
        return [(i, j, k), (i, k, j), (k, i, j)]
    if i < j == k:
        return [(i, j, k), (j, i, k), (j, k, i)]
    return [(i, j, k)]


def r1r2_add_t3_tri(nocc, nvir, t1_fock, t1_eris, t3_local, r1, r2):
    """distr_cc/rccsdt.py:253-321 (T3 -> r1, r2), single rank; r1, r2 updated in place."""
    o = nocc
    r1_loc = np.zeros_like(r1)
    r2_loc = np.zeros_like(r2)
    for (i, j, k) in canonical_triples(nocc):
        t = t3_spin_sum(take_t3(t3_local, i, j, k, nocc), "P3_422")  # :269-271
        if i < j and j < k:                                            # :272-278
            einsum('bc,abc->a', t1_eris[j, k, o:, o:], t, r1_loc[i, :], 0.5, 1.0)
            einsum('cb,acb->a', t1_eris[j, k, o:, o:], t, r1_loc[i, :], 0.5, 1.0)
            einsum('bc,bac->a', t1_eris[i, k, o:, o:], t, r1_loc[j, :], 0.5, 1.0)
            einsum('cb,cab->a', t1_eris[i, k, o:, o:], t, r1_loc[j, :], 0.5, 1.0)
            einsum('bc,bca->a', t1_eris[i, j, o:, o:], t, r1_loc[k, :], 0.5, 1.0)
            einsum('cb,cba->a', t1_eris[i, j, o:, o:], t, r1_loc[k, :], 0.5, 1.0)
        elif i == j and j < k:                                         # :279-282
            einsum('bc,abc->a', t1_eris[j, k, o:, o:], t, r1_loc[i, :], 0.5, 1.0)
            einsum('cb,acb->a', t1_eris[j, k, o:, o:], t, r1_loc[i, :], 0.5, 1.0)
            einsum('bc,bca->a', t1_eris[i, j, o:, o:], t, r1_loc[k, :], 0.5, 1.0)
        elif i < j and j == k:                                         # :283-286
            einsum('bc,abc->a', t1_eris[j, k, o:, o:], t, r1_loc[i, :], 0.5, 1.0)
            einsum('bc,bac->a', t1_eris[i, k, o:, o:], t, r1_loc[j, :], 0.5, 1.0)
            einsum('cb,cab->a', t1_eris[i, k, o:, o:], t, r1_loc[j, :], 0.5, 1.0)
    r1 += r1_loc                                                       # :293
    for (i, j, k) in canonical_triples(nocc):                          # :296-312
        for (pi, pj, pk) in distinct_perms(i, j, k):
            z = t3_spin_sum(take_t3(t3_local, pi, pj, pk, nocc), "P3_201")
            einsum("a,abc->bc", t1_fock[pi, o:], z, r2_loc[pj, pk], 0.5, 1.0)
            einsum("cad,dba->bc", t1_eris[o:, pi, o:, o:], z, r2_loc[pj, pk], 1.0, 1.0)
            einsum("la,abc->lbc", t1_eris[pk, pi, :o, o:], z, r2_loc[pj], -1.0, 1.0)
    r2 += r2_loc                                                       # :316
    return r1, r2


def intermediates_t3_add_t3_tri(nocc, nvir, t1_eris, t3_local, W_vooo, W_vvvo):
    """distr_cc/rccsdt.py:323-366 (T3 -> W_vooo, W_vvvo), single rank; in place."""
    o = nocc
    W_vooo_loc = np.zeros_like(W_vooo)
    W_vvvo_loc = np.zeros_like(W_vvvo)
    for (i, j, k) in canonical_triples(nocc):
        for (pi, pj, pk) in distinct_perms(i, j, k):
            z = t3_spin_sum(take_t3(t3_local, pi, pj, pk, nocc), "P3_201")
            einsum('mde,ead->am', t1_eris[:o, pi, o:, o:], z, W_vooo_loc[:, :, pj, pk], 1.0, 1.0)
            einsum('de,eba->abd', t1_eris[pk, pi, o:, o:], z, W_vvvo_loc[:, :, :, pj], -1.0, 1.0)
    W_vooo += W_vooo_loc
    W_vvvo += W_vvvo_loc
    return W_vooo, W_vvvo


def compute_r3_tri(nocc, nvir, imds, t2, t3_local, batch_size=100, targets=None):
    """distr_cc/rccsdt.py:368-433 -> r3 in canonical-block form (before projection).

    ``imds`` is a mapping with F_oo, F_vv, W_oooo, W_ovvo, W_ovov, W_vooo,
    W_vvvo, W_vvvv (the shapes of rccsdt.py:380,445,625).  ``targets``
    optionally restricts the computed target blocks to a subset of canonical
    indices (the other blocks of the result are left zero); it exists so a
    bounded sample of a large residual can be timed.
    """
    F_vv, W_vooo, W_vvvo, W_vvvv = imds['F_vv'], imds['W_vooo'], imds['W_vvvo'], imds['W_vvvv']
    r3 = np.zeros_like(t3_local)
    tset = None if targets is None else set(int(x) for x in targets)
    for idx, (i, j, k) in enumerate(canonical_triples(nocc)):       # :385-402
        if tset is not None and idx not in tset:
            continue
        r = r3[idx]
        t = t3_local[idx]
        einsum('abd,dc->abc', W_vvvo[..., j], t2[i, k], r, 1.0, 0.0)
        einsum('acd,db->abc', W_vvvo[..., k], t2[i, j], r, 1.0, 1.0)
        einsum('bad,dc->abc', W_vvvo[..., i], t2[j, k], r, 1.0, 1.0)
        einsum('bcd,da->abc', W_vvvo[..., k], t2[j, i], r, 1.0, 1.0)
        einsum('cad,db->abc', W_vvvo[..., i], t2[k, j], r, 1.0, 1.0)
        einsum('cbd,da->abc', W_vvvo[..., j], t2[k, i], r, 1.0, 1.0)
        einsum('al,lbc->abc', W_vooo[:, :, i, j], t2[:, k], r, -1.0, 1.0)
        einsum('al,lcb->abc', W_vooo[:, :, i, k], t2[:, j], r, -1.0, 1.0)
        einsum('bl,lac->abc', W_vooo[:, :, j, i], t2[:, k], r, -1.0, 1.0)
        einsum('bl,lca->abc', W_vooo[:, :, j, k], t2[:, i], r, -1.0, 1.0)
        einsum('cl,lab->abc', W_vooo[:, :, k, i], t2[:, j], r, -1.0, 1.0)
        einsum('cl,lba->abc', W_vooo[:, :, k, j], t2[:, i], r, -1.0, 1.0)
        einsum('ad,dbc->abc', F_vv, t, r, 1.0, 1.0)
        einsum('bd,adc->abc', F_vv, t, r, 1.0, 1.0)
        einsum('cd,abd->abc', F_vv, t, r, 1.0, 1.0)
    compute_oooo_oovv_contraction(nocc, nvir, imds, t3_local, r3, batch_size, tset)   # :414
    for idx in range(r3.shape[0]):                                  # :421-426
        if tset is not None and idx not in tset:
            continue
        t = t3_local[idx]
        r = r3[idx]
        einsum('abde,dec->abc', W_vvvv, t, r, 1.0, 1.0)
        einsum('acde,dbe->abc', W_vvvv, t, r, 1.0, 1.0)
        einsum('bcde,ade->abc', W_vvvv, t, r, 1.0, 1.0)
    return r3


def compute_oooo_oovv_contraction(nocc, nvir, imds, t3_local, r3, batch_size=100, tset=None):
    """distr_cc/rccsdt.py:435-614 at one rank (t3_data = t3_local[batch], :523-526)."""
    W_ovvo, W_ovov = np.ascontiguousarray(imds['W_ovvo']), imds['W_ovov']
    ijk_list = np.array(canonical_triples(nocc), dtype=np.int32).reshape(-1, 3)
    n3 = len(ijk_list)
    batch_size = min(batch_size, n3)
    targets = canonical_triples(nocc)
    # with a target subset, a source block is only used when one of its index pairs is an index pair
    # of a sampled target (every ring condition below compares two target indices with two source
    # indices); the others add exactly nothing, so their transposes are not formed
    pairs = None
    if tset is not None:
        pairs = set()
        for q in tset:
            i, j, k = targets[q]
            pairs.update(((j, k), (i, k), (i, j)))
    for b0 in range(0, n3, batch_size):
        b1 = min(b0 + batch_size, n3)
        batch_ijk = ijk_list[b0:b1]
        t3_data = t3_local[b0:b1]
        compute_r3_Foo_Woooo_contribution_batch(nocc, r3, t3_data, batch_ijk, imds, tset)
        for idx, (o00, o10, o20) in enumerate(batch_ijk):
            if pairs is not None and not pairs.intersection(((o00, o10), (o00, o20), (o10, o20))):
                continue
            t = t3_data[idx]
            t_2 = t3_transpose_add(t)                                 # :534
            p0, p1, p2 = t3_triple_sym(t)                             # :535
            for tgt, (i, j, k) in enumerate(targets):
                if tset is not None and tgt not in tset:
                    continue
                r = r3[tgt]
                # part 3 (:540-558)
                if j == o00 and k == o10:
                    einsum('bd,dca->abc', W_ovov[o20, :, i, :], t, r, -1.0, 1.0)
                    einsum('ad,dcb->abc', W_ovov[o20, :, i, :], t_2, r, -0.5, 1.0)
                    einsum('cd,bda->abc', W_ovov[o20, :, i, :], t, r, -1.0, 1.0)
                    einsum('ad,bcd->abc', W_ovvo[o20, :, :, i], p0, r, 0.5, 1.0)
                if o10 != o20 and j == o00 and k == o20:
                    einsum('bd,dac->abc', W_ovov[o10, :, i, :], t, r, -1.0, 1.0)
                    einsum('cd,bad->abc', W_ovov[o10, :, i, :], t, r, -1.0, 1.0)
                    einsum('ad,bcd->abc', W_ovov[o10, :, i, :], t_2, r, -0.5, 1.0)
                    einsum('ad,bdc->abc', W_ovvo[o10, :, :, i], p1, r, 0.5, 1.0)
                if o00 != o10 and j == o10 and k == o20:
                    einsum('bd,adc->abc', W_ovov[o00, :, i, :], t, r, -1.0, 1.0)
                    einsum('ad,bdc->abc', W_ovov[o00, :, i, :], t_2, r, -0.5, 1.0)
                    einsum('cd,abd->abc', W_ovov[o00, :, i, :], t, r, -1.0, 1.0)
                    einsum('ad,dbc->abc', W_ovvo[o00, :, :, i], p2, r, 0.5, 1.0)
                # part 4 (:559-577)
                if i == o00 and k == o10:
                    einsum('ad,dcb->abc', W_ovov[o20, :, j, :], t, r, -1.0, 1.0)
                    einsum('bd,dca->abc', W_ovov[o20, :, j, :], t_2, r, -0.5, 1.0)
                    einsum('cd,adb->abc', W_ovov[o20, :, j, :], t, r, -1.0, 1.0)
                    einsum('bd,acd->abc', W_ovvo[o20, :, :, j], p0, r, 0.5, 1.0)
                if o10 != o20 and i == o00 and k == o20:
                    einsum('ad,dbc->abc', W_ovov[o10, :, j, :], t, r, -1.0, 1.0)
                    einsum('cd,abd->abc', W_ovov[o10, :, j, :], t, r, -1.0, 1.0)
                    einsum('bd,acd->abc', W_ovov[o10, :, j, :], t_2, r, -0.5, 1.0)
                    einsum('bd,adc->abc', W_ovvo[o10, :, :, j], p1, r, 0.5, 1.0)
                if o00 != o10 and i == o10 and k == o20:
                    einsum('ad,bdc->abc', W_ovov[o00, :, j, :], t, r, -1.0, 1.0)
                    einsum('bd,adc->abc', W_ovov[o00, :, j, :], t_2, r, -0.5, 1.0)
                    einsum('cd,bad->abc', W_ovov[o00, :, j, :], t, r, -1.0, 1.0)
                    einsum('bd,dac->abc', W_ovvo[o00, :, :, j], p2, r, 0.5, 1.0)
                # part 5 (:578-596)
                if i == o00 and j == o10:
                    einsum('bd,adc->abc', W_ovov[o20, :, k, :], t, r, -1.0, 1.0)
                    einsum('ad,dbc->abc', W_ovov[o20, :, k, :], t, r, -1.0, 1.0)
                    einsum('cd,dba->abc', W_ovov[o20, :, k, :], t_2, r, -0.5, 1.0)
                    einsum('cd,abd->abc', W_ovvo[o20, :, :, k], p0, r, 0.5, 1.0)
                if o10 != o20 and i == o00 and j == o20:
                    einsum('cd,abd->abc', W_ovov[o10, :, k, :], t_2, r, -0.5, 1.0)
                    einsum('bd,acd->abc', W_ovov[o10, :, k, :], t, r, -1.0, 1.0)
                    einsum('ad,dcb->abc', W_ovov[o10, :, k, :], t, r, -1.0, 1.0)
                    einsum('cd,adb->abc', W_ovvo[o10, :, :, k], p1, r, 0.5, 1.0)
                if o00 != o10 and i == o10 and j == o20:
                    einsum('bd,cad->abc', W_ovov[o00, :, k, :], t, r, -1.0, 1.0)
                    einsum('ad,cdb->abc', W_ovov[o00, :, k, :], t, r, -1.0, 1.0)
                    einsum('cd,adb->abc', W_ovov[o00, :, k, :], t_2, r, -0.5, 1.0)
                    einsum('cd,dab->abc', W_ovvo[o00, :, :, k], p2, r, 0.5, 1.0)
    return r3


def compute_r3_Foo_Woooo_contribution_batch(nocc, r3, t3_data, batch_ijk, imds, tset=None):
    """distr_cc/rccsdt.py:616-732 (six prefactor matrices + six GEMMs)."""
    F_oo, W_oooo = imds['F_oo'], imds['W_oooo']
    targets = canonical_triples(nocc)
    pfs = np.zeros((6, len(batch_ijk), len(targets)))
    for idx, (o00, o10, o20) in enumerate(batch_ijk):
        for q, (i, j, k) in enumerate(targets):
            if tset is not None and q not in tset:
                continue
            # part 0 (:632-650)
            if k == o20:
                pfs[0, idx, q] += W_oooo[o00, o10, i, j]
            if o10 != o20 and k == o10:
                pfs[1, idx, q] += W_oooo[o00, o20, i, j]
            if o00 != o10 and k == o20:
                pfs[2, idx, q] += W_oooo[o10, o00, i, j]
            if o10 != o20 and k == o10:
                pfs[3, idx, q] += W_oooo[o20, o00, i, j]
            if o00 != o10 and k == o00:
                pfs[4, idx, q] += W_oooo[o10, o20, i, j]
            if o00 != o10 and o10 != o20 and k == o00:
                pfs[5, idx, q] += W_oooo[o20, o10, i, j]
            # part 1 (:652-670)
            if j == o20:
                pfs[1, idx, q] += W_oooo[o00, o10, i, k]
            if o10 != o20 and j == o10:
                pfs[0, idx, q] += W_oooo[o00, o20, i, k]
            if o00 != o10 and j == o20:
                pfs[4, idx, q] += W_oooo[o10, o00, i, k]
            if o10 != o20 and j == o10:
                pfs[5, idx, q] += W_oooo[o20, o00, i, k]
            if o00 != o10 and j == o00:
                pfs[2, idx, q] += W_oooo[o10, o20, i, k]
            if o00 != o10 and o10 != o20 and j == o00:
                pfs[3, idx, q] += W_oooo[o20, o10, i, k]
            # part 2 (:672-690)
            if i == o20:
                pfs[3, idx, q] += W_oooo[o00, o10, j, k]
            if o10 != o20 and i == o10:
                pfs[2, idx, q] += W_oooo[o00, o20, j, k]
            if o00 != o10 and i == o20:
                pfs[5, idx, q] += W_oooo[o10, o00, j, k]
            if o10 != o20 and i == o10:
                pfs[4, idx, q] += W_oooo[o20, o00, j, k]
            if o00 != o10 and i == o00:
                pfs[0, idx, q] += W_oooo[o10, o20, j, k]
            if o00 != o10 and o10 != o20 and i == o00:
                pfs[1, idx, q] += W_oooo[o20, o10, j, k]
            # part 3 (:692-701)
            if j == o00 and k == o10:
                pfs[3, idx, q] -= F_oo[o20, i]
            if o10 != o20 and j == o00 and k == o20:
                pfs[2, idx, q] -= F_oo[o10, i]
            if o00 != o10 and j == o10 and k == o20:
                pfs[0, idx, q] -= F_oo[o00, i]
            # part 4 (:703-712)
            if i == o00 and k == o10:
                pfs[1, idx, q] -= F_oo[o20, j]
            if o10 != o20 and i == o00 and k == o20:
                pfs[0, idx, q] -= F_oo[o10, j]
            if o00 != o10 and i == o10 and k == o20:
                pfs[2, idx, q] -= F_oo[o00, j]
            # part 5 (:714-723)
            if i == o00 and j == o10:
                pfs[0, idx, q] -= F_oo[o20, k]
            if o10 != o20 and i == o00 and j == o20:
                pfs[1, idx, q] -= F_oo[o10, k]
            if o00 != o10 and i == o10 and j == o20:
                pfs[4, idx, q] -= F_oo[o00, k]
    if tset is None:
        out, sel, src = r3, pfs, t3_data
    else:
        # only the sampled target columns of pfs are nonzero, and only the source rows that meet one
        # of them: the GEMMs run over those (the skipped terms are exact zeros)
        qs = np.array(sorted(tset), dtype=np.int64)
        ps = np.nonzero(np.any(pfs[:, :, qs] != 0.0, axis=(0, 2)))[0]
        if len(ps) == 0:
            return r3
        out, sel, src = r3[qs], pfs[:, ps][:, :, qs], t3_data[ps]
    einsum('pq,pabc->qabc', sel[0], src, out, 1.0, 1.0)            # :726-731
    einsum('pq,pacb->qabc', sel[1], src, out, 1.0, 1.0)
    einsum('pq,pbac->qabc', sel[2], src, out, 1.0, 1.0)
    einsum('pq,pabc->qcab', sel[3], src, out, 1.0, 1.0)
    einsum('pq,pabc->qbca', sel[4], src, out, 1.0, 1.0)
    einsum('pq,pabc->qcba', sel[5], src, out, 1.0, 1.0)
    if tset is not None:
        r3[qs] = out
    return r3


def project_p3_full(r3):
    """distr_cc/rccsdt.py:947 -> src/rccsdt.c:131-233 with P3_full, alpha=-1/6, beta=1."""
    return t3_spin_sum(r3, "P3_full", -1.0 / 6.0, 1.0)


def divide_e(nocc, r3, mo_energy, level_shift=0.0):
    """distr_cc/rccsdt.py:734-744."""
    eia = mo_energy[:nocc, None] - mo_energy[None, nocc:] - level_shift
    out = np.empty_like(r3)
    for idx, (i, j, k) in enumerate(canonical_triples(nocc)):
        d = eia[i, :, None, None] + eia[j, None, :, None] + eia[k, None, None, :]
        out[idx] = r3[idx] / d
    return out


def t3_update(nocc, r3_raw, t3_local, mo_energy, level_shift=0.0, nvir_diis=None):
    """distr_cc/rccsdt.py:947-971: projection, denominator, norms, t3 += r3.

    Returns (t3_new, r3_final, norm_sq, active_norm_sq).
    """
    r3 = divide_e(nocc, project_p3_full(r3_raw), mo_energy, level_shift)
    nvir = r3.shape[-1]
    norm_sq = float(np.linalg.norm(r3) ** 2)
    nd = nvir if nvir_diis is None else min(int(nvir_diis), nvir)
    if nd < nvir:
        act = float(np.linalg.norm(r3[:, :nd, :nd, :nd]) ** 2)
    else:
        act = norm_sq
    return t3_local + r3, r3, norm_sq, act


def init_amps_rhf(nocc, fock, mo_energy, pppp):
    """distr_cc/rccsdt.py:197-251: MP2 t1, t2 and energy; T3 = 0."""
    o = nocc
    eia = mo_energy[:o, None] - mo_energy[None, o:]
    eijab = eia[:, None, :, None] + eia[None, :, None, :]
    t1 = fock[:o, o:] / eia
    t2 = pppp[:o, :o, o:, o:] / eijab
    tau = t2 + np.einsum("ia,jb->ijab", t1, t1)
    e = 2.0 * np.einsum("ijab,ijab->", pppp[:o, :o, o:, o:], tau)
    e -= np.einsum("ijba,ijab->", pppp[:o, :o, o:, o:], tau)
    e += 2.0 * np.einsum("ai,ia->", fock[o:, :o], t1)
    return e, t1, t2


def energy_rhf(nocc, fock, pppp, t1, t2):
    """Closed-shell CC energy (same closed form as distr_cc/rccsdt.py:218-221)."""
    o = nocc
    tau = t2 + np.einsum("ia,jb->ijab", t1, t1)
    e = 2.0 * np.einsum("ijab,ijab->", pppp[:o, :o, o:, o:], tau)
    e -= np.einsum("ijba,ijab->", pppp[:o, :o, o:, o:], tau)
    e += 2.0 * np.einsum("ai,ia->", fock[o:, :o], t1)
    return float(e)


def n_blocks(nocc):
    return nocc3(nocc)
