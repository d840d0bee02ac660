"""Canonical T3 block layout (oracle restatement, test infrastructure only).

Follows
  * ``src/rccsdt.c:31-39``   permutation table ``tp_t3``
  * ``src/rccsdt.c:41-73``   ``get_canonical_case``
  * ``src/rccsdt.c:75-85``   ``ijk_to_linear`` (loop form)
  * ``src/rccsdt.c:87-103``  ``apply_perm_and_copy``
  * ``src/rccsdt.c:119-129`` ``take_t3_ijk_single_``
  * ``distr_cc/distribute_t3.py:62-83`` ``linear_to_ijk``
  * ``src/rccsdt_q.c:655-663`` abc triplet index (same form over nvir)

Pins: brute-force enumeration of i<=j<=k triples (tests/test_oracle_layout.py),
the reference C ``take_t3_ijk_single_`` compiled into oracle/_ref, and the
simultaneous-permutation identity T[s(ijk)][s(abc)] == T[ijk][abc] on a dense
symmetric tensor.
"""
import numpy as np

# src/rccsdt.c:32-39
TP_T3 = ((0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0))


def nocc3(nocc):
    return nocc * (nocc + 1) * (nocc + 2) // 6


def get_canonical_case(i, j, k):
    """src/rccsdt.c:41-73 -> (case, (ci, cj, ck))."""
    if i <= j and j <= k:
        return 0, (i, j, k)
    if i <= k and k < j:
        return 1, (i, k, j)
    if j < i and i <= k:
        return 2, (j, i, k)
    if k < i and i <= j:
        return 3, (k, i, j)
    if j <= k and k < i:
        return 4, (j, k, i)
    return 5, (k, j, i)


def ijk_to_linear(i, j, k, nocc):
    """src/rccsdt.c:75-85 (loop form)."""
    n_before_i = 0
    for ii in range(i):
        n_before_i += (nocc - ii) * (nocc - ii + 1) // 2
    n_before_j = 0
    for jj in range(i, j):
        n_before_j += nocc - jj
    return n_before_i + n_before_j + (k - j)


def linear_to_ijk(idx, nocc):
    """distr_cc/distribute_t3.py:62-83."""
    remaining = idx
    i = 0
    while i < nocc:
        ni = nocc - i
        n_i = ni * (ni + 1) // 2
        if remaining < n_i:
            break
        remaining -= n_i
        i += 1
    j = i
    while j < nocc:
        n_j = nocc - j
        if remaining < n_j:
            break
        remaining -= n_j
        j += 1
    return i, j, j + remaining


def canonical_triples(nocc):
    """Lexicographic i<=j<=k order (distribute_t3.py single-rank order)."""
    return [(i, j, k) for i in range(nocc) for j in range(i, nocc) for k in range(j, nocc)]


def apply_perm_and_copy(src, case):
    """src/rccsdt.c:87-103: dst[abc[p0], abc[p1], abc[p2]] = src[a, b, c]."""
    p = TP_T3[case]
    # dst[x0,x1,x2] = src[y] with y[p_n] = x_n, i.e. numpy transpose with axes = p
    return np.ascontiguousarray(src.transpose(p))


def take_t3(t3_local, i, j, k, nocc):
    """src/rccsdt.c:119-129 at one rank (global_to_local_offset = identity)."""
    case, (ci, cj, ck) = get_canonical_case(i, j, k)
    return apply_perm_and_copy(t3_local[ijk_to_linear(ci, cj, ck, nocc)], case)


def full_t3_from_tri(t3_local, nocc):
    """Dense (o,o,o,v,v,v) tensor built through take_t3 for every ordered triple."""
    nvir = t3_local.shape[-1]
    full = np.empty((nocc,) * 3 + (nvir,) * 3, dtype=t3_local.dtype)
    for i in range(nocc):
        for j in range(nocc):
            for k in range(nocc):
                full[i, j, k] = take_t3(t3_local, i, j, k, nocc)
    return full


def abc_to_linear(a, b, c, nvir):
    """src/rccsdt_q.c:655-663 (a<=b<=c triplet index)."""
    te_n = nvir * (nvir + 1) * (nvir + 2) // 6
    rem = nvir - a
    te_rem = rem * (rem + 1) * (rem + 2) // 6
    idx = te_n - te_rem
    idx += (b - a) * (2 * nvir - a - b + 1) // 2
    idx += c - b
    return idx


def t3_abc_from_tri(t3_local, nocc):
    """IJK -> ABC at one rank, as the reference's packing loop (distr_cc/distribute_t3.py:1228-1420):
    for every canonical block (i,j,k) and each of the six simultaneous permutations p of (a,b,c) and
    (i,j,k), the entries whose permuted virtual triple is sorted are stored at
    t3_abc[abc_to_linear(p(a,b,c)), p(i,j,k)]."""
    nvir = t3_local.shape[-1]
    out = np.zeros((nvir * (nvir + 1) * (nvir + 2) // 6, nocc, nocc, nocc), dtype=t3_local.dtype)
    a, b, c = np.mgrid[0:nvir, 0:nvir, 0:nvir]
    perms = [((a, b, c), (0, 1, 2)), ((a, c, b), (0, 2, 1)), ((b, a, c), (1, 0, 2)),
             ((b, c, a), (1, 2, 0)), ((c, a, b), (2, 0, 1)), ((c, b, a), (2, 1, 0))]
    maps = []
    for (pa, pb, pc), order in perms:
        mask = (pa <= pb) & (pb <= pc)
        idx = np.array([abc_to_linear(x, y, z, nvir) for x, y, z in zip(pa[mask], pb[mask], pc[mask])],
                       dtype=np.int64)
        maps.append((mask, idx, order))
    for n, (i, j, k) in enumerate(canonical_triples(nocc)):
        blk = t3_local[n]
        ijk = (i, j, k)
        for mask, idx, order in maps:
            out[idx, ijk[order[0]], ijk[order[1]], ijk[order[2]]] = blk[mask]
    return out
