"""Pins of the canonical block layout (oracle vs brute force, product closed form, reference C)."""
import itertools

import numpy as np
import pytest

from oracle import refc
from oracle import t3_layout as L


@pytest.mark.parametrize("nocc", [1, 2, 3, 5, 10, 16])
def test_linear_index_matches_enumeration(nocc):
    from distr_cc import distribute_t3 as D
    trip = [(i, j, k) for i in range(nocc) for j in range(i, nocc) for k in range(j, nocc)]
    assert len(trip) == L.nocc3(nocc) == D.nocc3(nocc)
    for n, (i, j, k) in enumerate(trip):
        assert L.ijk_to_linear(i, j, k, nocc) == n
        assert D.ijk_to_linear(i, j, k, nocc) == n
        assert L.linear_to_ijk(n, nocc) == (i, j, k)
        assert D.linear_to_ijk(n, nocc) == (i, j, k)


def test_canonical_case_matches_sorting():
    from distr_cc import distribute_t3 as D
    for i, j, k in itertools.product(range(4), repeat=3):
        case, can = L.get_canonical_case(i, j, k)
        assert tuple(sorted((i, j, k))) == can
        assert D.canonical_case(i, j, k) == (case, can)
        # the case permutation maps canonical positions onto (i,j,k)
        p = L.TP_T3[case]
        assert tuple(can[p[n]] for n in range(3)) == (i, j, k) or len({i, j, k}) < 3


def test_take_t3_gives_symmetric_dense_tensor():
    rng = np.random.default_rng(0)
    o, v = 4, 3
    full = rng.normal(size=(o, o, o, v, v, v))
    # symmetrize under simultaneous permutations, then compress to canonical blocks
    sym = sum(full.transpose(*p, *(3 + np.array(p))) for p in itertools.permutations(range(3))) / 6
    tri = np.stack([sym[i, j, k] for (i, j, k) in L.canonical_triples(o)])
    assert np.allclose(L.full_t3_from_tri(tri, o), sym, atol=1e-14, rtol=0)


@pytest.mark.skipif(not refc.available(), reason="oracle/_ref not built")
def test_take_t3_matches_reference_c():
    rng = np.random.default_rng(1)
    o, v = 4, 3
    t3 = rng.normal(size=(L.nocc3(o), v, v, v))
    for i, j, k in itertools.product(range(o), repeat=3):
        assert np.array_equal(refc.take_t3(t3, i, j, k, o), L.take_t3(t3, i, j, k, o))


def test_product_take_matches_oracle():
    from distr_cc.distribute_t3 import DistributedT3IJK
    rng = np.random.default_rng(2)
    o, v = 3, 4
    dt3 = DistributedT3IJK(o, v)
    t3 = rng.normal(size=(L.nocc3(o), v, v, v))
    blk = np.empty((v, v, v))
    for i, j, k in itertools.product(range(o), repeat=3):
        assert np.array_equal(dt3.take_t3_single_local(t3, blk, i, j, k), L.take_t3(t3, i, j, k, o))


def exact_symmetric_t3(nocc, nvir, seed):
    """A random T[i,j,k,a,b,c] that is exactly (bit for bit) invariant under simultaneous permutations
    of (i,j,k) and (a,b,c): a random table indexed by the sorted pairs (i,a), (j,b), (k,c)."""
    rng = np.random.default_rng(seed)
    M = nocc * nvir
    table = rng.standard_normal(M ** 3)
    T = np.empty((nocc,) * 3 + (nvir,) * 3)
    for i, j, k, a, b, c in itertools.product(range(nocc), range(nocc), range(nocc),
                                              range(nvir), range(nvir), range(nvir)):
        p = sorted((i * nvir + a, j * nvir + b, k * nvir + c))
        T[i, j, k, a, b, c] = table[(p[0] * M + p[1]) * M + p[2]]
    return T


@pytest.mark.parametrize("nocc,nvir", [(1, 3), (3, 4), (4, 5)])
def test_abc_layout_is_the_full_tensor(nocc, nvir):
    """t3_abc[abc_to_linear(a<=b<=c), i, j, k] == T[i,j,k,a,b,c] of the full tensor whose canonical blocks
    are packed (brute force over every index)."""
    T = exact_symmetric_t3(nocc, nvir, nocc * 10 + nvir)
    t3 = np.array([T[i, j, k] for (i, j, k) in L.canonical_triples(nocc)])
    abc = L.t3_abc_from_tri(t3, nocc)
    for a in range(nvir):
        for b in range(a, nvir):
            for c in range(b, nvir):
                assert np.array_equal(abc[L.abc_to_linear(a, b, c, nvir)], T[:, :, :, a, b, c])


def test_abc_redistribution_and_dense_gather_emulated(monkeypatch):
    """redistribute_ijk_to_abc (device gather) == the oracle packing; the ABC -> dense gather of the (Q)
    engine gives back the full tensor, zero-padded."""
    import abi_emulator
    abi_emulator.install(monkeypatch)
    from distr_cc import _lib
    from distr_cc.device import DeviceArray
    from distr_cc.distribute_t3 import DistributedT3IJK, redistribute_ijk_to_abc
    o, v, vp = 3, 5, 6
    T = exact_symmetric_t3(o, v, 5)
    t3 = np.array([T[i, j, k] for (i, j, k) in L.canonical_triples(o)])
    dt3_abc, t3_abc = redistribute_ijk_to_abc(DistributedT3IJK(o, v), t3)
    assert np.array_equal(t3_abc, L.t3_abc_from_tri(t3, o))
    assert dt3_abc.abc_to_global_idx[(1, 2, 4)] == L.abc_to_linear(1, 2, 4, v)
    assert dt3_abc.get_local_index(4, 1, 2, 2, 0, 1) == (L.abc_to_linear(1, 2, 4, v) * o ** 3 + 2 * o * o + 1)
    src, dst = DeviceArray.from_host(t3_abc), DeviceArray((vp, vp, vp, o, o, o))
    _lib.call("dcc_t3_abc_packed_to_dense_f64", src.vp, o, v, vp, dst.vp, None)
    dense = dst.download()
    assert np.array_equal(dense[:v, :v, :v], T.transpose(3, 4, 5, 0, 1, 2))
    assert not dense[v:].any() and not dense[:, v:].any() and not dense[:, :, v:].any()
