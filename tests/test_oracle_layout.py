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
