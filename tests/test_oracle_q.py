"""Pins of the (Q)/[Q] oracle: the reference's own invariance tests (blksize, job
splitting: tests/test_rccsdt_q.py:56-94), a single-block (unblocked) evaluation,
and the reference C helpers for division, factor and P4_444."""
import itertools

import numpy as np
import pytest

from oracle import q_energy as Q
from oracle import refc
from oracle import t3_layout as L


def _inputs(o, v, seed):
    from distr_cc.synthetic import SyntheticSystem
    s = SyntheticSystem(o, v, naux=20, seed=seed, sigma_t3=1e-3)
    return s, s.pppp(), s.t2(), L.full_t3_from_tri(s.t3_tri(), o)


@pytest.mark.parametrize("o,v", [(2, 5), (3, 4)])
def test_blksize_invariance_and_dense(o, v):
    s, pppp, t2, T = _inputs(o, v, 3 + o)
    ref = Q.kernel(o, v, pppp, t2, T, s.mo_energy, blksize=v)   # one task: unblocked
    for bs in (1, 2, 3):
        eb, ep = Q.kernel(o, v, pppp, t2, T, s.mo_energy, blksize=bs)
        assert abs(eb - ref[0]) < 1e-14 * abs(ref[0]) + 1e-18
        assert abs(ep - ref[1]) < 1e-14 * abs(ref[1]) + 1e-18
    assert ref[0] != 0.0 and ref[1] != 0.0


def test_job_split_additivity():
    o, v = 2, 5
    s, pppp, t2, T = _inputs(o, v, 8)
    full = Q.kernel(o, v, pppp, t2, T, s.mo_energy, blksize=2)
    for n_jobs in (2, 3, 7):
        parts = [Q.kernel(o, v, pppp, t2, T, s.mo_energy, blksize=2, job_idx=j, n_jobs=n_jobs)
                 for j in range(n_jobs)]
        assert abs(sum(p[0] for p in parts) - full[0]) < 1e-14 * abs(full[0])
        assert abs(sum(p[1] for p in parts) - full[1]) < 1e-14 * abs(full[1])


def test_task_list_covers_each_ordered_block_quadruple_once():
    v, bs = 11, 3
    tasks = Q.create_full_task_list(v, bs)
    nb = (v + bs - 1) // bs
    seen = sorted((t[0] // bs, t[3] // bs, t[6] // bs, t[9] // bs) for t in tasks)
    expect = sorted(q for q in itertools.product(range(nb), repeat=4) if q[0] <= q[1] <= q[2] <= q[3])
    assert seen == expect


def test_factor_brute_force():
    bs = 3
    for (a0, b0, c0, d0) in [(0, 0, 0, 0), (0, 0, 3, 3), (0, 3, 6, 9), (3, 3, 3, 6)]:
        f = Q.get_factor(a0, a0 + bs, b0, b0 + bs, c0, c0 + bs, d0, d0 + bs, bs)
        for ia, ib, ic, id_ in itertools.product(range(bs), repeat=4):
            a, b, c, d = a0 + ia, b0 + ib, c0 + ic, d0 + id_
            if a < b < c < d:
                want = 24.0
            elif (a == b < c < d) or (a < b == c < d) or (a < b < c == d):
                want = 12.0
            elif a == b < c == d:
                want = 6.0
            else:
                want = 0.0
            assert f[ia, ib, ic, id_] == want


@pytest.mark.skipif(not refc.available(), reason="oracle/_ref not built")
def test_division_and_factor_vs_reference_c():
    rng = np.random.default_rng(0)
    o, bs = 3, 2
    e_occ = np.sort(rng.uniform(-1, -0.2, o))
    e_vir = np.sort(rng.uniform(0.1, 2.0, 6))
    r4 = rng.normal(size=(bs,) * 4 + (o,) * 4)
    mine = Q.e_division(r4.copy(), e_occ, e_vir, 0, 2, 2, 4, 2, 4, 4, 6)
    theirs = refc.e_abcdijkl_division(r4, e_occ, e_vir, 0, 2, 2, 4, 2, 4, 4, 6, bs)
    assert np.allclose(mine, theirs, atol=1e-15, rtol=1e-14)
    fac = Q.get_factor(0, 2, 2, 4, 2, 4, 4, 6, bs)
    assert np.allclose(refc.t4_multiply_factor(r4, fac, bs, o), r4 * fac[..., None, None, None, None])
