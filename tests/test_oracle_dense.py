"""Block-loop oracle (follows the reference lines) against the independent dense restatement."""
import numpy as np
import pytest

from oracle import rccsdt_dense as dense
from oracle import rccsdt_tri as tri
from oracle import t3_layout as L


def _system(o, v, seed):
    from distr_cc.synthetic import SyntheticSystem
    s = SyntheticSystem(o, v, naux=20, seed=seed)
    rng = np.random.default_rng(seed)
    im = {k: x + 1e-3 * rng.normal(size=x.shape) for k, x in s.bare_intermediates().items()}
    return s, im, s.t2(), s.t3_tri()


@pytest.mark.parametrize("o,v,bs", [(1, 3, 100), (3, 5, 100), (4, 6, 3), (5, 4, 7), (4, 3, 1)])
def test_r3_block_loop_equals_dense(o, v, bs):
    s, im, t2, t3 = _system(o, v, 7 * o + v)
    T = L.full_t3_from_tri(t3, o)
    r3 = tri.compute_r3_tri(o, v, im, t2, t3, batch_size=bs)
    Rd = dense.r3_dense(im, t2, T)
    ref = np.stack([Rd[i, j, k] for (i, j, k) in L.canonical_triples(o)])
    assert abs(r3 - ref).max() <= 1e-12 * abs(ref).max()


def test_r3_batch_size_invariance():
    """The reference's own invariance test (tests/test_rccsdt.py:85-101) at the residual level."""
    o, v = 5, 4
    s, im, t2, t3 = _system(o, v, 3)
    r_a = tri.compute_r3_tri(o, v, im, t2, t3, batch_size=3)
    for bs in (11, 17, 100):
        assert abs(tri.compute_r3_tri(o, v, im, t2, t3, batch_size=bs) - r_a).max() < 1e-17


def test_dense_r3_has_simultaneous_permutation_symmetry():
    o, v = 3, 4
    s = __import__("distr_cc.synthetic", fromlist=["SyntheticSystem"]).SyntheticSystem(o, v, naux=20, seed=9)
    im, t2, t3 = s.bare_intermediates(), s.t2(), s.t3_tri()  # pair-symmetric ERI slices
    R = dense.r3_dense(im, t2, L.full_t3_from_tri(t3, o))
    for p in [(1, 0, 2), (2, 1, 0), (0, 2, 1), (1, 2, 0)]:
        assert np.allclose(R, R.transpose(*p, *(3 + np.array(p))), atol=1e-16, rtol=0)


@pytest.mark.parametrize("o,v", [(3, 4), (4, 3)])
def test_t3_to_r1_r2_and_w_block_equals_dense(o, v):
    s, im, t2, t3 = _system(o, v, 21)
    nmo = o + v
    rng = np.random.default_rng(4)
    t1_eris = s.pppp() + 0.01 * rng.normal(size=(nmo,) * 4)
    t1_eris = 0.5 * (t1_eris + t1_eris.transpose(1, 0, 3, 2))
    t1_fock = s.fock + 0.01 * rng.normal(size=(nmo, nmo))
    T = L.full_t3_from_tri(t3, o)
    r1, r2 = np.zeros((o, v)), np.zeros((o, o, v, v))
    tri.r1r2_add_t3_tri(o, v, t1_fock, t1_eris, t3, r1, r2)
    r1d, r2d = dense.r1r2_t3_dense(o, t1_fock, t1_eris, T)
    assert abs(r1 - r1d).max() <= 1e-12 * abs(r1d).max()
    assert abs(r2 - r2d).max() <= 1e-12 * abs(r2d).max()
    Wv, Wvv = np.zeros((v, o, o, o)), np.zeros((v, v, v, o))
    tri.intermediates_t3_add_t3_tri(o, v, t1_eris, t3, Wv, Wvv)
    Wvd, Wvvd = dense.w_t3_dense(o, t1_eris, T)
    assert abs(Wv - Wvd).max() <= 1e-12 * abs(Wvd).max()
    assert abs(Wvv - Wvvd).max() <= 1e-12 * abs(Wvvd).max()


def test_denominator_closed_form():
    o, v = 3, 4
    s, im, t2, t3 = _system(o, v, 5)
    r = np.random.default_rng(0).normal(size=t3.shape)
    out = tri.divide_e(o, r, s.mo_energy, level_shift=0.1)
    e = s.mo_energy
    for n, (i, j, k) in enumerate(L.canonical_triples(o)):
        for a, b, c in [(0, 1, 2), (3, 3, 0), (2, 0, 1)]:
            d = e[i] + e[j] + e[k] - e[o + a] - e[o + b] - e[o + c] - 0.3
            assert abs(out[n, a, b, c] - r[n, a, b, c] / d) < 1e-15 * abs(r[n, a, b, c] / d) + 1e-300


@pytest.mark.parametrize("o,v,bs", [(5, 4, 100), (6, 3, 7)])
def test_r3_target_subset_equals_full(o, v, bs):
    """compute_r3_tri(targets=S) (which skips the source blocks and W_oooo/F_oo rows that meet no
    sampled target) equals the rows S of the full residual; the other rows stay zero."""
    s, im, t2, t3 = _system(o, v, 5 * o + v)
    full = tri.compute_r3_tri(o, v, im, t2, t3, batch_size=bs)
    S = [0, 3, len(t3) // 2, len(t3) - 1]
    sub = tri.compute_r3_tri(o, v, im, t2, t3, batch_size=bs, targets=S)
    assert abs(sub[S] - full[S]).max() <= 1e-14 * abs(full).max()
    rest = np.setdiff1d(np.arange(len(t3)), S)
    assert not sub[rest].any()
