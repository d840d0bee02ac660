"""The oracle's spin summations against the reference C helpers compiled into oracle/_ref,
and against brute-force permutation sums / group-algebra identities."""
import itertools

import numpy as np
import pytest

from oracle import refc
from oracle import spin_sum as S

needs_ref = pytest.mark.skipif(not refc.available(), reason="oracle/_ref not built")
PERMS = list(itertools.permutations(range(3)))


@needs_ref
@pytest.mark.parametrize("v", [1, 2, 3, 7])
@pytest.mark.parametrize("pattern", ["P3_full", "P3_422", "P3_201"])
def test_t3_spin_sum_vs_reference(v, pattern):
    A = np.random.default_rng(v).normal(size=(v, v, v))
    for alpha, beta in [(1.0, 0.0), (-1.0 / 6.0, 1.0), (0.3, -0.7)]:
        assert np.allclose(refc.t3_spin_sum(A, pattern, alpha, beta), S.t3_spin_sum(A, pattern, alpha, beta),
                           atol=1e-14, rtol=0)


@needs_ref
@pytest.mark.parametrize("v", [1, 4, 9])
def test_triple_sym_and_transpose_add_vs_reference(v):
    A = np.random.default_rng(10 + v).normal(size=(v, v, v))
    for x, y in zip(refc.t3_triple_sym(A), S.t3_triple_sym(A)):
        assert np.allclose(x, y, atol=1e-14, rtol=0)
    assert np.allclose(refc.t3_transpose_add(A), S.t3_transpose_add(A), atol=1e-14, rtol=0)


@needs_ref
@pytest.mark.parametrize("v", [1, 2, 3, 5])
@pytest.mark.parametrize("pattern", ["P4_full", "P4_444", "P4_201", "P4_442"])
def test_t4_spin_sum_vs_reference(v, pattern):
    A = np.random.default_rng(v).normal(size=(2, v, v, v, v))
    assert np.allclose(refc.t4_spin_sum(A, pattern, 0.7, 0.2), S.t4_spin_sum(A, pattern, 0.7, 0.2),
                       atol=1e-12, rtol=0)


def test_p3_full_is_a_projector_and_p3_identities():
    A = np.random.default_rng(3).normal(size=(6, 6, 6))
    P = S.t3_spin_sum(A, "P3_full", -1.0 / 6.0, 1.0)
    assert np.allclose(S.t3_spin_sum(P, "P3_full", -1.0 / 6.0, 1.0), P, atol=1e-14)
    # the projected block has no totally symmetric part
    assert np.allclose(sum(P.transpose(p) for p in PERMS), 0.0, atol=1e-13)
    # brute force: P3_full = sum over S3
    assert np.allclose(S.t3_spin_sum(A, "P3_full"), sum(A.transpose(p) for p in PERMS), atol=1e-13)
    # P3_422 = (2 - P_bc) P3_201
    z = S.t3_spin_sum(A, "P3_201")
    assert np.allclose(S.t3_spin_sum(A, "P3_422"), 2 * z - z.transpose(0, 2, 1), atol=1e-13)
    # P3_201 = 2A - A[c,b,a] - A[b,a,c] (explicit)
    assert np.allclose(z, 2 * A - A.transpose(2, 1, 0) - A.transpose(1, 0, 2), atol=1e-14)
