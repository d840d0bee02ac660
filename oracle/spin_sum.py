"""Spin summations and block transposes (oracle restatement, test infrastructure only).

Follows
  * ``src/rccsdt.c:131-233`` ``t3_single_spin_summation_inplace_``:
        X = p0*A + p1*A[c,b,a] + p2*A[b,a,c];  A <- beta*A + alpha*(p3*X + p4*X[a,c,b])
    with (p0..p4) = P3_full (1,1,1,1,1), P3_422 (2,-1,-1,2,-1), P3_201 (2,-1,-1,1,0)
    (``src/rccsdt.c:135-156``).  The C code walks orbits a>=b>=c; the closed
    form above is the same linear map written with whole-block transposes.
  * ``src/rccsdt.c:245-310`` ``t3_spin_summation_triple_sym_``:
        B0 = 2T - T[a,c,b] - T[c,b,a],  B1 = 2T - T[b,a,c] - T[a,c,b],
        B2 = 2T - T[c,b,a] - T[b,a,c]
  * ``src/rccsdt.c:320-375`` ``t3_transpose_add_``: B[a,b,c] = A[a,b,c] + A[c,a,b]
  * ``src/rccsdt_q.c:40-472`` ``t4_spin_summation_inplace_`` (P4 patterns,
    ``src/rccsdt_q.c:35-38,49-81``):
        X1 = (p0 + p1 P_ab + p2 P_ac + p3 P_ad) A
        X2 = (p4 + p5 P_bc + p6 P_bd) X1
        A <- beta*A + alpha*(p7 + p8 P_cd) X2
    applied on the last four axes.

Pins: the reference C helpers compiled into oracle/_ref (tests/test_oracle_refc.py),
brute force over explicit permutation sums, idempotence of the P3_full
projector and the group-algebra identities P3_422 = (2 - P_bc) P3_201.
"""
import numpy as np

P3_PATTERNS = {
    "P3_full": (1.0, 1.0, 1.0, 1.0, 1.0),
    "P3_422": (2.0, -1.0, -1.0, 2.0, -1.0),
    "P3_201": (2.0, -1.0, -1.0, 1.0, 0.0),
}

P4_PATTERNS = {
    "P4_full": (1.0,) * 9,
    "P4_444": (2.0, -1.0, -1.0, -1.0, 2.0, -1.0, -1.0, 2.0, -1.0),
    "P4_201": (2.0, -1.0, -1.0, -1.0, 1.0, 0.0, 0.0, 1.0, 0.0),
    "P4_442": (2.0, -1.0, -1.0, -1.0, 2.0, -1.0, -1.0, 1.0, 0.0),
}


def t3_spin_sum(A, pattern, alpha=1.0, beta=0.0):
    """Return beta*A + alpha*P(A) on the last three axes (src/rccsdt.c:131-233)."""
    p0, p1, p2, p3, p4 = P3_PATTERNS[pattern]
    nd = A.ndim
    ax = list(range(nd - 3))
    a, b, c = nd - 3, nd - 2, nd - 1
    X = p0 * A + p1 * A.transpose(ax + [c, b, a]) + p2 * A.transpose(ax + [b, a, c])
    return beta * A + alpha * (p3 * X + p4 * X.transpose(ax + [a, c, b]))


def t3_triple_sym(T):
    """src/rccsdt.c:245-310 -> (B0, B1, B2)."""
    t_acb = T.transpose(0, 2, 1)
    t_cba = T.transpose(2, 1, 0)
    t_bac = T.transpose(1, 0, 2)
    B0 = 2.0 * T - t_acb - t_cba
    B1 = 2.0 * T - t_bac - t_acb
    B2 = 2.0 * T - t_cba - t_bac
    return B0, B1, B2


def t3_transpose_add(A):
    """src/rccsdt.c:320-375: B[a,b,c] = A[a,b,c] + A[c,a,b]."""
    return A + A.transpose(1, 2, 0)


def t4_spin_sum(A, pattern, alpha=1.0, beta=0.0):
    """src/rccsdt_q.c:40-472 on the last four axes of A."""
    p = P4_PATTERNS[pattern]
    nd = A.ndim
    ax = list(range(nd - 4))
    a, b, c, d = nd - 4, nd - 3, nd - 2, nd - 1

    def sw(X, x, y):
        perm = ax + [a, b, c, d]
        i, j = perm.index(x), perm.index(y)
        perm[i], perm[j] = perm[j], perm[i]
        return X.transpose(perm)

    X1 = p[0] * A + p[1] * sw(A, a, b) + p[2] * sw(A, a, c) + p[3] * sw(A, a, d)
    X2 = p[4] * X1 + p[5] * sw(X1, b, c) + p[6] * sw(X1, b, d)
    X3 = p[7] * X2 + p[8] * sw(X2, c, d)
    return beta * A + alpha * X3
