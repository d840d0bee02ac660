"""ctypes binding of the reference C helpers built into oracle/_ref (test infrastructure only).

The library is compiled by ``oracle/Makefile`` from the reference's
``src/rccsdt.c`` and ``src/rccsdt_q.c`` as they lie; it is used only to check
the oracle's NumPy restatement.  Signatures follow the reference's own
binding ``distr_cc/_c_rccsdt.py:65-141`` and ``distr_cc/_c_rccsdt_q.py``.
"""
import ctypes
import os

import numpy as np

_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_ref", "libdistr_cc_ref.so")


def available():
    return os.path.exists(_PATH)


def _lib():
    lib = ctypes.CDLL(_PATH)
    vp, i64, dbl = ctypes.c_void_p, ctypes.c_int64, ctypes.c_double
    lib.t3_single_spin_summation_inplace_.argtypes = [vp, i64, ctypes.c_char_p, dbl, dbl]
    lib.t3_single_spin_summation_inplace_.restype = None
    lib.t3_transpose_add_.argtypes = [vp, vp, i64]
    lib.t3_transpose_add_.restype = None
    lib.t3_spin_summation_triple_sym_.argtypes = [vp, vp, vp, vp, i64]
    lib.t3_spin_summation_triple_sym_.restype = None
    lib.take_t3_ijk_single_.argtypes = [vp, vp, vp, i64, i64, i64, i64, i64]
    lib.take_t3_ijk_single_.restype = None
    lib.t4_spin_summation_inplace_.argtypes = [vp, i64, i64, ctypes.c_char_p, dbl, dbl]
    lib.t4_spin_summation_inplace_.restype = None
    lib.e_abcdijkl_division_.argtypes = [vp, vp, vp] + [i64] * 13
    lib.e_abcdijkl_division_.restype = None
    lib.t4_multiply_factor_.argtypes = [vp, vp, i64, i64, i64, i64, i64]
    lib.t4_multiply_factor_.restype = None
    lib.fill_t3_from_ptr_array.argtypes = [vp, i64, i64, i64, i64, i64, i64, vp, i64, i64, i64]
    lib.fill_t3_from_ptr_array.restype = None
    return lib


_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        _LIB = _lib()
    return _LIB


def _p(a):
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.c_void_p)


def t3_spin_sum(A, pattern, alpha=1.0, beta=0.0):
    A = np.array(A, dtype=np.float64, order="C", copy=True)
    lib().t3_single_spin_summation_inplace_(_p(A), A.shape[-1], pattern.encode(), alpha, beta)
    return A


def t3_transpose_add(A):
    A = np.ascontiguousarray(A, dtype=np.float64)
    B = np.empty_like(A)
    lib().t3_transpose_add_(_p(A), _p(B), A.shape[-1])
    return B


def t3_triple_sym(A):
    A = np.ascontiguousarray(A, dtype=np.float64)
    B0, B1, B2 = np.empty_like(A), np.empty_like(A), np.empty_like(A)
    lib().t3_spin_summation_triple_sym_(_p(A), _p(B0), _p(B1), _p(B2), A.shape[-1])
    return B0, B1, B2


def take_t3(t3_local, i, j, k, nocc):
    t3_local = np.ascontiguousarray(t3_local, dtype=np.float64)
    nvir = t3_local.shape[-1]
    blk = np.zeros((nvir,) * 3)
    g2l = np.arange(t3_local.shape[0], dtype=np.int64)
    lib().take_t3_ijk_single_(_p(t3_local), _p(blk), _p(g2l), i, j, k, nocc, nvir)
    return blk


def t4_spin_sum(A, pattern, alpha=1.0, beta=0.0):
    """Apply on the last four axes; leading axes are the C function's ``nocc4``."""
    A = np.array(A, dtype=np.float64, order="C", copy=True)
    n = A.shape[-1]
    lead = int(np.prod(A.shape[:-4])) if A.ndim > 4 else 1
    lib().t4_spin_summation_inplace_(_p(A), lead, n, pattern.encode(), alpha, beta)
    return A


def e_abcdijkl_division(r4, e_occ, e_vir, a0, a1, b0, b1, c0, c1, d0, d1, blk):
    r4 = np.array(r4, dtype=np.float64, order="C", copy=True)
    e_occ = np.ascontiguousarray(e_occ, dtype=np.float64)
    e_vir = np.ascontiguousarray(e_vir, dtype=np.float64)
    nocc = e_occ.size
    lib().e_abcdijkl_division_(_p(r4), _p(e_occ), _p(e_vir), a0, a1, b0, b1, c0, c1, d0, d1,
                               blk, blk, blk, blk, nocc)
    return r4


def t4_multiply_factor(t4, factor, blk, nocc):
    t4 = np.array(t4, dtype=np.float64, order="C", copy=True)
    factor = np.ascontiguousarray(factor, dtype=np.float64)
    lib().t4_multiply_factor_(_p(t4), _p(factor), blk, blk, blk, blk, nocc)
    return t4


def fill_t3_from_abc(t3_abc, a0, a1, b0, b1, nvir, nocc, blk):
    """Gather t3_blk[a,b,c,i,j,k] (shape (blk,blk,nvir,o,o,o)) from ABC-stored blocks."""
    t3_abc = np.ascontiguousarray(t3_abc, dtype=np.float64)
    out = np.zeros((blk, blk, nvir, nocc, nocc, nocc))
    nblk = t3_abc.shape[0]
    ptrs = np.array([t3_abc.ctypes.data + n * t3_abc[0].nbytes for n in range(nblk)], dtype=np.uint64)
    lib().fill_t3_from_ptr_array(_p(out), a0, a1, b0, b1, nvir, nocc, _p(ptrs), blk, out.size, nblk)
    return out
