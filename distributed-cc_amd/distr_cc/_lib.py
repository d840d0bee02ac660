"""ctypes binding of libdcc.so (include/dcc.h), loaded at import.

Mirrors the reference loader ``distr_cc/_lib.py:19-42``: the shared library
sits in ``<package>/../build`` and a missing library fails the import with an
error that says how to build it.  There is no CPU fallback.
"""
import ctypes
import os

import numpy as np

# DCC_LIBRARY selects another build of the library (for example one compiled with different options)
_LIB_PATH = os.environ.get("DCC_LIBRARY") or os.path.join(os.path.dirname(os.path.abspath(__file__)), "..",
                                                          "build", "libdcc.so")

if not os.path.exists(_LIB_PATH):
    raise RuntimeError(
        "libdcc.so not found at %s; build it with `make -C distributed-cc_amd/csrc` "
        "(or __graft_entry__.build())" % os.path.normpath(_LIB_PATH))

_lib = ctypes.CDLL(_LIB_PATH)

c_i64 = ctypes.c_int64
c_dbl = ctypes.c_double
c_vp = ctypes.c_void_p


class IdxMap(ctypes.Structure):
    """dcc_idxmap: table lookup or three-level mixed-radix index map."""
    _fields_ = [("table", c_vp), ("table_bstride", c_i64), ("e1", c_i64), ("e2", c_i64),
                ("s0", c_i64), ("s1", c_i64), ("s2", c_i64)]


class Batch(ctypes.Structure):
    _fields_ = [("a_off", c_i64), ("b_off", c_i64), ("c_off", c_i64), ("m_begin", c_i64), ("m_end", c_i64)]


class S3Job(ctypes.Structure):
    _fields_ = [("dst", c_i64), ("src", c_i64), ("coef", c_dbl * 6)]


class Strided(ctypes.Structure):
    """dcc_strided: up to 6 axes with per-side element strides."""
    _fields_ = [("ndim", c_i64), ("shape", c_i64 * 6), ("dst_stride", c_i64 * 6), ("src_stride", c_i64 * 6)]


class GemmProblem(ctypes.Structure):
    """dcc_gemm_problem: one contraction of a grouped launch."""
    _fields_ = [("N", c_i64), ("K", c_i64), ("alpha", c_dbl), ("beta", c_dbl), ("A", c_vp), ("B", c_vp),
                ("C", c_vp), ("am", IdxMap), ("ak", IdxMap), ("bk", IdxMap), ("bn", IdxMap), ("cm", IdxMap),
                ("cn", IdxMap), ("batches", c_vp), ("nbatch", c_i64), ("max_rows", c_i64), ("tiles_m", c_i64),
                ("tiles_n", c_i64), ("tile_begin", c_i64), ("flags", c_i64)]


BATCH_DTYPE = np.dtype([("a_off", np.int64), ("b_off", np.int64), ("c_off", np.int64),
                        ("m_begin", np.int64), ("m_end", np.int64)])
S3JOB_DTYPE = np.dtype([("dst", np.int64), ("src", np.int64), ("coef", np.float64, (6,))])
S3GATHER_DTYPE = np.dtype([("dst", np.int64), ("src", np.int64, (6,)), ("coef", np.float64, (6,))])
assert BATCH_DTYPE.itemsize == ctypes.sizeof(Batch)
assert S3JOB_DTYPE.itemsize == ctypes.sizeof(S3Job)

_SIGS = {
    "dcc_abi_version": ([], ctypes.c_int),
    "dcc_last_error": ([], ctypes.c_char_p),
    "dcc_device_count": ([ctypes.POINTER(ctypes.c_int)], ctypes.c_int),
    "dcc_set_device": ([ctypes.c_int, ctypes.c_char_p, c_i64], ctypes.c_int),
    "dcc_mem_get_info": ([ctypes.POINTER(c_i64), ctypes.POINTER(c_i64)], ctypes.c_int),
    "dcc_malloc": ([ctypes.POINTER(c_vp), c_i64], ctypes.c_int),
    "dcc_free": ([c_vp], ctypes.c_int),
    "dcc_memcpy_htod": ([c_vp, c_vp, c_i64, c_vp], ctypes.c_int),
    "dcc_memcpy_dtoh": ([c_vp, c_vp, c_i64, c_vp], ctypes.c_int),
    "dcc_memcpy_dtod": ([c_vp, c_vp, c_i64, c_vp], ctypes.c_int),
    "dcc_fill_f64": ([c_vp, c_i64, c_dbl, c_vp], ctypes.c_int),
    "dcc_fill_f32": ([c_vp, c_i64, ctypes.c_float, c_vp], ctypes.c_int),
    "dcc_stream_create": ([ctypes.POINTER(c_vp)], ctypes.c_int),
    "dcc_stream_destroy": ([c_vp], ctypes.c_int),
    "dcc_stream_synchronize": ([c_vp], ctypes.c_int),
    "dcc_device_synchronize": ([], ctypes.c_int),
    "dcc_event_create": ([ctypes.POINTER(c_vp)], ctypes.c_int),
    "dcc_event_destroy": ([c_vp], ctypes.c_int),
    "dcc_event_record": ([c_vp, c_vp], ctypes.c_int),
    "dcc_event_elapsed_ms": ([ctypes.POINTER(ctypes.c_float), c_vp, c_vp], ctypes.c_int),
    "dcc_contract_f64": ([c_i64, c_i64, c_i64, c_dbl, c_dbl, c_vp, IdxMap, IdxMap, c_vp, IdxMap, IdxMap,
                          c_vp, IdxMap, IdxMap, c_vp, c_i64, c_i64, c_vp], ctypes.c_int),
    "dcc_contract_f32": ([c_i64, c_i64, c_i64, c_dbl, c_dbl, c_vp, IdxMap, IdxMap, c_vp, IdxMap, IdxMap,
                          c_vp, IdxMap, IdxMap, c_vp, c_i64, c_i64, c_vp], ctypes.c_int),
    "dcc_contract_tiled_f64": ([c_i64, c_i64, c_i64, c_dbl, c_dbl, c_vp, IdxMap, IdxMap, c_vp, IdxMap, IdxMap,
                                c_vp, IdxMap, IdxMap, c_vp, c_i64, c_i64, c_i64, c_vp], ctypes.c_int),
    "dcc_contract_tiled_f32": ([c_i64, c_i64, c_i64, c_dbl, c_dbl, c_vp, IdxMap, IdxMap, c_vp, IdxMap, IdxMap,
                                c_vp, IdxMap, IdxMap, c_vp, c_i64, c_i64, c_i64, c_vp], ctypes.c_int),
    "dcc_contract_f32_bf16x3": ([c_i64, c_i64, c_i64, c_dbl, c_dbl, c_vp, IdxMap, IdxMap, c_vp, IdxMap, IdxMap,
                                 c_vp, IdxMap, IdxMap, c_vp, c_i64, c_i64, c_vp], ctypes.c_int),
    "dcc_pack_bf16x3": ([c_vp, c_vp, c_i64, c_vp], ctypes.c_int),
    "dcc_t2t3_operands_f64": ([c_vp] * 5 + [c_i64] * 4 + [c_vp], ctypes.c_int),
    "dcc_t2t3_operands_f32": ([c_vp] * 5 + [c_i64] * 4 + [c_vp], ctypes.c_int),
    "dcc_t2t3_operands_packed_f32": ([c_vp] * 5 + [c_i64] * 4 + [c_vp], ctypes.c_int),
    "dcc_vvvv_pair_weights_f64": ([c_vp, c_i64, c_i64, c_vp, c_vp], ctypes.c_int),
    "dcc_vvvv_pair_weights_f32": ([c_vp, c_i64, c_i64, c_vp, c_vp], ctypes.c_int),
    "dcc_vvvv_pair_weights_packed_f32": ([c_vp, c_i64, c_i64, c_vp, c_vp], ctypes.c_int),
    "dcc_vvvv_pair_split_f64": ([c_vp, c_i64, c_i64, ctypes.c_int32, c_i64, c_vp, c_vp], ctypes.c_int),
    "dcc_vvvv_pair_split_f32": ([c_vp, c_i64, c_i64, ctypes.c_int32, c_i64, c_vp, c_vp], ctypes.c_int),
    "dcc_vvvv_pair_split_packed_f32": ([c_vp, c_i64, c_i64, ctypes.c_int32, c_i64, c_vp, c_vp], ctypes.c_int),
    "dcc_vvvv_pair_unfold_f64": ([c_vp, c_i64, c_i64, ctypes.c_int32, c_vp, c_vp], ctypes.c_int),
    "dcc_vvvv_pair_unfold_f32": ([c_vp, c_i64, c_i64, ctypes.c_int32, c_vp, c_vp], ctypes.c_int),
    "dcc_contract_packed_bf16x3": ([c_i64, c_i64, c_i64, c_dbl, c_dbl, c_vp, IdxMap, IdxMap, c_vp, IdxMap,
                                    IdxMap, c_vp, IdxMap, IdxMap, c_vp, c_i64, c_i64, c_i64, c_vp], ctypes.c_int),
    "dcc_contract_tiled_f32_bf16x3": ([c_i64, c_i64, c_i64, c_dbl, c_dbl, c_vp, IdxMap, IdxMap, c_vp, IdxMap,
                                       IdxMap, c_vp, IdxMap, IdxMap, c_vp, c_i64, c_i64, c_i64, c_vp], ctypes.c_int),
    "dcc_s3_transform_f64": ([c_vp, c_vp, c_i64, c_dbl, c_vp, c_i64, c_vp], ctypes.c_int),
    "dcc_s3_transform_f32": ([c_vp, c_vp, c_i64, c_dbl, c_vp, c_i64, c_vp], ctypes.c_int),
    "dcc_s3_transform_pack_f32": ([c_vp, c_vp, c_i64, c_vp, c_i64, c_vp], ctypes.c_int),
    "dcc_s3_gather_f64": ([c_vp, c_vp, c_i64, c_dbl, c_vp, c_i64, c_vp], ctypes.c_int),
    "dcc_s3_gather_f32": ([c_vp, c_vp, c_i64, c_dbl, c_vp, c_i64, c_vp], ctypes.c_int),
    "dcc_r3_finalize_tiles": ([c_i64], c_i64),
    "dcc_r3_finalize_f64": ([c_vp, c_vp, c_vp, c_i64, c_i64, c_vp, c_vp, c_i64, c_i64, c_vp, c_vp],
                            ctypes.c_int),
    "dcc_r3_finalize_f32": ([c_vp, c_vp, c_vp, c_i64, c_i64, c_vp, c_vp, c_i64, c_i64, c_vp, c_vp],
                            ctypes.c_int),
    "dcc_r3_finalize_noalloc_f64": ([c_vp, c_vp, c_vp, c_i64, c_i64, c_vp, c_vp, c_i64, c_i64, c_vp, c_vp],
                                    ctypes.c_int),
    "dcc_reduce_pairs_f64": ([c_vp, c_i64, c_vp, c_vp], ctypes.c_int),
    "dcc_oooo_prefactors_f64": ([c_vp, c_vp, c_i64, c_vp, c_i64, c_vp, c_vp], ctypes.c_int),
    "dcc_ring_weights_f64": ([c_vp, c_vp, c_i64, c_i64, c_vp, c_vp, c_vp, c_vp], ctypes.c_int),
    "dcc_ring_weights_f32": ([c_vp, c_vp, c_i64, c_i64, c_vp, c_vp, c_vp, c_vp], ctypes.c_int),
    "dcc_oooo_prefactors_f32": ([c_vp, c_vp, c_i64, c_vp, c_i64, c_vp, c_vp], ctypes.c_int),
    "dcc_t3_to_abc_f64": ([c_vp, c_i64, c_i64, c_i64, c_vp, c_vp], ctypes.c_int),
    "dcc_t3_tri_to_abc_packed_f64": ([c_vp, c_i64, c_i64, c_vp, c_vp], ctypes.c_int),
    "dcc_t3_abc_packed_to_dense_f64": ([c_vp, c_i64, c_i64, c_i64, c_vp, c_vp], ctypes.c_int),
    "dcc_q_finalize_f64": ([c_vp, c_vp, c_vp, c_i64, c_i64, c_i64, c_i64, c_i64, c_vp, c_vp, c_vp, c_vp, c_vp,
                            c_vp, c_vp], ctypes.c_int),
    "dcc_tensor_axpby_f64": ([c_vp, c_vp, ctypes.POINTER(Strided), c_dbl, c_dbl, c_vp], ctypes.c_int),
    "dcc_tensor_axpby_f32": ([c_vp, c_vp, ctypes.POINTER(Strided), c_dbl, c_dbl, c_vp], ctypes.c_int),
    "dcc_tensor_axpby_f32_f64": ([c_vp, c_vp, ctypes.POINTER(Strided), c_dbl, c_dbl, c_vp], ctypes.c_int),
    "dcc_tensor_axpby_f64_f32": ([c_vp, c_vp, ctypes.POINTER(Strided), c_dbl, c_dbl, c_vp], ctypes.c_int),
    "dcc_grouped_prepare": ([c_vp, c_i64, c_i64, ctypes.POINTER(c_i64), ctypes.POINTER(c_i64)], ctypes.c_int),
    "dcc_contract_grouped_f64": ([c_vp, c_i64, c_i64, c_i64, c_i64, c_vp], ctypes.c_int),
    "dcc_graph_capture_begin": ([c_vp], ctypes.c_int),
    "dcc_graph_capture_end": ([c_vp, ctypes.POINTER(c_vp)], ctypes.c_int),
    "dcc_graph_launch": ([c_vp, c_vp], ctypes.c_int),
    "dcc_graph_destroy": ([c_vp], ctypes.c_int),
    "dcc_sum_slices_f64": ([c_vp, c_vp, c_i64, c_i64, c_dbl, c_dbl, c_vp], ctypes.c_int),
    "dcc_sum_slices_f32": ([c_vp, c_vp, c_i64, c_i64, c_dbl, c_dbl, c_vp], ctypes.c_int),
    "dcc_dot_partials": ([], c_i64),
    "dcc_dot_f64": ([c_vp, c_vp, c_i64, c_vp, c_vp, c_vp], ctypes.c_int),
    "dcc_dot_f32": ([c_vp, c_vp, c_i64, c_vp, c_vp, c_vp], ctypes.c_int),
    "dcc_denom_divide_f64": ([c_vp, c_i64, c_i64, c_i64, c_vp, c_vp], ctypes.c_int),
}

for _name, (_args, _res) in _SIGS.items():
    _fn = getattr(_lib, _name)
    _fn.argtypes = _args
    _fn.restype = _res

EXPORTED = tuple(_SIGS)


class DCCError(RuntimeError):
    pass


def check(status, what=""):
    """Raise DCCError with the library's message when a call returns nonzero."""
    if status != 0:
        msg = _lib.dcc_last_error().decode(errors="replace")
        raise DCCError("%s failed (status %d): %s" % (what or "libdcc call", status, msg))
    return status


def call(name, *args):
    return check(getattr(_lib, name)(*args), name)


def _as_array(arr, dtype, name):
    """Same contract as the reference's distr_cc/_lib.py:21-27."""
    out = np.asarray(arr)
    if out.dtype != np.dtype(dtype):
        raise TypeError(f"{name} must have dtype {np.dtype(dtype)}")
    if not out.flags["C_CONTIGUOUS"]:
        raise ValueError(f"{name} must be C-contiguous")
    return out
