"""[Q] and (Q) corrections on the GPU with the interface of distr_cc/rccsdt_q.py.

``kernel(mycc, eris=None, tamps=None, blksize=4, comm=None, job_idx=0,
n_jobs=1, log_redistribution=False) -> (e_bracket, e_paren)`` as in
rccsdt_q.py:227-1191.  T3 may come in the canonical IJK layout or ABC-prepared
(``prepare_tamps_for_q``, rccsdt_q.py:53-96: the IJK -> ABC redistribution of
distribute_t3.py:1544-1642 is one device gather here); both forms of T2 are
accepted, as in the reference (rccsdt_q.py:42-50).
"""
import numpy as np

from distr_cc.device import DeviceArray
from distr_cc.distribute_t3 import redistribute_ijk_to_abc
from distr_cc.q_engine import QEngine


def prepare_t2_for_q(mycc, t2):
    """(o,o,v,v) -> (v,v,o,o); an already prepared (v,v,o,o) T2 passes (rccsdt_q.py:42-50)."""
    nocc, nmo = mycc.nocc, mycc.nmo
    nvir = nmo - nocc
    if t2.shape == (nvir, nvir, nocc, nocc):
        return np.ascontiguousarray(t2)
    if t2.shape == (nocc, nocc, nvir, nvir):
        return np.ascontiguousarray(t2.transpose(2, 3, 0, 1))
    raise ValueError("RCCSDT(Q) expected T2 with shape (nocc,nocc,nvir,nvir) "
                     "or prepared shape (nvir,nvir,nocc,nocc); got %s" % (t2.shape,))


def _is_abc_distributed_t3(t3):
    """rccsdt_q.py:75-79: an ABC-prepared T3 is recognised by its abc_to_global_idx table."""
    try:
        return hasattr(t3[0], "abc_to_global_idx")
    except Exception:
        return False


def prepare_t3_for_q(t3, comm=None, blksize=4, chunk_size=None, log_redistribution=True, verbose=None,
                     stdout=None):
    """(DistributedT3IJK, t3_local) -> (DistributedT3ABC, t3_abc) (rccsdt_q.py:59-66); ABC input passes."""
    if _is_abc_distributed_t3(t3):
        return t3
    dt3, t3_local = t3
    return redistribute_ijk_to_abc(dt3, t3_local, comm, blksize_abc=blksize, chunk_size=chunk_size,
                                   log_redistribution=log_redistribution, verbose=verbose, stdout=stdout)


def prepare_tamps_for_q(mycc, tamps=None, blksize=4, comm=None, chunk_size=None, log_redistribution=True):
    """[t1, t2 (v,v,o,o), (DistributedT3ABC, t3_abc)] (rccsdt_q.py:82-96); the redistribution is a device
    gather."""
    if tamps is None:
        tamps = mycc.tamps
    if len(tamps) < 3:
        raise ValueError("RCCSDT(Q) requires at least T1, T2, and T3 amplitudes")
    if comm is None:
        comm = getattr(mycc, "comm", None)
    t1, t2, t3 = tamps[:3]
    return [t1, prepare_t2_for_q(mycc, t2), prepare_t3_for_q(t3, comm=comm, blksize=blksize, chunk_size=chunk_size,
                                                             log_redistribution=log_redistribution)]


def kernel(mycc, eris=None, tamps=None, blksize=4, comm=None, job_idx=0, n_jobs=1, log_redistribution=False):
    """(e_[Q], e_(Q)) of job ``job_idx`` out of ``n_jobs`` (rccsdt_q.py:227).  ``tamps`` may hold T3 in
    the canonical IJK layout or ABC-prepared (prepare_tamps_for_q), T2 as (o,o,v,v) or (v,v,o,o)."""
    if comm is not None and comm.Get_size() != 1:
        raise ValueError("this build computes (Q) on one GPU per process; comm size must be 1")
    if eris is None:
        eris = getattr(mycc.engine, "_eris", None) if getattr(mycc, "engine", None) is not None else None
        if eris is None:
            eris = mycc.ao2mo(mycc.mo_coeff)
    if tamps is None:
        tamps = mycc.tamps
    nocc, nmo = mycc.nocc, mycc.nmo
    nvir = nmo - nocc
    t2 = tamps[1]
    t2_oovv = t2 if t2.shape == (nocc, nocc, nvir, nvir) else np.ascontiguousarray(t2.transpose(2, 3, 0, 1))
    abc = _is_abc_distributed_t3(tamps[2])
    t3_local = np.ascontiguousarray(tamps[2][1], dtype=np.float64)
    t3_dev = DeviceArray.from_host(t3_local)
    try:
        e_bracket, e_paren = QEngine(nocc, nvir, blksize=blksize).kernel(eris.pppp, t2_oovv, t3_dev,
                                                                         eris.mo_energy, job_idx=job_idx,
                                                                         n_jobs=n_jobs,
                                                                         t3_layout="abc" if abc else "ijk")
    finally:
        t3_dev.free()
    return e_bracket, e_paren
