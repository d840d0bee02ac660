"""[Q] and (Q) corrections on the GPU with the interface of distr_cc/rccsdt_q.py.

``kernel(mycc, eris=None, tamps=None, blksize=4, comm=None, job_idx=0,
n_jobs=1, log_redistribution=False) -> (e_bracket, e_paren)`` as in
rccsdt_q.py:227-1191.  The IJK -> ABC redistribution of T3
(rccsdt_q.py:53-80, distribute_t3.py:1544-1642) is a device gather inside
``QEngine``, so ``prepare_tamps_for_q`` only brings T2 to (v, v, o, o) and
leaves T3 in its canonical-block form; both forms of T2 are accepted, as in
the reference (rccsdt_q.py:63-65).
"""
import numpy as np

from distr_cc.device import DeviceArray
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


def prepare_tamps_for_q(mycc, tamps=None, blksize=4, comm=None, chunk_size=None, log_redistribution=True):
    """[t1, t2 (v,v,o,o), t3] (rccsdt_q.py:82-96); T3 keeps its canonical-block layout."""
    if tamps is None:
        tamps = mycc.tamps
    if len(tamps) < 3:
        raise ValueError("RCCSDT(Q) requires at least T1, T2, and T3 amplitudes")
    t1, t2, t3 = tamps[:3]
    return [t1, prepare_t2_for_q(mycc, t2), t3]


def kernel(mycc, eris=None, tamps=None, blksize=4, comm=None, job_idx=0, n_jobs=1, log_redistribution=False):
    """(e_[Q], e_(Q)) of job ``job_idx`` out of ``n_jobs`` (rccsdt_q.py:227)."""
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
    t3_local = np.ascontiguousarray(tamps[2][1], dtype=np.float64)
    t3_dev = DeviceArray.from_host(t3_local)
    try:
        e_bracket, e_paren = QEngine(nocc, nvir, blksize=blksize).kernel(eris.pppp, t2_oovv, t3_dev,
                                                                         eris.mo_energy, job_idx=job_idx,
                                                                         n_jobs=n_jobs)
    finally:
        t3_dev.free()
    return e_bracket, e_paren
