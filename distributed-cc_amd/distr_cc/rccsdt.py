"""RCCSDT driver with the public interface of the reference ``distr_cc.rccsdt``.

``RCCSDT(mf, comm, frozen=None, mo_coeff=None, mo_occ=None).kernel(tamps=None,
eris=None) -> (e_corr, tamps)`` as in distr_cc/rccsdt.py:1162-1308, with the
convergence loop of ``kernel`` (rccsdt.py:1069-1148) and the unified T1/T2/T3
DIIS of ``run_diis_t3`` (rccsdt.py:1026-1067).  Every amplitude, integral and
intermediate of the iteration lives on the GPU (``RCCSDTEngine``); the host
holds the loop, the DIIS B matrix and the log.

PySCF is not available to this package, so ``mf`` is duck-typed: it needs
``mo_coeff``, ``mo_occ`` and ``mo_energy``, and -- only when no ``eris`` is
passed -- ``get_hcore()``, ``_eri`` (AO chemist integrals, full (n,)*4 or
8-fold packed) and ``energy_nuc()``.  ``eris`` may be any object with
``fock``, ``mo_energy`` and ``pppp`` (physicist <pq|rs> over the active
orbitals, occupied first), the fields the reference reads
(rccsdt.py:203-221).  ``comm`` is required as in the reference
(rccsdt.py:1211-1213); this build runs on one GPU, so its size must be 1.
"""
import sys

import numpy as np

from distr_cc import _lib as L
from distr_cc.device import DeviceArray
from distr_cc.diis import DeviceDIIS
from distr_cc.distribute_t3 import DistributedT3IJK
from distr_cc.rccsdt_engine import RCCSDTEngine
from distr_cc.t3_engine import Operand, plan_einsum


class _PhysicistsERIs:
    """fock, mo_energy and pppp = <pq|rs> over the active orbitals (rccsdt.py:203-221)."""

    def __init__(self, fock, mo_energy, pppp, e_hf=None):
        self.fock = np.asarray(fock, dtype=np.float64)
        self.mo_energy = np.asarray(mo_energy, dtype=np.float64)
        self.pppp = np.asarray(pppp, dtype=np.float64)
        self.e_hf = e_hf


def _restore_eri(eri, nao):
    """Full (nao,)*4 chemist integrals from a full, 4-fold or 8-fold packed array."""
    eri = np.asarray(eri, dtype=np.float64)
    if eri.ndim == 4:
        return eri
    npair = nao * (nao + 1) // 2
    iu = np.triu_indices(nao)
    pair = np.zeros((nao, nao), dtype=np.int64)
    pair[iu] = np.arange(npair)
    pair = np.maximum(pair, pair.T)
    if eri.ndim == 2 and eri.shape == (npair, npair):
        return eri[pair[:, :, None, None], pair[None, None, :, :]]
    if eri.ndim == 1 and eri.size == npair * (npair + 1) // 2:
        full = np.zeros((npair, npair))
        ju = np.triu_indices(npair)
        full[ju] = eri
        full = full + full.T - np.diag(np.diag(full))
        return full[pair[:, :, None, None], pair[None, None, :, :]]
    raise ValueError("unrecognised _eri layout %r" % (eri.shape,))


def _device_ao2mo(h_ao, eri_ao, C, C_occ, stream=None):
    """Fock matrix, <pq|rs> and the HF electronic energy in the MO basis C, on the GPU.

    J/K of the closed-shell density 2 C_occ C_occ^T and the four quarter
    transformations are gather-GEMM launches."""
    nao, nmo = C.shape
    ext = {x: nao for x in "MNLS"}
    ext.update({x: nmo for x in "pqrs"})
    one = [({}, {}, {})]
    dev = {}

    def up(name, a):
        dev[name] = DeviceArray.from_host(np.ascontiguousarray(a, dtype=np.float64), stream)
        return dev[name]

    eri = up("eri", eri_ao)
    D = up("D", 2.0 * C_occ @ C_occ.T)
    h = up("h", h_ao)
    Cd = up("C", C)
    F = DeviceArray((nao, nao))
    F.upload(np.ascontiguousarray(h_ao, dtype=np.float64), stream)
    E = lambda buf, l: Operand(buf, l, buf.shape)
    plan_einsum("J", "MNLS,LS->MN", E(eri, "MNLS"), E(D, "LS"), E(F, "MN"), 1.0, 1.0, ext, one).run(stream)
    plan_einsum("K", "MLNS,LS->MN", E(eri, "MLNS"), E(D, "LS"), E(F, "MN"), -0.5, 1.0, ext, one).run(stream)
    T1 = DeviceArray((nmo, nao, nao, nao))
    T2 = DeviceArray((nmo, nmo, nao, nao))
    plan_einsum("q1", "Mp,MNLS->pNLS", E(Cd, "Mp"), E(eri, "MNLS"), E(T1, "pNLS"), 1.0, 0.0, ext, one).run(stream)
    plan_einsum("q2", "pNLS,Nq->pqLS", E(T1, "pNLS"), E(Cd, "Nq"), E(T2, "pqLS"), 1.0, 0.0, ext, one).run(stream)
    T1.free()
    T3 = DeviceArray((nmo, nmo, nmo, nao))
    plan_einsum("q3", "Lr,pqLS->pqrS", E(Cd, "Lr"), E(T2, "pqLS"), E(T3, "pqrS"), 1.0, 0.0, ext, one).run(stream)
    T2.free()
    chem = DeviceArray((nmo, nmo, nmo, nmo))
    plan_einsum("q4", "pqrS,Ss->pqrs", E(T3, "pqrS"), E(Cd, "Ss"), E(chem, "pqrs"), 1.0, 0.0, ext, one).run(stream)
    T3.free()
    # physicist <pq|rs> = (pr|qs)
    pppp = DeviceArray((nmo, nmo, nmo, nmo))
    from distr_cc.t3_engine import tensor_axpby
    tensor_axpby(Operand(pppp, "prqs", (nmo,) * 4), Operand(chem, "pqrs", (nmo,) * 4), "pqrs", (nmo,) * 4,
                 1.0, 0.0, stream)
    Fa = F.download(stream=stream)
    Fmo = C.T @ Fa @ C
    e_elec = 0.5 * float(np.sum(2.0 * C_occ @ C_occ.T * (h_ao + Fa)))
    out = pppp.download(stream=stream)
    for a in (pppp, chem, eri, D, h, Cd, F):
        a.free()
    return Fmo, out, e_elec


class RCCSDT:
    """Closed-shell CCSDT with triangular (i <= j <= k) T3 storage on one GPU."""

    # defaults of rccsdt.py:1164-1191 and of the CCSDBase attributes it inherits
    conv_tol = 1e-7
    conv_tol_normt = 1e-6
    cc_order = 3
    do_diis_max_t = True
    batch_size = 100
    einsum_backend = "numpy"
    nvir_diis = None
    max_cycle = 50
    diis = True
    diis_space = 6
    diis_start_cycle = 0
    diis_start_energy_diff = 1e9
    level_shift = 0.0
    iterative_damping = 1.0
    verbose = 4
    callback = None
    # "fp64" (the reference's arithmetic), "fp32" (T3 and the triples contractions in FP32 on MFMA)
    # or "bf16x3" (FP32 T3, contractions as split-BF16x3 on BF16 MFMA); everything else in FP64
    triples_precision = "fp64"
    # time the workgroup tiles of every triples contraction once, when the device engine is built for
    # a system, and keep the fastest (TriplesEngine.autotune); off: the library's default tiles
    triples_autotune = False

    def __init__(self, mf, comm=None, frozen=None, mo_coeff=None, mo_occ=None):
        if comm is None:
            raise ValueError("This %s implementation requires an MPI communicator." % self.__class__.__name__)
        self.comm = comm
        self.rank = comm.Get_rank()
        self.size = comm.Get_size()
        if self.size != 1:
            raise ValueError("this build runs RCCSDT on one GPU per process; comm size must be 1")
        self._scf = mf
        self.mol = getattr(mf, "mol", None)
        self.frozen = frozen
        self.mo_coeff = mf.mo_coeff if mo_coeff is None else mo_coeff
        self.mo_occ = mf.mo_occ if mo_occ is None else mo_occ
        self.stdout = sys.stdout
        self.tamps = [None, None, None]
        self.e_hf = getattr(mf, "e_tot", None)
        self.e_corr = None
        self.emp2 = None
        self.converged = False
        self.cycles = 0
        self.unique_tamps_map = None
        self.engine = None

    # -- orbital bookkeeping (pyscf.mp.mp2.get_frozen_mask semantics)
    def get_frozen_mask(self):
        nmo = len(self.mo_occ)
        mask = np.ones(nmo, dtype=bool)
        if self.frozen is None or (np.isscalar(self.frozen) and int(self.frozen) == 0):
            return mask
        if np.isscalar(self.frozen):
            mask[: int(self.frozen)] = False
        else:
            mask[np.asarray(self.frozen, dtype=int)] = False
        return mask

    @property
    def nocc(self):
        return int(np.count_nonzero(np.asarray(self.mo_occ)[self.get_frozen_mask()] > 0))

    @property
    def nmo(self):
        return int(np.count_nonzero(self.get_frozen_mask()))

    @property
    def t1(self):
        return self.tamps[0]

    @t1.setter
    def t1(self, val):
        self.tamps[0] = val

    @property
    def t2(self):
        return self.tamps[1]

    @t2.setter
    def t2(self, val):
        self.tamps[1] = val

    @property
    def t3(self):
        return self.tamps[2]

    @t3.setter
    def t3(self, val):
        self.tamps[2] = val

    @property
    def e_tot(self):
        return (self.e_hf or 0.0) + (self.e_corr or 0.0)

    do_tri_max_t = property(lambda self: True)

    def set_einsum_backend(self, backend):
        """Accepted for compatibility (rccsdt.py:1255); every contraction runs on the GPU."""
        self.einsum_backend = backend

    # -- integrals
    def ao2mo(self, mo_coeff=None):
        mf = self._scf
        C = np.asarray(self.mo_coeff if mo_coeff is None else mo_coeff, dtype=np.float64)
        if not hasattr(mf, "get_hcore") or getattr(mf, "_eri", None) is None:
            raise ValueError("no integrals: pass eris=... or an mf with get_hcore() and _eri")
        h_ao = np.asarray(mf.get_hcore(), dtype=np.float64)
        nao = h_ao.shape[0]
        eri = _restore_eri(mf._eri, nao)
        occ = np.asarray(self.mo_occ) > 0
        mask = self.get_frozen_mask()
        Fmo, pppp, e_elec = _device_ao2mo(h_ao, eri, C[:, mask], C[:, occ])
        e_nuc = float(mf.energy_nuc()) if hasattr(mf, "energy_nuc") else 0.0
        return _PhysicistsERIs(Fmo, np.diag(Fmo).copy(), pppp, e_elec + e_nuc)

    # -- amplitudes
    def _engine(self, eris):
        """The device engine of ``eris``: built once per (eris, triples_precision); the level shift and
        nvir_diis are pushed into it at every call, and the tiles are tuned once when
        triples_autotune is set."""
        if self.engine is None or self.engine._eris is not eris or \
                self.engine._precision != self.triples_precision:
            nocc, nmo = self.nocc, self.nmo
            t3_dtype, contract = {"fp64": (np.float64, "native"), "fp32": (np.float32, "native"),
                                  "bf16x3": (np.float32, "bf16x3")}[self.triples_precision]
            self.engine = None   # release the previous engine's device memory first
            self.engine = RCCSDTEngine(nocc, nmo - nocc, eris.fock, eris.pppp, eris.mo_energy,
                                       level_shift=self.level_shift, nvir_diis=self.nvir_diis, t3_dtype=t3_dtype,
                                       triples_contract=contract)
            self.engine._eris = eris
            self.engine._precision = self.triples_precision
            self.engine._tuned = False
        self.engine.configure(level_shift=self.level_shift, nvir_diis=self.nvir_diis)
        if self.triples_autotune and not self.engine._tuned:
            # runs the contractions on the engine's current buffers; every buffer they write is
            # cleared or overwritten before the first iteration reads it
            self.engine.tr.autotune()
            self.engine._tuned = True
        return self.engine

    def init_amps(self, eris=None):
        """MP2 t1, t2 and T3 = 0 (rccsdt.py:197-251); returns (emp2, tamps)."""
        if eris is None:
            eris = self.ao2mo(self.mo_coeff)
        eng = self._engine(eris)
        eng.init_amps()
        self.emp2 = eng.energy()
        return self.emp2, self._host_tamps(eng)

    def _host_tamps(self, eng):
        t1, t2, t3_local = eng.get_amplitudes()
        dt3 = DistributedT3IJK(eng.nocc, eng.nvir, self.comm)
        return [t1, t2, (dt3, t3_local)]

    def energy(self, tamps=None, eris=None):
        eng = self.engine
        if tamps is not None:
            eng.set_amplitudes(tamps[0], tamps[1], tamps[2][1])
        return eng.energy()

    def kernel(self, tamps=None, eris=None):
        return self.ccsdt(tamps, eris)

    def ccsdt(self, tamps=None, eris=None):
        if eris is None:
            eris = self.ao2mo(self.mo_coeff)
        if getattr(eris, "e_hf", None) is not None:
            self.e_hf = eris.e_hf
        eng = self._engine(eris)
        if tamps is None:
            eng.init_amps()
        else:
            tamps = list(tamps)
            if len(tamps) < 3:
                eng.init_amps()
                t1, t2, t3 = eng.get_amplitudes()
                tamps += [None] * (3 - len(tamps))
                tamps = [tamps[0] if tamps[0] is not None else t1, tamps[1] if tamps[1] is not None else t2,
                         tamps[2] if tamps[2] is not None else (None, t3)]
            eng.set_amplitudes(tamps[0], tamps[1], tamps[2][1])
        self.converged, self.e_corr = kernel(self, eng, tol=self.conv_tol, tolnormt=self.conv_tol_normt,
                                             max_cycle=self.max_cycle)
        self.tamps = self._host_tamps(eng)
        return self.e_corr, self.tamps


def kernel(mycc, eng, tol=1e-8, tolnormt=1e-6, max_cycle=50):
    """Convergence loop of rccsdt.py:1069-1148 over the device engine."""
    log = mycc.stdout if mycc.verbose >= 4 else None
    e_corr_old = 0.0
    e_corr = eng.energy()
    if log:
        print("Init E_corr(RCCSDT) = %.15g" % e_corr, file=log)
    # unified T1/T2/T3 DIIS (run_diis_t3) or, with do_diis_max_t = False, the standard DIIS over
    # T1/T2 only, T3 not extrapolated (rccsdt.py:1096-1099,1114-1122)
    with_t3 = bool(mycc.do_diis_max_t)
    if getattr(mycc, "iterative_damping", 1.0) < 1.0:
        raise NotImplementedError("Damping is not implemented")
    adiis = DeviceDIIS(eng.diis_size(with_t3), mycc.diis_space) if mycc.diis else None
    vec = DeviceArray((eng.diis_size(with_t3),)) if adiis is not None else None
    converged = False
    mycc.cycles = 0
    for istep in range(max_cycle):
        res_norm = eng.update_amps()
        if mycc.callback is not None:
            mycc.callback(locals())
        normt = float(np.linalg.norm(res_norm))
        if adiis is not None and istep >= mycc.diis_start_cycle and abs(e_corr - e_corr_old) < \
                mycc.diis_start_energy_diff:
            eng.pack(vec, with_t3)
            xnew = adiis.update(vec)
            eng.unpack(xnew, with_t3)
            xnew.free()
        e_corr_old, e_corr = e_corr, eng.energy()
        mycc.cycles = istep + 1
        if log:
            print("cycle = %2d  E_corr(RCCSDT) = % .12f  dE = % .12e  norm(res) = %.8e"
                  % (istep + 1, e_corr, e_corr - e_corr_old, normt), file=log)
        if abs(e_corr - e_corr_old) < tol and normt < tolnormt:
            converged = True
            break
    return converged, e_corr
