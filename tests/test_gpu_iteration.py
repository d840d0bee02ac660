"""GPU parity of the whole RCCSDT iteration (T1 dressing, lower order, T3, DIIS, loop)."""
import sys

import numpy as np
import pytest

from oracle import rccsdt_lower as lo
from oracle import rccsdt_tri as tri

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def dev():
    from distr_cc import device
    arch = device.init(0)
    _say("device arch %s" % arch)
    return device


@pytest.mark.parametrize("o,v,seed,split", [(1, 4, 3, False), (2, 3, 1, False), (3, 4, 2, False), (4, 7, 5, False),
                                          (5, 9, 6, False), (4, 7, 5, True), (5, 9, 6, True)])
def test_full_iteration_vs_oracle(dev, monkeypatch, o, v, seed, split):
    """``split`` lowers the split-K thresholds so the long-K lower-order contractions run as partial
    slices plus dcc_sum_slices (at production sizes they do so by themselves)."""
    from ccsdt_cases import molecule_like, random_amplitudes
    from distr_cc import rccsdt_engine as RE
    from distr_cc.rccsdt_engine import RCCSDTEngine
    if split:
        monkeypatch.setattr(RE, "SPLIT_K_MIN", 1)
        monkeypatch.setattr(RE, "SPLIT_MIN_SLICES", 2)
    _say("full iteration o=%d v=%d split-K %d" % (o, v, split))
    F, g, e, _ = molecule_like(o, v, seed)
    t1, t2, t3 = random_amplitudes(o, v, seed + 100)
    eng = RCCSDTEngine(o, v, F, g, e, level_shift=0.05, nvir_diis=max(1, v - 1))
    eng.set_amplitudes(t1, t2, t3)
    assert abs(eng.energy() - tri.energy_rhf(o, F, g, t1, t2)) < 1e-13
    res = eng.update_amps()
    a1, a2, a3 = eng.get_amplitudes()
    b1, b2, b3, rres = lo.update_amps(o, F, g, e, t1, t2, t3, level_shift=0.05, nvir_diis=max(1, v - 1))
    assert np.allclose(res, rres, rtol=1e-11, atol=1e-14)
    for a, b in ((a1, b1), (a2, b2), (a3, b3)):
        assert abs(a - b).max() <= 1e-11 * max(abs(b).max(), 1e-30)


def test_kernel_two_electrons_is_fci(dev):
    """Two electrons: RCCSDT == FCI (tests/test_rccsdt.py:221-230 uses He)."""
    from ccsdt_cases import DuckMF, molecule_like
    from distr_cc.distribute_t3 import SINGLE_RANK
    from distr_cc.rccsdt import RCCSDT, _PhysicistsERIs
    from oracle.fci_cc import DetSpace
    o, v = 1, 5
    _say("kernel two-electron o=%d v=%d" % (o, v))
    F, g, e, h = molecule_like(o, v, 31)
    cc = RCCSDT(DuckMF(o, v, e), SINGLE_RANK)
    cc.conv_tol, cc.conv_tol_normt, cc.verbose = 1e-12, 1e-10, 0
    e_cc, _ = cc.kernel(eris=_PhysicistsERIs(F, e, g))
    assert cc.converged
    ds = DetSpace(o + v, o)
    w, U = np.linalg.eigh(ds.hamiltonian(h, g).toarray())
    root = np.argmax(np.abs(U[ds.ref]))
    e_hf = 2 * h[0, 0] + g[0, 0, 0, 0]
    assert abs((w[root] - e_hf) - e_cc) < 1e-9


def test_kernel_matches_oracle_fixed_point(dev):
    from ccsdt_cases import DuckMF, molecule_like
    from distr_cc.distribute_t3 import SINGLE_RANK
    from distr_cc.rccsdt import RCCSDT, _PhysicistsERIs
    o, v = 3, 6
    _say("kernel o=%d v=%d" % (o, v))
    F, g, e, _ = molecule_like(o, v, 23)
    cc = RCCSDT(DuckMF(o, v, e), SINGLE_RANK)
    cc.conv_tol, cc.conv_tol_normt, cc.verbose = 1e-12, 1e-10, 0
    e_cc, tamps = cc.kernel(eris=_PhysicistsERIs(F, e, g))
    assert cc.converged
    _, t1, t2 = tri.init_amps_rhf(o, F, e, g)
    t3 = np.zeros_like(tamps[2][1])
    for _ in range(300):
        t1, t2, t3, res = lo.update_amps(o, F, g, e, t1, t2, t3)
        if max(res) < 1e-12:
            break
    assert abs(e_cc - tri.energy_rhf(o, F, g, t1, t2)) < 1e-10


def test_device_ao2mo(dev):
    from distr_cc.rccsdt import _device_ao2mo
    _say("ao2mo")
    rng = np.random.default_rng(12)
    nao, nocc = 9, 3
    h = rng.standard_normal((nao, nao))
    h = h + h.T
    x = rng.standard_normal((nao,) * 4)
    c = x + x.transpose(1, 0, 2, 3)
    c = c + c.transpose(0, 1, 3, 2)
    c = c + c.transpose(2, 3, 0, 1)
    C = np.linalg.qr(rng.standard_normal((nao, nao)))[0]
    Fmo, pppp, e_elec = _device_ao2mo(h, c, C[:, 1:], C[:, :nocc])
    D = 2 * C[:, :nocc] @ C[:, :nocc].T
    Fao = h + np.einsum("mnls,ls->mn", c, D) - 0.5 * np.einsum("mlns,ls->mn", c, D)
    ref_F = C[:, 1:].T @ Fao @ C[:, 1:]
    chem = np.einsum("mnls,mp,nq,lr,st->pqrt", c, C[:, 1:], C[:, 1:], C[:, 1:], C[:, 1:], optimize=True)
    assert abs(Fmo - ref_F).max() < 1e-11 * abs(ref_F).max()
    assert abs(pppp - chem.transpose(0, 2, 1, 3)).max() < 1e-11 * abs(chem).max()
    assert abs(e_elec - 0.5 * np.sum(D * (h + Fao))) < 1e-10


def test_rccsdt_then_q(dev):
    """RCCSDT(mf, comm).kernel() then rccsdt_q.kernel(): (Q) of the converged amplitudes
    equals the oracle's (Q) on the same amplitudes."""
    from ccsdt_cases import DuckMF, molecule_like
    from distr_cc import rccsdt_q
    from distr_cc.distribute_t3 import SINGLE_RANK
    from distr_cc.rccsdt import RCCSDT, _PhysicistsERIs
    from oracle import q_energy
    from oracle.t3_layout import full_t3_from_tri
    o, v = 3, 6
    _say("rccsdt + (Q) o=%d v=%d" % (o, v))
    F, g, e, _ = molecule_like(o, v, 41)
    cc = RCCSDT(DuckMF(o, v, e), SINGLE_RANK)
    cc.verbose = 0
    eris = _PhysicistsERIs(F, e, g)
    cc.kernel(eris=eris)
    assert cc.converged
    for bs in (2, 4):
        eb, ep = rccsdt_q.kernel(cc, eris=eris, blksize=bs, comm=SINGLE_RANK)
        rb, rp = q_energy.kernel(o, v, g, cc.t2, full_t3_from_tri(cc.t3[1], o), e, blksize=bs)
        assert abs(eb - rb) < 1e-12 and abs(ep - rp) < 1e-12


def test_mixed_iteration_vs_oracle(dev):
    from ccsdt_cases import molecule_like, random_amplitudes
    from distr_cc.rccsdt_engine import RCCSDTEngine
    o, v = 4, 9
    _say("mixed iteration o=%d v=%d" % (o, v))
    F, g, e, _ = molecule_like(o, v, 5)
    t1, t2, t3 = random_amplitudes(o, v, 105)
    eng = RCCSDTEngine(o, v, F, g, e, t3_dtype=np.float32)
    eng.set_amplitudes(t1, t2, t3)
    eng.update_amps()
    a1, a2, a3 = eng.get_amplitudes()
    b1, b2, b3, _ = lo.update_amps(o, F, g, e, t1, t2, t3)
    assert abs(a1 - b1).max() <= 1e-6 * abs(b1).max()
    assert abs(a2 - b2).max() <= 1e-6 * abs(b2).max()
    assert abs(a3 - b3).max() <= 1e-5 * abs(b3).max()


def test_mixed_kernel_energy_error(dev):
    """Converged E_corr with FP32 triples within 1 microhartree of the FP64 path.

    FP64 is converged tightly; the FP32 run uses the reference's default thresholds
    (conv_tol 1e-7, conv_tol_normt 1e-6, rccsdt.py:1164-1165), since FP32 accumulation puts a
    floor under the T3 residual norm."""
    from ccsdt_cases import DuckMF, molecule_like
    from distr_cc.distribute_t3 import SINGLE_RANK
    from distr_cc.rccsdt import RCCSDT, _PhysicistsERIs
    o, v = 4, 10
    _say("mixed kernel o=%d v=%d" % (o, v))
    F, g, e, _ = molecule_like(o, v, 51)
    res = {}
    for prec in ("fp64", "fp32"):
        cc = RCCSDT(DuckMF(o, v, e), SINGLE_RANK)
        cc.verbose, cc.triples_precision = 0, prec
        if prec == "fp64":
            cc.conv_tol, cc.conv_tol_normt = 1e-10, 1e-8
        res[prec] = cc.kernel(eris=_PhysicistsERIs(F, e, g))[0]
        _say("%s converged=%s cycles=%d" % (prec, cc.converged, cc.cycles))
        assert cc.converged
    _say("E fp64 %.12f fp32 %.12f" % (res["fp64"], res["fp32"]))
    assert abs(res["fp64"] - res["fp32"]) < 1e-6


def test_bf16x3_kernel_energy_error(dev):
    """Converged E_corr with split-BF16x3 triples contractions against the FP64 path."""
    from ccsdt_cases import DuckMF, molecule_like
    from distr_cc.distribute_t3 import SINGLE_RANK
    from distr_cc.rccsdt import RCCSDT, _PhysicistsERIs
    o, v = 4, 10
    _say("bf16x3 kernel o=%d v=%d" % (o, v))
    F, g, e, _ = molecule_like(o, v, 51)
    res = {}
    for prec in ("fp64", "bf16x3"):
        cc = RCCSDT(DuckMF(o, v, e), SINGLE_RANK)
        cc.verbose, cc.triples_precision = 0, prec
        if prec == "fp64":
            cc.conv_tol, cc.conv_tol_normt = 1e-10, 1e-8
        res[prec] = cc.kernel(eris=_PhysicistsERIs(F, e, g))[0]
        _say("%s converged=%s cycles=%d E=%.12f" % (prec, cc.converged, cc.cycles, res[prec]))
        assert cc.converged
    assert abs(res["fp64"] - res["bf16x3"]) < 1e-6


def test_kernel_with_autotuned_tiles(dev):
    """RCCSDT.triples_autotune: the tuned tiles give the energy of the default tiles (the tile changes
    neither the K order of a sum nor any result beyond round-off)."""
    from ccsdt_cases import DuckMF, molecule_like
    from distr_cc.distribute_t3 import SINGLE_RANK
    from distr_cc.rccsdt import RCCSDT, _PhysicistsERIs
    o, v = 4, 10
    _say("autotuned kernel o=%d v=%d" % (o, v))
    F, g, e, _ = molecule_like(o, v, 51)
    res = {}
    for tune in (False, True):
        cc = RCCSDT(DuckMF(o, v, e), SINGLE_RANK)
        cc.verbose, cc.triples_autotune = 0, tune
        cc.conv_tol, cc.conv_tol_normt = 1e-10, 1e-8
        res[tune] = cc.kernel(eris=_PhysicistsERIs(F, e, g))[0]
        assert cc.converged
        if tune:
            assert any(t is not None and t >= 0 for t in cc.engine.tr.tile_config())
    _say("E default %.14f tuned %.14f" % (res[False], res[True]))
    assert abs(res[False] - res[True]) < 1e-11


def test_init_amps_level_shift_vs_oracle(dev):
    """RCCSDT.init_amps with level_shift = 0.05 against the oracle's init_amps_rhf (unshifted
    denominators, rccsdt.py:197-251); t1 != 0 through a non-canonical f_ov."""
    from ccsdt_cases import DuckMF, molecule_like, with_fov
    from distr_cc.distribute_t3 import SINGLE_RANK
    from distr_cc.rccsdt import RCCSDT, _PhysicistsERIs
    o, v = 4, 9
    _say("init_amps o=%d v=%d level_shift 0.05" % (o, v))
    F, g, e, _ = molecule_like(o, v, 51)
    F = with_fov(F, o, 0.02, 52)
    cc = RCCSDT(DuckMF(o, v, e), SINGLE_RANK)
    cc.level_shift, cc.verbose = 0.05, 0
    emp2, tamps = cc.init_amps(_PhysicistsERIs(F, e, g))
    e_ref, t1, t2 = tri.init_amps_rhf(o, F, e, g)
    assert abs(tamps[0] - t1).max() <= 1e-14 * abs(t1).max()
    assert abs(tamps[1] - t2).max() <= 1e-14 * abs(t2).max()
    assert abs(emp2 - e_ref) <= 1e-13 * abs(e_ref) and not tamps[2][1].any()


@pytest.mark.parametrize("with_t3,nvir_diis,diis,shift", [(True, None, True, 0.0), (True, 3, True, 0.05),
                                                         (False, None, True, 0.0), (False, 3, True, 0.05),
                                                         (True, None, False, 0.0)])
def test_kernel_diis_trajectory_vs_oracle(dev, with_t3, nvir_diis, diis, shift):
    """Five cycles of RCCSDT.kernel against five cycles of the oracle loop with the oracle DIIS:
    unified T1/T2/T3 DIIS with and without an nvir_diis subspace (run_diis_t3, rccsdt.py:1026-1067),
    standard T1/T2 DIIS (do_diis_max_t = False, rccsdt.py:1096-1099,1119-1122) and none."""
    from ccsdt_cases import DuckMF, molecule_like, oracle_kernel, with_fov
    from distr_cc.distribute_t3 import SINGLE_RANK
    from distr_cc.rccsdt import RCCSDT, _PhysicistsERIs
    o, v = 4, 7
    _say("DIIS trajectory with_t3=%s nvir_diis=%s diis=%s shift=%g" % (with_t3, nvir_diis, diis, shift))
    F, g, e, _ = molecule_like(o, v, 61)
    F = with_fov(F, o, 0.01, 62)
    cc = RCCSDT(DuckMF(o, v, e), SINGLE_RANK)
    cc.do_diis_max_t, cc.nvir_diis, cc.diis, cc.level_shift = with_t3, nvir_diis, diis, shift
    cc.max_cycle, cc.verbose = 5, 0
    e_cc, tamps = cc.kernel(eris=_PhysicistsERIs(F, e, g))
    t1, t2, t3, energies = oracle_kernel(o, F, g, e, 5, level_shift=shift, nvir_diis=nvir_diis, diis=diis,
                                         with_t3=with_t3)
    assert cc.cycles == 5 and abs(e_cc - energies[-1]) <= 1e-12
    for a, b in ((tamps[0], t1), (tamps[1], t2), (tamps[2][1], t3)):
        assert abs(a - b).max() <= 1e-10 * abs(b).max()


def test_device_diis_vs_oracle(dev):
    """DeviceDIIS step by step against oracle.diis (distr_cc/diis.py:105-177): B matrix, coefficients
    and extrapolated vectors, through a wrap of the history (space 4, 9 steps)."""
    from distr_cc.device import DeviceArray
    from distr_cc.diis import DeviceDIIS
    from oracle.diis import DIIS
    n = 200003
    _say("DeviceDIIS n=%d" % n)
    rng = np.random.default_rng(71)
    dd, od = DeviceDIIS(n, space=4), DIIS(space=4)
    x = DeviceArray((n,))
    vec = rng.standard_normal(n)
    for step in range(9):
        # steps of unit size, so that the error vectors x - x_prev do not cancel to round-off
        vec = 0.7 * vec + rng.standard_normal(n)
        x.upload(vec)
        got = dd.update(x)
        want = od.update(vec)
        g = got.download()
        got.free()
        assert abs(g - want).max() <= 1e-12 * abs(want).max()
        if step > 0:
            # B-matrix entries are dot products summed in another order: compare each against the
            # Cauchy-Schwarz scale sqrt(B_ii B_jj) of its terms
            nd = min(step, 4)
            hd, ho = dd._H[1:nd + 1, 1:nd + 1], od._H[1:nd + 1, 1:nd + 1]
            scale = np.sqrt(np.outer(np.diag(ho), np.diag(ho)))
            assert np.all(np.abs(hd - ho) <= 1e-12 * scale)
            cd, co = dd.coefficients(nd), od.coefficients(nd)
            assert np.abs(cd - co).max() <= 1e-9 * np.abs(co).max()
