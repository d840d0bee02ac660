"""Host-side planning of the GPU engine (index maps, batch tables, S3 jobs) checked
against the oracle through the NumPy emulator of the C ABI (no GPU needed)."""
import numpy as np
import pytest

import abi_emulator
from oracle import rccsdt_tri as tri


@pytest.mark.parametrize("o,v", [(1, 2), (3, 4), (4, 5), (5, 3)])
def test_engine_residual_and_update_plan(monkeypatch, o, v):
    abi_emulator.install(monkeypatch)
    from distr_cc.synthetic import SyntheticSystem
    from distr_cc.t3_engine import TriplesEngine
    s = SyntheticSystem(o, v, naux=20, seed=11 + o)
    rng = np.random.default_rng(3)
    im = {k: x + 1e-3 * rng.normal(size=x.shape) for k, x in s.bare_intermediates().items()}
    t2, t3 = s.t2(), s.t3_tri()
    eng = TriplesEngine(o, v)
    eng.set_t3(t3)
    eng.set_intermediates(im, t2)
    eng.residual()
    Ro = tri.compute_r3_tri(o, v, im, t2, t3, batch_size=7)
    assert abs(eng.R.download() - Ro).max() <= 1e-12 * abs(Ro).max()
    n2, na = eng.finalize(s.mo_energy, level_shift=0.02, nvir_diis=max(1, v - 1))
    t3o, _, n2o, nao = tri.t3_update(o, Ro, t3, s.mo_energy, 0.02, max(1, v - 1))
    assert abs(eng.get_t3() - t3o).max() <= 1e-13 * abs(t3o).max()
    assert abs(n2 - n2o) <= 1e-12 * n2o


@pytest.mark.parametrize("o,v", [(1, 3), (3, 4), (4, 3), (2, 5)])
def test_engine_t3_to_r1r2_and_w_plan(monkeypatch, o, v):
    abi_emulator.install(monkeypatch)
    from distr_cc.synthetic import SyntheticSystem
    from distr_cc.t3_engine import TriplesEngine
    s = SyntheticSystem(o, v, naux=20, seed=17 + v)
    im, t2, t3 = s.bare_intermediates(), s.t2(), s.t3_tri()
    nmo = o + v
    rng = np.random.default_rng(4)
    E = s.pppp() + 0.01 * rng.normal(size=(nmo,) * 4)
    E = 0.5 * (E + E.transpose(1, 0, 3, 2))
    F = s.fock + 0.01 * rng.normal(size=(nmo, nmo))
    eng = TriplesEngine(o, v)
    eng.set_t3(t3)
    eng.set_intermediates(im, t2)
    eng.set_t1_dressed(F[:o, o:], E[:o, :o, o:, o:], E[:o, :o, :o, o:], E[o:, :o, o:, o:])
    r1, r2 = eng.t3_to_r1r2()
    r1o, r2o = np.zeros((o, v)), np.zeros((o, o, v, v))
    tri.r1r2_add_t3_tri(o, v, F, E, t3, r1o, r2o)
    floor = 1e-12 * abs(t3).max() * abs(E).max()
    assert abs(r1 - r1o).max() <= 1e-12 * abs(r1o).max() + floor
    assert abs(r2 - r2o).max() <= 1e-12 * abs(r2o).max() + floor
    eng.t3_to_w()
    Wv, Wvv = eng.get_w_t3_adjusted()
    Wvo, Wvvo = im['W_vooo'].copy(), im['W_vvvo'].copy()
    tri.intermediates_t3_add_t3_tri(o, v, E, t3, Wvo, Wvvo)
    assert abs(Wv - Wvo).max() <= 1e-12 * abs(Wvo).max()
    assert abs(Wvv - Wvvo).max() <= 1e-12 * abs(Wvvo).max()


@pytest.mark.parametrize("o,v,bs", [(2, 5, 2), (3, 4, 3), (2, 5, 5), (1, 3, 2)])
def test_q_engine_plan(monkeypatch, o, v, bs):
    abi_emulator.install(monkeypatch)
    from distr_cc.device import DeviceArray
    from distr_cc.q_engine import QEngine
    from distr_cc.synthetic import SyntheticSystem
    from oracle import q_energy as Q
    from oracle import t3_layout as TL
    s = SyntheticSystem(o, v, naux=20, seed=3 + o, sigma_t3=1e-3)
    pppp, t2, t3 = s.pppp(), s.t2(), s.t3_tri()
    ref = Q.kernel(o, v, pppp, t2, TL.full_t3_from_tri(t3, o), s.mo_energy, blksize=bs)
    got = QEngine(o, v, blksize=bs).kernel(pppp, t2, DeviceArray.from_host(t3), s.mo_energy)
    for g, r in zip(got, ref):
        assert abs(g - r) <= 1e-12 * abs(r) + 1e-25


@pytest.mark.parametrize("o,v", [(3, 4), (4, 5)])
def test_engine_fp32_plan(monkeypatch, o, v):
    """FP32 storage path: same plan, FP32 rounding only (relative 1e-5 of the residual scale)."""
    abi_emulator.install(monkeypatch)
    from distr_cc.synthetic import SyntheticSystem
    from distr_cc.t3_engine import TriplesEngine
    s = SyntheticSystem(o, v, naux=20, seed=40 + o)
    im, t2, t3 = s.bare_intermediates(), s.t2(), s.t3_tri()
    eng = TriplesEngine(o, v, dtype=np.float32)
    eng.set_t3(t3)
    eng.set_intermediates(im, t2)
    eng.residual()
    Ro = tri.compute_r3_tri(o, v, im, t2, t3)
    assert abs(eng.R.download().astype(np.float64) - Ro).max() <= 1e-5 * abs(Ro).max()
    eng.finalize(s.mo_energy)
    t3o = tri.t3_update(o, Ro, t3, s.mo_energy)[0]
    assert abs(eng.get_t3() - t3o).max() <= 1e-5 * abs(t3o).max()
