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
