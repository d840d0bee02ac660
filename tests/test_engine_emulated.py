"""Host-side planning of the GPU engine (index maps, batch tables, S3 jobs) checked
against the oracle through the NumPy emulator of the C ABI (no GPU needed)."""
import ctypes

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
    # (an i=j=k block is entirely abc-symmetric, so its projected update is round-off: absolute floor)
    assert abs(n2 - n2o) <= 1e-12 * n2o + 1e-26


@pytest.mark.parametrize("o,v,pairs,fused", [(4, 5, 3, True), (5, 3, 1, True), (4, 4, 4, False)])
def test_engine_ring_chunked(monkeypatch, o, v, pairs, fused):
    """The ring operator G computed for a few pairs at a time (the memory-bounded path of large
    systems), with and without the fused W_ovvo + Z GEMM."""
    abi_emulator.install(monkeypatch)
    from distr_cc.synthetic import SyntheticSystem
    from distr_cc.t3_engine import TriplesEngine
    s = SyntheticSystem(o, v, naux=20, seed=19 + o)
    rng = np.random.default_rng(4)
    im = {k: x + 1e-3 * rng.normal(size=x.shape) for k, x in s.bare_intermediates().items()}
    t2, t3 = s.t2(), s.t3_tri()
    eng = TriplesEngine(o, v, ring_fused=fused, ring_g_bytes=pairs * o * v ** 3 * 8)
    assert eng.Gs.shape[0] == pairs
    eng.set_t3(t3)
    eng.set_intermediates(im, t2)
    eng.residual()
    Ro = tri.compute_r3_tri(o, v, im, t2, t3, batch_size=7)
    assert abs(eng.R.download() - Ro).max() <= 1e-12 * abs(Ro).max()


def pair_symmetric_noise(im, rng, scale=1e-3):
    """Independent noise on every intermediate except W_vvvv, whose noise keeps its pair symmetry
    W[a,b,d,e] = W[b,a,e,d] (and nothing more: it is not symmetric in (a,b) or (d,e) alone)."""
    out = {k: x + scale * rng.normal(size=x.shape) for k, x in im.items()}
    n = scale * rng.normal(size=im['W_vvvv'].shape)
    out['W_vvvv'] = im['W_vvvv'] + n + n.transpose(1, 0, 3, 2)
    return out


@pytest.mark.parametrize("o,v,prec,chunk", [(3, 5, "fp64", None), (4, 4, "fp64", 7), (3, 6, "fp32", 5),
                                            (3, 7, "bf16x3", None), (2, 9, "bf16x3", 2)])
def test_engine_vvvv_pairs(monkeypatch, o, v, prec, chunk):
    """W_vvvv x T3 as symmetric / antisymmetric products over triangular pairs (split, one two-entry
    GEMM, unfold per term and chunk of blocks) against the oracle's plain einsums."""
    abi_emulator.install(monkeypatch)
    from distr_cc.synthetic import SyntheticSystem
    from distr_cc.t3_engine import GemmLaunch, TriplesEngine, pair_dims
    s = SyntheticSystem(o, v, naux=20, seed=41 + v)
    im = pair_symmetric_noise(s.bare_intermediates(), np.random.default_rng(6))
    t2, t3 = s.t2(), s.t3_tri()
    kw = {"fp64": {}, "fp32": dict(dtype=np.float32),
          "bf16x3": dict(dtype=np.float32, contract="bf16x3")}[prec]
    nS, nA, kp = pair_dims(v)
    assert (nS, nA) == (v * (v + 1) // 2, v * (v - 1) // 2) and kp % 64 == 0 and kp >= nS
    isz = 8 if prec == "fp64" else 4
    eng = TriplesEngine(o, v, vvvv_pairs=True,
                        pair_bytes=None if chunk is None else chunk * 2 * v * (kp + nS) * isz, **kw)
    # chunks of equal size within the bound (and within the T2 -> T3 / ring scratch they borrow)
    nbc = eng.Tp.shape[1]
    nchunk = -(-eng.n3 // nbc)
    assert nbc <= (chunk or eng.n3) and nbc == -(-eng.n3 // nchunk)
    assert eng.Tp.nbytes <= eng.Xt.nbytes and eng.Up.nbytes <= eng.Gs.nbytes
    pairs = [g for g in eng.launches if isinstance(g, GemmLaunch) and g.name.startswith("W_vvvv pairs")]
    assert len(pairs) == 3 * nchunk and not any(g.name.startswith("W_vvvv a") for g in eng.launches)
    if prec == "bf16x3":
        assert all(g.fn == "dcc_contract_packed_bf16x3" for g in pairs)
    eng.set_t3(t3)
    eng.set_intermediates(im, t2)
    eng.residual()
    Ro = tri.compute_r3_tri(o, v, im, t2, t3)
    tol = {"fp64": 1e-12, "fp32": 2e-5, "bf16x3": 3e-4}[prec]
    assert abs(eng.R.download().astype(np.float64) - Ro).max() <= tol * abs(Ro).max()


def test_engine_vvvv_pairs_rejects_asymmetric(monkeypatch):
    """An engine with pair-symmetric W_vvvv products refuses intermediates whose W_vvvv lacks the
    symmetry, instead of computing the symmetrised term."""
    abi_emulator.install(monkeypatch)
    from distr_cc.synthetic import SyntheticSystem
    from distr_cc.t3_engine import TriplesEngine
    s = SyntheticSystem(3, 4, naux=20, seed=5)
    rng = np.random.default_rng(8)
    im = {k: x + 1e-3 * rng.normal(size=x.shape) for k, x in s.bare_intermediates().items()}
    eng = TriplesEngine(3, 4, vvvv_pairs=True)
    with pytest.raises(ValueError, match="pair-symmetric"):
        eng.set_intermediates(im, s.t2())
    eng.set_intermediates(pair_symmetric_noise(s.bare_intermediates(), rng), s.t2())


def test_engine_oooo_chunk_rounding(monkeypatch):
    """W_oooo/F_oo target chunks: a bound of 52 targets is rounded up to one 64-row chunk (within
    5/4 of the bound), leaving a partial second chunk of 20 of the 84 targets."""
    abi_emulator.install(monkeypatch)
    from distr_cc.synthetic import SyntheticSystem
    from distr_cc.t3_engine import TriplesEngine
    o, v = 7, 3
    s = SyntheticSystem(o, v, naux=20, seed=23)
    rng = np.random.default_rng(5)
    im = {k: x + 1e-3 * rng.normal(size=x.shape) for k, x in s.bare_intermediates().items()}
    t2, t3 = s.t2(), s.t3_tri()
    eng = TriplesEngine(o, v, ring_g_bytes=52 * 5 * v ** 3 * 8)
    assert eng.n3 == 84 and eng.Xp.shape[1] == 64
    eng.set_t3(t3)
    eng.set_intermediates(im, t2)
    eng.residual()
    Ro = tri.compute_r3_tri(o, v, im, t2, t3, batch_size=11)
    assert abs(eng.R.download() - Ro).max() <= 1e-12 * abs(Ro).max()


@pytest.mark.parametrize("o,v", [(1, 3), (3, 4), (4, 3), (2, 5)])
def test_engine_t3_to_r1r2_and_w_plan(monkeypatch, o, v):
    _t3_to_r1r2_and_w(monkeypatch, o, v, None)


@pytest.mark.parametrize("o,v,qc", [(4, 3, 1), (5, 3, 2), (3, 4, 2)])
def test_engine_t3_to_r1r2_and_w_chunked(monkeypatch, o, v, qc):
    """Middle index q processed in chunks of qc (the memory-bounded path of large systems)."""
    _t3_to_r1r2_and_w(monkeypatch, o, v, qc)


def _t3_to_r1r2_and_w(monkeypatch, o, v, qc):
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
    if qc is not None:
        eng._setup_lower(z_budget_bytes=qc * o * o * v ** 3 * 8)
        assert eng._lower['qc'] == qc and len(eng._lower['chunks']) == -(-o // qc)
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
@pytest.mark.parametrize("w_mode", ["dense", "task"])
def test_q_engine_plan(monkeypatch, o, v, bs, w_mode):
    """Both W_vvvvoo forms: dense, and per-chunk slots of the distinct ordered block triples (with a
    one-task chunk size, so that every chunk builds its own slots)."""
    abi_emulator.install(monkeypatch)
    from distr_cc.device import DeviceArray
    from distr_cc.q_engine import QEngine
    from distr_cc.synthetic import SyntheticSystem
    from oracle import q_energy as Q
    from oracle import t3_layout as TL
    s = SyntheticSystem(o, v, naux=20, seed=3 + o, sigma_t3=1e-3)
    pppp, t2, t3 = s.pppp(), s.t2(), s.t3_tri()
    ref = Q.kernel(o, v, pppp, t2, TL.full_t3_from_tri(t3, o), s.mo_energy, blksize=bs)
    chunk_bytes = bs ** 4 * o ** 4 * 8 if w_mode == "task" else 1 << 30
    eng = QEngine(o, v, blksize=bs, w_mode=w_mode, max_chunk_bytes=chunk_bytes)
    got = eng.kernel(pppp, t2, DeviceArray.from_host(t3), s.mo_energy)
    for g, r in zip(got, ref):
        assert abs(g - r) <= 1e-12 * abs(r) + 1e-25
    if w_mode == "task":
        from distr_cc.q_engine import create_full_task_list
        assert "Wv" not in eng.dev
        assert len(next(iter(eng.task_plans.values()))[-1]) == len(create_full_task_list(v, bs))


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


@pytest.mark.parametrize("split", [False, True])
@pytest.mark.parametrize("o,v,seed", [(2, 3, 1), (3, 4, 2), (1, 4, 3)])
def test_full_iteration_plan(monkeypatch, o, v, seed, split):
    """RCCSDTEngine.update_amps (dressing, lower order, T3, update) == oracle update_amps; ``split``
    lowers the split-K thresholds so every long-K lower-order contraction takes the partial-slice path."""
    abi_emulator.install(monkeypatch)
    from ccsdt_cases import molecule_like, random_amplitudes
    from distr_cc import rccsdt_engine as RE
    from distr_cc.rccsdt_engine import RCCSDTEngine
    from oracle import rccsdt_lower as lo
    if split:
        monkeypatch.setattr(RE, "SPLIT_K_MIN", 1)
        monkeypatch.setattr(RE, "SPLIT_MIN_SLICES", 2)
    F, g, e, _ = molecule_like(o, v, seed)
    t1, t2, t3 = random_amplitudes(o, v, seed + 100)
    eng = RCCSDTEngine(o, v, F, g, e, level_shift=0.05, nvir_diis=max(1, v - 1))
    eng.set_amplitudes(t1, t2, t3)
    assert abs(eng.energy() - tri.energy_rhf(o, F, g, t1, t2)) < 1e-14
    res = eng.update_amps()
    a1, a2, a3 = eng.get_amplitudes()
    b1, b2, b3, rres = lo.update_amps(o, F, g, e, t1, t2, t3, level_shift=0.05, nvir_diis=max(1, v - 1))
    assert np.allclose(res, rres, rtol=1e-12, atol=1e-15)
    for a, b in ((a1, b1), (a2, b2), (a3, b3)):
        assert abs(a - b).max() <= 1e-12 * max(abs(b).max(), 1e-30)


def test_diis_packing_round_trip_and_steps(monkeypatch):
    abi_emulator.install(monkeypatch)
    from ccsdt_cases import molecule_like, random_amplitudes
    from distr_cc.device import DeviceArray
    from distr_cc.diis import DeviceDIIS
    from distr_cc.rccsdt_engine import RCCSDTEngine
    from oracle.diis import DIIS
    o, v = 3, 4
    F, g, e, _ = molecule_like(o, v, 7)
    eng = RCCSDTEngine(o, v, F, g, e, nvir_diis=3)
    t1, t2, t3 = random_amplitudes(o, v, 8)
    eng.set_amplitudes(t1, t2, t3)
    x = DeviceArray((eng.diis_size(),))
    eng.pack(x)
    host = x.download()
    pairs = [(i, j) for i in range(o) for j in range(i, o)]
    ref = np.concatenate([t1.ravel(), np.concatenate([t2[i, j].ravel() for i, j in pairs]),
                          t3[:, :3, :3, :3].ravel()])
    assert np.array_equal(host, ref)
    eng.set_amplitudes(0 * t1, 0 * t2, t3)
    eng.unpack(x)
    a1, a2, a3 = eng.get_amplitudes()
    assert np.array_equal(a1, t1) and np.array_equal(a2, t2) and np.array_equal(a3, t3)
    rng = np.random.default_rng(9)
    dd, od = DeviceDIIS(ref.size, space=3), DIIS(space=3)
    vec = ref.copy()
    for step in range(6):
        vec = vec + 0.5 ** step * rng.standard_normal(vec.size)
        x.upload(vec)
        got = dd.update(x).download()
        want = od.update(vec)
        assert abs(got - want).max() <= 1e-12 * abs(want).max()


def test_kernel_converges_to_oracle_fixed_point(monkeypatch):
    abi_emulator.install(monkeypatch)
    from ccsdt_cases import DuckMF, molecule_like
    from distr_cc.distribute_t3 import SINGLE_RANK
    from distr_cc.rccsdt import RCCSDT, _PhysicistsERIs
    from oracle import rccsdt_lower as lo
    o, v = 2, 3
    F, g, e, _ = molecule_like(o, v, 22)
    cc = RCCSDT(DuckMF(o, v, e), SINGLE_RANK)
    cc.conv_tol, cc.conv_tol_normt, cc.verbose = 1e-12, 1e-10, 0
    e_cc, tamps = cc.kernel(eris=_PhysicistsERIs(F, e, g))
    assert cc.converged
    _, t1, t2 = tri.init_amps_rhf(o, F, e, g)
    t3 = np.zeros_like(tamps[2][1])
    for _ in range(200):
        t1, t2, t3, res = lo.update_amps(o, F, g, e, t1, t2, t3)
        if max(res) < 1e-12:
            break
    assert abs(e_cc - tri.energy_rhf(o, F, g, t1, t2)) < 1e-10
    assert abs(tamps[1] - t2).max() < 1e-8


def test_mixed_iteration_plan(monkeypatch):
    """FP32 triples inside the FP64 iteration: one step stays within FP32 rounding of the oracle."""
    abi_emulator.install(monkeypatch)
    from ccsdt_cases import molecule_like, random_amplitudes
    from distr_cc.rccsdt_engine import RCCSDTEngine
    from oracle import rccsdt_lower as lo
    o, v = 3, 4
    F, g, e, _ = molecule_like(o, v, 2)
    t1, t2, t3 = random_amplitudes(o, v, 102)
    eng = RCCSDTEngine(o, v, F, g, e, t3_dtype=np.float32)
    eng.set_amplitudes(t1, t2, t3)
    eng.update_amps()
    a1, a2, a3 = eng.get_amplitudes()
    b1, b2, b3, _ = lo.update_amps(o, F, g, e, t1, t2, t3)
    assert abs(a1 - b1).max() <= 1e-6 * abs(b1).max()
    assert abs(a2 - b2).max() <= 1e-6 * abs(b2).max()
    assert abs(a3 - b3).max() <= 1e-5 * abs(b3).max()


@pytest.mark.parametrize("o,v", [(3, 5), (3, 7), (3, 14)])
def test_bf16x3_residual_plan(monkeypatch, o, v):
    """Split-BF16x3 residual plan; at nvir=5 an FP32 GEMM (K = o v <= 16) reads Y/Z, which then get a
    packed copy, at nvir=7 and 14 Y/Z are written as packed words by the transform itself; at nvir=14
    the fused T2 -> T3 GEMM (K = v + o = 17) runs on operands its operand pass packs."""
    abi_emulator.install(monkeypatch)
    from distr_cc.synthetic import SyntheticSystem
    from distr_cc.t3_engine import TriplesEngine
    s = SyntheticSystem(o, v, naux=20, seed=31)
    im, t2, t3 = s.bare_intermediates(), s.t2(), s.t3_tri()
    eng = TriplesEngine(o, v, dtype=np.float32, contract="bf16x3")
    assert eng._yz_direct == (o * v > 16)
    # every residual GEMM with K > 16 reads packed (pre-split) operands, the others stay FP32
    gemms = [g for g in eng.launches if hasattr(g, "args")]
    assert all(g.fn == ("dcc_contract_packed_bf16x3" if g.args[2] > 16 else "dcc_contract_f32") for g in gemms)
    assert any(g.fn == "dcc_contract_f32" for g in gemms) and any(g.args[2] > 16 for g in gemms)
    eng.set_t3(t3)
    eng.set_intermediates(im, t2)
    eng.residual()
    Ro = tri.compute_r3_tri(o, v, im, t2, t3)
    assert abs(eng.R.download().astype(np.float64) - Ro).max() <= 3e-4 * abs(Ro).max()


def test_bf16x3_tile_transfer(monkeypatch):
    """Tiles transferred by launch name: a 256-wide split-BF16x3 tile falls back to 128x128 on a
    launch whose extent on that side is at most 128; other tiles are applied as given."""
    abi_emulator.install(monkeypatch)
    from distr_cc.t3_engine import GemmLaunch, TriplesEngine
    eng = TriplesEngine(3, 5, dtype=np.float32, contract="bf16x3")
    gemms = [g for g in eng.launches if isinstance(g, GemmLaunch) and g.fn.endswith("bf16x3")]
    assert gemms and all(g.max_rows <= 128 and g.args[1] <= 128 for g in gemms)
    for t, want in ((4, 3), (5, 3), (2, 2)):
        eng.set_tiles_by_name({g.name: t for g in gemms})
        assert all(g.tile == want for g in gemms)


def test_batched_ensemble_plan(monkeypatch):
    """Grouped launches + graph replay == per-system engines, for an uneven ensemble."""
    abi_emulator.install(monkeypatch)
    from distr_cc.batched import BatchedTriples
    from distr_cc.device import Stream
    from distr_cc.synthetic import SyntheticSystem
    from distr_cc.t3_engine import TriplesEngine
    systems = [(2, 4), (3, 5), (1, 3), (3, 3)]
    st = Stream.__new__(Stream)
    st.handle, st._own = ctypes.c_void_p(7), False
    bt = BatchedTriples(systems, st)
    refs = []
    for n, (o, v) in enumerate(systems):
        s = SyntheticSystem(o, v, naux=15, seed=50 + n)
        im, t2, t3 = s.bare_intermediates(), s.t2(), s.t3_tri()
        bt.set_system(n, t3, im, t2, s.mo_energy)
        e = TriplesEngine(o, v)
        e.set_t3(t3)
        e.set_intermediates(im, t2)
        e.residual()
        refs.append((e.finalize(s.mo_energy), e.get_t3()))
    bt.capture()
    bt.step()
    for e, (nrm, t3ref) in zip(bt.engines, refs):
        assert abs(e.get_t3() - t3ref).max() <= 1e-13 * abs(t3ref).max()
        assert abs(e.read_norms()[0] - nrm[0]) <= 1e-12 * nrm[0]


def test_init_amps_unshifted_with_level_shift(monkeypatch):
    """The MP2 guess divides by e_i - e_a without the level shift (rccsdt.py:212-214), also when the
    iteration's denominators carry one; t1 != 0 through a non-canonical f_ov."""
    abi_emulator.install(monkeypatch)
    from ccsdt_cases import DuckMF, molecule_like, with_fov
    from distr_cc.distribute_t3 import SINGLE_RANK
    from distr_cc.rccsdt import RCCSDT, _PhysicistsERIs
    o, v = 3, 4
    F, g, e, _ = molecule_like(o, v, 31)
    F = with_fov(F, o, 0.02, 32)
    cc = RCCSDT(DuckMF(o, v, e), SINGLE_RANK)
    cc.level_shift, cc.verbose = 0.05, 0
    emp2, tamps = cc.init_amps(_PhysicistsERIs(F, e, g))
    e_ref, t1, t2 = tri.init_amps_rhf(o, F, e, g)
    assert abs(t1).max() > 1e-4
    assert abs(tamps[0] - t1).max() <= 1e-14 * abs(t1).max()
    assert abs(tamps[1] - t2).max() <= 1e-14 * abs(t2).max()
    assert abs(emp2 - e_ref) <= 1e-14 * abs(e_ref) and not tamps[2][1].any()


@pytest.mark.parametrize("with_t3,nvir_diis,diis,shift", [(True, None, True, 0.0), (True, 2, True, 0.05),
                                                         (False, None, True, 0.0), (False, 2, True, 0.05),
                                                         (True, None, False, 0.0)])
def test_kernel_diis_trajectory(monkeypatch, with_t3, nvir_diis, diis, shift):
    """RCCSDT.kernel after 4 cycles == the oracle loop after 4 cycles: unified T1/T2/T3 DIIS
    (do_diis_max_t = True, with an nvir_diis subspace), standard T1/T2 DIIS (do_diis_max_t = False)
    and no DIIS, with and without a level shift (distr_cc/rccsdt.py:1026-1148)."""
    abi_emulator.install(monkeypatch)
    from ccsdt_cases import DuckMF, molecule_like, oracle_kernel, with_fov
    from distr_cc.distribute_t3 import SINGLE_RANK
    from distr_cc.rccsdt import RCCSDT, _PhysicistsERIs
    o, v = 2, 4
    F, g, e, _ = molecule_like(o, v, 41)
    F = with_fov(F, o, 0.01, 42)
    cc = RCCSDT(DuckMF(o, v, e), SINGLE_RANK)
    cc.do_diis_max_t, cc.nvir_diis, cc.diis, cc.level_shift = with_t3, nvir_diis, diis, shift
    cc.max_cycle, cc.verbose = 4, 0
    e_cc, tamps = cc.kernel(eris=_PhysicistsERIs(F, e, g))
    t1, t2, t3, energies = oracle_kernel(o, F, g, e, 4, level_shift=shift, nvir_diis=nvir_diis, diis=diis,
                                         with_t3=with_t3)
    assert cc.cycles == 4
    assert abs(e_cc - energies[-1]) <= 1e-13
    for a, b in ((tamps[0], t1), (tamps[1], t2), (tamps[2][1], t3)):
        assert abs(a - b).max() <= 1e-11 * abs(b).max()


def test_engine_cache_follows_attributes(monkeypatch):
    """A second kernel() on the same eris picks up changed level_shift / nvir_diis / triples_precision."""
    abi_emulator.install(monkeypatch)
    from ccsdt_cases import DuckMF, molecule_like
    from distr_cc.distribute_t3 import SINGLE_RANK
    from distr_cc.rccsdt import RCCSDT, _PhysicistsERIs
    o, v = 2, 3
    F, g, e, _ = molecule_like(o, v, 43)
    eris = _PhysicistsERIs(F, e, g)
    cc = RCCSDT(DuckMF(o, v, e), SINGLE_RANK)
    cc.max_cycle, cc.verbose = 2, 0
    cc.kernel(eris=eris)
    eng = cc.engine
    cc.level_shift, cc.nvir_diis = 0.1, 2
    cc.kernel(eris=eris)
    assert cc.engine is eng and eng.level_shift == 0.1 and eng.nvir_diis == 2
    ref = e[:o, None] - e[None, o:] - 0.1
    assert abs(eng.eia.download() - ref).max() < 1e-15
    cc.triples_precision = "fp32"
    cc.kernel(eris=eris)
    assert cc.engine is not eng and cc.engine.mixed


def test_batched_rccsdt_plan(monkeypatch):
    """Whole RCCSDT iterations of an uneven ensemble (grouped triples launches, one graph per step) ==
    each system's own RCCSDTEngine.update_amps, energies and amplitudes after two steps."""
    abi_emulator.install(monkeypatch)
    from ccsdt_cases import molecule_like, random_amplitudes
    from distr_cc.batched import BatchedRCCSDT
    from distr_cc.device import Stream
    from distr_cc.rccsdt_engine import RCCSDTEngine
    shapes = [(2, 3, 81), (3, 4, 82), (1, 4, 83), (3, 3, 84)]
    st = Stream.__new__(Stream)
    st.handle, st._own = ctypes.c_void_p(7), False
    syss = [(o, v) + molecule_like(o, v, seed)[:3] for o, v, seed in shapes]
    amps = [random_amplitudes(o, v, seed + 1) for o, v, seed in shapes]
    bt = BatchedRCCSDT(syss, st)
    # the lower-order contractions of the four systems run grouped, operation by operation
    assert len(bt.lower_phases) >= 50 and all(len(ph.launches) == 4 for ph in bt.lower_phases)
    for n, a in enumerate(amps):
        bt.set_amplitudes(n, *a)
    bt.capture()
    for _ in range(2):
        bt.step()
        got = bt.energies()
    for n, ((o, v, F, g, e), a) in enumerate(zip(syss, amps)):
        ref = RCCSDTEngine(o, v, F, g, e)
        ref.set_amplitudes(*a)
        for _ in range(2):
            ref.update_amps()
        assert abs(got[n] - ref.energy()) <= 1e-14
        for x, y in zip(bt.engines[n].get_amplitudes(), ref.get_amplitudes()):
            assert abs(x - y).max() <= 1e-13 * max(abs(y).max(), 1e-30)
