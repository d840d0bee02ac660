"""The triples update at the full sizes of the measured workloads, against the oracle.

* nocc=10, nvir=90 (220 canonical blocks of 90^3, FP64): the residual on a sample of target blocks
  (every tie pattern i=j, j=k, i=j=k and distinct indices) against oracle.compute_r3_tri(targets=...);
  the fused projection/denominator/update and both norms on ALL blocks against oracle.t3_update
  applied to the device residual.
* nocc=16, nvir=160 (816 blocks of 160^3): the FP64 residual against the oracle on two target
  blocks; the FP32 and split-BF16x3 residuals against the FP64 device residual on all blocks (relative
  Frobenius error below 2e-6 for FP32 and 2e-5 for split-BF16x3, max error below 2e-5 and 1e-4 of
  max |R|: FP32 accumulation over K up to nvir^2 = 25600 with products rounded to 2^-24, or ~2^-16 per
  split-BF16x3 product); the split-BF16x3 update leaves no totally abc-symmetric part in any block.
* 96 systems (nocc 3..8, nvir 20..64, seeds 20261202+s) updated by one replay of the captured HIP
  graph of grouped persistent launches: the smallest, a middle and the largest system against the
  oracle (residual rows, and the update of every block from the device residual); and the same 96
  systems iterated as whole RCCSDT iterations (two graph replays) against every system's own ungrouped
  iterations and, for the smallest and a middle system, two oracle update_amps iterations.
Inputs are the seeded synthetic generator (B-factor sigma 0.05, naux 300, T3 noise 1e-4 / D).
"""
import gc
import sys

import numpy as np
import pytest

from oracle import rccsdt_tri as tri
from oracle.t3_layout import ijk_to_linear

pytestmark = pytest.mark.gpu
SEED = 20261202


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def dev():
    from distr_cc import device
    arch = device.init(0)
    _say("device arch %s" % arch)
    return device


def _system(o, v, seed=SEED):
    from distr_cc.synthetic import SyntheticSystem
    s = SyntheticSystem(o, v, naux=300, seed=seed)
    return s, s.bare_intermediates(), s.t2(), s.t3_tri()


def _rel_rows(a, b, rows):
    return float(np.abs(a[rows] - b[rows]).max() / np.abs(b[rows]).max())


def _check_update_all_blocks(o, R, t3_old, t3_new, mo_energy, norms, nvir_diis=None, tol=1e-13):
    """Device finalize against oracle.t3_update applied to the device residual R, on every block."""
    t3o, r3o, n2, na = tri.t3_update(o, R, t3_old, mo_energy, nvir_diis=nvir_diis)
    scale = np.abs(r3o).max()
    assert np.abs((t3_new - t3_old) - r3o).max() <= tol * scale
    assert abs(norms[0] - n2) <= 1e-12 * n2 and abs(norms[1] - na) <= 1e-12 * na
    return r3o


def test_c2_fp64_residual_and_update_vs_oracle(dev):
    """nocc=10, nvir=90 in FP64 (the plain-FP64 residual workload)."""
    from distr_cc.t3_engine import TriplesEngine
    o, v = 10, 90
    _say("C2 inputs nocc=%d nvir=%d" % (o, v))
    s, im, t2, t3 = _system(o, v)
    eng = TriplesEngine(o, v)
    eng.set_t3(t3)
    eng.set_intermediates(im, t2)
    eng.residual()
    R = eng.R.download()
    norms = eng.finalize(s.mo_energy, nvir_diis=60)
    t3_new = eng.get_t3()
    del eng
    gc.collect()
    S = [ijk_to_linear(*x, o) for x in ((0, 0, 0), (0, 0, 5), (2, 7, 7), (1, 4, 8), (9, 9, 9), (3, 6, 9))]
    _say("C2 oracle on target blocks %s" % S)
    Ro = tri.compute_r3_tri(o, v, im, t2, t3, targets=S)
    err = _rel_rows(R, Ro, S)
    _say("C2 residual rel err on sampled blocks %.2e" % err)
    assert err <= 1e-12
    _check_update_all_blocks(o, R, t3, t3_new, s.mo_energy, norms, nvir_diis=60)


def _residual(o, v, s, im, t2, t3, dtype, contract, finalize=False):
    from distr_cc.device import DeviceArray
    from distr_cc.t3_engine import TriplesEngine
    eng = TriplesEngine(o, v, dtype=dtype, contract=contract)
    eng.set_t3(t3)
    eng.set_intermediates(im, t2)
    eng.residual()
    R = eng.R.download()
    out = None
    if finalize:
        r3 = DeviceArray(eng.R.shape, eng.dtype)
        eng.set_denominators(s.mo_energy)
        eng.finalize_launch(r3_out=r3)
        out = (r3.download(), eng.read_norms())
        r3.free()
    del eng
    gc.collect()
    return R, out


def _frob_and_max(X, Y, chunk=64):
    """(|X - Y|_F / |Y|_F, max |X - Y| / max |Y|) over block chunks, in FP64."""
    num = den = mx = my = 0.0
    for b in range(0, len(Y), chunk):
        y = Y[b:b + chunk]
        d = X[b:b + chunk].astype(np.float64) - y
        num += float(np.vdot(d, d))
        den += float(np.vdot(y, y))
        mx = max(mx, float(np.abs(d).max()))
        my = max(my, float(np.abs(y).max()))
    return (num / den) ** 0.5, mx / my


@pytest.fixture(scope="module")
def c3(dev):
    """Inputs and the FP64 device residual at nocc=16, nvir=160 (shared by the three tests below)."""
    o, v = 16, 160
    _say("C3 inputs nocc=%d nvir=%d" % (o, v))
    s, im, t2, t3 = _system(o, v)
    _say("C3 FP64 residual")
    R64, _ = _residual(o, v, s, im, t2, t3, np.float64, "native")
    yield o, v, s, im, t2, t3, R64


def test_c3_fp64_residual_vs_oracle(c3):
    """nocc=16, nvir=160: the FP64 residual against the oracle on two target blocks (distinct indices
    and i=j)."""
    o, v, s, im, t2, t3, R64 = c3
    S = [ijk_to_linear(*x, o) for x in ((3, 7, 11), (5, 5, 12))]
    _say("C3 oracle on target blocks %s" % S)
    Ro = tri.compute_r3_tri(o, v, im, t2, t3, targets=S)
    err = _rel_rows(R64, Ro, S)
    _say("C3 FP64 residual rel err on sampled blocks %.2e" % err)
    assert err <= 1e-12


# FP32 products round to 2^-24; a split-BF16x3 product drops lo*lo and keeps 8 bits of each lo part,
# ~2^-16 relative per product, accumulated in FP32 over K up to nvir^2 = 25600
C3_BOUNDS = {"fp32": (2e-6, 2e-5), "bf16x3": (2e-5, 1e-4)}


def test_c3_fp32_residual_vs_fp64(c3):
    """nocc=16, nvir=160: the FP32 residual (FP32 MFMA) against the FP64 device residual, all blocks."""
    o, v, s, im, t2, t3, R64 = c3
    R, _ = _residual(o, v, s, im, t2, t3, np.float32, "native")
    frob, mx = _frob_and_max(R, R64)
    _say("C3 fp32 vs FP64: rel Frobenius %.2e, max %.2e" % (frob, mx))
    assert frob <= C3_BOUNDS["fp32"][0] and mx <= C3_BOUNDS["fp32"][1]


def test_c3_bf16x3_residual_and_update(c3):
    """nocc=16, nvir=160: the split-BF16x3 residual against the FP64 device residual on all blocks; its
    update r3 = P3_full(R)/D leaves no totally abc-symmetric part (checked on every 17th block) and its
    norm is the one the finalize kernel reports."""
    o, v, s, im, t2, t3, R64 = c3
    R, (r3, norms) = _residual(o, v, s, im, t2, t3, np.float32, "bf16x3", finalize=True)
    frob, mx = _frob_and_max(R, R64)
    _say("C3 bf16x3 vs FP64: rel Frobenius %.2e, max %.2e" % (frob, mx))
    assert frob <= C3_BOUNDS["bf16x3"][0] and mx <= C3_BOUNDS["bf16x3"][1]
    del R
    x = r3[::17].astype(np.float64)
    sym = (x + x.transpose(0, 1, 3, 2) + x.transpose(0, 2, 1, 3) + x.transpose(0, 2, 3, 1)
           + x.transpose(0, 3, 1, 2) + x.transpose(0, 3, 2, 1)) / 6.0
    worst = float(np.abs(sym).max() / np.abs(x).max())
    _say("C3 bf16x3 update: max symmetric part %.2e of max |r|; |r|^2 %.6e" % (worst, norms[0]))
    assert worst <= 1e-6
    n2 = sum(float(np.vdot(r3[b:b + 64].astype(np.float64), r3[b:b + 64].astype(np.float64)))
             for b in range(0, len(r3), 64))
    assert abs(norms[0] - n2) <= 1e-5 * n2


def test_c4_batched_ensemble_vs_oracle(dev):
    """96 systems in one captured graph of grouped launches: smallest, middle and largest system
    against the oracle."""
    from distr_cc.batched import BatchedTriples
    from distr_cc.device import Stream
    systems = []
    for k in range(96):
        r = np.random.default_rng(SEED + k)
        systems.append((int(r.integers(3, 9)), int(r.integers(20, 65))))
    _say("C4 building 96 systems")
    st = Stream()
    bt = BatchedTriples(systems, st)
    inputs = []
    for k, (o, v) in enumerate(systems):
        s, im, t2, t3 = _system(o, v, SEED + k)
        inputs.append((s, im, t2, t3))
        bt.set_system(k, t3, im, t2, s.mo_energy)
    bt.capture()
    bt.step()
    norms = bt.norms()
    cost = [o * (o + 1) * (o + 2) * v ** 5 for (o, v) in systems]
    order = np.argsort(cost)
    for k in (int(order[0]), int(order[len(order) // 2]), int(order[-1])):
        o, v = systems[k]
        s, im, t2, t3 = inputs[k]
        e = bt.engines[k]
        R, t3_new = e.R.download(), e.get_t3()
        n3 = len(t3)
        S = sorted({0, n3 // 3, n3 // 2, n3 - 1})
        _say("C4 system %d nocc=%d nvir=%d oracle on blocks %s" % (k, o, v, S))
        Ro = tri.compute_r3_tri(o, v, im, t2, t3, targets=S)
        err = _rel_rows(R, Ro, S)
        assert err <= 1e-12, (k, err)
        _check_update_all_blocks(o, R, t3, t3_new, s.mo_energy, norms[k])
    bt.close()


def test_c4_batched_rccsdt_iterations(dev):
    """96 systems iterated together (full RCCSDT iterations, grouped triples launches, one HIP graph per
    step): after two steps every system's energy and amplitudes equal its own ungrouped iterations, and
    the smallest and a middle system equal the oracle's two update_amps iterations."""
    from distr_cc.batched import BatchedRCCSDT
    from distr_cc.device import Stream
    from distr_cc.synthetic import SyntheticSystem
    from oracle import rccsdt_lower as lo
    syss, amps = [], []
    for k in range(96):
        r = np.random.default_rng(SEED + k)
        o, v = int(r.integers(3, 9)), int(r.integers(20, 65))
        s = SyntheticSystem(o, v, naux=300, seed=SEED + k)
        syss.append((o, v, s.fock, s.pppp(), s.mo_energy))
        amps.append((np.zeros((o, v)), s.t2(), s.t3_tri()))
    _say("C4 full iterations: building 96 engines")
    st = Stream()
    bt = BatchedRCCSDT(syss, st)
    for k, a in enumerate(amps):
        bt.set_amplitudes(k, *a)
    bt.capture()
    for step in range(2):
        bt.step()
        e_batched = bt.energies()
        _say("C4 step %d: E_corr of system 0 %.12f" % (step, e_batched[0]))
    got = [e.get_amplitudes() for e in bt.engines]
    worst = 0.0
    for k, e in enumerate(bt.engines):
        e.set_amplitudes(*amps[k])
        for _ in range(2):
            e.update_amps()
        assert abs(e.energy() - e_batched[k]) <= 1e-12 * max(1.0, abs(e_batched[k])), k
        for x, y in zip(got[k], e.get_amplitudes()):
            worst = max(worst, float(np.abs(x - y).max() / max(np.abs(y).max(), 1e-300)))
    _say("C4 grouped vs ungrouped iterations: worst amplitude rel diff %.2e" % worst)
    assert worst <= 1e-11
    cost = [o * (o + 1) * (o + 2) * v ** 5 for (o, v, _, _, _) in syss]
    order = np.argsort(cost)
    for k in (int(order[0]), int(order[len(order) // 2])):
        o, v, F, g, e = syss[k]
        _say("C4 system %d nocc=%d nvir=%d: two oracle iterations" % (k, o, v))
        t1, t2, t3 = amps[k]
        for _ in range(2):
            t1, t2, t3, _res = lo.update_amps(o, F, g, e, t1, t2, t3)
        assert abs(tri.energy_rhf(o, F, g, t1, t2) - e_batched[k]) <= 1e-10 * max(1.0, abs(e_batched[k]))
        for x, y in zip(got[k], (t1, t2, t3)):
            assert np.abs(x - y).max() <= 1e-10 * np.abs(y).max()
    bt.close()
