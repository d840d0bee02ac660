"""GPU parity of the libdcc kernels against NumPy / the oracle (FP64)."""
import sys

import numpy as np
import pytest

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


@pytest.mark.parametrize("M,N,K", [(16, 16, 4), (37, 53, 29), (64, 64, 16), (130, 70, 300)])
def test_contract_plain_gemm(dev, M, N, K):
    """C = 0.5*C + 1.5*A@B through dcc_contract_f64; pins the MFMA lane layout and masking."""
    from distr_cc import _lib as L
    from distr_cc.device import DeviceArray
    _say("contract plain %dx%dx%d" % (M, N, K))
    rng = np.random.default_rng(M + N + K)
    A = rng.normal(size=(M, K))
    B = rng.normal(size=(K, N))
    C0 = rng.normal(size=(M, N))
    dA, dB, dC = DeviceArray.from_host(A), DeviceArray.from_host(B), DeviceArray.from_host(C0)
    bt = np.zeros(1, dtype=L.BATCH_DTYPE)
    bt[0] = (0, 0, 0, 0, M)
    dbt = DeviceArray.from_host(bt)
    lin = lambda s: L.IdxMap(None, 0, 1, 1 << 20, 0, 0, s)
    ref = 0.5 * C0 + 1.5 * A @ B
    for tile in range(-1, 8):   # default, then the four shapes x two prefetch depths
        dC.upload(C0)
        if tile < 0:
            L.call("dcc_contract_f64", M, N, K, 1.5, 0.5, dA.vp, lin(K), lin(1), dB.vp, lin(N), lin(1),
                   dC.vp, lin(N), lin(1), dbt.vp, 1, M, None)
        else:
            L.call("dcc_contract_tiled_f64", M, N, K, 1.5, 0.5, dA.vp, lin(K), lin(1), dB.vp, lin(N), lin(1),
                   dC.vp, lin(N), lin(1), dbt.vp, 1, M, tile, None)
        C = dC.download()
        assert abs(C - ref).max() < 1e-12 * max(1.0, abs(ref).max()), tile


@pytest.mark.parametrize("M,N,K", [(75, 41, 33), (76, 42, 34), (130, 138, 64)])
def test_contract_transposed_operands(dev, M, N, K):
    """A stored K-major and B stored N-major-transposed exercise the other two loaders (odd sizes:
    per-element loads; even sizes: pair vector loads)."""
    from distr_cc import _lib as L
    from distr_cc.device import DeviceArray
    _say("contract transposed %dx%dx%d" % (M, N, K))
    rng = np.random.default_rng(5)
    At = rng.normal(size=(K, M))  # A[m,k] = At[k,m]
    Bt = rng.normal(size=(N, K))  # B[k,n] = Bt[n,k]
    dA, dB = DeviceArray.from_host(At), DeviceArray.from_host(Bt)
    dC = DeviceArray.zeros((M, N))
    bt = np.zeros(1, dtype=L.BATCH_DTYPE)
    bt[0] = (0, 0, 0, 0, M)
    dbt = DeviceArray.from_host(bt)
    lin = lambda s: L.IdxMap(None, 0, 1, 1 << 20, 0, 0, s)
    ref = At.T @ Bt.T
    for tile in range(-1, 8):   # default, then the four shapes x two prefetch depths
        dC.fill(0.0)
        if tile < 0:
            L.call("dcc_contract_f64", M, N, K, 1.0, 0.0, dA.vp, lin(1), lin(M), dB.vp, lin(1), lin(K),
                   dC.vp, lin(N), lin(1), dbt.vp, 1, M, None)
        else:
            L.call("dcc_contract_tiled_f64", M, N, K, 1.0, 0.0, dA.vp, lin(1), lin(M), dB.vp, lin(1), lin(K),
                   dC.vp, lin(N), lin(1), dbt.vp, 1, M, tile, None)
        assert abs(dC.download() - ref).max() < 1e-12 * abs(ref).max(), tile


@pytest.mark.parametrize("dt", ["f64", "f32"])
@pytest.mark.parametrize("v", [1, 5, 8, 13, 20, 33, 48, 50])
def test_s3_transform_all_perms(dev, v, dt):
    """dcc_s3_transform_*: every single permutation and a mixed combination, FP64 and FP32 storage
    (partial 16^3 tiles and several tile orbits)."""
    from distr_cc import _lib as L
    from distr_cc.device import DeviceArray
    from distr_cc.t3_engine import s3_jobs_array
    _say("s3 transform v=%d %s" % (v, dt))
    rng = np.random.default_rng(v)
    typ = np.float64 if dt == "f64" else np.float32
    X = rng.normal(size=(3, v, v, v)).astype(typ)
    perms = ((0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0))
    coefs = [np.eye(6)[p] for p in range(6)] + [np.array([2.0, -1.0, -1.0, 0.5, 0.25, -3.0])]
    jobs = []
    for n, c in enumerate(coefs):
        jobs.append((n * v ** 3, (n % 3) * v ** 3, c))
    dX = DeviceArray.from_host(X)
    dO = DeviceArray.zeros((len(coefs), v, v, v), typ)
    dj = DeviceArray.from_host(s3_jobs_array(jobs))
    L.call("dcc_s3_transform_" + dt, dO.vp, dX.vp, v, 0.0, dj.vp, len(jobs), None)
    out = dO.download().astype(np.float64)
    for n, c in enumerate(coefs):
        src = X[n % 3].astype(np.float64)
        ref = sum(ci * src.transpose(p) for ci, p in zip(c, perms))
        # FP64 exact up to the sum's rounding; FP32: one rounding of the FP64 sum to float
        assert abs(out[n] - ref).max() < (1e-13 if dt == "f64" else 1e-6 * max(1.0, abs(ref).max())), (n, v)


@pytest.mark.parametrize("v", [5, 8, 20, 50])
def test_s3_transform_pack_f32(dev, v):
    """dcc_s3_transform_pack_f32 writes exactly the words of dcc_s3_transform_f32 followed by
    dcc_pack_bf16x3."""
    from distr_cc import _lib as L
    from distr_cc.device import DeviceArray
    from distr_cc.t3_engine import s3_jobs_array
    _say("s3 transform pack v=%d" % v)
    rng = np.random.default_rng(70 + v)
    X = rng.normal(size=(2, v, v, v)).astype(np.float32)
    jobs = [(0, v ** 3, np.array([2.0, -1.0, 0.0, 0.0, 0.0, -1.0])), (v ** 3, 0, np.eye(6)[3])]
    dX, dj = DeviceArray.from_host(X), DeviceArray.from_host(s3_jobs_array(jobs))
    ref, got = DeviceArray.zeros((2, v, v, v), np.float32), DeviceArray((2, v, v, v), np.uint32)
    L.call("dcc_s3_transform_f32", ref.vp, dX.vp, v, 0.0, dj.vp, len(jobs), None)
    L.call("dcc_pack_bf16x3", ref.vp, ref.vp, ref.size, None)
    L.call("dcc_s3_transform_pack_f32", got.vp, dX.vp, v, dj.vp, len(jobs), None)
    assert np.array_equal(got.download(), ref.download().view(np.uint32))


@pytest.mark.parametrize("shift", [0, 1])
@pytest.mark.parametrize("dt", ["f64", "f32"])
@pytest.mark.parametrize("v", [1, 5, 8, 13, 20, 33, 48, 50])
def test_s3_gather_multi_source(dev, v, dt, shift):
    """dcc_s3_gather_*: a different source block per permutation, repeated sources, zero
    coefficients and beta accumulation, against NumPy; ``shift`` = 1 moves every block one element
    off alignment."""
    from distr_cc import _lib as L
    from distr_cc.device import DeviceArray
    _say("s3 gather v=%d %s shift %d" % (v, dt, shift))
    rng = np.random.default_rng(40 + v)
    typ = np.float64 if dt == "f64" else np.float32
    X = rng.normal(size=(5, v, v, v)).astype(typ)
    O0 = rng.normal(size=(3, v, v, v)).astype(typ)
    perms = ((0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0))
    v3 = v ** 3
    jobs = [(0, [0, 1, 2, 3, 4, 0], np.array([1.0, -1.0, 0.5, 2.0, 0.0, 3.0])),
            (2 * v3, [4, 4, 4, 4, 4, 4], np.ones(6)),
            (1 * v3, [2, 0, 0, 0, 0, 1], np.array([1.5, 0.0, 0.0, 0.0, 0.0, -2.0]))]
    arr = np.zeros(len(jobs), dtype=L.S3GATHER_DTYPE)
    for n, (dst, srcs, coef) in enumerate(jobs):
        arr[n]['dst'], arr[n]['src'], arr[n]['coef'] = dst + shift, [x * v3 + shift for x in srcs], coef
    pad = np.zeros(shift, typ)
    dX = DeviceArray.from_host(np.concatenate([pad, X.ravel()]))
    dO = DeviceArray.from_host(np.concatenate([pad, O0.ravel()]))
    dj = DeviceArray.from_host(arr)
    L.call("dcc_s3_gather_" + dt, dO.vp, dX.vp, v, 0.5, dj.vp, len(jobs), None)
    outf = dO.download()
    assert np.all(outf[:shift] == 0)
    out = outf[shift:].reshape(O0.shape).astype(np.float64)
    tol = 1e-13 if dt == "f64" else 1e-6
    for dst, srcs, coef in jobs:
        q = dst // v3
        ref = 0.5 * O0[q].astype(np.float64) + sum(c * X[sx].astype(np.float64).transpose(p)
                                                     for c, sx, p in zip(coef, srcs, perms) if c != 0.0)
        assert abs(out[q] - ref).max() < tol * max(1.0, abs(ref).max()), (q, v)


@pytest.mark.parametrize("prec", ["fp64", "bf16x3"])
def test_residual_chunked_vs_oracle(dev, prec):
    """Residual with the memory-bounded paths of large systems at nocc=10, nvir=9 (220 blocks): ring
    operator in chunks of 26 of the 55 pairs, W_oooo/F_oo targets in 64-row chunks (a bound of 52
    rounded up) and a partial one, terms 1..5 as one stacked GEMM; split-BF16x3 with Y/Z written as
    packed words."""
    from distr_cc.synthetic import SyntheticSystem
    from distr_cc.t3_engine import TriplesEngine
    from oracle import rccsdt_tri as tri
    o, v = 10, 9
    _say("residual chunked o=%d v=%d %s" % (o, v, prec))
    s = SyntheticSystem(o, v, naux=30, seed=77)
    rng = np.random.default_rng(78)
    im = {k: x + 1e-3 * rng.normal(size=x.shape) for k, x in s.bare_intermediates().items()}
    t2, t3 = s.t2(), s.t3_tri()
    kw = dict(dtype=np.float32, contract="bf16x3") if prec == "bf16x3" else {}
    eng = TriplesEngine(o, v, ring_g_bytes=52 * 5 * v ** 3 * (4 if prec == "bf16x3" else 8), **kw)
    assert eng.Xp.shape[1] == 64 and eng.Gs.shape[0] == 26 and eng.npairs == 55
    assert eng._yz_direct == (prec == "bf16x3")
    eng.set_t3(t3)
    eng.set_intermediates(im, t2)
    eng.residual()
    R = eng.R.download().astype(np.float64)
    Ro = tri.compute_r3_tri(o, v, im, t2, t3, batch_size=13)
    tol = 1e-12 if prec == "fp64" else 3e-4
    assert abs(R - Ro).max() <= tol * abs(Ro).max(), abs(R - Ro).max() / abs(Ro).max()


@pytest.mark.parametrize("o,v", [(1, 3), (3, 4), (4, 9), (5, 17), (3, 24)])
def test_residual_and_update_vs_oracle(dev, o, v):
    from distr_cc.synthetic import SyntheticSystem
    from distr_cc.t3_engine import TriplesEngine
    from oracle import rccsdt_tri as tri
    _say("residual o=%d v=%d" % (o, v))
    s = SyntheticSystem(o, v, naux=30, seed=100 + o * v)
    im = s.bare_intermediates()
    rng = np.random.default_rng(o * 31 + v)
    im = {k: x + 1e-3 * rng.normal(size=x.shape) for k, x in im.items()}
    t2, t3 = s.t2(), s.t3_tri()
    eng = TriplesEngine(o, v)
    eng.set_t3(t3)
    eng.set_intermediates(im, t2)
    eng.residual()
    R = eng.R.download()
    Ro = tri.compute_r3_tri(o, v, im, t2, t3, batch_size=5)
    scale = abs(Ro).max()
    assert abs(R - Ro).max() <= 1e-12 * scale, abs(R - Ro).max() / scale
    nd = max(1, v - 2)
    n2, na = eng.finalize(s.mo_energy, level_shift=0.05, nvir_diis=nd)
    t3_new = eng.get_t3()
    t3o, r3o, n2o, nao = tri.t3_update(o, Ro, t3, s.mo_energy, level_shift=0.05, nvir_diis=nd)
    assert abs(t3_new - t3o).max() <= 1e-13 * abs(t3o).max()
    # the projection cancels most of R when o is small (o = 1: n2 ~ 1e-10 of sum R^2), so the norms
    # also carry an absolute round-off floor relative to sum R^2
    floor = 1e-20 * float((Ro ** 2).sum()) + 1e-30
    assert abs(n2 - n2o) <= 1e-12 * n2o + floor and abs(na - nao) <= 1e-12 * nao + floor


@pytest.mark.parametrize("o,v", [(1, 3), (3, 7), (4, 12), (5, 19)])
def test_t3_to_r1r2_and_w_vs_oracle(dev, o, v):
    from distr_cc.synthetic import SyntheticSystem
    from distr_cc.t3_engine import TriplesEngine
    from oracle import rccsdt_tri as tri
    _say("t3 -> r1/r2/W o=%d v=%d" % (o, v))
    s = SyntheticSystem(o, v, naux=30, seed=300 + o * v)
    im, t2, t3 = s.bare_intermediates(), s.t2(), s.t3_tri()
    nmo = o + v
    rng = np.random.default_rng(o + v)
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
    # the residual sees the T3-adjusted W_vooo / W_vvvo (rccsdt.py:935-941)
    eng.residual()
    im2 = dict(im, W_vooo=Wvo, W_vvvo=Wvvo)
    Ro = tri.compute_r3_tri(o, v, im2, t2, t3)
    assert abs(eng.R.download() - Ro).max() <= 1e-12 * abs(Ro).max()


@pytest.mark.parametrize("o,v,bs,form", [(2, 5, 2, "dense"), (3, 7, 4, "dense"), (5, 9, 4, "dense"),
                                         (4, 11, 3, "dense"), (3, 7, 4, "task+ws"), (5, 9, 4, "task+ws"),
                                         (4, 11, 3, "ws"), (10, 5, 3, "task+ws"), (12, 4, 2, "dense")])
def test_q_energy_vs_oracle(dev, o, v, bs, form):
    """(Q)/[Q] against the oracle.  form: "dense" = dense W_vvvvoo (and the LDS finalize up to nocc 9,
    the workspace finalize beyond); "task" = W_vvvvoo slots per chunk of two tasks; "ws" = the workspace
    finalize at any nocc."""
    from distr_cc.device import DeviceArray
    from distr_cc.q_engine import QEngine
    from distr_cc.synthetic import SyntheticSystem
    from oracle import q_energy as Q
    from oracle import t3_layout as TL
    _say("(Q) o=%d v=%d blksize=%d %s" % (o, v, bs, form))
    s = SyntheticSystem(o, v, naux=30, seed=500 + o * v, sigma_t3=1e-3)
    pppp, t2, t3 = s.pppp(), s.t2(), s.t3_tri()
    ref = Q.kernel(o, v, pppp, t2, TL.full_t3_from_tri(t3, o), s.mo_energy, blksize=bs)
    opts = {}
    if "task" in form:
        opts.update(w_mode="task", max_chunk_bytes=2 * bs ** 4 * o ** 4 * 8)
    if "ws" in form:
        opts["finalize_ws"] = True
    got = QEngine(o, v, blksize=bs, **opts).kernel(pppp, t2, DeviceArray.from_host(t3), s.mo_energy)
    for g, r in zip(got, ref):
        assert abs(g - r) <= 1e-12 * abs(r)
    if form != "dense":
        return
    # job splitting is additive (tests/test_rccsdt_q.py:70-94)
    parts = [QEngine(o, v, blksize=bs).kernel(pppp, t2, DeviceArray.from_host(t3), s.mo_energy, job_idx=j,
                                              n_jobs=3) for j in range(3)]
    assert abs(sum(p[1] for p in parts) - got[1]) <= 1e-12 * abs(got[1])


@pytest.mark.parametrize("M,N,K", [(16, 16, 4), (37, 53, 29), (130, 70, 300)])
def test_contract_f32_plain_gemm(dev, M, N, K):
    """FP32 MFMA path; pins the v_mfma_f32_16x16x4f32 accumulator layout and masking."""
    from distr_cc import _lib as L
    from distr_cc.device import DeviceArray
    _say("contract f32 %dx%dx%d" % (M, N, K))
    rng = np.random.default_rng(M * N + K)
    A = rng.normal(size=(M, K)).astype(np.float32)
    B = rng.normal(size=(K, N)).astype(np.float32)
    C0 = rng.normal(size=(M, N)).astype(np.float32)
    dA, dB, dC = DeviceArray.from_host(A), DeviceArray.from_host(B), DeviceArray.from_host(C0)
    bt = np.zeros(1, dtype=L.BATCH_DTYPE)
    bt[0] = (0, 0, 0, 0, M)
    dbt = DeviceArray.from_host(bt)
    lin = lambda s: L.IdxMap(None, 0, 1, 1 << 20, 0, 0, s)
    L.call("dcc_contract_f32", M, N, K, 1.5, 0.5, dA.vp, lin(K), lin(1), dB.vp, lin(N), lin(1),
           dC.vp, lin(N), lin(1), dbt.vp, 1, M, None)
    ref = 0.5 * C0.astype(np.float64) + 1.5 * (A.astype(np.float64) @ B.astype(np.float64))
    err = abs(dC.download().astype(np.float64) - ref).max()
    assert err < 3e-6 * max(1.0, abs(ref).max()) * np.sqrt(K), err


@pytest.mark.parametrize("tile", [8, 9, 10, 11, 12, 13])
@pytest.mark.parametrize("c_mt", [False, True])
@pytest.mark.parametrize("pad", [8, 6])
def test_contract_bf16x3_dma_tiles(dev, tile, c_mt, pad):
    """Tiles 8..13 (operands staged by direct global -> LDS requests, two or three stages in flight; 12
    and 13 on four waves with AGPR accumulators) on packed, K-contiguous A and B with K a multiple of
    32 x stages, partial M/N tiles, both C layouts; batch offsets of 8 words take the 16-byte requests,
    6 words the 4-byte ones.  Checked against the same split evaluated exactly on the host, and against
    the register-staged kernel of the same tile."""
    from abi_emulator import bf16_split
    from distr_cc import _lib as L
    from distr_cc.device import DeviceArray
    M, N, K = 300, 260, 192
    _say("contract bf16x3 direct-to-LDS tile %d c_mt %d pad %d" % (tile, c_mt, pad))
    rng = np.random.default_rng(70 + tile)
    A = rng.normal(size=(M, K)).astype(np.float32)
    B = rng.normal(size=(K, N)).astype(np.float32)
    C0 = rng.normal(size=(M, N)).astype(np.float32)
    lin = lambda s: L.IdxMap(None, 0, 1, 1 << 20, 0, 0, s)
    hA = np.concatenate([np.zeros(pad, np.float32), A.ravel()])
    hB = np.concatenate([np.zeros(pad, np.float32), B.T.ravel()])        # K-contiguous B
    hC = np.concatenate([np.zeros(pad, np.float32), (C0.T if c_mt else C0).ravel()])
    am, ak, bk, bn = lin(K), lin(1), lin(1), lin(K)
    cm, cn = (lin(1), lin(M)) if c_mt else (lin(N), lin(1))
    bt = np.zeros(1, dtype=L.BATCH_DTYPE)
    bt[0] = (pad, pad, pad, 0, M)
    dbt = DeviceArray.from_host(bt)
    dA, dB = DeviceArray.from_host(hA), DeviceArray.from_host(hB)
    pA, pB = DeviceArray(hA.shape, np.uint32), DeviceArray(hB.shape, np.uint32)
    L.call("dcc_pack_bf16x3", dA.vp, pA.vp, hA.size, None)
    L.call("dcc_pack_bf16x3", dB.vp, pB.vp, hB.size, None)
    outs = []
    for t in (tile, {8: 6, 9: 4, 10: 5, 11: 7, 12: 6, 13: 7}[tile]):
        dC = DeviceArray.from_host(hC)
        L.call("dcc_contract_packed_bf16x3", M, N, K, 1.5, 0.5, pA.vp, am, ak, pB.vp, bk, bn, dC.vp, cm, cn,
               dbt.vp, 1, M, t, None)
        out = dC.download()
        assert np.all(out[:pad] == 0)
        out = out[pad:].astype(np.float64)
        outs.append(out.reshape(N, M).T if c_mt else out.reshape(M, N))
    ah, al = bf16_split(A)
    bh, bl = bf16_split(B)
    split_ref = 0.5 * C0.astype(np.float64) + 1.5 * (ah @ bh + ah @ bl + al @ bh)
    scale = max(1.0, abs(split_ref).max())
    assert abs(outs[0] - split_ref).max() < 3e-6 * scale * np.sqrt(K)
    assert abs(outs[0] - outs[1]).max() < 3e-6 * scale * np.sqrt(K)


@pytest.mark.parametrize("o,v", [(3, 9), (5, 17)])
def test_residual_fp32_vs_oracle(dev, o, v):
    from distr_cc.synthetic import SyntheticSystem
    from distr_cc.t3_engine import TriplesEngine
    from oracle import rccsdt_tri as tri
    _say("residual fp32 o=%d v=%d" % (o, v))
    s = SyntheticSystem(o, v, naux=30, seed=700 + o * v)
    im, t2, t3 = s.bare_intermediates(), s.t2(), s.t3_tri()
    eng = TriplesEngine(o, v, dtype=np.float32)
    eng.set_t3(t3)
    eng.set_intermediates(im, t2)
    eng.residual()
    Ro = tri.compute_r3_tri(o, v, im, t2, t3)
    assert abs(eng.R.download().astype(np.float64) - Ro).max() <= 2e-5 * abs(Ro).max()
    n2, _ = eng.finalize(s.mo_energy)
    t3o, _, n2o, _ = tri.t3_update(o, Ro, t3, s.mo_energy)
    assert abs(eng.get_t3() - t3o).max() <= 2e-5 * abs(t3o).max()
    assert abs(n2 - n2o) <= 1e-4 * n2o


_PAIR_PLANE = {0: (0, 3, 1, 2), 1: (0, 2, 1, 3), 2: (0, 1, 2, 3)}   # block -> [b, z, d, e]


@pytest.mark.parametrize("dt", ["f64", "f32", "packed_f32"])
@pytest.mark.parametrize("v,nb", [(5, 3), (33, 2), (40, 2), (64, 1), (100, 1)])
def test_vvvv_pair_kernels(dev, dt, v, nb):
    """dcc_vvvv_pair_weights / _split / _unfold element for element against NumPy for all three pair
    axes (packed words against dcc_pack_bf16x3 of the FP32 values); outputs start as NaN / junk words so
    padding the kernels leave unwritten shows.  The plane kernels tile the pair plane 32 x 32: nvir=100
    has off-diagonal and partial tiles."""
    from distr_cc import _lib as L
    from distr_cc.device import DeviceArray
    from distr_cc.t3_engine import pair_dims
    _say("vvvv pair kernels %s v=%d nb=%d" % (dt, v, nb))
    rng = np.random.default_rng(v * 10 + nb)
    typ = np.float64 if dt == "f64" else np.float32
    nS, nA, kp = pair_dims(v)
    dS, eS = np.triu_indices(v)
    dA, eA = np.triu_indices(v, 1)

    def junk(shape):
        if dt == "packed_f32":
            return DeviceArray.from_host(np.full(shape, 0xDEADBEEF, np.uint32))
        return DeviceArray.from_host(np.full(shape, np.nan, typ))

    def check(got, ref):
        if dt == "packed_f32":
            r = DeviceArray.from_host(np.ascontiguousarray(ref, np.float32))
            w = DeviceArray(ref.shape, np.uint32)
            L.call("dcc_pack_bf16x3", r.vp, w.vp, ref.size, None)
            assert np.array_equal(got.download(), w.download())
        else:
            assert np.array_equal(got.download(), ref)

    W1 = rng.normal(size=(v, v, v, v)).astype(typ)
    ref = np.zeros((nS + nA, kp), typ)
    WS, WA = W1[dS, eS], W1[dA, eA]
    ref[:nS, :nS] = np.where(dS == eS, WS[:, dS, eS], WS[:, dS, eS] + WS[:, eS, dS])
    ref[nS:, :nA] = WA[:, dA, eA] - WA[:, eA, dA]
    got, dW1 = junk(ref.shape), DeviceArray.from_host(W1)   # named: a temporary would be freed before the launch
    L.call("dcc_vvvv_pair_weights_" + dt, dW1.vp, v, kp, got.vp, None)
    check(got, ref)
    T = rng.normal(size=(nb, v, v, v)).astype(typ)
    dT = DeviceArray.from_host(T)
    for axes in range(3):
        P = T.transpose(_PAIR_PLANE[axes])
        ref = np.zeros((2, nb, v, kp), typ)
        ref[0, :, :, :nS] = P[:, :, dS, eS] + P[:, :, eS, dS]
        ref[1, :, :, :nA] = P[:, :, dA, eA] - P[:, :, eA, dA]
        if axes == 0:
            ref = np.ascontiguousarray(ref.transpose(0, 1, 3, 2))
        got = junk(ref.shape)
        L.call("dcc_vvvv_pair_split_" + dt, dT.vp, nb, v, axes, kp, got.vp, None)
        check(got, ref)
        if dt == "packed_f32":
            continue
        U = rng.normal(size=(2, nb, v, nS)).astype(typ)       # [part][b][z][p]
        R0 = rng.normal(size=(nb, v, v, v)).astype(typ)
        qS = dA * v - dA * (dA - 1) // 2 + (eA - dA)
        full = np.zeros((nb, v, v, v), typ)                   # [b, z, x, y]
        full[:, :, dS, eS] = typ(0.5) * U[0]
        full[:, :, dA, eA] = typ(0.5) * (U[0][:, :, qS] + U[1][:, :, :nA])
        full[:, :, eA, dA] = typ(0.5) * (U[0][:, :, qS] - U[1][:, :, :nA])
        want = R0.copy()
        want.transpose(_PAIR_PLANE[axes])[...] += full
        dU = DeviceArray.from_host(np.ascontiguousarray(U.transpose(0, 1, 3, 2)) if axes == 0 else U)
        dR = DeviceArray.from_host(R0)
        L.call("dcc_vvvv_pair_unfold_" + dt, dU.vp, nb, v, axes, dR.vp, None)
        assert np.array_equal(dR.download(), want), axes


@pytest.mark.parametrize("prec,o,v,chunk", [("fp64", 3, 33, None), ("fp64", 4, 20, 7), ("fp32", 3, 40, None),
                                            ("bf16x3", 3, 33, None), ("bf16x3", 4, 50, 9)])
def test_residual_vvvv_pairs_vs_oracle(dev, prec, o, v, chunk):
    """The residual with pair-symmetric W_vvvv products (chunks of blocks where given) against the oracle
    and against the same engine with the plain W_vvvv GEMMs, on intermediates with independent noise
    except a pair-symmetric one on W_vvvv."""
    from distr_cc.synthetic import SyntheticSystem
    from distr_cc.t3_engine import TriplesEngine, pair_dims
    from oracle import rccsdt_tri as tri
    _say("residual vvvv pairs %s o=%d v=%d chunk=%s" % (prec, o, v, chunk))
    s = SyntheticSystem(o, v, naux=30, seed=900 + o * v)
    rng = np.random.default_rng(901)
    im = {k: x + 1e-3 * rng.normal(size=x.shape) for k, x in s.bare_intermediates().items()}
    n = 1e-3 * rng.normal(size=im['W_vvvv'].shape)
    im['W_vvvv'] = s.bare_intermediates()['W_vvvv'] + n + n.transpose(1, 0, 3, 2)
    t2, t3 = s.t2(), s.t3_tri()
    kw = {"fp64": {}, "fp32": dict(dtype=np.float32), "bf16x3": dict(dtype=np.float32, contract="bf16x3")}[prec]
    nS, nA, kp = pair_dims(v)
    isz = 8 if prec == "fp64" else 4
    eng = TriplesEngine(o, v, vvvv_pairs=True,
                        pair_bytes=None if chunk is None else chunk * 2 * v * (kp + nS) * isz, **kw)
    # chunks of at most `chunk` blocks, made equal (20 blocks with a bound of 9: 3 chunks of 7), in the
    # T2 -> T3 and ring scratch buffers
    nbc = eng.Tp.shape[1]
    assert nbc <= (chunk or eng.n3) and nbc == -(-eng.n3 // -(-eng.n3 // nbc))
    assert chunk is None or nbc == 7 or eng.n3 != 20
    assert eng.Tp.ptr == eng.Xt.ptr and eng.Up.ptr == eng.Gs.ptr
    R = []
    for e in (eng, TriplesEngine(o, v, **kw)):
        e.set_t3(t3)
        e.set_intermediates(im, t2)
        e.residual()
        R.append(e.R.download().astype(np.float64))
    Ro = tri.compute_r3_tri(o, v, im, t2, t3)
    tol = {"fp64": 1e-12, "fp32": 2e-5, "bf16x3": 3e-4}[prec]
    assert abs(R[0] - Ro).max() <= tol * abs(Ro).max(), abs(R[0] - Ro).max() / abs(Ro).max()
    assert abs(R[0] - R[1]).max() <= tol * abs(Ro).max()


@pytest.mark.parametrize("M,N,K", [(16, 16, 32), (37, 53, 29), (130, 70, 300), (64, 64, 64)])
def test_contract_bf16x3_plain_gemm(dev, M, N, K):
    """Split-BF16x3 path: pins the v_mfma_f32_16x16x32_bf16 operand/accumulator layouts and the split
    (against the same split evaluated exactly on the host), and its error against FP64."""
    from abi_emulator import bf16_split
    from distr_cc import _lib as L
    from distr_cc.device import DeviceArray
    _say("contract bf16x3 %dx%dx%d" % (M, N, K))
    rng = np.random.default_rng(M * N + K + 1)
    A = rng.normal(size=(M, K)).astype(np.float32)
    B = rng.normal(size=(K, N)).astype(np.float32)
    C0 = rng.normal(size=(M, N)).astype(np.float32)
    dA, dB, dC = DeviceArray.from_host(A), DeviceArray.from_host(B), DeviceArray.from_host(C0)
    bt = np.zeros(1, dtype=L.BATCH_DTYPE)
    bt[0] = (0, 0, 0, 0, M)
    dbt = DeviceArray.from_host(bt)
    lin = lambda s: L.IdxMap(None, 0, 1, 1 << 20, 0, 0, s)
    L.call("dcc_contract_f32_bf16x3", M, N, K, 1.5, 0.5, dA.vp, lin(K), lin(1), dB.vp, lin(N), lin(1),
           dC.vp, lin(N), lin(1), dbt.vp, 1, M, None)
    got = dC.download().astype(np.float64)
    ah, al = bf16_split(A)
    bh, bl = bf16_split(B)
    split_ref = 0.5 * C0.astype(np.float64) + 1.5 * (ah @ bh + ah @ bl + al @ bh)
    exact = 0.5 * C0.astype(np.float64) + 1.5 * (A.astype(np.float64) @ B.astype(np.float64))
    scale = max(1.0, abs(exact).max())
    assert abs(got - split_ref).max() < 3e-6 * scale * np.sqrt(K)     # FP32 accumulation only
    assert abs(got - exact).max() < 1e-4 * scale * np.sqrt(K)         # split error ~2^-16 per product


@pytest.mark.parametrize("vec", ["none", "pairs", "quads"])
@pytest.mark.parametrize("packed", [False, True])
@pytest.mark.parametrize("tile", [0, 1, 2, 3, 4, 5, 6, 7])
@pytest.mark.parametrize("layout", ["kfast-nfast", "mfast-nfast", "kfast-kfast", "mfast-kfast-ct"])
def test_contract_bf16x3_layouts(dev, layout, tile, packed, vec):
    """Every staging layout of the split-BF16x3 kernel (A contiguous along k or m, B along n or k),
    the LDS-staged epilogue (C contiguous along m), partial M/N tiles, a K tail and a batch offset,
    for each workgroup tile, with FP32 operands or operands packed by dcc_pack_bf16x3; checked
    against the same split evaluated exactly on the host.  ``vec`` picks sizes and offsets that force
    per-element loads, allow 8-byte pair loads (extents 2 mod 4), or allow 16-byte quad loads of
    packed operands (extents and offsets divisible by 4)."""
    from abi_emulator import bf16_split
    from distr_cc import _lib as L
    from distr_cc.device import DeviceArray
    M, N, K = {"none": (150, 83, 75), "pairs": (150, 84, 74), "quads": (152, 84, 76)}[vec]
    even = vec != "none"
    _say("contract bf16x3 layout %s tile %d packed %d vec %s" % (layout, tile, packed, vec))
    rng = np.random.default_rng(11 + tile)
    A = rng.normal(size=(M, K)).astype(np.float32)
    B = rng.normal(size=(K, N)).astype(np.float32)
    C0 = rng.normal(size=(M, N)).astype(np.float32)
    lin = lambda s: L.IdxMap(None, 0, 1, 1 << 20, 0, 0, s)
    a_mfast, b_kfast, c_mt = layout.startswith("mfast"), "-kfast" in layout, layout.endswith("-ct")
    pad = 8 if even else 7   # batch offset of every operand
    hA = np.concatenate([np.zeros(pad, np.float32), (A.T if a_mfast else A).ravel()])
    hB = np.concatenate([np.zeros(pad, np.float32), (B.T if b_kfast else B).ravel()])
    hC = np.concatenate([np.zeros(pad, np.float32), (C0.T if c_mt else C0).ravel()])
    dA, dB, dC = DeviceArray.from_host(hA), DeviceArray.from_host(hB), DeviceArray.from_host(hC)
    am, ak = (lin(1), lin(M)) if a_mfast else (lin(K), lin(1))
    bk, bn = (lin(1), lin(K)) if b_kfast else (lin(N), lin(1))
    cm, cn = (lin(1), lin(M)) if c_mt else (lin(N), lin(1))
    bt = np.zeros(1, dtype=L.BATCH_DTYPE)
    bt[0] = (pad, pad, pad, 0, M)
    dbt = DeviceArray.from_host(bt)
    if packed:
        pA, pB = DeviceArray(hA.shape, np.uint32), DeviceArray(hB.shape, np.uint32)
        L.call("dcc_pack_bf16x3", dA.vp, pA.vp, hA.size, None)
        L.call("dcc_pack_bf16x3", dB.vp, pB.vp, hB.size, None)
        L.call("dcc_contract_packed_bf16x3", M, N, K, 1.5, 0.5, pA.vp, am, ak, pB.vp, bk, bn, dC.vp, cm, cn,
               dbt.vp, 1, M, tile, None)
        # the packing itself: hi in the low half, lo in the high half, as the host split
        w = pA.download()[pad:]
        hi, lo = bf16_split(hA[pad:])
        assert np.array_equal((w << 16).view(np.float32), hi.astype(np.float32))
        assert np.array_equal((w & 0xFFFF0000).view(np.float32), lo.astype(np.float32))
    else:
        L.call("dcc_contract_tiled_f32_bf16x3", M, N, K, 1.5, 0.5, dA.vp, am, ak, dB.vp, bk, bn, dC.vp, cm, cn,
               dbt.vp, 1, M, tile, None)
    out = dC.download()[pad:].astype(np.float64)
    got = out.reshape(N, M).T if c_mt else out.reshape(M, N)
    ah, al = bf16_split(A)
    bh, bl = bf16_split(B)
    split_ref = 0.5 * C0.astype(np.float64) + 1.5 * (ah @ bh + ah @ bl + al @ bh)
    scale = max(1.0, abs(split_ref).max())
    assert abs(got - split_ref).max() < 3e-6 * scale * np.sqrt(K)
    assert np.all(dC.download()[:pad] == 0)


@pytest.mark.parametrize("tile", [0, 1, 2, 3, 4, 5, 6, 7])
def test_contract_bf16x3_vector_epilogue_tail(dev, tile):
    """Packed operands, a short K and C contiguous along n with rows of stride 84 (16-byte vector rows)
    but N = 82: the row tail of the vector epilogue is written element by element, and the two padding
    columns of every row, the elements before the batch offset and after the last row keep their
    values."""
    from abi_emulator import bf16_split
    from distr_cc import _lib as L
    from distr_cc.device import DeviceArray
    M, N, K, ld, pad = 150, 82, 76, 84, 8
    _say("contract bf16x3 vector epilogue tail tile %d" % tile)
    rng = np.random.default_rng(90 + tile)
    A = rng.normal(size=(M, K)).astype(np.float32)
    B = rng.normal(size=(K, N)).astype(np.float32)
    C0 = rng.normal(size=(M, ld)).astype(np.float32)
    lin = lambda s: L.IdxMap(None, 0, 1, 1 << 20, 0, 0, s)
    hC = np.concatenate([np.full(pad, 7.0, np.float32), C0.ravel(), np.full(pad, 7.0, np.float32)])
    hA = np.concatenate([np.zeros(pad, np.float32), A.ravel()])
    hB = np.concatenate([np.zeros(pad, np.float32), B.ravel()])
    pA, pB = DeviceArray(hA.shape, np.uint32), DeviceArray(hB.shape, np.uint32)
    dA, dB = DeviceArray.from_host(hA), DeviceArray.from_host(hB)
    L.call("dcc_pack_bf16x3", dA.vp, pA.vp, hA.size, None)
    L.call("dcc_pack_bf16x3", dB.vp, pB.vp, hB.size, None)
    dC = DeviceArray.from_host(hC)
    bt = np.zeros(1, dtype=L.BATCH_DTYPE)
    bt[0] = (pad, pad, pad, 0, M)
    dbt = DeviceArray.from_host(bt)
    L.call("dcc_contract_packed_bf16x3", M, N, K, 1.5, 0.5, pA.vp, lin(K), lin(1), pB.vp, lin(N), lin(1),
           dC.vp, lin(ld), lin(1), dbt.vp, 1, M, tile, None)
    out = dC.download()
    got = out[pad:pad + M * ld].reshape(M, ld)
    ah, al = bf16_split(A)
    bh, bl = bf16_split(B)
    split_ref = 0.5 * C0[:, :N].astype(np.float64) + 1.5 * (ah @ bh + ah @ bl + al @ bh)
    scale = max(1.0, abs(split_ref).max())
    assert abs(got[:, :N] - split_ref).max() < 3e-6 * scale * np.sqrt(K)
    assert np.array_equal(got[:, N:], C0[:, N:])
    assert np.all(out[:pad] == 7.0) and np.all(out[pad + M * ld:] == 7.0)


@pytest.mark.parametrize("dt", ["f64", "f32", "packed_f32"])
@pytest.mark.parametrize("o,v,i0,n", [(3, 7, 1, 2), (4, 9, 0, 4), (2, 20, 1, 1)])
def test_t2t3_operands(dev, dt, o, v, i0, n):
    """dcc_t2t3_operands_*: the K = (d, l) operands of the fused T2 -> T3 base term, element for element
    against NumPy (the packed form against dcc_pack_bf16x3 of the FP32 operands)."""
    from distr_cc import _lib as L
    from distr_cc.device import DeviceArray
    _say("t2t3 operands %s o=%d v=%d i0=%d n=%d" % (dt, o, v, i0, n))
    rng = np.random.default_rng(7 * o + v)
    typ = np.float64 if dt == "f64" else np.float32
    Wv = rng.normal(size=(o, v, v, v)).astype(typ)     # [j,a,b,d]
    Wo = rng.normal(size=(o, o, v, o)).astype(typ)     # [i,j,a,l]
    t2 = rng.normal(size=(o, o, v, v)).astype(typ)     # [i,k,d,c]
    A = np.empty((n, v, o, v, v + o), typ)
    A[..., :v] = Wv.transpose(2, 0, 1, 3)[None]
    A[..., v:] = -Wo[i0:i0 + n][:, None]
    B = np.empty((n, v, v + o, o, v), typ)
    B[:, :, :v] = t2[i0:i0 + n].transpose(0, 2, 1, 3)[:, None]
    B[:, :, v:] = t2.transpose(2, 0, 1, 3)[None]
    dWv, dWo, dt2 = DeviceArray.from_host(Wv), DeviceArray.from_host(Wo), DeviceArray.from_host(t2)
    out_t = np.uint32 if dt == "packed_f32" else typ
    dA, dB = DeviceArray(A.shape, out_t), DeviceArray(B.shape, out_t)
    L.call("dcc_t2t3_operands_" + dt, dA.vp, dB.vp, dWv.vp, dWo.vp, dt2.vp, o, v, i0, n, None)
    for got, ref in ((dA, A), (dB, B)):
        if dt == "packed_f32":
            r = DeviceArray.from_host(ref)
            w = DeviceArray(ref.shape, np.uint32)
            L.call("dcc_pack_bf16x3", r.vp, w.vp, ref.size, None)
            assert np.array_equal(got.download(), w.download())
        else:
            assert np.array_equal(got.download(), ref)


@pytest.mark.parametrize("o,v", [(3, 9), (5, 17)])
def test_residual_bf16x3_vs_oracle(dev, o, v):
    from distr_cc.synthetic import SyntheticSystem
    from distr_cc.t3_engine import TriplesEngine
    from oracle import rccsdt_tri as tri
    _say("residual bf16x3 o=%d v=%d" % (o, v))
    s = SyntheticSystem(o, v, naux=30, seed=700 + o * v)
    im, t2, t3 = s.bare_intermediates(), s.t2(), s.t3_tri()
    eng = TriplesEngine(o, v, dtype=np.float32, contract="bf16x3")
    eng.set_t3(t3)
    eng.set_intermediates(im, t2)
    eng.residual()
    Ro = tri.compute_r3_tri(o, v, im, t2, t3)
    assert abs(eng.R.download().astype(np.float64) - Ro).max() <= 3e-4 * abs(Ro).max()


def test_batched_ensemble_vs_engines(dev):
    """Grouped persistent launches in a HIP graph == per-system engines (uneven nocc/nvir)."""
    from distr_cc.batched import BatchedTriples
    from distr_cc.device import Stream
    from distr_cc.synthetic import SyntheticSystem
    from distr_cc.t3_engine import TriplesEngine
    systems = [(3, 20), (5, 9), (1, 7), (4, 33), (2, 17)]
    _say("batched ensemble %s" % (systems,))
    st = Stream()
    bt = BatchedTriples(systems, st)
    refs = []
    for n, (o, v) in enumerate(systems):
        s = SyntheticSystem(o, v, naux=40, seed=90 + n)
        im, t2, t3 = s.bare_intermediates(), s.t2(), s.t3_tri()
        bt.set_system(n, t3, im, t2, s.mo_energy)
        e = TriplesEngine(o, v)
        e.set_t3(t3)
        e.set_intermediates(im, t2)
        e.residual()
        refs.append((e.finalize(s.mo_energy), e.get_t3()))
    tuned = bt.autotune()
    _say("autotune %s" % (tuned,))
    for n, (o, v) in enumerate(systems):   # autotune ran GEMMs on the inputs; restore t3 for a clean step
        bt.engines[n].set_t3(SyntheticSystem(o, v, naux=40, seed=90 + n).t3_tri())
    bt.capture()
    bt.step()
    for e, (nrm, t3ref) in zip(bt.engines, refs):
        assert abs(e.get_t3() - t3ref).max() <= 1e-12 * abs(t3ref).max()
        assert abs(e.read_norms()[0] - nrm[0]) <= 1e-11 * nrm[0]
    bt.close()
    # every grouped kernel variant (four tile shapes x one/two stages), on a persistent grid of 7
    # workgroups so each launch walks many windows, the last one partial
    for tile in range(8):
        _say("grouped variant %d" % tile)
        for ph in bt.phases:
            if hasattr(ph, "variant"):
                ph.prepare(tile, st)
                ph.nworkgroups = 7
        for n, (o, v) in enumerate(systems):
            bt.engines[n].set_t3(SyntheticSystem(o, v, naux=40, seed=90 + n).t3_tri())
        bt.issue_step()
        for e, (nrm, t3ref) in zip(bt.engines, refs):
            assert abs(e.get_t3() - t3ref).max() <= 1e-12 * abs(t3ref).max(), tile


@pytest.mark.parametrize("o,v", [(3, 20), (5, 17)])
def test_residual_autotuned_vs_oracle(dev, o, v):
    """Every contraction on its autotuned tile (all four tile shapes get exercised) == oracle."""
    from distr_cc.synthetic import SyntheticSystem
    from distr_cc.t3_engine import TriplesEngine
    from oracle import rccsdt_tri as tri
    _say("residual autotuned o=%d v=%d" % (o, v))
    s = SyntheticSystem(o, v, naux=30, seed=800 + o * v)
    im, t2, t3 = s.bare_intermediates(), s.t2(), s.t3_tri()
    eng = TriplesEngine(o, v)
    eng.set_t3(t3)
    eng.set_intermediates(im, t2)
    tuned = eng.autotune()
    _say("tiles %s" % sorted(set(t for t, _ in tuned.values())))
    for tile in range(4):   # force each shape once on every launch, then the tuned choice
        for g in eng.launches:
            if hasattr(g, "tile"):
                g.tile = tile
        eng.residual()
        Ro = tri.compute_r3_tri(o, v, im, t2, t3)
        assert abs(eng.R.download() - Ro).max() <= 1e-12 * abs(Ro).max(), tile


@pytest.mark.parametrize("o,v,vp", [(4, 9, 12), (5, 19, 20)])
def test_t3_abc_layout_kernels_vs_oracle(dev, o, v, vp):
    """IJK -> ABC (DistributedT3ABC layout) against the oracle's restatement of the reference packing,
    bit for bit, and the ABC -> dense (Q) gather against the full tensor."""
    import itertools
    from distr_cc import _lib as L
    from distr_cc.device import DeviceArray
    from oracle import t3_layout as TL
    _say("ABC layout kernels o=%d v=%d vpad=%d" % (o, v, vp))
    rng = np.random.default_rng(o * 100 + v)
    M = o * v
    table = rng.standard_normal(M ** 3)
    # exactly symmetric under simultaneous (ijk)(abc) permutations
    T = np.empty((o,) * 3 + (v,) * 3)
    for i, j, k in itertools.product(range(o), repeat=3):
        p = np.sort(np.stack(np.broadcast_arrays((i * v + np.arange(v))[:, None, None],
                                                 (j * v + np.arange(v))[None, :, None],
                                                 (k * v + np.arange(v))[None, None, :])), axis=0)
        T[i, j, k] = table[(p[0] * M + p[1]) * M + p[2]]
    t3 = np.ascontiguousarray([T[i, j, k] for (i, j, k) in TL.canonical_triples(o)])
    src = DeviceArray.from_host(t3)
    abc = DeviceArray((v * (v + 1) * (v + 2) // 6, o, o, o))
    L.call("dcc_t3_tri_to_abc_packed_f64", src.vp, o, v, abc.vp, None)
    got = abc.download()
    assert np.array_equal(got, TL.t3_abc_from_tri(t3, o))
    dense = DeviceArray((vp, vp, vp, o, o, o))
    L.call("dcc_t3_abc_packed_to_dense_f64", abc.vp, o, v, vp, dense.vp, None)
    d = dense.download()
    assert np.array_equal(d[:v, :v, :v], T.transpose(3, 4, 5, 0, 1, 2))
    assert not d[v:].any() and not d[:, v:].any() and not d[:, :, v:].any()
