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
    L.call("dcc_contract_f64", M, N, K, 1.5, 0.5, dA.vp, lin(K), lin(1), dB.vp, lin(N), lin(1),
           dC.vp, lin(N), lin(1), dbt.vp, 1, M, None)
    C = dC.download()
    ref = 0.5 * C0 + 1.5 * A @ B
    assert abs(C - ref).max() < 1e-12 * max(1.0, abs(ref).max())


def test_contract_transposed_operands(dev):
    """A stored K-major and B stored N-major-transposed exercise the other two loaders."""
    from distr_cc import _lib as L
    from distr_cc.device import DeviceArray
    M, N, K = 75, 41, 33
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
    L.call("dcc_contract_f64", M, N, K, 1.0, 0.0, dA.vp, lin(1), lin(M), dB.vp, lin(1), lin(K),
           dC.vp, lin(N), lin(1), dbt.vp, 1, M, None)
    ref = At.T @ Bt.T
    assert abs(dC.download() - ref).max() < 1e-12 * abs(ref).max()


@pytest.mark.parametrize("v", [1, 5, 8, 13, 20])
def test_s3_transform_all_perms(dev, v):
    from distr_cc import _lib as L
    from distr_cc.device import DeviceArray
    from distr_cc.t3_engine import s3_jobs_array
    _say("s3 transform v=%d" % v)
    rng = np.random.default_rng(v)
    X = rng.normal(size=(3, v, v, v))
    perms = ((0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0))
    coefs = [np.eye(6)[p] for p in range(6)] + [np.array([2.0, -1.0, -1.0, 0.5, 0.25, -3.0])]
    jobs = []
    for n, c in enumerate(coefs):
        jobs.append((n * v ** 3, (n % 3) * v ** 3, c))
    dX = DeviceArray.from_host(X)
    dO = DeviceArray.zeros((len(coefs), v, v, v))
    dj = DeviceArray.from_host(s3_jobs_array(jobs))
    L.call("dcc_s3_transform_f64", dO.vp, dX.vp, v, 0.0, dj.vp, len(jobs), None)
    out = dO.download()
    for n, c in enumerate(coefs):
        src = X[n % 3]
        ref = sum(ci * src.transpose(p) for ci, p in zip(c, perms))
        assert abs(out[n] - ref).max() < 1e-13, (n, v)


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
    assert abs(n2 - n2o) <= 1e-12 * n2o + 1e-30 and abs(na - nao) <= 1e-12 * nao + 1e-30
