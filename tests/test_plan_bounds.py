"""Launch plans of large engines (the residual and the T3 -> r1/r2/W chunks) checked on the host, where
buffer offsets pass 2^31: every GEMM operand range (index maps over each batch entry's rows), every S3
job's target and source blocks and every pair split/unfold range lies inside the buffer its launch points
into, and no offset arithmetic overflows.
A fake device allocator hands out address ranges: nothing is allocated or run."""
import warnings

import numpy as np
import pytest

import abi_emulator


def _fake_device(monkeypatch):
    abi_emulator.install(monkeypatch)
    from distr_cc import _lib as L
    from distr_cc import device as D
    allocs, nxt = [], [1 << 40]

    def call(name, *args):
        if name == "dcc_malloc":
            n = int(args[1])
            args[0]._obj.value = nxt[0]
            allocs.append((nxt[0], n))
            nxt[0] += -(-n // 4096) * 4096 + 4096
        return 0

    def from_host(cls, arr, stream=None, dtype=None):
        a = np.ascontiguousarray(arr, dtype=dtype if dtype is not None else np.asarray(arr).dtype)
        d = cls(a.shape, a.dtype)
        d._host = a
        return d

    monkeypatch.setattr(L, "call", call)
    monkeypatch.setattr(D.DeviceArray, "from_host", classmethod(from_host))
    return allocs


def _owner(allocs, p):
    for s, n in allocs:
        if s <= p < s + n:
            return s, n
    raise AssertionError("pointer %#x outside every allocation" % p)


def _map(mp, count):
    """Offsets of indices 0 .. count-1 under a three-level dcc_idxmap."""
    x = np.arange(count, dtype=np.int64)
    e2, e12 = max(1, int(mp.e2)), max(1, int(mp.e1) * int(mp.e2))
    q0 = x // e12
    r0 = x - q0 * e12
    q1 = r0 // e2
    return q0 * mp.s0 + q1 * mp.s1 + (r0 - q1 * e2) * mp.s2


def _check_gemm(g, allocs, isz):
    """Largest element offset of the launch after checking every operand range against its buffer."""
    M, N, K = g.args[:3]
    A, am, ak, B, bk, bn, C, cm, cn = g.args[5:14]
    assert not any(mp.table for mp in (am, ak, bk, bn, cm, cn)), g.name
    k_a, k_b, n_b, n_c = _map(ak, K), _map(bk, K), _map(bn, N), _map(cn, N)
    big = 0
    for row in g.batches._host:
        mb, me = int(row["m_begin"]), int(row["m_end"])
        m_a, m_c = _map(am, me)[mb:], _map(cm, me)[mb:]
        for what, ptr, off, r1, r2 in (("A", A.value, row["a_off"], m_a, k_a),
                                       ("B", B.value, row["b_off"], k_b, n_b),
                                       ("C", C.value, row["c_off"], m_c, n_c)):
            lo, hi = int(off) + int(r1.min()) + int(r2.min()), int(off) + int(r1.max()) + int(r2.max())
            s, n = _owner(allocs, ptr)
            assert lo >= 0 and ptr + (hi + 1) * isz <= s + n, (g.name, what, hi, (s + n - ptr) // isz)
            big = max(big, hi)
    return big


@pytest.mark.parametrize("o,v,prec,gbytes", [(16, 160, "bf16x3", 16), (16, 160, "fp64", 8), (10, 90, "fp64", 8)])
def test_residual_plan_in_bounds(monkeypatch, o, v, prec, gbytes):
    allocs = _fake_device(monkeypatch)
    from distr_cc.t3_engine import GemmLaunch, PairLaunch, S3GatherLaunch, S3Launch, TriplesEngine
    kw = dict(dtype=np.float32, contract="bf16x3") if prec == "bf16x3" else {}
    with warnings.catch_warnings():
        warnings.simplefilter("error", RuntimeWarning)     # NumPy integer overflow in offset arithmetic
        eng = TriplesEngine(o, v, ring_fused=True, ring_g_bytes=gbytes << 30, vvvv_pairs=True, **kw)
        eng._setup_lower()                                   # T3 -> r1/r2/W chunks (Z slices of T3)
    isz = eng.dtype.itemsize
    v3 = v ** 3
    big = 0
    for ch in eng._lower["chunks"]:
        jobs = ch[0]._host
        assert (jobs["dst"] >= 0).all() and (jobs["dst"] + v3 <= eng._lower["Z"].size).all()
        assert (jobs["src"] >= 0).all() and (jobs["src"] + v3 <= eng.t3.size).all()
        for g in ch[2] + ch[3]:
            if isinstance(g, GemmLaunch):
                big = max(big, _check_gemm(g, allocs, isz))
    for g in eng.launches:
        if isinstance(g, GemmLaunch):
            big = max(big, _check_gemm(g, allocs, isz))
        elif isinstance(g, (S3GatherLaunch, S3Launch)):
            jobs = g.jobs._host
            assert (jobs["dst"] >= 0).all() and (jobs["dst"] + v3 <= g.out.size).all(), g.name
            src = jobs["src"].reshape(len(jobs), -1)
            used = jobs["coef"] != 0 if src.shape[1] == 6 else np.ones_like(src, dtype=bool)
            assert (src[used] >= 0).all() and (src[used] + v3 <= g.src.size).all(), g.name
            big = max(big, int(src[used].max()) if used.any() else 0)
        elif isinstance(g, PairLaunch):
            first = (g.blk.value - g.blocks.ptr) // g.blocks.dtype.itemsize
            assert first >= 0 and first + g.nb * v3 <= g.blocks.size, g.name
            nS, nA, kp = eng.pair_dims
            assert g.buf.size >= 2 * g.nb * v * (kp if g.kpad is not None else nS), g.name
    if (o, v, gbytes) == (16, 160, 16):
        assert big >= 2 ** 31      # the case where 32-bit offset arithmetic would wrap
