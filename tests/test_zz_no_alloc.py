"""The allocation-free finalize entry point against the one with a caller workspace.

dcc_r3_finalize_noalloc_f64 keeps its partial norms in a file-scope device array of libdcc.  Two
calls in a row (the second on fewer, smaller blocks, so stale partials of the first call would show)
must give bit for bit the t3 update, the projected residual and the two norms of
dcc_r3_finalize_f64 + dcc_reduce_pairs_f64, and the norms must match the oracle's t3_update.
"""
import ctypes
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _case(nocc, nvir, seed):
    from oracle import t3_layout
    rng = np.random.default_rng(seed)
    trip = np.array(t3_layout.canonical_triples(nocc), dtype=np.int32)
    n3 = len(trip)
    R = rng.normal(size=(n3, nvir, nvir, nvir))
    t3 = 1e-3 * rng.normal(size=(n3, nvir, nvir, nvir))
    e = np.concatenate([np.sort(rng.uniform(-1.5, -0.3, nocc)), np.sort(rng.uniform(0.1, 3.0, nvir))])
    return trip, R, t3, e


def _run_both(trip, R, t3, e, nocc, nvir, nvir_diis, level_shift):
    from distr_cc import _lib as L
    from distr_cc.device import DeviceArray
    n3 = len(trip)
    eia = np.ascontiguousarray(e[:nocc, None] - e[None, nocc:] - level_shift)
    dR, dtrip, deia = DeviceArray.from_host(R), DeviceArray.from_host(trip), DeviceArray.from_host(eia)
    # with the caller's workspace
    ta, ra = DeviceArray.from_host(t3), DeviceArray(R.shape)
    part = DeviceArray((2 * n3 * L._lib.dcc_r3_finalize_tiles(nvir),))
    na = DeviceArray((2,))
    L.call("dcc_r3_finalize_f64", dR.vp, ta.vp, ra.vp, n3, nvir, dtrip.vp, deia.vp, nocc, nvir_diis, part.vp, None)
    L.call("dcc_reduce_pairs_f64", part.vp, n3 * L._lib.dcc_r3_finalize_tiles(nvir), na.vp, None)
    # allocation-free
    tb, rb, nb = DeviceArray.from_host(t3), DeviceArray(R.shape), DeviceArray((2,))
    nb.fill(-1.0)
    L.call("dcc_r3_finalize_noalloc_f64", dR.vp, tb.vp, rb.vp, n3, nvir, dtrip.vp, deia.vp, nocc, nvir_diis,
           nb.vp, None)
    return (ta.download(), ra.download(), na.download()), (tb.download(), rb.download(), nb.download())


def test_noalloc_finalize_matches_workspace_entry_point():
    from distr_cc import device
    from oracle import rccsdt_tri as tri
    device.init(0)
    for (nocc, nvir, nd, shift, seed) in ((6, 40, 23, 0.05, 1), (3, 13, 13, 0.0, 2)):
        _say("noalloc finalize nocc=%d nvir=%d nvir_diis=%d level_shift=%g" % (nocc, nvir, nd, shift))
        trip, R, t3, e = _case(nocc, nvir, seed)
        (ta, ra, na), (tb, rb, nb) = _run_both(trip, R, t3, e, nocc, nvir, nd, shift)
        assert np.array_equal(ta, tb) and np.array_equal(ra, rb)
        assert np.array_equal(na, nb), (na, nb)
        t_ref, r_ref, n2, act = tri.t3_update(nocc, R, t3, e, level_shift=shift, nvir_diis=nd)
        assert np.abs(tb - t_ref).max() <= 1e-13 * np.abs(t_ref).max()
        assert abs(nb[0] - n2) <= 1e-12 * n2 and abs(nb[1] - act) <= 1e-12 * act


def test_noalloc_finalize_rejects_oversized_call():
    """A size beyond the scratch is refused on the host, before any launch (no device pointer is read)."""
    from distr_cc import _lib as L
    from distr_cc import device
    device.init(0)
    tiles = L._lib.dcc_r3_finalize_tiles(160)
    nblk = (3 << 20) // (2 * tiles) + 1
    st = L._lib.dcc_r3_finalize_noalloc_f64(None, None, None, nblk, 160, None, None, 16, 160,
                                            ctypes.c_void_p(8), None)
    assert st == 2
    assert b"allocation-free scratch" in L._lib.dcc_last_error()
