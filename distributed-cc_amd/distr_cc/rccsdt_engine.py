"""A whole RCCSDT iteration resident on one GPU.

``RCCSDTEngine.update_amps`` performs update_amps_rccsdt_tri_
(distr_cc/rccsdt.py:890-977) with every tensor in HBM: the T1 dressing of the
integrals (call site :909), the singles/doubles intermediates and residual
(:912-915), T3 -> r1/r2 (:916), symmetrisation, denominators and the T1/T2
update (:919-933), the T2-dressed triples intermediates plus T3 -> W (:935-936),
the triples residual (:941) and the fused projection/denominator/update of
T3 (:947-971).  ``DeviceDIIS`` is the unified T1/T2/T3 DIIS of
run_diis_t3 (:1026-1067) and distr_cc/diis.py:25-177 on device vectors.

Index conventions: integrals are physicist <pq|rs> (creators p, q;
annihilators r, s) in the orbital order occupied-then-virtual, as eris.pppp
of the reference (rccsdt.py:216-221).  Letters i..n are occupied, a..h
virtual, P..S general orbitals.

The closed-shell equations of the lower-order part are those of
oracle/rccsdt_lower.py (the reference takes them from PySCF 2.13.1); the
device program below issues the same contractions.
"""
import ctypes

import numpy as np

from distr_cc import _lib as L
from distr_cc.device import DeviceArray
from distr_cc.t3_engine import Operand, TriplesEngine, plan_einsum, tensor_axpby

OCC = "ijklmn"
VIR = "abcdefgh"
GEN = "PQRS"


# split-K thresholds: contractions with K >= SPLIT_K_MIN whose output fills fewer than 64 tiles, peeled
# K letter of at least SPLIT_MIN_SLICES values
SPLIT_K_MIN = 4096
SPLIT_MIN_SLICES = 8


def _split_k_letter(sub, A, B, C, ext, tile=64, min_tiles=64):
    """For a contraction whose M x N would occupy fewer than ``min_tiles`` tiles of tile x tile against
    a long K (an r1, F_oo or Foo1-type term: one occupied or virtual index per side, K = o v^2), the K
    letter to peel into batch entries that write partial slices: the largest one that is not the
    contiguous index of A or B (so every slice still loads coalesced).  None when the contraction
    fills the chip anyway or C is not a contiguous array in its letter order."""
    sa, rest = sub.split(',')
    sb, sc = rest.split('->')
    kl = [x for x in sa if x in sb and x not in sc]
    m = int(np.prod([ext[x] for x in sc if x in sa]))
    n = int(np.prod([ext[x] for x in sc if x in sb]))
    k = int(np.prod([ext[x] for x in kl]))
    if (-(-m // tile) * -(-n // tile) >= min_tiles or k < SPLIT_K_MIN or C.shape is None
            or tuple(C.letters) != tuple(sc)):
        return None
    shape = [ext[x] for x in sc]
    acc, want = 1, []
    for e in reversed(shape):
        want.append(acc)
        acc *= e
    if [C.strides[x] for x in sc] != list(reversed(want)):
        return None
    cand = [x for x in kl if A.strides[x] != 1 and B.strides[x] != 1 and SPLIT_MIN_SLICES <= ext[x] <= 1024]
    return max(cand, key=lambda x: ext[x]) if cand else None


def split_k_partial_elems(sub, A, B, C, ext, tile=64, min_tiles=64):
    """Elements of the partial slices plan_einsum_steps would use for this contraction (0: no split)."""
    x = _split_k_letter(sub, A, B, C, ext, tile, min_tiles)
    if x is None:
        return 0
    sc = sub.split('->')[1]
    return int(ext[x] * np.prod([ext[c] for c in sc]))


def plan_einsum_steps(name, sub, A, B, C, alpha, beta, ext, stream, tile=64, min_tiles=64, scratch=None):
    """C = alpha * einsum(sub, A, B) + beta * C as device steps (callables of the stream, the first one
    the GemmLaunch itself) and the buffers they hold: one gather-GEMM, or, when the output fills fewer than ``min_tiles`` tiles against
    a long K, a split along one K letter x -- slice s fixes x = s and writes its own partial of C, and
    the slices are then summed in a fixed order into C (deterministic, no atomics).  The partials go
    to ``scratch`` (a DeviceArray of C's dtype, shared by steps that run one after another on one
    stream) when it is large enough, else to a buffer of their own."""
    x = _split_k_letter(sub, A, B, C, ext, tile, min_tiles)
    if x is None:
        return [plan_einsum(name, sub, A, B, C, alpha, beta, ext, [({}, {}, {})])], []
    sa, rest = sub.split(',')
    sb, sc = rest.split('->')
    S = ext[x]
    sl = next(c for c in "TUVWYZ" if c not in ext and c not in sub)
    shape = (S,) + tuple(ext[c] for c in sc)
    if scratch is not None and scratch.dtype == C.buf.dtype and scratch.size >= int(np.prod(shape)):
        P = DeviceArray(shape, C.buf.dtype, ptr=scratch.ptr, owner=False)
    else:
        P = DeviceArray(shape, C.buf.dtype)
    g = plan_einsum(name + " split-K", "%s,%s->%s" % (sa.replace(x, ""), sb.replace(x, ""), sc),
                    A, B, Operand(P, sl + sc, P.shape), alpha, 0.0, dict(ext, **{sl: S}),
                    [({x: s}, {x: s}, {sl: s}) for s in range(S)])
    count = int(np.prod(P.shape[1:]))
    fn = "dcc_sum_slices_" + ("f64" if C.buf.dtype == np.float64 else "f32")
    dst = C.buf.offset_ptr(C.base)
    return [g, lambda _s: L.call(fn, dst, P.vp, S, count, 1.0, float(beta),
                                 stream.handle if stream is not None else None)], [P]


class _Program:
    """An ordered list of device steps built once and replayed every iteration."""

    def __init__(self, engine):
        self.eng = engine
        self.steps = []
        self.keep = []
        # the operations in order, (kind, ...), for programs that are differentiated (adjoint.py)
        self.ops = []
        # the steps of every operation, one list per operation (a contraction's GemmLaunch first)
        self.op_steps = []

    def add(self, steps):
        """An operation given as its list of steps."""
        self.op_steps.append(list(steps))
        self.steps += steps

    def ein(self, name, sub, A, B, C, alpha, beta):
        self.ops.append(("ein", name, sub, A, B, C, alpha, beta))
        steps, keep = plan_einsum_steps(name, sub, A, B, C, alpha, beta, self.eng.ext, self.eng.stream)
        self.add(steps)
        self.keep += keep

    def axpby(self, dst, src, letters, alpha, beta):
        self.ops.append(("axpby", dst, src, letters, alpha, beta))
        ext = self.eng.ext
        shape = tuple(ext[x] for x in letters)
        stream = self.eng.stream
        self.add([lambda _s: tensor_axpby(dst, src, letters, shape, alpha, beta, stream)])

    def call(self, fn):
        self.ops.append(("call", fn))
        self.add([lambda _s: fn()])

    def extend(self, ops):
        """Operations given as lists of steps (TriplesEngine.t3_to_r1r2_ops)."""
        for op in ops:
            self.ops.append(("steps", op))
            self.add(op)

    def run(self, stream):
        for step in self.steps:
            step(stream)


class RCCSDTEngine:
    """Device state of the RCCSDT amplitude iteration for one closed-shell system."""

    def __init__(self, nocc, nvir, fock, pppp, mo_energy, level_shift=0.0, nvir_diis=None, stream=None,
                 device=0, t3_dtype=np.float64, triples_contract="native", vvvv_pairs="auto"):
        """``t3_dtype`` float32 selects the mixed-precision iteration: T3, the triples residual and its
        operands in FP32 with FP32-MFMA contractions, everything else (integrals, T1/T2, lower-order
        residual, denominators, projection arithmetic, norms, energy, DIIS) in FP64;
        ``triples_contract="bf16x3"`` (with float32) runs those contractions as split-BF16x3.
        ``vvvv_pairs`` is the triples engine's option: the W_vvvv built here (the T1-dressed <ab|de>
        plus <lm|de> t2[l,m,a,b]) is pair-symmetric, so "auto" uses the half-flop pair products
        whenever the device has room for them."""
        o, v = int(nocc), int(nvir)
        n = o + v
        fock = np.asarray(fock, dtype=np.float64)
        pppp = np.asarray(pppp, dtype=np.float64)
        if fock.shape != (n, n) or pppp.shape != (n, n, n, n):
            raise ValueError("fock must be (nmo, nmo) and pppp (nmo,)*4 with nmo = nocc + nvir")
        self.nocc, self.nvir, self.nmo = o, v, n
        self.stream = stream
        self.level_shift = float(level_shift)
        # resolve_nvir_diis (_runtime.py:121-128): None -> nvir, positive, at most nvir
        self.nvir_diis = v if nvir_diis is None else int(nvir_diis)
        if self.nvir_diis <= 0:
            raise ValueError("nvir_diis must be a positive integer")
        self.nvir_diis = min(self.nvir_diis, v)
        self.tr = tr = TriplesEngine(o, v, stream=stream, device=device, dtype=t3_dtype,
                                     contract=triples_contract, vvvv_pairs=vvvv_pairs)
        self.mixed = tr.dtype != np.float64
        self.ext = {x: o for x in OCC}
        self.ext.update({x: v for x in VIR})
        self.ext.update({x: n for x in GEN})
        self.ext["X"] = 1
        f64 = np.float64
        # resident inputs
        self.G0 = DeviceArray.from_host(pppp, stream)
        self.F0 = DeviceArray.from_host(fock, stream)
        self.mo_energy = np.asarray(mo_energy, dtype=f64).copy()
        eia0 = self.mo_energy[:o, None] - self.mo_energy[None, o:]
        # unshifted e_i - e_a for the MP2 guess (rccsdt.py:212-214); shifted for the residual division
        self.eia0 = DeviceArray.from_host(np.ascontiguousarray(eia0), stream)
        self.eia = DeviceArray.from_host(np.ascontiguousarray(eia0 - self.level_shift), stream)
        # work buffers
        self.G = DeviceArray((n, n, n, n), f64)
        self.Hc = DeviceArray((n, n), f64)
        self.Fd = DeviceArray((n, n), f64)
        self.ones = DeviceArray.from_host(np.ones((o, 1)), stream)
        self.t1 = DeviceArray.zeros((o, v), f64, stream)
        self.t2 = DeviceArray((o, o, v, v), f64) if self.mixed else tr.t2
        # FP64 targets of the triples intermediates (the triples engine's own buffers unless mixed)
        names = ("F_oo", "F_vv", "W_oooo", "W_vvvv", "W_ovov", "W_ovvo", "W_ovvv", "W_oovo")
        self.imd = {k: (DeviceArray(getattr(tr, k).shape, f64) if self.mixed else getattr(tr, k)) for k in names}
        self.r1 = DeviceArray((o, v), f64)
        self.r2 = DeviceArray((o, o, v, v), f64)
        self.r2s = DeviceArray((o, o, v, v), f64)
        self.Foo1 = DeviceArray((o, o), f64)
        self.Fvv1 = DeviceArray((v, v), f64)
        self.Woooo1 = DeviceArray((o, o, o, o), f64)
        self.Wvoov = DeviceArray((v, o, o, v), f64)
        self.Wvovo = DeviceArray((v, o, v, o), f64)
        self.h_ovvo = DeviceArray((o, v, v, o), f64)
        self.Loovv = DeviceArray((o, o, v, v), f64)
        self.fov0 = DeviceArray((o, v), f64)
        self.tau = DeviceArray((o, o, v, v), f64)
        self.partials = DeviceArray((int(L._lib.dcc_dot_partials()),), f64)
        self.scal = DeviceArray((8,), f64)
        self.t2.fill(0.0, stream)
        tr.t3.fill(0.0, stream)
        self._setup_constants()
        tr._setup_lower()
        self.prog = self._build_program()
        self._eprog = self._build_energy_program()   # planned here: its launch tables are uploaded once

    # ------------------------------------------------------------------ views
    def gview(self, buf, letters):
        """View of an nmo^4 (or nmo^2) orbital tensor through block letters."""
        n, o = self.nmo, self.nocc
        nd = len(letters)
        strides = [n ** (nd - 1 - d) for d in range(nd)]
        st, base = {}, 0
        for x, s in zip(letters, strides):
            if x in VIR:
                base += o * s
            st[x] = s
        return Operand.strided(buf, st, base)

    @staticmethod
    def nview(buf, letters):
        return Operand(buf, letters, buf.shape)

    def set_integrals(self, pppp):
        """Replace the integrals <pq|rs> (same nmo) and rebuild what is derived from them."""
        n = self.nmo
        pppp = np.asarray(pppp, dtype=np.float64)
        if pppp.shape != (n, n, n, n):
            raise ValueError("pppp must be (nmo,)*4")
        self.G0.upload(np.ascontiguousarray(pppp), self.stream)
        self._setup_constants()

    def _setup_constants(self):
        o, v, n = self.nocc, self.nvir, self.nmo
        s = self.stream
        p = _Program(self)
        # h = F - sum_k (2 <Pk|Qk> - <Pk|kQ>) from the bare integrals
        p.axpby(self.nview(self.Hc, "PQ"), self.nview(self.F0, "PQ"), "PQ", 1.0, 0.0)
        self._jk(p, self.G0, self.Hc, -1.0)
        # E_corr pieces: L_oovv = 2 <ij|ab> - <ij|ba>, f_ov
        p.axpby(self.nview(self.Loovv, "ijab"), self.gview(self.G0, "ijab"), "ijab", 2.0, 0.0)
        p.axpby(self.nview(self.Loovv, "ijab"), self.gview(self.G0, "ijba"), "ijab", -1.0, 1.0)
        p.axpby(self.nview(self.fov0, "ia"), self.gview(self.F0, "ia"), "ia", 1.0, 0.0)
        p.run(s)

    def _jk(self, p, G, F, sign):
        """F[P,Q] += sign * sum_k (2 <Pk|Qk> - <Pk|kQ>)."""
        n, o = self.nmo, self.nocc
        J = Operand.strided(G, {"P": n ** 3, "Q": n, "k": n * n + 1})
        K = Operand.strided(G, {"P": n ** 3, "Q": 1, "k": n * n + n})
        C = Operand.strided(F, {"P": n, "Q": 1, "X": 0})
        one = self.nview(self.ones, "kX")
        p.ein("jk J", "PQk,kX->PQX", J, one, C, 2.0 * sign, 1.0)
        p.ein("jk K", "PQk,kX->PQX", K, one, C, -1.0 * sign, 1.0)

    # ---------------------------------------------------------------- program
    def _build_program(self):
        o, v, n = self.nocc, self.nvir, self.nmo
        tr = self.tr
        G, gv, nv = self.G, self.gview, self.nview
        t1 = nv(self.t1, "ia")

        def t2v(letters):
            return Operand(self.t2, letters, (o, o, v, v))

        steps = {}

        # -- T1 dressing of the integrals and the Fock matrix (rccsdt.py:909)
        p = _Program(self)
        p.call(lambda: L.call("dcc_memcpy_dtod", self.G.vp, self.G0.vp, self.G.nbytes,
                              self.stream.handle if self.stream is not None else None))
        p.ein("dress c0", "ia,iQRS->aQRS", t1, gv(G, "iQRS"), gv(G, "aQRS"), -1.0, 1.0)
        p.ein("dress c1", "ia,PiRS->PaRS", t1, gv(G, "PiRS"), gv(G, "PaRS"), -1.0, 1.0)
        p.ein("dress a2", "PQaS,ia->PQiS", gv(G, "PQaS"), t1, gv(G, "PQiS"), 1.0, 1.0)
        p.ein("dress a3", "PQRa,ia->PQRi", gv(G, "PQRa"), t1, gv(G, "PQRi"), 1.0, 1.0)
        p.axpby(nv(self.Fd, "PQ"), nv(self.Hc, "PQ"), "PQ", 1.0, 0.0)
        p.ein("dress fock c", "ia,iQ->aQ", t1, gv(self.Fd, "iQ"), gv(self.Fd, "aQ"), -1.0, 1.0)
        p.ein("dress fock a", "Pa,ia->Pi", gv(self.Fd, "Pa"), t1, gv(self.Fd, "Pi"), 1.0, 1.0)
        self._jk(p, G, self.Fd, 1.0)
        steps["dress"] = p

        # -- singles/doubles intermediates and residual without T3 (:912-915)
        p = _Program(self)
        Fd = self.Fd
        Foo1, Fvv1 = nv(self.Foo1, "ki"), nv(self.Fvv1, "ac")
        p.axpby(Foo1, gv(Fd, "ki"), "ki", 1.0, 0.0)
        p.ein("Foo1 a", "klcd,ilcd->ki", gv(G, "klcd"), t2v("ilcd"), Foo1, 2.0, 1.0)
        p.ein("Foo1 b", "kldc,ilcd->ki", gv(G, "kldc"), t2v("ilcd"), Foo1, -1.0, 1.0)
        p.axpby(Fvv1, gv(Fd, "ac"), "ac", 1.0, 0.0)
        p.ein("Fvv1 a", "klcd,klad->ac", gv(G, "klcd"), t2v("klad"), Fvv1, -2.0, 1.0)
        p.ein("Fvv1 b", "kldc,klad->ac", gv(G, "kldc"), t2v("klad"), Fvv1, 1.0, 1.0)
        W4 = nv(self.Woooo1, "klij")
        p.axpby(W4, gv(G, "klij"), "klij", 1.0, 0.0)
        p.ein("Woooo1", "klcd,ijcd->klij", gv(G, "klcd"), t2v("ijcd"), W4, 1.0, 1.0)
        Wvoov = nv(self.Wvoov, "akic")
        p.axpby(Wvoov, gv(G, "akic"), "akic", 1.0, 0.0)
        p.ein("Wvoov a", "lkdc,ilda->akic", gv(G, "lkdc"), t2v("ilda"), Wvoov, -0.5, 1.0)
        p.ein("Wvoov b", "lkcd,ilad->akic", gv(G, "lkcd"), t2v("ilad"), Wvoov, -0.5, 1.0)
        p.ein("Wvoov c", "lkdc,ilad->akic", gv(G, "lkdc"), t2v("ilad"), Wvoov, 1.0, 1.0)
        Wvovo = nv(self.Wvovo, "akci")
        p.axpby(Wvovo, gv(G, "akci"), "akci", 1.0, 0.0)
        p.ein("Wvovo", "lkcd,ilda->akci", gv(G, "lkcd"), t2v("ilda"), Wvovo, -0.5, 1.0)

        r1, r2 = nv(self.r1, "ia"), nv(self.r2, "ijab")
        p.axpby(r1, gv(Fd, "ai"), "ia", 1.0, 0.0)
        p.ein("r1 fov a", "kc,kica->ia", gv(Fd, "kc"), t2v("kica"), r1, 2.0, 1.0)
        p.ein("r1 fov b", "kc,ikca->ia", gv(Fd, "kc"), t2v("ikca"), r1, -1.0, 1.0)
        p.ein("r1 vovv a", "akcd,ikcd->ia", gv(G, "akcd"), t2v("ikcd"), r1, 2.0, 1.0)
        p.ein("r1 vovv b", "akdc,ikcd->ia", gv(G, "akdc"), t2v("ikcd"), r1, -1.0, 1.0)
        p.ein("r1 oovo a", "lkci,klac->ia", gv(G, "lkci"), t2v("klac"), r1, -2.0, 1.0)
        p.ein("r1 oovo b", "klci,klac->ia", gv(G, "klci"), t2v("klac"), r1, 1.0, 1.0)
        p.axpby(r2, gv(G, "abij"), "ijab", 0.5, 0.0)
        p.ein("r2 oooo", "klij,klab->ijab", W4, t2v("klab"), r2, 0.5, 1.0)
        p.ein("r2 vvvv", "abcd,ijcd->ijab", gv(G, "abcd"), t2v("ijcd"), r2, 0.5, 1.0)
        p.ein("r2 Fvv", "ac,ijcb->ijab", Fvv1, t2v("ijcb"), r2, 1.0, 1.0)
        p.ein("r2 Foo", "ki,kjab->ijab", Foo1, t2v("kjab"), r2, -1.0, 1.0)
        p.ein("r2 ring a", "akic,kjcb->ijab", Wvoov, t2v("kjcb"), r2, 2.0, 1.0)
        p.ein("r2 ring b", "akci,kjcb->ijab", Wvovo, t2v("kjcb"), r2, -1.0, 1.0)
        p.ein("r2 ring c", "akic,kjbc->ijab", Wvoov, t2v("kjbc"), r2, -1.0, 1.0)
        p.ein("r2 ring d", "bkci,kjac->ijab", nv(self.Wvovo, "bkci"), t2v("kjac"), r2, -1.0, 1.0)
        steps["r1r2"] = p

        # -- T3 -> r1/r2 (:916), symmetrise (:919), divide (:924), norms, update (:929-931)
        p = _Program(self)
        p.call(lambda: tr.set_t1_dressed_device(self.G, self.Fd, n))
        lo_r1, lo_r2 = tr._lower["r1"], tr._lower["r2"]
        p.extend(tr.t3_to_r1r2_ops())
        p.axpby(r1, Operand(lo_r1, "iaX", (o, v, 1)), "ia", 1.0, 1.0)
        p.axpby(r2, nv(lo_r2, "ijab"), "ijab", 1.0, 1.0)
        r2s = nv(self.r2s, "ijab")
        p.axpby(r2s, r2, "ijab", 1.0, 0.0)
        p.axpby(r2s, nv(self.r2, "jiba"), "ijab", 1.0, 1.0)
        sh = lambda: self.stream.handle if self.stream is not None else None
        p.call(lambda: L.call("dcc_denom_divide_f64", self.r1.vp, o, v, 1, self.eia.vp, sh()))
        p.call(lambda: L.call("dcc_denom_divide_f64", self.r2s.vp, o, v, 2, self.eia.vp, sh()))
        p.call(lambda: L.call("dcc_dot_f64", self.r1.vp, self.r1.vp, self.r1.size, self.partials.vp,
                              self.scal.offset_ptr(0), sh()))
        p.call(lambda: L.call("dcc_dot_f64", self.r2s.vp, self.r2s.vp, self.r2s.size, self.partials.vp,
                              self.scal.offset_ptr(1), sh()))
        p.axpby(t1, r1, "ia", 1.0, 1.0)
        p.axpby(t2v("ijab"), r2s, "ijab", 1.0, 1.0)
        steps["update12"] = p

        # -- T2-dressed triples intermediates into the triples engine (:935)
        p = _Program(self)
        F_oo, F_vv = nv(self.imd["F_oo"], "li"), nv(self.imd["F_vv"], "ad")
        p.axpby(F_oo, gv(Fd, "li"), "li", 1.0, 0.0)
        p.ein("F_oo a", "lmde,imde->li", gv(G, "lmde"), t2v("imde"), F_oo, 2.0, 1.0)
        p.ein("F_oo b", "lmed,imde->li", gv(G, "lmed"), t2v("imde"), F_oo, -1.0, 1.0)
        p.axpby(F_vv, gv(Fd, "ad"), "ad", 1.0, 0.0)
        p.ein("F_vv a", "lmde,lmae->ad", gv(G, "lmde"), t2v("lmae"), F_vv, -2.0, 1.0)
        p.ein("F_vv b", "lmed,lmae->ad", gv(G, "lmed"), t2v("lmae"), F_vv, 1.0, 1.0)
        W_oooo = nv(self.imd["W_oooo"], "lmij")
        p.axpby(W_oooo, gv(G, "lmij"), "lmij", 1.0, 0.0)
        p.ein("W_oooo", "lmde,ijde->lmij", gv(G, "lmde"), t2v("ijde"), W_oooo, 1.0, 1.0)
        W_vvvv = nv(self.imd["W_vvvv"], "abde")
        p.axpby(W_vvvv, gv(G, "abde"), "abde", 1.0, 0.0)
        p.ein("W_vvvv", "lmde,lmab->abde", gv(G, "lmde"), t2v("lmab"), W_vvvv, 1.0, 1.0)
        h = nv(self.h_ovvo, "ladi")
        p.axpby(h, gv(G, "ladi"), "ladi", 1.0, 0.0)
        p.ein("h_ovvo a", "mled,imea->ladi", gv(G, "mled"), t2v("imea"), h, -1.0, 1.0)
        p.ein("h_ovvo b", "mlde,imae->ladi", gv(G, "mlde"), t2v("imae"), h, -1.0, 1.0)
        p.ein("h_ovvo c", "mled,imae->ladi", gv(G, "mled"), t2v("imae"), h, 2.0, 1.0)
        W_ovov = nv(self.imd["W_ovov"], "lbid")
        p.axpby(W_ovov, gv(G, "lbid"), "lbid", 1.0, 0.0)
        p.ein("W_ovov", "mlde,imeb->lbid", gv(G, "mlde"), t2v("imeb"), W_ovov, -1.0, 1.0)
        W_ovvo = nv(self.imd["W_ovvo"], "ladi")
        p.axpby(W_ovvo, h, "ladi", 2.0, 0.0)
        p.axpby(W_ovvo, nv(self.imd["W_ovov"], "laid"), "ladi", -1.0, 1.0)
        Wvvvo = nv(self.imd["W_ovvv"], "jabd")        # W_vvvo[a,b,d,j] stored as [j,a,b,d]
        p.axpby(Wvvvo, gv(G, "abdj"), "abdj", 1.0, 0.0)
        p.ein("W_vvvo a", "amde,jmbe->abdj", gv(G, "amde"), t2v("jmbe"), Wvvvo, 2.0, 1.0)
        p.ein("W_vvvo b", "amde,jmeb->abdj", gv(G, "amde"), t2v("jmeb"), Wvvvo, -1.0, 1.0)
        p.ein("W_vvvo c", "amed,jmbe->abdj", gv(G, "amed"), t2v("jmbe"), Wvvvo, -1.0, 1.0)
        p.ein("W_vvvo d", "bmed,jmea->abdj", gv(G, "bmed"), t2v("jmea"), Wvvvo, -1.0, 1.0)
        p.ein("W_vvvo e", "mndj,mnab->abdj", gv(G, "mndj"), t2v("mnab"), Wvvvo, 1.0, 1.0)
        p.ein("W_vvvo f", "md,mjab->abdj", gv(Fd, "md"), t2v("mjab"), Wvvvo, -1.0, 1.0)
        Wvooo = nv(self.imd["W_oovo"], "ijal")         # W_vooo[a,l,i,j] stored as [i,j,a,l]
        p.axpby(Wvooo, gv(G, "alij"), "alij", 1.0, 0.0)
        p.ein("W_vooo a", "alef,ijef->alij", gv(G, "alef"), t2v("ijef"), Wvooo, 1.0, 1.0)
        p.ein("W_vooo b", "lmje,imae->alij", gv(G, "lmje"), t2v("imae"), Wvooo, 2.0, 1.0)
        p.ein("W_vooo c", "lmje,imea->alij", gv(G, "lmje"), t2v("imea"), Wvooo, -1.0, 1.0)
        p.ein("W_vooo d", "lmej,imae->alij", gv(G, "lmej"), t2v("imae"), Wvooo, -1.0, 1.0)
        p.ein("W_vooo e", "lmei,jmea->alij", gv(G, "lmei"), t2v("jmea"), Wvooo, -1.0, 1.0)
        if self.mixed:   # round the FP64 intermediates and T2 into the FP32 triples engine
            for k, buf in self.imd.items():
                dst = getattr(tr, k)
                letters = "ABCD"[:len(buf.shape)]
                p.add([lambda _s, d=dst, b=buf, l=letters: tensor_axpby(
                    Operand(d, l, d.shape), Operand(b, l, b.shape), l, d.shape, 1.0, 0.0, self.stream)])
            p.add([lambda _s: tensor_axpby(Operand(tr.t2, "ijab", (o, o, v, v)), t2v("ijab"), "ijab",
                                           (o, o, v, v), 1.0, 0.0, self.stream)])
        steps["imds3"] = p
        return steps

    # ----------------------------------------------------------------- public
    def set_amplitudes(self, t1, t2, t3_local):
        s = self.stream
        self.t1.upload(np.ascontiguousarray(t1, dtype=np.float64), s)
        self.t2.upload(np.ascontiguousarray(t2, dtype=np.float64), s)
        self.tr.set_t3(t3_local)

    def get_amplitudes(self):
        s = self.stream
        return self.t1.download(stream=s), self.t2.download(stream=s), self.tr.get_t3()

    def update_amps(self, timer=None, launch_timer=None):
        """One iteration of update_amps_rccsdt_tri_ (rccsdt.py:890-977) on the device.

        Returns [|r1|, |r2|, |r3|] (the res_norm of :929 and :966).  ``timer`` (optional) is
        called as timer(phase, 'start'|'end') around each phase, ``launch_timer`` as
        launch_timer(launch, 'start'|'end') around each contraction of the triples residual, e.g. to
        record HIP events."""
        st = self.stream
        tick = timer if timer is not None else (lambda name, what: None)
        phases = [("dress", lambda: self.prog["dress"].run(st)),
                  ("lower r1/r2", lambda: self.prog["r1r2"].run(st)),
                  ("T3->r1/r2 + update t1/t2", lambda: self.prog["update12"].run(st)),
                  ("triples intermediates", lambda: self.prog["imds3"].run(st)),
                  ("T3->W", lambda: self.tr.t3_to_w(reuse_z=True)),
                  ("ring/oooo operands", self.tr.prepare_residual),
                  ("triples residual", lambda: self.tr.residual(timer=launch_timer))]
        for name, fn in phases:
            tick(name, "start")
            fn()
            tick(name, "end")
        tick("finalize", "start")
        n3, act = self.tr.finalize(self.mo_energy, self.level_shift, self.nvir_diis)
        tick("finalize", "end")
        self.last_t3_norm_sq, self.last_t3_active_norm_sq = n3, act
        sc = self.scal.download(stream=st)
        return [float(np.sqrt(max(sc[0], 0.0))), float(np.sqrt(max(sc[1], 0.0))), float(np.sqrt(max(n3, 0.0)))]

    def energy(self):
        """E_corr = sum (2<ij|ab> - <ij|ba>)(t2 + t1 t1) + 2 sum f_ia t1_ia (rccsdt.py:218-221 closed form)."""
        self.energy_launch()
        return self.read_energy()

    def read_energy(self):
        """The energy of the last energy_launch (downloads two device scalars)."""
        sc = self.scal.download(stream=self.stream)
        return float(sc[2] + 2.0 * sc[3])

    def lower_ops(self):
        """The device work of issue_lower as a list of operations, each a list of steps (callables of
        the stream; a contraction's GemmLaunch first), in issue order."""
        ops = []
        for name in ("dress", "r1r2", "update12", "imds3"):
            ops += self.prog[name].op_steps
        ops.append([lambda _s: self.tr.t3_to_w(reuse_z=True)])
        ops.append([lambda _s: self.tr.prepare_residual()])
        return ops

    def issue_lower(self):
        """Device work of update_amps before the triples residual, with no host synchronisation (so it
        can be captured in a HIP graph): T1 dressing, lower-order residual, T3 -> r1/r2 and the t1/t2
        update, the T2-dressed triples intermediates, T3 -> W, the residual's operands."""
        st = self.stream
        for name in ("dress", "r1r2", "update12", "imds3"):
            self.prog[name].run(st)
        self.tr.t3_to_w(reuse_z=True)
        self.tr.prepare_residual()

    def _build_energy_program(self):
        o, v = self.nocc, self.nvir
        st = self.stream
        p = _Program(self)
        tau = self.nview(self.tau, "ijab")
        p.axpby(tau, Operand(self.t2, "ijab", (o, o, v, v)), "ijab", 1.0, 0.0)
        p.ein("tau", "ia,jb->ijab", self.nview(self.t1, "ia"), self.nview(self.t1, "jb"), tau, 1.0, 1.0)
        sh = lambda: st.handle if st is not None else None
        p.call(lambda: L.call("dcc_dot_f64", self.Loovv.vp, self.tau.vp, self.tau.size, self.partials.vp,
                              self.scal.offset_ptr(2), sh()))
        p.call(lambda: L.call("dcc_dot_f64", self.fov0.vp, self.t1.vp, self.t1.size, self.partials.vp,
                              self.scal.offset_ptr(3), sh()))
        return p

    def energy_launch(self):
        """Issue the energy dot products into device scalars (no host synchronisation)."""
        self._eprog.run(self.stream)

    # ------------------------------------------------------------- DIIS packing
    def diis_size(self, with_t3=True):
        """Length of the unified DIIS vector of run_diis_t3 (rccsdt.py:1038-1049):
        t1, t2 over pairs i <= j, and t3[:, :nd, :nd, :nd]; with_t3=False: the T1/T2 vector of the
        standard DIIS (do_diis_max_t = False, rccsdt.py:1096-1099,1119-1122)."""
        o, v, nd = self.nocc, self.nvir, self.nvir_diis
        return o * v + (o * (o + 1) // 2) * v * v + (self.tr.n3 * nd ** 3 if with_t3 else 0)

    def _copy(self, dst, dst_off, src, src_off, shape, dst_strides, src_strides, alpha=1.0, beta=0.0):
        lay = L.Strided()
        lay.ndim = len(shape)
        for d, (e, sd, ss) in enumerate(zip(shape, dst_strides, src_strides)):
            lay.shape[d], lay.dst_stride[d], lay.src_stride[d] = int(e), int(sd), int(ss)
        sd = "f64" if dst.dtype == np.float64 else "f32"
        ss = "f64" if src.dtype == np.float64 else "f32"
        L.call("dcc_tensor_axpby_" + (sd if sd == ss else sd + "_" + ss), dst.offset_ptr(dst_off),
               src.offset_ptr(src_off), ctypes.byref(lay), float(alpha), float(beta),
               self.stream.handle if self.stream is not None else None)

    def pack(self, x, with_t3=True):
        o, v, nd, n3 = self.nocc, self.nvir, self.nvir_diis, self.tr.n3
        v2 = v * v
        self._copy(x, 0, self.t1, 0, (o * v,), (1,), (1,))
        off = o * v
        for i in range(o):
            cnt = (o - i) * v2
            self._copy(x, off, self.t2, (i * o + i) * v2, (cnt,), (1,), (1,))
            off += cnt
        if with_t3:
            self._copy(x, off, self.tr.t3, 0, (n3, nd, nd, nd), (nd ** 3, nd * nd, nd, 1), (v ** 3, v2, v, 1))

    def unpack(self, x, with_t3=True):
        o, v, nd, n3 = self.nocc, self.nvir, self.nvir_diis, self.tr.n3
        v2 = v * v
        self._copy(self.t1, 0, x, 0, (o * v,), (1,), (1,))
        off = o * v
        for i in range(o):
            cnt = (o - i) * v2
            self._copy(self.t2, (i * o + i) * v2, x, off, (cnt,), (1,), (1,))
            for j in range(i + 1, o):   # t2[j,i,b,a] = t2[i,j,a,b]
                self._copy(self.t2, (j * o + i) * v2, x, off + (j - i) * v2, (v, v), (1, v), (v, 1))
            off += cnt
        if with_t3:
            self._copy(self.tr.t3, 0, x, off, (n3, nd, nd, nd), (v ** 3, v2, v, 1), (nd ** 3, nd * nd, nd, 1))

    def configure(self, level_shift=None, nvir_diis=None):
        """Change the level shift of the denominators and the DIIS virtual subspace of an existing
        engine (the RCCSDT attributes of rccsdt.py:1164-1191 read again at each kernel() call)."""
        o = self.nocc
        if level_shift is not None and float(level_shift) != self.level_shift:
            self.level_shift = float(level_shift)
            eia = self.mo_energy[:o, None] - self.mo_energy[None, o:] - self.level_shift
            self.eia.upload(np.ascontiguousarray(eia), self.stream)
        nd = self.nvir if nvir_diis is None else int(nvir_diis)
        if nd <= 0:
            raise ValueError("nvir_diis must be a positive integer")
        self.nvir_diis = min(nd, self.nvir)

    def init_amps(self):
        """t1 = f_ov / eia, t2 = <ij|ab> / eijab, T3 = 0 (rccsdt.py:197-251), on the device; eia here
        is e_i - e_a without the level shift (rccsdt.py:212-214)."""
        o, v = self.nocc, self.nvir
        st = self.stream
        sh = st.handle if st is not None else None
        tensor_axpby(self.nview(self.t1, "ia"), self.nview(self.fov0, "ia"), "ia", (o, v), 1.0, 0.0, st)
        L.call("dcc_denom_divide_f64", self.t1.vp, o, v, 1, self.eia0.vp, sh)
        tensor_axpby(Operand(self.t2, "ijab", (o, o, v, v)), self.gview(self.G0, "ijab"), "ijab", (o, o, v, v),
                     1.0, 0.0, st)
        L.call("dcc_denom_divide_f64", self.t2.vp, o, v, 2, self.eia0.vp, sh)
        self.tr.t3.fill(0.0, st)
