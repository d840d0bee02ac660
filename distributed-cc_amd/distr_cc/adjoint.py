"""Adjoint (vector-Jacobian product) of the T2 -> T3 part of the triples residual.

Forward (distr_cc/rccsdt.py:388-399), per canonical block (i,j,k):

    r3[ijk][abc] = sum_d W_vvvo[a,b,d,j] t2[i,k,d,c] + (5 simultaneous permutations)
                 - sum_l W_vooo[a,l,i,j] t2[l,k,b,c] + (5 simultaneous permutations)

W_vvvo holds the (ov|vv)-type and W_vooo the (oo|ov)-type integral blocks (bare or T2-dressed).
For an adjoint lam (one v^3 block per canonical block) and L = sum lam . r3, write r3 as the
canonical blocks of S(X) with X the single terms above and S the sum over the six simultaneous
permutations.  With Lam the simultaneously symmetric full tensor that equals lam on canonical
blocks and is zero weight elsewhere, L = sum_{ijk canonical} lam . S(X) = 6 sum_{all ijk} Lam_w . X,
where Lam_w[ijk] = lam[canonical(ijk)] symmetrised over the block's stabiliser (degenerate i=j or
j=k blocks), permuted to ijk and divided by the number of distinct orderings of (i,j,k).  Hence every gradient is one dense contraction:

    dW_vvvo[a,b,d,j] =  6 sum_{i,k,c} Lam_w[i,j,k,a,b,c] t2[i,k,d,c]
    dW_vooo[a,l,i,j] = -6 sum_{k,b,c} Lam_w[i,j,k,a,b,c] t2[l,k,b,c]
    dt2[i,k,d,c]     =  6 sum_{j,a,b} Lam_w[i,j,k,a,b,c] W_vvvo[a,b,d,j]
                      - 6 sum_{i',j',a} Lam_w[i',j',i,a,d,c] W_vooo[a,k,i',j']   (relabelled)

Lam_w is rebuilt on the device from lam by S3 gather jobs for every backward call, one chunk of
first occupied indices i at a time, in the same scratch buffer the forward uses for its base term X
(lam is what autograd hands over; no dense T3-sized tensor is stored).  ``T2T3Residual`` wraps
forward and backward in a torch.autograd.Function (PyTorch only as the autograd wrapper; every
contraction is a libdcc gather-GEMM).
"""
import ctypes

import numpy as np

from distr_cc import _lib as L
from distr_cc.device import DeviceArray, init as device_init
from distr_cc.distribute_t3 import TP_T3, canonical_case, ijk_to_linear, nocc3
from distr_cc.rccsdt_engine import plan_einsum_steps, split_k_partial_elems
from distr_cc.t3_engine import Operand, plan_einsum, plan_t2t3, s3_jobs_array, tensor_axpby


def _stabiliser(i, j, k):
    """abc permutations that leave a canonical block with these occupied labels invariant."""
    if i == j == k:
        return list(TP_T3)
    if i == j:
        return [(0, 1, 2), (1, 0, 2)]
    if j == k:
        return [(0, 1, 2), (0, 2, 1)]
    return [(0, 1, 2)]


class T2T3Adjoint:
    """Device buffers and launch plans of the forward T2->T3 residual and its VJP."""

    def __init__(self, nocc, nvir, stream=None, device=0, scratch_bytes=2 << 30):
        """Holds t2, W_vvvo, W_vooo, the canonical r3 blocks, the adjoint lam and the gradients, plus
        ONE scratch buffer of at most ``scratch_bytes`` for a chunk of first occupied indices i: the
        forward's base term X and the backward's Lam_w use it in turn, so no T3-sized dense tensor
        is ever stored -- both are rebuilt chunk by chunk on every call."""
        device_init(device)
        o, v = int(nocc), int(nvir)
        self.nocc, self.nvir, self.stream = o, v, stream
        self.n3 = nocc3(o)
        f64 = np.float64
        self.R = DeviceArray((self.n3, v, v, v), f64)
        self.t2 = DeviceArray((o, o, v, v), f64)
        self.W_ovvv = DeviceArray((o, v, v, v), f64)     # W_vvvo[a,b,d,j] as [j,a,b,d]
        self.W_oovo = DeviceArray((o, o, v, o), f64)     # W_vooo[a,l,i,j] as [i,j,a,l]
        self.scratch, self.fwd = plan_t2t3(o, v, self.R, self.t2, self.W_ovvv, self.W_oovo, scratch_bytes)
        assert self.fwd
        self.lam = DeviceArray((self.n3, v, v, v), f64)
        self.dt2 = DeviceArray((o, o, v, v), f64)
        self.dW_ovvv = DeviceArray((o, v, v, v), f64)    # dW_vvvo[a,b,d,j] as [j,a,b,d]
        self.dW_oovo = DeviceArray((o, o, v, o), f64)    # dW_vooo[a,l,i,j] as [i,j,a,l]
        v3 = v ** 3
        ic = self.scratch.shape[0]
        # split-K partials of the two gradients whose M x N is one chunk of (i, occupied, virtual) rows
        # against one index, with K = o v^2: the peeled occupied index becomes a batch entry writing
        # its own slice, and the slices are summed in a fixed order (dcc_sum_slices)
        self.P_vooo = DeviceArray((o, ic, o, v, o), f64)
        self.P_t2 = DeviceArray((o, ic, o, v, v), f64)
        ext = {x: v for x in "abcdef"}
        ext.update({x: o for x in "ijklmn"})
        one = [({}, {}, {})]
        self.chunks = []
        for i0 in range(0, o, ic):
            n = min(ic, o - i0)
            jobs = []
            for i in range(i0, i0 + n):
                for j in range(o):
                    for k in range(o):
                        case, (ci, cj, ck) = canonical_case(i, j, k)
                        # lam of a degenerate block only enters L through its part symmetric under
                        # the block's stabiliser; symmetrising it makes Lam_w simultaneously symmetric:
                        # Lam_w[ijk] = (1/6) sum_{g in stab} lam[c].transpose(g).transpose(TP_T3[case])
                        coef = np.zeros(6)
                        t = TP_T3[case]
                        for g in _stabiliser(ci, cj, ck):
                            coef[TP_T3.index((g[t[0]], g[t[1]], g[t[2]]))] += 1.0 / 6.0
                        jobs.append(((((i - i0) * o + j) * o + k) * v3, ijk_to_linear(ci, cj, ck, o) * v3, coef))
            ext["i"] = n
            Lw = Operand(self.scratch, "ijkabc", (n, o, o, v, v, v))
            t2_i = Operand(self.t2, "ikdc", (n, o, v, v), base=i0 * o * v * v)
            # every gradient is zeroed first and accumulated over chunks; split-K partials are
            # written with beta 0 and summed into their gradient rows
            split_k = [({"k": k}, {"k": k}, {"S": k}) for k in range(o)]
            split_j = [({"j": j}, {"j": j}, {"S": j}) for j in range(o)]
            launches = [
                plan_einsum("vjp dW_vvvo", "ijkabc,ikdc->jabd", Lw, t2_i,
                            Operand(self.dW_ovvv, "jabd", (o, v, v, v)), 6.0, 1.0, ext, one),
                plan_einsum("vjp dW_vooo", "ijabc,lbc->ijal", Lw, Operand(self.t2, "lkbc", (o, o, v, v)),
                            Operand(self.P_vooo, "Sijal", (o, n, o, v, o)), -6.0, 0.0, ext, split_k),
                ("sum", self.dW_oovo, i0 * o * v * o, self.P_vooo, o, n * o * v * o),
                plan_einsum("vjp dt2 vvvo", "ikabc,abd->ikdc", Lw, Operand(self.W_ovvv, "jabd", (o, v, v, v)),
                            Operand(self.P_t2, "Sikdc", (o, n, o, v, v)), 6.0, 0.0, ext, split_j),
                ("sum", self.dt2, i0 * o * v * v, self.P_t2, o, n * o * v * v),
                plan_einsum("vjp dt2 vooo", "ijkabc,ijal->lkbc", Lw,
                            Operand(self.W_oovo, "ijal", (n, o, v, o), base=i0 * o * v * o),
                            Operand(self.dt2, "lkbc", (o, o, v, v)), -6.0, 1.0, ext, one),
            ]
            self.chunks.append((DeviceArray.from_host(s3_jobs_array(jobs)), len(jobs), launches))

    def _h(self):
        return self.stream.handle if self.stream is not None else None

    def set_inputs(self, t2, W_vvvo, W_vooo):
        """Host arrays: t2 (o,o,v,v), W_vvvo (v,v,v,o), W_vooo (v,o,o,o)."""
        self.t2.upload(np.ascontiguousarray(t2, dtype=np.float64), self.stream)
        self.W_ovvv.upload(np.ascontiguousarray(np.asarray(W_vvvo).transpose(3, 0, 1, 2)), self.stream)
        self.W_oovo.upload(np.ascontiguousarray(np.asarray(W_vooo).transpose(2, 3, 0, 1)), self.stream)

    def forward(self):
        """R <- T2->T3 part of r3 (canonical blocks) for the current inputs (the base-term GEMMs and
        the permutation passes of the triples engine, chunk by chunk)."""
        self.R.fill(0.0, self.stream)
        for g in self.fwd:
            g.run(self.stream)
        return self.R

    def backward(self):
        """dt2, dW_ovvv ([j,a,b,d] layout), dW_oovo ([i,j,a,l] layout) for the adjoint in ``lam``."""
        for grad in (self.dt2, self.dW_ovvv, self.dW_oovo):
            grad.fill(0.0, self.stream)
        for jobs, njobs, launches in self.chunks:
            L.call("dcc_s3_transform_f64", self.scratch.vp, self.lam.vp, self.nvir, 0.0, jobs.vp, njobs,
                   self._h())
            for g in launches:
                if isinstance(g, tuple):     # ('sum', gradient, row offset, partials, slices, count)
                    _, dst, off, part, ns, cnt = g
                    L.call("dcc_sum_slices_f64", dst.offset_ptr(off), part.vp, ns, cnt, 1.0, 1.0, self._h())
                else:
                    g.run(self.stream)
        return self.dt2, self.dW_ovvv, self.dW_oovo

    def vjp_host(self, lam):
        """Convenience: upload lam, run backward, return host (dt2, dW_vvvo, dW_vooo)."""
        self.lam.upload(np.ascontiguousarray(lam, dtype=np.float64), self.stream)
        dt2, dw1, dw2 = self.backward()
        return (dt2.download(stream=self.stream), dw1.download(stream=self.stream).transpose(1, 2, 3, 0),
                dw2.download(stream=self.stream).transpose(2, 3, 0, 1))


def make_autograd_function():
    """torch.autograd.Function over T2T3Adjoint: (t2, W_vvvo, W_vooo) -> r3 canonical blocks.

    Tensors are FP64; they are exchanged with the device engine through host copies, so the
    function works with CPU or GPU torch tensors alike."""
    import torch

    class T2T3Residual(torch.autograd.Function):
        @staticmethod
        def forward(ctx, t2, W_vvvo, W_vooo, engine):
            engine.set_inputs(t2.detach().cpu().numpy(), W_vvvo.detach().cpu().numpy(),
                              W_vooo.detach().cpu().numpy())
            r3 = engine.forward().download(stream=engine.stream)
            ctx.engine = engine
            ctx.device = t2.device
            return torch.from_numpy(r3).to(t2.device)

        @staticmethod
        def backward(ctx, grad_r3):
            eng = ctx.engine
            dt2, dw1, dw2 = eng.vjp_host(grad_r3.detach().cpu().numpy())
            to = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(ctx.device)
            return to(dt2), to(dw1), to(dw2), None

    return T2T3Residual


# ---------------------------------------------------------------------------------------------------
# Adjoint of the whole T3 residual
# ---------------------------------------------------------------------------------------------------

def _grad_view(op, grads):
    """The view ``op`` with its buffer replaced by that buffer's gradient (None if not differentiated)."""
    g = grads.get(id(op.buf))
    if g is None:
        return None
    out = Operand.__new__(Operand)
    out.buf, out.letters, out.shape, out.strides, out.base = g, op.letters, op.shape, dict(op.strides), op.base
    out.ext = getattr(op, "ext", {})
    return out


class TriplesResidualAdjoint:
    """The triples residual r3(t2, g) of one RCCSDT iteration at T1 = 0 and its vector-Jacobian product.

    r3 is computed as update_amps_rccsdt_tri_ does it (distr_cc/rccsdt.py:935-941): the T2-dressed
    intermediates of intermediates_t3 (F_oo, F_vv, W_oooo, W_vvvv, W_ovvo, W_ovov, W_vvvo, W_vooo), the T3
    terms added to W_vooo/W_vvvo (rccsdt.py:323-366), then compute_r3_tri (T2->T3, F_vv, W_vvvv, ring,
    W_oooo/F_oo terms) on the canonical T3 blocks, which are held fixed.  The Fock matrix is held fixed
    too (at T1 = 0 the dressed Fock equals it), so r3 is a function of t2 and the integrals <pq|rs>.

    For an adjoint lam of r3 (canonical blocks), L = sum lam . r3 = sum over ALL ordered (i,j,k) of
    Lam . R, with R the full simultaneously symmetric residual and Lam the weighted adjoint of the
    T2->T3 adjoint above (lam symmetrised over each block's stabiliser, divided among the orderings).
    backward() builds Lam and the full T3 from the canonical blocks on the device (S3 gathers, every
    call: no T3-sized tensor is kept between calls), then
      1. reverses every term of the residual into gradients of the intermediates and of t2 (each term
         is one gather-GEMM; the three simultaneous images of a term contribute equally because Lam
         and T are symmetric, so each is evaluated once and scaled);
      2. reverses the device program of the T2-dressed intermediates operation by operation (for
         C = beta C + alpha A.B: dA += alpha dC.B, dB += alpha A.dC, dC *= beta; copies likewise) into
         gradients of t2 and of <pq|rs>.
    The integral gradient is reported for the (ov|vv)- and (oo|ov)-type entries of <pq|rs> (exactly one,
    or exactly three, occupied indices); the others are held constant (the T3 -> W terms depend only
    on the <oo|vv> block).
    """

    def __init__(self, nocc, nvir, fock, pppp, t3_local, stream=None, device=0):
        from distr_cc.rccsdt_engine import RCCSDTEngine
        device_init(device)
        o, v = int(nocc), int(nvir)
        n = o + v
        self.nocc, self.nvir, self.nmo, self.stream = o, v, n, stream
        fock = np.asarray(fock, dtype=np.float64)
        self.eng = eng = RCCSDTEngine(o, v, fock, pppp, np.diag(fock).copy(), stream=stream, vvvv_pairs=False)
        self.tr = tr = eng.tr
        self.n3 = tr.n3
        tr.set_t3(t3_local)
        f64 = np.float64
        v3 = v ** 3
        self.lam = DeviceArray((self.n3, v, v, v), f64)
        self.Lam = DeviceArray((o, o, o, v, v, v), f64)
        self.T = DeviceArray((o, o, o, v, v, v), f64)
        self.dG = DeviceArray((n, n, n, n), f64)
        self.dt2 = DeviceArray((o, o, v, v), f64)
        self.dh = DeviceArray(eng.h_ovvo.shape, f64)
        self.d_imd = {k: DeviceArray(b.shape, f64) for k, b in eng.imd.items()}
        # S3 jobs: Lam[ijk] from lam (stabiliser-symmetrised, weighted) and T[ijk] = the canonical block
        lam_jobs, t_jobs = [], []
        for i in range(o):
            for j in range(o):
                for k in range(o):
                    case, (ci, cj, ck) = canonical_case(i, j, k)
                    dst, src = ((i * o + j) * o + k) * v3, ijk_to_linear(ci, cj, ck, o) * v3
                    coef = np.zeros(6)
                    tp = TP_T3[case]
                    for g in _stabiliser(ci, cj, ck):
                        coef[TP_T3.index((g[tp[0]], g[tp[1]], g[tp[2]]))] += 1.0 / 6.0
                    lam_jobs.append((dst, src, coef))
                    take = np.zeros(6)
                    take[case] = 1.0
                    t_jobs.append((dst, src, take))
        self.lam_jobs = DeviceArray.from_host(s3_jobs_array(lam_jobs))
        self.t_jobs = DeviceArray.from_host(s3_jobs_array(t_jobs))
        self.njobs = len(lam_jobs)
        self.keep = []
        self._plan_residual_adjoint()
        self._plan_program_adjoint()

    # ------------------------------------------------------------------ planning
    def _plan_residual_adjoint(self):
        o, v = self.nocc, self.nvir
        ext = {x: v for x in "abcdef"}
        ext.update({x: o for x in "ijklm"})
        six = (o, o, o, v, v, v)
        lam = lambda s: Operand(self.Lam, s, six)
        T = lambda s: Operand(self.T, s, six)
        d = self.d_imd
        dF_vv, dF_oo = Operand(d["F_vv"], "ad", (v, v)), Operand(d["F_oo"], "li", (o, o))
        dW_vvvv = Operand(d["W_vvvv"], "abde", (v, v, v, v))
        dW_oooo = Operand(d["W_oooo"], "lmij", (o, o, o, o))
        dW_ovov = lambda s: Operand(d["W_ovov"], s, (o, v, o, v))
        dW_ovvo = Operand(d["W_ovvo"], "ladi", (o, v, v, o))
        dW_ovvv = Operand(d["W_ovvv"], "jabd", (o, v, v, v))     # W_vvvo[a,b,d,j] as [j,a,b,d]
        dW_oovo = Operand(d["W_oovo"], "ijal", (o, o, v, o))     # W_vooo[a,l,i,j] as [i,j,a,l]
        W_ovvv = Operand(self.tr.W_ovvv, "jabd", (o, v, v, v))
        W_oovo = Operand(self.tr.W_oovo, "ijal", (o, o, v, o))
        t2 = lambda s: Operand(self.tr.t2, s, (o, o, v, v))
        dt2 = lambda s: Operand(self.dt2, s, (o, o, v, v))
        terms = []
        P = lambda name, sub, A, B, C, alpha: terms.append(("adj r3 " + name, sub, A, B, C, alpha))
        [
            # T2 -> T3: the twelve terms are the six simultaneous images of
            # X = W_vvvo[a,b,d,j] t2[i,k,d,c] - W_vooo[a,l,i,j] t2[l,k,b,c]   (rccsdt.py:388-399)
            P("dW_vvvo", "ijkabc,ikdc->jabd", lam("ijkabc"), t2("ikdc"), dW_ovvv, 6.0),
            P("dW_vooo", "ijkabc,lkbc->ijal", lam("ijkabc"), t2("lkbc"), dW_oovo, -6.0),
            P("dt2 vvvo", "ijkabc,jabd->ikdc", lam("ijkabc"), W_ovvv, dt2("ikdc"), 6.0),
            P("dt2 vooo", "ijkabc,ijal->lkbc", lam("ijkabc"), W_oovo, dt2("lkbc"), -6.0),
            # F_vv T3 and W_vvvv T3 (rccsdt.py:400-402, 421-427)
            P("dF_vv", "ijkabc,ijkdbc->ad", lam("ijkabc"), T("ijkdbc"), dF_vv, 3.0),
            P("dW_vvvv", "ijkabc,ijkdec->abde", lam("ijkabc"), T("ijkdec"), dW_vvvv, 3.0),
            # W_oooo T3 - F_oo T3 (rccsdt.py:616-732)
            P("dW_oooo", "lmkabc,ijkabc->lmij", T("lmkabc"), lam("ijkabc"), dW_oooo, 3.0),
            P("dF_oo", "ljkabc,ijkabc->li", T("ljkabc"), lam("ijkabc"), dF_oo, -3.0),
            # ring terms: r3 += G + G[jik,bac] + G[kji,cba] with (rccsdt.py:540-596)
            # G = - W_ovov[lbid] T[ljk,adc] - W_ovov[lcid] T[ljk,abd] - 1/2 W_ovov[laid] (T[ljk,bdc] + T[ljk,cbd])
            #     + 1/2 W_ovvo[ladi] (2 T[ljk,dbc] - T[ljk,cbd] - T[ljk,bdc])
            P("dW_ovov 1", "ljkadc,ijkabc->lbid", T("ljkadc"), lam("ijkabc"), dW_ovov("lbid"), -3.0),
            P("dW_ovov 2", "ljkabd,ijkabc->lcid", T("ljkabd"), lam("ijkabc"), dW_ovov("lcid"), -3.0),
            P("dW_ovov 3", "ljkbdc,ijkabc->laid", T("ljkbdc"), lam("ijkabc"), dW_ovov("laid"), -1.5),
            P("dW_ovov 4", "ljkcbd,ijkabc->laid", T("ljkcbd"), lam("ijkabc"), dW_ovov("laid"), -1.5),
            P("dW_ovvo 1", "ljkdbc,ijkabc->ladi", T("ljkdbc"), lam("ijkabc"), dW_ovvo, 3.0),
            P("dW_ovvo 2", "ljkcbd,ijkabc->ladi", T("ljkcbd"), lam("ijkabc"), dW_ovvo, -1.5),
            P("dW_ovvo 3", "ljkbdc,ijkabc->ladi", T("ljkbdc"), lam("ijkabc"), dW_ovvo, -1.5),
        ]
        # reductions whose output fills few tiles against a long K (dW_oooo, dF_vv, dF_oo, the ring
        # terms, dt2) are split along one K letter; their partial slices share one scratch buffer
        split = dict(tile=128, min_tiles=256)
        need = max(split_k_partial_elems(sub, A, B, C, ext, **split) for _, sub, A, B, C, _ in terms)
        self.split_scratch = DeviceArray((max(need, 1),))
        self.res_adj = []
        for name, sub, A, B, C, alpha in terms:
            steps, keep = plan_einsum_steps(name, sub, A, B, C, alpha, 1.0, ext, self.stream,
                                            scratch=self.split_scratch, **split)
            self.keep += keep
            self.res_adj.append((name, steps))

    def _plan_program_adjoint(self):
        """Reverse of the intermediates program (RCCSDTEngine 'imds3'), as planned launches."""
        eng = self.eng
        grads = {id(eng.G): self.dG, id(eng.t2): self.dt2, id(eng.h_ovvo): self.dh}
        grads.update({id(b): self.d_imd[k] for k, b in eng.imd.items()})
        ext = eng.ext
        st = self.stream
        steps = []
        for op in reversed(eng.prog["imds3"].ops):
            kind = op[0]
            if kind == "ein":
                _, name, sub, A, B, C, alpha, beta = op
                sa, rest = sub.split(",")
                sb, sc = rest.split("->")
                gC = _grad_view(C, grads)
                gA, gB = _grad_view(A, grads), _grad_view(B, grads)
                for nm, grad, rsub, X, Y in (("A", gA, "%s,%s->%s" % (sc, sb, sa), gC, B),
                                             ("B", gB, "%s,%s->%s" % (sa, sc, sb), A, gC)):
                    if grad is None:
                        continue
                    ps, keep = plan_einsum_steps("adj %s %s" % (name, nm), rsub, X, Y, grad, alpha, 1.0, ext, st,
                                                 tile=128, min_tiles=256, scratch=self.split_scratch)
                    self.keep += keep
                    steps += [(getattr(ps[0], "name", name), p) for p in ps]
                if beta != 1.0:
                    steps.append(("adj elementwise", self._scale_step(gC, sc, beta)))
            elif kind == "axpby":
                _, dst, src, letters, alpha, beta = op
                gd, gs = _grad_view(dst, grads), _grad_view(src, grads)
                shape = tuple(ext[x] for x in letters)
                if gs is not None:
                    if any(x not in src.strides for x in letters):
                        raise ValueError("broadcast copy in a differentiated program")
                    steps.append(("adj elementwise", lambda _s, gs=gs, gd=gd, l=letters, sh=shape, a=alpha:
                                  tensor_axpby(gs, gd, l, sh, a, 1.0, st)))
                if beta != 1.0:
                    steps.append(("adj elementwise", self._scale_step(gd, letters, beta)))
            else:
                raise ValueError("the intermediates program has a step that cannot be differentiated")
        self.prog_adj = steps

    def _scale_step(self, g, letters, beta):
        shape = tuple(self.eng.ext[x] for x in letters)
        st = self.stream
        if beta == 0.0:
            return lambda _s: tensor_axpby(g, g, letters, shape, 0.0, 0.0, st)
        return lambda _s: tensor_axpby(g, g, letters, shape, 0.0, float(beta), st)

    # ------------------------------------------------------------------ evaluation
    def set_inputs(self, t2, pppp=None):
        """t2 (o,o,v,v) and, optionally, new integrals <pq|rs> (nmo^4)."""
        if pppp is not None:
            self.eng.set_integrals(pppp)
        self.eng.set_amplitudes(np.zeros((self.nocc, self.nvir)), t2, self.tr.get_t3())

    def forward(self):
        """r3 (canonical blocks, before projection) for the current inputs; returns the device R."""
        eng, tr, st = self.eng, self.tr, self.stream
        eng.prog["dress"].run(st)
        eng.prog["imds3"].run(st)
        tr.set_t1_dressed_device(eng.G, eng.Fd, self.nmo)
        tr.t3_to_w(reuse_z=False)
        tr.prepare_residual()
        tr.residual()
        return tr.R

    def backward(self, timer=None):
        """(dt2, dG) for the adjoint in ``lam``; call after forward() on the same inputs (the
        intermediates of the forward are read).  ``timer(name, "start"|"end")``, if given, is called
        around every launch (to record events on the stream)."""
        h = self.stream.handle if self.stream is not None else None
        st = self.stream
        tm = timer if timer is not None else (lambda name, what: None)
        for buf in [self.dG, self.dt2, self.dh] + list(self.d_imd.values()):
            buf.fill(0.0, st)
        tm("dense Lambda and T3", "start")
        L.call("dcc_s3_transform_f64", self.Lam.vp, self.lam.vp, self.nvir, 0.0, self.lam_jobs.vp, self.njobs, h)
        L.call("dcc_s3_transform_f64", self.T.vp, self.tr.t3.vp, self.nvir, 0.0, self.t_jobs.vp, self.njobs, h)
        tm("dense Lambda and T3", "end")
        for name, steps in self.res_adj:
            tm(name, "start")
            for step in steps:
                step(st)
            tm(name, "end")
        for name, step in self.prog_adj:
            tm(name, "start")
            step(st)
            tm(name, "end")
        return self.dt2, self.dG

    def integral_mask(self):
        """Boolean (nmo,)*4 mask of the (ov|vv)- and (oo|ov)-type entries of <pq|rs>."""
        n, o = self.nmo, self.nocc
        occ = (np.arange(n) < o).astype(np.int8)
        cnt = occ[:, None, None, None] + occ[None, :, None, None] + occ[None, None, :, None] + occ[None, None, None, :]
        return (cnt == 1) | (cnt == 3)

    def vjp_host(self, lam):
        """Upload lam, run backward, return host (dt2, d<pq|rs>) as gradients on the symmetric inputs:
        t2 with t2[i,j,a,b] = t2[j,i,b,a] and real integrals with their 8-fold permutational symmetry
        (the program reads some entries through an equivalent permuted one, so only these projections
        are unique); d<pq|rs> is reported on the (ov|vv)/(oo|ov) entries and 0 elsewhere."""
        self.lam.upload(np.ascontiguousarray(lam, dtype=np.float64), self.stream)
        dt2, dG = self.backward()
        t = dt2.download(stream=self.stream)
        g = dG.download(stream=self.stream)
        return symmetrize_t2(t), symmetrize_eri(g) * self.integral_mask()


# the 8 index permutations of <pq|rs> = (pr|qs) that leave real integrals invariant
ERI_PERMS = ((0, 1, 2, 3), (2, 1, 0, 3), (0, 3, 2, 1), (2, 3, 0, 1), (1, 0, 3, 2), (3, 0, 1, 2), (1, 2, 3, 0),
             (3, 2, 1, 0))


def symmetrize_t2(g):
    return 0.5 * (g + g.transpose(1, 0, 3, 2))


def symmetrize_eri(g):
    return sum(g.transpose(p) for p in ERI_PERMS) / 8.0


def make_residual_autograd_function():
    """torch.autograd.Function over TriplesResidualAdjoint: (t2, pppp, engine) -> r3 canonical blocks,
    differentiable in t2 and in the (ov|vv)/(oo|ov) entries of pppp.  FP64 tensors are exchanged with
    the device engine through host copies (CPU or GPU torch tensors alike)."""
    import torch

    class TriplesResidual(torch.autograd.Function):
        @staticmethod
        def forward(ctx, t2, pppp, engine):
            engine.set_inputs(t2.detach().cpu().numpy(), pppp.detach().cpu().numpy())
            r3 = engine.forward().download(stream=engine.stream)
            ctx.engine, ctx.device = engine, t2.device
            return torch.from_numpy(r3).to(t2.device)

        @staticmethod
        def backward(ctx, grad_r3):
            dt2, dg = ctx.engine.vjp_host(grad_r3.detach().cpu().numpy())
            to = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(ctx.device)
            return to(dt2), to(dg), None

    return TriplesResidual
