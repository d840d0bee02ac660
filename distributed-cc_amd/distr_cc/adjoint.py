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
from distr_cc.t3_engine import Operand, plan_einsum, plan_t2t3, s3_jobs_array


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
