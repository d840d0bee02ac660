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

Lam_w is rebuilt on the device from lam by S3 gather jobs for every backward call (lam is what
autograd hands over; the dense form is never kept between calls).  ``T2T3Residual`` wraps
forward and backward in a torch.autograd.Function (PyTorch only as the autograd wrapper; every
contraction is a libdcc gather-GEMM).
"""
import ctypes

import numpy as np

from distr_cc import _lib as L
from distr_cc.device import DeviceArray, init as device_init
from distr_cc.distribute_t3 import TP_T3, canonical_case, ijk_to_linear, nocc3
from distr_cc.t3_engine import Operand, TriplesEngine, plan_einsum, s3_jobs_array


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

    def __init__(self, nocc, nvir, stream=None, device=0):
        device_init(device)
        o, v = int(nocc), int(nvir)
        self.nocc, self.nvir, self.stream = o, v, stream
        self.n3 = nocc3(o)
        self.eng = TriplesEngine(o, v, stream=stream, device=device)
        self.fwd = [g for g in self.eng.launches if g.name.startswith("T2T3:")]
        assert self.fwd
        f64 = np.float64
        self.lam = DeviceArray((self.n3, v, v, v), f64)
        self.lam_w = DeviceArray((o, o, o, v, v, v), f64)
        self.dt2 = DeviceArray((o, o, v, v), f64)
        self.dW_ovvv = DeviceArray((o, v, v, v), f64)    # dW_vvvo[a,b,d,j] as [j,a,b,d]
        self.dW_oovo = DeviceArray((o, o, v, o), f64)    # dW_vooo[a,l,i,j] as [i,j,a,l]
        v3 = v ** 3
        jobs = []
        for i in range(o):
            for j in range(o):
                for k in range(o):
                    case, (ci, cj, ck) = canonical_case(i, j, k)
                    # lam of a degenerate block only enters L through its part symmetric under the
                    # block's stabiliser; symmetrising it makes Lam_w simultaneously symmetric:
                    # Lam_w[ijk] = (1/6) sum_{g in stab} lam[c].transpose(g).transpose(TP_T3[case])
                    coef = np.zeros(6)
                    t = TP_T3[case]
                    for g in _stabiliser(ci, cj, ck):
                        coef[TP_T3.index((g[t[0]], g[t[1]], g[t[2]]))] += 1.0 / 6.0
                    jobs.append((((i * o + j) * o + k) * v3, ijk_to_linear(ci, cj, ck, o) * v3, coef))
        self.jobs = DeviceArray.from_host(s3_jobs_array(jobs))
        self.njobs = len(jobs)
        ext = {x: v for x in "abcdef"}
        ext.update({x: o for x in "ijklmn"})
        Lw = lambda letters: Operand(self.lam_w, letters, (o, o, o, v, v, v))
        t2 = lambda letters: Operand(self.eng.t2, letters, (o, o, v, v))
        one = [({}, {}, {})]
        self.bwd = [
            plan_einsum("vjp dW_vvvo", "ijkabc,ikdc->jabd", Lw("ijkabc"), t2("ikdc"),
                        Operand(self.dW_ovvv, "jabd", (o, v, v, v)), 6.0, 0.0, ext, one),
            plan_einsum("vjp dW_vooo", "ijkabc,lkbc->ijal", Lw("ijkabc"), t2("lkbc"),
                        Operand(self.dW_oovo, "ijal", (o, o, v, o)), -6.0, 0.0, ext, one),
            plan_einsum("vjp dt2 vvvo", "ijkabc,jabd->ikdc", Lw("ijkabc"),
                        Operand(self.eng.W_ovvv, "jabd", (o, v, v, v)),
                        Operand(self.dt2, "ikdc", (o, o, v, v)), 6.0, 0.0, ext, one),
            plan_einsum("vjp dt2 vooo", "ijkabc,ijal->lkbc", Lw("ijkabc"),
                        Operand(self.eng.W_oovo, "ijal", (o, o, v, o)),
                        Operand(self.dt2, "lkbc", (o, o, v, v)), -6.0, 1.0, ext, one),
        ]

    def _h(self):
        return self.stream.handle if self.stream is not None else None

    def set_inputs(self, t2, W_vvvo, W_vooo):
        """Host arrays: t2 (o,o,v,v), W_vvvo (v,v,v,o), W_vooo (v,o,o,o)."""
        self.eng.t2.upload(np.ascontiguousarray(t2, dtype=np.float64), self.stream)
        self.eng.W_ovvv.upload(np.ascontiguousarray(np.asarray(W_vvvo).transpose(3, 0, 1, 2)), self.stream)
        self.eng.W_oovo.upload(np.ascontiguousarray(np.asarray(W_vooo).transpose(2, 3, 0, 1)), self.stream)

    def forward(self):
        """R <- T2->T3 part of r3 (canonical blocks) for the current inputs (the engine's T2->T3
        launches: the base-term GEMMs and the permutation passes)."""
        self.eng.R.fill(0.0, self.stream)
        for g in self.fwd:
            g.run(self.stream)
        return self.eng.R

    def backward(self):
        """dt2, dW_ovvv ([j,a,b,d] layout), dW_oovo ([i,j,a,l] layout) for the adjoint in ``lam``."""
        L.call("dcc_s3_transform_f64", self.lam_w.vp, self.lam.vp, self.nvir, 0.0, self.jobs.vp, self.njobs,
               self._h())
        for g in self.bwd:
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
