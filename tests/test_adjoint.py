"""Adjoint of the T2 -> T3 part of the triples residual: central finite differences (h = 1e-5) and
an FP64 reference gradient from torch autograd over a dense restatement of rccsdt.py:388-399."""
import numpy as np
import pytest

import abi_emulator


def _case(o, v, seed):
    rng = np.random.default_rng(seed)
    t2 = rng.standard_normal((o, o, v, v))
    t2 = t2 + t2.transpose(1, 0, 3, 2)
    W1 = rng.standard_normal((v, v, v, o))
    W2 = rng.standard_normal((v, o, o, o))
    lam = rng.standard_normal((o * (o + 1) * (o + 2) // 6, v, v, v))
    return t2, W1, W2, lam


def _r3_oracle(o, v, t2, W1, W2):
    from oracle import rccsdt_tri as tri
    imds = {"F_oo": np.zeros((o, o)), "F_vv": np.zeros((v, v)), "W_oooo": np.zeros((o,) * 4),
            "W_ovvo": np.zeros((o, v, v, o)), "W_ovov": np.zeros((o, v, o, v)), "W_vvvv": np.zeros((v,) * 4),
            "W_vvvo": W1, "W_vooo": W2}
    return tri.compute_r3_tri(o, v, imds, t2, np.zeros((o * (o + 1) * (o + 2) // 6, v, v, v)))


def _torch_reference_grads(o, v, t2, W1, W2, lam):
    """Dense FP64 restatement of the 12 T2->T3 einsums on canonical blocks, differentiated by torch."""
    import torch
    from oracle.t3_layout import canonical_triples
    T2 = torch.tensor(t2, requires_grad=True)
    A = torch.tensor(W1, requires_grad=True)
    B = torch.tensor(W2, requires_grad=True)
    lam_t = torch.tensor(lam)
    loss = torch.zeros((), dtype=torch.float64)
    for n, (i, j, k) in enumerate(canonical_triples(o)):
        r = (torch.einsum("abd,dc->abc", A[..., j], T2[i, k]) + torch.einsum("acd,db->abc", A[..., k], T2[i, j])
             + torch.einsum("bad,dc->abc", A[..., i], T2[j, k]) + torch.einsum("bcd,da->abc", A[..., k], T2[j, i])
             + torch.einsum("cad,db->abc", A[..., i], T2[k, j]) + torch.einsum("cbd,da->abc", A[..., j], T2[k, i])
             - torch.einsum("al,lbc->abc", B[:, :, i, j], T2[:, k]) - torch.einsum("al,lcb->abc", B[:, :, i, k], T2[:, j])
             - torch.einsum("bl,lac->abc", B[:, :, j, i], T2[:, k]) - torch.einsum("bl,lca->abc", B[:, :, j, k], T2[:, i])
             - torch.einsum("cl,lab->abc", B[:, :, k, i], T2[:, j]) - torch.einsum("cl,lba->abc", B[:, :, k, j], T2[:, i]))
        loss = loss + (lam_t[n] * r).sum()
    loss.backward()
    return T2.grad.numpy(), A.grad.numpy(), B.grad.numpy()


def _check(adj, o, v, seed):
    t2, W1, W2, lam = _case(o, v, seed)
    adj.set_inputs(t2, W1, W2)
    R = adj.forward().download()
    assert np.abs(R - _r3_oracle(o, v, t2, W1, W2)).max() < 1e-12 * np.abs(R).max()
    dt2, dW1, dW2 = adj.vjp_host(lam)
    ref = _torch_reference_grads(o, v, t2, W1, W2, lam)
    for got, want in zip((dt2, dW1, dW2), ref):
        assert np.abs(got - want).max() < 1e-12 * np.abs(want).max()
    rng = np.random.default_rng(seed + 1)
    h = 1e-5
    loss = lambda a, b, c: float(np.sum(lam * _r3_oracle(o, v, a, b, c)))
    for which, grad in ((0, dt2), (1, dW1), (2, dW2)):
        args = [t2, W1, W2]
        d = rng.standard_normal(args[which].shape)
        plus = list(args)
        minus = list(args)
        plus[which] = args[which] + h * d
        minus[which] = args[which] - h * d
        fd = (loss(*plus) - loss(*minus)) / (2 * h)
        assert abs(fd - np.sum(grad * d)) < 1e-8 * max(1.0, abs(fd))


def _adjoint(o, v, ic):
    """T2T3Adjoint whose scratch holds ic first occupied indices (None: the default bound)."""
    from distr_cc.adjoint import T2T3Adjoint
    if ic is None:
        return T2T3Adjoint(o, v)
    adj = T2T3Adjoint(o, v, scratch_bytes=ic * o * o * v ** 3 * 8)
    assert adj.scratch.shape[0] == ic and len(adj.chunks) == -(-o // ic)
    return adj


@pytest.mark.parametrize("o,v,ic", [(1, 4, None), (3, 5, None), (4, 3, None), (4, 3, 1), (5, 3, 2)])
def test_adjoint_plan(monkeypatch, o, v, ic):
    abi_emulator.install(monkeypatch)
    _check(_adjoint(o, v, ic), o, v, 10 + o)


def test_autograd_function(monkeypatch):
    abi_emulator.install(monkeypatch)
    import torch
    from distr_cc.adjoint import T2T3Adjoint, make_autograd_function
    o, v = 3, 4
    t2, W1, W2, lam = _case(o, v, 3)
    fn = make_autograd_function()
    adj = T2T3Adjoint(o, v)
    T2 = torch.tensor(t2, requires_grad=True)
    A = torch.tensor(W1, requires_grad=True)
    B = torch.tensor(W2, requires_grad=True)
    r3 = fn.apply(T2, A, B, adj)
    (r3 * torch.tensor(lam)).sum().backward()
    ref = _torch_reference_grads(o, v, t2, W1, W2, lam)
    for got, want in zip((T2.grad, A.grad, B.grad), ref):
        assert np.abs(got.numpy() - want).max() < 1e-12 * np.abs(want).max()


@pytest.mark.gpu
@pytest.mark.parametrize("o,v,ic", [(1, 4, None), (3, 7, None), (6, 13, None), (6, 13, 4), (6, 40, None), (6, 40, 4)])
def test_adjoint_gpu(o, v, ic):
    """Forward against the oracle, VJP against the FP64 torch reference, central differences through the
    oracle; nocc=6, nvir=40 is the adjoint's reference-scale case (one and two scratch chunks)."""
    from distr_cc import device
    device.init(0)
    print("adjoint o=%d v=%d chunk=%s" % (o, v, ic), flush=True)
    _check(_adjoint(o, v, ic), o, v, 20 + o)
