"""Adjoints of the triples residual, each checked against central finite differences (h = 1e-5)
through the oracle and an FP64 reference gradient from torch autograd:

* T2T3Adjoint: the T2 -> T3 part (rccsdt.py:388-399) w.r.t. t2, W_vvvo and W_vooo as inputs;
* TriplesResidualAdjoint: the whole residual of one iteration at T1 = 0 (T2-dressed intermediates,
  T3 -> W terms, compute_r3_tri) w.r.t. t2 and the (ov|vv)/(oo|ov) integrals, against torch
  autograd over a dense restatement of oracle/rccsdt_lower.py intermediates_t3 and
  oracle/rccsdt_dense.py (whose forward is checked against the oracle's block loop)."""
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


# ----------------------------------------------------------------------------------------------------
# Adjoint of the whole T3 residual r3(t2, <pq|rs>) at T1 = 0 (T3 and the Fock matrix held fixed)
# ----------------------------------------------------------------------------------------------------

def _residual_case(o, v, seed):
    from distr_cc.synthetic import SyntheticSystem
    s = SyntheticSystem(o, v, naux=40, seed=seed)
    rng = np.random.default_rng(seed + 5)
    t3 = s.t3_tri() * 50.0
    lam = rng.standard_normal(t3.shape)
    return s.fock, s.pppp(), s.t2(), t3, lam


def _r3_oracle_full(o, v, fock, pppp, t2, t3):
    """oracle: T2-dressed intermediates at T1 = 0, the T3 terms of W_vooo/W_vvvo, compute_r3_tri."""
    from oracle import rccsdt_lower as lo, rccsdt_tri as tri
    imds = lo.intermediates_t3(o, fock, pppp, t2)
    tri.intermediates_t3_add_t3_tri(o, v, pppp, t3, imds["W_vooo"], imds["W_vvvo"])
    return tri.compute_r3_tri(o, v, imds, t2, t3)


def _torch_residual_grads(o, v, fock, pppp, t2, t3, lam, mask):
    """FP64 reference gradient: torch autograd over a dense restatement of intermediates_t3
    (oracle/rccsdt_lower.py), the T3 terms of W (oracle/rccsdt_dense.py w_t3_dense) and the dense r3
    (oracle/rccsdt_dense.py r3_dense), restricted to the canonical blocks.  Its forward is checked
    against the oracle's block loop first."""
    import torch
    from oracle.t3_layout import canonical_triples, full_t3_from_tri
    E = torch.einsum
    g = torch.tensor(pppp, requires_grad=True)
    t = torch.tensor(t2, requires_grad=True)
    f = torch.tensor(fock)
    T = torch.tensor(full_t3_from_tri(t3, o))
    O, V = slice(0, o), slice(o, o + v)
    oovv = g[O, O, V, V]
    F_oo = f[O, O] + 2.0 * E("lmde,imde->li", oovv, t) - E("lmed,imde->li", oovv, t)
    F_vv = f[V, V] - 2.0 * E("lmde,lmae->ad", oovv, t) + E("lmed,lmae->ad", oovv, t)
    W_oooo = g[O, O, O, O] + E("lmde,ijde->lmij", oovv, t)
    W_vvvv = g[V, V, V, V] + E("lmde,lmab->abde", oovv, t)
    h_ovvo = (g[O, V, V, O] - E("mled,imea->ladi", oovv, t) - E("mlde,imae->ladi", oovv, t)
              + 2.0 * E("mled,imae->ladi", oovv, t))
    W_ovov = g[O, V, O, V] - E("mlde,imeb->lbid", oovv, t)
    W_ovvo = 2.0 * h_ovvo - W_ovov.permute(0, 1, 3, 2)
    gvovv, gooov, goovo = g[V, O, V, V], g[O, O, O, V], g[O, O, V, O]
    W_vvvo = (g[V, V, V, O] + 2.0 * E("amde,jmbe->abdj", gvovv, t) - E("amde,jmeb->abdj", gvovv, t)
              - E("amed,jmbe->abdj", gvovv, t) - E("bmed,jmea->abdj", gvovv, t)
              + E("mndj,mnab->abdj", goovo, t) - E("md,mjab->abdj", f[O, V], t))
    W_vooo = (g[V, O, O, O] + E("alef,ijef->alij", gvovv, t) + 2.0 * E("lmje,imae->alij", gooov, t)
              - E("lmje,imea->alij", gooov, t) - E("lmej,imae->alij", goovo, t) - E("lmei,jmea->alij", goovo, t))
    # T3 -> W (P3_201 of every ordered block: 2T - T_cba - T_bac on the virtual indices)
    Z = 2.0 * T - T.permute(0, 1, 2, 5, 4, 3) - T.permute(0, 1, 2, 4, 3, 5)
    W_vooo = W_vooo + E("mpde,pqread->amqr", oovv, Z)
    W_vvvo = W_vvvo - E("rpde,pqreba->abdq", oovv, Z)
    R = E("abdj,ikdc->ijkabc", W_vvvo, t) + E("acdk,ijdb->ijkabc", W_vvvo, t)
    R = R + E("badi,jkdc->ijkabc", W_vvvo, t) + E("bcdk,jida->ijkabc", W_vvvo, t)
    R = R + E("cadi,kjdb->ijkabc", W_vvvo, t) + E("cbdj,kida->ijkabc", W_vvvo, t)
    R = R - E("alij,lkbc->ijkabc", W_vooo, t) - E("alik,ljcb->ijkabc", W_vooo, t)
    R = R - E("blji,lkac->ijkabc", W_vooo, t) - E("bljk,lica->ijkabc", W_vooo, t)
    R = R - E("clki,ljab->ijkabc", W_vooo, t) - E("clkj,liba->ijkabc", W_vooo, t)
    R = R + E("ad,ijkdbc->ijkabc", F_vv, T) + E("bd,ijkadc->ijkabc", F_vv, T) + E("cd,ijkabd->ijkabc", F_vv, T)
    R = R + E("abde,ijkdec->ijkabc", W_vvvv, T) + E("acde,ijkdbe->ijkabc", W_vvvv, T)
    R = R + E("bcde,ijkade->ijkabc", W_vvvv, T)
    R = R + E("lmij,lmkabc->ijkabc", W_oooo, T) + E("lmik,lmjacb->ijkabc", W_oooo, T)
    R = R + E("lmjk,lmibca->ijkabc", W_oooo, T)
    R = R - E("li,ljkabc->ijkabc", F_oo, T) - E("lj,ilkabc->ijkabc", F_oo, T) - E("lk,ijlabc->ijkabc", F_oo, T)
    G = (-E("lbid,ljkadc->ijkabc", W_ovov, T) - E("lcid,ljkabd->ijkabc", W_ovov, T)
         - 0.5 * E("laid,ljkbdc->ijkabc", W_ovov, T) - 0.5 * E("laid,ljkcbd->ijkabc", W_ovov, T)
         + E("ladi,ljkdbc->ijkabc", W_ovvo, T) - 0.5 * E("ladi,ljkcbd->ijkabc", W_ovvo, T)
         - 0.5 * E("ladi,ljkbdc->ijkabc", W_ovvo, T))
    R = R + G + G.permute(1, 0, 2, 4, 3, 5) + G.permute(2, 1, 0, 5, 4, 3)
    rc = torch.stack([R[i, j, k] for (i, j, k) in canonical_triples(o)])
    (rc * torch.tensor(lam)).sum().backward()
    from distr_cc.adjoint import symmetrize_eri, symmetrize_t2
    return rc.detach().numpy(), symmetrize_t2(t.grad.numpy()), symmetrize_eri(g.grad.numpy()) * mask


def _check_residual_adjoint(adj, o, v, fock, pppp, t2, t3, lam, fd_dirs=2, tol=1e-11):
    adj.set_inputs(t2, pppp)
    R = adj.forward().download()
    Ro = _r3_oracle_full(o, v, fock, pppp, t2, t3)
    assert np.abs(R - Ro).max() <= 1e-12 * np.abs(Ro).max()
    dt2, dg = adj.vjp_host(lam)
    mask = adj.integral_mask()
    rt, rdt2, rdg = _torch_residual_grads(o, v, fock, pppp, t2, t3, lam, mask)
    assert np.abs(rt - Ro).max() <= 1e-12 * np.abs(Ro).max()      # the torch restatement is the oracle
    assert np.abs(dt2 - rdt2).max() <= tol * np.abs(rdt2).max()
    assert np.abs(dg - rdg).max() <= tol * np.abs(rdg).max()
    # central differences, h = 1e-5, along random directions of t2 and of the (ov|vv)/(oo|ov) integrals
    rng = np.random.default_rng(7)
    h = 1e-5
    loss = lambda a, b: float(np.sum(lam * _r3_oracle_full(o, v, fock, b, a, t3)))
    for _ in range(fd_dirs):
        # a unit direction: r3 is cubic in t2 and the integrals, so the central difference carries an
        # h^2 / 6 third-derivative term that a unit direction keeps below the bound
        from distr_cc.adjoint import symmetrize_eri
        d2 = rng.standard_normal(t2.shape)
        d2 = d2 + d2.transpose(1, 0, 3, 2)
        dgd = symmetrize_eri(rng.standard_normal(pppp.shape)) * mask
        nrm = np.sqrt(np.sum(d2 * d2) + np.sum(dgd * dgd))
        d2, dgd = d2 / nrm, dgd / nrm
        fd = (loss(t2 + h * d2, pppp + h * dgd) - loss(t2 - h * d2, pppp - h * dgd)) / (2 * h)
        an = float(np.sum(dt2 * d2) + np.sum(dg * dgd))
        assert abs(fd - an) <= 1e-8 * max(1.0, abs(an)), (fd, an)


@pytest.mark.parametrize("o,v", [(1, 3), (2, 4), (3, 4)])
def test_residual_adjoint_plan(monkeypatch, o, v):
    abi_emulator.install(monkeypatch)
    from distr_cc.adjoint import TriplesResidualAdjoint
    fock, pppp, t2, t3, lam = _residual_case(o, v, 30 + o)
    adj = TriplesResidualAdjoint(o, v, fock, pppp, t3)
    _check_residual_adjoint(adj, o, v, fock, pppp, t2, t3, lam)


def test_residual_autograd_function(monkeypatch):
    abi_emulator.install(monkeypatch)
    import torch
    from distr_cc.adjoint import TriplesResidualAdjoint, make_residual_autograd_function
    o, v = 2, 4
    fock, pppp, t2, t3, lam = _residual_case(o, v, 41)
    adj = TriplesResidualAdjoint(o, v, fock, pppp, t3)
    fn = make_residual_autograd_function()
    T2 = torch.tensor(t2, requires_grad=True)
    G = torch.tensor(pppp, requires_grad=True)
    (fn.apply(T2, G, adj) * torch.tensor(lam)).sum().backward()
    _, rdt2, rdg = _torch_residual_grads(o, v, fock, pppp, t2, t3, lam, adj.integral_mask())
    assert np.abs(T2.grad.numpy() - rdt2).max() <= 1e-11 * np.abs(rdt2).max()
    assert np.abs(G.grad.numpy() - rdg).max() <= 1e-11 * np.abs(rdg).max()


@pytest.mark.gpu
@pytest.mark.parametrize("o,v,dirs", [(3, 7, 2), (6, 40, 1)])
def test_residual_adjoint_gpu(o, v, dirs):
    """The whole-residual VJP on the MI355X against torch autograd of the dense FP64 restatement and
    central differences (h = 1e-5) through the oracle; nocc=6, nvir=40 is the adjoint's reference case.
    The gradient is also taken through the torch.autograd.Function."""
    import torch
    from distr_cc import device
    from distr_cc.adjoint import TriplesResidualAdjoint, make_residual_autograd_function
    device.init(0)
    print("residual adjoint o=%d v=%d" % (o, v), flush=True)
    fock, pppp, t2, t3, lam = _residual_case(o, v, 50 + o)
    adj = TriplesResidualAdjoint(o, v, fock, pppp, t3)
    _check_residual_adjoint(adj, o, v, fock, pppp, t2, t3, lam, fd_dirs=dirs)
    fn = make_residual_autograd_function()
    T2 = torch.tensor(t2, requires_grad=True)
    G = torch.tensor(pppp, requires_grad=True)
    (fn.apply(T2, G, adj) * torch.tensor(lam)).sum().backward()
    dt2, dg = adj.vjp_host(lam)
    assert np.array_equal(T2.grad.numpy(), dt2) and np.array_equal(G.grad.numpy(), dg)
