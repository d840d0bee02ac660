"""GPU triples engine: the T3 residual, its projection/denominator/update and
the T3 -> r1/r2/W contributions, as launches of libdcc kernels.

Data in HBM (FP64, C order):
  t3, R          (n3, v, v, v)     canonical blocks, lexicographic i<=j<=k
  Y              (npairs, o, v, v, v)  Y[(p,q)][l] = T[l,p,q] for p<=q (ring B operand)
  Z              (npairs, o, v, v, v)  Y[l][b,d,c] + Y[l][c,b,d] stored as [d,b,c], right after Y
  WR1, WR3[2]    (o, v, o, v)      ring weights, rows (r,x), cols (l,d)
  Gs             (pairs per chunk, o, v, v, v)  G[r;pq][a,b,c] before its placement
  Xt             (i per chunk, o, o, v, v, v)   T2->T3 base term X over ordered (i,j,k)
  pfs            (6, n3, n3)       W_oooo/F_oo prefactors
  W1             (v, v, v, v)      W_vvvv with the F_vv terms folded in
  Wp             (nS + nA, kpad)   pair-symmetric weights from W1 (vvvv_pairs); Tp / Up, the split T3
                                   and the products of a chunk of blocks, are views of Xt / Gs
  Xp             (5, targets per chunk, v, v, v)  W_oooo/F_oo terms s = 1..5 before their permutation
  intermediates  shapes of distr_cc/rccsdt.py:380,445,625

The residual follows distr_cc/rccsdt.py:368-433 term by term, grouped into
batched gather-GEMMs (dcc_contract_f64):
  * T2->T3 (rccsdt.py:388-399): the base term X[ijk] for every ordered (i,j,k) (2 GEMMs), then one
    S3 gather pass: each canonical block adds its six simultaneously (ijk, abc)-permuted sources;
  * F_vv and W_vvvv x T3 (rccsdt.py:400-402,424-426): 3 launches, N = (block, ...); the F_vv terms
    are folded into one copy of W_vvvv with F_vv/2 added on two of its diagonals (prepare_residual);
    with ``vvvv_pairs`` (a pair-symmetric W_vvvv) each term is instead, per chunk of blocks, a split of
    T3 into T[de] +- T[ed] over triangular pairs, one GEMM of two batch entries (symmetric and
    antisymmetric weights, half the flops of the plain product) and an unfold adding (US +- UA)/2 to R;
  * W_ovov/W_ovvo x T3 (rccsdt.py:540-596): with G the part-3 operator
        G[r;pq][a,b,c] = sum_{l,d} ( -W_ovov[l,b,r,d] Y[l][a,d,c] - W_ovov[l,c,r,d] Y[l][a,b,d]
                        + W_ovvo[l,a,d,r] Y[l][d,b,c]
                        - 1/2 (W_ovov[l,a,r,d] + W_ovvo[l,a,d,r]) (Y[l][b,d,c] + Y[l][c,b,d]) )
    part 3 of target (i,j,k) is G[i;jk], part 4 is G[j;ik] with (a,b) swapped and
    part 5 is G[k;ij] read as [c,a,b]; G is computed once for all r (GEMMs batched over
    the pairs p<=q, rows (r, x)) into a scratch buffer and placed by one S3 gather pass
    (all three parts of a target at once); with Z the W_ovvo term and the two
    (W_ovov + W_ovvo) terms are one GEMM over K = (term, l, d), B = [Y; Z];
  * W_oooo/F_oo x T3 (rccsdt.py:616-732): 6 GEMMs M = n3 targets, K = n3 sources,
    N = v^3, with the prefactors built on the device; term 0 goes straight into R, terms
    1..5 into Xp per chunk of targets and are added, permuted, by one S3 gather pass.
Then dcc_r3_finalize_f64 fuses rccsdt.py:947-971.

T3 -> r1/r2 (rccsdt.py:253-321) and T3 -> W_vooo/W_vvvo (rccsdt.py:323-366)
run per middle occupied index q over a slice Zq[p][r] = P3_201(T[p,q,r])
built by one S3 transform, with all ordered (p,q,r) covered exactly once:
  r2[q,l,b,c] += sum_{r,p,a} (-E_ooov[r,p,l,a] + 1/2 d_lr f_ov[p,a]) Z[p,q,r][a,b,c]
  r2[q,r,b,c] += sum_{p,a,d} E_vovv[c,p,a,d] Z[p,q,r][d,b,a]
  r1[i,a]     += sum_{k,b,c} (E_oovv[q,k,b,c] - 1/2 E_oovv[q,k,c,b]) Z[i,q,k][a,b,c]
  W_vooo[a,m,q,r] += sum_{p,d,e} E_oovv[m,p,d,e] Z[p,q,r][e,a,d]
  W_vvvo[a,b,d,q] -= sum_{r,p,e} E_oovv[r,p,d,e] Z[p,q,r][e,b,a]
(E_* are slices of the T1-dressed <pq|rs>; r1 uses P3_422 = 2 P3_201 - (P3_201)^{acb}.)
"""
import ctypes

import numpy as np

from distr_cc import _lib as L
from distr_cc.device import DeviceArray, init as device_init
from distr_cc.distribute_t3 import TP_T3, canonical_case, canonical_triples, ijk_to_linear, nocc3


def c_strides(shape):
    st, acc = [], 1
    for s in reversed(shape):
        st.append(acc)
        acc *= int(s)
    return tuple(reversed(st))


class Operand:
    """A device tensor with one letter per axis (letters not in a term are fixed per batch)."""

    def __init__(self, buf, letters, shape, base=0):
        assert len(letters) == len(shape)
        self.buf, self.letters, self.shape = buf, letters, tuple(int(s) for s in shape)
        self.strides = dict(zip(letters, c_strides(self.shape)))
        self.ext = dict(zip(letters, self.shape))
        self.base = int(base)

    @classmethod
    def strided(cls, buf, strides, base=0):
        """A view with explicit per-letter element strides and a base offset."""
        op = cls.__new__(cls)
        op.buf, op.letters, op.shape, op.ext = buf, ''.join(strides), None, {}
        op.strides = {k: int(v) for k, v in strides.items()}
        op.base = int(base)
        return op

    def offset(self, fixed):
        return self.base + sum(self.strides[k] * int(v) for k, v in fixed.items() if k in self.strides)


def idxmap(letters, ext, strides):
    """Three-level map of the group `letters` (outermost first) into an operand."""
    ls = list(letters)
    if not 1 <= len(ls) <= 3:
        raise ValueError("index group must have 1-3 letters: %r" % (letters,))
    while len(ls) < 3:
        ls.insert(0, None)
    e1 = ext[ls[1]] if ls[1] else 1
    e2 = ext[ls[2]] if ls[2] else 1
    s = [strides[x] if x else 0 for x in ls]
    return L.IdxMap(None, 0, int(e1), int(e2), int(s[0]), int(s[1]), int(s[2]))


def table_map(table_dev, bstride):
    return L.IdxMap(table_dev.ptr, int(bstride), 1, 1, 0, 0, 0)


def _suffix(dtype):
    return {np.dtype(np.float64): "f64", np.dtype(np.float32): "f32"}[np.dtype(dtype)]


class GemmLaunch:
    """One dcc_contract_f64 / dcc_contract_f32 call with fixed operands, maps and batch table."""

    def __init__(self, name, M, N, K, alpha, beta, A, am, ak, B, bk, bn, C, cm, cn, batches, max_rows,
                 flops, keep=()):
        if not (A.dtype == B.dtype == C.dtype):
            raise TypeError("%s: operand dtypes differ (%s, %s, %s)" % (name, A.dtype, B.dtype, C.dtype))
        self.fn = "dcc_contract_" + _suffix(A.dtype)
        self.name = name
        self.args = (int(M), int(N), int(K), float(alpha), float(beta), ctypes.c_void_p(A.ptr), am, ak,
                     ctypes.c_void_p(B.ptr), bk, bn, ctypes.c_void_p(C.ptr), cm, cn)
        rows = np.asarray(batches, dtype=np.int64).reshape(-1, 5)
        bt = np.zeros(rows.shape[0], dtype=L.BATCH_DTYPE)
        for n, field in enumerate(L.BATCH_DTYPE.names):
            bt[field] = rows[:, n]
        self.batches = DeviceArray.from_host(bt)
        self.nbatch = len(batches)
        self.max_rows = int(max_rows)
        self.flops = float(flops)
        self.keep = keep  # device tables referenced by maps
        self.tile = -1    # -1: library default; 0..7: dcc_contract_tiled_* tile_shape (autotuned)

    def __call__(self, stream):
        self.run(stream)

    def run(self, stream):
        s = stream.handle if stream is not None else None
        if self.fn == "dcc_contract_packed_bf16x3":
            L.call(self.fn, *self.args, self.batches.vp, self.nbatch, self.max_rows, self.tile, s)
        elif self.tile >= 0:
            L.call(self.fn.replace("contract_", "contract_tiled_"), *self.args, self.batches.vp, self.nbatch,
                   self.max_rows, self.tile, s)
        else:
            L.call(self.fn, *self.args, self.batches.vp, self.nbatch, self.max_rows, s)


def _fuse(letters, ops, ext):
    """Split an ordered letter list into runs that are contiguous in every operand of ``ops``."""
    groups = []
    for x in letters:
        if groups:
            prev = groups[-1][-1]
            if all(op.strides[prev] == op.strides[x] * ext[x] for op in ops):
                groups[-1].append(x)
                continue
        groups.append([x])
    return groups


def _group_map(groups, ext, op):
    """dcc_idxmap of up to three fused groups (outermost first) into operand ``op``."""
    if not groups:
        return L.IdxMap(None, 0, 1, 1, 0, 0, 0), 1
    exts = [int(np.prod([ext[x] for x in g])) for g in groups]
    strides = [op.strides[g[-1]] for g in groups]
    while len(exts) < 3:
        exts.insert(0, 1)
        strides.insert(0, 0)
    return L.IdxMap(None, 0, exts[1], exts[2], strides[0], strides[1], strides[2]), int(np.prod(exts))


def plan_einsum(name, sub, A, B, C, alpha, beta, ext, batches, max_levels=3):
    """Plan C (+)= alpha * einsum(sub, A, B) as one batched gather-GEMM launch.

    ``sub`` names the free/contracted letters ('xy,yz->xz' style; no letter may
    appear in all three operands).  ``batches`` is a list of (fixA, fixB, fixC)
    dicts giving, per operand, values of letters that are fixed for that batch
    entry (block offsets, occupied labels, ...); they only shift base offsets.
    M = letters of C taken from A (in A's order), N = letters of C taken from B
    (in B's order), K = letters shared by A and B.  Letters contiguous in both
    operands of a dimension are fused; if a dimension still needs more than
    ``max_levels`` levels, its outermost groups are moved into the batch list.
    """
    sa, rest = sub.split(',')
    sb, sc = rest.split('->')
    assert not (set(sa) & set(sb) & set(sc)), sub
    m_l = [x for x in sa if x in sc]
    n_l = [x for x in sb if x in sc]
    k_a = [x for x in sa if x not in sc]
    k_b = [x for x in sb if x not in sc]
    assert sorted(k_a) == sorted(k_b) and sorted(m_l + n_l) == sorted(sc), sub
    k_l = k_b if (sb[-1] in k_b or sa[-1] not in k_a) else k_a
    gm, gn, gk = _fuse(m_l, (A, C), ext), _fuse(n_l, (B, C), ext), _fuse(k_l, (A, B), ext)
    if len(gk) > max_levels:
        raise ValueError("%s: contracted index needs %d levels" % (sub, len(gk)))
    peel = []
    while len(gm) > max_levels:
        peel += gm.pop(0)
    while len(gn) > max_levels:
        peel += gn.pop(0)
    # base offsets of the given entries, then every value of each peeled letter (entry-major order)
    offs = np.array([(A.offset(fa), B.offset(fb), C.offset(fc)) for fa, fb, fc in batches],
                    dtype=np.int64).reshape(-1, 3)
    for x in peel:
        sx = np.array([op.strides.get(x, 0) for op in (A, B, C)], dtype=np.int64)
        vals = np.arange(ext[x], dtype=np.int64)
        offs = (offs[:, None, :] + vals[None, :, None] * sx[None, None, :]).reshape(-1, 3)
    am, M = _group_map(gm, ext, A)
    ak, K = _group_map(gk, ext, A)
    bk, _ = _group_map(gk, ext, B)
    bn, N = _group_map(gn, ext, B)
    cm, _ = _group_map(gm, ext, C)
    cn, _ = _group_map(gn, ext, C)
    bt = np.zeros((offs.shape[0], 5), dtype=np.int64)
    bt[:, :3] = offs
    bt[:, 4] = M
    return GemmLaunch(name, M, N, K, alpha, beta, A.buf, am, ak, B.buf, bk, bn, C.buf, cm, cn, bt, M,
                      2.0 * M * N * K * len(bt))


def tensor_axpby(dst, src, letters, shape, alpha, beta, stream):
    """dst (+)= alpha * src over the index space ``letters`` x ``shape`` (Operand views;
    a letter missing from ``src`` broadcasts with stride 0)."""
    sd, ss = _suffix(dst.buf.dtype), _suffix(src.buf.dtype)
    fn = "dcc_tensor_axpby_" + (sd if sd == ss else sd + "_" + ss)
    lay = L.Strided()
    lay.ndim = len(letters)
    for d, (x, e) in enumerate(zip(letters, shape)):
        lay.shape[d] = int(e)
        lay.dst_stride[d] = dst.strides[x]
        lay.src_stride[d] = src.strides.get(x, 0)
    L.call(fn, ctypes.c_void_p(dst.buf.ptr + dst.base * dst.buf.dtype.itemsize),
           ctypes.c_void_p(src.buf.ptr + src.base * src.buf.dtype.itemsize), ctypes.byref(lay), float(alpha),
           float(beta), stream.handle if stream is not None else None)


def plan_term(name, sub, A, B, C, alpha, beta, batches_fixed, ext):
    """plan_einsum with the same fixed-letter values applied to all three operands."""
    return plan_einsum(name, sub, A, B, C, alpha, beta, ext, [(f, f, f) for f in batches_fixed])


class S3Launch:
    """One dcc_s3_transform call (out = beta*out + sum_s coef[s] src^{TP_T3[s]} per job) in a launch list."""

    def __init__(self, name, out, src, nvir, beta, jobs):
        self.name, self.flops = name, 0.0
        self.fn = "dcc_s3_transform_" + _suffix(out.dtype)
        self.out, self.src, self.nvir, self.beta = out, src, int(nvir), float(beta)
        self.jobs = DeviceArray.from_host(s3_jobs_array(jobs))
        self.njobs = len(jobs)

    def run(self, stream):
        L.call(self.fn, self.out.vp, self.src.vp, self.nvir, self.beta, self.jobs.vp, self.njobs,
               stream.handle if stream is not None else None)


class S3GatherLaunch:
    """One dcc_s3_gather call: out[dst] = beta*out[dst] + sum_p coef[p] src[src_p]^{TP_T3[p]} per job."""

    def __init__(self, name, out, src, nvir, beta, jobs):
        self.name, self.flops = name, 0.0
        self.fn = "dcc_s3_gather_" + _suffix(out.dtype)
        self.out, self.src, self.nvir, self.beta = out, src, int(nvir), float(beta)
        arr = np.zeros(len(jobs), dtype=L.S3GATHER_DTYPE)
        for n, (dst, srcs, coef) in enumerate(jobs):
            arr[n]['dst'] = dst
            arr[n]['src'] = srcs
            arr[n]['coef'] = coef
        self.jobs = DeviceArray.from_host(arr)
        self.njobs = len(jobs)

    def run(self, stream):
        L.call(self.fn, self.out.vp, self.src.vp, self.nvir, self.beta, self.jobs.vp, self.njobs,
               stream.handle if stream is not None else None)


class PairLaunch:
    """One dcc_vvvv_pair_split_* (kpad given) or dcc_vvvv_pair_unfold_* call over the canonical blocks
    b0 .. b0+nb-1 of ``blocks`` (T3 for a split, R for an unfold); ``buf`` is the pair buffer."""

    def __init__(self, name, fn, blocks, buf, b0, nb, nvir, axes, kpad=None):
        self.name, self.flops, self.fn = name, 0.0, fn
        v3 = int(nvir) ** 3
        self.blocks = blocks   # keeps the array behind the pointer alive
        self.blk = ctypes.c_void_p(blocks.ptr + int(b0) * v3 * blocks.dtype.itemsize)
        self.buf, self.nb, self.nvir, self.axes, self.kpad = buf, int(nb), int(nvir), int(axes), kpad

    def run(self, stream):
        s = stream.handle if stream is not None else None
        if self.kpad is not None:
            L.call(self.fn, self.blk, self.nb, self.nvir, self.axes, int(self.kpad), self.buf.vp, s)
        else:
            L.call(self.fn, self.buf.vp, self.nb, self.nvir, self.axes, self.blk, s)


def pair_dims(nvir):
    """(nS, nA, kpad) of the pair-symmetric W_vvvv products: symmetric and strict pair counts and
    their common K, padded to a multiple of 64 (the two-stage direct-to-LDS GEMM tiles step K by 64)."""
    v = int(nvir)
    nS, nA = v * (v + 1) // 2, v * (v - 1) // 2
    return nS, nA, -(-nS // 64) * 64


def check_pair_symmetric(W, samples=4096, rtol=1e-8):
    """Raise ValueError unless W[a,b,d,e] == W[b,a,e,d] (to rtol of the sampled magnitude) on a fixed
    random sample of index tuples: the pair-symmetric W_vvvv products assume it."""
    W = np.asarray(W)
    v = W.shape[0]
    if v == 0:
        return
    a, b, d, e = np.random.default_rng(12345).integers(0, v, size=(4, samples))
    x, y = W[a, b, d, e], W[b, a, e, d]
    scale = max(float(np.abs(x).max()), np.finfo(np.float64).tiny)
    dev = float(np.abs(x - y).max())
    if dev > rtol * scale:
        raise ValueError("W_vvvv is not pair-symmetric (max |W[a,b,d,e] - W[b,a,e,d]| = %.3g of %.3g); "
                         "build the engine with vvvv_pairs=False" % (dev, scale))


class T2T3Operands:
    """The operands of the fused T2 -> T3 base-term GEMM of one chunk of first occupied indices
    i0 .. i0+n-1 (rccsdt.py:388-399), with K = (d, then l) of extent v + o:
        Ap[i,b,j,a,K] = W_vvvo[a,b,K,j]        Bp[i,b,K,k,c] = t2[i0+i,k,K,c]      (K < v)
        Ap[i,b,j,a,K] = -W_vooo[a,K-v,i0+i,j]   Bp[i,b,K,k,c] = t2[K-v,k,b,c]      (K >= v)
    written by one dcc_t2t3_operands call (W_ovvv holds W_vvvo[a,b,d,j] as [j,a,b,d], W_oovo holds
    W_vooo[a,l,i,j] as [i,j,a,l], t2 is [i,k,d,c]), as split-BF16x3 words for split-BF16x3 products."""

    def __init__(self, name, Ap, Bp, W_ovvv, W_oovo, t2, o, v, i0, n, pack):
        self.name, self.flops = name, 0.0
        self.Ap, self.Bp, self.W_ovvv, self.W_oovo, self.t2 = Ap, Bp, W_ovvv, W_oovo, t2
        self.o, self.v, self.i0, self.n, self.pack = o, v, i0, n, pack

    def run(self, stream):
        fn = "dcc_t2t3_operands_" + ("packed_f32" if self.pack else _suffix(self.Ap.dtype))
        L.call(fn, self.Ap.vp, self.Bp.vp, self.W_ovvv.vp, self.W_oovo.vp, self.t2.vp, self.o, self.v, self.i0,
               self.n, stream.handle if stream is not None else None)


def plan_t2t3(o, v, R, t2, W_ovvv, W_oovo, scratch_bytes, packed=False):
    """T2 -> T3 terms (rccsdt.py:388-399) added into the canonical blocks R; returns (Xt, launches).

    The twelve terms are the six simultaneous (ijk, abc) permutations of one base term
        X[i,j,k][a,b,c] = sum_d W_vvvo[a,b,d,j] t2[i,k,d,c] - sum_l W_vooo[a,l,i,j] t2[l,k,b,c],
    r3[ijk][abc] += sum_pi X[pi(ijk)][pi(abc)].  With i and b fixed, both sums are one product over
    K = (d, l): X[i,.,.][.,b,.] = Ap[i,b] Bp[i,b] with M = (j,a), N = (k,c) (T2T3Operands builds Ap and
    Bp), so X is built for every ordered (i,j,k) by one batched GEMM over the (i,b) pairs that writes X
    once (two GEMMs -- K = d, then K = l adding into X -- wrote it and then read and rewrote it), and
    one multi-source S3 gather pass adds, for every canonical target, its (up to six) permuted source
    blocks in one read-modify-write.  Ordered triples are processed in chunks of the first index whose
    X fits ``scratch_bytes``; Xt is that chunk buffer (ic, o, o, v, v, v).  ``packed``: the GEMM runs
    as split-BF16x3 products on operands packed by T2T3Operands (when K = v + o > 16; a shorter K stays
    on FP32 like every such GEMM of the split-BF16x3 engine)."""
    v3 = v ** 3
    dt = R.dtype
    ic = int(max(1, min(o, scratch_bytes // max(1, o * o * v3 * dt.itemsize))))
    Xt = DeviceArray((ic, o, o, v, v, v), dt)
    K2 = v + o
    packed = packed and K2 > 16
    Ap = DeviceArray((ic, v, o, v, K2), dt)
    Bp = DeviceArray((ic, v, K2, o, v), dt)
    ext = {'j': o, 'a': v, 'K': K2, 'k': o, 'c': v}
    # source-position permutations of (i,j,k); the abc permutation of the S3 job is its inverse
    pis = [(0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0)]
    triples = canonical_triples(o)
    launches = []
    for i0 in range(0, o, ic):
        n = min(ic, o - i0)
        launches.append(T2T3Operands('T2T3: operands', Ap, Bp, W_ovvv, W_oovo, t2, o, v, i0, n, packed))
        A = Operand(Ap, 'ibjaK', (ic, v, o, v, K2))
        B = Operand(Bp, 'ibKkc', (ic, v, K2, o, v))
        X = Operand(Xt, 'ijkabc', (ic, o, o, v, v, v))
        fixed = [{'i': i, 'b': b} for i in range(n) for b in range(v)]
        g = plan_einsum('T2T3: X', 'jaK,Kkc->jkac', A, B, X, 1.0, 0.0, ext, [(f, f, f) for f in fixed])
        if packed:
            g.fn = "dcc_contract_packed_bf16x3"
        launches.append(g)
        # one gather pass: target q takes sum_pi X[pi(ijk)]^{pi^-1} over the sources in this chunk
        jobs = []
        for q, t in enumerate(triples):
            t = tuple(int(x) for x in t)   # Python integers: the offsets pass 2^31 for large chunks
            srcs, coef = [0] * 6, np.zeros(6)
            for pos in pis:
                src = (t[pos[0]], t[pos[1]], t[pos[2]])
                if i0 <= src[0] < i0 + n:
                    p = TP_T3.index(tuple(pos.index(u) for u in range(3)))
                    srcs[p] = (((src[0] - i0) * o + src[1]) * o + src[2]) * v3
                    coef[p] = 1.0
            if coef.any():
                jobs.append((q * v3, srcs, coef))
        launches.append(S3GatherLaunch('T2T3: symmetrise', R, Xt, v, 1.0, jobs))
    return Xt, launches


def s3_jobs_array(jobs):
    arr = np.zeros(len(jobs), dtype=L.S3JOB_DTYPE)
    for n, (dst, src, coef) in enumerate(jobs):
        arr[n]['dst'] = dst
        arr[n]['src'] = src
        arr[n]['coef'] = coef
    return arr


class TriplesEngine:
    """Holds the device state of one RCCSDT triples update for (nocc, nvir)."""

    def __init__(self, nocc, nvir, stream=None, device=0, dtype=np.float64, contract="native", ring_fused=None,
                 ring_g_bytes=8 << 30, vvvv_pairs=False, pair_bytes=None):
        """``dtype`` float64 is the plain FP64 path; float32 stores T3, R, Y and the
        intermediates in FP32 and contracts on FP32 MFMA, keeping the denominator,
        the projection arithmetic and the norms in FP64.  ``contract="bf16x3"`` (FP32 storage
        only) runs every contraction as split-BF16x3 products on BF16 MFMA instead.

        ``ring_fused`` (default: when the device has room for a second Y-sized buffer) stores
        Z[(p,q)][l][d,b,c] = Y[l][b,d,c] + Y[l][c,b,d] next to Y, so the two terms that share the
        W_ovov + W_ovvo weight become one, fused along K with the W_ovvo term: three ring
        GEMMs per part instead of five, and 4/5 of the ring flops.  ``ring_g_bytes`` bounds the
        scratch buffer of the ring operator G (pairs are processed in chunks that fit).

        ``vvvv_pairs`` runs the three W_vvvv terms as symmetric and antisymmetric products over
        triangular index pairs (half the flops; dcc_vvvv_pair_* in the library), which requires a
        pair-symmetric W_vvvv, W[a,b,d,e] = W[b,a,e,d], as every T1/T2-dressed one is; set_intermediates
        checks a sample of it.  "auto" turns it on when the device has room for its weights (its
        other buffers are views of the T2 -> T3 and ring scratch).  ``pair_bytes`` (optional) bounds
        that scratch further (blocks are processed in chunks that fit)."""
        device_init(device)
        o, v = int(nocc), int(nvir)
        self.nocc, self.nvir = o, v
        self.dtype = dt = np.dtype(dtype)
        self.sfx = _suffix(dt)
        self.n3 = n3 = nocc3(o)
        self.stream = stream
        self.triples = canonical_triples(o)
        self.blk_ijk = DeviceArray.from_host(np.ascontiguousarray(self.triples, dtype=np.int32))
        self.pairs = [(p, q) for p in range(o) for q in range(p, o)]
        self.npairs = len(self.pairs)
        v3 = v ** 3
        # --- device buffers ---
        self.t3 = DeviceArray((n3, v, v, v), dt)
        self.R = DeviceArray((n3, v, v, v), dt)
        ysize = self.npairs * o * v3
        if ring_fused is None:
            from distr_cc.device import mem_get_info
            free = mem_get_info()[0]
            # room for Z plus the rest of the engine (W_vvvv and its F_vv-folded copy, W_oooo scratch,
            # lower-order chunks) and a margin
            ring_fused = free > (ysize * 2 + n3 * v3 * 2 + v3 * v) * dt.itemsize + (24 << 30)
        self.ring_fused = bool(ring_fused)
        self.Yz = DeviceArray((2 if self.ring_fused else 1, self.npairs, o, v, v, v), dt)
        self.Y = DeviceArray((self.npairs, o, v, v, v), dt, ptr=self.Yz.ptr, owner=False)
        self.Z = (DeviceArray((self.npairs, o, v, v, v), dt, ptr=self.Yz.ptr + ysize * dt.itemsize, owner=False)
                  if self.ring_fused else None)
        self.t2 = DeviceArray((o, o, v, v), dt)
        self.F_oo = DeviceArray((o, o), dt)
        self.F_vv = DeviceArray((v, v), dt)
        self.W_oooo = DeviceArray((o, o, o, o), dt)
        self.W_ovvo = DeviceArray((o, v, v, o), dt)
        self.W_ovov = DeviceArray((o, v, o, v), dt)
        self.W_vooo = DeviceArray((v, o, o, o), dt)
        self.W_vvvo = DeviceArray((v, v, v, o), dt)
        self.W_ovvv = DeviceArray((o, v, v, v), dt)   # W_vvvo[a,b,d,j] stored as [j,a,b,d]
        self.W_oovo = DeviceArray((o, o, v, o), dt)   # W_vooo[a,l,i,j] stored as [i,j,a,l]
        self.W_vvvv = DeviceArray((v, v, v, v), dt)
        # W_vvvv with the F_vv terms folded in (built by prepare_residual), read by all three W_vvvv
        # terms: W1[p,q,d,e] = W_vvvv[p,q,d,e] + 1/2 F_vv[p,d] d_qe + 1/2 F_vv[q,e] d_pd.  The three
        # terms then add 1/2 (F1 + F2) + 1/2 (F1 + F3) + 1/2 (F2 + F3) = F1 + F2 + F3, the F_vv terms of
        # rccsdt.py:400-402, which need no launches of their own
        self.W1 = DeviceArray((v, v, v, v), dt)
        self.WR1 = DeviceArray((o, v, o, v), dt)
        self.WR3 = DeviceArray((2, o, v, o, v), dt)   # [W_ovvo term, (W_ovov + W_ovvo) term]
        w3 = (o * v) ** 2 * dt.itemsize
        self.WR3a = DeviceArray((o, v, o, v), dt, ptr=self.WR3.ptr, owner=False)
        self.WR3b = DeviceArray((o, v, o, v), dt, ptr=self.WR3.ptr + w3, owner=False)
        self.pfs = DeviceArray((6, n3, n3), dt)
        # W_oooo/F_oo terms s = 1..5 before their permutation, for a chunk of target blocks (the GEMMs'
        # M); a chunk is rounded up to a multiple of 64 rows when that stays within 5/4 of the bound,
        # so the GEMM tiles are not partly empty (e.g. 52 -> 64 at nocc=16, nvir=160 in FP64)
        per_target = max(1, 5 * v3 * dt.itemsize)
        qc = ring_g_bytes // per_target
        if qc >= 32 and -(-qc // 64) * 64 * per_target * 4 <= 5 * ring_g_bytes:
            qc = -(-qc // 64) * 64
        qc = int(max(1, min(n3, qc)))
        self.Xp = DeviceArray((5, qc, v, v, v), dt)
        self.eia = DeviceArray((o, v))
        self.norb = int(L._lib.dcc_r3_finalize_tiles(v))
        self.partials = DeviceArray((n3 * self.norb, 2))
        self.norms = DeviceArray((2,))
        # --- Y materialization jobs: Y[(p,q)][l] = T[l,p,q] ---
        jobs = []
        for pq, (p, q) in enumerate(self.pairs):
            for l in range(o):
                case, (ci, cj, ck) = canonical_case(l, p, q)
                coef = np.zeros(6)
                coef[case] = 1.0
                jobs.append(((pq * o + l) * v3, ijk_to_linear(ci, cj, ck, o) * v3, coef))
        self.y_jobs = DeviceArray.from_host(s3_jobs_array(jobs))
        self.n_y_jobs = len(jobs)
        # --- Z jobs: Z[(p,q)][l] = Y[(p,q)][l].transpose(1,0,2) + Y[(p,q)][l].transpose(2,1,0), from T ---
        if self.ring_fused:
            jobs = []
            for pq, (p, q) in enumerate(self.pairs):
                for l in range(o):
                    case, (ci, cj, ck) = canonical_case(l, p, q)
                    a = TP_T3[case]
                    coef = np.zeros(6)
                    for b in ((1, 0, 2), (2, 1, 0)):
                        coef[TP_T3.index((a[b[0]], a[b[1]], a[b[2]]))] += 1.0
                    jobs.append(((pq * o + l) * v3, ijk_to_linear(ci, cj, ck, o) * v3, coef))
            self.z_jobs = DeviceArray.from_host(s3_jobs_array(jobs))
        # --- pair-symmetric W_vvvv: weights [nS + nA][kpad], split T3 and products for a chunk of blocks
        # (the split T3 and the products of a chunk of blocks live in the T2 -> T3 and ring scratch
        # buffers, idle during the W_vvvv terms: only the weights take memory of their own)
        nS, nA, kp = pair_dims(v)
        if isinstance(vvvv_pairs, str):
            if vvvv_pairs != "auto":
                raise ValueError("vvvv_pairs must be True, False or 'auto'")
            from distr_cc.device import mem_get_info
            vvvv_pairs = mem_get_info()[0] > (nS + nA) * kp * dt.itemsize + (24 << 30)
        self.vvvv_pairs = bool(vvvv_pairs)
        self.pair_bytes = None if pair_bytes is None else int(pair_bytes)
        if self.vvvv_pairs:
            self.pair_dims = (nS, nA, kp)
            self.Wp = DeviceArray((nS + nA, kp), dt)
        if contract not in ("native", "bf16x3") or (contract == "bf16x3" and dt != np.float32):
            raise ValueError("contract must be 'native', or 'bf16x3' with float32 storage")
        self.contract = contract
        self._tables = []
        self.ring_g_bytes = int(ring_g_bytes)
        self.launches = self._plan_residual()
        self._retarget(self.launches)
        self._yz_direct = False   # Y/Z written as packed words (set by _setup_packing)
        self._packs = self._setup_packing() if contract == "bf16x3" else []
        self._lower = None

    def _setup_packing(self):
        """Split-BF16x3 residual: point every GEMM operand at a packed (hi | lo << 16) copy of its
        tensor, so each element is split once per residual instead of once per tile that reads it.
        Y and Z are packed in place (they are rebuilt from T3 for every residual and read by nothing
        else), by writing them as packed words in the first place; the others get packed copies.
        Returns the (source, packed, count) passes to run."""
        sources = [self.t3, self.Yz, self.W1, self.W_ovvv, self.W_oovo, self.t2, self.WR1, self.WR3, self.pfs]
        # the pair-symmetric W_vvvv operands are written as packed words by their own kernels
        in_place = [self.Wp, self.Tp] if self.vvvv_pairs else []
        sources += in_place
        packed = {}
        # Y/Z are packed in place unless a GEMM kept on FP32 (K <= 16, tiny systems) reads them
        yz = self.Yz.ptr, self.Yz.ptr + self.Yz.nbytes
        yz_in_place = not any(yz[0] <= g.args[i].value < yz[1] for g in self.launches
                              if isinstance(g, GemmLaunch) and g.fn != "dcc_contract_f32_bf16x3" for i in (5, 8))

        def find(ptr):
            for src in sources:
                if src.ptr <= ptr < src.ptr + src.nbytes:
                    if id(src) not in packed:
                        direct = (src is self.Yz and yz_in_place) or any(src is b for b in in_place)
                        dst = src if direct else DeviceArray(src.shape, np.uint32)
                        packed[id(src)] = (src, dst)
                    return packed[id(src)]
            return None

        for g in self.launches:
            if not isinstance(g, GemmLaunch) or g.fn != "dcc_contract_f32_bf16x3":
                continue
            a, b = find(g.args[5].value), find(g.args[8].value)
            if a is None or b is None:
                continue
            args = list(g.args)
            args[5] = ctypes.c_void_p(a[1].ptr + (g.args[5].value - a[0].ptr))
            args[8] = ctypes.c_void_p(b[1].ptr + (g.args[8].value - b[0].ptr))
            g.args = tuple(args)
            g.fn = "dcc_contract_packed_bf16x3"
        # Y/Z packed in place are written as packed words by materialize_y itself
        self._yz_direct = yz_in_place and id(self.Yz) in packed
        return [(src, dst, int(np.prod(src.shape))) for src, dst in packed.values()
                if not (self._yz_direct and src is self.Yz) and not any(src is b for b in in_place)]

    def pack_operands(self):
        """Run the split-BF16x3 packing passes (after T3, Y/Z and the intermediates are final)."""
        s = self.stream.handle if self.stream is not None else None
        for src, dst, n in self._packs:
            L.call("dcc_pack_bf16x3", src.vp, dst.vp, n, s)

    def _retarget(self, launches):
        """Split-BF16x3 engine: GEMMs with K > 16 run as split-BF16x3 products; those with K <= 16 (the
        occupied-index T2->T3 term: an HBM-bound rank-o update, whose k-step of 32 would be half
        padding) stay on FP32 MFMA, which is also the more accurate of the two."""
        if self.contract == "bf16x3":
            for g in launches:
                if isinstance(g, GemmLaunch) and g.args[2] > 16 and g.fn != "dcc_contract_packed_bf16x3":
                    g.fn = "dcc_contract_f32_bf16x3"

    # ------------------------------------------------------------------ plan
    def _plan_residual(self):
        o, v, n3 = self.nocc, self.nvir, self.n3
        ext = {x: v for x in 'abcdexyz'}
        ext.update({'l': o, 'm': o, 'r': o, 'B': n3, 'p': n3, 'q': n3})
        T = Operand(self.t3, 'Bxyz', (n3, v, v, v))
        launches = []
        blocks = [dict(I=i, J=j, K=k, B=n) for n, (i, j, k) in enumerate(self.triples)]
        # T2 -> T3 (rccsdt.py:388-399)
        Rb = Operand(self.R, 'Babc', (n3, v, v, v))
        launches += self._plan_t2t3()
        ring = self._plan_ring()
        # W_vvvv x T3 with the block index in N (rccsdt.py:424-426); the F_vv terms (rccsdt.py:400-402)
        # ride along in W1 (see __init__)
        one = [dict()]
        if self.vvvv_pairs:
            self._pair_scratch()
            launches += self._plan_vvvv_pairs()
        else:
            for sub, tl in [('abde,Bdec->Babc', 'Bdec'), ('acde,Bdbe->Babc', 'Bdbe'), ('bcde,Bade->Babc', 'Bade')]:
                A = Operand(self.W1, sub[:4], (v, v, v, v))
                B = Operand(self.t3, tl, (n3, v, v, v))
                launches.append(plan_term('W_vvvv ' + sub, sub, A, B, Rb, 1.0, 1.0, one, ext))
        launches += ring
        # W_oooo / F_oo (rccsdt.py:726-731).  The abc permutation of each of the six terms commutes
        # with the sum over source blocks p, so every term is the same natural GEMM
        # X_s[q] = sum_p pfs[s][p][q] T[p] (coalesced in and out) followed, for s > 0, by R[q] +=
        # X_s[q] transposed by the term's permutation (an S3 gather pass).
        # s: 0 'pq,pabc->qabc', 1 'pq,pacb->qabc', 2 'pq,pbac->qabc', 3 'pq,pabc->qcab',
        #    4 'pq,pabc->qbca', 5 'pq,pabc->qcba'; R[q] += X_s[q].transpose(perm[s]).
        A = Operand(self.pfs, 'Spq', (6, n3, n3))
        B = Operand(self.t3, 'pabc', (n3, v, v, v))
        Rq = Operand(self.R, 'qabc', (n3, v, v, v))
        launches.append(plan_term('W_oooo/F_oo s0', 'pq,pabc->qabc', A, B, Rq, 1.0, 1.0, [dict(S=0)], ext))
        # s = 1..5 for a chunk of targets q at a time into Xp[s-1][q - q0]: one GEMM whose rows are
        # (s, q), so each T3 panel is read by neighbouring row tiles of all five terms while it is in
        # L2 (five GEMMs of nq rows each re-read all of T3); then one gather pass
        v3 = v ** 3
        perm = {1: (0, 2, 1), 2: (1, 0, 2), 3: (2, 0, 1), 4: (1, 2, 0), 5: (2, 1, 0)}
        qc = self.Xp.shape[1]
        for q0 in range(0, n3, qc):
            nq = min(qc, n3 - q0)
            ext_c = dict(ext, q=nq, S=5)
            Aq = Operand.strided(self.pfs, {'S': n3 * n3, 'p': n3, 'q': 1}, base=n3 * n3 + q0)
            X = Operand.strided(self.Xp, {'S': qc * v3, 'q': v3, 'a': v * v, 'b': v, 'c': 1})
            launches.append(plan_term('W_oooo/F_oo s1-5', 'pSq,pabc->Sqabc', Aq, B, X, 1.0, 0.0, [dict()], ext_c))
            jobs = []
            for q in range(q0, q0 + nq):
                srcs, coef = [0] * 6, np.zeros(6)
                for s in range(1, 6):
                    p = TP_T3.index(perm[s])
                    srcs[p], coef[p] = ((s - 1) * qc + q - q0) * v3, 1.0
                jobs.append((q * v3, srcs, coef))
            launches.append(S3GatherLaunch('W_oooo/F_oo add', self.R, self.Xp, v, 1.0, jobs))
        return launches

    def _pair_scratch(self):
        """Tp (split T3) and Up (products) of the pair-symmetric W_vvvv terms for a chunk of blocks, as
        views of the T2 -> T3 base-term buffer Xt and the ring scratch Gs: the W_vvvv launches run after
        the T2 -> T3 ones have consumed Xt and before the ring writes Gs (one stream).  Chunks hold as
        many blocks as both buffers (and ``pair_bytes``, if given) allow, made equal in size."""
        v, n3, dt = self.nvir, self.n3, self.dtype
        nS, nA, kp = self.pair_dims
        t_blk, u_blk = 2 * v * kp, 2 * v * nS          # elements per block
        cap = min(int(np.prod(self.Xt.shape)) // t_blk, int(np.prod(self.Gs.shape)) // u_blk, n3)
        if self.pair_bytes is not None:
            cap = min(cap, self.pair_bytes // ((t_blk + u_blk) * dt.itemsize))
        if cap >= 1:
            nbc = -(-n3 // -(-n3 // cap))               # equal chunks: no short remainder
            self.Tp = DeviceArray((2, nbc, v, kp), dt, ptr=self.Xt.ptr, owner=False)
            self.Up = DeviceArray((2, nbc, v, nS), dt, ptr=self.Gs.ptr, owner=False)
        else:                                           # scratch bounds below one block: own buffers
            self.Tp = DeviceArray((2, 1, v, kp), dt)
            self.Up = DeviceArray((2, 1, v, nS), dt)

    def _plan_vvvv_pairs(self):
        """W_vvvv x T3 (rccsdt.py:424-426) as pair-symmetric products.  Term t has its output pair and
        T3 pair at block positions (0,1), (0,2), (1,2) (pair_axes t); per chunk of blocks: one split of
        T3 into TS/TA, one GEMM with two batch entries (WS x TS over the nS symmetric pairs, WA x TA over
        the nA strict ones, common K = kpad) into US/UA, one unfold adding (US +- UA)/2 into R.  Axes 0
        keeps the shared index innermost ([pair][z] rows: B is N-contiguous, as T3 is); axes 1 and 2 put
        the pairs innermost (B K-contiguous, for the direct-to-LDS tiles)."""
        v, n3 = self.nvir, self.n3
        nS, nA, kp = self.pair_dims
        nbc = self.Tp.shape[1]
        packed = self.contract == "bf16x3"
        split_fn = "dcc_vvvv_pair_split_" + ("packed_f32" if packed else self.sfx)
        unfold_fn = "dcc_vvvv_pair_unfold_" + self.sfx
        launches = []
        for axes, sub in enumerate(('abde,Bdec->Babc', 'acde,Bdbe->Babc', 'bcde,Bade->Babc')):
            for b0 in range(0, n3, nbc):
                nb = min(nbc, n3 - b0)
                launches.append(PairLaunch('W_vvvv pairs split', split_fn, self.t3, self.Tp, b0, nb, v, axes, kp))
                am, ak = L.IdxMap(None, 0, 1, nS, 0, 0, kp), L.IdxMap(None, 0, 1, kp, 0, 0, 1)
                if axes == 0:     # Tp [2][nb][kp][v], Up [2][nb][nS][v]
                    bk, bn = L.IdxMap(None, 0, 1, kp, 0, 0, v), L.IdxMap(None, 0, nb, v, 0, kp * v, 1)
                    cm, cn = L.IdxMap(None, 0, 1, nS, 0, 0, v), L.IdxMap(None, 0, nb, v, 0, nS * v, 1)
                else:             # Tp [2][nb][v][kp], Up [2][nb][v][nS]
                    bk, bn = L.IdxMap(None, 0, 1, kp, 0, 0, 1), L.IdxMap(None, 0, nb, v, 0, v * kp, kp)
                    cm, cn = L.IdxMap(None, 0, 1, nS, 0, 0, 1), L.IdxMap(None, 0, nb, v, 0, v * nS, nS)
                batches = [(0, 0, 0, 0, nS), (nS * kp, nb * v * kp, nb * v * nS, 0, nA)]
                launches.append(GemmLaunch('W_vvvv pairs ' + sub, nS, nb * v, kp, 1.0, 0.0, self.Wp, am, ak,
                                           self.Tp, bk, bn, self.Up, cm, cn, batches, nS,
                                           2.0 * (nS + nA) * nb * v * kp))
                launches.append(PairLaunch('W_vvvv pairs unfold', unfold_fn, self.R, self.Up, b0, nb, v, axes))
        return launches

    def _plan_t2t3(self):
        """T2 -> T3 terms (plan_t2t3), with the base-term scratch bounded by ``ring_g_bytes``."""
        self.Xt, launches = plan_t2t3(self.nocc, self.nvir, self.R, self.t2, self.W_ovvv, self.W_oovo,
                                      self.ring_g_bytes, packed=self.contract == "bf16x3")
        return launches

    def _plan_ring(self):
        """Ring terms (rccsdt.py:540-596).

        G[r;pq] (the part-3 operator of the module docstring) is the same for the three parts; only
        its placement differs (part 3 of target (r,p,q) is G[r;pq], part 4 of (p,r,q) is G[r;pq] with
        (a,b) swapped, part 5 of (p,q,r) is G[r;pq] read as [c,a,b]).  So G is computed once for every
        r into a scratch buffer Gs[pair][r][a,b,c] -- GEMMs with M = (r, x) over all o*v rows,
        batched over pairs, plain strided output -- and then added into R by one S3 placement pass
        per part.  Pairs are processed in chunks whose G fits ``ring_g_bytes``."""
        o, v = self.nocc, self.nvir
        v2, v3 = v * v, v ** 3
        ov = o * v
        dt = self.dtype
        pc = int(max(1, min(self.npairs, self.ring_g_bytes // max(1, o * v3 * dt.itemsize))))
        self.Gs = DeviceArray((pc, o, v, v, v), dt)
        Sg = (v2, v, 1)   # strides of [a,b,c] in a G block

        def rows(part, p, q):
            return {3: (0, p + 1), 4: (p, q + 1), 5: (q, o)}[part]

        def target(part, r, p, q):
            t = {3: (r, p, q), 4: (p, r, q), 5: (p, q, r)}[part]
            return ijk_to_linear(t[0], t[1], t[2], o)

        # placement of G[a,b,c] in the target block, as the S3 permutation of the source block
        perm = {3: (0, 1, 2), 4: (1, 0, 2), 5: (1, 2, 0)}
        # term: (weight, G axis carried by M, N axes (outer, inner), B strides of d, n_outer, n_inner)
        terms = []
        if self.ring_fused:
            # W_ovvo[l,a,d,r] Y[l][d,b,c] - 1/2(..) Z[l][d,b,c] as one GEMM over K = (term, l, d)
            terms.append(('WR3', 0, (1, 2), v2, v, 1))
        else:
            terms += [
                ('WR3a', 0, (1, 2), v2, v, 1),   #  W_ovvo[l,a,d,r] Y[l][d,b,c]
                ('WR3b', 0, (1, 2), v, v2, 1),   # -1/2(..) Y[l][b,d,c]
                ('WR3b', 0, (1, 2), 1, v, v2),   # -1/2(..) Y[l][c,b,d]
            ]
        terms += [
            ('WR1', 1, (0, 2), v, v2, 1),    # -W_ovov[l,b,r,d] Y[l][a,d,c]
            ('WR1', 2, (0, 1), 1, v2, v),    # -W_ovov[l,c,r,d] Y[l][a,b,d]
        ]
        launches = []
        for c0 in range(0, self.npairs, pc):
            chunk = list(enumerate(self.pairs))[c0:c0 + pc]
            batches = [(0, pq * o * v3, (pq - c0) * o * v3, 0, ov) for pq, _ in chunk]
            for n, (wname, mx, (n0, n1), sd, sn0, sn1) in enumerate(terms):
                W = getattr(self, wname)
                nt = 2 if wname == 'WR3' else 1                  # terms fused along K
                am = L.IdxMap(None, 0, 1, ov, 0, 0, ov)          # row (r,x) of the (ov x ov) weight
                ak = L.IdxMap(None, 0, 1, ov, ov * ov, 0, 1)     # (term, (l,d))
                zoff = self.npairs * o * v3                      # Z follows Y in one buffer
                bk = L.IdxMap(None, 0, o, v, zoff, v3, sd)       # (term, l, d)
                bn = L.IdxMap(None, 0, v, v, 0, sn0, sn1)        # (n_outer, n_inner)
                cm = L.IdxMap(None, 0, o, v, 0, v3, Sg[mx])      # (r, x)
                cn = L.IdxMap(None, 0, v, v, 0, Sg[n0], Sg[n1])
                launches.append(GemmLaunch('ring %s M%d' % (wname, mx), ov, v2, nt * ov, 1.0,
                                           0.0 if n == 0 else 1.0, W, am, ak, self.Y, bk, bn, self.Gs, cm, cn,
                                           batches, ov, 2.0 * len(chunk) * ov * v2 * ov * nt))
            # one gather pass: each target takes its (up to three) part contributions of this chunk
            contrib = {}
            for part in (3, 4, 5):
                pidx = TP_T3.index(perm[part])
                for pq, (p, q) in chunk:
                    r0, r1 = rows(part, p, q)
                    for r in range(r0, r1):
                        contrib.setdefault(target(part, r, p, q), []).append((pidx, ((pq - c0) * o + r) * v3))
            jobs = []
            for tq, lst in sorted(contrib.items()):
                srcs, coef = [0] * 6, np.zeros(6)
                for pidx, off in lst:
                    assert coef[pidx] == 0.0   # the three parts use three different permutations
                    srcs[pidx], coef[pidx] = off, 1.0
                jobs.append((tq * v3, srcs, coef))
            launches.append(S3GatherLaunch('ring place', self.R, self.Gs, v, 1.0, jobs))
        return launches


    # ------------------------------------------------- T3 -> r1/r2 and T3 -> W
    def _setup_lower(self, z_budget_bytes=16 << 30):
        """Plans of T3 -> r1/r2 (rccsdt.py:253-321) and T3 -> W (rccsdt.py:323-366).

        Z[Q][p][r] = P3_201(T[p, q0+Q, r]) is built for a chunk of middle indices q at a time (as many
        as ``z_budget_bytes`` allows; all of them at nocc=10, nvir=90), so every contraction folds the
        chunk into its M dimension.  Contractions whose M x N would leave the chip idle (M or N of one
        occupied or virtual index against K = o v^2) are split along K: the peeled K letters become
        batch entries writing partial slices, summed in a fixed order by dcc_sum_slices."""
        o, v = self.nocc, self.nvir
        v3 = v ** 3
        dt = self.dtype
        qc = int(max(1, min(o, z_budget_bytes // max(1, o * o * v3 * dt.itemsize))))
        lo = {'qc': qc}
        lo['Z'] = DeviceArray((qc, o, o, v, v, v), dt)
        lo['Ecomb'] = DeviceArray((o, o, o, v), dt)      # [l,r,p,a]
        lo['Evovv'] = DeviceArray((v, o, v, v), dt)      # [c,p,d,a] = E_vovv[c,p,a,d]
        lo['Eoovv'] = DeviceArray((o, o, v, v), dt)      # [r,p,d,e]
        lo['Eoovv_t'] = DeviceArray((o, o, v, v), dt)    # [m,p,e,d] = E_oovv[m,p,d,e]
        lo['Er1'] = DeviceArray((o, o, v, v, 1), dt)     # [j,k,b,c] = E - 1/2 E^(bc)
        lo['r1'] = DeviceArray((o, v, 1), dt)
        lo['r2'] = DeviceArray((o, o, v, v), dt)
        lo['P2'] = DeviceArray((o, qc, o, v, v), dt)     # split-K partials of r2 (E_vovv term)
        lo['P1'] = DeviceArray((qc * o, o, v, 1), dt)    # split-K partials of r1
        lo['P3'] = DeviceArray((o, qc, o, v, o), dt)     # split-K partials of W_vooo
        # T3 terms of W_vooo / W_vvvo, accumulated while Z is built for T3 -> r1/r2 and added to the
        # intermediates by t3_to_w, so each chunk of Z is built once per iteration
        lo['dWoovo'] = DeviceArray((o, o, v, o), dt)     # as W_oovo [i,j,a,l]
        lo['dWovvv'] = DeviceArray((o, v, v, v), dt)     # as W_ovvv [j,a,b,d]
        ext = {x: v for x in 'abcdexyz'}
        ext.update({'p': o, 'r': o, 'l': o, 'm': o, 'i': o, 'k': o, 'Q': qc, 'X': 1})
        Zs = (qc, o, o, v, v, v)
        Z = lambda letters: Operand(lo['Z'], letters, Zs)
        chunks = []
        for q0 in range(0, o, qc):
            nq = min(qc, o - q0)
            ext_c = dict(ext, Q=nq)
            jobs = []
            for Q in range(nq):
                for p in range(o):
                    for r in range(o):
                        case, (ci, cj, ck) = canonical_case(p, q0 + Q, r)
                        P = TP_T3[case]
                        coef = np.zeros(6)
                        for c, Qp in ((2.0, (0, 1, 2)), (-1.0, (2, 1, 0)), (-1.0, (1, 0, 2))):
                            comp = (P[Qp[0]], P[Qp[1]], P[Qp[2]])
                            coef[TP_T3.index(comp)] += c
                        jobs.append((((Q * o + p) * o + r) * v3, ijk_to_linear(ci, cj, ck, o) * v3, coef))
            # --- T3 -> r2, r1 ---
            a2 = []
            e_split = [({'p': p}, {'p': p}, {'S': p}) for p in range(o)]
            a2.append(plan_einsum('t3->r2 E_vovv q%d+%d' % (q0, nq), 'Qrdba,cda->Qrbc', Z('Qprdba'),
                                  Operand(lo['Evovv'], 'cpda', (v, o, v, v)),
                                  Operand(lo['P2'], 'SQrbc', (o, nq, o, v, v)), 1.0, 0.0, ext_c, e_split))
            a2.append(('sum', lo['r2'], q0 * o * v * v, lo['P2'], o, nq * o * v * v, 1.0))
            a2.append(plan_einsum('t3->r2 E_ooov q%d+%d' % (q0, nq), 'Qprabc,lrpa->Qlbc', Z('Qprabc'),
                                  Operand(lo['Ecomb'], 'lrpa', (o, o, o, v)),
                                  Operand(lo['r2'], 'Qlbc', (o, o, v, v)), 1.0, 1.0, ext_c,
                                  [({}, {}, {'Q': q0})]))
            r1_split = [({'Q': Q, 'k': k}, {'Q': q0 + Q, 'k': k}, {'S': Q * o + k})
                        for Q in range(nq) for k in range(o)]
            a2.append(plan_einsum('t3->r1 q%d+%d' % (q0, nq), 'iabc,bcX->iaX', Z('Qikabc'),
                                  Operand(lo['Er1'], 'QkbcX', (o, o, v, v, 1)),
                                  Operand(lo['P1'], 'SiaX', (qc * o, o, v, 1)), 1.0, 0.0, ext_c, r1_split))
            a2.append(('sum', lo['r1'], 0, lo['P1'], nq * o, o * v, 1.0))
            # --- T3 -> W_vooo, W_vvvo ---
            a3 = []
            a3.append(plan_einsum('t3->W_vooo q%d+%d' % (q0, nq), 'Qread,med->Qram', Z('Qpread'),
                                  Operand(lo['Eoovv_t'], 'mped', (o, o, v, v)),
                                  Operand(lo['P3'], 'SQram', (o, nq, o, v, o)), 1.0, 0.0, ext_c, e_split))
            a3.append(('sum', lo['dWoovo'], q0 * o * v * o, lo['P3'], o, nq * o * v * o, 1.0))
            a3.append(plan_einsum('t3->W_vvvo q%d+%d' % (q0, nq), 'Qpreba,rpde->Qabd', Z('Qpreba'),
                                  Operand(lo['Eoovv'], 'rpde', (o, o, v, v)),
                                  Operand(lo['dWovvv'], 'Qabd', (o, v, v, v)), -1.0, 1.0, ext_c,
                                  [({}, {}, {'Q': q0})]))
            self._retarget([g for g in a2 + a3 if isinstance(g, GemmLaunch)])
            chunks.append((DeviceArray.from_host(s3_jobs_array(jobs)), len(jobs), a2, a3))
        lo['chunks'] = chunks
        lo['dw_valid'] = False
        self._lower = lo

    def set_t1_dressed(self, t1_fock_ov, E_oovv, E_ooov, E_vovv):
        """Slices of the T1-dressed Fock/ERIs used by T3 -> r1/r2/W:
        t1_fock[:o, o:], t1_eris[:o,:o,o:,o:], t1_eris[:o,:o,:o,o:], t1_eris[o:,:o,o:,o:]."""
        if self._lower is None:
            self._setup_lower()
        lo, o = self._lower, self.nocc
        f = np.asarray(t1_fock_ov, dtype=np.float64)
        Eo = np.asarray(E_oovv, dtype=np.float64)
        comb = -np.asarray(E_ooov, dtype=np.float64).transpose(2, 0, 1, 3).copy()  # [l,r,p,a]
        for l in range(o):
            comb[l, l] += 0.5 * f
        dt = self.dtype
        lo['Ecomb'].upload(np.ascontiguousarray(comb, dtype=dt), self.stream)
        lo['Evovv'].upload(np.ascontiguousarray(np.asarray(E_vovv).transpose(0, 1, 3, 2), dtype=dt), self.stream)
        lo['Eoovv'].upload(np.ascontiguousarray(Eo, dtype=dt), self.stream)
        lo['Eoovv_t'].upload(np.ascontiguousarray(Eo.transpose(0, 1, 3, 2), dtype=dt), self.stream)
        lo['Er1'].upload(np.ascontiguousarray((Eo - 0.5 * Eo.transpose(0, 1, 3, 2))[..., None], dtype=dt),
                         self.stream)

    def set_t1_dressed_device(self, G, Fd, nmo):
        """Build the T3 -> r1/r2/W operands from device copies of the T1-dressed
        integrals G = t1_eris (nmo^4) and Fock Fd = t1_fock (nmo^2), both FP64."""
        if self._lower is None:
            self._setup_lower()
        lo, o, v, n = self._lower, self.nocc, self.nvir, int(nmo)
        g_str = {'P': n ** 3, 'Q': n ** 2, 'R': n, 'S': 1}
        fv = o  # first virtual orbital

        def gv(letters, starts):
            st = dict(zip(letters, (g_str[x] for x in 'PQRS')))
            return Operand.strided(G, st, sum(st[x] * starts.get(x, 0) for x in letters))

        def run(dst, src, letters, shape, alpha, beta):
            tensor_axpby(dst, src, letters, shape, alpha, beta, self.stream)

        Ecomb = Operand(lo['Ecomb'], 'lrpa', (o, o, o, v))
        run(Ecomb, gv('rpla', {'a': fv}), 'lrpa', (o, o, o, v), -1.0, 0.0)
        diag = Operand.strided(lo['Ecomb'], {'l': o * o * v + o * v, 'p': v, 'a': 1})
        run(diag, Operand.strided(Fd, {'l': 0, 'p': n, 'a': 1}, fv), 'lpa', (o, o, v), 0.5, 1.0)
        run(Operand(lo['Evovv'], 'cpda', (v, o, v, v)), gv('cpad', {'c': fv, 'a': fv, 'd': fv}), 'cpda',
            (v, o, v, v), 1.0, 0.0)
        oovv = gv('rpde', {'d': fv, 'e': fv})
        run(Operand(lo['Eoovv'], 'rpde', (o, o, v, v)), oovv, 'rpde', (o, o, v, v), 1.0, 0.0)
        run(Operand(lo['Eoovv_t'], 'rped', (o, o, v, v)), oovv, 'rpde', (o, o, v, v), 1.0, 0.0)
        Er1 = Operand(lo['Er1'], 'rpdeX', (o, o, v, v, 1))
        run(Er1, oovv, 'rpde', (o, o, v, v), 1.0, 0.0)
        run(Er1, Operand.strided(G, {'r': n ** 3, 'p': n ** 2, 'e': n, 'd': 1}, (fv * n + fv)), 'rpde',
            (o, o, v, v), -0.5, 1.0)

    def prepare_residual(self):
        """Ring weights, W_oooo/F_oo prefactors and the F_vv-folded W1 from the device intermediates."""
        o, v = self.nocc, self.nvir
        s = self.stream.handle if self.stream is not None else None
        v2, v3 = v * v, v ** 3
        L.call("dcc_memcpy_dtod", self.W1.vp, self.W_vvvv.vp, self.W1.nbytes, s)
        F = self.F_vv
        # W1[p,q,d,q] += 1/2 F[p,d];  W1[p,q,p,e] += 1/2 F[q,e]
        tensor_axpby(Operand.strided(self.W1, {'a': v3, 'b': v2 + 1, 'd': v}),
                     Operand.strided(F, {'a': v, 'd': 1}), 'abd', (v, v, v), 0.5, 1.0, self.stream)
        tensor_axpby(Operand.strided(self.W1, {'a': v3 + v, 'b': v2, 'e': 1}),
                     Operand.strided(F, {'b': v, 'e': 1}), 'abe', (v, v, v), 0.5, 1.0, self.stream)
        if self.vvvv_pairs:
            fn = "dcc_vvvv_pair_weights_" + ("packed_f32" if self.contract == "bf16x3" else self.sfx)
            L.call(fn, self.W1.vp, v, self.pair_dims[2], self.Wp.vp, s)
        L.call("dcc_ring_weights_" + self.sfx, self.W_ovov.vp, self.W_ovvo.vp, o, v, self.WR1.vp, self.WR3a.vp,
               self.WR3b.vp, s)
        L.call("dcc_oooo_prefactors_" + self.sfx, self.W_oooo.vp, self.F_oo.vp, o, self.blk_ijk.vp, self.n3,
               self.pfs.vp, s)

    def _build_z(self, c):
        jobs, njobs = self._lower['chunks'][c][:2]
        s = self.stream.handle if self.stream is not None else None
        L.call("dcc_s3_transform_" + self.sfx, self._lower['Z'].vp, self.t3.vp, self.nvir, 0.0, jobs.vp, njobs, s)

    def _run_lower(self, steps):
        s = self.stream.handle if self.stream is not None else None
        for g in steps:
            if isinstance(g, tuple):        # ('sum', dst, dst_offset, partials, nslices, n, alpha)
                _, dst, off, part, ns, n, alpha = g
                L.call("dcc_sum_slices_" + self.sfx, dst.offset_ptr(off), part.vp, ns, n, alpha, 1.0, s)
            else:
                g.run(self.stream)

    def t3_to_r1r2(self, download=True):
        """T3 contributions to r1 and the un-symmetrized r2 (rccsdt.py:253-321).

        The T3 terms of W_vooo / W_vvvo (rccsdt.py:323-366) are computed from the same chunks of Z
        and kept for t3_to_w.  Returns host arrays, or with ``download=False`` the device buffers
        r1 (o, v, 1) and r2 (o, o, v, v)."""
        lo = self._lower
        for buf in ('r1', 'r2', 'dWoovo', 'dWovvv'):
            lo[buf].fill(0.0, self.stream)
        for c, chunk in enumerate(lo['chunks']):
            self._build_z(c)
            self._run_lower(chunk[2])
            self._run_lower(chunk[3])
        lo['dw_valid'] = True
        if not download:
            return lo['r1'], lo['r2']
        return (lo['r1'].download(stream=self.stream).reshape(self.nocc, self.nvir).astype(np.float64),
                lo['r2'].download(stream=self.stream).astype(np.float64))

    def t3_to_r1r2_ops(self):
        """The device work of t3_to_r1r2(download=False) as a list of operations, each a list of steps
        (callables of the stream; the contractions are the GemmLaunch objects themselves), for callers
        that interleave the operations of several engines."""
        if self._lower is None:
            self._setup_lower()
        lo = self._lower
        st = self.stream
        s = lambda: st.handle if st is not None else None
        ops = [[lambda _s, b=lo[buf]: b.fill(0.0, st)] for buf in ('r1', 'r2', 'dWoovo', 'dWovvv')]
        for c, chunk in enumerate(lo['chunks']):
            ops.append([lambda _s, c=c: self._build_z(c)])
            for g in chunk[2] + chunk[3]:
                if isinstance(g, tuple):    # ('sum', dst, dst_offset, partials, nslices, n, alpha)
                    _, dst, off, part, ns, n, alpha = g
                    ops.append([lambda _s, dst=dst, off=off, part=part, ns=ns, n=n, alpha=alpha:
                                L.call("dcc_sum_slices_" + self.sfx, dst.offset_ptr(off), part.vp, ns, n, alpha, 1.0,
                                       s())])
                else:
                    ops.append([g])
        ops.append([lambda _s: lo.__setitem__('dw_valid', True)])
        return ops

    def _add_dw(self):
        s = self.stream.handle if self.stream is not None else None
        for dst, src in ((self.W_oovo, self._lower['dWoovo']), (self.W_ovvv, self._lower['dWovvv'])):
            lay = L.Strided()
            lay.ndim = 1
            lay.shape[0], lay.dst_stride[0], lay.src_stride[0] = src.size, 1, 1
            L.call("dcc_tensor_axpby_" + self.sfx, dst.vp, src.vp, ctypes.byref(lay), 1.0, 1.0, s)

    def t3_to_w(self, reuse_z=False):
        """Add the T3 terms of rccsdt.py:323-366 to the device W_vooo / W_vvvo (after set_intermediates).

        ``reuse_z`` adds the terms computed by the preceding t3_to_r1r2 when T3 and the dressed
        integrals have not changed since; otherwise they are computed here, chunk by chunk."""
        lo = self._lower
        if not (reuse_z and lo['dw_valid']):
            for buf in ('dWoovo', 'dWovvv'):
                lo[buf].fill(0.0, self.stream)
            for c, chunk in enumerate(lo['chunks']):
                self._build_z(c)
                self._run_lower(chunk[3])
        self._add_dw()
        lo['dw_valid'] = False

    def get_w_t3_adjusted(self):
        """Device W_vooo, W_vvvo (incl. T3 terms) back in the reference layouts."""
        w_oovo = self.W_oovo.download(stream=self.stream).astype(np.float64)
        w_ovvv = self.W_ovvv.download(stream=self.stream).astype(np.float64)
        return w_oovo.transpose(2, 3, 0, 1), w_ovvv.transpose(1, 2, 3, 0)

    # --------------------------------------------------------------- inputs
    def set_t3(self, t3_local):
        self.t3.upload(np.ascontiguousarray(t3_local, dtype=self.dtype), self.stream)

    def get_t3(self):
        return self.t3.download(stream=self.stream).astype(np.float64)

    def set_intermediates(self, imds, t2):
        o, v = self.nocc, self.nvir
        if self.vvvv_pairs:
            check_pair_symmetric(imds['W_vvvv'])
        for name in ('F_oo', 'F_vv', 'W_oooo', 'W_ovvo', 'W_ovov', 'W_vooo', 'W_vvvo', 'W_vvvv'):
            getattr(self, name).upload(np.ascontiguousarray(imds[name], dtype=self.dtype), self.stream)
        self.t2.upload(np.ascontiguousarray(t2, dtype=self.dtype), self.stream)
        self.W_ovvv.upload(np.ascontiguousarray(np.asarray(imds['W_vvvo']).transpose(3, 0, 1, 2), dtype=self.dtype),
                           self.stream)
        self.W_oovo.upload(np.ascontiguousarray(np.asarray(imds['W_vooo']).transpose(2, 3, 0, 1), dtype=self.dtype),
                           self.stream)
        self.prepare_residual()

    # -------------------------------------------------------------- compute
    def materialize_y(self):
        """Y (and Z) from T3; in the split-BF16x3 engine written directly as packed words."""
        s = self.stream.handle if self.stream is not None else None
        if self._yz_direct:
            call = lambda dst, jobs: L.call("dcc_s3_transform_pack_f32", dst.vp, self.t3.vp, self.nvir, jobs.vp,
                                            self.n_y_jobs, s)
        else:
            call = lambda dst, jobs: L.call("dcc_s3_transform_" + self.sfx, dst.vp, self.t3.vp, self.nvir, 0.0,
                                            jobs.vp, self.n_y_jobs, s)
        call(self.Y, self.y_jobs)
        if self.ring_fused:
            call(self.Z, self.z_jobs)

    def residual(self, timer=None):
        """R <- r3 of distr_cc/rccsdt.py:368-433 for the current t3 and intermediates.

        ``timer`` (optional) is called as timer(launch, 'start'|'end') around
        every GEMM launch, e.g. to record HIP events on the engine's stream.
        """
        self.R.fill(0.0, self.stream)
        self.materialize_y()
        if self._packs:
            self.pack_operands()
        for g in self.launches:
            if timer is not None:
                timer(g, 'start')
            g.run(self.stream)
            if timer is not None:
                timer(g, 'end')

    def set_denominators(self, mo_energy, level_shift=0.0, nvir_diis=None):
        """Upload eia = e_i - e_a - shift (rccsdt.py:737) and fix the active virtual count."""
        o, v = self.nocc, self.nvir
        e = np.asarray(mo_energy, dtype=np.float64)
        eia = e[:o, None] - e[None, o:] - level_shift
        self.eia.upload(np.ascontiguousarray(eia), self.stream)
        self.nvir_diis = v if nvir_diis is None else int(nvir_diis)

    def finalize_launch(self, r3_out=None):
        """Issue the fused projection/denominator/update and the norm reduction (no host sync)."""
        s = self.stream.handle if self.stream is not None else None
        L.call("dcc_r3_finalize_" + self.sfx, self.R.vp, self.t3.vp, r3_out.vp if r3_out is not None else None,
               self.n3, self.nvir, self.blk_ijk.vp, self.eia.vp, self.nocc, self.nvir_diis, self.partials.vp, s)
        L.call("dcc_reduce_pairs_f64", self.partials.vp, self.n3 * self.norb, self.norms.vp, s)

    def read_norms(self):
        n = self.norms.download(stream=self.stream)
        return float(n[0]), float(n[1])

    def finalize(self, mo_energy, level_shift=0.0, nvir_diis=None, r3_out=None):
        """rccsdt.py:947-971 fused: t3 += P3_full-projected R / D; returns (norm^2, active norm^2)."""
        self.set_denominators(mo_energy, level_shift, nvir_diis)
        self.finalize_launch(r3_out)
        return self.read_norms()

    def _tunable(self, include_lower=True):
        launches = list(self.launches)
        if include_lower and self._lower is not None:
            for chunk in self._lower['chunks']:
                launches += [g for g in chunk[2] + chunk[3] if isinstance(g, GemmLaunch)]
        return launches

    def tile_config(self):
        """Workgroup tile of every contraction launch, in autotune order (None for other launches)."""
        return [g.tile if isinstance(g, GemmLaunch) else None for g in self._tunable()]

    def set_tile_config(self, tiles):
        """Apply a tile_config() list (for example one saved from an earlier autotune)."""
        launches = self._tunable()
        if len(tiles) != len(launches):
            raise ValueError("tile list has %d entries, the engine %d launches" % (len(tiles), len(launches)))
        for g, t in zip(launches, tiles):
            if isinstance(g, GemmLaunch) and t is not None:
                g.tile = int(t)

    def tiles_by_name(self):
        """{launch name: tile} of the contraction launches (for transfer to an engine of another size);
        of several launches with one name (chunks of a term) the one with the most flops decides."""
        best = {}
        for g in self._tunable():
            if isinstance(g, GemmLaunch) and (g.name not in best or g.flops > best[g.name][0]):
                best[g.name] = (g.flops, g.tile)
        return {k: t for k, (_, t) in best.items()}

    def set_tiles_by_name(self, tiles):
        """Apply a tiles_by_name() map; launches without an entry keep their tile.  A split-BF16x3
        tile side of 256 (4 and 9: 256x128, 5 and 10: 128x256, 6, 7, 8, 11, 12 and 13: 256x256) is halved to 128
        on a launch whose
        extent on that side is at most 128 here (a chunked launch of a larger system), where it would
        be mostly padding."""
        for g in self._tunable():
            if isinstance(g, GemmLaunch) and g.name in tiles:
                t = int(tiles[g.name])
                if g.fn.endswith("bf16x3") and t >= 4:
                    wide_m = t in (4, 6, 7, 8, 9, 11, 12, 13) and g.max_rows > 128
                    wide_n = t in (5, 6, 7, 8, 10, 11, 12, 13) and g.args[1] > 128
                    t = {(True, True): t if t in (6, 7, 8, 11, 12, 13) else 6, (True, False): 9 if t == 9 else 4,
                         (False, True): 10 if t == 10 else 5, (False, False): 3}[(wide_m, wide_n)]
                g.tile = t

    def autotune(self, reps=2, include_lower=True):
        """Time the workgroup tiles of every contraction launch and keep the fastest: eight for the
        FP64/FP32 kernels (four shapes x one- or two-stage prefetch), twelve for split-BF16x3
        (the four of 256 threads and 256x128, 128x256, 256x256 of 512 threads on 16 x 16 x 32 MFMAs,
        256x256 on 32 x 32 x 16 MFMAs, and four of those staged by direct global -> LDS requests: 256x256
        with two stages, 256x128 and 128x256 with three, 256x256 on 32 x 32 x 16 with two, and the
        256x256 tile on four waves with AGPR accumulators on either MFMA shape; a launch whose operands
        do not allow the direct requests runs the register-staged tile instead).

        Run it before a residual (it accumulates into R and the scratch buffers); returns
        {launch name: (tile, ms)}."""
        from distr_cc.device import Event
        shapes = ("64x64", "128x64", "64x128", "128x128", "64x64s1", "128x64s1", "64x128s1", "128x128s1")
        shapes3 = ("64x64", "128x64", "64x128", "128x128", "256x128", "128x256", "256x256", "256x256m32",
                   "256x256lds2", "256x128lds3", "128x256lds3", "256x256m32lds2", "256x256w4", "256x256m32w4")
        launches = self._tunable(include_lower)
        out = {}
        st = self.stream
        for n, g in enumerate(launches):
            if not isinstance(g, GemmLaunch):
                continue
            best = None
            names = shapes3 if g.fn.endswith("bf16x3") else shapes
            for tile in range(len(names)):
                g.tile = tile
                g.run(st)
                a, b = Event(), Event()
                a.record(st)
                for _ in range(reps):
                    g.run(st)
                b.record(st)
                ms = a.elapsed_ms(b) / reps
                if best is None or ms < best[1]:
                    best = (tile, ms)
            g.tile = best[0]
            out["%d %s" % (n, g.name)] = (names[best[0]], best[1])
        return out

    def residual_flops(self):
        return sum(g.flops for g in self.launches)
