"""(Q)/[Q] energy of RCCSDT (oracle restatement, test infrastructure only).

Follows ``distr_cc/rccsdt_q.py`` at one rank:
  * task list over virtual blocks A<=B<=C<=D   (rccsdt_q.py:302-314)
  * job splitting job_idx / n_jobs              (rccsdt_q.py:364-383)
  * W_vvvvoo / W_vvoooo builders                (rccsdt_q.py:390-408)
  * factor 24/12/6/0                            (rccsdt_q.py:410-428)
  * per-task body: 24 z4_p einsums, 12 + 12 t4 einsums, t4 += z4_p,
    denominator with the 1e-15 -> 0 rule, P4_444 on the occupied indices,
    factor, paren/bracket dot products         (rccsdt_q.py:1001-1160)
The t3 blocks the reference gathers with ``fill_t3_from_ptr_array``
(src/rccsdt_q.c:665-828) are the dense T[i,j,k][a,b,c] for a in A, b in B and
every c; here they are slices of the dense tensor in (a,b,c,i,j,k) order.
``t4_add_`` (PySCF libccsdt, absent) is the elementwise t4 += r4 its CCSDTQ
call site implies (rccsdtq.py:1463).

Pins (tests/test_oracle_q.py): blksize invariance and job-split additivity
(the reference's own tests, tests/test_rccsdt_q.py:56-94), the reference C
helpers for division / factor / P4_444 (oracle/_ref), and an independent dense
restatement that sums over all (a,b,c,d) without task blocking.
"""
import numpy as np

from oracle.spin_sum import t4_spin_sum


def einsum(sub, a, b, out, alpha=1.0, beta=1.0):
    res = np.einsum(sub, a, b, optimize=True)
    if beta == 0.0:
        out[...] = alpha * res
    else:
        out *= beta
        out += alpha * res
    return out


def prange(start, stop, step):
    """pyscf.lib.prange semantics: (i, min(i + step, stop)) for i in range(start, stop, step)."""
    for i in range(start, stop, step):
        yield i, min(i + step, stop)


def create_full_task_list(nvir, blksize):
    """rccsdt_q.py:302-314."""
    tasks = []
    for d0, d1 in prange(0, nvir, blksize):
        for c0, c1 in prange(0, d1, blksize):
            for b0, b1 in prange(0, c1, blksize):
                for a0, a1 in prange(0, b1, blksize):
                    tasks.append((a0, a1, a1 - a0, b0, b1, b1 - b0, c0, c1, c1 - c0, d0, d1, d1 - d0))
    return tasks


def job_tasks(tasks, job_idx, n_jobs):
    """rccsdt_q.py:364-367 (one rank)."""
    per = (len(tasks) + n_jobs - 1) // n_jobs
    start = job_idx * per
    return tasks[start:min(start + per, len(tasks))]


def get_factor(a0, a1, b0, b1, c0, c1, d0, d1, blksize):
    """rccsdt_q.py:410-428."""
    a = np.arange(a0, a1)[:, None, None, None]
    b = np.arange(b0, b1)[None, :, None, None]
    c = np.arange(c0, c1)[None, None, :, None]
    d = np.arange(d0, d1)[None, None, None, :]
    f_ = np.zeros((a1 - a0, b1 - b0, c1 - c0, d1 - d0))
    f_[(a < b) & (b < c) & (c < d)] = 24.0
    f_[((a == b) & (b < c) & (c < d)) | ((a < b) & (b == c) & (c < d)) | ((a < b) & (b < c) & (c == d))] = 12.0
    f_[(a == b) & (b < c) & (c == d)] = 6.0
    f = np.zeros((blksize,) * 4)
    f[:a1 - a0, :b1 - b0, :c1 - c0, :d1 - d0] = f_
    return f


def e_division(r4, e_occ, e_vir, a0, a1, b0, b1, c0, c1, d0, d1):
    """src/rccsdt_q.c:474-588 on the valid (non-padded) sub-block."""
    ei = e_occ
    eo = ei[:, None, None, None] + ei[None, :, None, None] + ei[None, None, :, None] + ei[None, None, None, :]
    for ia, a in enumerate(range(a0, a1)):
        for ib, b in enumerate(range(b0, b1)):
            for ic, c in enumerate(range(c0, c1)):
                for id_, d in enumerate(range(d0, d1)):
                    den = eo - (e_vir[a] + e_vir[b] + e_vir[c] + e_vir[d])
                    blk = r4[ia, ib, ic, id_]
                    ok = np.abs(den) > 1e-15
                    blk[ok] = blk[ok] / den[ok]
                    blk[~ok] = 0.0
    return r4


def kernel(nocc, nvir, pppp, t2_oovv, T_dense, mo_energy, blksize=4, job_idx=0, n_jobs=1):
    """rccsdt_q.py:227-1191 at one rank -> (e_[Q], e_(Q)).

    ``pppp`` physicist <pq|rs>; ``t2_oovv`` (o,o,v,v); ``T_dense`` the dense
    ordered-triple T3, shape (o,o,o,v,v,v).
    """
    o, v, bs = nocc, nvir, blksize
    t2 = np.ascontiguousarray(t2_oovv.transpose(2, 3, 0, 1))      # rccsdt_q.py:42-50
    T = np.ascontiguousarray(T_dense.transpose(3, 4, 5, 0, 1, 2))  # T[a,b,c,i,j,k]
    e_occ = np.ascontiguousarray(mo_energy[:o])
    e_vir = np.ascontiguousarray(mo_energy[o:])
    eris_vvvo = pppp[o:, o:, o:, :o].copy()                      # rccsdt_q.py:288-294
    eris_vooo = pppp[o:, :o, :o, :o].copy()
    eris_vvoo = pppp[o:, o:, :o, :o].copy()
    eris_ovvo = pppp[:o, o:, o:, :o].copy()
    eris_ovov = pppp[:o, o:, :o, o:].copy()
    eris_vvvv = pppp[o:, o:, o:, o:].copy()
    eris_oooo = pppp[:o, :o, :o, :o].copy()

    def W_vvvvoo(W, a0, a1, b0, b1, c0, c1):                     # rccsdt_q.py:390-400
        ba, bb, bc = a1 - a0, b1 - b0, c1 - c0
        einsum('abef,fcjk->abcejk', eris_vvvv[a0:a1, b0:b1], t2[:, c0:c1], W[:ba, :bb, :bc], 0.5, 0.0)
        einsum('acef,fbkj->abcejk', eris_vvvv[a0:a1, c0:c1], t2[:, b0:b1], W[:ba, :bb, :bc], 0.5, 1.0)
        einsum('mbej,acmk->abcejk', eris_ovvo[:, b0:b1], t2[a0:a1, c0:c1], W[:ba, :bb, :bc], -1.0, 1.0)
        einsum('mcek,abmj->abcejk', eris_ovvo[:, c0:c1], t2[a0:a1, b0:b1], W[:ba, :bb, :bc], -1.0, 1.0)
        einsum('maje,bcmk->abcejk', eris_ovov[:, a0:a1], t2[b0:b1, c0:c1], W[:ba, :bb, :bc], -1.0, 1.0)
        einsum('make,cbmj->abcejk', eris_ovov[:, a0:a1], t2[c0:c1, b0:b1], W[:ba, :bb, :bc], -1.0, 1.0)
        return W

    def W_vvoooo(W, a0, a1, b0, b1):                             # rccsdt_q.py:402-408
        ba, bb = a1 - a0, b1 - b0
        einsum('mnki,abnj->abmijk', eris_oooo, t2[a0:a1, b0:b1], W[:ba, :bb], -0.5, 0.0)
        einsum('mnkj,bani->abmijk', eris_oooo, t2[b0:b1, a0:a1], W[:ba, :bb], -0.5, 1.0)
        return W

    def fill(a0, a1, b0, b1):
        blk = np.zeros((bs, bs, v, o, o, o))
        blk[:a1 - a0, :b1 - b0] = T[a0:a1, b0:b1]
        return blk

    e_bracket = 0.0
    e_paren = 0.0
    for task in job_tasks(create_full_task_list(v, bs), job_idx, n_jobs):
        (a0, a1, ba, b0, b1, bb, c0, c1, bc, d0, d1, bd) = task
        t4 = np.zeros((bs,) * 4 + (o,) * 4)
        z4 = np.zeros_like(t4)
        Wv = np.zeros((bs,) * 3 + (v,) + (o,) * 2)
        Wo = np.zeros((bs,) * 2 + (o,) * 4)
        zz = z4[:ba, :bb, :bc, :bd]
        t3b = fill(a0, a1, b0, b1)                               # rccsdt_q.py:1014-1025
        einsum('cdel,abeijk->abcdijkl', eris_vvvo[c0:c1, d0:d1], t3b[:ba, :bb], zz, 1.0, 0.0)
        einsum('dcek,abeijl->abcdijkl', eris_vvvo[d0:d1, c0:c1], t3b[:ba, :bb], zz, 1.0, 1.0)
        einsum('dmli,abcmjk->abcdijkl', eris_vooo[d0:d1], t3b[:ba, :bb, c0:c1], zz, -1.0, 1.0)
        einsum('dmlj,abcimk->abcdijkl', eris_vooo[d0:d1], t3b[:ba, :bb, c0:c1], zz, -1.0, 1.0)
        einsum('cmki,abdmjl->abcdijkl', eris_vooo[c0:c1], t3b[:ba, :bb, d0:d1], zz, -1.0, 1.0)
        einsum('cmkj,abdiml->abcdijkl', eris_vooo[c0:c1], t3b[:ba, :bb, d0:d1], zz, -1.0, 1.0)
        t3b = fill(a0, a1, c0, c1)                               # :1027-1036
        einsum('bdel,aceikj->abcdijkl', eris_vvvo[b0:b1, d0:d1], t3b[:ba, :bc], zz, 1.0, 1.0)
        einsum('dbej,aceikl->abcdijkl', eris_vvvo[d0:d1, b0:b1], t3b[:ba, :bc], zz, 1.0, 1.0)
        einsum('bmji,acdmkl->abcdijkl', eris_vooo[b0:b1], t3b[:ba, :bc, d0:d1], zz, -1.0, 1.0)
        einsum('bmjk,acdiml->abcdijkl', eris_vooo[b0:b1], t3b[:ba, :bc, d0:d1], zz, -1.0, 1.0)
        einsum('dmlk,acbimj->abcdijkl', eris_vooo[d0:d1], t3b[:ba, :bc, b0:b1], zz, -1.0, 1.0)
        t3b = fill(a0, a1, d0, d1)                               # :1038-1045
        einsum('bcek,adeilj->abcdijkl', eris_vvvo[b0:b1, c0:c1], t3b[:ba, :bd], zz, 1.0, 1.0)
        einsum('cbej,adeilk->abcdijkl', eris_vvvo[c0:c1, b0:b1], t3b[:ba, :bd], zz, 1.0, 1.0)
        einsum('bmjl,adcimk->abcdijkl', eris_vooo[b0:b1], t3b[:ba, :bd, c0:c1], zz, -1.0, 1.0)
        einsum('cmkl,adbimj->abcdijkl', eris_vooo[c0:c1], t3b[:ba, :bd, b0:b1], zz, -1.0, 1.0)
        t3b = fill(b0, b1, c0, c1)                               # :1048-1055
        einsum('adel,bcejki->abcdijkl', eris_vvvo[a0:a1, d0:d1], t3b[:bb, :bc], zz, 1.0, 1.0)
        einsum('daei,bcejkl->abcdijkl', eris_vvvo[d0:d1, a0:a1], t3b[:bb, :bc], zz, 1.0, 1.0)
        einsum('amij,bcdmkl->abcdijkl', eris_vooo[a0:a1], t3b[:bb, :bc, d0:d1], zz, -1.0, 1.0)
        einsum('amik,bcdjml->abcdijkl', eris_vooo[a0:a1], t3b[:bb, :bc, d0:d1], zz, -1.0, 1.0)
        t3b = fill(b0, b1, d0, d1)                               # :1058-1064
        einsum('acek,bdejli->abcdijkl', eris_vvvo[a0:a1, c0:c1], t3b[:bb, :bd], zz, 1.0, 1.0)
        einsum('caei,bdejlk->abcdijkl', eris_vvvo[c0:c1, a0:a1], t3b[:bb, :bd], zz, 1.0, 1.0)
        einsum('amil,bdcjmk->abcdijkl', eris_vooo[a0:a1], t3b[:bb, :bd, c0:c1], zz, -1.0, 1.0)
        t3b = fill(c0, c1, d0, d1)                               # :1066-1070
        einsum('abej,cdekli->abcdijkl', eris_vvvo[a0:a1, b0:b1], t3b[:bc, :bd], zz, 1.0, 1.0)
        einsum('baei,cdeklj->abcdijkl', eris_vvvo[b0:b1, a0:a1], t3b[:bc, :bd], zz, 1.0, 1.0)
        tt = t4[:ba, :bb, :bc, :bd]                               # :1074-1102
        W_vvvvoo(Wv, a0, a1, b0, b1, c0, c1)
        einsum('abcejk,edil->abcdijkl', Wv[:ba, :bb, :bc], t2[:, d0:d1], tt, 1.0, 0.0)
        W_vvvvoo(Wv, a0, a1, b0, b1, d0, d1)
        einsum('abdejl,ecik->abcdijkl', Wv[:ba, :bb, :bd], t2[:, c0:c1], tt, 1.0, 1.0)
        W_vvvvoo(Wv, a0, a1, c0, c1, d0, d1)
        einsum('acdekl,ebij->abcdijkl', Wv[:ba, :bc, :bd], t2[:, b0:b1], tt, 1.0, 1.0)
        W_vvvvoo(Wv, b0, b1, a0, a1, c0, c1)
        einsum('baceik,edjl->abcdijkl', Wv[:bb, :ba, :bc], t2[:, d0:d1], tt, 1.0, 1.0)
        W_vvvvoo(Wv, b0, b1, a0, a1, d0, d1)
        einsum('badeil,ecjk->abcdijkl', Wv[:bb, :ba, :bd], t2[:, c0:c1], tt, 1.0, 1.0)
        W_vvvvoo(Wv, b0, b1, c0, c1, d0, d1)
        einsum('bcdekl,eaji->abcdijkl', Wv[:bb, :bc, :bd], t2[:, a0:a1], tt, 1.0, 1.0)
        W_vvvvoo(Wv, c0, c1, a0, a1, b0, b1)
        einsum('cabeij,edkl->abcdijkl', Wv[:bc, :ba, :bb], t2[:, d0:d1], tt, 1.0, 1.0)
        W_vvvvoo(Wv, c0, c1, a0, a1, d0, d1)
        einsum('cadeil,ebkj->abcdijkl', Wv[:bc, :ba, :bd], t2[:, b0:b1], tt, 1.0, 1.0)
        W_vvvvoo(Wv, c0, c1, b0, b1, d0, d1)
        einsum('cbdejl,eaki->abcdijkl', Wv[:bc, :bb, :bd], t2[:, a0:a1], tt, 1.0, 1.0)
        W_vvvvoo(Wv, d0, d1, a0, a1, b0, b1)
        einsum('dabeij,eclk->abcdijkl', Wv[:bd, :ba, :bb], t2[:, c0:c1], tt, 1.0, 1.0)
        W_vvvvoo(Wv, d0, d1, a0, a1, c0, c1)
        einsum('daceik,eblj->abcdijkl', Wv[:bd, :ba, :bc], t2[:, b0:b1], tt, 1.0, 1.0)
        W_vvvvoo(Wv, d0, d1, b0, b1, c0, c1)
        einsum('dbcejk,eali->abcdijkl', Wv[:bd, :bb, :bc], t2[:, a0:a1], tt, 1.0, 1.0)
        W_vvoooo(Wo, a0, a1, b0, b1)                              # :1106-1129
        einsum('abmijk,cdml->abcdijkl', Wo[:ba, :bb], t2[c0:c1, d0:d1], tt, -1.0, 1.0)
        einsum('abmijl,dcmk->abcdijkl', Wo[:ba, :bb], t2[d0:d1, c0:c1], tt, -1.0, 1.0)
        W_vvoooo(Wo, a0, a1, c0, c1)
        einsum('acmikj,bdml->abcdijkl', Wo[:ba, :bc], t2[b0:b1, d0:d1], tt, -1.0, 1.0)
        einsum('acmikl,dbmj->abcdijkl', Wo[:ba, :bc], t2[d0:d1, b0:b1], tt, -1.0, 1.0)
        W_vvoooo(Wo, a0, a1, d0, d1)
        einsum('admilj,bcmk->abcdijkl', Wo[:ba, :bd], t2[b0:b1, c0:c1], tt, -1.0, 1.0)
        einsum('admilk,cbmj->abcdijkl', Wo[:ba, :bd], t2[c0:c1, b0:b1], tt, -1.0, 1.0)
        W_vvoooo(Wo, b0, b1, c0, c1)
        einsum('bcmjki,adml->abcdijkl', Wo[:bb, :bc], t2[a0:a1, d0:d1], tt, -1.0, 1.0)
        einsum('bcmjkl,dami->abcdijkl', Wo[:bb, :bc], t2[d0:d1, a0:a1], tt, -1.0, 1.0)
        W_vvoooo(Wo, b0, b1, d0, d1)
        einsum('bdmjli,acmk->abcdijkl', Wo[:bb, :bd], t2[a0:a1, c0:c1], tt, -1.0, 1.0)
        einsum('bdmjlk,cami->abcdijkl', Wo[:bb, :bd], t2[c0:c1, a0:a1], tt, -1.0, 1.0)
        W_vvoooo(Wo, c0, c1, d0, d1)
        einsum('cdmkli,abmj->abcdijkl', Wo[:bc, :bd], t2[a0:a1, b0:b1], tt, -1.0, 1.0)
        einsum('cdmklj,bami->abcdijkl', Wo[:bc, :bd], t2[b0:b1, a0:a1], tt, -1.0, 1.0)
        t4 += z4                                                  # t4_add_ (:1131)
        e_division(t4, e_occ, e_vir, a0, a1, b0, b1, c0, c1, d0, d1)   # :1132
        t4 = t4_spin_sum(t4, "P4_444", 1.0, 0.0)                   # :1133
        t4 *= get_factor(a0, a1, b0, b1, c0, c1, d0, d1, bs)[..., None, None, None, None]  # :1136-1137
        paren = np.dot(t4.ravel(), z4.ravel()) / 12.0             # :1140
        einsum('abij,cdkl->abcdijkl', eris_vvoo[a0:a1, b0:b1], t2[c0:c1, d0:d1], zz, 1.0, 0.0)   # :1143-1150
        einsum('acik,bdjl->abcdijkl', eris_vvoo[a0:a1, c0:c1], t2[b0:b1, d0:d1], zz, 1.0, 1.0)
        einsum('adil,bcjk->abcdijkl', eris_vvoo[a0:a1, d0:d1], t2[b0:b1, c0:c1], zz, 1.0, 1.0)
        einsum('bcjk,adil->abcdijkl', eris_vvoo[b0:b1, c0:c1], t2[a0:a1, d0:d1], zz, 1.0, 1.0)
        einsum('bdjl,acik->abcdijkl', eris_vvoo[b0:b1, d0:d1], t2[a0:a1, c0:c1], zz, 1.0, 1.0)
        einsum('cdkl,abij->abcdijkl', eris_vvoo[c0:c1, d0:d1], t2[a0:a1, b0:b1], zz, 1.0, 1.0)
        bracket = np.dot(t4.ravel(), z4.ravel()) / 12.0           # :1153
        e_bracket += bracket
        e_paren += paren + bracket
    return e_bracket, e_paren
