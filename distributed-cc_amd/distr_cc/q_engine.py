"""GPU (Q)/[Q] energy of RCCSDT (the computation of distr_cc/rccsdt_q.py:227-1191).

Layout in HBM (FP64, virtual dimension zero-padded to vp = ceil(v/bs)*bs):
  T_abc    (vp, vp, vp, o, o, o)   dense T3 in (a,b,c,i,j,k) order -- the
                                   one-GPU form of the IJK -> ABC redistribution
  W_vvoooo (vp, vp, o, o, o, o)    built once for all tasks
  W_vvvvoo either dense (vp, vp, vp, vp, o, o), built once, or -- when that
           would exceed ``dense_w_bytes`` -- per chunk of tasks as slots
           (nslot, bs, bs, bs, vp, o, o), one per distinct ordered triple of
           the chunk's virtual blocks (the reference builds the block slices
           per task, rccsdt_q.py:390-408,943-947)
  z4, t4   (ntask_chunk, bs, bs, bs, bs, o, o, o, o)
Each of the 24 z4_p einsums and 24 t4 einsums of a task (rccsdt_q.py:1014-1129)
is one batched gather-GEMM over a chunk of tasks; a fused kernel then applies
t4 += z4_p, the denominator, P4_444, the factor and both dot products
(rccsdt_q.py:1131-1153), writing per-workgroup partials that are summed in a
fixed order into one pair per chunk; the host reads the pairs once at the end.
For nocc > 9 the fused kernel keeps its o^4 intermediate in a global workspace
of the size of t4 instead of LDS.
"""
import numpy as np

from distr_cc import _lib as L
from distr_cc.device import DeviceArray, init as device_init
from distr_cc.t3_engine import Operand, plan_einsum

# (subscripts, A operand, B operand, alpha) in the order of rccsdt_q.py:1017-1070
Z4_TERMS = [
    ('cdel,abeijk->abcdijkl', 'vvvo', 'T', 1.0), ('dcek,abeijl->abcdijkl', 'vvvo', 'T', 1.0),
    ('dmli,abcmjk->abcdijkl', 'vooo', 'T', -1.0), ('dmlj,abcimk->abcdijkl', 'vooo', 'T', -1.0),
    ('cmki,abdmjl->abcdijkl', 'vooo', 'T', -1.0), ('cmkj,abdiml->abcdijkl', 'vooo', 'T', -1.0),
    ('bdel,aceikj->abcdijkl', 'vvvo', 'T', 1.0), ('dbej,aceikl->abcdijkl', 'vvvo', 'T', 1.0),
    ('bmji,acdmkl->abcdijkl', 'vooo', 'T', -1.0), ('bmjk,acdiml->abcdijkl', 'vooo', 'T', -1.0),
    ('dmlk,acbimj->abcdijkl', 'vooo', 'T', -1.0),
    ('bcek,adeilj->abcdijkl', 'vvvo', 'T', 1.0), ('cbej,adeilk->abcdijkl', 'vvvo', 'T', 1.0),
    ('bmjl,adcimk->abcdijkl', 'vooo', 'T', -1.0), ('cmkl,adbimj->abcdijkl', 'vooo', 'T', -1.0),
    ('adel,bcejki->abcdijkl', 'vvvo', 'T', 1.0), ('daei,bcejkl->abcdijkl', 'vvvo', 'T', 1.0),
    ('amij,bcdmkl->abcdijkl', 'vooo', 'T', -1.0), ('amik,bcdjml->abcdijkl', 'vooo', 'T', -1.0),
    ('acek,bdejli->abcdijkl', 'vvvo', 'T', 1.0), ('caei,bdejlk->abcdijkl', 'vvvo', 'T', 1.0),
    ('amil,bdcjmk->abcdijkl', 'vooo', 'T', -1.0),
    ('abej,cdekli->abcdijkl', 'vvvo', 'T', 1.0), ('baei,cdeklj->abcdijkl', 'vvvo', 'T', 1.0),
]
# rccsdt_q.py:1075-1129
T4_TERMS = [
    ('abcejk,edil->abcdijkl', 'Wv', 't2', 1.0), ('abdejl,ecik->abcdijkl', 'Wv', 't2', 1.0),
    ('acdekl,ebij->abcdijkl', 'Wv', 't2', 1.0), ('baceik,edjl->abcdijkl', 'Wv', 't2', 1.0),
    ('badeil,ecjk->abcdijkl', 'Wv', 't2', 1.0), ('bcdekl,eaji->abcdijkl', 'Wv', 't2', 1.0),
    ('cabeij,edkl->abcdijkl', 'Wv', 't2', 1.0), ('cadeil,ebkj->abcdijkl', 'Wv', 't2', 1.0),
    ('cbdejl,eaki->abcdijkl', 'Wv', 't2', 1.0), ('dabeij,eclk->abcdijkl', 'Wv', 't2', 1.0),
    ('daceik,eblj->abcdijkl', 'Wv', 't2', 1.0), ('dbcejk,eali->abcdijkl', 'Wv', 't2', 1.0),
    ('abmijk,cdml->abcdijkl', 'Wo', 't2', -1.0), ('abmijl,dcmk->abcdijkl', 'Wo', 't2', -1.0),
    ('acmikj,bdml->abcdijkl', 'Wo', 't2', -1.0), ('acmikl,dbmj->abcdijkl', 'Wo', 't2', -1.0),
    ('admilj,bcmk->abcdijkl', 'Wo', 't2', -1.0), ('admilk,cbmj->abcdijkl', 'Wo', 't2', -1.0),
    ('bcmjki,adml->abcdijkl', 'Wo', 't2', -1.0), ('bcmjkl,dami->abcdijkl', 'Wo', 't2', -1.0),
    ('bdmjli,acmk->abcdijkl', 'Wo', 't2', -1.0), ('bdmjlk,cami->abcdijkl', 'Wo', 't2', -1.0),
    ('cdmkli,abmj->abcdijkl', 'Wo', 't2', -1.0), ('cdmklj,bami->abcdijkl', 'Wo', 't2', -1.0),
]
# dense W builders (rccsdt_q.py:390-408) over the full virtual range
WV_TERMS = [('abef,fcjk->abcejk', 'vvvv', 't2', 0.5), ('acef,fbkj->abcejk', 'vvvv', 't2', 0.5),
            ('mbej,acmk->abcejk', 'ovvo', 't2', -1.0), ('mcek,abmj->abcejk', 'ovvo', 't2', -1.0),
            ('maje,bcmk->abcejk', 'ovov', 't2', -1.0), ('make,cbmj->abcejk', 'ovov', 't2', -1.0)]
WO_TERMS = [('mnki,abnj->abmijk', 'oooo', 't2', -0.5), ('mnkj,bani->abmijk', 'oooo', 't2', -0.5)]


def create_full_task_list(nvir, blksize):
    """Task order of rccsdt_q.py:302-314 (virtual block quadruples A<=B<=C<=D)."""
    tasks = []
    for d0 in range(0, nvir, blksize):
        d1 = min(d0 + blksize, nvir)
        for c0 in range(0, d1, blksize):
            c1 = min(c0 + blksize, d1)
            for b0 in range(0, c1, blksize):
                b1 = min(b0 + blksize, c1)
                for a0 in range(0, b1, blksize):
                    tasks.append((a0, b0, c0, d0))
    return tasks


def _pad(x, axes_v, vp, o):
    shape = tuple(vp if ax in axes_v else s for ax, s in enumerate(x.shape))
    out = np.zeros(shape)
    out[tuple(slice(0, s) for s in x.shape)] = x
    return np.ascontiguousarray(out)


class QEngine:
    """Device state of (Q)/[Q] evaluations for one (nocc, nvir, blksize).

    The device buffers and the launch plans (index maps and batch tables of every GEMM) are built on
    the first kernel() call and reused by later calls of the same engine, which only upload the new
    integrals and amplitudes and replay the launches."""

    def __init__(self, nocc, nvir, blksize=4, stream=None, device=0, max_chunk_bytes=1 << 30,
                 dense_w_bytes=8 << 30, w_mode=None, finalize_ws=None):
        device_init(device)
        self.o, self.v, self.bs = int(nocc), int(nvir), int(blksize)
        if self.bs < 1:
            raise ValueError("blksize must be positive")
        self.vp = -(-self.v // self.bs) * self.bs
        self.stream = stream
        o, vp, bs = self.o, self.vp, self.bs
        per_task = bs ** 4 * o ** 4 * 8
        self.chunk = max(1, min(int(max_chunk_bytes // per_task), 4096))
        if w_mode is None:
            w_mode = "dense" if vp ** 4 * o * o * 8 <= dense_w_bytes else "task"
        if w_mode not in ("dense", "task"):
            raise ValueError("w_mode must be 'dense' or 'task'")
        self.w_mode = w_mode
        # the fused finalize keeps two o^4 blocks in LDS up to nocc = 9, a global workspace beyond
        self.finalize_ws = (o > 9) if finalize_ws is None else bool(finalize_ws)
        if o > 9 and not self.finalize_ws:
            raise ValueError("nocc > 9 needs the workspace form of the (Q) finalize")
        self.dev = None
        self.wplans = None
        self.task_plans = {}

    def _s(self):
        return self.stream.handle if self.stream is not None else None

    def _op(self, name, letters):
        shapes = {'vvvo': 'vvvo', 'vooo': 'vooo', 'vvoo': 'vvoo', 'ovvo': 'ovvo', 'ovov': 'ovov',
                  'vvvv': 'vvvv', 'oooo': 'oooo', 't2': 'vvoo', 'T': 'vvvooo', 'Wv': 'vvvvoo', 'Wo': 'vvoooo'}
        shp = tuple(self.vp if ch == 'v' else self.o for ch in shapes[name])
        return Operand(self.dev[name], letters, shp)

    def _setup(self):
        o, vp = self.o, self.vp
        shp = {'vvvo': (vp, vp, vp, o), 'vooo': (vp, o, o, o), 'vvoo': (vp, vp, o, o), 'ovvo': (o, vp, vp, o),
               'ovov': (o, vp, o, vp), 'vvvv': (vp, vp, vp, vp), 'oooo': (o, o, o, o), 't2': (vp, vp, o, o),
               'e_occ': (o,), 'e_vir': (vp,), 'T': (vp, vp, vp, o, o, o), 'Wo': (vp, vp, o, o, o, o)}
        if self.w_mode == "dense":
            shp['Wv'] = (vp, vp, vp, vp, o, o)
        self.dev = {k: DeviceArray(s) for k, s in shp.items()}
        full = {x: vp for x in 'abcdef'}
        full.update({x: o for x in 'ijklmn'})
        # dense W_vvvvoo / W_vvoooo (beta = 0 on the first term of each)
        self.wplans = []
        for terms, out in ((WV_TERMS, 'Wv'), (WO_TERMS, 'Wo')):
            if out not in self.dev:
                continue
            for n, (sub, an, bn, alpha) in enumerate(terms):
                sa, rest = sub.split(',')
                sb, sc = rest.split('->')
                self.wplans.append(plan_einsum('Q W', sub, self._op(an, sa), self._op(bn, sb), self._op(out, sc),
                                               alpha, 0.0 if n == 0 else 1.0, full, [({}, {}, {})]))

    def _slot_launches(self, Ws, nslot_max, keys):
        """W_vvvvoo slots Ws[s] = W[X:X+bs, Y:Y+bs, Z:Z+bs, :, :, :] for keys[s] = (X, Y, Z): each builder
        term of rccsdt_q.py:394-399 as one batched launch over the slots."""
        o, vp, bs = self.o, self.vp, self.bs
        ext = {x: bs for x in 'abc'}
        ext.update({'e': vp, 'f': vp})
        ext.update({x: o for x in 'jkm'})
        C = Operand(Ws, 'Sabcejk', (nslot_max, bs, bs, bs, vp, o, o))
        launches = []
        for n, (sub, an, bn, alpha) in enumerate(WV_TERMS):
            sa, rest = sub.split(',')
            sb = rest.split('->')[0]
            batches = []
            for slot, (X, Y, Z) in enumerate(keys):
                blk = dict(a=X, b=Y, c=Z)
                batches.append(({x: blk[x] for x in sa if x in blk}, {x: blk[x] for x in sb if x in blk},
                                {'S': slot}))
            launches.append(plan_einsum('Q Wslot', sub, self._op(an, sa), self._op(bn, sb), C, alpha,
                                        0.0 if n == 0 else 1.0, ext, batches))
        return launches

    def _plan_tasks(self, tasks):
        """Launches of every chunk of ``tasks``; the chunks share one pair of z4/t4 buffers (and, per task,
        the W_vvvvoo slot buffer)."""
        o, vp, bs = self.o, self.vp, self.bs
        ext = {x: bs for x in 'abcd'}
        ext.update({'e': vp, 'f': vp})
        ext.update({x: o for x in 'ijklmn'})
        nmax = min(self.chunk, len(tasks))
        z4 = DeviceArray((nmax, bs, bs, bs, bs, o, o, o, o))
        t4 = DeviceArray((nmax, bs, bs, bs, bs, o, o, o, o))
        ws = DeviceArray((nmax, bs, bs, bs, bs, o, o, o, o)) if self.finalize_ws else None
        part = DeviceArray((nmax * bs ** 4, 2))
        nchunk = -(-len(tasks) // self.chunk)
        tot = DeviceArray((max(nchunk, 1), 2))
        # W_vvvvoo slots of each chunk: one per distinct ordered block triple its t4 terms read
        chunk_keys = []
        for c0 in range(0, len(tasks), self.chunk):
            keys = {}
            if self.w_mode == "task":
                for (a0, b0, cc, d0) in tasks[c0:c0 + self.chunk]:
                    blk = dict(a=a0, b=b0, c=cc, d=d0)
                    for sub, an, _, _ in T4_TERMS:
                        if an == 'Wv':
                            keys.setdefault(tuple(blk[x] for x in sub[:3]), len(keys))
            chunk_keys.append(keys)
        nslot_max = max((len(k) for k in chunk_keys), default=0)
        Ws = DeviceArray((nslot_max, bs, bs, bs, vp, o, o)) if nslot_max else None
        chunks = []
        for ci, c0 in enumerate(range(0, len(tasks), self.chunk)):
            chunk = tasks[c0:c0 + self.chunk]
            nt = len(chunk)
            ext['H'] = nt
            keys = chunk_keys[ci]
            blocks = [dict(a=a0, b=b0, c=cc, d=d0) for (a0, b0, cc, d0) in chunk]
            launches = self._slot_launches(Ws, nslot_max, list(keys)) if keys else []
            for out, terms in ((z4, Z4_TERMS), (t4, T4_TERMS)):
                C = Operand(out, 'Habcdijkl', (nmax, bs, bs, bs, bs, o, o, o, o))
                for n, (sub, an, bn, alpha) in enumerate(terms):
                    sa, rest = sub.split(',')
                    sb = rest.split('->')[0]
                    if an == 'Wv' and keys:
                        A = Operand(Ws, 'S' + sa, (nslot_max, bs, bs, bs, vp, o, o))
                        batches = [({'S': keys[tuple(blk[x] for x in sa[:3])]}, blk, {'H': h})
                                   for h, blk in enumerate(blocks)]
                    else:
                        A = self._op(an, sa)
                        batches = [(blk, blk, {'H': h}) for h, blk in enumerate(blocks)]
                    launches.append(plan_einsum('Q ' + sub, sub, A, self._op(bn, sb), C, alpha,
                                                0.0 if n == 0 else 1.0, ext, batches))
            tk = DeviceArray.from_host(np.array(chunk, dtype=np.int32).reshape(-1, 4))
            chunks.append((nt, launches, tk))
        return z4, t4, ws, part, tot, chunks

    def kernel(self, pppp, t2_oovv, t3_dev, mo_energy, job_idx=0, n_jobs=1, t3_layout="ijk"):
        """(e_[Q], e_(Q)) for the T3 already on the device (``t3_dev``): canonical i<=j<=k blocks
        (t3_layout "ijk") or the ABC array t3_abc[lin(a<=b<=c), i, j, k] ("abc")."""
        if t3_layout not in ("ijk", "abc"):
            raise ValueError("t3_layout must be 'ijk' or 'abc'")
        o, v, vp, bs = self.o, self.v, self.vp, self.bs
        if self.dev is None:
            self._setup()
        dev, st = self.dev, self.stream
        e = np.asarray(mo_energy, dtype=np.float64)
        pppp = np.asarray(pppp)
        t2q = np.asarray(t2_oovv).transpose(2, 3, 0, 1)
        dev['vvvo'].upload(_pad(pppp[o:, o:, o:, :o], (0, 1, 2), vp, o), st)
        dev['vooo'].upload(_pad(pppp[o:, :o, :o, :o], (0,), vp, o), st)
        dev['vvoo'].upload(_pad(pppp[o:, o:, :o, :o], (0, 1), vp, o), st)
        dev['ovvo'].upload(_pad(pppp[:o, o:, o:, :o], (1, 2), vp, o), st)
        dev['ovov'].upload(_pad(pppp[:o, o:, :o, o:], (1, 3), vp, o), st)
        dev['vvvv'].upload(_pad(pppp[o:, o:, o:, o:], (0, 1, 2, 3), vp, o), st)
        dev['oooo'].upload(np.ascontiguousarray(pppp[:o, :o, :o, :o], dtype=np.float64), st)
        dev['t2'].upload(_pad(t2q, (0, 1), vp, o), st)
        dev['e_occ'].upload(np.ascontiguousarray(e[:o]), st)
        dev['e_vir'].upload(np.concatenate([e[o:], np.zeros(vp - v)]), st)
        if t3_layout == "ijk":
            L.call("dcc_t3_to_abc_f64", t3_dev.vp, o, v, vp, dev['T'].vp, self._s())
        else:
            L.call("dcc_t3_abc_packed_to_dense_f64", t3_dev.vp, o, v, vp, dev['T'].vp, self._s())
        for g in self.wplans:
            g.run(st)
        key = (int(job_idx), int(n_jobs))
        if key not in self.task_plans:
            tasks = create_full_task_list(v, bs)
            per = (len(tasks) + n_jobs - 1) // n_jobs
            self.task_plans[key] = self._plan_tasks(tasks[job_idx * per: min(job_idx * per + per, len(tasks))])
        z4, t4, ws, part, tot, chunks = self.task_plans[key]
        if not chunks:
            return 0.0, 0.0
        for ci, (nt, launches, tk) in enumerate(chunks):
            for g in launches:
                g.run(st)
            nwg = nt * bs ** 4
            L.call("dcc_q_finalize_f64", t4.vp, z4.vp, tk.vp, nt, bs, o, v, vp, dev['e_occ'].vp, dev['e_vir'].vp,
                   dev['vvoo'].vp, dev['t2'].vp, ws.vp if ws is not None else None, part.vp, self._s())
            L.call("dcc_reduce_pairs_f64", part.vp, nwg, tot.offset_ptr(2 * ci), self._s())
        pairs = tot.download(stream=st).reshape(-1, 2)[:len(chunks)]
        paren = bracket = 0.0
        for p, b in pairs:      # chunk order, as the per-chunk host sums were
            paren += p
            bracket += b
        return bracket / 12.0, (paren + bracket) / 12.0
