"""Canonical i<=j<=k storage of T3 at one GPU (drop-in for ``distr_cc/distribute_t3.py``).

The reference spreads canonical blocks over MPI ranks (``DistributedT3IJK``,
``distr_cc/distribute_t3.py:86-1090``).  At one GPU every block is local and
the local order is the lexicographic i<=j<=k order, so ``t3_local[lin(i,j,k)]``
is block (i,j,k) (``distribute_t3.py:232-277`` at size 1).  The class keeps the
reference's public names and its save/load file format
(``distribute_t3.py:959-1090``) so checkpoints interoperate.
"""
import json

import numpy as np

# permutation table of the six orderings of (i,j,k) (reference src/rccsdt.c:32-39)
TP_T3 = ((0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0))


def ijk_to_linear(i, j, k, nocc):
    """Closed form of distribute_t3.py:51-60 (canonical i<=j<=k)."""
    n = nocc
    ni = nocc - i
    n_before_i = (n * (n + 1) * (n + 2) - ni * (ni + 1) * (ni + 2)) // 6
    n_before_j = (j - i) * (2 * nocc - j - i + 1) // 2
    return n_before_i + n_before_j + (k - j)


def linear_to_ijk(idx, nocc):
    """Inverse of ijk_to_linear by table search (distribute_t3.py:62-83 semantics)."""
    tbl = _triples_table(nocc)
    if not 0 <= idx < len(tbl):
        raise ValueError("linear index %d out of range for nocc=%d" % (idx, nocc))
    return tuple(int(x) for x in tbl[idx])


_TRIPLES = {}


def _triples_table(nocc):
    t = _TRIPLES.get(nocc)
    if t is None:
        t = np.array([(i, j, k) for i in range(nocc) for j in range(i, nocc) for k in range(j, nocc)],
                     dtype=np.int32).reshape(-1, 3)
        _TRIPLES[nocc] = t
    return t


def canonical_triples(nocc):
    return _triples_table(nocc)


def canonical_case(i, j, k):
    """Case number and canonical triple of an ordered (i,j,k) (reference src/rccsdt.c:41-73)."""
    s = sorted((i, j, k))
    # case c satisfies (i,j,k)[n] == s[TP_T3[c][n]] with ties resolved as the reference does
    if i <= j <= k:
        return 0, tuple(s)
    if i <= k < j:
        return 1, tuple(s)
    if j < i <= k:
        return 2, tuple(s)
    if k < i <= j:
        return 3, tuple(s)
    if j <= k < i:
        return 4, tuple(s)
    return 5, tuple(s)


def nocc3(nocc):
    return nocc * (nocc + 1) * (nocc + 2) // 6


class _SingleRankComm:
    """Stand-in communicator of size one (Get_rank 0, Get_size 1, identity bcast)."""

    def Get_rank(self):
        return 0

    def Get_size(self):
        return 1

    rank = property(Get_rank)
    size = property(Get_size)

    def bcast(self, obj, root=0):
        return obj

    def allreduce(self, obj, op=None):
        return obj

    def Allreduce(self, *args, **kwargs):
        return None

    def Barrier(self):
        return None


SINGLE_RANK = _SingleRankComm()


class DistributedT3IJK:
    """Canonical-block T3 layout at one rank; reference ``distribute_t3.py:86``."""

    def __init__(self, nocc, nvir, comm=None, distribution='balanced', batch_size=None, dtype=np.float64,
                 allow_python_fallback=None):
        self.nocc, self.nvir = int(nocc), int(nvir)
        self.dtype = np.dtype(dtype)
        self.comm = comm if comm is not None else SINGLE_RANK
        if self.comm.Get_size() != 1:
            raise ValueError("this build runs the triples path on one GPU; comm size must be 1")
        self.distribution = distribution
        self.batch_size = batch_size
        self.allow_python_fallback = allow_python_fallback
        self.nocc3 = nocc3(self.nocc)
        self.rank, self.size = 0, 1
        self.local_triples = [tuple(int(x) for x in t) for t in _triples_table(self.nocc)]
        self.local_nocc3 = len(self.local_triples)
        self.local_ijk_list = [(i, j, k, n) for n, (i, j, k) in enumerate(self.local_triples)]
        self.global_to_local_offset = np.arange(self.nocc3, dtype=np.int64)
        self.log = None

    def allocate_local(self):
        return np.zeros((self.local_nocc3, self.nvir, self.nvir, self.nvir), dtype=self.dtype)

    def iter_local_ijk(self):
        for i, j, k, idx in self.local_ijk_list:
            yield i, j, k, idx

    def get_local_offset(self, i, j, k):
        _, (ci, cj, ck) = canonical_case(i, j, k)
        return ijk_to_linear(ci, cj, ck, self.nocc)

    get_local_index = get_local_offset

    def get_owner(self, i, j, k):
        return 0

    def take_t3_single_local(self, t3_local, t3_blk, i0, j0, k0):
        """Host copy of block T[i0,j0,k0] in any order (distribute_t3.py:708-737)."""
        case, (ci, cj, ck) = canonical_case(i0, j0, k0)
        t3_blk[...] = t3_local[ijk_to_linear(ci, cj, ck, self.nocc)].transpose(TP_T3[case])
        return t3_blk

    def configure_communication_logging(self, enabled=None, log_dir=None):
        return None

    def print_distribution_info(self):
        return None

    def close_communication_log(self):
        return None

    def save_to_disk(self, t3_local, prefix):
        """Files of distribute_t3.py:959-996 (rank 0 of a size-1 run)."""
        np.save(f"{prefix}t3_local_0.npy", np.asarray(t3_local))
        np.save(f"{prefix}ijk_triples_0.npy", np.array(self.local_triples, dtype=np.int32))
        meta = {'nocc': self.nocc, 'nvir': self.nvir, 'distribution': str(self.distribution),
                'batch_size': None if self.batch_size is None else int(self.batch_size),
                'dtype': self.dtype.str, 'size': 1}
        with open(f"{prefix}metadata.json", 'w') as f:
            json.dump(meta, f)

    @classmethod
    def load_from_disk(cls, prefix, comm=None, batch_size=None, mmap_mode=None):
        """Reader of distribute_t3.py:998-1059; a size mismatch raises ValueError (:1017-1019)."""
        with open(f"{prefix}metadata.json") as f:
            meta = json.load(f)
        size = 1 if comm is None else comm.Get_size()
        if meta['size'] != size:
            raise ValueError(f"MPI size mismatch: saved with {meta['size']} ranks, "
                             f"but loading with {size} ranks. Redistribution not yet supported.")
        obj = cls(meta['nocc'], meta['nvir'], comm=comm, distribution=meta['distribution'],
                  batch_size=meta.get('batch_size', batch_size), dtype=np.dtype(meta['dtype']))
        saved = [tuple(int(x) for x in t) for t in np.load(f"{prefix}ijk_triples_0.npy")]
        t3 = np.load(f"{prefix}t3_local_0.npy", mmap_mode=mmap_mode)
        if saved != obj.local_triples:
            order = [ijk_to_linear(i, j, k, obj.nocc) for (i, j, k) in saved]
            full = obj.allocate_local()
            full[order] = t3
            t3 = full
        return obj, t3


class _AbcIndex:
    """Mapping (a, b, c) with a <= b <= c -> lexicographic canonical index over nvir (the closed form of
    ijk_to_linear), the abc_to_global_idx table of distribute_t3.py:1130-1133 without storing it."""

    def __init__(self, nvir):
        self.nvir = int(nvir)

    def __getitem__(self, abc):
        a, b, c = (int(x) for x in abc)
        if not 0 <= a <= b <= c < self.nvir:
            raise KeyError(abc)
        return ijk_to_linear(a, b, c, self.nvir)

    def get(self, abc, default=None):
        try:
            return self[abc]
        except KeyError:
            return default

    def __contains__(self, abc):
        return self.get(abc) is not None

    def __len__(self):
        return nocc3(self.nvir)


class DistributedT3ABC:
    """T3 stored as t3_local[lin(a<=b<=c), i, j, k] = T_{ijk}^{abc} for every ordered (i,j,k), at one rank
    (reference distribute_t3.py:1098-1221: with one rank every virtual triple is local, in lexicographic
    a<=b<=c order)."""

    def __init__(self, nocc, nvir, comm=None, blksize_abc=None, dtype=np.float64):
        self.nocc, self.nvir = int(nocc), int(nvir)
        self.dtype = np.dtype(dtype)
        self.comm = comm if comm is not None else SINGLE_RANK
        if self.comm.Get_size() != 1:
            raise ValueError("this build runs the triples path on one GPU; comm size must be 1")
        self.rank, self.size = 0, 1
        self.blksize_abc = 1 if blksize_abc is None else int(blksize_abc)
        self.n_abc_triplets = self.nvir3 = nocc3(self.nvir)
        self.local_nvir3 = self.n_abc_triplets
        self.abc_to_global_idx = _AbcIndex(self.nvir)

    @property
    def local_abc_triplets(self):
        return [tuple(int(x) for x in t) for t in _triples_table(self.nvir)]

    def get_owner_abc(self, a, b, c):
        return 0

    def abc_idx_to_abc(self, abc_idx):
        if 0 <= abc_idx < self.n_abc_triplets:
            return tuple(int(x) for x in _triples_table(self.nvir)[abc_idx])
        raise ValueError(f"Invalid abc_idx: {abc_idx}")

    def get_local_index(self, a, b, c, i, j, k):
        """Flat index of T_{ijk}^{abc} in t3_local (distribute_t3.py:1192-1199)."""
        off = self.abc_to_global_idx[tuple(sorted((a, b, c)))]
        return off * self.nocc ** 3 + (i * self.nocc + j) * self.nocc + k

    def allocate_local(self):
        return np.zeros((self.local_nvir3, self.nocc, self.nocc, self.nocc), dtype=self.dtype)

    def memory_usage_gb(self):
        return self.local_nvir3 * self.nocc ** 3 * self.dtype.itemsize / (1024 ** 3)


def redistribute_ijk_to_abc(dt3_ijk, t3_local_ijk, comm=None, blksize_abc=None, chunk_size=None,
                            log_redistribution=True, verbose=None, stdout=None):
    """IJK -> ABC layout (distribute_t3.py:1544-1642) as one device gather (dcc_t3_tri_to_abc_packed_f64);
    returns (DistributedT3ABC, t3_local_abc) on the host, like the reference."""
    from distr_cc import _lib as L
    from distr_cc.device import DeviceArray
    o, v = dt3_ijk.nocc, dt3_ijk.nvir
    dt3_abc = DistributedT3ABC(o, v, comm, blksize_abc=blksize_abc)
    src = DeviceArray.from_host(np.ascontiguousarray(t3_local_ijk, dtype=np.float64))
    dst = DeviceArray((dt3_abc.local_nvir3, o, o, o))
    try:
        L.call("dcc_t3_tri_to_abc_packed_f64", src.vp, o, v, dst.vp, None)
        out = dst.download()
    finally:
        src.free()
        dst.free()
    return dt3_abc, out
