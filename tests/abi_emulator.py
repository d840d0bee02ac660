"""NumPy emulator of the libdcc C ABI (include/dcc.h) -- CPU test infrastructure only.

It implements the documented semantics of each entry point on host memory so
that the host-side planning (index maps, batch tables, S3 jobs) can be checked
against the oracle without a GPU.  It says nothing about the HIP kernels
themselves; those are checked by the ``-m gpu`` tests.
"""
import ctypes
import itertools

import numpy as np

_PERMS = ((0, 1, 2), (0, 2, 1), (1, 0, 2), (1, 2, 0), (2, 0, 1), (2, 1, 0))


def bf16_rne(x):
    """float32 -> nearest-even bfloat16, returned as float64."""
    u = np.ascontiguousarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def bf16_split(x):
    """x = hi + lo with hi = bf16(x), lo = bf16(x - hi) (the split of dcc_contract_f32_bf16x3)."""
    x = np.asarray(x, dtype=np.float32)
    hi = bf16_rne(x)
    lo = bf16_rne((x.astype(np.float64) - hi).astype(np.float32))
    return hi, lo


class FakeDevice:
    def __init__(self):
        self.allocs = {}  # base -> bytearray-backed numpy uint8
        self.next = 1 << 40

    def malloc(self, nbytes):
        base = self.next
        self.next += ((nbytes + 4095) // 4096 + 1) * 4096
        self.allocs[base] = np.zeros(max(nbytes, 1), dtype=np.uint8)
        return base

    def resolve(self, ptr):
        for base, buf in self.allocs.items():
            if base <= ptr < base + buf.size:
                return buf, ptr - base
        raise KeyError("unknown device pointer %x" % ptr)

    def view(self, ptr, dtype, count=None):
        buf, off = self.resolve(int(ptr))
        arr = buf[off:].view(np.uint8)
        n = (arr.size // np.dtype(dtype).itemsize) if count is None else count
        return arr[: n * np.dtype(dtype).itemsize].view(dtype)


def _ptr(x):
    if x is None:
        return None
    if isinstance(x, ctypes.c_void_p):
        return x.value
    return int(x)


def _map(dev, m, idx, batch):
    idx = np.asarray(idx, dtype=np.int64)
    if m.table:
        tab = dev.view(m.table, np.int64)
        return tab[batch * m.table_bstride + idx]
    e12 = m.e1 * m.e2
    q0 = idx // e12
    r0 = idx - q0 * e12
    q1 = r0 // m.e2
    q2 = r0 - q1 * m.e2
    return q0 * m.s0 + q1 * m.s1 + q2 * m.s2


class Emulator:
    def __init__(self):
        self.dev = FakeDevice()
        self.error = b""
        self.calls = []

    # each handler returns a status
    def __call__(self, name, *args):
        self.calls.append(name)
        rec = getattr(self, "_recording", None)
        if rec is not None and name not in ("dcc_graph_capture_end", "dcc_graph_capture_begin"):
            rec.append((name, args))
            return 0
        return getattr(self, name)(*args)

    def dcc_graph_capture_begin(self, stream):
        self._recording = []
        return 0

    def dcc_graph_capture_end(self, stream, pgraph):
        self._graphs = getattr(self, "_graphs", {})
        gid = len(self._graphs) + 1
        self._graphs[gid] = self._recording
        self._recording = None
        pgraph._obj.value = gid
        return 0

    def dcc_graph_launch(self, graph, stream):
        for name, args in self._graphs[_ptr(graph)]:
            st = getattr(self, name)(*args)
            if st:
                return st
        return 0

    def dcc_graph_destroy(self, graph):
        return 0

    def dcc_grouped_prepare(self, probs, nprob, tile, ptotal, pvariant):
        from distr_cc import _lib as L
        if not 0 <= int(tile) <= 7:
            raise AssertionError("tile shape must be 0..7")
        bm, bn = {0: (64, 64), 1: (128, 64), 2: (64, 128), 3: (128, 128)}[int(tile) % 4]
        arr = (L.GemmProblem * int(nprob)).from_address(_ptr(probs))
        acc = 0
        for p in arr:
            p.tiles_m = (p.max_rows + bm - 1) // bm
            p.tiles_n = (p.N + bn - 1) // bn
            p.tile_begin = acc
            acc += p.nbatch * p.tiles_m * p.tiles_n
        ptotal._obj.value = acc
        pvariant._obj.value = int(tile)
        return 0

    def dcc_contract_grouped_f64(self, probs, nprob, total, variant, nwg, stream):
        if not 0 <= int(variant) < 64:
            raise AssertionError("variant must come from dcc_grouped_prepare")
        from distr_cc import _lib as L
        raw = self.dev.view(_ptr(probs), np.uint8, int(nprob) * ctypes.sizeof(L.GemmProblem))
        arr = (L.GemmProblem * int(nprob)).from_buffer_copy(raw.tobytes())
        for p in arr:
            st = self._contract(np.float64, 0, p.N, p.K, p.alpha, p.beta, p.A, p.am, p.ak, p.B, p.bk, p.bn,
                                p.C, p.cm, p.cn, p.batches, p.nbatch, p.max_rows, stream)
            if st:
                return st
        return 0

    def dcc_device_count(self, count):
        count._obj.value = 1
        return 0

    def dcc_set_device(self, device, buf, n):
        return 0

    def dcc_malloc(self, pptr, nbytes):
        pptr._obj.value = self.dev.malloc(int(nbytes))
        return 0

    def dcc_free(self, p):
        return 0

    def dcc_mem_get_info(self, pfree, ptotal):
        # an MI355X-sized device: 288 GB, all of it free
        pfree._obj.value = ptotal._obj.value = 288 << 30
        return 0

    def dcc_memcpy_htod(self, dst, src, nbytes, stream):
        d = self.dev.view(_ptr(dst), np.uint8, int(nbytes))
        d[:] = np.ctypeslib.as_array(ctypes.cast(src, ctypes.POINTER(ctypes.c_uint8)), (int(nbytes),))
        return 0

    def dcc_memcpy_dtoh(self, dst, src, nbytes, stream):
        s = self.dev.view(_ptr(src), np.uint8, int(nbytes))
        h = np.ctypeslib.as_array(ctypes.cast(dst, ctypes.POINTER(ctypes.c_uint8)), (int(nbytes),))
        h[:] = s
        return 0

    def dcc_memcpy_dtod(self, dst, src, nbytes, stream):
        self.dev.view(_ptr(dst), np.uint8, int(nbytes))[:] = self.dev.view(_ptr(src), np.uint8, int(nbytes))
        return 0

    def dcc_fill_f64(self, dst, n, value, stream):
        self.dev.view(_ptr(dst), np.float64, int(n))[:] = value
        return 0

    def dcc_fill_f32(self, dst, n, value, stream):
        self.dev.view(_ptr(dst), np.float32, int(n))[:] = value
        return 0

    def dcc_stream_synchronize(self, s):
        return 0

    def dcc_device_synchronize(self):
        return 0

    def dcc_contract_f64(self, *a):
        return self._contract(np.float64, *a)

    def dcc_contract_f32(self, *a):
        return self._contract(np.float32, *a)

    def dcc_contract_tiled_f64(self, *a):
        return self._contract(np.float64, *a[:-2], a[-1])

    def dcc_contract_tiled_f32(self, *a):
        return self._contract(np.float32, *a[:-2], a[-1])

    def dcc_contract_f32_bf16x3(self, *a):
        return self._contract("bf16x3", *a)

    def dcc_contract_tiled_f32_bf16x3(self, *a):
        return self._contract("bf16x3", *a[:-2], a[-1])

    def dcc_contract_packed_bf16x3(self, *a):
        return self._contract("packed", *a[:-2], a[-1])

    def dcc_pack_bf16x3(self, src, dst, n, stream):
        n = int(n)
        x = self.dev.view(_ptr(src), np.float32, n).copy()
        hi, lo = bf16_split(x)
        w = (hi.astype(np.float32).view(np.uint32) >> 16) | (lo.astype(np.float32).view(np.uint32) & 0xFFFF0000)
        self.dev.view(_ptr(dst), np.uint32, n)[:] = w
        return 0

    def _t2t3_operands(self, dt, packed, Ap, Bp, W_ovvv, W_oovo, t2, nocc, nvir, i0, n, stream):
        o, v, i0, n = int(nocc), int(nvir), int(i0), int(n)
        dev = self.dev
        Wv = dev.view(_ptr(W_ovvv), dt, o * v ** 3).reshape(o, v, v, v)          # [j,a,b,d]
        Wo = dev.view(_ptr(W_oovo), dt, o * o * v * o).reshape(o, o, v, o)      # [i,j,a,l]
        T2 = dev.view(_ptr(t2), dt, o * o * v * v).reshape(o, o, v, v)          # [i,k,d,c]
        A = np.empty((n, v, o, v, v + o), dt)
        A[..., :v] = Wv.transpose(2, 0, 1, 3)[None]                              # [b,j,a,d]
        A[..., v:] = -Wo[i0:i0 + n][:, None]                                     # [i,j,a,l] -> [i,b,j,a,l]
        B = np.empty((n, v, v + o, o, v), dt)
        B[:, :, :v] = T2[i0:i0 + n].transpose(0, 2, 1, 3)[:, None]               # [i,d,k,c]
        B[:, :, v:] = T2.transpose(2, 0, 1, 3)[None]                             # [b,l,k,c]
        for ptr, x in ((Ap, A), (Bp, B)):
            if packed:
                hi, lo = bf16_split(x.astype(np.float32).ravel())
                dev.view(_ptr(ptr), np.uint32, x.size)[:] = ((hi.astype(np.float32).view(np.uint32) >> 16) |
                                                             (lo.astype(np.float32).view(np.uint32) & 0xFFFF0000))
            else:
                dev.view(_ptr(ptr), dt, x.size)[:] = x.ravel()
        return 0

    def _store(self, ptr, x, dt, packed):
        """Write x at ptr as dt values, or as split-BF16x3 words when packed."""
        if packed:
            hi, lo = bf16_split(x.astype(np.float32).ravel())
            self.dev.view(_ptr(ptr), np.uint32, x.size)[:] = ((hi.astype(np.float32).view(np.uint32) >> 16) |
                                                              (lo.astype(np.float32).view(np.uint32) & 0xFFFF0000))
        else:
            self.dev.view(_ptr(ptr), dt, x.size)[:] = x.ravel()

    def _pair_weights(self, dt, packed, W1, nvir, kpad, Wp, stream):
        v, kp = int(nvir), int(kpad)
        if kp < v * (v + 1) // 2:
            return 1
        W = self.dev.view(_ptr(W1), dt, v ** 4).reshape(v, v, v, v)
        dS, eS = np.triu_indices(v)
        dA, eA = np.triu_indices(v, 1)
        out = np.zeros((len(dS) + len(dA), kp), dt)
        WS = W[dS, eS]                                   # (nS, v, v) rows x <= y
        out[:len(dS), :len(dS)] = WS[:, dS, eS] + WS[:, eS, dS] * (dS != eS)[None, :]
        WA = W[dA, eA]
        out[len(dS):, :len(dA)] = WA[:, dA, eA] - WA[:, eA, dA]
        self._store(Wp, out, dt, packed)
        return 0

    @staticmethod
    def _pair_plane(axes):
        """(nb, v, v, v) block view -> [b, z, d, e] with the pair on the last two axes."""
        return {0: (0, 3, 1, 2), 1: (0, 2, 1, 3), 2: (0, 1, 2, 3)}[axes]

    def _pair_split(self, dt, packed, t3, nblk, nvir, axes, kpad, Tp, stream):
        nb, v, kp, axes = int(nblk), int(nvir), int(kpad), int(axes)
        if kp < v * (v + 1) // 2 or axes not in (0, 1, 2):
            return 1
        P = self.dev.view(_ptr(t3), dt, nb * v ** 3).reshape(nb, v, v, v).transpose(self._pair_plane(axes))
        dS, eS = np.triu_indices(v)
        dA, eA = np.triu_indices(v, 1)
        out = np.zeros((2, nb, v, kp), dt)
        out[0, :, :, :len(dS)] = P[:, :, dS, eS] + P[:, :, eS, dS]
        out[1, :, :, :len(dA)] = P[:, :, dA, eA] - P[:, :, eA, dA]
        if axes == 0:
            out = out.transpose(0, 1, 3, 2)               # [2][nb][kp][v]
        self._store(Tp, np.ascontiguousarray(out), dt, packed)
        return 0

    def _pair_unfold(self, dt, Up, nblk, nvir, axes, R, stream):
        nb, v, axes = int(nblk), int(nvir), int(axes)
        if axes not in (0, 1, 2):
            return 1
        dS, eS = np.triu_indices(v)
        dA, eA = np.triu_indices(v, 1)
        mp = len(dS)
        U = self.dev.view(_ptr(Up), dt, 2 * nb * mp * v)
        U = U.reshape(2, nb, mp, v).transpose(0, 1, 3, 2) if axes == 0 else U.reshape(2, nb, v, mp)
        US, UA = U[0], U[1][:, :, :len(dA)]
        qS = dA * v - dA * (dA - 1) // 2 + (eA - dA)          # symmetric index of each strict pair
        full = np.zeros((nb, v, v, v), dt)                     # [b, z, x, y]
        full[:, :, dS, eS] = 0.5 * US
        full[:, :, dA, eA] = 0.5 * (US[:, :, qS] + UA)
        full[:, :, eA, dA] = 0.5 * (US[:, :, qS] - UA)
        Rv = self.dev.view(_ptr(R), dt, nb * v ** 3).reshape(nb, v, v, v).transpose(self._pair_plane(axes))
        Rv += full
        return 0

    def dcc_vvvv_pair_weights_f64(self, *a):
        return self._pair_weights(np.float64, False, *a)

    def dcc_vvvv_pair_weights_f32(self, *a):
        return self._pair_weights(np.float32, False, *a)

    def dcc_vvvv_pair_weights_packed_f32(self, *a):
        return self._pair_weights(np.float32, True, *a)

    def dcc_vvvv_pair_split_f64(self, *a):
        return self._pair_split(np.float64, False, *a)

    def dcc_vvvv_pair_split_f32(self, *a):
        return self._pair_split(np.float32, False, *a)

    def dcc_vvvv_pair_split_packed_f32(self, *a):
        return self._pair_split(np.float32, True, *a)

    def dcc_vvvv_pair_unfold_f64(self, *a):
        return self._pair_unfold(np.float64, *a)

    def dcc_vvvv_pair_unfold_f32(self, *a):
        return self._pair_unfold(np.float32, *a)

    def dcc_t2t3_operands_f64(self, *a):
        return self._t2t3_operands(np.float64, False, *a)

    def dcc_t2t3_operands_f32(self, *a):
        return self._t2t3_operands(np.float32, False, *a)

    def dcc_t2t3_operands_packed_f32(self, *a):
        return self._t2t3_operands(np.float32, True, *a)

    def _contract(self, dt, M, N, K, alpha, beta, A, am, ak, B, bk, bn, C, cm, cn, batches, nbatch, max_rows,
                  stream):
        split, packed = dt in ("bf16x3", "packed"), dt == "packed"
        if split:
            dt = np.float32
        dev = self.dev
        Av, Bv, Cv = dev.view(_ptr(A), dt), dev.view(_ptr(B), dt), dev.view(_ptr(C), dt)
        if packed:
            Av, Bv = dev.view(_ptr(A), np.uint32), dev.view(_ptr(B), np.uint32)

        def parts(w):   # hi, lo of packed words
            return ((w << 16).astype(np.uint32).view(np.float32).astype(np.float64),
                    (w & 0xFFFF0000).astype(np.uint32).view(np.float32).astype(np.float64))
        bt = dev.view(_ptr(batches), np.int64, 5 * int(nbatch)).reshape(-1, 5)
        ks = np.arange(K)
        ns = np.arange(N)
        for b in range(int(nbatch)):
            a_off, b_off, c_off, mb, me = (int(x) for x in bt[b])
            ms = np.arange(mb, me)
            if ms.size == 0:
                continue
            ai = a_off + _map(dev, am, ms, b)[:, None] + _map(dev, ak, ks, b)[None, :]
            bi = b_off + _map(dev, bk, ks, b)[:, None] + _map(dev, bn, ns, b)[None, :]
            ci = c_off + _map(dev, cm, ms, b)[:, None] + _map(dev, cn, ns, b)[None, :]
            if split and K > 0:
                ah, al = parts(Av[ai]) if packed else bf16_split(Av[ai])
                bh, bl = parts(Bv[bi]) if packed else bf16_split(Bv[bi])
                prod = ah @ bh + ah @ bl + al @ bh
            else:
                prod = (Av[ai].astype(np.float64) @ Bv[bi].astype(np.float64)) if K > 0 else np.zeros(ci.shape)
            old = Cv[ci].astype(np.float64) if beta != 0.0 else 0.0
            Cv[ci] = (alpha * prod + beta * old).astype(dt)
        return 0

    def dcc_tensor_axpby_f64(self, *a):
        return self._axpby(np.float64, *a)

    def dcc_tensor_axpby_f32(self, *a):
        return self._axpby(np.float32, *a)

    def dcc_tensor_axpby_f32_f64(self, *a):
        return self._axpby((np.float32, np.float64), *a)

    def dcc_tensor_axpby_f64_f32(self, *a):
        return self._axpby((np.float64, np.float32), *a)

    def _axpby(self, dt, dst, src, layout, alpha, beta, stream):
        dtd, dts = dt if isinstance(dt, tuple) else (dt, dt)
        lay = layout._obj if hasattr(layout, "_obj") else layout.contents
        nd = int(lay.ndim)
        shape = [int(lay.shape[d]) for d in range(nd)]
        if int(np.prod(shape)) == 0:
            return 0
        idx = np.indices(shape).reshape(nd, -1)
        od = sum(idx[d] * int(lay.dst_stride[d]) for d in range(nd))
        os_ = sum(idx[d] * int(lay.src_stride[d]) for d in range(nd))
        D, S = self.dev.view(_ptr(dst), dtd), self.dev.view(_ptr(src), dts)
        val = alpha * S[os_].astype(np.float64)
        if beta != 0.0:
            val = val + beta * D[od].astype(np.float64)
        D[od] = val.astype(dtd)
        return 0

    def dcc_sum_slices_f64(self, *a):
        return self._sum_slices(np.float64, *a)

    def dcc_sum_slices_f32(self, *a):
        return self._sum_slices(np.float32, *a)

    def _sum_slices(self, dt, dst, src, nslices, n, alpha, beta, stream):
        n, ns = int(n), int(nslices)
        D = self.dev.view(_ptr(dst), dt, n)
        acc = self.dev.view(_ptr(src), dt, ns * n).reshape(ns, n).astype(np.float64).sum(axis=0) if ns else 0.0
        D[:] = ((beta * D.astype(np.float64) if beta != 0.0 else 0.0) + alpha * acc).astype(dt)
        return 0

    def dcc_dot_partials(self):
        return 1024

    def dcc_dot_f64(self, *a):
        return self._dot(np.float64, *a)

    def dcc_dot_f32(self, *a):
        return self._dot(np.float32, *a)

    def _dot(self, dt, x, y, n, partials, out, stream):
        n = int(n)
        o = self.dev.view(_ptr(out), np.float64, 1)
        if n == 0:
            o[0] = 0.0
            return 0
        X = self.dev.view(_ptr(x), dt, n).astype(np.float64)
        Y = self.dev.view(_ptr(y), dt, n).astype(np.float64)
        o[0] = float(X @ Y)
        return 0

    def dcc_denom_divide_f64(self, x, nocc, nvir, rank, eia, stream):
        o, v = int(nocc), int(nvir)
        E = self.dev.view(_ptr(eia), np.float64, o * v).reshape(o, v)
        if rank == 1:
            X = self.dev.view(_ptr(x), np.float64, o * v).reshape(o, v)
            X /= E
        else:
            X = self.dev.view(_ptr(x), np.float64, o * o * v * v).reshape(o, o, v, v)
            X /= E[:, None, :, None] + E[None, :, None, :]
        return 0

    def dcc_s3_transform_f64(self, *a):
        return self._s3(np.float64, *a)

    def dcc_s3_transform_f32(self, *a):
        return self._s3(np.float32, *a)

    def dcc_s3_transform_pack_f32(self, out, src, nvir, jobs, njobs, stream):
        # float results in place, then packed in place (the words the fused kernel writes)
        self._s3(np.float32, out, src, nvir, 0.0, jobs, njobs, stream)
        v3 = int(nvir) ** 3
        J = self.dev.view(_ptr(jobs), np.int64, 8 * int(njobs)).reshape(-1, 8)
        for dst in J[:, 0]:
            self.dcc_pack_bf16x3(_ptr(out) + 4 * int(dst), _ptr(out) + 4 * int(dst), v3, stream)
        return 0

    def _s3(self, dt, out, src, nvir, beta, jobs, njobs, stream):
        v = int(nvir)
        O = self.dev.view(_ptr(out), dt)
        S = self.dev.view(_ptr(src), dt)
        raw = self.dev.view(_ptr(jobs), np.uint8, int(njobs) * 64)
        jb = raw.view(np.dtype([("dst", np.int64), ("src", np.int64), ("coef", np.float64, (6,))]))
        # the device runs the jobs of one call concurrently: destinations must be distinct
        assert len(set(int(d) for d in jb["dst"])) == len(jb), "S3 jobs of one call share a destination"
        for j in jb:
            x = S[int(j["src"]): int(j["src"]) + v ** 3].reshape(v, v, v).astype(np.float64)
            val = sum(c * x.transpose(p) for c, p in zip(j["coef"], _PERMS) if c != 0.0)
            o = O[int(j["dst"]): int(j["dst"]) + v ** 3].reshape(v, v, v)
            o[...] = ((beta * o.astype(np.float64) if beta != 0.0 else 0.0) + val).astype(dt)
        return 0

    def dcc_s3_gather_f64(self, *a):
        return self._s3_gather(np.float64, *a)

    def dcc_s3_gather_f32(self, *a):
        return self._s3_gather(np.float32, *a)

    def _s3_gather(self, dt, out, src, nvir, beta, jobs, njobs, stream):
        v = int(nvir)
        O = self.dev.view(_ptr(out), dt)
        S = self.dev.view(_ptr(src), dt)
        gdt = np.dtype([("dst", np.int64), ("src", np.int64, (6,)), ("coef", np.float64, (6,))])
        jb = self.dev.view(_ptr(jobs), np.uint8, int(njobs) * gdt.itemsize).view(gdt)
        assert len(set(int(d) for d in jb["dst"])) == len(jb), "S3 jobs of one call share a destination"
        for j in jb:
            val = 0.0
            for c, p, so in zip(j["coef"], _PERMS, j["src"]):
                if c != 0.0:
                    val = val + c * S[int(so): int(so) + v ** 3].reshape(v, v, v).astype(np.float64).transpose(p)
            o = O[int(j["dst"]): int(j["dst"]) + v ** 3].reshape(v, v, v)
            o[...] = ((beta * o.astype(np.float64) if beta != 0.0 else 0.0) + val).astype(dt)
        return 0

    def dcc_r3_finalize_f64(self, *a):
        return self._finalize(np.float64, *a)

    def dcc_r3_finalize_f32(self, *a):
        return self._finalize(np.float32, *a)

    def _finalize(self, dt, R, t3, r3_out, nblk, nvir, blk_ijk, eia, nocc, nvir_diis, partials, stream):
        v, n3, o = int(nvir), int(nblk), int(nocc)
        Rv = self.dev.view(_ptr(R), dt, n3 * v ** 3).reshape(n3, v, v, v).astype(np.float64)
        Tv = self.dev.view(_ptr(t3), dt, n3 * v ** 3).reshape(n3, v, v, v)
        ijk = self.dev.view(_ptr(blk_ijk), np.int32, 3 * n3).reshape(n3, 3)
        e = self.dev.view(_ptr(eia), np.float64, o * v).reshape(o, v)
        sym = sum(Rv.transpose(0, *(1 + np.array(p))) for p in _PERMS)
        r = Rv - sym / 6.0
        d = e[ijk[:, 0], :, None, None] + e[ijk[:, 1], None, :, None] + e[ijk[:, 2], None, None, :]
        r = r / d
        Tv[...] = (Tv.astype(np.float64) + r).astype(dt)
        if _ptr(r3_out):
            self.dev.view(_ptr(r3_out), dt, n3 * v ** 3)[:] = r.ravel().astype(dt)
        nd = int(nvir_diis)
        self._norms = (float((r * r).sum()), float((r[:, :nd, :nd, :nd] ** 2).sum()))
        P = self.dev.view(_ptr(partials), np.float64, 2)
        P[0], P[1] = self._norms
        self._npart = 1
        return 0

    def dcc_reduce_pairs_f64(self, x, n, out, stream):
        X = self.dev.view(_ptr(x), np.float64, 2 * int(n)).reshape(-1, 2)
        # finalize wrote only the first pair in the emulator
        o = self.dev.view(_ptr(out), np.float64, 2)
        o[:] = X[0] if getattr(self, "_npart", 0) == 1 else X.sum(0)
        return 0

    def dcc_r3_finalize_tiles(self, nvir):
        nt = (int(nvir) + 7) // 8
        return nt * (nt + 1) * (nt + 2) // 6

    def dcc_oooo_prefactors_f64(self, *a):
        return self._pfs(np.float64, *a)

    def dcc_oooo_prefactors_f32(self, *a):
        return self._pfs(np.float32, *a)

    def _pfs(self, dt, W, F, nocc, blk_ijk, nblk, pfs, stream):
        o, n3 = int(nocc), int(nblk)
        Wv = self.dev.view(_ptr(W), dt, o ** 4).reshape(o, o, o, o).astype(np.float64)
        Fv = self.dev.view(_ptr(F), dt, o * o).reshape(o, o).astype(np.float64)
        ijk = self.dev.view(_ptr(blk_ijk), np.int32, 3 * n3).reshape(n3, 3)
        Pd = self.dev.view(_ptr(pfs), dt, 6 * n3 * n3).reshape(6, n3, n3)
        P = np.zeros((6, n3, n3))
        # dense definition: coefficient of T[p]^s in R[q] collected from the
        # full-tensor W_oooo / F_oo terms by enumerating ordered (l,m) pairs
        lin = {tuple(int(x) for x in t): n for n, t in enumerate(ijk)}

        def add(q_abc_perm, srcs, coef, q):
            # srcs: ordered occupied triple of the source, q_abc_perm: how target abc reads source
            s_sorted = tuple(sorted(srcs))
            p = lin[s_sorted]
            # T[srcs][x] = Tc[p].transpose(case)[x]; combined with the target read permutation
            case = _case(*srcs)
            P[_compose(case, q_abc_perm), p, q] += coef

        for q, (i, j, k) in enumerate(ijk):
            for l, m in itertools.product(range(o), range(o)):
                add((0, 1, 2), (l, m, k), Wv[l, m, i, j], q)
                add((0, 2, 1), (l, m, j), Wv[l, m, i, k], q)
                add((2, 0, 1), (l, m, i), Wv[l, m, j, k], q)
            for l in range(o):
                add((0, 1, 2), (l, j, k), -Fv[l, i], q)
                add((0, 1, 2), (i, l, k), -Fv[l, j], q)
                add((0, 1, 2), (i, j, l), -Fv[l, k], q)
        Pd[...] = P.astype(dt)
        return 0

    def dcc_t3_tri_to_abc_packed_f64(self, t3, nocc, nvir, out, stream):
        o, v = int(nocc), int(nvir)
        n3 = o * (o + 1) * (o + 2) // 6
        T = self.dev.view(_ptr(t3), np.float64, n3 * v ** 3).reshape(n3, v, v, v)
        lin = {t: n for n, t in enumerate((i, j, k) for i in range(o) for j in range(i, o) for k in range(j, o))}
        full = np.zeros((v, v, v, o, o, o))
        for i, j, k in itertools.product(range(o), repeat=3):
            full[:, :, :, i, j, k] = T[lin[tuple(sorted((i, j, k)))]].transpose(_PERMS[_case(i, j, k)])
        abc = [(a, b, c) for a in range(v) for b in range(a, v) for c in range(b, v)]
        packed = np.array([full[a, b, c] for (a, b, c) in abc]).reshape(len(abc), o, o, o)
        self.dev.view(_ptr(out), np.float64, packed.size)[:] = packed.ravel()
        return 0

    def dcc_t3_abc_packed_to_dense_f64(self, t3_abc, nocc, nvir, vpad, out, stream):
        o, v, vp = int(nocc), int(nvir), int(vpad)
        abc = [(a, b, c) for a in range(v) for b in range(a, v) for c in range(b, v)]
        P = self.dev.view(_ptr(t3_abc), np.float64, len(abc) * o ** 3).reshape(len(abc), o, o, o)
        full = np.zeros((vp, vp, vp, o, o, o))
        for n, (a, b, c) in enumerate(abc):
            for perm in itertools.permutations(range(3)):
                x = (a, b, c)
                # T_{y[perm]}^{x[perm]} = T_y^x: place the block at the permuted virtual indices
                full[x[perm[0]], x[perm[1]], x[perm[2]]] = P[n].transpose(perm)
        self.dev.view(_ptr(out), np.float64, full.size)[:] = full.ravel()
        return 0

    def dcc_t3_to_abc_f64(self, t3, nocc, nvir, vpad, out, stream):
        o, v, vp = int(nocc), int(nvir), int(vpad)
        n3 = o * (o + 1) * (o + 2) // 6
        T = self.dev.view(_ptr(t3), np.float64, n3 * v ** 3).reshape(n3, v, v, v)
        lin = {}
        n = 0
        for i in range(o):
            for j in range(i, o):
                for k in range(j, o):
                    lin[(i, j, k)] = n
                    n += 1
        full = np.zeros((vp, vp, vp, o, o, o))
        for i, j, k in itertools.product(range(o), repeat=3):
            c = _case(i, j, k)
            blk = T[lin[tuple(sorted((i, j, k)))]].transpose(_PERMS[c])
            full[:v, :v, :v, i, j, k] = blk
        self.dev.view(_ptr(out), np.float64, full.size)[:] = full.ravel()
        return 0

    def dcc_q_finalize_f64(self, t4, z4, tasks, ntask, bs, nocc, nvir, vpad, e_occ, e_vir, vvoo, t2, ws, partials,
                           stream):
        nt, bs, o, v, vp = int(ntask), int(bs), int(nocc), int(nvir), int(vpad)
        n = nt * bs ** 4 * o ** 4
        T4 = self.dev.view(_ptr(t4), np.float64, n).reshape(nt, bs, bs, bs, bs, o, o, o, o)
        Z4 = self.dev.view(_ptr(z4), np.float64, n).reshape(nt, bs, bs, bs, bs, o, o, o, o)
        tk = self.dev.view(_ptr(tasks), np.int32, 4 * nt).reshape(nt, 4)
        eo = self.dev.view(_ptr(e_occ), np.float64, o)
        ev = self.dev.view(_ptr(e_vir), np.float64, vp)
        VV = self.dev.view(_ptr(vvoo), np.float64, vp * vp * o * o).reshape(vp, vp, o, o)
        T2 = self.dev.view(_ptr(t2), np.float64, vp * vp * o * o).reshape(vp, vp, o, o)
        P = self.dev.view(_ptr(partials), np.float64, 2 * nt * bs ** 4).reshape(nt, bs, bs, bs, bs, 2)
        eocc4 = eo[:, None, None, None] + eo[None, :, None, None] + eo[None, None, :, None] + eo[None, None, None, :]

        def fac(a, b, c, d):
            if a < b < c < d:
                return 24.0
            if (a == b < c < d) or (a < b == c < d) or (a < b < c == d):
                return 12.0
            if a == b < c == d:
                return 6.0
            return 0.0

        def sw(X, x, y):
            ax = [0, 1, 2, 3]
            ax[x], ax[y] = ax[y], ax[x]
            return X.transpose(ax)

        for t in range(nt):
            for ia, ib, ic, id_ in itertools.product(range(bs), repeat=4):
                a, b, c, d = (int(tk[t, 0]) + ia, int(tk[t, 1]) + ib, int(tk[t, 2]) + ic, int(tk[t, 3]) + id_)
                f = fac(a, b, c, d) if max(a, b, c, d) < v else 0.0
                if f == 0.0:
                    P[t, ia, ib, ic, id_] = 0.0
                    continue
                den = eocc4 - (ev[a] + ev[b] + ev[c] + ev[d])
                num = T4[t, ia, ib, ic, id_] + Z4[t, ia, ib, ic, id_]
                x = np.where(np.abs(den) > 1e-15, num / np.where(den == 0, 1, den), 0.0)
                x1 = 2 * x - sw(x, 0, 1) - sw(x, 0, 2) - sw(x, 0, 3)
                x2 = 2 * x1 - sw(x1, 1, 2) - sw(x1, 1, 3)
                x3 = f * (2 * x2 - sw(x2, 2, 3))
                zsq = (np.einsum('ij,kl->ijkl', VV[a, b], T2[c, d]) + np.einsum('ik,jl->ijkl', VV[a, c], T2[b, d])
                       + np.einsum('il,jk->ijkl', VV[a, d], T2[b, c]) + np.einsum('jk,il->ijkl', VV[b, c], T2[a, d])
                       + np.einsum('jl,ik->ijkl', VV[b, d], T2[a, c]) + np.einsum('kl,ij->ijkl', VV[c, d], T2[a, b]))
                P[t, ia, ib, ic, id_] = ((x3 * Z4[t, ia, ib, ic, id_]).sum(), (x3 * zsq).sum())
        self._npart = 0
        return 0

    def dcc_ring_weights_f64(self, *a):
        return self._ring(np.float64, *a)

    def dcc_ring_weights_f32(self, *a):
        return self._ring(np.float32, *a)

    def _ring(self, dt, Wovov, Wovvo, nocc, nvir, WR1, WR3a, WR3b, stream):
        o, v = int(nocc), int(nvir)
        A = self.dev.view(_ptr(Wovov), dt, (o * v) ** 2).reshape(o, v, o, v).astype(np.float64)
        B = self.dev.view(_ptr(Wovvo), dt, (o * v) ** 2).reshape(o, v, v, o).astype(np.float64)
        ovov = A.transpose(2, 1, 0, 3)     # [r,x,l,d] = W_ovov[l,x,r,d]
        ovvo = B.transpose(3, 1, 0, 2)     # [r,x,l,d] = W_ovvo[l,x,d,r]
        self.dev.view(_ptr(WR1), dt, (o * v) ** 2)[:] = (-ovov).ravel().astype(dt)
        self.dev.view(_ptr(WR3a), dt, (o * v) ** 2)[:] = ovvo.ravel().astype(dt)
        self.dev.view(_ptr(WR3b), dt, (o * v) ** 2)[:] = (-0.5 * (ovov + ovvo)).ravel().astype(dt)
        return 0


def _case(i, j, k):
    if i <= j <= k:
        return 0
    if i <= k < j:
        return 1
    if j < i <= k:
        return 2
    if k < i <= j:
        return 3
    if j <= k < i:
        return 4
    return 5


def _compose(case, read):
    """Index s with  Tc.transpose(P[case]).transpose(read) == Tc.transpose(P[s]) ... output side.

    R[q][a,b,c] += T_src.transpose(read)[a,b,c] where T_src = Tc.transpose(P[case]).
    The launches apply output permutations; the emulator maps onto the six
    (source-transpose) GEMM forms of rccsdt.py:726-731 by composing.
    """
    p = _PERMS[case]
    total = tuple(p[r] for r in read)
    # GEMM form s adds Tc.transpose(gemm_form[s]) to R[q]:
    # 'pq,pabc->qcab' puts Tc[a,b,c] at R[c,a,b], i.e. R[x,y,z] += Tc[y,z,x] = Tc.transpose(2,0,1)
    # 'pq,pabc->qbca' puts Tc[a,b,c] at R[b,c,a], i.e. R[x,y,z] += Tc[z,x,y] = Tc.transpose(1,2,0)
    gemm_form = {0: (0, 1, 2), 1: (0, 2, 1), 2: (1, 0, 2), 3: (2, 0, 1), 4: (1, 2, 0), 5: (2, 1, 0)}
    for s, f in gemm_form.items():
        if f == total:
            return s
    raise AssertionError(total)


def install(monkeypatch):
    """Route distr_cc._lib.call (and the raw finalize-tiles query) through a fresh emulator."""
    from distr_cc import _lib as L
    from distr_cc import device

    emu = Emulator()
    monkeypatch.setattr(L, "call", lambda name, *a: L.check(emu(name, *a), name))

    class _Raw:
        def __getattr__(self, name):
            return getattr(emu, name)

    monkeypatch.setattr(L, "_lib", _Raw())
    monkeypatch.setattr(device, "_DEVICE_READY", {0: "gfx950-emulated"})
    return emu
