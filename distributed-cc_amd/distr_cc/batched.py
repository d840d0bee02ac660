"""Many small RCCSDT triples updates at once (an ensemble of uneven nocc/nvir systems).

Every system keeps its own device buffers (a TriplesEngine each); the launch lists of all systems
have the same structure, so launch p of every system is one residual term of that system.  Those
contractions write disjoint outputs (each system's own R), so they run as ONE persistent grouped
GEMM launch (dcc_contract_grouped_f64) per term; the bandwidth passes (Y gather, W_oooo/F_oo
permuted adds, fused finalize) stay per system.  ``capture()`` records a whole step -- the residual
operands from the intermediates, R reset, Y gather, all grouped GEMMs, all finalizes -- into one HIP
graph, so a step is one graph launch.
``autotune()`` times the eight tile variants of the grouped kernel for every term and keeps the fastest.
"""
import ctypes

import numpy as np

from distr_cc import _lib as L
from distr_cc.device import DeviceArray, Event
from distr_cc.t3_engine import GemmLaunch, TriplesEngine

TILE_SHAPES = {0: "64x64", 1: "128x64", 2: "64x128", 3: "128x128",
               4: "64x64s1", 5: "128x64s1", 6: "64x128s1", 7: "128x128s1"}


class _GroupedPhase:
    def __init__(self, launches):
        self.launches = launches
        self.name = launches[0].name
        self.flops = sum(g.flops for g in launches)
        self.tile = 0
        self.variant = 0
        self.dev = None
        self.total = 0
        self.nworkgroups = 0   # persistent grid size; 0 = as many workgroups as are resident at once

    def problems(self):
        arr = (L.GemmProblem * len(self.launches))()
        for p, g in zip(arr, self.launches):
            M, N, K, alpha, beta, A, am, ak, B, bk, bn, C, cm, cn = g.args
            p.N, p.K, p.alpha, p.beta = N, K, alpha, beta
            p.A, p.B, p.C = A.value, B.value, C.value
            p.am, p.ak, p.bk, p.bn, p.cm, p.cn = am, ak, bk, bn, cm, cn
            p.batches, p.nbatch, p.max_rows = g.batches.ptr, g.nbatch, g.max_rows
        return arr

    def prepare(self, tile, stream):
        arr = self.problems()
        total, variant = ctypes.c_int64(), ctypes.c_int64()
        L.call("dcc_grouped_prepare", ctypes.cast(arr, ctypes.c_void_p), len(arr), int(tile), ctypes.byref(total),
               ctypes.byref(variant))
        self.tile, self.total, self.variant = int(tile), int(total.value), int(variant.value)
        host = np.frombuffer(bytes(arr), dtype=np.uint8).copy()
        if self.dev is not None:
            self.dev.free()
        self.dev = DeviceArray.from_host(host, stream)

    def run(self, stream):
        L.call("dcc_contract_grouped_f64", self.dev.vp, len(self.launches), self.total, self.variant,
               self.nworkgroups,
               stream.handle if stream is not None else None)


class BatchedTriples:
    """FP64 triples updates of an ensemble ``systems = [(nocc, nvir), ...]`` on one stream."""

    def __init__(self, systems, stream, device=0):
        if stream is None:
            raise ValueError("BatchedTriples needs a non-default stream (graph capture)")
        self.stream = stream
        self.systems = [(int(o), int(v)) for o, v in systems]
        self.engines = [TriplesEngine(o, v, stream=stream, device=device) for o, v in self.systems]
        nl = len(self.engines[0].launches)
        if any(len(e.launches) != nl for e in self.engines):
            raise ValueError("launch lists differ between systems")
        self.phases = []
        for p in range(nl):
            items = [e.launches[p] for e in self.engines]
            if all(isinstance(g, GemmLaunch) for g in items):
                ph = _GroupedPhase(items)
                ph.prepare(0, stream)
                self.phases.append(ph)
            else:
                self.phases.append(items)
        self.graph = None

    def set_system(self, s, t3, imds, t2, mo_energy, level_shift=0.0, nvir_diis=None):
        e = self.engines[s]
        e.set_t3(t3)
        e.set_intermediates(imds, t2)
        e.set_denominators(mo_energy, level_shift, nvir_diis)

    def issue_step(self):
        """Issue one triples update of every system on the stream (no host synchronisation)."""
        st = self.stream
        for e in self.engines:
            e.prepare_residual()
            e.R.fill(0.0, st)
            e.materialize_y()
        for ph in self.phases:
            if isinstance(ph, _GroupedPhase):
                ph.run(st)
            else:
                for g in ph:
                    g.run(st)
        for e in self.engines:
            e.finalize_launch()

    def capture(self):
        if self.graph is not None:
            L.call("dcc_graph_destroy", self.graph)
        L.call("dcc_graph_capture_begin", self.stream.handle)
        self.issue_step()
        g = ctypes.c_void_p()
        L.call("dcc_graph_capture_end", self.stream.handle, ctypes.byref(g))
        self.graph = g

    def step(self):
        if self.graph is not None:
            L.call("dcc_graph_launch", self.graph, self.stream.handle)
        else:
            self.issue_step()

    def norms(self):
        self.stream.synchronize()
        return [e.read_norms() for e in self.engines]

    def autotune(self, reps=2):
        """Pick the tile shape of every grouped phase by timing; returns {phase: (shape, ms)}.
        Run before a step: it accumulates into R and overwrites the W_oooo/F_oo scratch."""
        st = self.stream
        out = {}
        for n, ph in enumerate(self.phases):
            if not isinstance(ph, _GroupedPhase):
                continue
            best = None
            for tile in TILE_SHAPES:
                ph.prepare(tile, st)
                ph.run(st)
                a, b = Event(), Event()
                a.record(st)
                for _ in range(reps):
                    ph.run(st)
                b.record(st)
                ms = a.elapsed_ms(b) / reps
                if best is None or ms < best[1]:
                    best = (tile, ms)
            ph.prepare(best[0], st)
            out["%d %s" % (n, ph.name)] = (TILE_SHAPES[best[0]], best[1])
        if self.graph is not None:
            self.capture()
        return out

    def flops(self):
        return sum(ph.flops if isinstance(ph, _GroupedPhase) else 0.0 for ph in self.phases)

    def close(self):
        if self.graph is not None:
            L.call("dcc_graph_destroy", self.graph)
            self.graph = None
