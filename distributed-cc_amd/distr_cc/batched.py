"""Many small RCCSDT triples updates at once (an ensemble of uneven nocc/nvir systems).

Every system keeps its own device buffers (a TriplesEngine each); the launch lists of all systems
have the same structure, so launch p of every system is one residual term of that system.  Those
contractions write disjoint outputs (each system's own R), so they run as ONE persistent grouped
GEMM launch (dcc_contract_grouped_f64) per term; the bandwidth passes (Y gather, W_oooo/F_oo
permuted adds, fused finalize) stay per system.  ``capture()`` records a whole step -- the residual
operands from the intermediates, R reset, Y gather, all grouped GEMMs, all finalizes -- into one HIP
graph, so a step is one graph launch.  ``BatchedRCCSDT`` does the same for whole RCCSDT iterations
(lower order, T3 -> r1/r2/W, triples residual, update and energy of every system); there the
contractions of the lower-order part are grouped the same way, operation by operation.
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


class _Grouped:
    """Grouping of the triples residual launches of several TriplesEngines (launch p of every engine is the
    same term), the graph of one step, and the tile autotuner of the grouped launches."""

    def _group(self, tr_engines, stream):
        if stream is None:
            raise ValueError("a batched ensemble needs a non-default stream (graph capture)")
        self.stream = stream
        nl = len(tr_engines[0].launches)
        if any(len(e.launches) != nl for e in tr_engines):
            raise ValueError("launch lists differ between systems")
        self.phases = []
        for p in range(nl):
            items = [e.launches[p] for e in tr_engines]
            if all(isinstance(g, GemmLaunch) for g in items):
                ph = _GroupedPhase(items)
                ph.prepare(0, stream)
                self.phases.append(ph)
            else:
                self.phases.append(items)
        self.graph = None

    def _run_phases(self):
        for ph in self.phases:
            if isinstance(ph, _GroupedPhase):
                ph.run(self.stream)
            else:
                for g in ph:
                    g.run(self.stream)

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

    def _tunable_phases(self):
        return [ph for ph in self.phases if isinstance(ph, _GroupedPhase)]

    def autotune(self, reps=2):
        """Pick the tile shape of every grouped phase by timing; returns {phase: (shape, ms)}.
        Run before a step (and set the amplitudes afterwards): it accumulates into R, into the
        intermediates and overwrites the W_oooo/F_oo scratch."""
        st = self.stream
        out = {}
        for n, ph in enumerate(self._tunable_phases()):
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
        for ph in self._tunable_phases():
            if ph.dev is not None:
                ph.dev.free()
                ph.dev = None
        if self.graph is not None:
            L.call("dcc_graph_destroy", self.graph)
            self.graph = None


class BatchedTriples(_Grouped):
    """FP64 triples updates of an ensemble ``systems = [(nocc, nvir), ...]`` on one stream."""

    def __init__(self, systems, stream, device=0):
        self.systems = [(int(o), int(v)) for o, v in systems]
        self.engines = [TriplesEngine(o, v, stream=stream, device=device) for o, v in self.systems]
        self._group(self.engines, stream)

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
        self._run_phases()
        for e in self.engines:
            e.finalize_launch()

    def norms(self):
        self.stream.synchronize()
        return [e.read_norms() for e in self.engines]


class BatchedRCCSDT(_Grouped):
    """Full FP64 RCCSDT iterations of an ensemble of systems on one stream: every system has its own
    RCCSDTEngine (T1 dressing, lower order, T3 -> r1/r2/W, triples residual, update, energy), the
    triples residual contractions of all systems run as one grouped persistent launch per term, and a
    whole iteration of the ensemble -- followed by every system's energy -- is one HIP graph.

    ``systems`` is a list of (nocc, nvir, fock, pppp, mo_energy).  No DIIS (its B-matrix solve is a host
    step per iteration): the iteration is the Jacobi update of update_amps_rccsdt_tri_
    (distr_cc/rccsdt.py:890-977)."""

    def __init__(self, systems, stream, level_shift=0.0, device=0, group_lower=True):
        from distr_cc.rccsdt_engine import RCCSDTEngine
        self.systems = [(int(o), int(v)) for o, v, _, _, _ in systems]
        # plain W_vvvv GEMMs: the phases are grouped launch by launch across the systems
        self.engines = [RCCSDTEngine(o, v, F, g, e, level_shift=level_shift, stream=stream, device=device,
                                     vvvv_pairs=False) for o, v, F, g, e in systems]
        for e in self.engines:
            e.tr.set_denominators(e.mo_energy, e.level_shift, e.nvir_diis)
        self._group([e.tr for e in self.engines], stream)
        self._group_lower(group_lower)

    def _group_lower(self, group):
        """Operation q of every engine's issue_lower (RCCSDTEngine.lower_ops): when each engine's
        operation starts with an FP64 contraction, those contractions run as one grouped persistent
        launch, followed by the rest of every engine's operation (split-K sums, ...).  Operations
        without a contraction run engine by engine.  If the engines' operation lists differ in length
        (or ``group`` is False), every engine's lower-order work runs on its own."""
        per = [e.lower_ops() for e in self.engines]
        self.lower_phases = []
        if not group or any(len(p) != len(per[0]) for p in per):
            self.lower_plan = [(None, [[s for op in p for s in op] for p in per])]
            return
        self.lower_plan = []
        for q in range(len(per[0])):
            ops = [p[q] for p in per]
            heads = [op[0] for op in ops]
            if all(isinstance(h, GemmLaunch) and h.fn == "dcc_contract_f64" for h in heads):
                ph = _GroupedPhase(heads)
                ph.prepare(0, self.stream)
                self.lower_phases.append(ph)
                self.lower_plan.append((ph, [op[1:] for op in ops]))
            else:
                self.lower_plan.append((None, ops))

    def _tunable_phases(self):
        return super()._tunable_phases() + self.lower_phases

    def issue_lower(self):
        """issue_lower of every engine, with the contractions of each operation grouped."""
        st = self.stream
        for ph, lists in self.lower_plan:
            if ph is not None:
                ph.run(st)
            for steps in lists:
                for step in steps:
                    step(st)

    def set_amplitudes(self, s, t1, t2, t3):
        self.engines[s].set_amplitudes(t1, t2, t3)

    def issue_step(self, timer=None):
        """One iteration of every system and then its energy, with no host synchronisation.
        ``timer(phase, "start"|"end")``, if given, is called around the four phases (not under capture)."""
        st = self.stream
        tm = timer if timer is not None else (lambda name, what: None)
        tm("lower order, T3 -> r1/r2/W, residual operands", "start")
        self.issue_lower()
        tm("lower order, T3 -> r1/r2/W, residual operands", "end")
        tm("R reset and Y gather", "start")
        for e in self.engines:
            e.tr.R.fill(0.0, st)
            e.tr.materialize_y()
        tm("R reset and Y gather", "end")
        tm("grouped triples residual", "start")
        self._run_phases()
        tm("grouped triples residual", "end")
        tm("finalize and energy", "start")
        for e in self.engines:
            e.tr.finalize_launch()
            e.energy_launch()
        tm("finalize and energy", "end")

    def energies(self):
        self.stream.synchronize()
        return [e.read_energy() for e in self.engines]
