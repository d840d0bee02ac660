#!/usr/bin/env python
"""Benchmark of the FP64 RCCSDT triples update on one MI355X.

Workload (default): the synthetic plain-FP64 T3 residual at nocc=10, nvir=90
(220 canonical blocks of 90^3, naux=300 B-factor ERIs, seed 20261202).  One
step = one full triples update on the device: the T3 residual (T2->T3,
F_vv, W_vvvv, W_ovov/W_ovvo, W_oooo/F_oo contractions) followed by the fused
P3_full projection, denominator, norms and t3 += r3, with every input resident
in HBM before the timed region.  Prints one JSON line.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu-baseline]
"""
import argparse
import gc
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "distributed-cc_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

METRIC = "RCCSDT iteration time, mixed vs FP64 at |ΔE|≤1 µEh; % MFMA peak; (Q) time"
PEAK_FP64_MFMA_TFLOPS = 78.6
PEAK_HBM_GBS = 8000.0


def algorithmic_flops(o, v):
    """Flops of one T3 residual with canonical-block savings (reference einsum count)."""
    o3u = o * (o + 1) * (o + 2) // 6
    return o3u * (6 * v ** 5 + 24 * o * v ** 4 + 18 * v ** 4 + 12 * o * v ** 3 + 6 * o * o * v ** 3
                  + 6 * o * v ** 3) + o ** 3 * (4 * v ** 4 + (4 * o + 2) * v ** 3)


def cpu_baseline(o, v, sysm, im, t2, t3, budget_s=25.0):
    """Time the oracle (reference algorithm, NumPy/BLAS FP64) on a bounded sample of the
    same residual: n and 2n target blocks, extrapolated linearly to all blocks."""
    from oracle import rccsdt_tri as tri
    n3 = t3.shape[0]
    try:
        from threadpoolctl import threadpool_info
        threads = max([d.get("num_threads", 1) for d in threadpool_info()] or [1])
    except Exception:
        threads = len(os.sched_getaffinity(0))

    eia = sysm.mo_energy[:o, None] - sysm.mo_energy[None, o:]
    triples = tri.canonical_triples(o)

    def run(n):
        t0 = time.perf_counter()
        r3 = tri.compute_r3_tri(o, v, im, t2, t3, targets=range(n))
        proj = tri.project_p3_full(r3[:n])
        for idx in range(n):
            i, j, k = triples[idx]
            d = eia[i, :, None, None] + eia[j, None, :, None] + eia[k, None, None, :]
            _ = t3[idx] + proj[idx] / d
        return time.perf_counter() - t0

    n1 = 1
    t1 = run(n1)
    n2 = max(3, min(n3, int(n1 + max(1.0, (budget_s * 0.7 - t1) / max(t1, 1e-3)))))
    t2s = run(n2)
    slope = max((t2s - t1) / (n2 - n1), 1e-9)
    fixed = max(t1 - slope * n1, 0.0)
    est = fixed + slope * n3
    return {"value": est, "unit": "s", "cores": int(threads), "kind": "port",
            "sample": "oracle block-loop residual+update timed on %d and %d of %d target blocks "
                      "(%.1f s, %.1f s), extrapolated linearly to all blocks" % (n1, n2, n3, t1, t2s)}


def measured_traffic(o, v):
    """HBM bytes per W_vvvv launch from the committed PMC summary of this workload, if any."""
    path = os.path.join(ROOT, "profiles", "contract_f64_w_vvvv_o%d_v%d.md" % (o, v))
    try:
        for line in open(path):
            if line.startswith("traffic_bytes_per_launch"):
                return float(line.split("=")[1]), os.path.relpath(path, ROOT)
    except OSError:
        pass
    return None, None


def _tiles_from_small(dt, contract, seed, sigma_b, o=10, v=90):
    """Autotune the engine of one precision at nocc=10, nvir=90 and return its {launch name: tile}."""
    from distr_cc.device import Stream
    from distr_cc.rccsdt_engine import RCCSDTEngine
    from distr_cc.synthetic import SyntheticSystem
    sysm = SyntheticSystem(o, v, naux=300, seed=seed, sigma_b=sigma_b)
    st = Stream()
    eng = RCCSDTEngine(o, v, sysm.fock, sysm.pppp(), sysm.mo_energy, stream=st, t3_dtype=dt,
                       triples_contract=contract)
    eng.set_amplitudes(np.zeros((o, v)), sysm.t2(), sysm.t3_tri())
    eng.update_amps()
    eng.tr.autotune()
    st.synchronize()
    tiles = eng.tr.tiles_by_name()
    del eng
    gc.collect()
    return tiles


def mixed_vs_fp64(o, v, iters, seed=20261202, sigma_b=0.01, autotune="direct"):
    """Full device RCCSDT iterations (T1 dressing, lower order, T3 -> r1/r2/W, triples residual,
    update; no DIIS) from the synthetic amplitudes with FP64, FP32 and split-BF16x3 triples
    contractions (FP64 elsewhere).  Returns per-iteration |E_mixed - E_fp64| and the mean time of
    one iteration in each precision (HIP events on the engine stream).  ``autotune``: "direct"
    tunes each engine, "transfer" applies the tiles tuned at nocc=10, nvir=90 in the same
    precision (by launch name), "none" keeps the library defaults."""
    from distr_cc.device import Event, Stream
    from distr_cc.rccsdt_engine import RCCSDTEngine
    from distr_cc.synthetic import SyntheticSystem
    sysm = SyntheticSystem(o, v, naux=300, seed=seed, sigma_b=sigma_b)
    pppp = sysm.pppp()
    t1, t2, t3 = np.zeros((o, v)), sysm.t2(), sysm.t3_tri()
    print("bench: mixed inputs ready (T3 %.1f GB)" % (t3.nbytes / 1e9), file=sys.stderr, flush=True)
    out = {}
    for name, dt, contract in (("fp64", np.float64, "native"), ("fp32", np.float32, "native"),
                               ("bf16x3", np.float32, "bf16x3")):
        print("bench: mixed %s engine nocc=%d nvir=%d" % (name, o, v), file=sys.stderr, flush=True)
        # tiles tuned on the small system first, so that the two engines never share HBM
        tiles = _tiles_from_small(dt, contract, seed, sigma_b) if autotune == "transfer" else None
        st = Stream()
        eng = RCCSDTEngine(o, v, sysm.fock, pppp, sysm.mo_energy, stream=st, t3_dtype=dt,
                           triples_contract=contract)
        eng.set_amplitudes(t1, t2, t3)
        if autotune == "direct":
            eng.tr.autotune()
        elif tiles is not None:
            eng.tr.set_tiles_by_name(tiles)
        st.synchronize()
        eng.update_amps()       # warm-up iteration (first-launch costs), then restart
        print("bench: mixed %s warm-up done" % name, file=sys.stderr, flush=True)
        eng.set_amplitudes(t1, t2, t3)
        energies, times = [eng.energy()], []
        phase_ev = []

        def timer(ph, what):
            e = Event()
            e.record(st)
            phase_ev.append((ph, what, e))

        for it in range(iters):
            a, b = Event(), Event()
            a.record(st)
            eng.update_amps(timer=timer)
            b.record(st)
            times.append(a.elapsed_ms(b))
            energies.append(eng.energy())
            print("bench: mixed %s iteration %d: %.1f ms, E_corr %.12f" % (name, it, times[-1], energies[-1]),
                  file=sys.stderr, flush=True)
        phases = {}
        starts = {}
        for ph, what, e in phase_ev:
            if what == "start":
                starts[ph] = e
            else:
                phases.setdefault(ph, []).append(starts[ph].elapsed_ms(e))
        out[name] = (energies, times, {k: float(np.mean(x)) for k, x in phases.items()})
        del eng, timer, phase_ev
        gc.collect()   # the engine's programs hold reference cycles; free its device memory now
    e64 = out["fp64"][0]
    t64 = float(np.mean(out["fp64"][1]))
    res = {"workload": "synthetic RCCSDT iterations nocc=%d nvir=%d, B-factor sigma=%g (weakly correlated, "
                       "so the Jacobi iteration converges), no DIIS, %d iterations, tiles %s"
                       % (o, v, sigma_b, iters, {"direct": "autotuned", "none": "library defaults",
                                                 "transfer": "autotuned at nocc=10 nvir=90"}[autotune]),
           "fp64_ms_per_iteration": t64, "e_corr_fp64": e64, "fp64_phase_ms": out["fp64"][2]}
    for name in ("fp32", "bf16x3"):
        de = [abs(x - y) for x, y in zip(out[name][0], e64)]
        tm = float(np.mean(out[name][1]))
        res.update({name + "_ms_per_iteration": tm, "speedup_" + name: t64 / tm,
                    name + "_abs_dE_per_iteration": de, name + "_max_abs_dE": max(de),
                    name + "_phase_ms": out[name][2]})
    return res


def q_time(o, v, blksize=4, seed=20261202):
    """(Q)/[Q] of a synthetic canonical-block T3 already on the device (rccsdt_q.kernel's work:
    integral slices to HBM, IJK -> ABC gather, W_vvvvoo/W_vvoooo, all tasks, energies)."""
    from distr_cc.device import DeviceArray, Event, Stream
    from distr_cc.q_engine import QEngine
    from distr_cc.synthetic import SyntheticSystem
    sysm = SyntheticSystem(o, v, naux=300, seed=seed)
    pppp, t2, t3 = sysm.pppp(), sysm.t2(), sysm.t3_tri()
    st = Stream()
    t3d = DeviceArray.from_host(t3, st)
    q = QEngine(o, v, blksize=blksize, stream=st)
    a0, b0 = Event(), Event()
    a0.record(st)
    q.kernel(pppp, t2, t3d, sysm.mo_energy)     # first call: allocates the buffers and plans the launches
    b0.record(st)
    st.synchronize()
    a, b = Event(), Event()
    w0 = time.perf_counter()
    a.record(st)
    eb, ep = q.kernel(pppp, t2, t3d, sysm.mo_energy)
    b.record(st)
    st.synchronize()
    return {"workload": "synthetic (Q) nocc=%d nvir=%d blksize=%d" % (o, v, blksize),
            "q_time_s": a.elapsed_ms(b) / 1e3, "wall_s": time.perf_counter() - w0,
            "first_call_s": a0.elapsed_ms(b0) / 1e3, "e_bracket": eb, "e_paren": ep,
            "note": "second call of one QEngine: host->device copies of the integral slices (PCIe), IJK->ABC "
                    "gather, W builders, all task GEMMs, finalize and reductions; the first call also "
                    "allocates the buffers and plans every launch (first_call_s)"}


def adjoint_cost(o, v, reps=3, seed=20261202):
    """Forward T2->T3 residual vs its hand-written VJP (distr_cc.adjoint) on the synthetic system:
    mean device time of each, their ratio, and the device memory the adjoint engine holds."""
    from distr_cc.adjoint import T2T3Adjoint
    from distr_cc.device import Event, Stream, mem_get_info
    from distr_cc.synthetic import SyntheticSystem
    sysm = SyntheticSystem(o, v, naux=300, seed=seed)
    im = sysm.bare_intermediates()
    st = Stream()
    free0, _ = mem_get_info()
    adj = T2T3Adjoint(o, v, stream=st)
    adj.set_inputs(sysm.t2(), im["W_vvvo"], im["W_vooo"])
    adj.lam.upload(np.ascontiguousarray(sysm.t3_tri()), st)
    free1, _ = mem_get_info()

    def timed(fn):
        fn()
        st.synchronize()
        out = []
        for _ in range(reps):
            a, b = Event(), Event()
            a.record(st)
            fn()
            b.record(st)
            out.append(a.elapsed_ms(b))
        return float(np.mean(out))

    fwd, bwd = timed(adj.forward), timed(adj.backward)
    return {"workload": "T2->T3 residual and its VJP (dt2, dW_vvvo, dW_vooo), nocc=%d nvir=%d" % (o, v),
            "forward_ms": fwd, "backward_ms": bwd, "backward_over_forward": bwd / fwd,
            "adjoint_engine_bytes": int(free0 - free1),
            "dense_adjoint_bytes": int(o ** 3 * v ** 3 * 8)}


def batched_ensemble(nsys, steps=3, seed=20261202):
    """C4-style ensemble: nsys synthetic systems, nocc ~ U{3..8}, nvir ~ U{20..64} (system s drawn
    from seed + s), triples updates of all systems per step as grouped persistent launches captured in
    one HIP graph, tile shapes autotuned per term.  Reports aggregate system-updates per second."""
    from distr_cc.batched import BatchedTriples
    from distr_cc.device import Event, Stream
    from distr_cc.synthetic import SyntheticSystem
    systems = []
    for s in range(nsys):
        r = np.random.default_rng(seed + s)
        systems.append((int(r.integers(3, 9)), int(r.integers(20, 65))))
    st = Stream()
    bt = BatchedTriples(systems, st)
    inputs = []
    for s, (o, v) in enumerate(systems):
        sysm = SyntheticSystem(o, v, naux=300, seed=seed + s)
        inputs.append((sysm.t3_tri(), sysm.bare_intermediates(), sysm.t2(), sysm.mo_energy))
        bt.set_system(s, *inputs[-1])
    tuned = bt.autotune()
    for s, inp in enumerate(inputs):
        bt.engines[s].set_t3(inp[0])
    bt.capture()
    bt.step()                          # warm-up replay
    st.synchronize()
    for s, inp in enumerate(inputs):
        bt.engines[s].set_t3(inp[0])
    a, b = Event(), Event()
    a.record(st)
    for _ in range(steps):
        bt.step()
    b.record(st)
    ms = a.elapsed_ms(b) / steps
    norms = bt.norms()
    # parity at this scale: one replay from the inputs against every system's own launches (one GEMM
    # per term and system), compared through the update norm ||r3||^2 of each system
    for s, inp in enumerate(inputs):
        bt.engines[s].set_t3(inp[0])
    bt.step()
    graph_norms = bt.norms()
    rel = []
    for s, inp in enumerate(inputs):
        e = bt.engines[s]
        e.set_t3(inp[0])
        e.residual()
        own = e.finalize(inp[3])[0]
        rel.append(abs(graph_norms[s][0] - own) / abs(own))
    shapes = {}
    for k, (shape, _) in tuned.items():
        shapes[shape] = shapes.get(shape, 0) + 1
    out = {"workload": "%d synthetic systems, nocc~U{3..8}, nvir~U{20..64}, FP64 triples update per system "
                       "per step, grouped persistent GEMMs in one HIP graph" % nsys,
           "ms_per_step": ms, "system_updates_per_s": nsys / (ms * 1e-3),
           "contraction_tflops": bt.flops() / (ms * 1e-3) / 1e12,
           "autotuned_tile_shapes": shapes, "finite_norms": bool(np.all(np.isfinite(norms))),
           "max_rel_norm_diff_vs_per_system_launches": float(max(rel))}
    bt.close()
    return out


def probe_residual(o, v, precision, tiles, seed=20261202):
    """One triples engine of the given precision on the synthetic system (nocc=o, nvir=v): tiles loaded
    from ``tiles`` if it exists (a tile list of this size or a {launch name: tile} map), else autotuned
    (and saved there by launch name when a path is given); the residual is
    run twice and the second one is timed, per launch.  Used to profile one precision's kernels."""
    from distr_cc.device import Event, Stream
    from distr_cc.synthetic import SyntheticSystem
    from distr_cc.t3_engine import TriplesEngine
    dt, contract = {"fp64": (np.float64, "native"), "fp32": (np.float32, "native"),
                    "bf16x3": (np.float32, "bf16x3")}[precision]
    sysm = SyntheticSystem(o, v, naux=300, seed=seed)
    st = Stream()
    eng = TriplesEngine(o, v, stream=st, dtype=dt, contract=contract)
    eng.set_t3(sysm.t3_tri().astype(dt))
    eng.set_intermediates(sysm.bare_intermediates(), sysm.t2())
    if tiles and os.path.exists(tiles):
        cfg = json.load(open(tiles))
        # a tile_config() list of this size, or a tiles_by_name() map (from any size)
        eng.set_tiles_by_name(cfg) if isinstance(cfg, dict) else eng.set_tile_config(cfg)
    else:
        eng.autotune()
        if tiles:
            json.dump(eng.tiles_by_name(), open(tiles, "w"))
    eng.residual()
    ev = {}

    def timer(g, what):
        e = Event()
        e.record(st)
        ev.setdefault(id(g), []).append((what, e))

    a, b = Event(), Event()
    a.record(st)
    eng.residual(timer=timer)
    b.record(st)
    st.synchronize()
    launches = []
    for g in eng.launches:
        lst = ev.get(id(g), [])
        starts = [e for w, e in lst if w == "start"]
        ends = [e for w, e in lst if w == "end"]
        if starts and ends:
            launches.append({"name": g.name, "ms": starts[0].elapsed_ms(ends[0]), "flops": getattr(g, "flops", 0.0)})
    print(json.dumps({"probe": "triples residual", "precision": precision, "nocc": o, "nvir": v,
                      "residual_ms": a.elapsed_ms(b), "launches": launches}))


def calibrate_copy(mib):
    """dst = src over ``mib`` MiB of FP64 with dcc_tensor_axpby_f64 (reads and writes mib MiB each)."""
    import ctypes
    from distr_cc import _lib as L
    from distr_cc.device import DeviceArray, Stream
    n = mib * (1 << 20) // 8
    src, dst = DeviceArray((n,)), DeviceArray((n,))
    s = Stream()
    src.fill(1.0, s)
    lay = L.Strided()
    lay.ndim = 1
    lay.shape[0], lay.dst_stride[0], lay.src_stride[0] = n, 1, 1
    for _ in range(3):
        L.call("dcc_tensor_axpby_f64", dst.vp, src.vp, ctypes.byref(lay), 1.0, 0.0, s.handle)
    s.synchronize()
    print(json.dumps({"copy_mib": mib, "bytes_read_per_launch": n * 8, "bytes_written_per_launch": n * 8}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--nocc", type=int, default=10)
    ap.add_argument("--nvir", type=int, default=90)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--breakdown", action="store_true", help="print per-launch timings to stderr")
    ap.add_argument("--autotune", type=int, default=1, help="autotune the contraction tiles first (0 = no)")
    ap.add_argument("--tiles", metavar="JSON", default=None,
                    help="tile choice file: loaded (no autotune) if it exists, else written after autotuning")
    ap.add_argument("--mixed-iters", type=int, default=2,
                    help="full mixed-vs-FP64 iterations to time and compare (0 = skip)")
    ap.add_argument("--mixed-nocc", type=int, default=16)
    ap.add_argument("--mixed-nvir", type=int, default=160)
    ap.add_argument("--mixed-autotune", default="transfer", choices=("direct", "transfer", "none"),
                    help="tiles of the mixed-vs-FP64 engines: tuned on them, tuned at nocc=10 nvir=90, or defaults")
    ap.add_argument("--adjoint", type=int, default=1, help="time the T2->T3 VJP (0 = skip)")
    ap.add_argument("--batched", type=int, default=96, help="systems in the batched ensemble (0 = skip)")
    ap.add_argument("--q-nocc", type=int, default=5)
    ap.add_argument("--q-nvir", type=int, default=19, help="(Q) timing size (0 = skip)")
    ap.add_argument("--probe-residual", choices=("fp64", "fp32", "bf16x3"), default=None,
                    help="only time one engine's residual at --nocc/--nvir (tiles via --tiles) and exit")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (HBM counter calibration) and exit")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs the triples path on one GPU; --gpus %d is not supported "
              "(the canonical i<=j<=k blocks would shard over GPUs as DistributedT3IJK distributes them; "
              "no part of this path uses a second GPU)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    from distr_cc import _lib as L
    from distr_cc.device import DeviceArray, Event, Stream, init
    from distr_cc.synthetic import SyntheticSystem
    from distr_cc.t3_engine import TriplesEngine

    o, v = args.nocc, args.nvir
    init(0)
    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)
        return
    if args.probe_residual:
        probe_residual(o, v, args.probe_residual, args.tiles)
        return
    print("bench: building synthetic system o=%d v=%d" % (o, v), file=sys.stderr, flush=True)
    sysm = SyntheticSystem(o, v, naux=300, seed=20261202)
    im = sysm.bare_intermediates()
    t2, t3 = sysm.t2(), sysm.t3_tri()
    stream = Stream()
    eng = TriplesEngine(o, v, stream=stream)
    eng.set_t3(t3)
    eng.set_intermediates(im, t2)
    nmo = o + v
    # T1 = 0 for the synthetic system, so the T1-dressed integrals are the bare ones
    O, V = slice(0, o), slice(o, nmo)
    eng.set_t1_dressed(sysm.fock[:o, o:], sysm.eri_phys(O, O, V, V), sysm.eri_phys(O, O, O, V),
                       sysm.eri_phys(V, O, V, V))
    w_saved = [DeviceArray(w.shape, w.dtype) for w in (eng.W_oovo, eng.W_ovvv)]
    for dst, src in zip(w_saved, (eng.W_oovo, eng.W_ovvv)):
        L.call("dcc_memcpy_dtod", dst.vp, src.vp, src.nbytes, stream.handle)
    stream.synchronize()
    phase_ev = []

    def step(timer=None):
        """A2 + A3 + A4-A6 + A7-A9: T3 -> r1/r2, T3 -> W (on W rebuilt from T2), the residual's operands
        from the intermediates (ring weights, W_oooo/F_oo prefactors, F_vv folded into W_vvvv),
        residual, update."""
        for src, dst in zip(w_saved, (eng.W_oovo, eng.W_ovvv)):
            L.call("dcc_memcpy_dtod", dst.vp, src.vp, src.nbytes, stream.handle)
        eng.prepare_residual()
        if timer is not None:
            e = Event()
            e.record(stream)
            phase_ev.append(("start", e))
        eng.t3_to_r1r2(download=False)
        if timer is not None:
            e = Event()
            e.record(stream)
            phase_ev.append(("A2", e))
        eng.t3_to_w(reuse_z=True)
        if timer is not None:
            e = Event()
            e.record(stream)
            phase_ev.append(("A3", e))
        eng.residual(timer=timer)
        return eng.finalize(sysm.mo_energy)

    # dominant kernel: the three W_vvvv x T3 GEMMs (65% of the algorithmic flops)
    dom = [g for g in eng.launches if g.name.startswith("W_vvvv")]
    ev = {}

    def timer(g, what):
        if g in dom or args.breakdown:
            e = Event()
            e.record(stream)
            ev.setdefault(id(g), []).append((what, e))

    tuned = None
    if args.tiles and os.path.exists(args.tiles):
        eng.set_tile_config(json.load(open(args.tiles)))
        print("bench: tiles loaded from %s" % args.tiles, file=sys.stderr, flush=True)
    elif args.autotune:
        print("bench: autotuning contraction tiles", file=sys.stderr, flush=True)
        tuned = eng.autotune()
        if args.tiles:
            json.dump(eng.tile_config(), open(args.tiles, "w"))
    print("bench: warmup %d" % args.warmup, file=sys.stderr, flush=True)
    for _ in range(args.warmup):
        step()
    stream.synchronize()
    ev.clear()
    e0, e1 = Event(), Event()
    print("bench: timing %d steps" % args.steps, file=sys.stderr, flush=True)
    w0 = time.perf_counter()
    e0.record(stream)
    norms = None
    for _ in range(args.steps):
        norms = step(timer=timer)
    e1.record(stream)
    stream.synchronize()
    wall = time.perf_counter() - w0
    ms = e0.elapsed_ms(e1) / args.steps

    durs = {}
    for g in eng.launches:
        lst = ev.get(id(g), [])
        starts = [e for w, e in lst if w == "start"]
        ends = [e for w, e in lst if w == "end"]
        durs[id(g)] = [a.elapsed_ms(b) for a, b in zip(starts, ends)]
    dom_ms = [d for g in dom for d in durs[id(g)]]
    dom_avg_ms = float(np.mean(dom_ms))
    dom_flops = float(np.mean([g.flops for g in dom]))
    achieved = dom_flops / (dom_avg_ms * 1e-3) / 1e12
    if args.breakdown:
        print("bench: launch order per step: " + " | ".join(g.name for g in eng.launches), file=sys.stderr)
        tot = 0.0
        for g in eng.launches:
            d = np.mean(durs[id(g)]) if durs[id(g)] else float("nan")
            tot += d
            print("  %-36s %9.3f ms  %7.2f TF/s" % (g.name, d, g.flops / (d * 1e-3) / 1e12), file=sys.stderr)
        print("  sum of GEMM launches %.3f ms of %.3f ms per step" % (tot, ms), file=sys.stderr)
        a2 = [b.elapsed_ms(c) for (w1, b), (w2, c) in zip(phase_ev[0::3], phase_ev[1::3])]
        a3 = [b.elapsed_ms(c) for (w1, b), (w2, c) in zip(phase_ev[1::3], phase_ev[2::3])]
        print("  T3->r1/r2 with the T3->W terms (A2, A3 contractions) %.3f ms, adding the T3->W terms %.3f ms"
              % (np.mean(a2), np.mean(a3)), file=sys.stderr)

    # the HBM-bound fused projection/denominator/update (A7-A9): algorithmic bytes 3 S (read R, read and
    # write t3), S = n3 v^3 8 bytes; timed on its own after the steps (it keeps updating t3)
    fin = []
    for _ in range(5):
        a, b = Event(), Event()
        a.record(stream)
        eng.finalize_launch()
        b.record(stream)
        fin.append((a, b))
    stream.synchronize()
    fin_ms = float(np.mean([a.elapsed_ms(b) for a, b in fin[1:]]))
    fin_bytes = 3 * eng.n3 * v ** 3 * 8
    F = algorithmic_flops(o, v)
    out = {
        "metric": METRIC,
        "value": ms / 1e3,
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "plain-FP64 RCCSDT triples update: T3->r1/r2, T3->W, T3 residual, "
                               "projection/denominator/update, nocc=%d nvir=%d, %d canonical blocks"
                               % (o, v, eng.n3),
                   "nocc": o, "nvir": v, "naux": 300, "seed": 20261202, "precision": "fp64",
                   "residual_tflops": F / (ms * 1e-3) / 1e12,
                   "residual_frac_fp64_peak": F / (ms * 1e-3) / 1e12 / PEAK_FP64_MFMA_TFLOPS,
                   "r3_norm_sq": norms[0], "wall_s": wall,
                   "r3_finalize": {"launches": "dcc_r3_finalize_f64 + its norm reduction dcc_reduce_pairs_f64",
                                   "ms": fin_ms, "algorithmic_bytes": fin_bytes,
                                   "GB_per_s": fin_bytes / (fin_ms * 1e-3) / 1e9,
                                   "frac_hbm_peak": fin_bytes / (fin_ms * 1e-3) / 1e9 / PEAK_HBM_GBS},
                   "autotuned_tiles": None if tuned is None else
                   {k: v[0] for k, v in tuned.items() if k.split(" ", 1)[1].startswith(("W_vvvv", "ring"))}},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": PEAK_FP64_MFMA_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP64_MFMA_TFLOPS, "traffic": measured_traffic(o, v)[0],
                     "traffic_source": measured_traffic(o, v)[1],
                     "kernel": "contract_f64_kernel (W_vvvv x T3 launches)",
                     "flops_per_launch": dom_flops, "avg_launch_ms": dom_avg_ms},
    }
    if args.mixed_iters > 0:
        print("bench: mixed vs FP64 full iterations", file=sys.stderr, flush=True)
        del eng, step, dom, timer
        gc.collect()
        out["config"]["mixed_vs_fp64"] = mixed_vs_fp64(args.mixed_nocc, args.mixed_nvir, args.mixed_iters, autotune=args.mixed_autotune)
    if args.batched > 0:
        print("bench: batched ensemble of %d systems" % args.batched, file=sys.stderr, flush=True)
        out["config"]["batched"] = batched_ensemble(args.batched)
    if args.adjoint:
        print("bench: adjoint cost", file=sys.stderr, flush=True)
        out["config"]["adjoint"] = adjoint_cost(o, v)
    if args.q_nvir > 0:
        print("bench: (Q) timing", file=sys.stderr, flush=True)
        out["config"]["q"] = q_time(args.q_nocc, args.q_nvir)
    if not args.no_cpu_baseline:
        print("bench: CPU baseline (oracle, bounded sample)", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(o, v, sysm, im, t2, t3)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
