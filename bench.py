#!/usr/bin/env python
"""Benchmark of the RCCSDT iteration on one MI355X: mixed precision against FP64.

Headline workload: the synthetic system nocc=16, nvir=160 (816 canonical T3 blocks of 160^3; B-factor
ERIs with naux=300, seed 20261202).  One step = one full RCCSDT iteration on the device with the
triples contractions in split-BF16x3 (FP32 accumulation) and everything else in FP64: T1 dressing,
lower-order residual, T3 -> r1/r2/W, the T3 residual, projection/denominator/update, the unified
T1/T2/T3 DIIS extrapolation and the energy, with every input resident in HBM.  The same DIIS trajectory
is run with FP64 and with FP32 contractions for the per-iteration energy differences and speedups.
Then: the plain-FP64 residual at nocc=10, nvir=90 with its roofline numbers, H2O/cc-pVDZ RCCSDT + (Q)
end to end against the oracle on the host, the 96-system batched ensemble, the adjoint and (Q) timings.
Prints one JSON line.

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
PEAK_FP32_MFMA_TFLOPS = 157.3
PEAK_BF16_MFMA_TFLOPS = 2500.0
PEAK_HBM_GBS = 8000.0
SIGMA_B_DEFAULT = 0.01


def algorithmic_flops(o, v):
    """Flops of one T3 residual with canonical-block savings (reference einsum count)."""
    o3u = o * (o + 1) * (o + 2) // 6
    return o3u * (6 * v ** 5 + 24 * o * v ** 4 + 18 * v ** 4 + 12 * o * v ** 3 + 6 * o * o * v ** 3
                  + 6 * o * v ** 3) + o ** 3 * (4 * v ** 4 + (4 * o + 2) * v ** 3)


def host_cores():
    """(cores this process may run on, physical cores of the host per lscpu or None)."""
    avail = len(os.sched_getaffinity(0))
    phys = None
    try:
        import subprocess
        out = subprocess.run(["lscpu", "-p=CORE,SOCKET"], capture_output=True, text=True, timeout=10).stdout
        phys = len({ln for ln in out.splitlines() if ln and not ln.startswith("#")})
    except Exception:
        pass
    return avail, phys


def cpu_baseline(inp, reps=3, blocks=1):
    """The oracle (the reference's block loop, NumPy/BLAS FP64) on a bounded sample of the same
    iteration: the triples residual of ``blocks`` target block(s) (every source block that feeds it)
    and their projection/denominator/update, median of ``reps`` runs, scaled by the number of
    canonical blocks to one iteration's triples residual.  BLAS uses every core this process may run
    on."""
    from oracle import rccsdt_tri as tri
    o, v = inp.sysm.nocc, inp.sysm.nvir
    avail, phys = host_cores()
    try:
        from threadpoolctl import threadpool_limits
        limiter = threadpool_limits(limits=avail)
    except Exception:
        limiter = None
    im = inp.sysm.bare_intermediates()
    n3 = inp.t3.shape[0]
    S = list(range(blocks))
    times = []
    for _ in range(reps):
        t0 = time.perf_counter()
        r3 = tri.compute_r3_tri(o, v, im, inp.t2, inp.t3, targets=S)
        proj = tri.project_p3_full(r3[S])
        eia = inp.sysm.mo_energy[:o, None] - inp.sysm.mo_energy[None, o:]
        for n, idx in enumerate(S):
            i, j, k = tri.canonical_triples(o)[idx]
            _ = inp.t3[idx] + proj[n] / (eia[i, :, None, None] + eia[j, None, :, None] + eia[k, None, None, :])
        times.append(time.perf_counter() - t0)
        del r3, proj
    if limiter is not None:
        limiter.unregister() if hasattr(limiter, "unregister") else None
    med = float(np.median(times))
    return {"value": med * n3 / len(S) * 1e3, "unit": "ms", "cores": int(avail), "kind": "port",
            "host_physical_cores": phys, "ranks_x_threads": "1 x %d" % avail,
            "sample": "oracle triples residual + update of %d of %d canonical blocks at nocc=%d nvir=%d "
                      "(median of %d runs: %s s), scaled by %d/%d to one iteration's triples residual "
                      "(the FP64 lower-order part of the iteration is not in the sample)"
                      % (len(S), n3, o, v, reps, ", ".join("%.1f" % t for t in times), n3, len(S))}


def measured_traffic(precision, o, v, pairs=False):
    """HBM bytes per W_vvvv launch (the pair-symmetric launches if ``pairs``) from the committed PMC
    summary of this workload, if any."""
    path = os.path.join(ROOT, "profiles", "contract_%s_w_vvvv%s_o%d_v%d.md" % (precision, "_pairs" if pairs else "",
                                                                              o, v))
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
                       triples_contract=contract, vvvv_pairs=VVVV_PAIRS["mode"])
    eng.set_amplitudes(np.zeros((o, v)), sysm.t2(), sysm.t3_tri())
    eng.update_amps()
    eng.tr.autotune()
    st.synchronize()
    tiles = eng.tr.tiles_by_name()
    del eng
    gc.collect()
    return tiles


PRECISIONS = {"fp64": (np.float64, "native"), "fp32": (np.float32, "native"), "bf16x3": (np.float32, "bf16x3")}
TILE_DIR = os.path.join(ROOT, "distributed-cc_amd", "distr_cc", "tiles")
# W_vvvv terms as pair-symmetric products ("auto": when the device has room for their weights) or as the
# plain GEMMs (False); set from --vvvv-pairs
VVVV_PAIRS = {"mode": "auto"}


def tuned_tiles(precision, o, v):
    """{launch name: tile} saved by `bench.py --tune-tiles` for this precision and size, if any."""
    path = os.path.join(TILE_DIR, "%s_o%d_v%d.json" % (precision, o, v))
    if os.path.exists(path):
        return json.load(open(path)), os.path.relpath(path, ROOT)
    return None, None


class _Inputs:
    """Host arrays of one synthetic system, built once and shared by the engines of every precision."""

    def __init__(self, o, v, seed, sigma_b):
        from distr_cc.synthetic import SyntheticSystem
        t0 = time.perf_counter()
        self.sigma_b = sigma_b
        self.sysm = SyntheticSystem(o, v, naux=300, seed=seed, sigma_b=sigma_b)
        self.pppp = self.sysm.pppp()
        self.t2 = self.sysm.t2()
        self.t3 = self.sysm.t3_tri()
        self.build_s = time.perf_counter() - t0


def iterate(inp, precision, iters, warmup, nvir_diis, tiles=None, launch_prefix=None, stop_de=None, log=""):
    """RCCSDT iterations of the synthetic system on the device in one precision of the triples
    contractions (FP64 everywhere else), with the reference's unified T1/T2/T3 DIIS (space 6, start at
    cycle 0, T3 part on the first nvir_diis virtuals; rccsdt.py:1026-1148): ``warmup`` untimed
    iterations, then ``iters`` timed ones (HIP events on the engine stream around each iteration:
    update_amps, DIIS extrapolation and the energy).  ``stop_de``: stop once |dE| < stop_de after the
    warm-up.  Launches whose name starts with ``launch_prefix`` are timed one by one."""
    from distr_cc.device import DeviceArray, Event, Stream
    from distr_cc.diis import DeviceDIIS
    from distr_cc.rccsdt_engine import RCCSDTEngine
    o, v = inp.sysm.nocc, inp.sysm.nvir
    dt, contract = PRECISIONS[precision]
    st = Stream()
    eng = RCCSDTEngine(o, v, inp.sysm.fock, inp.pppp, inp.sysm.mo_energy, stream=st, t3_dtype=dt,
                       triples_contract=contract, nvir_diis=nvir_diis, vvvv_pairs=VVVV_PAIRS["mode"])
    if tiles:
        eng.tr.set_tiles_by_name(tiles)
    eng.set_amplitudes(np.zeros((o, v)), inp.t2, inp.t3)
    adiis = DeviceDIIS(eng.diis_size(), 6, stream=st)
    vec = DeviceArray((eng.diis_size(),))
    ev = {}

    def ltimer(g, what):
        if launch_prefix and g.name.startswith(launch_prefix):
            e = Event()
            e.record(st)
            ev.setdefault(id(g), (g, []))[1].append((what, e))

    phase_ev = []

    def ptimer(ph, what):
        e = Event()
        e.record(st)
        phase_ev.append((ph, what, e))

    energies, ms, res_norms = [eng.energy()], [], []
    w0 = time.perf_counter()
    for it in range(warmup + iters):
        timed = it >= warmup
        a, b = Event(), Event()
        a.record(st)
        res = eng.update_amps(timer=ptimer if timed else None, launch_timer=ltimer if timed else None)
        eng.pack(vec)
        xnew = adiis.update(vec)
        eng.unpack(xnew)
        xnew.free()
        energies.append(eng.energy())
        b.record(st)
        st.synchronize()
        res_norms.append(float(np.linalg.norm(res)))
        if timed:
            ms.append(a.elapsed_ms(b))
        print("bench: %s%s iteration %2d%s: %8.1f ms  E_corr % .12f  dE % .3e  |res| %.3e"
              % (log, precision, it + 1, "" if timed else " (warm-up)", a.elapsed_ms(b), energies[-1],
                 energies[-1] - energies[-2], res_norms[-1]), file=sys.stderr, flush=True)
        if stop_de is not None and timed and abs(energies[-1] - energies[-2]) < stop_de:
            break
    wall = time.perf_counter() - w0
    phases = {}
    starts = {}
    for ph, what, e in phase_ev:
        if what == "start":
            starts[ph] = e
        else:
            phases.setdefault(ph, []).append(starts[ph].elapsed_ms(e))
    launches = []
    for g, lst in ev.values():
        ss = [e for w, e in lst if w == "start"]
        es = [e for w, e in lst if w == "end"]
        launches += [(g.name, g.flops, x.elapsed_ms(y)) for x, y in zip(ss, es)]
    out = {"energies": energies, "ms": ms, "res_norms": res_norms, "wall_s": wall,
           "phase_ms": {k: float(np.mean(x)) for k, x in phases.items()}, "launches": launches,
           "residual_flops": eng.tr.residual_flops() if hasattr(eng.tr, "residual_flops") else None}
    del eng, adiis, vec, ltimer, ptimer
    gc.collect()   # the engine's programs hold reference cycles; free its device memory now
    return out


def _converged_at(energies, tol):
    for n in range(1, len(energies)):
        if abs(energies[n] - energies[n - 1]) < tol:
            return n
    return None


def mixed_vs_fp64(inp, steps, warmup, nvir_diis, fp64_iters, fp32_iters, tiles_mode):
    """The headline: split-BF16x3 triples contractions (FP64 denominators, energy, DIIS) for ``warmup``
    + ``steps`` DIIS iterations, the timed ones giving ms/iteration; then the same trajectory with FP64
    contractions (``fp64_iters`` iterations) and FP32 ones (``fp32_iters``): per-iteration |E - E_fp64|
    and ms/iteration of each."""
    o, v = inp.sysm.nocc, inp.sysm.nvir
    runs, tile_src = {}, {}
    plan = [("bf16x3", steps, warmup)]
    if fp64_iters > 1:
        plan.append(("fp64", fp64_iters - 1, 1))
    if fp32_iters > 1:
        plan.append(("fp32", fp32_iters - 1, 1))
    for name, n_it, n_w in plan:
        tiles, src = tuned_tiles(name, o, v)
        if tiles is None and tiles_mode == "transfer":
            dt, contract = PRECISIONS[name]
            tiles, src = _tiles_from_small(dt, contract, 20261202, inp.sigma_b), "autotuned at nocc=10 nvir=90"
        tile_src[name] = src or "library defaults"
        print("bench: %s engine nocc=%d nvir=%d (%d + %d iterations, tiles: %s)"
              % (name, o, v, n_w, n_it, tile_src[name]), file=sys.stderr, flush=True)
        runs[name] = iterate(inp, name, n_it, n_w, nvir_diis, tiles=tiles,
                             launch_prefix="W_vvvv" if name in ("bf16x3", "fp64") else None)
    return runs, tile_src


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
    """The whole triples residual r3(t2, <pq|rs>) of one iteration at T1 = 0 (T2-dressed intermediates,
    T3 -> W terms, residual) and its hand-written VJP (distr_cc.adjoint.TriplesResidualAdjoint) on the
    synthetic system: mean device time of each, their ratio, and the device memory the adjoint holds."""
    from distr_cc.adjoint import TriplesResidualAdjoint
    from distr_cc.device import Event, Stream, mem_get_info
    from distr_cc.synthetic import SyntheticSystem
    sysm = SyntheticSystem(o, v, naux=300, seed=seed)
    pppp, t2, t3 = sysm.pppp(), sysm.t2(), sysm.t3_tri()
    st = Stream()
    gc.collect()
    free0, _ = mem_get_info()
    adj = TriplesResidualAdjoint(o, v, sysm.fock, pppp, t3, stream=st)
    adj.set_inputs(t2, pppp)
    adj.lam.upload(np.ascontiguousarray(t3), st)
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

    fwd = timed(adj.forward)
    bwd = timed(adj.backward)
    free2, _ = mem_get_info()
    # per-launch breakdown of one more backward
    ev = []

    def timer(name, what):
        e = Event()
        e.record(st)
        ev.append((name, what, e))

    adj.backward(timer=timer)
    st.synchronize()
    parts, start = {}, {}
    for name, what, e in ev:
        if what == "start":
            start[name] = e
        else:
            parts[name] = parts.get(name, 0.0) + start.pop(name).elapsed_ms(e)
    top = dict(sorted(parts.items(), key=lambda kv: -kv[1])[:10])
    out = {"workload": "whole triples residual of one iteration (T2-dressed intermediates, T3->W terms, "
                       "residual; T1 = 0) and its VJP w.r.t. t2 and the (ov|vv)/(oo|ov) integrals, "
                       "nocc=%d nvir=%d" % (o, v),
           "forward_ms": fwd, "backward_ms": bwd, "backward_over_forward": bwd / fwd,
           "device_bytes_held": int(free0 - min(free1, free2)),
           "dense_workspace_bytes": int(2 * o ** 3 * v ** 3 * 8),
           "backward_top_launches_ms": top,
           "note": "backward rebuilds the full Lambda and T3 from the canonical blocks on every call "
                   "(dense_workspace_bytes of the held memory); no allocation inside forward or backward"}
    del adj
    gc.collect()
    return out


def _fp64_solve(eng, amps, tol=1e-11, max_iter=60):
    """One system's own FP64 RCCSDT solve: its engine's ungrouped launches with the unified T1/T2/T3
    DIIS (space 6), from the given amplitudes until |dE| < tol; returns (energy, iterations)."""
    from distr_cc.device import DeviceArray
    from distr_cc.diis import DeviceDIIS
    eng.set_amplitudes(*amps)
    adiis = DeviceDIIS(eng.diis_size(), 6, stream=eng.stream)
    vec = DeviceArray((eng.diis_size(),))
    e_old, e = 0.0, eng.energy()
    it = 0
    for it in range(1, max_iter + 1):
        eng.update_amps()
        eng.pack(vec)
        x = adiis.update(vec)
        eng.unpack(x)
        x.free()
        e_old, e = e, eng.energy()
        if abs(e - e_old) < tol:
            break
    vec.free()
    del adiis
    return e, it


def batched_ensemble(nsys, steps=20, seed=20261202, sigma_b=SIGMA_B_DEFAULT, solves=True, graph=True):
    """The batched ensemble: nsys synthetic systems, nocc ~ U{3..8}, nvir ~ U{20..64} (system s drawn
    from seed + s, B-factor sigma ``sigma_b``), whole FP64 RCCSDT iterations of all systems per step
    (lower order, T3 -> r1/r2/W, triples residual as grouped persistent launches with autotuned tiles,
    update, energy) captured in one HIP graph and replayed ``steps`` times from the MP2/T3 guess (the
    timed region).  Then every system is solved on its own (its engine's ungrouped launches with DIIS,
    to |dE| < 1e-11) and the ensemble's energies after the timed steps are compared with those solves
    (skipped with ``solves=False``, for profiling the ensemble's own launches; ``graph=False`` issues the
    same steps eagerly instead of replaying the graph, because rocprofv3's kernel tracing aborts the queue
    on graph-launched kernels with a malformed-packet error).  With sigma 0.05 the
    plain Jacobi steps diverge for every system (and the MP2 guess energies are tens of Eh), so the
    ensemble runs at sigma 0.01, like the nocc=16, nvir=160 workload."""
    from distr_cc.batched import BatchedRCCSDT
    from distr_cc.device import Event, Stream
    from distr_cc.synthetic import SyntheticSystem
    syss, amps = [], []
    t0 = time.perf_counter()
    for k in range(nsys):
        r = np.random.default_rng(seed + k)
        o, v = int(r.integers(3, 9)), int(r.integers(20, 65))
        sysm = SyntheticSystem(o, v, naux=300, seed=seed + k, sigma_b=sigma_b)
        syss.append((o, v, sysm.fock, sysm.pppp(), sysm.mo_energy))
        amps.append((np.zeros((o, v)), sysm.t2(), sysm.t3_tri()))
    st = Stream()
    bt = BatchedRCCSDT(syss, st)
    build_s = time.perf_counter() - t0
    print("bench: %d engines built in %.1f s; autotuning the grouped tiles" % (nsys, build_s), file=sys.stderr,
          flush=True)
    tuned = bt.autotune()

    def restart():
        for k, a in enumerate(amps):
            bt.set_amplitudes(k, *a)

    # one step issued without the graph, with events around its phases
    restart()
    ev = []

    def timer(name, what):
        e = Event()
        e.record(st)
        ev.append((name, what, e))

    bt.issue_step(timer=timer)
    st.synchronize()
    phase_ms, t0s = {}, {}
    for name, what, e in ev:
        if what == "start":
            t0s[name] = e
        else:
            phase_ms[name] = t0s.pop(name).elapsed_ms(e)
    restart()
    if graph:
        bt.capture()
    bt.step()                          # warm-up replay
    st.synchronize()
    restart()
    # steps - 1 replays, the energies read on the host, then the last replay: both segments are timed
    # with events, the host read between them is not
    a, b, c, d = Event(), Event(), Event(), Event()
    a.record(st)
    for _ in range(steps - 1):
        bt.step()
    b.record(st)
    e_prev = bt.energies()
    c.record(st)
    bt.step()
    d.record(st)
    e_batched = bt.energies()
    ms = (a.elapsed_ms(b) + c.elapsed_ms(d)) / steps
    last_de = [abs(x - y) for x, y in zip(e_batched, e_prev)]
    print("bench: ensemble %.1f ms per step; per-system FP64 DIIS solves" % ms, file=sys.stderr, flush=True)
    t1 = time.perf_counter()
    sol = [_fp64_solve(e, amps[k]) for k, e in enumerate(bt.engines)] if solves else []
    solve_s = time.perf_counter() - t1
    de = [abs(e_batched[k] - sol[k][0]) for k in range(len(sol))] or [float("nan")]
    shapes = {}
    for k, (shape, _) in tuned.items():
        shapes[shape] = shapes.get(shape, 0) + 1
    out = {"workload": "%d synthetic systems, nocc~U{3..8}, nvir~U{20..64}, B-factor sigma %g, whole FP64 RCCSDT "
                       "iterations of every system per step (grouped persistent triples GEMMs with autotuned tiles, "
                       "everything in one HIP graph per step), %d steps from the MP2/T3 guess%s"
                       % (nsys, sigma_b, steps, "" if graph else " (steps issued eagerly, not replayed as a graph)"),
           "ms_per_step": ms, "system_iterations_per_s": nsys / (ms * 1e-3), "steps": steps,
           "triples_contraction_tflops": bt.flops() / (ms * 1e-3) / 1e12,
           "phase_ms_one_step_without_graph": phase_ms,
           "autotuned_tile_shapes": shapes, "engines_build_s": build_s,
           "e_corr_min_max": [min(e_batched), max(e_batched)],
           "max_abs_dE_last_step": float(max(last_de)),
           "max_abs_dE_vs_per_system_fp64_solves": float(max(de)),
           "median_abs_dE_vs_per_system_fp64_solves": float(np.median(de)),
           "per_system_solve_iterations_min_max": [min(x[1] for x in sol), max(x[1] for x in sol)] if sol else None,
           "per_system_solves_s": solve_s,
           "finite_energies": bool(np.all(np.isfinite(e_batched)))}
    bt.close()
    del bt
    gc.collect()
    return out


def probe_residual(o, v, precision, tiles, seed=20261202, vvvv_pairs=False):
    """One triples engine of the given precision on the synthetic system (nocc=o, nvir=v): tiles loaded
    from ``tiles`` if it exists (a tile list of this size or a {launch name: tile} map), else autotuned
    (and saved there by launch name when a path is given); the residual is run twice and the second one
    is timed, per launch, followed by one fused projection/denominator/update.  Used to profile one
    precision's kernels."""
    from distr_cc.device import Event, Stream
    from distr_cc.synthetic import SyntheticSystem
    from distr_cc.t3_engine import TriplesEngine
    dt, contract = {"fp64": (np.float64, "native"), "fp32": (np.float32, "native"),
                    "bf16x3": (np.float32, "bf16x3")}[precision]
    sysm = SyntheticSystem(o, v, naux=300, seed=seed)
    st = Stream()
    eng = TriplesEngine(o, v, stream=st, dtype=dt, contract=contract, vvvv_pairs=vvvv_pairs)
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
    # then the fused projection/denominator/update of the same R (denominators from the orbital energies)
    eng.set_denominators(sysm.mo_energy)
    c, d = Event(), Event()
    c.record(st)
    eng.finalize_launch()
    d.record(st)
    st.synchronize()
    launches = []
    for g in eng.launches:
        lst = ev.get(id(g), [])
        starts = [e for w, e in lst if w == "start"]
        ends = [e for w, e in lst if w == "end"]
        if starts and ends:
            launches.append({"name": g.name, "ms": starts[0].elapsed_ms(ends[0]), "flops": getattr(g, "flops", 0.0)})
    print(json.dumps({"probe": "triples residual", "precision": precision, "nocc": o, "nvir": v,
                      "residual_ms": a.elapsed_ms(b), "finalize_ms": c.elapsed_ms(d), "launches": launches}))


def tune_tiles(o, v, precision, sigma_b, seed=20261202):
    """Autotune every contraction of one precision's engine on the synthetic system (nocc=o, nvir=v)
    after one iteration, and save {launch name: tile} where the bench looks for it."""
    from distr_cc.device import Stream
    from distr_cc.rccsdt_engine import RCCSDTEngine
    inp = _Inputs(o, v, seed, sigma_b)
    dt, contract = PRECISIONS[precision]
    st = Stream()
    eng = RCCSDTEngine(o, v, inp.sysm.fock, inp.pppp, inp.sysm.mo_energy, stream=st, t3_dtype=dt,
                       triples_contract=contract, vvvv_pairs=VVVV_PAIRS["mode"])
    eng.set_amplitudes(np.zeros((o, v)), inp.t2, inp.t3)
    eng.update_amps()
    print("bench: autotuning %s tiles at nocc=%d nvir=%d" % (precision, o, v), file=sys.stderr, flush=True)
    tuned = eng.tr.autotune()
    st.synchronize()
    os.makedirs(TILE_DIR, exist_ok=True)
    path = os.path.join(TILE_DIR, "%s_o%d_v%d.json" % (precision, o, v))
    json.dump(eng.tr.tiles_by_name(), open(path, "w"), indent=0, sort_keys=True)
    print(json.dumps({"tuned": precision, "nocc": o, "nvir": v, "file": os.path.relpath(path, ROOT),
                      "launches": {k: list(x) for k, x in tuned.items()} if isinstance(tuned, dict) else None}))


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


def plain_fp64_residual(o, v, steps, warmup, autotune=True, tiles=None, breakdown=False):
    """The synthetic plain-FP64 T3 residual (nocc=10, nvir=90): T3 -> r1/r2, T3 -> W, the residual's
    operands from the intermediates, the residual and the fused projection/denominator/update, with t3
    restored from a device copy before every step (outside the timed region) so that every step runs on
    the same inputs.  Returns the config entry and the FP64 W_vvvv launch timings."""
    from distr_cc import _lib as L
    from distr_cc.device import DeviceArray, Event, Stream
    from distr_cc.synthetic import SyntheticSystem
    from distr_cc.t3_engine import TriplesEngine
    sysm = SyntheticSystem(o, v, naux=300, seed=20261202)
    im = sysm.bare_intermediates()
    t2, t3 = sysm.t2(), sysm.t3_tri()
    stream = Stream()
    eng = TriplesEngine(o, v, stream=stream, vvvv_pairs=VVVV_PAIRS["mode"])
    eng.set_t3(t3)
    eng.set_intermediates(im, t2)
    eng.set_denominators(sysm.mo_energy)
    nmo = o + v
    # T1 = 0 for the synthetic system, so the T1-dressed integrals are the bare ones
    O, V = slice(0, o), slice(o, nmo)
    eng.set_t1_dressed(sysm.fock[:o, o:], sysm.eri_phys(O, O, V, V), sysm.eri_phys(O, O, O, V),
                       sysm.eri_phys(V, O, V, V))
    saved = [DeviceArray(w.shape, w.dtype) for w in (eng.W_oovo, eng.W_ovvv, eng.t3)]
    for dst, src in zip(saved, (eng.W_oovo, eng.W_ovvv, eng.t3)):
        L.call("dcc_memcpy_dtod", dst.vp, src.vp, src.nbytes, stream.handle)
    stream.synchronize()
    dom = [g for g in eng.launches if g.name.startswith("W_vvvv") and g.flops > 0]
    ev = {}

    def timer(g, what):
        if g in dom or breakdown:
            e = Event()
            e.record(stream)
            ev.setdefault(id(g), []).append((what, e))

    def restore():
        for src, dst in zip(saved, (eng.W_oovo, eng.W_ovvv, eng.t3)):
            L.call("dcc_memcpy_dtod", dst.vp, src.vp, src.nbytes, stream.handle)

    def step(tm=None):
        eng.prepare_residual()
        eng.t3_to_r1r2(download=False)
        eng.t3_to_w(reuse_z=True)
        eng.residual(timer=tm)
        eng.finalize_launch()

    tuned = None
    if tiles and os.path.exists(tiles):
        eng.set_tile_config(json.load(open(tiles)))
    elif autotune:
        print("bench: autotuning the FP64 contraction tiles at nocc=%d nvir=%d" % (o, v), file=sys.stderr,
              flush=True)
        tuned = eng.autotune()
        if tiles:
            json.dump(eng.tile_config(), open(tiles, "w"))
    for _ in range(warmup):
        restore()
        step()
    stream.synchronize()
    ev.clear()
    ms, norms = [], []
    for _ in range(steps):
        restore()
        a, b = Event(), Event()
        a.record(stream)
        step(timer)
        b.record(stream)
        norms.append(eng.read_norms()[0])
        ms.append(a.elapsed_ms(b))
    durs = {}
    for g in eng.launches:
        lst = ev.get(id(g), [])
        starts = [e for w, e in lst if w == "start"]
        ends = [e for w, e in lst if w == "end"]
        durs[id(g)] = [x.elapsed_ms(y) for x, y in zip(starts, ends)]
    if breakdown:
        for g in eng.launches:
            d = np.mean(durs[id(g)]) if durs[id(g)] else float("nan")
            print("  %-36s %9.3f ms  %7.2f TF/s" % (g.name, d, g.flops / (d * 1e-3) / 1e12), file=sys.stderr)
    # the HBM-bound fused projection/denominator/update alone (algorithmic bytes 3 S: read R, read and
    # write t3, S = n3 v^3 8 bytes)
    fin = []
    for _ in range(6):
        a, b = Event(), Event()
        a.record(stream)
        eng.finalize_launch()
        b.record(stream)
        fin.append((a, b))
    stream.synchronize()
    fin_ms = float(np.mean([x.elapsed_ms(y) for x, y in fin[1:]]))
    fin_bytes = 3 * eng.n3 * v ** 3 * 8
    step_ms = float(np.mean(ms))
    F = algorithmic_flops(o, v)
    dom_ms = [d for g in dom for d in durs[id(g)]]
    entry = {"workload": "plain-FP64 RCCSDT triples update: T3->r1/r2, T3->W, T3 residual, "
                         "projection/denominator/update, nocc=%d nvir=%d, %d canonical blocks, t3 restored "
                         "before every step" % (o, v, eng.n3),
             "steps": steps, "ms_per_step": step_ms, "ms_per_step_min_max": [min(ms), max(ms)],
             # the reference's einsum flop count per time (the pair-symmetric W_vvvv products execute
             # about half of its W_vvvv share): work delivered, not the rate of the matrix cores
             "algorithmic_tflops": F / (step_ms * 1e-3) / 1e12,
             "algorithmic_frac_fp64_peak": F / (step_ms * 1e-3) / 1e12 / PEAK_FP64_MFMA_TFLOPS,
             "vvvv_pairs": bool(eng.vvvv_pairs),
             "r3_norm_sq_every_step": [norms[0], max(abs(x - norms[0]) for x in norms)],
             "w_vvvv": {"launches": len(dom_ms), "avg_ms": float(np.mean(dom_ms)),
                        "tflops": float(np.mean([g.flops for g in dom])) / (np.mean(dom_ms) * 1e-3) / 1e12,
                        "frac_fp64_peak": float(np.mean([g.flops for g in dom])) / (np.mean(dom_ms) * 1e-3)
                        / 1e12 / PEAK_FP64_MFMA_TFLOPS},
             "r3_finalize": {"launches": "dcc_r3_finalize_f64 + its norm reduction dcc_reduce_pairs_f64",
                             "ms": fin_ms, "algorithmic_bytes": fin_bytes,
                             "GB_per_s": fin_bytes / (fin_ms * 1e-3) / 1e9,
                             "frac_hbm_peak": fin_bytes / (fin_ms * 1e-3) / 1e9 / PEAK_HBM_GBS},
             "autotuned_tiles": None if tuned is None else
             {k: x[0] for k, x in tuned.items() if k.split(" ", 1)[1].startswith(("W_vvvv", "ring"))}}
    del eng, saved, step, timer
    gc.collect()
    return entry


def c1_end_to_end(cpu=True):
    """H2O, O (0,0,0), H (0,+-0.757,0.587) Angstrom, cc-pVDZ, RHF, all electrons (nocc=5, nvir=19):
    RCCSDT to |dE| < 1e-9 (reference defaults otherwise) and (Q) at blksize 4, timed on the GPU
    (RCCSDT(mf, comm).kernel() including its device ao2mo, then rccsdt_q.kernel), and the oracle's
    FP64 CPU run of the same (update_amps + the reference DIIS + (Q)) timed beside it."""
    from distr_cc import rccsdt_q
    from distr_cc.distribute_t3 import SINGLE_RANK
    from distr_cc.molecule import RHF, Molecule
    from distr_cc.rccsdt import RCCSDT
    mol = Molecule([("O", (0., 0., 0.)), ("H", (0., -0.757, 0.587)), ("H", (0., 0.757, 0.587))], "cc-pvdz")
    mf = RHF(mol, conv_tol=1e-12)
    mf.kernel()
    out = {"workload": "H2O/cc-pVDZ all-electron (nocc=5, nvir=19): RCCSDT conv_tol 1e-9, then (Q) blksize 4"}
    for rep in range(2):   # the second run is reported (the first loads and plans every launch)
        w0 = time.perf_counter()
        cc = RCCSDT(mf, SINGLE_RANK)
        cc.conv_tol, cc.verbose = 1e-9, 0
        cc.kernel()
        w1 = time.perf_counter()
        eb, ep = rccsdt_q.kernel(cc, comm=SINGLE_RANK, blksize=4)
        w2 = time.perf_counter()
    out.update({"gpu_rccsdt_s": w1 - w0, "gpu_q_s": w2 - w1, "gpu_total_s": w2 - w0, "cycles": cc.cycles,
                "e_corr": cc.e_corr, "e_bracket": eb, "e_paren": ep})
    if cpu:
        from oracle import q_energy, rccsdt_lower as lo, rccsdt_tri as tri
        from oracle.diis import DIIS
        from oracle.t3_layout import full_t3_from_tri
        C, o = mf.mo_coeff, mol.nelectron // 2
        w0 = time.perf_counter()
        pppp = np.einsum("pqrs,pP,qQ,rR,sS->PRQS", mf._eri, C, C, C, C, optimize=True)
        e = mf.mo_energy
        fock, v = np.diag(e), len(e) - o
        _, t1, t2 = tri.init_amps_rhf(o, fock, e, pppp)
        t3 = np.zeros((o * (o + 1) * (o + 2) // 6, v, v, v))
        d, e_old, e_cc = DIIS(6), 0.0, tri.energy_rhf(o, fock, pppp, t1, t2)
        pairs = [(i, j) for i in range(o) for j in range(i, o)]
        for it in range(100):
            t1, t2, t3, res = lo.update_amps(o, fock, pppp, e, t1, t2, t3)
            x = d.update(np.concatenate([t1.ravel()] + [t2[i, j].ravel() for i, j in pairs] + [t3.ravel()]))
            t1 = x[:o * v].reshape(o, v).copy()
            off, t2 = o * v, t2.copy()
            for i, j in pairs:
                t2[i, j] = x[off:off + v * v].reshape(v, v)
                t2[j, i] = t2[i, j].T
                off += v * v
            t3 = x[off:].reshape(t3.shape).copy()
            e_old, e_cc = e_cc, tri.energy_rhf(o, fock, pppp, t1, t2)
            if abs(e_cc - e_old) < 1e-9 and np.linalg.norm(res) < 1e-6:
                break
        w1 = time.perf_counter()
        ceb, cep = q_energy.kernel(o, v, pppp, t2, full_t3_from_tri(t3, o), e, blksize=4)
        w2 = time.perf_counter()
        out.update({"cpu_oracle_rccsdt_s": w1 - w0, "cpu_oracle_q_s": w2 - w1, "cpu_oracle_total_s": w2 - w0,
                    "cpu_cores": host_cores()[0], "cpu_cycles": it + 1,
                    "abs_dE_gpu_vs_cpu": [abs(cc.e_corr - e_cc), abs(eb - ceb), abs(ep - cep)]})
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20, help="timed split-BF16x3 iterations at nocc=16 nvir=160")
    ap.add_argument("--warmup", type=int, default=5, help="untimed split-BF16x3 iterations before them")
    ap.add_argument("--nocc", type=int, default=16)
    ap.add_argument("--nvir", type=int, default=160)
    ap.add_argument("--sigma-b", type=float, default=SIGMA_B_DEFAULT,
                    help="B-factor sigma of the synthetic integrals of the headline system")
    ap.add_argument("--nvir-diis", type=int, default=40, help="virtuals of the T3 part of the DIIS vector")
    ap.add_argument("--fp64-iters", type=int, default=14, help="iterations of the FP64 reference trajectory")
    ap.add_argument("--fp32-iters", type=int, default=3, help="iterations with FP32 contractions (0 = skip)")
    ap.add_argument("--tiles-mode", default="transfer", choices=("transfer", "none"),
                    help="tiles of a precision without a saved file: tuned at nocc=10 nvir=90, or defaults")
    ap.add_argument("--tune-tiles", choices=("fp64", "fp32", "bf16x3"), default=None,
                    help="only autotune that precision's engine at --nocc/--nvir and save its tiles, then exit")
    ap.add_argument("--c2", type=int, default=1, help="plain-FP64 residual at nocc=10 nvir=90 (0 = skip)")
    ap.add_argument("--c2-steps", type=int, default=10)
    ap.add_argument("--c1", type=int, default=1, help="H2O/cc-pVDZ RCCSDT + (Q) end to end (0 = skip)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--breakdown", action="store_true", help="print per-launch timings of the C2 step to stderr")
    ap.add_argument("--adjoint", type=int, default=1, help="time the T2->T3 VJP (0 = skip)")
    ap.add_argument("--batched", type=int, default=96, help="systems in the batched ensemble (0 = skip)")
    ap.add_argument("--q-nocc", type=int, default=5)
    ap.add_argument("--q-nvir", type=int, default=19, help="(Q) timing size (0 = skip)")
    ap.add_argument("--probe-residual", choices=("fp64", "fp32", "bf16x3"), default=None,
                    help="only time one engine's residual at --nocc/--nvir (tiles via --tiles) and exit")
    ap.add_argument("--tiles", metavar="JSON", default=None, help="tile file of --probe-residual")
    ap.add_argument("--vvvv-pairs", type=int, default=1,
                    help="pair-symmetric W_vvvv products in the triples engines when the device has room "
                         "(0 = plain GEMMs everywhere)")
    ap.add_argument("--probe-batched", type=int, default=0, metavar="NSYS",
                    help="only the batched ensemble of NSYS systems (--steps steps issued eagerly, no "
                         "per-system solves); prints its JSON (for profiling)")
    ap.add_argument("--probe-q", type=int, default=0, metavar="BLKSIZE",
                    help="only the (Q) timing at --nocc/--nvir with this block size; prints its JSON "
                         "(for profiling)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (HBM counter calibration) and exit")
    args = ap.parse_args()
    VVVV_PAIRS["mode"] = "auto" if args.vvvv_pairs else False
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs the triples path on one GPU; --gpus %d is not supported "
              "(the canonical i<=j<=k blocks would shard over GPUs as DistributedT3IJK distributes them; "
              "no part of this path uses a second GPU)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    from distr_cc.device import init
    o, v = args.nocc, args.nvir
    init(0)
    if args.calibrate_copy:
        calibrate_copy(args.calibrate_copy)
        return
    if args.probe_residual:
        probe_residual(o, v, args.probe_residual, args.tiles, vvvv_pairs=VVVV_PAIRS["mode"])
        return
    if args.tune_tiles:
        tune_tiles(o, v, args.tune_tiles, args.sigma_b)
        return
    if args.probe_batched:
        print(json.dumps(batched_ensemble(args.probe_batched, steps=args.steps, solves=False, graph=False)))
        return
    if args.probe_q:
        print(json.dumps(q_time(o, v, blksize=args.probe_q)))
        return

    t_start = time.perf_counter()
    print("bench: building the synthetic system nocc=%d nvir=%d (B-factor sigma %g)" % (o, v, args.sigma_b),
          file=sys.stderr, flush=True)
    inp = _Inputs(o, v, 20261202, args.sigma_b)
    runs, tile_src = mixed_vs_fp64(inp, args.steps, args.warmup, args.nvir_diis, args.fp64_iters,
                                   args.fp32_iters, args.tiles_mode)
    head = runs["bf16x3"]
    ms = float(np.mean(head["ms"]))
    e_mixed = head["energies"]
    mixed = {"bf16x3_ms_per_iteration": ms, "bf16x3_ms_min_max": [min(head["ms"]), max(head["ms"])],
             "bf16x3_e_corr": e_mixed, "bf16x3_res_norms": head["res_norms"],
             "bf16x3_converged_iteration_dE_1e-6": _converged_at(e_mixed, 1e-6),
             "bf16x3_phase_ms": head["phase_ms"], "tiles": tile_src}
    for name in ("fp64", "fp32"):
        if name not in runs:
            continue
        r = runs[name]
        mixed[name + "_ms_per_iteration"] = float(np.mean(r["ms"]))
        mixed[name + "_e_corr"] = r["energies"]
        mixed[name + "_phase_ms"] = r["phase_ms"]
    if "fp64" in runs:
        e64 = runs["fp64"]["energies"]
        n = len(e64)
        for name in ("bf16x3", "fp32"):
            if name in runs:
                e = runs[name]["energies"]
                de = [abs(a - b) for a, b in zip(e[:n], e64)]
                mixed[name + "_abs_dE_vs_fp64_per_iteration"] = de
                mixed[name + "_max_abs_dE_vs_fp64"] = max(de)
                mixed["speedup_" + name] = mixed["fp64_ms_per_iteration"] / mixed[name + "_ms_per_iteration"]
        mixed["fp64_converged_iteration_dE_1e-6"] = _converged_at(e64, 1e-6)
    # dominant kernel of the headline: the three split-BF16x3 W_vvvv x T3 launches per iteration
    # (the GEMMs only: with the pair-symmetric products the split / unfold passes carry no flops and are
    # reported beside them)
    lw = [x for x in head["launches"] if x[1] > 0]
    pair_pass_ms = sum(x[2] for x in head["launches"] if x[1] == 0) / max(1, len(head["ms"]))
    avg_ms = float(np.mean([x[2] for x in lw]))
    flops = float(np.mean([x[1] for x in lw]))
    achieved = 3.0 * flops / (avg_ms * 1e-3) / 1e12
    if "fp64" in runs and runs["fp64"]["launches"]:
        l64 = [x for x in runs["fp64"]["launches"] if x[1] > 0]
        a64 = float(np.mean([x[2] for x in l64]))
        mixed["fp64_w_vvvv"] = {"avg_ms": a64, "tflops": float(np.mean([x[1] for x in l64])) / (a64 * 1e-3) / 1e12}
        mixed["fp64_w_vvvv"]["frac_fp64_peak"] = mixed["fp64_w_vvvv"]["tflops"] / PEAK_FP64_MFMA_TFLOPS
    traffic, traffic_src = measured_traffic("bf16x3", o, v, pairs=pair_pass_ms > 0)
    out = {
        "metric": METRIC,
        "value": ms,
        "unit": "ms",
        "n_gpus": 1,
        "steps": len(head["ms"]),
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16x3 contractions (fp32 accumulate), fp64 denominators/energy/DIIS",
        "data": "synthetic",
        "config": {"workload": "RCCSDT iteration (T1 dressing, lower order, T3->r1/r2/W, triples residual, "
                               "projection/denominator/update, unified T1/T2/T3 DIIS, energy) of the synthetic "
                               "system nocc=%d nvir=%d (%d canonical blocks of %d^3), split-BF16x3 triples "
                               "contractions" % (o, v, inp.t3.shape[0], v),
                   "nocc": o, "nvir": v, "naux": 300, "seed": 20261202, "sigma_b": args.sigma_b,
                   "nvir_diis": args.nvir_diis, "inputs_build_s": inp.build_s,
                   "mixed_vs_fp64": mixed},
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": PEAK_BF16_MFMA_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_BF16_MFMA_TFLOPS, "traffic": traffic, "traffic_source": traffic_src,
                     "kernel": "contract_bf16x3_dma_kernel / contract_bf16x3_kernel: the W_vvvv x T3 GEMMs, "
                               "as symmetric + antisymmetric products over triangular index pairs (one launch "
                               "with two batch entries per term and chunk of blocks; the two terms whose "
                               "operands are both K-contiguous staged by direct global->LDS requests, the "
                               "first register-staged)" if pair_pass_ms > 0 else
                               "contract_bf16x3_dma_kernel / contract_bf16x3_kernel (the three W_vvvv x T3 "
                               "launches: two with both operands staged by direct global->LDS requests, one "
                               "register-staged)",
                     "flops_per_launch": 3.0 * flops, "fp32_equivalent_flops_per_launch": flops,
                     "avg_launch_ms": avg_ms, "launches_per_iteration": len(lw) / max(1, len(head["ms"])),
                     "pair_split_unfold_ms_per_iteration": pair_pass_ms,
                     "note": "achieved counts the three BF16 products (hi*hi + hi*lo + lo*hi) of every FP32 "
                             "multiply-add the launch executes (pair form: 2 (nS + nA) kpad v per block, "
                             "nS = v(v+1)/2, nA = v(v-1)/2, kpad = nS rounded up to 64: about half the "
                             "2 v^5 of the plain product); the FP32-equivalent rate is achieved/3 = %.0f "
                             "TFLOP/s (%.2f x the FP32 MFMA peak)"
                             % (achieved / 3, achieved / 3 / PEAK_FP32_MFMA_TFLOPS)},
    }
    if not args.no_cpu_baseline:
        print("bench: CPU baseline (oracle, bounded sample of nocc=%d nvir=%d)" % (o, v), file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(inp)
    else:
        out["cpu_baseline"] = None
    del inp, runs, head
    gc.collect()
    if args.c2:
        print("bench: plain-FP64 residual nocc=10 nvir=90", file=sys.stderr, flush=True)
        out["config"]["plain_fp64_c2"] = plain_fp64_residual(10, 90, args.c2_steps, 2, breakdown=args.breakdown)
    if args.c1:
        print("bench: H2O/cc-pVDZ RCCSDT + (Q) end to end", file=sys.stderr, flush=True)
        out["config"]["c1_end_to_end"] = c1_end_to_end(cpu=not args.no_cpu_baseline)
    if args.batched > 0:
        print("bench: batched ensemble of %d systems" % args.batched, file=sys.stderr, flush=True)
        out["config"]["batched"] = batched_ensemble(args.batched)
    if args.adjoint:
        print("bench: adjoint cost", file=sys.stderr, flush=True)
        out["config"]["adjoint"] = adjoint_cost(10, 90)
    if args.q_nvir > 0:
        print("bench: (Q) timing", file=sys.stderr, flush=True)
        out["config"]["q"] = q_time(args.q_nocc, args.q_nvir)
    out["config"]["bench_wall_s"] = time.perf_counter() - t_start
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
