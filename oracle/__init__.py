"""CPU oracle for the RCCSDT triples path -- TEST INFRASTRUCTURE ONLY.

This package restates, in plain FP64 NumPy, the algorithm of the reference
Distributed-CC triples update (``distr_cc/rccsdt.py``, ``src/rccsdt.c``,
``distr_cc/rccsdt_q.py``, ``src/rccsdt_q.c``).  Every function cites the
reference file:line it follows.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import anything from here, and only as the checker / the
timed CPU baseline.  The product path (``distributed-cc_amd/distr_cc``) never
imports this package and raises when its HIP library is missing.

Modules
-------
t3_layout   canonical i<=j<=k block indexing and the six permutation cases
spin_sum    P3_* spin summations, triple_sym, transpose_add (closed forms)
rccsdt_tri  block-loop restatement of the T3 residual, T3->r1/r2, T3->W
rccsdt_dense independent full-tensor restatement used to cross-check the
            block loop (different derivation, same mathematics)
q_energy    restatement of the (Q)/[Q] task loop
refc        ctypes binding of the reference C helpers compiled from
            /root/reference/src into oracle/_ref (cross-check only)

Pinning status of each function is listed in the module docstrings; a
function marked "parity unpinned" has no independent check.
"""
