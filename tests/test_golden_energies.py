"""End-to-end energies against the reference's own known values.

The molecules, basis sets, thresholds and expected numbers are those of the reference's tests:
  * RCCSDT, H2O/6-31G, all electrons: -0.1364767434621007, tolerance 5e-8 (tests/test_rccsdt.py:26-27,51-56);
  * [Q] and (Q), H2O/cc-pVDZ at an asymmetric geometry: -0.00044374834015582527 and
    -0.0004917163848923114 to 10 decimals, for blksize 2, 3, 4, 7 and for jobs split 3, 9 and 13 ways
    (tests/test_rccsdt_q.py:20-21,27-34,55-94).
Integrals and RHF come from distr_cc.molecule; the CPU tests run the device program through the
NumPy emulator of the C ABI, the GPU tests run it on the MI355X.  The C1 case (H2O/cc-pVDZ,
symmetric geometry) has no reference value; the GPU result is checked against the oracle's FP64 CPU
energies in tests/golden/c1_h2o_ccpvdz.json (written by tests/golden/make_c1_energies.py) to 1e-9.
"""
import json
import os

import numpy as np
import pytest

import abi_emulator

HERE = os.path.dirname(os.path.abspath(__file__))
E_RCCSDT_631G = -0.1364767434621007
E_BRACKET_Q = -0.00044374834015582527
E_PAREN_Q = -0.0004917163848923114


def _h2o(basis, geometry):
    from distr_cc.molecule import Molecule, RHF
    h1 = {"sym": (0.587, 0.587), "asym": (0.487, 0.687)}[geometry]
    mol = Molecule([("O", (0., 0., 0.)), ("H", (0., -0.757, h1[0])), ("H", (0., 0.757, h1[1]))], basis)
    mf = RHF(mol, conv_tol=1e-14)
    mf.kernel()
    return mf


def _rccsdt(mf, conv_tol):
    from distr_cc.distribute_t3 import SINGLE_RANK
    from distr_cc.rccsdt import RCCSDT
    cc = RCCSDT(mf, SINGLE_RANK)
    cc.conv_tol, cc.verbose = conv_tol, 0
    cc.kernel()
    assert cc.converged
    return cc


def test_h2o_631g_rccsdt_emulated(monkeypatch):
    abi_emulator.install(monkeypatch)
    cc = _rccsdt(_h2o("6-31g", "sym"), 1e-8)
    assert abs(cc.e_corr - E_RCCSDT_631G) < 5e-8


def test_h2o_ccpvdz_q_emulated(monkeypatch):
    abi_emulator.install(monkeypatch)
    from distr_cc import rccsdt_q
    from distr_cc.distribute_t3 import SINGLE_RANK
    cc = _rccsdt(_h2o("cc-pvdz", "asym"), 1e-10)
    eb, ep = rccsdt_q.kernel(cc, comm=SINGLE_RANK)
    assert abs(eb - E_BRACKET_Q) < 1e-10 and abs(ep - E_PAREN_Q) < 1e-10


@pytest.mark.gpu
def test_h2o_631g_rccsdt_gpu():
    from distr_cc import device
    device.init(0)
    cc = _rccsdt(_h2o("6-31g", "sym"), 1e-8)
    print("E_corr 6-31G %.15f" % cc.e_corr, flush=True)
    assert abs(cc.e_corr - E_RCCSDT_631G) < 5e-8


@pytest.mark.gpu
def test_h2o_ccpvdz_q_gpu():
    from distr_cc import device, rccsdt_q
    from distr_cc.distribute_t3 import SINGLE_RANK
    device.init(0)
    cc = _rccsdt(_h2o("cc-pvdz", "asym"), 1e-10)
    for bs in (2, 3, 4, 7):
        eb, ep = rccsdt_q.kernel(cc, blksize=bs, comm=SINGLE_RANK)
        print("blksize %d [Q] %.17f (Q) %.17f" % (bs, eb, ep), flush=True)
        assert abs(eb - E_BRACKET_Q) < 1e-10 and abs(ep - E_PAREN_Q) < 1e-10
    for n_jobs, bs in ((3, 4), (9, 6), (13, 3)):
        parts = [rccsdt_q.kernel(cc, blksize=bs, comm=SINGLE_RANK, job_idx=i, n_jobs=n_jobs)
                 for i in range(n_jobs)]
        eb, ep = (sum(x[0] for x in parts), sum(x[1] for x in parts))
        assert abs(eb - E_BRACKET_Q) < 1e-10 and abs(ep - E_PAREN_Q) < 1e-10


@pytest.mark.gpu
def test_c1_matches_cpu_oracle_gpu():
    """H2O/cc-pVDZ (the reference's CPU-runnable configuration): GPU FP64 == CPU FP64 to 1e-9 Eh."""
    from distr_cc import device, rccsdt_q
    from distr_cc.distribute_t3 import SINGLE_RANK
    device.init(0)
    ref = json.load(open(os.path.join(HERE, "golden", "c1_h2o_ccpvdz.json")))
    mf = _h2o("cc-pvdz", "sym")
    assert abs(mf.e_tot - ref["e_hf"]) < 1e-9
    cc = _rccsdt(mf, 1e-11)
    eb, ep = rccsdt_q.kernel(cc, comm=SINGLE_RANK)
    print("C1 E_corr %.15f [Q] %.15f (Q) %.15f" % (cc.e_corr, eb, ep), flush=True)
    assert abs(cc.e_corr - ref["e_corr_rccsdt"]) < 1e-9
    assert abs(eb - ref["e_bracket_q"]) < 1e-9 and abs(ep - ref["e_paren_q"]) < 1e-9
