"""Write tests/golden/c1_h2o_ccpvdz.json: the oracle's (CPU, FP64) RCCSDT and [Q]/(Q) energies for
H2O, O (0,0,0), H (0,+-0.757,0.587) Angstrom, cc-pVDZ, RHF, all electrons (nocc=5, nvir=19).

The integrals and RHF come from distr_cc.molecule; the amplitude iteration is oracle.rccsdt_lower.update_amps
with the oracle DIIS (reference defaults: space 6, start at cycle 0), converged to |dE| < 1e-12 and
|res| < 1e-10; (Q) is oracle.q_energy with blksize 4.  Run from the repository root:

    python tests/golden/make_c1_energies.py
"""
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path[:0] = [ROOT, os.path.join(ROOT, "distributed-cc_amd")]

from distr_cc.molecule import Molecule, RHF  # noqa: E402
from oracle import q_energy, rccsdt_lower as lo, rccsdt_tri as tri  # noqa: E402
from oracle.diis import DIIS  # noqa: E402
from oracle.t3_layout import full_t3_from_tri  # noqa: E402


def main():
    mol = Molecule([("O", (0., 0., 0.)), ("H", (0., -0.757, 0.587)), ("H", (0., 0.757, 0.587))], "cc-pvdz")
    mf = RHF(mol, conv_tol=1e-14)
    e_hf = mf.kernel()
    C = mf.mo_coeff
    o = mol.nelectron // 2
    pppp = np.einsum("pqrs,pP,qQ,rR,sS->PRQS", mf._eri, C, C, C, C, optimize=True)
    fock = np.diag(mf.mo_energy)
    e = mf.mo_energy
    v = len(e) - o
    _, t1, t2 = tri.init_amps_rhf(o, fock, e, pppp)
    t3 = np.zeros((o * (o + 1) * (o + 2) // 6, v, v, v))
    d = DIIS(6)
    e_old = 0.0
    e_cc = tri.energy_rhf(o, fock, pppp, t1, t2)
    for it in range(200):
        t1, t2, t3, res = lo.update_amps(o, fock, pppp, e, t1, t2, t3)
        sizes = [t1.size, t2.size]
        x = d.update(np.concatenate([t1.ravel(), t2.ravel(), t3.ravel()]))
        t1 = x[:sizes[0]].reshape(t1.shape)
        t2 = x[sizes[0]:sizes[0] + sizes[1]].reshape(t2.shape)
        t3 = x[sizes[0] + sizes[1]:].reshape(t3.shape)
        e_old, e_cc = e_cc, tri.energy_rhf(o, fock, pppp, t1, t2)
        print("cycle %2d  E_corr %.14f  dE %.2e  res %.2e" % (it + 1, e_cc, e_cc - e_old, np.linalg.norm(res)),
              flush=True)
        if abs(e_cc - e_old) < 1e-12 and np.linalg.norm(res) < 1e-10:
            break
    eb, ep = q_energy.kernel(o, v, pppp, t2, full_t3_from_tri(t3, o), e, blksize=4)
    out = {"system": "H2O O(0,0,0) H(0,+-0.757,0.587) A, cc-pVDZ (spherical), RHF, all electrons",
           "nocc": o, "nvir": v, "e_hf": e_hf, "e_corr_rccsdt": e_cc, "e_bracket_q": eb, "e_paren_q": ep,
           "made_by": "tests/golden/make_c1_energies.py (oracle update_amps + oracle DIIS, oracle (Q))"}
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "c1_h2o_ccpvdz.json")
    json.dump(out, open(path, "w"), indent=1)
    print(json.dumps(out, indent=1))


if __name__ == "__main__":
    main()
