"""Writes tests/golden/planewave_L4_A28.json: the oracle's per-step correlation energies for
plane-wave L=4, A=28 (256 orbitals, h=28, p=228), SU(4) contact C0=-4.0, three-body contact
D0=+1.0 normal-ordered at the plane-wave reference, legacy damped Jacobi (mixing 0.5, no DIIS)
the first 20 steps (CPU, numpy FP64; about 15 minutes on 8 cores).

    python tests/golden/make_planewave_L4.py
"""
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from nulattice_amd.planewave import planewave_hamiltonian  # noqa: E402
from oracle import ccsd_ref as cc  # noqa: E402


def main():
    L, A, C0, D0 = 4, 28, -4.0, 1.0
    E0, fock, v = planewave_hamiltonian(L, A, C0, D0)
    coo = lambda blk: np.concatenate([blk[0], blk[1][:, None]], 1).tolist()
    vl = [coo(v[0]), coo(v[1])] + list(v[2:])
    hist = []
    t0 = time.time()
    out = cc.ccsd_solver(fock, vl, eps=1e-12, maxSteps=20, max_diis=0, history=hist,
                         verbose=True)
    e = out[0] if out is not None else hist[-1]
    res = {"L": L, "A": A, "C0": C0, "D0": D0, "E_ref": float(E0), "E_corr": float(e),
           "energies": [float(x) for x in hist], "steps": len(hist) - 1, "mixing": 0.5,
           "max_diis": 0,
           "seconds": time.time() - t0,
           "generator": "oracle.ccsd_ref.ccsd_solver on nulattice_amd.planewave blocks"}
    with open(os.path.join(HERE, "planewave_L4_A28.json"), "w") as f:
        json.dump(res, f, indent=1)
    print(f"E_corr={e!r} after {len(hist) - 1} steps, {res['seconds']:.0f} s")


if __name__ == "__main__":
    main()
