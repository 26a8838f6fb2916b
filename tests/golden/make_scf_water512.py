"""Writes tests/golden/scf_water512.npz: the oracle's plain AM1 RHF SCF of the 512-water cluster
(8x8x8 grid, seed 2; 1536 atoms, 3072 basis functions) with the default options (SAD guess,
A-DIIS -> CDIIS, e_tol 1e-8 eV, p_tol 1e-7). The values come from the CPU restatement (oracle/),
not from the reference; the GPU test compares the large-system path against them at full size.

With --tight it writes tests/golden/scf_water512_tight.npz instead: the same SCF with p_tol
1e-9, the density-change test of a mixed-precision run (100x the default), which ends within
~1e-8 e of the fixed point. Mixed-precision runs, whose trajectory differs from the FP64 one, are
checked against that.

Run from the repository root: OMP_NUM_THREADS=8 python tests/golden/make_scf_water512.py [--tight]
"""
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "am1-rs_amd"))

from am1g import geometry as G  # noqa: E402
from oracle import oracle as O  # noqa: E402


def main():
    threads = int(os.environ.get("OMP_NUM_THREADS", "8"))
    O.set_lapack(True, threads)
    mol = G.water_cluster_config(2)
    fixed = "--tight" in sys.argv[1:]
    t = time.time()
    r = O.scf(mol.z, mol.xyz_bohr, full=True, **({"p_tol": 1e-9} if fixed else {}))
    dt = time.time() - t
    print({k: r[k] for k in ("status", "iterations", "total_ev", "homo_ev", "lumo_ev")}, f"{dt:.1f} s", flush=True)
    assert r["status"] == 0, r.get("message")
    np.savez_compressed(os.path.join(os.path.dirname(os.path.abspath(__file__)), "scf_water512_tight.npz" if fixed else "scf_water512.npz"),
                        total_ev=r["total_ev"], electronic_ev=r["electronic_ev"], core_ev=r["core_ev"],
                        heat_of_formation_kcal=r["heat_of_formation_kcal"], homo_ev=r["homo_ev"],
                        lumo_ev=r["lumo_ev"], iterations=r["iterations"], charges=r["charges"],
                        mo_energies=r["mo_energies"], dipole_debye=r["dipole_debye"], oracle_seconds=dt,
                        threads=threads)


if __name__ == "__main__":
    main()
