"""Writes tests/golden/dc_water2000.npz: the oracle's divide-and-conquer result on the
2000-water cluster (20x10x10 grid, seed 3) with one-molecule cores, a 4.0 A buffer and
kT = 1000 K, converged tightly (e_tol 1e-9 eV, p_tol 1e-10) so that the result no longer
depends on where the iteration stops (the default p_tol is an rms over all nao^2 = 1.44e8
entries, loose enough at this size to leave 1e-5 eV between two runs that stop one iteration
apart). The values come from the CPU restatement (oracle/), not from the reference; the GPU
test compares the device path against them at full size.

Run from the repository root: OMP_NUM_THREADS=8 python tests/golden/make_dc_water2000.py
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

KT_1000K = 0.08617333262
BUFFER_BOHR = 4.0 / 0.529167
E_TOL, P_TOL, MAX_SCF = 1e-9, 1e-10, 300


def main():
    threads = int(os.environ.get("OMP_NUM_THREADS", "8"))
    O.set_lapack(True, 1)  # one BLAS thread per OpenMP worker
    mol = G.water_cluster_config(3)
    t = time.time()
    r = O.dc(mol.z, mol.xyz_bohr, core_of_atom=np.arange(len(mol.z)) // 3, buffer_radius=BUFFER_BOHR,
             kt=KT_1000K, e_tol=E_TOL, p_tol=P_TOL, max_scf=MAX_SCF)
    dt = time.time() - t
    print({k: r[k] for k in ("status", "iterations", "total_ev", "fermi_ev", "entropy_ev", "n_subsystems",
                             "largest_subsystem_aos")}, f"{dt:.1f} s", flush=True)
    assert r["status"] == 0, r["message"]
    np.savez_compressed(os.path.join(os.path.dirname(os.path.abspath(__file__)), "dc_water2000.npz"),
                        total_ev=r["total_ev"], electronic_ev=r["electronic_ev"], core_ev=r["core_ev"],
                        heat_of_formation_kcal=r["heat_of_formation_kcal"], fermi_ev=r["fermi_ev"],
                        entropy_ev=r["entropy_ev"], iterations=r["iterations"], n_subsystems=r["n_subsystems"],
                        largest_subsystem_aos=r["largest_subsystem_aos"], charges=r["charges"],
                        oracle_seconds=dt, threads=threads)


if __name__ == "__main__":
    main()
