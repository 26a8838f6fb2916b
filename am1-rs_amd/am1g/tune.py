"""Autotuner for the device-dependent launch choices of libam1g (``am1g_set_tunable``).

Measures on the current GPU, through the C ABI, and writes the table the library reads at
``am1g_create`` (``tuned.txt`` beside ``libam1g.so``, or the file ``AM1G_TUNE_FILE`` names):

* ``route_large_ao``: the size from which one molecule runs faster on the grid-level kernels
  (the whole device) than in one workgroup of the batched kernel. Water clusters of 16 to 42
  molecules (96 to 252 basis functions) are timed both ways; the crossover is the smallest size
  from which the grid path wins at every larger size measured.
* ``tri_max``: the largest order the batched kernel diagonalizes by the tridiagonal method
  (above 128 with the matrix in global memory), Jacobi beyond it; 128 and 256 are timed on a
  batch of clusters of 144 to 192 basis functions.
* ``jacobi_lds_n``: whether the batched Jacobi solver (the fallback, and the solver above
  ``tri_max``) keeps its matrix in LDS up to 192 (the LDS capacity) or works in global memory;
  timed on the same batch with ``tri_max`` 128.
* ``eig_graphs``: whether replaying the large eigensolver's tridiagonalization as a HIP graph
  pays; timed on a 1200-order eigensolve.
* ``tri_blocked_min``: the order from which the large eigensolver's tridiagonalization runs in
  panels (two grid kernels per column reading the trailing matrix once, a rank-2k matrix-core
  update per panel) instead of one column at a time (one grid pass that reads and writes the
  trailing matrix, then a one-workgroup reflector step); orders 600, 1200 and 3072 are timed
  both ways, and the blocked form is kept from the smallest order from which it wins at every
  larger order measured (otherwise off).
* ``batch_own_pairs``: whether a batched launch computes each molecule's pair integrals in the
  molecule's own workgroup (one launch) or runs the pair-integral kernel over the whole batch
  first; timed on 1024 chains of the benchmark workload.
* ``pairs_block``: the workgroup size of the pair-integral kernel (64, 128 or 256 threads);
  timed by the core-Hamiltonian build of 2000 waters (18 M atom pairs), which that kernel
  dominates.

Usage: ``python -m am1g.tune [path]``. Each choice is a size- or machine-keyed constant; the
measured times are written as comments beside it.
"""
from __future__ import annotations

import os
import sys
import time

import numpy as np

import am1g
from am1g import geometry as G

_TRI_MAX = 128  # largest order the batched kernel diagonalizes by the tridiagonal method


def _time_scf(eng, mol, reps=3):
    best = float("inf")
    for _ in range(reps):
        r = eng.scf(mol.z, mol.xyz_bohr, full=False)
        if r["status"] != 0:
            raise RuntimeError(r.get("message", r["status_name"]))
        best = min(best, r["latency_ms"])
    return best


def tune_route(eng, log):
    sizes = list(range(16, 25)) + list(range(26, 43, 2))
    rows = []
    for k in sizes:
        mol = G.small_water_cluster(k, np.random.default_rng(100 + k))
        eng.set_tunable("route_large_ao", 100000)  # one workgroup
        t_wg = _time_scf(eng, mol)
        eng.set_tunable("route_large_ao", 0)  # grid-level kernels
        t_grid = _time_scf(eng, mol)
        rows.append((mol.nao, t_wg, t_grid))
        log(f"route: {mol.nao} AOs: workgroup {t_wg:.2f} ms, grid {t_grid:.2f} ms")
    route = rows[-1][0]
    for i in range(len(rows)):
        if all(t_grid < t_wg for _, t_wg, t_grid in rows[i:]):
            # The workgroup path changes solver above _TRI_MAX functions (the matrix in
            # registers up to it, in global memory or Jacobi above), which is where its time
            # jumps: a crossover found between two measured sizes that bracket that switch is
            # placed at the switch itself.
            lo = rows[i - 1][0] if i > 0 else rows[i][0] - 1
            route = _TRI_MAX if lo <= _TRI_MAX < rows[i][0] else lo
            break
    eng.set_tunable("route_large_ao", route)
    return route, rows


def _cluster_batch():
    mols = [G.small_water_cluster(k, np.random.default_rng(200 + k)) for k in (24, 26, 28, 30, 32) for _ in range(8)]
    return [(m.z, m.xyz_bohr) for m in mols]


def _batch_ms(eng, batch):
    eng.scf_batch(batch)
    ts = []
    for _ in range(2):
        eng.scf_batch(batch)
        ts.append(eng.last_kernel_ms())
    return min(ts)


def tune_jacobi_lds(eng, log):
    batch = _cluster_batch()
    times = {}
    old_tri = eng.tunables()["tri_max"]
    eng.set_tunable("tri_max", 128)
    for lds_n in (0, 192):
        eng.set_tunable("jacobi_lds_n", lds_n)
        times[lds_n] = _batch_ms(eng, batch)
        log(f"jacobi_lds_n {lds_n}: batch of {len(batch)} clusters {times[lds_n]:.1f} ms")
    eng.set_tunable("tri_max", old_tri)
    best = min(times, key=times.get)
    eng.set_tunable("jacobi_lds_n", best)
    return best, times


def tune_tri_max(eng, log):
    batch = _cluster_batch()
    times = {}
    for tm in (128, 256):
        eng.set_tunable("tri_max", tm)
        times[tm] = _batch_ms(eng, batch)
        log(f"tri_max {tm}: batch of {len(batch)} clusters {times[tm]:.1f} ms")
    best = min(times, key=times.get)
    eng.set_tunable("tri_max", best)
    return best, times


def tune_graphs(eng, log):
    n = 1200
    rng = np.random.default_rng(5)
    a = rng.standard_normal((n, n))
    a = a + a.T
    times = {}
    for g in (0, 1):
        eng.set_tunable("eig_graphs", g)
        eng.eigh(a)
        ts = []
        for _ in range(3):
            t = time.perf_counter()
            eng.eigh(a)
            ts.append((time.perf_counter() - t) * 1e3)
        times[g] = min(ts)
        log(f"eig_graphs {g}: {n}-order eigensolve {times[g]:.1f} ms")
    best = 1 if times[1] < 0.98 * times[0] else 0
    eng.set_tunable("eig_graphs", best)
    return best, times


def tune_own_pairs(eng, log):
    ms = [(m.z, m.xyz_bohr) for m in G.chain_batch(1024, seed=1)]
    times = {}
    for own in (0, 1):
        eng.set_tunable("batch_own_pairs", own)
        eng.scf_batch(ms)
        ts = []
        for _ in range(3):
            eng.scf_batch(ms)
            ts.append(eng.last_kernel_ms())
        times[own] = min(ts)
        log(f"batch_own_pairs {own}: 1024 chains {times[own]:.2f} ms")
    best = min(times, key=times.get)
    eng.set_tunable("batch_own_pairs", best)
    return best, times


_BLOCKED_OFF = 1 << 30


def tune_blocked(eng, log):
    rng = np.random.default_rng(6)
    rows = []
    for n in (600, 1200, 3072):
        a = rng.standard_normal((n, n))
        a = a + a.T
        t = {}
        for mode, bmin in (("column", _BLOCKED_OFF), ("blocked", 8)):
            eng.set_tunable("tri_blocked_min", bmin)
            eng.eigh(a)
            ts = []
            for _ in range(3):
                eng.eigh(a)
                ts.append(eng.last_kernel_ms())
            t[mode] = min(ts)
        rows.append((n, t["column"], t["blocked"]))
        log(f"tri_blocked_min: {n}-order eigensolve: by column {t['column']:.1f} ms, blocked {t['blocked']:.1f} ms")
    best = _BLOCKED_OFF
    for n, tc, tb in reversed(rows):
        if tb < tc:
            best = n
        else:
            break
    eng.set_tunable("tri_blocked_min", best)
    return best, rows


def _device_name(device: int) -> str:
    try:
        import torch

        props = torch.cuda.get_device_properties(device)
        return f"{props.name} ({props.gcnArchName})"
    except Exception as exc:  # noqa: BLE001 - the name is informational only
        return f"device {device} (name unavailable: {type(exc).__name__})"


def tune_pairs_block(eng, log):
    m = G.water_cluster_config(3)
    times = {}
    for b in (64, 128, 256):
        eng.set_tunable("pairs_block", b)
        eng.core_hamiltonian(m.z, m.xyz_bohr, with_h=False, with_pairs=False)
        ts = []
        for _ in range(3):
            t = time.perf_counter()
            eng.core_hamiltonian(m.z, m.xyz_bohr, with_h=False, with_pairs=False)
            ts.append((time.perf_counter() - t) * 1e3)
        times[b] = min(ts)
        log(f"pairs_block {b}: core Hamiltonian of 2000 waters {times[b]:.1f} ms")
    best = min(times, key=times.get)
    eng.set_tunable("pairs_block", best)
    return best, times


def autotune(path: str | None = None, device: int = 0, log=print) -> dict:
    """Measure every tunable on `device` and write the table to `path` (default: tuned.txt
    beside the library). Returns the chosen values."""
    path = path or os.path.join(os.path.dirname(am1g.LIB_PATH), "tuned.txt")
    os.environ["AM1G_TUNE_FILE"] = ""  # measure from the defaults, not from an earlier table
    name = _device_name(device)  # before the library touches the device (torch's query fails after)
    eng = am1g.Engine(device)
    try:
        tri, tt = tune_tri_max(eng, log)
        route, rows = tune_route(eng, log)
        lds_n, jt = tune_jacobi_lds(eng, log)
        graphs, gt = tune_graphs(eng, log)
        pblock, pt = tune_pairs_block(eng, log)
        bmin, bt = tune_blocked(eng, log)
        own, ot = tune_own_pairs(eng, log)
    finally:
        eng.close()
    lines = [f"# libam1g tunables measured by am1g.tune on {name}, {time.strftime('%Y-%m-%d')}",
             "# route_large_ao: one molecule, best of 3 (ms): AOs workgroup grid"]
    lines += [f"#   {nao} {a:.2f} {b:.2f}" for nao, a, b in rows]
    lines.append(f"route_large_ao {route}")
    lines.append("# tri_max: batch of 40 clusters of 144-192 AOs (ms): " + ", ".join(f"{k}: {v:.1f}" for k, v in tt.items()))
    lines.append(f"tri_max {tri}")
    lines.append("# jacobi_lds_n: batched Jacobi (tri_max 128) on 40 clusters of 144-192 AOs (ms): " +
                 ", ".join(f"{k}: {v:.1f}" for k, v in jt.items()))
    lines.append(f"jacobi_lds_n {lds_n}")
    lines.append("# eig_graphs: 1200-order eigensolve (ms): " + ", ".join(f"{k}: {v:.1f}" for k, v in gt.items()))
    lines.append(f"eig_graphs {graphs}")
    lines.append("# pairs_block: core Hamiltonian of 2000 waters (ms): " + ", ".join(f"{k}: {v:.1f}" for k, v in pt.items()))
    lines.append(f"pairs_block {pblock}")
    lines.append("# tri_blocked_min: large eigensolve (ms): order by-column blocked")
    lines += [f"#   {n} {a:.1f} {b:.1f}" for n, a, b in bt]
    lines.append(f"tri_blocked_min {bmin}")
    lines.append("# batch_own_pairs: 1024 chains (ms): " + ", ".join(f"{k}: {v:.2f}" for k, v in ot.items()))
    lines.append(f"batch_own_pairs {own}")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
    return {"route_large_ao": route, "tri_max": tri, "jacobi_lds_n": lds_n, "eig_graphs": graphs, "pairs_block": pblock,
            "tri_blocked_min": bmin, "batch_own_pairs": own, "path": path}


if __name__ == "__main__":
    out = autotune(sys.argv[1] if len(sys.argv) > 1 else None)
    print(out)
