#!/usr/bin/env python3
"""Benchmark: converged AM1 SCFs/s on one MI355X, batched persistent launch vs the plain path.

Workload: 4096 branched saturated chains C_nH_(2n+2), n ~ U{3..20}, with OH/NH2 terminal
substitutions (seed 1), neutral closed-shell AM1 RHF in FP64, SAD guess, A-DIIS -> CDIIS,
e_tol 1e-8 eV. One step = one batched launch over all 4096 molecules. An SCF counts as
converged only when it converged and its total energy is within 1e-6 eV of the FP64 CPU
oracle on the same geometry.

Prints one JSON line (the last line of stdout). Progress goes to stderr.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (ROOT, os.path.join(ROOT, "am1-rs_amd")):
    if p not in sys.path:
        sys.path.insert(0, p)

FP64_PEAK_TFLOPS = 78.6  # MI355X dense FP64 (vector and matrix)
METRIC = "converged AM1 SCFs/s (ΔE≤1e-6 eV) batched vs plain, 1 GPU; p50/p99 ms"


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def scf_flops(nao, n_heavy, n_h, nocc):
    """Algorithmic FP64 flops of one SCF iteration: Fock pair contractions, eigensolve
    (10/3 n^3), commutator (2 n^3), density (2 n^2 n_occ), DIIS (16 n^2)."""
    hh = n_heavy * (n_heavy - 1) / 2
    hH = n_heavy * n_h
    HH = n_h * (n_h - 1) / 2
    fock = 2.0 * (768 * hh + 48 * hH + 3 * HH)
    n = float(nao)
    return fock + (10.0 / 3.0) * n ** 3 + 2.0 * n ** 3 + 2.0 * n * n * nocc + 16.0 * n * n


def pmc_traffic():
    """HBM bytes per batched launch from the committed FETCH_SIZE / WRITE_SIZE profile of the same
    workload (counters scaled by a copy-kernel calibration), or None when it is absent."""
    path = os.path.join(ROOT, "profiles", "k_scf_batch_chains4096_pmc.md")
    try:
        for line in open(path):
            if line.startswith("traffic_bytes_per_launch:"):
                return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def percentile(x, q):
    return float(np.percentile(np.asarray(x, dtype=float), q)) if len(x) else None


def serve_stream(am1g, stream, rate, lanes, max_batch, precision):
    """Submit the requests at their Poisson arrival times to a continuous-batching server and
    wait for all of them: throughput over the whole run and per-request latency percentiles."""
    sv = am1g.Server(0, lanes=lanes, max_batch=max_batch)
    tickets = []
    t0 = time.perf_counter()
    for req in stream:
        wait = req.arrival_s - (time.perf_counter() - t0)
        if wait > 0:
            time.sleep(wait)
        tickets.append(sv.submit(req.molecule.z, req.molecule.xyz_bohr, method=req.method,
                                 multiplicity=req.molecule.multiplicity, precision=precision))
    res = [sv.poll(t, wait_ms=-1) for t in tickets]
    elapsed = time.perf_counter() - t0
    stats = sv.stats()
    sv.close()
    lat = [r["latency_ms"] for r in res]
    ok = sum(1 for r in res if r["status"] == 0)
    return {"offered_per_s": rate, "requests": len(stream), "ok": ok, "scfs_per_s": round(ok / elapsed, 2),
            "p50_ms": percentile(lat, 50), "p95_ms": percentile(lat, 95), "p99_ms": percentile(lat, 99),
            "lanes": lanes, "max_batch": max_batch, "precision": precision, "launches": stats["launches"],
            "uhf": sum(1 for r in res if r["unrestricted"]),
            "rm1": sum(1 for q in stream if q.method == "rm1"),
            "bcc_requested": sum(1 for q in stream if q.bcc)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--molecules", type=int, default=4096)
    ap.add_argument("--plain-sample", type=int, default=64)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-mixed", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-large", action="store_true")
    ap.add_argument("--serve-requests", type=int, default=1000)
    ap.add_argument("--serve-rate", type=float, default=2000.0)
    ap.add_argument("--copy-probe", action="store_true",
                    help="only copy 1 GiB device to device (calibrates the FETCH_SIZE / WRITE_SIZE counters)")
    ap.add_argument("--dc-probe", action="store_true",
                    help="only the 2000-water divide-and-conquer SCF of the large section (for counter runs)")
    args = ap.parse_args()
    if args.copy_probe:
        import torch
        n = 256 * 1024 * 1024  # 2^28 float32 = 1 GiB
        x = torch.ones(n, device="cuda")
        y = torch.empty_like(x)
        torch.cuda.synchronize()
        y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copied_bytes": n * 4}), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (every box in this setup has one); --gpus {args.gpus} "
              "is not supported. Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import am1g
    from am1g import geometry as G

    if args.dc_probe:
        eng = am1g.Engine(0)
        w2000 = G.water_cluster_config(3)
        rd = eng.dc(w2000.z, w2000.xyz_bohr, core_of_atom=np.arange(len(w2000.z)) // 3,
                    buffer_radius=4.0 / 0.529167, kt=1000.0 * 8.617333262e-5)
        print(json.dumps({"dc_status": rd["status"], "iterations": rd["iterations"],
                          "wall_ms": round(rd["wall_ms"], 1)}), flush=True)
        return

    t0 = time.time()
    mols = G.chain_batch(args.molecules, seed=1)
    log(f"generated {len(mols)} molecules in {time.time() - t0:.1f}s "
        f"(nao {min(m.nao for m in mols)}-{max(m.nao for m in mols)}, mean {np.mean([m.nao for m in mols]):.1f})")
    ms = [(m.z, m.xyz_bohr) for m in mols]
    eng = am1g.Engine(0)

    # ---------------------------------------------------------------- batched FP64
    for w in range(args.warmup):
        eng.scf_batch(ms)
        log(f"warmup {w + 1}/{args.warmup}: kernel {eng.last_kernel_ms():.1f} ms")
    walls, calls, kernels = [], [], []
    res = None
    for s in range(args.steps):
        t = time.perf_counter()
        res = eng.scf_batch(ms)
        walls.append(time.perf_counter() - t)
        calls.append(eng.last_call_ms)
        kernels.append(eng.last_kernel_ms())
        log(f"step {s + 1}/{args.steps}: kernel {kernels[-1]:.1f} ms, wall {walls[-1] * 1e3:.1f} ms")
    kernel_s = float(np.mean(kernels)) / 1e3
    wall_s = float(np.mean(walls))
    lat = [r["latency_ms"] for r in res]

    # ---------------------------------------------------------------- CPU oracle: baseline + verification
    cpu = None
    energies_ref = None
    if not args.no_cpu:
        from oracle import oracle as O

        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        O.set_lapack(True, threads=1)
        t = time.perf_counter()
        ob = O.scf_many(ms, threads=threads)
        cpu_s = time.perf_counter() - t
        energies_ref = np.array([o["total_ev"] for o in ob])
        ok_ref = np.array([o["status"] == 0 for o in ob])
        cpu_model = platform.processor() or "unknown"
        try:
            for line in open("/proc/cpuinfo"):
                if line.startswith("model name"):
                    cpu_model = line.split(":", 1)[1].strip()
                    break
        except OSError:
            pass
        cpu = {"value": float(ok_ref.sum() / cpu_s), "unit": "converged SCFs/s", "cores": threads,
               "kind": "port",
               "sample": f"all {len(ms)} molecules of the workload, FP64 C++ restatement of am1-rs 0.2.3 "
                         f"(OpenMP over molecules, LAPACK dsyevd/dgemm via numpy's OpenBLAS: "
                         f"{'yes' if O.lapack_active() else 'no'}), {cpu_s:.1f} s on {cpu_model}"}
        log(f"cpu oracle: {len(ms)} SCFs in {cpu_s:.1f}s on {threads} threads")

    energies = np.array([r["total_ev"] for r in res])
    conv = np.array([r["status"] == 0 and r["converged"] for r in res])
    if energies_ref is not None:
        de = np.abs(energies - energies_ref)
        verified = conv & (de <= 1e-6)
        max_de = float(de[conv].max()) if conv.any() else None
    else:
        verified = conv
        max_de = None
    n_ok = int(verified.sum())
    value = n_ok / kernel_s

    # roofline of the batched kernel: algorithmic flops of every executed iteration
    flops = 0.0
    for m, r in zip(mols, res):
        nh = int((m.z != 1).sum())
        nH = int((m.z == 1).sum())
        nocc = r["n_occ"]
        flops += (r["iterations"] + 1) * scf_flops(m.nao, nh, nH, nocc)  # +1: final Fock/energy pass
    achieved = flops / kernel_s / 1e12
    roofline = {"bound": "mfma", "achieved": round(achieved, 4), "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": round(achieved / FP64_PEAK_TFLOPS, 6), "traffic": pmc_traffic(),
                "kernel": "k_scf_batch", "flops_per_launch": flops}

    # ---------------------------------------------------------------- plain path on a stratified sample
    step = max(1, len(ms) // max(1, args.plain_sample))
    idx = list(range(0, len(ms), step))[: args.plain_sample]
    plain_ms = []
    plain_ok = 0
    for i in idx:
        r = eng.scf(ms[i][0], ms[i][1], full=False)
        plain_ms.append(r["latency_ms"])
        ok = r["status"] == 0 and (energies_ref is None or abs(r["total_ev"] - energies_ref[i]) <= 1e-6)
        plain_ok += int(ok)
    plain_rate = plain_ok / (sum(plain_ms) / 1e3)
    log(f"plain: {len(idx)} molecules, {plain_rate:.1f} SCFs/s, p50 {percentile(plain_ms, 50):.2f} ms")

    # ---------------------------------------------------------------- mixed precision batched
    mixed = None
    if not args.no_mixed:
        rm = eng.scf_batch(ms, precision="mixed")
        km = eng.last_kernel_ms() / 1e3
        em = np.array([r["total_ev"] for r in rm])
        convm = np.array([r["status"] == 0 and r["converged"] for r in rm])
        ref = energies_ref if energies_ref is not None else energies
        dem = np.abs(em - ref)
        okm = convm & (dem <= 1e-6)
        dq = max(float(np.abs(a["charges"] - b["charges"]).max()) for a, b in zip(rm, res))
        mixed = {"value": float(okm.sum() / km), "kernel_ms": km * 1e3, "converged_verified": int(okm.sum()),
                 "max_abs_dE_ev": float(dem[convm].max()) if convm.any() else None, "max_abs_dq_vs_fp64": dq,
                 "mean_fp32_iterations": float(np.mean([r["fp32_iterations"] for r in rm]))}
        log(f"mixed: {mixed['value']:.0f} SCFs/s")

    # ---------------------------------------------------------------- serving stream
    serving = None
    if not args.no_serve:
        slow_rate, slow_n = 20.0, 60
        fast = G.serving_stream(args.serve_requests, args.serve_rate)
        slow = G.serving_stream(slow_n, slow_rate)
        mid_rate = 200.0
        mid = G.serving_stream(args.serve_requests, mid_rate)
        serving = {
            "plain": serve_stream(am1g, slow, slow_rate, lanes=1, max_batch=1, precision="fp64"),
            "batched_mixed_same_load": serve_stream(am1g, slow, slow_rate, lanes=4, max_batch=512,
                                                    precision="mixed"),
            "batched_mixed_200": serve_stream(am1g, mid, mid_rate, lanes=4, max_batch=512, precision="mixed"),
            "batched_mixed": serve_stream(am1g, fast, args.serve_rate, lanes=4, max_batch=512, precision="mixed"),
            "note": "AM1-BCC charges are not computed (the SCF each such request needs is); "
                    "latency = submit to completion on the host clock",
        }
        log(f"serving at {mid_rate:g}/s: {serving['batched_mixed_200']['scfs_per_s']} SCFs/s "
            f"p50 {serving['batched_mixed_200']['p50_ms']:.1f} ms p99 {serving['batched_mixed_200']['p99_ms']:.1f} ms")
        log(f"serving: plain {serving['plain']['scfs_per_s']} SCFs/s p99 {serving['plain']['p99_ms']:.1f} ms; "
            f"batched mixed at {args.serve_rate:g}/s: {serving['batched_mixed']['scfs_per_s']} SCFs/s "
            f"p99 {serving['batched_mixed']['p99_ms']:.1f} ms")

    # ---------------------------------------------------------------- one large system; divide and conquer
    large = None
    if not args.no_large:
        w512 = G.water_cluster_config(2)
        large = {"workload": "512 H2O (1536 atoms, 3072 basis functions, seed 2), AM1 RHF, A-DIIS->CDIIS, "
                             "plain path, one molecule on the whole device"}
        for prec in ("fp64", "mixed"):
            r = eng.scf(w512.z, w512.xyz_bohr, full=False, precision=prec)
            large[prec] = {"status": r["status"], "iterations": r["iterations"], "fp32_iterations": r["fp32_iterations"],
                           "total_ev": r["total_ev"], "homo_ev": r["homo_ev"], "lumo_ev": r["lumo_ev"],
                           "wall_ms": round(r["latency_ms"], 1),
                           "phase_ms": {k: round(v, 1) for k, v in r["phase_ms"].items()}}
            log(f"512 waters {prec}: {r['iterations']} iterations in {r['latency_ms']:.0f} ms")
        w2000 = G.water_cluster_config(3)
        rd = eng.dc(w2000.z, w2000.xyz_bohr, core_of_atom=np.arange(len(w2000.z)) // 3,
                    buffer_radius=4.0 / 0.529167, kt=1000.0 * 8.617333262e-5)
        large["divide_conquer"] = {
            "workload": "2000 H2O (6000 atoms, 12000 basis functions, seed 3), one-molecule cores, 4.0 A buffer, "
                        "Fermi level at 1000 K, all subsystem eigensolves in one launch per iteration",
            "status": rd["status"], "iterations": rd["iterations"], "total_ev": rd["total_ev"],
            "n_subsystems": rd["n_subsystems"], "largest_subsystem_aos": rd["largest_subsystem_aos"],
            "wall_ms": round(rd["wall_ms"], 1), "phase_ms": {k: round(v, 1) for k, v in rd["phase_ms"].items()}}
        log(f"2000 waters D&C: {rd['iterations']} iterations in {rd['wall_ms']:.0f} ms")

    out = {
        "metric": METRIC,
        "value": round(value, 2),
        "unit": "converged SCFs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(kernel_s * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"{len(ms)} branched C3-C20 chains with OH/NH2 ends (seed 1), AM1 RHF FP64, "
                               "A-DIIS->CDIIS, e_tol 1e-8 eV, one persistent launch",
                   "molecules": len(ms), "nao_mean": float(np.mean([m.nao for m in mols]))},
        "batched": {"p50_ms": percentile(lat, 50), "p99_ms": percentile(lat, 99),
                    "converged_verified": n_ok, "max_abs_dE_ev": max_de,
                    "mean_iterations": float(np.mean([r["iterations"] for r in res])),
                    "wall_ms_per_step": round(wall_s * 1e3, 3),
                    "wall_inclusive_scfs_per_s": round(n_ok / wall_s, 2),
                    "c_call_ms_per_step": round(float(np.mean(calls)), 3),
                    "c_call_inclusive_scfs_per_s": round(n_ok / (float(np.mean(calls)) / 1e3), 2)},
        "plain": {"value": round(plain_rate, 2), "sample": len(idx), "p50_ms": percentile(plain_ms, 50),
                  "p99_ms": percentile(plain_ms, 99)},
        "batched_over_plain": round(value / plain_rate, 2) if plain_rate > 0 else None,
        "mixed": mixed,
        "serving": serving,
        "large_system": large,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    eng.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
