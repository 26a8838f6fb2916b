#!/usr/bin/env python3
"""Benchmark: converged AM1 SCFs/s on one MI355X, batched persistent launch vs the plain path.

Workload: 4096 branched saturated chains C_nH_(2n+2), n ~ U{3..20}, with OH/NH2 terminal
substitutions (seed 1), neutral closed-shell AM1 RHF in FP64, SAD guess, A-DIIS -> CDIIS,
e_tol 1e-8 eV. One step = one batched launch over all 4096 molecules. An SCF counts as
converged only when it converged and its total energy (1e-6 eV), every Mulliken charge (1e-6 e)
and its HOMO/LUMO (1e-6 eV) agree with the FP64 CPU oracle on the same geometry.

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


E_TOL, Q_TOL, ORB_TOL = 1e-6, 1e-6, 1e-6  # eV, e, eV: the parity contract against the FP64 oracle


def verify(results, refs):
    """Per-SCF contract check against reference results (oracle dicts or fixture values): an SCF
    counts only when it converged and its total energy, every Mulliken charge and the HOMO/LUMO
    (both spin channels) are within tolerance. Returns (ok mask, largest errors)."""
    ok = np.zeros(len(results), dtype=bool)
    worst = {"dE_ev": 0.0, "dq_e": 0.0, "dorb_ev": 0.0, "not_converged": 0, "oracle_not_converged": 0}
    for i, (r, o) in enumerate(zip(results, refs)):
        if o["status"] != 0:
            worst["oracle_not_converged"] += 1
        if r["status"] != 0 or not r["converged"]:
            worst["not_converged"] += 1
        if r["status"] != 0 or not r["converged"] or o["status"] != 0:
            continue
        de = abs(r["total_ev"] - o["total_ev"])
        dq = float(np.abs(np.asarray(r["charges"]) - np.asarray(o["charges"])).max())
        dorb = 0.0
        good = True
        for k in ("homo_ev", "lumo_ev", "homo_beta_ev", "lumo_beta_ev"):
            a, b = r.get(k), o.get(k)
            if b is None and a is None:
                continue
            if a is None or b is None:
                good = False
                continue
            dorb = max(dorb, abs(a - b))
        worst["dE_ev"] = max(worst["dE_ev"], de)
        worst["dq_e"] = max(worst["dq_e"], dq)
        worst["dorb_ev"] = max(worst["dorb_ev"], dorb)
        ok[i] = good and de <= E_TOL and dq <= Q_TOL and dorb <= ORB_TOL
    return ok, worst


def serve_stream(am1g, stream, rate, lanes, max_batch, precision, continuous=True):
    """Submit the requests at their Poisson arrival times to a continuous-batching server and
    wait for all of them: throughput over the whole run and per-request latency percentiles.
    A request that asks for AM1-BCC charges gets them from its own SCF when it is an AM1 request;
    an RM1 request that asks for them also submits the AM1 SCF the charges need (the corrections
    are fitted to AM1 Mulliken charges), and finishes when both have."""
    sv = am1g.Server(0, lanes=lanes, max_batch=max_batch, continuous=continuous)
    tickets = []
    t0 = time.perf_counter()
    for req in stream:
        wait = req.arrival_s - (time.perf_counter() - t0)
        if wait > 0:
            time.sleep(wait)
        m = req.molecule
        main = sv.submit(m.z, m.xyz_bohr, method=req.method, multiplicity=m.multiplicity, precision=precision,
                         bcc=req.bcc and req.method == "am1")
        extra = None
        if req.bcc and req.method != "am1":
            extra = sv.submit(m.z, m.xyz_bohr, method="am1", multiplicity=m.multiplicity, precision=precision,
                              bcc=True)
        tickets.append((main, extra))
    res = []
    for main, extra in tickets:
        r = sv.poll(main, wait_ms=-1)
        if extra is not None:
            rb = sv.poll(extra, wait_ms=-1)
            r["latency_ms"] = max(r["latency_ms"], rb["latency_ms"])
            r["_bcc"] = rb
        res.append(r)
    elapsed = time.perf_counter() - t0
    stats = sv.stats()
    sv.close()
    lat = [r["latency_ms"] for r in res]
    ok = sum(1 for r in res if r["status"] == 0)
    bcc_done = sum(1 for q, r in zip(stream, res) if q.bcc and (r.get("_bcc") or r)["status"] == 0)
    bcc_clean = sum(1 for q, r in zip(stream, res) if q.bcc and (r.get("_bcc") or r)["status"] == 0
                    and (r.get("_bcc") or r)["bcc_warnings"] == 0)
    n_scf = len(stream) + sum(1 for q in stream if q.bcc and q.method != "am1")
    return {"offered_per_s": rate, "requests": len(stream), "ok": ok, "scfs_per_s": round(ok / elapsed, 2),
            "scfs_run": n_scf, "elapsed_s": elapsed, "_results": res,
            "p50_ms": percentile(lat, 50), "p95_ms": percentile(lat, 95), "p99_ms": percentile(lat, 99),
            "mode": "service launch" if continuous else f"{lanes} lanes of batched launches",
            "lanes": lanes, "max_batch": max_batch, "precision": precision, "launches": stats["launches"],
            "uhf": sum(1 for r in res if r["unrestricted"]),
            "rm1": sum(1 for q in stream if q.method == "rm1"),
            "bcc_requested": sum(1 for q in stream if q.bcc), "bcc_computed": bcc_done,
            "bcc_without_warnings": bcc_clean}


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def large_flops(nao, n_heavy, n_h, nocc):
    """Algorithmic FP64 flops of one SCF iteration of one large system (same terms as
    scf_flops)."""
    return scf_flops(nao, n_heavy, n_h, nocc)


def fock_bytes(n_heavy, n_h, nao):
    """Algorithmic HBM bytes of one Fock build: the pair blocks plus P, H read and F written."""
    hh = n_heavy * (n_heavy - 1) / 2
    hH = n_heavy * n_h
    HH = n_h * (n_h - 1) / 2
    return 800.0 * hh + 80.0 * hH + 8.0 * HH + 24.0 * nao * nao


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--molecules", type=int, default=4096)
    ap.add_argument("--plain-sample", type=int, default=64)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--cpu-single-sample", type=int, default=192,
                    help="molecules of the batch timed on one host thread")
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
    ap.add_argument("--large-probe", action="store_true",
                    help="only the 512-water FP64 SCF of the large section (for trace and counter runs)")
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
    if args.large_probe:
        eng = am1g.Engine(0)
        w512 = G.water_cluster_config(2)
        r = eng.scf(w512.z, w512.xyz_bohr, full=False)
        print(json.dumps({"status": r["status"], "iterations": r["iterations"],
                          "wall_ms": round(r["latency_ms"], 1)}), flush=True)
        return

    t0 = time.time()
    mols = G.chain_batch(args.molecules, seed=1)
    log(f"generated {len(mols)} molecules in {time.time() - t0:.1f}s "
        f"(nao {min(m.nao for m in mols)}-{max(m.nao for m in mols)}, mean {np.mean([m.nao for m in mols]):.1f})")
    ms = [(m.z, m.xyz_bohr) for m in mols]
    eng = am1g.Engine(0)

    # ---------------------------------------------------------------- batched FP64 (the timed steps)
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
    from oracle import oracle as O

    threads = args.cpu_threads or min(16, os.cpu_count() or 1)
    O.set_lapack(True, threads=1)
    cpu = None
    cpu_rows = {}
    t = time.perf_counter()
    ob = O.scf_many(ms, threads=threads)
    cpu_s = time.perf_counter() - t
    ok_ref = np.array([o["status"] == 0 for o in ob])
    lapack = "yes" if O.lapack_active() else "no"
    if not args.no_cpu:
        cpu = {"value": float(ok_ref.sum() / cpu_s), "unit": "converged SCFs/s", "cores": threads,
               "kind": "port",
               "sample": f"all {len(ms)} molecules of the workload, FP64 C++ restatement of am1-rs 0.2.3 "
                         f"(OpenMP over molecules, LAPACK dsyevd/dgemm via numpy's OpenBLAS: {lapack}), "
                         f"{cpu_s:.1f} s on {threads} threads of {cpu_model()} (host reports {os.cpu_count()} CPUs)"}
        log(f"cpu oracle: {len(ms)} SCFs in {cpu_s:.1f}s on {threads} threads")
        ns = min(args.cpu_single_sample, len(ms))
        sidx = list(range(0, len(ms), max(1, len(ms) // ns)))[:ns]
        t = time.perf_counter()
        o1 = O.scf_many([ms[i] for i in sidx], threads=1)
        s1 = time.perf_counter() - t
        cpu_rows["chains_single_thread"] = {
            "value": float(sum(o["status"] == 0 for o in o1) / s1), "unit": "converged SCFs/s", "cores": 1,
            "sample": f"{ns} molecules spread over the batch (every {len(ms) // ns}th), {s1:.1f} s"}
        log(f"cpu oracle, one thread: {ns} SCFs in {s1:.1f}s")

    # ---------------------------------------------------------------- verification of the timed result
    verified, worst = verify(res, ob)
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

    # ---------------------------------------------------------------- plain paths on a stratified sample
    step = max(1, len(ms) // max(1, args.plain_sample))
    idx = list(range(0, len(ms), step))[: args.plain_sample]
    plain = {}
    for name, force in (("workgroup", "0"), ("grid", "1")):
        os.environ["AM1G_FORCE_LARGE_PATH"] = force
        eng.scf(ms[idx[0]][0], ms[idx[0]][1], full=False)  # first call of a path sets up its buffers
        lat_p, rs_p = [], []
        for i in idx:
            r = eng.scf(ms[i][0], ms[i][1], full=False)
            lat_p.append(r["latency_ms"])
            rs_p.append(r)
        okp, _ = verify(rs_p, [ob[i] for i in idx])
        plain[name] = {"lat": np.array(lat_p), "ok": okp}
    os.environ["AM1G_FORCE_LARGE_PATH"] = "0"
    best = np.minimum(plain["workgroup"]["lat"], plain["grid"]["lat"])
    best_ok = plain["workgroup"]["ok"] & plain["grid"]["ok"]
    plain_rate = float(best_ok.sum() / (best.sum() / 1e3))
    plain_out = {"value": round(plain_rate, 2), "sample": len(idx), "verified": int(best_ok.sum()),
                 "p50_ms": percentile(best, 50), "p99_ms": percentile(best, 99),
                 "path": "per molecule the faster of the one-workgroup launch and the grid-level kernels"}
    for name in ("workgroup", "grid"):
        p = plain[name]
        plain_out[name] = {"value": round(float(p["ok"].sum() / (p["lat"].sum() / 1e3)), 2),
                           "p50_ms": percentile(p["lat"], 50), "verified": int(p["ok"].sum())}
    log(f"plain: {len(idx)} molecules, {plain_rate:.1f} SCFs/s (workgroup {plain_out['workgroup']['value']}, "
        f"grid {plain_out['grid']['value']})")

    # ---------------------------------------------------------------- mixed precision batched
    mixed = None
    if not args.no_mixed:
        rm = eng.scf_batch(ms, precision="mixed")
        km = eng.last_kernel_ms() / 1e3
        # another trajectory than FP64, ended by a density-change test 100x tighter than p_tol:
        # held to the oracle run with that same test (p_tol 1e-9)
        t = time.perf_counter()
        ofp = O.scf_many(ms, threads=threads, p_tol=1e-9)
        log(f"oracle at p_tol 1e-9: {len(ms)} SCFs in {time.perf_counter() - t:.1f}s")
        okm, wm = verify(rm, ofp)
        _, wm_default = verify(rm, ob)
        mixed = {"value": float(okm.sum() / km), "kernel_ms": km * 1e3, "converged_verified": int(okm.sum()),
                 "max_abs_dE_ev": wm["dE_ev"], "max_abs_dq_e": wm["dq_e"], "max_abs_dorb_ev": wm["dorb_ev"],
                 "checked_against": "oracle with the mixed run's density test (p_tol 1e-9)",
                 "errors_vs_default_tolerance_oracle": wm_default,
                 "mean_fp32_iterations": float(np.mean([r["fp32_iterations"] for r in rm])),
                 "mean_iterations": float(np.mean([r["iterations"] for r in rm]))}
        log(f"mixed: {mixed['value']:.0f} SCFs/s, {int(okm.sum())} verified")

    # ---------------------------------------------------------------- C0: 32 waters, plain GPU and CPU
    w32 = G.water_cluster_config(0)
    c0 = {"workload": "32 H2O (96 atoms, 192 basis functions, 4x4x2 grid, seed 0), AM1 RHF, SAD, CDIIS, "
                      "e_tol 1e-8 eV, plain path"}
    eng.scf(w32.z, w32.xyz_bohr, full=False, accelerator="cdiis")
    lat0 = []
    for _ in range(5):
        r0 = eng.scf(w32.z, w32.xyz_bohr, full=False, accelerator="cdiis")
        lat0.append(r0["latency_ms"])
    o0 = None
    for th in (threads, 1):
        O.set_lapack(True, threads=th)
        t = time.perf_counter()
        o0 = O.scf(w32.z, w32.xyz_bohr, full=False, accelerator="cdiis")
        cpu_rows[f"water32_{th}_threads"] = {"value_ms": round((time.perf_counter() - t) * 1e3, 1), "cores": th,
                                             "sample": "one SCF of the 32-water cluster (oracle, LAPACK "
                                                       f"dsyevd/dgemm on {th} threads)"}
    O.set_lapack(True, threads=1)
    ok0, w0 = verify([r0], [o0])
    c0.update(gpu_ms=percentile(lat0, 50), iterations=r0["iterations"], oracle_iterations=o0["iterations"],
              verified=bool(ok0[0]), errors=w0)
    log(f"32 waters: GPU {c0['gpu_ms']:.1f} ms, CPU {cpu_rows[f'water32_{threads}_threads']['value_ms']} ms "
        f"({threads} threads), verified {bool(ok0[0])}")

    # ---------------------------------------------------------------- serving stream
    serving = None
    if not args.no_serve:
        slow_rate, slow_n = 20.0, 60
        fast = G.serving_stream(args.serve_requests, args.serve_rate)
        slow = G.serving_stream(slow_n, slow_rate)
        mid_rate = 200.0
        mid = G.serving_stream(args.serve_requests, mid_rate)
        serving = {
            "plain": serve_stream(am1g, slow, slow_rate, lanes=1, max_batch=1, precision="fp64", continuous=False),
            "batched_mixed_same_load": serve_stream(am1g, slow, slow_rate, lanes=4, max_batch=512,
                                                    precision="mixed"),
            "batched_mixed_200": serve_stream(am1g, mid, mid_rate, lanes=4, max_batch=512, precision="mixed"),
            "batched_mixed": serve_stream(am1g, fast, args.serve_rate, lanes=4, max_batch=512, precision="mixed"),
            "batched_fp64": serve_stream(am1g, fast, args.serve_rate, lanes=4, max_batch=512, precision="fp64"),
            "lanes_mixed": serve_stream(am1g, fast, args.serve_rate, lanes=4, max_batch=512, precision="mixed",
                                        continuous=False),
            "note": "AM1-BCC requests get the corrections on the host after their AM1 SCF (an RM1 request "
                    "that asks for them also runs the AM1 SCF); latency = submit to completion on the host "
                    "clock; max_abs_dbcc_e compares the served BCC charges of AM1 requests with the "
                    "corrections applied to the oracle's Mulliken charges",
        }
        # every served request against the oracle (the streams share their molecules by position)
        longest = max((slow, mid, fast), key=len)
        t = time.perf_counter()
        sreq = [(q.molecule.z, q.molecule.xyz_bohr, {"method": q.method, "multiplicity": q.molecule.multiplicity})
                for q in longest]
        osv = {"fp64": O.scf_many(sreq, threads=threads), "mixed": O.scf_many(sreq, threads=threads, p_tol=1e-9)}
        log(f"serving oracle: 2 x {len(longest)} SCFs in {time.perf_counter() - t:.1f}s")
        for key in ("plain", "batched_mixed_same_load", "batched_mixed_200", "batched_mixed", "batched_fp64",
                    "lanes_mixed"):
            leg = serving[key]
            rs = leg.pop("_results")
            okv, wv = verify(rs, osv[leg["precision"]][: len(rs)])
            dbcc = 0.0
            for q, r, o in zip(longest, rs, osv[leg["precision"]]):
                if q.bcc and q.method == "am1" and r["status"] == 0 and o["status"] == 0:
                    ref = am1g.bcc_correct(q.molecule.z, q.molecule.xyz_bohr, o["charges"])["charges"]
                    dbcc = max(dbcc, float(np.abs(np.asarray(r["bcc_charges"]) - np.asarray(ref)).max()))
            leg["max_abs_dbcc_e"] = dbcc
            leg["verified"] = int(okv.sum())
            leg["verified_scfs_per_s"] = round(float(okv.sum()) / leg.pop("elapsed_s"), 2)
            leg["max_abs_dE_ev"] = wv["dE_ev"]
            leg["max_abs_dq_e"] = wv["dq_e"]
            leg["max_abs_dorb_ev"] = wv["dorb_ev"]
            leg["not_converged"] = wv["not_converged"]
            leg["oracle_not_converged"] = wv["oracle_not_converged"]
        log(f"serving at {mid_rate:g}/s: {serving['batched_mixed_200']['verified_scfs_per_s']} verified SCFs/s "
            f"p50 {serving['batched_mixed_200']['p50_ms']:.1f} ms p99 {serving['batched_mixed_200']['p99_ms']:.1f} ms")
        log(f"serving: plain {serving['plain']['verified_scfs_per_s']} SCFs/s p99 {serving['plain']['p99_ms']:.1f} ms; "
            f"batched mixed at {args.serve_rate:g}/s: {serving['batched_mixed']['verified_scfs_per_s']} SCFs/s "
            f"p99 {serving['batched_mixed']['p99_ms']:.1f} ms")

    # ---------------------------------------------------------------- one large system; divide and conquer
    large = None
    if not args.no_large:
        golden = os.path.join(ROOT, "tests", "golden")
        w512 = G.water_cluster_config(2)
        g512 = np.load(os.path.join(golden, "scf_water512.npz"))
        ref512 = {"status": 0, "total_ev": float(g512["total_ev"]), "charges": g512["charges"],
                  "homo_ev": float(g512["homo_ev"]), "lumo_ev": float(g512["lumo_ev"]),
                  "homo_beta_ev": None, "lumo_beta_ev": None}
        nh512 = int((w512.z != 1).sum())
        nH512 = int((w512.z == 1).sum())
        large = {"workload": "512 H2O (1536 atoms, 3072 basis functions, seed 2), AM1 RHF, A-DIIS->CDIIS, "
                             "plain path, one molecule on the whole device; FP64 checked against the oracle's "
                             "committed result (tests/golden/scf_water512.npz), mixed precision against the "
                             "oracle at p_tol 1e-9 (tests/golden/scf_water512_tight.npz)"}
        gfp = np.load(os.path.join(golden, "scf_water512_tight.npz"))
        refs512 = {"fp64": ref512,
                   "mixed": {"status": 0, "total_ev": float(gfp["total_ev"]), "charges": gfp["charges"],
                             "homo_ev": float(gfp["homo_ev"]), "lumo_ev": float(gfp["lumo_ev"]),
                             "homo_beta_ev": None, "lumo_beta_ev": None}}
        for prec in ("fp64", "mixed"):
            r = eng.scf(w512.z, w512.xyz_bohr, full=False, precision=prec)
            okl, wl = verify([r], [refs512[prec]])
            fl = (r["iterations"] + 1) * large_flops(w512.nao, nh512, nH512, r["n_occ"])
            ach = fl / (r["latency_ms"] / 1e3) / 1e12
            # the matrix-core work actually run: the SP2 squares (upper 64 x 64 tiles of n^3) and
            # the commutator F P of every iteration
            tn = (w512.nao + 63) // 64
            mc = r["eigen_sweeps"] * 2.0 * w512.nao ** 3 * (tn + 1) / (2.0 * tn) + \
                (r["iterations"] + 1) * 2.0 * w512.nao ** 3
            large[prec] = {"status": r["status"], "iterations": r["iterations"], "fp32_iterations": r["fp32_iterations"],
                           "total_ev": r["total_ev"], "homo_ev": r["homo_ev"], "lumo_ev": r["lumo_ev"],
                           "verified": bool(okl[0]), "errors": wl,
                           "wall_ms": round(r["latency_ms"], 1),
                           "sp2_steps": r["eigen_sweeps"],
                           "matrix_core": {"flops": mc, "tflops_over_wall": round(mc / (r["latency_ms"] / 1e3) / 1e12, 2),
                                           "note": "SP2 squares and commutators, FP64 (the FP32 ones of the "
                                                   "mixed run counted alike), over the SCF's wall time"},
                           "roofline": {"bound": "mfma", "achieved": round(ach, 4), "peak": FP64_PEAK_TFLOPS,
                                        "unit": "TFLOP/s", "frac": round(ach / FP64_PEAK_TFLOPS, 6),
                                        "flops": fl},
                           "phase_ms": {k: round(v, 1) for k, v in r["phase_ms"].items()}}
            log(f"512 waters {prec}: {r['iterations']} iterations in {r['latency_ms']:.0f} ms, verified {bool(okl[0])}")
        w2000 = G.water_cluster_config(3)
        g2000 = np.load(os.path.join(golden, "dc_water2000.npz"))
        rd = eng.dc(w2000.z, w2000.xyz_bohr, core_of_atom=np.arange(len(w2000.z)) // 3,
                    buffer_radius=4.0 / 0.529167, kt=1000.0 * 8.617333262e-5)
        de_dc = abs(rd["total_ev"] - float(g2000["total_ev"]))
        dq_dc = float(np.abs(rd["charges"] - g2000["charges"]).max())
        nh2 = int((w2000.z != 1).sum())
        nH2 = int((w2000.z == 1).sum())
        fb = fock_bytes(nh2, nH2, w2000.nao) * (rd["iterations"] + 1)
        fock_ms = rd["phase_ms"]["fock"]
        large["divide_conquer"] = {
            "workload": "2000 H2O (6000 atoms, 12000 basis functions, seed 3), one-molecule cores, 4.0 A buffer, "
                        "Fermi level at 1000 K, all subsystem eigensolves in one launch per iteration; checked "
                        "against the oracle's committed result (tests/golden/dc_water2000.npz)",
            "status": rd["status"], "iterations": rd["iterations"], "total_ev": rd["total_ev"],
            "verified": bool(rd["status"] == 0 and de_dc <= E_TOL and dq_dc <= Q_TOL),
            "errors": {"dE_ev": de_dc, "dq_e": dq_dc},
            "n_subsystems": rd["n_subsystems"], "largest_subsystem_aos": rd["largest_subsystem_aos"],
            "wall_ms": round(rd["wall_ms"], 1), "phase_ms": {k: round(v, 1) for k, v in rd["phase_ms"].items()},
            "fock_roofline": {"bound": "hbm", "achieved": round(fb / (fock_ms / 1e3) / 1e9, 1), "peak": 8000.0,
                              "unit": "GB/s", "frac": round(fb / (fock_ms / 1e3) / 8e12, 4),
                              "algorithmic_bytes": fb, "note": "Fock builds of all iterations over the "
                              "phase's wall time"}}
        log(f"2000 waters D&C: {rd['iterations']} iterations in {rd['wall_ms']:.0f} ms, "
            f"|dE| {de_dc:.1e} eV, |dq| {dq_dc:.1e}")

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
                    "converged_verified": n_ok, "max_abs_dE_ev": worst["dE_ev"], "max_abs_dq_e": worst["dq_e"],
                    "max_abs_dorb_ev": worst["dorb_ev"],
                    "mean_iterations": float(np.mean([r["iterations"] for r in res])),
                    "oracle_mean_iterations": float(np.mean([o["iterations"] for o in ob])),
                    "wall_ms_per_step": round(wall_s * 1e3, 3),
                    "wall_inclusive_scfs_per_s": round(n_ok / wall_s, 2),
                    "c_call_ms_per_step": round(float(np.mean(calls)), 3),
                    "c_call_inclusive_scfs_per_s": round(n_ok / (float(np.mean(calls)) / 1e3), 2)},
        "plain": plain_out,
        "batched_over_plain": round(value / plain_rate, 2) if plain_rate > 0 else None,
        "mixed": mixed,
        "mixed_over_plain": round(mixed["value"] / plain_rate, 2) if mixed and plain_rate > 0 else None,
        "water32": c0,
        "serving": serving,
        "large_system": large,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "cpu_rows": cpu_rows,
    }
    eng.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
