"""Continuous-batching server (``am1g_serve_*``): every result is the same alone and in the mix,
requests that arrive while others run are admitted, cancellation and the memory budget work, in
both modes: the persistent service launch (``continuous``, the default), where a request joins the
running launch, and lanes of batched launches."""
import time

import numpy as np
import pytest

import am1g
from am1g import geometry as G
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _mix(seed=7, n_chain=24, n_water=8):
    """Chains, water clusters, doublet radicals (UHF) and RM1 requests."""
    rng = np.random.default_rng(seed)
    out = []
    for i, m in enumerate(G.chain_batch(n_chain, seed=seed)):
        z, x = m.z, m.xyz_bohr
        kw = {"method": "rm1" if i % 5 == 0 else "am1"}
        if i % 6 == 3:  # drop one hydrogen: neutral doublet, UHF
            keep = np.ones(len(z), dtype=bool)
            keep[np.nonzero(z == 1)[0][0]] = False
            z, x = z[keep], x[keep]
            kw["multiplicity"] = 2
        out.append((z, x, kw))
    for _ in range(n_water):
        w = G.small_water_cluster(int(rng.integers(2, 9)), rng)
        out.append((w.z, w.xyz_bohr, {"method": "am1"}))
    return out


def test_results_in_the_mix_equal_results_alone(engine):
    reqs = _mix()
    sv = am1g.Server(0, lanes=3, max_batch=8, continuous=False)
    try:
        tickets = [sv.submit(z, x, **kw) for z, x, kw in reqs]
        got = [sv.poll(t, wait_ms=-1) for t in tickets]
        stats = sv.stats()
    finally:
        sv.close()
    assert stats["completed"] == len(reqs) and stats["launches"] >= len(reqs) // 8
    for (z, x, kw), r in zip(reqs, got):
        alone = engine.scf_batch([(z, x, kw)])[0]
        assert r["status"] == alone["status"] == 0
        assert r["total_ev"] == alone["total_ev"]
        assert np.array_equal(r["charges"], alone["charges"])
        assert r["iterations"] == alone["iterations"]
        assert r["unrestricted"] == (kw.get("multiplicity", 1) == 2)
    for i in (0, 3, len(reqs) - 1):  # and against the oracle
        z, x, kw = reqs[i]
        o = O.scf(z, x, full=False, **kw)
        assert abs(got[i]["total_ev"] - o["total_ev"]) <= 1e-6
        assert np.abs(got[i]["charges"] - o["charges"]).max() <= 1e-6


def test_requests_arriving_during_a_launch_are_admitted(engine):
    mols = G.chain_batch(400, seed=3)
    sv = am1g.Server(0, lanes=2, max_batch=256, continuous=False)
    try:
        first = [sv.submit(m.z, m.xyz_bohr) for m in mols[:300]]
        time.sleep(0.005)
        late = [sv.submit(m.z, m.xyz_bohr, precision="mixed") for m in mols[300:]]
        res = [sv.poll(t, wait_ms=-1) for t in first + late]
        stats = sv.stats()
    finally:
        sv.close()
    assert all(r["status"] == 0 for r in res)
    assert stats["launches"] >= 2 and stats["completed"] == 400
    # mixed-precision requests stay within the FP64 contract
    ref = engine.scf_batch([(m.z, m.xyz_bohr) for m in mols[300:310]])
    for a, b in zip(res[300:310], ref):
        assert abs(a["total_ev"] - b["total_ev"]) <= 1e-6


def test_cancel_and_unknown_ticket():
    mols = G.chain_batch(64, seed=5)
    sv = am1g.Server(0, lanes=1, max_batch=4, continuous=False)
    try:
        ts = [sv.submit(m.z, m.xyz_bohr) for m in mols]
        sv.cancel(ts[-1])
        last = sv.poll(ts[-1], wait_ms=-1)
        assert last["status"] == am1g.CANCELLED
        rest = [sv.poll(t, wait_ms=-1) for t in ts[:-1]]
        assert all(r["status"] == 0 for r in rest)
        with pytest.raises(KeyError):
            sv.poll(ts[0])
    finally:
        sv.close()


def test_memory_budget_bounds_admission():
    mols = G.chain_batch(48, seed=9)
    sv = am1g.Server(0, lanes=4, max_batch=512, memory_budget_mb=2.0, continuous=False)
    try:
        ts = [sv.submit(m.z, m.xyz_bohr) for m in mols]
        res = [sv.poll(t, wait_ms=-1) for t in ts]
        stats = sv.stats()
    finally:
        sv.close()
    assert all(r["status"] == 0 for r in res)
    # a closed-shell molecule holds at least its six dense nao x nao FP64 matrices on the device,
    # so a 2 MB budget needs at least this many launches however many lanes are free
    dense_mb = sum(6 * 8 * (4 * int((m.z != 1).sum()) + int((m.z == 1).sum())) ** 2 for m in mols) / 2**20
    assert stats["launches"] >= int(dense_mb / 2.0)
    assert stats["launches"] > 1


def test_results_stream_out_before_the_launch_ends():
    """One launch holding a 21-water cluster (126 basis functions, the largest that is batched)
    held for exactly 200 iterations (zero tolerances, so it ends unconverged) and 40 small
    chains: the chains are handed out as each finishes, long before the cluster is."""
    rng = np.random.default_rng(12)
    big = G.small_water_cluster(21, rng)
    small = G.chain_batch(40, seed=13)
    sv = am1g.Server(0, lanes=1, max_batch=64, max_wait_ms=500.0, continuous=False)
    try:
        tb = sv.submit(big.z, big.xyz_bohr, e_tol=0.0, p_tol=0.0, max_scf=200)
        ts = [sv.submit(m.z, m.xyz_bohr) for m in small]
        rs = [sv.poll(t, wait_ms=-1) for t in ts]
        rb = sv.poll(tb, wait_ms=-1)
        stats = sv.stats()
    finally:
        sv.close()
    assert stats["launches"] == 1
    assert rb["status"] == 4 and rb["iterations"] == 200  # ScfNotConverged
    assert all(r["status"] == 0 for r in rs)
    assert max(r["latency_ms"] for r in rs) < rb["latency_ms"]


def test_large_requests_run_alone_on_the_grid_path(engine):
    """A water cluster above the grid-path route (192 basis functions by the committed table)
    runs alone through the plain entry point's grid-level route when it is the only such request
    of its precision in the queue, so its result is the one ``Engine.scf`` gives, while the chains
    submitted around it are batched as usual."""
    rng = np.random.default_rng(21)
    bigs = [G.small_water_cluster(k, rng) for k in (34, 40)]
    small = G.chain_batch(30, seed=22)
    sv = am1g.Server(0, lanes=2, max_batch=16, continuous=False)
    try:
        tb = [sv.submit(bigs[0].z, bigs[0].xyz_bohr)]
        ts = [sv.submit(m.z, m.xyz_bohr) for m in small[:15]]
        tb.append(sv.submit(bigs[1].z, bigs[1].xyz_bohr, precision="mixed"))
        ts += [sv.submit(m.z, m.xyz_bohr) for m in small[15:]]
        rb = [sv.poll(t, wait_ms=-1) for t in tb]
        rs = [sv.poll(t, wait_ms=-1) for t in ts]
        stats = sv.stats()
    finally:
        sv.close()
    assert all(r["status"] == 0 for r in rb + rs)
    assert stats["completed"] == 32 and stats["launches"] >= 4  # two alone, at least two batches
    for m, r, kw in zip(bigs, rb, ({}, {"precision": "mixed"})):
        alone = engine.scf(m.z, m.xyz_bohr, full=False, **kw)
        assert r["total_ev"] == alone["total_ev"]
        assert np.array_equal(r["charges"], alone["charges"])
        assert r["iterations"] == alone["iterations"]
    o = O.scf(bigs[0].z, bigs[0].xyz_bohr, full=False)
    assert abs(rb[0]["total_ev"] - o["total_ev"]) <= 1e-6
    o1 = O.scf(bigs[1].z, bigs[1].xyz_bohr, full=False)
    assert abs(rb[1]["total_ev"] - o1["total_ev"]) <= 1e-6
    alone = engine.scf_batch([(m.z, m.xyz_bohr) for m in small])
    for a, b in zip(rs, alone):
        assert a["total_ev"] == b["total_ev"]


def test_many_large_requests_share_one_launch(engine):
    """When at least serve_large_batch (default 8) grid-path requests are queued, a lane runs them
    in one batched launch: fewer launches than requests, every result within the contract of the
    oracle and bit-identical to what that path gives the molecule alone (the grid path or the
    batched kernel, whichever ran it)."""
    rng = np.random.default_rng(31)
    bigs = [G.small_water_cluster(int(k), rng) for k in rng.integers(33, 42, size=12)]
    sv = am1g.Server(0, lanes=2, max_batch=64, continuous=False)
    try:
        ts = [sv.submit(m.z, m.xyz_bohr) for m in bigs]
        rs = [sv.poll(t, wait_ms=-1) for t in ts]
        stats = sv.stats()
    finally:
        sv.close()
    assert all(r["status"] == 0 for r in rs)
    assert stats["completed"] == 12 and stats["launches"] < 12, stats
    for m, r in zip(bigs, rs):
        o = O.scf(m.z, m.xyz_bohr, full=False)
        assert abs(r["total_ev"] - o["total_ev"]) <= 1e-6
        assert np.abs(r["charges"] - o["charges"]).max() <= 1e-6
        grid = engine.scf(m.z, m.xyz_bohr, full=False)
        batched = engine.scf_batch([(m.z, m.xyz_bohr)])[0]
        assert r["total_ev"] in (grid["total_ev"], batched["total_ev"])


# ------------------------------------------------------------------ persistent service launch
def test_service_results_equal_results_alone(engine):
    """Chains, water clusters, UHF radicals and RM1 requests through the service launch: each
    result is bit for bit the molecule's batched result alone (same arithmetic, pair integrals
    computed by the request's own workgroup), and within the contract of the oracle."""
    reqs = _mix(seed=17)
    sv = am1g.Server(0)
    try:
        tickets = [sv.submit(z, x, **kw) for z, x, kw in reqs]
        got = [sv.poll(t, wait_ms=-1) for t in tickets]
        stats = sv.stats()
    finally:
        sv.close()
    assert stats["completed"] == len(reqs)
    for (z, x, kw), r in zip(reqs, got):
        alone = engine.scf_batch([(z, x, kw)])[0]
        assert r["status"] == alone["status"] == 0
        assert r["total_ev"] == alone["total_ev"]
        assert np.array_equal(r["charges"], alone["charges"])
        assert r["iterations"] == alone["iterations"]
        assert r["homo_ev"] == alone["homo_ev"] and r["lumo_ev"] == alone["lumo_ev"]
    for i in (0, 3, len(reqs) - 1):
        z, x, kw = reqs[i]
        o = O.scf(z, x, full=False, **kw)
        assert abs(got[i]["total_ev"] - o["total_ev"]) <= 1e-6
        assert np.abs(got[i]["charges"] - o["charges"]).max() <= 1e-6


def test_service_request_joins_the_running_launch():
    """A 21-water cluster held for exactly 200 iterations keeps the service launch running; chains
    submitted after it started run in that same launch (one launch in all) and finish first."""
    rng = np.random.default_rng(12)
    big = G.small_water_cluster(21, rng)
    small = G.chain_batch(40, seed=13)
    sv = am1g.Server(0)
    try:
        tb = sv.submit(big.z, big.xyz_bohr, e_tol=0.0, p_tol=0.0, max_scf=200)
        t0 = time.time()
        while sv.stats()["running"] < 1 and time.time() - t0 < 10.0:
            time.sleep(0.001)
        time.sleep(0.01)  # the cluster is on the device now
        ts = [sv.submit(m.z, m.xyz_bohr) for m in small]
        rs = [sv.poll(t, wait_ms=-1) for t in ts]
        rb = sv.poll(tb, wait_ms=-1)
        stats = sv.stats()
    finally:
        sv.close()
    assert rb["status"] == 4 and rb["iterations"] == 200  # ScfNotConverged
    assert all(r["status"] == 0 for r in rs)
    assert stats["launches"] == 1, stats
    assert max(r["latency_ms"] for r in rs) < rb["latency_ms"]


def test_service_memory_budget_and_cancel(engine):
    """A 2 MB budget admits a few requests at a time (the rest wait for memory), a cancelled
    queued request is dropped, every other result is the one the molecule gets alone, and the
    device memory held never exceeds the budget."""
    mols = G.chain_batch(48, seed=9)
    sv = am1g.Server(0, memory_budget_mb=2.0)
    try:
        ts = [sv.submit(m.z, m.xyz_bohr) for m in mols]
        sv.cancel(ts[-1])
        res = [sv.poll(t, wait_ms=-1) for t in ts]
        stats = sv.stats()
    finally:
        sv.close()
    assert res[-1]["status"] == am1g.CANCELLED
    assert all(r["status"] == 0 for r in res[:-1])
    assert stats["peak_admitted_mb"] <= 2.0
    alone = engine.scf_batch([(m.z, m.xyz_bohr) for m in mols[:-1]])
    for a, b in zip(res[:-1], alone):
        assert a["total_ev"] == b["total_ev"]


def test_service_beyond_one_workgroup_runs_on_the_grid_lane(engine):
    """A 50-water cluster (300 basis functions) is beyond the service launch's reach: it runs on
    the grid-level path and gives what ``Engine.scf`` gives, while chains go through the service.
    It is submitted while a long request keeps the service launch busy and chains keep arriving:
    the service launch pauses after its running requests, so the cluster does not wait for the
    stream of chains to end."""
    rng = np.random.default_rng(41)
    big = G.small_water_cluster(50, rng)
    hold = G.small_water_cluster(21, rng)
    small = G.chain_batch(60, seed=42)
    sv = am1g.Server(0)
    try:
        th = sv.submit(hold.z, hold.xyz_bohr, e_tol=0.0, p_tol=0.0, max_scf=200)
        ts = [sv.submit(m.z, m.xyz_bohr) for m in small[:20]]
        tb = sv.submit(big.z, big.xyz_bohr)
        ts += [sv.submit(m.z, m.xyz_bohr) for m in small[20:]]
        rb = sv.poll(tb, wait_ms=-1)
        rs = [sv.poll(t, wait_ms=-1) for t in ts]
        rh = sv.poll(th, wait_ms=-1)
        stats = sv.stats()
    finally:
        sv.close()
    assert rb["status"] == 0 and all(r["status"] == 0 for r in rs)
    assert rh["status"] == 4 and rh["iterations"] == 200
    assert stats["completed"] == 62
    alone = engine.scf(big.z, big.xyz_bohr, full=False)
    assert rb["total_ev"] == alone["total_ev"] and np.array_equal(rb["charges"], alone["charges"])
    batched = engine.scf_batch([(m.z, m.xyz_bohr) for m in small])
    for a, b in zip(rs, batched):
        assert a["total_ev"] == b["total_ev"]
