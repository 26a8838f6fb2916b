"""Device-dependent launch choices (am1g_set_tunable / am1g_get_tunable and the table read at
am1g_create): the table is read, a tunable changes routing but not results, unknown names are
refused, and the large eigensolver gives the same eigenpairs with and without its graph replay
(several replays, eager work between them)."""
import os

import numpy as np
import pytest

import am1g

pytestmark = pytest.mark.gpu


def test_table_is_read_at_create(tmp_path, monkeypatch):
    p = tmp_path / "tuned.txt"
    p.write_text("# test table\nroute_large_ao 77  # comment\njacobi_lds_n 150\nunknown_name 3\n")
    monkeypatch.setenv("AM1G_TUNE_FILE", str(p))
    eng = am1g.Engine(0)
    try:
        t = eng.tunables()
        assert t["route_large_ao"] == 77 and t["jacobi_lds_n"] == 150
        with pytest.raises(ValueError):
            eng.set_tunable("no_such_tunable", 1)
    finally:
        eng.close()


def test_routing_changes_the_path_not_the_result(engine):
    from am1g import geometry as G

    m = G.small_water_cluster(12, np.random.default_rng(1))  # 72 basis functions
    old = engine.tunables()["route_large_ao"]
    try:
        engine.set_tunable("route_large_ao", 1000)
        a = engine.scf(m.z, m.xyz_bohr, full=False)
        engine.set_tunable("route_large_ao", 0)
        b = engine.scf(m.z, m.xyz_bohr, full=False)
    finally:
        engine.set_tunable("route_large_ao", old)
    assert a["status"] == b["status"] == 0
    assert abs(a["total_ev"] - b["total_ev"]) <= 1e-6
    assert np.abs(a["charges"] - b["charges"]).max() <= 1e-6


def test_eigensolver_graph_replays_match_eager(engine):
    rng = np.random.default_rng(9)
    n = 300
    old = engine.tunables()["eig_graphs"]
    try:
        for rep in range(4):
            a = rng.standard_normal((n, n))
            a = a + a.T
            engine.set_tunable("eig_graphs", 1)
            wg, vg = engine.eigh(a)
            engine.set_tunable("eig_graphs", 0)
            we, ve = engine.eigh(a)
            assert np.array_equal(wg, we) and np.array_equal(vg, ve), rep
            wn = np.linalg.eigvalsh(a)
            assert np.abs(wg - wn).max() < 1e-10 * max(1.0, np.abs(wn).max())
    finally:
        engine.set_tunable("eig_graphs", old)
