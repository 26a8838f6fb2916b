"""AM1-BCC (am1g_bcc_charges, host code in libam1g) against the reference's own pins: the atom
types of tests/bcc_atom_types.rs and src/bcc/mod.rs:341-383, the nitro correction arithmetic
(tests/bcc_atom_types.rs:135-173), the parameter-file facts about bond types 8, 10 and 11
(:241-277), charge conservation, and the external antechamber values for ethanol
(tests/topology_bcc.rs:444-475). Mulliken charges come from the oracle, so these run on the CPU;
the GPU test at the end feeds the GPU's own charges."""
import math
import os
import re

import numpy as np
import pytest

import am1g
from oracle import oracle as O

ANG = 1.0 / 0.529167
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _lib_or_skip():
    if not os.path.exists(am1g.LIB_PATH):
        pytest.skip("libam1g.so not built")


def mol(atoms):
    z = [a[0] for a in atoms]
    x = np.array([a[1] for a in atoms], dtype=float) * ANG
    return z, x


def types_of(z, x):
    return [int(t) for t in am1g.bcc_correct(z, x, np.zeros(len(z)))["atom_types"]]


def table():
    text = open(os.path.join(ROOT, "am1-rs_amd", "data", "bcc_tables.inc")).read()
    rows = re.findall(r"\{(\d+), (\d+), (\d+), (-?[\d.]+)\}", text)
    return {(int(a), int(b), int(c)): float(d) for a, b, c, d in rows}


WATER = [(8, [0.0, 0.0, 0.0]), (1, [0.9584, 0.0, 0.0]), (1, [-0.24, 0.9278, 0.0])]
METHANOL = [(6, [-0.36, 0.0, 0.0]), (8, [1.06, 0.0, 0.0]), (1, [-0.74, 1.02, 0.0]), (1, [-0.74, -0.51, 0.88]),
            (1, [-0.74, -0.51, -0.88]), (1, [1.36, 0.90, 0.0])]
NITROMETHANE = [(7, [0.0, 0.0, 0.0]), (8, [1.10, 0.53, 0.0]), (8, [-1.10, 0.53, 0.0]), (6, [0.0, -1.49, 0.0]),
                (1, [0.90, -1.90, 0.45]), (1, [-0.90, -1.90, 0.45]), (1, [0.0, -1.85, -1.03])]
ETHANOL = [(6, [-1.20, 0.32, 0.0]), (6, [0.0, -0.57, 0.0]), (8, [1.19, 0.20, 0.0]), (1, [1.95, -0.37, 0.0]),
           (1, [-0.03, -1.22, 0.88]), (1, [-0.03, -1.22, -0.88]), (1, [-2.13, -0.25, 0.0]),
           (1, [-1.20, 0.96, 0.88]), (1, [-1.20, 0.96, -0.88])]


def five_ring(hetero):
    bond = 1.71 if hetero == 16 else 1.37
    radius = bond / (2.0 * math.sin(math.pi / 5.0))
    atoms = []
    for k in range(5):
        t = 2.0 * math.pi * k / 5.0 + math.pi / 2
        atoms.append((hetero if k == 0 else 6, [radius * math.cos(t), radius * math.sin(t), 0.0]))
    for k in range(5):
        if k == 0 and hetero != 7:
            continue
        t = 2.0 * math.pi * k / 5.0 + math.pi / 2
        r = radius + 1.02
        atoms.append((1, [r * math.cos(t), r * math.sin(t), 0.0]))
    return atoms


def pyridine():
    radius = 1.39 / (2.0 * math.sin(math.pi / 6.0))
    atoms = []
    for k in range(6):
        t = 2.0 * math.pi * k / 6.0
        atoms.append((7 if k == 0 else 6, [radius * math.cos(t), radius * math.sin(t), 0.0]))
    for k in range(1, 6):
        t = 2.0 * math.pi * k / 6.0
        atoms.append((1, [(radius + 1.08) * math.cos(t), (radius + 1.08) * math.sin(t), 0.0]))
    return atoms


def test_water_and_methanol_types():
    # src/bcc/mod.rs:351-383
    _lib_or_skip()
    t = types_of(*mol(WATER))
    assert t[0] == 31 and t[1] == 91
    t = types_of(*mol(METHANOL))
    assert t[0] == 11 and t[1] == 31 and all(v == 91 for v in t[2:])


def test_carbonyl_acid_amide_nitro_types():
    # tests/bcc_atom_types.rs:53-131
    _lib_or_skip()
    formaldehyde = [(6, [0.0, 0.0, 0.0]), (8, [0.0, 1.21, 0.0]), (1, [0.94, -0.59, 0.0]), (1, [-0.94, -0.59, 0.0])]
    t = types_of(*mol(formaldehyde))
    assert t[0] == 14 and t[1] == 31
    formic = [(6, [0.0, 0.0, 0.0]), (8, [0.0, 1.21, 0.0]), (8, [1.163, -0.672, 0.0]), (1, [-1.03, -0.42, 0.0]),
              (1, [1.10, -1.61, 0.0])]
    assert types_of(*mol(formic))[1] == 32
    formamide = [(6, [0.0, 0.0, 0.0]), (8, [0.0, 1.22, 0.0]), (7, [1.20, -0.72, 0.0]), (1, [-1.03, -0.43, 0.0]),
                 (1, [2.19, -0.55, 0.0]), (1, [1.15, -1.72, 0.0])]
    t = types_of(*mol(formamide))
    assert t[2] == 22 and t[1] == 31
    assert types_of(*mol(NITROMETHANE))[0] == 23


def test_aromatic_ring_types():
    # tests/bcc_atom_types.rs:175-226
    _lib_or_skip()
    t = types_of(*mol(five_ring(7)))
    assert t[0] == 23 and t[1:5] == [16, 16, 16, 16]
    t = types_of(*mol(pyridine()))
    assert t[1] == 17 and t[5] == 17 and t[2] == 16


def test_parameter_file_facts():
    # tests/bcc_atom_types.rs:241-277 and src/bcc/mod.rs:341-349
    tab = table()
    assert len(tab) == 405
    assert tab[(11, 12, 1)] == pytest.approx(0.0042, abs=1e-12)
    for dup in (8, 10):
        shared = [(i, j) for (i, j, b) in tab if b == dup and (i, j, 7) in tab]
        assert shared
        for i, j in shared:
            assert tab[(i, j, dup)] == tab[(i, j, 7)]
    same = [(k, v) for k, v in tab.items() if k[2] == 11]
    assert len(same) == 26 and all(v == 0.0 for _, v in same)


def test_nitro_correction_is_the_sum_of_its_bonds():
    # tests/bcc_atom_types.rs:135-173: the N moves by 2 d(23,31,6) - d(11,23,1)
    _lib_or_skip()
    z, x = mol(NITROMETHANE)
    q = O.scf(z, x, full=False)["charges"]
    r = am1g.bcc_correct(z, x, q)
    tab = table()
    right = 2.0 * tab[(23, 31, 6)] - tab[(11, 23, 1)]
    assert not r["warnings"], r["warnings"]
    assert r["charges"][0] - q[0] == pytest.approx(right, abs=1e-9)
    assert abs(right - (2.0 * tab[(21, 31, 6)] - tab[(11, 21, 1)])) > 0.6


def test_ethanol_matches_antechamber_and_conserves_charge():
    # tests/topology_bcc.rs:444-475: O -0.60 and hydroxyl H +0.40 within 0.03 (antechamber)
    _lib_or_skip()
    z, x = mol(ETHANOL)
    q = O.scf(z, x, full=False)["charges"]
    r = am1g.bcc_correct(z, x, q)
    assert not r["warnings"], r["warnings"]
    assert r["charges"][2] == pytest.approx(-0.60, abs=0.03)
    assert r["charges"][3] == pytest.approx(0.40, abs=0.03)
    assert abs(sum(r["charges"])) < 1e-9


def test_unparameterized_element_warns():
    _lib_or_skip()
    z, x = mol([(30, [0.0, 0.0, 0.0]), (1, [1.5, 0.0, 0.0])])
    r = am1g.bcc_correct(z, x, [0.1, -0.1])
    assert any("no AM1-BCC atom type" in w for w in r["warnings"])
    assert r["atom_types"][0] == "0"


@pytest.mark.gpu
def test_am1_bcc_through_the_gpu_scf():
    z, x = mol(ETHANOL)
    r = am1g.am1_bcc(z, [a[1] for a in ETHANOL])
    o = O.scf(z, x, full=False)
    assert np.abs(np.array(r["mulliken"]) - o["charges"]).max() <= 1e-6
    ref = am1g.bcc_correct(z, x, o["charges"])
    assert np.abs(np.array(r["charges"]) - np.array(ref["charges"])).max() <= 1e-6
    assert r["atom_types"] == ref["atom_types"] and not r["warnings"]


@pytest.mark.gpu
def test_bcc_in_the_batch_and_through_the_server(engine):
    """bcc=1 on a request: the batched launch and the server return the corrected charges next
    to the Mulliken ones; an RM1 request that asks for them is refused (the corrections are
    fitted to AM1 charges)."""
    mols = [mol(ETHANOL), mol(METHANOL), mol(NITROMETHANE)]
    rb = engine.scf_batch([(z, x, {"bcc": True}) for z, x in mols])
    for (z, x), r in zip(mols, rb):
        assert r["status"] == 0 and r["bcc_warnings"] == 0
        ref = am1g.bcc_correct(z, x, r["charges"])["charges"]
        assert np.abs(r["bcc_charges"] - np.array(ref)).max() <= 1e-12
    bad = engine.scf(mols[0][0], mols[0][1], full=False, method="rm1", bcc=True)
    assert bad["status"] == 1
    sv = am1g.Server(0, lanes=2, max_batch=8)
    try:
        ts = [sv.submit(z, x, bcc=True) for z, x in mols]
        rs = [sv.poll(t, wait_ms=-1) for t in ts]
    finally:
        sv.close()
    for a, b in zip(rs, rb):
        assert a["status"] == 0
        assert np.array_equal(a["bcc_charges"], b["bcc_charges"])
