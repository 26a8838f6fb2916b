"""The C-ABI library loads on a machine without a GPU and exports every symbol its header
declares; the engine's own parameter derivation agrees with the oracle's (CPU only)."""
import os
import re

import numpy as np
import pytest

import am1g
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_functions():
    text = open(os.path.join(ROOT, "include", "am1g.h")).read()
    return sorted(set(re.findall(r"\b(am1g_[a-z_]+)\s*\(", text)))


def test_header_declares_the_python_export_list():
    assert _declared_functions() == sorted(am1g.EXPORTS)


def test_library_exports_every_declared_symbol():
    if not os.path.exists(am1g.LIB_PATH):
        pytest.skip("libam1g.so not built")
    L = am1g.load()
    for name in _declared_functions():
        assert hasattr(L, name), name


@pytest.mark.parametrize("method", ["am1", "rm1"])
def test_engine_parameters_match_oracle(method):
    if not os.path.exists(am1g.LIB_PATH):
        pytest.skip("libam1g.so not built")
    for z in range(1, 87):
        a = am1g.element_info(z, method)
        b = O.element(z, method)
        assert (a is None) == (b is None), z
        if a is None:
            continue
        for k in a:
            assert a[k] == pytest.approx(b[k], rel=1e-14, abs=1e-14), (z, k)


def test_request_defaults_match_am1options():
    if not os.path.exists(am1g.LIB_PATH):
        pytest.skip("libam1g.so not built")
    import ctypes

    rq = am1g.Request()
    am1g.load().am1g_request_defaults(ctypes.byref(rq))
    # scf.rs:174-197
    assert rq.max_scf == 200 and rq.e_tol == 1e-8 and rq.p_tol == 1e-7 and rq.adiis_switch == 0.1
    assert rq.accelerator == 2 and rq.use_diis == 1 and rq.multiplicity == 1 and rq.reference == 0
