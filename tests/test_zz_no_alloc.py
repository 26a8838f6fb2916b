"""am1g_scf_batch_noalloc (through the C ABI): the allocation-free batched entry point gives the
same results as am1g_scf_batch, bit for bit (same kernels, fixed-order reductions), on two calls
in a row, the second with a smaller input, and refuses a batch beyond its scratch array."""
import numpy as np
import pytest

import am1g
from am1g import geometry as G

pytestmark = pytest.mark.gpu


def _same(a, b):
    assert a["status"] == b["status"] == 0
    assert a["total_ev"] == b["total_ev"]
    assert a["iterations"] == b["iterations"]
    assert np.array_equal(a["charges"], b["charges"])
    for k in ("homo_ev", "lumo_ev", "homo_beta_ev", "lumo_beta_ev"):
        assert a[k] == b[k], k
    for k in ("density", "mo_coeff", "mo_energies"):
        if k in b:
            assert np.array_equal(a[k], b[k]), k


def test_noalloc_matches_the_allocating_entry_point():
    eng = am1g.Engine(0)
    try:
        assert eng._L.am1g_noalloc_capacity_bytes() == 64 << 20
        mols = G.chain_batch(40, seed=31)
        big = [(m.z, m.xyz_bohr) for m in mols]
        # a radical (UHF) and an RM1 molecule in the same batch
        z, x = mols[0].z, mols[0].xyz_bohr
        h = int(np.nonzero(z == 1)[0][0])
        keep = np.arange(len(z)) != h
        big.append((z[keep], x[keep], {"multiplicity": 2}))
        big.append((mols[1].z, mols[1].xyz_bohr, {"method": "rm1"}))
        ref = eng.scf_batch(big, full=True)
        got = eng.scf_batch(big, full=True, noalloc=True)
        for a, b in zip(got, ref):
            _same(a, b)
        small = big[5:12]
        ref_s = eng.scf_batch(small, full=True)
        got_s = eng.scf_batch(small, full=True, noalloc=True)
        for a, b in zip(got_s, ref_s):
            _same(a, b)
        # a batch beyond the 64 MiB array is refused with a message, not run
        many = [(m.z, m.xyz_bohr) for m in G.chain_batch(400, seed=32)]
        with pytest.raises(RuntimeError, match="allocation-free"):
            eng.scf_batch(many, noalloc=True)
        # and the entry point still works after the refusal
        again = eng.scf_batch(small, noalloc=True)
        for a, b in zip(again, ref_s):
            assert a["total_ev"] == b["total_ev"]
    finally:
        eng.close()
