"""ctypes front end of the FP64 CPU oracle (``liboracle.so``).

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, by ``__graft_entry__.smoke()`` and by the
CPU-baseline leg of ``bench.py``. The GPU library never imports, calls or falls back to it.

The oracle restates am1-rs 0.2.3's ``run_am1`` (src/scf.rs:294-468) and its callees in plain
C++ double precision; see ``am1_oracle.cpp`` for the file:line each routine follows.
"""
from __future__ import annotations

import ctypes
import glob
import math
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

METHODS = {"am1": 0, "rm1": 1}
REFERENCES = {"auto": 0, "rhf": 1, "uhf": 2}
ACCELERATORS = {"none": 0, "cdiis": 1, "adiis_cdiis": 2}
ANGSTROM_TO_BOHR = 1.0 / 0.529167  # constants.rs:18


class _Request(ctypes.Structure):
    _fields_ = [
        ("n_atoms", ctypes.c_int),
        ("z", ctypes.POINTER(ctypes.c_int)),
        ("xyz_bohr", ctypes.POINTER(ctypes.c_double)),
        ("charge", ctypes.c_double),
        ("multiplicity", ctypes.c_int),
        ("method", ctypes.c_int),
        ("reference", ctypes.c_int),
        ("accelerator", ctypes.c_int),
        ("max_scf", ctypes.c_int),
        ("e_tol", ctypes.c_double),
        ("p_tol", ctypes.c_double),
        ("adiis_switch", ctypes.c_double),
        ("use_diis", ctypes.c_int),
        ("electric_field", ctypes.POINTER(ctypes.c_double)),
        ("multipole_cutoff", ctypes.c_double),
    ]


class _Result(ctypes.Structure):
    _fields_ = [
        ("status", ctypes.c_int),
        ("nao", ctypes.c_int),
        ("n_occ", ctypes.c_int),
        ("n_occ_beta", ctypes.c_int),
        ("iterations", ctypes.c_int),
        ("converged", ctypes.c_int),
        ("unrestricted", ctypes.c_int),
        ("electronic_ev", ctypes.c_double),
        ("core_ev", ctypes.c_double),
        ("total_ev", ctypes.c_double),
        ("heat_of_formation_kcal", ctypes.c_double),
        ("homo_ev", ctypes.c_double),
        ("lumo_ev", ctypes.c_double),
        ("homo_beta_ev", ctypes.c_double),
        ("lumo_beta_ev", ctypes.c_double),
        ("dipole_debye", ctypes.c_double * 3),
        ("residual", ctypes.c_double),
        ("field_nuclear_ev", ctypes.c_double),
        ("charges", ctypes.POINTER(ctypes.c_double)),
        ("mo_energies", ctypes.POINTER(ctypes.c_double)),
        ("density", ctypes.POINTER(ctypes.c_double)),
        ("mo_coeff", ctypes.POINTER(ctypes.c_double)),
        ("mo_energies_beta", ctypes.POINTER(ctypes.c_double)),
        ("message", ctypes.c_char * 256),
        ("density_alpha", ctypes.POINTER(ctypes.c_double)),
        ("density_beta", ctypes.POINTER(ctypes.c_double)),
    ]


_lib = None


def build() -> str:
    """Compile the oracle with its Makefile (gcc, OpenMP)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


def _find_openblas() -> str | None:
    import numpy

    base = os.path.dirname(numpy.__file__)
    cands = glob.glob(os.path.join(base, "..", "numpy.libs", "libscipy_openblas*.so")) + glob.glob(
        os.path.join(base, ".libs", "libscipy_openblas*.so")
    )
    return os.path.abspath(cands[0]) if cands else None


def lib(use_lapack: bool = True, threads: int = 0):
    """Load ``liboracle.so`` (building it if absent); optionally bind numpy's OpenBLAS LAPACK."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        L.oracle_scf.argtypes = [ctypes.POINTER(_Request), ctypes.POINTER(_Result)]
        L.oracle_scf_many.argtypes = [ctypes.c_int, ctypes.POINTER(_Request), ctypes.POINTER(_Result), ctypes.c_int]
        L.oracle_set_lapack.argtypes = [ctypes.c_char_p, ctypes.c_int]
        L.oracle_element.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
        L.oracle_bintgs.argtypes = [ctypes.c_double, ctypes.POINTER(ctypes.c_double)]
        L.oracle_aintgs.argtypes = [ctypes.c_double, ctypes.POINTER(ctypes.c_double)]
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int)
        L.oracle_overlap.argtypes = [ctypes.c_int, dp, ctypes.c_int, dp, ctypes.c_int, dp]
        L.oracle_pair.argtypes = [ctypes.c_int, ctypes.c_int, dp, ctypes.c_int, dp, dp, dp, ip]
        L.oracle_core.argtypes = [ctypes.c_int, ip, dp, ctypes.c_int, dp, dp, ip]
        L.oracle_fock.argtypes = [ctypes.c_int, ip, dp, ctypes.c_int, dp, dp, ctypes.c_double, dp]
        L.oracle_eigen.argtypes = [ctypes.c_int, dp, dp, dp]
        L.oracle_diis_solve.argtypes = [ctypes.c_int, dp, dp]
        L.oracle_adiis_simplex.argtypes = [ctypes.c_int, dp, dp, dp]
        L.oracle_overlap_locals.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int, ctypes.c_int, dp]
        L.oracle_gradient.argtypes = [ctypes.POINTER(_Request), ctypes.POINTER(_Result), ctypes.c_double, dp]
        _lib = L
        if use_lapack:
            path = _find_openblas()
            if path:
                L.oracle_set_lapack(path.encode(), threads)
    return _lib


def lapack_active() -> bool:
    return bool(lib().oracle_lapack_active())


def set_lapack(enabled: bool, threads: int = 0) -> bool:
    L = lib()
    path = _find_openblas() if enabled else None
    L.oracle_set_lapack(path.encode() if path else b"", threads)
    return bool(L.oracle_lapack_active())


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def _make_request(z, xyz_bohr, charge=0.0, multiplicity=1, method="am1", reference="auto",
                  accelerator="adiis_cdiis", max_scf=200, e_tol=1e-8, p_tol=1e-7, adiis_switch=0.1,
                  use_diis=True, electric_field=None, multipole_cutoff=None, keep=None):
    """``electric_field``: three floats in eV per (e*Bohr), the unit of Am1Options::electric_field
    (scf.rs:141-153), or None. ``multipole_cutoff``: Am1Options::multipole_cutoff (Bohr,
    scf.rs:130-140), None or a value that is not > 0 for none."""
    z = np.ascontiguousarray(z, dtype=np.int32)
    xyz = np.ascontiguousarray(xyz_bohr, dtype=np.float64).reshape(-1, 3)
    if keep is not None:
        keep.extend([z, xyz])
    rq = _Request()
    rq.n_atoms = len(z)
    rq.z = _ip(z)
    rq.xyz_bohr = _dp(xyz)
    rq.charge = float(charge)
    rq.multiplicity = int(multiplicity)
    rq.method = METHODS[method] if isinstance(method, str) else int(method)
    rq.reference = REFERENCES[reference] if isinstance(reference, str) else int(reference)
    rq.accelerator = ACCELERATORS[accelerator] if isinstance(accelerator, str) else int(accelerator)
    rq.max_scf = int(max_scf)
    rq.e_tol = float(e_tol)
    rq.p_tol = float(p_tol)
    rq.adiis_switch = float(adiis_switch)
    rq.use_diis = 1 if use_diis else 0
    rq.multipole_cutoff = float(multipole_cutoff) if multipole_cutoff is not None else 0.0
    if electric_field is not None:
        f = np.ascontiguousarray(electric_field, dtype=np.float64).reshape(3)
        if keep is not None:
            keep.append(f)
        rq.electric_field = _dp(f)
    return rq


def _result_dict(rs, extra):
    out = {
        "status": rs.status,
        "message": rs.message.decode(errors="replace"),
        "nao": rs.nao,
        "n_occ": rs.n_occ,
        "n_occ_beta": rs.n_occ_beta,
        "iterations": rs.iterations,
        "converged": bool(rs.converged),
        "unrestricted": bool(rs.unrestricted),
        "electronic_ev": rs.electronic_ev,
        "core_ev": rs.core_ev,
        "total_ev": rs.total_ev,
        "heat_of_formation_kcal": rs.heat_of_formation_kcal,
        "homo_ev": None if math.isnan(rs.homo_ev) else rs.homo_ev,
        "lumo_ev": None if math.isnan(rs.lumo_ev) else rs.lumo_ev,
        "homo_beta_ev": None if math.isnan(rs.homo_beta_ev) else rs.homo_beta_ev,
        "lumo_beta_ev": None if math.isnan(rs.lumo_beta_ev) else rs.lumo_beta_ev,
        "dipole_debye": np.array(list(rs.dipole_debye)),
        "residual": rs.residual,
        "field_nuclear_ev": rs.field_nuclear_ev,
    }
    out.update(extra)
    return out


def scf(z, xyz_bohr, full=True, **kw):
    """One SCF (``run_am1``); returns a dict of observables (arrays when ``full``)."""
    L = lib()
    keep = []
    rq = _make_request(z, xyz_bohr, keep=keep, **kw)
    n = len(z)
    nao_max = 4 * n
    rs = _Result()
    charges = np.zeros(n)
    rs.charges = _dp(charges)
    extra = {"charges": charges}
    if full:
        mo = np.zeros(nao_max)
        mob = np.zeros(nao_max)
        dens = np.zeros(nao_max * nao_max)
        coeff = np.zeros(nao_max * nao_max)
        rs.mo_energies = _dp(mo)
        rs.mo_energies_beta = _dp(mob)
        rs.density = _dp(dens)
        rs.mo_coeff = _dp(coeff)
        da = np.zeros(nao_max * nao_max)
        db = np.zeros(nao_max * nao_max)
        rs.density_alpha = _dp(da)
        rs.density_beta = _dp(db)
    L.oracle_scf(ctypes.byref(rq), ctypes.byref(rs))
    if full:
        nao = rs.nao
        extra.update(
            mo_energies=mo[:nao].copy(),
            mo_energies_beta=mob[:nao].copy() if rs.unrestricted else None,
            density=dens[: nao * nao].reshape(nao, nao).copy(),
            mo_coeff=coeff[: nao * nao].reshape(nao, nao).copy(),
            spin_density=(da[: nao * nao] - db[: nao * nao]).reshape(nao, nao) if rs.unrestricted else None,
        )
    return _result_dict(rs, extra)


def scf_many(molecules, threads=0, **kw):
    """Independent SCFs, OpenMP over molecules. ``molecules`` = list of (z, xyz_bohr) or
    (z, xyz_bohr, per_molecule_kwargs)."""
    L = lib()
    keep = []
    n = len(molecules)
    rqs = (_Request * n)()
    rss = (_Result * n)()
    charges = []
    for i, mol in enumerate(molecules):
        z, xyz = mol[0], mol[1]
        opts = dict(kw)
        if len(mol) > 2 and mol[2]:
            opts.update(mol[2])
        rqs[i] = _make_request(z, xyz, keep=keep, **opts)
        q = np.zeros(len(z))
        charges.append(q)
        rss[i].charges = _dp(q)
    L.oracle_scf_many(n, rqs, rss, int(threads))
    return [_result_dict(rss[i], {"charges": charges[i]}) for i in range(n)]


def overlap_locals(za, zb, r, method="am1", numeric=False):
    """The five local overlaps of the ordered pair (za the higher n) at distance r (Bohr): the
    closed forms, or with ``numeric`` the n >= 4 quadrature."""
    out = np.zeros(5)
    st = lib().oracle_overlap_locals(int(za), int(zb), float(r), METHODS[method], 1 if numeric else 0, _dp(out))
    if st:
        raise ValueError("no closed form for this pair")
    return out


def gradient(z, xyz_bohr, step=1.0e-4, **kw):
    """Nuclear gradient (eV/Bohr, [n][3]) of the converged SCF by central differences of the energy
    at the converged densities (``oracle_gradient``), with the SCF's result dict."""
    L = lib()
    keep = []
    rq = _make_request(z, xyz_bohr, keep=keep, **kw)
    n = len(z)
    rs = _Result()
    charges = np.zeros(n)
    rs.charges = _dp(charges)
    g = np.zeros((n, 3))
    L.oracle_gradient(ctypes.byref(rq), ctypes.byref(rs), float(step), _dp(g))
    out = _result_dict(rs, {"charges": charges})
    out["gradient"] = g
    return out


def gradient_density(z, xyz_bohr, density, spin_density=None, step=1.0e-4, **kw):
    """Fixed-density gradient (eV/Bohr, [n][3]) at a given total density and, for an open shell,
    spin density P_alpha - P_beta (``oracle_gradient_density``)."""
    L = lib()
    L.oracle_gradient_density.argtypes = [ctypes.POINTER(_Request), ctypes.POINTER(ctypes.c_double),
                                          ctypes.POINTER(ctypes.c_double), ctypes.c_double,
                                          ctypes.POINTER(ctypes.c_double)]
    keep = []
    rq = _make_request(z, xyz_bohr, keep=keep, **kw)
    p = np.ascontiguousarray(density, dtype=np.float64)
    sp = np.zeros_like(p) if spin_density is None else np.ascontiguousarray(spin_density, dtype=np.float64)
    pa = np.ascontiguousarray(0.5 * (p + sp))
    pb = np.ascontiguousarray(0.5 * (p - sp))
    g = np.zeros((len(z), 3))
    st = L.oracle_gradient_density(ctypes.byref(rq), _dp(pa), _dp(pb), float(step), _dp(g))
    if st:
        raise ValueError(f"oracle_gradient_density status {st}")
    return g


def element(z, method="am1"):
    out = np.zeros(23)
    st = lib().oracle_element(int(z), METHODS[method], _dp(out))
    if st:
        return None
    keys = ["n", "norb", "uss", "upp", "zs", "zp", "bs", "bp", "gss", "gsp", "gpp", "gp2", "hsp", "alpha",
            "core", "eheat_ev", "e_isol", "dd", "qq", "rho0", "rho1", "rho2", "ng"]
    return dict(zip(keys, out.tolist()))


def bintgs(x):
    out = np.zeros(13)
    lib().oracle_bintgs(float(x), _dp(out))
    return out


def aintgs(x):
    out = np.zeros(13)
    lib().oracle_aintgs(float(x), _dp(out))
    return out


def overlap(zi, pi, zj, pj, method="am1"):
    out = np.zeros(16)
    pi = np.ascontiguousarray(pi, dtype=np.float64)
    pj = np.ascontiguousarray(pj, dtype=np.float64)
    st = lib().oracle_overlap(int(zi), _dp(pi), int(zj), _dp(pj), METHODS[method], _dp(out))
    if st:
        raise ValueError("overlap not available for these elements")
    return out.reshape(4, 4)


def pair(za, zb, dvec, method="am1"):
    """Two-centre integrals of the ordered pair (a at origin, b at dvec, Bohr)."""
    w = np.zeros(100)
    e1b = np.zeros(16)
    e2a = np.zeros(16)
    dims = np.zeros(3, dtype=np.int32)
    d = np.ascontiguousarray(dvec, dtype=np.float64)
    lib().oracle_pair(int(za), int(zb), _dp(d), METHODS[method], _dp(w), _dp(e1b), _dp(e2a), _ip(dims))
    na, nb, wcols = dims.tolist()
    rows = 10 if na == 4 else 1
    return {"w": w[: rows * wcols].reshape(rows, wcols), "e1b": e1b.reshape(4, 4), "e2a": e2a.reshape(4, 4),
            "na": na, "nb": nb}


def core(z, xyz_bohr, method="am1"):
    z = np.ascontiguousarray(z, dtype=np.int32)
    xyz = np.ascontiguousarray(xyz_bohr, dtype=np.float64)
    nao_max = 4 * len(z)
    H = np.zeros(nao_max * nao_max)
    e = np.zeros(1)
    nao = np.zeros(1, dtype=np.int32)
    st = lib().oracle_core(len(z), _ip(z), _dp(xyz), METHODS[method], _dp(H), _dp(e), _ip(nao))
    if st:
        raise ValueError(f"oracle_core status {st}")
    n = int(nao[0])
    return H[: n * n].reshape(n, n).copy(), float(e[0])


def fock(z, xyz_bohr, p_tot, p_spin=None, spin_scale=0.5, method="am1"):
    z = np.ascontiguousarray(z, dtype=np.int32)
    xyz = np.ascontiguousarray(xyz_bohr, dtype=np.float64)
    pt = np.ascontiguousarray(p_tot, dtype=np.float64)
    ps = pt if p_spin is None else np.ascontiguousarray(p_spin, dtype=np.float64)
    F = np.zeros_like(pt)
    st = lib().oracle_fock(len(z), _ip(z), _dp(xyz), METHODS[method], _dp(pt), _dp(ps), float(spin_scale), _dp(F))
    if st:
        raise ValueError(f"oracle_fock status {st}")
    return F


def eigen(a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    n = a.shape[0]
    w = np.zeros(n)
    v = np.zeros((n, n))
    st = lib().oracle_eigen(n, _dp(a), _dp(w), _dp(v))
    if st:
        raise ValueError("eigensolver rejected the input")
    return w, v


def fill(energies, weights, n_elec, kt):
    """``fermi::fill`` (fermi.rs:124-269): fractions, chemical potential and T*S for levels with
    the given weights; ``kt <= 0`` is aufbau. Raises ``ValueError`` when the electrons do not fit."""
    e = np.ascontiguousarray(energies, dtype=np.float64)
    w = np.ascontiguousarray(weights, dtype=np.float64)
    f = np.zeros(len(e))
    mu = ctypes.c_double()
    ts = ctypes.c_double()
    L = lib()
    L.oracle_fill.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                              ctypes.c_double, ctypes.c_double, ctypes.POINTER(ctypes.c_double),
                              ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
    if L.oracle_fill(len(e), _dp(e), _dp(w), float(n_elec), float(kt), _dp(f), ctypes.byref(mu), ctypes.byref(ts)):
        raise ValueError("cannot fill: the electron count does not fit the levels")
    return f, mu.value, ts.value


def diis_solve(B):
    B = np.ascontiguousarray(B, dtype=np.float64)
    m = B.shape[0]
    c = np.zeros(m + 1)
    st = lib().oracle_diis_solve(m, _dp(B), _dp(c))
    return None if st else c


def adiis_simplex(d, s):
    d = np.ascontiguousarray(d, dtype=np.float64)
    s = np.ascontiguousarray(s, dtype=np.float64)
    c = np.zeros(len(d))
    lib().oracle_adiis_simplex(len(d), _dp(d), _dp(s), _dp(c))
    return c


# ---------------------------------------------------------------------------- divide and conquer
class _DcOptions(ctypes.Structure):
    _fields_ = [
        ("core_size", ctypes.c_int),
        ("buffer_radius", ctypes.c_double),
        ("kt", ctypes.c_double),
        ("max_scf", ctypes.c_int),
        ("e_tol", ctypes.c_double),
        ("p_tol", ctypes.c_double),
        ("mixing", ctypes.c_double),
        ("core_of_atom", ctypes.POINTER(ctypes.c_int)),
    ]


class _DcResult(ctypes.Structure):
    _fields_ = [
        ("status", ctypes.c_int),
        ("iterations", ctypes.c_int),
        ("converged", ctypes.c_int),
        ("unrestricted", ctypes.c_int),
        ("n_subsystems", ctypes.c_int),
        ("largest_subsystem_aos", ctypes.c_int),
        ("electronic_ev", ctypes.c_double),
        ("core_ev", ctypes.c_double),
        ("total_ev", ctypes.c_double),
        ("heat_of_formation_kcal", ctypes.c_double),
        ("fermi_ev", ctypes.c_double),
        ("fermi_beta_ev", ctypes.c_double),
        ("entropy_ev", ctypes.c_double),
        ("gap_ev", ctypes.c_double),
        ("charges", ctypes.POINTER(ctypes.c_double)),
        ("density", ctypes.POINTER(ctypes.c_double)),
        ("message", ctypes.c_char * 256),
    ]


def _dc_options(core_size, buffer_radius, kt, max_scf, e_tol, p_tol, mixing, core_of_atom, keep):
    op = _DcOptions()
    op.core_size = int(core_size)
    op.buffer_radius = float(buffer_radius)
    op.kt = float(kt)
    op.max_scf = int(max_scf)
    op.e_tol = float(e_tol)
    op.p_tol = float(p_tol)
    op.mixing = float(mixing)
    if core_of_atom is not None:
        c = np.ascontiguousarray(core_of_atom, dtype=np.int32)
        keep.append(c)
        op.core_of_atom = _ip(c)
    return op


def dc(z, xyz_bohr, core_size=12, buffer_radius=11.0, kt=0.05, max_scf=300, e_tol=1e-7, p_tol=1e-6,
       mixing=0.4, core_of_atom=None, density=False, **kw):
    """``run_divide_conquer`` (divide_conquer.rs:406-572). ``buffer_radius`` in Bohr, ``kt`` in eV
    (``kt <= 0`` is aufbau filling). ``core_of_atom`` gives caller-chosen cores."""
    L = lib()
    L.oracle_dc.argtypes = [ctypes.POINTER(_Request), ctypes.POINTER(_DcOptions), ctypes.POINTER(_DcResult)]
    keep = []
    rq = _make_request(z, xyz_bohr, keep=keep, **kw)
    op = _dc_options(core_size, buffer_radius, kt, max_scf, e_tol, p_tol, mixing, core_of_atom, keep)
    rs = _DcResult()
    q = np.zeros(len(z))
    rs.charges = _dp(q)
    P = None
    if density:
        nao = sum(1 if int(v) == 1 else 4 for v in z)
        P = np.zeros(nao * nao)
        rs.density = _dp(P)
    L.oracle_dc(ctypes.byref(rq), ctypes.byref(op), ctypes.byref(rs))
    out = {k: getattr(rs, k) for k, _ in _DcResult._fields_ if k not in ("charges", "density", "message")}
    out["message"] = rs.message.decode(errors="replace")
    out["charges"] = q
    out["converged"] = bool(rs.converged)
    if density:
        n = int(round(len(P) ** 0.5))
        out["density"] = P.reshape(n, n)
    return out


def dc_weight_sum(z, xyz_bohr, core_size=12, buffer_radius=11.0, core_of_atom=None, method="am1"):
    L = lib()
    L.oracle_dc_weight_sum.argtypes = [ctypes.POINTER(_Request), ctypes.POINTER(_DcOptions),
                                       ctypes.POINTER(ctypes.c_double)]
    keep = []
    rq = _make_request(z, xyz_bohr, keep=keep, method=method)
    op = _dc_options(core_size, buffer_radius, 0.0, 1, 0.0, 0.0, 0.0, core_of_atom, keep)
    n = len(z)
    out = np.zeros(n * n)
    L.oracle_dc_weight_sum(ctypes.byref(rq), ctypes.byref(op), _dp(out))
    return out.reshape(n, n)
