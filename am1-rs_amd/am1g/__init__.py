"""am1g — MI355X-native AM1/RM1 SCF engine (Python front end over the C ABI in include/am1g.h).

``single_point`` mirrors am1-rs's Python entry point ``am1_rs.native.single_point``
(python/am1_rs/native.py:52-97 -> src/python.rs:147-199): coordinates in Angstrom, result dict
in eV and Hartree, every error raised as ``ValueError`` with the engine's message
(python.rs:27-29). ``Engine.scf_batch`` runs many molecules in one persistent GPU launch.

The engine is ``libam1g.so`` built from ``csrc/`` for gfx950. There is no CPU fallback: when the
library or a GPU is missing, loading fails with an error that says so.
"""
from __future__ import annotations

import ctypes
import time
import math
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libam1g.so")
CSRC = os.path.join(os.path.dirname(HERE), "csrc")

AM1_EV = 27.21            # constants.rs:13
ANGSTROM_TO_BOHR = 1.0 / 0.529167  # constants.rs:18

METHODS = {"am1": 0, "rm1": 1}
REFERENCES = {"auto": 0, "rhf": 1, "uhf": 2, "restricted": 1, "unrestricted": 2}
ACCELERATORS = {"none": 0, "cdiis": 1, "adiis_cdiis": 2, "adiis": 2}
PRECISIONS = {"fp64": 0, "mixed": 1}
STATUS_NAMES = {0: "Ok", 1: "InvalidInput", 2: "ElementNotParameterized", 3: "LinearAlgebra",
                4: "ScfNotConverged", 5: "ScfDiverged", 6: "DeviceError", 7: "Pending", 8: "Cancelled"}
PENDING = 7
CANCELLED = 8
PHASES = ["setup", "fock", "commutator+diis", "eigen", "density", "final"]
SYMBOLS = {"H": 1, "Be": 4, "B": 5, "C": 6, "N": 7, "O": 8, "F": 9, "Al": 13, "Si": 14, "P": 15, "S": 16,
           "Cl": 17, "Zn": 30, "Ge": 32, "As": 33, "Se": 34, "Br": 35, "Sb": 51, "Te": 52, "I": 53, "Hg": 80}


class Request(ctypes.Structure):
    _fields_ = [
        ("n_atoms", ctypes.c_int),
        ("z", ctypes.POINTER(ctypes.c_uint8)),
        ("xyz_bohr", ctypes.POINTER(ctypes.c_double)),
        ("charge", ctypes.c_double),
        ("multiplicity", ctypes.c_int),
        ("method", ctypes.c_int),
        ("reference", ctypes.c_int),
        ("accelerator", ctypes.c_int),
        ("use_diis", ctypes.c_int),
        ("max_scf", ctypes.c_int),
        ("e_tol", ctypes.c_double),
        ("p_tol", ctypes.c_double),
        ("adiis_switch", ctypes.c_double),
        ("precision", ctypes.c_int),
        ("electric_field", ctypes.POINTER(ctypes.c_double)),
        ("bcc", ctypes.c_int),
        ("multipole_cutoff", ctypes.c_double),
    ]


class Result(ctypes.Structure):
    _fields_ = [
        ("status", ctypes.c_int),
        ("nao", ctypes.c_int),
        ("n_occ", ctypes.c_int),
        ("n_occ_beta", ctypes.c_int),
        ("iterations", ctypes.c_int),
        ("converged", ctypes.c_int),
        ("unrestricted", ctypes.c_int),
        ("has_homo", ctypes.c_int),
        ("has_lumo", ctypes.c_int),
        ("has_homo_beta", ctypes.c_int),
        ("has_lumo_beta", ctypes.c_int),
        ("electronic_ev", ctypes.c_double),
        ("core_ev", ctypes.c_double),
        ("total_ev", ctypes.c_double),
        ("heat_of_formation_kcal", ctypes.c_double),
        ("homo_ev", ctypes.c_double),
        ("lumo_ev", ctypes.c_double),
        ("homo_beta_ev", ctypes.c_double),
        ("lumo_beta_ev", ctypes.c_double),
        ("dipole_debye", ctypes.c_double * 3),
        ("dipole_magnitude", ctypes.c_double),
        ("field_nuclear_ev", ctypes.c_double),
        ("residual", ctypes.c_double),
        ("fp32_iterations", ctypes.c_int),
        ("latency_ms", ctypes.c_double),
        ("phase_ms", ctypes.c_double * 6),
        ("eigen_sweeps", ctypes.c_int),
        ("bcc_warnings", ctypes.c_int),
        ("charges", ctypes.POINTER(ctypes.c_double)),
        ("mo_energies", ctypes.POINTER(ctypes.c_double)),
        ("mo_energies_beta", ctypes.POINTER(ctypes.c_double)),
        ("density", ctypes.POINTER(ctypes.c_double)),
        ("mo_coeff", ctypes.POINTER(ctypes.c_double)),
        ("mo_coeff_beta", ctypes.POINTER(ctypes.c_double)),
        ("spin_density", ctypes.POINTER(ctypes.c_double)),
        ("bcc_charges", ctypes.POINTER(ctypes.c_double)),
    ]


class ServeOptions(ctypes.Structure):
    _fields_ = [
        ("lanes", ctypes.c_int),
        ("max_batch", ctypes.c_int),
        ("memory_budget_mb", ctypes.c_double),
        ("max_wait_ms", ctypes.c_double),
        ("continuous", ctypes.c_int),
    ]


class ServeStats(ctypes.Structure):
    _fields_ = [
        ("submitted", ctypes.c_longlong),
        ("completed", ctypes.c_longlong),
        ("cancelled", ctypes.c_longlong),
        ("queued", ctypes.c_longlong),
        ("running", ctypes.c_longlong),
        ("launches", ctypes.c_longlong),
        ("admitted_mb", ctypes.c_double),
        ("peak_admitted_mb", ctypes.c_double),
    ]


class DcOptions(ctypes.Structure):
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


class DcResult(ctypes.Structure):
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
        ("homo_lumo_gap_ev", ctypes.c_double),
        ("diagonalization_work", ctypes.c_double),
        ("retained_pattern", ctypes.c_longlong),
        ("phase_ms", ctypes.c_double * 5),
        ("charges", ctypes.POINTER(ctypes.c_double)),
        ("density", ctypes.POINTER(ctypes.c_double)),
        ("spin_density", ctypes.POINTER(ctypes.c_double)),
    ]


DC_PHASES = ["fock", "eigen", "fermi", "assembly", "diis"]

# Every symbol include/am1g.h declares (checked by the CPU test-suite against the built library).
EXPORTS = ["am1g_request_defaults", "am1g_create", "am1g_destroy", "am1g_scf", "am1g_scf_batch",
           "am1g_dc_defaults", "am1g_dc", "am1g_serve_defaults", "am1g_serve_create", "am1g_serve_destroy",
           "am1g_serve_submit", "am1g_serve_poll", "am1g_serve_cancel", "am1g_serve_get_stats",
           "am1g_serve_last_error", "am1g_symmetric_eigen", "am1g_last_error", "am1g_element_info", "am1g_last_kernel_ms",
           "am1g_scf_batch_noalloc", "am1g_noalloc_capacity_bytes", "am1g_bcc_charges", "am1g_gradient",
           "am1g_set_tunable", "am1g_get_tunable", "am1g_core_hamiltonian", "am1g_gradient_density",
           "am1g_matmul"]

_lib = None


def build(verbose: bool = False) -> str:
    """Compile ``libam1g.so`` for gfx950 with the Makefile in ``csrc/``."""
    subprocess.run(["make", "-C", CSRC] + ([] if verbose else ["-s"]), check=True)
    return LIB_PATH


def load():
    """Load ``libam1g.so``. Raises ``RuntimeError`` when it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"am1g: {LIB_PATH} is missing; build it with __graft_entry__.build() "
                           "(there is no CPU fallback)")
    L = ctypes.CDLL(LIB_PATH)
    L.am1g_create.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]
    L.am1g_destroy.argtypes = [ctypes.c_void_p]
    L.am1g_scf.argtypes = [ctypes.c_void_p, ctypes.POINTER(Request), ctypes.POINTER(Result)]
    L.am1g_scf_batch.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(Request), ctypes.POINTER(Result)]
    L.am1g_scf_batch_noalloc.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(Request),
                                         ctypes.POINTER(Result)]
    L.am1g_bcc_charges.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_uint8), ctypes.POINTER(ctypes.c_double),
                                   ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                                   ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int), ctypes.c_char_p,
                                   ctypes.c_int]
    L.am1g_set_tunable.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_double]
    L.am1g_get_tunable.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.POINTER(ctypes.c_double)]
    L.am1g_gradient.argtypes = [ctypes.c_void_p, ctypes.POINTER(Request), ctypes.POINTER(Result),
                                ctypes.POINTER(ctypes.c_double)]
    L.am1g_core_hamiltonian.argtypes = [ctypes.c_void_p, ctypes.POINTER(Request), ctypes.POINTER(ctypes.c_double),
                                        ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                                        ctypes.c_longlong]
    L.am1g_gradient_density.argtypes = [ctypes.c_void_p, ctypes.POINTER(Request), ctypes.POINTER(ctypes.c_double),
                                        ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
    L.am1g_noalloc_capacity_bytes.argtypes = []
    L.am1g_noalloc_capacity_bytes.restype = ctypes.c_longlong
    L.am1g_dc_defaults.argtypes = [ctypes.POINTER(DcOptions)]
    L.am1g_dc.argtypes = [ctypes.c_void_p, ctypes.POINTER(Request), ctypes.POINTER(DcOptions),
                          ctypes.POINTER(DcResult)]
    L.am1g_serve_defaults.argtypes = [ctypes.POINTER(ServeOptions)]
    L.am1g_serve_create.argtypes = [ctypes.c_int, ctypes.POINTER(ServeOptions), ctypes.POINTER(ctypes.c_void_p)]
    L.am1g_serve_destroy.argtypes = [ctypes.c_void_p]
    L.am1g_serve_submit.argtypes = [ctypes.c_void_p, ctypes.POINTER(Request), ctypes.POINTER(ctypes.c_longlong)]
    L.am1g_serve_poll.argtypes = [ctypes.c_void_p, ctypes.c_longlong, ctypes.c_double, ctypes.POINTER(Result)]
    L.am1g_serve_cancel.argtypes = [ctypes.c_void_p, ctypes.c_longlong]
    L.am1g_serve_get_stats.argtypes = [ctypes.c_void_p, ctypes.POINTER(ServeStats)]
    L.am1g_serve_last_error.argtypes = [ctypes.c_void_p]
    L.am1g_serve_last_error.restype = ctypes.c_char_p
    L.am1g_symmetric_eigen.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_double),
                                       ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]
    L.am1g_matmul.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                              ctypes.c_double, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                              ctypes.c_double, ctypes.POINTER(ctypes.c_double), ctypes.c_int]
    L.am1g_last_error.argtypes = [ctypes.c_void_p]
    L.am1g_last_error.restype = ctypes.c_char_p
    L.am1g_request_defaults.argtypes = [ctypes.POINTER(Request)]
    L.am1g_element_info.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
    L.am1g_last_kernel_ms.argtypes = [ctypes.c_void_p]
    L.am1g_last_kernel_ms.restype = ctypes.c_double
    _lib = L
    return L


def element_info(z: int, method: str = "am1"):
    out = (ctypes.c_double * 23)()
    st = load().am1g_element_info(int(z), METHODS[method], out)
    if st:
        return None
    keys = ["n", "norb", "uss", "upp", "zs", "zp", "bs", "bp", "gss", "gsp", "gpp", "gp2", "hsp", "alpha",
            "core", "eheat_ev", "e_isol", "dd", "qq", "rho0", "rho1", "rho2", "ng"]
    return dict(zip(keys, list(out)))


class Engine:
    """One am1g context on one GPU (owns the device pools and stream)."""

    def __init__(self, device: int = 0):
        self._L = load()
        self._ctx = ctypes.c_void_p()
        st = self._L.am1g_create(int(device), ctypes.byref(self._ctx))
        if st != 0:
            msg = self._L.am1g_last_error(None)
            raise RuntimeError(f"am1g_create failed ({STATUS_NAMES.get(st, st)}): {msg.decode() if msg else ''}")

    def close(self):
        if self._ctx:
            self._L.am1g_destroy(self._ctx)
            self._ctx = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def last_error(self) -> str:
        m = self._L.am1g_last_error(self._ctx)
        return m.decode() if m else ""

    def last_kernel_ms(self) -> float:
        return float(self._L.am1g_last_kernel_ms(self._ctx))

    TUNABLES = ("route_large_ao", "jacobi_lds_n", "eig_graphs", "serve_large_batch", "tri_max", "pairs_block")

    def tunables(self) -> dict:
        """The context's device-dependent launch choices (``am1g_get_tunable``)."""
        out = {}
        for k in self.TUNABLES:
            v = ctypes.c_double()
            self._L.am1g_get_tunable(self._ctx, k.encode(), ctypes.byref(v))
            out[k] = v.value
        return out

    def set_tunable(self, name: str, value: float) -> None:
        st = self._L.am1g_set_tunable(self._ctx, name.encode(), float(value))
        if st != 0:
            raise ValueError(self.last_error())

    def eigh(self, a):
        """``symmetric_eigen`` on the device: (eigenvalues ascending, eigenvectors as columns)."""
        a = np.ascontiguousarray(a, dtype=np.float64)
        n = a.shape[0]
        w = np.zeros(n)
        v = np.zeros((n, n))
        dp = ctypes.POINTER(ctypes.c_double)
        st = self._L.am1g_symmetric_eigen(self._ctx, n, a.ctypes.data_as(dp), w.ctypes.data_as(dp),
                                          v.ctypes.data_as(dp))
        if st != 0:
            raise ValueError(f"{STATUS_NAMES.get(st, st)}: {self.last_error()}")
        return w, v

    def matmul(self, a, b, c=None, trans_a=False, trans_b=False, alpha=1.0, beta=0.0, precision="fp64"):
        """``alpha op(a) op(b) + beta c`` on the device (``Matrix::matmul`` and its transposed and
        accumulating forms); returns the product as a new array."""
        a = np.ascontiguousarray(a, dtype=np.float64)
        b = np.ascontiguousarray(b, dtype=np.float64)
        m, k = (a.shape[1], a.shape[0]) if trans_a else a.shape
        kb, n = (b.shape[1], b.shape[0]) if trans_b else b.shape
        if kb != k:
            raise ValueError(f"matmul: inner dimensions differ ({k} and {kb})")
        out = np.zeros((m, n)) if c is None else np.array(c, dtype=np.float64, order="C", copy=True)
        if out.shape != (m, n):
            raise ValueError("matmul: c has the wrong shape")
        dp = ctypes.POINTER(ctypes.c_double)
        st = self._L.am1g_matmul(self._ctx, int(trans_a), int(trans_b), m, n, k, float(alpha), a.ctypes.data_as(dp),
                                 b.ctypes.data_as(dp), float(beta), out.ctypes.data_as(dp), PRECISIONS[precision])
        if st != 0:
            raise ValueError(f"{STATUS_NAMES.get(st, st)}: {self.last_error()}")
        return out

    # -- request marshalling -------------------------------------------------------------
    def _request(self, z, xyz_bohr, keep, charge=0.0, multiplicity=1, method="am1", reference="auto",
                 accelerator="adiis_cdiis", use_diis=True, max_scf=200, e_tol=1e-8, p_tol=1e-7,
                 adiis_switch=0.1, precision="fp64", electric_field=None, bcc=False, multipole_cutoff=None):
        """``electric_field``: three floats in eV per (e*Bohr) (Am1Options::electric_field) or None.
        ``multipole_cutoff``: Bohr (Am1Options::multipole_cutoff), None for every pair exact."""
        rq = Request()
        self._L.am1g_request_defaults(ctypes.byref(rq))
        zi = np.asarray(z)
        if zi.size and (not np.issubdtype(zi.dtype, np.integer) or zi.min() < 1 or zi.max() > 255):
            # an out-of-range atomic number would wrap in the uint8 cast and name another element
            raise ValueError(f"atomic numbers must be integers in 1..255, got {zi.ravel().tolist()[:8]}")
        za = np.ascontiguousarray(zi, dtype=np.uint8)
        xa = np.ascontiguousarray(xyz_bohr, dtype=np.float64).reshape(-1, 3)
        keep.extend([za, xa])
        rq.n_atoms = len(za)
        rq.z = za.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8))
        rq.xyz_bohr = xa.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        rq.charge = float(charge)
        rq.multiplicity = int(multiplicity)
        rq.method = METHODS[method] if isinstance(method, str) else int(method)
        rq.reference = REFERENCES[reference] if isinstance(reference, str) else int(reference)
        rq.accelerator = ACCELERATORS[accelerator] if isinstance(accelerator, str) else int(accelerator)
        rq.use_diis = 1 if use_diis else 0
        rq.max_scf = int(max_scf)
        rq.e_tol = float(e_tol)
        rq.p_tol = float(p_tol)
        rq.adiis_switch = float(adiis_switch)
        rq.precision = PRECISIONS[precision] if isinstance(precision, str) else int(precision)
        rq.bcc = 1 if bcc else 0
        rq.multipole_cutoff = float(multipole_cutoff) if multipole_cutoff is not None else 0.0
        if electric_field is not None:
            fa = np.ascontiguousarray(electric_field, dtype=np.float64).reshape(-1)
            if fa.size != 3:
                raise ValueError("electric_field must have three components")
            keep.append(fa)
            rq.electric_field = fa.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        return rq

    @staticmethod
    def _attach(rs, n_atoms, full, keep):
        q = np.zeros(n_atoms)
        qb = np.zeros(n_atoms)
        keep.extend([q, qb])
        rs.charges = q.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        rs.bcc_charges = qb.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        arrays = {"charges": q, "bcc": qb}
        if full:
            nao_max = 4 * n_atoms
            mo = np.zeros(nao_max)
            mob = np.zeros(nao_max)
            dens = np.zeros(nao_max * nao_max)
            coeff = np.zeros(nao_max * nao_max)
            keep.extend([mo, mob, dens, coeff])
            rs.mo_energies = mo.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
            rs.mo_energies_beta = mob.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
            rs.density = dens.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
            rs.mo_coeff = coeff.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
            coeff_b = np.zeros(nao_max * nao_max)
            spin = np.zeros(nao_max * nao_max)
            keep.extend([coeff_b, spin])
            rs.mo_coeff_beta = coeff_b.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
            rs.spin_density = spin.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
            arrays.update(mo=mo, mob=mob, dens=dens, coeff=coeff, coeff_b=coeff_b, spin=spin)
        return arrays

    @staticmethod
    def _to_dict(rs, arrays, full):
        d = {
            "status": rs.status,
            "status_name": STATUS_NAMES.get(rs.status, str(rs.status)),
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
            "homo_ev": rs.homo_ev if rs.has_homo else None,
            "lumo_ev": rs.lumo_ev if rs.has_lumo else None,
            "homo_beta_ev": rs.homo_beta_ev if rs.has_homo_beta else None,
            "lumo_beta_ev": rs.lumo_beta_ev if rs.has_lumo_beta else None,
            "dipole_debye": np.array(list(rs.dipole_debye)),
            "dipole_magnitude": rs.dipole_magnitude,
            "field_nuclear_ev": rs.field_nuclear_ev,
            "residual": rs.residual,
            "fp32_iterations": rs.fp32_iterations,
            "latency_ms": rs.latency_ms,
            "phase_ms": dict(zip(PHASES, list(rs.phase_ms))),
            "eigen_sweeps": rs.eigen_sweeps,
            "charges": arrays["charges"].copy(),
            "bcc_charges": arrays["bcc"].copy() if "bcc" in arrays else None,
            "bcc_warnings": rs.bcc_warnings,
        }
        if full and rs.status in (0, 4):
            n = rs.nao
            d["mo_energies"] = arrays["mo"][:n].copy()
            d["mo_energies_beta"] = arrays["mob"][:n].copy() if rs.unrestricted else None
            d["density"] = arrays["dens"][: n * n].reshape(n, n).copy()
            d["mo_coeff"] = arrays["coeff"][: n * n].reshape(n, n).copy()
            d["mo_coeff_beta"] = arrays["coeff_b"][: n * n].reshape(n, n).copy() if rs.unrestricted else None
            d["spin_density"] = arrays["spin"][: n * n].reshape(n, n).copy() if rs.unrestricted else None
        return d

    def scf(self, z, xyz_bohr, full: bool = True, **kw) -> dict:
        """One SCF on the plain path (``am1g_scf``)."""
        keep = []
        rq = self._request(z, xyz_bohr, keep, **kw)
        rs = Result()
        arrays = self._attach(rs, rq.n_atoms, full, keep)
        self._L.am1g_scf(self._ctx, ctypes.byref(rq), ctypes.byref(rs))
        d = self._to_dict(rs, arrays, full)
        if rs.status not in (0, 4):
            d["message"] = self.last_error()
        return d

    def gradient(self, z, xyz_bohr, **kw) -> dict:
        """SCF plus nuclear gradient (``am1g_gradient``): the ``scf`` dict (without matrices) with
        ``gradient`` [n][3] in eV/Bohr."""
        keep = []
        rq = self._request(z, xyz_bohr, keep, **kw)
        rs = Result()
        arrays = self._attach(rs, rq.n_atoms, False, keep)
        g = np.zeros((rq.n_atoms, 3))
        self._L.am1g_gradient(self._ctx, ctypes.byref(rq), ctypes.byref(rs), g.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
        d = self._to_dict(rs, arrays, False)
        d["gradient"] = g
        if rs.status not in (0, 4):
            d["message"] = self.last_error()
        return d

    def core_hamiltonian(self, z, xyz_bohr, with_h: bool = True, with_pairs: bool = True, **kw) -> dict:
        """The device's core Hamiltonian (``am1g_core_hamiltonian``): ``h_core`` [nao][nao],
        ``core_ev`` and ``pairs``, a dict {(a, b): w} of every pair's two-electron block with atom
        indices in input order, a the heavy atom (or the lower index of two like atoms) and w of
        shape (10, 10), (10, 1) or (1, 1). ``with_h`` / ``with_pairs`` False skip copying those
        out (``h_core`` / ``pairs`` are then None)."""
        keep = []
        rq = self._request(z, xyz_bohr, keep, **kw)
        zz = np.asarray(z, dtype=np.int64).ravel()
        heavy = [a for a in range(len(zz)) if zz[a] != 1]
        hyd = [a for a in range(len(zz)) if zz[a] == 1]
        nh, nH = len(heavy), len(hyd)
        nao = 4 * nh + nH
        w_len = 100 * nh * (nh - 1) // 2 + 10 * nh * nH + nH * (nH - 1) // 2
        H = np.zeros((nao, nao)) if with_h else None
        w = np.zeros(max(w_len, 1)) if with_pairs else None
        ec = ctypes.c_double(0.0)
        dp = ctypes.POINTER(ctypes.c_double)
        st = self._L.am1g_core_hamiltonian(self._ctx, ctypes.byref(rq), H.ctypes.data_as(dp) if with_h else None,
                                           ctypes.byref(ec), w.ctypes.data_as(dp) if with_pairs else None,
                                           len(w) if with_pairs else 0)
        if st != 0:
            raise ValueError(f"{STATUS_NAMES.get(st, st)}: {self.last_error()}")
        if not with_pairs:
            return {"h_core": H, "core_ev": ec.value, "pairs": None}
        pairs = {}
        k = 0
        for i in range(nh):
            for j in range(i + 1, nh):
                pairs[(heavy[i], heavy[j])] = w[k:k + 100].reshape(10, 10).copy()
                k += 100
        for i in range(nh):
            for j in range(nH):
                pairs[(heavy[i], hyd[j])] = w[k:k + 10].reshape(10, 1).copy()
                k += 10
        for i in range(nH):
            for j in range(i + 1, nH):
                pairs[(hyd[i], hyd[j])] = w[k:k + 1].reshape(1, 1).copy()
                k += 1
        return {"h_core": H, "core_ev": ec.value, "pairs": pairs}

    def scf_batch(self, molecules, full: bool = False, noalloc: bool = False, **kw) -> list:
        """Many SCFs in one persistent launch (``am1g_scf_batch``; ``am1g_scf_batch_noalloc`` with
        ``noalloc``). ``molecules`` is a list of ``(z, xyz_bohr)`` or
        ``(z, xyz_bohr, per_molecule_kwargs)``."""
        keep = []
        n = len(molecules)
        rqs = (Request * n)()
        rss = (Result * n)()
        arrays = []
        for i, mol in enumerate(molecules):
            z, xyz = mol[0], mol[1]
            opts = dict(kw)
            if len(mol) > 2 and mol[2]:
                opts.update(mol[2])
            rqs[i] = self._request(z, xyz, keep, **opts)
            arrays.append(self._attach(rss[i], len(z), full, keep))
        t0 = time.perf_counter()
        entry = self._L.am1g_scf_batch_noalloc if noalloc else self._L.am1g_scf_batch
        st = entry(self._ctx, n, rqs, rss)
        # the C call alone (host tables, copies, launches, result fill): what a native caller waits
        self.last_call_ms = (time.perf_counter() - t0) * 1e3
        if st != 0:
            raise RuntimeError(f"am1g_scf_batch failed ({STATUS_NAMES.get(st, st)}): {self.last_error()}")
        return [self._to_dict(rss[i], arrays[i], full) for i in range(n)]

    def dc(self, z, xyz_bohr, core_size=None, buffer_radius=None, kt=None, max_scf=None, e_tol=None,
           p_tol=None, mixing=None, core_of_atom=None, density=False, **kw) -> dict:
        """Divide-and-conquer SCF (``am1g_dc``; run_divide_conquer, divide_conquer.rs:406-572).
        ``buffer_radius`` in Bohr, ``kt`` in eV (``kt <= 0`` is aufbau filling); options left
        as None take the library defaults. ``core_of_atom`` gives caller-chosen cores."""
        keep = []
        rq = self._request(z, xyz_bohr, keep, **kw)
        op = DcOptions()
        self._L.am1g_dc_defaults(ctypes.byref(op))
        for k, v in dict(core_size=core_size, buffer_radius=buffer_radius, kt=kt, max_scf=max_scf, e_tol=e_tol,
                         p_tol=p_tol, mixing=mixing).items():
            if v is not None:
                setattr(op, k, v)
        if core_of_atom is not None:
            ca = np.ascontiguousarray(core_of_atom, dtype=np.int32)
            keep.append(ca)
            op.core_of_atom = ca.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
        rs = DcResult()
        q = np.zeros(rq.n_atoms)
        rs.charges = q.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        P = S = None
        if density:
            nao = sum(1 if int(v) == 1 else 4 for v in np.asarray(z).ravel())
            P = np.zeros(nao * nao)
            S = np.zeros(nao * nao)
            rs.density = P.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
            rs.spin_density = S.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        self._L.am1g_dc(self._ctx, ctypes.byref(rq), ctypes.byref(op), ctypes.byref(rs))
        out = {k: getattr(rs, k) for k, _ in DcResult._fields_
               if k not in ("charges", "density", "spin_density", "phase_ms")}
        out["converged"] = bool(rs.converged)
        out["status_name"] = STATUS_NAMES.get(rs.status, str(rs.status))
        out["phase_ms"] = dict(zip(DC_PHASES, list(rs.phase_ms)))
        out["wall_ms"] = self.last_kernel_ms()
        out["charges"] = q
        if density:
            n = int(round(len(P) ** 0.5))
            out["density"] = P.reshape(n, n)
            out["spin_density"] = S.reshape(n, n) if rs.unrestricted else None
        if rs.status not in (0, 4):
            out["message"] = self.last_error()
        return out

    def gradient_density(self, z, xyz_bohr, density, spin_density=None, **kw):
        """Fixed-density nuclear gradient (``am1g_gradient_density``) at a given total density and,
        for an open shell, spin density P_alpha - P_beta: [n][3] in eV/Bohr."""
        keep = []
        rq = self._request(z, xyz_bohr, keep, **kw)
        p = np.ascontiguousarray(density, dtype=np.float64)
        sp = None if spin_density is None else np.ascontiguousarray(spin_density, dtype=np.float64)
        g = np.zeros((rq.n_atoms, 3))
        dp = ctypes.POINTER(ctypes.c_double)
        st = self._L.am1g_gradient_density(self._ctx, ctypes.byref(rq), p.ctypes.data_as(dp),
                                           sp.ctypes.data_as(dp) if sp is not None else None, g.ctypes.data_as(dp))
        if st != 0:
            raise ValueError(f"{STATUS_NAMES.get(st, st)}: {self.last_error()}")
        return g


class Server:
    """Continuous-batching server (``am1g_serve_*``): ``submit`` queues a request and returns a
    ticket; ``poll`` returns None while it runs, else the result dict (``latency_ms`` = submit to
    completion)."""

    def __init__(self, device: int = 0, lanes: int | None = None, max_batch: int | None = None,
                 memory_budget_mb: float | None = None, max_wait_ms: float | None = None,
                 continuous: bool | None = None):
        """``continuous`` (default True): one persistent service launch takes requests as they
        arrive; False: lanes of batched launches (``lanes``, ``max_batch``, ``max_wait_ms``)."""
        self._L = load()
        op = ServeOptions()
        self._L.am1g_serve_defaults(ctypes.byref(op))
        if lanes is not None:
            op.lanes = lanes
        if max_batch is not None:
            op.max_batch = max_batch
        if memory_budget_mb is not None:
            op.memory_budget_mb = memory_budget_mb
        if max_wait_ms is not None:
            op.max_wait_ms = max_wait_ms
        if continuous is not None:
            op.continuous = int(bool(continuous))
        self._sv = ctypes.c_void_p()
        st = self._L.am1g_serve_create(int(device), ctypes.byref(op), ctypes.byref(self._sv))
        if st != 0:
            msg = self._L.am1g_last_error(None)
            raise RuntimeError(f"am1g_serve_create failed ({STATUS_NAMES.get(st, st)}): "
                               f"{msg.decode() if msg else ''}")
        self._req = Engine.__new__(Engine)
        self._req._L = self._L
        self._n_atoms = {}

    def close(self):
        if self._sv:
            self._L.am1g_serve_destroy(self._sv)
            self._sv = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def submit(self, z, xyz_bohr, **kw) -> int:
        keep = []
        rq = self._req._request(z, xyz_bohr, keep, **kw)
        t = ctypes.c_longlong()
        st = self._L.am1g_serve_submit(self._sv, ctypes.byref(rq), ctypes.byref(t))
        if st != 0:
            raise ValueError(self._L.am1g_serve_last_error(self._sv).decode())
        self._n_atoms[t.value] = rq.n_atoms
        return t.value

    def poll(self, ticket: int, wait_ms: float = 0.0):
        keep = []
        rs = Result()
        n = self._n_atoms[ticket]
        q = np.zeros(n)
        qb = np.zeros(n)
        mo = np.zeros(4 * n)
        mob = np.zeros(4 * n)
        keep.extend([q, qb, mo, mob])
        rs.charges = q.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        rs.bcc_charges = qb.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        rs.mo_energies = mo.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        rs.mo_energies_beta = mob.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        st = self._L.am1g_serve_poll(self._sv, int(ticket), float(wait_ms), ctypes.byref(rs))
        if st == PENDING:
            return None
        del self._n_atoms[ticket]
        d = Engine._to_dict(rs, {"charges": q, "bcc": qb}, False)
        d["status"] = st
        d["status_name"] = STATUS_NAMES.get(st, str(st))
        if st in (0, 4):
            d["mo_energies"] = mo[: rs.nao].copy()
            d["mo_energies_beta"] = mob[: rs.nao].copy() if rs.unrestricted else None
        return d

    def cancel(self, ticket: int) -> None:
        self._L.am1g_serve_cancel(self._sv, int(ticket))

    def stats(self) -> dict:
        s = ServeStats()
        self._L.am1g_serve_get_stats(self._sv, ctypes.byref(s))
        return {k: getattr(s, k) for k, _ in ServeStats._fields_}


_default_engine = None


def engine() -> Engine:
    global _default_engine
    if _default_engine is None:
        _default_engine = Engine(0)
    return _default_engine


def _as_numbers(numbers):
    out = []
    for v in numbers:
        if isinstance(v, str):
            if v not in SYMBOLS:
                raise ValueError(f"unknown element symbol {v!r}")
            out.append(SYMBOLS[v])
        else:
            iv = int(v)
            if iv != v or not 1 <= iv <= 255:
                raise ValueError(f"atomic number {v!r} is not an integer in 1..255")
            out.append(iv)
    return out


def single_point(numbers, positions, charge=0.0, multiplicity=1, reference="auto", method="am1",
                 electric_field=None) -> dict:
    """``am1_rs.native.single_point`` counterpart (python/am1_rs/native.py:52-97 ->
    src/python.rs:147-199): same signature and positional order, positions in Angstrom, and the
    same result keys: ``method``, ``energy_hartree``, ``energy_ev``, ``heat_of_formation_kcal``,
    ``electronic_ev``, ``core_ev``, ``charges``, ``dipole_debye``, ``dipole_magnitude``,
    ``homo_ev``, ``lumo_ev``, ``homo_beta_ev``, ``lumo_beta_ev``, ``field_nuclear_ev``,
    ``converged``, ``iterations``, ``unrestricted``. ``electric_field`` is three floats in atomic
    units (Hartree per e*Bohr) or None. Every error, non-convergence included, raises
    ``ValueError`` (python.rs:27-29)."""
    z, pos, field = _molecular_input(numbers, positions, method, reference, electric_field)
    r = engine().scf(z, pos, full=False, charge=float(charge), multiplicity=max(1, int(multiplicity)),
                     method=str(method).lower(), reference=str(reference).lower(), electric_field=field)
    if r["status"] == 4:
        raise ValueError(f"SCF did not converge in {r['iterations']} iterations "
                         f"(residual {r['residual']:.3e})")
    if r["status"] != 0:
        raise ValueError(r.get("message", r["status_name"]))
    return {
        "method": str(method).lower(),
        "energy_hartree": r["total_ev"] / AM1_EV,
        "energy_ev": r["total_ev"],
        "heat_of_formation_kcal": r["heat_of_formation_kcal"],
        "electronic_ev": r["electronic_ev"],
        "core_ev": r["core_ev"],
        "charges": r["charges"].tolist(),
        "dipole_debye": r["dipole_debye"].tolist(),
        "dipole_magnitude": r["dipole_magnitude"],
        "homo_ev": r["homo_ev"],
        "lumo_ev": r["lumo_ev"],
        "homo_beta_ev": r["homo_beta_ev"],
        "lumo_beta_ev": r["lumo_beta_ev"],
        "field_nuclear_ev": r["field_nuclear_ev"],
        "converged": r["converged"],
        "iterations": r["iterations"],
        "unrestricted": r["unrestricted"],
    }


def _molecular_input(numbers, positions, method, reference, electric_field):
    """Validated (z, positions in Bohr, field in eV per e*Bohr or None) for the module-level
    functions, with the reference front end's checks and errors (python.rs:27-70)."""
    z = _as_numbers(numbers)
    pos = np.asarray(positions, dtype=np.float64)
    if pos.ndim != 2 or pos.shape[1] != 3:
        raise ValueError("each position must have 3 components")
    if len(z) != pos.shape[0]:
        raise ValueError("numbers and positions length mismatch")
    field = None
    if electric_field is not None:
        f = np.asarray(electric_field, dtype=float).reshape(-1)
        if f.size != 3:
            raise ValueError("electric_field must have three components")
        field = f * AM1_EV  # atomic units (Hartree per e*Bohr) -> eV per e*Bohr
    if str(method).lower() not in METHODS:
        raise ValueError(f"unknown method {method!r} (expected 'am1' or 'rm1')")
    if str(reference).lower() not in REFERENCES:
        raise ValueError(f"unknown reference {reference!r} (expected 'auto', 'rhf' or 'uhf')")
    return z, pos * ANGSTROM_TO_BOHR, field


def orbitals(numbers, positions, charge=0.0, multiplicity=1, reference="auto", method="am1",
             electric_field=None) -> dict:
    """``am1_rs`` ``orbitals`` counterpart (src/python.rs:1905-1992): orbital energies in Hartree
    and eV, the coefficient matrix (rows atomic orbitals, columns orbitals), the occupied count,
    HOMO/LUMO, both spin channels for an unrestricted run (``beta_*`` keys), the AO labels
    ``(atom, "s" | "px" | "py" | "pz")`` and the method."""
    z, pos, field = _molecular_input(numbers, positions, method, reference, electric_field)
    r = engine().scf(z, pos, full=True, charge=float(charge), multiplicity=max(1, int(multiplicity)),
                     method=str(method).lower(), reference=str(reference).lower(), electric_field=field)
    if r["status"] == 4:
        raise ValueError(f"SCF did not converge in {r['iterations']} iterations (residual {r['residual']:.3e})")
    if r["status"] != 0:
        raise ValueError(r.get("message", r["status_name"]))
    d = {
        "energies_hartree": (r["mo_energies"] / AM1_EV).tolist(),
        "energies_ev": r["mo_energies"].tolist(),
        "coefficients": r["mo_coeff"].tolist(),
        "n_occupied": r["n_occ"],
        "homo_ev": r["homo_ev"],
        "lumo_ev": r["lumo_ev"],
        "unrestricted": r["unrestricted"],
    }
    if r["unrestricted"]:
        d.update({
            "beta_energies_hartree": (r["mo_energies_beta"] / AM1_EV).tolist(),
            "beta_energies_ev": r["mo_energies_beta"].tolist(),
            "beta_coefficients": r["mo_coeff_beta"].tolist(),
            "beta_n_occupied": r["n_occ_beta"],
            "homo_beta_ev": r["homo_beta_ev"],
            "lumo_beta_ev": r["lumo_beta_ev"],
        })
    labels = []
    for a, zz in enumerate(z):
        labels.extend([(a, o) for o in (("s",) if int(zz) == 1 else ("s", "px", "py", "pz"))])
    d["ao_labels"] = labels
    d["method"] = str(method).lower()
    return d


def _block_nonzero(density, z):
    """Atom x atom mask of the density blocks that are not identically zero."""
    nz = np.asarray(density) != 0.0
    offs = np.concatenate([[0], np.cumsum([1 if int(v) == 1 else 4 for v in z])[:-1]])
    return np.logical_or.reduceat(np.logical_or.reduceat(nz, offs, axis=0), offs, axis=1)


def divide_conquer(numbers, positions, charge=0.0, multiplicity=1, reference="auto", method="am1",
                   core_size=12, buffer_radius=11.0, smearing_ev=0.05, e_tol=1.0e-7, p_tol=1.0e-6, max_scf=300,
                   mixing=0.4, gap_warn_ev=0.5, forces=True, multipole_cutoff=None, electric_field=None, cell=None,
                   pbc=None, realspace_cutoff=40.0, exchange_cutoff=20.0, optimize=False, opt_max_iter=200,
                   opt_gtol=1.0e-3) -> dict:
    """``am1_rs`` ``divide_conquer`` counterpart (src/python.rs:448-663): the same signature and
    defaults (positions in Angstrom, ``buffer_radius`` and ``multipole_cutoff`` in Bohr,
    ``smearing_ev`` the Fermi-Dirac kT, 0 for aufbau filling) and the same result keys, forces in
    eV/Angstrom from the fixed-density gradient at the divide-and-conquer density
    (divide_conquer.rs:1448-1515). Non-convergence raises ``ValueError``. Periodic cells
    (``cell``/``pbc``) and geometry optimization (``optimize``) are outside the GPU engine and
    raise ``NotImplementedError``; ``realspace_cutoff``, ``exchange_cutoff``, ``opt_max_iter`` and
    ``opt_gtol`` only apply to those."""
    if (cell is None) != (pbc is None):
        raise ValueError("cell and pbc must be given together")
    if cell is not None:
        raise NotImplementedError("periodic divide-and-conquer is not part of the am1g GPU engine")
    if optimize:
        raise NotImplementedError("divide-and-conquer geometry optimization is not part of the am1g GPU engine")
    z, pos, field = _molecular_input(numbers, positions, method, reference, electric_field)
    eng = engine()
    kw = dict(charge=float(charge), multiplicity=max(1, int(multiplicity)), method=str(method).lower(),
              reference=str(reference).lower(), electric_field=field, multipole_cutoff=multipole_cutoff)
    r = eng.dc(z, pos, core_size=int(core_size), buffer_radius=float(buffer_radius), kt=float(smearing_ev),
               max_scf=int(max_scf), e_tol=float(e_tol), p_tol=float(p_tol), mixing=float(mixing), density=True, **kw)
    if r["status"] not in (0, 4):
        raise ValueError(r.get("message", r["status_name"]))
    if not r["converged"]:
        raise ValueError(f"divide-and-conquer SCF did not converge in {r['iterations']} iterations")
    unrestricted = bool(r["unrestricted"])
    gap = r["homo_lumo_gap_ev"]
    warning = None
    if gap < gap_warn_ev:
        warning = (f"smallest subsystem gap {gap:.3f} eV is below gap_warn_ev = {gap_warn_ev:.3f} eV: the density "
                   "of a small-gap system decays slowly with distance, so the buffer may not contain what the "
                   "cores need; compare with a full SCF or enlarge buffer_radius")
    blocks = _block_nonzero(r["density"], z)
    iu = np.triu_indices(len(z), 1)
    near = np.ones(len(iu[0]), dtype=bool)
    if multipole_cutoff is not None and multipole_cutoff > 0:
        dist = np.linalg.norm(pos[iu[1]] - pos[iu[0]], axis=1)
        near = dist <= multipole_cutoff
    d = {
        "energy_ev": r["total_ev"],
        "energy_hartree": r["total_ev"] / AM1_EV,
        "free_energy_ev": r["total_ev"] - r["entropy_ev"],
        "heat_of_formation_kcal": r["heat_of_formation_kcal"],
        "electronic_ev": r["electronic_ev"],
        "core_ev": r["core_ev"],
        "converged": True,
        "positions_angstrom": (pos / ANGSTROM_TO_BOHR).tolist(),
    }
    if forces:
        g = eng.gradient_density(z, pos, r["density"], r["spin_density"] if unrestricted else None, **kw)
        d["forces_ev_per_angstrom"] = (-g * ANGSTROM_TO_BOHR).tolist()
    d.update({
        "charges": np.asarray(r["charges"]).tolist(),
        "fermi_energy_ev": r["fermi_ev"],
        "fermi_energies_ev": [r["fermi_ev"], r["fermi_beta_ev"]] if unrestricted else [r["fermi_ev"]],
        "entropy_ev": r["entropy_ev"],
        "iterations": r["iterations"],
        "unrestricted": unrestricted,
        "homo_lumo_gap_ev": gap,
        "small_gap_warning": warning,
        "subsystems": r["n_subsystems"],
        "largest_subsystem_aos": r["largest_subsystem_aos"],
        "diagonalization_work": r["diagonalization_work"],
        "coulomb_work": int(near.sum()),
        "exchange_work": int((blocks[iu] & near).sum()),
        "retained_density_blocks": int(blocks.sum()),
        "diis_pattern_elements": int(r["retained_pattern"]),
        "dense_triangle_elements": int(len(r["density"]) * (len(r["density"]) + 1) // 2),
        "method": str(method).lower(),
    })
    return d


def bcc_correct(numbers, xyz_bohr, mulliken) -> dict:
    """AM1-BCC corrections (``am1g_bcc_charges``, host code) applied to given AM1 Mulliken charges:
    ``charges``, ``atom_types`` (antechamber codes as strings, "0" = none) and ``warnings``."""
    L = load()
    z = np.ascontiguousarray(_as_numbers(numbers), dtype=np.uint8)
    x = np.ascontiguousarray(xyz_bohr, dtype=np.float64).reshape(-1, 3)
    q = np.ascontiguousarray(mulliken, dtype=np.float64).reshape(-1)
    if not (len(z) == len(x) == len(q)):
        raise ValueError("numbers, positions and charges differ in length")
    out = np.zeros(len(z))
    types = np.zeros(len(z), dtype=np.int32)
    nw = ctypes.c_int(0)
    buf = ctypes.create_string_buffer(8192)
    st = L.am1g_bcc_charges(len(z), z.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
                            x.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                            q.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                            out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                            types.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), ctypes.byref(nw), buf, len(buf))
    if st != 0:
        raise ValueError(f"am1g_bcc_charges failed ({STATUS_NAMES.get(st, st)})")
    warnings = [w for w in buf.value.decode(errors="replace").split("\n") if w]
    return {"charges": out.tolist(), "atom_types": [str(int(t)) for t in types], "warnings": warnings}


def am1_bcc(numbers, positions, charge=0.0, multiplicity=1) -> dict:
    """``am1_rs.native.am1_bcc`` counterpart (python/am1_rs/native.py:834-857, src/python.rs:1785-1815):
    the AM1 SCF on the GPU for the Mulliken charges, then the AM1-BCC corrections. Positions in
    Angstrom. Returns ``charges``, ``mulliken``, ``atom_types`` and ``warnings``."""
    z = _as_numbers(numbers)
    pos = np.asarray(positions, dtype=np.float64)
    if pos.ndim != 2 or pos.shape[1] != 3 or pos.shape[0] != len(z):
        raise ValueError("numbers and positions length mismatch")
    xyz = pos * ANGSTROM_TO_BOHR
    r = engine().scf(z, xyz, full=False, charge=float(charge), multiplicity=max(1, int(multiplicity)), method="am1")
    if r["status"] == 4:
        raise ValueError(f"SCF did not converge in {r['iterations']} iterations (residual {r['residual']:.3e})")
    if r["status"] != 0:
        raise ValueError(r.get("message", r["status_name"]))
    out = bcc_correct(z, xyz, r["charges"])
    out["mulliken"] = r["charges"].tolist()
    return out


def gradient(numbers, positions, charge=0.0, multiplicity=1, reference="auto", method="am1",
             electric_field=None) -> dict:
    """``am1_rs.native.gradient`` counterpart (src/python.rs:202-275): the SCF and its nuclear
    gradient in one call, positions in Angstrom. Keys as the reference: ``energy_hartree``,
    ``energy_ev``, ``heat_of_formation_kcal``, ``electronic_ev``, ``core_ev``, ``homo_ev``,
    ``lumo_ev``, ``gradient_hartree_per_bohr``, ``gradient_ev_per_angstrom``, ``charges``,
    ``dipole_debye``, ``dipole_magnitude``, ``max_gradient_hartree_per_bohr``, ``converged``,
    ``iterations``, ``unrestricted``, ``method``. Errors raise ``ValueError``."""
    z, pos, field = _molecular_input(numbers, positions, method, reference, electric_field)
    r = engine().gradient(z, pos, charge=float(charge), multiplicity=max(1, int(multiplicity)),
                          method=str(method).lower(), reference=str(reference).lower(), electric_field=field)
    if r["status"] == 4:
        raise ValueError(f"SCF did not converge in {r['iterations']} iterations (residual {r['residual']:.3e})")
    if r["status"] != 0:
        raise ValueError(r.get("message", r["status_name"]))
    g = r["gradient"]  # eV/Bohr
    return {
        "energy_hartree": r["total_ev"] / AM1_EV,
        "energy_ev": r["total_ev"],
        "heat_of_formation_kcal": r["heat_of_formation_kcal"],
        "electronic_ev": r["electronic_ev"],
        "core_ev": r["core_ev"],
        "homo_ev": r["homo_ev"],
        "lumo_ev": r["lumo_ev"],
        "gradient_hartree_per_bohr": (g / AM1_EV).tolist(),
        "gradient_ev_per_angstrom": (g * ANGSTROM_TO_BOHR).tolist(),
        "charges": r["charges"].tolist(),
        "dipole_debye": r["dipole_debye"].tolist(),
        "dipole_magnitude": r["dipole_magnitude"],
        "max_gradient_hartree_per_bohr": float(np.abs(g).max()) / AM1_EV if g.size else 0.0,
        "converged": r["converged"],
        "iterations": r["iterations"],
        "unrestricted": r["unrestricted"],
        "method": str(method).lower(),
    }
