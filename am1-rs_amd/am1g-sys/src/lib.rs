//! Raw bindings to `libam1g` (include/am1g.h) and a thin safe layer.
//!
//! am1-rs forbids `unsafe` in its own crate (`lib.rs:39`), so the `extern "C"` calls live here.
//! The `raw` items mirror the C header field for field (`#[repr(C)]`, same order and types); the
//! `Context` type owns one `am1g_ctx` and turns status codes into `Error` values that carry the
//! same variants as am1-rs's `Am1Error` (`error.rs:213-293`).
#![allow(non_camel_case_types)]

pub mod raw {
    use std::os::raw::{c_char, c_double, c_int, c_longlong};

    pub const AM1G_OK: c_int = 0;
    pub const AM1G_INVALID_INPUT: c_int = 1;
    pub const AM1G_ELEMENT_NOT_PARAMETERIZED: c_int = 2;
    pub const AM1G_LINEAR_ALGEBRA: c_int = 3;
    pub const AM1G_SCF_NOT_CONVERGED: c_int = 4;
    pub const AM1G_SCF_DIVERGED: c_int = 5;
    pub const AM1G_DEVICE_ERROR: c_int = 6;
    pub const AM1G_PENDING: c_int = 7;
    pub const AM1G_CANCELLED: c_int = 8;

    #[repr(C)]
    pub struct am1g_ctx {
        _private: [u8; 0],
    }

    #[repr(C)]
    #[derive(Clone, Copy)]
    pub struct am1g_request {
        pub n_atoms: c_int,
        pub z: *const u8,
        pub xyz_bohr: *const c_double,
        pub charge: c_double,
        pub multiplicity: c_int,
        pub method: c_int,
        pub reference: c_int,
        pub accelerator: c_int,
        pub use_diis: c_int,
        pub max_scf: c_int,
        pub e_tol: c_double,
        pub p_tol: c_double,
        pub adiis_switch: c_double,
        pub precision: c_int,
        pub electric_field: *const c_double,
        pub bcc: c_int,
        pub multipole_cutoff: c_double,
    }

    #[repr(C)]
    #[derive(Clone, Copy)]
    pub struct am1g_result {
        pub status: c_int,
        pub nao: c_int,
        pub n_occ: c_int,
        pub n_occ_beta: c_int,
        pub iterations: c_int,
        pub converged: c_int,
        pub unrestricted: c_int,
        pub has_homo: c_int,
        pub has_lumo: c_int,
        pub has_homo_beta: c_int,
        pub has_lumo_beta: c_int,
        pub electronic_ev: c_double,
        pub core_ev: c_double,
        pub total_ev: c_double,
        pub heat_of_formation_kcal: c_double,
        pub homo_ev: c_double,
        pub lumo_ev: c_double,
        pub homo_beta_ev: c_double,
        pub lumo_beta_ev: c_double,
        pub dipole_debye: [c_double; 3],
        pub dipole_magnitude: c_double,
        pub field_nuclear_ev: c_double,
        pub residual: c_double,
        pub fp32_iterations: c_int,
        pub latency_ms: c_double,
        pub phase_ms: [c_double; 6],
        pub eigen_sweeps: c_int,
        pub bcc_warnings: c_int,
        pub charges: *mut c_double,
        pub mo_energies: *mut c_double,
        pub mo_energies_beta: *mut c_double,
        pub density: *mut c_double,
        pub mo_coeff: *mut c_double,
        pub mo_coeff_beta: *mut c_double,
        pub spin_density: *mut c_double,
        pub bcc_charges: *mut c_double,
    }

    #[repr(C)]
    #[derive(Clone, Copy)]
    pub struct am1g_dc_options {
        pub core_size: c_int,
        pub buffer_radius: c_double,
        pub kt: c_double,
        pub max_scf: c_int,
        pub e_tol: c_double,
        pub p_tol: c_double,
        pub mixing: c_double,
        pub core_of_atom: *const c_int,
    }

    #[repr(C)]
    #[derive(Clone, Copy)]
    pub struct am1g_dc_result {
        pub status: c_int,
        pub iterations: c_int,
        pub converged: c_int,
        pub unrestricted: c_int,
        pub n_subsystems: c_int,
        pub largest_subsystem_aos: c_int,
        pub electronic_ev: c_double,
        pub core_ev: c_double,
        pub total_ev: c_double,
        pub heat_of_formation_kcal: c_double,
        pub fermi_ev: c_double,
        pub fermi_beta_ev: c_double,
        pub entropy_ev: c_double,
        pub homo_lumo_gap_ev: c_double,
        pub diagonalization_work: c_double,
        pub retained_pattern: c_longlong,
        pub phase_ms: [c_double; 5],
        pub charges: *mut c_double,
        pub density: *mut c_double,
        pub spin_density: *mut c_double,
    }

    #[repr(C)]
    pub struct am1g_server {
        _private: [u8; 0],
    }

    #[repr(C)]
    #[derive(Clone, Copy)]
    pub struct am1g_serve_options {
        pub lanes: c_int,
        pub max_batch: c_int,
        pub memory_budget_mb: c_double,
        pub max_wait_ms: c_double,
        pub continuous: c_int,
    }

    #[repr(C)]
    #[derive(Clone, Copy)]
    pub struct am1g_serve_stats {
        pub submitted: c_longlong,
        pub completed: c_longlong,
        pub cancelled: c_longlong,
        pub queued: c_longlong,
        pub running: c_longlong,
        pub launches: c_longlong,
        pub admitted_mb: c_double,
        pub peak_admitted_mb: c_double,
    }

    extern "C" {
        pub fn am1g_request_defaults(rq: *mut am1g_request);
        pub fn am1g_create(device: c_int, out: *mut *mut am1g_ctx) -> c_int;
        pub fn am1g_destroy(ctx: *mut am1g_ctx);
        pub fn am1g_scf(ctx: *mut am1g_ctx, rq: *const am1g_request, res: *mut am1g_result) -> c_int;
        pub fn am1g_scf_batch(ctx: *mut am1g_ctx, n: c_int, rqs: *const am1g_request, results: *mut am1g_result)
            -> c_int;
        pub fn am1g_scf_batch_noalloc(ctx: *mut am1g_ctx, n: c_int, rqs: *const am1g_request,
                                      results: *mut am1g_result) -> c_int;
        pub fn am1g_noalloc_capacity_bytes() -> c_longlong;
        pub fn am1g_dc_defaults(op: *mut am1g_dc_options);
        pub fn am1g_dc(ctx: *mut am1g_ctx, rq: *const am1g_request, op: *const am1g_dc_options,
                       res: *mut am1g_dc_result) -> c_int;
        pub fn am1g_serve_defaults(op: *mut am1g_serve_options);
        pub fn am1g_serve_create(device: c_int, op: *const am1g_serve_options, out: *mut *mut am1g_server)
            -> c_int;
        pub fn am1g_serve_destroy(sv: *mut am1g_server);
        pub fn am1g_serve_submit(sv: *mut am1g_server, rq: *const am1g_request, ticket: *mut c_longlong) -> c_int;
        pub fn am1g_serve_poll(sv: *mut am1g_server, ticket: c_longlong, wait_ms: c_double, out: *mut am1g_result)
            -> c_int;
        pub fn am1g_serve_cancel(sv: *mut am1g_server, ticket: c_longlong) -> c_int;
        pub fn am1g_serve_get_stats(sv: *mut am1g_server, out: *mut am1g_serve_stats) -> c_int;
        pub fn am1g_serve_last_error(sv: *const am1g_server) -> *const c_char;
        pub fn am1g_symmetric_eigen(ctx: *mut am1g_ctx, n: c_int, a: *const c_double, eigenvalues: *mut c_double,
                                    eigenvectors: *mut c_double) -> c_int;
        pub fn am1g_matmul(ctx: *mut am1g_ctx, trans_a: c_int, trans_b: c_int, m: c_int, n: c_int, k: c_int,
                           alpha: c_double, a: *const c_double, b: *const c_double, beta: c_double, c: *mut c_double,
                           precision: c_int) -> c_int;
        pub fn am1g_last_error(ctx: *const am1g_ctx) -> *const c_char;
        pub fn am1g_element_info(z: c_int, method: c_int, out: *mut c_double) -> c_int;
        pub fn am1g_last_kernel_ms(ctx: *const am1g_ctx) -> c_double;
        pub fn am1g_set_tunable(ctx: *mut am1g_ctx, name: *const c_char, value: c_double) -> c_int;
        pub fn am1g_get_tunable(ctx: *const am1g_ctx, name: *const c_char, value: *mut c_double) -> c_int;
        pub fn am1g_gradient(ctx: *mut am1g_ctx, rq: *const am1g_request, res: *mut am1g_result,
                             gradient: *mut c_double) -> c_int;
        pub fn am1g_gradient_density(ctx: *mut am1g_ctx, rq: *const am1g_request, density: *const c_double,
                                     spin_density: *const c_double, gradient: *mut c_double) -> c_int;
        pub fn am1g_core_hamiltonian(ctx: *mut am1g_ctx, rq: *const am1g_request, h_core: *mut c_double,
                                     core_ev: *mut c_double, pair_blocks: *mut c_double,
                                     pair_blocks_len: c_longlong) -> c_int;
        pub fn am1g_bcc_charges(n_atoms: c_int, z: *const u8, xyz_bohr: *const c_double, mulliken: *const c_double,
                                charges: *mut c_double, atom_types: *mut c_int, n_warnings: *mut c_int,
                                warnings: *mut c_char, warnings_len: c_int) -> c_int;
    }
}

use std::ffi::CStr;

/// The library's failures, one variant per status code (`Am1Error` of am1-rs, plus the device).
#[derive(Debug, Clone, PartialEq)]
pub enum Error {
    InvalidInput(String),
    ElementNotParameterized(String),
    LinearAlgebra(String),
    ScfNotConverged { iterations: usize, residual: f64 },
    ScfDiverged { iterations: usize },
    Device(String),
}

impl std::fmt::Display for Error {
    fn fmt(&self, f: &mut std::fmt::Formatter<'_>) -> std::fmt::Result {
        match self {
            Error::InvalidInput(m) | Error::ElementNotParameterized(m) | Error::LinearAlgebra(m) | Error::Device(m) => {
                f.write_str(m)
            }
            Error::ScfNotConverged { iterations, residual } => {
                write!(f, "SCF did not converge in {iterations} iterations (residual {residual:.3e})")
            }
            Error::ScfDiverged { iterations } => write!(f, "SCF diverged after {iterations} iterations"),
        }
    }
}

impl std::error::Error for Error {}

fn last_error(ctx: *const raw::am1g_ctx) -> String {
    // SAFETY: am1g_last_error returns a NUL-terminated thread-local string that stays valid until
    // the next failing call on this thread; it is copied at once.
    unsafe {
        let p = raw::am1g_last_error(ctx);
        if p.is_null() {
            String::new()
        } else {
            CStr::from_ptr(p).to_string_lossy().into_owned()
        }
    }
}

fn status_to_error(st: i32, res: &raw::am1g_result, ctx: *const raw::am1g_ctx) -> Error {
    match st {
        raw::AM1G_INVALID_INPUT => Error::InvalidInput(last_error(ctx)),
        raw::AM1G_ELEMENT_NOT_PARAMETERIZED => Error::ElementNotParameterized(last_error(ctx)),
        raw::AM1G_LINEAR_ALGEBRA => Error::LinearAlgebra(last_error(ctx)),
        raw::AM1G_SCF_NOT_CONVERGED => {
            Error::ScfNotConverged { iterations: res.iterations.max(0) as usize, residual: res.residual }
        }
        raw::AM1G_SCF_DIVERGED => Error::ScfDiverged { iterations: res.iterations.max(0) as usize },
        _ => Error::Device(last_error(ctx)),
    }
}

/// Options of one SCF (the Am1Options fields the molecular SCF reads, scf.rs:53-197).
#[derive(Debug, Clone)]
pub struct ScfOptions {
    pub charge: f64,
    pub multiplicity: u32,
    pub method: i32,
    pub reference: i32,
    pub accelerator: i32,
    pub use_diis: bool,
    pub max_scf: u32,
    pub e_tol: f64,
    pub p_tol: f64,
    pub adiis_switch: f64,
    pub mixed_precision: bool,
    /// eV per (e*Bohr)
    pub electric_field: Option<[f64; 3]>,
    /// Bohr; pairs farther apart are net-charge monopoles (Am1Options::multipole_cutoff)
    pub multipole_cutoff: Option<f64>,
}

impl Default for ScfOptions {
    fn default() -> Self {
        ScfOptions {
            charge: 0.0,
            multiplicity: 1,
            method: 0,
            reference: 0,
            accelerator: 2,
            use_diis: true,
            max_scf: 200,
            e_tol: 1.0e-8,
            p_tol: 1.0e-7,
            adiis_switch: 0.1,
            mixed_precision: false,
            electric_field: None,
            multipole_cutoff: None,
        }
    }
}

/// What am1-rs's Am1Result carries (scf.rs:200-240); matrices row-major, MOs in columns.
#[derive(Debug, Clone)]
pub struct ScfOutput {
    pub electronic_ev: f64,
    pub core_ev: f64,
    pub field_nuclear_ev: f64,
    pub total_ev: f64,
    pub heat_of_formation_kcal: f64,
    pub charges: Vec<f64>,
    pub dipole_debye: [f64; 3],
    pub dipole_magnitude: f64,
    pub homo_ev: Option<f64>,
    pub lumo_ev: Option<f64>,
    pub homo_beta_ev: Option<f64>,
    pub lumo_beta_ev: Option<f64>,
    pub mo_energies: Vec<f64>,
    pub mo_energies_beta: Option<Vec<f64>>,
    pub density: Vec<f64>,
    pub mo_coeff: Vec<f64>,
    pub iterations: usize,
    pub converged: bool,
    pub unrestricted: bool,
    pub nao: usize,
}

/// One library context on one GPU. Calls on a context are serialized inside the library, so it
/// may be shared between threads.
pub struct Context {
    ptr: *mut raw::am1g_ctx,
}

unsafe impl Send for Context {}
unsafe impl Sync for Context {}

impl Context {
    pub fn new(device: i32) -> Result<Self, Error> {
        let mut ptr = std::ptr::null_mut();
        // SAFETY: `ptr` is a valid out-pointer; on success the library hands over ownership.
        let st = unsafe { raw::am1g_create(device, &mut ptr) };
        if st != raw::AM1G_OK {
            return Err(Error::Device(last_error(std::ptr::null())));
        }
        Ok(Context { ptr })
    }

    /// run_am1 (scf.rs:294-298) on the device: `z` atomic numbers, `xyz_bohr` [n][3] in Bohr.
    pub fn scf(&self, z: &[u8], xyz_bohr: &[[f64; 3]], o: &ScfOptions) -> Result<ScfOutput, Error> {
        if z.len() != xyz_bohr.len() {
            return Err(Error::InvalidInput("atomic numbers and positions differ in length".into()));
        }
        let nao: usize = z.iter().map(|&v| if v == 1 { 1 } else { 4 }).sum();
        let flat: Vec<f64> = xyz_bohr.iter().flat_map(|p| p.iter().copied()).collect();
        let field = o.electric_field;
        // SAFETY: am1g_request is plain data; defaults fill every field before it is used.
        let mut rq: raw::am1g_request = unsafe {
            let mut r = std::mem::zeroed();
            raw::am1g_request_defaults(&mut r);
            r
        };
        rq.n_atoms = z.len() as i32;
        rq.z = z.as_ptr();
        rq.xyz_bohr = flat.as_ptr();
        rq.charge = o.charge;
        rq.multiplicity = o.multiplicity as i32;
        rq.method = o.method;
        rq.reference = o.reference;
        rq.accelerator = o.accelerator;
        rq.use_diis = o.use_diis as i32;
        rq.max_scf = o.max_scf as i32;
        rq.e_tol = o.e_tol;
        rq.p_tol = o.p_tol;
        rq.adiis_switch = o.adiis_switch;
        rq.precision = o.mixed_precision as i32;
        rq.electric_field = field.as_ref().map_or(std::ptr::null(), |f| f.as_ptr());
        rq.multipole_cutoff = o.multipole_cutoff.unwrap_or(0.0);
        let mut q = vec![0.0; z.len()];
        let mut eps = vec![0.0; nao];
        let mut eps_b = vec![0.0; nao];
        let mut p = vec![0.0; nao * nao];
        let mut c = vec![0.0; nao * nao];
        // SAFETY: zeroed pointers are NULL (skipped outputs) until set below.
        let mut res: raw::am1g_result = unsafe { std::mem::zeroed() };
        res.charges = q.as_mut_ptr();
        res.mo_energies = eps.as_mut_ptr();
        res.mo_energies_beta = eps_b.as_mut_ptr();
        res.density = p.as_mut_ptr();
        res.mo_coeff = c.as_mut_ptr();
        // SAFETY: every buffer lives until the call returns and has the size the header states.
        let st = unsafe { raw::am1g_scf(self.ptr, &rq, &mut res) };
        if st != raw::AM1G_OK {
            return Err(status_to_error(st, &res, self.ptr));
        }
        let opt = |has: i32, v: f64| if has != 0 { Some(v) } else { None };
        Ok(ScfOutput {
            electronic_ev: res.electronic_ev,
            core_ev: res.core_ev,
            field_nuclear_ev: res.field_nuclear_ev,
            total_ev: res.total_ev,
            heat_of_formation_kcal: res.heat_of_formation_kcal,
            charges: q,
            dipole_debye: res.dipole_debye,
            dipole_magnitude: res.dipole_magnitude,
            homo_ev: opt(res.has_homo, res.homo_ev),
            lumo_ev: opt(res.has_lumo, res.lumo_ev),
            homo_beta_ev: opt(res.has_homo_beta, res.homo_beta_ev),
            lumo_beta_ev: opt(res.has_lumo_beta, res.lumo_beta_ev),
            mo_energies: eps,
            mo_energies_beta: if res.unrestricted != 0 { Some(eps_b) } else { None },
            density: p,
            mo_coeff: c,
            iterations: res.iterations.max(0) as usize,
            converged: res.converged != 0,
            unrestricted: res.unrestricted != 0,
            nao,
        })
    }

    pub fn as_ptr(&self) -> *mut raw::am1g_ctx {
        self.ptr
    }
}

impl Drop for Context {
    fn drop(&mut self) {
        // SAFETY: the pointer came from am1g_create and is destroyed once.
        unsafe { raw::am1g_destroy(self.ptr) }
    }
}
