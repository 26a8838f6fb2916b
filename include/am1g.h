/*
 * am1g — MI355X-native AM1/RM1 NDDO SCF engine, C ABI.
 *
 * This header is the drop-in boundary for am1-rs 0.2.3's molecular SCF entry points. All
 * buffers are caller-owned; coordinates are Bohr, energies eV, heats of formation kcal/mol,
 * dipoles Debye. Matrices are dense row-major nao x nao; MO coefficients hold one orbital
 * per column, orbitals in ascending energy (the layout of am1-rs Am1Result).
 *
 * Reference interfaces replaced:
 *   am1g_scf        <- pub fn run_am1(&Molecule, &Am1Parameters, &Am1Options)
 *                      -> Result<Am1Result>                    (src/scf.rs:294-298)
 *                      and Am1Calculator::calculate            (src/scf.rs:257-259)
 *   am1g_scf_batch  <- many run_am1 calls (no single reference counterpart: the reference runs
 *                      them one at a time, e.g. python.rs:147-199 per molecule)
 *   am1g_request    <- Molecule (src/system.rs:19-39) + Am1Options (src/scf.rs:53-197)
 *   am1g_result     <- Am1Result (src/scf.rs:200-240)
 *   status codes    <- Am1Error (src/error.rs:213-293)
 */
#ifndef AM1G_H
#define AM1G_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes; same order as am1-rs's Am1Error variants used on this path. */
enum {
    AM1G_OK = 0,
    AM1G_INVALID_INPUT = 1,        /* Am1Error::InvalidInput */
    AM1G_ELEMENT_NOT_PARAMETERIZED = 2, /* Am1Error::ElementNotParameterized */
    AM1G_LINEAR_ALGEBRA = 3,       /* Am1Error::LinearAlgebra */
    AM1G_SCF_NOT_CONVERGED = 4,    /* Am1Error::ScfNotConverged; result is filled, residual set */
    AM1G_SCF_DIVERGED = 5,         /* Am1Error::ScfDiverged; iterations set */
    AM1G_DEVICE_ERROR = 6,         /* HIP runtime failure (no reference counterpart) */
    AM1G_PENDING = 7,              /* serving: the request has not finished yet */
    AM1G_CANCELLED = 8             /* serving: the request was cancelled */
};

enum { AM1G_METHOD_AM1 = 0, AM1G_METHOD_RM1 = 1 };                       /* NddoMethod */
enum { AM1G_REF_AUTO = 0, AM1G_REF_RHF = 1, AM1G_REF_UHF = 2 };           /* ScfReference */
enum { AM1G_ACCEL_NONE = 0, AM1G_ACCEL_CDIIS = 1, AM1G_ACCEL_ADIIS_CDIIS = 2 }; /* ScfAccelerator */
enum { AM1G_PRECISION_FP64 = 0, AM1G_PRECISION_MIXED = 1 };

typedef struct am1g_ctx am1g_ctx;

/* One SCF request: Molecule + the Am1Options fields the molecular SCF reads. */
typedef struct {
    int n_atoms;
    const uint8_t* z;        /* [n_atoms] atomic numbers */
    const double* xyz_bohr;  /* [n_atoms][3] */
    double charge;
    int multiplicity;        /* >= 1 */
    int method;              /* AM1G_METHOD_* */
    int reference;           /* AM1G_REF_* */
    int accelerator;         /* AM1G_ACCEL_* */
    int use_diis;            /* 0 forces AM1G_ACCEL_NONE (Am1Options::use_diis) */
    int max_scf;             /* default 200 */
    double e_tol;            /* default 1e-8 eV */
    double p_tol;            /* default 1e-7 */
    double adiis_switch;     /* default 0.1 */
    int precision;           /* AM1G_PRECISION_*; FP32 GEMM phase until the DIIS error < 1e-4 */
    const double* electric_field; /* [3] uniform external field in eV per (e*Bohr), or NULL for none
                                     (Am1Options::electric_field, scf.rs:141-153; every SCF
                                     entry point, am1g_dc included) */
    int bcc;                 /* nonzero: also AM1-BCC charges from this SCF's Mulliken charges into
                                am1g_result.bcc_charges (am1_bcc_charges, bcc/mod.rs:99-150); AM1
                                method only, and the charges output must be given */
    double multipole_cutoff; /* Bohr; atom pairs farther apart leave the pair list and interact
                                through their net charges only (Am1Options::multipole_cutoff,
                                scf.rs:130-140, farfield.rs:80-236); 0 (the default), a negative
                                value or NaN keeps every pair exact */
} am1g_request;

/* Result of one SCF. Pointer members are caller-provided outputs; NULL skips that output. */
typedef struct {
    int status;
    int nao, n_occ, n_occ_beta;
    int iterations, converged, unrestricted;
    int has_homo, has_lumo, has_homo_beta, has_lumo_beta;
    double electronic_ev, core_ev, total_ev, heat_of_formation_kcal;
    double homo_ev, lumo_ev, homo_beta_ev, lumo_beta_ev;
    double dipole_debye[3];
    double dipole_magnitude;
    double field_nuclear_ev;  /* -F . sum_a Z_a R_a, part of total_ev (Am1Result::field_nuclear_ev) */
    double residual;          /* ||[F,P]|| of the last iteration */
    int fp32_iterations;      /* iterations run with FP32 GEMMs (mixed precision) */
    double latency_ms;        /* batched: device time from the launch's first molecule start to this
                                 molecule's completion; plain: wall time of the call */
    double phase_ms[6];       /* device time per phase: setup (integrals, H_core), Fock + energy,
                                 commutator + DIIS, eigensolve, density + convergence, final
                                 energy and properties (am1-rs AM1_TIMING phases, timing.rs:34-85) */
    int eigen_sweeps;         /* Jacobi sweeps summed over all iterations; on the grid-level path
                                 the SP2 purification steps (diagnostic) */
    int bcc_warnings;         /* AM1-BCC: perception or parameter warnings (am1g_bcc_charges) */
    double* charges;          /* [n_atoms] Mulliken charges */
    double* mo_energies;      /* [nao] (alpha for UHF) */
    double* mo_energies_beta; /* [nao] UHF only */
    double* density;          /* [nao*nao] total density */
    double* mo_coeff;         /* [nao*nao] orbitals as columns (alpha for UHF) */
    double* mo_coeff_beta;    /* [nao*nao] UHF only: beta orbitals as columns (BetaOrbitals::coeff) */
    double* spin_density;     /* [nao*nao] UHF only: P_alpha - P_beta (Am1Result::spin_density) */
    double* bcc_charges;      /* [n_atoms] AM1-BCC charges when the request set bcc */
} am1g_result;

/* Fill a request with am1-rs's Am1Options defaults (scf.rs:174-197). */
void am1g_request_defaults(am1g_request* rq);

/* Create a context on HIP device `device`: uploads the AM1 and RM1 parameter tables and owns
 * the device pools and stream. No reference counterpart (the reference caches parameters in a
 * process-wide OnceLock, params.rs:100-115). */
int am1g_create(int device, am1g_ctx** out);
void am1g_destroy(am1g_ctx* ctx);

/* Plain path: one molecule. Up to 128 basis functions (and 256 atoms) one workgroup runs the
 * whole SCF; larger molecules run on grid-level kernels over the whole device. Replaces run_am1
 * (src/scf.rs:294-298). */
int am1g_scf(am1g_ctx* ctx, const am1g_request* rq, am1g_result* res);

/* Batched path: n independent molecules (each at most 256 atoms and 256 basis functions) in one
 * persistent launch, one workgroup per molecule. Each result is
 * the same as the molecule's am1g_scf result within the stated tolerance. Returns AM1G_OK when
 * the launch ran; per-molecule outcomes are in results[i].status. */
int am1g_scf_batch(am1g_ctx* ctx, int n, const am1g_request* rqs, am1g_result* results);

/* Allocation-free batched path. No counterpart in the reference: it is for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc). Same kernels and results as am1g_scf_batch, but every device buffer of the
 * launch comes from one fixed scratch array per GPU (am1g_noalloc_capacity_bytes(), 64 MiB),
 * reserved when the library's device code is loaded and held until the process ends; the call
 * makes no hipMalloc / hipFree. A batch whose buffers exceed the array is refused with
 * AM1G_INVALID_INPUT and a message giving both sizes (as a guide, about 25 molecules of 128
 * basis functions or 90 of 70). The array keeps the last call's contents; calls that overlap on
 * one GPU (from two contexts or two threads without a shared context) race on it. Density and
 * orbital matrices are returned as with am1g_scf_batch. */
int am1g_scf_batch_noalloc(am1g_ctx* ctx, int n, const am1g_request* rqs, am1g_result* results);
/* Bytes of the per-GPU scratch array of am1g_scf_batch_noalloc. */
long long am1g_noalloc_capacity_bytes(void);

/* Divide-and-conquer controls: DcOptions (src/divide_conquer.rs:66-119). */
typedef struct {
    int core_size;            /* target atoms per core when core_of_atom is NULL (default 12) */
    double buffer_radius;     /* Bohr (default 11.0) */
    double kt;                /* Fermi-Dirac electronic temperature in eV (default 0.05); <= 0 is aufbau */
    int max_scf;              /* default 300 */
    double e_tol;             /* default 1e-7 eV */
    double p_tol;             /* default 1e-6 */
    double mixing;            /* linear mixing fraction when DIIS declines (default 0.4) */
    const int* core_of_atom;  /* optional [n_atoms]: caller-chosen core index of every atom */
} am1g_dc_options;

/* Divide-and-conquer result: DcResult (src/divide_conquer.rs:167-220). */
typedef struct {
    int status, iterations, converged, unrestricted;
    int n_subsystems, largest_subsystem_aos;
    double electronic_ev, core_ev, total_ev, heat_of_formation_kcal;
    double fermi_ev;                 /* chemical potential (alpha channel when unrestricted) */
    double fermi_beta_ev;            /* beta channel's chemical potential (= fermi_ev for RHF) */
    double entropy_ev, homo_lumo_gap_ev;
    double diagonalization_work;     /* sum over subsystems of n^3 */
    long long retained_pattern;      /* AO pairs of the retained density pattern (upper triangle) */
    double phase_ms[5];              /* Fock, subsystem eigensolves, Fermi level, assembly, DIIS */
    double* charges;                 /* [n_atoms] */
    double* density;                 /* [nao*nao] optional */
    double* spin_density;            /* [nao*nao] optional, unrestricted runs only: P_alpha - P_beta
                                        (DcResult::spin_density) */
} am1g_dc_result;

void am1g_dc_defaults(am1g_dc_options* op);

/* Divide-and-conquer SCF on the GPU: run_divide_conquer (src/divide_conquer.rs:406-572), RHF and
 * UHF (dc_loop_restricted / dc_loop_unrestricted, one chemical potential per spin). All subsystem
 * eigensolves of an iteration run as one batched launch per spin channel. */
int am1g_dc(am1g_ctx* ctx, const am1g_request* rq, const am1g_dc_options* op, am1g_dc_result* res);

/* ---------------------------------------------------------------- serving
 * Continuous batching. No reference counterpart: am1-rs's callers run run_am1 concurrently,
 * one molecule per call, with the GIL released (python.rs:163, 551). A server admits requests
 * as they arrive into concurrent batched launches ("lanes", one HIP stream and one set of
 * device pools each): whenever a lane is free it takes every queued request of one precision
 * that fits the device-memory budget (up to max_batch), so requests that arrive while others
 * are on the GPU start as soon as a lane and memory free up, and launches of different lanes
 * overlap on the device. Each result equals the molecule's am1g_scf_batch result alone. */
typedef struct am1g_server am1g_server;

typedef struct {
    int lanes;                /* concurrent launches (default 4, at most 16) */
    int max_batch;            /* molecules per launch (default 512) */
    double memory_budget_mb;  /* device memory the running launches may hold (default 32768) */
    double max_wait_ms;       /* batching window: a free lane waits this long after the first queued
                                 request for more to arrive before launching (default 0) */
    int continuous;           /* 1 (default): one persistent service launch takes requests as they
                                 arrive (a request joins the running launch as soon as a workgroup is
                                 free); lanes and max_batch then apply only to requests beyond one
                                 workgroup's reach, which run alone on the grid-level path.
                                 0: lanes of batched launches as described above */
} am1g_serve_options;

typedef struct {
    long long submitted, completed, cancelled, queued, running, launches;
    double admitted_mb;       /* device memory held by running launches now */
    double peak_admitted_mb;
} am1g_serve_stats;

void am1g_serve_defaults(am1g_serve_options* op);
int am1g_serve_create(int device, const am1g_serve_options* op, am1g_server** out);
/* Stops the lanes after their running launches; queued requests are dropped. */
void am1g_serve_destroy(am1g_server* sv);
/* Copies the request (atoms included) and queues it; *ticket identifies it. */
int am1g_serve_submit(am1g_server* sv, const am1g_request* rq, long long* ticket);
/* Waits up to wait_ms (0: do not wait, < 0: until done). Returns AM1G_PENDING while the request
 * runs; otherwise fills *out (scalars, and charges / mo_energies / mo_energies_beta when those
 * pointers are set; latency_ms = submit to completion), returns the request's status and forgets
 * the ticket. */
int am1g_serve_poll(am1g_server* sv, long long ticket, double wait_ms, am1g_result* out);
/* A queued request is dropped at once; a running one finishes on the device and is reported as
 * AM1G_CANCELLED. */
int am1g_serve_cancel(am1g_server* sv, long long ticket);
int am1g_serve_get_stats(am1g_server* sv, am1g_serve_stats* out);
const char* am1g_serve_last_error(const am1g_server* sv);

/* Dense symmetric eigensolver on the device: symmetric_eigen (src/linalg.rs:276-331). `a` is
 * n x n row-major and symmetric; eigenvalues come out ascending, eigenvectors as the columns of
 * the row-major n x n `eigenvectors`. Non-finite input is AM1G_LINEAR_ALGEBRA, as in the
 * reference. Householder tridiagonalization, tridiagonal divide and conquer and a blocked
 * back-transformation; the same solver serves the SCF of systems above 256 basis functions. */
int am1g_symmetric_eigen(am1g_ctx* ctx, int n, const double* a, double* eigenvalues, double* eigenvectors);

/* Dense matrix product on the device: Matrix::matmul, transpose_matmul, matmul_transpose and
 * their _acc forms (src/linalg.rs:92-181), the products of the SCF's commutator and density
 * steps. c = alpha * op(a) * op(b) + beta * c, all row-major: op(a) is m x k, op(b) k x n, c m x n;
 * trans_a / trans_b = 1 means the operand is stored transposed (a is k x m, b is n x k). beta = 0
 * does not read c. AM1G_PRECISION_FP64 runs FP64 matrix cores; AM1G_PRECISION_MIXED rounds the
 * operands to FP32 and accumulates in FP32 (the mixed-precision density product). The library's
 * own grid-level GEMM: the large-system SCF and eigensolver use the same kernels.
 * am1g_last_kernel_ms gives the device time of the product alone. */
int am1g_matmul(am1g_ctx* ctx, int trans_a, int trans_b, int m, int n, int k, double alpha, const double* a,
                const double* b, double beta, double* c, int precision);

/* Device-dependent launch choices of a context, measured by the project's autotuner (am1g.tune)
 * and read at am1g_create from tuned.txt beside the library (or the file AM1G_TUNE_FILE names):
 *   route_large_ao     single molecules above this many basis functions run on the grid-level
 *                      path (am1g_scf, am1g_gradient, the server); default 192
 *   tri_max            the batched kernel diagonalizes orders up to this by the tridiagonal
 *                      method (above 128 with the matrix in global memory), larger ones by
 *                      Jacobi; default 256 (at most 256)
 *   jacobi_lds_n       the batched Jacobi solver works in LDS up to this order, in global memory
 *                      above it; default 192 (at most 192)
 *   eig_graphs         1: the large eigensolver's tridiagonalization replays as a HIP graph;
 *                      default 1
 *   serve_large_batch  a server lane runs the queued grid-path requests (within one
 *                      workgroup's reach) in one batched launch once this many are queued;
 *                      default 2
 *   pairs_block        threads per workgroup of the pair-integral kernel (64, 128 or 256);
 *                      default 256
 *   batch_own_pairs    1: in a batched launch each workgroup computes its molecule's pair
 *                      integrals before its SCF (one launch); 0: a k_pairs launch over every pair
 *                      of the batch first; the same bits either way; default 1
 *   tri_blocked_min    the large eigensolver reduces orders from this one to tridiagonal form
 *                      in panels of 32 columns (the trailing matrix read once per column, a
 *                      matrix-core rank-64 update per panel) instead of one column at a time;
 *                      default 2^30 (off)
 *   batch_small_wg     1: in a batch with at least two molecules of at most 64 basis functions
 *                      per CU, those run two per CU in a 256-thread build of the batched kernel
 *                      (its own launch after the larger molecules'); their results then differ
 *                      from the same molecule's alone in the last bits (another summation order);
 *                      0: every molecule in the 512-thread build; default 1
 *   sp2                1: the grid-level SCF forms each iteration's density by SP2 purification
 *                      of the extrapolated Fock matrix (a chain of n x n x n products on the
 *                      matrix cores) and diagonalizes once, after the loop, for the orbital
 *                      energies and orbitals; 0: it diagonalizes in every iteration; default 1
 * No reference counterpart. AM1G_INVALID_INPUT for an unknown name. */
int am1g_set_tunable(am1g_ctx* ctx, const char* name, double value);
int am1g_get_tunable(const am1g_ctx* ctx, const char* name, double* value);

/* Nuclear gradient of the SCF energy: closed_form_gradient (src/gradient.rs:173-254) behind the
 * Python gradient (src/python.rs:202-275). Runs the am1g_scf SCF (res as there), then on the
 * device the derivative of every atom pair's energy at the converged densities (resonance,
 * electron-core, two-electron and core-core terms; central differences of the pair's own
 * integrals, the scheme of analytic_gradient, gradient.rs:41-90), RHF and UHF, plus -F q_a for
 * an external field. gradient[n_atoms][3] in eV/Bohr. */
int am1g_gradient(am1g_ctx* ctx, const am1g_request* rq, am1g_result* res, double* gradient);

/* Nuclear gradient at given densities (the fixed-density, Hellmann-Feynman gradient):
 * electronic_gradient_fixed_density(_spin) plus the core-core, far-field and field terms
 * (src/gradient.rs, src/divide_conquer.rs:1448-1515 divide_conquer_gradient, which applies it to
 * the divide-and-conquer density). density[nao][nao] is the total density; spin_density (NULL
 * for a closed shell) is P_alpha - P_beta. gradient[n_atoms][3] in eV/Bohr. The request's
 * geometry, method, field and multipole_cutoff are used; no SCF is run. */
int am1g_gradient_density(am1g_ctx* ctx, const am1g_request* rq, const double* density,
                          const double* spin_density, double* gradient);

/* AM1-BCC partial charges (am1_bcc_charges, src/bcc/mod.rs:99-150) from AM1 Mulliken charges:
 * the molecular graph is perceived from the geometry (src/topology.rs: bonds by covalent radii,
 * rings, aromaticity, Kekule structure, delocalized groups), every atom gets its antechamber BCC
 * atom type (ATOMTYPE_BCC.DEF rules, src/bcc/atomtype.rs) and every bond its correction from
 * BCCPARM.DAT (+delta onto the first atom, -delta onto the second). Host code, no device.
 * charges[n_atoms] receives the corrected charges; atom_types[n_atoms] (optional) the type codes
 * (11..91, 0 = none); *n_warnings (optional) the number of warnings, whose texts go,
 * newline-separated, into warnings[warnings_len] (optional). AM1G_INVALID_INPUT for bad
 * arguments. */
int am1g_bcc_charges(int n_atoms, const uint8_t* z, const double* xyz_bohr, const double* mulliken,
                     double* charges, int* atom_types, int* n_warnings, char* warnings, int warnings_len);

/* The core Hamiltonian of one molecule as the device builds it: build_core_with_neighbors
 * (src/hamiltonian.rs:211-366, CoreHamiltonian{h_core, pairs}) and
 * core_core_energy_with_neighbors (src/repulsion.rs:38-56). h_core[nao][nao] (row-major; NULL =
 * skip) receives H_core, including an external field's dipole term; *core_ev (NULL = skip) the
 * core-core repulsion in eV; pair_blocks[pair_blocks_len] (NULL = skip) the two-electron pair
 * integrals (PairIntegral::w, src/hamiltonian.rs:58-83) grouped by pair class, each pair
 * oriented heavy atom first with b - a as the separation: heavy-heavy pairs (i < j over the heavy
 * atoms in input order) as 10x10 row-major blocks, then heavy-hydrogen pairs (heavy-major, each
 * heavy atom against every hydrogen in input order) as 10 values, then hydrogen-hydrogen pairs
 * (i < j) as one value; it needs 100 hh + 10 hH + HH doubles. For per-block checks of the
 * pair-integral kernel. AM1G_INVALID_INPUT for a short pair_blocks buffer. */
int am1g_core_hamiltonian(am1g_ctx* ctx, const am1g_request* rq, double* h_core, double* core_ev,
                          double* pair_blocks, long long pair_blocks_len);

/* Text of the last error raised on the calling thread by any am1g_* call that takes a context
 * (the Display text of the matching Am1Error). Thread-local: concurrent callers sharing a
 * context each read their own message. */
const char* am1g_last_error(const am1g_ctx* ctx);

/* Derived element parameters as the library uses them (for parity checks):
 * out[23] = n, norb, uss, upp, zs, zp, bs, bp, gss, gsp, gpp, gp2, hsp, alpha, core,
 *           eheat_ev, e_isol, dd, qq, rho0, rho1, rho2, n_gaussians.
 * Mirrors Am1Element (src/params.rs:24-55). */
int am1g_element_info(int z, int method, double* out);

/* Kernel time of the last launch on the context, in milliseconds (HIP events on the
 * context's stream). */
double am1g_last_kernel_ms(const am1g_ctx* ctx);

#ifdef __cplusplus
}
#endif

#endif /* AM1G_H */
