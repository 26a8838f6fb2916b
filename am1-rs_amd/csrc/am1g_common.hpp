// Shared host/device definitions of the am1g engine: element parameters as the kernels see them,
// per-molecule descriptors for the batched kernel, and result records.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__) || defined(__HIP__)
#define AM1G_HD __host__ __device__
#else
#define AM1G_HD
#endif

namespace am1g {

// MOPAC7 model constants the AM1 and RM1 parameters were fitted with (am1-rs constants.rs:13-32).
constexpr double kEv = 27.21;
constexpr double kA0 = 0.529167;
constexpr double kEvToKcal = 23.061;
constexpr double kAuDipoleToDebye = 2.541746473;

constexpr int kMaxZ = 87;
constexpr int kMethods = 2;  // 0 AM1, 1 RM1
constexpr int kDiisDepth = 8;

// One element as the kernels read it: the parameter row plus the derived multipole terms.
struct DevElem {
    double uss, upp, zs, zp, bs, bp;
    double gss, gsp, gpp, gp2, hsp, alpha;
    double core, dd, qq, rho0, rho1, rho2;
    double eheat_ev, e_isol;
    double gk[4], gl[4], gm[4];
    int n, norb, ng, valid;
};

// Status codes; the numbering follows am1-rs's Am1Error variants (error.rs:213-293).
enum Status : int {
    kOk = 0,
    kInvalidInput = 1,
    kElementNotParameterized = 2,
    kLinearAlgebra = 3,
    kScfNotConverged = 4,
    kScfDiverged = 5,
    kDeviceError = 6,
};

// Per-molecule record of the batched SCF. Offsets are element offsets into the shared pools.
struct MolDesc {
    int nat, nao, n_alpha, n_beta;
    int method, uhf, accel, max_scf;
    int precision;   // 0 fp64, 1 mixed
    int n_heavy, n_h, use_diis;
    int all_vecs;    // 1: every eigenvector in each iteration (the caller takes the coefficients);
                     // 0: the occupied ones, all the density needs
    int pad0;
    double e_tol, p_tol, adiis_switch;
    double field[3];    // uniform external field, eV per (e*Bohr); zero for none
    double field_nuc;   // its nuclear energy -F . sum_a Z_a R_a (eV)
    double far_cut;     // far-field cutoff (Bohr): pairs beyond it are net-charge monopoles; 0 = off
    int64_t atom_off;   // into z[] / xyz[] (xyz: 3 per atom)
    int64_t ao_off;     // into per-AO arrays (mo energies)
    int64_t mat_off;    // start of this molecule's dense matrices
    int64_t w_off;      // start of this molecule's pair-integral pool
    int64_t hist_off;   // start of this molecule's DIIS history
};

// Scalars each SCF reports. Mirrors the scalar fields of am1-rs Am1Result (scf.rs:200-240).
struct MolResult {
    int status, iterations, converged, nao;
    double electronic_ev, core_ev, total_ev, heat_of_formation_kcal;
    double homo_ev, lumo_ev, homo_beta_ev, lumo_beta_ev;
    double dipole[3];
    double residual;
    double fp32_iterations;  // iterations spent with fp32 GEMMs (mixed precision)
    uint64_t t_start, t_done;  // device real-time clock (100 MHz) when the molecule was taken / finished
    uint64_t phase[6];         // device clock ticks per SCF phase (see kPhaseNames)
    int eigen_sweeps;          // Jacobi sweeps summed over the SCF
    int pad1;
};

// Phase names of MolResult::phase, in the spirit of am1-rs's AM1_TIMING timers
// (scf.rs:315, 592-660: basis+core, scf:fock, scf:commutator + scf:accel, scf:eigen, scf:density).
constexpr int kPhases = 6;

// Dense matrices kept per molecule (each nao*nao doubles), in this order from mat_off.
enum MatSlot : int {
    kH = 0,   // core Hamiltonian
    kF,       // Fock (alpha)
    kP,       // total density (RHF) / alpha density (UHF)
    kC,       // eigenvectors, transposed: row k = MO k (alpha)
    kT0,      // scratch
    kT1,      // scratch
    kFb,      // UHF: beta Fock
    kPb,      // UHF: beta density
    kCb,      // UHF: beta eigenvectors (transposed)
    kPt,      // UHF: total density
    kMatSlotsRhf = kT1 + 1,
    kMatSlotsUhf = kPt + 1,
};

AM1G_HD inline int64_t w_pool_size(int n_heavy, int n_h) {
    int64_t hh = (int64_t)n_heavy * (n_heavy - 1) / 2;
    int64_t hH = (int64_t)n_heavy * n_h;
    int64_t HH = (int64_t)n_h * (n_h - 1) / 2;
    return hh * 100 + hH * 10 + HH;
}

AM1G_HD inline int64_t tri_len(int n) { return (int64_t)n * (n + 1) / 2; }

// DIIS history per molecule: depth x (packed F + packed E) for each spin channel.
AM1G_HD inline int64_t hist_size(int nao, int uhf) {
    int64_t per = 2 * tri_len(nao);
    return (int64_t)kDiisDepth * per * (uhf ? 2 : 1) + 64 * 4;
}

}  // namespace am1g
