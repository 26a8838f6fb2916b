// Launch interface between the host driver and the SCF kernels.
#pragma once
#include <hip/hip_runtime.h>

#include "am1g_common.hpp"

namespace am1g {

constexpr int kMaxAtoms = 256;
// Largest basis the per-molecule kernel accepts: the Jacobi round holds at most 128 rotations.
constexpr int kMaxBatchAo = 256;
// A single molecule with more basis functions than this runs faster on the grid-level kernels
// (the whole device on one molecule) than in one workgroup. Measured on water clusters, one
// workgroup against the grid path: 6.1 / 25.5 ms at 126 functions, 15.8 / 23.1 ms at 132 (the
// tridiagonal method with the matrix in global memory from 129 on), 28.1 / 31.0 ms at 192,
// 36.1 / 33.3 ms at 204, 55.0 / 43.5 ms at 252. Inside a full batch the workgroup path still
// gives more molecules per second (it holds 1/256 of the device), so am1g_scf_batch does not
// route; am1g_scf and the server do, for latency. The context's tunable "route_large_ao" starts
// at this value and can be re-measured per device.
constexpr int kRouteLargeAo = 192;
// Dynamic LDS of the per-molecule kernel (the static part is the molecule tables).
constexpr int kLdsBytes = 19200 * 8;

// Continuous admission (the server's service launch): one request of the host-pinned ring,
// written by the host before it publishes the request's ticket. The pool offsets in md point
// into the service launch's device pools.
struct SvcSlot {
    MolDesc md;
    int z[kMaxAtoms];
    int aoff[kMaxAtoms];
    int acls[kMaxAtoms];
    int alist[kMaxAtoms];
    double xyz[3 * kMaxAtoms];
};
// One request's results, written by the service launch into the host-pinned ring.
struct SvcOut {
    MolResult res;
    double charges[kMaxAtoms];
    double mo[kMaxBatchAo];
    double mob[kMaxBatchAo];
};

struct BatchArgs {
    const MolDesc* mols;
    int n_mols;
    int lds_n;  // matrices up to this order are diagonalized in LDS, larger ones in global memory
    int tri_max;  // orders up to this take the tridiagonal method, larger ones Jacobi
    const int* z;
    const double* xyz;
    const DevElem* elems;  // [kMethods][kMaxZ]
    double* mats;
    double* wpool;
    double* hist;
    double* mo_e;       // per-AO orbital energies (alpha)
    double* mo_e_beta;  // per-AO orbital energies (beta, UHF)
    double* charges;    // per atom
    MolResult* results;
    int* queue;         // [0]: next molecule
    const int* order;   // molecule processing order (largest first)
    // per-atom tables built on the host (indexed from MolDesc::atom_off)
    const int* alist;   // heavy atoms (ascending), then hydrogens (ascending)
    const int* aoff;    // AO offset of each atom inside its molecule
    const int* acls;    // index of each atom inside the heavy list or the hydrogen list
    const int64_t* pair_off;  // [n_mols + 1] first global pair index of each molecule
    int64_t n_pairs;
    double* ecore_pair;  // core-core repulsion of each pair, summed per molecule in fixed order
    int own_pairs;       // 1: each workgroup computes its molecule's pairs (no k_pairs launch)
    // Streaming (optional, all null otherwise): as each molecule finishes, its result, charges and
    // orbital energies are copied into host-visible memory and done[m] is set with system scope,
    // so the host can hand results out while the launch still runs.
    int* done;
    MolResult* results_h;
    double* charges_h;
    double* mo_e_h;
    double* mo_eb_h;
    // Service launch (all null otherwise): workgroups take requests from the host-pinned ring as
    // the host publishes them (ticket t in slot t % svc_cap) and compute each request's pair
    // integrals themselves; they leave once the host sets *svc_stop, or once the launch has been
    // idle for svc_idle_us (a deadline all workgroups share, so they leave together). The atom
    // tables (z, xyz, aoff, acls, alist) are then written by the kernel.
    const SvcSlot* svc_in;
    SvcOut* svc_out;
    long long* svc_done;             // per slot: ticket + 1 once that ticket's results are out
    const long long* svc_published;  // tickets below this one are written
    const int* svc_stop;
    long long* svc_next;             // device: the next ticket to take
    int* svc_busy;                   // device: workgroups working on a request now
    unsigned long long* svc_last;    // device: device clock of the last start, take or finish
    int* svc_alive;                  // host-pinned: workgroups of service launches that have not left
    int svc_cap;
    int svc_idle_us;                 // workgroups leave together once nothing has started, been taken or
                                     // finished for this long and none is busy
};

hipError_t launch_scf_batch(const BatchArgs& a, int precision, int grid, size_t lds_bytes, hipStream_t stream);
// The same kernel built for 256-thread workgroups (two per CU) and molecules of at most
// kSmallWgAo basis functions (am1g_scf_small.hip); grid: workgroups (two per CU fill the device).
constexpr int kSmallWgAo = 64;
hipError_t launch_scf_batch_small(const BatchArgs& a, int precision, int grid, hipStream_t stream);
// Size of the per-device scratch array of the allocation-free batched entry point.
constexpr size_t kScratchBytes = (size_t)64 << 20;
constexpr int kMaxDevices = 64;
// Looks up the device address of the scratch array on `device`, which must be current; once per
// device, later calls return at once.
hipError_t scratch_init(int device);
// Device address of the scratch array on `device` (null before scratch_init succeeded there).
void* scratch_ptr(int device);
// Zeroes n ints with a kernel (no memset node, so the sequence stays capturable).
hipError_t launch_zero_ints(int* p, int n, hipStream_t stream);
// Largest order for the tridiagonal eigensolver: kTriMax, or 0 when AM1G_EIGEN_JACOBI=1 (every
// eigen-step then takes the Jacobi solver; for testing that solver on the same molecules).
int eigen_tri_max();
// block: threads per workgroup of k_pairs (64, 128 or 256; the context tunable "pairs_block")
hipError_t launch_pairs(const BatchArgs& a, hipStream_t stream, int block = 256);
// The pairs of one molecule (a.n_mols == 1, md = its descriptor) on the whole device: heavy-heavy
// and heavy-H blocks staged through LDS and written 64 pairs at a time, hydrogen pairs directly.
hipError_t launch_pairs_one(const BatchArgs& a, const MolDesc& md, hipStream_t stream);
void build_element_tables(DevElem* tab);

}  // namespace am1g

namespace am1g {

// One large system handled by grid-level kernels (divide-and-conquer and the plain path).
struct SysArgs {
    MolDesc md;
    const int* z;        // [nat]
    const double* xyz;   // [nat][3]
    const int* alist;    // heavy atoms then hydrogens
    const int* aoff;     // AO offset per atom
    const int* acls;     // index inside the heavy / hydrogen list
    const int* later;    // [nat] class order: where the other class's partners later in AO order begin
    const DevElem* elems;
    const double* wpool;
    double* H;           // [nao][nao]
    double* pdiag;       // [nat][10] packed diagonal density blocks (off-diagonal entries doubled),
                         // rows in class order: heavy atoms, then hydrogens
    double* tcol;        // [n_hh][10] heavy pair (i < j): Coulomb of i's diagonal block on j's, at
                         // j (j - 1) / 2 + i (each column contiguous)
    double* rsum;        // [n_heavy][10] Coulomb on heavy atom i from its heavy partners j > i
    double* far;         // [2 nat + 1] far-field terms (md.far_cut > 0): per atom -p_a V_a, then
                         // Z_a V_a, then sum Z_a V_a of the last Fock build; null when off
};

// Divide-and-conquer subsystem tables (all device pointers).
struct DcArgs {
    int n_sub;
    int lds_n;
    int tri_max;             // orders up to this take the tridiagonal method, larger ones Jacobi
    const int* sub_n;        // [n_sub] AO count
    const int64_t* sub_ao;   // [n_sub + 1] offset into ao_list / level arrays
    const int64_t* sub_mat;  // [n_sub] offset of the subsystem's n*n matrices
    const int* ao_list;      // global AO index of every subsystem AO (ascending within a subsystem)
    const unsigned char* ao_core;  // 1 when the AO sits on a core atom
    double* fbuf;   // per subsystem: gathered Fock block / density block
    double* cbuf;   // per subsystem: eigenvectors (rows)
    double* t0;     // per subsystem scratch
    double* t1;     // per subsystem scratch
    double* eps;    // [levels]
    double* proj;   // [levels]
    double* frac;   // [levels]
    int warm;
    double cap;     // 2 restricted, 1 per spin channel
    double tol_abs; // Jacobi tolerance floor
    int* status;    // [n_sub] eigensolver status (sweeps, or -1)
    double* red;    // reduction scratch for the Fermi search: fermi_scratch_doubles()
};

hipError_t launch_hcore_diag(const SysArgs& a, hipStream_t st);
// Fock matrix F of the total density Pt and spin density Ps (exchange scaled by `scale`). With
// e_out given, also writes sum Ps (H + F) there, from per-workgroup partials in epart
// (2 * fock_energy_parts doubles).
// full = false leaves F's lower triangle unwritten (off-diagonal atom blocks are written in the
// upper triangle only; a reader then takes F[min(i,j)][max(i,j)]).
hipError_t launch_fock(const SysArgs& a, const double* Pt, const double* Ps, double scale, double* F, hipStream_t st,
                       double* epart = nullptr, double* e_out = nullptr, bool full = true);
inline int64_t fock_energy_parts(const MolDesc& md) {
    return 2 * (int64_t)md.n_heavy + (md.n_h + 7) / 8 + md.n_h;
}
hipError_t launch_dot3(int64_t n, const double* a, const double* b, const double* c, double* partial, double* out,
                       hipStream_t st);
hipError_t launch_dc_eigen(const DcArgs& d, const double* F, int nao, hipStream_t st);
int64_t fermi_scratch_doubles();  // size of DcArgs::red
hipError_t launch_fermi(const DcArgs& d, int64_t levels, double n_elec, double kt, double* out, hipStream_t st);
hipError_t launch_dc_density(const DcArgs& d, hipStream_t st);
hipError_t launch_dc_gather(const DcArgs& d, int64_t npat, const int64_t* pat, const int* ao_atom,
                            const int* core_sub, int nao, double* pout, hipStream_t st);
hipError_t launch_dc_hist(int64_t npat, const int64_t* pat, const double* mult, const double* P, const double* pout,
                          double* hin, double* hres, const double* hres_all, const int* ring, int m, int64_t stride,
                          double* partial, double* out, hipStream_t st);
hipError_t launch_dc_update(int64_t npat, const int64_t* pat, int nao, const double* coef, const double* hin_all,
                            const double* hres_all, const int* ring, int m, int64_t stride, double mix,
                            const double* P_in, const double* pout, double* P, hipStream_t st);
hipError_t launch_sad(const SysArgs& a, double scale, double* P, hipStream_t st);

// Large-system SCF (one molecule on the whole device). `partial` holds large_partials(n)
// doubles; `out` receives the reduced scalars.
int64_t large_partials(int n);
hipError_t launch_push(int n, const void* T, bool t_f32, const double* F, const double* P, double* hf, double* he,
                       double* hd, double* partial, double* out, hipStream_t st);
hipError_t launch_diis_dots(int n, const double* hf, const double* he, const double* hd, const int* ring, int m,
                            int newest, double* partial, double* out, hipStream_t st);
hipError_t launch_combine(int n, const double* hf, const int* ring, const double* coef, int m, double* Fout,
                          double* partial, double* out, hipStream_t st);
hipError_t launch_density_update(int n, const void* Pn, bool f32, double* P, double* partial, double* out,
                                 hipStream_t st);
// SP2 density purification (large-system SCF): Gershgorin rows (3 doubles per row: lower bound,
// upper bound, diagonal), X_0 = scale (e_max I - F), the trace of X, X2 <- 2X - X2, X <- w X.
hipError_t launch_gersh(int n, const double* F, double* out, hipStream_t st);
template <typename T>
hipError_t launch_sp2_init(int n, const double* F, double emax, double scale, T* X, hipStream_t st);
template <typename T>
hipError_t launch_trace(int n, const T* X, double* out, hipStream_t st);
template <typename T>
hipError_t launch_sp2_combine(int64_t len, const T* X, T* X2, hipStream_t st);
template <typename T>
hipError_t launch_scale(int64_t len, double w, T* X, hipStream_t st);
hipError_t launch_to_f32(int64_t len, const double* x, float* y, hipStream_t st);
hipError_t launch_add(int64_t len, const double* x, const double* y, double* z, hipStream_t st);
hipError_t launch_atom_props(const SysArgs& a, const double* P, double* out, hipStream_t st);
hipError_t launch_sum(int64_t n, const double* x, double* partial, double* out, hipStream_t st);

}  // namespace am1g
