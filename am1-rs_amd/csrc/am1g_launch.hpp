// Launch interface between the host driver and the SCF kernels.
#pragma once
#include <hip/hip_runtime.h>

#include "am1g_common.hpp"

namespace am1g {

constexpr int kMaxAtoms = 512;
// Largest basis the per-molecule kernel accepts: the Jacobi round holds at most 128 rotations.
constexpr int kMaxBatchAo = 256;

struct BatchArgs {
    const MolDesc* mols;
    int n_mols;
    int lds_n;  // matrices up to this order are diagonalized in LDS, larger ones in global memory
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
};

hipError_t launch_scf_batch(const BatchArgs& a, int precision, int grid, size_t lds_bytes, hipStream_t stream);
void build_element_tables(DevElem* tab);

}  // namespace am1g
