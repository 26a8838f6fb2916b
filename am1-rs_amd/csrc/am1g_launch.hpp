// Launch interface between the host driver and the SCF kernels.
#pragma once
#include <hip/hip_runtime.h>

#include "am1g_common.hpp"

namespace am1g {

constexpr int kMaxAtoms = 256;
// Largest basis the per-molecule kernel accepts: the Jacobi round holds at most 128 rotations.
constexpr int kMaxBatchAo = 256;
// Dynamic LDS of the per-molecule kernel (the static part is the molecule tables).
constexpr int kLdsBytes = 19200 * 8;

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
    // per-atom tables built on the host (indexed from MolDesc::atom_off)
    const int* alist;   // heavy atoms (ascending), then hydrogens (ascending)
    const int* aoff;    // AO offset of each atom inside its molecule
    const int* acls;    // index of each atom inside the heavy list or the hydrogen list
    const int64_t* pair_off;  // [n_mols + 1] first global pair index of each molecule
    int64_t n_pairs;
    double* ecore_pair;  // core-core repulsion of each pair, summed per molecule in fixed order
};

hipError_t launch_scf_batch(const BatchArgs& a, int precision, int grid, size_t lds_bytes, hipStream_t stream);
hipError_t launch_pairs(const BatchArgs& a, hipStream_t stream);
void build_element_tables(DevElem* tab);

}  // namespace am1g
