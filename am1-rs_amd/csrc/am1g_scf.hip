// am1g SCF kernels for gfx950.
//
// k_scf_batch: one persistent launch solves a whole batch of molecules of uneven size. Each
// workgroup takes the next molecule from a device work counter (largest first) and runs its
// entire SCF — pair integrals, core Hamiltonian, every Fock/DIIS/eigen/density iteration and
// the result epilogue — without returning to the host. All of a molecule's sums are reduced in
// a fixed order inside its own workgroup, so results do not depend on the rest of the batch.
//
// Reference behaviour: am1-rs 0.2.3 run_am1 / rhf_loop / uhf_loop (src/scf.rs:294-858),
// build_core_with_neighbors (src/hamiltonian.rs:211-366), build_fock_spin_with
// (src/fock.rs:249-511), packed CDIIS / A-DIIS (src/scf.rs:876-1204).
#include <hip/hip_runtime.h>

#include <mutex>

#include <algorithm>
#include <cfloat>
#include <cstdlib>

#include "am1g_common.hpp"
#include "am1g_pair.hpp"
#include "am1g_launch.hpp"
#include "am1g_wg.hpp"
#include "am1g_tri.hpp"

namespace am1g {



struct MolShared {
    int z[kMaxAtoms];
    int off[kMaxAtoms];
    int cls[kMaxAtoms];    // index inside the heavy list or the H list
    int heavy[kMaxAtoms];  // atom indices of heavy atoms, ascending
    int hyd[kMaxAtoms];    // atom indices of hydrogens, ascending
    double B[kDiisDepth * kDiisDepth];   // CDIIS Gram of the error history (logical order)
    double G[kDiisDepth * kDiisDepth];   // A-DIIS Gram <D_i, F_j>
    double hd[kDiisDepth];               // A-DIIS <D_i, H>
    double coef[kDiisDepth + 1];
    int ring[kDiisDepth];  // logical -> physical history slot
    int nh, nH, hcount, head;
    int flags[8];
};

// Packed index k (0..9) of a 4-orbital block -> orbital pair (row >= col).
__device__ __constant__ int kPackRow[10] = {0, 1, 1, 2, 2, 2, 3, 3, 3, 3};
__device__ __constant__ int kPackCol[10] = {0, 0, 1, 0, 1, 2, 0, 1, 2, 3};
// The same pairs as compile-time constants, for fully unrolled loops.
constexpr int kPackRowH[10] = {0, 1, 1, 2, 2, 2, 3, 3, 3, 3};
constexpr int kPackColH[10] = {0, 0, 1, 0, 1, 2, 0, 1, 2, 3};

// Where the integral block of atoms x and y lives, and whether x is its bra (row) atom.
struct BlockRef {
    const double* w;
    int wcols;
    bool x_is_bra;
};

__device__ __forceinline__ BlockRef block_of(const MolShared& ms, const double* wpool, int x, int y) {
    const int nh = ms.nh, nH = ms.nH;
    const int64_t hh = (int64_t)nh * (nh - 1) / 2, hH = (int64_t)nh * nH;
    const bool hx = ms.z[x] != 1, hy = ms.z[y] != 1;
    BlockRef r;
    if (hx && hy) {
        int i = ms.cls[x], j = ms.cls[y];
        bool xa = i < j;
        r.w = wpool + tri_index(min(i, j), max(i, j), nh) * 100;
        r.wcols = 10;
        r.x_is_bra = xa;
    } else if (hx != hy) {
        int hi = hx ? ms.cls[x] : ms.cls[y];
        int mi = hx ? ms.cls[y] : ms.cls[x];
        r.w = wpool + hh * 100 + ((int64_t)hi * nH + mi) * 10;
        r.wcols = 1;
        r.x_is_bra = hx;
    } else {
        int i = ms.cls[x], j = ms.cls[y];
        r.w = wpool + hh * 100 + hH * 10 + tri_index(min(i, j), max(i, j), nH);
        r.wcols = 1;
        r.x_is_bra = i < j;
    }
    return r;
}

// ---------------------------------------------------------------------------------------------
// k_pairs: one thread per atom pair of the whole batch. Writes the pair's rotated two-electron
// block into the integral pool, its resonance block 1/2 (beta_mu + beta_nu) S_mu_nu into both
// triangles of H_core (each off-diagonal atom block belongs to exactly one pair), and its
// core-core repulsion (hamiltonian.rs:281-344, repulsion.rs:144-167).
// A resonance entry (i, j) of H_core (i and j on different atoms): both places, or with UPPER the
// one above the diagonal only.
template <bool UPPER>
__device__ __forceinline__ void res_store(double* H, int nao, int i, int j, double v) {
    if (UPPER) {
        H[(size_t)min(i, j) * nao + max(i, j)] = v;
    } else {
        H[(size_t)i * nao + j] = v;
        H[(size_t)j * nao + i] = v;
    }
}

// One atom pair k_all (molecule-local: heavy-heavy, heavy-H, H-H in class order) of molecule md:
// writes its integral block (100, 10 or 1 entries) to w (the pool, or a staging row in LDS), its
// resonance block into H_core (both triangles; with UPPER only the entries above the diagonal,
// for a later mirror pass), and returns its core-core term. CLS >= 0: the pair is known to be of
// that class (0 heavy-heavy, 1 heavy-H, 2 H-H), so only that class's code is compiled in (the
// light classes then need far fewer registers).
template <bool UPPER, int CLS = -1>
__device__ __forceinline__ double pair_one_to(const BatchArgs& g, const MolDesc& md, int64_t k_all, double* w) {
    const DevElem* el = g.elems + md.method * kMaxZ;
    const double* xyz = g.xyz + md.atom_off * 3;
    const int* z = g.z + md.atom_off;
    const int* alist = g.alist + md.atom_off;
    const int* aoff = g.aoff + md.atom_off;
    const int nh = md.n_heavy, nH = md.n_h, nao = md.nao;
    const int64_t n_hh = (int64_t)nh * (nh - 1) / 2, n_hH = (int64_t)nh * nH;
    double* H = g.mats + md.mat_off + kH * (size_t)nao * nao;
    const int cls = CLS >= 0 ? CLS : (k_all < n_hh ? 0 : (k_all < n_hh + n_hH ? 1 : 2));
    int a, b;
    if (cls == 0) {
        int i, j;
        tri_decode(k_all, nh, i, j);
        a = alist[i];
        b = alist[j];
    } else if (cls == 1) {
        int64_t k = k_all - n_hh;
        a = alist[(int)(k / nH)];
        b = alist[nh + (int)(k % nH)];
    } else {
        int i, j;
        tri_decode(k_all - n_hh - n_hH, nH, i, j);
        a = alist[nh + i];
        b = alist[nh + j];
    }
    const DevElem& ea = el[z[a]];
    const DevElem& eb = el[z[b]];
    double d[3] = {xyz[3 * b] - xyz[3 * a], xyz[3 * b + 1] - xyz[3 * a + 1], xyz[3 * b + 2] - xyz[3 * a + 2]};
    double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    const int oa = aoff[a], ob = aoff[b];
    if (md.far_cut > 0.0 && r > md.far_cut) {
        // beyond the far-field cutoff the pair is out of the pair list (neighbors.rs:90-101):
        // no integrals, no resonance, no core-core term; its net charges interact through
        // the far-field potential instead
        if (cls == 0) {
            for (int q = 0; q < 100; ++q) w[q] = 0.0;
            for (int i2 = 0; i2 < 4; ++i2)
                for (int j2 = 0; j2 < 4; ++j2) res_store<UPPER>(H, nao, oa + i2, ob + j2, 0.0);
        } else if (cls == 1) {
            for (int q = 0; q < 10; ++q) w[q] = 0.0;
            for (int i2 = 0; i2 < 4; ++i2) res_store<UPPER>(H, nao, oa + i2, ob, 0.0);
        } else {
            w[0] = 0.0;
            res_store<UPPER>(H, nao, oa, ob, 0.0);
        }
        return 0.0;
    }
    double S[16];
    overlap_block(ea, eb, d, S);
    if (cls == 0) {
        heavy_heavy_block(ea, eb, d, w);
#pragma unroll
        for (int i2 = 0; i2 < 4; ++i2)
#pragma unroll
            for (int j2 = 0; j2 < 4; ++j2) {
                double v = 0.5 * ((i2 == 0 ? ea.bs : ea.bp) + (j2 == 0 ? eb.bs : eb.bp)) * S[i2 * 4 + j2];
                res_store<UPPER>(H, nao, oa + i2, ob + j2, v);
            }
    } else if (cls == 1) {
        double wh[10];
        heavy_h_block(ea, eb, d, wh);
#pragma unroll
        for (int q = 0; q < 10; ++q) w[q] = wh[q];
#pragma unroll
        for (int i2 = 0; i2 < 4; ++i2) {
            double v = 0.5 * ((i2 == 0 ? ea.bs : ea.bp) + eb.bs) * S[i2 * 4];
            res_store<UPPER>(H, nao, oa + i2, ob, v);
        }
    } else {
        w[0] = hh_integral(ea, eb, r);
        double v = 0.5 * (ea.bs + eb.bs) * S[0];
        res_store<UPPER>(H, nao, oa, ob, v);
    }
    // the core-core term is symmetric in the pair; evaluate it in index order (i < j)
    const int lo_a = min(a, b), hi_a = max(a, b);
    return core_core_pair(el[z[lo_a]], el[z[hi_a]], z[lo_a], z[hi_a], r);
}

// Where pair k_all's integral block starts in the molecule's pool (its class offset plus its index).
__device__ __forceinline__ double* pair_block(const BatchArgs& g, const MolDesc& md, int64_t k_all) {
    const int nh = md.n_heavy, nH = md.n_h;
    const int64_t n_hh = (int64_t)nh * (nh - 1) / 2, n_hH = (int64_t)nh * nH;
    double* wpool = g.wpool + md.w_off;
    return k_all < n_hh ? wpool + k_all * 100
                        : (k_all < n_hh + n_hH ? wpool + n_hh * 100 + (k_all - n_hh) * 10
                                               : wpool + n_hh * 100 + n_hH * 10 + (k_all - n_hh - n_hH));
}

// The pair's integral block straight into the pool, its resonance block at both places.
__device__ __forceinline__ double pair_one(const BatchArgs& g, const MolDesc& md, int64_t k_all) {
    return pair_one_to<false>(g, md, k_all, pair_block(g, md, k_all));
}

__global__ void k_sym_lower(int n, double* F);

// One molecule's pairs of one class (0 heavy-heavy, 1 heavy-H) on the whole device, 64
// consecutive pairs per one-wavefront workgroup: each lane forms its pair's block in LDS, then
// the wavefront writes the 64 blocks, which are contiguous in the pool, with consecutive lanes on
// consecutive doubles (one lane per pair wrote 800- or 80-byte blocks 64 apart, a cache line
// per lane and store). The resonance entries go above the diagonal, the core-core terms as before.
template <int CLS>
__global__ __launch_bounds__(64) void k_pairs_cls(BatchArgs g) {
    constexpr int W = CLS == 0 ? 100 : 10, LD = W + 1;  // odd row stride: lanes on distinct banks
    __shared__ double st[64 * LD];
    const MolDesc& md = g.mols[0];
    const int nh = md.n_heavy, nH = md.n_h;
    const int64_t n_hh = (int64_t)nh * (nh - 1) / 2, n_hH = (int64_t)nh * nH;
    const int64_t n_cls = CLS == 0 ? n_hh : n_hH, first = CLS == 0 ? 0 : n_hh;
    double* pool = g.wpool + md.w_off + (CLS == 0 ? 0 : n_hh * 100);
    const int lane = threadIdx.x;
    for (int64_t k0 = (int64_t)blockIdx.x * 64; k0 < n_cls; k0 += (int64_t)gridDim.x * 64) {
        const int64_t k = k0 + lane;
        if (k < n_cls) {
            g.ecore_pair[first + k] = pair_one_to<true, CLS>(g, md, first + k, st + lane * LD);
        }
        __syncthreads();
        const int cnt = (int)min((int64_t)64, n_cls - k0);
        double* dst = pool + k0 * W;
        for (int e = lane; e < cnt * W; e += 64) dst[e] = st[(e / W) * LD + e % W];
        __syncthreads();
    }
}

// One molecule's pairs [from, to) of class CLS (one thread each), resonance above the diagonal.
template <int CLS>
__global__ __launch_bounds__(256) void k_pairs_range(BatchArgs g, int64_t from, int64_t to) {
    const MolDesc& md = g.mols[0];
    for (int64_t p = from + (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < to;
         p += (int64_t)gridDim.x * blockDim.x)
        g.ecore_pair[p] = pair_one_to<true, CLS>(g, md, p, pair_block(g, md, p));
}

__global__ __launch_bounds__(256) void k_pairs(BatchArgs g) {
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < g.n_pairs;
         p += (int64_t)gridDim.x * blockDim.x) {
        // molecule of this pair: last m with pair_off[m] <= p
        int lo = 0, hi = g.n_mols - 1;
        while (lo < hi) {
            int mid = (lo + hi + 1) >> 1;
            if (g.pair_off[mid] <= p) lo = mid;
            else hi = mid - 1;
        }
        g.ecore_pair[p] = pair_one(g, g.mols[lo], p - g.pair_off[lo]);
    }
}

// Service launch: every pair of one molecule by its own workgroup. Returns the core-core sum,
// accumulated per thread in the order of mol_setup's sum over k_pairs' results, then reduced the
// same way, so it equals the batched launch's value bit for bit.
__device__ __noinline__ double pairs_wg(const BatchArgs& g, const MolDesc& md, WgScratch& s) {
    const int64_t np = (int64_t)md.nat * (md.nat - 1) / 2;
    double ecore = 0.0;
    for (int64_t k = threadIdx.x; k < np; k += kWg) ecore += pair_one(g, md, k);
    return wg_sum(ecore, s);
}

// Per-molecule setup inside the SCF kernel: atom tables into LDS, H_core diagonal blocks
// U + sum over partners of (-Z_y)(mu nu|s_y s_y) (hamiltonian.rs:229-332), core-core sum.
__device__ double mol_setup(const MolDesc& md, int mi, const BatchArgs& g, MolShared& ms, WgScratch& s, double* H,
                            const double* wpool, bool own_pairs, double ecore_own) {
    const int t = threadIdx.x;
    const int nat = md.nat, nao = md.nao;
    const DevElem* el = g.elems + md.method * kMaxZ;
    for (int a = t; a < nat; a += kWg) {
        ms.z[a] = g.z[md.atom_off + a];
        ms.off[a] = g.aoff[md.atom_off + a];
        ms.cls[a] = g.acls[md.atom_off + a];
        int x = g.alist[md.atom_off + a];
        if (a < md.n_heavy) ms.heavy[a] = x;
        else ms.hyd[a - md.n_heavy] = x;
    }
    if (t == 0) {
        ms.nh = md.n_heavy;
        ms.nH = md.n_h;
    }
    __syncthreads();
    const int nh = ms.nh, nH = ms.nH;
    const int n_items = nh * 10 + nH;
    for (int it = t; it < n_items; it += kWg) {
        int x, k;
        if (it < nh * 10) {
            x = ms.heavy[it / 10];
            k = it % 10;
        } else {
            x = ms.hyd[it - nh * 10];
            k = 0;
        }
        const DevElem& ex = el[ms.z[x]];
        int mu = kPackRow[k], nu = kPackCol[k];
        double acc = 0.0;
        for (int y = 0; y < nat; ++y) {
            if (y == x) continue;
            BlockRef br = block_of(ms, wpool, x, y);
            double wv = br.x_is_bra ? br.w[k * br.wcols] : br.w[k];
            acc += -el[ms.z[y]].core * wv;
        }
        double u = (mu == nu) ? (mu == 0 ? ex.uss : ex.upp) : 0.0;
        // uniform external field F . M (dipole.rs:126-144): F . R_x on the diagonal, F_alpha dd on
        // the s-p_alpha entries (mu = 1 + alpha, nu = 0 in the packed order)
        const double* rx = g.xyz + (md.atom_off + x) * 3;
        if (mu == nu) u += md.field[0] * rx[0] + md.field[1] * rx[1] + md.field[2] * rx[2];
        else if (nu == 0) u += md.field[mu - 1] * ex.dd;
        int o = ms.off[x];
        H[(size_t)(o + mu) * nao + o + nu] = u + acc;
        H[(size_t)(o + nu) * nao + o + mu] = u + acc;
    }
    if (own_pairs) return ecore_own;  // the workgroup computed its pairs itself (service launch)
    double ecore = 0.0;
    const int64_t p0 = g.pair_off[mi], p1 = g.pair_off[mi + 1];
    for (int64_t p = p0 + t; p < p1; p += kWg) ecore += g.ecore_pair[p];
    ecore = wg_sum(ecore, s);
    return ecore;
}

// Far-field potential at atom x from the atoms y = y0, y0 + dy, ... (FarField::potential,
// farfield.rs:150-176): 27.21 / sqrt(r^2 + (rho0_x + rho0_y)^2) * Q_y summed over the atoms
// farther than md.far_cut, with the net charge Q_y = Z_y minus y's population on Pt's diagonal
// (pbc::ewald::net_charges).
__device__ double far_potential(const MolDesc& md, const DevElem* el, const MolShared& ms, const double* xyz,
                                const double* Pt, int x, int y0, int dy) {
    const int nao = md.nao;
    const double cut2 = md.far_cut * md.far_cut;
    const double rx = el[ms.z[x]].rho0;
    const double x0 = xyz[3 * x], x1 = xyz[3 * x + 1], x2 = xyz[3 * x + 2];
    double acc = 0.0;
    for (int y = y0; y < md.nat; y += dy) {
        const double d0 = xyz[3 * y] - x0, d1 = xyz[3 * y + 1] - x1, d2 = xyz[3 * y + 2] - x2;
        const double r2 = d0 * d0 + d1 * d1 + d2 * d2;
        if (y == x || r2 <= cut2) continue;
        const DevElem& ey = el[ms.z[y]];
        const int oy = ms.off[y];
        double pop = 0.0;
        for (int k = 0; k < ey.norb; ++k) pop += Pt[(size_t)(oy + k) * nao + oy + k];
        const double eta = rx + ey.rho0;
        acc += kEv / sqrt(r2 + eta * eta) * (ey.core - pop);
    }
    return acc;
}

// F = H + one-centre(P, P_spin) + J(P) - K(scale * P_spin) for one spin channel
// (fock.rs:249-511). Diagonal atom blocks: one thread per (atom, packed orbital pair), summing
// over partner atoms in index order. Off-diagonal blocks: one thread per pair.
__device__ __noinline__ void mol_fock(const MolDesc& md, const DevElem* el, const MolShared& ms, const double* wpool,
                         const double* H, const double* Pt, const double* Ps, double scale, double* F,
                         double* pdg, const double* xyz) {
    const int t = threadIdx.x;
    const int nao = md.nao;
    const int nh = ms.nh, nH = ms.nH;
    const int n_items = nh * 10 + nH;
    const int64_t n_hh = (int64_t)nh * (nh - 1) / 2, n_hH = (int64_t)nh * nH, n_HH = (int64_t)nH * (nH - 1) / 2;
    // every atom's diagonal density block (pdg: n_items doubles of shared scratch): heavy atoms'
    // 10 packed entries with the off-diagonal ones doubled, in heavy-list order, then one entry
    // per hydrogen; read by all of the atom's partners below
    for (int idx = t; idx < n_items; idx += kWg) {
        double v;
        if (idx < nh * 10) {
            const int y = ms.heavy[idx / 10], q = idx % 10, oy = ms.off[y];
            const int l = kPackRow[q], c = kPackCol[q];
            v = Pt[(size_t)(oy + l) * nao + oy + c] * (l == c ? 1.0 : 2.0);
        } else {
            const int oy = ms.off[ms.hyd[idx - nh * 10]];
            v = Pt[(size_t)oy * nao + oy];
        }
        pdg[idx] = v;
    }
    __syncthreads();
    const double* pdh = pdg + nh * 10;
    const double* wHh = wpool + n_hh * 100;  // heavy-hydrogen blocks (10 each), heavy atom as bra
    const double* wHH = wHh + n_hH * 10;     // hydrogen-hydrogen (1 each)
    // diagonal blocks, G threads per entry (as many as the workgroup allows, up to 8): the
    // partners are split among them and their sums combined in a fixed order. Partners are
    // walked by class (heavy atoms below and above, then hydrogens), so each loop has a fixed
    // shape and its loads issue together.
    int G = 1;
    while (G < 8 && n_items * (2 * G) <= kWg) G *= 2;
    const int gl = t & (G - 1);
    for (int it = t / G; it < n_items; it += kWg / G) {
        int x, k;
        double acc = 0.0;
        if (it < nh * 10) {
            const int ix = it / 10;
            x = ms.heavy[ix];
            k = it % 10;
#pragma unroll 4
            for (int iy = gl; iy < ix; iy += G) {  // x is the ket: column k of the block
                const double* w = wpool + tri_index(iy, ix, nh) * 100 + k;
                const double* py = pdg + iy * 10;
#pragma unroll
                for (int q = 0; q < 10; ++q) acc += py[q] * w[q * 10];
            }
#pragma unroll 4
            for (int iy = ix + 1 + gl; iy < nh; iy += G) {  // x is the bra: row k
                const double* w = wpool + tri_index(ix, iy, nh) * 100 + k * 10;
                const double* py = pdg + iy * 10;
#pragma unroll
                for (int q = 0; q < 10; ++q) acc += py[q] * w[q];
            }
            const double* w = wHh + (int64_t)ix * nH * 10 + k;
#pragma unroll 4
            for (int m = gl; m < nH; m += G) acc += pdh[m] * w[m * 10];
        } else {
            const int mx = it - nh * 10;
            x = ms.hyd[mx];
            k = 0;
#pragma unroll 4
            for (int iy = gl; iy < nh; iy += G) {
                const double* w = wHh + ((int64_t)iy * nH + mx) * 10;
                const double* py = pdg + iy * 10;
#pragma unroll
                for (int q = 0; q < 10; ++q) acc += py[q] * w[q];
            }
#pragma unroll 4
            for (int my = gl; my < mx; my += G) acc += pdh[my] * wHH[tri_index(my, mx, nH)];
#pragma unroll 4
            for (int my = mx + 1 + gl; my < nH; my += G) acc += pdh[my] * wHH[tri_index(mx, my, nH)];
        }
        const DevElem& ex = el[ms.z[x]];
        const int ox = ms.off[x], nx = ex.norb;
        const int mu = kPackRow[k], nu = kPackCol[k];
        // far-field shift -V_x on the atom's diagonal (fock.rs:278-318), partners split the same way
        if (md.far_cut > 0.0 && mu == nu) acc -= far_potential(md, el, ms, xyz, Pt, x, gl, G);
        for (int off = 1; off < G; off <<= 1) acc += __shfl_xor(acc, off, 64);
        if (gl == 0) {
            acc += H[(size_t)(ox + mu) * nao + ox + nu];
            // one-centre (the atom's density block is loaded at once)
            if (nx == 4) {
                double pt[16], ps[16];
#pragma unroll
                for (int q = 0; q < 16; ++q) {
                    pt[q] = Pt[(size_t)(ox + (q >> 2)) * nao + ox + (q & 3)];
                    ps[q] = Ps[(size_t)(ox + (q >> 2)) * nao + ox + (q & 3)];
                }
#pragma unroll
                for (int la = 0; la < 4; ++la)
#pragma unroll
                    for (int si = 0; si < 4; ++si) {
                        acc += pt[la * 4 + si] * one_centre(mu, nu, la, si, ex);
                        acc -= scale * ps[la * 4 + si] * one_centre(mu, la, nu, si, ex);
                    }
            } else {
                acc += Pt[(size_t)ox * nao + ox] * one_centre(mu, nu, 0, 0, ex);
                acc -= scale * Ps[(size_t)ox * nao + ox] * one_centre(mu, 0, nu, 0, ex);
            }
            F[(size_t)(ox + mu) * nao + ox + nu] = acc;
            F[(size_t)(ox + nu) * nao + ox + mu] = acc;
        }
    }
    // exchange blocks, one pair per thread, grouped by class
    for (int64_t k = t; k < n_hh; k += kWg) {
        int i, j;
        tri_decode(k, nh, i, j);
        int a = ms.heavy[i], b = ms.heavy[j];
        int oa = ms.off[a], ob = ms.off[b];
        const double* w = wpool + k * 100;
        double pab[16], acc[16];
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                pab[u * 4 + v] = scale * Ps[(size_t)(oa + u) * nao + ob + v];
                acc[u * 4 + v] = 0.0;
            }
        // K[mu][la] = sum_{nu,si} Ps[nu][si] (mu nu | la si): walk the 10 packed bra rows once,
        // each serving (mu, nu) and, off the diagonal, (nu, mu). The rows are loaded five at a
        // time, so each half of the block costs one memory round trip.
#pragma unroll
        for (int half = 0; half < 2; ++half) {
            double rows[50];
#pragma unroll
            for (int q = 0; q < 50; ++q) rows[q] = w[half * 50 + q];
#pragma unroll
            for (int r = 0; r < 5; ++r) {
                const int kb = half * 5 + r;
                const int mu = kPackRowH[kb], nu = kPackColH[kb];
#pragma unroll
                for (int la = 0; la < 4; ++la) {
                    double x = 0.0, y = 0.0;
#pragma unroll
                    for (int si = 0; si < 4; ++si) {
                        const double wv = rows[r * 10 + pack4(la, si)];
                        x += pab[nu * 4 + si] * wv;
                        y += pab[mu * 4 + si] * wv;
                    }
                    acc[mu * 4 + la] += x;
                    if (mu != nu) acc[nu * 4 + la] += y;
                }
            }
        }
#pragma unroll
        for (int mu = 0; mu < 4; ++mu)
#pragma unroll
            for (int la = 0; la < 4; ++la) {
                F[(size_t)(oa + mu) * nao + ob + la] = H[(size_t)(oa + mu) * nao + ob + la] - acc[mu * 4 + la];
                F[(size_t)(ob + la) * nao + oa + mu] = H[(size_t)(ob + la) * nao + oa + mu] - acc[mu * 4 + la];
            }
    }
    // heavy-hydrogen and hydrogen-hydrogen blocks: a few pairs per thread per round, every load
    // of a round issued before its stores (for the compiler F may alias the inputs, so loads
    // after a store would wait for it)
    for (int64_t k0 = t; k0 < n_hH; k0 += 2 * kWg) {
        double pab[2][4], wv[2][10], hv[2][8];
        int oa[2], ob[2];
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            const int64_t k = min(k0 + (int64_t)u * kWg, n_hH - 1);
            const int i = (int)(k / nH), m = (int)(k - (int64_t)i * nH);
            oa[u] = ms.off[ms.heavy[i]];
            ob[u] = ms.off[ms.hyd[m]];
            const double* w = wpool + n_hh * 100 + k * 10;
#pragma unroll
            for (int q = 0; q < 10; ++q) wv[u][q] = w[q];
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                pab[u][v] = scale * Ps[(size_t)(oa[u] + v) * nao + ob[u]];
                hv[u][v] = H[(size_t)(oa[u] + v) * nao + ob[u]];
                hv[u][4 + v] = H[(size_t)ob[u] * nao + oa[u] + v];
            }
        }
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            if (k0 + (int64_t)u * kWg >= n_hH) continue;
#pragma unroll
            for (int mu = 0; mu < 4; ++mu) {
                double acc = 0.0;
#pragma unroll
                for (int nu = 0; nu < 4; ++nu) acc += pab[u][nu] * wv[u][pack4(mu, nu)];
                F[(size_t)(oa[u] + mu) * nao + ob[u]] = hv[u][mu] - acc;
                F[(size_t)ob[u] * nao + oa[u] + mu] = hv[u][4 + mu] - acc;
            }
        }
    }
    for (int64_t k0 = t; k0 < n_HH; k0 += 4 * kWg) {
        double pv[4], wv[4], h1[4], h2[4];
        int oa[4], ob[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int64_t k = min(k0 + (int64_t)u * kWg, n_HH - 1);
            int i, j;
            tri_decode(k, nH, i, j);
            oa[u] = ms.off[ms.hyd[i]];
            ob[u] = ms.off[ms.hyd[j]];
            pv[u] = Ps[(size_t)oa[u] * nao + ob[u]];
            wv[u] = wpool[n_hh * 100 + n_hH * 10 + k];
            h1[u] = H[(size_t)oa[u] * nao + ob[u]];
            h2[u] = H[(size_t)ob[u] * nao + oa[u]];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            if (k0 + (int64_t)u * kWg >= n_HH) continue;
            const double acc = scale * pv[u] * wv[u];
            F[(size_t)oa[u] * nao + ob[u]] = h1[u] - acc;
            F[(size_t)ob[u] * nao + oa[u]] = h2[u] - acc;
        }
    }
}

// LDS of one workgroup: 150 KiB of dynamic space (the static part holds the molecule tables).
// The eigensolver keeps A (packed lower triangle) there and, while it fits, the eigenvector
// matrix too; the GEMMs use the whole region as panel scratch when the eigensolver is idle.
// Dynamic LDS of the batched kernel in doubles: 150 KiB beside the static part (one workgroup per
// CU); the small-molecule build takes 68 KiB (two per CU).
#ifndef AM1G_LDS_DOUBLES
#define AM1G_LDS_DOUBLES 19200
#endif
constexpr int kLdsDoubles = AM1G_LDS_DOUBLES;
constexpr int kALds = 192;  // largest order diagonalized with A in LDS (n(n+1)/2 <= kLdsDoubles)

// P = w * C_occ C_occ^T from transposed eigenvectors Ct (row k = MO k): P = w Ct_occ^T Ct_occ.
template <typename T>
__device__ void mol_density(int nao, int nocc, double w, const double* Ct, double* P, double* lds) {
    if (nocc == 0) {
        for (int idx = threadIdx.x; idx < nao * nao; idx += kWg) P[idx] = 0.0;
        return;
    }
    wg_gemm_any<T, true, false>(nao, nao, nocc, w, Ct, nao, Ct, nao, P, nao, lds, kLdsDoubles);
}

// One spin channel's eigen-step. Orders up to kTriMax: the tridiagonal method (wg_eigh_tri).
// Larger orders, and the rare matrix whose close eigenvalues that method cannot separate: A =
// Ct F Ct^T (warm start in the previous orbitals, or F itself on the first pass), Jacobi, sort
// ascending. Ct is replaced by the sorted eigenvectors (transposed) and eps by the sorted
// eigenvalues. warm: Ct holds the previous orbitals, all of them (the Jacobi warm start).
// tol_abs: Jacobi's rotation threshold; bracket: half-width around the previous eigenvalues in
// eps for a warm eigenvalue search (tridiagonal method; 0 = none). nvec: the tridiagonal method
// forms only the nvec lowest eigenvectors (rows nvec.. of Ct are left as they were; every
// eigenvalue is still computed), and partial tells whether it did (nvec < nao). Returns the
// Jacobi sweep count (0 for the tridiagonal method), or -1.
template <typename T>
__device__ int mol_eigen(int nao, const double* Fuse, double* Ct, double* T0, double* T1, double* lds, int lds_n,
                         int tri_max, bool warm, double* eps, WgScratch& s, double tol_abs, double bracket, int nvec,
                         bool& partial) {
    const int t = threadIdx.x;
    const int n = nao;
    const size_t nn = (size_t)n * n;
    const size_t tri = (size_t)n * (n + 1) / 2;
    const bool a_lds = n <= kALds && n <= lds_n;
    // eigenvector rows [0, v_rows) beside A in LDS, the rest in place in Ct
    const int v_rows = a_lds ? (int)min((size_t)n, ((size_t)kLdsDoubles - tri) / (size_t)n) : 0;
    const bool v_lds = a_lds && v_rows == n;
    if (n <= tri_max && n <= kTriMax && n <= lds_n && tri_lds_doubles(n) + n <= kLdsDoubles) {
        // direct method (tridiagonalization); it declines only for unresolved close eigenvalues
        if (wg_eigh_tri(n, Fuse, Ct, eps, T1, lds, kLdsDoubles, s, bracket, nvec) == 0) {
            partial = nvec < n;
            return 0;
        }
    } else if (n > kTriMax && n <= tri_max && n <= kTriBig) {
        // the same method with the matrix in global memory (T0: the matrix, then the
        // eigenvectors of T; T1: the reflectors)
        if (wg_eigh_tri_g(n, Fuse, Ct, eps, T0, T1, lds, kLdsDoubles, s, bracket) == 0) {
            partial = false;
            return 0;
        }
    }
    const double* src = Fuse;
    if (warm) {
        // T0 = Ct F ; T1 = T0 Ct^T = the Fock matrix in the previous orbital basis
        wg_gemm_any<T, false, false>(n, n, n, 1.0, Ct, n, Fuse, n, T0, n, lds, kLdsDoubles);
        __syncthreads();
        wg_gemm_any<T, false, true>(n, n, n, 1.0, T0, n, Ct, n, T1, n, lds, kLdsDoubles);
        __syncthreads();
        src = T1;
    }
    int sweeps;
    if (a_lds) {
        // packed lower triangle, symmetrized; eigenvectors start from the previous ones
        for (size_t k = t; k < tri; k += kWg) {
            int i = (int)((sqrt(8.0 * (double)k + 1.0) - 1.0) * 0.5);
            while ((size_t)i * (i + 1) / 2 > k) --i;
            while ((size_t)(i + 1) * (i + 2) / 2 <= k) ++i;
            int j = (int)(k - (size_t)i * (i + 1) / 2);
            lds[k] = (i == j) ? src[(size_t)i * n + i] : 0.5 * (src[(size_t)i * n + j] + src[(size_t)j * n + i]);
        }
        for (size_t idx = t; idx < nn; idx += kWg) {
            int i = (int)(idx / n), j = (int)(idx - (size_t)i * n);
            const double v = warm ? Ct[idx] : (i == j ? 1.0 : 0.0);
            if (i < v_rows) lds[tri + idx] = v;
            else if (!warm) Ct[idx] = v;
        }
        __syncthreads();
        sweeps = wg_jacobi_lds(n, (int)tri, v_rows, Ct, s, 60, tol_abs);
    } else {
        double* A = T1;
        if (!warm)
            for (size_t idx = t; idx < nn; idx += kWg) {
                int i = (int)(idx / n), j = (int)(idx - (size_t)i * n);
                A[idx] = Fuse[idx];
                Ct[idx] = (i == j) ? 1.0 : 0.0;
            }
        else
            for (size_t idx = t; idx < nn; idx += kWg) {
                int i = (int)(idx / n), j = (int)(idx - (size_t)i * n);
                if (i < j) {
                    double v = 0.5 * (A[idx] + A[(size_t)j * n + i]);
                    A[idx] = v;
                    A[(size_t)j * n + i] = v;
                }
            }
        __syncthreads();
        sweeps = wg_jacobi_big(n, A, n, Ct, n, s, 60, tol_abs);
    }
    __syncthreads();
    // sort by rank (stable on ties) and write the permuted eigenvectors out
    double* dst = v_lds ? Ct : T0;
    for (int i = t / 16; i < n; i += kWg / 16) {
        const double ei = a_lds ? lds[(size_t)i * (i + 1) / 2 + i] : T1[(size_t)i * n + i];
        int rank = 0;
        for (int j = t % 16; j < n; j += 16) {
            const double ej = a_lds ? lds[(size_t)j * (j + 1) / 2 + j] : T1[(size_t)j * n + j];
            rank += (ej < ei) || (ej == ei && j < i);
        }
        for (int off = 8; off > 0; off >>= 1) rank += __shfl_xor(rank, off, 16);
        if (t % 16 == 0) eps[rank] = ei;
        const double* vi = i < v_rows ? lds + tri + (size_t)i * n : Ct + (size_t)i * n;
        for (int c = t % 16; c < n; c += 16) dst[(size_t)rank * n + c] = vi[c];
    }
    __syncthreads();
    if (!v_lds) {
        for (size_t idx = t; idx < nn; idx += kWg) Ct[idx] = T0[idx];
        __syncthreads();
    }
    partial = false;
    return sweeps;
}

__device__ __forceinline__ int64_t tri_pos(int i, int j, int n) {  // upper incl. diagonal, i <= j
    return (int64_t)i * n - (int64_t)i * (i - 1) / 2 + (j - i);
}

// Row and column of entry p of the packed upper triangle of order n (diagonal included,
// row-major): the row from a single-precision estimate, corrected exactly in integers.
__device__ __forceinline__ void tri_rowcol(int64_t p, int n, int& i, int& j) {
    const float b = 2.0f * (float)n + 1.0f;
    int r = (int)((b - sqrtf(fmaxf(b * b - 8.0f * (float)p, 0.0f))) * 0.5f);
    r = max(0, min(r, n - 1));
    while (r > 0 && tri_pos(r, r, n) > p) --r;
    while (r + 1 < n && tri_pos(r + 1, r + 1, n) <= p) ++r;
    i = r;
    j = r + (int)(p - tri_pos(r, r, n));
}

// The same for the strict upper triangle of order n (the packed error vectors), which is laid
// out like the full upper triangle of order n - 1 with columns shifted by one.
__device__ __forceinline__ void stri_rowcol(int64_t p, int n, int& i, int& j) {
    tri_rowcol(p, n - 1, i, j);
    j += 1;
}

// ---------------------------------------------------------------------------------------------
// DIIS bookkeeping for the SCF loops. History slots hold packed F (n(n+1)/2) and the packed
// error (strict upper triangle, n(n-1)/2). The A-DIIS products with past densities come from
// F and cached scalars instead of stored densities (see the RHF Gram update).
struct DiisView {
    double* f;
    double* e;
    int64_t len_s, len_a;
};

__device__ __forceinline__ DiisView diis_view(double* hist, int nao, int channel) {
    DiisView v;
    v.len_s = tri_len(nao);
    v.len_a = (int64_t)nao * (nao - 1) / 2;
    double* base = hist + (int64_t)channel * kDiisDepth * 2 * v.len_s;
    v.f = base;
    v.e = base + kDiisDepth * v.len_s;
    return v;
}

// Uniform values into scalar registers (a device function receives its arguments in vector
// registers, and values read from LDS are vector values too).
__device__ __forceinline__ int uni(int x) { return __builtin_amdgcn_readfirstlane(x); }
template <typename T>
__device__ __forceinline__ T* uni(T* p) {
    const uint64_t v = (uint64_t)p;
    const uint64_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)v);
    const uint64_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(v >> 32));
    return (T*)((hi << 32) | lo);
}

// Walk of a packed upper triangle of order n by wavefronts: wavefront w takes rows w, w + kWaves,
// ... and each row in chunks of 64 columns (lane l on the chunk's column l); the rounds below take
// kTriSlots chunks at a time and issue all of their loads before any arithmetic, so a wavefront
// waits for memory once per round. With the diagonal (DIAG), row i holds columns i..n-1 at packed
// offset i n - i (i - 1) / 2; without it (the strict triangle) columns i+1..n-1 at offset
// i (n - 1) - i (i - 1) / 2. The cursor (row, chunk start) is uniform across the wavefront.
constexpr int kTriSlots = 4;
template <bool DIAG>
__device__ __forceinline__ int tri_row_len(int i, int n) { return DIAG ? n - i : n - 1 - i; }
template <bool DIAG>
__device__ __forceinline__ void tri_cursor_next(int& ri, int& ci, int n) {
    if (ri < n - (DIAG ? 0 : 1)) {
        ci += 64;
        if (ci >= tri_row_len<DIAG>(ri, n)) {
            ri += kWaves;
            ci = 0;
        }
    }
}
// Slot entry: packed offset and (row, column); an idle slot (past the end) points at entry (0, 0)
// of the triangle ((0, 1) without the diagonal), so its loads stay in bounds.
template <bool DIAG>
__device__ __forceinline__ bool tri_slot(int ri, int ci, int n, int& pk, int& i, int& j) {
    const int off = ci + (threadIdx.x & 63);
    const bool ok = ri < n - (DIAG ? 0 : 1) && off < tri_row_len<DIAG>(ri, n);
    i = ok ? ri : 0;
    const int o = ok ? off : 0;
    j = i + o + (DIAG ? 0 : 1);
    pk = (DIAG ? i * n : i * (n - 1)) - i * (i - 1) / 2 + o;
    return ok;
}

// RHF DIIS push (the history of scf.rs:560-606): writes the new entry (the packed error
// e_ij = T_ij - T_ji of T = FP over the strict upper triangle, and the packed F) into history slot
// (head + newest) % kDiisDepth and leaves, for each of the 2 kDiisDepth + 2 sums below, one partial
// per wavefront in buf[q * kWaves + w] (the caller adds the kWaves partials in order):
// q < D: <e_new, e_q> (CDIIS Gram row; q = newest gives ||e||^2); with need_d, D + q:
// <D_new, F_q> and 2 D: <D_new, H>; 2 D + 1: E_elec = 1/2 sum P (H + F). Sums over the packed
// triangle count off-diagonal entries once, so diagonal ones enter at half weight (the caller
// doubles). History slot q is (head + q) % D for q < newest; the ones past the history read the
// oldest slot and are discarded (in the first iteration that is the new entry's own slot).
__device__ __noinline__ void diis_push_rhf(int nao_, const double* T0_, const double* F_, const double* H_,
                                           const double* P_, double* he_base_, double* hf_base_, int len_a_,
                                           int len_s_, int head_, int newest_, int need_d_, double* buf) {
    constexpr int D = kDiisDepth, S = kTriSlots;
    const int nao = uni(nao_), head = uni(head_), newest = uni(newest_), len_a = uni(len_a_), len_s = uni(len_s_);
    const bool need_d = uni(need_d_) != 0;
    const double* T0 = uni(T0_);
    const double* F = uni(F_);
    const double* H = uni(H_);
    const double* P = uni(P_);
    double* he_base = uni(he_base_);
    double* hf_base = uni(hf_base_);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const double* ep[D];
    const double* fp[D];
#pragma unroll
    for (int q = 0; q < D; ++q) {
        const int r = (head + (q < newest ? q : 0)) % D;
        ep[q] = he_base + (int64_t)r * len_a;
        fp[q] = hf_base + (int64_t)r * len_s;
    }
    const int rn = (head + newest) % D;
    double* __restrict__ he = he_base + (int64_t)rn * len_a;
    double* __restrict__ hf = hf_base + (int64_t)rn * len_s;
    double part[2 * D + 2];
#pragma unroll
    for (int q = 0; q < 2 * D + 2; ++q) part[q] = 0.0;
    for (int ri = w, ci = 0; ri < nao - 1;) {
        int pk[S];
        bool ok[S];
        double ta[S], tb[S], hv[S][D];
#pragma unroll
        for (int u = 0; u < S; ++u) {
            int i, j;
            ok[u] = tri_slot<false>(ri, ci, nao, pk[u], i, j);
            tri_cursor_next<false>(ri, ci, nao);
            ta[u] = T0[i * nao + j];
            tb[u] = T0[j * nao + i];
#pragma unroll
            for (int q = 0; q < D; ++q) hv[u][q] = ep[q][pk[u]];
        }
#pragma unroll
        for (int u = 0; u < S; ++u) {
            if (ok[u]) {
                const double e = ta[u] - tb[u];
                he[pk[u]] = e;
#pragma unroll
                for (int q = 0; q < D; ++q) part[q] = fma(e, q < newest ? hv[u][q] : (q == newest ? e : 0.0), part[q]);
            }
        }
    }
    for (int ri = w, ci = 0; ri < nao;) {
        int pk[S];
        bool ok[S], dg[S];
        double fv[S], hv[S], pv[S], qv[S][D];
#pragma unroll
        for (int u = 0; u < S; ++u) {
            int i, j;
            ok[u] = tri_slot<true>(ri, ci, nao, pk[u], i, j);
            tri_cursor_next<true>(ri, ci, nao);
            dg[u] = i == j;
            fv[u] = F[i * nao + j];
            hv[u] = H[i * nao + j];
            pv[u] = P[i * nao + j];
            if (need_d) {
#pragma unroll
                for (int q = 0; q < D; ++q) qv[u][q] = fp[q][pk[u]];
            }
        }
#pragma unroll
        for (int u = 0; u < S; ++u) {
            if (ok[u]) {
                const double f = fv[u], h = hv[u];
                const double pw = dg[u] ? 0.5 * pv[u] : pv[u];
                hf[pk[u]] = f;
                part[2 * D + 1] = fma(pw, h + f, part[2 * D + 1]);
                if (need_d) {
                    part[2 * D] = fma(pw, h, part[2 * D]);
#pragma unroll
                    for (int q = 0; q < D; ++q)
                        part[D + q] = fma(pw, q < newest ? qv[u][q] : (q == newest ? f : 0.0), part[D + q]);
                }
            }
        }
    }
#pragma unroll
    for (int q = 0; q < 2 * D + 2; ++q) {
        const double x = wave_sum(part[q]);
        if (lane == 0) buf[q * kWaves + w] = x;
    }
    __syncthreads();
}

// F <- sum_{q < m} c_q F_q (scf.rs:627-653) from the packed history (slots ring[q]), written to
// both triangles of the dense F; returns 1.0 on a thread that produced a non-finite entry.
__device__ __noinline__ double diis_combine(int nao_, double* F_, const double* hf_base_, int len_s_,
                                            const int* ring, const double* coef, int m_) {
    constexpr int D = kDiisDepth, S = kTriSlots;
    const int nao = uni(nao_), len_s = uni(len_s_), m = uni(m_);
    double* F = uni(F_);
    const double* hf_base = uni(hf_base_);
    const int w = threadIdx.x >> 6;
    const double* fq[D];
    double cq[D];
#pragma unroll
    for (int q = 0; q < D; ++q) {
        fq[q] = hf_base + (int64_t)uni(ring[q < m ? q : m - 1]) * len_s;
        cq[q] = q < m ? coef[q] : 0.0;
    }
    double bad = 0.0;
    for (int ri = w, ci = 0; ri < nao;) {
        int pk[S], ii[S], jj[S];
        bool ok[S];
        double v[S][D];
#pragma unroll
        for (int u = 0; u < S; ++u) {
            ok[u] = tri_slot<true>(ri, ci, nao, pk[u], ii[u], jj[u]);
            tri_cursor_next<true>(ri, ci, nao);
#pragma unroll
            for (int q = 0; q < D; ++q) v[u][q] = fq[q][pk[u]];
        }
#pragma unroll
        for (int u = 0; u < S; ++u) {
            double acc = 0.0;
#pragma unroll
            for (int q = 0; q < D; ++q) acc = fma(cq[q], q < m ? v[u][q] : 0.0, acc);
            if (ok[u]) {
                F[ii[u] * nao + jj[u]] = acc;
                F[jj[u] * nao + ii[u]] = acc;
                if (!isfinite(acc)) bad = 1.0;
            }
        }
    }
    return bad;
}

// P <- Pn over n entries; returns this thread's part of sum (Pn - P)^2. Eight entries per thread
// per round, all loaded before any store, so each round costs one memory round trip.
__device__ __forceinline__ double density_update(const double* __restrict__ Pn, double* __restrict__ P, size_t n) {
    constexpr int U = 8;
    double dd = 0.0;
    for (size_t base = threadIdx.x; base < n; base += (size_t)U * kWg) {
        double a[U], b[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const size_t idx = base + (size_t)u * kWg;
            a[u] = idx < n ? Pn[idx] : 0.0;
            b[u] = idx < n ? P[idx] : 0.0;
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const size_t idx = base + (size_t)u * kWg;
            const double x = a[u] - b[u];
            dd += x * x;
            if (idx < n) P[idx] = a[u];
        }
    }
    return dd;
}

// Per-phase device clock, kept by thread 0 (phases: 0 setup, 1 Fock + energy, 2 commutator +
// DIIS, 3 eigensolve, 4 density + convergence, 5 final energy and properties).
struct PhaseClock {
    uint64_t last;
    uint64_t acc[kPhases];
};

__device__ __forceinline__ void phase_mark(PhaseClock& pc, int k) {
    if (threadIdx.x == 0) {
        uint64_t now = __builtin_amdgcn_s_memrealtime();
        pc.acc[k] += now - pc.last;
        pc.last = now;
    }
}

// ---------------------------------------------------------------------------------------------
template <int PREC>
__device__ __noinline__ void scf_rhf(const MolDesc& md, const BatchArgs& g, MolShared& ms, WgScratch& s, double* lds,
                        double ecore, MolResult& res, PhaseClock& pc) {
    const int t = threadIdx.x;
    const int nao = md.nao, nat = md.nat, nocc = md.n_alpha;
    const DevElem* el = g.elems + md.method * kMaxZ;
    double* M = g.mats + md.mat_off;
    const size_t nn = (size_t)nao * nao;
    double* H = M + kH * nn;
    double* F = M + kF * nn;
    double* P = M + kP * nn;
    double* Ct = M + kC * nn;
    double* T0 = M + kT0 * nn;
    double* T1 = M + kT1 * nn;
    double* wpool = g.wpool + md.w_off;
    double* eps = g.mo_e + md.ao_off;
    DiisView dv = diis_view(g.hist + md.hist_off, nao, 0);
    const int accel = md.accel;
    const bool need_d = accel == 2;

    // SAD guess (scf.rs:475-494)
    for (size_t idx = t; idx < nn; idx += kWg) P[idx] = 0.0;
    __syncthreads();
    for (int a = t; a < nat; a += kWg) {
        const DevElem& e = el[ms.z[a]];
        double ns = fmin(e.core, 2.0), np = fmax(e.core - ns, 0.0);
        int o = ms.off[a];
        P[(size_t)o * nao + o] = ns;
        if (e.norb == 4)
            for (int k = 1; k < 4; ++k) P[(size_t)(o + k) * nao + o + k] = np / 3.0;
    }
    if (t == 0) {
        ms.hcount = 0;
        ms.head = 0;
    }
    __syncthreads();

    double e_old = 0.0, residual = INFINITY;
    int iterations = 0, converged = 0, status = kOk;
    bool have_c = false;
    int fp32_iters = 0, total_sweeps = 0;
    // mixed precision: FP32 GEMMs until ||[F,P]|| < 1e-4, or until the FP32-measured error stops
    // improving for two iterations (its rounding floor); FP64 from then on
    bool fp32_phase = (PREC == 1) && md.precision == 1;
    double fp32_best = INFINITY;
    int fp32_stall = 0;
    // iterations run with FP64 GEMMs; convergence needs two of them, so that the energy change and
    // the density change compare two FP64 densities (after an FP32 phase as well)
    int fp64_iters = 0;
    int loose_iters = 0;  // FP64 iterations that met the p_tol test
    for (int iter = 0; iter < md.max_scf; ++iter) {
        iterations = iter + 1;
        mol_fock(md, el, ms, wpool, H, P, P, 0.5, F, lds, g.xyz + md.atom_off * 3);
        __syncthreads();
        phase_mark(pc, 1);
        // error = FP - (FP)^T; F and P are symmetric so FP = F^T P
        const bool low = fp32_phase;
        if (!low) ++fp64_iters;
        if (low) {
            wg_gemm_any<float, true, false>(nao, nao, nao, 1.0, F, nao, P, nao, T0, nao, lds, kLdsDoubles);
            ++fp32_iters;
        } else {
            wg_gemm_any<double, true, false>(nao, nao, nao, 1.0, F, nao, P, nao, T0, nao, lds, kLdsDoubles);
        }
        __syncthreads();
        // History ring after this push (every thread derives it; the shared copy follows the
        // sums): the new entry takes logical position m - 1, replacing the oldest when full.
        const bool full = ms.hcount == kDiisDepth;
        const int head = full ? (ms.head + 1) % kDiisDepth : ms.head;
        const int m = full ? kDiisDepth : ms.hcount + 1;
        const int newest = m - 1;
        // Two passes over the packed triangles push the new entry and form, in one reduction:
        // <e_new, e_q> (CDIIS Gram row; q = newest gives ||e||^2), and for A-DIIS <D_new, F_q>,
        // <D_new, H>; and E_elec = 1/2 sum P (H + F). A matrix product <X, Y> over the packed upper
        // triangle counts off-diagonal entries twice, so diagonal ones enter at half weight and the
        // sum is doubled. The remaining A-DIIS entries need no density history: F_j = H + G(D_j)
        // with G linear and symmetric (<A, G(B)> = <B, G(A)>), so <D_j, F_new> = h_j + <D_new, F_j>
        // - h_new.
        diis_push_rhf(nao, T0, F, H, P, dv.e, dv.f, (int)dv.len_a, (int)dv.len_s, head, newest, need_d, lds);
        // The cached Gram entries move one step when the oldest entry is dropped: wavefront 0
        // reads them here, before the reduction's barriers, and writes them after, lane l
        // moving entry (l / 8, l % 8) from (l / 8 + 1, l % 8 + 1). Only stores follow the
        // barriers, so no lane reads an entry another lane rewrites.
        const int gi = (t & 63) >> 3, gj = t & 7;
        const bool gmove = full && t < 64 && gi < kDiisDepth - 1 && gj < kDiisDepth - 1;
        const double bsh = gmove ? ms.B[(gi + 1) * kDiisDepth + gj + 1] : 0.0;
        const double gsh = gmove ? ms.G[(gi + 1) * kDiisDepth + gj + 1] : 0.0;
        double hsh[kDiisDepth];  // <D_j, H> in the new logical order
#pragma unroll
        for (int j = 0; j < kDiisDepth; ++j) hsh[j] = full ? (j + 1 < kDiisDepth ? ms.hd[j + 1] : 0.0) : ms.hd[j];
        // the totals of the push's partial sums (LDS is free outside the eigensolver)
        double tot[2 * kDiisDepth + 2];
#pragma unroll
        for (int q = 0; q < 2 * kDiisDepth + 2; ++q) {
            double x = 0.0;
#pragma unroll
            for (int w = 0; w < kWaves; ++w) x += lds[q * kWaves + w];
            tot[q] = x;
        }
        __syncthreads();
        double bnn = 0.0;
#pragma unroll
        for (int q = 0; q < kDiisDepth; ++q)
            if (q == newest) bnn = tot[q];
        const double err = sqrt(2.0 * bnn);
        const double e_elec = tot[2 * kDiisDepth + 1];
        if (gmove) {
            ms.B[gi * kDiisDepth + gj] = bsh;
            ms.G[gi * kDiisDepth + gj] = gsh;
        }
        if (t == 0) {
            ms.hcount = m;
            ms.head = head;
            const double hn = 2.0 * tot[2 * kDiisDepth];
#pragma unroll
            for (int i = 0; i < kDiisDepth; ++i) {
                if (i >= m) continue;
                ms.ring[i] = (head + i) % kDiisDepth;
                ms.B[newest * kDiisDepth + i] = 2.0 * tot[i];
                ms.B[i * kDiisDepth + newest] = 2.0 * tot[i];
                if (need_d) {
                    const double ti = 2.0 * tot[kDiisDepth + i];
                    ms.hd[i] = i == newest ? hn : hsh[i];
                    ms.G[newest * kDiisDepth + i] = ti;                                      // <D_new, F_i>
                    ms.G[i * kDiisDepth + newest] = i == newest ? ti : hsh[i] + ti - hn;  // <D_i, F_new>
                }
            }
        }
        __syncthreads();
        if (!isfinite(e_elec) || !isfinite(err)) {
            status = kScfDiverged;
            break;
        }
        // extrapolation coefficients, scf.rs:627-653: CDIIS and A-DIIS each on wavefront 0
        const bool cdiis = accel == 1 || (accel == 2 && err <= md.adiis_switch);
        if (cdiis) {
            if (t < 64) {
                const bool use = m >= 2 && cdiis_solve_wave(ms.B, m, ms.coef, true);
                if (t == 0) ms.flags[0] = use ? 1 : 0;
            }
        } else if (accel == 2 && m >= 2) {
            if (t < 64) adiis_solve_wave(ms.G, m, ms.coef);
            if (t == 0) ms.flags[0] = 1;
        } else if (t == 0) {
            ms.flags[0] = 0;
        }
        __syncthreads();
        if (ms.flags[0]) {
            // F <- sum_i c_i F_i (unpacked) in place: F itself is history slot `newest` and is
            // rebuilt next pass; a non-finite combination puts that slot back
            double bad = diis_combine(nao, F, dv.f, (int)dv.len_s, ms.ring, ms.coef, m);
            bad = wg_max(bad, s);
            if (bad != 0.0) {
                const double* fn = dv.f + (int64_t)ms.ring[m - 1] * dv.len_s;
                for (int64_t pk = t; pk < dv.len_s; pk += kWg) {
                    int i, j;
                    tri_rowcol(pk, nao, i, j);
                    F[(size_t)i * nao + j] = fn[pk];
                    F[(size_t)j * nao + i] = fn[pk];
                }
            }
        }
        const double* Fuse = F;
        __syncthreads();
        phase_mark(pc, 2);
        // the occupied orbitals are all the density needs; every one when the caller takes the
        // coefficients
        bool partial = false;
        int sweeps = mol_eigen<double>(nao, Fuse, Ct, T0, T1, lds, g.lds_n, g.tri_max, have_c, eps, s, 1.0e-7 * err,
                                       iter > 0 ? 100.0 * err : 0.0, md.all_vecs ? nao : nocc, partial);
        phase_mark(pc, 3);
        total_sweeps += sweeps;
        have_c = !partial;
        if (sweeps < 0) {
            status = kLinearAlgebra;
            break;
        }
        // new density into T0, rms change, then P <- T0
        if (low) mol_density<float>(nao, nocc, 2.0, Ct, T0, lds);
        else mol_density<double>(nao, nocc, 2.0, Ct, T0, lds);
        __syncthreads();
        const double dd = density_update(T0, P, nn);
        double dp = sqrt(wg_sum(dd, s) / (double)max((size_t)1, nn));
        phase_mark(pc, 4);
        double de = fabs(e_elec - e_old);
        e_old = e_elec;
        residual = err;
        if (fp32_phase) {
            if (err < fp32_best) {
                fp32_best = err;
                fp32_stall = 0;
            } else {
                ++fp32_stall;
            }
            if (err <= 1.0e-4 || fp32_stall >= 2) fp32_phase = false;
        }
        // scf.rs:693 (RHF), :820 (UHF). A mixed-precision run takes another trajectory than the FP64 one, and the
        // density-change test at p_tol stops up to ~1e-6 e short of the fixed point, so it ends by
        // a test 100x tighter; should that one stall (a near-degenerate open shell can hold the
        // error on a plateau), ten iterations after the p_tol test is met end the run as well.
        const bool conv_loose = de < md.e_tol && (dp < md.p_tol || err < 1.0e-7);
        const bool conv_tight = de < md.e_tol && (dp < 1.0e-2 * md.p_tol || err < 1.0e-7);
        if (conv_loose && fp64_iters >= 2) ++loose_iters;
        if (fp64_iters >= 2 && (md.precision == 1 ? (conv_tight || loose_iters > 10) : conv_loose)) {
            converged = 1;
            break;
        }
        __syncthreads();
    }
    __syncthreads();
    // final energy from the final density (scf.rs:698-701)
    mol_fock(md, el, ms, wpool, H, P, P, 0.5, F, lds, g.xyz + md.atom_off * 3);
    __syncthreads();
    double pe = 0.0;
    for (size_t idx = t; idx < nn; idx += kWg) pe += P[idx] * (H[idx] + F[idx]);
    double electronic = 0.5 * wg_sum(pe, s);
    // Mulliken charges and dipole (scf.rs:400-427)
    double dip[3] = {0.0, 0.0, 0.0};
    double eiso = 0.0, eheat = 0.0, ezv = 0.0;
    const double* xyz = g.xyz + md.atom_off * 3;
    for (int a = t; a < nat; a += kWg) {
        const DevElem& e = el[ms.z[a]];
        if (md.far_cut > 0.0) ezv += e.core * far_potential(md, el, ms, xyz, P, a, 0, 1);
        int o = ms.off[a];
        double pop = 0.0;
        for (int k = 0; k < e.norb; ++k) pop += P[(size_t)(o + k) * nao + o + k];
        double q = e.core - pop;
        g.charges[md.atom_off + a] = q;
        for (int k = 0; k < 3; ++k) dip[k] += xyz[3 * a + k] * q;
        if (e.norb == 4)
            for (int k = 0; k < 3; ++k) dip[k] += -2.0 * e.dd * P[(size_t)o * nao + o + 1 + k];
        eiso += e.e_isol;
        eheat += e.eheat_ev;
    }
    for (int k = 0; k < 3; ++k) dip[k] = wg_sum(dip[k], s);
    eiso = wg_sum(eiso, s);
    eheat = wg_sum(eheat, s);
    // the far-field energy term +1/2 sum_a Z_a V_a of the final energy (scf.rs:699-701)
    if (md.far_cut > 0.0) electronic += 0.5 * wg_sum(ezv, s);
    if (t == 0) {
        res.status = (status == kOk && !converged) ? kScfNotConverged : status;
        res.iterations = iterations;
        res.converged = converged;
        res.nao = nao;
        res.electronic_ev = electronic;
        res.core_ev = ecore;
        res.total_ev = electronic + ecore + md.field_nuc;
        res.heat_of_formation_kcal = (electronic + ecore + md.field_nuc - eiso + eheat) * kEvToKcal;
        res.homo_ev = nocc >= 1 ? eps[nocc - 1] : NAN;
        res.lumo_ev = nocc < nao ? eps[nocc] : NAN;
        res.homo_beta_ev = NAN;
        res.lumo_beta_ev = NAN;
        for (int k = 0; k < 3; ++k) res.dipole[k] = dip[k] * kAuDipoleToDebye;
        res.residual = residual;
        res.fp32_iterations = fp32_iters;
        res.eigen_sweeps = total_sweeps;
    }
    phase_mark(pc, 5);
}


// Unrestricted SCF (am1-rs uhf_loop, scf.rs:719-858): F_sigma = H + J(Pa + Pb) - K(P_sigma),
// E = 1/2 (Pt.H + Pa.Fa + Pb.Fb); CDIIS on the stacked error [Ea ; Eb] with one coefficient set
// for both spins (no A-DIIS); the divergence check comes after the extrapolation; rms changes of
// the two spin densities are summed.
template <int PREC>
__device__ __noinline__ void scf_uhf(const MolDesc& md, const BatchArgs& g, MolShared& ms, WgScratch& s,
                                     double* lds, double ecore, MolResult& res, PhaseClock& pc) {
    const int t = threadIdx.x;
    const int nao = md.nao, nat = md.nat, na = md.n_alpha, nb = md.n_beta;
    const DevElem* el = g.elems + md.method * kMaxZ;
    double* M = g.mats + md.mat_off;
    const size_t nn = (size_t)nao * nao;
    double* H = M + kH * nn;
    double* T0 = M + kT0 * nn;
    double* T1 = M + kT1 * nn;
    double* Pt = M + kPt * nn;
    double* Fs[2] = {M + kF * nn, M + kFb * nn};
    double* Ps[2] = {M + kP * nn, M + kPb * nn};
    double* Cs[2] = {M + kC * nn, M + kCb * nn};
    double* eps[2] = {g.mo_e + md.ao_off, g.mo_e_beta + md.ao_off};
    const int nocc[2] = {na, nb};
    double* wpool = g.wpool + md.w_off;
    DiisView dv[2] = {diis_view(g.hist + md.hist_off, nao, 0), diis_view(g.hist + md.hist_off, nao, 1)};
    const bool diis = md.use_diis != 0;  // UHF runs CDIIS whenever DIIS is on (scf.rs:777-797)

    // SAD guess split by the alpha / beta electron fractions (scf.rs:730-740)
    const double ntot = (double)max(na + nb, 1);
    const double frac[2] = {na / ntot, nb / ntot};
    for (size_t idx = t; idx < nn; idx += kWg) Ps[0][idx] = Ps[1][idx] = 0.0;
    __syncthreads();
    for (int a = t; a < nat; a += kWg) {
        const DevElem& e = el[ms.z[a]];
        double nsv = fmin(e.core, 2.0), npv = fmax(e.core - nsv, 0.0);
        int o = ms.off[a];
        for (int c = 0; c < 2; ++c) {
            Ps[c][(size_t)o * nao + o] = nsv * frac[c];
            if (e.norb == 4)
                for (int k = 1; k < 4; ++k) Ps[c][(size_t)(o + k) * nao + o + k] = npv / 3.0 * frac[c];
        }
    }
    if (t == 0) {
        ms.hcount = 0;
        ms.head = 0;
    }
    __syncthreads();
    double e_old = 0.0, residual = INFINITY, err = INFINITY;
    int iterations = 0, converged = 0, status = kOk, fp32_iters = 0, total_sweeps = 0;
    bool have_c = false;
    bool fp32_phase = (PREC == 1) && md.precision == 1;
    double fp32_best = INFINITY;
    int fp32_stall = 0;
    // iterations run with FP64 GEMMs; convergence needs two of them, so that the energy change and
    // the density change compare two FP64 densities (after an FP32 phase as well)
    int fp64_iters = 0;
    int loose_iters = 0;  // FP64 iterations that met the p_tol test
    for (int iter = 0; iter < md.max_scf; ++iter) {
        iterations = iter + 1;
        for (size_t idx = t; idx < nn; idx += kWg) Pt[idx] = Ps[0][idx] + Ps[1][idx];
        __syncthreads();
        mol_fock(md, el, ms, wpool, H, Pt, Ps[0], 1.0, Fs[0], lds, g.xyz + md.atom_off * 3);
        mol_fock(md, el, ms, wpool, H, Pt, Ps[1], 1.0, Fs[1], lds, g.xyz + md.atom_off * 3);
        __syncthreads();
        double pe = 0.0;
        for (size_t idx = t; idx < nn; idx += kWg)
            pe += Pt[idx] * H[idx] + Ps[0][idx] * Fs[0][idx] + Ps[1][idx] * Fs[1][idx];
        const double e_elec = 0.5 * wg_sum(pe, s);
        phase_mark(pc, 1);
        const bool low = fp32_phase;
        if (!low) ++fp64_iters;
        if (low) ++fp32_iters;
        int slot = 0;
        if (diis) slot = ms.hcount < kDiisDepth ? (ms.head + ms.hcount) % kDiisDepth : ms.head;
        double en = 0.0;
        for (int c = 0; c < 2; ++c) {
            if (low)
                wg_gemm_any<float, true, false>(nao, nao, nao, 1.0, Fs[c], nao, Ps[c], nao, T0, nao, lds, kLdsDoubles);
            else
                wg_gemm_any<double, true, false>(nao, nao, nao, 1.0, Fs[c], nao, Ps[c], nao, T0, nao, lds, kLdsDoubles);
            __syncthreads();
            double* hf = dv[c].f + (int64_t)slot * dv[c].len_s;
            double* he = dv[c].e + (int64_t)slot * dv[c].len_a;
#pragma unroll 4
            for (int64_t pk = t; pk < dv[c].len_a; pk += kWg) {
                int i, j;
                stri_rowcol(pk, nao, i, j);
                const double e = T0[(size_t)i * nao + j] - T0[(size_t)j * nao + i];
                en += e * e;
                if (diis) he[pk] = e;
            }
            if (diis)
                for (int64_t pk = t; pk < dv[c].len_s; pk += kWg) {
                    int i, j;
                    tri_rowcol(pk, nao, i, j);
                    hf[pk] = Fs[c][(size_t)i * nao + j];
                }
            __syncthreads();
        }
        err = sqrt(2.0 * wg_sum(en, s));
        const double* Fuse[2] = {Fs[0], Fs[1]};
        if (diis) {
            // ring update; when full the cached Gram matrix moves one step: wavefront 0 reads the
            // moved entries (lane l: entry (l / 8 + 1, l % 8 + 1)) before the barrier and writes
            // them after it
            const bool full = ms.hcount == kDiisDepth;
            const int gi = (t & 63) >> 3, gj = t & 7;
            const bool gmove = full && t < 64 && gi < kDiisDepth - 1 && gj < kDiisDepth - 1;
            const double bsh = gmove ? ms.B[(gi + 1) * kDiisDepth + gj + 1] : 0.0;
            __syncthreads();
            if (t == 0) {
                if (!full) ms.hcount++;
                else ms.head = (ms.head + 1) % kDiisDepth;
                for (int i = 0; i < ms.hcount; ++i) ms.ring[i] = (ms.head + i) % kDiisDepth;
            }
            if (gmove) ms.B[gi * kDiisDepth + gj] = bsh;
            __syncthreads();
            const int m = ms.hcount, newest = m - 1;
            double part[kDiisDepth];
#pragma unroll
            for (int q = 0; q < kDiisDepth; ++q) part[q] = 0.0;
            for (int c = 0; c < 2; ++c) {
                const double* ep[kDiisDepth];
#pragma unroll
                for (int i = 0; i < kDiisDepth; ++i) ep[i] = dv[c].e + (int64_t)ms.ring[i < m ? i : newest] * dv[c].len_a;
                const double* en_p = ep[newest];
                // every slot pointer is valid (past the history: the newest), so all loads issue
                // without branches and the selects discard the extra ones
#pragma unroll 4
                for (int64_t pk = t; pk < dv[c].len_a; pk += kWg) {
                    const double x = en_p[pk];
#pragma unroll
                    for (int i = 0; i < kDiisDepth; ++i) {
                        const double v = ep[i][pk];
                        part[i] = fma(x, i < m ? v : 0.0, part[i]);
                    }
                }
            }
            double tot[kDiisDepth];
            wg_sum_many(part, lds, tot);
            if (t == 0) {
#pragma unroll
                for (int i = 0; i < kDiisDepth; ++i) {
                    if (i >= m) continue;
                    ms.B[newest * kDiisDepth + i] = 2.0 * tot[i];
                    ms.B[i * kDiisDepth + newest] = 2.0 * tot[i];
                }
            }
            __syncthreads();
            if (t < 64) {  // CDIIS on one wavefront (the serial solve's pivoting; any finite or not)
                const bool use = m >= 2 && cdiis_solve_wave(ms.B, m, ms.coef, false);
                if (t == 0) ms.flags[0] = use ? 1 : 0;
            }
            __syncthreads();
            if (ms.flags[0]) {
                // combined Focks in place of F (each channel's own F is its newest history slot
                // and is rebuilt next pass)
                double cq[kDiisDepth];
#pragma unroll
                for (int q = 0; q < kDiisDepth; ++q) cq[q] = q < m ? ms.coef[q] : 0.0;
                for (int c = 0; c < 2; ++c) {
                    const double* fq[kDiisDepth];
#pragma unroll
                    for (int q = 0; q < kDiisDepth; ++q) fq[q] = dv[c].f + (int64_t)ms.ring[q < m ? q : m - 1] * dv[c].len_s;
                    // four entries per thread per round, every load before any store
                    constexpr int U = 4;
                    for (int64_t base = t; base < dv[c].len_s; base += (int64_t)U * kWg) {
                        double v[U][kDiisDepth];
#pragma unroll
                        for (int u = 0; u < U; ++u) {
                            const int64_t pk = min(base + (int64_t)u * kWg, dv[c].len_s - 1);
#pragma unroll
                            for (int q = 0; q < kDiisDepth; ++q) v[u][q] = fq[q][pk];
                        }
#pragma unroll
                        for (int u = 0; u < U; ++u) {
                            const int64_t pk = base + (int64_t)u * kWg;
                            if (pk >= dv[c].len_s) continue;
                            double acc = 0.0;
#pragma unroll
                            for (int q = 0; q < kDiisDepth; ++q) acc = fma(cq[q], q < m ? v[u][q] : 0.0, acc);
                            int i, j;
                            tri_rowcol(pk, nao, i, j);
                            Fs[c][(size_t)i * nao + j] = acc;
                            Fs[c][(size_t)j * nao + i] = acc;
                        }
                    }
                }
                __syncthreads();
            }
        }
        if (!isfinite(e_elec) || !isfinite(err)) {
            status = kScfDiverged;
            break;
        }
        phase_mark(pc, 2);
        bool full_c = true;
        for (int c = 0; c < 2; ++c) {
            bool partial = false;
            int sw = mol_eigen<double>(nao, Fuse[c], Cs[c], T0, T1, lds, g.lds_n, g.tri_max, have_c, eps[c], s,
                                      1.0e-7 * err, iter > 0 ? 100.0 * err : 0.0,
                                      md.all_vecs ? nao : (c == 0 ? md.n_alpha : md.n_beta), partial);
            if (sw < 0) status = kLinearAlgebra;
            else total_sweeps += sw;
            full_c = full_c && !partial;
        }
        have_c = full_c;
        phase_mark(pc, 3);
        if (status != kOk) break;
        double dsum[2];
        for (int c = 0; c < 2; ++c) {
            if (low) mol_density<float>(nao, nocc[c], 1.0, Cs[c], T0, lds);
            else mol_density<double>(nao, nocc[c], 1.0, Cs[c], T0, lds);
            __syncthreads();
            const double dd = density_update(T0, Ps[c], nn);
            dsum[c] = sqrt(wg_sum(dd, s) / (double)max((size_t)1, nn));
        }
        const double dp = dsum[0] + dsum[1];
        phase_mark(pc, 4);
        const double de = fabs(e_elec - e_old);
        e_old = e_elec;
        residual = err;
        if (fp32_phase) {
            if (err < fp32_best) {
                fp32_best = err;
                fp32_stall = 0;
            } else {
                ++fp32_stall;
            }
            if (err <= 1.0e-4 || fp32_stall >= 2) fp32_phase = false;
        }
        // scf.rs:693 (RHF), :820 (UHF). A mixed-precision run takes another trajectory than the FP64 one, and the
        // density-change test at p_tol stops up to ~1e-6 e short of the fixed point, so it ends by
        // a test 100x tighter; should that one stall (a near-degenerate open shell can hold the
        // error on a plateau), ten iterations after the p_tol test is met end the run as well.
        const bool conv_loose = de < md.e_tol && (dp < md.p_tol || err < 1.0e-7);
        const bool conv_tight = de < md.e_tol && (dp < 1.0e-2 * md.p_tol || err < 1.0e-7);
        if (conv_loose && fp64_iters >= 2) ++loose_iters;
        if (fp64_iters >= 2 && (md.precision == 1 ? (conv_tight || loose_iters > 10) : conv_loose)) {
            converged = 1;
            break;
        }
        __syncthreads();
    }
    __syncthreads();
    for (size_t idx = t; idx < nn; idx += kWg) Pt[idx] = Ps[0][idx] + Ps[1][idx];
    __syncthreads();
    mol_fock(md, el, ms, wpool, H, Pt, Ps[0], 1.0, Fs[0], lds, g.xyz + md.atom_off * 3);
    mol_fock(md, el, ms, wpool, H, Pt, Ps[1], 1.0, Fs[1], lds, g.xyz + md.atom_off * 3);
    __syncthreads();
    double pe = 0.0;
    for (size_t idx = t; idx < nn; idx += kWg) pe += Pt[idx] * H[idx] + Ps[0][idx] * Fs[0][idx] + Ps[1][idx] * Fs[1][idx];
    double electronic = 0.5 * wg_sum(pe, s);
    double dip[3] = {0.0, 0.0, 0.0}, eiso = 0.0, eheat = 0.0, ezv = 0.0;
    const double* xyz = g.xyz + md.atom_off * 3;
    for (int a = t; a < nat; a += kWg) {
        const DevElem& e = el[ms.z[a]];
        if (md.far_cut > 0.0) ezv += e.core * far_potential(md, el, ms, xyz, Pt, a, 0, 1);
        int o = ms.off[a];
        double pop = 0.0;
        for (int k = 0; k < e.norb; ++k) pop += Pt[(size_t)(o + k) * nao + o + k];
        double q = e.core - pop;
        g.charges[md.atom_off + a] = q;
        for (int k = 0; k < 3; ++k) dip[k] += xyz[3 * a + k] * q;
        if (e.norb == 4)
            for (int k = 0; k < 3; ++k) dip[k] += -2.0 * e.dd * Pt[(size_t)o * nao + o + 1 + k];
        eiso += e.e_isol;
        eheat += e.eheat_ev;
    }
    for (int k = 0; k < 3; ++k) dip[k] = wg_sum(dip[k], s);
    eiso = wg_sum(eiso, s);
    eheat = wg_sum(eheat, s);
    // the far-field energy term +1/2 sum_a Z_a V_a of the final energy (scf.rs:699-701)
    if (md.far_cut > 0.0) electronic += 0.5 * wg_sum(ezv, s);
    if (t == 0) {
        res.status = (status == kOk && !converged) ? kScfNotConverged : status;
        res.iterations = iterations;
        res.converged = converged;
        res.nao = nao;
        res.electronic_ev = electronic;
        res.core_ev = ecore;
        res.total_ev = electronic + ecore + md.field_nuc;
        res.heat_of_formation_kcal = (electronic + ecore + md.field_nuc - eiso + eheat) * kEvToKcal;
        res.homo_ev = na >= 1 ? eps[0][na - 1] : NAN;
        res.lumo_ev = na < nao ? eps[0][na] : NAN;
        res.homo_beta_ev = nb >= 1 ? eps[1][nb - 1] : NAN;
        res.lumo_beta_ev = nb < nao ? eps[1][nb] : NAN;
        for (int k = 0; k < 3; ++k) res.dipole[k] = dip[k] * kAuDipoleToDebye;
        res.residual = residual;
        res.fp32_iterations = fp32_iters;
        res.eigen_sweeps = total_sweeps;
    }
    phase_mark(pc, 5);
}

// Service launch: the next published ticket for this workgroup, or -1 once the host sets the stop
// flag, or once the launch has been idle for svc_idle_us: no workgroup busy and nothing started,
// taken or finished since (one deadline for all workgroups, so the launch empties at once rather
// than thinning out while requests still come). Thread 0 only. Every wait is bounded: the
// compare-and-swap fails only when another workgroup took a ticket, and the busy workgroups finish
// their requests (each SCF is bounded by max_scf).
__device__ long long svc_take(const BatchArgs& g) {
    const uint64_t limit = (uint64_t)g.svc_idle_us * 100u;  // 100 MHz device clock
    for (;;) {
        if (__hip_atomic_load(g.svc_stop, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM)) return -1;
        const long long pub = __hip_atomic_load(g.svc_published, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_SYSTEM);
        long long cur = __hip_atomic_load(g.svc_next, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        while (cur < pub) {
            if (__hip_atomic_compare_exchange_strong(g.svc_next, &cur, cur + 1, __ATOMIC_ACQ_REL, __ATOMIC_RELAXED,
                                                     __HIP_MEMORY_SCOPE_AGENT)) {
                __hip_atomic_fetch_max(g.svc_last, (unsigned long long)__builtin_amdgcn_s_memrealtime(),
                                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                return cur;
            }
        }
        const unsigned long long last = __hip_atomic_load(g.svc_last, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const unsigned long long now = __builtin_amdgcn_s_memrealtime();
        if (now > last && now - last > limit &&
            __hip_atomic_load(g.svc_busy, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0)
            return -1;
        // about 14 us between polls of the pinned host flags
        __builtin_amdgcn_s_sleep(127);
        __builtin_amdgcn_s_sleep(127);
        __builtin_amdgcn_s_sleep(127);
        __builtin_amdgcn_s_sleep(127);
    }
}

template <int PREC>
// Two wavefronts per SIMD: one 512-thread workgroup per CU, or two 256-thread ones (the
// small-molecule build), each wavefront within 256 registers.
__global__ __launch_bounds__(kWg, 2) void k_scf_batch(BatchArgs g) {
    double* lds = lds_dyn;
    __shared__ MolShared ms;
    __shared__ WgScratch s;
    __shared__ int next;
    __shared__ long long ticket;
    __shared__ MolDesc smd;
    const bool svc = g.svc_in != nullptr;
    if (svc && threadIdx.x == 0) {
        __hip_atomic_fetch_add(g.svc_alive, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
        __hip_atomic_fetch_max(g.svc_last, (unsigned long long)__builtin_amdgcn_s_memrealtime(), __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
    }
    for (;;) {
        int mi;
        MolDesc md;
        double ecore_own = 0.0;
        if (svc) {
            if (threadIdx.x == 0) {
                ticket = svc_take(g);
                if (ticket >= 0) {
                    __hip_atomic_fetch_add(g.svc_busy, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    smd = g.svc_in[ticket % g.svc_cap].md;
                }
            }
            __syncthreads();
            const long long tk = ticket;
            if (tk < 0) break;
            mi = (int)(tk % g.svc_cap);
            md = smd;
            __syncthreads();
            // Pool regions are reused by later requests on other XCDs: drop this XCD's cached
            // lines of them before use (the previous user wrote its lines back when it finished).
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
            const SvcSlot& sl = g.svc_in[mi];
            int* z = const_cast<int*>(g.z) + md.atom_off;
            int* aoff = const_cast<int*>(g.aoff) + md.atom_off;
            int* acls = const_cast<int*>(g.acls) + md.atom_off;
            int* alist = const_cast<int*>(g.alist) + md.atom_off;
            double* xyz = const_cast<double*>(g.xyz) + md.atom_off * 3;
            for (int a = threadIdx.x; a < md.nat; a += kWg) {
                z[a] = sl.z[a];
                aoff[a] = sl.aoff[a];
                acls[a] = sl.acls[a];
                alist[a] = sl.alist[a];
            }
            for (int k = threadIdx.x; k < 3 * md.nat; k += kWg) xyz[k] = sl.xyz[k];
            __syncthreads();
            ecore_own = pairs_wg(g, md, s);
            __syncthreads();
        } else {
            if (threadIdx.x == 0) next = atomicAdd(g.queue, 1);
            __syncthreads();
            const int slot = next;
            __syncthreads();
            if (slot >= g.n_mols) break;
            mi = g.order ? g.order[slot] : slot;
            md = g.mols[mi];
            if (g.own_pairs) ecore_own = pairs_wg(g, md, s);  // instead of a k_pairs launch before
        }
        MolResult res;
        PhaseClock pc;
        res.t_start = __builtin_amdgcn_s_memrealtime();
        pc.last = res.t_start;
        for (int k = 0; k < kPhases; ++k) pc.acc[k] = 0;
        double* M = g.mats + md.mat_off;
        double ecore = mol_setup(md, mi, g, ms, s, M + kH * (size_t)md.nao * md.nao, g.wpool + md.w_off,
                                 svc || g.own_pairs, ecore_own);
        __syncthreads();
        phase_mark(pc, 0);
        if (md.uhf) scf_uhf<PREC>(md, g, ms, s, lds, ecore, res, pc);
        else scf_rhf<PREC>(md, g, ms, s, lds, ecore, res, pc);
        if (threadIdx.x == 0) {
            for (int k = 0; k < kPhases; ++k) res.phase[k] = pc.acc[k];
            res.t_done = __builtin_amdgcn_s_memrealtime();
            g.results[mi] = res;
        }
        __syncthreads();
        if (svc) {
            // results into the host-pinned ring, then the slot's done flag (system scope, after the
            // data and after this XCD's cached pool lines are written back)
            SvcOut& out = g.svc_out[mi];
            for (int a = threadIdx.x; a < md.nat; a += kWg) out.charges[a] = g.charges[md.atom_off + a];
            for (int k = threadIdx.x; k < md.nao; k += kWg) {
                out.mo[k] = g.mo_e[md.ao_off + k];
                if (md.uhf) out.mob[k] = g.mo_e_beta[md.ao_off + k];
            }
            if (threadIdx.x == 0) out.res = res;
            __threadfence_system();
            __syncthreads();
            if (threadIdx.x == 0) {
                __hip_atomic_store(g.svc_done + mi, ticket + 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
                __hip_atomic_fetch_max(g.svc_last, (unsigned long long)__builtin_amdgcn_s_memrealtime(),
                                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                __hip_atomic_fetch_add(g.svc_busy, -1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            __syncthreads();
        } else if (g.done) {
            // publish this molecule to the host while the launch goes on
            for (int a = threadIdx.x; a < md.nat; a += kWg)
                g.charges_h[md.atom_off + a] = g.charges[md.atom_off + a];
            for (int k = threadIdx.x; k < md.nao; k += kWg) {
                g.mo_e_h[md.ao_off + k] = g.mo_e[md.ao_off + k];
                if (md.uhf) g.mo_eb_h[md.ao_off + k] = g.mo_e_beta[md.ao_off + k];
            }
            if (threadIdx.x == 0) g.results_h[mi] = res;
            __threadfence_system();
            __syncthreads();
            if (threadIdx.x == 0) __hip_atomic_store(g.done + mi, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
        }
    }
    if (svc && threadIdx.x == 0) __hip_atomic_fetch_add(g.svc_alive, -1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
}

template __global__ void k_scf_batch<0>(BatchArgs);
template __global__ void k_scf_batch<1>(BatchArgs);

}  // namespace am1g

namespace am1g {
#ifndef AM1G_WG_VARIANT  // the small-molecule build needs only the batched kernel and its launch
hipError_t launch_pairs_one(const BatchArgs& a, const MolDesc& md, hipStream_t stream) {
    const int64_t n_hh = (int64_t)md.n_heavy * (md.n_heavy - 1) / 2, n_hH = (int64_t)md.n_heavy * md.n_h;
    auto blocks = [](int64_t n, int per) { return dim3((unsigned)std::min<int64_t>((n + per - 1) / per, 65536)); };
    // one kernel per class, each compiled for its class alone; the heavy-H blocks (80 bytes) are
    // staged through LDS so that their stores coalesce (measured: staging the 800-byte
    // heavy-heavy blocks as well cost more in occupancy than it saved)
    if (n_hh > 0) hipLaunchKernelGGL(k_pairs_range<0>, blocks(n_hh, 256), dim3(256), 0, stream, a, (int64_t)0, n_hh);
    if (n_hH > 0) hipLaunchKernelGGL(k_pairs_cls<1>, blocks(n_hH, 64), dim3(64), 0, stream, a);
    if (a.n_pairs > n_hh + n_hH)
        hipLaunchKernelGGL(k_pairs_range<2>, blocks(a.n_pairs - n_hh - n_hH, 256), dim3(256), 0, stream, a,
                           n_hh + n_hH, a.n_pairs);
    // the resonance blocks were written above the diagonal only: mirror them (a tiled transpose,
    // coalesced on both sides, instead of a scattered store per entry)
    const int nt = (md.nao + 31) / 32;
    if (md.nao > 1)
        hipLaunchKernelGGL(k_sym_lower, dim3((unsigned)(nt * (nt + 1) / 2)), dim3(256), 0, stream, md.nao,
                           a.mats + md.mat_off + kH * (size_t)md.nao * md.nao);
    return hipGetLastError();
}

hipError_t launch_pairs(const BatchArgs& a, hipStream_t stream, int block) {
    if (a.n_pairs == 0) return hipSuccess;
    block = block == 64 || block == 128 ? block : 256;  // the kernel's grid-stride loop takes any
    int64_t blocks = (a.n_pairs + block - 1) / block;
    if (blocks > 65536) blocks = 65536;
    hipLaunchKernelGGL(k_pairs, dim3((unsigned)blocks), dim3(block), 0, stream, a);
    return hipGetLastError();
}

// Device scratch of the allocation-free batched entry point: one array per GPU, reserved when
// the code object is loaded and held until the process ends. Host code never passes `scratch`
// itself to a kernel (on the host it names a stand-in, not device memory); it uses the device
// address that hipGetSymbolAddress returns, looked up once per device by scratch_init.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

hipError_t scratch_init(int device) {
    if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[device]) return hipSuccess;
    int cur = -1;
    hipError_t e = hipGetDevice(&cur);
    if (e != hipSuccess) return e;
    if (cur != device) return hipErrorInvalidDevice;  // the caller makes `device` current first
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    if (!p) return hipErrorInvalidSymbol;
    g_scratch_ptr[device] = p;
    return hipSuccess;
}

void* scratch_ptr(int device) {
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    return g_scratch_ptr[device];
}

__global__ void k_zero_ints(int* p, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

hipError_t launch_zero_ints(int* p, int n, hipStream_t stream) {
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_zero_ints, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, stream, p, n);
    return hipGetLastError();
}

int eigen_tri_max() {
    const char* v = std::getenv("AM1G_EIGEN_JACOBI");
    return (v && v[0] == '1') ? 0 : kTriBig;
}

#endif  // AM1G_WG_VARIANT

hipError_t launch_scf_batch(const BatchArgs& a, int precision, int grid, size_t lds_bytes, hipStream_t stream) {
    static bool attr_done = false;
    if (!attr_done) {
        // allow the dynamic LDS region to take what the static part leaves of the 160 KiB
        const int dyn = kLdsDoubles * (int)sizeof(double);
        hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_scf_batch<0>),
                                           hipFuncAttributeMaxDynamicSharedMemorySize, dyn);
        if (e != hipSuccess) return e;
        e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_scf_batch<1>),
                                hipFuncAttributeMaxDynamicSharedMemorySize, dyn);
        if (e != hipSuccess) return e;
        attr_done = true;
    }
    if (precision == 1)
        hipLaunchKernelGGL(k_scf_batch<1>, dim3(grid), dim3(kWg), lds_bytes, stream, a);
    else
        hipLaunchKernelGGL(k_scf_batch<0>, dim3(grid), dim3(kWg), lds_bytes, stream, a);
    return hipGetLastError();
}
}  // namespace am1g

#ifndef AM1G_WG_VARIANT

// =============================================================================================
// Grid-level kernels for one large system: core Hamiltonian diagonal blocks, Fock build and
// reductions (the plain path and divide-and-conquer), plus the divide-and-conquer stages.
namespace am1g {



// Compensated (sum, error) pairs. The large reductions of one big system (E_core over ~1e7
// pairs, Tr P(H+F) over ~1e8 entries) are +-1e8 eV terms that cancel to ~1e6 eV; plain FP64
// summation in any order moves the result by ~1e-5 eV at 6000 atoms, so these sums carry their
// rounding error (Neumaier within a thread, error-free TwoSum when combining) and are accurate to
// a few ulps of the result whatever the grid shape.
struct KSum {
    double s = 0.0, c = 0.0;
};

__device__ __forceinline__ void ksum_add(KSum& k, double x) {
    const double t = k.s + x;
    k.c += fabs(k.s) >= fabs(x) ? (k.s - t) + x : (x - t) + k.s;
    k.s = t;
}

__device__ __forceinline__ KSum ksum_join(KSum a, KSum b) {
    KSum r;
    r.s = a.s + b.s;
    const double bp = r.s - a.s;
    r.c = a.c + b.c + ((a.s - (r.s - bp)) + (b.s - bp));
    return r;
}

__device__ __forceinline__ KSum ksum_block(KSum v) {  // 256 threads, fixed order; result in thread 0
    __shared__ double sh[2][4];
    for (int off = 32; off > 0; off >>= 1) {
        KSum o;
        o.s = __shfl_xor(v.s, off, 64);
        o.c = __shfl_xor(v.c, off, 64);
        v = ksum_join(v, o);
    }
    __syncthreads();
    if ((threadIdx.x & 63) == 0) {
        sh[0][threadIdx.x >> 6] = v.s;
        sh[1][threadIdx.x >> 6] = v.c;
    }
    __syncthreads();
    KSum r;
    if (threadIdx.x == 0) {
        KSum w[4];
        for (int q = 0; q < 4; ++q) {
            w[q].s = sh[0][q];
            w[q].c = sh[1][q];
        }
        r = ksum_join(ksum_join(w[0], w[1]), ksum_join(w[2], w[3]));
    }
    return r;
}

// Diagonal blocks of H_core and of the Fock matrix. Each block needs a sum over every partner y
// of y's 10 packed factors (pdiag, rows in class order: heavy atoms, then hydrogens) against the
// pair's integrals: the Coulomb term sum_q P_y[q] (mu nu | q_y) for the Fock matrix, or the core
// attraction -Z_y (mu nu | s_y s_y) when pdiag holds (-Z_y, 0, ..., 0). Workgroups
// 0..n_heavy-1 take one heavy atom each (its class index); the rest take 8 consecutive hydrogens
// of the hydrogen list. Lanes read consecutive doubles of the pool, so one load instruction
// covers a partner's whole 100-entry block (heavy atom) or the 8 hydrogens' 80 contiguous
// entries against one heavy partner (hydrogen group); the four wavefronts split the partners and
// meet in LDS. For the Fock matrix a heavy atom's heavy partners are not read here:
// k_fock_rows_heavy, which stages those blocks anyway, leaves their terms in rsum and tcol.
// Stores entry k of x's diagonal block and returns its share of sum Ps (H + F) over the whole
// matrix (both (mu, nu) and (nu, mu); zero for H_core).
__device__ __forceinline__ double diag_block_store(const SysArgs& a, int x, int k, double v, bool fock,
                                                   const double* Pt, const double* Ps, double scale, double* F) {
    const DevElem& ex = a.elems[a.md.method * kMaxZ + a.z[x]];
    const int nx = ex.norb, nao = a.md.nao, ox = a.aoff[x];
    const int mu = kPackRow[k], nu = kPackCol[k];
    if (!fock) {
        const double u = (mu == nu) ? (mu == 0 ? ex.uss : ex.upp) : 0.0;
        a.H[(size_t)(ox + mu) * nao + ox + nu] = u + v;
        a.H[(size_t)(ox + nu) * nao + ox + mu] = u + v;
        return 0.0;
    }
    double oc = 0.0;
    for (int la = 0; la < nx; ++la)
        for (int si = 0; si < nx; ++si) {
            oc += Pt[(size_t)(ox + la) * nao + ox + si] * one_centre(mu, nu, la, si, ex);
            oc -= scale * Ps[(size_t)(ox + la) * nao + ox + si] * one_centre(mu, la, nu, si, ex);
        }
    const double h = a.H[(size_t)(ox + mu) * nao + ox + nu], f = h + oc + v;
    F[(size_t)(ox + mu) * nao + ox + nu] = f;
    F[(size_t)(ox + nu) * nao + ox + mu] = f;
    return (mu == nu ? 1.0 : 2.0) * Ps[(size_t)(ox + mu) * nao + ox + nu] * (h + f);
}

// Per-workgroup partial of sum Ps (H + F): the workgroup's compensated sum at slot `slot`.
__device__ __forceinline__ void energy_part(const KSum& v, double* epart, int64_t slot) {
    const KSum r = ksum_block(v);
    if (threadIdx.x == 0) {
        epart[2 * slot] = r.s;
        epart[2 * slot + 1] = r.c;
    }
}

template <bool kFock>
__global__ __launch_bounds__(256) void k_diag_blocks(SysArgs a, const double* Pt, const double* Ps, double scale,
                                                     double* F, double* epart) {
    __shared__ double red[4][260];
    const int nh = a.md.n_heavy, nH = a.md.n_h;
    const int lane = threadIdx.x & 63, w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int64_t n_hh = (int64_t)nh * (nh - 1) / 2, n_hH = (int64_t)nh * nH;
    const double* wHh = a.wpool + n_hh * 100;  // heavy-hydrogen blocks, heavy atom as bra
    const double* wHH = wHh + n_hH * 10;       // hydrogen-hydrogen scalars
    const double* pd = a.pdiag;
    const double* pdH = pd + (size_t)nh * 10;  // hydrogens' rows
    const int* al = a.alist;
    double* r = red[w];
    KSum ev;
    if ((int)blockIdx.x < nh) {
        const int cx = blockIdx.x;
        // hydrogen partners: x's heavy-hydrogen row is nH * 10 contiguous doubles, entry
        // 10 m + k; lanes below 60 take 6 partners per load, lane -> (m offset lane / 10, k lane % 10)
        const int lm = lane / 10, lk = lane % 10;
        const double* wx = wHh + (int64_t)cx * nH * 10;
        double h = 0.0;
        if (lane < 60) {
#pragma unroll 4
            for (int m = w * 6 + lm; m < nH; m += 24) h += pdH[(size_t)m * 10] * wx[(int64_t)m * 10 + lk];
        }
        if (lane < 60) r[200 + lane] = h;
        if constexpr (kFock) {
            // heavy partners: k_fock_rows_heavy has left the terms from partners above x in
            // rsum[x] and those from partners below x in x's column of tcol (cx rows of 10
            // contiguous doubles), read here like the hydrogen row
            const double* tc = a.tcol + (int64_t)cx * (cx - 1) / 2 * 10;
            double c = 0.0;
            if (lane < 60) {
#pragma unroll 4
                for (int m = w * 6 + lm; m < cx; m += 24) c += tc[(int64_t)m * 10 + lk];
                r[lane] = c;
            }
        } else {
            // Block entry e = 10 * (bra pair) + (ket pair). Lane holds e0 = lane and e1 = 64 + lane
            // (lanes below 36; the others repeat e0 and are dropped). With x as the ket (partners
            // below x in the heavy list) the factor index is e / 10 and the target e % 10; with x
            // as the bra (partners above) the reverse.
            const int e0 = lane, e1 = lane < 36 ? 64 + lane : lane;
            const int qb0 = e0 / 10, qb1 = e1 / 10, qa0 = e0 % 10, qa1 = e1 % 10;
            double b0 = 0.0, b1 = 0.0, u0 = 0.0, u1 = 0.0;
#pragma unroll 4
            for (int iy = w; iy < cx; iy += 4) {
                const double* blk = a.wpool + tri_index(iy, cx, nh) * 100;
                const double* py = pd + (size_t)iy * 10;
                b0 += py[qb0] * blk[e0];
                b1 += py[qb1] * blk[e1];
            }
#pragma unroll 4
            for (int iy = cx + 1 + w; iy < nh; iy += 4) {
                const double* blk = a.wpool + tri_index(cx, iy, nh) * 100;
                const double* py = pd + (size_t)iy * 10;
                u0 += py[qa0] * blk[e0];
                u1 += py[qa1] * blk[e1];
            }
            if (lane >= 36) b1 = u1 = 0.0;
            r[lane] = b0;
            r[100 + lane] = u0;
            if (lane < 36) {
                r[64 + lane] = b1;
                r[164 + lane] = u1;
            }
        }
        __syncthreads();
        if (threadIdx.x < 40) {
            const int k = threadIdx.x % 10;
            const double* s = red[threadIdx.x / 10];
            double v = 0.0;
            if constexpr (kFock) {
#pragma unroll
                for (int g = 0; g < 6; ++g) v += s[g * 10 + k] + s[200 + g * 10 + k];
            } else {
#pragma unroll
                for (int q = 0; q < 10; ++q) v += s[q * 10 + k] + s[100 + k * 10 + q];
#pragma unroll
                for (int g = 0; g < 6; ++g) v += s[200 + g * 10 + k];
            }
            v += __shfl(v, threadIdx.x + 10, 64);
            v += __shfl(v, threadIdx.x + 20, 64);
            if constexpr (kFock) v += a.rsum[cx * 10 + k];
            const int x = al[cx];
            const int nk = a.elems[a.md.method * kMaxZ + a.z[x]].norb == 4 ? 10 : 1;
            if (threadIdx.x < nk) ksum_add(ev, diag_block_store(a, x, k, v, kFock, Pt, Ps, scale, F));
        }
    } else {
        // hydrogen group: class indices cx0 .. cx0 + gsz - 1
        const int cx0 = ((int)blockIdx.x - nh) * 8;
        const int gsz = min(8, nH - cx0);
        // heavy partners: the group's entries against heavy iy are 10 * gsz contiguous doubles,
        // entry 10 j + q (hydrogen cx0 + j, factor q of the heavy partner)
        const bool v0 = lane < 10 * gsz, v1 = lane < 16 && 64 + lane < 10 * gsz;
        const int e0 = v0 ? lane : 0, e1 = v1 ? 64 + lane : 0;
        const int q0 = e0 % 10, q1 = e1 % 10;
        double s0 = 0.0, s1 = 0.0, hh = 0.0;
#pragma unroll 4
        for (int iy = w; iy < nh; iy += 4) {
            const double* seg = wHh + ((int64_t)iy * nH + cx0) * 10;
            const double* py = pd + (size_t)iy * 10;
            s0 += py[q0] * seg[e0];
            s1 += py[q1] * seg[e1];
        }
        // hydrogen partners: lane -> (hydrogen cx0 + (lane & 7), partner m = base + (lane >> 3));
        // partners below the group read 8 consecutive entries of row m, partners above it 8
        // consecutive entries of each of the 8 rows
        const int j = lane & 7, x = cx0 + j;
        if (nH > 1) {
#pragma unroll 4
            for (int mb = w * 8; mb < nH; mb += 32) {
                const int m = mb + (lane >> 3);
                const bool ok = j < gsz && m < nH && m != x;
                const int64_t t = ok ? tri_index(min(m, x), max(m, x), nH) : 0;
                const double c = pdH[(size_t)(ok ? m : cx0) * 10];
                const double g = wHH[t];
                hh += ok ? c * g : 0.0;
            }
        }
        if (!v0) s0 = 0.0;
        if (!v1) s1 = 0.0;
        r[lane] = s0;
        if (lane < 16) r[64 + lane] = s1;
        r[80 + lane] = hh;
        __syncthreads();
        if (threadIdx.x < 32) {
            const int jj = threadIdx.x & 7;
            const double* s = red[threadIdx.x >> 3];
            double v = 0.0;
#pragma unroll
            for (int q = 0; q < 10; ++q) v += s[jj * 10 + q];
#pragma unroll
            for (int g = 0; g < 8; ++g) v += s[80 + g * 8 + jj];
            v += __shfl(v, threadIdx.x + 8, 64);
            v += __shfl(v, threadIdx.x + 16, 64);
            if (threadIdx.x < gsz)
                ksum_add(ev, diag_block_store(a, al[nh + cx0 + jj], 0, v, kFock, Pt, Ps, scale, F));
        }
    }
    if (kFock && epart) energy_part(ev, epart, blockIdx.x);
}

// Row of atom x in pdiag: heavy atoms first, then hydrogens, each in class order.
__device__ __forceinline__ double* pdiag_row(const SysArgs& a, int x) {
    return a.pdiag + (size_t)(a.z[x] == 1 ? a.md.n_heavy + a.acls[x] : a.acls[x]) * 10;
}

// Partner factors of the core attraction: (-Z_y, 0, ..., 0) per atom.
__global__ __launch_bounds__(256) void k_pack_core(SysArgs a) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= a.md.nat) return;
    const double c = -a.elems[a.md.method * kMaxZ + a.z[x]].core;
    double* row = pdiag_row(a, x);
#pragma unroll
    for (int q = 0; q < 10; ++q) row[q] = q == 0 ? c : 0.0;
}

// Packed diagonal density blocks, off-diagonal entries doubled (the 2 of the symmetric sum).
__global__ __launch_bounds__(256) void k_pack_diag(SysArgs a, const double* Pt) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= a.md.nat) return;
    const int nao = a.md.nao, o = a.aoff[x];
    double* row = pdiag_row(a, x);
    if (a.z[x] == 1) {
        row[0] = Pt[(size_t)o * nao + o];
        return;
    }
#pragma unroll
    for (int q = 0; q < 10; ++q) {
        const int l = kPackRow[q], c = kPackCol[q];
        row[q] = Pt[(size_t)(o + l) * nao + o + c] * (l == c ? 1.0 : 2.0);
    }
}

// Off-diagonal Fock blocks, H - K, written in the upper triangle only (atom AO ranges are
// disjoint, so a block lies wholly on one side); k_sym_lower then copies the upper triangle to
// the lower in coalesced tiles. P and H are read at the same upper positions. One workgroup per
// row atom walks that atom's matrix rows left to right, so every cache line of P, H and F in the
// rows is fetched once. Both class lists keep the input order, so AO offsets grow along each and
// the partners later in AO order are a suffix of each list (`later` gives where it begins).
constexpr int kHhPairs = 16;

// Orders one wavefront's LDS accesses around this point (lanes read what other lanes wrote).
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Rows of heavy atom ci (workgroup ci). The later heavy partners are taken 16 at a time, one chunk
// per wavefront: their 100-entry integral blocks (contiguous in the pool) are staged through LDS
// with coalesced loads, lane (pair lane / 4, row mu = lane % 4) forms row mu of the exchange
// block, K(mu, la) = sum_{nu, si} P(nu, si) (mu nu | la si), and writes F(mu, la) = H - K. The
// hydrogens whose columns lie among the chunk's follow at once, one lane per hydrogen (4 x 1
// blocks). From the staged block the wavefront also takes both atoms' Coulomb terms for their
// diagonal blocks: ci's (row k of the block against j's packed factors, summed over the row into
// rsum[ci]) and j's (column k against ci's factors, into tcol), so the diagonal-block kernel
// need not read the heavy pair blocks again.
__global__ __launch_bounds__(256) void k_fock_rows_heavy(SysArgs a, const double* Ps, double scale, double* F,
                                                         double* epart, int64_t slot0) {
    __shared__ double stage[4][kHhPairs * 100];
    const int nh = a.md.n_heavy, nH = a.md.n_h, nao = a.md.nao;
    const int64_t n_hh = (int64_t)nh * (nh - 1) / 2;
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    double* st = stage[wv];
    const int pr = lane >> 2, mu = lane & 3;
    const int ci = blockIdx.x, oa = a.aoff[a.alist[ci]];
    const size_t r0 = (size_t)oa * nao;
    const double* wi = a.wpool + n_hh * 100 + (int64_t)ci * nH * 10;  // ci's heavy-hydrogen row
    const int nj = nh - 1 - ci;
    const int nck = max(1, (nj + kHhPairs - 1) / kHhPairs);
    double pci[10], arow[3] = {0.0, 0.0, 0.0};  // ci's factors; lane's terms for k = mu, mu + 4, mu + 8
#pragma unroll
    for (int q = 0; q < 10; ++q) pci[q] = a.pdiag[(size_t)ci * 10 + q];
    KSum ev;
    for (int k = wv; k < nck; k += 4) {
        const int j0 = ci + 1 + k * kHhPairs, j1 = min(j0 + kHhPairs, nh);
        if (j1 > j0) {
            const int nd = (j1 - j0) * 100;
            const double* src = a.wpool + tri_index(ci, j0, nh) * 100;
            double v[kHhPairs * 100 / 64];
#pragma unroll
            for (int t = 0; t < kHhPairs * 100 / 64; ++t) {
                const int e = lane + 64 * t;
                v[t] = src[e < nd ? e : 0];
            }
            wave_lds_sync();  // the previous chunk's LDS reads come first
#pragma unroll
            for (int t = 0; t < kHhPairs * 100 / 64; ++t) {
                const int e = lane + 64 * t;
                if (e < nd) st[e] = v[t];
            }
            wave_lds_sync();
            if (pr < j1 - j0) {
                const int ob = a.aoff[a.alist[j0 + pr]];
                double pab[16], hv[4], kk[4];
#pragma unroll
                for (int nu = 0; nu < 4; ++nu)
#pragma unroll
                    for (int si = 0; si < 4; ++si) pab[nu * 4 + si] = Ps[r0 + (size_t)nu * nao + ob + si];
#pragma unroll
                for (int la = 0; la < 4; ++la) {
                    hv[la] = a.H[r0 + (size_t)mu * nao + ob + la];
                    kk[la] = 0.0;
                }
                const double* blk = st + pr * 100;
#pragma unroll
                for (int nu = 0; nu < 4; ++nu) {
                    const double* row = blk + pack4(mu, nu) * 10;
                    double r[10];
#pragma unroll
                    for (int q = 0; q < 10; ++q) r[q] = row[q];
#pragma unroll
                    for (int la = 0; la < 4; ++la)
#pragma unroll
                        for (int si = 0; si < 4; ++si) kk[la] += pab[nu * 4 + si] * r[pack4(la, si)];
                }
#pragma unroll
                for (int la = 0; la < 4; ++la) {
                    const double f = hv[la] - scale * kk[la];
                    F[r0 + (size_t)mu * nao + ob + la] = f;
                    ksum_add(ev, 2.0 * pab[mu * 4 + la] * (hv[la] + f));
                }
                const int j = j0 + pr;
                double pj[10];
#pragma unroll
                for (int q = 0; q < 10; ++q) pj[q] = a.pdiag[(size_t)j * 10 + q];
                double* tj = a.tcol + ((int64_t)j * (j - 1) / 2 + ci) * 10;
#pragma unroll 1
                for (int t = 0; t < 3; ++t) {
                    const int kq = mu + 4 * t;
                    if (kq < 10) {
                        double sr = 0.0, sc = 0.0;
#pragma unroll
                        for (int q = 0; q < 10; ++q) {
                            sr += pj[q] * blk[kq * 10 + q];
                            sc += pci[q] * blk[q * 10 + kq];
                        }
                        arow[t] += sr;  // t is uniform in the loop; arow stays in registers
                        tj[kq] = sc;
                    }
                }
            }
        }
        // hydrogens after heavy j0 - 1 (ci itself for the first chunk) and before heavy j1
        const int m0 = a.later[j0 - 1], m1 = j1 >= nh ? nH : a.later[j1 - 1];
        for (int m = m0 + lane; m < m1; m += 64) {
            const int ob = a.aoff[a.alist[nh + m]];
            const double* w = wi + (int64_t)m * 10;
            double pab[4], hv[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                pab[u] = Ps[r0 + (size_t)u * nao + ob];
                hv[u] = a.H[r0 + (size_t)u * nao + ob];
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                double acc = 0.0;
#pragma unroll
                for (int nu = 0; nu < 4; ++nu) acc += pab[nu] * w[pack4(u, nu)];
                const double f = hv[u] - scale * acc;
                F[r0 + (size_t)u * nao + ob] = f;
                ksum_add(ev, 2.0 * pab[u] * (hv[u] + f));
            }
        }
    }
    // rsum[ci][k]: lanes l with l % 4 == k % 4 hold k's terms in arow[k / 4]
    __syncthreads();
    double* flat = &stage[0][0];
#pragma unroll
    for (int t = 0; t < 3; ++t) flat[wv * 192 + lane * 3 + t] = arow[t];
    __syncthreads();
    if (threadIdx.x < 40) {
        const int kq = threadIdx.x % 10, g = threadIdx.x / 10;
        double v = 0.0;
        for (int p = 0; p < 16; ++p) v += flat[g * 192 + (p * 4 + kq % 4) * 3 + kq / 4];
        v += __shfl(v, threadIdx.x + 10, 64);
        v += __shfl(v, threadIdx.x + 20, 64);
        if (threadIdx.x < 10) a.rsum[(size_t)ci * 10 + kq] = v;
    }
    if (epart) energy_part(ev, epart, slot0 + blockIdx.x);
}

// Row of hydrogen cm (workgroup cm): its heavy partners later in AO order, lane per (partner,
// orbital), the 1 x 4 blocks; then its hydrogen partners later in the list.
__global__ __launch_bounds__(256) void k_fock_rows_h(SysArgs a, const double* Ps, double scale, double* F,
                                                     double* epart, int64_t slot0) {
    const int nh = a.md.n_heavy, nH = a.md.n_h, nao = a.md.nao;
    const int64_t n_hh = (int64_t)nh * (nh - 1) / 2, n_hH = (int64_t)nh * nH;
    const double* wHh = a.wpool + n_hh * 100;
    const double* wHH = wHh + n_hH * 10;
    const int cm = blockIdx.x, ob = a.aoff[a.alist[nh + cm]];
    const size_t row = (size_t)ob * nao;
    const int mu = threadIdx.x & 3;
    KSum ev;
    for (int i = a.later[nh + cm] + (threadIdx.x >> 2); i < nh; i += 64) {
        const int oa = a.aoff[a.alist[i]];
        const double* w = wHh + ((int64_t)i * nH + cm) * 10;
        double acc = 0.0;
#pragma unroll
        for (int nu = 0; nu < 4; ++nu) acc += Ps[row + oa + nu] * w[pack4(mu, nu)];
        const double h = a.H[row + oa + mu], f = h - scale * acc;
        F[row + oa + mu] = f;
        ksum_add(ev, 2.0 * Ps[row + oa + mu] * (h + f));
    }
    const double* wm = wHH + tri_index(cm, cm + 1, nH) - (cm + 1);  // entry of partner m > cm: wm[m]
    for (int m = cm + 1 + threadIdx.x; m < nH; m += 256) {
        const size_t ix = row + a.aoff[a.alist[nh + m]];
        const double p = Ps[ix], h = a.H[ix], f = h - scale * p * wm[m];
        F[ix] = f;
        ksum_add(ev, 2.0 * p * (h + f));
    }
    if (epart) energy_part(ev, epart, slot0 + blockIdx.x);
}

// F[i][j] = F[j][i] for i > j, in 32 x 32 tiles through LDS (grid: the tiles on and below the
// diagonal, tile (bi, bj) with bi >= bj; 32 x 8 threads, four rows each).
__global__ __launch_bounds__(256) void k_sym_lower(int n, double* F) {
    __shared__ double tile[32][33];
    const int nt = (n + 31) / 32;
    // linear block index -> (bi, bj), bj <= bi
    const int64_t lb = blockIdx.x;
    int bi = (int)((sqrt(8.0 * (double)lb + 1.0) - 1.0) * 0.5);
    while ((int64_t)bi * (bi + 1) / 2 > lb) --bi;
    while ((int64_t)(bi + 1) * (bi + 2) / 2 <= lb) ++bi;
    const int bj = (int)(lb - (int64_t)bi * (bi + 1) / 2);
    if (bi >= nt) return;
    const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;
    // read the upper tile (rows of block bj, columns of block bi)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int row = bj * 32 + ty + 8 * r, col = bi * 32 + tx;
        if (row < n && col < n && row < col) tile[ty + 8 * r][tx] = F[(size_t)row * n + col];
    }
    __syncthreads();
    // write its transpose (rows of block bi, columns of block bj) below the diagonal
#pragma unroll
    for (int r = 0; r < 4; ++r) {
        const int row = bi * 32 + ty + 8 * r, col = bj * 32 + tx;
        if (row < n && col < n && row > col) F[(size_t)row * n + col] = tile[tx][ty + 8 * r];
    }
}

// Deterministic two-stage sums: fixed-size grid of block partials, then one block in order.
constexpr int kRedBlocks = 1024;

// Block partials of sum a * (b + c) (b, c optional), as (sum, error) pairs.
__global__ __launch_bounds__(256) void k_dot3_partial(int64_t n, const double* a, const double* b, const double* c,
                                                      double* partial) {
    KSum v;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)kRedBlocks * 256)
        ksum_add(v, a[i] * (b ? (c ? b[i] + c[i] : b[i]) : 1.0));
    const KSum r = ksum_block(v);
    if (threadIdx.x == 0) {
        partial[2 * blockIdx.x] = r.s;
        partial[2 * blockIdx.x + 1] = r.c;
    }
}

__global__ __launch_bounds__(256) void k_final_sum(int nparts, const double* partial, double* out) {
    KSum v;
    for (int i = threadIdx.x; i < nparts; i += 256) {
        KSum p;
        p.s = partial[2 * i];
        p.c = partial[2 * i + 1];
        v = ksum_join(v, p);
    }
    const KSum r = ksum_block(v);
    if (threadIdx.x == 0) out[0] = r.s + r.c;
}

__global__ __launch_bounds__(256) void k_sad(SysArgs a, double scale, double* P) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= a.md.nat) return;
    const DevElem& e = a.elems[a.md.method * kMaxZ + a.z[x]];
    const int nao = a.md.nao, o = a.aoff[x];
    const double nsv = fmin(e.core, 2.0), npv = fmax(e.core - nsv, 0.0);
    P[(size_t)o * nao + o] = nsv * scale;
    if (e.norb == 4)
        for (int k = 1; k < 4; ++k) P[(size_t)(o + k) * nao + o + k] = npv / 3.0 * scale;
}

// ---------------------------------------------------------------- divide and conquer stages
// One workgroup per subsystem: gather its Fock block, diagonalize it (warm-started from the
// previous iteration's subsystem orbitals), sort, and project every level on the core AOs
// (divide_conquer.rs:588-618).
__global__ __launch_bounds__(kWg) void k_dc_eigen(DcArgs d, const double* F, int nao) {
    __shared__ WgScratch s;
    const int sub = blockIdx.x;
    const int n = d.sub_n[sub];
    const int t = threadIdx.x;
    const int* ao = d.ao_list + d.sub_ao[sub];
    double* Fs = d.fbuf + d.sub_mat[sub];
    double* Ct = d.cbuf + d.sub_mat[sub];
    // F holds its upper triangle (and the diagonal atom blocks): the D&C Fock build skips the
    // copy to the lower triangle. The subsystem's AO list ascends, so i <= j reads the upper
    // triangle (row ao[i], consecutive threads along it) and fills both mirror entries.
    for (int idx = t; idx < n * n; idx += kWg) {
        const int i = idx / n, j = idx - i * n;
        if (j < i) continue;
        const double v = F[(size_t)ao[i] * nao + ao[j]];
        Fs[(size_t)i * n + j] = v;
        Fs[(size_t)j * n + i] = v;
    }
    __syncthreads();
    bool partial = false;
    const int sw = mol_eigen<double>(n, Fs, Ct, d.t0 + d.sub_mat[sub], d.t1 + d.sub_mat[sub], lds_dyn, d.lds_n, d.tri_max,
                                     d.warm != 0, d.eps + d.sub_ao[sub], s, d.tol_abs, 0.0, n, partial);
    const unsigned char* core = d.ao_core + d.sub_ao[sub];
    for (int i = t / 16; i < n; i += kWg / 16) {
        double acc = 0.0;
        for (int p = t % 16; p < n; p += 16) {
            const double c = Ct[(size_t)i * n + p];
            if (core[p]) acc += c * c;
        }
        for (int off = 8; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 16);
        if (t % 16 == 0) d.proj[d.sub_ao[sub] + i] = acc;
    }
    if (t == 0) d.status[sub] = sw;
}

__device__ __forceinline__ double fermi_dirac(double x) {  // fermi.rs:110-119
    if (x > 40.0) return 0.0;
    if (x < -40.0) return 1.0;
    return 1.0 / (1.0 + exp(x));
}

// Common chemical potential (fermi.rs:208-269): the reference bisects the electron count
// N(mu) = sum_i cap proj_i f((e_i - mu) / kT) on [min - 50 kT - 1, max + 50 kT + 1], at most 200
// steps, stopping at |N - N_e| < 1e-13. Here each pass evaluates the 31 midpoints of the next
// five bisection steps (a complete binary tree of intervals) over all levels at once, and one
// thread then replays those five steps exactly as the reference takes them; the replay also
// stops when the interval no longer shrinks, which is where 200 reference steps end up.
// State in red[0..4): lo, hi, done, mu, steps; block partials from red[8], midpoint-major.
constexpr int kFermiDepth = 5;
constexpr int kFermiK = (1 << kFermiDepth) - 1;
constexpr int kFermiBlocks = 256;
constexpr int kFermiPasses = 40;  // 200 steps
constexpr int kFermiFirst = 12;   // passes launched before the host looks at the done flag

int64_t fermi_scratch_doubles() { return 8 + (int64_t)kFermiBlocks * kFermiK; }

__device__ __forceinline__ void fermi_tree(double lo, double hi, double (&mid)[kFermiK]) {
    double l[kFermiK], h[kFermiK];
    l[0] = lo;
    h[0] = hi;
#pragma unroll
    for (int q = 0; q < kFermiK; ++q) {
        mid[q] = 0.5 * (l[q] + h[q]);
        if (2 * q + 2 < kFermiK) {
            l[2 * q + 1] = l[q];
            h[2 * q + 1] = mid[q];
            l[2 * q + 2] = mid[q];
            h[2 * q + 2] = h[q];
        }
    }
}

// Block partials of the lowest and highest level (red[8 + 2 b], red[9 + 2 b]).
__global__ __launch_bounds__(256) void k_fermi_minmax(DcArgs d, int64_t L) {
    __shared__ double sh[2][4];
    double lo = INFINITY, hi = -INFINITY;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < L; i += (int64_t)gridDim.x * 256) {
        lo = fmin(lo, d.eps[i]);
        hi = fmax(hi, d.eps[i]);
    }
    for (int off = 32; off > 0; off >>= 1) {
        lo = fmin(lo, __shfl_xor(lo, off, 64));
        hi = fmax(hi, __shfl_xor(hi, off, 64));
    }
    if ((threadIdx.x & 63) == 0) {
        sh[0][threadIdx.x >> 6] = lo;
        sh[1][threadIdx.x >> 6] = hi;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        d.red[8 + 2 * blockIdx.x] = fmin(fmin(sh[0][0], sh[0][1]), fmin(sh[0][2], sh[0][3]));
        d.red[9 + 2 * blockIdx.x] = fmax(fmax(sh[1][0], sh[1][1]), fmax(sh[1][2], sh[1][3]));
    }
}

__global__ __launch_bounds__(kWg) void k_fermi_bounds(DcArgs d, double kt) {
    __shared__ WgScratch s;
    double lo = INFINITY, hi = -INFINITY;
    for (int b = threadIdx.x; b < kFermiBlocks; b += kWg) {
        lo = fmin(lo, d.red[8 + 2 * b]);
        hi = fmax(hi, d.red[9 + 2 * b]);
    }
    lo = -wg_max(-lo, s) - 50.0 * kt - 1.0;
    hi = wg_max(hi, s) + 50.0 * kt + 1.0;
    if (threadIdx.x == 0) {
        d.red[0] = lo;
        d.red[1] = hi;
        d.red[2] = 0.0;
        d.red[3] = 0.5 * (lo + hi);
        d.red[4] = 0.0;
    }
}

__global__ __launch_bounds__(256) void k_fermi_count(DcArgs d, int64_t L, double kt) {
    if (d.red[2] != 0.0) return;
    double mu[kFermiK], c[kFermiK];
    const double lo = d.red[0], hi = d.red[1];
    fermi_tree(lo, hi, mu);
#pragma unroll
    for (int q = 0; q < kFermiK; ++q) c[q] = 0.0;
    double below = 0.0;  // weight of the levels below every midpoint by more than 40 kT
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < L; i += (int64_t)gridDim.x * 256) {
        const double w = d.cap * d.proj[i], e = d.eps[i];
        // every midpoint lies in [lo, hi] and rounded subtraction and division are monotone, so
        // beyond these bounds fermi_dirac is exactly 0 (above) or 1 (below) for all of them
        if ((e - hi) / kt > 40.0) continue;
        if ((e - lo) / kt < -40.0) {
            below += w;
            continue;
        }
#pragma unroll
        for (int q = 0; q < kFermiK; ++q) c[q] += w * fermi_dirac((e - mu[q]) / kt);
    }
#pragma unroll
    for (int q = 0; q < kFermiK; ++q) c[q] += below;
    __shared__ double sh[4][kFermiK];
#pragma unroll
    for (int q = 0; q < kFermiK; ++q) {
        double x = c[q];
        for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
        if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6][q] = x;
    }
    __syncthreads();
    double* part = d.red + 8;
    if (threadIdx.x < kFermiK)
        part[threadIdx.x * kFermiBlocks + blockIdx.x] =
            (sh[0][threadIdx.x] + sh[1][threadIdx.x]) + (sh[2][threadIdx.x] + sh[3][threadIdx.x]);
}

constexpr int kPickGroups = 32;  // block partials per midpoint summed by one thread: kFermiBlocks / 32
static_assert(kFermiBlocks % kPickGroups == 0 && kPickGroups * kFermiK <= 1024, "pick layout");
__global__ __launch_bounds__(1024) void k_fermi_pick(DcArgs d, double n_elec) {
    if (d.red[2] != 0.0) return;
    __shared__ double tot[kFermiK];
    __shared__ double sh[kPickGroups][kFermiK];
    const double* part = d.red + 8;
    // thread (q, g) sums the g-th slice of midpoint q's block partials; then thread q the slices
    if (threadIdx.x < kPickGroups * kFermiK) {
        const int q = threadIdx.x % kFermiK, g = threadIdx.x / kFermiK;
        const double* p = part + (size_t)q * kFermiBlocks + g * (kFermiBlocks / kPickGroups);
        double v = 0.0;
#pragma unroll
        for (int b = 0; b < kFermiBlocks / kPickGroups; ++b) v += p[b];
        sh[g][q] = v;
    }
    __syncthreads();
    if (threadIdx.x < kFermiK) {
        const int q = threadIdx.x;
        double v = 0.0;
        for (int g = 0; g < kPickGroups; ++g) v += sh[g][q];
        tot[q] = v;
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    double mid[kFermiK];
    double lo = d.red[0], hi = d.red[1];
    fermi_tree(lo, hi, mid);
    int steps = (int)d.red[4];
    int node = 0;
    for (int depth = 0; depth < kFermiDepth; ++depth) {
        const double m = mid[node];
        ++steps;
        d.red[3] = m;
        if (fabs(tot[node] - n_elec) < 1.0e-13 || steps >= 200 || m <= lo || m >= hi) {
            d.red[2] = 1.0;
            break;
        }
        if (tot[node] < n_elec) {
            lo = m;
            node = 2 * node + 2;
        } else {
            hi = m;
            node = 2 * node + 1;
        }
    }
    d.red[0] = lo;
    d.red[1] = hi;
    d.red[4] = (double)steps;
}

// Fractions at the chemical potential, block partials of the entropy and of the count.
__global__ __launch_bounds__(256) void k_fermi_final(DcArgs d, int64_t L, double kt) {
    const double mu = d.red[3];
    double sk = 0.0, cnt = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < L; i += (int64_t)gridDim.x * 256) {
        const double f = fermi_dirac((d.eps[i] - mu) / kt);
        const double w = d.cap * d.proj[i];
        d.frac[i] = f;
        cnt += w * f;
        if (f > 1.0e-14 && f < 1.0 - 1.0e-14) sk -= w * (f * log(f) + (1.0 - f) * log(1.0 - f));
    }
    __shared__ double sh[2][4];
    for (int off = 32; off > 0; off >>= 1) {
        sk += __shfl_xor(sk, off, 64);
        cnt += __shfl_xor(cnt, off, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        sh[0][threadIdx.x >> 6] = sk;
        sh[1][threadIdx.x >> 6] = cnt;
    }
    __syncthreads();
    double* part = d.red + 8;
    if (threadIdx.x < 2)
        part[blockIdx.x * 2 + threadIdx.x] =
            (sh[threadIdx.x][0] + sh[threadIdx.x][1]) + (sh[threadIdx.x][2] + sh[threadIdx.x][3]);
}

__global__ __launch_bounds__(256) void k_fermi_out(DcArgs d, double kt, double* out) {
    __shared__ double sh[2][4];
    const double* part = d.red + 8;
    double sk = 0.0, cnt = 0.0;
    for (int b = threadIdx.x; b < kFermiBlocks; b += 256) {
        sk += part[2 * b];
        cnt += part[2 * b + 1];
    }
    for (int off = 32; off > 0; off >>= 1) {
        sk += __shfl_xor(sk, off, 64);
        cnt += __shfl_xor(cnt, off, 64);
    }
    if ((threadIdx.x & 63) == 0) {
        sh[0][threadIdx.x >> 6] = sk;
        sh[1][threadIdx.x >> 6] = cnt;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        out[0] = d.red[3];
        out[1] = kt * ((sh[0][0] + sh[0][1]) + (sh[0][2] + sh[0][3]));
        out[2] = (sh[1][0] + sh[1][1]) + (sh[1][2] + sh[1][3]);
    }
}

// Subsystem density blocks D_s = sum_i cap f_i c_i c_i^T over levels with f_i > 1e-14
// (divide_conquer.rs:686-713), written over the gathered Fock block.
__global__ __launch_bounds__(kWg) void k_dc_density(DcArgs d) {
    const int sub = blockIdx.x;
    const int n = d.sub_n[sub];
    const int t = threadIdx.x;
    const double* Ct = d.cbuf + d.sub_mat[sub];
    const double* f = d.frac + d.sub_ao[sub];
    double* T0 = d.t0 + d.sub_mat[sub];
    // levels come in ascending order and f falls with the level, so the rows that contribute are
    // a prefix: the product runs over the rows up to the last one with f > 1e-14 (the rest are
    // exact zeros)
    __shared__ int k_used;
    if (t == 0) k_used = 1;
    __syncthreads();
    int last = 0;
    for (int i = t; i < n; i += kWg)
        if (f[i] > 1.0e-14) last = i + 1;
    if (last > 0) atomicMax(&k_used, last);
    for (int idx = t; idx < n * n; idx += kWg) {
        const int i = idx / n;
        T0[idx] = f[i] > 1.0e-14 ? Ct[idx] * sqrt(d.cap * f[i]) : 0.0;
    }
    __syncthreads();
    wg_gemm_any<double, true, false>(n, n, k_used, 1.0, T0, n, T0, n, d.fbuf + d.sub_mat[sub], n, lds_dyn,
                                     kLdsDoubles);
}

__device__ __forceinline__ int local_ao(const DcArgs& d, int sub, int g) {
    const int* ao = d.ao_list + d.sub_ao[sub];
    int lo = 0, hi = d.sub_n[sub] - 1;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (ao[mid] < g) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// Yang assembly as a gather over the retained pattern: an AO pair (r, c) on atoms (a, b) within
// the buffer radius gets weight 1/2 from the subsystem whose core holds a and 1/2 from the one
// whose core holds b (1 when that is one subsystem); every other subsystem weighs it 0.
__global__ __launch_bounds__(256) void k_dc_gather(DcArgs d, int64_t npat, const int64_t* pat, const int* ao_atom,
                                                   const int* core_sub, int nao, double* pout) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= npat) return;
    const int r = (int)(pat[k] / nao), c = (int)(pat[k] % nao);
    const int s1 = core_sub[ao_atom[r]], s2 = core_sub[ao_atom[c]];
    const int n1 = d.sub_n[s1];
    const double v1 = d.fbuf[d.sub_mat[s1] + (int64_t)local_ao(d, s1, r) * n1 + local_ao(d, s1, c)];
    if (s1 == s2) {
        pout[k] = v1;
        return;
    }
    const int n2 = d.sub_n[s2];
    const double v2 = d.fbuf[d.sub_mat[s2] + (int64_t)local_ao(d, s2, r) * n2 + local_ao(d, s2, c)];
    pout[k] = 0.5 * v1 + 0.5 * v2;
}

// Push one density-DIIS entry (input and residual on the pattern) and form, in the same pass,
// the multiplicity-weighted sum of squared residuals and the new residual's dots with every
// history entry (divide_conquer.rs:800-1019). Block partials: [grid][1 + m].
__global__ __launch_bounds__(256) void k_dc_hist(int64_t npat, const int64_t* pat, const double* mult, const double* P,
                                                 const double* pout, double* hin, double* hres, const double* hres_all,
                                                 const int* ring, int m, int64_t stride, double* partial) {
    __shared__ double sh[4][kDiisDepth + 1];
    double v[kDiisDepth + 1];
#pragma unroll
    for (int q = 0; q <= kDiisDepth; ++q) v[q] = 0.0;
    for (int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x; k < npat; k += (int64_t)kRedBlocks * 256) {
        const double pin = P[pat[k]];
        const double r = pout[k] - pin;
        if (hin) {
            hin[k] = pin;
            hres[k] = r;
        }
        const double w = mult[k];
        v[0] += w * r * r;
#pragma unroll
        for (int j = 0; j < kDiisDepth; ++j)
            if (j < m) v[1 + j] += w * r * (ring[j] < 0 ? r : hres_all[(int64_t)ring[j] * stride + k]);
    }
#pragma unroll
    for (int q = 0; q <= kDiisDepth; ++q) {
        double x = v[q];
        for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
        if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6][q] = x;
    }
    __syncthreads();
    if (threadIdx.x <= kDiisDepth)
        partial[(int64_t)blockIdx.x * (kDiisDepth + 1) + threadIdx.x] =
            (sh[0][threadIdx.x] + sh[1][threadIdx.x]) + (sh[2][threadIdx.x] + sh[3][threadIdx.x]);
}

__global__ __launch_bounds__(256) void k_final_sum_many(int nparts, int width, const double* partial, double* out) {
    __shared__ double sh[4];
    for (int q = 0; q < width; ++q) {
        double v = 0.0;
        for (int i = threadIdx.x; i < nparts; i += 256) v += partial[(int64_t)i * width + q];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
        __syncthreads();
        if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
        __syncthreads();
        if (threadIdx.x == 0) out[q] = (sh[0] + sh[1]) + (sh[2] + sh[3]);
    }
}

// Next input density on the pattern (both triangles): sum_i c_i (in_i + res_i) when coef is given,
// linear mixing in + mix (out - in) when mix >= 0, or the output density itself when mix < 0.
__global__ __launch_bounds__(256) void k_dc_update(int64_t npat, const int64_t* pat, int nao, const double* coef,
                                                   const double* hin_all, const double* hres_all, const int* ring,
                                                   int m, int64_t stride, double mix, const double* pout, double* P) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= npat) return;
    double v;
    if (coef) {
        v = 0.0;
        for (int i = 0; i < m; ++i) {
            const int64_t o = (int64_t)ring[i] * stride + k;
            v += coef[i] * (hin_all[o] + hres_all[o]);
        }
    } else if (mix >= 0.0) {
        const double pin = P[pat[k]];
        v = pin + mix * (pout[k] - pin);
    } else {
        v = pout[k];
    }
    const int r = (int)(pat[k] / nao), c = (int)(pat[k] % nao);
    P[(size_t)r * nao + c] = v;
    P[(size_t)c * nao + r] = v;
}

// ---------------------------------------------------------------- launchers
static inline unsigned blocks_for(int64_t n, int per = 256) { return (unsigned)((n + per - 1) / per); }

// Uniform external field on H_core (dipole.rs:126-144): F . R_a on each AO's diagonal entry and
// F_alpha dd on the s-p_alpha entries of each heavy atom. One thread per atom.
__global__ void k_field_hcore(SysArgs a) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.md.nat) return;
    const int nao = a.md.nao, o = a.aoff[i];
    const DevElem& e = a.elems[a.md.method * kMaxZ + a.z[i]];
    const double* r = a.xyz + 3 * (size_t)i;
    const double fr = a.md.field[0] * r[0] + a.md.field[1] * r[1] + a.md.field[2] * r[2];
    for (int k = 0; k < e.norb; ++k) a.H[(size_t)(o + k) * nao + o + k] += fr;
    if (e.norb == 4)
        for (int k = 0; k < 3; ++k) {
            a.H[(size_t)o * nao + o + 1 + k] += a.md.field[k] * e.dd;
            a.H[(size_t)(o + 1 + k) * nao + o] += a.md.field[k] * e.dd;
        }
}

hipError_t launch_hcore_diag(const SysArgs& a, hipStream_t st) {
    if (a.md.nat > 0) {
        hipLaunchKernelGGL(k_pack_core, dim3(blocks_for(a.md.nat)), dim3(256), 0, st, a);
        hipLaunchKernelGGL(k_diag_blocks<false>, dim3(a.md.n_heavy + (a.md.n_h + 7) / 8), dim3(256), 0, st, a,
                           nullptr, nullptr, 0.0, nullptr, nullptr);
        if (a.md.field[0] != 0.0 || a.md.field[1] != 0.0 || a.md.field[2] != 0.0)
            hipLaunchKernelGGL(k_field_hcore, dim3(blocks_for(a.md.nat)), dim3(256), 0, st, a);
    }
    return hipGetLastError();
}

// Far-field shift of one spin channel's Fock matrix (fock.rs:278-318): F_mumu -= V_a on every AO
// of atom a, with V_a = sum over atoms b farther than the cutoff of 27.21 / sqrt(r^2 + (rho0_a +
// rho0_b)^2) * Q_b (FarField::potential, farfield.rs:150-176) and the net charges Q_b from the
// packed diagonal blocks of Pt (k_pack_diag). One thread per atom; also the atom's energy terms
// -ps_a V_a (ps_a: its population in Ps, the density this channel's energy sum contracts with F)
// and Z_a V_a.
__global__ __launch_bounds__(256) void k_far_shift(SysArgs a, const double* Ps, double* F) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int nat = a.md.nat, nao = a.md.nao;
    if (i >= nat) return;
    const DevElem* el = a.elems + a.md.method * kMaxZ;
    const double cut2 = a.md.far_cut * a.md.far_cut;
    const double x0 = a.xyz[3 * (size_t)i], x1 = a.xyz[3 * (size_t)i + 1], x2 = a.xyz[3 * (size_t)i + 2];
    const DevElem& ei = el[a.z[i]];
    double v = 0.0;
    for (int b = 0; b < nat; ++b) {
        const double d0 = a.xyz[3 * (size_t)b] - x0, d1 = a.xyz[3 * (size_t)b + 1] - x1,
                     d2 = a.xyz[3 * (size_t)b + 2] - x2;
        const double r2 = d0 * d0 + d1 * d1 + d2 * d2;
        if (b == i || r2 <= cut2) continue;
        const DevElem& eb = el[a.z[b]];
        const double* pd = pdiag_row(a, b);
        const double pop = a.z[b] == 1 ? pd[0] : pd[0] + pd[2] + pd[5] + pd[9];
        const double eta = ei.rho0 + eb.rho0;
        v += kEv / sqrt(r2 + eta * eta) * (eb.core - pop);
    }
    const int o = a.aoff[i];
    double ps = 0.0;
    for (int k = 0; k < ei.norb; ++k) {
        F[(size_t)(o + k) * nao + o + k] -= v;
        ps += Ps[(size_t)(o + k) * nao + o + k];
    }
    a.far[i] = -ps * v;
    a.far[nat + i] = ei.core * v;
}

// Sums of k_far_shift's per-atom terms in a fixed order: the channel's energy sum gains
// sum -ps_a V_a (its Fock sum was taken before the shift) and far[2 nat] = sum Z_a V_a, for the
// final energy's +1/2 sum Z_a V_a (long_range_energy_term, fock.rs:200-232).
__global__ __launch_bounds__(256) void k_far_energy(int nat, double* far, double* e_out) {
    KSum pv, zv;
    for (int i = threadIdx.x; i < nat; i += 256) {
        ksum_add(pv, far[i]);
        ksum_add(zv, far[nat + i]);
    }
    const KSum rp = ksum_block(pv);
    const KSum rz = ksum_block(zv);
    if (threadIdx.x == 0) {
        if (e_out) e_out[0] += rp.s + rp.c;
        far[2 * nat] = rz.s + rz.c;
    }
}

hipError_t launch_fock(const SysArgs& a, const double* Pt, const double* Ps, double scale, double* F, hipStream_t st,
                       double* epart, double* e_out, bool full) {
    const int64_t nh = a.md.n_heavy, nH = a.md.n_h, ndiag = nh + (nH + 7) / 8;
    if (!e_out) epart = nullptr;
    hipLaunchKernelGGL(k_pack_diag, dim3(blocks_for(a.md.nat)), dim3(256), 0, st, a, Pt);
    if (nh > 0)
        hipLaunchKernelGGL(k_fock_rows_heavy, dim3((unsigned)nh), dim3(256), 0, st, a, Ps, scale, F, epart, ndiag);
    if (nH > 0)
        hipLaunchKernelGGL(k_fock_rows_h, dim3((unsigned)nH), dim3(256), 0, st, a, Ps, scale, F, epart, ndiag + nh);
    // after k_fock_rows_heavy, which leaves the heavy partners' terms in rsum and tcol
    hipLaunchKernelGGL(k_diag_blocks<true>, dim3((unsigned)ndiag), dim3(256), 0, st, a, Pt, Ps, scale, F, epart);
    if (full && a.md.nat > 1) {
        const int64_t nt = ((int64_t)a.md.nao + 31) / 32;
        hipLaunchKernelGGL(k_sym_lower, dim3((unsigned)(nt * (nt + 1) / 2)), dim3(256), 0, st, a.md.nao, F);
    }
    if (e_out) hipLaunchKernelGGL(k_final_sum, dim3(1), dim3(256), 0, st, (int)fock_energy_parts(a.md), epart, e_out);
    if (a.md.far_cut > 0.0) {
        if (!a.far) return hipErrorInvalidValue;
        hipLaunchKernelGGL(k_far_shift, dim3(blocks_for(a.md.nat)), dim3(256), 0, st, a, Ps, F);
        hipLaunchKernelGGL(k_far_energy, dim3(1), dim3(256), 0, st, a.md.nat, a.far, e_out);
    }
    return hipGetLastError();
}

hipError_t launch_dot3(int64_t n, const double* a, const double* b, const double* c, double* partial, double* out,
                       hipStream_t st) {
    hipLaunchKernelGGL(k_dot3_partial, dim3(kRedBlocks), dim3(256), 0, st, n, a, b, c, partial);
    hipLaunchKernelGGL(k_final_sum, dim3(1), dim3(256), 0, st, kRedBlocks, partial, out);
    return hipGetLastError();
}

hipError_t launch_sum(int64_t n, const double* x, double* partial, double* out, hipStream_t st) {
    return launch_dot3(n, x, nullptr, nullptr, partial, out, st);
}

hipError_t launch_sad(const SysArgs& a, double scale, double* P, hipStream_t st) {
    hipLaunchKernelGGL(k_sad, dim3(blocks_for(a.md.nat)), dim3(256), 0, st, a, scale, P);
    return hipGetLastError();
}

hipError_t launch_dc_eigen(const DcArgs& d, const double* F, int nao, hipStream_t st) {
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_dc_eigen),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, kLdsBytes);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_dc_eigen, dim3(d.n_sub), dim3(kWg), kLdsBytes, st, d, F, nao);
    return hipGetLastError();
}

hipError_t launch_fermi(const DcArgs& d, int64_t levels, double n_elec, double kt, double* out, hipStream_t st) {
    hipLaunchKernelGGL(k_fermi_minmax, dim3(kFermiBlocks), dim3(256), 0, st, d, levels);
    hipLaunchKernelGGL(k_fermi_bounds, dim3(1), dim3(kWg), 0, st, d, kt);
    for (int pass = 0; pass < kFermiPasses; ++pass) {
        if (pass == kFermiFirst) {
            // 60 steps end nearly every search (the interval reaches the spacing of doubles
            // near mu); look at the done flag before launching the rest
            double done = 0.0;
            hipError_t e = hipMemcpyAsync(&done, d.red + 2, sizeof(double), hipMemcpyDeviceToHost, st);
            if (e == hipSuccess) e = hipStreamSynchronize(st);
            if (e != hipSuccess) return e;
            if (done != 0.0) break;
        }
        hipLaunchKernelGGL(k_fermi_count, dim3(kFermiBlocks), dim3(256), 0, st, d, levels, kt);
        hipLaunchKernelGGL(k_fermi_pick, dim3(1), dim3(1024), 0, st, d, n_elec);
    }
    hipLaunchKernelGGL(k_fermi_final, dim3(kFermiBlocks), dim3(256), 0, st, d, levels, kt);
    hipLaunchKernelGGL(k_fermi_out, dim3(1), dim3(256), 0, st, d, kt, out);
    return hipGetLastError();
}

hipError_t launch_dc_density(const DcArgs& d, hipStream_t st) {
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&k_dc_density),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, kLdsBytes);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(k_dc_density, dim3(d.n_sub), dim3(kWg), kLdsBytes, st, d);
    return hipGetLastError();
}

hipError_t launch_dc_gather(const DcArgs& d, int64_t npat, const int64_t* pat, const int* ao_atom, const int* core_sub,
                            int nao, double* pout, hipStream_t st) {
    if (npat > 0) hipLaunchKernelGGL(k_dc_gather, dim3(blocks_for(npat)), dim3(256), 0, st, d, npat, pat, ao_atom, core_sub, nao, pout);
    return hipGetLastError();
}

hipError_t launch_dc_hist(int64_t npat, const int64_t* pat, const double* mult, const double* P, const double* pout,
                          double* hin, double* hres, const double* hres_all, const int* ring, int m, int64_t stride,
                          double* partial, double* out, hipStream_t st) {
    hipLaunchKernelGGL(k_dc_hist, dim3(kRedBlocks), dim3(256), 0, st, npat, pat, mult, P, pout, hin, hres, hres_all,
                       ring, m, stride, partial);
    hipLaunchKernelGGL(k_final_sum_many, dim3(1), dim3(256), 0, st, kRedBlocks, kDiisDepth + 1, partial, out);
    return hipGetLastError();
}

hipError_t launch_dc_update(int64_t npat, const int64_t* pat, int nao, const double* coef, const double* hin_all,
                            const double* hres_all, const int* ring, int m, int64_t stride, double mix,
                            const double* P_in, const double* pout, double* P, hipStream_t st) {
    (void)P_in;
    if (npat > 0)
        hipLaunchKernelGGL(k_dc_update, dim3(blocks_for(npat)), dim3(256), 0, st, npat, pat, nao, coef, hin_all,
                           hres_all, ring, m, stride, mix, pout, P);
    return hipGetLastError();
}

}  // namespace am1g
#endif  // AM1G_WG_VARIANT
