// Divide-and-conquer SCF driver (host side of am1g_dc).
//
// Follows am1-rs run_divide_conquer with dc_loop_restricted / dc_loop_unrestricted
// (src/divide_conquer.rs:406-572, 1033-1254): the full Fock build over every atom pair, one
// batched eigensolve of all subsystem blocks per spin channel, a common Fermi level per channel
// by bisection, Yang-weight assembly, and Pulay DIIS on each channel's density residual over the
// retained pattern (linear mixing when DIIS declines). The heavy steps are device kernels; the
// host keeps the partition, the pattern and the tiny DIIS solves.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

#include "am1g_ctx.hpp"

namespace {

// Recursive proportional bisection along the widest axis (divide_conquer.rs:249-297).
void bisect_cores(const double* xyz, std::vector<int> idx, int core_size, std::vector<std::vector<int>>& out) {
    if ((int)idx.size() <= core_size) {
        out.push_back(idx);
        return;
    }
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int i : idx)
        for (int k = 0; k < 3; ++k) {
            lo[k] = std::min(lo[k], xyz[3 * i + k]);
            hi[k] = std::max(hi[k], xyz[3 * i + k]);
        }
    int axis = 0;
    for (int a = 1; a < 3; ++a)
        if (hi[a] - lo[a] > hi[axis] - lo[axis]) axis = a;
    std::sort(idx.begin(), idx.end(), [&](int a, int b) {
        const double ka = xyz[3 * a + axis], kb = xyz[3 * b + axis];
        return ka < kb || (ka == kb && a < b);
    });
    const int n = (int)idx.size();
    const int parts = (n + core_size - 1) / core_size;
    const int split = std::min(std::max(n * (parts / 2) / parts, 1), n - 1);
    std::vector<int> right(idx.begin() + split, idx.end());
    idx.resize(split);
    bisect_cores(xyz, idx, core_size, out);
    bisect_cores(xyz, right, core_size, out);
}

// Partial-pivot LU solve that refuses non-finite or huge solutions (linalg.rs:334-361).
bool lu_solve(std::vector<double> a, std::vector<double> b, int n, std::vector<double>& x) {
    for (int col = 0; col < n; ++col) {
        int piv = col;
        for (int r = col + 1; r < n; ++r)
            if (std::fabs(a[r * n + col]) > std::fabs(a[piv * n + col])) piv = r;
        if (piv != col) {
            for (int j = 0; j < n; ++j) std::swap(a[col * n + j], a[piv * n + j]);
            std::swap(b[col], b[piv]);
        }
        for (int r = col + 1; r < n; ++r) {
            const double f = a[r * n + col] / a[col * n + col];
            for (int j = col; j < n; ++j) a[r * n + j] -= f * a[col * n + j];
            b[r] -= f * b[col];
        }
    }
    x.assign(n, 0.0);
    for (int col = n - 1; col >= 0; --col) {
        double s = b[col];
        for (int j = col + 1; j < n; ++j) s -= a[col * n + j] * x[j];
        x[col] = s / a[col * n + col];
    }
    double bmax = 1.0;
    for (double v : b) bmax = std::max(bmax, std::fabs(v));
    for (double v : x)
        if (!std::isfinite(v) || std::fabs(v) > 1e8 * bmax) return false;
    return true;
}

// Uniform grid of cubic cells of side at least `cell` (the buffer radius) over the atoms, so the
// atoms within that distance of an atom are found among the 27 cells around it. The side grows
// when the grid would have more than 8 cells per atom (a very large radius ends in one cell).
struct CellGrid {
    double lo[3], side;
    int dim[3];
    std::vector<int> start, atoms;  // atoms of cell k: atoms[start[k] .. start[k + 1])
    int cell_of(const double* p) const {
        int c[3];
        for (int k = 0; k < 3; ++k) c[k] = std::min(dim[k] - 1, std::max(0, (int)((p[k] - lo[k]) / side)));
        return (c[2] * dim[1] + c[1]) * dim[0] + c[0];
    }
    CellGrid(const double* xyz, int nat, double cell) {
        double hi[3];
        for (int k = 0; k < 3; ++k) {
            lo[k] = INFINITY;
            hi[k] = -INFINITY;
        }
        for (int a = 0; a < nat; ++a)
            for (int k = 0; k < 3; ++k) {
                lo[k] = std::min(lo[k], xyz[3 * a + k]);
                hi[k] = std::max(hi[k], xyz[3 * a + k]);
            }
        side = std::max(cell, 1.0e-6);
        for (;;) {
            double cells = 1.0;
            for (int k = 0; k < 3; ++k) {
                dim[k] = (int)std::min(1.0e6, std::floor((hi[k] - lo[k]) / side) + 1.0);
                cells *= dim[k];
            }
            if (cells <= 8.0 * std::max(nat, 1)) break;
            side *= 2.0;
        }
        const int nc = dim[0] * dim[1] * dim[2];
        start.assign(nc + 1, 0);
        std::vector<int> cid(nat);
        for (int a = 0; a < nat; ++a) {
            cid[a] = cell_of(xyz + 3 * a);
            ++start[cid[a] + 1];
        }
        for (int k = 0; k < nc; ++k) start[k + 1] += start[k];
        atoms.resize(nat);
        std::vector<int> fill(start.begin(), start.end() - 1);
        for (int a = 0; a < nat; ++a) atoms[fill[cid[a]]++] = a;  // ascending within a cell
    }
    // calls f(b) for every atom b in the 27 cells around point p
    template <typename Fn>
    void around(const double* p, Fn&& f) const {
        int c[3];
        for (int k = 0; k < 3; ++k) c[k] = std::min(dim[k] - 1, std::max(0, (int)((p[k] - lo[k]) / side)));
        for (int z = std::max(0, c[2] - 1); z <= std::min(dim[2] - 1, c[2] + 1); ++z)
            for (int y = std::max(0, c[1] - 1); y <= std::min(dim[1] - 1, c[1] + 1); ++y)
                for (int x = std::max(0, c[0] - 1); x <= std::min(dim[0] - 1, c[0] + 1); ++x) {
                    const int k = (z * dim[1] + y) * dim[0] + x;
                    for (int q = start[k]; q < start[k + 1]; ++q) f(atoms[q]);
                }
    }
};

}  // namespace

void am1g_dc_release(am1g_ctx* ctx) {
    delete ctx->dc;
    ctx->dc = nullptr;
}

extern "C" {

void am1g_dc_defaults(am1g_dc_options* op) {
    std::memset(op, 0, sizeof(*op));
    op->core_size = 12;
    op->buffer_radius = 11.0;
    op->kt = 0.05;
    op->max_scf = 300;
    op->e_tol = 1.0e-7;
    op->p_tol = 1.0e-6;
    op->mixing = 0.4;
}

int am1g_dc(am1g_ctx* ctx, const am1g_request* rq, const am1g_dc_options* op, am1g_dc_result* res) {
    if (!ctx || !rq || !op || !res) return am1g_set_error(ctx, AM1G_INVALID_INPUT, "null argument");
    std::lock_guard<std::mutex> lock(ctx->mu);
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    res->iterations = 0;
    res->converged = 0;
    if (!(op->buffer_radius > 0.0))
        return res->status = am1g_set_error(ctx, AM1G_INVALID_INPUT, "divide-and-conquer buffer_radius must be positive");
    Prepared prep;
    std::string msg;
    int st = am1g_prepare(ctx, *rq, prep, msg, INT_MAX);
    if (st != AM1G_OK) return res->status = am1g_set_error(ctx, st, "%s", msg.c_str());
    if (!ctx->dc) ctx->dc = new DcBuffers();
    DcBuffers& B = *ctx->dc;
    hipStream_t s = ctx->stream;
    MolDesc md = prep.md;
    const int nat = md.nat, nao = md.nao;
    const size_t nn = (size_t)nao * nao;
    const DevElem* tab = ctx->host_tab + md.method * kMaxZ;

    // ------------------------------------------------------------------ atom tables, integrals, H_core
    std::vector<int> zs(nat), alist, aoff(nat), acls(nat), ao_atom(nao);
    std::vector<double> xyz(rq->xyz_bohr, rq->xyz_bohr + 3 * (size_t)nat);
    {
        int off = 0, nh = 0, nH = 0;
        std::vector<int> heavy, hyd;
        for (int a = 0; a < nat; ++a) {
            zs[a] = rq->z[a];
            aoff[a] = off;
            for (int k = 0; k < tab[zs[a]].norb; ++k) ao_atom[off + k] = a;
            off += tab[zs[a]].norb;
            if (zs[a] != 1) {
                acls[a] = nh++;
                heavy.push_back(a);
            } else {
                acls[a] = nH++;
                hyd.push_back(a);
            }
        }
        alist = heavy;
        alist.insert(alist.end(), hyd.begin(), hyd.end());
    }
    const int64_t npairs = (int64_t)nat * (nat - 1) / 2;
    std::vector<int64_t> pair_off = {0, npairs};
    HIP_TRY(ctx, ctx->mols.reserve(sizeof(MolDesc)));
    HIP_TRY(ctx, ctx->z.reserve(sizeof(int) * nat));
    HIP_TRY(ctx, ctx->xyz.reserve(sizeof(double) * 3 * nat));
    HIP_TRY(ctx, ctx->alist.reserve(sizeof(int) * nat));
    HIP_TRY(ctx, ctx->aoff.reserve(sizeof(int) * nat));
    HIP_TRY(ctx, ctx->acls.reserve(sizeof(int) * nat));
    HIP_TRY(ctx, ctx->later.reserve(sizeof(int) * nat));
    HIP_TRY(ctx, ctx->pair_off.reserve(sizeof(int64_t) * 2));
    HIP_TRY(ctx, ctx->ecore_pair.reserve(sizeof(double) * std::max<int64_t>(npairs, 1)));
    HIP_TRY(ctx, ctx->wpool.reserve(sizeof(double) * std::max<int64_t>(w_pool_size(md.n_heavy, md.n_h), 1)));
    HIP_TRY(ctx, ctx->mats.reserve(sizeof(double) * nn));  // H
    HIP_TRY(ctx, hipMemcpyAsync(ctx->mols.p, &md, sizeof(MolDesc), hipMemcpyHostToDevice, s));
    HIP_TRY(ctx, hipMemcpyAsync(ctx->z.p, zs.data(), sizeof(int) * nat, hipMemcpyHostToDevice, s));
    HIP_TRY(ctx, hipMemcpyAsync(ctx->xyz.p, xyz.data(), sizeof(double) * 3 * nat, hipMemcpyHostToDevice, s));
    HIP_TRY(ctx, hipMemcpyAsync(ctx->alist.p, alist.data(), sizeof(int) * nat, hipMemcpyHostToDevice, s));
    HIP_TRY(ctx, hipMemcpyAsync(ctx->aoff.p, aoff.data(), sizeof(int) * nat, hipMemcpyHostToDevice, s));
    HIP_TRY(ctx, hipMemcpyAsync(ctx->acls.p, acls.data(), sizeof(int) * nat, hipMemcpyHostToDevice, s));
    const std::vector<int> later = later_partners(alist, aoff, md.n_heavy);
    HIP_TRY(ctx, hipMemcpyAsync(ctx->later.p, later.data(), sizeof(int) * nat, hipMemcpyHostToDevice, s));
    HIP_TRY(ctx, hipMemcpyAsync(ctx->pair_off.p, pair_off.data(), sizeof(int64_t) * 2, hipMemcpyHostToDevice, s));

    BatchArgs ba;
    std::memset(&ba, 0, sizeof(ba));
    ba.mols = ctx->mols.as<MolDesc>();
    ba.n_mols = 1;
    ba.z = ctx->z.as<int>();
    ba.xyz = ctx->xyz.as<double>();
    ba.elems = ctx->elems.as<DevElem>();
    ba.mats = ctx->mats.as<double>();  // kH == 0: H_core at the start of the matrix pool
    ba.wpool = ctx->wpool.as<double>();
    ba.alist = ctx->alist.as<int>();
    ba.aoff = ctx->aoff.as<int>();
    ba.acls = ctx->acls.as<int>();
    ba.pair_off = ctx->pair_off.as<int64_t>();
    ba.n_pairs = npairs;
    ba.ecore_pair = ctx->ecore_pair.as<double>();

    HIP_TRY(ctx, B.pdiag.reserve(sizeof(double) * 10 * nat));
    SysArgs sa;
    sa.md = md;
    sa.z = ba.z;
    sa.xyz = ba.xyz;
    sa.alist = ba.alist;
    sa.aoff = ba.aoff;
    sa.acls = ba.acls;
    sa.later = ctx->later.as<int>();
    sa.elems = ba.elems;
    sa.wpool = ba.wpool;
    sa.H = ba.mats;
    sa.pdiag = B.pdiag.as<double>();
    HIP_TRY(ctx, B.tcol.reserve(sizeof(double) * 10 * std::max<int64_t>(1, (int64_t)md.n_heavy * (md.n_heavy - 1) / 2)));
    HIP_TRY(ctx, B.rsum.reserve(sizeof(double) * 10 * std::max(1, md.n_heavy)));
    sa.tcol = B.tcol.as<double>();
    sa.rsum = B.rsum.as<double>();
    sa.far = nullptr;
    if (md.far_cut > 0.0) {
        HIP_TRY(ctx, B.far.reserve(sizeof(double) * (2 * (size_t)nat + 1)));
        sa.far = B.far.as<double>();
    }

    HIP_TRY(ctx, B.partial.reserve(sizeof(double) * std::max<int64_t>({1024 * (kDiisDepth + 1), 2 * fock_energy_parts(md),
                                                                       fermi_scratch_doubles()})));
    HIP_TRY(ctx, B.scal.reserve(sizeof(double) * 64));
    double* scal = B.scal.as<double>();
    auto t_start = std::chrono::steady_clock::now();
    HIP_TRY(ctx, launch_pairs_one(ba, md, s));
    HIP_TRY(ctx, launch_hcore_diag(sa, s));
    HIP_TRY(ctx, launch_sum(npairs, ba.ecore_pair, B.partial.as<double>(), scal + 20, s));

    // ------------------------------------------------------------------ partition and subsystems
    std::vector<std::vector<int>> cores;
    if (op->core_of_atom) {
        int nc = 0;
        for (int a = 0; a < nat; ++a) {
            if (op->core_of_atom[a] < 0)
                return res->status = am1g_set_error(ctx, AM1G_INVALID_INPUT, "negative core index");
            nc = std::max(nc, op->core_of_atom[a] + 1);
        }
        cores.assign(nc, {});
        for (int a = 0; a < nat; ++a) cores[op->core_of_atom[a]].push_back(a);
        cores.erase(std::remove_if(cores.begin(), cores.end(), [](const std::vector<int>& c) { return c.empty(); }),
                    cores.end());
    } else {
        std::vector<int> all(nat);
        std::iota(all.begin(), all.end(), 0);
        bisect_cores(xyz.data(), all, std::max(op->core_size, 1), cores);
    }
    const double r2 = op->buffer_radius * op->buffer_radius;
    auto d2 = [&](int a, int b) {
        double acc = 0.0;
        for (int k = 0; k < 3; ++k) {
            const double d = xyz[3 * b + k] - xyz[3 * a + k];
            acc += d * d;
        }
        return acc;
    };
    const int n_sub = (int)cores.size();
    std::vector<int> core_sub(nat, -1), sub_n(n_sub), ao_list;
    std::vector<unsigned char> ao_core;
    std::vector<int64_t> sub_ao(n_sub + 1, 0), sub_mat(n_sub, 0);
    int64_t mat_total = 0;
    double diag_work = 0.0;
    int largest = 0;
    // subsystem c: its core atoms and every atom within the buffer radius of one of them, in
    // ascending atom order (neighbours from a cell grid instead of all atom pairs)
    const CellGrid grid(xyz.data(), nat, op->buffer_radius);
    std::vector<int> stamp(nat, -1), cand;
    std::vector<char> in_core(nat, 0);
    for (int c = 0; c < n_sub; ++c) {
        for (int a : cores[c]) {
            in_core[a] = 1;
            core_sub[a] = c;
        }
        cand.clear();
        for (int a : cores[c])
            grid.around(xyz.data() + 3 * a, [&](int j) {
                if (stamp[j] != c) {
                    stamp[j] = c;
                    cand.push_back(j);
                }
            });
        std::sort(cand.begin(), cand.end());
        int n = 0;
        for (int j : cand) {
            bool take = in_core[j];
            for (size_t q = 0; !take && q < cores[c].size(); ++q) take = d2(cores[c][q], j) <= r2;
            if (!take) continue;
            for (int k = 0; k < tab[zs[j]].norb; ++k) {
                ao_list.push_back(aoff[j] + k);
                ao_core.push_back(in_core[j] ? 1 : 0);
            }
            n += tab[zs[j]].norb;
        }
        if (n > kMaxBatchAo)
            return res->status = am1g_set_error(ctx, AM1G_INVALID_INPUT,
                                                "subsystem %d has %d basis functions; the batched eigensolver takes at "
                                                "most %d (reduce buffer_radius or the core size)",
                                                c, n, kMaxBatchAo);
        sub_n[c] = n;
        sub_ao[c + 1] = sub_ao[c] + n;
        sub_mat[c] = mat_total;
        mat_total += (int64_t)n * n;
        diag_work += (double)n * n * n;
        largest = std::max(largest, n);
        for (int a : cores[c]) in_core[a] = 0;
    }
    for (int a = 0; a < nat; ++a)
        if (core_sub[a] < 0) return res->status = am1g_set_error(ctx, AM1G_INVALID_INPUT, "atom %d has no core", a);
    const int64_t L = sub_ao[n_sub];
    // retained pattern: AO pairs (r <= c) of atom pairs within the buffer radius
    std::vector<int64_t> pat;
    std::vector<double> mult;
    std::vector<int> nbr;
    for (int a = 0; a < nat; ++a) {
        nbr.clear();
        grid.around(xyz.data() + 3 * a, [&](int b) {
            if (b >= a && d2(a, b) <= r2) nbr.push_back(b);
        });
        for (int b : nbr) {
            for (int i = 0; i < tab[zs[a]].norb; ++i)
                for (int j = 0; j < tab[zs[b]].norb; ++j) {
                    int r = aoff[a] + i, c = aoff[b] + j;
                    if (r > c) std::swap(r, c);
                    if (a == b && r != aoff[a] + i) continue;  // one triangle of a diagonal block
                    pat.push_back((int64_t)r * nao + c);
                }
        }
    }
    std::sort(pat.begin(), pat.end());
    pat.erase(std::unique(pat.begin(), pat.end()), pat.end());
    mult.resize(pat.size());
    for (size_t k = 0; k < pat.size(); ++k) mult[k] = (pat[k] / nao == pat[k] % nao) ? 1.0 : 2.0;
    const int64_t npat = (int64_t)pat.size();
    // DIIS depth under a 512 MiB history budget (divide_conquer.rs:861-880)
    const size_t per_entry = 2 * std::max<size_t>((size_t)npat, 1) * sizeof(double);
    const int max_depth = (int)std::min<size_t>(kDiisDepth, std::max<size_t>((size_t(512) << 20) / per_entry, 2));

    const bool uhf = prep.uhf;
    const int nch = uhf ? 2 : 1;
    HIP_TRY(ctx, B.fbuf.reserve(sizeof(double) * mat_total));
    HIP_TRY(ctx, B.t0.reserve(sizeof(double) * mat_total));
    HIP_TRY(ctx, B.t1.reserve(sizeof(double) * mat_total));
    HIP_TRY(ctx, B.status.reserve(sizeof(int) * 2 * n_sub));
    HIP_TRY(ctx, B.sub_n.reserve(sizeof(int) * n_sub));
    HIP_TRY(ctx, B.sub_ao.reserve(sizeof(int64_t) * (n_sub + 1)));
    HIP_TRY(ctx, B.sub_mat.reserve(sizeof(int64_t) * n_sub));
    HIP_TRY(ctx, B.ao_list.reserve(sizeof(int) * L));
    HIP_TRY(ctx, B.ao_core.reserve(L));
    HIP_TRY(ctx, B.pat.reserve(sizeof(int64_t) * std::max<int64_t>(npat, 1)));
    HIP_TRY(ctx, B.mult.reserve(sizeof(double) * std::max<int64_t>(npat, 1)));
    HIP_TRY(ctx, B.coef.reserve(sizeof(double) * (kDiisDepth + 1)));
    HIP_TRY(ctx, B.ao_atom.reserve(sizeof(int) * nao));
    HIP_TRY(ctx, B.core_sub.reserve(sizeof(int) * nat));
    if (uhf) HIP_TRY(ctx, B.Pt.reserve(sizeof(double) * nn));
    for (int c = 0; c < nch; ++c) {
        DcChannel& ch = B.ch[c];
        HIP_TRY(ctx, ch.F.reserve(sizeof(double) * nn));
        HIP_TRY(ctx, ch.P.reserve(sizeof(double) * nn));
        HIP_TRY(ctx, ch.cbuf.reserve(sizeof(double) * mat_total));
        HIP_TRY(ctx, ch.eps.reserve(sizeof(double) * L));
        HIP_TRY(ctx, ch.proj.reserve(sizeof(double) * L));
        HIP_TRY(ctx, ch.frac.reserve(sizeof(double) * L));
        HIP_TRY(ctx, ch.pout.reserve(sizeof(double) * std::max<int64_t>(npat, 1)));
        HIP_TRY(ctx, ch.hin.reserve(sizeof(double) * kDiisDepth * std::max<int64_t>(npat, 1)));
        HIP_TRY(ctx, ch.hres.reserve(sizeof(double) * kDiisDepth * std::max<int64_t>(npat, 1)));
        HIP_TRY(ctx, ch.ring.reserve(sizeof(int) * (kDiisDepth + 1)));
    }
    HIP_TRY(ctx, hipMemcpyAsync(B.sub_n.p, sub_n.data(), sizeof(int) * n_sub, hipMemcpyHostToDevice, s));
    HIP_TRY(ctx, hipMemcpyAsync(B.sub_ao.p, sub_ao.data(), sizeof(int64_t) * (n_sub + 1), hipMemcpyHostToDevice, s));
    HIP_TRY(ctx, hipMemcpyAsync(B.sub_mat.p, sub_mat.data(), sizeof(int64_t) * n_sub, hipMemcpyHostToDevice, s));
    HIP_TRY(ctx, hipMemcpyAsync(B.ao_list.p, ao_list.data(), sizeof(int) * L, hipMemcpyHostToDevice, s));
    HIP_TRY(ctx, hipMemcpyAsync(B.ao_core.p, ao_core.data(), L, hipMemcpyHostToDevice, s));
    HIP_TRY(ctx, hipMemcpyAsync(B.pat.p, pat.data(), sizeof(int64_t) * npat, hipMemcpyHostToDevice, s));
    HIP_TRY(ctx, hipMemcpyAsync(B.mult.p, mult.data(), sizeof(double) * npat, hipMemcpyHostToDevice, s));
    HIP_TRY(ctx, hipMemcpyAsync(B.ao_atom.p, ao_atom.data(), sizeof(int) * nao, hipMemcpyHostToDevice, s));
    HIP_TRY(ctx, hipMemcpyAsync(B.core_sub.p, core_sub.data(), sizeof(int) * nat, hipMemcpyHostToDevice, s));

    // initial density: SAD, split between the spins in proportion to their electron counts
    // (divide_conquer.rs:1143-1159)
    const double n_alpha = md.n_alpha, n_beta = md.n_beta;
    const double n_tot = std::max(n_alpha + n_beta, 1.0);
    double* P[2] = {B.ch[0].P.as<double>(), uhf ? B.ch[1].P.as<double>() : nullptr};
    double* F[2] = {B.ch[0].F.as<double>(), uhf ? B.ch[1].F.as<double>() : nullptr};
    for (int c = 0; c < nch; ++c) {
        HIP_TRY(ctx, hipMemsetAsync(P[c], 0, sizeof(double) * nn, s));
        HIP_TRY(ctx, launch_sad(sa, uhf ? (c == 0 ? n_alpha : n_beta) / n_tot : 1.0, P[c], s));
    }
    double* Pt = uhf ? B.Pt.as<double>() : P[0];

    DcArgs d0;
    d0.n_sub = n_sub;
    d0.lds_n = ctx->tun.jacobi_lds_n;
    d0.tri_max = std::min(eigen_tri_max(), ctx->tun.tri_max);
    d0.sub_n = B.sub_n.as<int>();
    d0.sub_ao = B.sub_ao.as<int64_t>();
    d0.sub_mat = B.sub_mat.as<int64_t>();
    d0.ao_list = B.ao_list.as<int>();
    d0.ao_core = B.ao_core.as<unsigned char>();
    d0.fbuf = B.fbuf.as<double>();
    d0.t0 = B.t0.as<double>();
    d0.t1 = B.t1.as<double>();
    d0.cap = uhf ? 1.0 : 2.0;
    d0.tol_abs = 0.0;
    d0.red = B.partial.as<double>();  // at least fermi_scratch_doubles()
    DcArgs dch[2];
    for (int c = 0; c < 2; ++c) {
        dch[c] = d0;
        DcChannel& ch = B.ch[c < nch ? c : 0];
        dch[c].cbuf = ch.cbuf.as<double>();
        dch[c].eps = ch.eps.as<double>();
        dch[c].proj = ch.proj.as<double>();
        dch[c].frac = ch.frac.as<double>();
        dch[c].status = B.status.as<int>() + c * n_sub;
    }
    const double n_elec_ch[2] = {uhf ? n_alpha : n_alpha + n_beta, n_beta};

    // scal layout: 0 sum Pt.H (+ P.F for a closed shell), 3/4 sum P_s.F_s, 20 E_core,
    // per channel c at 32 + 16 c: mu, T*S, count, -, and from +4 the DIIS sums (sum r^2, dots)
    auto cbase = [](int c) { return 32 + 16 * c; };
    hipEvent_t ev[6];
    for (auto& e : ev) HIP_TRY(ctx, hipEventCreate(&e));
    double phase[5] = {0, 0, 0, 0, 0};
    struct HostDiis {
        std::vector<int> ring;                  // physical slots in logical order
        std::vector<std::vector<double>> dots;  // lower triangle of <r_i, r_j>
    } hd[2];
    std::vector<int> sub_status(2 * (size_t)n_sub);
    std::vector<double> host_scal(64);
    // host staging for per-iteration uploads; every iteration synchronizes before reusing them
    std::vector<double> coef_host(kDiisDepth + 1);
    std::vector<int> ring_host(2 * (kDiisDepth + 1));
    double e_old = 0.0, mu[2] = {0.0, 0.0}, ts[2] = {0.0, 0.0};
    bool converged = false;
    int iterations = 0, status = AM1G_OK;
    double* partial = B.partial.as<double>();

    // Fock matrices of the current densities and the electronic energy sums
    auto fock_energy = [&]() -> int {
        if (uhf) {
            HIP_TRY(ctx, launch_add((int64_t)nn, P[0], P[1], Pt, s));
            HIP_TRY(ctx, launch_fock(sa, Pt, P[0], 1.0, F[0], s, partial, scal + 3, false));
            HIP_TRY(ctx, launch_fock(sa, Pt, P[1], 1.0, F[1], s, partial, scal + 4, false));
        } else {
            HIP_TRY(ctx, launch_fock(sa, P[0], P[0], 0.5, F[0], s, partial, scal + 0, false));
        }
        return AM1G_OK;
    };
    // E_elec = 1/2 sum P (H + F) closed shell, 1/2 sum_s P_s (H + F_s) open shell
    auto e_of = [&](const std::vector<double>& h) { return uhf ? 0.5 * (h[3] + h[4]) : 0.5 * h[0]; };

    // ------------------------------------------------------------------ SCF loop (divide_conquer.rs:1033-1254)
    for (int iter = 0; iter < op->max_scf; ++iter) {
        iterations = iter + 1;
        HIP_TRY(ctx, hipEventRecord(ev[0], s));
        if ((st = fock_energy()) != AM1G_OK) return res->status = st;
        HIP_TRY(ctx, hipEventRecord(ev[1], s));
        for (int c = 0; c < nch; ++c) {
            dch[c].warm = iter > 0 ? 1 : 0;
            HIP_TRY(ctx, launch_dc_eigen(dch[c], F[c], nao, s));
        }
        HIP_TRY(ctx, hipEventRecord(ev[2], s));
        for (int c = 0; c < nch; ++c) {
            const DcArgs& d = dch[c];
            if (op->kt > 0.0) {
                HIP_TRY(ctx, launch_fermi(d, L, n_elec_ch[c], op->kt, scal + cbase(c), s));
                continue;
            }
            // aufbau filling (fermi.rs:153-206) on the host: sort the levels, fill degenerate
            // groups evenly
            std::vector<double> e(L), w(L), f(L, 0.0);
            HIP_TRY(ctx, hipMemcpyAsync(e.data(), d.eps, sizeof(double) * L, hipMemcpyDeviceToHost, s));
            HIP_TRY(ctx, hipMemcpyAsync(w.data(), d.proj, sizeof(double) * L, hipMemcpyDeviceToHost, s));
            HIP_TRY(ctx, hipStreamSynchronize(s));
            for (auto& x : w) x *= d.cap;
            std::vector<int64_t> order(L);
            std::iota(order.begin(), order.end(), 0);
            std::stable_sort(order.begin(), order.end(), [&](int64_t a, int64_t b) { return e[a] < e[b]; });
            double left = n_elec_ch[c], fermi = L ? e[order.back()] : 0.0;
            int64_t i = 0;
            while (i < L) {
                if (left <= 1e-14) {
                    fermi = 0.5 * (fermi + e[order[i]]);
                    break;
                }
                const double ei = e[order[i]];
                int64_t j = i;
                double gc = 0.0;
                while (j < L && std::fabs(e[order[j]] - ei) < 1e-10) gc += w[order[j++]];
                const double take = std::min(left, gc);
                const double fr = gc > 0.0 ? take / gc : 0.0;
                for (int64_t k = i; k < j; ++k) f[order[k]] = fr;
                left -= take;
                fermi = ei;
                i = j;
            }
            HIP_TRY(ctx, hipMemcpyAsync(d.frac, f.data(), sizeof(double) * L, hipMemcpyHostToDevice, s));
            host_scal[cbase(c)] = fermi;
            host_scal[cbase(c) + 1] = 0.0;
            HIP_TRY(ctx, hipMemcpyAsync(scal + cbase(c), &host_scal[cbase(c)], sizeof(double) * 2,
                                        hipMemcpyHostToDevice, s));
            HIP_TRY(ctx, hipStreamSynchronize(s));
        }
        HIP_TRY(ctx, hipEventRecord(ev[3], s));
        // subsystem densities, assembly and the DIIS sums per channel
        int slot[2] = {0, 0}, mdots[2] = {0, 0};
        std::vector<int> ring_next[2];
        for (int c = 0; c < nch; ++c) {
            DcChannel& ch = B.ch[c];
            HIP_TRY(ctx, launch_dc_density(dch[c], s));
            HIP_TRY(ctx, launch_dc_gather(dch[c], npat, B.pat.as<int64_t>(), B.ao_atom.as<int>(),
                                          B.core_sub.as<int>(), nao, ch.pout.as<double>(), s));
            // next history slot: evict the oldest when full
            const std::vector<int>& ring = hd[c].ring;
            slot[c] = (int)ring.size() < max_depth ? (int)ring.size() : ring.front();
            ring_next[c] = ring;
            if ((int)ring_next[c].size() >= max_depth) ring_next[c].erase(ring_next[c].begin());
            int* rh = ring_host.data() + c * (kDiisDepth + 1);
            std::copy(ring_next[c].begin(), ring_next[c].end(), rh);
            rh[ring_next[c].size()] = -1;  // the new entry, dotted with itself
            mdots[c] = (int)ring_next[c].size() + 1;
            HIP_TRY(ctx, hipMemcpyAsync(ch.ring.p, rh, sizeof(int) * mdots[c], hipMemcpyHostToDevice, s));
            HIP_TRY(ctx, launch_dc_hist(npat, B.pat.as<int64_t>(), B.mult.as<double>(), P[c], ch.pout.as<double>(),
                                        ch.hin.as<double>() + (int64_t)slot[c] * npat,
                                        ch.hres.as<double>() + (int64_t)slot[c] * npat, ch.hres.as<double>(),
                                        ch.ring.as<int>(), mdots[c], npat, partial, scal + cbase(c) + 4, s));
        }
        HIP_TRY(ctx, hipEventRecord(ev[4], s));
        HIP_TRY(ctx, hipMemcpyAsync(host_scal.data(), scal, sizeof(double) * 64, hipMemcpyDeviceToHost, s));
        HIP_TRY(ctx, hipMemcpyAsync(sub_status.data(), B.status.p, sizeof(int) * nch * n_sub, hipMemcpyDeviceToHost,
                                    s));
        HIP_TRY(ctx, hipStreamSynchronize(s));
        for (int q = 0; q < nch * n_sub; ++q)
            if (sub_status[q] < 0) status = AM1G_LINEAR_ALGEBRA;
        if (status != AM1G_OK) break;
        const double e_elec = e_of(host_scal);
        double dp = 0.0;
        for (int c = 0; c < nch; ++c) {
            mu[c] = host_scal[cbase(c)];
            ts[c] = host_scal[cbase(c) + 1];
            dp = std::max(dp, std::sqrt(host_scal[cbase(c) + 4] / (double)std::max<size_t>(nn, 1)));
        }
        const double de = std::fabs(e_elec - e_old);
        e_old = e_elec;
        if (!std::isfinite(e_elec) || !std::isfinite(dp)) {
            status = AM1G_SCF_DIVERGED;
            break;
        }
        if (iter > 0 && de < op->e_tol && dp < op->p_tol) {
            for (int c = 0; c < nch; ++c)
                HIP_TRY(ctx, launch_dc_update(npat, B.pat.as<int64_t>(), nao, nullptr, nullptr, nullptr, nullptr, 0,
                                              npat, -1.0, P[c], B.ch[c].pout.as<double>(), P[c], s));
            converged = true;
        } else {
            for (int c = 0; c < nch; ++c) {
                DcChannel& ch = B.ch[c];
                HostDiis& h = hd[c];
                // DIIS bookkeeping: the new entry takes `slot`; its dots extend the cached Gram matrix
                h.ring = ring_next[c];
                if ((int)h.dots.size() >= max_depth) {
                    h.dots.erase(h.dots.begin());
                    for (auto& row : h.dots) row.erase(row.begin());
                }
                std::vector<double> row(mdots[c]);
                for (int j = 0; j < mdots[c]; ++j) row[j] = host_scal[cbase(c) + 5 + j];
                h.dots.push_back(row);
                h.ring.push_back(slot[c]);
                const int n = (int)h.ring.size();
                bool used = false;
                if (n >= 2) {
                    const int dim = n + 1;
                    std::vector<double> Bm((size_t)dim * dim, 0.0), rhs(dim, 0.0), cf;
                    double scale = 0.0;
                    for (int i = 0; i < n; ++i) {
                        for (int j = 0; j <= i; ++j) {
                            Bm[i * dim + j] = Bm[j * dim + i] = h.dots[i][j];
                            scale = std::max(scale, std::fabs(h.dots[i][j]));
                        }
                        Bm[i * dim + n] = Bm[n * dim + i] = -1.0;
                    }
                    if (scale > 0.0) {
                        for (int i = 0; i < n; ++i) {
                            for (int j = 0; j < n; ++j) Bm[i * dim + j] /= scale;
                            Bm[i * dim + i] += 1e-10;
                        }
                        rhs[n] = -1.0;
                        if (lu_solve(Bm, rhs, dim, cf)) {
                            bool finite = true;
                            for (int i = 0; i < n; ++i) finite = finite && std::isfinite(cf[i]);
                            if (finite) {
                                std::copy(cf.begin(), cf.begin() + n, coef_host.begin());
                                int* rh = ring_host.data() + c * (kDiisDepth + 1);
                                std::copy(h.ring.begin(), h.ring.end(), rh);
                                HIP_TRY(ctx, hipMemcpyAsync(B.coef.p, coef_host.data(), sizeof(double) * n,
                                                            hipMemcpyHostToDevice, s));
                                HIP_TRY(ctx, hipMemcpyAsync(ch.ring.p, rh, sizeof(int) * n, hipMemcpyHostToDevice, s));
                                HIP_TRY(ctx, launch_dc_update(npat, B.pat.as<int64_t>(), nao, B.coef.as<double>(),
                                                              ch.hin.as<double>(), ch.hres.as<double>(),
                                                              ch.ring.as<int>(), n, npat, 0.0, P[c],
                                                              ch.pout.as<double>(), P[c], s));
                                // coef_host is reused by the other channel: finish this upload first
                                HIP_TRY(ctx, hipStreamSynchronize(s));
                                used = true;
                            }
                        }
                    }
                }
                if (!used)
                    HIP_TRY(ctx, launch_dc_update(npat, B.pat.as<int64_t>(), nao, nullptr, nullptr, nullptr, nullptr,
                                                  0, npat, op->mixing, P[c], ch.pout.as<double>(), P[c], s));
            }
        }
        HIP_TRY(ctx, hipEventRecord(ev[5], s));
        HIP_TRY(ctx, hipEventSynchronize(ev[5]));
        for (int k = 0; k < 5; ++k) {
            float ms = 0.f;
            HIP_TRY(ctx, hipEventElapsedTime(&ms, ev[k], ev[k + 1]));
            phase[k] += ms;
        }
        if (converged) break;
    }
    for (auto& e : ev) (void)hipEventDestroy(e);

    // ------------------------------------------------------------------ final energy and properties
    if ((st = fock_energy()) != AM1G_OK) return res->status = st;
    HIP_TRY(ctx, hipMemcpyAsync(host_scal.data(), scal, sizeof(double) * 64, hipMemcpyDeviceToHost, s));
    std::vector<double> pd(10 * (size_t)nat), eps[2];
    HIP_TRY(ctx, hipMemcpyAsync(pd.data(), sa.pdiag, sizeof(double) * 10 * nat, hipMemcpyDeviceToHost, s));
    for (int c = 0; c < nch; ++c) {
        eps[c].resize(L);
        HIP_TRY(ctx, hipMemcpyAsync(eps[c].data(), dch[c].eps, sizeof(double) * L, hipMemcpyDeviceToHost, s));
    }
    double zv = 0.0;  // sum Z_a V_a of the far field (long_range_energy_term, fock.rs:200-232)
    if (sa.far) HIP_TRY(ctx, hipMemcpyAsync(&zv, sa.far + 2 * (size_t)nat, sizeof(double), hipMemcpyDeviceToHost, s));
    HIP_TRY(ctx, hipStreamSynchronize(s));
    const double electronic = e_of(host_scal) + 0.5 * zv;
    const double ecore = host_scal[20];
    double eiso = 0.0, eheat = 0.0;
    for (int a = 0; a < nat; ++a) {
        eiso += tab[zs[a]].e_isol;
        eheat += tab[zs[a]].eheat_ev;
        // Mulliken population from the packed diagonal block of the total density (entries 0,
        // 2, 5, 9 are the diagonal; rows in class order, heavy atoms first)
        const double* row = pd.data() + 10 * (size_t)(zs[a] == 1 ? md.n_heavy + acls[a] : acls[a]);
        double pop = row[0];
        if (zs[a] != 1) pop += row[2] + row[5] + row[9];
        if (res->charges) res->charges[a] = tab[zs[a]].core - pop;
    }
    double gap = INFINITY;  // smallest_gap per channel (divide_conquer.rs:659-683), the smaller one
    for (int c = 0; c < nch; ++c)
        for (int q = 0; q < n_sub; ++q) {
            double below = -INFINITY, above = INFINITY;
            for (int64_t i = sub_ao[q]; i < sub_ao[q + 1]; ++i) {
                if (eps[c][i] <= mu[c]) below = std::max(below, eps[c][i]);
                else above = std::min(above, eps[c][i]);
            }
            if (std::isfinite(below) && std::isfinite(above)) gap = std::min(gap, above - below);
        }
    if (res->density) HIP_TRY(ctx, hipMemcpy(res->density, Pt, sizeof(double) * nn, hipMemcpyDeviceToHost));
    if (res->spin_density && uhf) {  // P_alpha - P_beta
        std::vector<double> pb(nn);
        HIP_TRY(ctx, hipMemcpy(res->spin_density, P[0], sizeof(double) * nn, hipMemcpyDeviceToHost));
        HIP_TRY(ctx, hipMemcpy(pb.data(), P[1], sizeof(double) * nn, hipMemcpyDeviceToHost));
        for (size_t k = 0; k < nn; ++k) res->spin_density[k] -= pb[k];
    }
    res->iterations = iterations;
    res->converged = converged;
    res->unrestricted = uhf ? 1 : 0;
    res->n_subsystems = n_sub;
    res->largest_subsystem_aos = largest;
    res->electronic_ev = electronic;
    res->core_ev = ecore;
    // the field's nuclear half (divide_conquer.rs:494-500); its electronic half is in H_core
    res->total_ev = electronic + ecore + md.field_nuc;
    res->heat_of_formation_kcal = (electronic + ecore + md.field_nuc - eiso + eheat) * kEvToKcal;
    res->fermi_ev = mu[0];
    res->fermi_beta_ev = uhf ? mu[1] : mu[0];
    res->entropy_ev = ts[0] + (uhf ? ts[1] : 0.0);
    res->homo_lumo_gap_ev = std::isfinite(gap) ? gap : 0.0;
    res->diagonalization_work = diag_work * nch;
    res->retained_pattern = npat;
    for (int k = 0; k < 5; ++k) res->phase_ms[k] = phase[k];
    ctx->last_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
    if (status == AM1G_OK && !converged) status = AM1G_SCF_NOT_CONVERGED;
    res->status = status;
    if (status == AM1G_SCF_NOT_CONVERGED) am1g_set_error(ctx, status, "divide-and-conquer SCF did not converge");
    if (status == AM1G_LINEAR_ALGEBRA) am1g_set_error(ctx, status, "a subsystem eigensolve did not converge");
    if (status == AM1G_SCF_DIVERGED) am1g_set_error(ctx, status, "divide-and-conquer SCF diverged");
    return status;
}

}  // extern "C"
