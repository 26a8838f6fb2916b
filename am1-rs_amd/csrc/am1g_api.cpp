// am1g host runtime: context, device pools, request validation and the C ABI of include/am1g.h.
//
// Validation and result assembly follow am1-rs run_am1 (src/scf.rs:294-468): electron count and
// multiplicity parity, RHF/UHF selection, ScfNotConverged reported after every output is filled.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <climits>
#include <cstring>
#include <dlfcn.h>
#include <fstream>
#include <sstream>
#include <mutex>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

#include "../../include/am1g.h"
#include "am1g_common.hpp"
#include "am1g_launch.hpp"
#include "am1g_ctx.hpp"

using namespace am1g;

// Error text of the last failing call made on this thread. Each thread owns its copy, so threads
// that share a context never read a buffer another thread is rewriting.
thread_local std::string g_tls_error;

int am1g_set_error(am1g_ctx* ctx, int code, const char* fmt, ...) {
    (void)ctx;
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_tls_error = buf;
    return code;
}

namespace {
bool set_tunable(Tunables& t, const std::string& name, double v) {
    if (name == "route_large_ao") t.route_large_ao = (int)v;
    else if (name == "jacobi_lds_n") t.jacobi_lds_n = std::min(192, std::max(0, (int)v));
    else if (name == "eig_graphs") t.eig_graphs = v != 0.0;
    else if (name == "serve_large_batch") t.serve_large_batch = std::max(1, (int)v);
    else if (name == "tri_max") t.tri_max = std::min(256, std::max(0, (int)v));
    else if (name == "pairs_block") t.pairs_block = (v == 64.0 || v == 128.0) ? (int)v : 256;
    else if (name == "tri_blocked_min") t.tri_blocked_min = std::max(3, (int)v);
    else if (name == "batch_own_pairs") t.batch_own_pairs = v != 0.0;
    else if (name == "sp2") t.sp2 = v != 0.0;
    else if (name == "batch_small_wg") t.batch_small_wg = v != 0.0;
    else return false;
    return true;
}

bool get_tunable(const Tunables& t, const std::string& name, double& v) {
    if (name == "route_large_ao") v = t.route_large_ao;
    else if (name == "jacobi_lds_n") v = t.jacobi_lds_n;
    else if (name == "eig_graphs") v = t.eig_graphs;
    else if (name == "serve_large_batch") v = t.serve_large_batch;
    else if (name == "tri_max") v = t.tri_max;
    else if (name == "pairs_block") v = t.pairs_block;
    else if (name == "tri_blocked_min") v = t.tri_blocked_min;
    else if (name == "batch_own_pairs") v = t.batch_own_pairs;
    else if (name == "sp2") v = t.sp2;
    else if (name == "batch_small_wg") v = t.batch_small_wg;
    else return false;
    return true;
}
}  // namespace

void am1g_load_tunables(am1g_ctx* ctx) {
    std::string path;
    if (const char* env = std::getenv("AM1G_TUNE_FILE")) {
        path = env;
    } else {
        Dl_info info;
        if (dladdr(reinterpret_cast<void*>(&am1g_load_tunables), &info) && info.dli_fname) {
            path = info.dli_fname;
            const size_t slash = path.find_last_of('/');
            path = (slash == std::string::npos ? std::string() : path.substr(0, slash + 1)) + "tuned.txt";
        }
    }
    if (path.empty()) return;
    std::ifstream in(path);
    if (!in) return;
    std::string line;
    bool any = false;
    while (std::getline(in, line)) {
        const size_t hash = line.find('#');
        if (hash != std::string::npos) line.resize(hash);
        std::istringstream ls(line);
        std::string name;
        double v;
        if (ls >> name >> v) any = set_tunable(ctx->tun, name, v) || any;
    }
    if (any) ctx->tune_source = path;
}

// run_am1's input checks (scf.rs:300-366) plus the element checks of Basis::build.
int am1g_prepare(am1g_ctx* ctx, const am1g_request& rq, Prepared& out, std::string& msg, int max_ao) {
    char buf[256];
    if (rq.n_atoms < 1 || !rq.z || !rq.xyz_bohr) {
        msg = "molecule has no atoms";
        return AM1G_INVALID_INPUT;
    }
    if (max_ao <= kMaxBatchAo && rq.n_atoms > kMaxAtoms) {
        std::snprintf(buf, sizeof(buf), "molecule has %d atoms; this entry point takes at most %d", rq.n_atoms,
                      kMaxAtoms);
        msg = buf;
        return AM1G_INVALID_INPUT;
    }
    if (rq.multiplicity < 1) {
        msg = "multiplicity must be >= 1";
        return AM1G_INVALID_INPUT;
    }
    if (rq.method != 0 && rq.method != 1) {
        msg = "unknown method";
        return AM1G_INVALID_INPUT;
    }
    if (rq.bcc && rq.method != AM1G_METHOD_AM1) {
        // the bond charge corrections were fitted to AM1 Mulliken charges (native.py:841-842)
        msg = "AM1-BCC charges need the AM1 method";
        return AM1G_INVALID_INPUT;
    }
    const DevElem* tab = ctx->host_tab + rq.method * kMaxZ;
    int nao = 0, nh = 0, nH = 0;
    double n_elec = 0.0;
    for (int i = 0; i < rq.n_atoms; ++i) {
        int zz = rq.z[i];
        if (zz <= 0 || zz >= kMaxZ || !tab[zz].valid) {
            std::snprintf(buf, sizeof(buf), "%s has no parameters for element Z=%d",
                          rq.method == 0 ? "AM1" : "RM1", zz);
            msg = buf;
            return AM1G_ELEMENT_NOT_PARAMETERIZED;
        }
        nao += tab[zz].norb;
        (zz == 1 ? nH : nh)++;
        n_elec += tab[zz].core;
        for (int k = 0; k < 3; ++k)
            if (!std::isfinite(rq.xyz_bohr[3 * i + k])) {
                msg = "non-finite coordinate";
                return AM1G_INVALID_INPUT;
            }
    }
    n_elec -= rq.charge;
    long long ne = std::llround(n_elec);
    if (std::fabs(n_elec - (double)ne) > 1.0e-6 || ne < 0) {
        std::snprintf(buf, sizeof(buf), "invalid electron count %g", n_elec);
        msg = buf;
        return AM1G_INVALID_INPUT;
    }
    long long nun = rq.multiplicity - 1;
    if (ne - nun < 0 || (ne - nun) % 2 != 0) {
        std::snprintf(buf, sizeof(buf),
                      "electron count %lld is incompatible with multiplicity %d (need same parity)", ne,
                      rq.multiplicity);
        msg = buf;
        return AM1G_INVALID_INPUT;
    }
    int na = (int)((ne + nun) / 2), nb = (int)((ne - nun) / 2);
    bool closed = na == nb;
    bool uhf;
    if (rq.reference == AM1G_REF_AUTO) uhf = !closed;
    else if (rq.reference == AM1G_REF_UHF) uhf = true;
    else {
        if (!closed) {
            std::snprintf(buf, sizeof(buf),
                          "restricted (RHF) reference requires a closed shell, but multiplicity %d gives %d alpha "
                          "and %d beta electrons; use the unrestricted (UHF) reference for open-shell systems",
                          rq.multiplicity, na, nb);
            msg = buf;
            return AM1G_INVALID_INPUT;
        }
        uhf = false;
    }
    if (nao > max_ao) {
        std::snprintf(buf, sizeof(buf), "%d basis functions; this entry point takes at most %d", nao, max_ao);
        msg = buf;
        return AM1G_INVALID_INPUT;
    }
    if (na > nao || nb > nao) {
        msg = "more electrons than the basis can hold";
        return AM1G_INVALID_INPUT;
    }
    MolDesc& md = out.md;
    std::memset(&md, 0, sizeof(md));
    md.nat = rq.n_atoms;
    md.nao = nao;
    md.n_alpha = na;
    md.n_beta = nb;
    md.method = rq.method;
    md.uhf = uhf ? 1 : 0;
    md.accel = rq.use_diis ? rq.accelerator : 0;
    md.use_diis = rq.use_diis ? 1 : 0;
    md.max_scf = rq.max_scf;
    md.precision = rq.precision;
    md.n_heavy = nh;
    md.n_h = nH;
    md.e_tol = rq.e_tol;
    md.p_tol = rq.p_tol;
    md.adiis_switch = rq.adiis_switch;
    // a cutoff that is not > 0 (NaN included) keeps every pair (farfield.rs:97-104)
    md.far_cut = rq.multipole_cutoff > 0.0 ? rq.multipole_cutoff : 0.0;
    if (rq.electric_field) {
        // the nuclear half of -mu.F (dipole.rs:146-153); the electronic half enters H_core
        double nuc = 0.0;
        for (int k = 0; k < 3; ++k) {
            if (!std::isfinite(rq.electric_field[k])) {
                msg = "non-finite electric field";
                return AM1G_INVALID_INPUT;
            }
            md.field[k] = rq.electric_field[k];
        }
        for (int i = 0; i < rq.n_atoms; ++i)
            for (int k = 0; k < 3; ++k) nuc += tab[rq.z[i]].core * rq.xyz_bohr[3 * i + k] * md.field[k];
        md.field_nuc = -nuc;
    }
    out.nat = rq.n_atoms;
    out.uhf = uhf;
    return AM1G_OK;
}

namespace {

}  // namespace

void am1g_apply_bcc(const am1g_request& rq, const double* mulliken, am1g_result& r) {
    if (!rq.bcc || !r.bcc_charges || !mulliken || (r.status != AM1G_OK && r.status != AM1G_SCF_NOT_CONVERGED)) return;
    int nw = 0;
    am1g_bcc_charges(rq.n_atoms, rq.z, rq.xyz_bohr, mulliken, r.bcc_charges, nullptr, &nw, nullptr, 0);
    r.bcc_warnings = nw;
}

void am1g_fill_result(const MolResult& r, const MolDesc& md, am1g_result& out) {
    out.status = r.status;
    out.nao = md.nao;
    out.n_occ = md.n_alpha;
    out.n_occ_beta = md.n_beta;
    out.iterations = r.iterations;
    out.converged = r.converged;
    out.unrestricted = md.uhf;
    out.electronic_ev = r.electronic_ev;
    out.core_ev = r.core_ev;
    out.total_ev = r.total_ev;
    out.heat_of_formation_kcal = r.heat_of_formation_kcal;
    out.homo_ev = r.homo_ev;
    out.lumo_ev = r.lumo_ev;
    out.homo_beta_ev = r.homo_beta_ev;
    out.lumo_beta_ev = r.lumo_beta_ev;
    out.has_homo = std::isfinite(r.homo_ev);
    out.has_lumo = std::isfinite(r.lumo_ev);
    out.has_homo_beta = md.uhf && std::isfinite(r.homo_beta_ev);
    out.has_lumo_beta = md.uhf && std::isfinite(r.lumo_beta_ev);
    for (int k = 0; k < 3; ++k) out.dipole_debye[k] = r.dipole[k];
    out.dipole_magnitude = std::sqrt(r.dipole[0] * r.dipole[0] + r.dipole[1] * r.dipole[1] + r.dipole[2] * r.dipole[2]);
    out.field_nuclear_ev = md.field_nuc;
    out.residual = r.residual;
    out.fp32_iterations = (int)r.fp32_iterations;
}

static int batch_impl(am1g_ctx* ctx, int n, const am1g_request* rqs, am1g_result* results,
                      const std::function<void(int)>& on_done, bool noalloc);

extern "C" {

void am1g_request_defaults(am1g_request* rq) {
    std::memset(rq, 0, sizeof(*rq));
    rq->multiplicity = 1;
    rq->method = AM1G_METHOD_AM1;
    rq->reference = AM1G_REF_AUTO;
    rq->accelerator = AM1G_ACCEL_ADIIS_CDIIS;
    rq->use_diis = 1;
    rq->max_scf = 200;
    rq->e_tol = 1.0e-8;
    rq->p_tol = 1.0e-7;
    rq->adiis_switch = 0.1;
    rq->precision = AM1G_PRECISION_FP64;
}

int am1g_create(int device, am1g_ctx** out) {
    if (!out) return am1g_set_error(nullptr, AM1G_INVALID_INPUT, "null output pointer");
    *out = nullptr;
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count == 0)
        return am1g_set_error(nullptr, AM1G_DEVICE_ERROR, "no HIP device available (%s)", hipGetErrorString(e));
    if (device < 0 || device >= count)
        return am1g_set_error(nullptr, AM1G_INVALID_INPUT, "device %d out of range (%d devices)", device, count);
    am1g_ctx* ctx = new am1g_ctx();
    ctx->device = device;
    HIP_TRY(ctx, hipSetDevice(device));
    hipDeviceProp_t prop;
    HIP_TRY(ctx, hipGetDeviceProperties(&prop, device));
    ctx->cus = prop.multiProcessorCount;
    HIP_TRY(ctx, hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
    // device address of this GPU's scratch array for am1g_scf_batch_noalloc (once per device)
    HIP_TRY(ctx, scratch_init(device));
    am1g_load_tunables(ctx);
    HIP_TRY(ctx, hipEventCreate(&ctx->ev0));
    HIP_TRY(ctx, hipEventCreate(&ctx->ev1));
    HIP_TRY(ctx, hipEventCreateWithFlags(&ctx->ev2, hipEventDisableTiming));
    HIP_TRY(ctx, hipStreamCreateWithFlags(&ctx->stream2, hipStreamNonBlocking));
    build_element_tables(ctx->host_tab);
    HIP_TRY(ctx, ctx->elems.reserve(sizeof(ctx->host_tab)));
    HIP_TRY(ctx, hipMemcpy(ctx->elems.p, ctx->host_tab, sizeof(ctx->host_tab), hipMemcpyHostToDevice));
    *out = ctx;
    return AM1G_OK;
}

void am1g_destroy(am1g_ctx* ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    am1g_dc_release(ctx);
    am1g_large_release(ctx);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
    if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
    if (ctx->ev2) (void)hipEventDestroy(ctx->ev2);
    if (ctx->stream2) (void)hipStreamDestroy(ctx->stream2);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

const char* am1g_last_error(const am1g_ctx* ctx) {
    (void)ctx;
    return g_tls_error.c_str();
}

double am1g_last_kernel_ms(const am1g_ctx* ctx) { return ctx ? ctx->last_ms : 0.0; }

int am1g_element_info(int z, int method, double* out) {
    static DevElem tab[kMethods * kMaxZ];
    static std::once_flag once;
    std::call_once(once, [] { build_element_tables(tab); });
    if (method < 0 || method >= kMethods || z <= 0 || z >= kMaxZ || !tab[method * kMaxZ + z].valid)
        return AM1G_ELEMENT_NOT_PARAMETERIZED;
    const DevElem& e = tab[method * kMaxZ + z];
    double v[23] = {(double)e.n, (double)e.norb, e.uss, e.upp, e.zs, e.zp, e.bs, e.bp, e.gss, e.gsp, e.gpp, e.gp2,
                    e.hsp, e.alpha, e.core, e.eheat_ev, e.e_isol, e.dd, e.qq, e.rho0, e.rho1, e.rho2, (double)e.ng};
    std::memcpy(out, v, sizeof(v));
    return AM1G_OK;
}

int am1g_scf_batch(am1g_ctx* ctx, int n, const am1g_request* rqs, am1g_result* results) {
    return am1g_scf_batch_stream(ctx, n, rqs, results, std::function<void(int)>());
}

int am1g_set_tunable(am1g_ctx* ctx, const char* name, double value) {
    if (!ctx || !name) return am1g_set_error(ctx, AM1G_INVALID_INPUT, "null argument");
    std::lock_guard<std::mutex> lock(ctx->mu);
    if (!set_tunable(ctx->tun, name, value)) return am1g_set_error(ctx, AM1G_INVALID_INPUT, "unknown tunable %s", name);
    if (ctx->eig) {
        ctx->eig->use_graphs = ctx->tun.eig_graphs != 0;
        ctx->eig->blocked_min = ctx->tun.tri_blocked_min;
    }
    return AM1G_OK;
}

int am1g_get_tunable(const am1g_ctx* ctx, const char* name, double* value) {
    if (!ctx || !name || !value) return am1g_set_error(nullptr, AM1G_INVALID_INPUT, "null argument");
    if (!get_tunable(ctx->tun, name, *value))
        return am1g_set_error(nullptr, AM1G_INVALID_INPUT, "unknown tunable %s", name);
    return AM1G_OK;
}

int am1g_scf_batch_noalloc(am1g_ctx* ctx, int n, const am1g_request* rqs, am1g_result* results) {
    return batch_impl(ctx, n, rqs, results, std::function<void(int)>(), true);
}

long long am1g_noalloc_capacity_bytes(void) { return (long long)kScratchBytes; }

}  // extern "C"

int am1g_scf_batch_stream(am1g_ctx* ctx, int n, const am1g_request* rqs, am1g_result* results,
                          const std::function<void(int)>& on_done) {
    return batch_impl(ctx, n, rqs, results, on_done, false);
}

// The batched launch. Its device pools come from the context's grow-only buffers, or with
// `noalloc` from the per-device scratch array (no allocation at all; sizes beyond it are refused).
static int batch_impl(am1g_ctx* ctx, int n, const am1g_request* rqs, am1g_result* results,
                      const std::function<void(int)>& on_done, bool noalloc) {
    const bool streaming = (bool)on_done;
    if (noalloc && streaming) return am1g_set_error(ctx, AM1G_INVALID_INPUT, "streaming needs the allocating entry point");
    if (!ctx) return am1g_set_error(nullptr, AM1G_INVALID_INPUT, "null context");
    if (n < 0 || (n > 0 && (!rqs || !results))) return am1g_set_error(ctx, AM1G_INVALID_INPUT, "bad batch arguments");
    std::lock_guard<std::mutex> lock(ctx->mu);
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    std::vector<Prepared> prep(n);
    std::vector<int> ok_idx;
    ok_idx.reserve(n);
    for (int i = 0; i < n; ++i) {
        std::string msg;
        int st = am1g_prepare(ctx, rqs[i], prep[i], msg, kMaxBatchAo);
        am1g_result& r = results[i];
        if (st != AM1G_OK) {
            r.status = st;
            r.iterations = 0;
            r.converged = 0;
            am1g_set_error(ctx, st, "%s", msg.c_str());
            continue;
        }
        ok_idx.push_back(i);
    }
    const int m = (int)ok_idx.size();
    if (streaming)  // refused requests are final at once
        for (int i = 0, j = 0; i < n; ++i) {
            if (j < m && ok_idx[j] == i) ++j;
            else on_done(i);
        }
    if (m == 0) return AM1G_OK;
    // pool layout
    std::vector<MolDesc> descs(m);
    std::vector<int> zs, alist, aoff, acls;
    std::vector<double> xyz;
    std::vector<int64_t> pair_off(1, 0);
    int64_t atom_off = 0, ao_off = 0, mat_off = 0, w_off = 0, hist_off = 0;
    int max_nao = 0;
    for (int j = 0; j < m; ++j) {
        const am1g_request& rq = rqs[ok_idx[j]];
        MolDesc md = prep[ok_idx[j]].md;
        const am1g_result& rj = results[ok_idx[j]];
        md.all_vecs = (rj.mo_coeff || rj.mo_coeff_beta) ? 1 : 0;
        md.atom_off = atom_off;
        md.ao_off = ao_off;
        md.mat_off = mat_off;
        md.w_off = w_off;
        md.hist_off = hist_off;
        {
            // heavy atoms first, then hydrogens, each in input order; AO offsets s,px,py,pz per
            // heavy atom and s per hydrogen in input order (basis.rs:37-60)
            const DevElem* tab = ctx->host_tab + md.method * kMaxZ;
            int off = 0, nh = 0, nH = 0;
            std::vector<int> heavy, hyd;
            for (int a = 0; a < rq.n_atoms; ++a) {
                int zz = rq.z[a];
                zs.push_back(zz);
                for (int k = 0; k < 3; ++k) xyz.push_back(rq.xyz_bohr[3 * a + k]);
                aoff.push_back(off);
                off += tab[zz].norb;
                if (zz != 1) {
                    acls.push_back(nh++);
                    heavy.push_back(a);
                } else {
                    acls.push_back(nH++);
                    hyd.push_back(a);
                }
            }
            alist.insert(alist.end(), heavy.begin(), heavy.end());
            alist.insert(alist.end(), hyd.begin(), hyd.end());
            pair_off.push_back(pair_off.back() + (int64_t)rq.n_atoms * (rq.n_atoms - 1) / 2);
        }
        atom_off += rq.n_atoms;
        ao_off += md.nao;
        mat_off += (int64_t)(md.uhf ? kMatSlotsUhf : kMatSlotsRhf) * md.nao * md.nao;
        w_off += w_pool_size(md.n_heavy, md.n_h);
        hist_off += hist_size(md.nao, md.uhf);
        max_nao = std::max(max_nao, md.nao);
        descs[j] = md;
    }
    // largest molecules first so the tail of the persistent launch is short
    std::vector<int> order(m);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return descs[a].nao > descs[b].nao; });

    // device pools: (buffer, bytes) in one fixed order
    struct Pool {
        DevBuf* buf;
        size_t bytes;
        void* p;
    };
    Pool pools[] = {
        {&ctx->mols, sizeof(MolDesc) * m, nullptr},
        {&ctx->z, sizeof(int) * zs.size(), nullptr},
        {&ctx->xyz, sizeof(double) * xyz.size(), nullptr},
        {&ctx->mats, sizeof(double) * mat_off, nullptr},
        {&ctx->wpool, sizeof(double) * std::max<int64_t>(w_off, 1), nullptr},
        {&ctx->hist, sizeof(double) * hist_off, nullptr},
        {&ctx->mo_e, sizeof(double) * ao_off, nullptr},
        {&ctx->mo_eb, sizeof(double) * ao_off, nullptr},
        {&ctx->charges, sizeof(double) * atom_off, nullptr},
        {&ctx->results, sizeof(MolResult) * m, nullptr},
        {&ctx->queue, sizeof(int) * 4, nullptr},
        {&ctx->order, sizeof(int) * m, nullptr},
        {&ctx->alist, sizeof(int) * alist.size(), nullptr},
        {&ctx->aoff, sizeof(int) * aoff.size(), nullptr},
        {&ctx->acls, sizeof(int) * acls.size(), nullptr},
        {&ctx->pair_off, sizeof(int64_t) * pair_off.size(), nullptr},
        {&ctx->ecore_pair, sizeof(double) * std::max<int64_t>(pair_off.back(), 1), nullptr},
    };
    constexpr int kPools = (int)(sizeof(pools) / sizeof(pools[0]));
    enum { pMols, pZ, pXyz, pMats, pW, pHist, pMo, pMob, pQ, pRes, pQueue, pOrder, pAlist, pAoff, pAcls, pPairOff,
           pEcore };
    if (noalloc) {
        char* base = static_cast<char*>(scratch_ptr(ctx->device));
        if (!base)
            return am1g_set_error(ctx, AM1G_DEVICE_ERROR, "device scratch of device %d is not set up", ctx->device);
        size_t used = 0;
        for (int k = 0; k < kPools; ++k) used = ((used + 255) & ~(size_t)255) + pools[k].bytes;
        if (used > kScratchBytes)
            return am1g_set_error(ctx, AM1G_INVALID_INPUT,
                                  "batch needs %.1f MiB of device memory; the allocation-free entry point has %.1f MiB",
                                  used / 1048576.0, kScratchBytes / 1048576.0);
        used = 0;
        for (int k = 0; k < kPools; ++k) {
            used = (used + 255) & ~(size_t)255;
            pools[k].p = base + used;
            used += pools[k].bytes;
        }
    } else {
        for (int k = 0; k < kPools; ++k) {
            HIP_TRY(ctx, pools[k].buf->reserve(pools[k].bytes));
            pools[k].p = pools[k].buf->p;
        }
    }
    hipStream_t st = ctx->stream;
    HIP_TRY(ctx, hipMemcpyAsync(pools[pMols].p, descs.data(), sizeof(MolDesc) * m, hipMemcpyHostToDevice, st));
    HIP_TRY(ctx, hipMemcpyAsync(pools[pZ].p, zs.data(), sizeof(int) * zs.size(), hipMemcpyHostToDevice, st));
    HIP_TRY(ctx, hipMemcpyAsync(pools[pXyz].p, xyz.data(), sizeof(double) * xyz.size(), hipMemcpyHostToDevice, st));
    HIP_TRY(ctx, hipMemcpyAsync(pools[pOrder].p, order.data(), sizeof(int) * m, hipMemcpyHostToDevice, st));
    HIP_TRY(ctx, hipMemcpyAsync(pools[pAlist].p, alist.data(), sizeof(int) * alist.size(), hipMemcpyHostToDevice, st));
    HIP_TRY(ctx, hipMemcpyAsync(pools[pAoff].p, aoff.data(), sizeof(int) * aoff.size(), hipMemcpyHostToDevice, st));
    HIP_TRY(ctx, hipMemcpyAsync(pools[pAcls].p, acls.data(), sizeof(int) * acls.size(), hipMemcpyHostToDevice, st));
    HIP_TRY(ctx, hipMemcpyAsync(pools[pPairOff].p, pair_off.data(), sizeof(int64_t) * pair_off.size(),
                                hipMemcpyHostToDevice, st));
    HIP_TRY(ctx, launch_zero_ints(static_cast<int*>(pools[pQueue].p), 4, st));

    BatchArgs a;
    std::memset(&a, 0, sizeof(a));
    a.mols = static_cast<MolDesc*>(pools[pMols].p);
    a.n_mols = m;
    a.lds_n = ctx->tun.jacobi_lds_n;
    a.tri_max = std::min(eigen_tri_max(), ctx->tun.tri_max);
    a.z = static_cast<int*>(pools[pZ].p);
    a.xyz = static_cast<double*>(pools[pXyz].p);
    a.elems = ctx->elems.as<DevElem>();
    a.mats = static_cast<double*>(pools[pMats].p);
    a.wpool = static_cast<double*>(pools[pW].p);
    a.hist = static_cast<double*>(pools[pHist].p);
    a.mo_e = static_cast<double*>(pools[pMo].p);
    a.mo_e_beta = static_cast<double*>(pools[pMob].p);
    a.charges = static_cast<double*>(pools[pQ].p);
    a.results = static_cast<MolResult*>(pools[pRes].p);
    a.queue = static_cast<int*>(pools[pQueue].p);
    a.order = static_cast<int*>(pools[pOrder].p);
    a.alist = static_cast<int*>(pools[pAlist].p);
    a.aoff = static_cast<int*>(pools[pAoff].p);
    a.acls = static_cast<int*>(pools[pAcls].p);
    a.pair_off = static_cast<int64_t*>(pools[pPairOff].p);
    a.n_pairs = pair_off.back();
    a.ecore_pair = static_cast<double*>(pools[pEcore].p);
    size_t lds = (size_t)kLdsBytes;
    int precision = 0;
    for (int j = 0; j < m; ++j) precision |= descs[j].precision;
    int grid = std::min(m, ctx->cus);
    if (streaming) {
        HIP_TRY(ctx, ctx->h_done.reserve(sizeof(int) * m));
        HIP_TRY(ctx, ctx->h_results.reserve(sizeof(MolResult) * m));
        HIP_TRY(ctx, ctx->h_charges.reserve(sizeof(double) * atom_off));
        HIP_TRY(ctx, ctx->h_mo.reserve(sizeof(double) * ao_off));
        HIP_TRY(ctx, ctx->h_mob.reserve(sizeof(double) * ao_off));
        std::memset(ctx->h_done.p, 0, sizeof(int) * m);
        a.done = ctx->h_done.as<int>();
        a.results_h = ctx->h_results.as<MolResult>();
        a.charges_h = ctx->h_charges.as<double>();
        a.mo_e_h = ctx->h_mo.as<double>();
        a.mo_eb_h = ctx->h_mob.as<double>();
    }
    HIP_TRY(ctx, hipEventRecord(ctx->ev0, st));
    a.own_pairs = ctx->tun.batch_own_pairs;
    if (!a.own_pairs) HIP_TRY(ctx, launch_pairs(a, st, ctx->tun.pairs_block));
    // Molecules of at most kSmallWgAo basis functions (the tail of the largest-first order) run
    // two per CU in the 256-thread build when there are enough of them to fill the device twice;
    // the rest (all of a smaller batch) in the 512-thread one. The split depends on the batch, and
    // the two builds sum in different orders (results within rounding of each other), so it is
    // made only for batches that large: a molecule alone, in a small batch or in the server gets
    // the 512-thread build's bits.
    int n_small = 0;
    if (ctx->tun.batch_small_wg)
        for (int j = 0; j < m; ++j) n_small += descs[j].nao <= kSmallWgAo ? 1 : 0;
    if (n_small >= 2 * ctx->cus) {
        // The small-molecule launch runs on a second stream after everything queued so far
        // (ev0): its workgroups take each CU that a workgroup of the larger molecules' launch
        // leaves, so the tail of that launch fills with small molecules.
        const int n_big = m - n_small;
        if (n_big > 0) {
            BatchArgs big = a;
            big.n_mols = n_big;
            HIP_TRY(ctx, launch_scf_batch(big, precision, std::min(n_big, ctx->cus), lds, st));
        }
        BatchArgs small = a;
        small.order = a.order + n_big;
        small.n_mols = n_small;
        small.queue = a.queue + 1;
        HIP_TRY(ctx, hipStreamWaitEvent(ctx->stream2, ctx->ev0, 0));
        HIP_TRY(ctx, launch_scf_batch_small(small, precision, std::min(n_small, 2 * ctx->cus), ctx->stream2));
        HIP_TRY(ctx, hipEventRecord(ctx->ev2, ctx->stream2));
        HIP_TRY(ctx, hipStreamWaitEvent(st, ctx->ev2, 0));
    } else {
        HIP_TRY(ctx, launch_scf_batch(a, precision, grid, lds, st));
    }
    HIP_TRY(ctx, hipEventRecord(ctx->ev1, st));
    if (streaming) {
        // hand out each molecule as its done flag appears, while the launch runs
        std::vector<char> reported(m, 0);
        int left = m;
        const int* done = ctx->h_done.as<int>();
        const MolResult* hr = ctx->h_results.as<MolResult>();
        auto publish = [&](int j) {
            const MolDesc& md = descs[j];
            am1g_result& r = results[ok_idx[j]];
            am1g_fill_result(hr[j], md, r);
            r.latency_ms = (double)(hr[j].t_done - hr[j].t_start) / 1.0e5;
            for (int k = 0; k < kPhases; ++k) r.phase_ms[k] = (double)hr[j].phase[k] / 1.0e5;
            r.eigen_sweeps = hr[j].eigen_sweeps;
            if (r.charges) std::memcpy(r.charges, ctx->h_charges.as<double>() + md.atom_off, sizeof(double) * md.nat);
            am1g_apply_bcc(rqs[ok_idx[j]], ctx->h_charges.as<double>() + md.atom_off, r);
            if (r.mo_energies) std::memcpy(r.mo_energies, ctx->h_mo.as<double>() + md.ao_off, sizeof(double) * md.nao);
            if (r.mo_energies_beta && md.uhf)
                std::memcpy(r.mo_energies_beta, ctx->h_mob.as<double>() + md.ao_off, sizeof(double) * md.nao);
            on_done(ok_idx[j]);
        };
        for (;;) {
            const hipError_t q = hipStreamQuery(st);
            if (q != hipSuccess && q != hipErrorNotReady)
                return am1g_set_error(ctx, AM1G_DEVICE_ERROR, "HIP error %s in the batched launch", hipGetErrorString(q));
            for (int j = 0; j < m; ++j)
                if (!reported[j] && __atomic_load_n(done + j, __ATOMIC_ACQUIRE)) {
                    reported[j] = 1;
                    --left;
                    publish(j);
                }
            if (left == 0 || q == hipSuccess) break;
            std::this_thread::sleep_for(std::chrono::microseconds(50));
        }
        HIP_TRY(ctx, hipStreamSynchronize(st));
        for (int j = 0; j < m; ++j)  // the launch has ended: every flag is set
            if (!reported[j]) publish(j);
        float ms = 0.f;
        HIP_TRY(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
        ctx->last_ms = ms;
        return AM1G_OK;
    }

    std::vector<MolResult> res(m);
    std::vector<double> mo(ao_off), mob(ao_off), q(atom_off);
    HIP_TRY(ctx, hipMemcpyAsync(res.data(), pools[pRes].p, sizeof(MolResult) * m, hipMemcpyDeviceToHost, st));
    HIP_TRY(ctx, hipMemcpyAsync(mo.data(), pools[pMo].p, sizeof(double) * ao_off, hipMemcpyDeviceToHost, st));
    HIP_TRY(ctx, hipMemcpyAsync(q.data(), pools[pQ].p, sizeof(double) * atom_off, hipMemcpyDeviceToHost, st));
    HIP_TRY(ctx, hipMemcpyAsync(mob.data(), pools[pMob].p, sizeof(double) * ao_off, hipMemcpyDeviceToHost, st));
    HIP_TRY(ctx, hipStreamSynchronize(st));
    float ms = 0.f;
    HIP_TRY(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    ctx->last_ms = ms;
    uint64_t t0 = UINT64_MAX;
    for (int j = 0; j < m; ++j) t0 = std::min(t0, res[j].t_start);
    std::vector<double> mat;
    for (int j = 0; j < m; ++j) {
        const int i = ok_idx[j];
        const MolDesc& md = descs[j];
        am1g_result& r = results[i];
        am1g_fill_result(res[j], md, r);
        r.latency_ms = (double)(res[j].t_done - t0) / 1.0e5;  // 100 MHz device clock
        for (int k = 0; k < kPhases; ++k) r.phase_ms[k] = (double)res[j].phase[k] / 1.0e5;
        r.eigen_sweeps = res[j].eigen_sweeps;
        if (r.charges) std::memcpy(r.charges, q.data() + md.atom_off, sizeof(double) * md.nat);
        am1g_apply_bcc(rqs[i], q.data() + md.atom_off, r);
        if (r.mo_energies) std::memcpy(r.mo_energies, mo.data() + md.ao_off, sizeof(double) * md.nao);
        if (r.mo_energies_beta && md.uhf)
            std::memcpy(r.mo_energies_beta, mob.data() + md.ao_off, sizeof(double) * md.nao);
        const size_t nn = (size_t)md.nao * md.nao;
        if (r.density) {
            const int slot = md.uhf ? kPt : kP;
            HIP_TRY(ctx, hipMemcpy(r.density, static_cast<double*>(pools[pMats].p) + md.mat_off + slot * nn, nn * sizeof(double),
                                   hipMemcpyDeviceToHost));
        }
        const double* M = static_cast<double*>(pools[pMats].p) + md.mat_off;
        // eigenvectors are stored transposed (row k = MO k); the result holds MOs as columns
        auto columns = [&](int slot, double* out) -> hipError_t {
            mat.resize(nn);
            hipError_t e = hipMemcpy(mat.data(), M + slot * nn, nn * sizeof(double), hipMemcpyDeviceToHost);
            if (e != hipSuccess) return e;
            for (int mu = 0; mu < md.nao; ++mu)
                for (int k = 0; k < md.nao; ++k) out[(size_t)mu * md.nao + k] = mat[(size_t)k * md.nao + mu];
            return hipSuccess;
        };
        if (r.mo_coeff) HIP_TRY(ctx, columns(kC, r.mo_coeff));
        if (md.uhf && r.mo_coeff_beta) HIP_TRY(ctx, columns(kCb, r.mo_coeff_beta));
        if (md.uhf && r.spin_density) {
            std::vector<double> pb(nn);
            HIP_TRY(ctx, hipMemcpy(r.spin_density, M + kP * nn, nn * sizeof(double), hipMemcpyDeviceToHost));
            HIP_TRY(ctx, hipMemcpy(pb.data(), M + kPb * nn, nn * sizeof(double), hipMemcpyDeviceToHost));
            for (size_t k = 0; k < nn; ++k) r.spin_density[k] -= pb[k];
        }
    }
    return AM1G_OK;
}

extern "C" {

int am1g_scf(am1g_ctx* ctx, const am1g_request* rq, am1g_result* res) {
    if (!ctx || !rq || !res) return am1g_set_error(ctx, AM1G_INVALID_INPUT, "null argument");
    auto w0 = std::chrono::steady_clock::now();
    // Molecules above kRouteLargeAo basis functions (or beyond one workgroup's atom tables) take
    // the large-system path; AM1G_FORCE_LARGE_PATH=1 sends every molecule there (for testing it
    // on sizes the oracle finishes quickly).
    Prepared prep;
    std::string msg;
    int st = am1g_prepare(ctx, *rq, prep, msg, INT_MAX);
    if (st != AM1G_OK) {
        res->status = st;
        res->iterations = 0;
        res->converged = 0;
        return am1g_set_error(ctx, st, "%s", msg.c_str());
    }
    const char* force = std::getenv("AM1G_FORCE_LARGE_PATH");
    if (prep.md.nao > ctx->tun.route_large_ao || rq->n_atoms > kMaxAtoms || (force && force[0] == '1')) {
        std::lock_guard<std::mutex> lock(ctx->mu);
        HIP_TRY(ctx, hipSetDevice(ctx->device));
        st = am1g_large_scf(ctx, *rq, prep, res);
        am1g_apply_bcc(*rq, res->charges, *res);
        res->latency_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - w0).count();
        ctx->last_ms = res->latency_ms;
        return st;
    }
    st = am1g_scf_batch(ctx, 1, rq, res);
    res->latency_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - w0).count();
    if (st != AM1G_OK) return st;
    return res->status;
}

}  // extern "C"
