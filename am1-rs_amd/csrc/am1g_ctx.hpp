// Internal host-side context of the am1g library (not part of the C ABI).
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <functional>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/am1g.h"
#include "am1g_common.hpp"
#include "am1g_eig.hpp"
#include "am1g_launch.hpp"

using namespace am1g;  // internal header of the library

// Grow-only pinned host buffer (device-visible), for results streamed out of a running launch.
struct HostBuf {
    void* p = nullptr;
    size_t bytes = 0;
    hipError_t reserve(size_t need) {
        if (need <= bytes) return hipSuccess;
        if (p) (void)hipHostFree(p);
        p = nullptr;
        bytes = 0;
        size_t want = std::max(need, (size_t)4096);
        hipError_t e = hipHostMalloc(&p, want, hipHostMallocDefault);
        if (e == hipSuccess) bytes = want;
        return e;
    }
    template <typename T>
    T* as() const {
        return static_cast<T*>(p);
    }
    ~HostBuf() {
        if (p) (void)hipHostFree(p);
    }
};

// Grow-only device buffer.
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    hipError_t reserve(size_t need) {
        if (need <= bytes) return hipSuccess;
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
        size_t want = std::max(need, (size_t)4096);
        hipError_t e = hipMalloc(&p, want);
        if (e == hipSuccess) bytes = want;
        return e;
    }
    template <typename T>
    T* as() const {
        return static_cast<T*>(p);
    }
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
};

// Launch choices that depend on the device rather than on the inputs; defaults here, replaced at
// context creation by the table the autotuner (am1g.tune) measured, when one is found.
struct Tunables {
    int route_large_ao = kRouteLargeAo;  // single molecules above this many AOs take the grid path
    int jacobi_lds_n = 192;              // batched Jacobi (129-256 AOs) keeps orders up to this in LDS
    int eig_graphs = 1;                  // replay the large eigensolver's tridiagonalization as a graph
    int serve_large_batch = 2;           // a server lane batches queued grid-path requests from this many
    int tri_max = 256;                   // batched kernel: tridiagonal method up to this order, Jacobi above
    int pairs_block = 256;               // threads per workgroup of the pair-integral kernel
    int tri_blocked_min = 1 << 30;       // large eigensolver: tridiagonalization in panels from this order
    int batch_own_pairs = 1;             // batched launch: workgroups compute their own pair integrals
    int sp2 = 1;                         // large-system SCF: densities by SP2 purification (0: eigenvectors)
    int batch_small_wg = 1;              // large batches: molecules <= 64 AOs two per CU (256-thread build)
};

struct am1g_ctx {
    int device = 0;
    Tunables tun;
    std::string tune_source;  // the table the tunables came from (empty: defaults)
    int cus = 256;
    hipStream_t stream = nullptr;
    hipStream_t stream2 = nullptr;  // the small-molecule launch of a split batch, beside `stream`
    hipEvent_t ev0 = nullptr, ev1 = nullptr, ev2 = nullptr;
    double last_ms = 0.0;
    std::mutex mu;
    DevElem host_tab[kMethods * kMaxZ];
    DevBuf elems, mols, z, xyz, mats, wpool, hist, mo_e, mo_eb, charges, results, queue, order;
    DevBuf alist, aoff, acls, later, pair_off, ecore_pair;
    HostBuf h_done, h_results, h_charges, h_mo, h_mob;  // streaming outputs of the batched launch
    struct DcBuffers* dc = nullptr;
    struct LargeBuffers* large = nullptr;  // one large system: dense matrices, DIIS history
    EigWork* eig = nullptr;                // large-order eigensolver workspace
    DevBuf grad_buf;                       // nuclear gradient: tables, densities, pair terms
};

// Buffers of the divide-and-conquer driver, kept on the context so repeated calls reuse them.
struct DcChannel {  // one spin channel (the only one for a closed shell)
    DevBuf F, P, cbuf, eps, proj, frac, pout, hin, hres, ring;
};
struct DcBuffers {
    DcChannel ch[2];
    DevBuf far;  // far-field terms (SysArgs::far)
    DevBuf Pt, pdiag, tcol, rsum, fbuf, t0, t1, status, sub_n, sub_ao, sub_mat, ao_list, ao_core, pat, mult, coef, ao_atom,
        core_sub, partial, scal;
};

struct Prepared {
    MolDesc md;
    int nat;
    bool uhf;
};


int am1g_set_error(am1g_ctx* ctx, int code, const char* fmt, ...);
// Reads the tunables table: the file AM1G_TUNE_FILE names (empty: none), else tuned.txt beside the
// library; lines "name value", '#' starts a comment. Unknown names are ignored.
void am1g_load_tunables(am1g_ctx* ctx);
int am1g_prepare(am1g_ctx* ctx, const am1g_request& rq, Prepared& out, std::string& msg, int max_ao);
void am1g_dc_release(am1g_ctx* ctx);
// Batched launch that hands each molecule's result to on_done(request index) as soon as the
// molecule finishes on the device (charges and orbital energies filled), before the launch ends.
// Density and orbital matrices are not streamed.
int am1g_scf_batch_stream(am1g_ctx* ctx, int n, const am1g_request* rqs, am1g_result* results,
                          const std::function<void(int)>& on_done);
// Scalars of a device-side MolResult into the C-ABI result.
void am1g_fill_result(const MolResult& r, const MolDesc& md, am1g_result& out);
// AM1-BCC charges into r.bcc_charges from the given Mulliken charges when the request asks for
// them and the SCF produced a result.
void am1g_apply_bcc(const am1g_request& rq, const double* mulliken, am1g_result& r);
void am1g_large_release(am1g_ctx* ctx);
// Creates the context's rocBLAS handle and eigensolver workspace on first use.
int am1g_large_setup(am1g_ctx* ctx);
// Restricted SCF of one molecule on the whole device (no size limit beyond device memory).
int am1g_large_scf(am1g_ctx* ctx, const am1g_request& rq, const Prepared& prep, am1g_result* res);

#define HIP_TRY(ctx, expr)                                                                                    \
    do {                                                                                                      \
        hipError_t e_ = (expr);                                                                               \
        if (e_ != hipSuccess)                                                                                 \
            return am1g_set_error(ctx, AM1G_DEVICE_ERROR, "HIP error %s at %s", hipGetErrorString(e_), #expr); \
    } while (0)// Per class-list entry, where the other class's partners later in AO order begin: for heavy atom
// j (entry j), the first hydrogen-list index whose AO offset exceeds j's; for hydrogen m (entry
// n_heavy + m), the first heavy-list index whose AO offset exceeds m's. Both lists keep the input
// order, so AO offsets grow along each.
inline std::vector<int> later_partners(const std::vector<int>& alist, const std::vector<int>& aoff, int nh) {
    const int nat = (int)alist.size(), nH = nat - nh;
    std::vector<int> out(nat);
    for (int j = 0, m = 0; j < nh; ++j) {
        while (m < nH && aoff[alist[nh + m]] < aoff[alist[j]]) ++m;
        out[j] = m;
    }
    for (int m = 0, j = 0; m < nH; ++m) {
        while (j < nh && aoff[alist[j]] < aoff[alist[nh + m]]) ++j;
        out[nh + m] = j;
    }
    return out;
}


