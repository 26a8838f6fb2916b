// Dense symmetric eigensolver for large orders (one matrix on the whole device): Householder
// tridiagonalization, tridiagonal divide-and-conquer, blocked back-transformation. Internal
// header of the library.
#pragma once
#include <hip/hip_runtime.h>

#include <string>
#include <vector>

struct EigBuf {
    void* p = nullptr;
    size_t bytes = 0;
    hipError_t reserve(size_t need) {
        if (need <= bytes) return hipSuccess;
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
        hipError_t e = hipMalloc(&p, need < 4096 ? 4096 : need);
        if (e == hipSuccess) bytes = need < 4096 ? 4096 : need;
        return e;
    }
    template <typename T>
    T* as() const {
        return static_cast<T*>(p);
    }
    ~EigBuf() {
        if (p) (void)hipFree(p);
    }
};

// A captured launch sequence of the tridiagonalization (2 n kernels) for one order and one set of
// buffers, replayed as one graph launch.
struct TridiagGraph {
    int n = 0;
    const void* key[16] = {};  // every device buffer the captured launches read or write
    hipGraphExec_t exec = nullptr;
};

// Device and host workspace, grown on demand and kept between calls.
struct EigWork {
    EigBuf d, e, tau, vbuf, pbuf, scal;       // tridiagonal and Householder vectors
    EigBuf Q;                                 // n x n column-major: eigenvectors of T, then of A
    EigBuf W, Qt, U, Qn;                      // merge scratch
    EigBuf V, Tm, X, Y, G;                    // back-transformation blocks
    EigBuf rot, idx, flt, lam;                // merge tables
    EigBuf gw;                                // split-K partial products of the GEMMs
    EigBuf bV, bW, bybuf, bvpart, brow, brpart, bscal;  // blocked tridiagonalization
    std::vector<double> h_d, h_e, h_lam;
    bool use_graphs = true;                   // replay the tridiagonalization as a hipGraph
    int blocked_min = 1 << 30;                // orders from which the tridiagonalization runs in panels
    std::vector<TridiagGraph> graphs;         // at most kMaxGraphs, oldest dropped first
    static constexpr size_t kMaxGraphs = 4;
    ~EigWork() {
        for (auto& g : graphs)
            if (g.exec) (void)hipGraphExecDestroy(g.exec);
    }
};

namespace am1g {

// Eigen-decomposition of the symmetric n x n matrix A (row-major; overwritten). On return
// evals[0..n) holds the eigenvalues in ascending order and Ct (row-major n x n) holds the
// eigenvectors as rows (row k = eigenvector of evals[k]). Returns hipSuccess or the first
// device error; `err` names a numerical failure (then the return is hipErrorUnknown).
hipError_t eig_large(EigWork& w, int n, double* A, double* evals, double* Ct, hipStream_t st,
                     std::string& err);

// Only the eigenvalues of ascending indices k_lo..k_hi of the symmetric n x n A (row-major;
// overwritten), into vals[0..k_hi - k_lo] on the host: the same tridiagonalization, then bisection
// on Sturm counts of the tridiagonal to rounding level instead of the divide and conquer and the
// back-transformation.
hipError_t eig_large_values(EigWork& w, int n, double* A, int k_lo, int k_hi, double* vals, hipStream_t st,
                            std::string& err);
hipError_t eig_large_impl(EigWork& w, int n, double* A, double* evals, double* Ct, hipStream_t st,
                          std::string& err, int k_lo, int k_hi, double* vals);

}  // namespace am1g
