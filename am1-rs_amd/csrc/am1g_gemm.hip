// Grid-level matrix-core GEMM (FP64 and FP32) for the large-system SCF and the large eigensolver:
// the commutator F P, the density C_occ C_occ^T, the eigensolver's merge products and its
// blocked back-transformation. Column-major operands, C = alpha op(A) op(B) + beta C.
//
// Each 256-thread workgroup owns a 64 x 64 tile of C; its four wavefronts own 32 x 32 quarters,
// four 16 x 16 MFMA accumulators each. K advances in panels of 16: both operand panels are
// staged through LDS (double-buffered, one barrier per panel) while the next panel's loads are
// in flight in registers. The MFMA computes each tile transposed (D = op(B)^T op(A)^T), so that
// the sixteen lanes sharing a result register hold sixteen consecutive rows of column-major C and
// the stores coalesce.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "am1g_gemm.hpp"

namespace am1g {
namespace {

constexpr int kBM = 64, kBN = 64, kBK = 16, kThreads = 256;
constexpr int kPad = 4;  // LDS row padding (elements), shifts the banks of consecutive K rows

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

// 16x16x4 MFMA: operand a(i, k) in lane i + 16k, operand b(k, j) in lane j + 16k; result lane l,
// register r holds row `row(l, r)`, column l % 16 of the 16 x 16 product.
template <typename T>
struct Mfma;
template <>
struct Mfma<double> {
    typedef v4d V;
    static __device__ __forceinline__ V mma(double a, double b, V c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int l, int r) { return (l >> 4) + 4 * r; }
};
template <>
struct Mfma<float> {
    typedef v4f V;
    static __device__ __forceinline__ V mma(float a, float b, V c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int l, int r) { return 4 * (l >> 4) + r; }
};

// Tile grid over C plus, in blockIdx.y, the split of K. `part` (split-K only) receives the raw
// partial product of each split, [split][n][m] column-major with leading dimension m. sym: the
// product is known to be symmetric (m == n); only the tiles on and above the diagonal are formed,
// and each off-diagonal one is also stored transposed below it.
template <typename T, bool TA, bool TB>
__global__ __launch_bounds__(kThreads, 2) void k_gemm(int m, int n, int k, int k_chunk, T alpha,
                                                      const T* __restrict__ A, int lda, const T* __restrict__ B,
                                                      int ldb, T beta, T* __restrict__ C, int ldc,
                                                      T* __restrict__ part, int tiles_m, int tiles_n, int sym) {
    typedef typename Mfma<T>::V V;
    __shared__ T As[2][kBK][kBM + kPad];  // op(A) panel, [k][i]
    __shared__ T Bs[2][kBK][kBN + kPad];  // op(B) panel, [k][j]
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    // Consecutive workgroup ids go to the eight XCDs in turn; when the tile count allows, each
    // XCD takes a contiguous run of tiles (one column block of C: the same B panel in its L2).
    const int ntiles = sym ? tiles_n * (tiles_n + 1) / 2 : tiles_m * tiles_n;
    int b = blockIdx.x;
    if ((ntiles & 7) == 0) b = (b & 7) * (ntiles >> 3) + (b >> 3);
    int ti = b % tiles_m, tj = b / tiles_m;
    if (sym) {
        // upper-triangle tile b -> (ti, tj), ti <= tj, column by column: column tj starts at
        // tj (tj + 1) / 2 (a single-precision estimate, corrected in integers)
        tj = (int)((sqrtf(8.0f * (float)b + 1.0f) - 1.0f) * 0.5f);
        while (tj * (tj + 1) / 2 > b) --tj;
        while ((tj + 1) * (tj + 2) / 2 <= b) ++tj;
        ti = b - tj * (tj + 1) / 2;
    }
    const int i0 = ti * kBM, j0 = tj * kBN;
    const bool mirror = sym && ti != tj;
    const int kb = blockIdx.y * k_chunk, ke = min(k, kb + k_chunk);

    T ra[4], rb[4];
    auto load = [&](int k0) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int idx = t + kThreads * u;
            int i, ka, j, kbq;
            if (TA) {
                ka = idx & 15;
                i = idx >> 4;
            } else {
                i = idx & 63;
                ka = idx >> 6;
            }
            if (TB) {
                j = idx & 63;
                kbq = idx >> 6;
            } else {
                kbq = idx & 15;
                j = idx >> 4;
            }
            const int gi = i0 + i, gka = k0 + ka, gj = j0 + j, gkb = k0 + kbq;
            ra[u] = (gi < m && gka < ke) ? (TA ? A[(size_t)gi * lda + gka] : A[(size_t)gka * lda + gi]) : T(0);
            rb[u] = (gj < n && gkb < ke) ? (TB ? B[(size_t)gkb * ldb + gj] : B[(size_t)gj * ldb + gkb]) : T(0);
        }
    };
    auto store = [&](int buf) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int idx = t + kThreads * u;
            if (TA) As[buf][idx & 15][idx >> 4] = ra[u];
            else As[buf][idx >> 6][idx & 63] = ra[u];
            if (TB) Bs[buf][idx >> 6][idx & 63] = rb[u];
            else Bs[buf][idx & 15][idx >> 4] = rb[u];
        }
    };

    V acc[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int c = 0; c < 2; ++c) acc[a][c] = V{0, 0, 0, 0};
    const int wi = (w & 1) * 32, wj = (w >> 1) * 32;
    if (kb < ke) {
        load(kb);
        store(0);
    }
    __syncthreads();
    int buf = 0;
    for (int k0 = kb; k0 < ke; k0 += kBK) {
        const bool more = k0 + kBK < ke;
        if (more) load(k0 + kBK);
#pragma unroll
        for (int kk = 0; kk < kBK; kk += 4) {
            const int kq = kk + (lane >> 4);
            T xb[2], ya[2];
#pragma unroll
            for (int r = 0; r < 2; ++r) {
                xb[r] = Bs[buf][kq][wj + 16 * r + (lane & 15)];
                ya[r] = As[buf][kq][wi + 16 * r + (lane & 15)];
            }
#pragma unroll
            for (int ri = 0; ri < 2; ++ri)
#pragma unroll
                for (int rj = 0; rj < 2; ++rj) acc[ri][rj] = Mfma<T>::mma(xb[rj], ya[ri], acc[ri][rj]);
        }
        if (more) store(buf ^ 1);
        __syncthreads();
        buf ^= 1;
    }
    // acc[ri][rj] holds the transposed 16 x 16 block: lane l, register r -> C row l % 16, column row(l, r)
#pragma unroll
    for (int ri = 0; ri < 2; ++ri)
#pragma unroll
        for (int rj = 0; rj < 2; ++rj)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int gi = i0 + wi + 16 * ri + (lane & 15);
                const int gj = j0 + wj + 16 * rj + Mfma<T>::row(lane, r);
                if (gi >= m || gj >= n) continue;
                const T v = acc[ri][rj][r];
                if (part) {
                    part[(size_t)blockIdx.y * m * n + (size_t)gj * m + gi] = v;
                } else {
                    T* c = C + (size_t)gj * ldc + gi;
                    *c = beta != T(0) ? alpha * v + beta * *c : alpha * v;
                    if (mirror) {
                        T* c2 = C + (size_t)gi * ldc + gj;
                        *c2 = beta != T(0) ? alpha * v + beta * *c2 : alpha * v;
                    }
                }
            }
}

// Split-K epilogue: C = alpha * sum over splits (in split order) + beta C.
template <typename T>
__global__ __launch_bounds__(256) void k_splitk_sum(int m, int n, int splits, T alpha, const T* __restrict__ part,
                                                    T beta, T* __restrict__ C, int ldc) {
    const size_t mn = (size_t)m * n;
    for (size_t idx = (size_t)blockIdx.x * 256 + threadIdx.x; idx < mn; idx += (size_t)gridDim.x * 256) {
        T s = part[idx];
        for (int q = 1; q < splits; ++q) s += part[(size_t)q * mn + idx];
        const int gi = (int)(idx % m), gj = (int)(idx / m);
        T* c = C + (size_t)gj * ldc + gi;
        *c = beta != T(0) ? alpha * s + beta * *c : alpha * s;
    }
}

// K split for an m x n x k product: enough workgroups for the device (about 512 tiles in
// flight), each split at least 256 deep.
void plan_split(int m, int n, int k, int& split, int& chunk) {
    const long tiles = (long)((m + kBM - 1) / kBM) * ((n + kBN - 1) / kBN);
    split = 1;
    if (tiles < 256 && k >= 512) split = (int)std::min<long>(std::max<long>(1, 512 / tiles), std::min(16, k / 256));
    chunk = ((k + split - 1) / split + kBK - 1) / kBK * kBK;
    split = chunk > 0 ? (k + chunk - 1) / chunk : 1;
    if (split < 1) split = 1;
}

template <typename T>
hipError_t gemm_t(hipStream_t st, bool ta, bool tb, int m, int n, int k, T alpha, const T* A, int lda, const T* B,
                  int ldb, T beta, T* C, int ldc, T* work, size_t work_elems, bool sym = false) {
    if (m <= 0 || n <= 0) return hipSuccess;
    if (sym && m != n) return hipErrorInvalidValue;
    const int tiles_m = (m + kBM - 1) / kBM, tiles_n = (n + kBN - 1) / kBN;
    int split = 1, chunk = k;
    if (k > 0 && !sym) plan_split(m, n, k, split, chunk);
    if (split > 1 && (!work || work_elems < (size_t)split * m * n)) {
        split = 1;
        chunk = k;
    }
    if (k <= 0) chunk = 0;
    T* part = split > 1 ? work : nullptr;
    const int ntiles = sym ? tiles_n * (tiles_n + 1) / 2 : tiles_m * tiles_n;
    const dim3 grid((unsigned)ntiles, (unsigned)split);
#define AM1G_GEMM_LAUNCH(TA_, TB_)                                                                                \
    hipLaunchKernelGGL((k_gemm<T, TA_, TB_>), grid, dim3(kThreads), 0, st, m, n, k, chunk, alpha, A, lda, B, ldb, \
                       beta, C, ldc, part, tiles_m, tiles_n, sym ? 1 : 0)
    if (ta && tb) AM1G_GEMM_LAUNCH(true, true);
    else if (ta) AM1G_GEMM_LAUNCH(true, false);
    else if (tb) AM1G_GEMM_LAUNCH(false, true);
    else AM1G_GEMM_LAUNCH(false, false);
#undef AM1G_GEMM_LAUNCH
    hipError_t e = hipGetLastError();
    if (e != hipSuccess || split == 1) return e;
    const size_t mn = (size_t)m * n;
    const unsigned blocks = (unsigned)std::min<size_t>((mn + 255) / 256, 4096);
    hipLaunchKernelGGL(k_splitk_sum<T>, dim3(blocks), dim3(256), 0, st, m, n, split, alpha, part, beta, C, ldc);
    return hipGetLastError();
}

}  // namespace

size_t gemm_splitk_elems(int m, int n, int k) {
    if (m <= 0 || n <= 0 || k <= 0) return 0;
    int split, chunk;
    plan_split(m, n, k, split, chunk);
    return split > 1 ? (size_t)split * m * n : 0;
}

hipError_t gemm_f64(hipStream_t st, bool trans_a, bool trans_b, int m, int n, int k, double alpha, const double* A,
                    int lda, const double* B, int ldb, double beta, double* C, int ldc, double* work,
                    size_t work_elems) {
    return gemm_t<double>(st, trans_a, trans_b, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, work, work_elems);
}

hipError_t gemm_f32(hipStream_t st, bool trans_a, bool trans_b, int m, int n, int k, float alpha, const float* A,
                    int lda, const float* B, int ldb, float beta, float* C, int ldc, float* work,
                    size_t work_elems) {
    return gemm_t<float>(st, trans_a, trans_b, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, work, work_elems);
}

hipError_t gemm_sym_f64(hipStream_t st, bool trans_a, bool trans_b, int n, int k, double alpha, const double* A,
                        int lda, const double* B, int ldb, double beta, double* C, int ldc) {
    return gemm_t<double>(st, trans_a, trans_b, n, n, k, alpha, A, lda, B, ldb, beta, C, ldc, nullptr, 0, true);
}

hipError_t gemm_sym_f32(hipStream_t st, bool trans_a, bool trans_b, int n, int k, float alpha, const float* A,
                        int lda, const float* B, int ldb, float beta, float* C, int ldc) {
    return gemm_t<float>(st, trans_a, trans_b, n, n, k, alpha, A, lda, B, ldb, beta, C, ldc, nullptr, 0, true);
}

}  // namespace am1g
