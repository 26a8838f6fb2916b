// Grid kernels of the large-system SCF (one molecule on the whole device): packing the DIIS
// history with the commutator, the history dot products, the extrapolated Fock matrix, the
// density update with its rms change, per-atom populations, and FP32 staging for the mixed
// precision GEMMs. Every reduction is a fixed-shape two-stage sum (block partials, then one
// block in order), so results do not depend on scheduling.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "am1g_common.hpp"
#include "am1g_launch.hpp"

namespace am1g {
namespace {

__device__ __forceinline__ int64_t up_pos(int i, int j, int n) {  // upper incl. diagonal, i <= j
    return (int64_t)i * n - (int64_t)i * (i - 1) / 2 + (j - i);
}
__device__ __forceinline__ int64_t sa_pos(int i, int j, int n) {  // strict upper, i < j
    return (int64_t)i * n - (int64_t)i * (i + 1) / 2 + (j - i - 1);
}

template <int W>
__device__ __forceinline__ void block_partials(double (&v)[W], double* partial) {
    __shared__ double sh[4][W];
#pragma unroll
    for (int q = 0; q < W; ++q) {
        double x = v[q];
        for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
        if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6][q] = x;
    }
    __syncthreads();
    const int64_t b = (int64_t)blockIdx.y * gridDim.x + blockIdx.x;
    if (threadIdx.x < W) partial[b * W + threadIdx.x] = (sh[0][threadIdx.x] + sh[1][threadIdx.x]) +
                                                        (sh[2][threadIdx.x] + sh[3][threadIdx.x]);
}

// History slot: packed F and D (upper triangle), packed E = FP - (FP)^T (strict upper), with
// FP = T (double or float); block partials of sum E^2.
template <typename TT>
__global__ __launch_bounds__(256) void k_push(int n, const TT* T, const double* F, const double* P, double* hf,
                                              double* he, double* hd, double* partial) {
    const int i = blockIdx.y;
    const int j = blockIdx.x * 256 + threadIdx.x;
    double v[1] = {0.0};
    if (j < n && j >= i) {
        const int64_t u = up_pos(i, j, n);
        hf[u] = F[(size_t)i * n + j];
        if (hd) hd[u] = P[(size_t)i * n + j];
        if (j > i) {
            const double e = (double)T[(size_t)i * n + j] - (double)T[(size_t)j * n + i];
            he[sa_pos(i, j, n)] = e;
            v[0] = e * e;
        }
    }
    block_partials<1>(v, partial);
}

// <e_new, e_r> over the packed commutators of the ring's slots (m <= 8).
__global__ __launch_bounds__(256) void k_dots_e(int64_t len, const double* he, int64_t stride, const int* ring, int m,
                                                int newest, double* partial) {
    double v[kDiisDepth];
#pragma unroll
    for (int q = 0; q < kDiisDepth; ++q) v[q] = 0.0;
    const double* en = he + (int64_t)ring[newest] * stride;
    for (int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x; k < len; k += (int64_t)gridDim.x * 256) {
        const double x = en[k];
#pragma unroll
        for (int q = 0; q < kDiisDepth; ++q)
            if (q < m) v[q] += x * he[(int64_t)ring[q] * stride + k];
    }
    block_partials<kDiisDepth>(v, partial);
}

// A-DIIS Gram entries <D_new, F_r> and <D_r, F_new> over packed symmetric matrices (diagonal
// entries once, off-diagonal twice).
__global__ __launch_bounds__(256) void k_dots_df(int n, const double* hf, const double* hd, int64_t stride,
                                                 const int* ring, int m, int newest, double* partial) {
    const int i = blockIdx.y;
    const int j = blockIdx.x * 256 + threadIdx.x;
    double v[2 * kDiisDepth];
#pragma unroll
    for (int q = 0; q < 2 * kDiisDepth; ++q) v[q] = 0.0;
    if (j < n && j >= i) {
        const int64_t u = up_pos(i, j, n);
        const double w = (i == j) ? 1.0 : 2.0;
        const double dn = w * hd[(int64_t)ring[newest] * stride + u];
        const double fn = w * hf[(int64_t)ring[newest] * stride + u];
#pragma unroll
        for (int q = 0; q < kDiisDepth; ++q)
            if (q < m) {
                v[q] += dn * hf[(int64_t)ring[q] * stride + u];
                v[kDiisDepth + q] += fn * hd[(int64_t)ring[q] * stride + u];
            }
    }
    block_partials<2 * kDiisDepth>(v, partial);
}

// out[q] = sum over the nparts block partials of column q, one block per column (launched with
// `width` blocks), each summed in a fixed order.
__global__ __launch_bounds__(256) void k_reduce_cols(int64_t nparts, int width, const double* partial, double* out) {
    __shared__ double sh[4];
    const int q = blockIdx.x;
    double v = 0.0;
    for (int64_t i = threadIdx.x; i < nparts; i += 256) v += partial[i * width + q];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) out[q] = (sh[0] + sh[1]) + (sh[2] + sh[3]);
}

// Fout = sum_q c_q F_q (unpacked); block partials count non-finite entries.
__global__ __launch_bounds__(256) void k_combine(int n, const double* hf, int64_t stride, const int* ring,
                                                 const double* coef, int m, double* Fout, double* partial) {
    const int i = blockIdx.y;
    const int j = blockIdx.x * 256 + threadIdx.x;
    double v[1] = {0.0};
    if (j < n) {
        const int64_t u = i <= j ? up_pos(i, j, n) : up_pos(j, i, n);
        double acc = 0.0;
        for (int q = 0; q < m; ++q) acc += coef[q] * hf[(int64_t)ring[q] * stride + u];
        Fout[(size_t)i * n + j] = acc;
        if (!isfinite(acc)) v[0] = 1.0;
    }
    block_partials<1>(v, partial);
}

// P <- (Pn + Pn^T) / 2 with block partials of sum (P_new - P_old)^2.
template <typename TT>
__global__ __launch_bounds__(256) void k_density_update(int n, const TT* Pn, double* P, double* partial) {
    const int i = blockIdx.y;
    const int j = blockIdx.x * 256 + threadIdx.x;
    double v[1] = {0.0};
    if (j < n) {
        const double x = 0.5 * ((double)Pn[(size_t)i * n + j] + (double)Pn[(size_t)j * n + i]);
        const double dlt = x - P[(size_t)i * n + j];
        v[0] = dlt * dlt;
        P[(size_t)i * n + j] = x;
    }
    block_partials<1>(v, partial);
}

__global__ __launch_bounds__(256) void k_add(int64_t len, const double* x, const double* y, double* z) {
    for (int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x; k < len; k += (int64_t)gridDim.x * 256) z[k] = x[k] + y[k];
}

__global__ __launch_bounds__(256) void k_to_f32(int64_t len, const double* x, float* y) {
    for (int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x; k < len; k += (int64_t)gridDim.x * 256)
        y[k] = (float)x[k];
}

// ---- Density matrix by second-order spectral projection (SP2; A. M. N. Niklasson, Phys. Rev. B
// 66, 155115 (2002)): the projector on the nocc lowest eigenvectors of F from a sequence of
// squarings X <- X^2 or X <- 2X - X^2 of X_0 = (e_max I - F) / (e_max - e_min), no eigenvectors.

// Gershgorin interval of each row of the symmetric F (one wavefront per row): out[3i] = F_ii -
// sum_{j != i} |F_ij|, out[3i + 1] = F_ii + sum_{j != i} |F_ij|, out[3i + 2] = F_ii.
__global__ __launch_bounds__(256) void k_gersh(int n, const double* F, double* out) {
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
    if (i >= n) return;
    const double* row = F + (size_t)i * n;
    double r = 0.0;
    for (int j = lane; j < n; j += 64) r += j == i ? 0.0 : fabs(row[j]);
    for (int off = 32; off > 0; off >>= 1) r += __shfl_xor(r, off, 64);
    if (lane == 0) {
        const double d = row[i];
        out[3 * i] = d - r;
        out[3 * i + 1] = d + r;
        out[3 * i + 2] = d;
    }
}

// X = scale (e_max I - F)
template <typename T>
__global__ __launch_bounds__(256) void k_sp2_init(int n, const double* F, double emax, double scale, T* X) {
    const int i = blockIdx.y;
    const int j = blockIdx.x * 256 + threadIdx.x;
    if (j < n) X[(size_t)i * n + j] = (T)(scale * ((i == j ? emax : 0.0) - F[(size_t)i * n + j]));
}

// out[0] = trace of X, summed in FP64 in a fixed order (one block)
template <typename T>
__global__ __launch_bounds__(256) void k_trace(int n, const T* X, double* out) {
    __shared__ double sh[4];
    double v = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) v += (double)X[(size_t)i * n + i];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) out[0] = (sh[0] + sh[1]) + (sh[2] + sh[3]);
}

// X2 <- 2 X - X2
template <typename T>
__global__ __launch_bounds__(256) void k_sp2_combine(int64_t len, const T* X, T* X2) {
    for (int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x; k < len; k += (int64_t)gridDim.x * 256)
        X2[k] = (T)2 * X[k] - X2[k];
}

// X <- w X
template <typename T>
__global__ __launch_bounds__(256) void k_scale(int64_t len, double w, T* X) {
    for (int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x; k < len; k += (int64_t)gridDim.x * 256)
        X[k] = (T)(w * (double)X[k]);
}

// out[a] = (population, P[s][px], P[s][py], P[s][pz]) of atom a.
__global__ __launch_bounds__(256) void k_atom_props(int nat, int n, const int* z, const int* aoff, const DevElem* el,
                                                    const double* P, double* out) {
    const int a = blockIdx.x * 256 + threadIdx.x;
    if (a >= nat) return;
    const int o = aoff[a], norb = el[z[a]].norb;
    double pop = 0.0;
    for (int k = 0; k < norb; ++k) pop += P[(size_t)(o + k) * n + o + k];
    out[4 * a] = pop;
    for (int k = 0; k < 3; ++k) out[4 * a + 1 + k] = norb == 4 ? P[(size_t)o * n + o + 1 + k] : 0.0;
}

inline unsigned nblk(int64_t n, int per) { return (unsigned)((n + per - 1) / per); }

}  // namespace

int64_t large_partials(int n) { return (int64_t)nblk(n, 256) * n * 2 * kDiisDepth + 1024 * kDiisDepth + 64; }

hipError_t launch_push(int n, const void* T, bool t_f32, const double* F, const double* P, double* hf, double* he,
                       double* hd, double* partial, double* out, hipStream_t st) {
    const dim3 grid(nblk(n, 256), n);
    if (t_f32)
        hipLaunchKernelGGL(k_push<float>, grid, dim3(256), 0, st, n, (const float*)T, F, P, hf, he, hd, partial);
    else
        hipLaunchKernelGGL(k_push<double>, grid, dim3(256), 0, st, n, (const double*)T, F, P, hf, he, hd, partial);
    hipLaunchKernelGGL(k_reduce_cols, dim3(1), dim3(256), 0, st, (int64_t)grid.x * grid.y, 1, partial, out);
    return hipGetLastError();
}

hipError_t launch_diis_dots(int n, const double* hf, const double* he, const double* hd, const int* ring, int m,
                            int newest, double* partial, double* out, hipStream_t st) {
    const int64_t len_s = (int64_t)n * (n + 1) / 2, len_a = (int64_t)n * (n - 1) / 2;
    const unsigned ge = 1024;
    hipLaunchKernelGGL(k_dots_e, dim3(ge), dim3(256), 0, st, len_a, he, len_a, ring, m, newest, partial);
    hipLaunchKernelGGL(k_reduce_cols, dim3(kDiisDepth), dim3(256), 0, st, (int64_t)ge, kDiisDepth, partial, out);
    if (hd) {
        const dim3 grid(nblk(n, 256), n);
        double* p2 = partial + (int64_t)ge * kDiisDepth;
        hipLaunchKernelGGL(k_dots_df, grid, dim3(256), 0, st, n, hf, hd, len_s, ring, m, newest, p2);
        hipLaunchKernelGGL(k_reduce_cols, dim3(2 * kDiisDepth), dim3(256), 0, st, (int64_t)grid.x * grid.y,
                           2 * kDiisDepth, p2, out + kDiisDepth);
    }
    return hipGetLastError();
}

hipError_t launch_combine(int n, const double* hf, const int* ring, const double* coef, int m, double* Fout,
                          double* partial, double* out, hipStream_t st) {
    const dim3 grid(nblk(n, 256), n);
    hipLaunchKernelGGL(k_combine, grid, dim3(256), 0, st, n, hf, (int64_t)n * (n + 1) / 2, ring, coef, m, Fout,
                       partial);
    hipLaunchKernelGGL(k_reduce_cols, dim3(1), dim3(256), 0, st, (int64_t)grid.x * grid.y, 1, partial, out);
    return hipGetLastError();
}

hipError_t launch_density_update(int n, const void* Pn, bool f32, double* P, double* partial, double* out,
                                 hipStream_t st) {
    const dim3 grid(nblk(n, 256), n);
    if (f32)
        hipLaunchKernelGGL(k_density_update<float>, grid, dim3(256), 0, st, n, (const float*)Pn, P, partial);
    else
        hipLaunchKernelGGL(k_density_update<double>, grid, dim3(256), 0, st, n, (const double*)Pn, P, partial);
    hipLaunchKernelGGL(k_reduce_cols, dim3(1), dim3(256), 0, st, (int64_t)grid.x * grid.y, 1, partial, out);
    return hipGetLastError();
}

hipError_t launch_add(int64_t len, const double* x, const double* y, double* z, hipStream_t st) {
    hipLaunchKernelGGL(k_add, dim3((unsigned)std::min<int64_t>(nblk(len, 256), 65536)), dim3(256), 0, st, len, x, y, z);
    return hipGetLastError();
}

hipError_t launch_to_f32(int64_t len, const double* x, float* y, hipStream_t st) {
    hipLaunchKernelGGL(k_to_f32, dim3((unsigned)std::min<int64_t>(nblk(len, 256), 65536)), dim3(256), 0, st, len, x, y);
    return hipGetLastError();
}

hipError_t launch_gersh(int n, const double* F, double* out, hipStream_t st) {
    hipLaunchKernelGGL(k_gersh, dim3(nblk(n, 4)), dim3(256), 0, st, n, F, out);
    return hipGetLastError();
}

template <typename T>
hipError_t launch_sp2_init(int n, const double* F, double emax, double scale, T* X, hipStream_t st) {
    hipLaunchKernelGGL(k_sp2_init<T>, dim3(nblk(n, 256), n), dim3(256), 0, st, n, F, emax, scale, X);
    return hipGetLastError();
}
template hipError_t launch_sp2_init<double>(int, const double*, double, double, double*, hipStream_t);
template hipError_t launch_sp2_init<float>(int, const double*, double, double, float*, hipStream_t);

template <typename T>
hipError_t launch_trace(int n, const T* X, double* out, hipStream_t st) {
    hipLaunchKernelGGL(k_trace<T>, dim3(1), dim3(256), 0, st, n, X, out);
    return hipGetLastError();
}
template hipError_t launch_trace<double>(int, const double*, double*, hipStream_t);
template hipError_t launch_trace<float>(int, const float*, double*, hipStream_t);

template <typename T>
hipError_t launch_sp2_combine(int64_t len, const T* X, T* X2, hipStream_t st) {
    hipLaunchKernelGGL(k_sp2_combine<T>, dim3((unsigned)std::min<int64_t>(nblk(len, 256), 65536)), dim3(256), 0, st,
                       len, X, X2);
    return hipGetLastError();
}
template hipError_t launch_sp2_combine<double>(int64_t, const double*, double*, hipStream_t);
template hipError_t launch_sp2_combine<float>(int64_t, const float*, float*, hipStream_t);

template <typename T>
hipError_t launch_scale(int64_t len, double w, T* X, hipStream_t st) {
    hipLaunchKernelGGL(k_scale<T>, dim3((unsigned)std::min<int64_t>(nblk(len, 256), 65536)), dim3(256), 0, st, len, w,
                       X);
    return hipGetLastError();
}
template hipError_t launch_scale<double>(int64_t, double, double*, hipStream_t);
template hipError_t launch_scale<float>(int64_t, double, float*, hipStream_t);

hipError_t launch_atom_props(const SysArgs& a, const double* P, double* out, hipStream_t st) {
    if (a.md.nat > 0)
        hipLaunchKernelGGL(k_atom_props, dim3(nblk(a.md.nat, 256)), dim3(256), 0, st, a.md.nat, a.md.nao, a.z, a.aoff,
                           a.elems + a.md.method * kMaxZ, P, out);
    return hipGetLastError();
}

}  // namespace am1g
