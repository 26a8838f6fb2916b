// Workgroup-level building blocks of the SCF: deterministic reductions, a register-tiled GEMM,
// the parallel cyclic Jacobi eigensolver and the DIIS small solves. Every routine is called by
// all threads of one workgroup and orders its floating-point sums independently of what other
// workgroups are doing, so a molecule's result does not depend on the batch it runs in.
#pragma once
#include <hip/hip_runtime.h>

#include "am1g_common.hpp"

namespace am1g {

// Threads per workgroup of the per-molecule SCF: 512 (8 wavefronts); the small-molecule build of
// the batched kernel compiles the same code with 256 (two molecules per CU).
#ifndef AM1G_WG_THREADS
#define AM1G_WG_THREADS 512
#endif
constexpr int kWg = AM1G_WG_THREADS;
constexpr int kWaves = kWg / 64;

struct WgScratch {
    double red[kWaves * 4];  // reduction scratch
    double bcast[16];     // broadcast slots
    double2 rot[128];     // Jacobi rotations of one round: (cos, sin)
    int2 pq[128];         // and their index pairs (q = -1: padding slot of an odd order)
};

// Data-parallel-primitive moves of a double within rows of 16 lanes (no LDS traffic); CTRL is
// the DPP control word (quad permutations 0xB1 = [1,0,3,2] and 0x4E = [2,3,0,1], row mirrors
// 0x141 = lanes 0-7 reversed in each half row and 0x140 = each row of 16 reversed).
template <int CTRL>
__device__ __forceinline__ double dpp_move(double v) {
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}

__device__ __forceinline__ double read_lane(double v, int l) {
    return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), l),
                            __builtin_amdgcn_readlane(__double2loint(v), l));
}

// Sum over the 64 lanes of a wavefront (all lanes active), the same bits in every lane: four
// DPP steps give every lane its row of 16's sum (each step adds a lane's value to its
// partner's, and a + b == b + a), then the four row sums are read as scalars and added in a
// fixed order.
__device__ __forceinline__ double wave_sum(double v) {
    v += dpp_move<0xB1>(v);
    v += dpp_move<0x4E>(v);
    v += dpp_move<0x141>(v);
    v += dpp_move<0x140>(v);
    return (read_lane(v, 0) + read_lane(v, 16)) + (read_lane(v, 32) + read_lane(v, 48));
}

// Deterministic sum over the workgroup: fixed-shape tree over thread slots.
__device__ __forceinline__ double wg_sum(double v, WgScratch& s) {
    const int t = threadIdx.x;
    // wavefront-level butterfly first (fixed order), then across the four wavefronts
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    __syncthreads();
    if ((t & 63) == 0) s.red[t >> 6] = v;
    __syncthreads();
    double tot = 0.0;
#pragma unroll
    for (int w = 0; w < kWaves; ++w) tot += s.red[w];
    __syncthreads();
    return tot;
}

__device__ __forceinline__ double wg_max(double v, WgScratch& s) {
    const int t = threadIdx.x;
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    __syncthreads();
    if ((t & 63) == 0) s.red[t >> 6] = v;
    __syncthreads();
    double tot = s.red[0];
#pragma unroll
    for (int w = 1; w < kWaves; ++w) tot = fmax(tot, s.red[w]);
    __syncthreads();
    return tot;
}

// Several independent sums at once, reduced through `buf` (kWaves * NV doubles of shared
// scratch). Every thread receives all NV totals in out.
template <int NV>
__device__ __forceinline__ void wg_sum_many(const double (&v)[NV], double* buf, double (&out)[NV]) {
    const int t = threadIdx.x;
#pragma unroll
    for (int i = 0; i < NV; ++i) {
        const double x = wave_sum(v[i]);
        if ((t & 63) == 0) buf[i * kWaves + (t >> 6)] = x;
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < NV; ++i) {
        double tot = 0.0;
#pragma unroll
        for (int w = 0; w < kWaves; ++w) tot += buf[i * kWaves + w];
        out[i] = tot;
    }
    __syncthreads();
}

// C[i][j] = alpha * sum_k opA(i,k) * opB(k,j), with opA(i,k) = TA ? A[k*lda+i] : A[i*lda+k] and
// opB(k,j) = TB ? B[j*ldb+k] : B[k*ldb+j]. Each thread owns 4x4 output tiles; operands are
// read through the cache hierarchy (they are the molecule's own, L2-resident matrices).
template <typename T, bool TA, bool TB>
__device__ __noinline__ void wg_gemm(int M, int N, int K, double alpha, const double* __restrict__ A, int lda,
                        const double* __restrict__ B, int ldb, double* __restrict__ C, int ldc) {
    const int tm = (M + 3) >> 2, tn = (N + 3) >> 2;
    for (int tile = threadIdx.x; tile < tm * tn; tile += kWg) {
        const int i0 = (tile / tn) << 2, j0 = (tile % tn) << 2;
        T acc[4][4];
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) acc[a][b] = (T)0;
        for (int k = 0; k < K; ++k) {
            T av[4], bv[4];
#pragma unroll
            for (int a = 0; a < 4; ++a) {
                int i = min(i0 + a, M - 1);
                av[a] = (T)(TA ? A[(size_t)k * lda + i] : A[(size_t)i * lda + k]);
            }
#pragma unroll
            for (int b = 0; b < 4; ++b) {
                int j = min(j0 + b, N - 1);
                bv[b] = (T)(TB ? B[(size_t)j * ldb + k] : B[(size_t)k * ldb + j]);
            }
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int b = 0; b < 4; ++b) acc[a][b] = fma(av[a], bv[b], acc[a][b]);
        }
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b)
                if (i0 + a < M && j0 + b < N) C[(size_t)(i0 + a) * ldc + j0 + b] = alpha * (double)acc[a][b];
    }
}

// C = alpha * opA (M x K) * opB (K x N) with k-panels staged through LDS scratch `sm`
// (`sm_doubles` available). Tiles of 4x4 outputs, at most two per thread, so M and N are at
// most 128; callers use wg_gemm beyond that. Panels are loaded with the contiguous direction of
// each operand across consecutive threads, so global reads are coalesced.
template <typename T, bool TA, bool TB>
__device__ __noinline__ void wg_gemm_staged(int M, int N, int K, double alpha, const double* __restrict__ A, int lda,
                               const double* __restrict__ B, int ldb, double* __restrict__ C, int ldc, double* sm,
                               int sm_doubles) {
    const int t = threadIdx.x;
    const int tm = (M + 3) >> 2, tn = (N + 3) >> 2, ntile = tm * tn;
    const int Mp = tm * 4, Np = tn * 4;
    // panels are staged in the arithmetic type (a float panel takes half the space and traffic)
    int kp = (int)(sm_doubles * (sizeof(double) / sizeof(T))) / (Mp + Np);
    kp = kp > 64 ? 64 : kp;
    T* As = (T*)sm;
    T* Bs = As + kp * Mp;
    T acc[2][4][4];
#pragma unroll
    for (int q = 0; q < 2; ++q)
#pragma unroll
        for (int a = 0; a < 4; ++a)
#pragma unroll
            for (int b = 0; b < 4; ++b) acc[q][a][b] = (T)0;
    for (int k0 = 0; k0 < K; k0 += kp) {
        const int kc = min(kp, K - k0);
        for (int idx = t; idx < kc * Mp; idx += kWg) {
            int kk, i;
            if (TA) {
                kk = idx / Mp;
                i = idx - kk * Mp;
            } else {
                i = idx / kc;
                kk = idx - i * kc;
            }
            int ic = min(i, M - 1);
            As[kk * Mp + i] = (T)(TA ? A[(size_t)(k0 + kk) * lda + ic] : A[(size_t)ic * lda + k0 + kk]);
        }
        for (int idx = t; idx < kc * Np; idx += kWg) {
            int kk, j;
            if (TB) {
                j = idx / kc;
                kk = idx - j * kc;
            } else {
                kk = idx / Np;
                j = idx - kk * Np;
            }
            int jc = min(j, N - 1);
            Bs[kk * Np + j] = (T)(TB ? B[(size_t)jc * ldb + k0 + kk] : B[(size_t)(k0 + kk) * ldb + jc]);
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const int tile = t + q * kWg;
            if (tile < ntile) {
                const int i0 = (tile / tn) << 2, j0 = (tile % tn) << 2;
                for (int kk = 0; kk < kc; ++kk) {
                    T av[4], bv[4];
#pragma unroll
                    for (int a = 0; a < 4; ++a) av[a] = As[kk * Mp + i0 + a];
#pragma unroll
                    for (int b = 0; b < 4; ++b) bv[b] = Bs[kk * Np + j0 + b];
#pragma unroll
                    for (int a = 0; a < 4; ++a)
#pragma unroll
                        for (int b = 0; b < 4; ++b) acc[q][a][b] = fma(av[a], bv[b], acc[q][a][b]);
                }
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int q = 0; q < 2; ++q) {
        const int tile = t + q * kWg;
        if (tile < ntile) {
            const int i0 = (tile / tn) << 2, j0 = (tile % tn) << 2;
#pragma unroll
            for (int a = 0; a < 4; ++a)
#pragma unroll
                for (int b = 0; b < 4; ++b)
                    if (i0 + a < M && j0 + b < N) C[(size_t)(i0 + a) * ldc + j0 + b] = alpha * (double)acc[q][a][b];
        }
    }
}

// Matrix-core GEMM for M, N <= 128: C = alpha * opA (M x K) * opB (K x N) on 16 x 16 output
// tiles, each an accumulation of 16x16x4 MFMA products (FP64, or FP32 with FP32 accumulation);
// wavefront w owns tiles w, w + 8, ... (row-major over the tile grid). K-panels of both operands
// are staged through LDS in the arithmetic type, zero-padded to whole tiles and to a multiple
// of four in K. Operand lanes: A(i, k) in lane i + 16k, B(k, j) in lane j + 16k; result lane l,
// register r: row l/16 + 4r (FP64) or 4(l/16) + r (FP32), column l%16.
typedef double mf64x4 __attribute__((ext_vector_type(4)));
typedef float mf32x4 __attribute__((ext_vector_type(4)));

template <typename T>
struct MfmaAcc;
template <>
struct MfmaAcc<double> {
    typedef mf64x4 V;
    static __device__ __forceinline__ V mma(double a, double b, V c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int l, int r) { return (l >> 4) + 4 * r; }
};
template <>
struct MfmaAcc<float> {
    typedef mf32x4 V;
    static __device__ __forceinline__ V mma(float a, float b, V c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int l, int r) { return 4 * (l >> 4) + r; }
};

template <typename T, bool TA, bool TB>
__device__ __noinline__ void wg_gemm_mfma(int M, int N, int K, double alpha, const double* __restrict__ A, int lda,
                                          const double* __restrict__ B, int ldb, double* __restrict__ C, int ldc,
                                          double* sm, int sm_doubles) {
    typedef typename MfmaAcc<T>::V V;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int TM = (M + 15) >> 4, TN = (N + 15) >> 4, ntile = TM * TN;
    const int Mp = TM * 16, Np = TN * 16;
    int kp = (int)(sm_doubles * (sizeof(double) / sizeof(T))) / (Mp + Np);
    kp = (kp > 64 ? 64 : kp) & ~3;
    T* As = (T*)sm;
    T* Bs = As + kp * Mp;
    V acc[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) acc[u] = V{0, 0, 0, 0};
    // Panel staging: each thread takes a fixed position along the operand's contiguous index and
    // steps over the other (one division per panel instead of one per element, so the loads of
    // a thread issue back to back).
    for (int k0 = 0; k0 < K; k0 += kp) {
        const int kc = min(kp, K - k0), kc4 = (kc + 3) & ~3;
        {
            // op(A) panel [kk][i]: contiguous along i (TA) or along kk
            const int fast = TA ? Mp : kc4, slow = TA ? kc4 : Mp;
            const int f = t % fast, s0 = t / fast, step = kWg / fast;
            if (s0 < step) {
#pragma unroll 4
                for (int sl = s0; sl < slow; sl += step) {
                    const int kk = TA ? sl : f, i = TA ? f : sl;
                    As[kk * Mp + i] = (kk < kc && i < M)
                                          ? (T)(TA ? A[(size_t)(k0 + kk) * lda + i] : A[(size_t)i * lda + k0 + kk])
                                          : (T)0;
                }
            }
        }
        {
            // op(B) panel [kk][j]: contiguous along kk (TB) or along j
            const int fast = TB ? kc4 : Np, slow = TB ? Np : kc4;
            const int f = t % fast, s0 = t / fast, step = kWg / fast;
            if (s0 < step) {
#pragma unroll 4
                for (int sl = s0; sl < slow; sl += step) {
                    const int kk = TB ? f : sl, j = TB ? sl : f;
                    Bs[kk * Np + j] = (kk < kc && j < N)
                                          ? (T)(TB ? B[(size_t)j * ldb + k0 + kk] : B[(size_t)(k0 + kk) * ldb + j])
                                          : (T)0;
                }
            }
        }
        __syncthreads();
        for (int kk = 0; kk < kc4; kk += 4) {
            const int ka = (kk + (lane >> 4));
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int tile = w + kWaves * u;
                if (tile < ntile) {
                    const int ti = tile / TN, tj = tile - ti * TN;
                    const T a = As[ka * Mp + ti * 16 + (lane & 15)];
                    const T b = Bs[ka * Np + tj * 16 + (lane & 15)];
                    acc[u] = MfmaAcc<T>::mma(a, b, acc[u]);
                }
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
        const int tile = w + kWaves * u;
        if (tile < ntile) {
            const int ti = tile / TN, tj = tile - ti * TN;
            const int col = tj * 16 + (lane & 15);
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = ti * 16 + MfmaAcc<T>::row(lane, r);
                if (row < M && col < N) C[(size_t)row * ldc + col] = alpha * (double)acc[u][r];
            }
        }
    }
}

// Dispatch: matrix cores when the output has at most eight 16 x 16 tiles per wavefront (8 x 8
// tiles with 8 wavefronts), LDS-staged register tiles when it fits two 4 x 4 tiles per thread
// (orders up to 128 with 4 wavefronts), register/global otherwise.
template <typename T, bool TA, bool TB>
__device__ __forceinline__ void wg_gemm_any(int M, int N, int K, double alpha, const double* A, int lda,
                                            const double* B, int ldb, double* C, int ldc, double* sm,
                                            int sm_doubles) {
    if (M <= 128 && N <= 128 && ((M + 15) / 16) * ((N + 15) / 16) <= 8 * kWaves)
        wg_gemm_mfma<T, TA, TB>(M, N, K, alpha, A, lda, B, ldb, C, ldc, sm, sm_doubles);
    else if (M <= 128 && N <= 128 && ((M + 3) / 4) * ((N + 3) / 4) <= 2 * kWg)
        wg_gemm_staged<T, TA, TB>(M, N, K, alpha, A, lda, B, ldb, C, ldc, sm, sm_doubles);
    else
        wg_gemm<T, TA, TB>(M, N, K, alpha, A, lda, B, ldb, C, ldc);
}

// Round-robin partner schedule for m (even) players, L = m - 1 rounds per sweep: round r in
// [0, L), slot k in [0, m/2) -> (p, q), p < q. Slot 0 pairs player 0 with r + 1; slot k > 0
// pairs players x + 1 and y + 1 with x = r + k and y = r - k (mod L). Both operands of each
// wrap are below 2L, so one conditional subtraction replaces the modulo.
__device__ __forceinline__ int rr_wrap(int x, int L) { return x >= L ? x - L : x; }

__device__ __forceinline__ void rr_pair(int m, int r, int k, int& p, int& q) {
    const int L = m - 1;
    int a, b;
    if (k == 0) {
        a = 0;
        b = r + 1;
    } else {
        a = rr_wrap(r + k, L) + 1;
        b = rr_wrap(r + L - k, L) + 1;
    }
    p = min(a, b);
    q = max(a, b);
}

// 1/x and 1/sqrt(x) for positive normal x: the hardware estimates refined by two Newton steps
// each (y <- y + y(1 - xy); y <- y + y(1 - xy^2)/2), which leaves them within a few units in the
// last place, ample for rotation parameters (their orthogonality is restored by c = 1/sqrt(1+t^2)).
__device__ __forceinline__ double fast_rcp(double x) {
    double y = __builtin_amdgcn_rcp(x);
    double e = fma(-x, y, 1.0);
    y = fma(y, e, y);
    e = fma(-x, y, 1.0);
    return fma(y, e, y);
}

__device__ __forceinline__ double fast_rsqrt(double x) {
    double y = __builtin_amdgcn_rsq(x);
    double h = 0.5 * x;
    double e = fma(-h * y, y, 0.5);
    y = fma(y, e, y);
    e = fma(-h * y, y, 0.5);
    return fma(y, e, y);
}

// Parallel cyclic Jacobi on a symmetric n x n matrix, rotating the rows of Vt (Vt = V^T, row k =
// eigenvector k) along with it. A is either packed (PACKED: lower triangle, a(i, j) at
// i(i+1)/2 + j for i >= j, any address space) or full row-major with stride lda. On return the
// eigenvalues are on A's diagonal (unsorted).
//
// Each round applies the m/2 disjoint rotations of a round-robin schedule (m = n rounded up to
// even). The matrix is cut into 2x2 blocks (row pair a, column pair b); a block's new values
// depend only on its old values and the two rotations, so each block is updated in place by
// one thread without races, and with packed storage the four entries of an off-diagonal block
// are four distinct stored numbers (nothing to mirror). A thread's blocks and eigenvector
// elements are the same slots every round, so they are decoded once per solve.
//
// Converged when the largest off-diagonal element is below tau = max(1e-13 x (rms of the
// diagonal), tol_abs); the same tau decides which rotations are applied, so a converged sweep is
// exactly one in which nothing would rotate (a tighter target sits at the rotations' own
// rounding floor). The SCF passes tol_abs = 1e-7 x its current commutator error: an intermediate
// Fock matrix needs its eigenvectors only to well inside the error the next iteration removes.
// Returns the number of sweeps, or -1 when max_sweeps did not suffice.
// Rotation blocks per thread: ceil(h(h+1)/2 / kWg) with h = n/2 pairs; 5 covers n <= 128 and 10
// covers n <= 192 (the larger count costs registers, so it is a separate instantiation).
constexpr int kBlkSmall = 5;
constexpr int kBlkLarge = 10;

template <bool PACKED>
__device__ __forceinline__ size_t jidx(int i, int j, int lda) {
    if (PACKED) return i >= j ? (size_t)i * (i + 1) / 2 + j : (size_t)j * (j + 1) / 2 + i;
    return (size_t)i * lda + j;
}

// One Jacobi rotation for the 2x2 block (app, apq; apq, aqq): (cos, sin), or (1, 0) when
// |apq| <= tau. With d = aqq - app, the tangent of the smaller rotation is
// t = sign(d) 2 apq / (|d| + sqrt(d^2 + 4 apq^2)) (sign(0) = sign(apq), so t = +1 at d = 0), the
// same number as sign(theta) / (|theta| + sqrt(theta^2 + 1)) with theta = d / (2 apq), without
// the division by apq; c = 1 / sqrt(1 + t^2), s = t c.
__device__ __forceinline__ double2 jacobi_angle(double app, double apq, double aqq, double tau) {
    double2 r = make_double2(1.0, 0.0);
    if (fabs(apq) > tau) {
        const double d = aqq - app;
        const double b = 2.0 * apq;
        const double w = fma(d, d, b * b);
        // the estimates need a normal, finite argument; outside 1e-280..1e280 use the library
        const double root = (w > 1.0e-280 && w < 1.0e280) ? w * fast_rsqrt(w) : sqrt(w);
        const double num = d > 0.0 ? b : (d < 0.0 ? -b : fabs(b));
        const double den = fabs(d) + root;
        const double tt = (den > 1.0e-280 && den < 1.0e280) ? num * fast_rcp(den) : num / den;
        r.x = fast_rsqrt(fma(tt, tt, 1.0));  // 1 + t^2 lies in [1, 2]
        r.y = tt * r.x;
    }
    return r;
}

// LDS-resident Jacobi: A packed at lds_dyn[0, n(n+1)/2). Eigenvector rows (rows of Vt) below
// v_rows live in LDS at lds_dyn[v_off + row * n], the others in global memory at Vg + row * n, so
// orders whose full eigenvector matrix does not fit beside A still keep most rows on chip.
// The eigenvector update gives each rotation pair a fixed group of threads that splits the
// pair's two rows by columns, so a thread reads its rotation once per round and then streams
// contiguous elements.
// The dynamic area starts after the kernel's static LDS; the 16-byte alignment keeps the 16-byte
// (double2) accesses of the eigensolvers aligned whatever the static size is (a misaligned base
// made every such access slower and the in-workgroup eigensolver about 1.5x slower).
extern __shared__ __align__(16) double lds_dyn[];

__device__ __forceinline__ void rotate_rows(double* __restrict__ xp, double* __restrict__ xq, int i0, int n, int step,
                                            double c, double sn) {
#pragma unroll 4
    for (int i = i0; i < n; i += step) {
        const double vp = xp[i], vq = xq[i];
        xp[i] = c * vp - sn * vq;
        xq[i] = sn * vp + c * vq;
    }
}

template <int kBlkPer>
__device__ __noinline__ int wg_jacobi_lds_t(int n, int v_off, int v_rows, double* __restrict__ Vg, WgScratch& s,
                                            int max_sweeps, double tol_abs) {
    const int t = threadIdx.x;
    const int m = (n + 1) & ~1;
    const int h = m >> 1;
    const int nblk = h * (h + 1) / 2;
    double* A = lds_dyn;
    int bslot[kBlkPer];
#pragma unroll
    for (int u = 0; u < kBlkPer; ++u) {
        int blk = t + u * kWg;
        bslot[u] = -1;
        if (blk < nblk) {
            int a = (int)((sqrt(8.0 * blk + 1.0) - 1.0) * 0.5);
            while (a * (a + 1) / 2 > blk) --a;
            while ((a + 1) * (a + 2) / 2 <= blk) ++a;
            bslot[u] = (a << 8) | (blk - a * (a + 1) / 2);
        }
    }
    // eigenvector work: pair va, columns vc0, vc0 + g, ... (g threads per pair)
    const int g = kWg / max(h, 1);
    const int va = t / g, vc0 = t - va * g;
    const bool vown = va < h;
    for (int sweep = 0; sweep <= max_sweeps; ++sweep) {
        double offmax = 0.0, dg = 0.0;
        for (int i = t / 32; i < n; i += kWg / 32) {
            const int rb = i * (i + 1) / 2;
            for (int j = t % 32; j <= i; j += 32) {
                double v = A[rb + j];
                if (i == j) dg += v * v;
                else offmax = fmax(offmax, fabs(v));
            }
        }
        offmax = wg_max(offmax, s);
        dg = wg_sum(dg, s);
        const double tau = fmax(1.0e-13 * sqrt(dg / (double)max(n, 1)), tol_abs);
        if (offmax <= tau) return sweep;
        if (sweep == max_sweeps) break;
        for (int r = 0; r < m - 1; ++r) {
            if (t < h) {
                int p, q;
                rr_pair(m, r, t, p, q);
                double2 cs = make_double2(1.0, 0.0);
                if (q < n) cs = jacobi_angle(A[p * (p + 1) / 2 + p], A[q * (q + 1) / 2 + p], A[q * (q + 1) / 2 + q], tau);
                s.rot[t] = cs;
                s.pq[t] = make_int2(p, q < n ? q : -1);
            }
            __syncthreads();
#pragma unroll
            for (int u = 0; u < kBlkPer; ++u) {
                if (bslot[u] < 0) continue;
                const int a = bslot[u] >> 8, b = bslot[u] & 255;
                const double2 ra = s.rot[a], rb = s.rot[b];
                const double ca = ra.x, sa = ra.y, cb = rb.x, sb = rb.y;
                // a slot with s == 0 still owns rows the other side's rotation moves
                if (sa == 0.0 && (a == b || sb == 0.0)) continue;
                const int2 ia = s.pq[a], ib = s.pq[b];
                const int pa = ia.x, qa = ia.y, pb = ib.x, qb = ib.y;
                if (a == b) {
                    const int i00 = pa * (pa + 1) / 2 + pa, i11 = qa * (qa + 1) / 2 + qa, i01 = qa * (qa + 1) / 2 + pa;
                    const double x00 = A[i00], x01 = A[i01], x11 = A[i11];
                    A[i00] = ca * (ca * x00 - sa * x01) - sa * (ca * x01 - sa * x11);
                    A[i11] = sa * (sa * x00 + ca * x01) + ca * (sa * x01 + ca * x11);
                    A[i01] = ca * (sa * x00 + ca * x01) - sa * (sa * x01 + ca * x11);
                } else {
                    // rows {pa, qa} x columns {pb, qb}; a missing (padding) index reads as zero
                    const int i00 = (int)jidx<true>(pa, pb, 0);
                    const int i01 = qb >= 0 ? (int)jidx<true>(pa, qb, 0) : -1;
                    const int i10 = qa >= 0 ? (int)jidx<true>(qa, pb, 0) : -1;
                    const int i11 = (qa >= 0 && qb >= 0) ? (int)jidx<true>(qa, qb, 0) : -1;
                    const double x00 = A[i00], x01 = i01 >= 0 ? A[i01] : 0.0, x10 = i10 >= 0 ? A[i10] : 0.0,
                                 x11 = i11 >= 0 ? A[i11] : 0.0;
                    // rows: J_a^T = [[ca, -sa], [sa, ca]]; columns: J_b = [[cb, sb], [-sb, cb]]
                    const double r00 = ca * x00 - sa * x10, r01 = ca * x01 - sa * x11;
                    const double r10 = sa * x00 + ca * x10, r11 = sa * x01 + ca * x11;
                    A[i00] = r00 * cb - r01 * sb;
                    if (i01 >= 0) A[i01] = r00 * sb + r01 * cb;
                    if (i10 >= 0) A[i10] = r10 * cb - r11 * sb;
                    if (i11 >= 0) A[i11] = r10 * sb + r11 * cb;
                }
            }
            // Vt rows p, q: (v_p, v_q) <- (c v_p - s v_q, s v_p + c v_q)
            if (vown) {
                const double2 rv = s.rot[va];
                if (rv.y != 0.0) {
                    const int2 iv = s.pq[va];
                    const bool pl = iv.x < v_rows, ql = iv.y < v_rows;
                    if (pl && ql) {
                        rotate_rows(lds_dyn + v_off + iv.x * n, lds_dyn + v_off + iv.y * n, vc0, n, g, rv.x, rv.y);
                    } else if (!pl && !ql) {
                        rotate_rows(Vg + iv.x * n, Vg + iv.y * n, vc0, n, g, rv.x, rv.y);
                    } else {
                        // one row on chip, one in global memory (p < q, so p is the LDS one)
                        double* xp = lds_dyn + v_off + iv.x * n;
                        double* xq = Vg + iv.y * n;
#pragma unroll 4
                        for (int i = vc0; i < n; i += g) {
                            const double vp = xp[i], vq = xq[i];
                            xp[i] = rv.x * vp - rv.y * vq;
                            xq[i] = rv.y * vp + rv.x * vq;
                        }
                    }
                }
            }
            __syncthreads();
        }
    }
    return -1;
}

__device__ __forceinline__ int wg_jacobi_lds(int n, int v_off, int v_rows, double* __restrict__ Vg, WgScratch& s,
                                             int max_sweeps, double tol_abs) {
    if (n <= 128) return wg_jacobi_lds_t<kBlkSmall>(n, v_off, v_rows, Vg, s, max_sweeps, tol_abs);
    return wg_jacobi_lds_t<kBlkLarge>(n, v_off, v_rows, Vg, s, max_sweeps, tol_abs);
}

// Jacobi for orders above 192 (global memory, full storage): same algorithm, slots decoded on
// the fly because a thread owns more of them than the register budget of the fast path holds.
__device__ __noinline__ int wg_jacobi_big(int n, double* A, int lda, double* Vt, int ldv, WgScratch& s, int max_sweeps,
                                          double tol_abs) {
    const int t = threadIdx.x;
    const int m = (n + 1) & ~1;
    const int h = m >> 1;
    const int nblk = h * (h + 1) / 2;
    for (int sweep = 0; sweep <= max_sweeps; ++sweep) {
        double offmax = 0.0, dg = 0.0;
        for (int i = t / 32; i < n; i += kWg / 32)
            for (int j = t % 32; j <= i; j += 32) {
                double v = A[(size_t)i * lda + j];
                if (i == j) dg += v * v;
                else offmax = fmax(offmax, fabs(v));
            }
        offmax = wg_max(offmax, s);
        dg = wg_sum(dg, s);
        const double tau = fmax(1.0e-13 * sqrt(dg / (double)max(n, 1)), tol_abs);
        if (offmax <= tau) return sweep;
        if (sweep == max_sweeps) break;
        for (int r = 0; r < m - 1; ++r) {
            if (t < h) {
                int p, q;
                rr_pair(m, r, t, p, q);
                double2 cs = make_double2(1.0, 0.0);
                if (q < n)
                    cs = jacobi_angle(A[(size_t)p * lda + p], A[(size_t)p * lda + q], A[(size_t)q * lda + q], tau);
                s.rot[t] = cs;
                s.pq[t] = make_int2(p, q < n ? q : -1);
            }
            __syncthreads();
            for (int blk = t; blk < nblk; blk += kWg) {
                int a = (int)((sqrt(8.0 * blk + 1.0) - 1.0) * 0.5);
                while (a * (a + 1) / 2 > blk) --a;
                while ((a + 1) * (a + 2) / 2 <= blk) ++a;
                const int b = blk - a * (a + 1) / 2;
                const double2 ra = s.rot[a], rb = s.rot[b];
                if (!((ra.y != 0.0) || (a != b && rb.y != 0.0))) continue;
                const int pa = s.pq[a].x, qa = s.pq[a].y, pb = s.pq[b].x, qb = s.pq[b].y;
                const double ca = ra.x, sa = ra.y, cb = rb.x, sb = rb.y;
                if (a == b) {
                    double x00 = A[(size_t)pa * lda + pa], x01 = A[(size_t)pa * lda + qa];
                    double x11 = A[(size_t)qa * lda + qa];
                    double y00 = ca * (ca * x00 - sa * x01) - sa * (ca * x01 - sa * x11);
                    double y11 = sa * (sa * x00 + ca * x01) + ca * (sa * x01 + ca * x11);
                    double y01 = ca * (sa * x00 + ca * x01) - sa * (sa * x01 + ca * x11);
                    A[(size_t)pa * lda + pa] = y00;
                    A[(size_t)qa * lda + qa] = y11;
                    A[(size_t)pa * lda + qa] = y01;
                    A[(size_t)qa * lda + pa] = y01;
                } else {
                    double x00 = A[(size_t)pa * lda + pb];
                    double x01 = qb >= 0 ? A[(size_t)pa * lda + qb] : 0.0;
                    double x10 = qa >= 0 ? A[(size_t)qa * lda + pb] : 0.0;
                    double x11 = (qa >= 0 && qb >= 0) ? A[(size_t)qa * lda + qb] : 0.0;
                    double r00 = ca * x00 - sa * x10, r01 = ca * x01 - sa * x11;
                    double r10 = sa * x00 + ca * x10, r11 = sa * x01 + ca * x11;
                    double y00 = r00 * cb - r01 * sb, y01 = r00 * sb + r01 * cb;
                    double y10 = r10 * cb - r11 * sb, y11 = r10 * sb + r11 * cb;
                    A[(size_t)pa * lda + pb] = y00;
                    A[(size_t)pb * lda + pa] = y00;
                    if (qb >= 0) {
                        A[(size_t)pa * lda + qb] = y01;
                        A[(size_t)qb * lda + pa] = y01;
                    }
                    if (qa >= 0) {
                        A[(size_t)qa * lda + pb] = y10;
                        A[(size_t)pb * lda + qa] = y10;
                    }
                    if (qa >= 0 && qb >= 0) {
                        A[(size_t)qa * lda + qb] = y11;
                        A[(size_t)qb * lda + qa] = y11;
                    }
                }
            }
            for (int idx = t; idx < h * n; idx += kWg) {
                const int a = idx / n, i = idx - a * n;
                const double sa = s.rot[a].y;
                if (sa == 0.0) continue;
                const double ca = s.rot[a].x;
                const int p = s.pq[a].x, q = s.pq[a].y;
                const double vp = Vt[(size_t)p * ldv + i], vq = Vt[(size_t)q * ldv + i];
                Vt[(size_t)p * ldv + i] = ca * vp - sa * vq;
                Vt[(size_t)q * ldv + i] = sa * vp + ca * vq;
            }
            __syncthreads();
        }
    }
    return -1;
}

// Gaussian elimination with partial pivoting on the bordered DIIS system (am1-rs
// scf.rs:1095-1141): `m` is dim x dim row-major, modified in place; false on a pivot below 1e-12.
__host__ __device__ inline bool solve_bordered(double* m, double* rhs, int n, double* x) {
    for (int col = 0; col < n; ++col) {
        int pivot = col;
        double best = fabs(m[col * n + col]);
        for (int row = col + 1; row < n; ++row) {
            double v = fabs(m[row * n + col]);
            if (v > best) {
                best = v;
                pivot = row;
            }
        }
        if (best < 1.0e-12) return false;
        if (pivot != col) {
            for (int j = 0; j < n; ++j) {
                double tmp = m[col * n + j];
                m[col * n + j] = m[pivot * n + j];
                m[pivot * n + j] = tmp;
            }
            double tmp = rhs[col];
            rhs[col] = rhs[pivot];
            rhs[pivot] = tmp;
        }
        for (int row = col + 1; row < n; ++row) {
            double factor = m[row * n + col] / m[col * n + col];
            if (factor == 0.0) continue;
            for (int j = col; j < n; ++j) m[row * n + j] -= factor * m[col * n + j];
            rhs[row] -= factor * rhs[col];
        }
    }
    for (int col = n - 1; col >= 0; --col) {
        double sum = rhs[col];
        for (int j = col + 1; j < n; ++j) sum -= m[col * n + j] * x[j];
        x[col] = sum / m[col * n + col];
    }
    return true;
}

// Euclidean projection onto the probability simplex (Duchi et al.; am1-rs scf.rs:1181-1204).
__host__ __device__ inline void simplex_project(const double* v, int n, double* out) {
    double u[kDiisDepth];
    for (int i = 0; i < n; ++i) u[i] = v[i];
    for (int i = 1; i < n; ++i) {  // insertion sort, descending
        double x = u[i];
        int j = i - 1;
        while (j >= 0 && u[j] < x) {
            u[j + 1] = u[j];
            --j;
        }
        u[j + 1] = x;
    }
    double css = 0.0, theta = 0.0;
    for (int j = 0; j < n; ++j) {
        css += u[j];
        double tt = (css - 1.0) / (double)(j + 1);
        if (u[j] - tt > 0.0) theta = tt;
    }
    for (int i = 0; i < n; ++i) out[i] = fmax(v[i] - theta, 0.0);
}

// A-DIIS coefficients from the Gram matrix g[i][j] = <D_i, F_j> (logical order, newest last):
// projected-gradient descent on the simplex (am1-rs scf.rs:1061-1091, 1144-1179).
__host__ __device__ inline void adiis_solve(const double* g, int m, double* c) {
    const int last = m - 1;
    double d[kDiisDepth], S[kDiisDepth * kDiisDepth];
    for (int i = 0; i < m; ++i) d[i] = g[i * kDiisDepth + last] - g[last * kDiisDepth + last];
    for (int i = 0; i < m; ++i)
        for (int j = 0; j < m; ++j)
            S[i * kDiisDepth + j] = g[i * kDiisDepth + j] - g[i * kDiisDepth + last] - g[last * kDiisDepth + j] +
                                    g[last * kDiisDepth + last];
    double l = 1.0e-12;
    for (int i = 0; i < m; ++i) {
        double row = 0.0;
        for (int j = 0; j < m; ++j) row += fabs(S[i * kDiisDepth + j] + S[j * kDiisDepth + i]);
        l = fmax(l, row);
    }
    double lr = 1.0 / l;
    for (int i = 0; i < m; ++i) c[i] = 0.0;
    c[m - 1] = 1.0;
    for (int it = 0; it < 400; ++it) {
        double trial[kDiisDepth], proj[kDiisDepth];
        for (int k = 0; k < m; ++k) {
            double acc = 2.0 * d[k];
            for (int j = 0; j < m; ++j) acc += (S[k * kDiisDepth + j] + S[j * kDiisDepth + k]) * c[j];
            trial[k] = c[k] - lr * acc;
        }
        simplex_project(trial, m, proj);
        double delta = 0.0;
        for (int i = 0; i < m; ++i) {
            delta += fabs(proj[i] - c[i]);
            c[i] = proj[i];
        }
        if (delta < 1.0e-12) break;
    }
}

// CDIIS coefficients on one wavefront: the bordered system [B -1; -1 0] [c; lambda] = [0; -1] of
// solve_bordered (Gaussian elimination with partial pivoting, the first largest pivot winning,
// then back substitution) with lane r holding row r, so each column's eliminations run in
// parallel and every lane takes the same pivot decisions. B: logical Gram matrix (row stride
// kDiisDepth), m >= 2 entries. Returns false when a pivot falls below 1e-12 or, with
// check_finite (the RHF loop's rule), when the coefficients are not finite; otherwise lane 0
// writes c_out[0..m). Called by all 64 lanes of one wavefront.
__device__ inline bool cdiis_solve_wave(const double* B, int m, double* c_out, bool check_finite) {
    constexpr int kD = kDiisDepth + 1;
    const int lane = threadIdx.x & 63;
    const int dim = m + 1;
    double a[kD];
#pragma unroll
    for (int j = 0; j < kD; ++j) {
        double v = 0.0;
        if (lane < m && j < m) v = B[lane * kDiisDepth + j];
        else if ((lane < m && j == m) || (lane == m && j < m)) v = -1.0;
        a[j] = v;
    }
    double rhs = lane == m ? -1.0 : 0.0;
    bool ok = true;  // (no early exits: the loops must unroll so that a[] stays in registers)
#pragma unroll
    for (int col = 0; col < kD; ++col) {
        if (col >= dim || !ok) continue;
        double best = -1.0;
        int piv = col;
        for (int row = col; row < dim; ++row) {
            const double v = fabs(read_lane(a[col], row));
            if (v > best) {
                best = v;
                piv = row;
            }
        }
        if (best < 1.0e-12) {
            ok = false;
            continue;
        }
        if (piv != col) {
            const int src = lane == col ? piv : (lane == piv ? col : lane);
#pragma unroll
            for (int j = col; j < kD; ++j) a[j] = __shfl(a[j], src, 64);
            rhs = __shfl(rhs, src, 64);
        }
        double prow[kD];
#pragma unroll
        for (int j = col; j < kD; ++j) prow[j] = read_lane(a[j], col);
        const double prhs = read_lane(rhs, col);
        if (lane > col && lane < dim) {
            const double factor = a[col] / prow[col];
            if (factor != 0.0) {
#pragma unroll
                for (int j = col; j < kD; ++j) a[j] -= factor * prow[j];
                rhs -= factor * prhs;
            }
        }
    }
    double x[kD];
#pragma unroll
    for (int j = 0; j < kD; ++j) x[j] = 0.0;
#pragma unroll
    for (int col = kD - 1; col >= 0; --col) {
        if (col >= dim || !ok) continue;
        double sum = rhs;
#pragma unroll
        for (int j = col + 1; j < kD; ++j)
            if (j < dim) sum -= a[j] * x[j];
        x[col] = read_lane(sum / a[col], col);
    }
    if (!ok) return false;
    double wsum = 0.0;
#pragma unroll
    for (int i = 0; i < kDiisDepth; ++i)
        if (i < m) wsum += fabs(x[i]);
    if (check_finite && !isfinite(wsum)) return false;
    if (lane == 0) {
#pragma unroll
        for (int i = 0; i < kDiisDepth; ++i)
            if (i < m) c_out[i] = x[i];
    }
    return true;
}

// A-DIIS coefficients on one wavefront: the projected-gradient iteration of adiis_solve with
// lanes 0..m-1 carrying one coefficient each (called by all 64 lanes of one wavefront). Each
// step's matrix-vector product, simplex projection (ranks instead of a sort; ties keep the lower
// index first, as the stable sort does) and convergence test read the other lanes' values as
// scalars, so every lane takes the same decisions. Writes c_out[0..m).
__device__ inline void adiis_solve_wave(const double* g, int m, double* c_out) {
    const int lane = threadIdx.x & 63;
    const int last = m - 1;
    const bool act = lane < m;
    const int k = act ? lane : 0;
    const double gll = g[last * kDiisDepth + last];
    const double dk = g[k * kDiisDepth + last] - gll;
    double Mk[kDiisDepth];
    double row = 0.0;
#pragma unroll
    for (int j = 0; j < kDiisDepth; ++j) {
        double v = 0.0;
        if (j < m) {
            const double skj = g[k * kDiisDepth + j] - g[k * kDiisDepth + last] - g[last * kDiisDepth + j] + gll;
            const double sjk = g[j * kDiisDepth + k] - g[j * kDiisDepth + last] - g[last * kDiisDepth + k] + gll;
            v = skj + sjk;
        }
        Mk[j] = v;
        row += fabs(v);
    }
    double l = 1.0e-12;
    for (int i = 0; i < m; ++i) l = fmax(l, read_lane(row, i));
    const double lr = 1.0 / l;
    double ck = (act && k == last) ? 1.0 : 0.0;
    for (int it = 0; it < 400; ++it) {
        double cj[kDiisDepth];
#pragma unroll
        for (int j = 0; j < kDiisDepth; ++j) cj[j] = j < m ? read_lane(ck, j) : 0.0;
        double acc = 2.0 * dk;
#pragma unroll
        for (int j = 0; j < kDiisDepth; ++j)
            if (j < m) acc += Mk[j] * cj[j];
        const double trial = ck - lr * acc;
        int rank = 0;
        double css = 0.0;
#pragma unroll
        for (int j = 0; j < kDiisDepth; ++j) {
            if (j >= m) continue;
            const double tj = read_lane(trial, j);
            const bool above = tj > trial || (tj == trial && j < k);
            rank += above;
            if (above || j == k) css += tj;
        }
        const double tt = (css - 1.0) / (double)(rank + 1);
        const int key = (act && trial - tt > 0.0) ? rank : -1;
        // theta: tt at the largest rank whose condition holds (0 when none holds)
        double theta = 0.0;
        int best = -1;
        for (int j = 0; j < m; ++j) {
            const int rj = __builtin_amdgcn_readlane(key, j);
            if (rj > best) {
                best = rj;
                theta = read_lane(tt, j);
            }
        }
        const double out = act ? fmax(trial - theta, 0.0) : 0.0;
        double delta = 0.0;
        for (int j = 0; j < m; ++j) delta += fabs(read_lane(out, j) - cj[j]);
        ck = out;
        if (delta < 1.0e-12) break;
    }
    if (act) c_out[k] = ck;
}

}  // namespace am1g
