// Workgroup-level building blocks of the SCF: deterministic reductions, a register-tiled GEMM,
// the parallel cyclic Jacobi eigensolver and the DIIS small solves. Every routine is called by
// all threads of one workgroup and orders its floating-point sums independently of what other
// workgroups are doing, so a molecule's result does not depend on the batch it runs in.
#pragma once
#include <hip/hip_runtime.h>

#include "am1g_common.hpp"

namespace am1g {

constexpr int kWg = 256;  // threads per workgroup for the per-molecule SCF

struct WgScratch {
    double red[kWg];      // reduction scratch
    double bcast[16];     // broadcast slots
    double jc[128], js[128];  // Jacobi rotations of one round
    int jp[128], jq[128];
};

// Deterministic sum over the workgroup: fixed-shape tree over thread slots.
__device__ __forceinline__ double wg_sum(double v, WgScratch& s) {
    const int t = threadIdx.x;
    // wavefront-level butterfly first (fixed order), then across the four wavefronts
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    __syncthreads();
    if ((t & 63) == 0) s.red[t >> 6] = v;
    __syncthreads();
    double tot = (s.red[0] + s.red[1]) + (s.red[2] + s.red[3]);
    __syncthreads();
    return tot;
}

__device__ __forceinline__ double wg_max(double v, WgScratch& s) {
    const int t = threadIdx.x;
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    __syncthreads();
    if ((t & 63) == 0) s.red[t >> 6] = v;
    __syncthreads();
    double tot = fmax(fmax(s.red[0], s.red[1]), fmax(s.red[2], s.red[3]));
    __syncthreads();
    return tot;
}

// Several independent sums at once, reduced through `buf` (4 * nv doubles of shared scratch).
// Every thread receives all nv totals in out[0..nv).
__device__ void wg_sum_many(const double* v, int nv, double* buf, double* out) {
    const int t = threadIdx.x;
    for (int i = 0; i < nv; ++i) {
        double x = v[i];
        for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off, 64);
        if ((t & 63) == 0) buf[i * 4 + (t >> 6)] = x;
    }
    __syncthreads();
    for (int i = 0; i < nv; ++i) out[i] = (buf[i * 4 + 0] + buf[i * 4 + 1]) + (buf[i * 4 + 2] + buf[i * 4 + 3]);
    __syncthreads();
}

// C[i][j] = alpha * sum_k opA(i,k) * opB(k,j), with opA(i,k) = TA ? A[k*lda+i] : A[i*lda+k] and
// opB(k,j) = TB ? B[j*ldb+k] : B[k*ldb+j]. Each thread owns 4x4 output tiles; operands are
// read through the cache hierarchy (they are the molecule's own, L2-resident matrices).
template <typename T, bool TA, bool TB>
__device__ void wg_gemm(int M, int N, int K, double alpha, const double* __restrict__ A, int lda,
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

// Round-robin partner schedule for m (even) players: round r, slot k -> (p, q), p < q.
__device__ __forceinline__ void rr_pair(int m, int r, int k, int& p, int& q) {
    const int L = m - 1;
    int a, b;
    if (k == 0) {
        a = 0;
        b = (r % L) + 1;
    } else {
        a = ((r + k) % L) + 1;
        b = ((r - k + L) % L) + 1;
    }
    p = min(a, b);
    q = max(a, b);
}

// Parallel cyclic Jacobi on the symmetric n x n matrix A (row stride lda; LDS or global),
// rotating the rows of Vt (Vt = V^T, row k = eigenvector k) along with it. On return the
// eigenvalues are on A's diagonal (unsorted). Each round applies m/2 disjoint rotations: one
// thread per 2x2 block of the (row pair, column pair) grid updates it in place, which is
// race-free because a block's new values depend only on the block itself.
// Returns the number of sweeps, or -1 when it did not converge in max_sweeps.
__device__ int wg_jacobi(int n, double* A, int lda, double* Vt, int ldv, WgScratch& s, int max_sweeps) {
    const int t = threadIdx.x;
    const int m = (n + 1) & ~1;
    const int h = m >> 1;
    const int nblk = h * (h + 1) / 2;
    for (int sweep = 0; sweep <= max_sweeps; ++sweep) {
        // convergence: off-diagonal Frobenius norm against the diagonal's
        double off = 0.0, dg = 0.0;
        for (int idx = t; idx < n * n; idx += kWg) {
            int i = idx / n, j = idx - i * n;
            double v = A[(size_t)i * lda + j];
            if (i == j) dg += v * v;
            else off += v * v;
        }
        off = wg_sum(off, s);
        dg = wg_sum(dg, s);
        if (off <= 1.0e-28 * dg || off == 0.0) return sweep;
        if (sweep == max_sweeps) break;
        const double skip = 1.0e-15 * sqrt(dg / (double)max(n, 1));
        for (int r = 0; r < m - 1; ++r) {
            if (t < h) {
                int p, q;
                rr_pair(m, r, t, p, q);
                double c = 1.0, sn = 0.0;
                if (q < n) {
                    double apq = A[(size_t)p * lda + q];
                    if (fabs(apq) > skip) {
                        double app = A[(size_t)p * lda + p], aqq = A[(size_t)q * lda + q];
                        double theta = (aqq - app) / (2.0 * apq);
                        double tt = (theta >= 0.0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                        c = 1.0 / sqrt(tt * tt + 1.0);
                        sn = tt * c;
                    }
                }
                s.jc[t] = c;
                s.js[t] = sn;
                s.jp[t] = p;
                s.jq[t] = q;
            }
            __syncthreads();
            // A <- J^T A J over (a >= b) blocks, mirrored
            for (int blk = t; blk < nblk; blk += kWg) {
                int a = (int)((sqrt(8.0 * blk + 1.0) - 1.0) * 0.5);
                while (a * (a + 1) / 2 > blk) --a;
                while ((a + 1) * (a + 2) / 2 <= blk) ++a;
                int b = blk - a * (a + 1) / 2;
                int pa = s.jp[a], qa = s.jq[a], pb = s.jp[b], qb = s.jq[b];
                double ca = s.jc[a], sa = s.js[a], cb = s.jc[b], sb = s.js[b];
                bool va = qa < n, vb = qb < n;  // padding player never touches memory
                if (a == b) {
                    if (!va || sa == 0.0) continue;
                    double x00 = A[(size_t)pa * lda + pa], x01 = A[(size_t)pa * lda + qa];
                    double x11 = A[(size_t)qa * lda + qa];
                    // J^T X J for the symmetric 2x2 block
                    double y00 = ca * (ca * x00 - sa * x01) - sa * (ca * x01 - sa * x11);
                    double y11 = sa * (sa * x00 + ca * x01) + ca * (sa * x01 + ca * x11);
                    double y01 = ca * (sa * x00 + ca * x01) - sa * (sa * x01 + ca * x11);
                    A[(size_t)pa * lda + pa] = y00;
                    A[(size_t)qa * lda + qa] = y11;
                    A[(size_t)pa * lda + qa] = y01;
                    A[(size_t)qa * lda + pa] = y01;
                } else {
                    // rows {pa, qa} (or the single row pa when a is the padding slot) x cols {pb, qb}
                    if ((!va || sa == 0.0) && (!vb || sb == 0.0)) continue;
                    double x00 = A[(size_t)pa * lda + pb];
                    double x01 = vb ? A[(size_t)pa * lda + qb] : 0.0;
                    double x10 = va ? A[(size_t)qa * lda + pb] : 0.0;
                    double x11 = (va && vb) ? A[(size_t)qa * lda + qb] : 0.0;
                    if (!va) { ca = 1.0; sa = 0.0; }
                    if (!vb) { cb = 1.0; sb = 0.0; }
                    // rows: J_a^T = [[ca, -sa], [sa, ca]]
                    double r00 = ca * x00 - sa * x10, r01 = ca * x01 - sa * x11;
                    double r10 = sa * x00 + ca * x10, r11 = sa * x01 + ca * x11;
                    // cols: J_b = [[cb, sb], [-sb, cb]]
                    double y00 = r00 * cb - r01 * sb, y01 = r00 * sb + r01 * cb;
                    double y10 = r10 * cb - r11 * sb, y11 = r10 * sb + r11 * cb;
                    A[(size_t)pa * lda + pb] = y00;
                    A[(size_t)pb * lda + pa] = y00;
                    if (vb) {
                        A[(size_t)pa * lda + qb] = y01;
                        A[(size_t)qb * lda + pa] = y01;
                    }
                    if (va) {
                        A[(size_t)qa * lda + pb] = y10;
                        A[(size_t)pb * lda + qa] = y10;
                    }
                    if (va && vb) {
                        A[(size_t)qa * lda + qb] = y11;
                        A[(size_t)qb * lda + qa] = y11;
                    }
                }
            }
            // Vt rows p, q: (v_p, v_q) <- (c v_p - s v_q, s v_p + c v_q)
            for (int idx = t; idx < h * n; idx += kWg) {
                int a = idx / n, i = idx - a * n;
                double sa = s.js[a];
                if (s.jq[a] >= n || sa == 0.0) continue;
                double ca = s.jc[a];
                int p = s.jp[a], q = s.jq[a];
                double vp = Vt[(size_t)p * ldv + i], vq = Vt[(size_t)q * ldv + i];
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
__device__ bool solve_bordered(double* m, double* rhs, int n, double* x) {
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
__device__ void simplex_project(const double* v, int n, double* out) {
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
__device__ void adiis_solve(const double* g, int m, double* c) {
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

}  // namespace am1g
