// Dense symmetric eigensolver for one large matrix on the whole device.
//
// Replaces faer's self_adjoint_eigen behind am1-rs symmetric_eigen (linalg.rs:276-331) for
// orders beyond what one workgroup holds on chip: eigenvalues ascending, eigenvectors
// orthonormal. Three stages:
//   1. Householder tridiagonalization (lower, unblocked): per column, one pass over the trailing
//      rows that applies the previous column's rank-2 update and the new column's matrix-vector
//      product, then a one-workgroup kernel that builds the next reflector.
//   2. Tridiagonal divide and conquer (Cuppen's splitting with the deflation rules of
//      Dongarra-Sorensen and Gu-Eisenstat's Loewner correction for orthogonality). Leaves of at
//      most 64 are diagonalized by the workgroup Jacobi solver; each merge solves its secular
//      equation one wavefront per root and multiplies the children's eigenvectors by the
//      merge's eigenvector matrix with a library GEMM.
//   3. Back-transformation by the Householder reflectors in blocks of 64 (compact WY form,
//      three GEMMs per block).
// The deflation scan and the final ordering of each merge run on the host (O(m) each).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>
#include <numeric>

#include "am1g_common.hpp"
#include "am1g_eig.hpp"
#include "am1g_gemm.hpp"
#include "am1g_wg.hpp"

namespace am1g {
namespace {

constexpr int kLeaf = 64;  // largest tridiagonal block diagonalized directly
constexpr int kNb = 64;    // reflectors per back-transformation block
// ---------------------------------------------------------------- tridiagonalization
// Column c's reflector H_c = I - tau_c v_c v_c^T (v_c[0] = 1, indices relative to row c+1) maps
// the updated row c, A[c, c+1:n], to beta e_1; v_c[1:] is stored back into A[c, c+2:n].
// Per column c, one grid kernel does, for every trailing row i > c, the deferred rank-2 update
// of column c-1 (A -= v w^T + w v^T) followed at once by p_c[i] = tau_c A[i, c+1:] v_c, so the
// trailing block is read and written once per column; a one-workgroup kernel then forms w_c,
// applies the rank-2 update of column c to row c+1 and builds reflector c+1 from it.
__device__ __forceinline__ double block_sum_256(double v, double* red) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    const double tot = (red[0] + red[1]) + (red[2] + red[3]);
    __syncthreads();
    return tot;
}

// Reflector of row r (entries r..n-1 up to date), by one 256-thread workgroup.
__device__ void reflect_row(int n, int r, double* A, double* d, double* e, double* tau, double* v, double* red) {
    const int m = n - r - 1;
    double* x = A + (size_t)r * n + r + 1;
    double ss = 0.0;
    for (int i = 1 + threadIdx.x; i < m; i += 256) ss += x[i] * x[i];
    ss = block_sum_256(ss, red);
    const double alpha = x[0];
    double beta, t, scale;
    if (ss == 0.0) {
        beta = alpha;
        t = 0.0;
        scale = 0.0;
    } else {
        beta = -copysign(sqrt(alpha * alpha + ss), alpha);
        t = (beta - alpha) / beta;
        scale = 1.0 / (alpha - beta);
    }
    for (int i = 1 + threadIdx.x; i < m; i += 256) {
        const double vi = x[i] * scale;
        x[i] = vi;
        v[i] = vi;
    }
    if (threadIdx.x == 0) {
        v[0] = 1.0;
        d[r] = A[(size_t)r * n + r];
        e[r] = beta;
        tau[r] = t;
    }
}

__global__ __launch_bounds__(256) void k_reflect0(int n, double* A, double* d, double* e, double* tau, double* v) {
    __shared__ double red[4];
    reflect_row(n, 0, A, d, e, tau, v, red);
    if (threadIdx.x == 0 && n == 2) d[1] = A[(size_t)n + 1];
}

// Column c (0 <= c <= n-3): vb/wb hold v and p/w of columns c (slot c&1) and c-1 (other slot).
// Trailing rows i > c: deferred rank-2 update of column c-1, then p_c[i] = tau_c A[i, c+1:] v_c;
// one wavefront per row, four elements in flight per lane.
__global__ __launch_bounds__(256) void k_tridiag_rows(int n, int c, double* A, const double* tau, const double* vb,
                                                      double* wb) {
    const int m = n - c - 1;
    const int lane = threadIdx.x & 63;
    const int t = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (t >= m) return;
    const double* vc = vb + (size_t)(c & 1) * n;
    double* pc = wb + (size_t)(c & 1) * n;
    const double* vp = vb + (size_t)((c + 1) & 1) * n;  // column c-1 (indices relative to row c)
    const double* wp = wb + (size_t)((c + 1) & 1) * n;
    const int i = c + 1 + t;
    double* row = A + (size_t)i * n;
    double acc = 0.0;
    if (c > 0) {
        const double vi = vp[t + 1], wi = wp[t + 1];
        int j = c + 1 + lane;
        for (; j + 192 < n; j += 256) {
            double a[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) a[u] = row[j + 64 * u];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int jj = j + 64 * u;
                a[u] -= vi * wp[jj - c] + wi * vp[jj - c];
                row[jj] = a[u];
                acc += a[u] * vc[jj - c - 1];
            }
        }
        for (; j < n; j += 64) {
            const double a = row[j] - (vi * wp[j - c] + wi * vp[j - c]);
            row[j] = a;
            acc += a * vc[j - c - 1];
        }
    } else {
        for (int j = c + 1 + lane; j < n; j += 64) acc += row[j] * vc[j - c - 1];
    }
    for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
    if (lane == 0) pc[t] = tau[c] * acc;
}

// One workgroup: w_c = p_c - (tau_c / 2)(p_c . v_c) v_c, the rank-2 update of column c on row
// c+1 (and on the last diagonal element when c+1 = n-2), then reflector c+1.
__global__ __launch_bounds__(256) void k_tridiag_next(int n, int c, double* A, double* d, double* e, double* tau,
                                                      double* vb, double* wb) {
    __shared__ double red[4];
    const int m = n - c - 1;
    const double* vc = vb + (size_t)(c & 1) * n;
    double* pc = wb + (size_t)(c & 1) * n;
    double pv = 0.0;
    for (int q = threadIdx.x; q < m; q += 256) pv += pc[q] * vc[q];
    const double alpha = 0.5 * tau[c] * block_sum_256(pv, red);
    for (int q = threadIdx.x; q < m; q += 256) pc[q] -= alpha * vc[q];  // pc now holds w_c
    __syncthreads();
    const int r = c + 1;
    double* rowr = A + (size_t)r * n;
    const double v0 = vc[0], w0 = pc[0];
    for (int jj = r + threadIdx.x; jj < n; jj += 256) rowr[jj] -= v0 * pc[jj - r] + w0 * vc[jj - r];
    if (threadIdx.x == 0 && r == n - 2) d[n - 1] = A[(size_t)(n - 1) * n + n - 1] - 2.0 * vc[1] * pc[1];
    __syncthreads();
    reflect_row(n, r, A, d, e, tau, vb + (size_t)(r & 1) * n, red);
}

// k_tridiag_next for m = n - c - 1 <= 4096 with 1024 threads: each thread keeps its entries
// q = t + 1024 u of v_c, p_c and row c+1 in registers, so the kernel makes one pass over global
// memory and two block reductions (the 256-thread version makes five dependent passes).
constexpr int kNextT = 1024, kNextU = 4;

__device__ __forceinline__ double block_sum_1024(double v, double* red) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    __syncthreads();  // red may still be read from the previous sum
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    double tot = 0.0;
#pragma unroll
    for (int w = 0; w < kNextT / 64; ++w) tot += red[w];
    return tot;
}

__global__ __launch_bounds__(kNextT) void k_tridiag_next_wide(int n, int c, double* A, double* d, double* e,
                                                              double* tau, double* vb, double* wb) {
    __shared__ double red[kNextT / 64];
    __shared__ double x0;
    const int m = n - c - 1, t = threadIdx.x, r = c + 1;
    const double* vc = vb + (size_t)(c & 1) * n;
    double* pc = wb + (size_t)(c & 1) * n;
    double* rowr = A + (size_t)r * n + r;  // rowr[q] = A[r][r + q], q < m
    double vq[kNextU], pq[kNextU], aq[kNextU];
    const double v0 = vc[0], p0 = pc[0];  // read before the reduction's barriers; pc[0] is rewritten
    double pv = 0.0;
#pragma unroll
    for (int u = 0; u < kNextU; ++u) {
        const int q = t + kNextT * u;
        vq[u] = q < m ? vc[q] : 0.0;
        pq[u] = q < m ? pc[q] : 0.0;
        aq[u] = q < m ? rowr[q] : 0.0;
        pv += pq[u] * vq[u];
    }
    const double alpha = 0.5 * tau[c] * block_sum_1024(pv, red);
    // w_c = p_c - alpha v_c (the next column's deferred update reads it), then the rank-2 update
    // of column c on row c+1; every thread forms w_c[0] itself
    const double w0 = p0 - alpha * v0;
    double ss = 0.0;
#pragma unroll
    for (int u = 0; u < kNextU; ++u) {
        const int q = t + kNextT * u;
        if (q < m) {
            pq[u] -= alpha * vq[u];
            pc[q] = pq[u];
            aq[u] -= v0 * pq[u] + w0 * vq[u];
            if (q >= 2) ss += aq[u] * aq[u];
        }
    }
    if (t == 1) x0 = aq[0];  // A[r][r+1], the reflector's leading entry
    if (t == 1 && r == n - 2) d[n - 1] = A[(size_t)(n - 1) * n + n - 1] - 2.0 * vq[0] * pq[0];
    ss = block_sum_1024(ss, red);
    // reflector of row r, as reflect_row
    const double ax = x0;
    double beta, tt, scale;
    if (ss == 0.0) {
        beta = ax;
        tt = 0.0;
        scale = 0.0;
    } else {
        beta = -copysign(sqrt(ax * ax + ss), ax);
        tt = (beta - ax) / beta;
        scale = 1.0 / (ax - beta);
    }
    double* vr = vb + (size_t)(r & 1) * n;
#pragma unroll
    for (int u = 0; u < kNextU; ++u) {
        const int q = t + kNextT * u;
        if (q < m) {
            double val = aq[u];
            if (q >= 2) {
                val *= scale;
                vr[q - 1] = val;
            }
            rowr[q] = val;
        }
    }
    if (t == 0) {
        vr[0] = 1.0;
        d[r] = aq[0];
        e[r] = beta;
        tau[r] = tt;
    }
}

// ---------------------------------------------------------------- blocked tridiagonalization
// The same reduction in panels of kPanel columns (the dsytrd/dlatrd scheme; rows stand for
// columns since A is symmetric). Within a panel the trailing matrix stays as it was at the panel
// start (A_ps): the current matrix is A_ps - V W^T - W V^T over the panel's columns so far, with
// V the reflectors and W the vectors w_q = p_q - alpha_q v_q (column-major n x kPanel, global
// row indices). Column c's reflector is v_c = e_{c+1} + scale_c r, r the current row c past c+1.
// Per column two grid kernels and no other synchronization:
//   k_btri_symv  (one wavefront per trailing row x) finishes the sums the previous k_btri_row
//                left per workgroup (every workgroup the same way, so all agree): |r|^2, the
//                reflector, and (workgroup 0) s_V = V^T v_c, s_W = W^T v_c; y_x = (A_ps v_c)_x from
//                the unscaled r (the row loads need none of those sums), and v_x y_x per workgroup;
//   k_btri_row   (one thread per index x) alpha_c = tau_c^2 / 2 (v^T A_ps v - 2 s_V . s_W) from
//                those partials, p_x = tau_c (y_x - V_x s_W - W_x s_V), w_x = p_x - alpha_c v_x,
//                the current row c+1 (x >= c+1), and per workgroup |r|^2, V^T r, W^T r for the
//                next column.
// After each panel one rank-2k update of the trailing matrix (two matrix-core GEMMs). The
// trailing matrix is read once per column instead of read and written, and every cross-workgroup
// sum is a fixed-order sum of per-workgroup partials, so results do not depend on scheduling.
constexpr int kPanel = 32;
constexpr int kRowPart = 2 * kPanel + 2;  // per k_btri_row workgroup: V^T r, W^T r, |r|^2
constexpr int kRowWg = 128;

// Scalars of column c written by workgroup 0 of k_btri_symv: tau_c, scale_c, s_V, s_W.
enum BtriScal { kBsTau = 0, kBsScale = 1, kBsV = 8, kBsW = 8 + kPanel, kBsLen = 8 + 2 * kPanel };

// Reflector of row c from |r|^2, summed over the previous k_btri_row's workgroups in a fixed
// order by every wavefront that needs it (no barrier): returns tau, sets scale and beta.
__device__ __forceinline__ double btri_reflector(int c, int nb_prev, const double* rbuf, const double* rpart,
                                                 double& scale, double& beta) {
    const int lane = threadIdx.x & 63;
    double ss = 0.0;
    for (int b = lane; b < nb_prev; b += 64) ss += rpart[(size_t)b * kRowPart + 2 * kPanel];
    for (int off = 32; off > 0; off >>= 1) ss += __shfl_xor(ss, off, 64);
    const double ax = rbuf[c + 1];
    double tt;
    if (ss == 0.0) {
        beta = ax;
        tt = 0.0;
        scale = 0.0;
    } else {
        beta = -copysign(sqrt(ax * ax + ss), ax);
        tt = (beta - ax) / beta;
        scale = 1.0 / (ax - beta);
    }
    return tt;
}

// One wavefront per trailing row x in (c, n), four per workgroup. jj = c's column in its panel.
__global__ __launch_bounds__(256) void k_btri_symv(int n, int c, int jj, int nb_prev, double* A, const double* rbuf,
                                                   const double* rpart, double* V, const double* W, double* ybuf,
                                                   double* vpart, double* scal, double* d, double* e, double* tau) {
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int x = c + 1 + blockIdx.x * 4 + w;
    __shared__ double red[4];
    // the row's sums first: they need neither the reflector's scale nor the panel sums
    double yr = 0.0, ax = 0.0;
    if (x < n) {
        const double* row = A + (size_t)x * n;
        int j = c + 2 + lane;
        for (; j + 15 * 64 < n; j += 16 * 64) {
            double a[16], r[16];
#pragma unroll
            for (int u = 0; u < 16; ++u) {
                a[u] = row[j + 64 * u];
                r[u] = rbuf[j + 64 * u];
            }
#pragma unroll
            for (int u = 0; u < 16; ++u) yr = fma(a[u], r[u], yr);
        }
        for (; j < n; j += 64) yr = fma(row[j], rbuf[j], yr);
        ax = row[c + 1];
    }
    double scale, beta;
    const double tc = btri_reflector(c, nb_prev, rbuf, rpart, scale, beta);
    if (blockIdx.x == 0 && w == 0) {
        // s_V = V^T v_c, s_W = W^T v_c of v_c = e_{c+1} + scale r (the partials hold V^T r and
        // W^T r from row c+2 on), lane q < kPanel for V and kPanel + q for W
        double acc = 0.0;
        for (int b = 0; b < nb_prev; ++b) acc += rpart[(size_t)b * kRowPart + lane];
        const int q = lane < kPanel ? lane : lane - kPanel;
        const double* X = lane < kPanel ? V : W;
        const double sq = q < jj ? X[(size_t)q * n + c + 1] + scale * acc : 0.0;
        scal[(lane < kPanel ? kBsV : kBsW) + q] = sq;
        if (lane == 0) {
            scal[kBsTau] = tc;
            scal[kBsScale] = scale;
            d[c] = rbuf[c];
            e[c] = beta;
            tau[c] = tc;
        }
    }
    for (int off = 32; off > 0; off >>= 1) yr += __shfl_xor(yr, off, 64);
    double vy = 0.0;
    if (x < n) {
        const double y = scale * yr + ax;
        const double vx = x == c + 1 ? 1.0 : scale * rbuf[x];
        if (lane == 0) {
            ybuf[x] = y;
            V[(size_t)jj * n + x] = vx;
            if (x >= c + 2) A[(size_t)c * n + x] = vx;
        }
        vy = vx * y;
    }
    if (lane == 0) red[w] = vy;
    __syncthreads();
    if (t == 0) vpart[blockIdx.x] = (red[0] + red[1]) + (red[2] + red[3]);
}

// One thread per index x in [c+1, n) of row c+1 (first: c = -1, row 0 of A as it is). Every
// wavefront forms alpha_c and w_c at row c+1 itself (fixed-order sums, so all agree).
__global__ __launch_bounds__(kRowWg) void k_btri_row(int n, int c, int jj, int first, int nb_sym, const double* A,
                                                     const double* ybuf, const double* vpart, const double* scal,
                                                     const double* V, double* W, double* rbuf, double* rpart) {
    constexpr int kWaves = kRowWg / 64;
    const int r = c + 1;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    __shared__ double red[kWaves];
    __shared__ double wsum[kWaves][2 * kPanel];
    const int x = r + blockIdx.x * kRowWg + t;
    // every load first, without branches on the panel column
    double vq[kPanel], wq[kPanel];
#pragma unroll
    for (int q = 0; q < kPanel; ++q) {
        vq[q] = 0.0;
        wq[q] = 0.0;
    }
    double rx = 0.0, yx = 0.0;
    const bool in = x < n;
    if (in) {
        rx = A[(size_t)r * n + x];
        if (!first) {
            yx = ybuf[x];
#pragma unroll
            for (int q = 0; q < kPanel; ++q) {
                vq[q] = V[(size_t)q * n + x];
                wq[q] = W[(size_t)q * n + x];
            }
        }
    }
    double alpha = 0.0, tc = 0.0, w_r = 0.0, vrl = 0.0, wrl = 0.0, svl = 0.0, swl = 0.0;
    if (!first) {
        // lane q < kPanel holds column q's s_V, s_W and row c+1's V, W entries
        const int q = lane & (kPanel - 1);
        svl = q < jj ? scal[kBsV + q] : 0.0;
        swl = q < jj ? scal[kBsW + q] : 0.0;
        vrl = q < jj ? V[(size_t)q * n + r] : (q == jj ? 1.0 : 0.0);
        wrl = q < jj ? W[(size_t)q * n + r] : 0.0;
        tc = scal[kBsTau];
        const double yr = ybuf[r];
        double acc = 0.0;
        int b = lane;
        for (; b + 7 * 64 < nb_sym; b += 8 * 64) {
            double v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) v[u] = vpart[b + 64 * u];
#pragma unroll
            for (int u = 0; u < 8; ++u) acc += v[u];
        }
        for (; b < nb_sym; b += 64) acc += vpart[b];
        double svw = lane < kPanel ? svl * swl : 0.0;
        double cqr = lane < kPanel ? vrl * swl + wrl * svl : 0.0;
        for (int off = 32; off > 0; off >>= 1) {
            acc += __shfl_xor(acc, off, 64);
            svw += __shfl_xor(svw, off, 64);
            cqr += __shfl_xor(cqr, off, 64);
        }
        alpha = 0.5 * tc * tc * (acc - 2.0 * svw);
        w_r = tc * (yr - cqr) - alpha;  // w_c at row c+1, where v_c = 1
    }
#pragma unroll
    for (int q = 0; q < kPanel; ++q) {
        vq[q] = q <= jj ? vq[q] : 0.0;
        wq[q] = q < jj ? wq[q] : 0.0;
    }
    double ss = 0.0;
    if (!first) {
        // every lane takes part (the shuffles read lanes 0..kPanel-1); lanes past n hold zeros
        double cq = 0.0;
#pragma unroll
        for (int q = 0; q < kPanel; ++q) {
            const double svq = __shfl(svl, q, 64), swq = __shfl(swl, q, 64);
            cq += vq[q] * swq + wq[q] * svq;
        }
        double vx = 0.0;
#pragma unroll
        for (int q = 0; q < kPanel; ++q)
            if (q == jj) vx = vq[q];
        const double wx = x == r ? w_r : tc * (yx - cq) - alpha * vx;
        if (in) W[(size_t)jj * n + x] = wx;
        double rc = wx + w_r * vx;
#pragma unroll
        for (int q = 0; q < kPanel; ++q) {
            const double vrq = __shfl(vrl, q, 64), wrq = __shfl(wrl, q, 64);
            if (q < jj) rc += vrq * wq[q] + wrq * vq[q];
            if (q == jj) wq[q] = wx;
        }
        rx -= rc;
    }
    if (in) rbuf[x] = rx;
    if (in && x >= r + 2) {
        ss = rx * rx;
#pragma unroll
        for (int q = 0; q < kPanel; ++q) {
            vq[q] *= rx;
            wq[q] *= rx;
        }
    } else {
#pragma unroll
        for (int q = 0; q < kPanel; ++q) {
            vq[q] = 0.0;
            wq[q] = 0.0;
        }
    }
    // per-workgroup partials [V^T r | W^T r] (64 entries, recursive halving: lane l ends with
    // entry l), then across the wavefronts; and |r|^2
#pragma unroll
    for (int k = 0; k < kPanel; ++k) {
        const bool hi = (lane & 32) != 0;
        const double recv = __shfl_xor(hi ? vq[k] : wq[k], 32, 64);
        vq[k] = (hi ? wq[k] : vq[k]) + recv;
    }
#pragma unroll
    for (int step = 16; step >= 1; step >>= 1) {
#pragma unroll
        for (int k = 0; k < step; ++k) {
            const bool hi = (lane & step) != 0;
            const double recv = __shfl_xor(hi ? vq[k] : vq[k + step], step, 64);
            vq[k] = (hi ? vq[k + step] : vq[k]) + recv;
        }
    }
    for (int off = 32; off > 0; off >>= 1) ss += __shfl_xor(ss, off, 64);
    wsum[w][lane] = vq[0];
    if (lane == 0) red[w] = ss;
    __syncthreads();
    double* bp = rpart + (size_t)blockIdx.x * kRowPart;
    if (t < 2 * kPanel) {
        double v = 0.0;
        for (int k = 0; k < kWaves; ++k) v += wsum[k][t];
        bp[t] = v;
    }
    if (t == 2 * kPanel) {
        double v = 0.0;
        for (int k = 0; k < kWaves; ++k) v += red[k];
        bp[t] = v;
    }
    if (t == 2 * kPanel + 1) bp[t] = 0.0;
}

// After the last column (c = n-3): the last two diagonal entries and the last off-diagonal one,
// from row n-2 (rbuf) and the panel's updates of A[n-1][n-1].
__global__ __launch_bounds__(64) void k_btri_tail(int n, int jj, const double* A, const double* rbuf, const double* V,
                                                  const double* W, double* d, double* e, double* tau) {
    if (threadIdx.x != 0) return;
    d[n - 2] = rbuf[n - 2];
    e[n - 2] = rbuf[n - 1];
    tau[n - 2] = 0.0;
    double dl = A[(size_t)(n - 1) * n + n - 1];
    for (int q = 0; q <= jj; ++q) dl -= 2.0 * V[(size_t)q * n + n - 1] * W[(size_t)q * n + n - 1];
    d[n - 1] = dl;
}

// ---------------------------------------------------------------- divide and conquer
// Leaves: the tridiagonal block [s, s + m) with the split-modified diagonal, diagonalized by the
// workgroup Jacobi solver; eigenvectors go to the columns of Q (column-major, ld n) in ascending
// order of their eigenvalues, eigenvalues to lam[s..s+m).
__global__ __launch_bounds__(kWg) void k_leaf(int n, const int* leaf_s, const int* leaf_m, const double* d,
                                              const double* e, double* Q, double* lam, int* status) {
    __shared__ WgScratch sc;
    const int s = leaf_s[blockIdx.x], m = leaf_m[blockIdx.x];
    const int t = threadIdx.x;
    const int tri = m * (m + 1) / 2;
    for (int k = t; k < tri; k += kWg) {
        int i = (int)((sqrt(8.0 * (double)k + 1.0) - 1.0) * 0.5);
        while (i * (i + 1) / 2 > k) --i;
        while ((i + 1) * (i + 2) / 2 <= k) ++i;
        const int j = k - i * (i + 1) / 2;
        lds_dyn[k] = (i == j) ? d[s + i] : (i == j + 1 ? e[s + j] : 0.0);
    }
    for (int idx = t; idx < m * m; idx += kWg) lds_dyn[tri + idx] = (idx / m == idx % m) ? 1.0 : 0.0;
    __syncthreads();
    const int sweeps = wg_jacobi_lds(m, tri, m, nullptr, sc, 60, 0.0);
    __syncthreads();
    for (int i = t / 16; i < m; i += kWg / 16) {
        const double ei = lds_dyn[i * (i + 1) / 2 + i];
        int rank = 0;
        for (int j = t % 16; j < m; j += 16) {
            const double ej = lds_dyn[j * (j + 1) / 2 + j];
            rank += (ej < ei) || (ej == ei && j < i);
        }
        for (int off = 8; off > 0; off >>= 1) rank += __shfl_xor(rank, off, 16);
        if (t % 16 == 0) lam[s + rank] = ei;
        for (int c = t % 16; c < m; c += 16) Q[(size_t)(s + c) + (size_t)(s + rank) * n] = lds_dyn[tri + i * m + c];
    }
    if (t == 0 && sweeps < 0) atomicExch(status, 1);
}

// z = (last row of Q1, sg * first row of Q2) for the merge of [s, s+m1) and [s+m1, s+m).
__global__ __launch_bounds__(256) void k_gather_z(int n, const double* Q, int s, int m, int m1, double sg, double* z) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= m) return;
    z[i] = i < m1 ? Q[(size_t)(s + m1 - 1) + (size_t)(s + i) * n] : sg * Q[(size_t)(s + m1) + (size_t)(s + i) * n];
}

struct Rot {
    int a, b;       // local columns
    double c, s;    // col_a <- c col_a + s col_b ; col_b <- c col_b - s col_a
};

// Deflation rotations, in order, one thread per row of the merge block.
__global__ __launch_bounds__(256) void k_givens(int n, double* Q, int s, int m, int nrot, const Rot* rot) {
    const int r = blockIdx.x * 256 + threadIdx.x;
    if (r >= m) return;
    double* row = Q + (size_t)(s + r);
    for (int q = 0; q < nrot; ++q) {
        const Rot g = rot[q];
        const double x = row[(size_t)(s + g.a) * n], y = row[(size_t)(s + g.b) * n];
        row[(size_t)(s + g.a) * n] = g.c * x + g.s * y;
        row[(size_t)(s + g.b) * n] = g.c * y - g.s * x;
    }
}

__device__ __forceinline__ double wave_sum(double v) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Secular equation 1 + rho sum_i z_i^2 / (d_i - lambda) = 0 for the roots of D + rho z z^T
// (rho > 0, d strictly increasing): root j lies in (d_j, d_{j+1}), the last in
// (d_{k-1}, d_{k-1} + rho |z|^2]. Each root is kept as origin + offset (lambda = d_o + tau with
// o the nearer pole), so differences to the poles stay accurate. Bisection on the offset, one
// wavefront per root.
__global__ __launch_bounds__(256) void k_secular(int k, const double* dk, const double* zk, double rho, double znorm2,
                                                 int* orig, double* toff) {
    const int j = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (j >= k) return;
    auto g = [&](int o, double tau) {
        double acc = 0.0;
        for (int i = lane; i < k; i += 64) acc += zk[i] * zk[i] / ((dk[i] - dk[o]) - tau);
        return 1.0 + rho * wave_sum(acc);
    };
    int o;
    double lo, hi;
    if (j < k - 1) {
        const double half = 0.5 * (dk[j + 1] - dk[j]);
        if (g(j, half) >= 0.0) {
            o = j;
            lo = 0.0;
            hi = half;
        } else {
            o = j + 1;
            lo = -half;
            hi = 0.0;
        }
    } else {
        o = j;
        lo = 0.0;
        hi = rho * znorm2;
    }
    for (int it = 0; it < 300; ++it) {
        const double mid = 0.5 * (lo + hi);
        if (mid <= lo || mid >= hi) break;
        if (g(o, mid) > 0.0) hi = mid;
        else lo = mid;
    }
    if (lane == 0) {
        orig[j] = o;
        toff[j] = 0.5 * (lo + hi);
    }
}

// Loewner: zhat_i^2 = (lambda_{k-1} - d_i)/rho * prod_{j<i} (lambda_j - d_i)/(d_j - d_i)
//                     * prod_{i<=j<k-1} (lambda_j - d_i)/(d_{j+1} - d_i), with the sign of z_i.
__global__ __launch_bounds__(256) void k_lowner(int k, const double* dk, const double* zk, double rho, const int* orig,
                                                const double* toff, double* zhat) {
    const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (i >= k) return;
    double prod = 1.0;
    for (int j = lane; j < k - 1; j += 64) {
        const double lj = (dk[orig[j]] - dk[i]) + toff[j];
        prod *= j < i ? lj / (dk[j] - dk[i]) : lj / (dk[j + 1] - dk[i]);
    }
    for (int off = 32; off > 0; off >>= 1) prod *= __shfl_xor(prod, off, 64);
    if (lane == 0) {
        const double last = ((dk[orig[k - 1]] - dk[i]) + toff[k - 1]) / rho;
        const double z2 = fabs(prod * last);
        zhat[i] = copysign(sqrt(z2), zk[i]);
    }
}

// Eigenvector j of D + rho zhat zhat^T: u_i = zhat_i / (d_i - lambda_j), normalized; U is k x k
// column-major.
__global__ __launch_bounds__(256) void k_build_u(int k, const double* dk, const double* zhat, const int* orig,
                                                 const double* toff, double* U, double* lamk) {
    const int j = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (j >= k) return;
    const int o = orig[j];
    const double tj = toff[j];
    double ss = 0.0;
    for (int i = lane; i < k; i += 64) {
        const double u = zhat[i] / ((dk[i] - dk[o]) - tj);
        U[(size_t)i + (size_t)j * k] = u;
        ss += u * u;
    }
    ss = wave_sum(ss);
    const double inv = 1.0 / sqrt(ss);
    for (int i = lane; i < k; i += 64) U[(size_t)i + (size_t)j * k] *= inv;
    if (lane == 0) lamk[j] = dk[o] + tj;
}

// W[:, j] = Q[s:s+m, s + col[j]] (m x k, column-major).
__global__ __launch_bounds__(256) void k_gather_cols(int n, const double* Q, int s, int m, int k, const int* col,
                                                     double* W) {
    const int r = blockIdx.x * 256 + threadIdx.x;
    const int j = blockIdx.y;
    if (r >= m || j >= k) return;
    W[(size_t)r + (size_t)j * m] = Q[(size_t)(s + r) + (size_t)(s + col[j]) * n];
}

// Qn[:, c] = src[c] >= 0 ? Qt[:, src[c]] : Q[s:s+m, s + (-1 - src[c])]  (m x m, column-major).
__global__ __launch_bounds__(256) void k_assemble(int n, const double* Q, const double* Qt, int s, int m,
                                                  const int* src, double* Qn) {
    const int r = blockIdx.x * 256 + threadIdx.x;
    const int c = blockIdx.y;
    if (r >= m) return;
    const int sc = src[c];
    Qn[(size_t)r + (size_t)c * m] = sc >= 0 ? Qt[(size_t)r + (size_t)sc * m] : Q[(size_t)(s + r) + (size_t)(s - 1 - sc) * n];
}

__global__ __launch_bounds__(256) void k_copy_block(int n, double* Q, int s, int m, const double* Qn) {
    const int r = blockIdx.x * 256 + threadIdx.x;
    const int c = blockIdx.y;
    if (r >= m) return;
    Q[(size_t)(s + r) + (size_t)(s + c) * n] = Qn[(size_t)r + (size_t)c * m];
}

// ---------------------------------------------------------------- back-transformation
// Vb (rows r0 = k0+1 .. n-1, nb columns, column-major): reflector k0+j has 1 at row k0+j+1 and
// the stored A[k0+j, i] below it.
__global__ __launch_bounds__(256) void k_build_v(int n, int k0, int nb, const double* A, double* Vb) {
    const int r = blockIdx.x * 256 + threadIdx.x;
    const int j = blockIdx.y;
    const int r0 = k0 + 1, mr = n - r0;
    if (r >= mr || j >= nb) return;
    const int row = r0 + r, k = k0 + j;
    double v = 0.0;
    if (row == k + 1) v = 1.0;
    else if (row > k + 1) v = A[(size_t)k * n + row];
    Vb[(size_t)r + (size_t)j * mr] = v;
}

// Compact WY factor T (nb x nb upper triangular, column-major) from G = Vb^T Vb (forward,
// columnwise), by recursive doubling: T_jj = tau_j, and two adjacent diagonal blocks A, B of a
// finished level combine as T_AB = -T_AA G_AB T_BB (the product (I - V_A T_AA V_A^T)(I - V_B T_BB
// V_B^T) = I - [V_A V_B] [[T_AA, T_AB], [0, T_BB]] [V_A V_B]^T). log2(nb) levels of two small
// products each instead of nb dependent column steps.
__global__ __launch_bounds__(256) void k_larft(int nb, int k0, const double* tau, const double* G, double* T) {
    __shared__ double sT[kNb * kNb], sG[kNb * kNb], sX[kNb * kNb];
    const int t = threadIdx.x;
    for (int idx = t; idx < nb * nb; idx += 256) {
        const int i = idx % nb, j = idx / nb;
        sT[idx] = i == j ? tau[k0 + j] : 0.0;
        sG[idx] = G[idx];
    }
    __syncthreads();
    for (int sz = 1; sz < nb; sz *= 2) {
        // X_AB = G_AB T_BB for every pair (A = [i0, i0+sz), B = [i0+sz, i0+2sz) cut at nb)
        for (int idx = t; idx < nb * nb; idx += 256) {
            const int a = idx % nb, b = idx / nb;
            const int i0 = a / (2 * sz) * (2 * sz), b0 = i0 + sz;
            if (a >= b0 || b < b0 || b >= min(b0 + sz, nb)) continue;
            double acc = 0.0;
            for (int l = b0; l <= b; ++l) acc += sG[a + l * nb] * sT[l + b * nb];
            sX[idx] = acc;
        }
        __syncthreads();
        // T_AB = -T_AA X_AB
        for (int idx = t; idx < nb * nb; idx += 256) {
            const int a = idx % nb, b = idx / nb;
            const int i0 = a / (2 * sz) * (2 * sz), b0 = i0 + sz;
            if (a >= b0 || b < b0 || b >= min(b0 + sz, nb)) continue;
            double acc = 0.0;
            for (int l = a; l < b0; ++l) acc += sT[a + l * nb] * sX[l + b * nb];
            sT[idx] = -acc;
        }
        __syncthreads();
    }
    for (int idx = t; idx < nb * nb; idx += 256) T[idx] = sT[idx];
}

inline unsigned blocks(int64_t n, int per) { return (unsigned)((n + per - 1) / per); }

#define EIG_TRY(x)                                   \
    do {                                             \
        hipError_t e_ = (x);                         \
        if (e_ != hipSuccess) return e_;             \
    } while (0)

struct Node {
    int s, m, m1;
};

void split(int s, int m, std::vector<Node>& merges, std::vector<int>& ls, std::vector<int>& lm) {
    if (m <= kLeaf) {
        ls.push_back(s);
        lm.push_back(m);
        return;
    }
    const int m1 = m / 2;
    split(s, m1, merges, ls, lm);
    split(s + m1, m - m1, merges, ls, lm);
    merges.push_back({s, m, m1});
}

}  // namespace

namespace {
// Number of eigenvalues of the symmetric tridiagonal T = (d, e) below x: the sign changes of the
// LDL^T pivots q_i = (d_i - x) - e_{i-1}^2 / q_{i-1}, a pivot below pivmin in size taken as
// -pivmin (as in LAPACK's dstebz).
int sturm_count(const std::vector<double>& d, const std::vector<double>& e, int n, double x, double pivmin) {
    int c = 0;
    double q = d[0] - x;
    if (std::fabs(q) < pivmin) q = -pivmin;
    if (q < 0.0) ++c;
    for (int i = 1; i < n; ++i) {
        q = (d[i] - x) - e[i - 1] * e[i - 1] / q;
        if (std::fabs(q) < pivmin) q = -pivmin;
        if (q < 0.0) ++c;
    }
    return c;
}
}  // namespace

hipError_t eig_large(EigWork& w, int n, double* A, double* evals, double* Ct, hipStream_t st,
                     std::string& err) {
    return eig_large_impl(w, n, A, evals, Ct, st, err, -1, -1, nullptr);
}

hipError_t eig_large_values(EigWork& w, int n, double* A, int k_lo, int k_hi, double* vals, hipStream_t st,
                            std::string& err) {
    return eig_large_impl(w, n, A, nullptr, nullptr, st, err, k_lo, k_hi, vals);
}

hipError_t eig_large_impl(EigWork& w, int n, double* A, double* evals, double* Ct, hipStream_t st,
                          std::string& err, int k_lo, int k_hi, double* vals) {
    err.clear();
    if (n < 1) return hipSuccess;
    const size_t nn = (size_t)n * n;
    EIG_TRY(w.d.reserve(sizeof(double) * n));
    EIG_TRY(w.e.reserve(sizeof(double) * n));
    EIG_TRY(w.tau.reserve(sizeof(double) * n));
    EIG_TRY(w.vbuf.reserve(sizeof(double) * 2 * (size_t)n));
    EIG_TRY(w.pbuf.reserve(sizeof(double) * 2 * (size_t)n));
    EIG_TRY(w.scal.reserve(sizeof(int) * 16));
    EIG_TRY(w.Q.reserve(sizeof(double) * nn));
    double* d = w.d.as<double>();
    double* e = w.e.as<double>();
    double* tau = w.tau.as<double>();
    double* v = w.vbuf.as<double>();
    double* p = w.pbuf.as<double>();
    double* Q = w.Q.as<double>();

    // ---- 1. tridiagonalization: 2-3 n dependent kernels; captured once per order and buffer set
    // as a hipGraph and replayed (kernels only, so the graph holds no memset or copy nodes)
    const bool blocked = n >= std::max(w.blocked_min, 8);
    double *bV = nullptr, *bW = nullptr, *bybuf = nullptr, *bvpart = nullptr, *brow = nullptr, *brpart = nullptr,
           *bscal = nullptr;
    if (blocked) {
        EIG_TRY(w.bV.reserve(sizeof(double) * (size_t)n * kPanel));
        EIG_TRY(w.bW.reserve(sizeof(double) * (size_t)n * kPanel));
        EIG_TRY(w.bybuf.reserve(sizeof(double) * n));
        EIG_TRY(w.bvpart.reserve(sizeof(double) * ((size_t)n / 4 + 1)));
        EIG_TRY(w.brow.reserve(sizeof(double) * n));
        EIG_TRY(w.brpart.reserve(sizeof(double) * ((size_t)n / kRowWg + 1) * kRowPart));
        EIG_TRY(w.bscal.reserve(sizeof(double) * kBsLen));
        bV = w.bV.as<double>();
        bW = w.bW.as<double>();
        bybuf = w.bybuf.as<double>();
        bvpart = w.bvpart.as<double>();
        brow = w.brow.as<double>();
        brpart = w.brpart.as<double>();
        bscal = w.bscal.as<double>();
    }
    auto blocked_launches = [&]() {
        int nb_prev = (int)blocks(n, kRowWg);
        hipLaunchKernelGGL(k_btri_row, dim3(nb_prev), dim3(kRowWg), 0, st, n, -1, 0, 1, 0, A, bybuf, bvpart, bscal, bV,
                           bW, brow, brpart);
        int jj = 0;
        for (int c = 0; c + 2 < n; ++c) {
            jj = c % kPanel;
            const int m = n - c - 1;
            const int nb_sym = (int)blocks(m, 4);
            hipLaunchKernelGGL(k_btri_symv, dim3(nb_sym), dim3(256), 0, st, n, c, jj, nb_prev, A, brow, brpart, bV, bW,
                               bybuf, bvpart, bscal, d, e, tau);
            nb_prev = (int)blocks(m, kRowWg);
            hipLaunchKernelGGL(k_btri_row, dim3(nb_prev), dim3(kRowWg), 0, st, n, c, jj, 0, nb_sym, A, bybuf, bvpart,
                               bscal, bV, bW, brow, brpart);
            if (jj == kPanel - 1 && c + 3 < n) {
                // panel done: A[k1:, k1:] -= V W^T + W V^T (column-major view of the symmetric block)
                const int k1 = c + 1, mt = n - k1;
                double* At = A + (size_t)k1 * n + k1;
                (void)gemm_f64(st, false, true, mt, mt, kPanel, -1.0, bV + k1, n, bW + k1, n, 1.0, At, n);
                (void)gemm_f64(st, false, true, mt, mt, kPanel, -1.0, bW + k1, n, bV + k1, n, 1.0, At, n);
            }
        }
        hipLaunchKernelGGL(k_btri_tail, dim3(1), dim3(64), 0, st, n, jj, A, brow, bV, bW, d, e, tau);
    };
    auto tridiag_launches = [&]() {
        if (blocked) {
            blocked_launches();
            return;
        }
        hipLaunchKernelGGL(k_reflect0, dim3(1), dim3(256), 0, st, n, A, d, e, tau, v);
        for (int c = 0; c + 2 < n; ++c) {
            hipLaunchKernelGGL(k_tridiag_rows, dim3(blocks(n - c - 1, 4)), dim3(256), 0, st, n, c, A, tau, v, p);
            if (n - c - 1 <= kNextT * kNextU)
                hipLaunchKernelGGL(k_tridiag_next_wide, dim3(1), dim3(kNextT), 0, st, n, c, A, d, e, tau, v, p);
            else
                hipLaunchKernelGGL(k_tridiag_next, dim3(1), dim3(256), 0, st, n, c, A, d, e, tau, v, p);
        }
    };
    if (n == 1) {
        EIG_TRY(hipMemcpyAsync(d, A, sizeof(double), hipMemcpyDeviceToDevice, st));
    } else if (!w.use_graphs || n < 64) {
        tridiag_launches();
    } else {
        const void* key[16] = {A, d, e, tau, v, p, bV, bW, bybuf, bvpart, brow, brpart, bscal, nullptr, nullptr,
                               nullptr};
        TridiagGraph* hit = nullptr;
        for (auto& g : w.graphs)
            if (g.n == n && std::equal(key, key + 16, g.key)) hit = &g;
        if (!hit) {
            hipGraph_t graph = nullptr;
            EIG_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            tridiag_launches();
            const hipError_t ec = hipStreamEndCapture(st, &graph);
            if (ec != hipSuccess) return ec;
            hipGraphExec_t exec = nullptr;
            const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            if (ei != hipSuccess) return ei;
            if (w.graphs.size() >= EigWork::kMaxGraphs) {
                (void)hipGraphExecDestroy(w.graphs.front().exec);
                w.graphs.erase(w.graphs.begin());
            }
            TridiagGraph g;
            g.n = n;
            std::copy(key, key + 16, g.key);
            g.exec = exec;
            w.graphs.push_back(g);
            hit = &w.graphs.back();
        }
        EIG_TRY(hipGraphLaunch(hit->exec, st));
    }
    EIG_TRY(hipGetLastError());
    w.h_d.resize(n);
    w.h_e.resize(n);
    EIG_TRY(hipMemcpyAsync(w.h_d.data(), d, sizeof(double) * n, hipMemcpyDeviceToHost, st));
    if (n > 1) EIG_TRY(hipMemcpyAsync(w.h_e.data(), e, sizeof(double) * (n - 1), hipMemcpyDeviceToHost, st));
    EIG_TRY(hipStreamSynchronize(st));
    for (int i = 0; i < n; ++i)
        if (!std::isfinite(w.h_d[i]) || (i + 1 < n && !std::isfinite(w.h_e[i]))) {
            err = "symmetric eigensolver: non-finite matrix";
            return hipErrorUnknown;
        }
    if (vals) {
        // only the eigenvalues k_lo..k_hi: bisection on Sturm counts of T to rounding level
        double lo = INFINITY, hi = -INFINITY, tnorm = 0.0;
        for (int i = 0; i < n; ++i) {
            const double r = (i > 0 ? std::fabs(w.h_e[i - 1]) : 0.0) + (i + 1 < n ? std::fabs(w.h_e[i]) : 0.0);
            lo = std::min(lo, w.h_d[i] - r);
            hi = std::max(hi, w.h_d[i] + r);
            tnorm = std::max(tnorm, std::fabs(w.h_d[i]) + r);
        }
        const double pivmin = 1.0e-280 * std::max(1.0, tnorm * tnorm);
        for (int k = k_lo; k <= k_hi; ++k) {
            double a = lo - 1.0e-12 * (tnorm + 1.0), b = hi + 1.0e-12 * (tnorm + 1.0);
            for (int it = 0; it < 200; ++it) {
                const double m = 0.5 * (a + b);
                if (!(m > a && m < b)) break;  // the interval is down to adjacent doubles
                if (sturm_count(w.h_d, w.h_e, n, m, pivmin) > k) b = m;
                else a = m;
            }
            vals[k - k_lo] = 0.5 * (a + b);
        }
        return hipSuccess;
    }

    // ---- 2. tridiagonal divide and conquer
    std::vector<Node> merges;
    std::vector<int> ls, lm;
    split(0, n, merges, ls, lm);
    std::vector<double> hd = w.h_d;
    for (const Node& nd : merges) {  // T = diag(T1', T2') + |beta| u u^T
        const double b = std::fabs(w.h_e[nd.s + nd.m1 - 1]);
        hd[nd.s + nd.m1 - 1] -= b;
        hd[nd.s + nd.m1] -= b;
    }
    EIG_TRY(hipMemcpyAsync(d, hd.data(), sizeof(double) * n, hipMemcpyHostToDevice, st));
    const int nl = (int)ls.size();
    EIG_TRY(w.idx.reserve(sizeof(int) * std::max<size_t>(2 * nl + 4 * (size_t)n + 8, 64)));
    EIG_TRY(w.lam.reserve(sizeof(double) * 4 * (size_t)n));
    int* ileaf = w.idx.as<int>();
    double* lam = w.lam.as<double>();
    EIG_TRY(hipMemcpyAsync(ileaf, ls.data(), sizeof(int) * nl, hipMemcpyHostToDevice, st));
    EIG_TRY(hipMemcpyAsync(ileaf + nl, lm.data(), sizeof(int) * nl, hipMemcpyHostToDevice, st));
    EIG_TRY(hipMemsetAsync(Q, 0, sizeof(double) * nn, st));
    int* status = w.scal.as<int>();
    EIG_TRY(hipMemsetAsync(status, 0, sizeof(int), st));
    const size_t leaf_lds = sizeof(double) * (kLeaf * (kLeaf + 1) / 2 + kLeaf * kLeaf);
    EIG_TRY(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_leaf), hipFuncAttributeMaxDynamicSharedMemorySize,
                                (int)leaf_lds));
    hipLaunchKernelGGL(k_leaf, dim3(nl), dim3(kWg), leaf_lds, st, n, ileaf, ileaf + nl, d, e, Q, lam, status);
    EIG_TRY(hipGetLastError());
    w.h_lam.resize(n);
    int h_status = 0;
    EIG_TRY(hipMemcpyAsync(w.h_lam.data(), lam, sizeof(double) * n, hipMemcpyDeviceToHost, st));
    EIG_TRY(hipMemcpyAsync(&h_status, status, sizeof(int), hipMemcpyDeviceToHost, st));
    EIG_TRY(hipStreamSynchronize(st));
    if (h_status) {
        err = "symmetric eigensolver: a leaf block did not converge";
        return hipErrorUnknown;
    }

    if (!merges.empty()) {
        EIG_TRY(w.W.reserve(sizeof(double) * nn));
        EIG_TRY(w.Qt.reserve(sizeof(double) * nn));
        EIG_TRY(w.U.reserve(sizeof(double) * nn));
        EIG_TRY(w.Qn.reserve(sizeof(double) * nn));
        EIG_TRY(w.rot.reserve(sizeof(Rot) * n));
        EIG_TRY(w.flt.reserve(sizeof(double) * 6 * (size_t)n));
    }
    const double eps = DBL_EPSILON;
    std::vector<double> hz, ds, zs, dk, zk, lamk;
    std::vector<int> perm, kept, src;
    std::vector<Rot> rots;
    std::vector<std::pair<double, int>> fin;
    for (const Node& nd : merges) {
        const int s = nd.s, m = nd.m, m1 = nd.m1;
        const double beta = w.h_e[s + m1 - 1];
        const double rho = std::fabs(beta);
        double* zbuf = w.flt.as<double>();
        hipLaunchKernelGGL(k_gather_z, dim3(blocks(m, 256)), dim3(256), 0, st, n, Q, s, m, m1, beta >= 0.0 ? 1.0 : -1.0,
                           zbuf);
        hz.resize(m);
        EIG_TRY(hipMemcpyAsync(hz.data(), zbuf, sizeof(double) * m, hipMemcpyDeviceToHost, st));
        EIG_TRY(hipStreamSynchronize(st));
        double z2 = 0.0;
        for (double x : hz) z2 += x * x;
        const double zn = std::sqrt(z2);
        const double rho_eff = rho * z2;
        // sort the children's eigenvalues (each child's list is already ascending)
        perm.resize(m);
        std::iota(perm.begin(), perm.end(), 0);
        std::stable_sort(perm.begin(), perm.end(),
                         [&](int a, int b) { return w.h_lam[s + a] < w.h_lam[s + b]; });
        ds.resize(m);
        zs.resize(m);
        double dmax = 0.0, zmax = 0.0;
        for (int i = 0; i < m; ++i) {
            ds[i] = w.h_lam[s + perm[i]];
            zs[i] = zn > 0.0 ? hz[perm[i]] / zn : 0.0;
            dmax = std::max(dmax, std::fabs(ds[i]));
            zmax = std::max(zmax, std::fabs(zs[i]));
        }
        const double tol = 8.0 * eps * std::max(dmax, zmax);
        // deflation scan (small z components; nearly equal poles combined by a rotation)
        kept.clear();
        rots.clear();
        fin.clear();
        int pj = -1;
        for (int j = 0; j < m; ++j) {
            if (rho_eff * std::fabs(zs[j]) <= tol) {
                fin.push_back({ds[j], -1 - perm[j]});
                continue;
            }
            if (pj < 0) {
                pj = j;
                continue;
            }
            double sv = zs[pj], cv = zs[j];
            const double tv = std::hypot(cv, sv);
            const double gap = ds[j] - ds[pj];
            cv /= tv;
            sv = -sv / tv;
            if (std::fabs(gap * cv * sv) <= tol) {
                zs[j] = tv;
                zs[pj] = 0.0;
                rots.push_back({perm[pj], perm[j], cv, sv});
                const double t1 = ds[pj] * cv * cv + ds[j] * sv * sv;
                ds[j] = ds[pj] * sv * sv + ds[j] * cv * cv;
                ds[pj] = t1;
                fin.push_back({ds[pj], -1 - perm[pj]});
                pj = j;
            } else {
                kept.push_back(pj);
                pj = j;
            }
        }
        if (pj >= 0) kept.push_back(pj);
        if (!rots.empty()) {
            Rot* drot = w.rot.as<Rot>();
            EIG_TRY(hipMemcpyAsync(drot, rots.data(), sizeof(Rot) * rots.size(), hipMemcpyHostToDevice, st));
            hipLaunchKernelGGL(k_givens, dim3(blocks(m, 256)), dim3(256), 0, st, n, Q, s, m, (int)rots.size(), drot);
        }
        const int k = (int)kept.size();
        if (k > 0) {
            dk.resize(k);
            zk.resize(k);
            std::vector<int> colk(k);
            double zk2 = 0.0;
            for (int j = 0; j < k; ++j) {
                dk[j] = ds[kept[j]];
                zk[j] = zs[kept[j]];
                colk[j] = perm[kept[j]];
                zk2 += zk[j] * zk[j];
            }
            double* f = w.flt.as<double>();
            double* ddk = f;
            double* dzk = f + n;
            double* dtoff = f + 2 * (size_t)n;
            double* dzhat = f + 3 * (size_t)n;
            double* dlamk = f + 4 * (size_t)n;
            int* iorig = w.idx.as<int>() + 2 * nl;
            int* icol = iorig + n;
            int* isrc = icol + n;
            EIG_TRY(hipMemcpyAsync(ddk, dk.data(), sizeof(double) * k, hipMemcpyHostToDevice, st));
            EIG_TRY(hipMemcpyAsync(dzk, zk.data(), sizeof(double) * k, hipMemcpyHostToDevice, st));
            EIG_TRY(hipMemcpyAsync(icol, colk.data(), sizeof(int) * k, hipMemcpyHostToDevice, st));
            hipLaunchKernelGGL(k_secular, dim3(blocks(k, 4)), dim3(256), 0, st, k, ddk, dzk, rho_eff, zk2, iorig, dtoff);
            hipLaunchKernelGGL(k_lowner, dim3(blocks(k, 4)), dim3(256), 0, st, k, ddk, dzk, rho_eff, iorig, dtoff,
                               dzhat);
            double* U = w.U.as<double>();
            hipLaunchKernelGGL(k_build_u, dim3(blocks(k, 4)), dim3(256), 0, st, k, ddk, dzhat, iorig, dtoff, U, dlamk);
            double* W = w.W.as<double>();
            double* Qt = w.Qt.as<double>();
            hipLaunchKernelGGL(k_gather_cols, dim3(blocks(m, 256), k), dim3(256), 0, st, n, Q, s, m, k, icol, W);
            EIG_TRY(hipGetLastError());
            const double one = 1.0, zero = 0.0;
            EIG_TRY(gemm_f64(st, false, false, m, k, k, one, W, m, U, k, zero, Qt, m));
            lamk.resize(k);
            EIG_TRY(hipMemcpyAsync(lamk.data(), dlamk, sizeof(double) * k, hipMemcpyDeviceToHost, st));
            EIG_TRY(hipStreamSynchronize(st));
            for (int j = 0; j < k; ++j) {
                if (!std::isfinite(lamk[j])) {
                    err = "symmetric eigensolver: secular equation failed";
                    return hipErrorUnknown;
                }
                fin.push_back({lamk[j], j});
            }
            std::stable_sort(fin.begin(), fin.end(),
                             [](const std::pair<double, int>& a, const std::pair<double, int>& b) {
                                 return a.first < b.first;
                             });
            src.resize(m);
            for (int c = 0; c < m; ++c) {
                src[c] = fin[c].second;
                w.h_lam[s + c] = fin[c].first;
            }
            EIG_TRY(hipMemcpyAsync(isrc, src.data(), sizeof(int) * m, hipMemcpyHostToDevice, st));
            double* Qn = w.Qn.as<double>();
            hipLaunchKernelGGL(k_assemble, dim3(blocks(m, 256), m), dim3(256), 0, st, n, Q, Qt, s, m, isrc, Qn);
            hipLaunchKernelGGL(k_copy_block, dim3(blocks(m, 256), m), dim3(256), 0, st, n, Q, s, m, Qn);
            EIG_TRY(hipGetLastError());
        } else {
            // fully deflated: the block is already diagonal in the rotated columns; order them
            std::stable_sort(fin.begin(), fin.end(),
                             [](const std::pair<double, int>& a, const std::pair<double, int>& b) {
                                 return a.first < b.first;
                             });
            src.resize(m);
            for (int c = 0; c < m; ++c) {
                src[c] = fin[c].second;
                w.h_lam[s + c] = fin[c].first;
            }
            int* isrc = w.idx.as<int>() + 2 * nl + 2 * n;
            EIG_TRY(hipMemcpyAsync(isrc, src.data(), sizeof(int) * m, hipMemcpyHostToDevice, st));
            double* Qn = w.Qn.as<double>();
            hipLaunchKernelGGL(k_assemble, dim3(blocks(m, 256), m), dim3(256), 0, st, n, Q, Q, s, m, isrc, Qn);
            hipLaunchKernelGGL(k_copy_block, dim3(blocks(m, 256), m), dim3(256), 0, st, n, Q, s, m, Qn);
            EIG_TRY(hipGetLastError());
            EIG_TRY(hipStreamSynchronize(st));
        }
    }

    // ---- 3. back-transformation: Z <- H_0 ... H_{n-3} Z, last block first
    if (n > 2) {
        EIG_TRY(w.V.reserve(sizeof(double) * (size_t)n * kNb));
        EIG_TRY(w.Tm.reserve(sizeof(double) * kNb * kNb));
        EIG_TRY(w.G.reserve(sizeof(double) * kNb * kNb));
        EIG_TRY(w.X.reserve(sizeof(double) * (size_t)n * kNb));
        EIG_TRY(w.Y.reserve(sizeof(double) * (size_t)n * kNb));
        const int nref = n - 2;  // reflectors 0 .. n-3 (the one of column n-2 is the identity)
        const int nblk = (nref + kNb - 1) / kNb;
        double* Vb = w.V.as<double>();
        double* T = w.Tm.as<double>();
        double* G = w.G.as<double>();
        double* X = w.X.as<double>();
        double* Y = w.Y.as<double>();
        const double one = 1.0, zero = 0.0, mone = -1.0;
        for (int b = nblk - 1; b >= 0; --b) {
            const int k0 = b * kNb;
            const int nb = std::min(kNb, nref - k0);
            const int r0 = k0 + 1, mr = n - r0;
            hipLaunchKernelGGL(k_build_v, dim3(blocks(mr, 256), nb), dim3(256), 0, st, n, k0, nb, A, Vb);
            EIG_TRY(hipGetLastError());
            EIG_TRY(w.gw.reserve(sizeof(double) * gemm_splitk_elems(nb, nb, mr)));
            EIG_TRY(gemm_f64(st, true, false, nb, nb, mr, one, Vb, mr, Vb, mr, zero, G, nb, w.gw.as<double>(),
                             w.gw.bytes / sizeof(double)));
            hipLaunchKernelGGL(k_larft, dim3(1), dim3(256), 0, st, nb, k0, tau, G, T);
            EIG_TRY(hipGetLastError());
            EIG_TRY(w.gw.reserve(sizeof(double) * gemm_splitk_elems(nb, n, mr)));
            EIG_TRY(gemm_f64(st, true, false, nb, n, mr, one, Vb, mr, Q + r0, n, zero, X, nb, w.gw.as<double>(),
                             w.gw.bytes / sizeof(double)));
            EIG_TRY(gemm_f64(st, false, false, nb, n, nb, one, T, nb, X, nb, zero, Y, nb));
            EIG_TRY(gemm_f64(st, false, false, mr, n, nb, mone, Vb, mr, Y, nb, one, Q + r0, n));
        }
    }
    EIG_TRY(hipMemcpyAsync(evals, w.h_lam.data(), sizeof(double) * n, hipMemcpyHostToDevice, st));
    EIG_TRY(hipMemcpyAsync(Ct, Q, sizeof(double) * nn, hipMemcpyDeviceToDevice, st));
    EIG_TRY(hipStreamSynchronize(st));
    return hipSuccess;
}

}  // namespace am1g
