// Workgroup symmetric eigensolver for orders up to kTriMax, one 512-thread workgroup per matrix:
//   1. Householder tridiagonalization in LDS. Step k builds reflector k from column k (every
//      wavefront computes it, so no barrier is needed to share it) and then, in one pass over
//      the trailing rows, applies the previous step's rank-2 update and forms the product for
//      the new reflector: one barrier and one read-modify-write of the trailing block per step.
//      The reflectors go to global scratch for step 5.
//   2. The tridiagonal T is split where an off-diagonal entry is below 1e-10 ||T||.
//   3. Eigenvalues of each unreduced block by multisection on Sturm counts (signs of the
//      leading principal minors), a group of threads per eigenvalue, two points per thread.
//   4. Eigenvectors of T from the twisted factorization at each eigenvalue (one thread per
//      eigenvector, O(n)); runs of eigenvalues closer than 1e-6 ||T|| inside one block take
//      distinct twist indices and are re-orthogonalized by Gram-Schmidt.
//   5. Sort, and back transformation x = H_0 ... H_{n-3} z row by row, the reflectors staged
//      through LDS.
// It replaces the Jacobi rounds of the reference eigensolver (linalg.rs) for the orders that fit
// in LDS: a Jacobi sweep needs n-1 barrier-separated rounds over the whole matrix; this needs
// about n barriers in all. The per-thread sequential loops load their operands four at a time,
// so LDS latency is paid once per four steps.
#pragma once
#include <cfloat>

#include "am1g_wg.hpp"

namespace am1g {

constexpr int kTriMax = 128;

__device__ __forceinline__ int tri_ld(int n) { return n | 1; }  // odd row stride: rows spread over LDS banks

// LDS layout for order n (in doubles): the matrix (n x ld, later the eigenvectors of T), eleven
// n-vectors (d, e, tau, lam, two reflectors, two products, block starts and ends as ints), then
// the (d_i, e_{i-1}^2) pairs at an even offset.
__host__ __device__ __forceinline__ int tri_de_offset(int n) { return (n * (n | 1) + 10 * n + 1) & ~1; }
__host__ __device__ __forceinline__ int tri_lds_doubles(int n) { return tri_de_offset(n) + 2 * n; }

__device__ __forceinline__ double tri_guard(double q, double pivmin) { return fabs(q) < pivmin ? -pivmin : q; }

// ---------------------------------------------------------------------------------------------
// Step 1 helper: row r of the trailing block, columns c, c+4, ... < n. Applies the previous
// reflector's update a = A - v_r w_c - w_r v_c (w = p - K v) and returns sum_c a * vn_c.
__device__ __forceinline__ double tri_fused_row(double* __restrict__ row, const double* __restrict__ vo,
                                                const double* __restrict__ po, const double* __restrict__ vn,
                                                double K, double vr, double wr, int c, int n) {
    double s0 = 0.0, s1 = 0.0;
    for (; c + 12 < n; c += 16) {
        double x[4], v[4], p[4], u[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            x[q] = row[c + 4 * q];
            v[q] = vo[c + 4 * q];
            p[q] = po[c + 4 * q];
            u[q] = vn[c + 4 * q];
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const double w = fma(-K, v[q], p[q]);
            const double a = fma(-vr, w, fma(-wr, v[q], x[q]));
            row[c + 4 * q] = a;
            if (q & 1) s1 = fma(a, u[q], s1);
            else s0 = fma(a, u[q], s0);
        }
    }
    for (; c < n; c += 4) {
        const double v = vo[c], w = fma(-K, v, po[c]);
        const double a = fma(-vr, w, fma(-wr, v, row[c]));
        row[c] = a;
        s0 = fma(a, vn[c], s0);
    }
    return s0 + s1;
}

// ---------------------------------------------------------------------------------------------
// Step 3 helpers: sign counts of the leading principal minors p_i = (d_i - x) p_{i-1} -
// e_{i-1}^2 p_{i-2} of T - x. A zero minor counts as a change (the sign it takes in the limit of
// a pivot -> -0, the convention of the pivot recurrence q_i = p_i / p_{i-1}).
__device__ __forceinline__ void tri_step(double2 q, double x, double& pm, double& pc, int& cnt) {
    double pn = fma(q.x - x, pc, -q.y * pm);
    if (pn == 0.0) pn = -pc * 0x1p-60;
    cnt += (pn < 0.0) != (pc < 0.0);
    pm = pc;
    pc = pn;
}

__device__ __forceinline__ void tri_rescale(double& pm, double& pc) {
    int ex;
    frexp(pc, &ex);
    pc = ldexp(pc, -ex);
    pm = ldexp(pm, -ex);
}

// Numbers of eigenvalues of the block [b0, b1) of T below xa and below xb (p_{-1} = 1), two
// points interleaved, the minors rescaled by a power of two every four steps.
__device__ __forceinline__ void tri_count2(const double2* __restrict__ de, int b0, int b1, double xa, double xb,
                                           int& ca, int& cb) {
    double am = 1.0, ac = de[b0].x - xa, bm = 1.0, bc = de[b0].x - xb;
    if (ac == 0.0) ac = -0x1p-60;
    if (bc == 0.0) bc = -0x1p-60;
    int na = ac < 0.0, nb = bc < 0.0;
    int i = b0 + 1;
    for (; i + 4 <= b1; i += 4) {
        double2 q[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) q[u] = de[i + u];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            tri_step(q[u], xa, am, ac, na);
            tri_step(q[u], xb, bm, bc, nb);
        }
        tri_rescale(am, ac);
        tri_rescale(bm, bc);
    }
    for (; i < b1; ++i) {
        const double2 q = de[i];
        tri_step(q, xa, am, ac, na);
        tri_step(q, xb, bm, bc, nb);
    }
    ca = na;
    cb = nb;
}

// ---------------------------------------------------------------------------------------------
// Step 4: eigenvector of the block [b0, b1) at eigenvalue lm into z (normalized), through the
// twisted factorization T - lm = N_r D_r N_r^T: z_r = 1, z_i = -(e_i / D+_i) z_{i+1} below the
// twist and z_{i+1} = -(e_i / D-_{i+1}) z_i above it. The twist is the index of the q-th
// smallest |gamma_i| = |D+_i + D-_i - (d_i - lm)| (q = 0: the smallest). Returns false when the
// result is not finite.
__device__ __forceinline__ bool tri_vector(double* __restrict__ z, const double2* __restrict__ de,
                                           const double* __restrict__ e, int b0, int b1, double lm, int q,
                                           double pivmin) {
    // D- from the bottom, stored in z: D-_i = (d_i - lm) - e_i^2 / D-_{i+1}
    double dm = tri_guard(de[b1 - 1].x - lm, pivmin);
    z[b1 - 1] = dm;
    int i = b1 - 2;
    for (; i - 3 >= b0; i -= 4) {
        double2 w[5];
#pragma unroll
        for (int u = 0; u < 5; ++u) w[u] = de[i + 1 - u];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            dm = tri_guard((w[u + 1].x - lm) - w[u].y * fast_rcp(dm), pivmin);
            z[i - u] = dm;
        }
    }
    for (; i >= b0; --i) {
        dm = tri_guard((de[i].x - lm) - de[i + 1].y * fast_rcp(dm), pivmin);
        z[i] = dm;
    }
    // D+ from the top and the four smallest |gamma| (ties keep the earlier index first)
    double gb[4] = {fabs(z[b0]), INFINITY, INFINITY, INFINITY};
    int ix[4] = {b0, b0, b0, b0};
    double dp = tri_guard(de[b0].x - lm, pivmin);
    auto consider = [&](double g, int gi) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            if (g < gb[u]) {
                const double tg = gb[u];
                const int ti = ix[u];
                gb[u] = g;
                ix[u] = gi;
                g = tg;
                gi = ti;
            }
        }
    };
    i = b0 + 1;
    for (; i + 4 <= b1; i += 4) {
        double2 w[4];
        double zm[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            w[u] = de[i + u];
            zm[u] = z[i + u];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            dp = tri_guard((w[u].x - lm) - w[u].y * fast_rcp(dp), pivmin);
            consider(fabs(dp + zm[u] - (w[u].x - lm)), i + u);
        }
    }
    for (; i < b1; ++i) {
        const double2 w = de[i];
        dp = tri_guard((w.x - lm) - w.y * fast_rcp(dp), pivmin);
        consider(fabs(dp + z[i] - (w.x - lm)), i);
    }
    const int r = q <= 0 ? ix[0] : (q == 1 ? ix[1] : (q == 2 ? ix[2] : ix[3]));
    // D+ again, stored below the twist
    dp = tri_guard(de[b0].x - lm, pivmin);
    i = b0;
    for (; i + 4 <= r; i += 4) {
        double2 w[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) w[u] = de[i + 1 + u];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            z[i + u] = dp;
            dp = tri_guard((w[u].x - lm) - w[u].y * fast_rcp(dp), pivmin);
        }
    }
    for (; i < r; ++i) {
        z[i] = dp;
        dp = tri_guard((de[i + 1].x - lm) - de[i + 1].y * fast_rcp(dp), pivmin);
    }
    z[r] = 1.0;
    // below the twist (z_i holds D+_i)
    double zc = 1.0;
    i = r - 1;
    for (; i - 3 >= b0; i -= 4) {
        double ee[4], dd[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            ee[u] = e[i - u];
            dd[u] = z[i - u];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            zc = -ee[u] * zc * fast_rcp(dd[u]);
            z[i - u] = zc;
        }
    }
    for (; i >= b0; --i) {
        zc = -e[i] * zc * fast_rcp(z[i]);
        z[i] = zc;
    }
    // above the twist (z_{i+1} holds D-_{i+1})
    zc = 1.0;
    i = r;
    for (; i + 4 < b1; i += 4) {
        double ee[4], dd[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            ee[u] = e[i + u];
            dd[u] = z[i + 1 + u];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            zc = -ee[u] * zc * fast_rcp(dd[u]);
            z[i + 1 + u] = zc;
        }
    }
    for (; i + 1 < b1; ++i) {
        zc = -e[i] * zc * fast_rcp(z[i + 1]);
        z[i + 1] = zc;
    }
    double s0 = 0.0, s1 = 0.0;
    i = b0;
    for (; i + 1 < b1; i += 2) {
        s0 = fma(z[i], z[i], s0);
        s1 = fma(z[i + 1], z[i + 1], s1);
    }
    if (i < b1) s0 = fma(z[i], z[i], s0);
    const double ss = s0 + s1;
    const bool ok = ss > 0.0 && ss < INFINITY;
    const double inv = ok ? 1.0 / sqrt(ss) : 0.0;
    for (i = b0; i < b1; ++i) z[i] *= inv;
    return ok;
}

// ---------------------------------------------------------------------------------------------
// Step 5 helpers: z . v over c, c+4, ... < n, and z -= f v over the same columns.
__device__ __forceinline__ double tri_dot4(const double* __restrict__ z, const double* __restrict__ v, int c, int n) {
    double s0 = 0.0, s1 = 0.0;
    for (; c + 12 < n; c += 16) {
        double a[4], b[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            a[q] = z[c + 4 * q];
            b[q] = v[c + 4 * q];
        }
        s0 = fma(a[0], b[0], s0);
        s1 = fma(a[1], b[1], s1);
        s0 = fma(a[2], b[2], s0);
        s1 = fma(a[3], b[3], s1);
    }
    for (; c < n; c += 4) s0 = fma(z[c], v[c], s0);
    return s0 + s1;
}

__device__ __forceinline__ void tri_axpy4(double* __restrict__ z, const double* __restrict__ v, double f, int c,
                                          int n) {
    for (; c + 12 < n; c += 16) {
        double a[4], b[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            a[q] = z[c + 4 * q];
            b[q] = v[c + 4 * q];
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) z[c + 4 * q] = fma(-f, b[q], a[q]);
    }
    for (; c < n; c += 4) z[c] = fma(-f, v[c], z[c]);
}

// ---------------------------------------------------------------------------------------------
// Eigen-decomposition of the symmetric n x n matrix F (global, row-major; symmetrized on load).
// On success writes the eigenvalues ascending to eps and the eigenvectors as the rows of Ct (row
// k belongs to eps[k]) and returns 0. Returns -1, with Ct and eps untouched, when a run of close
// eigenvalues could not be separated or a non-finite number appeared; the caller then uses
// another method. Vg: n*n doubles of global scratch (the reflectors). lds: lds_cap doubles, at
// least tri_lds_doubles(n) + n.
__device__ __noinline__ int wg_eigh_tri(int n, const double* __restrict__ F, double* __restrict__ Ct,
                                        double* __restrict__ eps_out, double* __restrict__ Vg,
                                        double* __restrict__ lds, int lds_cap, WgScratch& s) {
    const int t = threadIdx.x, lane = t & 63;
    const int ld = tri_ld(n);
    double* A = lds;
    double* d = A + n * ld;
    double* e = d + n;
    double* tau = e + n;
    double* lam = tau + n;
    double* vb = lam + n;          // two reflectors (by parity), indexed by absolute row
    double* pb = vb + 2 * n;       // their products p = tau A v
    int* bs = (int*)(pb + 2 * n);  // block start of each row
    int* be = bs + n;              // block end (exclusive)
    int* rk = (int*)vb;            // rank of each eigenvalue (vb is free after step 1)
    double2* de = (double2*)(lds + tri_de_offset(n));  // (d_i, e_{i-1}^2)
    const int rg = t >> 2, rl = t & 3;  // row-parallel passes: four threads per row

    for (int idx = t; idx < n * n; idx += kWg) {
        const int r = idx / n, c = idx - r * n;
        A[r * ld + c] = 0.5 * (F[idx] + F[c * n + r]);
    }
    for (int idx = t; idx < 4 * n; idx += kWg) vb[idx] = 0.0;  // both parities of v and p
    if (t == 0) {
        s.bcast[2] = 0.0;  // failure flag
        s.bcast[3] = 0.0;  // close eigenvalues present
    }
    __syncthreads();

    // ---- 1. tridiagonalization
    int par = 0;        // parity of the reflector built in this step; the previous one is par ^ 1
    double tauo = 0.0;  // tau of the previous reflector (0 before the first)
    for (int k = 0; k + 2 < n; ++k) {
        const double* vo = vb + (par ^ 1) * n;
        const double* po = pb + (par ^ 1) * n;
        double* vn = vb + par * n;
        double* pn = pb + par * n;
        // K = (tau/2) p.v of the previous reflector (its support starts at row k)
        double pv = 0.0;
        for (int i = k + lane; i < n; i += 64) pv = fma(po[i], vo[i], pv);
        for (int off = 32; off > 0; off >>= 1) pv += __shfl_xor(pv, off, 64);
        const double Ko = 0.5 * tauo * pv;
        // column k after the previous update: x_r = A[r][k] - v_r w_k - w_r v_k (rows k + lane,
        // k + 64 + lane; lane 0 holds the diagonal, lane 1 the reflector's leading entry)
        const double vk = vo[k], wk = fma(-Ko, vk, po[k]);
        const int ra = k + lane, rb = k + 64 + lane;
        double xa = 0.0, xb = 0.0;
        if (ra < n) {
            const double v = vo[ra];
            xa = fma(-v, wk, fma(-fma(-Ko, v, po[ra]), vk, A[ra * ld + k]));
        }
        if (rb < n) {
            const double v = vo[rb];
            xb = fma(-v, wk, fma(-fma(-Ko, v, po[rb]), vk, A[rb * ld + k]));
        }
        const double dk = __shfl(xa, 0, 64), x0 = __shfl(xa, 1, 64);
        double sig = ra >= k + 2 ? xa * xa : 0.0;
        sig = fma(xb, xb, sig);
        for (int off = 32; off > 0; off >>= 1) sig += __shfl_xor(sig, off, 64);
        double tk = 0.0, beta = x0, scale = 0.0;
        if (sig > 0.0) {
            beta = -copysign(sqrt(fma(x0, x0, sig)), x0);
            tk = (beta - x0) / beta;
            scale = 1.0 / (x0 - beta);
        }
        const double va = ra == k + 1 ? 1.0 : (ra >= k + 2 ? xa * scale : 0.0), vbv = xb * scale;
        // every wavefront writes the same reflector, so each reads back what it wrote
        if (ra < n) vn[ra] = va;
        if (rb < n) vn[rb] = vbv;
        if (t < 64) {
            if (ra > k && ra < n) Vg[(size_t)k * n + ra] = va;
            if (rb < n) Vg[(size_t)k * n + rb] = vbv;
            if (lane == 0) {
                d[k] = dk;
                e[k] = beta;
                tau[k] = tk;
            }
        }
        // rows k+1.. of the trailing block: previous update, and p = tau A v for this reflector
        const int r = k + 1 + rg;
        if (r < n) {
            const double vr = vo[r], wr = fma(-Ko, vr, po[r]);
            double acc = tri_fused_row(A + r * ld, vo, po, vn, Ko, vr, wr, k + 1 + rl, n);
            acc += __shfl_xor(acc, 1, 64);
            acc += __shfl_xor(acc, 2, 64);
            if (rl == 0) pn[r] = tk * acc;
        }
        tauo = tk;
        par ^= 1;
        __syncthreads();
    }
    if (t == 0) {  // the last 2 x 2 block, after the last reflector's update
        const double* vo = vb + (par ^ 1) * n;
        const double* po = pb + (par ^ 1) * n;
        const int k0 = n >= 2 ? n - 2 : 0;
        double pv = 0.0;
        for (int i = k0; i < n; ++i) pv = fma(po[i], vo[i], pv);
        const double Ko = 0.5 * tauo * pv;
        auto upd = [&](int r, int c) {
            const double wr = fma(-Ko, vo[r], po[r]), wc = fma(-Ko, vo[c], po[c]);
            return fma(-vo[r], wc, fma(-wr, vo[c], A[r * ld + c]));
        };
        if (n >= 2) {
            d[n - 2] = upd(n - 2, n - 2);
            e[n - 2] = upd(n - 1, n - 2);
        }
        d[n - 1] = upd(n - 1, n - 1);
        e[n - 1] = 0.0;
    }
    __syncthreads();

    // ---- 2. norm, splitting, blocks
    if (t < 64) {
        double nrm = 0.0;
        for (int i = lane; i < n; i += 64) {
            const double a = fabs(d[i]) + (i > 0 ? fabs(e[i - 1]) : 0.0) + (i + 1 < n ? fabs(e[i]) : 0.0);
            nrm = fmax(nrm, a);
        }
        for (int off = 32; off > 0; off >>= 1) nrm = fmax(nrm, __shfl_xor(nrm, off, 64));
        // entries below 1e-10 ||T|| are dropped: that moves eigenvalues by less than 1e-10 ||T||,
        // and it separates the copies of a degenerate eigenvalue that the reduction leaves
        // coupled only through rounding-level entries
        const double tol = 1.0e-10 * nrm;
        for (int i = lane; i < n; i += 64) {
            double ei = i + 1 < n ? e[i] : 0.0;
            if (fabs(ei) <= tol) ei = 0.0;
            e[i] = ei;
        }
        for (int i = lane; i < n; i += 64) {  // reads entries the loop above settled (same wavefront)
            const double ep = i > 0 ? e[i - 1] : 0.0;
            de[i] = make_double2(d[i], ep * ep);
        }
        // Gershgorin bounds of T (valid for every block)
        double glo = INFINITY, ghi = -INFINITY;
        for (int i = lane; i < n; i += 64) {
            const double r = (i > 0 ? fabs(e[i - 1]) : 0.0) + fabs(e[i]);
            glo = fmin(glo, d[i] - r);
            ghi = fmax(ghi, d[i] + r);
        }
        for (int off = 32; off > 0; off >>= 1) {
            glo = fmin(glo, __shfl_xor(glo, off, 64));
            ghi = fmax(ghi, __shfl_xor(ghi, off, 64));
        }
        if (lane == 0) {
            s.bcast[0] = nrm;
            s.bcast[5] = glo;
            s.bcast[6] = ghi;
        }
    }
    __syncthreads();
    if (t == 0) {
        int cur = 0;
        for (int j = 0; j < n; ++j) {
            if (j > 0 && e[j - 1] == 0.0) cur = j;
            bs[j] = cur;
        }
        cur = n;
        for (int j = n - 1; j >= 0; --j) {
            be[j] = cur;
            if (j > 0 && e[j - 1] == 0.0) cur = j;
        }
    }
    __syncthreads();
    const double nrm = s.bcast[0];
    const double pivmin = 1.0e-280 * fmax(1.0, nrm * nrm);

    // ---- 3. eigenvalues: G threads per eigenvalue (about four wavefronts busy), two points each
    // per pass, so each pass divides the interval by 2G + 1
    {
        int G = 1;
        while (G < 8 && n * (2 * G) <= 256) G *= 2;
        const int j = t / G, l = t - j * G;
        if (j < n) {
            const int b0 = bs[j], b1 = be[j], kk = j - b0;
            double lo = d[b0], hi = d[b0];
            if (b1 - b0 > 1) {
                lo = s.bcast[5];
                hi = s.bcast[6];
                const double pad = 4.0 * DBL_EPSILON * fmax(fabs(lo), fabs(hi)) + pivmin;
                lo -= pad;
                hi += pad;
                const double atol = DBL_EPSILON * nrm;
                for (int it = 0; it < 80; ++it) {
                    const double w = hi - lo;
                    if (!(w > fmax(2.0 * DBL_EPSILON * fmax(fabs(lo), fabs(hi)), atol))) break;
                    const double step = w / (double)(2 * G + 1);
                    const double xa = lo + (double)(2 * l + 1) * step, xb = lo + (double)(2 * l + 2) * step;
                    int ca, cb;
                    tri_count2(de, b0, b1, xa, xb, ca, cb);
                    // a point at or below the kk-th eigenvalue raises lo, one above it lowers hi
                    double nlo = -INFINITY, nhi = INFINITY;
                    if (ca <= kk) nlo = xa;
                    else nhi = xa;
                    if (cb <= kk) nlo = fmax(nlo, xb);
                    else nhi = fmin(nhi, xb);
                    for (int off = 1; off < G; off <<= 1) {
                        nlo = fmax(nlo, __shfl_xor(nlo, off, 64));
                        nhi = fmin(nhi, __shfl_xor(nhi, off, 64));
                    }
                    lo = fmax(lo, nlo);
                    hi = fmin(hi, nhi);
                }
            }
            if (l == 0) lam[j] = 0.5 * (lo + hi);
        }
    }
    __syncthreads();

    // ---- 4. eigenvectors of T (row j of the LDS matrix)
    if (t < n) {
        const int j = t;
        double* z = A + j * ld;
        const int b0 = bs[j], b1 = be[j];
        for (int i = 0; i < b0; ++i) z[i] = 0.0;
        for (int i = b1; i < n; ++i) z[i] = 0.0;
        bool ok = true;
        if (b1 - b0 == 1) {
            z[b0] = 1.0;
        } else {
            // position inside a run of close eigenvalues: the q-th member takes the q-th best
            // twist, so the run's vectors are independent before Gram-Schmidt
            int q = 0;
            while (q < 3 && j - q - 1 >= b0 && lam[j - q] - lam[j - q - 1] < 1.0e-6 * nrm) ++q;
            ok = tri_vector(z, de, e, b0, b1, lam[j], q, pivmin);
        }
        if (!ok) s.bcast[2] = 1.0;
    }
    __syncthreads();

    // ---- close eigenvalues inside one block: Gram-Schmidt over each run (one wavefront; rare)
    if (t + 1 < n && bs[t + 1] == bs[t] && lam[t + 1] - lam[t] < 1.0e-6 * nrm) s.bcast[3] = 1.0;
    __syncthreads();
    if (s.bcast[3] != 0.0 && t < 64) {
        const double gap = 1.0e-6 * nrm;
        for (int j0 = 0; j0 < n;) {
            int j1 = j0 + 1;
            while (j1 < n && bs[j1] == bs[j0] && lam[j1] - lam[j1 - 1] < gap) ++j1;
            if (j1 - j0 > 1) {
                const int b0 = bs[j0], b1 = be[j0];
                for (int a = j0 + 1; a < j1; ++a) {
                    double* za = A + a * ld;
                    for (int q = j0; q < a; ++q) {
                        const double* zq = A + q * ld;
                        double dot = 0.0;
                        for (int i = b0 + lane; i < b1; i += 64) dot = fma(zq[i], za[i], dot);
                        for (int off = 32; off > 0; off >>= 1) dot += __shfl_xor(dot, off, 64);
                        for (int i = b0 + lane; i < b1; i += 64) za[i] = fma(-dot, zq[i], za[i]);
                    }
                    double ss = 0.0;
                    for (int i = b0 + lane; i < b1; i += 64) ss = fma(za[i], za[i], ss);
                    for (int off = 32; off > 0; off >>= 1) ss += __shfl_xor(ss, off, 64);
                    if (!(ss > 1.0e-6)) {  // |remainder| < 1e-3: the run is not resolved
                        if (lane == 0) s.bcast[2] = 1.0;
                        ss = 1.0;
                    }
                    const double inv = 1.0 / sqrt(ss);
                    for (int i = b0 + lane; i < b1; i += 64) za[i] *= inv;
                }
            }
            j0 = j1;
        }
    }
    __syncthreads();
    const bool failed = s.bcast[2] != 0.0;
    __syncthreads();
    if (failed) return -1;

    // ---- 5. ranks (ascending, stable), eigenvalues out
    for (int i = t / 16; i < n; i += kWg / 16) {
        const double ei = lam[i];
        int rank = 0;
        for (int j = t % 16; j < n; j += 16) rank += (lam[j] < ei) || (lam[j] == ei && j < i);
        for (int off = 8; off > 0; off >>= 1) rank += __shfl_xor(rank, off, 16);
        if (t % 16 == 0) {
            eps_out[rank] = ei;
            rk[i] = rank;
        }
    }
    __syncthreads();
    // back transformation x^T = z^T H_{n-3} ... H_0, four threads per row; the reflectors pass
    // through the LDS left over beside the matrix, cb at a time
    double* vst = lds + tri_lds_doubles(n);
    const int cb = (lds_cap - tri_lds_doubles(n)) / n;
    for (int khi = n - 3; khi >= 0; khi -= cb) {
        const int klo = max(0, khi - cb + 1);
        for (int idx = t; idx < (khi - klo + 1) * n; idx += kWg) {
            const int q = idx / n, c = idx - q * n, k = klo + q;
            vst[idx] = c > k ? Vg[(size_t)k * n + c] : 0.0;
        }
        __syncthreads();
        if (rg < n) {
            double* z = A + rg * ld;
            for (int k = khi; k >= klo; --k) {
                const double tk = tau[k];
                if (tk == 0.0) continue;
                const double* v = vst + (k - klo) * n;
                double dot = tri_dot4(z, v, k + 1 + rl, n);
                dot += __shfl_xor(dot, 1, 64);
                dot += __shfl_xor(dot, 2, 64);
                tri_axpy4(z, v, tk * dot, k + 1 + rl, n);
            }
        }
        __syncthreads();
    }
    if (rg < n) {
        const double* z = A + rg * ld;
        double* dst = Ct + (size_t)rk[rg] * n;
        for (int c = rl; c < n; c += 4) dst[c] = z[c];
    }
    __syncthreads();
    return 0;
}

}  // namespace am1g
