// Workgroup symmetric eigensolver for orders up to kTriMax, one 512-thread workgroup per matrix:
//   1. Householder tridiagonalization with the matrix in registers: eight threads own two rows
//      (16 columns each, in pairs), so only vectors go through LDS. Step k builds reflector k
//      from row k, which its owners published in the previous step (every wavefront computes
//      the reflector, so no barrier is needed to share it), and then each owner applies the
//      previous step's rank-2 update to its row and forms its entry of the product for the new
//      reflector: one barrier per step. The reflectors go to global scratch for step 5.
//   2. The tridiagonal T is split where an off-diagonal entry is below 1e-10 ||T||.
//   3. Eigenvalues of each unreduced block by multisection on Sturm counts (signs of the
//      leading principal minors), a group of threads per eigenvalue, two points per thread.
//   4. Eigenvectors of T from the twisted factorization at each eigenvalue (one thread per
//      eigenvector, O(n), in LDS); runs of eigenvalues closer than 1e-6 ||T|| inside one block
//      take distinct twist indices and are re-orthogonalized by Gram-Schmidt.
//   5. Sort, and back transformation x = H_0 ... H_{n-3} z with each row in registers and all
//      reflectors in LDS (no barriers between reflectors).
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

// LDS layout for order n (in doubles): an n x ld area (the eigenvectors of T, then the
// reflectors), thirteen n-vectors (d, e, tau, lam, two reflectors, two products, two published
// rows, block starts and ends as ints; the vector pairs start at even offsets), then the
// (d_i, e_{i-1}^2) pairs at an even offset.
__host__ __device__ __forceinline__ int tri_vec_offset(int n) { return (n * (n | 1) + 4 * n + 1) & ~1; }
// Every vector has kTriMax entries (zero beyond n), so the register-blocked passes load all of
// their columns without bounds tests and the compiler can issue the loads together.
__host__ __device__ __forceinline__ int tri_vec_stride(int) { return kTriMax; }
__host__ __device__ __forceinline__ int tri_de_offset(int n) {
    return (tri_vec_offset(n) + 6 * tri_vec_stride(n) + n + 1) & ~1;
}
__host__ __device__ __forceinline__ int tri_lds_doubles(int n) { return tri_de_offset(n) + 2 * n; }

__device__ __forceinline__ double tri_guard(double q, double pivmin) { return fabs(q) < pivmin ? -pivmin : q; }

// ---------------------------------------------------------------------------------------------
// Step 3 helpers: sign changes along the leading principal minors p_i = (d_i - x) p_{i-1} -
// e_{i-1}^2 p_{i-2} of T - x, counted from the sign bits. A zero minor reads as positive; inside
// an unreduced block (e_{i-1} != 0) the next minor is then -e^2 p_{i-1}, so the pair of steps
// counts exactly one change, as the pivot recurrence q_i = p_i / p_{i-1} does with a zero pivot
// replaced by -0.
__device__ __forceinline__ void tri_step(double2 q, double x, double& pm, double& pc, int& cnt) {
    const double pn = fma(q.x - x, pc, -q.y * pm);
    cnt += (int)((unsigned)(__double2hiint(pn) ^ __double2hiint(pc)) >> 31);
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

// tri_count2 that also returns the last minors: p_{b1-1}(x) = f * 2^e for both points (the
// characteristic polynomial of the block at x, up to its sign convention), for the
// regula-falsi steps of step 3.
__device__ __forceinline__ void tri_eval2(const double2* __restrict__ de, int b0, int b1, double xa, double xb,
                                          int& ca, int& cb, double& fa, int& ea, double& fb, int& eb) {
    double am = 1.0, ac = de[b0].x - xa, bm = 1.0, bc = de[b0].x - xb;
    int na = ac < 0.0, nb = bc < 0.0;
    int sa = 0, sb = 0;
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
        int ex;
        frexp(ac, &ex);
        ac = ldexp(ac, -ex);
        am = ldexp(am, -ex);
        sa += ex;
        frexp(bc, &ex);
        bc = ldexp(bc, -ex);
        bm = ldexp(bm, -ex);
        sb += ex;
    }
    for (; i < b1; ++i) {
        const double2 q = de[i];
        tri_step(q, xa, am, ac, na);
        tri_step(q, xb, bm, bc, nb);
    }
    ca = na;
    cb = nb;
    fa = ac;
    ea = sa;
    fb = bc;
    eb = sb;
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

// Steps 2 and 3 of the tridiagonal method for T = (d, e) of order n, all in LDS: the norm, the
// split into unreduced blocks (bs, be: block start and end of each row), the (d_i, e_{i-1}^2)
// pairs (de) and every eigenvalue into lam (unsorted, block order). cnt: 2 kWg ints of LDS
// scratch for the shared first pass, used when cnt_ok. Returns ||T|| (max row sum).
__device__ __forceinline__ double tri_spectrum(int n, double* __restrict__ d, double* __restrict__ e,
                                               double* __restrict__ lam, int* __restrict__ bs, int* __restrict__ be,
                                               double2* __restrict__ de, int* __restrict__ cnt, bool cnt_ok,
                                               const double* __restrict__ eps_out, double bracket, WgScratch& s) {
    const int t = threadIdx.x, lane = t & 63;
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
        // blocks: bs[j] = the last i <= j with e[i-1] = 0 (else 0), a running maximum over 64-entry
        // chunks; be[j] = the first i > j with e[i-1] = 0 (else n), a running minimum from the end
        int carry = 0;
        for (int c = 0; c < n; c += 64) {
            const int j = c + lane;
            int v = (j > 0 && j < n && e[j - 1] == 0.0) ? j : 0;
            for (int off = 1; off < 64; off <<= 1) {
                const int u = __shfl_up(v, off, 64);
                if (lane >= off) v = max(v, u);
            }
            v = max(v, carry);
            if (j < n) bs[j] = v;
            carry = __shfl(v, 63, 64);
        }
        const int last_start = carry;
        carry = n;
        for (int c = (n - 1) & ~63; c >= 0; c -= 64) {
            const int j = c + lane;
            int v = (j + 1 < n && e[j] == 0.0) ? j + 1 : n;
            for (int off = 1; off < 64; off <<= 1) {
                const int u = __shfl_down(v, off, 64);
                if (lane + off < 64) v = min(v, u);
            }
            v = min(v, carry);
            if (j < n) be[j] = v;
            carry = __shfl(v, 0, 64);
        }
        if (lane == 0) {
            s.bcast[0] = nrm;
            s.bcast[5] = glo;
            s.bcast[6] = ghi;
            s.bcast[7] = last_start == 0 ? 1.0 : 0.0;  // T is one unreduced block
        }
    }
    __syncthreads();
    const double nrm = s.bcast[0];
    const double pivmin = 1.0e-280 * fmax(1.0, nrm * nrm);

    // ---- 3. eigenvalues: G threads per eigenvalue, two points each per pass, so each pass
    // divides the interval by 2G + 1. Up to four wavefronts (one per SIMD): the counts are
    // issue-bound, and more points per pass cost more issue than the passes they save.
    // An unreduced T of order n >= 23 starts with one shared pass: each thread counts at two of
    // 2 kWg + 1 equal divisions of the Gershgorin interval (the counts go to the A area, free
    // until step 4), and every eigenvalue takes the division that holds it.
    const bool one_block = s.bcast[7] != 0.0;
    const bool shared_pass = one_block && cnt_ok;
    double glo = s.bcast[5], ghi = s.bcast[6];
    {
        const double pad = 4.0 * DBL_EPSILON * fmax(fabs(glo), fabs(ghi)) + pivmin;
        glo -= pad;
        ghi += pad;
    }
    const double sstep = (ghi - glo) / (double)(2 * kWg + 1);
    if (shared_pass) {
        int ca, cb;
        tri_count2(de, 0, n, glo + (double)(2 * t + 1) * sstep, glo + (double)(2 * t + 2) * sstep, ca, cb);
        cnt[2 * t] = ca;
        cnt[2 * t + 1] = cb;
        __syncthreads();
    }
    {
        int G = 1;
        while (G < 8 && n * (2 * G) <= kWg / 2) G *= 2;
        const int j = t / G, l = t - j * G;
        if (j < n) {
            const int b0 = bs[j], b1 = be[j], kk = j - b0;
            double lo = d[b0], hi = d[b0];
            if (b1 - b0 > 1) {
                lo = glo;
                hi = ghi;
                // Bracket state: block-local counts at the ends (clo <= kk < chi) and, once known,
                // the characteristic polynomial there (f * 2^e).
                int clo = 0, chi = b1 - b0, elo = 0, ehi = 0;
                double flo = 0.0, fhi = 0.0;
                bool klo = false, khi = false;
                if (shared_pass) {
                    // cnt[p] is the count at glo + (p + 1) sstep; keep cnt[plo] <= kk < cnt[phi]
                    int plo = -1, phi = 2 * kWg;
                    while (phi - plo > 1) {
                        const int mid = (plo + phi) >> 1;
                        if (cnt[mid] <= kk) plo = mid;
                        else phi = mid;
                    }
                    if (plo >= 0) {
                        lo = glo + (double)(plo + 1) * sstep;
                        clo = cnt[plo];
                    }
                    if (phi < 2 * kWg) {
                        hi = glo + (double)(phi + 1) * sstep;
                        chi = cnt[phi];
                    }
                }
                // a verified bracket around the previous eigenvalue (one block spanning T, so
                // the block-local index is the global one) when it is the narrower one: its
                // evaluation also gives the polynomial at both ends, so regula falsi can start
                if (bracket > 0.0 && one_block && 2.0 * bracket < hi - lo) {
                    const double wl = eps_out[j] - bracket, wh = eps_out[j] + bracket;
                    int cl, ch, el, eh;
                    double fl, fh;
                    tri_eval2(de, 0, n, wl, wh, cl, ch, fl, el, fh, eh);
                    if (cl <= kk && ch > kk) {
                        if (wl > lo) {
                            lo = wl;
                            clo = cl;
                            flo = fl;
                            elo = el;
                            klo = true;
                        }
                        if (wh < hi) {
                            hi = wh;
                            chi = ch;
                            fhi = fh;
                            ehi = eh;
                            khi = true;
                        }
                    }
                }
                // Passes: multisection while the bracket holds more than this eigenvalue or an
                // end's polynomial value is unknown; then regula falsi: the secant point xr of
                // the polynomial and points within +-eps of it, eps from the error the previous
                // secant point turned out to have (error ~ K w^2). Counts decide every move, so
                // the bracket always holds the eigenvalue; a pass that fails to halve it makes
                // the next one a multisection.
                const double atol = DBL_EPSILON * nrm;
                double xr_prev = 0.0, w_prev = 0.0;
                bool have_prev = false, force = false;
                for (int it = 0; it < 120; ++it) {
                    const double w = hi - lo;
                    if (!(w > fmax(2.0 * DBL_EPSILON * fmax(fabs(lo), fabs(hi)), atol))) break;
                    const bool rf = chi - clo == 1 && klo && khi && !force;
                    double xa, xb;
                    if (rf) {
                        const double r = ldexp(fhi / flo, max(-1000, min(1000, ehi - elo)));
                        const double tiny = fmax(1.0e-9 * w, 0.25 * atol);
                        double xr = lo + w / (1.0 - r);
                        xr = fmin(fmax(xr, lo + tiny), hi - tiny);
                        double ep = w * (1.0 / 64.0);
                        if (have_prev) ep = 8.0 * fabs(xr - xr_prev) * (w / w_prev) * (w / w_prev);
                        ep = fmax(ep, 0.5 * atol);
                        const double sp = 2.0 * ep / (double)(2 * G - 1);
                        xa = fmin(fmax(xr - ep + (double)(2 * l) * sp, lo + tiny), hi - tiny);
                        xb = fmin(fmax(xr - ep + (double)(2 * l + 1) * sp, lo + tiny), hi - tiny);
                        xr_prev = xr;
                        w_prev = w;
                        have_prev = true;
                    } else {
                        const double step = w / (double)(2 * G + 1);
                        xa = lo + (double)(2 * l + 1) * step;
                        xb = lo + (double)(2 * l + 2) * step;
                    }
                    int ca, cb, ea, eb;
                    double fa, fb;
                    tri_eval2(de, b0, b1, xa, xb, ca, cb, fa, ea, fb, eb);
                    // a point at or below the kk-th eigenvalue raises lo, one above it lowers hi
                    double nlo = -INFINITY, nhi = INFINITY, nflo = 0.0, nfhi = 0.0;
                    int nclo = 0, nchi = 0, nelo = 0, nehi = 0;
                    if (ca <= kk) {
                        nlo = xa; nclo = ca; nflo = fa; nelo = ea;
                    } else {
                        nhi = xa; nchi = ca; nfhi = fa; nehi = ea;
                    }
                    if (cb <= kk) {
                        if (xb > nlo) { nlo = xb; nclo = cb; nflo = fb; nelo = eb; }
                    } else if (xb < nhi) {
                        nhi = xb; nchi = cb; nfhi = fb; nehi = eb;
                    }
                    for (int off = 1; off < G; off <<= 1) {
                        const double olo = __shfl_xor(nlo, off, 64), ohi = __shfl_xor(nhi, off, 64);
                        const double oflo = __shfl_xor(nflo, off, 64), ofhi = __shfl_xor(nfhi, off, 64);
                        const int oclo = __shfl_xor(nclo, off, 64), ochi = __shfl_xor(nchi, off, 64);
                        const int oelo = __shfl_xor(nelo, off, 64), oehi = __shfl_xor(nehi, off, 64);
                        if (olo > nlo) { nlo = olo; nclo = oclo; nflo = oflo; nelo = oelo; }
                        if (ohi < nhi) { nhi = ohi; nchi = ochi; nfhi = ofhi; nehi = oehi; }
                    }
                    if (nlo > lo) {
                        lo = nlo; clo = nclo; flo = nflo; elo = nelo; klo = true;
                    }
                    if (nhi < hi) {
                        hi = nhi; chi = nchi; fhi = nfhi; ehi = nehi; khi = true;
                    }
                    force = rf && hi - lo > 0.5 * w;
                    if (force) have_prev = false;
                }
            }
            if (l == 0) lam[j] = 0.5 * (lo + hi);
        }
    }
    __syncthreads();

    return nrm;
}

// ---------------------------------------------------------------------------------------------
// Eigen-decomposition of the symmetric n x n matrix F (global, row-major; symmetrized on load).
// On success writes the eigenvalues ascending to eps and the eigenvectors as the rows of Ct (row
// k belongs to eps[k]) and returns 0. Returns -1, with Ct and eps untouched, when a run of close
// eigenvalues could not be separated or a non-finite number appeared; the caller then uses
// another method. Vg: n*n doubles of global scratch (the reflectors). lds: tri_lds_doubles(n)
// doubles (lds_cap is checked by the caller). Eigenvalues are resolved to rounding level (a
// looser stop moved the SCF trajectory enough to change converged orbital energies by about
// 1e-6 eV). bracket > 0: eps holds
// the eigenvalues of a nearby matrix (the previous SCF iteration's), and the search for
// eigenvalue j starts from [eps_j - bracket, eps_j + bracket] when Sturm counts confirm that it
// contains exactly that one. nvec: only the eigenvectors of the nvec lowest eigenvalues are formed
// and written (rows nvec.. of Ct are left as they were); the others cost no back transformation.
// An eigenvector is computed the same way whatever nvec is, so the rows written do not depend on it.
__device__ __noinline__ int wg_eigh_tri(int n, const double* __restrict__ F, double* __restrict__ Ct,
                                        double* __restrict__ eps_out, double* __restrict__ Vg,
                                        double* __restrict__ lds, int lds_cap, WgScratch& s, double bracket,
                                        int nvec) {
    const int t = threadIdx.x, lane = t & 63;
    const int ld = tri_ld(n), sv = tri_vec_stride(n);
    double* A = lds;               // eigenvectors of T (step 4), reflectors (step 5)
    double* d = A + n * ld;
    double* e = d + n;
    double* tau = e + n;
    double* lam = tau + n;
    double* vb = lds + tri_vec_offset(n);  // two reflectors (by parity), indexed by absolute row
    double* pb = vb + 2 * sv;              // their products p = tau A v
    double* pub = pb + 2 * sv;             // row k of the matrix, published for step k (by parity)
    int* bs = (int*)(pub + 2 * sv);        // block start of each row
    int* be = bs + n;                      // block end (exclusive)
    int* rk = (int*)vb;                    // rank of each eigenvalue (vb is free after step 1)
    int* inv = (int*)pb;                   // the eigenvalue of each rank (pb is free after step 1)
    int* need = inv + kTriMax;             // whether eigenvector j of T is formed
    double2* de = (double2*)(lds + tri_de_offset(n));  // (d_i, e_{i-1}^2)
    // Register blocking: a group of eight threads owns rows 2g and 2g + 1 (g = t / 8); thread q
    // of the group owns columns 2q + 16j + {0, 1}, j = 0..7, of both rows, in zr[b][2j + h].
    const int g8 = t >> 3, q8 = t & 7;
    double zr[2][16];
#pragma unroll
    for (int b = 0; b < 2; ++b) {
        const int r = 2 * g8 + b;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int c = 2 * q8 + 16 * j + h;
                zr[b][2 * j + h] = (r < n && c < n) ? 0.5 * (F[(size_t)r * n + c] + F[(size_t)c * n + r]) : 0.0;
            }
        }
    }
    for (int idx = t; idx < 4 * sv; idx += kWg) vb[idx] = 0.0;  // both parities of v and p
    if (g8 == 0) {  // row 0 is published for step 0
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const int c = 2 * q8 + 16 * j;
            if (c < n) pub[c] = zr[0][2 * j];
            if (c + 1 < n) pub[c + 1] = zr[0][2 * j + 1];
        }
    }
    if (t == 0) {
        s.bcast[2] = 0.0;  // failure flag
        s.bcast[3] = 0.0;  // close eigenvalues present
    }
    __syncthreads();

    // ---- 1. tridiagonalization
    int par = 0;        // parity of the reflector built in this step; the previous one is par ^ 1
    double tauo = 0.0;  // tau of the previous reflector (0 before the first)
    double vown[16];    // the previous reflector on the owned columns
#pragma unroll
    for (int u = 0; u < 16; ++u) vown[u] = 0.0;
    const int wlast = (n - 1) >> 4;  // the wavefront that owns row n - 1 (active in every step)
    for (int k = 0; k + 2 < n; ++k) {
        // Rows and columns below k are finished: there the previous v and p vanish, so the update
        // leaves them unchanged and the new reflector adds nothing to the product. A wavefront
        // whose rows are all <= k has nothing left to do (row k is not read again) and skips the
        // step, reflector included, which leaves the SIMD's issue slots to the others; so does
        // every quarter of the columns that lies below k (both tests are uniform across a
        // wavefront).
        if (16 * (t >> 6) + 15 > k && 16 * (t >> 6) < n) {
            const double* vo = vb + (par ^ 1) * sv;
            const double* po = pb + (par ^ 1) * sv;
            double* vn = vb + par * sv;
            double* pn = pb + par * sv;
            const double* rowk = pub + (k & 1) * sv;
            const int ra = lane, rb = 64 + lane;
            const int r0 = 2 * g8, r1 = r0 + 1;
            // The step's LDS reads that do not depend on its reflector, issued together ahead of
            // the arithmetic. The vectors hold kTriMax entries (zero beyond n), so the reads
            // need no bounds tests; entries of the published row beyond n are masked.
            const int i0 = min(k + lane, sv - 1), i1 = min(k + 64 + lane, sv - 1);
            const double pv0 = po[i0], vv0 = vo[i0], pv1 = po[i1], vv1 = vo[i1];
            const double vra = vo[ra], pra = po[ra], ara = rowk[ra];
            const double vrb = vo[rb], prb = po[rb], arb = rowk[rb];
            const double vk = vo[k], pk = po[k], akk = rowk[k];
            const double v1 = vo[k + 1], p1 = po[k + 1], ak1 = rowk[k + 1];
            const double vr0 = vo[r0], pr0 = po[r0], vr1 = vo[r1], pr1 = po[r1];
            double2 p2[8];  // the previous product on the owned columns (16-byte aligned pairs)
#pragma unroll
            for (int j = 0; j < 8; ++j) p2[j] = *(const double2*)(po + 2 * q8 + 16 * j);
            // K = (tau/2) p.v of the previous reflector (its support starts at row k)
            double pv = k + lane < n ? fma(pv0, vv0, 0.0) : 0.0;
            if (k + 64 + lane < n) pv = fma(pv1, vv1, pv);
            const double Ko = 0.5 * tauo * wave_sum(pv);
            // column k (= row k) after the previous update: x_r = A[k][r] - v_r w_k - w_r v_k
            // for r >= k, rows lane and 64 + lane
            const double wk = fma(-Ko, vk, pk);
            const double xa = (ra >= k && ra < n) ? fma(-vra, wk, fma(-fma(-Ko, vra, pra), vk, ara)) : 0.0;
            const double xb = (rb >= k && rb < n) ? fma(-vrb, wk, fma(-fma(-Ko, vrb, prb), vk, arb)) : 0.0;
            // the diagonal entry and the reflector's leading entry, by the same formula
            const double dk = fma(-vk, wk, fma(-fma(-Ko, vk, pk), vk, akk));
            const double x0 = fma(-v1, wk, fma(-fma(-Ko, v1, p1), vk, ak1));
            double sig = ra >= k + 2 ? xa * xa : 0.0;
            if (rb >= k + 2) sig = fma(xb, xb, sig);
            sig = wave_sum(sig);
            // beta = -sign(x0) |x|, tau = (beta - x0) / beta = 1 + |x0| / |x|, v = x / (x0 - beta);
            // 1/|x| and the last reciprocal from refined hardware estimates (a few ulps)
            double tk = 0.0, beta = x0, scale = 0.0;
            if (sig > 0.0) {
                const double nx2 = fma(x0, x0, sig);
                const double rn = (nx2 > 1.0e-280 && nx2 < 1.0e280) ? fast_rsqrt(nx2) : 1.0 / sqrt(nx2);
                beta = -copysign(nx2 * rn, x0);
                tk = fma(fabs(x0), rn, 1.0);
                const double den = x0 - beta;
                scale = (fabs(den) > 1.0e-280 && fabs(den) < 1.0e280) ? fast_rcp(den) : 1.0 / den;
            }
            const double va = ra == k + 1 ? 1.0 : (ra >= k + 2 ? xa * scale : 0.0);
            const double vbv = rb == k + 1 ? 1.0 : (rb >= k + 2 ? xb * scale : 0.0);
            // every active wavefront writes the same reflector (zero above its support), so each
            // reads back what it wrote; rows <= k of the new product are cleared the same way
            if (ra < n) {
                vn[ra] = va;
                if (ra <= k) pn[ra] = 0.0;
            }
            if (rb < n) {
                vn[rb] = vbv;
                if (rb <= k) pn[rb] = 0.0;
            }
            if ((t >> 6) == wlast) {
                if (ra > k && ra < n) Vg[(size_t)k * n + ra] = va;
                if (rb > k && rb < n) Vg[(size_t)k * n + rb] = vbv;
                if (lane == 0) {
                    d[k] = dk;
                    e[k] = beta;
                    tau[k] = tk;
                }
            }
            // owned rows k+1..: previous update (zero on the columns it does not touch, since
            // the previous v and p vanish there) and this reflector's product
            // p_r = tau sum_c a_rc v_c
            const bool on0 = r0 > k && r0 < n, on1 = r1 > k && r1 < n;
            const double wr0 = fma(-Ko, vr0, pr0), wr1 = fma(-Ko, vr1, pr1);
            double2 u2[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) u2[j] = *(const double2*)(vn + 2 * q8 + 16 * j);
            auto fused = [&](int j, double& acc0, double& acc1) {
                const double2 p = p2[j], u = u2[j];
                const double w0 = fma(-Ko, vown[2 * j], p.x), w1 = fma(-Ko, vown[2 * j + 1], p.y);
                const double a00 = fma(-vr0, w0, fma(-wr0, vown[2 * j], zr[0][2 * j]));
                const double a01 = fma(-vr0, w1, fma(-wr0, vown[2 * j + 1], zr[0][2 * j + 1]));
                const double a10 = fma(-vr1, w0, fma(-wr1, vown[2 * j], zr[1][2 * j]));
                const double a11 = fma(-vr1, w1, fma(-wr1, vown[2 * j + 1], zr[1][2 * j + 1]));
                zr[0][2 * j] = a00;
                zr[0][2 * j + 1] = a01;
                zr[1][2 * j] = a10;
                zr[1][2 * j + 1] = a11;
                acc0 = fma(a00, u.x, fma(a01, u.y, acc0));
                acc1 = fma(a10, u.x, fma(a11, u.y, acc1));
                vown[2 * j] = u.x;  // this reflector is the next step's previous one
                vown[2 * j + 1] = u.y;
            };
            // odd column blocks accumulate separately (shorter chains)
            double s0 = 0.0, s1 = 0.0, u0 = 0.0, u1 = 0.0;
#pragma unroll
            for (int jq = 0; jq < 4; ++jq) {
                if (32 * jq + 32 > k && 32 * jq < n) {
                    fused(2 * jq, s0, s1);
                    fused(2 * jq + 1, u0, u1);
                }
            }
            s0 += u0;
            s1 += u1;
            s0 += dpp_move<0xB1>(s0);  // sums over the group of eight
            s1 += dpp_move<0xB1>(s1);
            s0 += dpp_move<0x4E>(s0);
            s1 += dpp_move<0x4E>(s1);
            s0 += dpp_move<0x141>(s0);
            s1 += dpp_move<0x141>(s1);
            if (q8 == 0) {
                if (on0) pn[r0] = tk * s0;
                if (on1) pn[r1] = tk * s1;
            }
            if (r0 == k + 1 || r1 == k + 1) {  // publish row k+1 for the next step
                const int b = r1 == k + 1;
                double* nxt = pub + ((k + 1) & 1) * sv;
#pragma unroll
                for (int j = 0; j < 8; ++j) {
                    const int c = 2 * q8 + 16 * j;
                    *(double2*)(nxt + c) = b ? make_double2(zr[1][2 * j], zr[1][2 * j + 1])
                                             : make_double2(zr[0][2 * j], zr[0][2 * j + 1]);
                }
            }
            tauo = tk;
        }
        par ^= 1;
        __syncthreads();
    }
    // the last 2 x 2 block (rows and columns n-2, n-1) after the last reflector's update
    {
        const double* vo = vb + (par ^ 1) * sv;
        const double* po = pb + (par ^ 1) * sv;
        const int k0 = n >= 2 ? n - 2 : 0;
        double pv = 0.0;
        for (int i = k0 + lane; i < n; i += 64) pv = fma(po[i], vo[i], pv);
        const double Ko = 0.5 * tauo * wave_sum(pv);
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            const int r = 2 * g8 + b;
            if (r >= k0 && r < n) {
                const double vr = vo[r], wr = fma(-Ko, vr, po[r]);
#pragma unroll
                for (int j = 0; j < 8; ++j) {
#pragma unroll
                    for (int h = 0; h < 2; ++h) {
                        const int c = 2 * q8 + 16 * j + h;
                        if (c >= k0 && c < n) {
                            const double a = fma(-vr, fma(-Ko, vo[c], po[c]), fma(-wr, vo[c], zr[b][2 * j + h]));
                            s.bcast[8 + 2 * (r - k0) + (c - k0)] = a;
                        }
                    }
                }
            }
        }
    }
    __syncthreads();
    if (t == 0) {
        if (n >= 2) {
            d[n - 2] = s.bcast[8];
            e[n - 2] = s.bcast[10];  // row n-1, column n-2
            d[n - 1] = s.bcast[11];
        } else {
            d[0] = s.bcast[8];
        }
        e[n - 1] = 0.0;
    }
    __syncthreads();

    const double nrm = tri_spectrum(n, d, e, lam, bs, be, de, (int*)A, n * ld >= kWg, eps_out, bracket, s);
    const double pivmin = 1.0e-280 * fmax(1.0, nrm * nrm);

    // ---- ranks (ascending, stable) of the eigenvalues, and the eigenvectors of T to form: those
    // of the nvec lowest eigenvalues, and every member of a run of close eigenvalues (inside one
    // block) that holds one of them, since Gram-Schmidt below treats a run as a whole
    for (int i = t / 16; i < n; i += kWg / 16) {
        const double ei = lam[i];
        int rank = 0;
        for (int j = t % 16; j < n; j += 16) rank += (lam[j] < ei) || (lam[j] == ei && j < i);
        for (int off = 8; off > 0; off >>= 1) rank += __shfl_xor(rank, off, 16);
        if (t % 16 == 0) {
            rk[i] = rank;
            inv[rank] = i;
        }
    }
    __syncthreads();
    const double gap = 1.0e-6 * nrm;
    if (t < n) {
        bool nd = rk[t] < nvec;
        for (int a = t; a > bs[t] && lam[a] - lam[a - 1] < gap; --a) nd = nd || rk[a - 1] < nvec;
        for (int a = t + 1; a < be[t] && lam[a] - lam[a - 1] < gap; ++a) nd = nd || rk[a] < nvec;
        need[t] = nd ? 1 : 0;
    }
    __syncthreads();

    // ---- 4. eigenvectors of T (row j of the LDS matrix)
    if (t < n && need[t]) {
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
            while (q < 3 && j - q - 1 >= b0 && lam[j - q] - lam[j - q - 1] < gap) ++q;
            ok = tri_vector(z, de, e, b0, b1, lam[j], q, pivmin);
        }
        if (!ok) s.bcast[2] = 1.0;
    }
    __syncthreads();

    // ---- close eigenvalues inside one block: Gram-Schmidt over each run (one wavefront; rare)
    if (t + 1 < n && need[t] && bs[t + 1] == bs[t] && lam[t + 1] - lam[t] < gap) s.bcast[3] = 1.0;
    __syncthreads();
    if (s.bcast[3] != 0.0 && t < 64) {
        for (int j0 = 0; j0 < n;) {
            int j1 = j0 + 1;
            while (j1 < n && bs[j1] == bs[j0] && lam[j1] - lam[j1 - 1] < gap) ++j1;
            if (j1 - j0 > 1 && need[j0]) {
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
                    const double rs = 1.0 / sqrt(ss);
                    for (int i = b0 + lane; i < b1; i += 64) za[i] *= rs;
                }
            }
            j0 = j1;
        }
    }
    __syncthreads();
    const bool failed = s.bcast[2] != 0.0;
    __syncthreads();
    if (failed) return -1;

    // ---- 5. eigenvalues out (ascending)
    for (int i = t; i < n; i += kWg) eps_out[rk[i]] = lam[i];
    // back transformation x^T = z^T H_{n-3} ... H_0: rows 2g, 2g + 1 of the eigenvectors of T
    // into the owners' registers, then all reflectors into the freed LDS area (even stride sr;
    // (n - 2) sr <= n ld)
    const int sr = (n + 1) & ~1;
    // (group g takes the eigenvectors of ranks 2g and 2g + 1, below nvec)
#pragma unroll
    for (int b = 0; b < 2; ++b) {
        const int r = 2 * g8 + b;
        const int jr = r < nvec ? inv[r] : 0;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int c = 2 * q8 + 16 * j + h;
                zr[b][2 * j + h] = (r < nvec && c < n) ? A[jr * ld + c] : 0.0;
            }
        }
    }
    __syncthreads();
#pragma unroll 8
    for (int idx = t; idx < (n - 2) * sr; idx += kWg) {
        const int k = idx / sr, c = idx - k * sr;
        A[idx] = (c > k && c < n) ? Vg[(size_t)k * n + c] : 0.0;
    }
    __syncthreads();
    if (2 * g8 < nvec) {
        for (int k = n - 3; k >= 0; --k) {
            // tau = 0 (a column that needed no reflection) makes the step a no-op
            const double tk = tau[k];
            const double* v = A + k * sr;
            // All loads first, so they cost one LDS round trip: columns >= n hold zeros in the
            // rows and in the owners' registers; beyond the row stride the load is clamped into
            // the row and its value replaced by zero.
            double2 vv[8];
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int c = 2 * q8 + 16 * j;
                const double2 ld2 = *(const double2*)(v + min(c, sr - 2));
                vv[j] = c < n ? ld2 : make_double2(0.0, 0.0);
            }
            // The reflector vanishes on columns <= k, so column quarters that lie there are
            // skipped (a uniform test); even and odd column blocks accumulate separately
            // (shorter dependency chains).
            double s0 = 0.0, s1 = 0.0, u0 = 0.0, u1 = 0.0;
            auto dot = [&](int j, double& a0, double& a1) {
                a0 = fma(zr[0][2 * j], vv[j].x, fma(zr[0][2 * j + 1], vv[j].y, a0));
                a1 = fma(zr[1][2 * j], vv[j].x, fma(zr[1][2 * j + 1], vv[j].y, a1));
            };
#pragma unroll
            for (int jq = 0; jq < 4; ++jq) {
                if (32 * jq + 31 > k && 32 * jq < n) {
                    dot(2 * jq, s0, s1);
                    dot(2 * jq + 1, u0, u1);
                }
            }
            s0 += u0;
            s1 += u1;
            s0 += dpp_move<0xB1>(s0);
            s1 += dpp_move<0xB1>(s1);
            s0 += dpp_move<0x4E>(s0);
            s1 += dpp_move<0x4E>(s1);
            s0 += dpp_move<0x141>(s0);
            s1 += dpp_move<0x141>(s1);
            const double f0 = tk * s0, f1 = tk * s1;
#pragma unroll
            for (int jq = 0; jq < 4; ++jq) {
                if (32 * jq + 31 > k && 32 * jq < n) {
#pragma unroll
                    for (int j = 2 * jq; j < 2 * jq + 2; ++j) {
                        zr[0][2 * j] = fma(-f0, vv[j].x, zr[0][2 * j]);
                        zr[0][2 * j + 1] = fma(-f0, vv[j].y, zr[0][2 * j + 1]);
                        zr[1][2 * j] = fma(-f1, vv[j].x, zr[1][2 * j]);
                        zr[1][2 * j + 1] = fma(-f1, vv[j].y, zr[1][2 * j + 1]);
                    }
                }
            }
        }
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            const int r = 2 * g8 + b;
            if (r >= nvec) continue;
            double* dst = Ct + (size_t)r * n;
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int c = 2 * q8 + 16 * j;
                if (c < n) dst[c] = zr[b][2 * j];
                if (c + 1 < n) dst[c + 1] = zr[b][2 * j + 1];
            }
        }
    }
    __syncthreads();
    return 0;
}

// ---------------------------------------------------------------------------------------------
// Orders kTriMax < n <= kTriBig: the same five steps with the matrix in global memory (too large
// for the registers of one workgroup). A (n x n, row-major, in the caller's scratch) holds the
// symmetrized matrix during the reduction and then the eigenvectors of T; Vg (n x n) the
// reflectors. The vectors of the reduction, T, the ranks and the eigenvalue counts live in LDS;
// the eigenvectors of T are built in LDS a chunk at a time and the reflectors are staged
// through LDS a block at a time for the back transformation.
//
// Step 1 keeps the scheme of the register version: step k builds reflector k from row k, which
// its owners published in the previous step, and each owner then applies the previous step's
// rank-2 update to its rows and forms its entry of the new product in the same pass over the
// rows (one read and one write of the trailing rows per step, one barrier). A group of eight
// threads owns rows 2g, 2g + 1, 128 + 2g and 129 + 2g; thread q of the group owns columns
// 2q + 16j + {0, 1}, j = 0..15, so eight consecutive threads cover 128 consecutive bytes of a row.
constexpr int kTriBig = 256;

// LDS layout (doubles) of the global variant: d, e, tau, lam, two reflectors, two products, two
// published rows (kTriBig each), block starts and ends and ranks (ints), the (d_i, e_{i-1}^2)
// pairs, the shared-pass counts (2 kWg ints), then the staging area.
constexpr int kTgD = 0, kTgE = 256, kTgTau = 512, kTgLam = 768, kTgV = 1024, kTgP = 1536, kTgPub = 2048;
constexpr int kTgBs = 2560, kTgBe = 2688, kTgRk = 2816, kTgDe = 2944, kTgCnt = 3456;
constexpr int kTgStage = kTgCnt + kWg;  // 3968

__device__ __noinline__ int wg_eigh_tri_g(int n, const double* __restrict__ F, double* __restrict__ Ct,
                                          double* __restrict__ eps_out, double* __restrict__ A,
                                          double* __restrict__ Vg, double* __restrict__ lds, int lds_cap,
                                          WgScratch& s, double bracket) {
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    double* d = lds + kTgD;
    double* e = lds + kTgE;
    double* tau = lds + kTgTau;
    double* lam = lds + kTgLam;
    double* vb = lds + kTgV;    // reflectors by parity, indexed by absolute row
    double* pb = lds + kTgP;    // their products p = tau A v
    double* pub = lds + kTgPub; // row k of the matrix, published for step k (by parity)
    int* bs = (int*)(lds + kTgBs);
    int* be = (int*)(lds + kTgBe);
    int* rk = (int*)(lds + kTgRk);
    double2* de = (double2*)(lds + kTgDe);
    int* cnt = (int*)(lds + kTgCnt);
    double* stage = lds + kTgStage;
    const int stage_cap = lds_cap - kTgStage;
    const int g8 = t >> 3, q8 = t & 7;

    // A = (F + F^T) / 2; row 0 published; vectors cleared
    for (int idx = t; idx < n * n; idx += kWg) {
        const int r = idx / n, c = idx - r * n;
        A[idx] = 0.5 * (F[(size_t)r * n + c] + F[(size_t)c * n + r]);
    }
    for (int idx = t; idx < 2 * kTriBig; idx += kWg) {
        vb[idx] = 0.0;
        pb[idx] = 0.0;
    }
    for (int c = t; c < kTriBig; c += kWg) pub[c] = c < n ? 0.5 * (F[c] + F[(size_t)c * n]) : 0.0;
    if (t == 0) {
        s.bcast[2] = 0.0;  // failure flag
        s.bcast[3] = 0.0;  // close eigenvalues present
    }
    __syncthreads();

    // ---- 1. tridiagonalization
    int par = 0;
    double tauo = 0.0;
    for (int k = 0; k + 2 < n; ++k) {
        const double* vo = vb + (par ^ 1) * kTriBig;
        const double* po = pb + (par ^ 1) * kTriBig;
        double* vn = vb + par * kTriBig;
        double* pn = pb + par * kTriBig;
        const double* rowk = pub + (k & 1) * kTriBig;
        double pv = 0.0;
        for (int i = k + lane; i < n; i += 64) pv = fma(po[i], vo[i], pv);
        const double Ko = 0.5 * tauo * wave_sum(pv);
        // column k after the previous update, rows lane + 64 u
        const double vk = vo[k], wk = fma(-Ko, vk, po[k]);
        double x[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int r = lane + 64 * u;
            x[u] = 0.0;
            if (r >= k && r < n) {
                const double v = vo[r];
                x[u] = fma(-v, wk, fma(-fma(-Ko, v, po[r]), vk, rowk[r]));
            }
        }
        const double dk = fma(-vk, wk, fma(-fma(-Ko, vk, po[k]), vk, rowk[k]));
        const double v1 = vo[k + 1];
        const double x0 = fma(-v1, wk, fma(-fma(-Ko, v1, po[k + 1]), vk, rowk[k + 1]));
        double sig = 0.0;
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (lane + 64 * u >= k + 2) sig = fma(x[u], x[u], sig);
        sig = wave_sum(sig);
        double tk = 0.0, beta = x0, scale = 0.0;
        if (sig > 0.0) {
            const double nx2 = fma(x0, x0, sig);
            const double rn = (nx2 > 1.0e-280 && nx2 < 1.0e280) ? fast_rsqrt(nx2) : 1.0 / sqrt(nx2);
            beta = -copysign(nx2 * rn, x0);
            tk = fma(fabs(x0), rn, 1.0);
            const double den = x0 - beta;
            scale = (fabs(den) > 1.0e-280 && fabs(den) < 1.0e280) ? fast_rcp(den) : 1.0 / den;
        }
        // every wavefront writes the same reflector (zero outside its support) and reads back
        // what it wrote; rows <= k of the new product are cleared the same way
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int r = lane + 64 * u;
            const double v = r == k + 1 ? 1.0 : (r >= k + 2 && r < n ? x[u] * scale : 0.0);
            vn[r] = v;
            if (r <= k) pn[r] = 0.0;
            if (wv == 0 && r > k && r < n) Vg[(size_t)k * n + r] = v;
        }
        if (t == 0) {
            d[k] = dk;
            e[k] = beta;
            tau[k] = tk;
        }
        // owned rows k+1..n-1: the previous update and this reflector's product, a half of the
        // rows (h) and a quarter of the columns (jq) at a time; finished rows and columns (<= k)
        // are skipped with wavefront-uniform tests
#pragma unroll 1
        for (int h = 0; h < 2; ++h) {
            const int r0 = 2 * g8 + 128 * h, r1 = r0 + 1;
            if (16 * wv + 128 * h + 15 <= k || 16 * wv + 128 * h >= n) continue;
            const bool on0 = r0 > k && r0 < n, on1 = r1 > k && r1 < n;
            const double vr0 = vo[r0 < n ? r0 : 0], wr0 = fma(-Ko, vr0, po[r0 < n ? r0 : 0]);
            const double vr1 = vo[r1 < n ? r1 : 0], wr1 = fma(-Ko, vr1, po[r1 < n ? r1 : 0]);
            double* a0 = A + (size_t)(on0 ? r0 : 0) * n;
            double* a1 = A + (size_t)(on1 ? r1 : 0) * n;
            const bool pub0 = r0 == k + 1, pub1 = r1 == k + 1;
            double* nxt = pub + ((k + 1) & 1) * kTriBig;
            double s0 = 0.0, s1 = 0.0;
#pragma unroll 1
            for (int jq = 0; jq < 4; ++jq) {
                if (64 * jq + 63 <= k || 64 * jq >= n) continue;
                double av0[8], av1[8], vv[8], pp[8], uu[8];
#pragma unroll
                for (int i = 0; i < 8; ++i) {
                    const int c = 2 * q8 + 16 * (4 * jq + (i >> 1)) + (i & 1);
                    const bool cin = c < n;
                    av0[i] = (on0 && cin) ? a0[c] : 0.0;
                    av1[i] = (on1 && cin) ? a1[c] : 0.0;
                    vv[i] = vo[c];
                    pp[i] = po[c];
                    uu[i] = vn[c];
                }
#pragma unroll
                for (int i = 0; i < 8; ++i) {
                    const int c = 2 * q8 + 16 * (4 * jq + (i >> 1)) + (i & 1);
                    const double w = fma(-Ko, vv[i], pp[i]);
                    const double b0 = fma(-vr0, w, fma(-wr0, vv[i], av0[i]));
                    const double b1 = fma(-vr1, w, fma(-wr1, vv[i], av1[i]));
                    s0 = fma(b0, uu[i], s0);
                    s1 = fma(b1, uu[i], s1);
                    if (c < n) {
                        if (on0) a0[c] = b0;
                        if (on1) a1[c] = b1;
                        if (pub0) nxt[c] = b0;
                        if (pub1) nxt[c] = b1;
                    }
                }
            }
            s0 += dpp_move<0xB1>(s0);  // sums over the group of eight
            s1 += dpp_move<0xB1>(s1);
            s0 += dpp_move<0x4E>(s0);
            s1 += dpp_move<0x4E>(s1);
            s0 += dpp_move<0x141>(s0);
            s1 += dpp_move<0x141>(s1);
            if (q8 == 0) {
                if (on0) pn[r0] = tk * s0;
                if (on1) pn[r1] = tk * s1;
            }
        }
        tauo = tk;
        par ^= 1;
        __syncthreads();
    }
    // the last 2 x 2 block (rows and columns n-2, n-1) after the last reflector's update
    if (t < 4) {
        const double* vo = vb + (par ^ 1) * kTriBig;
        const double* po = pb + (par ^ 1) * kTriBig;
        double pv = 0.0;
        for (int i = n - 2; i < n; ++i) pv = fma(po[i], vo[i], pv);
        const double Ko = 0.5 * tauo * pv;
        const int r = n - 2 + (t >> 1), c = n - 2 + (t & 1);
        const double vr = vo[r], wr = fma(-Ko, vr, po[r]);
        s.bcast[8 + t] = fma(-vr, fma(-Ko, vo[c], po[c]), fma(-wr, vo[c], A[(size_t)r * n + c]));
    }
    __syncthreads();
    if (t == 0) {
        d[n - 2] = s.bcast[8];
        e[n - 2] = s.bcast[10];  // row n-1, column n-2
        d[n - 1] = s.bcast[11];
        e[n - 1] = 0.0;
    }
    __syncthreads();

    // ---- 2, 3. split and eigenvalues
    const double nrm = tri_spectrum(n, d, e, lam, bs, be, de, cnt, true, eps_out, bracket, s);
    const double pivmin = 1.0e-280 * fmax(1.0, nrm * nrm);

    // ---- 4. eigenvectors of T, built in LDS a chunk of rows at a time, then rows of A
    const int ld = n | 1;
    const int chunk = min(kWg, stage_cap / ld);
    for (int j0 = 0; j0 < n; j0 += chunk) {
        const int j = j0 + t;
        if (t < chunk && j < n) {
            double* z = stage + t * ld;
            const int b0 = bs[j], b1 = be[j];
            for (int i = 0; i < b0; ++i) z[i] = 0.0;
            for (int i = b1; i < n; ++i) z[i] = 0.0;
            bool ok = true;
            if (b1 - b0 == 1) {
                z[b0] = 1.0;
            } else {
                int q = 0;
                while (q < 3 && j - q - 1 >= b0 && lam[j - q] - lam[j - q - 1] < 1.0e-6 * nrm) ++q;
                ok = tri_vector(z, de, e, b0, b1, lam[j], q, pivmin);
            }
            if (!ok) s.bcast[2] = 1.0;
        }
        __syncthreads();
        const int rows = min(chunk, n - j0);
        for (int idx = t; idx < rows * n; idx += kWg) {
            const int r = idx / n, c = idx - r * n;
            A[(size_t)(j0 + r) * n + c] = stage[r * ld + c];
        }
        __syncthreads();
    }
    // close eigenvalues inside one block: Gram-Schmidt over each run (one wavefront; rare)
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
                    double* za = A + (size_t)a * n;
                    for (int q = j0; q < a; ++q) {
                        const double* zq = A + (size_t)q * n;
                        double dot = 0.0;
                        for (int i = b0 + lane; i < b1; i += 64) dot = fma(zq[i], za[i], dot);
                        for (int off = 32; off > 0; off >>= 1) dot += __shfl_xor(dot, off, 64);
                        for (int i = b0 + lane; i < b1; i += 64) za[i] = fma(-dot, zq[i], za[i]);
                    }
                    double ss = 0.0;
                    for (int i = b0 + lane; i < b1; i += 64) ss = fma(za[i], za[i], ss);
                    for (int off = 32; off > 0; off >>= 1) ss += __shfl_xor(ss, off, 64);
                    if (!(ss > 1.0e-6)) {
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
    // back transformation x^T = z^T H_{n-3} ... H_0, a half of the rows at a time (rows 2g,
    // 2g + 1 of the half in the owners' registers); the reflectors go through LDS in blocks,
    // the last ones first (even stride sr, zero outside each reflector's support)
    const int sr = (n + 1) & ~1;
    const int blk = max(1, min(n - 2, stage_cap / sr));
#pragma unroll 1
    for (int h = 0; h < 2; ++h) {
        const int r0 = 2 * g8 + 128 * h, r1 = r0 + 1;
        double z0[32], z1[32];
#pragma unroll
        for (int i = 0; i < 32; ++i) {
            const int c = 2 * q8 + 16 * (i >> 1) + (i & 1);
            z0[i] = (r0 < n && c < n) ? A[(size_t)r0 * n + c] : 0.0;
            z1[i] = (r1 < n && c < n) ? A[(size_t)r1 * n + c] : 0.0;
        }
        const bool active = 16 * wv + 128 * h < n;
#pragma unroll 1
        for (int kb = n - 3; kb >= 0; kb -= blk) {
            const int k_lo = max(0, kb - blk + 1);
            __syncthreads();  // the previous block is consumed
            for (int idx = t; idx < (kb - k_lo + 1) * sr; idx += kWg) {
                const int kk = idx / sr, c = idx - kk * sr, k = k_lo + kk;
                stage[idx] = (c > k && c < n) ? Vg[(size_t)k * n + c] : 0.0;
            }
            __syncthreads();
            if (!active) continue;
#pragma unroll 1
            for (int k = kb; k >= k_lo; --k) {
                const double tk = tau[k];
                const double* v = stage + (k - k_lo) * sr;
                double s0 = 0.0, s1 = 0.0, u0 = 0.0, u1 = 0.0;
#pragma unroll
                for (int jh = 0; jh < 2; ++jh) {  // columns 0..127, then 128..255
                    if (128 * jh + 127 <= k || 128 * jh >= n) continue;
                    double2 vv[8];
#pragma unroll
                    for (int j = 0; j < 8; ++j) {
                        const int c = 2 * q8 + 16 * (8 * jh + j);
                        vv[j] = *(const double2*)(v + min(c, sr - 2));
                        if (c >= n) vv[j] = make_double2(0.0, 0.0);
                    }
#pragma unroll
                    for (int j = 0; j < 8; ++j) {
                        const int i = 2 * (8 * jh + j);
                        double& a0 = (j & 1) ? u0 : s0;
                        double& a1 = (j & 1) ? u1 : s1;
                        a0 = fma(z0[i], vv[j].x, fma(z0[i + 1], vv[j].y, a0));
                        a1 = fma(z1[i], vv[j].x, fma(z1[i + 1], vv[j].y, a1));
                    }
                }
                s0 += u0;
                s1 += u1;
                s0 += dpp_move<0xB1>(s0);
                s1 += dpp_move<0xB1>(s1);
                s0 += dpp_move<0x4E>(s0);
                s1 += dpp_move<0x4E>(s1);
                s0 += dpp_move<0x141>(s0);
                s1 += dpp_move<0x141>(s1);
                const double f0 = tk * s0, f1 = tk * s1;
#pragma unroll
                for (int jh = 0; jh < 2; ++jh) {
                    if (128 * jh + 127 <= k || 128 * jh >= n) continue;
#pragma unroll
                    for (int j = 0; j < 8; ++j) {
                        const int c = 2 * q8 + 16 * (8 * jh + j);
                        double2 vj = *(const double2*)(v + min(c, sr - 2));
                        if (c >= n) vj = make_double2(0.0, 0.0);
                        const int i = 2 * (8 * jh + j);
                        z0[i] = fma(-f0, vj.x, z0[i]);
                        z0[i + 1] = fma(-f0, vj.y, z0[i + 1]);
                        z1[i] = fma(-f1, vj.x, z1[i]);
                        z1[i + 1] = fma(-f1, vj.y, z1[i + 1]);
                    }
                }
            }
        }
        __syncthreads();  // every row of this half is read before any result row is written
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            const int r = b ? r1 : r0;
            if (r >= n) continue;
            double* dst = Ct + (size_t)rk[r] * n;
#pragma unroll
            for (int i = 0; i < 32; ++i) {
                const int c = 2 * q8 + 16 * (i >> 1) + (i & 1);
                if (c < n) dst[c] = b ? z1[i] : z0[i];
            }
        }
    }
    __syncthreads();
    return 0;
}

}  // namespace am1g
