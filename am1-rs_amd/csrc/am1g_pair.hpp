// Per-atom-pair device routines: Slater overlap, the Dewar-Sabelli-Klopman two-centre integrals
// in their frame-free rotated form, and the core-core repulsion. One thread evaluates one pair;
// the callers group pairs by class (heavy-heavy, heavy-H, H-H) so a wavefront runs one code path.
//
// Behaviour follows am1-rs 0.2.3: overlap.rs:16-488 (closed forms for n <= 3),
// integrals.rs:207-569 (22/4/1 local integrals, axis n = -(R_b - R_a)/r, projector I - n n^T),
// repulsion.rs:144-188. Distances in Bohr, energies in eV.
#pragma once
#include <hip/hip_runtime.h>

#include "am1g_common.hpp"

namespace am1g {

// Lower-triangle pack index of an orbital pair inside a 4-orbital block (integrals.rs:23).
__host__ __device__ constexpr int pack4(int a, int b) {
    return a >= b ? a * (a + 1) / 2 + b : b * (b + 1) / 2 + a;
}

__device__ __forceinline__ double rsq(double x) { return 1.0 / sqrt(x); }

// A_k(x) = int_1^inf t^k e^{-xt} dt for k <= K (overlap.rs:357-368).
template <int K>
__device__ __forceinline__ void aux_a(double x, double* a) {
    if (fabs(x) < 1.0e-12) {
#pragma unroll
        for (int i = 0; i <= K; ++i) a[i] = 0.0;
        return;
    }
    double inv = 1.0 / x;
    a[0] = exp(-x) * inv;
#pragma unroll
    for (int i = 1; i <= K; ++i) a[i] = a[0] + a[i - 1] * inv * (double)i;
}

// B_k(x) = int_-1^1 t^k e^{-xt} dt for k <= K: upward closed form beyond |x| = 5, the
// all-same-sign Taylor series inside (overlap.rs:455-488).
template <int K>
__device__ __forceinline__ void aux_b(double x, double* b) {
#pragma unroll
    for (int i = 0; i <= K; ++i) b[i] = 0.0;
    if (fabs(x) > 5.0) {
        double inv = 1.0 / x;
        double tx = exp(x) * inv;
        double tmx = -exp(-x) * inv;
        b[0] = tx + tmx;
#pragma unroll
        for (int i = 1; i <= K; ++i) b[i] = ((i & 1) ? -tx : tx) + tmx + b[i - 1] * inv * (double)i;
        return;
    }
    double term = 1.0, magnitude = 1.0, absx = fabs(x);
    for (int n = 0; n < 64; ++n) {
#pragma unroll
        for (int k = 0; k <= K; ++k)
            if (((k ^ n) & 1) == 0) b[k] += term * (2.0 / (double)(k + n + 1));
        term = term * x * (-1.0 / (double)(n + 1));
        magnitude *= absx / (double)(n + 1);
        if (n >= 4 && magnitude < 1.0e-18) break;
    }
}

#include "../data/gauss_legendre.inc"

// Overlap of two normalized Slater functions (n_a, l_a, m) and (n_b, l_b, m), l <= 1 and m = 0
// (sigma) or 1 (pi), at distance r, by quadrature in prolate spheroidal coordinates (the general
// path of overlap_numeric.rs:25-105 for principal quantum numbers n >= 4): 48 Gauss-Legendre
// points in xi on [1, 1 + 60 / max(p, 0.05)], p = (za + zb) r / 2, and 40 in eta on [-1, 1];
// r_a = r (xi + eta) / 2, r_b = r (xi - eta) / 2; Jacobian (r / 2)^3 (xi^2 - eta^2).
__device__ double sto_overlap_numeric(int na, double za, int la, int m, int nb, double zb, int lb, double r) {
    double c;
    if (la == 0 && lb == 0) c = 0.5;
    else if (la != lb && m == 0) c = 0.86602540378443865;  // sqrt(3) / 2
    else if (la == 1 && lb == 1) c = m == 0 ? 1.5 : 0.75;
    else return 0.0;
    auto norm = [](int n, double z) {
        double f = 1.0;
        for (int k = 2; k <= 2 * n; ++k) f *= k;
        return pow(2.0 * z, n + 0.5) / sqrt(f);
    };
    const double h = 0.5 * r;
    const double p = (za + zb) * h;
    const double xi_max = 1.0 + 60.0 / fmax(p, 0.05);
    const double mid = 0.5 * (1.0 + xi_max), half = 0.5 * (xi_max - 1.0);
    double sum = 0.0;
    for (int i = 0; i < 48; ++i) {
        const double xi = mid + half * kGlXiNode[i];
        const double wx = half * kGlXiWeight[i];
        for (int j = 0; j < 40; ++j) {
            const double eta = kGlEtaNode[j];
            const double ra = h * (xi + eta), rb = h * (xi - eta);
            if (ra <= 0.0 || rb <= 0.0) continue;
            double ang = 1.0;
            if (la == 1) {
                const double ca = (1.0 + xi * eta) / (xi + eta);
                ang *= m == 0 ? ca : sqrt(fmax(0.0, 1.0 - ca * ca));
            }
            if (lb == 1) {
                const double cb = (xi * eta - 1.0) / (xi - eta);
                ang *= m == 0 ? cb : sqrt(fmax(0.0, 1.0 - cb * cb));
            }
            double radial = exp(-za * ra - zb * rb);
            for (int k = 1; k < na; ++k) radial *= ra;
            for (int k = 1; k < nb; ++k) radial *= rb;
            sum += radial * (wx * kGlEtaWeight[j] * ang * (xi * xi - eta * eta));
        }
    }
    return h * h * h * sum * norm(na, za) * norm(nb, zb) * c;
}

// Local overlaps (s111, s211, s121, s221, s222) for the ordered pair with ea.n >= eb.n: the
// closed forms for n <= 3 (overlap.rs:120-296), the quadrature above for n >= 4
// (overlap_numeric.rs:8-24: ss, p_sigma|s, -(s|p_sigma), -(p_sigma|p_sigma), p_pi|p_pi).
__device__ bool overlap_locals(const DevElem& ea, const DevElem& eb, double r, double s[5]) {
    if (ea.n > 3 || eb.n > 3) {
        s[0] = sto_overlap_numeric(ea.n, ea.zs, 0, 0, eb.n, eb.zs, 0, r);
        s[1] = sto_overlap_numeric(ea.n, ea.zp, 1, 0, eb.n, eb.zs, 0, r);
        s[2] = -sto_overlap_numeric(ea.n, ea.zs, 0, 0, eb.n, eb.zp, 1, r);
        s[3] = -sto_overlap_numeric(ea.n, ea.zp, 1, 0, eb.n, eb.zp, 1, r);
        s[4] = sto_overlap_numeric(ea.n, ea.zp, 1, 1, eb.n, eb.zp, 1, r);
        return true;
    }
    const double sq3 = 1.7320508075688772;
    const double za0 = ea.zs, za1 = ea.zp, zb0 = eb.zs, zb1 = eb.zp;
    s[0] = s[1] = s[2] = s[3] = s[4] = 0.0;
    int na = ea.n, nb = eb.n;
    if (na == 1 && nb == 1) {
        double a[3], b[3];
        aux_a<2>(r * 0.5 * (za0 + zb0), a);
        aux_b<2>(r * 0.5 * (za0 - zb0), b);
        s[0] = pow(r * r * (za0 * zb0), 1.5) * (a[2] * b[0] - b[2] * a[0]) / 4.0;
        return true;
    }
    if (na == 2 && nb == 1) {
        double a[4], b[4];
        double r4 = r * r * r * r;
        aux_a<3>(r * 0.5 * (za0 + zb0), a);
        aux_b<3>(r * 0.5 * (za0 - zb0), b);
        s[0] = r4 * (pow(zb0, 1.5) * pow(za0, 2.5)) * (a[3] * b[0] - b[3] * a[0] + a[2] * b[1] - b[2] * a[1]) /
               (sq3 * 8.0);
        aux_a<3>(r * 0.5 * (za1 + zb0), a);
        aux_b<3>(r * 0.5 * (za1 - zb0), b);
        s[1] = r4 * (pow(zb0, 1.5) * pow(za1, 2.5)) * (a[2] * b[0] - b[2] * a[0] + a[3] * b[1] - b[3] * a[1]) / 8.0;
        return true;
    }
    if (na == 2 && nb == 2) {
        double a[5], b[5];
        double r5 = r * r * r * r * r;
        aux_a<4>(r * 0.5 * (za0 + zb0), a);
        aux_b<4>(r * 0.5 * (za0 - zb0), b);
        s[0] = r5 * pow(zb0 * za0, 2.5) * (a[4] * b[0] + b[4] * a[0] - a[2] * b[2] * 2.0) / 48.0;
        aux_a<4>(r * 0.5 * (za1 + zb0), a);
        aux_b<4>(r * 0.5 * (za1 - zb0), b);
        s[1] = r5 * pow(zb0 * za1, 2.5) *
               (a[3] * (b[0] - b[2]) - a[1] * (b[2] - b[4]) + b[3] * (a[0] - a[2]) - b[1] * (a[2] - a[4])) /
               (16.0 * sq3);
        aux_a<4>(r * 0.5 * (za0 + zb1), a);
        aux_b<4>(r * 0.5 * (za0 - zb1), b);
        s[2] = r5 * pow(zb1 * za0, 2.5) *
               (a[3] * (b[0] - b[2]) - a[1] * (b[2] - b[4]) - b[3] * (a[0] - a[2]) + b[1] * (a[2] - a[4])) /
               (16.0 * sq3);
        aux_a<4>(r * 0.5 * (za1 + zb1), a);
        aux_b<4>(r * 0.5 * (za1 - zb1), b);
        double w = r5 * pow(zb1 * za1, 2.5) / 16.0;
        s[3] = -(w * (b[2] * (a[4] + a[0]) - a[2] * (b[4] + b[0])));
        s[4] = w * (a[4] * (b[0] - b[2]) - b[4] * (a[0] - a[2]) - a[2] * b[0] + b[2] * a[0]) * 0.5;
        return true;
    }
    if (na == 3 && nb == 1) {
        double a[5], b[5];
        double r5 = r * r * r * r * r;
        aux_a<4>(r * 0.5 * (za0 + zb0), a);
        aux_b<4>(r * 0.5 * (za0 - zb0), b);
        s[0] = r5 * (pow(zb0, 1.5) * pow(za0, 3.5)) *
               (a[4] * b[0] + b[1] * a[3] * 2.0 - a[1] * b[3] * 2.0 - b[4] * a[0]) / (sqrt(10.0) * 24.0);
        aux_a<4>(r * 0.5 * (za1 + zb0), a);
        aux_b<4>(r * 0.5 * (za1 - zb0), b);
        s[1] = r5 * (pow(zb0, 1.5) * pow(za1, 3.5)) *
               (a[3] * (b[0] + b[2]) - a[1] * (b[4] + b[2]) + b[1] * (a[2] + a[4]) - b[3] * (a[2] + a[0])) /
               (8.0 * sqrt(30.0));
        return true;
    }
    if (na == 3 && nb == 2) {
        double a[6], b[6];
        double r6 = r * r * r * r * r * r;
        aux_a<5>(r * 0.5 * (za0 + zb0), a);
        aux_b<5>(r * 0.5 * (za0 - zb0), b);
        s[0] = r6 * (pow(zb0, 2.5) * pow(za0, 3.5)) *
               (a[5] * b[0] + b[1] * a[4] - b[2] * a[3] * 2.0 - a[2] * b[3] * 2.0 + b[4] * a[1] + b[5] * a[0]) /
               (sqrt(30.0) * 48.0);
        aux_a<5>(r * 0.5 * (za1 + zb0), a);
        aux_b<5>(r * 0.5 * (za1 - zb0), b);
        s[1] = r6 * (pow(zb0, 2.5) * pow(za1, 3.5)) *
               (a[4] * b[0] + b[1] * a[5] - b[3] * a[3] * 2.0 - a[2] * b[2] * 2.0 + a[1] * b[5] + a[0] * b[4]) /
               (48.0 * sqrt(10.0));
        aux_a<5>(r * 0.5 * (za0 + zb1), a);
        aux_b<5>(r * 0.5 * (za0 - zb1), b);
        s[2] = r6 * (pow(zb1, 2.5) * pow(za0, 3.5)) *
               ((a[4] * b[0] - a[5] * b[1]) + (a[3] * b[1] - a[4] * b[2]) * 2.0 - (a[1] * b[3] - a[2] * b[4]) * 2.0 -
                (a[0] * b[4] - a[1] * b[5])) /
               (48.0 * sqrt(10.0));
        aux_a<5>(r * 0.5 * (za1 + zb1), a);
        aux_b<5>(r * 0.5 * (za1 - zb1), b);
        double pre = r6 * (pow(zb1, 2.5) * pow(za1, 3.5));
        s[3] = pre *
               ((a[3] * b[0] - a[5] * b[2]) + (a[2] * b[1] - a[4] * b[3]) - (a[1] * b[2] - a[3] * b[4]) -
                (a[0] * b[3] - a[2] * b[5])) /
               (16.0 * sqrt(30.0));
        s[4] = pre *
               ((a[5] - a[3]) * (b[0] - b[2]) + (a[4] - a[2]) * (b[1] - b[3]) - (a[3] - a[1]) * (b[2] - b[4]) -
                (a[2] - a[0]) * (b[3] - b[5])) /
               (32.0 * sqrt(30.0));
        return true;
    }
    if (na == 3 && nb == 3) {
        double a[7], b[7];
        double r7 = r * r * r * r * r * r * r;
        aux_a<6>(r * 0.5 * (za0 + zb0), a);
        aux_b<6>(r * 0.5 * (za0 - zb0), b);
        s[0] = r7 * pow(zb0 * za0, 3.5) * (a[6] * b[0] - b[2] * a[4] * 3.0 + a[2] * b[4] * 3.0 - a[0] * b[6]) /
               1440.0;
        aux_a<6>(r * 0.5 * (za1 + zb0), a);
        aux_b<6>(r * 0.5 * (za1 - zb0), b);
        s[1] = r7 * pow(zb0 * za1, 3.5) *
               ((a[5] * b[0] + a[6] * b[1]) + (-a[4] * b[1] - a[5] * b[2]) - (a[3] * b[2] + a[4] * b[3]) * 2.0 -
                (-a[2] * b[3] - a[3] * b[4]) * 2.0 + (a[1] * b[4] + a[2] * b[5]) + (-a[0] * b[5] - a[1] * b[6])) /
               (480.0 * sq3);
        aux_a<6>(r * 0.5 * (za0 + zb1), a);
        aux_b<6>(r * 0.5 * (za0 - zb1), b);
        s[2] = r7 * pow(zb1 * za0, 3.5) *
               ((a[5] * b[0] - a[6] * b[1]) + (a[4] * b[1] - a[5] * b[2]) - (a[3] * b[2] - a[4] * b[3]) * 2.0 -
                (a[2] * b[3] - a[3] * b[4]) * 2.0 + (a[1] * b[4] - a[2] * b[5]) + (a[0] * b[5] - a[1] * b[6])) /
               (480.0 * sq3);
        aux_a<6>(r * 0.5 * (za1 + zb1), a);
        aux_b<6>(r * 0.5 * (za1 - zb1), b);
        double pre = r7 * (pow(zb1, 3.5) * pow(za1, 3.5));
        s[3] = pre *
               ((a[4] * b[0] - a[6] * b[2]) - (a[2] * b[2] - a[4] * b[4]) * 2.0 + (a[0] * b[4] - a[2] * b[6])) /
               480.0;
        s[4] = pre *
               ((a[6] - a[4]) * (b[0] - b[2]) - (a[4] - a[2]) * (b[2] - b[4]) * 2.0 + (a[2] - a[0]) * (b[4] - b[6])) /
               960.0;
        return true;
    }
    return false;
}

// 4x4 overlap block S[i][j] (row: orbital of atom a, column: orbital of atom b) for the
// displacement d = R_b - R_a (overlap.rs:16-40, build_di_g :301-322).
__device__ bool overlap_block(const DevElem& ea, const DevElem& eb, const double d[3], double S[16]) {
    double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
#pragma unroll
    for (int k = 0; k < 16; ++k) S[k] = 0.0;
    if (r < 1.0e-8) {
        S[0] = S[5] = S[10] = S[15] = 1.0;
        return true;
    }
    bool swap = ea.n < eb.n;
    double loc[5];
    double sgn = swap ? -1.0 : 1.0;
    double dir[3] = {sgn * d[0] / r, sgn * d[1] / r, sgn * d[2] / r};
    if (!(swap ? overlap_locals(eb, ea, r, loc) : overlap_locals(ea, eb, r, loc))) return false;
    // block of the higher-n atom (rows) against the other (columns)
    double di[16];
    di[0] = loc[0];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        di[(1 + k) * 4] = loc[1] * dir[k];
        di[1 + k] = -loc[2] * dir[k];
    }
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int l = 0; l < 3; ++l)
            di[(1 + k) * 4 + 1 + l] = -loc[3] * dir[k] * dir[l] + loc[4] * ((k == l ? 1.0 : 0.0) - dir[k] * dir[l]);
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) S[i * 4 + j] = swap ? di[j * 4 + i] : di[i * 4 + j];
    return true;
}

// H-H two-electron integral (ss|ss) = 27.21 / sqrt(r^2 + (rho0_a + rho0_b)^2) (integrals.rs:226-239).
__device__ __forceinline__ double hh_integral(const DevElem& ea, const DevElem& eb, double r) {
    double s = ea.rho0 + eb.rho0;
    return kEv * rsq(r * r + s * s);
}

// Heavy-H block: w[pack(a,b)] (10 values) for the heavy atom a (integrals.rs:309-329, 487-497).
__device__ void heavy_h_block(const DevElem& ea, const DevElem& eb, const double d[3], double w[10]) {
    const double ev1 = kEv / 2.0, ev2 = kEv / 4.0;
    double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    double n[3] = {-d[0] / r, -d[1] / r, -d[2] / r};
    double da = ea.dd, qa = ea.qq * 2.0;
    double t;
    t = ea.rho0 + eb.rho0; double aee = t * t;
    t = ea.rho1 + eb.rho0; double ade = t * t;
    t = ea.rho2 + eb.rho0; double aqe = t * t;
    double ee = rsq(r * r + aee) * kEv;
    double ev1dsqr6 = rsq(r * r + aqe) * ev1;
    double ri1 = rsq((r + da) * (r + da) + ade) * ev1 - rsq((r - da) * (r - da) + ade) * ev1;
    double ri2 = ee + rsq((r + qa) * (r + qa) + aqe) * ev2 + rsq((r - qa) * (r - qa) + aqe) * ev2 - ev1dsqr6;
    double ri3 = ee + rsq(r * r + qa * qa + aqe) * ev1 - ev1dsqr6;
    w[0] = ee;
#pragma unroll
    for (int k = 0; k < 3; ++k) w[pack4(k + 1, 0)] = ri1 * n[k];
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int l = 0; l <= k; ++l) w[pack4(k + 1, l + 1)] = ri2 * n[k] * n[l] + ri3 * ((k == l ? 1.0 : 0.0) - n[k] * n[l]);
}

// The 22 local heavy-heavy integrals (integrals.rs:331-484).
__device__ void local_heavy_heavy(const DevElem& ea, const DevElem& eb, double r, double ri[22]) {
    const double ev1 = kEv / 2.0, ev2 = kEv / 4.0, ev3 = kEv / 8.0, ev4 = kEv / 16.0;
    const double da = ea.dd, db = eb.dd, qa = ea.qq * 2.0, qb = eb.qq * 2.0, qa1 = ea.qq, qb1 = eb.qq;
    double t;
    t = ea.rho0 + eb.rho0; const double aee = t * t;
    t = ea.rho1 + eb.rho0; const double ade = t * t;
    t = ea.rho2 + eb.rho0; const double aqe = t * t;
    t = ea.rho0 + eb.rho1; const double aed = t * t;
    t = ea.rho0 + eb.rho2; const double aeq = t * t;
    t = ea.rho1 + eb.rho1; const double axx = t * t;
    t = ea.rho1 + eb.rho2; const double adq = t * t;
    t = ea.rho2 + eb.rho1; const double aqd = t * t;
    t = ea.rho2 + eb.rho2; const double aqq = t * t;
    const double r2 = r * r;
#define SQ2(x, c) rsq((x) * (x) + (c))
    const double ee = rsq(r2 + aee) * kEv;
    const double dze = SQ2(r - da, ade) * ev1 - SQ2(r + da, ade) * ev1;
    const double ev1dsqr6 = rsq(r2 + aqe) * ev1;
    const double qzze = SQ2(r - qa, aqe) * ev2 + SQ2(r + qa, aqe) * ev2 - ev1dsqr6;
    const double qxxe = rsq(r2 + qa * qa + aqe) * ev1 - ev1dsqr6;
    const double edz = SQ2(r + db, aed) * ev1 - SQ2(r - db, aed) * ev1;
    const double ev1dsqr12 = rsq(r2 + aeq) * ev1;
    const double eqzz = SQ2(r - qb, aeq) * ev2 + SQ2(r + qb, aeq) * ev2 - ev1dsqr12;
    const double eqxx = rsq(r2 + qb * qb + aeq) * ev1 - ev1dsqr12;
    const double ev2dsqr20 = SQ2(r + da, adq) * ev2;
    const double ev2dsqr22 = SQ2(r - da, adq) * ev2;
    const double ev2dsqr24 = SQ2(r - db, aqd) * ev2;
    const double ev2dsqr26 = SQ2(r + db, aqd) * ev2;
    const double ev2dsqr36 = rsq(r2 + aqq) * ev2;
    const double ev2dsqr39 = rsq(r2 + qa * qa + aqq) * ev2;
    const double ev2dsqr40 = rsq(r2 + qb * qb + aqq) * ev2;
    const double ev3dsqr42 = SQ2(r - qb, aqq) * ev3;
    const double ev3dsqr44 = SQ2(r + qb, aqq) * ev3;
    const double ev3dsqr46 = SQ2(r + qa, aqq) * ev3;
    const double ev3dsqr48 = SQ2(r - qa, aqq) * ev3;
    ri[0] = ee;
    ri[1] = -dze;
    ri[2] = ee + qzze;
    ri[3] = ee + qxxe;
    ri[4] = -edz;
    ri[5] = SQ2(r + da - db, axx) * ev2 + SQ2(r - da + db, axx) * ev2 - SQ2(r - da - db, axx) * ev2 -
            SQ2(r + da + db, axx) * ev2;
    ri[6] = rsq(r2 + (da - db) * (da - db) + axx) * ev1 - rsq(r2 + (da + db) * (da + db) + axx) * ev1;
    ri[7] = -edz + SQ2(r + qa - db, aqd) * ev3 - SQ2(r + qa + db, aqd) * ev3 + SQ2(r - qa - db, aqd) * ev3 -
            SQ2(r - qa + db, aqd) * ev3 - ev2dsqr24 + ev2dsqr26;
    ri[8] = -edz - ev2dsqr24 + rsq((r - db) * (r - db) + qa * qa + aqd) * ev2 + ev2dsqr26 -
            rsq((r + db) * (r + db) + qa * qa + aqd) * ev2;
    ri[9] = rsq((qa1 - db) * (qa1 - db) + (r + qa1) * (r + qa1) + aqd) * ev2 -
            rsq((qa1 - db) * (qa1 - db) + (r - qa1) * (r - qa1) + aqd) * ev2 -
            rsq((qa1 + db) * (qa1 + db) + (r + qa1) * (r + qa1) + aqd) * ev2 +
            rsq((qa1 + db) * (qa1 + db) + (r - qa1) * (r - qa1) + aqd) * ev2;
    ri[10] = ee + eqzz;
    ri[11] = ee + eqxx;
    ri[12] = -dze + SQ2(r + da - qb, adq) * ev3 - SQ2(r - da - qb, adq) * ev3 + SQ2(r + da + qb, adq) * ev3 -
             SQ2(r - da + qb, adq) * ev3 + ev2dsqr22 - ev2dsqr20;
    ri[13] = -dze - ev2dsqr20 + rsq((r + da) * (r + da) + qb * qb + adq) * ev2 + ev2dsqr22 -
             rsq((r - da) * (r - da) + qb * qb + adq) * ev2;
    ri[14] = rsq((da - qb1) * (da - qb1) + (r - qb1) * (r - qb1) + adq) * ev2 -
             rsq((da - qb1) * (da - qb1) + (r + qb1) * (r + qb1) + adq) * ev2 -
             rsq((da + qb1) * (da + qb1) + (r - qb1) * (r - qb1) + adq) * ev2 +
             rsq((da + qb1) * (da + qb1) + (r + qb1) * (r + qb1) + adq) * ev2;
    ri[15] = ee + eqzz + qzze + SQ2(r + qa - qb, aqq) * ev4 + SQ2(r + qa + qb, aqq) * ev4 +
             SQ2(r - qa - qb, aqq) * ev4 + SQ2(r - qa + qb, aqq) * ev4 - ev3dsqr48 - ev3dsqr46 - ev3dsqr42 -
             ev3dsqr44 + ev2dsqr36;
    ri[16] = ee + eqzz + qxxe + rsq((r - qb) * (r - qb) + qa * qa + aqq) * ev3 +
             rsq((r + qb) * (r + qb) + qa * qa + aqq) * ev3 - ev3dsqr42 - ev3dsqr44 - ev2dsqr39 + ev2dsqr36;
    ri[17] = ee + eqxx + qzze + rsq((r + qa) * (r + qa) + qb * qb + aqq) * ev3 +
             rsq((r - qa) * (r - qa) + qb * qb + aqq) * ev3 - ev3dsqr46 - ev3dsqr48 - ev2dsqr40 + ev2dsqr36;
    const double qxxqxx = rsq(r2 + (qa - qb) * (qa - qb) + aqq) * ev3 + rsq(r2 + (qa + qb) * (qa + qb) + aqq) * ev3 -
                          ev2dsqr39 - ev2dsqr40 + ev2dsqr36;
    ri[18] = ee + eqxx + qxxe + qxxqxx;
    const double s1 = (qa1 - qb1) * (qa1 - qb1), s2 = (qa1 + qb1) * (qa1 + qb1);
    ri[19] = rsq((r + qa1 - qb1) * (r + qa1 - qb1) + s1 + aqq) * ev3 -
             rsq((r + qa1 + qb1) * (r + qa1 + qb1) + s1 + aqq) * ev3 -
             rsq((r - qa1 - qb1) * (r - qa1 - qb1) + s1 + aqq) * ev3 +
             rsq((r - qa1 + qb1) * (r - qa1 + qb1) + s1 + aqq) * ev3 -
             rsq((r + qa1 - qb1) * (r + qa1 - qb1) + s2 + aqq) * ev3 +
             rsq((r + qa1 + qb1) * (r + qa1 + qb1) + s2 + aqq) * ev3 +
             rsq((r - qa1 - qb1) * (r - qa1 - qb1) + s2 + aqq) * ev3 -
             rsq((r - qa1 + qb1) * (r - qa1 + qb1) + s2 + aqq) * ev3;
    const double qxxqyy = rsq(r2 + qa * qa + qb * qb + aqq) * ev2 - ev2dsqr39 - ev2dsqr40 + ev2dsqr36;
    ri[20] = ee + eqxx + qxxe + qxxqyy;
    ri[21] = (qxxqxx - qxxqyy) * 0.5;
#undef SQ2
}

// Rotated heavy-heavy block (integrals.rs:515-569 in axis/projector form), each entry handed to
// sink(pack_a * 10 + pack_b, value). Fully unrolled: every branch is resolved at compile time,
// so one pair costs the same straight-line code for every orientation.
template <class Sink>
__device__ __forceinline__ void heavy_heavy_visit(const DevElem& ea, const DevElem& eb, const double d[3], Sink&& sink) {
    double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    double n[3] = {-d[0] / r, -d[1] / r, -d[2] / r};
    double p[3][3];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int b = 0; b < 3; ++b) p[a][b] = (a == b ? 1.0 : 0.0) - n[a] * n[b];
    double ri[22];
    local_heavy_heavy(ea, eb, r, ri);
#pragma unroll
    for (int kk = 0; kk < 4; ++kk)
#pragma unroll
        for (int ll = 0; ll <= kk; ++ll)
#pragma unroll
            for (int mm = 0; mm < 4; ++mm)
#pragma unroll
                for (int nn = 0; nn <= mm; ++nn) {
                    const int k = kk - 1, l = ll - 1, m = mm - 1, q = nn - 1;
                    double v;
                    if (kk == 0) {
                        if (mm == 0) v = ri[0];
                        else if (nn == 0) v = ri[4] * n[m];
                        else v = ri[10] * n[m] * n[q] + ri[11] * p[m][q];
                    } else if (ll == 0) {
                        if (mm == 0) v = ri[1] * n[k];
                        else if (nn == 0) v = ri[5] * n[k] * n[m] + ri[6] * p[k][m];
                        else
                            v = ri[12] * n[k] * n[m] * n[q] + ri[13] * p[m][q] * n[k] +
                                ri[14] * (p[k][q] * n[m] + p[k][m] * n[q]);
                    } else if (mm == 0) {
                        v = ri[2] * n[k] * n[l] + ri[3] * p[k][l];
                    } else if (nn == 0) {
                        v = ri[7] * n[k] * n[l] * n[m] + ri[8] * p[k][l] * n[m] + ri[9] * (p[l][m] * n[k] + p[k][m] * n[l]);
                    } else {
                        v = ri[15] * n[k] * n[l] * n[m] * n[q] + ri[16] * p[k][l] * n[m] * n[q] +
                            ri[17] * p[m][q] * n[k] * n[l] +
                            ri[19] * (n[k] * (p[l][q] * n[m] + p[l][m] * n[q]) + n[l] * (p[k][q] * n[m] + p[k][m] * n[q])) +
                            ri[20] * p[k][l] * p[m][q] + ri[21] * (p[k][m] * p[l][q] + p[k][q] * p[l][m]);
                    }
                    sink(pack4(kk, ll) * 10 + pack4(mm, nn), v);
                }
}

// The rotated heavy-heavy block written straight to global memory, w[pack_a * 10 + pack_b].
__device__ void heavy_heavy_block(const DevElem& ea, const DevElem& eb, const double d[3], double* __restrict__ w) {
    heavy_heavy_visit(ea, eb, d, [&](int idx, double v) { w[idx] = v; });
}

// Core-core repulsion of one atom pair, eV (repulsion.rs:144-188); za/zb are atomic numbers.
__device__ double core_core_pair(const DevElem& ea, const DevElem& eb, int za, int zb, double r) {
    double s = r * kA0;
    double rho = ea.rho0 + eb.rho0;
    double gam = kEv * rsq(r * r + rho * rho);
    double zz = ea.core * eb.core;
    double fa = exp(-ea.alpha * s);
    if ((za == 7 || za == 8) && zb == 1) fa *= s;
    double fb = exp(-eb.alpha * s);
    if ((zb == 7 || zb == 8) && za == 1) fb *= s;
    double g = 0.0;
    for (int k = 0; k < ea.ng; ++k) g += ea.gk[k] * exp(-ea.gl[k] * (s - ea.gm[k]) * (s - ea.gm[k]));
    for (int k = 0; k < eb.ng; ++k) g += eb.gk[k] * exp(-eb.gl[k] * (s - eb.gm[k]) * (s - eb.gm[k]));
    return gam * (fa + fb + 1.0) * zz + g * zz / s;
}

// One-centre two-electron integral (ab|cd) on one atom (fock.rs:27-65).
__device__ __forceinline__ double one_centre(int a, int b, int c, int d, const DevElem& e) {
    if (a == b && c == d) {
        if (a == 0 && c == 0) return e.gss;
        if (a == 0 || c == 0) return e.gsp;
        return a == c ? e.gpp : e.gp2;
    }
    int ba = min(a, b), bb = max(a, b), kc = min(c, d), kd = max(c, d);
    if (ba == 0 && bb != 0 && kc == 0 && kd != 0 && bb == kd) return e.hsp;
    if (ba != 0 && bb != 0 && ba != bb && ba == kc && bb == kd) return 0.5 * (e.gpp - e.gp2);
    return 0.0;
}

// ------------------------------------------------------------------ pair enumeration
// Pairs are grouped by class inside a molecule's integral pool:
//   [heavy-heavy, 100 each][heavy-H, 10 each][H-H, 1 each]
// heavy-heavy pairs (i < j in the heavy list) and H-H pairs (i < j in the H list) in row-major
// triangular order, heavy-H pairs as heavy_index * n_h + h_index.

__host__ __device__ __forceinline__ int64_t tri_index(int i, int j, int n) {  // i < j < n
    return (int64_t)i * n - (int64_t)i * (i + 1) / 2 + (j - i - 1);
}

// Inverse of tri_index: k -> (i, j), i < j.
__device__ __forceinline__ void tri_decode(int64_t k, int n, int& i, int& j) {
    double nn = (double)(2 * n - 1);
    int ii = (int)floor((nn - sqrt(nn * nn - 8.0 * (double)k)) * 0.5);
    if (ii < 0) ii = 0;
    while (ii > 0 && tri_index(ii, ii + 1, n) > k) --ii;
    while (ii + 1 < n - 1 && tri_index(ii + 1, ii + 2, n) <= k) ++ii;
    i = ii;
    j = (int)(k - ((int64_t)ii * n - (int64_t)ii * (ii + 1) / 2)) + ii + 1;
}

}  // namespace am1g
