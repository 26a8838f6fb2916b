// Nuclear gradient of the converged AM1/RM1 SCF energy on the device.
//
// At a converged (variational) density the gradient is the derivative of the energy at fixed
// density, which for NDDO is a sum over atom pairs (gradient.rs:336-470): for the pair (A, B)
// with d = R_B - R_A,
//   E_AB(d) = sum_ij Pt_ij (beta_i + beta_j) S_ij(d)                       resonance
//           + sum_{mu nu in A} Pt_mu_nu e1b_mu_nu(d) + sum_{la si in B} Pt_la_si e2a_la_si(d)
//           + sum (Pt_mu_nu Pt_la_si - Pa_mu_la Pa_nu_si - Pb_mu_la Pb_nu_si) (mu nu|la si)(d)
//           + E_core,AB(|d|),
// and dE/dR_B += dE_AB/dd, dE/dR_A -= dE_AB/dd. The reference's analytic_gradient
// (gradient.rs:41-90) differentiates the fixed-density energy by central differences; here each
// pair's energy is differentiated along each axis by the five-point central difference
// (8 (E(+h) - E(-h)) - (E(+2h) - E(-2h))) / 12h, h = kGradStep Bohr (error O(h^4)), every
// displaced pair block evaluated by the same device routines as the SCF's integrals and
// contracted on the fly with density coefficients prepared once per pair.
//
// Kernels: k_grad_coef (one thread per pair: the coefficients C[p][q] of the packed two-electron
// block, with the e1b / e2a terms folded into its (., ss) and (ss, .) entries), k_grad_pairs (one
// thread per pair: dE_AB/dd) and k_grad_atoms (one thread per atom: its pairs summed in a fixed
// order, so the result does not depend on scheduling).
#include <hip/hip_runtime.h>

#include <chrono>
#include <climits>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "am1g_ctx.hpp"
#include "am1g_pair.hpp"

namespace am1g {

constexpr double kGradStep = 1.0e-4;  // Bohr

struct GradArgs {
    MolDesc md;
    const int* z;
    const double* xyz;
    const int* alist;  // heavy atoms then hydrogens, input order within each
    const int* aoff;   // AO offset of each atom
    const int* acls;   // index inside the heavy or the hydrogen list
    const DevElem* elems;
    const double* pa;  // alpha density (RHF: half the total), nao x nao
    const double* pb;  // beta density
    double* coef;      // per pair, in the integral-pool layout (100 / 10 / 1 doubles)
    double* fpair;     // [n_pairs][3]: dE_AB / dd
    double* grad;      // [nat][3]
    int64_t n_pairs;
};

__device__ __forceinline__ void grad_decode(const GradArgs& g, int64_t k_all, int& a, int& b, int& cls) {
    const int nh = g.md.n_heavy, nH = g.md.n_h;
    const int64_t n_hh = (int64_t)nh * (nh - 1) / 2, n_hH = (int64_t)nh * nH;
    if (k_all < n_hh) {
        int i, j;
        tri_decode(k_all, nh, i, j);
        a = g.alist[i];
        b = g.alist[j];
        cls = 0;
    } else if (k_all < n_hh + n_hH) {
        const int64_t k = k_all - n_hh;
        a = g.alist[(int)(k / nH)];
        b = g.alist[nh + (int)(k % nH)];
        cls = 1;
    } else {
        int i, j;
        tri_decode(k_all - n_hh - n_hH, nH, i, j);
        a = g.alist[nh + i];
        b = g.alist[nh + j];
        cls = 2;
    }
}

__device__ __forceinline__ int64_t grad_coef_off(const GradArgs& g, int64_t k_all) {
    const int nh = g.md.n_heavy, nH = g.md.n_h;
    const int64_t n_hh = (int64_t)nh * (nh - 1) / 2, n_hH = (int64_t)nh * nH;
    if (k_all < n_hh) return k_all * 100;
    if (k_all < n_hh + n_hH) return n_hh * 100 + (k_all - n_hh) * 10;
    return n_hh * 100 + n_hH * 10 + (k_all - n_hh - n_hH);
}

__device__ __constant__ int kGradPackRow[10] = {0, 1, 1, 2, 2, 2, 3, 3, 3, 3};
__device__ __constant__ int kGradPackCol[10] = {0, 0, 1, 0, 1, 2, 0, 1, 2, 3};

// Coefficient of packed pair p = (m1 >= m2) on atom A and q = (l1 >= l2) on atom B: the sum over
// both orderings of each (the block is symmetric in mu nu and in la si) of
// Pt_AA Pt_BB - Pa_AB Pa_AB - Pb_AB Pb_AB.
__device__ double grad_c2(const GradArgs& g, int oa, int ob, int m1, int m2, int l1, int l2) {
    const int n = g.md.nao;
    const double* pa = g.pa;
    const double* pb = g.pb;
    auto pt = [&](int r, int c) { return pa[(size_t)r * n + c] + pb[(size_t)r * n + c]; };
    double acc = 0.0;
    for (int o1 = 0; o1 < (m1 == m2 ? 1 : 2); ++o1) {
        const int mu = o1 ? m2 : m1, nu = o1 ? m1 : m2;
        for (int o2 = 0; o2 < (l1 == l2 ? 1 : 2); ++o2) {
            const int la = o2 ? l2 : l1, si = o2 ? l1 : l2;
            acc += pt(oa + mu, oa + nu) * pt(ob + la, ob + si);
            acc -= pa[(size_t)(oa + mu) * n + ob + la] * pa[(size_t)(oa + nu) * n + ob + si];
            acc -= pb[(size_t)(oa + mu) * n + ob + la] * pb[(size_t)(oa + nu) * n + ob + si];
        }
    }
    return acc;
}

// sum over both orderings of Pt on one atom's packed pair (m1 >= m2)
__device__ double grad_c1(const GradArgs& g, int o, int m1, int m2) {
    const int n = g.md.nao;
    const double v = g.pa[(size_t)(o + m1) * n + o + m2] + g.pb[(size_t)(o + m1) * n + o + m2];
    return m1 == m2 ? v : 2.0 * v;
}

__global__ __launch_bounds__(256) void k_grad_coef(GradArgs g) {
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < g.n_pairs;
         p += (int64_t)gridDim.x * blockDim.x) {
        int a, b, cls;
        grad_decode(g, p, a, b, cls);
        const DevElem* el = g.elems + g.md.method * kMaxZ;
        const double za = el[g.z[a]].core, zb = el[g.z[b]].core;
        const int oa = g.aoff[a], ob = g.aoff[b];
        double* c = g.coef + grad_coef_off(g, p);
        if (cls == 0) {
            for (int pp = 0; pp < 10; ++pp)
                for (int qq = 0; qq < 10; ++qq) {
                    const int m1 = kGradPackRow[pp], m2 = kGradPackCol[pp];
                    const int l1 = kGradPackRow[qq], l2 = kGradPackCol[qq];
                    double v = grad_c2(g, oa, ob, m1, m2, l1, l2);
                    if (qq == 0) v -= zb * grad_c1(g, oa, m1, m2);  // e1b = -Z_B (mu nu|ss)
                    if (pp == 0) v -= za * grad_c1(g, ob, l1, l2);  // e2a = -Z_A (ss|la si)
                    c[pp * 10 + qq] = v;
                }
        } else if (cls == 1) {
            for (int pp = 0; pp < 10; ++pp) {
                const int m1 = kGradPackRow[pp], m2 = kGradPackCol[pp];
                double v = grad_c2(g, oa, ob, m1, m2, 0, 0) - zb * grad_c1(g, oa, m1, m2);
                if (pp == 0) v -= za * grad_c1(g, ob, 0, 0);
                c[pp] = v;
            }
        } else {
            c[0] = grad_c2(g, oa, ob, 0, 0, 0, 0) - zb * grad_c1(g, oa, 0, 0) - za * grad_c1(g, ob, 0, 0);
        }
    }
}

// Energy of the pair at fixed density for the displacement d = R_B - R_A.
__device__ double grad_pair_energy(const DevElem& ea, const DevElem& eb, int za, int zb, int cls, const double d[3],
                                   const double* __restrict__ c, const double cres[16]) {
    double S[16];
    overlap_block(ea, eb, d, S);
    double e = 0.0;
#pragma unroll
    for (int k = 0; k < 16; ++k) e = fma(cres[k], S[k], e);
    const double r = sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
    if (cls == 0) {
        double acc = 0.0;
        heavy_heavy_visit(ea, eb, d, [&](int idx, double v) { acc = fma(v, c[idx], acc); });
        e += acc;
    } else if (cls == 1) {
        double w[10];
        heavy_h_block(ea, eb, d, w);
#pragma unroll
        for (int q = 0; q < 10; ++q) e = fma(w[q], c[q], e);
    } else {
        e = fma(hh_integral(ea, eb, r), c[0], e);
    }
    return e + core_core_pair(ea, eb, za, zb, r);
}

__global__ __launch_bounds__(128) void k_grad_pairs(GradArgs g) {
    for (int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; p < g.n_pairs;
         p += (int64_t)gridDim.x * blockDim.x) {
        int a, b, cls;
        grad_decode(g, p, a, b, cls);
        const DevElem* el = g.elems + g.md.method * kMaxZ;
        const int za = g.z[a], zb = g.z[b];
        const DevElem& ea = el[za];
        const DevElem& eb = el[zb];
        const int oa = g.aoff[a], ob = g.aoff[b], n = g.md.nao;
        const double* xa = g.xyz + 3 * (size_t)a;
        const double* xb = g.xyz + 3 * (size_t)b;
        {
            const double d0 = xb[0] - xa[0], d1 = xb[1] - xa[1], d2 = xb[2] - xa[2];
            const double r2 = d0 * d0 + d1 * d1 + d2 * d2;
            if (g.md.far_cut > 0.0 && r2 > g.md.far_cut * g.md.far_cut) {
                // a far-field pair: its energy is the monopole Q_A Q_B 27.21 / sqrt(r^2 + eta^2)
                // at the converged net charges (FarField::gradient, farfield.rs:178-206)
                double pop_a = 0.0, pop_b = 0.0;
                for (int k = 0; k < ea.norb; ++k) pop_a += g.pa[(size_t)(oa + k) * (n + 1)] + g.pb[(size_t)(oa + k) * (n + 1)];
                for (int k = 0; k < eb.norb; ++k) pop_b += g.pa[(size_t)(ob + k) * (n + 1)] + g.pb[(size_t)(ob + k) * (n + 1)];
                const double eta = ea.rho0 + eb.rho0;
                const double sr = sqrt(r2 + eta * eta);
                const double f = -kEv * (ea.core - pop_a) * (eb.core - pop_b) / (sr * sr * sr);
                g.fpair[3 * p] = f * d0;
                g.fpair[3 * p + 1] = f * d1;
                g.fpair[3 * p + 2] = f * d2;
                continue;
            }
        }
        // resonance coefficients: Pt_ij (beta_i + beta_j), both triangles of H_core
        double cres[16];
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                double v = 0.0;
                if (i < ea.norb && j < eb.norb) {
                    const size_t ij = (size_t)(oa + i) * n + ob + j;
                    v = (g.pa[ij] + g.pb[ij]) * ((i == 0 ? ea.bs : ea.bp) + (j == 0 ? eb.bs : eb.bp));
                }
                cres[i * 4 + j] = v;
            }
        const double* c = g.coef + grad_coef_off(g, p);
        for (int k = 0; k < 3; ++k) {
            double e[4];
#pragma unroll 1
            for (int q = 0; q < 4; ++q) {
                double dd[3] = {xb[0] - xa[0], xb[1] - xa[1], xb[2] - xa[2]};
                dd[k] += (q & 1 ? -1.0 : 1.0) * (q < 2 ? 1.0 : 2.0) * kGradStep;
                e[q] = grad_pair_energy(ea, eb, za, zb, cls, dd, c, cres);
            }
            g.fpair[3 * p + k] = (8.0 * (e[0] - e[1]) - (e[2] - e[3])) / (12.0 * kGradStep);
        }
    }
}

__global__ __launch_bounds__(256) void k_grad_atoms(GradArgs g) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int nat = g.md.nat;
    if (x >= nat) return;
    const int nh = g.md.n_heavy, nH = g.md.n_h;
    const int64_t n_hh = (int64_t)nh * (nh - 1) / 2, n_hH = (int64_t)nh * nH;
    const bool hx = g.z[x] != 1;
    const int cx = g.acls[x];
    double acc[3] = {0.0, 0.0, 0.0};
    for (int y = 0; y < nat; ++y) {
        if (y == x) continue;
        const bool hy = g.z[y] != 1;
        const int cy = g.acls[y];
        int64_t k;
        bool x_first;
        if (hx && hy) {
            k = tri_index(min(cx, cy), max(cx, cy), nh);
            x_first = cx < cy;
        } else if (hx) {
            k = n_hh + (int64_t)cx * nH + cy;
            x_first = true;
        } else if (hy) {
            k = n_hh + (int64_t)cy * nH + cx;
            x_first = false;
        } else {
            k = n_hh + n_hH + tri_index(min(cx, cy), max(cx, cy), nH);
            x_first = cx < cy;
        }
        const double s = x_first ? -1.0 : 1.0;
#pragma unroll
        for (int q = 0; q < 3; ++q) acc[q] = fma(s, g.fpair[3 * k + q], acc[q]);
    }
#pragma unroll
    for (int q = 0; q < 3; ++q) g.grad[3 * x + q] = acc[q];
}

}  // namespace am1g

using namespace am1g;

// Nuclear gradient (closed_form_gradient, gradient.rs:173-254; its Python front end python.rs:
// 202-275): the SCF of am1g_scf, then the pair derivatives at the converged density, plus the
// external field's -F q_a (gradient.rs:243-270).
// The gradient at fixed alpha and beta densities pab = [Pa | Pb] (host, 2 nao^2 doubles) with the
// Mulliken charges q (for the field term): pair derivatives on the device, fixed-order sums per
// atom, then -F q_a. Takes the context lock.
static int gradient_at(am1g_ctx* ctx, const am1g_request* rq, const MolDesc& md0, const std::vector<double>& pab,
                       const double* q, double* gradient) {
    const int nat = rq->n_atoms, nao = md0.nao;
    const size_t nn = (size_t)nao * nao;
    std::lock_guard<std::mutex> lock(ctx->mu);
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    // atom tables: heavy atoms then hydrogens, input order within each (the pair layout of k_pairs)
    const DevElem* tab = ctx->host_tab + md0.method * kMaxZ;
    std::vector<int> zs(nat), aoff(nat), acls(nat), alist;
    std::vector<int> heavy, hyd;
    for (int a = 0, off = 0, nh = 0, nH = 0; a < nat; ++a) {
        zs[a] = rq->z[a];
        aoff[a] = off;
        off += tab[zs[a]].norb;
        if (zs[a] != 1) {
            acls[a] = nh++;
            heavy.push_back(a);
        } else {
            acls[a] = nH++;
            hyd.push_back(a);
        }
    }
    alist = heavy;
    alist.insert(alist.end(), hyd.begin(), hyd.end());
    const int64_t npairs = (int64_t)nat * (nat - 1) / 2;
    const int64_t ncoef = std::max<int64_t>(w_pool_size(md0.n_heavy, md0.n_h), 1);
    DevBuf& B = ctx->grad_buf;
    const size_t off_z = 0, off_xyz = off_z + 256 * ((sizeof(int) * 4 * nat + 255) / 256);
    const size_t off_p = off_xyz + 256 * ((sizeof(double) * 3 * nat + 255) / 256);
    const size_t off_c = off_p + 256 * ((sizeof(double) * 2 * nn + 255) / 256);
    const size_t off_f = off_c + 256 * ((sizeof(double) * ncoef + 255) / 256);
    const size_t off_g = off_f + 256 * ((sizeof(double) * 3 * std::max<int64_t>(npairs, 1) + 255) / 256);
    const size_t total = off_g + sizeof(double) * 3 * nat;
    HIP_TRY(ctx, B.reserve(total));
    char* base = static_cast<char*>(B.p);
    std::vector<int> itab(4 * (size_t)nat);
    std::memcpy(itab.data(), zs.data(), sizeof(int) * nat);
    std::memcpy(itab.data() + nat, alist.data(), sizeof(int) * nat);
    std::memcpy(itab.data() + 2 * nat, aoff.data(), sizeof(int) * nat);
    std::memcpy(itab.data() + 3 * nat, acls.data(), sizeof(int) * nat);
    hipStream_t s = ctx->stream;
    HIP_TRY(ctx, hipMemcpyAsync(base + off_z, itab.data(), sizeof(int) * 4 * nat, hipMemcpyHostToDevice, s));
    HIP_TRY(ctx, hipMemcpyAsync(base + off_xyz, rq->xyz_bohr, sizeof(double) * 3 * nat, hipMemcpyHostToDevice, s));
    HIP_TRY(ctx, hipMemcpyAsync(base + off_p, pab.data(), sizeof(double) * 2 * nn, hipMemcpyHostToDevice, s));
    GradArgs g;
    std::memset(&g, 0, sizeof(g));
    g.md = md0;
    const int* dint = reinterpret_cast<const int*>(base + off_z);
    g.z = dint;
    g.alist = dint + nat;
    g.aoff = dint + 2 * nat;
    g.acls = dint + 3 * nat;
    g.xyz = reinterpret_cast<const double*>(base + off_xyz);
    g.elems = ctx->elems.as<DevElem>();
    g.pa = reinterpret_cast<const double*>(base + off_p);
    g.pb = g.pa + nn;
    g.coef = reinterpret_cast<double*>(base + off_c);
    g.fpair = reinterpret_cast<double*>(base + off_f);
    g.grad = reinterpret_cast<double*>(base + off_g);
    g.n_pairs = npairs;
    if (npairs > 0) {
        const unsigned blocks = (unsigned)std::min<int64_t>((npairs + 127) / 128, 65536);
        hipLaunchKernelGGL(k_grad_coef, dim3((unsigned)std::min<int64_t>((npairs + 255) / 256, 65536)), dim3(256), 0,
                           s, g);
        hipLaunchKernelGGL(k_grad_pairs, dim3(blocks), dim3(128), 0, s, g);
    }
    hipLaunchKernelGGL(k_grad_atoms, dim3((nat + 255) / 256), dim3(256), 0, s, g);
    HIP_TRY(ctx, hipGetLastError());
    HIP_TRY(ctx, hipMemcpyAsync(gradient, g.grad, sizeof(double) * 3 * nat, hipMemcpyDeviceToHost, s));
    HIP_TRY(ctx, hipStreamSynchronize(s));
    if (rq->electric_field) {  // -F q_a (dipole.rs:155-162)
        for (int a = 0; a < nat; ++a)
            for (int k = 0; k < 3; ++k) gradient[3 * a + k] -= rq->electric_field[k] * q[a];
    }
    return AM1G_OK;
}

extern "C" int am1g_gradient(am1g_ctx* ctx, const am1g_request* rq, am1g_result* res, double* gradient) {
    if (!ctx || !rq || !res || !gradient) return am1g_set_error(ctx, AM1G_INVALID_INPUT, "null argument");
    Prepared prep;
    std::string msg;
    int st = am1g_prepare(ctx, *rq, prep, msg, INT_MAX);
    if (st != AM1G_OK) {
        res->status = st;
        return am1g_set_error(ctx, st, "%s", msg.c_str());
    }
    const MolDesc md0 = prep.md;
    const int nat = rq->n_atoms, nao = md0.nao;
    const size_t nn = (size_t)nao * nao;
    // the SCF, with the total (and spin) density brought back if the caller did not ask for them
    std::vector<double> ptot, pspin, qbuf;
    am1g_result r = *res;
    if (!r.density) {
        ptot.resize(nn);
        r.density = ptot.data();
    }
    if (prep.uhf && !r.spin_density) {
        pspin.resize(nn);
        r.spin_density = pspin.data();
    }
    if (!r.charges) {
        qbuf.resize(nat);
        r.charges = qbuf.data();
    }
    st = am1g_scf(ctx, rq, &r);
    {
        double* keep[] = {res->charges, res->density, res->spin_density};
        *res = r;
        res->charges = keep[0];
        res->density = keep[1];
        res->spin_density = keep[2];
    }
    if (st != AM1G_OK) return st;
    const auto w0 = std::chrono::steady_clock::now();
    // alpha and beta densities: halves of the total for RHF, (P +- P_spin) / 2 for UHF
    std::vector<double> pab(2 * nn);
    for (size_t k = 0; k < nn; ++k) {
        const double s = prep.uhf ? r.spin_density[k] : 0.0;
        pab[k] = 0.5 * (r.density[k] + s);
        pab[nn + k] = 0.5 * (r.density[k] - s);
    }
    st = gradient_at(ctx, rq, md0, pab, r.charges, gradient);
    res->phase_ms[5] += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - w0).count();
    return st;
}

extern "C" int am1g_gradient_density(am1g_ctx* ctx, const am1g_request* rq, const double* density,
                                     const double* spin_density, double* gradient) {
    if (!ctx || !rq || !density || !gradient) return am1g_set_error(ctx, AM1G_INVALID_INPUT, "null argument");
    Prepared prep;
    std::string msg;
    const int st = am1g_prepare(ctx, *rq, prep, msg, INT_MAX);
    if (st != AM1G_OK) return am1g_set_error(ctx, st, "%s", msg.c_str());
    const MolDesc& md0 = prep.md;
    const int nat = rq->n_atoms, nao = md0.nao;
    const size_t nn = (size_t)nao * nao;
    std::vector<double> pab(2 * nn);
    for (size_t k = 0; k < nn; ++k) {
        const double s = spin_density ? spin_density[k] : 0.0;
        pab[k] = 0.5 * (density[k] + s);
        pab[nn + k] = 0.5 * (density[k] - s);
    }
    // Mulliken charges of the given density, for the field term
    const DevElem* tab = ctx->host_tab + md0.method * kMaxZ;
    std::vector<double> q(nat);
    for (int a = 0, off = 0; a < nat; ++a) {
        const DevElem& el = tab[rq->z[a]];
        double pop = 0.0;
        for (int k = 0; k < el.norb; ++k) pop += density[(size_t)(off + k) * nao + off + k];
        q[a] = el.core - pop;
        off += el.norb;
    }
    return gradient_at(ctx, rq, md0, pab, q.data(), gradient);
}
