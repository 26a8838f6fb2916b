// One large system on the whole device: the dense symmetric eigensolver entry point and the
// plain-path SCF for molecules beyond one workgroup's reach (more than 256 atoms or basis
// functions).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstring>
#include <vector>

#include "am1g_ctx.hpp"
#include "am1g_gemm.hpp"
#include "am1g_wg.hpp"

// Device buffers of the large-system SCF, kept on the context so repeated calls reuse them.
struct LargeBuffers {
    DevBuf A, evals, Ct;                       // eigensolver entry point
    DevBuf H, F, P, T0, T1, Pn, f32a, f32b;    // dense n x n (FP32 staging for mixed precision)
    DevBuf hist_f, hist_e, hist_d, ring, coef, partial, scal, pdiag, tcol, rsum, props, far;
    DevBuf Fb, Pb, Pt, Tb, Ctb, evals_b, hist_fb, hist_eb;  // beta channel (UHF)
    DevBuf mm_a, mm_b, mm_c, mm_w;                         // am1g_matmul operands and split-K partials
};

int am1g_large_setup(am1g_ctx* ctx) {
    if (!ctx->eig) ctx->eig = new EigWork();
    ctx->eig->blocked_min = ctx->tun.tri_blocked_min;
    ctx->eig->use_graphs = ctx->tun.eig_graphs != 0;
    if (!ctx->large) ctx->large = new LargeBuffers();
    return AM1G_OK;
}

void am1g_large_release(am1g_ctx* ctx) {
    delete ctx->large;
    ctx->large = nullptr;
    delete ctx->eig;
    ctx->eig = nullptr;
}

namespace {

using Clock = std::chrono::steady_clock;

double ms_since(Clock::time_point& t) {
    const auto now = Clock::now();
    const double d = std::chrono::duration<double, std::milli>(now - t).count();
    t = now;
    return d;
}

// Density of the nocc lowest orbitals of the symmetric n x n matrix F (the extrapolated Fock
// matrix of one channel) by second-order spectral projection (SP2; A. M. N. Niklasson, Phys. Rev.
// B 66, 155115 (2002)), without eigenvectors: X_0 = (e_max I - F) / (e_max - e_min) from the
// Gershgorin bounds of F, then X <- X^2 when that brings tr X closer to nocc and X <- 2X - X^2
// otherwise. Each step is one n x n x n product on the matrix cores. The idempotency error
// tr X - tr X^2 falls quadratically once the eigenvalues of X have left the middle of [0, 1]; the
// run stops when it no longer falls below its value two steps back (it has reached rounding) or
// at zero. The result w X (w = 2 for a closed shell, 1 for one spin) is left in *out (X or X2).
// T = float runs the products in FP32 (the FP32 phase of mixed precision). ok is false when the
// trace did not reach nocc (a degenerate Fermi level, or no convergence in kMaxSteps): the
// caller diagonalizes instead. rows: 3n doubles of device scratch; dscal: one device double.
template <typename T>
hipError_t sp2_density(hipStream_t s, int n, const double* F, int nocc, double w, T* X, T* X2, double* rows,
                       double* dscal, T** out, bool& ok, int& steps) {
    constexpr int kMaxSteps = 100;
    ok = false;
    steps = 0;
    *out = X;
    hipError_t e;
    std::vector<double> hr(3 * (size_t)n);
    if ((e = launch_gersh(n, F, rows, s)) != hipSuccess) return e;
    if ((e = hipMemcpyAsync(hr.data(), rows, sizeof(double) * 3 * n, hipMemcpyDeviceToHost, s)) != hipSuccess) return e;
    if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
    double emin = INFINITY, emax = -INFINITY, trf = 0.0;
    for (int i = 0; i < n; ++i) {
        emin = std::min(emin, hr[3 * (size_t)i]);
        emax = std::max(emax, hr[3 * (size_t)i + 1]);
        trf += hr[3 * (size_t)i + 2];
    }
    if (!(emax > emin) || !std::isfinite(emax - emin)) return hipSuccess;
    const double scale = 1.0 / (emax - emin);
    if ((e = launch_sp2_init<T>(n, F, emax, scale, X, s)) != hipSuccess) return e;
    double trx = scale * ((double)n * emax - trf);
    double e1 = INFINITY, e2 = INFINITY;  // idempotency errors one and two steps back
    bool stopped = false;
    for (int k = 0; k < kMaxSteps && !stopped; ++k) {
        // X^2 of the symmetric X: the upper tiles only, mirrored
        if constexpr (sizeof(T) == 4)
            e = gemm_sym_f32(s, false, false, n, n, 1.0f, X, n, X, n, 0.0f, X2, n);
        else
            e = gemm_sym_f64(s, false, false, n, n, 1.0, X, n, X, n, 0.0, X2, n);
        if (e != hipSuccess) return e;
        double trx2 = 0.0;
        if ((e = launch_trace<T>(n, X2, dscal, s)) != hipSuccess) return e;
        if ((e = hipMemcpyAsync(&trx2, dscal, sizeof(double), hipMemcpyDeviceToHost, s)) != hipSuccess) return e;
        if ((e = hipStreamSynchronize(s)) != hipSuccess) return e;
        if (!std::isfinite(trx2)) return hipSuccess;
        const double idem = trx - trx2;
        if (std::fabs(trx2 - nocc) < std::fabs(2.0 * trx - trx2 - nocc)) {
            trx = trx2;
        } else {
            if ((e = launch_sp2_combine<T>((int64_t)n * n, X, X2, s)) != hipSuccess) return e;
            trx = 2.0 * trx - trx2;
        }
        std::swap(X, X2);
        ++steps;
        // the error of an eigenvalue above 0.618 grows under squaring, so the stall test applies
        // only once the error is small
        stopped = idem <= 0.0 || (k >= 2 && idem >= e2 && idem < 1.0e-3 * n);
        e2 = e1;
        e1 = idem;
    }
    ok = stopped && std::fabs(trx - nocc) < (sizeof(T) == 4 ? 1.0e-1 : 1.0e-4);
    if (ok && (e = launch_scale<T>((int64_t)n * n, w, X, s)) != hipSuccess) return e;
    *out = X;
    return hipSuccess;
}

}  // namespace

// Unrestricted SCF of one large molecule (uhf_loop, scf.rs:719-858): F_s = H + J(Pa + Pb) - K(P_s),
// E = 1/2 (Pt.H + Pa.Fa + Pb.Fb), CDIIS on the stacked error [Ea ; Eb] with one coefficient set
// for both spins whenever DIIS is on, the divergence check after the extrapolation, and the sum
// of the two channels' rms density changes. The system tables, pair pool, H_core and the
// core-core energy are already on the device.
static int large_uhf(am1g_ctx* ctx, const am1g_request& rq, const MolDesc& md, const SysArgs& sa, LargeBuffers& L,
                     double ecore, double* phase, Clock::time_point& t_ph, am1g_result* res) {
    hipStream_t s = ctx->stream;
    const int nat = md.nat, n = md.nao, na = md.n_alpha, nb = md.n_beta;
    const size_t nn = (size_t)n * n;
    const DevElem* tab = ctx->host_tab + md.method * kMaxZ;
    const int64_t len_s = (int64_t)n * (n + 1) / 2, len_a = (int64_t)n * (n - 1) / 2;
    const bool diis = md.use_diis != 0;
    HIP_TRY(ctx, L.F.reserve(sizeof(double) * nn));
    HIP_TRY(ctx, L.Fb.reserve(sizeof(double) * nn));
    HIP_TRY(ctx, L.P.reserve(sizeof(double) * nn));
    HIP_TRY(ctx, L.Pb.reserve(sizeof(double) * nn));
    HIP_TRY(ctx, L.Pt.reserve(sizeof(double) * nn));
    HIP_TRY(ctx, L.T0.reserve(sizeof(double) * nn));
    HIP_TRY(ctx, L.T1.reserve(sizeof(double) * nn));
    HIP_TRY(ctx, L.Tb.reserve(sizeof(double) * nn));
    HIP_TRY(ctx, L.Pn.reserve(sizeof(double) * nn));
    HIP_TRY(ctx, L.Ct.reserve(sizeof(double) * nn));
    HIP_TRY(ctx, L.Ctb.reserve(sizeof(double) * nn));
    HIP_TRY(ctx, L.evals.reserve(sizeof(double) * n));
    HIP_TRY(ctx, L.evals_b.reserve(sizeof(double) * n));
    HIP_TRY(ctx, L.hist_f.reserve(sizeof(double) * kDiisDepth * len_s));
    HIP_TRY(ctx, L.hist_fb.reserve(sizeof(double) * kDiisDepth * len_s));
    HIP_TRY(ctx, L.hist_e.reserve(sizeof(double) * kDiisDepth * std::max<int64_t>(len_a, 1)));
    HIP_TRY(ctx, L.hist_eb.reserve(sizeof(double) * kDiisDepth * std::max<int64_t>(len_a, 1)));
    if (md.precision == 1) {
        HIP_TRY(ctx, L.f32a.reserve(sizeof(float) * nn));
        HIP_TRY(ctx, L.f32b.reserve(sizeof(float) * nn));
    }
    double* F[2] = {L.F.as<double>(), L.Fb.as<double>()};
    double* P[2] = {L.P.as<double>(), L.Pb.as<double>()};
    double* Ct[2] = {L.Ct.as<double>(), L.Ctb.as<double>()};
    double* eps[2] = {L.evals.as<double>(), L.evals_b.as<double>()};
    double* hf[2] = {L.hist_f.as<double>(), L.hist_fb.as<double>()};
    double* he[2] = {L.hist_e.as<double>(), L.hist_eb.as<double>()};
    double* comb[2] = {L.T1.as<double>(), L.Tb.as<double>()};
    double* Pt = L.Pt.as<double>();
    double* T0 = L.T0.as<double>();
    double* Pn = L.Pn.as<double>();
    double* partial = L.partial.as<double>();
    double* scal = L.scal.as<double>();
    int* dring = L.ring.as<int>();
    double* dcoef = L.coef.as<double>();
    const int nocc[2] = {na, nb};
    // SAD guess split by the alpha / beta electron fractions (scf.rs:730-740)
    const double ntot = (double)std::max(na + nb, 1);
    for (int c = 0; c < 2; ++c) {
        HIP_TRY(ctx, hipMemsetAsync(P[c], 0, sizeof(double) * nn, s));
        HIP_TRY(ctx, launch_sad(sa, (c == 0 ? na : nb) / ntot, P[c], s));
    }
    // scal layout: 0 sum Pt.H, 1/2 sum Ps.Fs, 3/4 sum E_s^2, 5..12 / 13..20 DIIS dots, 21/22 sum dP^2
    double h_scal[64];
    double B[kDiisDepth * kDiisDepth] = {0};
    int hcount = 0, head = 0, ring[kDiisDepth];
    double e_old = 0.0, residual = INFINITY;
    int iterations = 0, converged = 0, status = AM1G_OK, fp32_iters = 0;
    bool fp32_phase = md.precision == 1;
    double fp32_best = INFINITY;
    int fp32_stall = 0;
    // iterations run with FP64 GEMMs; convergence needs two of them, so that the energy change and
    // the density change compare two FP64 densities (after an FP32 phase as well)
    int fp64_iters = 0;
    int loose_iters = 0;  // FP64 iterations that met the p_tol test
    const double one = 1.0, zero = 0.0;
    const float onef = 1.0f, zerof = 0.0f;
    std::string eig_err;
    const double* last_f[2] = {nullptr, nullptr};  // Fock matrices whose SP2 densities the loop ended with
    int sp2_steps = 0;
    auto fock_energy = [&]() -> int {
        HIP_TRY(ctx, launch_add((int64_t)nn, P[0], P[1], Pt, s));
        HIP_TRY(ctx, launch_fock(sa, Pt, P[0], 1.0, F[0], s, partial, scal + 1));
        HIP_TRY(ctx, launch_fock(sa, Pt, P[1], 1.0, F[1], s, partial, scal + 2));
        return AM1G_OK;
    };
    int st;
    for (int iter = 0; iter < md.max_scf; ++iter) {
        iterations = iter + 1;
        if ((st = fock_energy()) != AM1G_OK) return res->status = st;
        phase[1] += ms_since(t_ph);
        const bool low = fp32_phase;
        if (!low) ++fp64_iters;
        const int slot = diis ? (hcount < kDiisDepth ? (head + hcount) % kDiisDepth : head) : 0;
        for (int c = 0; c < 2; ++c) {
            HIP_TRY(ctx, gemm_f64(s, false, false, n, n, n, one, P[c], n, F[c], n, zero, T0, n));
            HIP_TRY(ctx, launch_push(n, T0, false, F[c], P[c], hf[c] + slot * len_s, he[c] + slot * len_a, nullptr,
                                     partial, scal + 3 + c, s));
        }
        int m = 0;
        if (diis) {
            if (hcount < kDiisDepth) {
                hcount++;
            } else {
                head = (head + 1) % kDiisDepth;
                for (int i = 0; i < kDiisDepth - 1; ++i)
                    for (int j = 0; j < kDiisDepth - 1; ++j) B[i * kDiisDepth + j] = B[(i + 1) * kDiisDepth + j + 1];
            }
            for (int i = 0; i < hcount; ++i) ring[i] = (head + i) % kDiisDepth;
            m = hcount;
            HIP_TRY(ctx, hipMemcpyAsync(dring, ring, sizeof(int) * m, hipMemcpyHostToDevice, s));
            for (int c = 0; c < 2; ++c)
                HIP_TRY(ctx, launch_diis_dots(n, hf[c], he[c], nullptr, dring, m, m - 1, partial, scal + 5 + 8 * c, s));
        }
        HIP_TRY(ctx, hipMemcpyAsync(h_scal, scal, sizeof(double) * 24, hipMemcpyDeviceToHost, s));
        HIP_TRY(ctx, hipStreamSynchronize(s));
        const double e_elec = 0.5 * (h_scal[1] + h_scal[2]);
        const double err = std::sqrt(2.0 * (h_scal[3] + h_scal[4]));
        const double* Fuse[2] = {F[0], F[1]};
        if (diis) {
            const int newest = m - 1;
            for (int i = 0; i < m; ++i) {
                const double bi = 2.0 * (h_scal[5 + i] + h_scal[13 + i]);
                B[newest * kDiisDepth + i] = bi;
                B[i * kDiisDepth + newest] = bi;
            }
            if (m >= 2) {
                double Bm[(kDiisDepth + 1) * (kDiisDepth + 1)], rhs[kDiisDepth + 1], c[kDiisDepth + 1];
                const int dim = m + 1;
                for (int i = 0; i < m; ++i) {
                    for (int j = 0; j < m; ++j) Bm[i * dim + j] = B[i * kDiisDepth + j];
                    Bm[i * dim + m] = -1.0;
                    Bm[m * dim + i] = -1.0;
                    rhs[i] = 0.0;
                }
                Bm[m * dim + m] = 0.0;
                rhs[m] = -1.0;
                if (solve_bordered(Bm, rhs, dim, c)) {
                    HIP_TRY(ctx, hipMemcpyAsync(dcoef, c, sizeof(double) * m, hipMemcpyHostToDevice, s));
                    for (int q = 0; q < 2; ++q) {
                        HIP_TRY(ctx, launch_combine(n, hf[q], dring, dcoef, m, comb[q], partial, scal + 23, s));
                        Fuse[q] = comb[q];
                    }
                    HIP_TRY(ctx, hipStreamSynchronize(s));  // c is a stack array
                }
            }
        }
        if (!std::isfinite(e_elec) || !std::isfinite(err)) {
            status = AM1G_SCF_DIVERGED;
            break;
        }
        phase[2] += ms_since(t_ph);
        // each channel's new density by SP2 purification of Fuse[c] (kept for the orbital energies
        // at the end), or from its eigenvectors when that declines or is switched off
        bool by_sp2[2] = {false, false};
        for (int c = 0; c < 2; ++c) {
            if (!ctx->tun.sp2 || nocc[c] <= 0 || nocc[c] >= n) continue;
            int steps = 0;
            if (low) {
                float* X = nullptr;
                HIP_TRY(ctx, sp2_density<float>(s, n, Fuse[c], nocc[c], 1.0, L.f32a.as<float>(), L.f32b.as<float>(),
                                                partial, scal + 48, &X, by_sp2[c], steps));
                if (by_sp2[c]) HIP_TRY(ctx, launch_density_update(n, X, true, P[c], partial, scal + 21 + c, s));
            } else {
                double* X = nullptr;
                HIP_TRY(ctx, sp2_density<double>(s, n, Fuse[c], nocc[c], 1.0, T0, Pn, partial, scal + 48, &X,
                                                 by_sp2[c], steps));
                if (by_sp2[c]) HIP_TRY(ctx, launch_density_update(n, X, false, P[c], partial, scal + 21 + c, s));
            }
            sp2_steps += steps;
        }
        for (int c = 0; c < 2; ++c) {
            last_f[c] = by_sp2[c] ? Fuse[c] : nullptr;
            if (by_sp2[c]) continue;
            hipError_t e = eig_large(*ctx->eig, n, const_cast<double*>(Fuse[c]), eps[c], Ct[c], s, eig_err);
            if (e != hipSuccess) {
                if (eig_err.empty())
                    return res->status = am1g_set_error(ctx, AM1G_DEVICE_ERROR, "HIP error %s in the eigensolver",
                                                        hipGetErrorString(e));
                status = AM1G_LINEAR_ALGEBRA;
            }
        }
        phase[3] += ms_since(t_ph);
        if (status != AM1G_OK) break;
        if (low) ++fp32_iters;
        double dsum = 0.0;
        for (int c = 0; c < 2; ++c) {
            if (by_sp2[c]) {
                HIP_TRY(ctx, hipMemcpyAsync(h_scal + 21 + c, scal + 21 + c, sizeof(double), hipMemcpyDeviceToHost, s));
                HIP_TRY(ctx, hipStreamSynchronize(s));
                dsum += std::sqrt(h_scal[21 + c] / (double)std::max<size_t>(1, nn));
                continue;
            }
            if (nocc[c] > 0 && low) {
                float* Cf = L.f32a.as<float>();
                float* Pf = L.f32b.as<float>();
                HIP_TRY(ctx, launch_to_f32((int64_t)n * nocc[c], Ct[c], Cf, s));
                HIP_TRY(ctx, gemm_f32(s, false, true, n, n, nocc[c], onef, Cf, n, Cf, n, zerof, Pf, n));
                HIP_TRY(ctx, launch_density_update(n, Pf, true, P[c], partial, scal + 21 + c, s));
            } else {
                if (nocc[c] > 0) {
                    HIP_TRY(ctx, gemm_f64(s, false, true, n, n, nocc[c], one, Ct[c], n, Ct[c], n, zero, Pn, n));
                } else {
                    HIP_TRY(ctx, hipMemsetAsync(Pn, 0, sizeof(double) * nn, s));
                }
                HIP_TRY(ctx, launch_density_update(n, Pn, false, P[c], partial, scal + 21 + c, s));
            }
            HIP_TRY(ctx, hipMemcpyAsync(h_scal + 21 + c, scal + 21 + c, sizeof(double), hipMemcpyDeviceToHost, s));
            HIP_TRY(ctx, hipStreamSynchronize(s));
            dsum += std::sqrt(h_scal[21 + c] / (double)std::max<size_t>(1, nn));
        }
        phase[4] += ms_since(t_ph);
        const double de = std::fabs(e_elec - e_old);
        e_old = e_elec;
        residual = err;
        if (fp32_phase) {
            if (err < fp32_best) {
                fp32_best = err;
                fp32_stall = 0;
            } else {
                ++fp32_stall;
            }
            if (err <= 1.0e-4 || fp32_stall >= 2) {
                // the FP64 phase starts a fresh history: entries built from FP32-rounded densities
                // let the extrapolation settle on a density whose own commutator is still ~1e-5
                fp32_phase = false;
                hcount = 0;
                head = 0;
            }
        }
        // scf.rs:693 (RHF), :820 (UHF). A mixed-precision run takes another trajectory than the FP64 one, and the
        // density-change test at p_tol stops up to ~1e-6 e short of the fixed point, so it ends by
        // a test 100x tighter; should that one stall (a near-degenerate open shell can hold the
        // error on a plateau), ten iterations after the p_tol test is met end the run as well.
        const bool conv_loose = de < md.e_tol && (dsum < md.p_tol || err < 1.0e-7);
        // a density test averaged over n^2 entries weakens with n: at hundreds of basis functions the
        // mixed run also needs the commutator below 1e-6 before it stops (orbital energies within
        // ~1e-7 eV of the fixed point)
        const bool conv_tight = de < md.e_tol && ((dsum < 1.0e-2 * md.p_tol && err < 1.0e-6) || err < 1.0e-7);
        if (conv_loose && fp64_iters >= 2) ++loose_iters;
        if (fp64_iters >= 2 && (md.precision == 1 ? (conv_tight || loose_iters > 10) : conv_loose)) {
            converged = 1;
            break;
        }
    }
    // orbital energies and orbitals of each channel's last extrapolated Fock matrix, when its
    // density came from SP2 (the eigensolve of the loop's last iteration otherwise)
    const bool frontier_only = !res->mo_energies && !res->mo_energies_beta && !res->mo_coeff && !res->mo_coeff_beta;
    double frontier[2][2] = {{NAN, NAN}, {NAN, NAN}};  // per channel: eigenvalues nocc - 1 and nocc
    for (int c = 0; c < 2 && status == AM1G_OK; ++c) {
        if (!last_f[c]) continue;
        hipError_t e;
        if (frontier_only) {
            const int k_lo = std::max(nocc[c] - 1, 0), k_hi = std::min(nocc[c], n - 1);
            double v[2] = {NAN, NAN};
            e = eig_large_values(*ctx->eig, n, const_cast<double*>(last_f[c]), k_lo, k_hi, v, s, eig_err);
            for (int k = k_lo; k <= k_hi; ++k) frontier[c][k - (nocc[c] - 1)] = v[k - k_lo];
        } else {
            e = eig_large(*ctx->eig, n, const_cast<double*>(last_f[c]), eps[c], Ct[c], s, eig_err);
        }
        if (e != hipSuccess) {
            if (eig_err.empty())
                return res->status = am1g_set_error(ctx, AM1G_DEVICE_ERROR, "HIP error %s in the eigensolver",
                                                    hipGetErrorString(e));
            status = AM1G_LINEAR_ALGEBRA;
        }
    }
    phase[3] += ms_since(t_ph);
    // ---- final energy, populations, dipole, both channels' orbitals (scf.rs:327-468)
    if ((st = fock_energy()) != AM1G_OK) return res->status = st;
    HIP_TRY(ctx, launch_atom_props(sa, Pt, L.props.as<double>(), s));
    std::vector<double> props(4 * (size_t)nat), heps(n), hepsb(n);
    HIP_TRY(ctx, hipMemcpyAsync(h_scal, scal, sizeof(double) * 24, hipMemcpyDeviceToHost, s));
    HIP_TRY(ctx, hipMemcpyAsync(props.data(), L.props.p, sizeof(double) * 4 * nat, hipMemcpyDeviceToHost, s));
    HIP_TRY(ctx, hipMemcpyAsync(heps.data(), eps[0], sizeof(double) * n, hipMemcpyDeviceToHost, s));
    HIP_TRY(ctx, hipMemcpyAsync(hepsb.data(), eps[1], sizeof(double) * n, hipMemcpyDeviceToHost, s));
    double zv = 0.0;  // sum Z_a V_a of the far field (long_range_energy_term, fock.rs:200-232)
    if (sa.far) HIP_TRY(ctx, hipMemcpyAsync(&zv, sa.far + 2 * (size_t)nat, sizeof(double), hipMemcpyDeviceToHost, s));
    HIP_TRY(ctx, hipStreamSynchronize(s));
    const double electronic = 0.5 * (h_scal[1] + h_scal[2]) + 0.5 * zv;
    MolResult r;
    std::memset(&r, 0, sizeof(r));
    double dip[3] = {0, 0, 0}, eiso = 0.0, eheat = 0.0;
    for (int a = 0; a < nat; ++a) {
        const DevElem& el = tab[rq.z[a]];
        const double q = el.core - props[4 * a];
        if (res->charges) res->charges[a] = q;
        for (int k = 0; k < 3; ++k) dip[k] += rq.xyz_bohr[3 * a + k] * q;
        if (el.norb == 4)
            for (int k = 0; k < 3; ++k) dip[k] += -2.0 * el.dd * props[4 * a + 1 + k];
        eiso += el.e_isol;
        eheat += el.eheat_ev;
    }
    r.status = (status == AM1G_OK && !converged) ? AM1G_SCF_NOT_CONVERGED : status;
    r.iterations = iterations;
    r.converged = converged;
    r.nao = n;
    r.electronic_ev = electronic;
    r.core_ev = ecore;
    r.total_ev = electronic + ecore + md.field_nuc;
    r.heat_of_formation_kcal = (electronic + ecore + md.field_nuc - eiso + eheat) * kEvToKcal;
    const bool fa = frontier_only && last_f[0], fb = frontier_only && last_f[1];
    r.homo_ev = na >= 1 ? (fa ? frontier[0][0] : heps[na - 1]) : NAN;
    r.lumo_ev = na < n ? (fa ? frontier[0][1] : heps[na]) : NAN;
    r.homo_beta_ev = nb >= 1 ? (fb ? frontier[1][0] : hepsb[nb - 1]) : NAN;
    r.lumo_beta_ev = nb < n ? (fb ? frontier[1][1] : hepsb[nb]) : NAN;
    for (int k = 0; k < 3; ++k) r.dipole[k] = dip[k] * kAuDipoleToDebye;
    r.residual = residual;
    r.fp32_iterations = fp32_iters;
    am1g_fill_result(r, md, *res);
    if (res->mo_energies) std::memcpy(res->mo_energies, heps.data(), sizeof(double) * n);
    if (res->mo_energies_beta) std::memcpy(res->mo_energies_beta, hepsb.data(), sizeof(double) * n);
    if (res->density) HIP_TRY(ctx, hipMemcpy(res->density, Pt, sizeof(double) * nn, hipMemcpyDeviceToHost));
    std::vector<double> tmp;
    auto columns = [&](const double* ct, double* out) -> int {
        tmp.resize(nn);
        HIP_TRY(ctx, hipMemcpy(tmp.data(), ct, sizeof(double) * nn, hipMemcpyDeviceToHost));
        for (int mu = 0; mu < n; ++mu)
            for (int k = 0; k < n; ++k) out[(size_t)mu * n + k] = tmp[(size_t)k * n + mu];
        return AM1G_OK;
    };
    if (res->mo_coeff && (st = columns(Ct[0], res->mo_coeff)) != AM1G_OK) return st;
    if (res->mo_coeff_beta && (st = columns(Ct[1], res->mo_coeff_beta)) != AM1G_OK) return st;
    if (res->spin_density) {
        std::vector<double> pb(nn);
        HIP_TRY(ctx, hipMemcpy(res->spin_density, P[0], sizeof(double) * nn, hipMemcpyDeviceToHost));
        HIP_TRY(ctx, hipMemcpy(pb.data(), P[1], sizeof(double) * nn, hipMemcpyDeviceToHost));
        for (size_t k = 0; k < nn; ++k) res->spin_density[k] -= pb[k];
    }
    phase[5] += ms_since(t_ph);
    for (int k = 0; k < 6; ++k) res->phase_ms[k] = phase[k];
    res->eigen_sweeps = sp2_steps;
    if (status == AM1G_LINEAR_ALGEBRA) am1g_set_error(ctx, status, "%s", eig_err.c_str());
    else if (status == AM1G_SCF_DIVERGED) am1g_set_error(ctx, status, "SCF diverged after %d iterations", iterations);
    else if (r.status == AM1G_SCF_NOT_CONVERGED)
        am1g_set_error(ctx, r.status, "SCF did not converge in %d iterations (residual %.3e)", iterations, residual);
    return r.status;
}

// Restricted SCF of one molecule on the whole device (rhf_loop, scf.rs:557-716): grid Fock
// kernels, library GEMMs for the commutator and the density (FP32 while the commutator norm is
// above 1e-4 in mixed precision), packed DIIS histories on the device with the small CDIIS and
// A-DIIS solves on the host, and the large-order eigensolver.
// The core Hamiltonian of one large molecule on the device (build_core_with_neighbors,
// hamiltonian.rs:211-366; core_core_energy_with_neighbors, repulsion.rs:38-56): atom tables, the
// pair pool (k_pairs), H_core in L.H and the core-core energy in L.scal[30] (queued on the
// context's stream, not yet synchronized). Fills `sa` for the kernels that follow and reserves
// L.partial and L.scal for the SCF.
static int large_core(am1g_ctx* ctx, const am1g_request& rq, const MolDesc& md, LargeBuffers& L, SysArgs& sa) {
    hipStream_t s = ctx->stream;
    const int nat = md.nat, n = md.nao;
    const size_t nn = (size_t)n * n;
    const DevElem* tab = ctx->host_tab + md.method * kMaxZ;
    std::vector<int> zs(nat), alist, aoff(nat), acls(nat);
    {
        int off = 0, nh = 0, nH = 0;
        std::vector<int> heavy, hyd;
        for (int a = 0; a < nat; ++a) {
            zs[a] = rq.z[a];
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
    }
    const int64_t npairs = (int64_t)nat * (nat - 1) / 2;
    std::vector<int64_t> pair_off = {0, npairs};
    HIP_TRY(ctx, ctx->mols.reserve(sizeof(MolDesc)));
    HIP_TRY(ctx, ctx->z.reserve(sizeof(int) * nat));
    HIP_TRY(ctx, ctx->xyz.reserve(sizeof(double) * 3 * nat));
    HIP_TRY(ctx, ctx->alist.reserve(sizeof(int) * nat));
    HIP_TRY(ctx, ctx->aoff.reserve(sizeof(int) * nat));
    HIP_TRY(ctx, ctx->acls.reserve(sizeof(int) * nat));
    HIP_TRY(ctx, ctx->later.reserve(sizeof(int) * nat));
    HIP_TRY(ctx, ctx->pair_off.reserve(sizeof(int64_t) * 2));
    HIP_TRY(ctx, ctx->ecore_pair.reserve(sizeof(double) * std::max<int64_t>(npairs, 1)));
    HIP_TRY(ctx, ctx->wpool.reserve(sizeof(double) * std::max<int64_t>(w_pool_size(md.n_heavy, md.n_h), 1)));
    HIP_TRY(ctx, L.H.reserve(sizeof(double) * nn));
    HIP_TRY(ctx, hipMemcpyAsync(ctx->mols.p, &md, sizeof(MolDesc), hipMemcpyHostToDevice, s));
    HIP_TRY(ctx, hipMemcpyAsync(ctx->z.p, zs.data(), sizeof(int) * nat, hipMemcpyHostToDevice, s));
    HIP_TRY(ctx, hipMemcpyAsync(ctx->xyz.p, rq.xyz_bohr, sizeof(double) * 3 * nat, hipMemcpyHostToDevice, s));
    HIP_TRY(ctx, hipMemcpyAsync(ctx->alist.p, alist.data(), sizeof(int) * nat, hipMemcpyHostToDevice, s));
    HIP_TRY(ctx, hipMemcpyAsync(ctx->aoff.p, aoff.data(), sizeof(int) * nat, hipMemcpyHostToDevice, s));
    HIP_TRY(ctx, hipMemcpyAsync(ctx->acls.p, acls.data(), sizeof(int) * nat, hipMemcpyHostToDevice, s));
    const std::vector<int> later = later_partners(alist, aoff, md.n_heavy);
    HIP_TRY(ctx, hipMemcpyAsync(ctx->later.p, later.data(), sizeof(int) * nat, hipMemcpyHostToDevice, s));
    HIP_TRY(ctx, hipMemcpyAsync(ctx->pair_off.p, pair_off.data(), sizeof(int64_t) * 2, hipMemcpyHostToDevice, s));
    BatchArgs ba;
    std::memset(&ba, 0, sizeof(ba));
    ba.mols = ctx->mols.as<MolDesc>();
    ba.n_mols = 1;
    ba.z = ctx->z.as<int>();
    ba.xyz = ctx->xyz.as<double>();
    ba.elems = ctx->elems.as<DevElem>();
    ba.mats = L.H.as<double>();  // kH == 0: H_core at the start
    ba.wpool = ctx->wpool.as<double>();
    ba.alist = ctx->alist.as<int>();
    ba.aoff = ctx->aoff.as<int>();
    ba.acls = ctx->acls.as<int>();
    ba.pair_off = ctx->pair_off.as<int64_t>();
    ba.n_pairs = npairs;
    ba.ecore_pair = ctx->ecore_pair.as<double>();
    HIP_TRY(ctx, L.pdiag.reserve(sizeof(double) * 10 * nat));
    sa.md = md;
    sa.z = ba.z;
    sa.xyz = ba.xyz;
    sa.alist = ba.alist;
    sa.aoff = ba.aoff;
    sa.acls = ba.acls;
    sa.later = ctx->later.as<int>();
    sa.elems = ba.elems;
    sa.wpool = ba.wpool;
    sa.H = L.H.as<double>();
    sa.pdiag = L.pdiag.as<double>();
    HIP_TRY(ctx, L.tcol.reserve(sizeof(double) * 10 * std::max<int64_t>(1, (int64_t)md.n_heavy * (md.n_heavy - 1) / 2)));
    HIP_TRY(ctx, L.rsum.reserve(sizeof(double) * 10 * std::max(1, md.n_heavy)));
    sa.tcol = L.tcol.as<double>();
    sa.rsum = L.rsum.as<double>();
    sa.far = nullptr;
    if (md.far_cut > 0.0) {
        HIP_TRY(ctx, L.far.reserve(sizeof(double) * (2 * (size_t)nat + 1)));
        sa.far = L.far.as<double>();
    }

    const int64_t n_part = std::max<int64_t>({large_partials(n), 4096, 2 * fock_energy_parts(md)});
    HIP_TRY(ctx, L.partial.reserve(sizeof(double) * n_part));
    HIP_TRY(ctx, L.scal.reserve(sizeof(double) * 64));
    HIP_TRY(ctx, launch_pairs_one(ba, md, s));
    HIP_TRY(ctx, launch_hcore_diag(sa, s));
    HIP_TRY(ctx, launch_sum(npairs, ba.ecore_pair, L.partial.as<double>(), L.scal.as<double>() + 30, s));
    return AM1G_OK;
}

int am1g_large_scf(am1g_ctx* ctx, const am1g_request& rq, const Prepared& prep, am1g_result* res) {
    const MolDesc& md0 = prep.md;
    const bool uhf = prep.uhf;
    int st = am1g_large_setup(ctx);
    if (st != AM1G_OK) return res->status = st;
    LargeBuffers& L = *ctx->large;
    hipStream_t s = ctx->stream;
    MolDesc md = md0;
    const int nat = md.nat, n = md.nao, nocc = md.n_alpha;
    const size_t nn = (size_t)n * n;
    const DevElem* tab = ctx->host_tab + md.method * kMaxZ;
    auto t_all = Clock::now(), t_ph = t_all;
    double phase[6] = {0, 0, 0, 0, 0, 0};

    // ---- atom tables, pair integrals, H_core, core-core energy
    SysArgs sa;
    st = large_core(ctx, rq, md, L, sa);
    if (st != AM1G_OK) return res->status = st;
    const bool need_d = md.accel == 2;
    const int64_t len_s = (int64_t)n * (n + 1) / 2, len_a = (int64_t)n * (n - 1) / 2;
    HIP_TRY(ctx, L.F.reserve(sizeof(double) * nn));
    HIP_TRY(ctx, L.P.reserve(sizeof(double) * nn));
    HIP_TRY(ctx, L.T0.reserve(sizeof(double) * nn));
    HIP_TRY(ctx, L.T1.reserve(sizeof(double) * nn));
    HIP_TRY(ctx, L.Pn.reserve(sizeof(double) * nn));
    HIP_TRY(ctx, L.Ct.reserve(sizeof(double) * nn));
    HIP_TRY(ctx, L.evals.reserve(sizeof(double) * n));
    HIP_TRY(ctx, L.hist_f.reserve(sizeof(double) * kDiisDepth * len_s));
    HIP_TRY(ctx, L.hist_e.reserve(sizeof(double) * kDiisDepth * std::max<int64_t>(len_a, 1)));
    if (need_d) HIP_TRY(ctx, L.hist_d.reserve(sizeof(double) * kDiisDepth * len_s));
    HIP_TRY(ctx, L.ring.reserve(sizeof(int) * kDiisDepth));
    HIP_TRY(ctx, L.coef.reserve(sizeof(double) * kDiisDepth));
    HIP_TRY(ctx, L.props.reserve(sizeof(double) * 4 * nat));
    if (md.precision == 1) {
        HIP_TRY(ctx, L.f32a.reserve(sizeof(float) * nn));
        HIP_TRY(ctx, L.f32b.reserve(sizeof(float) * nn));
    }
    double* F = L.F.as<double>();
    double* P = L.P.as<double>();
    double* T0 = L.T0.as<double>();
    double* T1 = L.T1.as<double>();
    double* Pn = L.Pn.as<double>();
    double* Ct = L.Ct.as<double>();
    double* eps = L.evals.as<double>();
    double* hf = L.hist_f.as<double>();
    double* he = L.hist_e.as<double>();
    double* hd = need_d ? L.hist_d.as<double>() : nullptr;
    double* partial = L.partial.as<double>();
    double* scal = L.scal.as<double>();
    int* dring = L.ring.as<int>();
    double* dcoef = L.coef.as<double>();
    // scal layout: 0 sum P(H+F), 1 sum E^2, 2..25 DIIS dots, 26 non-finite count, 27 sum dP^2, 30 E_core
    double h_scal[64];
    if (uhf) {
        HIP_TRY(ctx, hipMemcpyAsync(h_scal, scal + 30, sizeof(double), hipMemcpyDeviceToHost, s));
        HIP_TRY(ctx, hipStreamSynchronize(s));
        phase[0] += ms_since(t_ph);
        return large_uhf(ctx, rq, md, sa, L, h_scal[0], phase, t_ph, res);
    }
    HIP_TRY(ctx, hipMemsetAsync(P, 0, sizeof(double) * nn, s));
    HIP_TRY(ctx, launch_sad(sa, 1.0, P, s));
    HIP_TRY(ctx, hipMemcpyAsync(h_scal, scal, sizeof(double) * 64, hipMemcpyDeviceToHost, s));
    HIP_TRY(ctx, hipStreamSynchronize(s));
    const double ecore = h_scal[30];
    phase[0] += ms_since(t_ph);

    // ---- SCF loop
    double B[kDiisDepth * kDiisDepth] = {0}, G[kDiisDepth * kDiisDepth] = {0};
    int hcount = 0, head = 0, ring[kDiisDepth];
    double e_old = 0.0, residual = INFINITY;
    int iterations = 0, converged = 0, status = AM1G_OK, fp32_iters = 0;
    // Mixed precision: the density GEMM runs in FP32 (MFMA, FP32 accumulation) until the DIIS
    // error ||[F,P]|| drops below 1e-4, then FP64 for the rest of the SCF. The commutator itself
    // stays FP64 (noisy error vectors slow the DIIS down). An FP32 density carries a rounding
    // floor of about eps_32 |F| sqrt(n^2) into the error, above 1e-4 for hundreds of basis
    // functions, so the FP32 phase also ends once the error has not improved for two iterations.
    bool fp32_phase = md.precision == 1;
    double fp32_best = INFINITY;
    int fp32_stall = 0;
    // iterations run with FP64 GEMMs; convergence needs two of them, so that the energy change and
    // the density change compare two FP64 densities (after an FP32 phase as well)
    int fp64_iters = 0;
    int loose_iters = 0;  // FP64 iterations that met the p_tol test
    const double one = 1.0, zero = 0.0, two = 2.0;
    const float zerof = 0.0f, twof = 2.0f;
    std::string eig_err;
    double* last_f = nullptr;  // the Fock matrix whose SP2 density the loop ended with
    int sp2_steps = 0;
    for (int iter = 0; iter < md.max_scf; ++iter) {
        iterations = iter + 1;
        HIP_TRY(ctx, launch_fock(sa, P, P, 0.5, F, s, partial, scal + 0));
        phase[1] += ms_since(t_ph);
        const bool low = fp32_phase;
        if (!low) ++fp64_iters;
        // commutator FP - PF: column-major P*F read row-major is F*P
        HIP_TRY(ctx, gemm_f64(s, false, false, n, n, n, one, P, n, F, n, zero, T0, n));
        const int slot = hcount < kDiisDepth ? (head + hcount) % kDiisDepth : head;
        HIP_TRY(ctx, launch_push(n, T0, false, F, P, hf + slot * len_s, he + slot * len_a,
                                 hd ? hd + slot * len_s : nullptr, partial, scal + 1, s));
        if (hcount < kDiisDepth) {
            hcount++;
        } else {
            head = (head + 1) % kDiisDepth;
            for (int i = 0; i < kDiisDepth - 1; ++i)
                for (int j = 0; j < kDiisDepth - 1; ++j) {
                    B[i * kDiisDepth + j] = B[(i + 1) * kDiisDepth + j + 1];
                    G[i * kDiisDepth + j] = G[(i + 1) * kDiisDepth + j + 1];
                }
        }
        for (int i = 0; i < hcount; ++i) ring[i] = (head + i) % kDiisDepth;
        const int m = hcount, newest = m - 1;
        HIP_TRY(ctx, hipMemcpyAsync(dring, ring, sizeof(int) * m, hipMemcpyHostToDevice, s));
        HIP_TRY(ctx, launch_diis_dots(n, hf, he, hd, dring, m, newest, partial, scal + 2, s));
        HIP_TRY(ctx, hipMemcpyAsync(h_scal, scal, sizeof(double) * 30, hipMemcpyDeviceToHost, s));
        HIP_TRY(ctx, hipStreamSynchronize(s));
        const double e_elec = 0.5 * h_scal[0];
        const double err = std::sqrt(2.0 * h_scal[1]);
        for (int i = 0; i < m; ++i) {
            B[newest * kDiisDepth + i] = 2.0 * h_scal[2 + i];
            B[i * kDiisDepth + newest] = 2.0 * h_scal[2 + i];
        }
        if (need_d)
            for (int j = 0; j < m; ++j) {
                G[newest * kDiisDepth + j] = h_scal[2 + kDiisDepth + j];
                G[j * kDiisDepth + newest] = h_scal[2 + 2 * kDiisDepth + j];
            }
        if (!std::isfinite(e_elec) || !std::isfinite(err)) {
            status = AM1G_SCF_DIVERGED;
            break;
        }
        // extrapolation coefficients (scf.rs:627-653)
        bool use = false;
        double c[kDiisDepth + 1];
        if (md.accel == 1 || (md.accel == 2 && err <= md.adiis_switch)) {
            if (m >= 2) {
                double Bm[(kDiisDepth + 1) * (kDiisDepth + 1)], rhs[kDiisDepth + 1];
                const int dim = m + 1;
                for (int i = 0; i < m; ++i) {
                    for (int j = 0; j < m; ++j) Bm[i * dim + j] = B[i * kDiisDepth + j];
                    Bm[i * dim + m] = -1.0;
                    Bm[m * dim + i] = -1.0;
                    rhs[i] = 0.0;
                }
                Bm[m * dim + m] = 0.0;
                rhs[m] = -1.0;
                if (solve_bordered(Bm, rhs, dim, c)) {
                    double wsum = 0.0;
                    for (int i = 0; i < m; ++i) wsum += std::fabs(c[i]);
                    use = std::isfinite(wsum);
                }
            }
        } else if (md.accel == 2 && m >= 2) {
            adiis_solve(G, m, c);
            use = true;
        }
        double* Fuse = F;
        if (use) {
            HIP_TRY(ctx, hipMemcpyAsync(dcoef, c, sizeof(double) * m, hipMemcpyHostToDevice, s));
            HIP_TRY(ctx, launch_combine(n, hf, dring, dcoef, m, T1, partial, scal + 26, s));
            double bad = 0.0;
            HIP_TRY(ctx, hipMemcpyAsync(&bad, scal + 26, sizeof(double), hipMemcpyDeviceToHost, s));
            HIP_TRY(ctx, hipStreamSynchronize(s));
            if (bad == 0.0) Fuse = T1;
        }
        phase[2] += ms_since(t_ph);
        // New density: by SP2 purification of Fuse, which stays intact for the orbital energies at
        // the end, or from the eigenvectors when that declines or is switched off.
        bool by_sp2 = false;
        if (ctx->tun.sp2 && nocc > 0 && nocc < n) {
            int steps = 0;
            if (low) {
                float* X = nullptr;
                HIP_TRY(ctx, sp2_density<float>(s, n, Fuse, nocc, 2.0, L.f32a.as<float>(), L.f32b.as<float>(), partial,
                                                scal + 48, &X, by_sp2, steps));
                phase[3] += ms_since(t_ph);
                if (by_sp2) {
                    ++fp32_iters;
                    HIP_TRY(ctx, launch_density_update(n, X, true, P, partial, scal + 27, s));
                }
            } else {
                double* X = nullptr;
                HIP_TRY(ctx, sp2_density<double>(s, n, Fuse, nocc, 2.0, T0, Pn, partial, scal + 48, &X, by_sp2, steps));
                phase[3] += ms_since(t_ph);
                if (by_sp2) HIP_TRY(ctx, launch_density_update(n, X, false, P, partial, scal + 27, s));
            }
            sp2_steps += steps;
        }
        last_f = by_sp2 ? Fuse : nullptr;
        if (!by_sp2) {
            // eigensolve (destroys Fuse; F is rebuilt by the next Fock pass)
            hipError_t e = eig_large(*ctx->eig, n, Fuse, eps, Ct, s, eig_err);
            if (e != hipSuccess) {
                if (!eig_err.empty()) {
                    status = AM1G_LINEAR_ALGEBRA;
                    break;
                }
                return res->status = am1g_set_error(ctx, AM1G_DEVICE_ERROR, "HIP error %s in the eigensolver",
                                                    hipGetErrorString(e));
            }
            phase[3] += ms_since(t_ph);
            // density 2 C_occ C_occ^T: rows 0..nocc-1 of Ct are the occupied orbitals (column-major n x nocc)
            if (nocc > 0) {
                if (low) {
                    ++fp32_iters;
                    float* Cf = L.f32a.as<float>();
                    float* Pf = L.f32b.as<float>();
                    HIP_TRY(ctx, launch_to_f32((int64_t)n * nocc, Ct, Cf, s));
                    HIP_TRY(ctx, gemm_f32(s, false, true, n, n, nocc, twof, Cf, n, Cf, n, zerof, Pf, n));
                    HIP_TRY(ctx, launch_density_update(n, Pf, true, P, partial, scal + 27, s));
                } else {
                    HIP_TRY(ctx, gemm_f64(s, false, true, n, n, nocc, two, Ct, n, Ct, n, zero, Pn, n));
                    HIP_TRY(ctx, launch_density_update(n, Pn, false, P, partial, scal + 27, s));
                }
            } else {
                HIP_TRY(ctx, hipMemsetAsync(Pn, 0, sizeof(double) * nn, s));
                HIP_TRY(ctx, launch_density_update(n, Pn, false, P, partial, scal + 27, s));
            }
        }
        double dd = 0.0;
        HIP_TRY(ctx, hipMemcpyAsync(&dd, scal + 27, sizeof(double), hipMemcpyDeviceToHost, s));
        HIP_TRY(ctx, hipStreamSynchronize(s));
        const double dp = std::sqrt(dd / (double)std::max<size_t>(1, nn));
        phase[4] += ms_since(t_ph);
        const double de = std::fabs(e_elec - e_old);
        e_old = e_elec;
        residual = err;
        if (fp32_phase) {
            if (err < fp32_best) {
                fp32_best = err;
                fp32_stall = 0;
            } else {
                ++fp32_stall;
            }
            if (err <= 1.0e-4 || fp32_stall >= 2) {
                // the FP64 phase starts a fresh history: entries built from FP32-rounded densities
                // let the extrapolation settle on a density whose own commutator is still ~1e-5
                fp32_phase = false;
                hcount = 0;
                head = 0;
            }
        }
        // scf.rs:693 (RHF), :820 (UHF). A mixed-precision run takes another trajectory than the FP64 one, and the
        // density-change test at p_tol stops up to ~1e-6 e short of the fixed point, so it ends by
        // a test 100x tighter; should that one stall (a near-degenerate open shell can hold the
        // error on a plateau), ten iterations after the p_tol test is met end the run as well.
        const bool conv_loose = de < md.e_tol && (dp < md.p_tol || err < 1.0e-7);
        // see the UHF loop: the mixed run also needs the commutator below 1e-6
        const bool conv_tight = de < md.e_tol && ((dp < 1.0e-2 * md.p_tol && err < 1.0e-6) || err < 1.0e-7);
        if (conv_loose && fp64_iters >= 2) ++loose_iters;
        if (fp64_iters >= 2 && (md.precision == 1 ? (conv_tight || loose_iters > 10) : conv_loose)) {
            converged = 1;
            break;
        }
    }
    // orbital energies and orbitals of the last extrapolated Fock matrix, when its density came
    // from SP2 (the eigensolve of the loop's last iteration otherwise); when the caller takes
    // neither orbital energies nor orbitals, only HOMO and LUMO, by bisection on its tridiagonal form
    const bool frontier_only = !res->mo_energies && !res->mo_coeff;
    double frontier[2] = {NAN, NAN};  // eigenvalues nocc - 1 and nocc (frontier_only)
    if (last_f && status == AM1G_OK) {
        hipError_t e;
        if (frontier_only) {
            const int k_lo = std::max(nocc - 1, 0), k_hi = std::min(nocc, n - 1);
            double v[2] = {NAN, NAN};
            e = eig_large_values(*ctx->eig, n, last_f, k_lo, k_hi, v, s, eig_err);
            for (int k = k_lo; k <= k_hi; ++k) frontier[k - (nocc - 1)] = v[k - k_lo];
        } else {
            e = eig_large(*ctx->eig, n, last_f, eps, Ct, s, eig_err);
        }
        if (e != hipSuccess) {
            if (eig_err.empty())
                return res->status = am1g_set_error(ctx, AM1G_DEVICE_ERROR, "HIP error %s in the eigensolver",
                                                    hipGetErrorString(e));
            status = AM1G_LINEAR_ALGEBRA;
        }
        phase[3] += ms_since(t_ph);
    }
    const bool use_frontier = last_f && frontier_only;
    // ---- final energy, populations, dipole (scf.rs:327-468)
    HIP_TRY(ctx, launch_fock(sa, P, P, 0.5, F, s, partial, scal + 0));
    HIP_TRY(ctx, launch_atom_props(sa, P, L.props.as<double>(), s));
    std::vector<double> props(4 * (size_t)nat), heps(n);
    HIP_TRY(ctx, hipMemcpyAsync(h_scal, scal, sizeof(double) * 30, hipMemcpyDeviceToHost, s));
    HIP_TRY(ctx, hipMemcpyAsync(props.data(), L.props.p, sizeof(double) * 4 * nat, hipMemcpyDeviceToHost, s));
    HIP_TRY(ctx, hipMemcpyAsync(heps.data(), eps, sizeof(double) * n, hipMemcpyDeviceToHost, s));
    double zv = 0.0;  // sum Z_a V_a of the far field (long_range_energy_term, fock.rs:200-232)
    if (sa.far) HIP_TRY(ctx, hipMemcpyAsync(&zv, sa.far + 2 * (size_t)nat, sizeof(double), hipMemcpyDeviceToHost, s));
    HIP_TRY(ctx, hipStreamSynchronize(s));
    const double electronic = 0.5 * h_scal[0] + 0.5 * zv;
    MolResult r;
    std::memset(&r, 0, sizeof(r));
    double dip[3] = {0, 0, 0}, eiso = 0.0, eheat = 0.0;
    for (int a = 0; a < nat; ++a) {
        const DevElem& el = tab[rq.z[a]];
        const double q = el.core - props[4 * a];
        if (res->charges) res->charges[a] = q;
        for (int k = 0; k < 3; ++k) dip[k] += rq.xyz_bohr[3 * a + k] * q;
        if (el.norb == 4)
            for (int k = 0; k < 3; ++k) dip[k] += -2.0 * el.dd * props[4 * a + 1 + k];
        eiso += el.e_isol;
        eheat += el.eheat_ev;
    }
    r.status = (status == AM1G_OK && !converged) ? AM1G_SCF_NOT_CONVERGED : status;
    r.iterations = iterations;
    r.converged = converged;
    r.nao = n;
    r.electronic_ev = electronic;
    r.core_ev = ecore;
    r.total_ev = electronic + ecore + md.field_nuc;
    r.heat_of_formation_kcal = (electronic + ecore + md.field_nuc - eiso + eheat) * kEvToKcal;
    r.homo_ev = nocc >= 1 ? (use_frontier ? frontier[0] : heps[nocc - 1]) : NAN;
    r.lumo_ev = nocc < n ? (use_frontier ? frontier[1] : heps[nocc]) : NAN;
    r.homo_beta_ev = NAN;
    r.lumo_beta_ev = NAN;
    for (int k = 0; k < 3; ++k) r.dipole[k] = dip[k] * kAuDipoleToDebye;
    r.residual = residual;
    r.fp32_iterations = fp32_iters;
    am1g_fill_result(r, md, *res);
    if (res->mo_energies) std::memcpy(res->mo_energies, heps.data(), sizeof(double) * n);
    if (res->density) HIP_TRY(ctx, hipMemcpy(res->density, P, sizeof(double) * nn, hipMemcpyDeviceToHost));
    if (res->mo_coeff) {
        std::vector<double> ct(nn);
        HIP_TRY(ctx, hipMemcpy(ct.data(), Ct, sizeof(double) * nn, hipMemcpyDeviceToHost));
        for (int mu = 0; mu < n; ++mu)
            for (int k = 0; k < n; ++k) res->mo_coeff[(size_t)mu * n + k] = ct[(size_t)k * n + mu];
    }
    phase[5] += ms_since(t_ph);
    for (int k = 0; k < 6; ++k) res->phase_ms[k] = phase[k];
    res->eigen_sweeps = sp2_steps;
    if (status == AM1G_LINEAR_ALGEBRA) am1g_set_error(ctx, status, "%s", eig_err.c_str());
    else if (status == AM1G_SCF_DIVERGED) am1g_set_error(ctx, status, "SCF diverged after %d iterations", iterations);
    else if (r.status == AM1G_SCF_NOT_CONVERGED)
        am1g_set_error(ctx, r.status, "SCF did not converge in %d iterations (residual %.3e)", iterations, residual);
    return r.status;
}

extern "C" {

int am1g_core_hamiltonian(am1g_ctx* ctx, const am1g_request* rq, double* h_core, double* core_ev,
                          double* pair_blocks, long long pair_blocks_len) {
    if (!ctx || !rq) return am1g_set_error(ctx, AM1G_INVALID_INPUT, "null argument");
    Prepared prep;
    std::string msg;
    int st = am1g_prepare(ctx, *rq, prep, msg, INT_MAX);
    if (st != AM1G_OK) return am1g_set_error(ctx, st, "%s", msg.c_str());
    const MolDesc& md = prep.md;
    const int64_t w_len = w_pool_size(md.n_heavy, md.n_h);
    if (pair_blocks && pair_blocks_len < w_len)
        return am1g_set_error(ctx, AM1G_INVALID_INPUT, "pair_blocks holds %lld doubles; this molecule needs %lld",
                              pair_blocks_len, (long long)w_len);
    std::lock_guard<std::mutex> lock(ctx->mu);
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    st = am1g_large_setup(ctx);
    if (st != AM1G_OK) return st;
    LargeBuffers& L = *ctx->large;
    SysArgs sa;
    st = large_core(ctx, *rq, md, L, sa);
    if (st != AM1G_OK) return st;
    hipStream_t s = ctx->stream;
    const size_t nn = (size_t)md.nao * md.nao;
    if (h_core) HIP_TRY(ctx, hipMemcpyAsync(h_core, L.H.p, sizeof(double) * nn, hipMemcpyDeviceToHost, s));
    double ec = 0.0;
    HIP_TRY(ctx, hipMemcpyAsync(&ec, L.scal.as<double>() + 30, sizeof(double), hipMemcpyDeviceToHost, s));
    if (pair_blocks && w_len > 0)
        HIP_TRY(ctx, hipMemcpyAsync(pair_blocks, ctx->wpool.p, sizeof(double) * w_len, hipMemcpyDeviceToHost, s));
    HIP_TRY(ctx, hipStreamSynchronize(s));
    if (core_ev) *core_ev = ec;
    return AM1G_OK;
}

int am1g_symmetric_eigen(am1g_ctx* ctx, int n, const double* a, double* eigenvalues, double* eigenvectors) {
    if (!ctx || n < 0 || (n > 0 && (!a || !eigenvalues || !eigenvectors)))
        return am1g_set_error(ctx, AM1G_INVALID_INPUT, "null argument");
    if (n == 0) return AM1G_OK;
    std::lock_guard<std::mutex> lock(ctx->mu);
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    for (size_t k = 0; k < (size_t)n * n; ++k)
        if (!std::isfinite(a[k]))
            return am1g_set_error(ctx, AM1G_LINEAR_ALGEBRA, "symmetric_eigen: non-finite matrix element");
    int st = am1g_large_setup(ctx);
    if (st != AM1G_OK) return st;
    LargeBuffers& L = *ctx->large;
    const size_t nn = (size_t)n * n;
    HIP_TRY(ctx, L.A.reserve(sizeof(double) * nn));
    HIP_TRY(ctx, L.evals.reserve(sizeof(double) * n));
    HIP_TRY(ctx, L.Ct.reserve(sizeof(double) * nn));
    hipStream_t s = ctx->stream;
    HIP_TRY(ctx, hipMemcpyAsync(L.A.p, a, sizeof(double) * nn, hipMemcpyHostToDevice, s));
    HIP_TRY(ctx, hipEventRecord(ctx->ev0, s));
    std::string err;
    hipError_t e = eig_large(*ctx->eig, n, L.A.as<double>(), L.evals.as<double>(), L.Ct.as<double>(), s, err);
    if (e != hipSuccess) {
        if (!err.empty()) return am1g_set_error(ctx, AM1G_LINEAR_ALGEBRA, "%s", err.c_str());
        return am1g_set_error(ctx, AM1G_DEVICE_ERROR, "HIP error %s in the eigensolver", hipGetErrorString(e));
    }
    HIP_TRY(ctx, hipEventRecord(ctx->ev1, s));
    std::vector<double> ct(nn);
    HIP_TRY(ctx, hipMemcpyAsync(eigenvalues, L.evals.p, sizeof(double) * n, hipMemcpyDeviceToHost, s));
    HIP_TRY(ctx, hipMemcpyAsync(ct.data(), L.Ct.p, sizeof(double) * nn, hipMemcpyDeviceToHost, s));
    HIP_TRY(ctx, hipStreamSynchronize(s));
    float ms = 0.f;
    HIP_TRY(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    ctx->last_ms = ms;
    // rows of Ct are eigenvectors; the result holds them as columns
    for (int i = 0; i < n; ++i)
        for (int k = 0; k < n; ++k) eigenvectors[(size_t)i * n + k] = ct[(size_t)k * n + i];
    return AM1G_OK;
}

int am1g_matmul(am1g_ctx* ctx, int trans_a, int trans_b, int m, int n, int k, double alpha, const double* a,
                const double* b, double beta, double* c, int precision) {
    if (!ctx || m < 0 || n < 0 || k < 0 || (m > 0 && n > 0 && (!c || (k > 0 && (!a || !b)))))
        return am1g_set_error(ctx, AM1G_INVALID_INPUT, "null argument");
    if (precision != AM1G_PRECISION_FP64 && precision != AM1G_PRECISION_MIXED)
        return am1g_set_error(ctx, AM1G_INVALID_INPUT, "matmul: unknown precision %d", precision);
    if (m == 0 || n == 0) return AM1G_OK;
    std::lock_guard<std::mutex> lock(ctx->mu);
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    int st = am1g_large_setup(ctx);
    if (st != AM1G_OK) return st;
    LargeBuffers& L = *ctx->large;
    hipStream_t s = ctx->stream;
    const size_t na = (size_t)m * k, nb = (size_t)k * n, nc = (size_t)m * n;
    // Row-major c = op(a) op(b) is column-major c^T = op(b)^T op(a)^T: the column-major product
    // takes b first; a row-major operand read column-major is its transpose.
    const int lda = trans_b ? k : n, ldb = trans_a ? m : k;
    const size_t wn = gemm_splitk_elems(n, m, k);
    const bool f32 = precision == AM1G_PRECISION_MIXED;
    const size_t esz = f32 ? sizeof(float) : sizeof(double);
    HIP_TRY(ctx, L.mm_a.reserve(esz * std::max<size_t>(na, 1)));
    HIP_TRY(ctx, L.mm_b.reserve(esz * std::max<size_t>(nb, 1)));
    HIP_TRY(ctx, L.mm_c.reserve(esz * nc));
    HIP_TRY(ctx, L.mm_w.reserve(esz * std::max<size_t>(wn, 1)));
    if (f32) {
        std::vector<float> ha(na), hb(nb), hc(nc);
        for (size_t i = 0; i < na; ++i) ha[i] = (float)a[i];
        for (size_t i = 0; i < nb; ++i) hb[i] = (float)b[i];
        for (size_t i = 0; i < nc; ++i) hc[i] = beta != 0.0 ? (float)c[i] : 0.0f;
        if (na) HIP_TRY(ctx, hipMemcpyAsync(L.mm_a.p, ha.data(), esz * na, hipMemcpyHostToDevice, s));
        if (nb) HIP_TRY(ctx, hipMemcpyAsync(L.mm_b.p, hb.data(), esz * nb, hipMemcpyHostToDevice, s));
        HIP_TRY(ctx, hipMemcpyAsync(L.mm_c.p, hc.data(), esz * nc, hipMemcpyHostToDevice, s));
        HIP_TRY(ctx, hipEventRecord(ctx->ev0, s));
        HIP_TRY(ctx, gemm_f32(s, trans_b != 0, trans_a != 0, n, m, k, (float)alpha, L.mm_b.as<float>(), lda,
                              L.mm_a.as<float>(), ldb, (float)beta, L.mm_c.as<float>(), n, L.mm_w.as<float>(), wn));
        HIP_TRY(ctx, hipEventRecord(ctx->ev1, s));
        HIP_TRY(ctx, hipMemcpyAsync(hc.data(), L.mm_c.p, esz * nc, hipMemcpyDeviceToHost, s));
        HIP_TRY(ctx, hipStreamSynchronize(s));
        for (size_t i = 0; i < nc; ++i) c[i] = hc[i];
    } else {
        if (na) HIP_TRY(ctx, hipMemcpyAsync(L.mm_a.p, a, esz * na, hipMemcpyHostToDevice, s));
        if (nb) HIP_TRY(ctx, hipMemcpyAsync(L.mm_b.p, b, esz * nb, hipMemcpyHostToDevice, s));
        if (beta != 0.0) HIP_TRY(ctx, hipMemcpyAsync(L.mm_c.p, c, esz * nc, hipMemcpyHostToDevice, s));
        HIP_TRY(ctx, hipEventRecord(ctx->ev0, s));
        HIP_TRY(ctx, gemm_f64(s, trans_b != 0, trans_a != 0, n, m, k, alpha, L.mm_b.as<double>(), lda,
                              L.mm_a.as<double>(), ldb, beta, L.mm_c.as<double>(), n, L.mm_w.as<double>(), wn));
        HIP_TRY(ctx, hipEventRecord(ctx->ev1, s));
        HIP_TRY(ctx, hipMemcpyAsync(c, L.mm_c.p, esz * nc, hipMemcpyDeviceToHost, s));
        HIP_TRY(ctx, hipStreamSynchronize(s));
    }
    float ms = 0.f;
    HIP_TRY(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    ctx->last_ms = ms;
    return AM1G_OK;
}

}  // extern "C"
