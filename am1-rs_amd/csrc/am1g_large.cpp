// One large system on the whole device: the dense symmetric eigensolver entry point and the
// plain-path SCF for molecules beyond one workgroup's reach (more than 256 atoms or basis
// functions).
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstring>
#include <vector>

#include "am1g_ctx.hpp"

// Device buffers of the large-system SCF, kept on the context so repeated calls reuse them.
struct LargeBuffers {
    DevBuf A, evals, Ct;
};

int am1g_large_setup(am1g_ctx* ctx) {
    if (!ctx->blas) {
        if (rocblas_create_handle(&ctx->blas) != rocblas_status_success) {
            ctx->blas = nullptr;
            return am1g_set_error(ctx, AM1G_DEVICE_ERROR, "rocBLAS handle creation failed");
        }
        if (rocblas_set_stream(ctx->blas, ctx->stream) != rocblas_status_success)
            return am1g_set_error(ctx, AM1G_DEVICE_ERROR, "rocBLAS stream binding failed");
    }
    if (!ctx->eig) ctx->eig = new EigWork();
    if (!ctx->large) ctx->large = new LargeBuffers();
    return AM1G_OK;
}

void am1g_large_release(am1g_ctx* ctx) {
    delete ctx->large;
    ctx->large = nullptr;
    delete ctx->eig;
    ctx->eig = nullptr;
    if (ctx->blas) (void)rocblas_destroy_handle(ctx->blas);
    ctx->blas = nullptr;
}

extern "C" {

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
    hipError_t e = eig_large(*ctx->eig, ctx->blas, n, L.A.as<double>(), L.evals.as<double>(), L.Ct.as<double>(), s, err);
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

}  // extern "C"
