// Grid-level matrix-core GEMM for the large-system path (internal header of the library).
//
// C = alpha * op(A) * op(B) + beta * C on column-major operands (the BLAS convention, so the
// call sites read like the dgemm calls they replace). FP64 runs on 16x16x4 FP64 MFMA tiles; FP32
// on 16x16x4 FP32 MFMA tiles with FP32 accumulation. Outputs too small to fill the device split
// the K range over several workgroups when a workspace is given; the partial products are then
// summed in a fixed order, so results do not depend on scheduling.
#pragma once
#include <hip/hip_runtime.h>

#include <stddef.h>

namespace am1g {

// Workspace (in elements of the operand type) that lets gemm split K for an m x n output.
size_t gemm_splitk_elems(int m, int n, int k);

// Column-major C (m x n) = alpha op(A) op(B) + beta C; op(A) is m x k, op(B) is k x n. beta == 0
// does not read C. `work` may be null (no split-K).
hipError_t gemm_f64(hipStream_t st, bool trans_a, bool trans_b, int m, int n, int k, double alpha, const double* A,
                    int lda, const double* B, int ldb, double beta, double* C, int ldc, double* work = nullptr,
                    size_t work_elems = 0);
hipError_t gemm_f32(hipStream_t st, bool trans_a, bool trans_b, int m, int n, int k, float alpha, const float* A,
                    int lda, const float* B, int ldb, float beta, float* C, int ldc, float* work = nullptr,
                    size_t work_elems = 0);

// The same for a product known to be symmetric (n x n; e.g. X X with X symmetric, or A A^T):
// only the 64 x 64 tiles on and above the diagonal are computed (about half the work), and each
// off-diagonal tile is stored at both places. Off-diagonal pairs are then exactly symmetric;
// inside a diagonal tile the two triangles are computed separately.
hipError_t gemm_sym_f64(hipStream_t st, bool trans_a, bool trans_b, int n, int k, double alpha, const double* A,
                        int lda, const double* B, int ldb, double beta, double* C, int ldc);
hipError_t gemm_sym_f32(hipStream_t st, bool trans_a, bool trans_b, int n, int k, float alpha, const float* A,
                        int lda, const float* B, int ldb, float beta, float* C, int ldc);

}  // namespace am1g
