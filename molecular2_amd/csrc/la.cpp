// Dense FP64 linear algebra of the SCF on the device: symmetric eigensolver
// (rocSOLVER dsyevd), GEMM (rocBLAS dgemm), Cholesky (rocSOLVER dpotrf).
// These replace the Eigen calls of the reference's SCF driver:
// SelfAdjointEigenSolver in Fock::formOrthog / Fock::diagonalise
// (src/fock.cpp:124-142,228-233) and the matrix products of SCF::calcErr /
// Fock::transform (src/scf.cpp:52-69, src/fock.cpp:156-172).
#include "ctx.hpp"

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

namespace m2 {

static m2_status blas_handle(m2_ctx* ctx, rocblas_handle* h) {
    if (!ctx->blas) {
        rocblas_handle hh = nullptr;
        if (rocblas_create_handle(&hh) != rocblas_status_success) {
            ctx->err = "rocblas_create_handle failed";
            return M2_ERR_HIP;
        }
        rocblas_set_stream(hh, ctx->stream);
        rocblas_set_pointer_mode(hh, rocblas_pointer_mode_host);
        ctx->blas = (void*)hh;
    }
    *h = (rocblas_handle)ctx->blas;
    return M2_OK;
}

m2_status la_handle(m2_ctx* ctx, void** h) {
    rocblas_handle hh;
    m2_status st = blas_handle(ctx, &hh);
    *h = (void*)hh;
    return st;
}

void la_release(m2_ctx* ctx) {
    if (ctx->blas) rocblas_destroy_handle((rocblas_handle)ctx->blas);
    ctx->blas = nullptr;
    ctx->la_a.release();
    ctx->la_b.release();
    ctx->la_c.release();
    ctx->la_w.release();
    ctx->la_info.release();
}

// device-pointer eigensolver: A (n x n, column-major) overwritten by the
// eigenvectors, w ascending eigenvalues; E is n doubles of workspace.
m2_status eigh_device(m2_ctx* ctx, int n, double* dA, double* dw, double* dE, int* dinfo) {
    rocblas_handle h;
    m2_status st = blas_handle(ctx, &h);
    if (st) return st;
    if (rocsolver_dsyevd(h, rocblas_evect_original, rocblas_fill_lower, n, dA, n, dw, dE, dinfo) !=
        rocblas_status_success) {
        ctx->err = "rocsolver_dsyevd failed";
        return M2_ERR_HIP;
    }
    return M2_OK;
}

m2_status gemm_device(m2_ctx* ctx, char ta, char tb, int m, int n, int k, double alpha, const double* A, int lda,
                      const double* B, int ldb, double beta, double* C, int ldc) {
    rocblas_handle h;
    m2_status st = blas_handle(ctx, &h);
    if (st) return st;
    const rocblas_operation oa = (ta == 'T' || ta == 't') ? rocblas_operation_transpose : rocblas_operation_none;
    const rocblas_operation ob = (tb == 'T' || tb == 't') ? rocblas_operation_transpose : rocblas_operation_none;
    if (rocblas_dgemm(h, oa, ob, m, n, k, &alpha, A, lda, B, ldb, &beta, C, ldc) != rocblas_status_success) {
        ctx->err = "rocblas_dgemm failed";
        return M2_ERR_HIP;
    }
    return M2_OK;
}

m2_status potrf_device(m2_ctx* ctx, int n, double* dA, int* dinfo) {
    rocblas_handle h;
    m2_status st = blas_handle(ctx, &h);
    if (st) return st;
    if (rocsolver_dpotrf(h, rocblas_fill_lower, n, dA, n, dinfo) != rocblas_status_success) {
        ctx->err = "rocsolver_dpotrf failed";
        return M2_ERR_HIP;
    }
    return M2_OK;
}

// solve L X = B in place (L lower, n x n; B n x nrhs), column-major
m2_status trsm_lower_device(m2_ctx* ctx, int n, int nrhs, const double* dL, double* dB, int ldb) {
    rocblas_handle h;
    m2_status st = blas_handle(ctx, &h);
    if (st) return st;
    const double one = 1.0;
    if (rocblas_dtrsm(h, rocblas_side_left, rocblas_fill_lower, rocblas_operation_none, rocblas_diagonal_non_unit, n,
                      nrhs, &one, dL, n, dB, ldb) != rocblas_status_success) {
        ctx->err = "rocblas_dtrsm failed";
        return M2_ERR_HIP;
    }
    return M2_OK;
}

}  // namespace m2

using namespace m2;

extern "C" {

m2_status m2_eigh(m2_ctx* ctx, int n, const double* A, double* w, double* V) {
    if (!ctx) return M2_ERR_ARG;
    ARGCHK(n > 0 && A && w && V, "m2_eigh: bad arguments");
    HIPCHK(hipSetDevice(ctx->device));
    const size_t n2 = (size_t)n * n;
    HIPCHK(ctx->la_a.reserve(n2));
    HIPCHK(ctx->la_w.reserve(2 * (size_t)n));
    HIPCHK(ctx->la_info.reserve(1));
    HIPCHK(hipMemcpyAsync(ctx->la_a.p, A, sizeof(double) * n2, hipMemcpyHostToDevice, ctx->stream));
    m2_status st = eigh_device(ctx, n, ctx->la_a.p, ctx->la_w.p, ctx->la_w.p + n, ctx->la_info.p);
    if (st) return st;
    int info = 0;
    HIPCHK(hipMemcpyAsync(V, ctx->la_a.p, sizeof(double) * n2, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipMemcpyAsync(w, ctx->la_w.p, sizeof(double) * n, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipMemcpyAsync(&info, ctx->la_info.p, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if (info != 0) {
        ctx->err = "m2_eigh: rocsolver_dsyevd did not converge (info " + std::to_string(info) + ")";
        return M2_ERR_NUMERIC;
    }
    return M2_OK;
}

m2_status m2_gemm(m2_ctx* ctx, char ta, char tb, int m, int n, int k, double alpha, const double* A, int lda,
                  const double* B, int ldb, double beta, double* C, int ldc) {
    if (!ctx) return M2_ERR_ARG;
    ARGCHK(m > 0 && n > 0 && k > 0 && A && B && C, "m2_gemm: bad arguments");
    HIPCHK(hipSetDevice(ctx->device));
    const bool tA = (ta == 'T' || ta == 't'), tB = (tb == 'T' || tb == 't');
    const size_t sa = (size_t)lda * (tA ? m : k), sb = (size_t)ldb * (tB ? k : n), sc = (size_t)ldc * n;
    HIPCHK(ctx->la_a.reserve(sa));
    HIPCHK(ctx->la_b.reserve(sb));
    HIPCHK(ctx->la_c.reserve(sc));
    HIPCHK(hipMemcpyAsync(ctx->la_a.p, A, sizeof(double) * sa, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(ctx->la_b.p, B, sizeof(double) * sb, hipMemcpyHostToDevice, ctx->stream));
    if (beta != 0.0) HIPCHK(hipMemcpyAsync(ctx->la_c.p, C, sizeof(double) * sc, hipMemcpyHostToDevice, ctx->stream));
    m2_status st = gemm_device(ctx, ta, tb, m, n, k, alpha, ctx->la_a.p, lda, ctx->la_b.p, ldb, beta, ctx->la_c.p, ldc);
    if (st) return st;
    HIPCHK(hipMemcpyAsync(C, ctx->la_c.p, sizeof(double) * sc, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return M2_OK;
}

}  // extern "C"
