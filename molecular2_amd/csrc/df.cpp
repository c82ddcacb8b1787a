// Density-fitted J/K (the reference's Fock::formJKdf / compute_2body_fock_df,
// src/fockbuild.cpp:171-174,585-660).
//
// Set-up (once per geometry, FP64 on the device):
//   M[K][x][y] = (K|xy)       3-centre integrals, generated class kernels
//   V = (K|L), V = L L^T      2-centre metric, rocSOLVER dpotrf
//   B = M L^-T                rocBLAS dtrsm (row (x,y), column K)
//   stored as Bx[x][K][y] in FP64 and FP32 (the fitted tensor; n*n*naux).
// Per call, with C = C_occ (n x nocc, column-major) and Bx read as the
// column-major n x (n*naux) matrix B(y, (x,K)):
//   X  = C^T B                  GEMM  nocc x (n*naux), inner n         (xiK)
//   K  = Z^T Z, Z = X viewed as (naux*nocc) x n                        (k)
//   g_K = sum_{x,i} X(i,(x,K)) C(x,i)                                  (Jtmp)
//   J(:,x) = B(:, (x,.)) g      strided-batched GEMV over x            (j)
// The GEMMs are FP32 with FP32 accumulation on the MFMA units (rocBLAS
// sgemm) in the default mode, or FP64 (dgemm) when asked for FP64.
#include "ctx.hpp"
#include "df_kernel.hpp"

#include <rocblas/rocblas.h>

#include <chrono>
#include <cmath>

namespace m2 {
void launch_df3_0(int, int, const DF3Args&, const BoysTables&, const Task*, int, hipStream_t);
void launch_df3_1(int, int, const DF3Args&, const BoysTables&, const Task*, int, hipStream_t);
void launch_df3_2(int, int, const DF3Args&, const BoysTables&, const Task*, int, hipStream_t);
void launch_df3_3(int, int, const DF3Args&, const BoysTables&, const Task*, int, hipStream_t);
void launch_df2(int lp, int lq, int pure, const DF2Args& a, const BoysTables& tab, hipStream_t st);
m2_status potrf_device(m2_ctx* ctx, int n, double* dA, int* dinfo);

typedef void (*df3_launcher)(int, int, const DF3Args&, const BoysTables&, const Task*, int, hipStream_t);
static const df3_launcher DF3[4] = {launch_df3_0, launch_df3_1, launch_df3_2, launch_df3_3};

// [K][x][y] -> [x][K][y] in FP64 and FP32
__global__ void k_df_reorder(const double* __restrict__ M, double* __restrict__ B64, float* __restrict__ B32, int n,
                             int naux) {
    const size_t tot = (size_t)n * n * naux;
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < tot; t += (size_t)gridDim.x * blockDim.x) {
        const size_t y = t % n;
        const size_t K = (t / n) % naux;
        const size_t x = t / ((size_t)n * naux);
        const double v = M[(K * n + x) * n + y];
        B64[t] = v;
        B32[t] = (float)v;
    }
}

// g[K] = sum_{x,i} X[x][K][i] C[x + i*n]   (X as written by the X = C^T B GEMM),
// accumulated and stored in FP64 for either input precision
template <typename R>
__global__ __launch_bounds__(256) void k_df_gamma(const R* __restrict__ X, const R* __restrict__ C,
                                                  double* __restrict__ g, int n, int naux, int nocc) {
    const int K = blockIdx.x;
    __shared__ double red[256];
    double s = 0.0;
    for (int t = threadIdx.x; t < n * nocc; t += blockDim.x) {
        const int x = t / nocc, i = t % nocc;
        s += double(X[((size_t)x * naux + K) * nocc + i]) * double(C[(size_t)i * n + x]);
    }
    red[threadIdx.x] = s;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
        __syncthreads();
    }
    if (threadIdx.x == 0) g[K] = red[0];
}

// out[i] = sum_s part[s * m + i], summed and written in FP64
template <typename R>
__global__ void k_sum_slices(const R* __restrict__ part, double* __restrict__ out, size_t m, int ns) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= m) return;
    double v = 0.0;
    for (int s = 0; s < ns; ++s) v += double(part[(size_t)s * m + i]);
    out[i] = v;
}

// Coulomb part of compute_2body_fock_df (src/fockbuild.cpp:652-659),
// J(x, y) = sum_K B(xy, K) g_K, over the lower triangle y <= x of the fitted
// tensor stored Bx[x][K][y] (each unique pair read once), mirrored into J.
// One workgroup per (x, 256-wide y chunk); lanes over y read B coalesced,
// the loop runs over K with g staged in LDS; FP64 accumulation whatever the
// storage precision of B.
constexpr int DF_J_CHUNK = 256;
template <typename R>
__global__ __launch_bounds__(DF_J_CHUNK) void k_df_coulomb(const R* __restrict__ B, const double* __restrict__ g,
                                                           const int* __restrict__ blk_x, double* __restrict__ J,
                                                           int n, int naux) {
    extern __shared__ double sg[];
    const int x = blk_x[2 * blockIdx.x], y0 = blk_x[2 * blockIdx.x + 1];
    for (int k = threadIdx.x; k < naux; k += blockDim.x) sg[k] = g[k];
    __syncthreads();
    const int y = y0 + threadIdx.x;
    if (y > x) return;
    const R* b = B + (size_t)x * naux * n + y;
    double acc0 = 0.0, acc1 = 0.0, acc2 = 0.0, acc3 = 0.0;
    int k = 0;
    for (; k + 4 <= naux; k += 4) {
        acc0 += double(b[(size_t)k * n]) * sg[k];
        acc1 += double(b[(size_t)(k + 1) * n]) * sg[k + 1];
        acc2 += double(b[(size_t)(k + 2) * n]) * sg[k + 2];
        acc3 += double(b[(size_t)(k + 3) * n]) * sg[k + 3];
    }
    for (; k < naux; ++k) acc0 += double(b[(size_t)k * n]) * sg[k];
    const double v = (acc0 + acc1) + (acc2 + acc3);
    J[(size_t)x * n + y] = v;
    J[(size_t)y * n + x] = v;
}

// chunks of the auxiliary index of the exchange GEMM
constexpr int DF_KSPLIT = 48;

__global__ void k_to_float(const double* __restrict__ a, float* __restrict__ b, size_t n) {
    for (size_t t = blockIdx.x * (size_t)blockDim.x + threadIdx.x; t < n; t += (size_t)gridDim.x * blockDim.x)
        b[t] = float(a[t]);
}

}  // namespace m2

using namespace m2;

namespace {

BoysTables df_tables(m2_ctx* ctx) { return boys_tables(ctx); }

}  // namespace

namespace m2 {
// The raw 3-centre integrals M[K][x][y] = (K|xy) (n * n * naux doubles) and
// the Coulomb metric V = (K|L) (naux * naux), FP64, on the context's stream.
m2_status df_raw(m2_ctx* ctx, double* Mout, double* Vout) {
    BasisData& ob = ctx->obs;
    BasisData& ab = ctx->aux;
    ARGCHK(ob.nbf > 0, "m2_df_setup: no orbital basis set");
    ARGCHK(ab.nbf > 0, "m2_df_setup: no auxiliary basis set");
    for (const auto& s : ab.sh) ARGCHK(s.l <= 3, "m2_df_setup: auxiliary shells must have l <= 3");
    const int n = ob.nbf, naux = ab.nbf;
    // aux "pairs": each aux primitive with the unit s function
    const double sqrt2pi52 = std::sqrt(2.0 * std::pow(M_PI, 2.5));
    std::vector<PrimPair> pp;
    std::vector<int> off(ab.nsh), np(ab.nsh);
    for (int s = 0; s < ab.nsh; ++s) {
        off[s] = (int)pp.size();
        np[s] = (int)ab.sh[s].a.size();
        for (size_t k = 0; k < ab.sh[s].a.size(); ++k) {
            PrimPair q{};
            const double a = ab.sh[s].a[k];
            q.p = a;
            q.inv2p = 0.5 / a;
            q.K = sqrt2pi52 * ab.sh[s].c[k] / a;
            for (int d = 0; d < 3; ++d) {
                q.P[d] = ab.sh[s].O[d];
                q.PA[d] = 0.0;
            }
            q.a = a;
            pp.push_back(q);
        }
    }
    std::vector<int> lists;
    int lbeg[5] = {0};
    for (int l = 0; l <= 3; ++l) {
        lbeg[l] = (int)lists.size();
        for (int s = 0; s < ab.nsh; ++s)
            if (ab.sh[s].l == l) lists.push_back(s);
    }
    lbeg[4] = (int)lists.size();
    HIPCHK(ctx->d_aux_prims.reserve(pp.size()));
    HIPCHK(ctx->d_aux_off.reserve(ab.nsh));
    HIPCHK(ctx->d_aux_np.reserve(ab.nsh));
    HIPCHK(ctx->d_aux_lists.reserve(lists.size()));
    HIPCHK(hipMemcpy(ctx->d_aux_prims.p, pp.data(), sizeof(PrimPair) * pp.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->d_aux_off.p, off.data(), sizeof(int) * ab.nsh, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->d_aux_np.p, np.data(), sizeof(int) * ab.nsh, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->d_aux_lists.p, lists.data(), sizeof(int) * lists.size(), hipMemcpyHostToDevice));

    // 3-centre integrals into M[K][x][y]
    std::vector<Task> tasks;
    struct R3 {
        int lp, Y, t0, t1;
    };
    std::vector<R3> rr;
    for (int l = 0; l <= 3; ++l)
        for (int Y = 0; Y < NPAIRCLASS; ++Y) {
            const PairClassSeg& seg = ctx->cls[Y];
            if (seg.end <= seg.begin) continue;
            R3 r{l, Y, (int)tasks.size(), 0};
            for (int k = lbeg[l]; k < lbeg[l + 1]; ++k)
                for (int j = seg.begin; j < seg.end; j += 64)
                    tasks.push_back(Task{lists[k], j, std::min(64, seg.end - j), 0});
            r.t1 = (int)tasks.size();
            if (r.t1 > r.t0) rr.push_back(r);
        }
    DevBuf<Task> dt;
    HIPCHK(dt.reserve(tasks.size()));
    HIPCHK(hipMemcpy(dt.p, tasks.data(), sizeof(Task) * tasks.size(), hipMemcpyHostToDevice));
    DF3Args a3{};
    a3.aux_prims = ctx->d_aux_prims.p;
    a3.aux_off = ctx->d_aux_off.p;
    a3.aux_np = ctx->d_aux_np.p;
    a3.aux_bf = ab.d_bf.p;
    a3.prims = ctx->d_prims.p;
    a3.pair_off = ctx->d_pair_off.p;
    a3.pair_np = ctx->d_pair_np.p;
    a3.pair_sh = ctx->d_pair_sh.p;
    a3.sh_O = ob.d_O.p;
    a3.sh_bf = ob.d_bf.p;
    a3.M = Mout;
    a3.n = n;
    a3.naux = naux;
    for (const R3& r : rr) {
        DF3[r.lp](r.Y, ab.pure, a3, df_tables(ctx), dt.p + r.t0, r.t1 - r.t0, ctx->stream);
        HIPCHK(hipGetLastError());
    }
    // 2-centre metric
    for (int lp = 0; lp <= 3; ++lp)
        for (int lq = 0; lq <= lp; ++lq) {
            DF2Args a2{};
            a2.aux_prims = ctx->d_aux_prims.p;
            a2.aux_off = ctx->d_aux_off.p;
            a2.aux_np = ctx->d_aux_np.p;
            a2.aux_bf = ab.d_bf.p;
            a2.P_list = ctx->d_aux_lists.p + lbeg[lp];
            a2.nP = lbeg[lp + 1] - lbeg[lp];
            a2.Q_list = ctx->d_aux_lists.p + lbeg[lq];
            a2.nQ = lbeg[lq + 1] - lbeg[lq];
            a2.V = Vout;
            a2.naux = naux;
            launch_df2(lp, lq, ab.pure, a2, df_tables(ctx), ctx->stream);
            HIPCHK(hipGetLastError());
        }
    // the task list is freed once the kernels that read it have run
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return M2_OK;
}
}  // namespace m2

namespace {

m2_status df_setup(m2_ctx* ctx) {
    BasisData& ob = ctx->obs;
    BasisData& ab = ctx->aux;
    ARGCHK(ob.nbf > 0, "m2_df_setup: no orbital basis set");
    ARGCHK(ab.nbf > 0, "m2_df_setup: no auxiliary basis set");
    const int n = ob.nbf, naux = ab.nbf;
    const size_t nB = (size_t)n * n * naux;
    DevBuf<double> M;
    HIPCHK(M.reserve(nB));
    HIPCHK(ctx->d_dfV.reserve((size_t)naux * naux));
    m2_status st0 = df_raw(ctx, M.p, ctx->d_dfV.p);
    if (st0) return st0;
    // Cholesky and fitting
    DevBuf<int> info;
    HIPCHK(info.reserve(1));
    m2_status st = potrf_device(ctx, naux, ctx->d_dfV.p, info.p);
    if (st) return st;
    int hinfo = 0;
    HIPCHK(hipMemcpyAsync(&hinfo, info.p, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if (hinfo != 0) {
        ctx->err = "m2_df_setup: the Coulomb metric is not positive definite (dpotrf info " + std::to_string(hinfo) + ")";
        return M2_ERR_NUMERIC;
    }
    void* hv = nullptr;
    st = la_handle(ctx, &hv);
    if (st) return st;
    rocblas_handle h = (rocblas_handle)hv;
    const double one = 1.0;
    // B(xy, K): X L^T = M  (right side, lower L, transposed)
    if (rocblas_dtrsm(h, rocblas_side_right, rocblas_fill_lower, rocblas_operation_transpose, rocblas_diagonal_non_unit,
                      n * n, naux, &one, ctx->d_dfV.p, naux, M.p, n * n) != rocblas_status_success) {
        ctx->err = "m2_df_setup: rocblas_dtrsm failed";
        return M2_ERR_HIP;
    }
    HIPCHK(ctx->d_B64.reserve(nB));
    HIPCHK(ctx->d_B32.reserve(nB));
    hipLaunchKernelGGL(k_df_reorder, dim3(8192), dim3(256), 0, ctx->stream, M.p, ctx->d_B64.p, ctx->d_B32.p, n, naux);
    HIPCHK(hipGetLastError());
    // workgroups of the Coulomb contraction: (x, first y) over y <= x
    std::vector<int> jb;
    for (int x = 0; x < n; ++x)
        for (int y0 = 0; y0 <= x; y0 += DF_J_CHUNK) {
            jb.push_back(x);
            jb.push_back(y0);
        }
    ctx->df_jblocks = (int)jb.size() / 2;
    HIPCHK(ctx->d_df_jblk.reserve(jb.size()));
    HIPCHK(hipMemcpy(ctx->d_df_jblk.p, jb.data(), sizeof(int) * jb.size(), hipMemcpyHostToDevice));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    M.release();
    info.release();
    ctx->df_ready = true;
    return M2_OK;
}

template <typename R>
struct Blas;
template <>
struct Blas<float> {
    static rocblas_status gemm(rocblas_handle h, rocblas_operation a, rocblas_operation b, int m, int n, int k,
                               const float* al, const float* A, int lda, const float* B, int ldb, const float* be,
                               float* C, int ldc) {
        return rocblas_sgemm(h, a, b, m, n, k, al, A, lda, B, ldb, be, C, ldc);
    }
    static rocblas_status gemm_sb(rocblas_handle h, rocblas_operation a, rocblas_operation b, int m, int n, int k,
                                  const float* al, const float* A, int lda, rocblas_stride sA, const float* B, int ldb,
                                  rocblas_stride sB, const float* be, float* C, int ldc, rocblas_stride sC, int batch) {
        return rocblas_sgemm_strided_batched(h, a, b, m, n, k, al, A, lda, sA, B, ldb, sB, be, C, ldc, sC, batch);
    }
    static float* partial(m2_ctx* ctx, size_t cnt, hipError_t& e) {
        e = ctx->d_dfKp32.reserve(cnt);
        return ctx->d_dfKp32.p;
    }
};
template <>
struct Blas<double> {
    static rocblas_status gemm(rocblas_handle h, rocblas_operation a, rocblas_operation b, int m, int n, int k,
                               const double* al, const double* A, int lda, const double* B, int ldb, const double* be,
                               double* C, int ldc) {
        return rocblas_dgemm(h, a, b, m, n, k, al, A, lda, B, ldb, be, C, ldc);
    }
    static rocblas_status gemm_sb(rocblas_handle h, rocblas_operation a, rocblas_operation b, int m, int n, int k,
                                  const double* al, const double* A, int lda, rocblas_stride sA, const double* B,
                                  int ldb, rocblas_stride sB, const double* be, double* C, int ldc, rocblas_stride sC,
                                  int batch) {
        return rocblas_dgemm_strided_batched(h, a, b, m, n, k, al, A, lda, sA, B, ldb, sB, be, C, ldc, sC, batch);
    }
    static double* partial(m2_ctx* ctx, size_t cnt, hipError_t& e) {
        e = ctx->d_dfKp.reserve(cnt);
        return ctx->d_dfKp.p;
    }
};

template <typename R>
m2_status jk_df_core(m2_ctx* ctx, const R* B, const R* C, int nocc, R* X, double* g, double* J, double* K) {
    const int n = ctx->obs.nbf, naux = ctx->aux.nbf;
    void* hv = nullptr;
    m2_status st = la_handle(ctx, &hv);
    if (st) return st;
    rocblas_handle h = (rocblas_handle)hv;
    const R one = R(1), zero = R(0);
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    if (ctx->prof)
        for (auto& e : ev) HIPCHK(hipEventCreate(&e));
    if (ctx->prof) HIPCHK(hipEventRecord(ev[0], ctx->stream));
    // X = C^T B : (nocc x n) (n x n*naux)
    if (Blas<R>::gemm(h, rocblas_operation_transpose, rocblas_operation_none, nocc, n * naux, n, &one, C, n, B, n, &zero,
                      X, nocc) != rocblas_status_success) {
        ctx->err = "m2_jk_df: GEMM X = C^T B failed";
        return M2_ERR_HIP;
    }
    if (ctx->prof) HIPCHK(hipEventRecord(ev[1], ctx->stream));
    // K = Z^T Z with Z = X as (naux*nocc) x n, row i + nocc*K.  The output is
    // only n x n while the inner dimension is naux*nocc, so the product is
    // split over chunks of the auxiliary index (contiguous rows of Z): one
    // strided-batched GEMM into partial K slices, then their sum.
    const int ld = naux * nocc;
    const int nsplit = std::max(1, std::min(DF_KSPLIT, naux / 16));
    const int ks = (naux + nsplit - 1) / nsplit;       // aux functions per chunk
    const int nfull = naux / ks, krem = naux - nfull * ks;
    const int nslice = nfull + (krem > 0 ? 1 : 0);
    hipError_t pe;
    R* Kp = Blas<R>::partial(ctx, (size_t)nslice * n * n, pe);
    HIPCHK(pe);
    if (Blas<R>::gemm_sb(h, rocblas_operation_transpose, rocblas_operation_none, n, n, ks * nocc, &one, X, ld,
                         (rocblas_stride)ks * nocc, X, ld, (rocblas_stride)ks * nocc, &zero, Kp, n,
                         (rocblas_stride)n * n, nfull) != rocblas_status_success) {
        ctx->err = "m2_jk_df: split GEMM K = X^T X failed";
        return M2_ERR_HIP;
    }
    if (krem > 0) {
        const R* Xr = X + (size_t)nfull * ks * nocc;
        if (Blas<R>::gemm(h, rocblas_operation_transpose, rocblas_operation_none, n, n, krem * nocc, &one, Xr, ld, Xr,
                          ld, &zero, Kp + (size_t)nfull * n * n, n) != rocblas_status_success) {
            ctx->err = "m2_jk_df: GEMM K = X^T X (last chunk) failed";
            return M2_ERR_HIP;
        }
    }
    hipLaunchKernelGGL(k_sum_slices<R>, dim3(((size_t)n * n + 255) / 256), dim3(256), 0, ctx->stream, Kp, K,
                       (size_t)n * n, nslice);
    HIPCHK(hipGetLastError());
    if (ctx->prof) HIPCHK(hipEventRecord(ev[2], ctx->stream));
    hipLaunchKernelGGL(k_df_gamma<R>, dim3(naux), dim3(256), 0, ctx->stream, X, C, g, n, naux, nocc);
    HIPCHK(hipGetLastError());
    // J over the lower triangle of B, mirrored
    hipLaunchKernelGGL(k_df_coulomb<R>, dim3(ctx->df_jblocks), dim3(DF_J_CHUNK), sizeof(double) * naux, ctx->stream, B,
                       g, ctx->d_df_jblk.p, J, n, naux);
    HIPCHK(hipGetLastError());
    if (ctx->prof) {
        HIPCHK(hipEventRecord(ev[3], ctx->stream));
        HIPCHK(hipEventSynchronize(ev[3]));
        float ms[3];
        for (int i = 0; i < 3; ++i) HIPCHK(hipEventElapsedTime(&ms[i], ev[i], ev[i + 1]));
        ctx->df_last_ms[0] = ms[0];
        ctx->df_last_ms[1] = ms[1];
        ctx->df_last_ms[2] = ms[2];
        ctx->df_last_ms[3] = ms[0] + ms[1] + ms[2];
        for (auto& e : ev) (void)hipEventDestroy(e);
    }
    return M2_OK;
}

m2_status jk_df_device(m2_ctx* ctx, const double* dC, int nocc, int fp32, double* dJ, double* dK) {
    const int n = ctx->obs.nbf, naux = ctx->aux.nbf;
    if (!ctx->df_ready) {
        m2_status st = df_setup(ctx);
        if (st) return st;
    }
    const size_t nx = (size_t)nocc * n * naux;
    HIPCHK(ctx->d_dfG.reserve(naux));
    if (fp32) {
        // FP32 storage and MFMA GEMMs; gamma, the slice sum of K and the
        // Coulomb contraction accumulate in FP64 and write J, K directly
        HIPCHK(ctx->d_dfC32.reserve((size_t)n * nocc));
        HIPCHK(ctx->d_dfX32.reserve(nx));
        hipLaunchKernelGGL(k_to_float, dim3(256), dim3(256), 0, ctx->stream, dC, ctx->d_dfC32.p, (size_t)n * nocc);
        m2_status st = jk_df_core<float>(ctx, ctx->d_B32.p, ctx->d_dfC32.p, nocc, ctx->d_dfX32.p, ctx->d_dfG.p, dJ, dK);
        if (st) return st;
    } else {
        HIPCHK(ctx->d_dfX.reserve(nx));
        m2_status st = jk_df_core<double>(ctx, ctx->d_B64.p, dC, nocc, ctx->d_dfX.p, ctx->d_dfG.p, dJ, dK);
        if (st) return st;
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return M2_OK;
}

}  // namespace

extern "C" {

m2_status m2_df_setup(m2_ctx* ctx) {
    if (!ctx) return M2_ERR_ARG;
    HIPCHK(hipSetDevice(ctx->device));
    return df_setup(ctx);
}

m2_status m2_jk_df(m2_ctx* ctx, const double* Cocc, int nocc, int fp32, double* J, double* K) {
    if (!ctx) return M2_ERR_ARG;
    ARGCHK(Cocc && J && K && nocc > 0, "m2_jk_df: bad arguments");
    ARGCHK(ctx->obs.nbf > 0 && ctx->aux.nbf > 0, "m2_jk_df: orbital and auxiliary bases must be set");
    ARGCHK(nocc <= ctx->obs.nbf, "m2_jk_df: nocc exceeds the number of basis functions");
    HIPCHK(hipSetDevice(ctx->device));
    const int n = ctx->obs.nbf;
    HIPCHK(ctx->d_dfC.reserve((size_t)n * nocc));
    HIPCHK(ctx->d_dfJ.reserve((size_t)n * n));
    HIPCHK(ctx->d_dfK.reserve((size_t)n * n));
    HIPCHK(hipMemcpyAsync(ctx->d_dfC.p, Cocc, sizeof(double) * n * nocc, hipMemcpyHostToDevice, ctx->stream));
    m2_status st = jk_df_device(ctx, ctx->d_dfC.p, nocc, fp32, ctx->d_dfJ.p, ctx->d_dfK.p);
    if (st) return st;
    HIPCHK(hipMemcpyAsync(J, ctx->d_dfJ.p, sizeof(double) * n * n, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipMemcpyAsync(K, ctx->d_dfK.p, sizeof(double) * n * n, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return M2_OK;
}

m2_status m2_jk_df_device(m2_ctx* ctx, const double* dCocc, int nocc, int fp32, double* dJ, double* dK) {
    if (!ctx) return M2_ERR_ARG;
    ARGCHK(dCocc && dJ && dK && nocc > 0, "m2_jk_df_device: bad arguments");
    ARGCHK(ctx->obs.nbf > 0 && ctx->aux.nbf > 0, "m2_jk_df_device: orbital and auxiliary bases must be set");
    HIPCHK(hipSetDevice(ctx->device));
    return jk_df_device(ctx, dCocc, nocc, fp32, dJ, dK);
}

m2_status m2_df_timing(m2_ctx* ctx, double* ms4) {
    if (!ctx || !ms4) return M2_ERR_ARG;
    for (int i = 0; i < 4; ++i) ms4[i] = ctx->df_last_ms[i];
    return M2_OK;
}

}  // extern "C"
