// Fragment-local density-fitted J/K: Fock::compute_2body_fock_df_local with
// build_blocked_eris and build_domains (src/fockbuild.cpp:662-1047), the
// Fock build of the reference's ALMO SCF.
//
// Set-up (first call; the reference builds its blocks and domains while its
// xyK is still empty, src/fockbuild.cpp:810-825):
//   * the raw 3-centre tensor M[K][x][y] = (K|xy) and the metric V = (K|L)
//     on the device (the DF kernels), with every fragment block
//     (frag(x), frag(y), frag(K)) that build_blocked_eris screens out set to
//     zero: the screened tensor equals the reference's block store;
//   * the LMO, AO and fit domains on the host (build_domains: pure index
//     logic over Cocc sigmainv^(1/2) and S^(1/2) Cocc sigmainv^(1/2));
//   * per LMO the Cholesky factor L_i of V restricted to its fit domain
//     (rocSOLVER dpotrf), and the factor of the whole metric.
// Per call, with C' = Cocc sigmainv^(1/2) (rocSOLVER eigen-decomposition):
//   exchange: C'_i masked to the LMO's domain fragments; X = C'^T M (one
//     GEMM for all LMOs, X(i, (K, x)) = sum_y C'_i(y) (K|xy)); U_i = X
//     restricted to (AO domain x fit domain); Y_i = U_i L_i^-T (dtrsm);
//     K[AO_i, AO_i] += Y_i Y_i^T (dgemm, then a scatter-add);
//   Coulomb: Pk = 1/2 M^T vec(Pt) (every kept block once, diagonal blocks
//     halved, as the reference's f2 <= f1 loop with its 1/2 on f1 == f2),
//     Pk = V^-1 Pk (two triangular solves), J = 4 M Pk.
// Everything runs in FP64, as the reference does.
#include "ctx.hpp"

#include <rocblas/rocblas.h>

#include <cmath>
#include <cstring>
#include <string>
#include <vector>

namespace m2 {
m2_status df_raw(m2_ctx* ctx, double* M, double* V);
m2_status potrf_device(m2_ctx* ctx, int n, double* dA, int* dinfo);
m2_status eigh_device(m2_ctx* ctx, int n, double* dA, double* dw, double* dE, int* dinfo);
m2_status gemm_device(m2_ctx* ctx, char ta, char tb, int m, int n, int k, double alpha, const double* A, int lda,
                      const double* B, int ldb, double beta, double* C, int ldc);
m2_status la_handle(m2_ctx* ctx, void** h);

// Domain (inc/fock.hpp:37-57): fragment centres and their (start, size) ranges
struct LocalDomain {
    std::vector<int> starts, sizes, centres;
    int total = 0;
    void add(int c, int start, int size) {
        centres.push_back(c);
        starts.push_back(start);
        sizes.push_back(size);
        total += size;
    }
};

struct LocalDF {
    std::vector<m2_fragment> frag;
    double mo_thresh = 1e-6, fit_thresh = 0.05, r_thresh = 15.0, thrint = 1e-12;
    int nocc = 0;
    std::vector<unsigned char> keep;   // nf^3, symmetric in the first two
    std::vector<LocalDomain> lmo, ao, fit;
    DevBuf<double> M;                  // screened raw tensor [K][x][y]
    DevBuf<double> Lv;                 // Cholesky factor of the whole metric
    DevBuf<double> Lfit;               // per-LMO fit-metric factors
    std::vector<size_t> lfit_off, u_off, kw_off;
    std::vector<int> ao_off, fit_off;
    DevBuf<int> ao_idx, fit_idx, mo_mask;
    DevBuf<double> X, U, Kw, Cp, Cm, P, Pk, J, K, sig, w, E;
    DevBuf<int> info;
};

void local_df_release(m2_ctx* ctx) {
    delete ctx->ldf;
    ctx->ldf = nullptr;
}
}  // namespace m2

using namespace m2;

namespace {

// M[K][x][y] = 0 where build_blocked_eris keeps no block (frag x, frag y, frag K)
__global__ void k_mask_blocks(double* __restrict__ M, const int* __restrict__ fbf, const int* __restrict__ faux,
                              const unsigned char* __restrict__ keep, int n, int naux, int nf) {
    const size_t total = (size_t)n * n * naux;
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
        const int y = (int)(e % n), x = (int)((e / n) % n), K = (int)(e / ((size_t)n * n));
        if (!keep[((size_t)fbf[x] * nf + fbf[y]) * nf + faux[K]]) M[e] = 0.0;
    }
}

// Cm(y, i) = C'(y, i) when frag(y) lies in LMO i's domain, else 0 (column-major n x nocc)
__global__ void k_mask_mo(const double* __restrict__ C, double* __restrict__ Cm, const int* __restrict__ fbf,
                          const int* __restrict__ mask, int n, int nocc, int nf) {
    const size_t total = (size_t)n * nocc;
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
        const int y = (int)(e % n), i = (int)(e / n);
        Cm[e] = mask[(size_t)i * nf + fbf[y]] ? C[e] : 0.0;
    }
}

// U_i(r, c) = X(i, (fit[c], ao[r])) for LMO i = blockIdx.y, column-major a_i x f_i
__global__ void k_gather_u(const double* __restrict__ X, double* __restrict__ U, const int* __restrict__ ao_idx,
                           const int* __restrict__ fit_idx, const int* __restrict__ ao_off,
                           const int* __restrict__ fit_off, const size_t* __restrict__ u_off, int n, int nocc) {
    const int i = blockIdx.y;
    const int a = ao_off[i + 1] - ao_off[i], f = fit_off[i + 1] - fit_off[i];
    const int* ao = ao_idx + ao_off[i];
    const int* ft = fit_idx + fit_off[i];
    double* Ui = U + u_off[i];
    const size_t total = (size_t)a * f;
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
        const int r = (int)(e % a), c = (int)(e / a);
        Ui[e] = X[i + (size_t)nocc * ((size_t)ft[c] * n + ao[r])];
    }
}

// K[ao[r], ao[s]] += Kw_i(r, s) for LMO i = blockIdx.y (several LMOs share elements: FP64 atomics)
__global__ void k_scatter_k(const double* __restrict__ Kw, double* __restrict__ K, const int* __restrict__ ao_idx,
                            const int* __restrict__ ao_off, const size_t* __restrict__ kw_off, int n) {
    const int i = blockIdx.y;
    const int a = ao_off[i + 1] - ao_off[i];
    const int* ao = ao_idx + ao_off[i];
    const double* Ki = Kw + kw_off[i];
    const size_t total = (size_t)a * a;
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
        const int r = (int)(e % a), s = (int)(e / a);
        atomicAdd(&K[(size_t)ao[r] * n + ao[s]], Ki[e]);
    }
}

// V restricted to the fit domain of LMO i = blockIdx.y, column-major f_i x f_i
__global__ void k_gather_vfit(const double* __restrict__ V, double* __restrict__ L, const int* __restrict__ fit_idx,
                              const int* __restrict__ fit_off, const size_t* __restrict__ l_off, int naux) {
    const int i = blockIdx.y;
    const int f = fit_off[i + 1] - fit_off[i];
    const int* ft = fit_idx + fit_off[i];
    double* Li = L + l_off[i];
    const size_t total = (size_t)f * f;
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
        const int r = (int)(e % f), c = (int)(e / f);
        Li[e] = V[(size_t)ft[c] * naux + ft[r]];
    }
}

// E(:, j) *= sqrt(w_j): the columns of the eigenvectors scaled for V sqrt(D) V^T
__global__ void k_scale_sqrt(double* __restrict__ E, const double* __restrict__ w, int n) {
    const size_t total = (size_t)n * n;
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x)
        E[e] *= sqrt(w[e / n]);
}

__global__ void k_zero(double* __restrict__ p, size_t cnt) {
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < cnt; e += (size_t)gridDim.x * blockDim.x)
        p[e] = 0.0;
}

int grid_for(size_t cnt) { return (int)std::max<size_t>(1, std::min<size_t>((cnt + 255) / 256, 4096)); }

// A^(1/2) of a symmetric n x n device matrix (the reference's operatorSqrt):
// V sqrt(D) V^T from the eigen-decomposition, into out
m2_status sym_sqrt(m2_ctx* ctx, LocalDF& L, const double* dA, int n, double* out) {
    HIPCHK(L.sig.reserve((size_t)n * n));
    HIPCHK(L.E.reserve((size_t)n * n));
    HIPCHK(L.w.reserve((size_t)2 * n));
    HIPCHK(L.info.reserve(1));
    HIPCHK(hipMemcpyAsync(L.sig.p, dA, sizeof(double) * n * n, hipMemcpyDeviceToDevice, ctx->stream));
    m2_status st = eigh_device(ctx, n, L.sig.p, L.w.p, L.w.p + n, L.info.p);
    if (st) return st;
    int hinfo = 0;
    HIPCHK(hipMemcpyAsync(&hinfo, L.info.p, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    ARGCHK(hinfo == 0, "m2_jk_df_local: eigen-decomposition failed");
    std::vector<double> w(n);
    HIPCHK(hipMemcpy(w.data(), L.w.p, sizeof(double) * n, hipMemcpyDeviceToHost));
    for (double x : w) ARGCHK(x >= 0.0, "m2_jk_df_local: matrix square root of a matrix with a negative eigenvalue");
    HIPCHK(hipMemcpyAsync(L.E.p, L.sig.p, sizeof(double) * n * n, hipMemcpyDeviceToDevice, ctx->stream));
    hipLaunchKernelGGL(k_scale_sqrt, dim3(grid_for((size_t)n * n)), dim3(256), 0, ctx->stream, L.E.p, L.w.p, n);
    HIPCHK(hipGetLastError());
    return gemm_device(ctx, 'N', 'T', n, n, n, 1.0, L.E.p, n, L.sig.p, n, 0.0, out, n);
}

double dist(const m2_fragment& a, const m2_fragment& b) {
    const double dx = a.com[0] - b.com[0], dy = a.com[1] - b.com[1], dz = a.com[2] - b.com[2];
    return std::sqrt(dx * dx + dy * dy + dz * dz);
}

// build_blocked_eris (src/fockbuild.cpp:985-1047): which (f1, f2, fk) blocks are kept
void blocked_eris_mask(m2_ctx* ctx, LocalDF& L, const std::vector<double>& Pt, const std::vector<double>& V) {
    const int nf = (int)L.frag.size(), n = ctx->obs.nbf, naux = ctx->aux.nbf, nsh = ctx->obs.nsh;
    L.keep.assign((size_t)nf * nf * nf, 0);
    std::vector<int> sh0(nf + 1, 0);
    for (int f = 0; f < nf; ++f) sh0[f + 1] = sh0[f] + L.frag[f].nshells;
    for (int f1 = 0; f1 < nf; ++f1) {
        const m2_fragment& a = L.frag[f1];
        for (int f2 = 0; f2 <= f1; ++f2) {
            const m2_fragment& b = L.frag[f2];
            double infnorm = 0.0, ptnorm = 0.0;
            for (int s = sh0[f1]; s < sh0[f1 + 1]; ++s)
                for (int t = sh0[f2]; t < sh0[f2 + 1]; ++t) infnorm = std::max(infnorm, std::fabs(ctx->Qsh[(size_t)s * nsh + t]));
            for (int x = a.start; x < a.start + a.nbfs; ++x)
                for (int y = b.start; y < b.start + b.nbfs; ++y) ptnorm = std::max(ptnorm, std::fabs(Pt[(size_t)x * n + y]));
            for (int fk = 0; fk < nf; ++fk) {
                const m2_fragment& k = L.frag[fk];
                double vmax = 0.0;
                for (int P = k.auxstart; P < k.auxstart + k.naux; ++P)
                    for (int Q = k.auxstart; Q < k.auxstart + k.naux; ++Q)
                        vmax = std::max(vmax, std::fabs(V[(size_t)P * naux + Q]));
                const double infjpq = std::sqrt(vmax);
                const double rxyz = 0.5 * (dist(a, k) + dist(b, k));
                const double est = rxyz < 1.0 ? 1.0 : 2.0 * infnorm * ptnorm * infjpq / rxyz;
                if (est > L.thrint) {
                    L.keep[((size_t)f1 * nf + f2) * nf + fk] = 1;
                    L.keep[((size_t)f2 * nf + f1) * nf + fk] = 1;
                }
            }
        }
    }
}

// build_domains (src/fockbuild.cpp:662-768): C = Cocc sigmainv^(1/2) and
// SC = S^(1/2) C on the host, column-major n x nocc
void build_domains(LocalDF& L, int n, const std::vector<double>& C, const std::vector<double>& SC) {
    const int nf = (int)L.frag.size();
    L.lmo.clear();
    L.ao.clear();
    L.fit.clear();
    int i = 0;
    for (int B = 0; B < nf; ++B)
        for (int o = 0; o < L.frag[B].occ; ++o, ++i) {
            LocalDomain d, dao;
            std::vector<int> notc;
            int mu_off = 0;
            for (int A = 0; A < nf; ++A) {
                bool add = A == B;
                if (!add) {
                    // the reference sums occ(A) rows from the fragment's
                    // first function (src/fockbuild.cpp:675-677)
                    double sum = 0.0;
                    for (int mu = mu_off; mu < mu_off + L.frag[A].occ; ++mu) sum += C[(size_t)i * n + mu] * C[(size_t)i * n + mu];
                    add = sum > L.mo_thresh;
                }
                if (add) d.add(A, L.frag[A].start, L.frag[A].nbfs);
                else notc.push_back(A);
                mu_off += L.frag[A].nbfs;
            }
            dao = d;
            for (int A : d.centres) {
                std::vector<int> newnc;
                for (int Cc : notc) {
                    if (L.frag[Cc].radius > dist(L.frag[Cc], L.frag[A])) dao.add(Cc, L.frag[Cc].start, L.frag[Cc].nbfs);
                    else newnc.push_back(Cc);
                }
                notc.swap(newnc);
            }
            L.lmo.push_back(d);
            L.ao.push_back(dao);
        }
    i = 0;
    for (int f1 = 0; f1 < nf; ++f1)
        for (int o = 0; o < L.frag[f1].occ; ++o, ++i) {
            LocalDomain d;
            for (int f2 = 0; f2 < nf; ++f2) {
                const double sep = dist(L.frag[f2], L.frag[f1]);
                double INi = 0.0;
                for (int mu = L.frag[f2].start; mu < L.frag[f2].start + L.frag[f2].nbfs; ++mu)
                    INi += SC[(size_t)i * n + mu] * SC[(size_t)i * n + mu];
                if (INi > L.fit_thresh || sep < L.r_thresh) d.add(f2, L.frag[f2].auxstart, L.frag[f2].naux);
            }
            L.fit.push_back(d);
        }
}

template <typename T>
m2_status upload(m2_ctx* ctx, DevBuf<T>& b, const std::vector<T>& h) {
    HIPCHK(b.reserve(std::max<size_t>(h.size(), 1)));
    if (!h.empty()) HIPCHK(hipMemcpy(b.p, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice));
    return M2_OK;
}

m2_status local_setup(m2_ctx* ctx, LocalDF& L, const double* dCp, const double* dPt) {
    const int n = ctx->obs.nbf, naux = ctx->aux.nbf, nocc = L.nocc, nf = (int)L.frag.size();
    // raw tensor and metric
    HIPCHK(L.M.reserve((size_t)n * n * naux));
    HIPCHK(L.Lv.reserve((size_t)naux * naux));
    m2_status st = df_raw(ctx, L.M.p, L.Lv.p);
    if (st) return st;
    std::vector<double> V((size_t)naux * naux), Pt((size_t)n * n);
    HIPCHK(hipMemcpy(V.data(), L.Lv.p, sizeof(double) * V.size(), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(Pt.data(), dPt, sizeof(double) * Pt.size(), hipMemcpyDeviceToHost));
    blocked_eris_mask(ctx, L, Pt, V);
    std::vector<int> fbf(n), faux(naux);
    for (int f = 0; f < nf; ++f) {
        for (int x = L.frag[f].start; x < L.frag[f].start + L.frag[f].nbfs; ++x) fbf[x] = f;
        for (int P = L.frag[f].auxstart; P < L.frag[f].auxstart + L.frag[f].naux; ++P) faux[P] = f;
    }
    DevBuf<int> dfbf, dfaux;
    DevBuf<unsigned char> dkeep;
    if ((st = upload(ctx, dfbf, fbf)) || (st = upload(ctx, dfaux, faux)) || (st = upload(ctx, dkeep, L.keep))) return st;
    hipLaunchKernelGGL(k_mask_blocks, dim3(8192), dim3(256), 0, ctx->stream, L.M.p, dfbf.p, dfaux.p, dkeep.p, n, naux, nf);
    HIPCHK(hipGetLastError());
    // domains from C' and S^(1/2) C'
    std::vector<double> S((size_t)n * n), T((size_t)n * n), Vn((size_t)n * n);
    st = m2_one_body(ctx, 0, nullptr, nullptr, S.data(), T.data(), Vn.data());
    if (st) return st;
    DevBuf<double> dS, dS2, dSC;
    HIPCHK(dS.reserve((size_t)n * n));
    HIPCHK(dS2.reserve((size_t)n * n));
    HIPCHK(dSC.reserve((size_t)n * nocc));
    HIPCHK(hipMemcpy(dS.p, S.data(), sizeof(double) * S.size(), hipMemcpyHostToDevice));
    if ((st = sym_sqrt(ctx, L, dS.p, n, dS2.p))) return st;
    if ((st = gemm_device(ctx, 'N', 'N', n, nocc, n, 1.0, dS2.p, n, dCp, n, 0.0, dSC.p, n))) return st;
    std::vector<double> C((size_t)n * nocc), SC((size_t)n * nocc);
    HIPCHK(hipMemcpyAsync(C.data(), dCp, sizeof(double) * C.size(), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipMemcpyAsync(SC.data(), dSC.p, sizeof(double) * SC.size(), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    build_domains(L, n, C, SC);
    // index tables
    std::vector<int> ao_idx, fit_idx, mo_mask((size_t)nocc * nf, 0);
    L.ao_off.assign(1, 0);
    L.fit_off.assign(1, 0);
    L.lfit_off.assign(1, 0);
    L.u_off.assign(1, 0);
    L.kw_off.assign(1, 0);
    for (int i = 0; i < nocc; ++i) {
        for (size_t c = 0; c < L.ao[i].centres.size(); ++c)
            for (int x = 0; x < L.ao[i].sizes[c]; ++x) ao_idx.push_back(L.ao[i].starts[c] + x);
        for (size_t c = 0; c < L.fit[i].centres.size(); ++c)
            for (int P = 0; P < L.fit[i].sizes[c]; ++P) fit_idx.push_back(L.fit[i].starts[c] + P);
        for (int A : L.lmo[i].centres) mo_mask[(size_t)i * nf + A] = 1;
        const size_t a = L.ao[i].total, f = L.fit[i].total;
        L.ao_off.push_back((int)ao_idx.size());
        L.fit_off.push_back((int)fit_idx.size());
        L.lfit_off.push_back(L.lfit_off.back() + f * f);
        L.u_off.push_back(L.u_off.back() + a * f);
        L.kw_off.push_back(L.kw_off.back() + a * a);
    }
    if ((st = upload(ctx, L.ao_idx, ao_idx)) || (st = upload(ctx, L.fit_idx, fit_idx)) ||
        (st = upload(ctx, L.mo_mask, mo_mask)))
        return st;
    // per-LMO fit-metric factors, then the whole metric's
    DevBuf<int> dao_off, dfit_off;
    DevBuf<size_t> dl_off;
    if ((st = upload(ctx, dao_off, L.ao_off)) || (st = upload(ctx, dfit_off, L.fit_off)) ||
        (st = upload(ctx, dl_off, L.lfit_off)))
        return st;
    HIPCHK(L.Lfit.reserve(std::max<size_t>(L.lfit_off.back(), 1)));
    hipLaunchKernelGGL(k_gather_vfit, dim3(64, nocc), dim3(256), 0, ctx->stream, L.Lv.p, L.Lfit.p, L.fit_idx.p,
                       dfit_off.p, dl_off.p, naux);
    HIPCHK(hipGetLastError());
    HIPCHK(L.info.reserve(1));
    for (int i = 0; i < nocc; ++i) {
        const int f = L.fit_off[i + 1] - L.fit_off[i];
        if (f == 0) continue;   // an empty fit domain: the LMO adds no exchange
        if ((st = potrf_device(ctx, f, L.Lfit.p + L.lfit_off[i], L.info.p))) return st;
        int hinfo = 0;
        HIPCHK(hipMemcpyAsync(&hinfo, L.info.p, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
        ARGCHK(hinfo == 0, "m2_jk_df_local: a fit-domain metric is not positive definite");
    }
    if ((st = potrf_device(ctx, naux, L.Lv.p, L.info.p))) return st;
    int hinfo = 0;
    HIPCHK(hipMemcpyAsync(&hinfo, L.info.p, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    ARGCHK(hinfo == 0, "m2_jk_df_local: the Coulomb metric is not positive definite");
    return M2_OK;
}

m2_status local_jk(m2_ctx* ctx, LocalDF& L, const double* dCp, const double* dPt, double* dJ, double* dK) {
    const int n = ctx->obs.nbf, naux = ctx->aux.nbf, nocc = L.nocc, nf = (int)L.frag.size();
    void* hv = nullptr;
    m2_status st = la_handle(ctx, &hv);
    if (st) return st;
    rocblas_handle h = (rocblas_handle)hv;
    const double one = 1.0, zero = 0.0;
    // exchange
    std::vector<int> fbf(n);
    for (int f = 0; f < nf; ++f)
        for (int x = L.frag[f].start; x < L.frag[f].start + L.frag[f].nbfs; ++x) fbf[x] = f;
    DevBuf<int> dfbf, dao_off, dfit_off;
    DevBuf<size_t> du_off, dkw_off;
    if ((st = upload(ctx, dfbf, fbf)) || (st = upload(ctx, dao_off, L.ao_off)) || (st = upload(ctx, dfit_off, L.fit_off)) ||
        (st = upload(ctx, du_off, L.u_off)) || (st = upload(ctx, dkw_off, L.kw_off)))
        return st;
    HIPCHK(L.Cm.reserve((size_t)n * nocc));
    hipLaunchKernelGGL(k_mask_mo, dim3(grid_for((size_t)n * nocc)), dim3(256), 0, ctx->stream, dCp, L.Cm.p, dfbf.p,
                       L.mo_mask.p, n, nocc, nf);
    HIPCHK(hipGetLastError());
    HIPCHK(L.X.reserve((size_t)nocc * n * naux));
    if (rocblas_dgemm(h, rocblas_operation_transpose, rocblas_operation_none, nocc, n * naux, n, &one, L.Cm.p, n, L.M.p,
                      n, &zero, L.X.p, nocc) != rocblas_status_success) {
        ctx->err = "m2_jk_df_local: GEMM X = C^T M failed";
        return M2_ERR_HIP;
    }
    HIPCHK(L.U.reserve(std::max<size_t>(L.u_off.back(), 1)));
    HIPCHK(L.Kw.reserve(std::max<size_t>(L.kw_off.back(), 1)));
    hipLaunchKernelGGL(k_gather_u, dim3(64, nocc), dim3(256), 0, ctx->stream, L.X.p, L.U.p, L.ao_idx.p, L.fit_idx.p,
                       dao_off.p, dfit_off.p, du_off.p, n, nocc);
    HIPCHK(hipGetLastError());
    for (int i = 0; i < nocc; ++i) {
        const int a = L.ao_off[i + 1] - L.ao_off[i], f = L.fit_off[i + 1] - L.fit_off[i];
        double* Ui = L.U.p + L.u_off[i];
        if (f == 0) {   // no fit functions: Y_i Y_i^T = 0
            hipLaunchKernelGGL(k_zero, dim3(grid_for((size_t)a * a)), dim3(256), 0, ctx->stream, L.Kw.p + L.kw_off[i],
                               (size_t)a * a);
            continue;
        }
        // Y_i = U_i L_i^-T  (Y L^T = U)
        if (rocblas_dtrsm(h, rocblas_side_right, rocblas_fill_lower, rocblas_operation_transpose,
                          rocblas_diagonal_non_unit, a, f, &one, L.Lfit.p + L.lfit_off[i], f, Ui, a) !=
                rocblas_status_success ||
            rocblas_dgemm(h, rocblas_operation_none, rocblas_operation_transpose, a, a, f, &one, Ui, a, Ui, a, &zero,
                          L.Kw.p + L.kw_off[i], a) != rocblas_status_success) {
            ctx->err = "m2_jk_df_local: exchange block of an LMO failed";
            return M2_ERR_HIP;
        }
    }
    hipLaunchKernelGGL(k_zero, dim3(grid_for((size_t)n * n)), dim3(256), 0, ctx->stream, dK, (size_t)n * n);
    hipLaunchKernelGGL(k_scatter_k, dim3(64, nocc), dim3(256), 0, ctx->stream, L.Kw.p, dK, L.ao_idx.p, dao_off.p,
                       dkw_off.p, n);
    HIPCHK(hipGetLastError());
    // Coulomb
    HIPCHK(L.Pk.reserve(naux));
    const double half = 0.5, four = 4.0;
    if (rocblas_dgemv(h, rocblas_operation_transpose, n * n, naux, &half, L.M.p, n * n, dPt, 1, &zero, L.Pk.p, 1) !=
            rocblas_status_success ||
        rocblas_dtrsv(h, rocblas_fill_lower, rocblas_operation_none, rocblas_diagonal_non_unit, naux, L.Lv.p, naux,
                      L.Pk.p, 1) != rocblas_status_success ||
        rocblas_dtrsv(h, rocblas_fill_lower, rocblas_operation_transpose, rocblas_diagonal_non_unit, naux, L.Lv.p,
                      naux, L.Pk.p, 1) != rocblas_status_success ||
        rocblas_dgemv(h, rocblas_operation_none, n * n, naux, &four, L.M.p, n * n, L.Pk.p, 1, &zero, dJ, 1) !=
            rocblas_status_success) {
        ctx->err = "m2_jk_df_local: Coulomb contraction failed";
        return M2_ERR_HIP;
    }
    HIPCHK(hipStreamSynchronize(ctx->stream));   // the index tables above are freed on return
    return M2_OK;
}

}  // namespace

extern "C" {

m2_status m2_df_local_reset(m2_ctx* ctx) {
    if (!ctx) return M2_ERR_ARG;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    local_df_release(ctx);
    return M2_OK;
}

m2_status m2_jk_df_local(m2_ctx* ctx, int nfrag, const m2_fragment* frags, double mo_thresh, double fit_thresh,
                         double r_thresh, double thrint, const double* Cocc, int nocc, const double* sigmainv,
                         const double* Pt, double* J, double* K) {
    if (!ctx) return M2_ERR_ARG;
    ARGCHK(ctx->obs.nbf > 0 && ctx->aux.nbf > 0, "m2_jk_df_local: orbital and auxiliary bases must be set");
    ARGCHK(nfrag > 0 && frags && Cocc && sigmainv && Pt && J && K && nocc > 0, "m2_jk_df_local: bad arguments");
    const int n = ctx->obs.nbf, naux = ctx->aux.nbf;
    // the fragments tile both bases in order and own the LMOs in order
    int bf = 0, sh = 0, ax = 0, ash = 0, occ = 0;
    for (int f = 0; f < nfrag; ++f) {
        const m2_fragment& g = frags[f];
        ARGCHK(g.start == bf && g.auxstart == ax && g.nbfs > 0 && g.naux > 0 && g.nshells > 0 && g.ndfshells > 0 &&
                   g.occ >= 0,
               "m2_jk_df_local: fragments must tile the orbital and auxiliary bases in order");
        int nb = 0, na = 0;
        for (int s = sh; s < sh + g.nshells && s < ctx->obs.nsh; ++s) nb += ctx->obs.h_nbf[s];
        for (int s = ash; s < ash + g.ndfshells && s < ctx->aux.nsh; ++s) na += ctx->aux.h_nbf[s];
        ARGCHK(nb == g.nbfs && na == g.naux, "m2_jk_df_local: fragment shell counts do not match its functions");
        bf += g.nbfs;
        ax += g.naux;
        sh += g.nshells;
        ash += g.ndfshells;
        occ += g.occ;
    }
    ARGCHK(bf == n && ax == naux && sh == ctx->obs.nsh && ash == ctx->aux.nsh,
           "m2_jk_df_local: fragments must cover both bases");
    ARGCHK(occ == nocc, "m2_jk_df_local: the fragments' occupied counts must add up to nocc");
    HIPCHK(hipSetDevice(ctx->device));
    if (ctx->ldf && (ctx->ldf->nocc != nocc || (int)ctx->ldf->frag.size() != nfrag)) local_df_release(ctx);
    const bool fresh = !ctx->ldf;
    if (fresh) {
        ctx->ldf = new LocalDF();
        LocalDF& L = *ctx->ldf;
        L.frag.assign(frags, frags + nfrag);
        L.mo_thresh = mo_thresh;
        L.fit_thresh = fit_thresh;
        L.r_thresh = r_thresh;
        L.thrint = thrint;
        L.nocc = nocc;
    }
    LocalDF& L = *ctx->ldf;
    HIPCHK(L.P.reserve((size_t)n * n));
    HIPCHK(L.J.reserve((size_t)n * n));
    HIPCHK(L.K.reserve((size_t)n * n));
    HIPCHK(L.Cp.reserve((size_t)n * nocc));
    DevBuf<double> dC, dsig, dsq;
    HIPCHK(dC.reserve((size_t)n * nocc));
    HIPCHK(dsig.reserve((size_t)nocc * nocc));
    HIPCHK(dsq.reserve((size_t)nocc * nocc));
    HIPCHK(hipMemcpyAsync(dC.p, Cocc, sizeof(double) * n * nocc, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(dsig.p, sigmainv, sizeof(double) * nocc * nocc, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(L.P.p, Pt, sizeof(double) * n * n, hipMemcpyHostToDevice, ctx->stream));
    // C' = Cocc sigmainv^(1/2)
    m2_status st = sym_sqrt(ctx, L, dsig.p, nocc, dsq.p);
    if (!st) st = gemm_device(ctx, 'N', 'N', n, nocc, nocc, 1.0, dC.p, n, dsq.p, nocc, 0.0, L.Cp.p, n);
    if (!st && fresh) st = local_setup(ctx, L, L.Cp.p, L.P.p);
    if (!st) st = local_jk(ctx, L, L.Cp.p, L.P.p, L.J.p, L.K.p);
    if (st) {
        if (fresh) local_df_release(ctx);
        return st;
    }
    HIPCHK(hipMemcpyAsync(J, L.J.p, sizeof(double) * n * n, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipMemcpyAsync(K, L.K.p, sizeof(double) * n * n, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return M2_OK;
}

m2_status m2_df_local_domains(m2_ctx* ctx, int which, int* centres, int* counts) {
    if (!ctx) return M2_ERR_ARG;
    ARGCHK(ctx->ldf, "m2_df_local_domains: no fragment-local DF set up");
    ARGCHK(which >= 0 && which <= 2 && centres && counts, "m2_df_local_domains: bad arguments");
    const LocalDF& L = *ctx->ldf;
    const int nf = (int)L.frag.size();
    const std::vector<LocalDomain>& D = which == 0 ? L.lmo : which == 1 ? L.ao : L.fit;
    for (int i = 0; i < L.nocc; ++i) {
        counts[i] = (int)D[i].centres.size();
        for (int c = 0; c < nf; ++c) centres[(size_t)i * nf + c] = c < counts[i] ? D[i].centres[c] : -1;
    }
    return M2_OK;
}

}  // extern "C"
