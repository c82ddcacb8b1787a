// Small device kernels: one-electron integrals, shell-block norms,
// symmetrisation of J/K, fills.
#include "m2_common.hpp"
#include "one_body.hpp"
#include "boys.hpp"

namespace m2 {



constexpr int OB_LMAX = 2;                 // orbital shells
constexpr int OB_NC = (OB_LMAX + 1) * (OB_LMAX + 2) / 2;
constexpr int OB_CUM = (OB_LMAX + 1) * (OB_LMAX + 2) * (OB_LMAX + 3) / 6;   // components with |a| <= LMAX

__device__ __forceinline__ int ob_ncart(int l) { return (l + 1) * (l + 2) / 2; }
__device__ __forceinline__ int ob_cum(int l) { return l * (l + 1) * (l + 2) / 6; }
__device__ __forceinline__ int ob_cidx(int x, int y, int z) {
    const int L = x + y + z;
    return (L - x) * (L - x + 1) / 2 + (L - x - y);
}
__device__ __forceinline__ void ob_comp(int l, int i, int* c) {
    int k = 0;
    for (int x = l; x >= 0; --x)
        for (int y = l - x; y >= 0; --y) {
            if (k == i) {
                c[0] = x;
                c[1] = y;
                c[2] = l - x - y;
                return;
            }
            ++k;
        }
}

// Boys F_0..F_M with runtime M (same scheme as the templated version)
__device__ void boys_rt(double T, int M, double* F, const BoysTables& tab) {
    if (T < BOYS_TMAX) {
        const int k = __double2int_rn(T * (1.0 / BOYS_STEP));
        const double d = k * BOYS_STEP - T;
        const double* r = tab.d + k * BOYS_MTAB + M;
        double s = r[7];
        for (int j = 6; j >= 0; --j) s = r[j] + d * s / (j + 1);
        F[M] = s;
        const double e = exp(-T);
        for (int m = M - 1; m >= 0; --m) F[m] = (2.0 * T * F[m + 1] + e) / (2 * m + 1);
    } else {
        F[0] = 0.88622692545275801365 * rsqrt(T);
        for (int m = 0; m < M; ++m) F[m + 1] = (2 * m + 1) * F[m] * (0.5 / T);
    }
}

// One thread per shell pair (s1 >= s2): S, T (1-D Obara-Saika overlaps) and
// V (two-centre Obara-Saika nuclear-attraction recurrence), Cartesian, then
// the real-solid-harmonic transform for l >= 2, written to both blocks.
// One wavefront per shell pair s1 >= s2: lane 0 takes the overlap and kinetic
// integrals, the lanes share the nuclei of the attraction integrals (c = lane,
// lane + 64, ...) and the wavefront sums them.
__global__ __launch_bounds__(64) void k_one_body(OneBodyArgs a) {
    const int t = blockIdx.x;
    const int lane = threadIdx.x;
    const int npairs = a.nsh * (a.nsh + 1) / 2;
    if (t >= npairs) return;
    int s1 = (int)((sqrt(8.0 * t + 1.0) - 1.0) * 0.5);
    while (s1 * (s1 + 1) / 2 > t) --s1;
    while ((s1 + 1) * (s1 + 2) / 2 <= t) ++s1;
    const int s2 = t - s1 * (s1 + 1) / 2;
    const int la = a.sh_l[s1], lb = a.sh_l[s2];
    const int na = ob_ncart(la), nb = ob_ncart(lb);
    const double* A = a.sh_O + 3 * s1;
    const double* B = a.sh_O + 3 * s2;
    double AB2 = 0.0;
    for (int k = 0; k < 3; ++k) AB2 += (A[k] - B[k]) * (A[k] - B[k]);
    double sS[OB_NC * OB_NC], sT[OB_NC * OB_NC], sV[OB_NC * OB_NC];
    for (int i = 0; i < na * nb; ++i) sS[i] = sT[i] = sV[i] = 0.0;
    const int Lab = la + lb;
    for (int i = 0; i < a.sh_np[s1]; ++i)
        for (int j = 0; j < a.sh_np[s2]; ++j) {
            const double al = a.exps[a.sh_poff[s1] + i], be = a.exps[a.sh_poff[s2] + j];
            const double p = al + be, ip2 = 0.5 / p;
            const double cc = a.coefs[a.sh_poff[s1] + i] * a.coefs[a.sh_poff[s2] + j];
            const double Kab = exp(-al * be / p * AB2);
            double P[3], PA[3], PB[3];
            for (int k = 0; k < 3; ++k) {
                P[k] = (al * A[k] + be * B[k]) / p;
                PA[k] = P[k] - A[k];
                PB[k] = P[k] - B[k];
            }
            if (lane == 0) {
            // 1-D overlaps, i <= la, j <= lb + 2
            double O1[3][OB_LMAX + 1][OB_LMAX + 3];
            for (int k = 0; k < 3; ++k) {
                for (int ii = 0; ii <= la; ++ii)
                    for (int jj = 0; jj <= lb + 2; ++jj) {
                        double v;
                        if (ii == 0 && jj == 0) v = 1.0;
                        else if (ii > 0) {
                            v = PA[k] * O1[k][ii - 1][jj];
                            if (ii > 1) v += (ii - 1) * ip2 * O1[k][ii - 2][jj];
                            if (jj > 0) v += jj * ip2 * O1[k][ii - 1][jj - 1];
                        } else {
                            v = PB[k] * O1[k][0][jj - 1];
                            if (jj > 1) v += (jj - 1) * ip2 * O1[k][0][jj - 2];
                        }
                        O1[k][ii][jj] = v;
                    }
            }
            const double pref = pow(M_PI / p, 1.5) * Kab * cc;
            for (int ia = 0; ia < na; ++ia) {
                int ea[3];
                ob_comp(la, ia, ea);
                for (int ib = 0; ib < nb; ++ib) {
                    int eb[3];
                    ob_comp(lb, ib, eb);
                    const double sx = O1[0][ea[0]][eb[0]], sy = O1[1][ea[1]][eb[1]], sz = O1[2][ea[2]][eb[2]];
                    double tk[3];
                    for (int k = 0; k < 3; ++k) {
                        const int ii = ea[k], jj = eb[k];
                        double v = -2.0 * be * (2 * jj + 1) * O1[k][ii][jj] + 4.0 * be * be * O1[k][ii][jj + 2];
                        if (jj >= 2) v += jj * (jj - 1) * O1[k][ii][jj - 2];
                        tk[k] = -0.5 * v;
                    }
                    sS[ia * nb + ib] += pref * sx * sy * sz;
                    sT[ia * nb + ib] += pref * (tk[0] * sy * sz + sx * tk[1] * sz + sx * sy * tk[2]);
                }
            }
            }
            // nuclear attraction: [a|b]^(m), |a| <= la, |b| <= lb
            const double vpref = 2.0 * M_PI / p * Kab * cc;
            for (int c = lane; c < (a.want_v ? a.natom : 0); c += 64) {
                const double* C = a.xyz + 3 * c;
                double PC[3], PC2 = 0.0;
                for (int k = 0; k < 3; ++k) {
                    PC[k] = P[k] - C[k];
                    PC2 += PC[k] * PC[k];
                }
                double F[2 * OB_LMAX + 1];
                boys_rt(p * PC2, Lab, F, a.tab);
                double W[OB_CUM][OB_CUM][2 * OB_LMAX + 1];
                for (int m = 0; m <= Lab; ++m) W[0][0][m] = vpref * F[m];
                for (int lbb = 0; lbb <= lb; ++lbb)
                    for (int ib = ob_cum(lbb); ib < ob_cum(lbb + 1); ++ib) {
                        int eb[3];
                        ob_comp(lbb, ib - ob_cum(lbb), eb);
                        for (int laa = 0; laa <= la; ++laa)
                            for (int ia = ob_cum(laa); ia < ob_cum(laa + 1); ++ia) {
                                if (ia == 0 && ib == 0) continue;
                                int ea[3];
                                ob_comp(laa, ia - ob_cum(laa), ea);
                                const int mmax = Lab - laa - lbb;
                                if (lbb > 0) {
                                    int ax = eb[0] > 0 ? 0 : (eb[1] > 0 ? 1 : 2);
                                    int bm[3] = {eb[0], eb[1], eb[2]};
                                    bm[ax]--;
                                    const int b1 = ob_cum(lbb - 1) + ob_cidx(bm[0], bm[1], bm[2]);
                                    int b2 = -1, a1 = -1;
                                    if (bm[ax] > 0) {
                                        int t2[3] = {bm[0], bm[1], bm[2]};
                                        t2[ax]--;
                                        b2 = ob_cum(lbb - 2) + ob_cidx(t2[0], t2[1], t2[2]);
                                    }
                                    if (ea[ax] > 0) {
                                        int t2[3] = {ea[0], ea[1], ea[2]};
                                        t2[ax]--;
                                        a1 = ob_cum(laa - 1) + ob_cidx(t2[0], t2[1], t2[2]);
                                    }
                                    for (int m = 0; m <= mmax; ++m) {
                                        double v = PB[ax] * W[ia][b1][m] - PC[ax] * W[ia][b1][m + 1];
                                        if (b2 >= 0) v += bm[ax] * ip2 * (W[ia][b2][m] - W[ia][b2][m + 1]);
                                        if (a1 >= 0) v += ea[ax] * ip2 * (W[a1][b1][m] - W[a1][b1][m + 1]);
                                        W[ia][ib][m] = v;
                                    }
                                } else {
                                    int ax = ea[0] > 0 ? 0 : (ea[1] > 0 ? 1 : 2);
                                    int am[3] = {ea[0], ea[1], ea[2]};
                                    am[ax]--;
                                    const int a1 = ob_cum(laa - 1) + ob_cidx(am[0], am[1], am[2]);
                                    int a2 = -1;
                                    if (am[ax] > 0) {
                                        int t2[3] = {am[0], am[1], am[2]};
                                        t2[ax]--;
                                        a2 = ob_cum(laa - 2) + ob_cidx(t2[0], t2[1], t2[2]);
                                    }
                                    for (int m = 0; m <= mmax; ++m) {
                                        double v = PA[ax] * W[a1][0][m] - PC[ax] * W[a1][0][m + 1];
                                        if (a2 >= 0) v += am[ax] * ip2 * (W[a2][0][m] - W[a2][0][m + 1]);
                                        W[ia][0][m] = v;
                                    }
                                }
                            }
                    }
                for (int ia = 0; ia < na; ++ia)
                    for (int ib = 0; ib < nb; ++ib)
                        sV[ia * nb + ib] -= a.Z[c] * W[ob_cum(la) + ia][ob_cum(lb) + ib][0];
            }
        }
    if (a.want_v)
        for (int i = 0; i < na * nb; ++i) {
            double v = sV[i];
            for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
            sV[i] = v;
        }
    if (lane != 0) return;
    // transform and store
    const int fa = a.pure && la >= 2 ? 2 * la + 1 : na;
    const int fb = a.pure && lb >= 2 ? 2 * lb + 1 : nb;
    const int offa = la == 0 ? 0 : (la == 1 ? 1 : 10);
    const int offb = lb == 0 ? 0 : (lb == 1 ? 1 : 10);
    double* outs[3] = {a.S, a.T, a.V};
    double* ins[3] = {sS, sT, sV};
    for (int w = 0; w < 3; ++w) {
        const double* in = ins[w];
        for (int i = 0; i < fa; ++i)
            for (int j = 0; j < fb; ++j) {
                double s = 0.0;
                for (int ia = 0; ia < na; ++ia) {
                    const double ca = (fa == na) ? (ia == i ? 1.0 : 0.0) : a.sph[offa + i * na + ia];
                    if (ca == 0.0) continue;
                    for (int ib = 0; ib < nb; ++ib) {
                        const double cb = (fb == nb) ? (ib == j ? 1.0 : 0.0) : a.sph[offb + j * nb + ib];
                        if (cb == 0.0) continue;
                        s += ca * cb * in[ia * nb + ib];
                    }
                }
                const int I = a.sh_bf[s1] + i, Jj = a.sh_bf[s2] + j;
                outs[w][(size_t)I * a.n + Jj] = s;
                outs[w][(size_t)Jj * a.n + I] = s;
            }
    }
}

void launch_one_body(const OneBodyArgs& a, hipStream_t st) {
    const int npairs = a.nsh * (a.nsh + 1) / 2;
    hipLaunchKernelGGL(k_one_body, dim3(npairs), dim3(64), 0, st, a);
}

// shell-block inf-norms: one thread per (s1, s2)
__global__ void k_blocknorm(const double* __restrict__ A, int n, int nsh, const int* __restrict__ bf,
                            const int* __restrict__ nbf, double* __restrict__ out) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nsh * nsh) return;
    const int s1 = t / nsh, s2 = t % nsh;
    double mx = 0.0;
    for (int i = 0; i < nbf[s1]; ++i)
        for (int j = 0; j < nbf[s2]; ++j) mx = fmax(mx, fabs(A[(size_t)(bf[s1] + i) * n + bf[s2] + j]));
    out[t] = mx;
}

void launch_blocknorm(const double* A, int n, int nsh, const int* sh_bf, const int* sh_nbf, double* out, hipStream_t st) {
    const int nt = nsh * nsh;
    hipLaunchKernelGGL(k_blocknorm, dim3((nt + 255) / 256), dim3(256), 0, st, A, n, nsh, sh_bf, sh_nbf, out);
}

// X_out = (X + X^T) / 4 (src/fockbuild.cpp:161-162)
__global__ void k_symmetrize(const double* __restrict__ J, const double* __restrict__ K, double* __restrict__ Jo,
                             double* __restrict__ Ko, int n) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * n) return;
    const int i = t / n, j = t % n;
    Jo[t] = 0.25 * (J[t] + J[(size_t)j * n + i]);
    Ko[t] = 0.25 * (K[t] + K[(size_t)j * n + i]);
}

void launch_symmetrize(double* J, double* K, double* Jo, double* Ko, int n, hipStream_t st) {
    const int nt = n * n;
    hipLaunchKernelGGL(k_symmetrize, dim3((nt + 255) / 256), dim3(256), 0, st, J, K, Jo, Ko, n);
}

__global__ void k_zero(double* __restrict__ p, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = 0.0;
}

void launch_zero(double* p, size_t n, hipStream_t st) {
    const int blocks = (int)std::min<size_t>((n + 255) / 256, 4096);
    hipLaunchKernelGGL(k_zero, dim3(blocks > 0 ? blocks : 1), dim3(256), 0, st, p, n);
}

}  // namespace m2
