// RHF nuclear gradient on the device (the reference's Fock::compute_forces,
// src/rhfgrad.cpp:15-89, with compute_2body_fock_deriv<1>, :418-598).
//
// Two-electron part, adjoint form: for every unique shell quartet that the
// reference's gradient loop visits (pairs in the overlap shell-pair list,
// src/integrals.cpp:515-566; Schwarz x density screen against precision) the
// derivative integrals d/dA, d/dB, d/dC (ab|cd) are recomputed and contracted
// at once with
//   Gamma_abcd = deg (1/2 D_ab D_cd - 1/8 (D_ac D_bd + D_ad D_bc)),
// which equals 1/4 G1:D of src/rhfgrad.cpp:49-57 per quartet; d/dD follows
// from translational invariance.  Nothing is stored: small classes run one
// quartet per lane from generated register code, the others one quartet per
// wavefront from the generated LDS program.  The contraction is done in the
// Cartesian basis with the density blocks transformed to Cartesian.
//
// One-electron part: one wavefront per ordered shell pair (s1, s2), one lane
// per nucleus, adds twice d/dA of D:(T+V) - 2 W:S (A = centre of s1; the
// (s2, s1) wavefront supplies the other centre) and the Hellmann-Feynman term
// of each nucleus (dC = -(dA + dB)).
#include <mutex>
#include "grad_reg.hpp"
#include "generated/deriv_programs.inc"

namespace m2 {

// ------------------------------------------------------- program classes
struct DerivProg {
    int L, nslots, base, nacc;
    const int* lev;        // per VRR level: ins0, ins1, acc0, acc1
    int nvlev;
    const int* plev;       // c0, c1, then per post level: ins0, ins1, c0, c1
    int nplev;
    const int* ins;
    const double* coef;
    const int* acc;        // (acc slot, vrr slot, scalar id)
    const int* contr;      // (slot, gamma index, g index)
    const double* ccoef;
    int la, lb, lc, ld;
    const int* vw;         // VRR phase in pass order (DVRR_WORDS of the class)
    const int* vkind;      // per pass: 0 instructions, 1 accumulation, +4 barrier after
    int nvpass;
};

// The VRR phase of one primitive quartet from its pass-ordered words, the
// next pass's words loaded while the current one computes.
template <typename R>
__device__ __forceinline__ void vrr_phase_passes(const DerivProg& P, R* __restrict__ X, const R* __restrict__ S) {
    constexpr int NW = 7;
    const int lane = threadIdx.x;
    const int* wp = P.vw + lane;
    int cur[NW], nxt[NW];
#pragma unroll
    for (int k = 0; k < NW; ++k) cur[k] = wp[k * 64];
    for (int p = 0; p < P.nvpass; ++p) {
        if (p + 1 < P.nvpass) {
#pragma unroll
            for (int k = 0; k < NW; ++k) nxt[k] = wp[((p + 1) * NW + k) * 64];
        }
        const int kind = P.vkind[p];
        const int h = cur[0];
        if (h >= 0) {
            if ((kind & 3) == 0) {
                const int nt = h >> 16;
                R v = R(0);
#pragma unroll
                for (int t = 0; t < NW - 1; ++t)
                    if (t < nt) {
                        const int tw = cur[1 + t];
                        const int c6 = (tw >> 19) & 63;
                        v += R(c6 >= 32 ? c6 - 64 : c6) * S[(tw >> 14) & 31] * X[tw & 16383];
                    }
                X[h & 65535] = v;
            } else {
                X[h & 65535] += S[h >> 16] * X[cur[1]];
            }
        }
        if (kind & 4) __syncthreads();
#pragma unroll
        for (int k = 0; k < NW; ++k) cur[k] = nxt[k];
    }
}

template <typename R>
__device__ __forceinline__ void prog_level(const DerivProg& P, int i0, int i1, R* X, const R* S) {
    for (int j = i0 + (int)threadIdx.x; j < i1; j += 64) {
        const int* in = P.ins + 14 * j;
        const double* cf = P.coef + 6 * j;
        const int nt = in[1];
        R v = R(0);
        for (int t = 0; t < nt; ++t) v += R(cf[t]) * S[in[8 + t]] * X[in[2 + t]];
        X[in[0]] = v;
    }
}

template <typename R>
__device__ __forceinline__ void prog_contract(const DerivProg& P, int c0, int c1, const R* X, const R* Gam,
                                              double* g) {
    for (int j = c0 + (int)threadIdx.x; j < c1; j += 64) {
        const int* c = P.contr + 3 * j;
        g[c[2]] += double(R(P.ccoef[j]) * Gam[c[1]] * X[c[0]]);
    }
}

// R: type of the LDS slots and of the VRR / HRR arithmetic (float for the
// quartets a mixed-precision gradient leaves below its FP32 threshold); the
// Boys values, scalars and density blocks are formed in double.
template <typename R>
__global__ __launch_bounds__(64) void k_grad_prog(GradArgs a, BoysTables tab, const Task* __restrict__ tasks, int ntasks,
                                                  DerivProg P) {
    extern __shared__ __align__(16) unsigned char smem[];
    const int NA = ncart_d(P.la), NB = ncart_d(P.lb), NC = ncart_d(P.lc), ND = ncart_d(P.ld);
    const int ngam = NA * NB * NC * ND;
    double* Dt = reinterpret_cast<double*>(smem);   // 6 Cartesian density blocks, 36 each
    R* X = reinterpret_cast<R*>(Dt + 6 * 36);
    R* S = X + P.nslots;
    R* Gam = S + NSCALAR_D;
    // GRAD_PROG_SPLIT workgroups per task, each walking GRAD_PROG_KETS of its kets
    const int w = blockIdx.x / GRAD_PROG_SPLIT, part = blockIdx.x - w * GRAD_PROG_SPLIT;
    if (w >= ntasks) return;
    const int lane = threadIdx.x;
    const Task tk = tasks[w];
    const int2 sab = a.pair_sh[tk.bra];
    const int sa = sab.x, sb = sab.y;
    const double* O = a.sh_O;
    const double ABx = O[3 * sa] - O[3 * sb], ABy = O[3 * sa + 1] - O[3 * sb + 1], ABz = O[3 * sa + 2] - O[3 * sb + 2];
    const int kk1 = min(tk.nket, (part + 1) * GRAD_PROG_KETS);
    for (int kk = part * GRAD_PROG_KETS; kk < kk1; ++kk) {
        const int ket = tk.ket0 + kk;
        const int2 scd = a.pair_sh[ket];
        const int sc = scd.x, sd = scd.y;
        if (!grad_screen(a, tk.bra, ket, sa, sb, sc, sd)) continue;
        const double CDx = O[3 * sc] - O[3 * sd], CDy = O[3 * sc + 1] - O[3 * sd + 1], CDz = O[3 * sc + 2] - O[3 * sd + 2];
        const int sh[4] = {sa, sb, sc, sd};
        const int ls[4] = {P.la, P.lb, P.lc, P.ld};
        // Cartesian density blocks ab, cd, ac, bd, ad, bc (6 x 36 entries)
        const int bx[6] = {0, 2, 0, 1, 0, 1}, by[6] = {1, 3, 2, 3, 3, 2};
        for (int t = lane; t < 6 * 36; t += 64) {
            const int blk = t / 36, e = t - blk * 36;
            const int x = bx[blk], y = by[blk];
            const int ny = ncart_d(ls[y]);
            const int i = e / ny, j = e - i * ny;
            if (i < ncart_d(ls[x])) Dt[t] = a.Dc[(size_t)(a.sh_cf[sh[x]] + i) * a.nc + a.sh_cf[sh[y]] + j];
        }
        __syncthreads();
        const double deg = (sa == sb ? 1.0 : 2.0) * (sc == sd ? 1.0 : 2.0) * (tk.bra == ket ? 1.0 : 2.0);
        for (int t = lane; t < ngam; t += 64) {
            const int l = t % ND, k = (t / ND) % NC, j = (t / (ND * NC)) % NB, i = t / (ND * NC * NB);
            Gam[t] = R(deg * (0.5 * Dt[0 * 36 + i * NB + j] * Dt[1 * 36 + k * ND + l] -
                              0.125 * (Dt[2 * 36 + i * NC + k] * Dt[3 * 36 + j * ND + l] +
                                       Dt[4 * 36 + i * ND + l] * Dt[5 * 36 + j * NC + k])));
        }
        for (int t = lane; t < P.nacc; t += 64) X[t] = R(0);
        const PrimPair* bp = a.prims + a.pair_off[tk.bra];
        const PrimPair* kp = a.prims + a.pair_off[ket];
        const int nbp = a.pair_np[tk.bra], nkp = a.pair_np[ket];
        // primitive screening (uniform over the workgroup: one quartet)
        const float* bq = a.prim_q + a.pair_off[tk.bra];
        const float* kq = a.prim_q + a.pair_off[ket];
        const float cut = float(a.prim_tau / grad_dn(a, sa, sb, sc, sd));
        for (int ib = 0; ib < nbp; ++ib) {
            if (bq[ib] * kq[0] < cut) break;
            const PrimPair B = bp[ib];
            for (int ik = 0; ik < nkp; ++ik) {
                if (bq[ib] * kq[ik] < cut) break;
                const PrimPair K = kp[ik];
                const double p = B.p, q = K.p;
                const double dx = B.P[0] - K.P[0], dy = B.P[1] - K.P[1], dz = B.P[2] - K.P[2];
                const double rs = rsqrt(p + q);
                const double ipq = rs * rs;
                const double T = p * q * ipq * (dx * dx + dy * dy + dz * dz);
                __syncthreads();
                if (lane <= P.L) {
                    double F;
                    if (T < BOYS_TMAX) {
                        const int k = __double2int_rn(T * (1.0 / BOYS_STEP));
                        const double d = k * BOYS_STEP - T;
                        const double* r = tab.d + k * BOYS_MTAB + lane;
                        double s = r[6] + d * (1.0 / 7.0) * r[7];
                        s = r[5] + d * (1.0 / 6.0) * s;
                        s = r[4] + d * 0.2 * s;
                        s = r[3] + d * 0.25 * s;
                        s = r[2] + d * (1.0 / 3.0) * s;
                        s = r[1] + d * 0.5 * s;
                        F = r[0] + d * s;
                    } else {
                        F = 0.88622692545275801365 * rsqrt(T);
                        const double h = 0.5 / T;
                        for (int i = 0; i < lane; ++i) F = (2 * i + 1) * F * h;
                    }
                    X[P.base + lane] = R(B.K * K.K * rs * F);
                }
                if (lane == 0) {
                    const double rp = q * ipq, rq = p * ipq;
                    const double sv[NSCALAR_D] = {1.0, B.PA[0], B.PA[1], B.PA[2], -rp * dx, -rp * dy, -rp * dz,
                                                  K.PA[0], K.PA[1], K.PA[2], rq * dx, rq * dy, rq * dz,
                                                  B.inv2p, K.inv2p, 0.5 * ipq, B.inv2p * rp, K.inv2p * rq,
                                                  ABx, ABy, ABz, CDx, CDy, CDz,
                                                  2.0 * B.a, 2.0 * (B.p - B.a), 2.0 * K.a};
#pragma unroll
                    for (int q = 0; q < NSCALAR_D; ++q) S[q] = R(sv[q]);
                }
                __syncthreads();
                if (P.vw) {
                    vrr_phase_passes(P, X, S);
                } else {
                    for (int l = 0; l < P.nvlev; ++l) {
                        const int* L4 = P.lev + 4 * l;
                        prog_level(P, L4[0], L4[1], X, S);
                        __syncthreads();
                        for (int j = L4[2] + lane; j < L4[3]; j += 64) {
                            const int* ac = P.acc + 3 * j;
                            X[ac[0]] += S[ac[2]] * X[ac[1]];
                        }
                        __syncthreads();
                    }
                }
            }
        }
        __syncthreads();
        double g[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        prog_contract(P, P.plev[0], P.plev[1], X, Gam, g);
        for (int l = 0; l < P.nplev; ++l) {
            const int* L4 = P.plev + 2 + 4 * l;
            prog_level(P, L4[0], L4[1], X, S);
            __syncthreads();
            prog_contract(P, L4[2], L4[3], X, Gam, g);
            __syncthreads();
        }
        double tot[9];
#pragma unroll
        for (int i = 0; i < 9; ++i) tot[i] = wave_sum_g(g[i]);
        if (lane == 0) {
            const int at[4] = {a.sh_atom[sa], a.sh_atom[sb], a.sh_atom[sc], a.sh_atom[sd]};
            for (int i = 0; i < 3; ++i) {
                atomicAdd(&a.grad[at[0] * 3 + i], tot[i]);
                atomicAdd(&a.grad[at[1] * 3 + i], tot[3 + i]);
                atomicAdd(&a.grad[at[2] * 3 + i], tot[6 + i]);
                atomicAdd(&a.grad[at[3] * 3 + i], -(tot[i] + tot[3 + i] + tot[6 + i]));
            }
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------ one-electron
struct Grad1Args {
    const double* sh_O;
    const int* sh_l;
    const int* sh_bf;
    const int* sh_np;
    const int* sh_poff;
    const int* sh_atom;
    const double* exps;
    const double* coefs;
    const double* D;
    const double* W;
    const double* sph2;
    const double* Z;
    const double* xyz;
    int nsh, n, pure, natom;
    BoysTables tab;
    double* grad;
};

__device__ __forceinline__ void comp_of(int l, int i, int* c) {
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
__device__ __forceinline__ int cum_of(int l) { return l * (l + 1) * (l + 2) / 6; }
__device__ __forceinline__ int cidx_of(int x, int y, int z) {
    const int L = x + y + z;
    return (L - x) * (L - x + 1) / 2 + (L - x - y);
}

// Cartesian block of D (or W): rows components of s1 (l1), cols of s2 (l2)
__device__ void cart_block_rt(const double* M, int n, const double* sph2, int pure, int l1, int f1, int l2, int f2,
                              double* out) {
    const int n1 = ncart_d(l1), n2 = ncart_d(l2);
    const bool p1 = pure && l1 >= 2, p2 = pure && l2 >= 2;
    const int m1 = p1 ? 2 * l1 + 1 : n1, m2 = p2 ? 2 * l2 + 1 : n2;
    for (int i = 0; i < n1; ++i)
        for (int j = 0; j < n2; ++j) {
            double s = 0.0;
            for (int u = 0; u < m1; ++u) {
                const double cu = p1 ? sph2[u * 6 + i] : (u == i ? 1.0 : 0.0);
                if (cu == 0.0) continue;
                for (int v = 0; v < m2; ++v) {
                    const double cv = p2 ? sph2[v * 6 + j] : (v == j ? 1.0 : 0.0);
                    if (cv == 0.0) continue;
                    s += cu * cv * M[(size_t)(f1 + u) * n + f2 + v];
                }
            }
            out[i * n2 + j] = s;
        }
}

// One wavefront per ordered shell pair (s1, s2) of angular momenta (LA, LB)
// (a list per class, so that every array below has its exact compile-time
// size); lane c takes the nuclei c, c + 64, ...  Lane 0 also does the kinetic
// / Pulay part.
template <int LA, int LB>
__global__ __launch_bounds__(64) void k_grad_1e(Grad1Args a, const int2* __restrict__ pairs, int npairs) {
    constexpr int la = LA, lb = LB;
    constexpr int G1_CUMA = (la + 2) * (la + 3) * (la + 4) / 6;   // components with |a| <= la + 1
    constexpr int G1_CUMB = (lb + 1) * (lb + 2) * (lb + 3) / 6;   // components with |b| <= lb
    const int t = blockIdx.x;
    if (t >= npairs) return;
    const int s1 = pairs[t].x, s2 = pairs[t].y;
    const int lane = threadIdx.x;
    const int na = ncart_d(la), nb = ncart_d(lb);
    double Dt[36], Wt[36];
    cart_block_rt(a.D, a.n, a.sph2, a.pure, la, a.sh_bf[s1], lb, a.sh_bf[s2], Dt);
    cart_block_rt(a.W, a.n, a.sph2, a.pure, la, a.sh_bf[s1], lb, a.sh_bf[s2], Wt);
    const double* A = a.sh_O + 3 * s1;
    const double* B = a.sh_O + 3 * s2;
    double AB2 = 0.0;
    for (int k = 0; k < 3; ++k) AB2 += (A[k] - B[k]) * (A[k] - B[k]);
    double gA[3] = {0, 0, 0};   // T, Pulay and V (basis-centre) part on atom(s1)
    double gC[3] = {0, 0, 0};   // Hellmann-Feynman part on the lane's nucleus (natom <= 64)
    const int Lab = la + 1 + lb;
    for (int ip = 0; ip < a.sh_np[s1]; ++ip)
        for (int jp = 0; jp < a.sh_np[s2]; ++jp) {
            const double al = a.exps[a.sh_poff[s1] + ip], be = a.exps[a.sh_poff[s2] + jp];
            const double p = al + be, ip2 = 0.5 / p;
            const double cc = a.coefs[a.sh_poff[s1] + ip] * a.coefs[a.sh_poff[s2] + jp];
            const double Kab = exp(-al * be / p * AB2);
            double P[3], PA[3], PB[3];
            for (int k = 0; k < 3; ++k) {
                P[k] = (al * A[k] + be * B[k]) / p;
                PA[k] = P[k] - A[k];
                PB[k] = P[k] - B[k];
            }
            double O1[3][la + 2][lb + 3];
            for (int k = 0; k < 3; ++k)
                for (int ii = 0; ii <= la + 1; ++ii)
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
            const double spref = pow(M_PI / p, 1.5) * Kab * cc;
            // S and T of a Cartesian pair (exponents ex of a, eb of b)
            auto ST = [&](const int* ex, const int* eb, double& s, double& tk) {
                const double sx = O1[0][ex[0]][eb[0]], sy = O1[1][ex[1]][eb[1]], sz = O1[2][ex[2]][eb[2]];
                double t3[3];
                for (int k = 0; k < 3; ++k) {
                    const int ii = ex[k], jj = eb[k];
                    double v = -2.0 * be * (2 * jj + 1) * O1[k][ii][jj] + 4.0 * be * be * O1[k][ii][jj + 2];
                    if (jj >= 2) v += jj * (jj - 1) * O1[k][ii][jj - 2];
                    t3[k] = -0.5 * v;
                }
                s = spref * sx * sy * sz;
                tk = spref * (t3[0] * sy * sz + sx * t3[1] * sz + sx * sy * t3[2]);
            };
            for (int ia = 0; ia < (lane == 0 ? na : 0); ++ia) {
                int ea[3];
                comp_of(la, ia, ea);
                for (int ib = 0; ib < nb; ++ib) {
                    int eb[3];
                    comp_of(lb, ib, eb);
                    const double dd = Dt[ia * nb + ib], ww = Wt[ia * nb + ib];
                    for (int ax = 0; ax < 3; ++ax) {
                        int ep[3] = {ea[0], ea[1], ea[2]};
                        ep[ax]++;
                        double sp, tp;
                        ST(ep, eb, sp, tp);
                        double ds = 2.0 * al * sp, dt = 2.0 * al * tp;
                        if (ea[ax] > 0) {
                            int em[3] = {ea[0], ea[1], ea[2]};
                            em[ax]--;
                            double sm, tm;
                            ST(em, eb, sm, tm);
                            ds -= ea[ax] * sm;
                            dt -= ea[ax] * tm;
                        }
                        gA[ax] += dd * dt - 2.0 * ww * ds;
                    }
                }
            }
            // nuclear attraction: [a|b]^(m) with |a| <= la+1, |b| <= lb
            const double vpref = 2.0 * M_PI / p * Kab * cc;
            for (int c = lane; c < a.natom; c += 64) {
                const double* Cc = a.xyz + 3 * c;
                double PC[3], PC2 = 0.0;
                for (int k = 0; k < 3; ++k) {
                    PC[k] = P[k] - Cc[k];
                    PC2 += PC[k] * PC[k];
                }
                double F[la + lb + 2];
                {
                    const double T = p * PC2;
                    if (T < BOYS_TMAX) {
                        const int k = __double2int_rn(T * (1.0 / BOYS_STEP));
                        const double d = k * BOYS_STEP - T;
                        const double* r = a.tab.d + k * BOYS_MTAB + Lab;
                        double s = r[6] + d * (1.0 / 7.0) * r[7];
                        s = r[5] + d * (1.0 / 6.0) * s;
                        s = r[4] + d * 0.2 * s;
                        s = r[3] + d * 0.25 * s;
                        s = r[2] + d * (1.0 / 3.0) * s;
                        s = r[1] + d * 0.5 * s;
                        F[Lab] = r[0] + d * s;
                        const double e = exp(-T);
                        constexpr double inv_odd[6] = {1.0, 1.0 / 3, 1.0 / 5, 1.0 / 7, 1.0 / 9, 1.0 / 11};
                        for (int m = Lab - 1; m >= 0; --m) F[m] = (2.0 * T * F[m + 1] + e) * inv_odd[m];
                    } else {
                        F[0] = 0.88622692545275801365 * rsqrt(T);
                        for (int m = 0; m < Lab; ++m) F[m + 1] = (2 * m + 1) * F[m] * (0.5 / T);
                    }
                }
                double Wv[G1_CUMA][G1_CUMB][la + lb + 2];
                for (int m = 0; m <= Lab; ++m) Wv[0][0][m] = vpref * F[m];
#pragma unroll
                for (int lbb = 0; lbb <= lb; ++lbb)
#pragma unroll
                    for (int ib = cum_of(lbb); ib < cum_of(lbb + 1); ++ib) {
                        int eb[3];
                        comp_of(lbb, ib - cum_of(lbb), eb);
#pragma unroll
                        for (int laa = 0; laa <= la + 1; ++laa)
#pragma unroll
                            for (int ia = cum_of(laa); ia < cum_of(laa + 1); ++ia) {
                                if (ia == 0 && ib == 0) continue;
                                int ea[3];
                                comp_of(laa, ia - cum_of(laa), ea);
                                const int mmax = Lab - laa - lbb;
                                if (lbb > 0) {
                                    const int ax = eb[0] > 0 ? 0 : (eb[1] > 0 ? 1 : 2);
                                    int bm[3] = {eb[0], eb[1], eb[2]};
                                    bm[ax]--;
                                    const int b1 = cum_of(lbb - 1) + cidx_of(bm[0], bm[1], bm[2]);
                                    int b2 = -1, a1 = -1;
                                    if (bm[ax] > 0) {
                                        int t2[3] = {bm[0], bm[1], bm[2]};
                                        t2[ax]--;
                                        b2 = cum_of(lbb - 2) + cidx_of(t2[0], t2[1], t2[2]);
                                    }
                                    if (ea[ax] > 0) {
                                        int t2[3] = {ea[0], ea[1], ea[2]};
                                        t2[ax]--;
                                        a1 = cum_of(laa - 1) + cidx_of(t2[0], t2[1], t2[2]);
                                    }
                                    for (int m = 0; m <= mmax; ++m) {
                                        double v = PB[ax] * Wv[ia][b1][m] - PC[ax] * Wv[ia][b1][m + 1];
                                        if (b2 >= 0) v += bm[ax] * ip2 * (Wv[ia][b2][m] - Wv[ia][b2][m + 1]);
                                        if (a1 >= 0) v += ea[ax] * ip2 * (Wv[a1][b1][m] - Wv[a1][b1][m + 1]);
                                        Wv[ia][ib][m] = v;
                                    }
                                } else {
                                    const int ax = ea[0] > 0 ? 0 : (ea[1] > 0 ? 1 : 2);
                                    int am[3] = {ea[0], ea[1], ea[2]};
                                    am[ax]--;
                                    const int a1 = cum_of(laa - 1) + cidx_of(am[0], am[1], am[2]);
                                    int a2 = -1;
                                    if (am[ax] > 0) {
                                        int t2[3] = {am[0], am[1], am[2]};
                                        t2[ax]--;
                                        a2 = cum_of(laa - 2) + cidx_of(t2[0], t2[1], t2[2]);
                                    }
                                    for (int m = 0; m <= mmax; ++m) {
                                        double v = PA[ax] * Wv[a1][0][m] - PC[ax] * Wv[a1][0][m + 1];
                                        if (a2 >= 0) v += am[ax] * ip2 * (Wv[a2][0][m] - Wv[a2][0][m + 1]);
                                        Wv[ia][0][m] = v;
                                    }
                                }
                            }
                    }
                double gv[3] = {0, 0, 0};
                for (int ia = 0; ia < na; ++ia) {
                    int ea[3];
                    comp_of(la, ia, ea);
                    for (int ib = 0; ib < nb; ++ib) {
                        const double dd = Dt[ia * nb + ib];
                        const int bpos = cum_of(lb) + ib;
                        for (int ax = 0; ax < 3; ++ax) {
                            int ep[3] = {ea[0], ea[1], ea[2]};
                            ep[ax]++;
                            double dv = 2.0 * al * Wv[cum_of(la + 1) + cidx_of(ep[0], ep[1], ep[2])][bpos][0];
                            if (ea[ax] > 0) {
                                int em[3] = {ea[0], ea[1], ea[2]};
                                em[ax]--;
                                dv -= ea[ax] * Wv[cum_of(la - 1) + cidx_of(em[0], em[1], em[2])][bpos][0];
                            }
                            gv[ax] += dd * dv;
                        }
                    }
                }
                for (int ax = 0; ax < 3; ++ax) {
                    gA[ax] -= a.Z[c] * gv[ax];
                    // -(dA) of the point-charge term, on nucleus c
                    if (a.natom <= 64) gC[ax] += 2.0 * a.Z[c] * gv[ax];
                    else atomicAdd(&a.grad[c * 3 + ax], 2.0 * a.Z[c] * gv[ax]);
                }
            }
        }
    if (a.natom <= 64 && lane < a.natom)
        for (int ax = 0; ax < 3; ++ax)
            if (gC[ax] != 0.0) atomicAdd(&a.grad[lane * 3 + ax], gC[ax]);
    // the ket-centre derivative of block (s1, s2) is the bra-centre derivative
    // of block (s2, s1) (symmetric D and W): every ordered pair counts twice
    const int at = a.sh_atom[s1];
    for (int ax = 0; ax < 3; ++ax) {
        const double v = wave_sum_g(gA[ax]);
        if (lane == 0 && v != 0.0) atomicAdd(&a.grad[at * 3 + ax], 2.0 * v);
    }
}

// --------------------------------------------------------------- host side
struct DerivTables {
    int* ins = nullptr;
    double* coef = nullptr;
    int* lev = nullptr;
    int* acc = nullptr;
    int* contr = nullptr;
    double* ccoef = nullptr;
    int* vw = nullptr;
    int* vkind = nullptr;
};
static DerivTables g_dt[64];

hipError_t deriv_upload(int device) {
    if (device < 0 || device >= 64) return hipErrorInvalidDevice;
    static std::mutex mu;   // contexts may be created from several host threads
    std::lock_guard<std::mutex> lk(mu);
    DerivTables& t = g_dt[device];
    if (t.ins) return hipSuccess;
    hipError_t e;
    if ((e = hipMalloc((void**)&t.ins, sizeof(DPROG_INS)))) return e;
    if ((e = hipMalloc((void**)&t.coef, sizeof(DPROG_COEF)))) return e;
    if ((e = hipMalloc((void**)&t.lev, sizeof(DPROG_LEV)))) return e;
    if ((e = hipMalloc((void**)&t.acc, sizeof(DPROG_ACC)))) return e;
    if ((e = hipMalloc((void**)&t.contr, sizeof(DPROG_CONTR)))) return e;
    if ((e = hipMalloc((void**)&t.ccoef, sizeof(DPROG_CCOEF)))) return e;
    if ((e = hipMalloc((void**)&t.vw, sizeof(DVRR_WORDS)))) return e;
    if ((e = hipMalloc((void**)&t.vkind, sizeof(DVRR_KIND)))) return e;
    if ((e = hipMemcpy(t.vw, DVRR_WORDS, sizeof(DVRR_WORDS), hipMemcpyHostToDevice))) return e;
    if ((e = hipMemcpy(t.vkind, DVRR_KIND, sizeof(DVRR_KIND), hipMemcpyHostToDevice))) return e;
    if ((e = hipMemcpy(t.ins, DPROG_INS, sizeof(DPROG_INS), hipMemcpyHostToDevice))) return e;
    if ((e = hipMemcpy(t.coef, DPROG_COEF, sizeof(DPROG_COEF), hipMemcpyHostToDevice))) return e;
    if ((e = hipMemcpy(t.lev, DPROG_LEV, sizeof(DPROG_LEV), hipMemcpyHostToDevice))) return e;
    if ((e = hipMemcpy(t.acc, DPROG_ACC, sizeof(DPROG_ACC), hipMemcpyHostToDevice))) return e;
    if ((e = hipMemcpy(t.contr, DPROG_CONTR, sizeof(DPROG_CONTR), hipMemcpyHostToDevice))) return e;
    return hipMemcpy(t.ccoef, DPROG_CCOEF, sizeof(DPROG_CCOEF), hipMemcpyHostToDevice);
}

#include "generated/deriv_dispatch.inc"

// fp32: FP32 arithmetic for the tasks' integrals
hipError_t launch_grad_2e(int device, int la, int lb, int lc, int ld, const GradArgs& a, const BoysTables& tab,
                          const Task* t, int nt, int natom, int fp32, hipStream_t st) {
    if (nt <= 0) return hipSuccess;
    if (launch_grad_reg(la, lb, lc, ld, a, tab, t, nt, natom, fp32, st)) {
        // register code (grad_reg_*.hip)
    } else {
        const DerivTables& T = g_dt[device];
        int k = 0;
        for (; k < DPROG_NCLASS; ++k) {
            const int* d = DPROG_DESC + k * DPROG_DESC_FIELDS;
            if (d[0] == la && d[1] == lb && d[2] == lc && d[3] == ld) break;
        }
        if (k == DPROG_NCLASS) return hipErrorInvalidValue;
        // fields: la lb lc ld L nslots base nacc lev_off nvlev post_off nplev contr_off ncontr na nb nc nd
        const int* d = DPROG_DESC + k * DPROG_DESC_FIELDS;
        DerivProg P;
        P.la = la;
        P.lb = lb;
        P.lc = lc;
        P.ld = ld;
        P.L = d[4];
        P.nslots = d[5];
        P.base = d[6];
        P.nacc = d[7];
        P.lev = T.lev + d[8];
        P.nvlev = d[9];
        P.plev = T.lev + d[10];
        P.nplev = d[11];
        P.ins = T.ins;
        P.coef = T.coef;
        P.acc = T.acc;
        P.contr = T.contr;
        P.ccoef = T.ccoef;
        P.vw = nullptr;
        P.vkind = nullptr;
        P.nvpass = 0;
        for (int v = 0; v < DVRR_NCLASS; ++v) {
            const int* q = DVRR_DESC + 6 * v;
            if (q[0] == la && q[1] == lb && q[2] == lc && q[3] == ld) {
                P.vw = T.vw + (size_t)q[4] * 7 * 64;
                P.vkind = T.vkind + q[4];
                P.nvpass = q[5];
            }
        }
        const int ngam = d[14] * d[15] * d[16] * d[17];
        const size_t slots = (size_t)P.nslots + NSCALAR_D + ngam;
        if (fp32)
            hipLaunchKernelGGL(k_grad_prog<float>, dim3((unsigned)nt * GRAD_PROG_SPLIT), dim3(64),
                               sizeof(double) * 6 * 36 + sizeof(float) * slots, st, a, tab, t, nt, P);
        else
            hipLaunchKernelGGL(k_grad_prog<double>, dim3((unsigned)nt * GRAD_PROG_SPLIT), dim3(64),
                               sizeof(double) * (6 * 36 + slots), st, a, tab, t, nt, P);
    }
    return hipGetLastError();
}

hipError_t launch_grad_1e(const Grad1Args& a, const int2* pairs, const int* count, hipStream_t st) {
    // pairs grouped by class la * 3 + lb, count[9] pairs per class
    int off = 0;
#define M2_G1(LA, LB)                                                                                   \
    if (count[LA * 3 + LB] > 0)                                                                       \
        hipLaunchKernelGGL((k_grad_1e<LA, LB>), dim3(count[LA * 3 + LB]), dim3(64), 0, st, a, pairs + off, \
                           count[LA * 3 + LB]);                                                       \
    off += count[LA * 3 + LB];
    M2_G1(0, 0) M2_G1(0, 1) M2_G1(0, 2) M2_G1(1, 0) M2_G1(1, 1) M2_G1(1, 2) M2_G1(2, 0) M2_G1(2, 1) M2_G1(2, 2)
#undef M2_G1
    return hipGetLastError();
}

}  // namespace m2
