/*
 * m2_oracle.c -- CPU FP64 restatement of the reference's Fock-build path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the checker
 * or the reported CPU baseline.  The product (molecular2_amd/, libm2) never
 * links or calls it.
 *
 * What it restates (reference file:line):
 *   - libint2::Shell normalisation of raw contraction coefficients, the
 *     standard Cartesian order and the m=-l..l solid harmonics used by the
 *     reference's shells (src/basis.cpp:33-66; libint2 2.3.0-beta.3, absent
 *     from the reference tree, restated from its published conventions);
 *   - Boys function F_m(T);
 *   - contracted Coulomb ERIs (ab|cd) with Obara-Saika VRR + HGP HRR, which is
 *     the algorithm libint2's Engine(Operator::coulomb) implements
 *     (called at src/fockbuild.cpp:82-83,126);
 *   - IntegralEngine::compute_schwarz_ints (src/integrals.cpp:471-513),
 *     compute_shellblock_norm (src/integrals.cpp:451-469);
 *   - Fock::formJKdirect loop for loop (src/fockbuild.cpp:64-164);
 *   - the in-core path: compute_eris into a full tensor + Fock::formJK
 *     (src/eris.cpp:18-87, src/fockbuild.cpp:36-51);
 *   - one-electron S, T, V (src/integrals.cpp:394-449);
 *   - 3-centre (Q|xy) packed and 2-centre (P|Q) (src/eris.cpp:89-156,304-356);
 *   - Fock::compute_forces two-electron, one-electron and Pulay terms
 *     (src/rhfgrad.cpp:15-89,418-598) with compute_shellpair_list
 *     (src/integrals.cpp:515-566).
 *
 * Parity pins for each function are in tests/test_oracle_*.py (mpmath Boys,
 * closed-form (ss|ss), an independent numpy McMurchie-Davidson ERI code,
 * brute-force J/K, recorded reference energies, finite differences).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define OR_LMAX 6            /* max l of one (possibly derivative-shifted) shell */
#define OR_MAXPRIM 32
#define OR_LTOT 4 * OR_LMAX

typedef struct {
    int l, nprim, pure, atom;
    int bf, nbf, ncart;
    double O[3];
    double a[OR_MAXPRIM];
    double c[OR_MAXPRIM];    /* normalised (or raw for the unit s shell) */
} or_shell;

typedef struct {
    int nsh, nbf, pure;
    or_shell* sh;
} or_basis;

/* ------------------------------------------------------------------------ */
/* small math helpers                                                        */
/* ------------------------------------------------------------------------ */

static double dfact(int n) { /* n!! with (-1)!! = 1 */
    double r = 1.0;
    for (int k = n; k > 1; k -= 2) r *= k;
    return r;
}
static double fact(int n) {
    double r = 1.0;
    for (int k = 2; k <= n; ++k) r *= k;
    return r;
}
static double binom(int n, int k) {
    if (k < 0 || k > n) return 0.0;
    return fact(n) / (fact(k) * fact(n - k));
}

static inline int ncart_l(int l) { return (l + 1) * (l + 2) / 2; }
static inline int cum_l(int l) { return l * (l + 1) * (l + 2) / 6; } /* #components with |e| < l */
/* position of (x,y,z), x+y+z=L, inside its level: x descending, then y descending */
static inline int cidx(int x, int y, int z) {
    int L = x + y + z;
    return (L - x) * (L - x + 1) / 2 + (L - x - y);
    (void)z;
}
static inline int cpos(int x, int y, int z) { return cum_l(x + y + z) + cidx(x, y, z); }

/* component tables, filled lazily */
static int g_comp[4096][3];
static int g_comp_ready = 0;
static void init_comp(void) {
    if (g_comp_ready) return;
    int n = 0;
    for (int L = 0; L <= 2 * OR_LMAX + 4; ++L)
        for (int x = L; x >= 0; --x)
            for (int y = L - x; y >= 0; --y) {
                g_comp[n][0] = x; g_comp[n][1] = y; g_comp[n][2] = L - x - y; ++n;
            }
    g_comp_ready = 1;
}

/* ------------------------------------------------------------------------ */
/* Boys function                                                             */
/* ------------------------------------------------------------------------ */

void or_boys(int mmax, double T, double* F) {
    if (T < 1e-14) {
        for (int m = 0; m <= mmax; ++m) F[m] = 1.0 / (2 * m + 1) - T / (2 * m + 3);
        return;
    }
    double et = exp(-T);
    if (T > 45.0) { /* asymptotic start, upward recursion is stable here */
        F[0] = 0.5 * sqrt(M_PI / T) * erf(sqrt(T));
        for (int m = 0; m < mmax; ++m) F[m + 1] = ((2 * m + 1) * F[m] - et) / (2 * T);
        return;
    }
    /* series F_M(T) = e^-T sum_k (2T)^k / ((2M+1)(2M+3)...(2M+2k+1)) */
    double term = 1.0 / (2 * mmax + 1), sum = term;
    for (int k = 1; k < 400; ++k) {
        term *= 2.0 * T / (2 * mmax + 2 * k + 1);
        sum += term;
        if (term < 1e-18 * sum) break;
    }
    F[mmax] = et * sum;
    for (int m = mmax - 1; m >= 0; --m) F[m] = (2 * T * F[m + 1] + et) / (2 * m + 1);
}

/* ------------------------------------------------------------------------ */
/* normalisation and solid harmonics                                         */
/* ------------------------------------------------------------------------ */

/* libint2 convention: every Cartesian component of a shell carries the
 * normalisation of x^l exp(-a r^2); the contracted x^l function has unit
 * norm. */
static void normalise_shell(or_shell* s) {
    const int l = s->l;
    const double pi32 = pow(M_PI, 1.5);
    for (int k = 0; k < s->nprim; ++k) {
        double ta = 2.0 * s->a[k];
        s->c[k] *= sqrt(pow(2.0, l) * pow(ta, l + 1.5) / (pi32 * dfact(2 * l - 1)));
    }
    double nrm = 0.0;
    for (int p = 0; p < s->nprim; ++p)
        for (int q = 0; q < s->nprim; ++q) {
            double g = s->a[p] + s->a[q];
            nrm += s->c[p] * s->c[q] * dfact(2 * l - 1) * pi32 / (pow(2.0, l) * pow(g, l + 1.5));
        }
    nrm = 1.0 / sqrt(nrm);
    for (int k = 0; k < s->nprim; ++k) s->c[k] *= nrm;
}

/* Real solid harmonics from the closed form in Helgaker, Joergensen & Olsen,
 * Molecular Electronic-Structure Theory, eq. (6.4.47)-(6.4.50), then scaled so
 * that the combination of x^l-normalised Cartesians has unit norm. Rows are
 * m = -l..l, columns the Cartesian components in standard order. */
static double g_sph[OR_LMAX + 1][2 * OR_LMAX + 1][ (OR_LMAX + 1) * (OR_LMAX + 2) / 2];
static int g_sph_ready = 0;
static void init_sph(void) {
    if (g_sph_ready) return;
    init_comp();
    for (int l = 0; l <= OR_LMAX; ++l) {
        int nc = ncart_l(l);
        for (int m = -l; m <= l; ++m) {
            double* row = g_sph[l][m + l];
            for (int i = 0; i < nc; ++i) row[i] = 0.0;
            int am = abs(m);
            double vm = (m < 0) ? 0.5 : 0.0;
            for (int t = 0; t <= (l - am) / 2; ++t)
                for (int u = 0; u <= t; ++u) {
                    /* v runs over vm, vm+1, ..., floor(am/2 - vm) + vm */
                    int vmax2 = (int)floor(am / 2.0 - vm);
                    for (int vi = 0; vi <= vmax2; ++vi) {
                        double v = vi + vm;
                        int two_v = (int)lround(2 * v);
                        double sgn = ((t + (int)lround(v - vm)) % 2) ? -1.0 : 1.0;
                        double C = sgn * pow(0.25, t) * binom(l, t) * binom(l - t, am + t) *
                                   binom(t, u) * binom(am, two_v);
                        int px = 2 * t + am - 2 * u - two_v;
                        int py = 2 * u + two_v;
                        int pz = l - 2 * t - am;
                        if (px < 0 || pz < 0) continue;
                        row[cidx(px, py, pz)] += C;
                    }
                }
            /* normalise w.r.t. the x^l-normalised Cartesian metric */
            double nrm = 0.0;
            for (int i = 0; i < nc; ++i)
                for (int j = 0; j < nc; ++j) {
                    if (row[i] == 0.0 || row[j] == 0.0) continue;
                    const int* ci = g_comp[cum_l(l) + i];
                    const int* cj = g_comp[cum_l(l) + j];
                    if ((ci[0] + cj[0]) % 2 || (ci[1] + cj[1]) % 2 || (ci[2] + cj[2]) % 2) continue;
                    nrm += row[i] * row[j] * dfact(ci[0] + cj[0] - 1) * dfact(ci[1] + cj[1] - 1) *
                           dfact(ci[2] + cj[2] - 1) / dfact(2 * l - 1);
                }
            nrm = 1.0 / sqrt(nrm);
            for (int i = 0; i < nc; ++i) row[i] *= nrm;
        }
    }
    g_sph_ready = 1;
}

/* exported for tests: coefficient of Cartesian component ic in harmonic m */
double or_sph_coef(int l, int m, int ic) {
    init_sph();
    return g_sph[l][m + l][ic];
}

/* ------------------------------------------------------------------------ */
/* basis                                                                     */
/* ------------------------------------------------------------------------ */

or_basis* or_basis_create(int nsh, const int* l, const int* nprim, const double* exps,
                          const double* coefs, const double* centers, const int* atoms, int pure) {
    init_comp();
    init_sph();
    or_basis* b = (or_basis*)calloc(1, sizeof(or_basis));
    b->nsh = nsh;
    b->pure = pure;
    b->sh = (or_shell*)calloc(nsh, sizeof(or_shell));
    int off = 0, bf = 0;
    for (int s = 0; s < nsh; ++s) {
        or_shell* S = &b->sh[s];
        S->l = l[s];
        S->nprim = nprim[s];
        S->pure = pure;
        S->atom = atoms ? atoms[s] : -1;
        for (int k = 0; k < 3; ++k) S->O[k] = centers[3 * s + k];
        for (int k = 0; k < nprim[s]; ++k) { S->a[k] = exps[off + k]; S->c[k] = coefs[off + k]; }
        off += nprim[s];
        normalise_shell(S);
        S->ncart = ncart_l(S->l);
        S->nbf = pure ? 2 * S->l + 1 : S->ncart;
        S->bf = bf;
        bf += S->nbf;
    }
    b->nbf = bf;
    return b;
}
void or_basis_free(or_basis* b) { if (b) { free(b->sh); free(b); } }
int or_basis_nbf(const or_basis* b) { return b->nbf; }
/* normalised coefficients, for tests */
void or_basis_coefs(const or_basis* b, int s, double* out) {
    for (int k = 0; k < b->sh[s].nprim; ++k) out[k] = b->sh[s].c[k];
}

/* ------------------------------------------------------------------------ */
/* Cartesian ERI for one shell quartet: OS VRR + HGP HRR                     */
/* ------------------------------------------------------------------------ */

typedef struct {
    double* vrr;   /* [cum(Lab+1)][cum(Lcd+1)][L+1] */
    double* cont;  /* [cum(Lab+1)][cum(Lcd+1)] contracted m=0 */
    double* h1;    /* hrr scratch */
    double* h2;
    double* h3;
    size_t cap_vrr, cap_cont, cap_h;
} or_ws;

static void ws_reserve(or_ws* w, size_t nv, size_t nc, size_t nh) {
    if (nv > w->cap_vrr) { free(w->vrr); w->vrr = (double*)malloc(nv * sizeof(double)); w->cap_vrr = nv; }
    if (nc > w->cap_cont) { free(w->cont); w->cont = (double*)malloc(nc * sizeof(double)); w->cap_cont = nc; }
    if (nh > w->cap_h) {
        free(w->h1); free(w->h2); free(w->h3);
        w->h1 = (double*)malloc(nh * sizeof(double));
        w->h2 = (double*)malloc(nh * sizeof(double));
        w->h3 = (double*)malloc(nh * sizeof(double));
        w->cap_h = nh;
    }
}
static void ws_free(or_ws* w) { free(w->vrr); free(w->cont); free(w->h1); free(w->h2); free(w->h3); }

/* HRR on one side.  in[e][stride] for |e| in [la, la+lb] (indexed by cum
 * position, all lower levels unused), out[a][b][stride] with a,b Cartesian
 * components of la, lb.  AB = A - B. */
static void hrr(int la, int lb, const double* AB, const double* in, int in_rows, int stride,
                double* out, double* t1, double* t2) {
    /* level t: array over (e with |e| in [la, la+lb-t]) x (b with |b| = t) */
    int Lmax = la + lb;
    /* level 0 = in (rows cum positions) */
    const int nb0 = 1;
    (void)in_rows;
    /* copy level 0 into t1 with layout [epos - cum(la)][b][stride] */
    int ne0 = cum_l(Lmax + 1) - cum_l(la);
    double* cur = t1;
    double* nxt = t2;
    for (int e = 0; e < ne0; ++e)
        memcpy(cur + (size_t)e * nb0 * stride, in + (size_t)(cum_l(la) + e) * stride, stride * sizeof(double));
    int nb = nb0;
    for (int t = 1; t <= lb; ++t) {
        int ne = cum_l(Lmax - t + 1) - cum_l(la);
        int nbn = ncart_l(t);
        for (int e = 0; e < ne; ++e) {
            const int* ec = g_comp[cum_l(la) + e];
            for (int bi = 0; bi < nbn; ++bi) {
                const int* bc = g_comp[cum_l(t) + bi];
                int ax = bc[0] > 0 ? 0 : (bc[1] > 0 ? 1 : 2);
                int bm[3] = {bc[0], bc[1], bc[2]};
                bm[ax] -= 1;
                int ep[3] = {ec[0], ec[1], ec[2]};
                ep[ax] += 1;
                int bpos = cidx(bm[0], bm[1], bm[2]);
                int e1 = cpos(ep[0], ep[1], ep[2]) - cum_l(la);
                const double* x1 = cur + ((size_t)e1 * nb + bpos) * stride;
                const double* x0 = cur + ((size_t)e * nb + bpos) * stride;
                double* y = nxt + ((size_t)e * nbn + bi) * stride;
                for (int s = 0; s < stride; ++s) y[s] = x1[s] + AB[ax] * x0[s];
            }
        }
        double* tmp = cur; cur = nxt; nxt = tmp;
        nb = nbn;
    }
    /* rows of cur with |e| == la are the first ncart(la) entries */
    int na = ncart_l(la);
    memcpy(out, cur, (size_t)na * nb * stride * sizeof(double));
}

/* contracted Cartesian (ab|cd), out[na][nb][nc][nd].  Shell coefficient
 * arrays may be overridden (ca..cd) for derivative-weighted evaluations. */
static void eri_cart(const or_shell* A, const or_shell* B, const or_shell* C, const or_shell* D,
                     const double* ca, const double* cb, const double* cc, const double* cd,
                     or_ws* w, double* out) {
    const int la = A->l, lb = B->l, lc = C->l, ld = D->l;
    const int Lab = la + lb, Lcd = lc + ld, L = Lab + Lcd;
    const int NE = cum_l(Lab + 1), NF = cum_l(Lcd + 1), NM = L + 1;
    const size_t nvrr = (size_t)NE * NF * NM;
    const int na = ncart_l(la), nb = ncart_l(lb), nc = ncart_l(lc), nd = ncart_l(ld);
    size_t nh1 = (size_t)NE * ncart_l(lb) * nc * nd, nh2 = (size_t)NF * ncart_l(ld);
    size_t nh = (nh1 > nh2 ? nh1 : nh2) + (size_t)NE * nc * nd + 64;
    ws_reserve(w, nvrr, (size_t)NE * NF, nh);
    double* V = w->vrr;
    double* cont = w->cont;
    memset(cont, 0, sizeof(double) * NE * NF);
    if (!ca) ca = A->c;
    if (!cb) cb = B->c;
    if (!cc) cc = C->c;
    if (!cd) cd = D->c;
    double AB[3], CD[3], AB2 = 0, CD2 = 0;
    for (int k = 0; k < 3; ++k) {
        AB[k] = A->O[k] - B->O[k];
        CD[k] = C->O[k] - D->O[k];
        AB2 += AB[k] * AB[k];
        CD2 += CD[k] * CD[k];
    }
    double F[OR_LTOT + 8];
    const double twopi52 = 2.0 * pow(M_PI, 2.5);
#define VI(e, f, m) V[((size_t)(e) * NF + (f)) * NM + (m)]
    for (int i = 0; i < A->nprim; ++i)
        for (int j = 0; j < B->nprim; ++j) {
            double a = A->a[i], b = B->a[j], p = a + b;
            double P[3], PA[3];
            for (int k = 0; k < 3; ++k) {
                P[k] = (a * A->O[k] + b * B->O[k]) / p;
                PA[k] = P[k] - A->O[k];
            }
            double Kab = exp(-a * b / p * AB2) * ca[i] * cb[j];
            for (int k2 = 0; k2 < C->nprim; ++k2)
                for (int l2 = 0; l2 < D->nprim; ++l2) {
                    double c = C->a[k2], d = D->a[l2], q = c + d;
                    double Q[3], QC[3], W[3], WP[3], WQ[3], PQ2 = 0;
                    for (int k = 0; k < 3; ++k) {
                        Q[k] = (c * C->O[k] + d * D->O[k]) / q;
                        QC[k] = Q[k] - C->O[k];
                        W[k] = (p * P[k] + q * Q[k]) / (p + q);
                        WP[k] = W[k] - P[k];
                        WQ[k] = W[k] - Q[k];
                        PQ2 += (P[k] - Q[k]) * (P[k] - Q[k]);
                    }
                    double rho = p * q / (p + q);
                    double Kcd = exp(-c * d / q * CD2) * cc[k2] * cd[l2];
                    double pref = twopi52 / (p * q * sqrt(p + q)) * Kab * Kcd;
                    or_boys(L, rho * PQ2, F);
                    for (int m = 0; m <= L; ++m) VI(0, 0, m) = pref * F[m];
                    /* bra VRR: [e|0]^(m) */
                    for (int le = 1; le <= Lab; ++le)
                        for (int ei = cum_l(le); ei < cum_l(le + 1); ++ei) {
                            const int* e = g_comp[ei];
                            int ax = e[0] > 0 ? 0 : (e[1] > 0 ? 1 : 2);
                            int em[3] = {e[0], e[1], e[2]};
                            em[ax] -= 1;
                            int e1 = cpos(em[0], em[1], em[2]);
                            int e2 = -1;
                            double nfac = em[ax];
                            if (em[ax] > 0) { int t[3] = {em[0], em[1], em[2]}; t[ax] -= 1; e2 = cpos(t[0], t[1], t[2]); }
                            for (int m = 0; m <= L - le; ++m) {
                                double v = PA[ax] * VI(e1, 0, m) + WP[ax] * VI(e1, 0, m + 1);
                                if (e2 >= 0) v += nfac / (2 * p) * (VI(e2, 0, m) - rho / p * VI(e2, 0, m + 1));
                                VI(ei, 0, m) = v;
                            }
                        }
                    /* ket VRR: [e|f]^(m) */
                    for (int lf = 1; lf <= Lcd; ++lf)
                        for (int fi = cum_l(lf); fi < cum_l(lf + 1); ++fi) {
                            const int* f = g_comp[fi];
                            int ax = f[0] > 0 ? 0 : (f[1] > 0 ? 1 : 2);
                            int fm[3] = {f[0], f[1], f[2]};
                            fm[ax] -= 1;
                            int f1 = cpos(fm[0], fm[1], fm[2]);
                            int f2 = -1;
                            double nf = fm[ax];
                            if (fm[ax] > 0) { int t[3] = {fm[0], fm[1], fm[2]}; t[ax] -= 1; f2 = cpos(t[0], t[1], t[2]); }
                            for (int ei = 0; ei < cum_l(Lab + 1); ++ei) {
                                const int* e = g_comp[ei];
                                int le = e[0] + e[1] + e[2];
                                int e2 = -1;
                                if (e[ax] > 0) { int t[3] = {e[0], e[1], e[2]}; t[ax] -= 1; e2 = cpos(t[0], t[1], t[2]); }
                                for (int m = 0; m <= L - le - lf; ++m) {
                                    double v = QC[ax] * VI(ei, f1, m) + WQ[ax] * VI(ei, f1, m + 1);
                                    if (f2 >= 0) v += nf / (2 * q) * (VI(ei, f2, m) - rho / q * VI(ei, f2, m + 1));
                                    if (e2 >= 0) v += e[ax] / (2 * (p + q)) * VI(e2, f1, m + 1);
                                    VI(ei, fi, m) = v;
                                }
                            }
                        }
                    for (int ei = cum_l(la); ei < NE; ++ei)
                        for (int fi = cum_l(lc); fi < NF; ++fi) cont[ei * NF + fi] += VI(ei, fi, 0);
                }
        }
#undef VI
    /* ket HRR for every bra row e: cont[e][f] -> K[e][c][d] */
    int ncd = nc * nd;
    double* K = w->h3; /* [NE][nc*nd] */
    for (int e = cum_l(la); e < NE; ++e)
        hrr(lc, ld, CD, cont + (size_t)e * NF, NF, 1, K + (size_t)e * ncd, w->h1, w->h2);
    /* bra HRR over rows, stride ncd */
    hrr(la, lb, AB, K, NE, ncd, out, w->h1, w->h2);
    (void)na; (void)nb;
}

/* transform out[na][nb][nc][nd] Cartesian -> pure where requested */
static void cart2pure_dim(double* buf, int pre, int n_in, int post, int l, double* tmp) {
    int n_out = 2 * l + 1;
    for (int i = 0; i < pre; ++i)
        for (int m = 0; m < n_out; ++m)
            for (int k = 0; k < post; ++k) {
                double s = 0.0;
                for (int c = 0; c < n_in; ++c) s += g_sph[l][m][c] * buf[((size_t)i * n_in + c) * post + k];
                tmp[((size_t)i * n_out + m) * post + k] = s;
            }
    memcpy(buf, tmp, sizeof(double) * (size_t)pre * n_out * post);
}

static void to_pure4(const or_shell* const S[4], double* buf, double* tmp) {
    int dims[4];
    for (int i = 0; i < 4; ++i) dims[i] = S[i]->ncart;
    for (int i = 0; i < 4; ++i) {
        if (!S[i]->pure || S[i]->l < 2) { if (S[i]->pure) dims[i] = 2 * S[i]->l + 1; continue; }
        int pre = 1, post = 1;
        for (int k = 0; k < i; ++k) pre *= dims[k];
        for (int k = i + 1; k < 4; ++k) post *= dims[k];
        cart2pure_dim(buf, pre, dims[i], post, S[i]->l, tmp);
        dims[i] = 2 * S[i]->l + 1;
    }
}

/* contracted (ab|cd) in the basis's function convention */
static void eri_quartet(const or_shell* A, const or_shell* B, const or_shell* C, const or_shell* D,
                        or_ws* w, double* out, double* tmp) {
    eri_cart(A, B, C, D, NULL, NULL, NULL, NULL, w, out);
    const or_shell* S[4] = {A, B, C, D};
    to_pure4(S, out, tmp);
}

/* exported: one shell quartet in the basis convention (row-major f1..f4) */
void or_shell_quartet(const or_basis* b, int s1, int s2, int s3, int s4, double* out) {
    or_ws w = {0};
    const or_shell *A = &b->sh[s1], *B = &b->sh[s2], *C = &b->sh[s3], *D = &b->sh[s4];
    size_t n = (size_t)A->ncart * B->ncart * C->ncart * D->ncart;
    double* buf = (double*)malloc(sizeof(double) * n);
    double* tmp = (double*)malloc(sizeof(double) * n);
    eri_quartet(A, B, C, D, &w, buf, tmp);
    memcpy(out, buf, sizeof(double) * (size_t)A->nbf * B->nbf * C->nbf * D->nbf);
    free(buf); free(tmp); ws_free(&w);
}

/* ------------------------------------------------------------------------ */
/* Schwarz, shell-block norms                                                */
/* ------------------------------------------------------------------------ */

/* IntegralEngine::compute_schwarz_ints (src/integrals.cpp:471-513):
 * Q(s1,s2) = sqrt(max_ab |(ab|ab)|), symmetric fill. */
void or_schwarz(const or_basis* b, double* Q) {
    int nsh = b->nsh;
#pragma omp parallel
    {
        or_ws w = {0};
        double* buf = (double*)malloc(sizeof(double) * 50625 * 4);
        double* tmp = (double*)malloc(sizeof(double) * 50625 * 4);
#pragma omp for schedule(dynamic)
        for (int s1 = 0; s1 < nsh; ++s1)
            for (int s2 = 0; s2 <= s1; ++s2) {
                const or_shell *A = &b->sh[s1], *B = &b->sh[s2];
                eri_quartet(A, B, A, B, &w, buf, tmp);
                int n12 = A->nbf * B->nbf;
                double mx = 0.0;
                for (int i = 0; i < n12; ++i) {
                    double v = fabs(buf[(size_t)i * n12 + i]);
                    if (v > mx) mx = v;
                }
                Q[s1 * nsh + s2] = sqrt(mx);
                Q[s2 * nsh + s1] = sqrt(mx);
            }
        free(buf); free(tmp); ws_free(&w);
    }
}

/* compute_shellblock_norm (src/integrals.cpp:451-469): inf-norm (max |x|)
 * of each shell block.  A is n x n (symmetric use; row-major here). */
void or_shellblock_norm(const or_basis* b, const double* A, double* out) {
    int nsh = b->nsh, n = b->nbf;
    for (int s1 = 0; s1 < nsh; ++s1)
        for (int s2 = 0; s2 < nsh; ++s2) {
            double mx = 0.0;
            const or_shell *S1 = &b->sh[s1], *S2 = &b->sh[s2];
            for (int i = 0; i < S1->nbf; ++i)
                for (int j = 0; j < S2->nbf; ++j) {
                    double v = fabs(A[(size_t)(S1->bf + i) * n + S2->bf + j]);
                    if (v > mx) mx = v;
                }
            out[s1 * nsh + s2] = mx;
        }
}

/* ------------------------------------------------------------------------ */
/* Fock::formJKdirect restated (src/fockbuild.cpp:64-164)                   */
/* ------------------------------------------------------------------------ */

/* D, J, K: n x n row-major (matrices are symmetric so storage order only
 * matters for the unsymmetrised intermediates, which we symmetrise as the
 * reference does).  Q may be NULL (no Schwarz screening, as when the
 * reference's Schwarz matrix is empty).  nthreads <= 0: all.
 * Returns the number of quartets computed. */
/* Quartets with s1 in [s1lo, s1hi) only (a bounded sample of one build when
 * the range is partial); counts[0] += quartets, counts[1] += primitive
 * quartets.  OpenMP over the (s1, s2) pairs of the range. */
/* Dscreen: the density whose shell-block norms screen the quartets (NULL:
 * D itself).  UnrestrictedFock::formJKdirect (src/fockbuild.cpp:267-364)
 * screens both spin builds with Da + Db and digests each density like the
 * restricted loop. */
long or_jk_direct_range_s(const or_basis* b, const double* D, const double* Dscreen, const double* Q,
                          double precision, double multiplier, double* J, double* K, int nthreads, int s1lo,
                          int s1hi, long* counts) {
    const int nsh = b->nsh, n = b->nbf;
    const int do_screen = (Q != NULL);
    double* Dn = (double*)malloc(sizeof(double) * nsh * nsh);
    or_shellblock_norm(b, Dscreen ? Dscreen : D, Dn);
    long count = 0, pcount = 0;
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#else
    nthreads = 1;
#endif
    if (s1lo < 0) s1lo = 0;
    if (s1hi > nsh) s1hi = nsh;
    /* flat list of (s1, s2) pairs of the range, heaviest first */
    long npairs = 0;
    for (int s1 = s1lo; s1 < s1hi; ++s1) npairs += s1 + 1;
    int* pl = (int*)malloc(sizeof(int) * 2 * (npairs > 0 ? npairs : 1));
    long k = 0;
    for (int s1 = s1hi - 1; s1 >= s1lo; --s1)
        for (int s2 = 0; s2 <= s1; ++s2) { pl[2 * k] = s1; pl[2 * k + 1] = s2; ++k; }
    double* Jt = (double*)calloc((size_t)nthreads * n * n, sizeof(double));
    double* Kt = (double*)calloc((size_t)nthreads * n * n, sizeof(double));
#pragma omp parallel num_threads(nthreads) reduction(+ : count, pcount)
    {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        double* Jl = Jt + (size_t)tid * n * n;
        double* Kl = Kt + (size_t)tid * n * n;
        or_ws w = {0};
        double* buf = (double*)malloc(sizeof(double) * 50625 * 4);
        double* tmp = (double*)malloc(sizeof(double) * 50625 * 4);
#pragma omp for schedule(dynamic)
        for (long pi = 0; pi < npairs; ++pi) {
            const int s1 = pl[2 * pi], s2 = pl[2 * pi + 1];
            const or_shell* S1 = &b->sh[s1];
            const or_shell* S2 = &b->sh[s2];
            double Dn12 = do_screen ? Dn[s1 * nsh + s2] : 0.0;
            for (int s3 = 0; s3 <= s1; ++s3) {
                const or_shell* S3 = &b->sh[s3];
                double Dn123 = do_screen ? fmax(Dn[s1 * nsh + s3], fmax(Dn[s2 * nsh + s3], Dn12)) : 0.0;
                int s4max = (s1 == s3) ? s2 : s3;
                for (int s4 = 0; s4 <= s4max; ++s4) {
                    double Dn1234 = do_screen ? fmax(Dn[s1 * nsh + s4], fmax(Dn[s2 * nsh + s4],
                                                     fmax(Dn[s3 * nsh + s4], Dn123))) : 0.0;
                    if (do_screen && Dn1234 * Q[s1 * nsh + s2] * Q[s3 * nsh + s4] < precision) continue;
                    const or_shell* S4 = &b->sh[s4];
                    double deg = (s1 == s2 ? 1.0 : 2.0) * (s3 == s4 ? 1.0 : 2.0) *
                                 ((s1 == s3) ? (s2 == s4 ? 1.0 : 2.0) : 2.0);
                    eri_quartet(S1, S2, S3, S4, &w, buf, tmp);
                    ++count;
                    pcount += (long)S1->nprim * S2->nprim * S3->nprim * S4->nprim;
                    int n1 = S1->nbf, n2 = S2->nbf, n3 = S3->nbf, n4 = S4->nbf;
                    for (int f1 = 0, f = 0; f1 < n1; ++f1) {
                        int b1 = S1->bf + f1;
                        for (int f2 = 0; f2 < n2; ++f2) {
                            int b2 = S2->bf + f2;
                            for (int f3 = 0; f3 < n3; ++f3) {
                                int b3 = S3->bf + f3;
                                for (int f4 = 0; f4 < n4; ++f4, ++f) {
                                    int b4 = S4->bf + f4;
                                    double v = buf[f] * deg;
                                    Jl[b1 * n + b2] += multiplier * D[b3 * n + b4] * v;
                                    Jl[b3 * n + b4] += multiplier * D[b1 * n + b2] * v;
                                    Kl[b1 * n + b3] += 0.5 * multiplier * D[b2 * n + b4] * v;
                                    Kl[b2 * n + b4] += 0.5 * multiplier * D[b1 * n + b3] * v;
                                    Kl[b1 * n + b4] += 0.5 * multiplier * D[b2 * n + b3] * v;
                                    Kl[b2 * n + b3] += 0.5 * multiplier * D[b1 * n + b4] * v;
                                }
                            }
                        }
                    }
                }
            }
        }
        free(buf); free(tmp); ws_free(&w);
    }
    /* reduce thread-private partials, then symmetrise: X = (X + X^T)/4 */
    for (int t = 1; t < nthreads; ++t)
        for (size_t i = 0; i < (size_t)n * n; ++i) { Jt[i] += Jt[(size_t)t * n * n + i]; Kt[i] += Kt[(size_t)t * n * n + i]; }
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            J[i * n + j] = 0.25 * (Jt[i * n + j] + Jt[j * n + i]);
            K[i * n + j] = 0.25 * (Kt[i * n + j] + Kt[j * n + i]);
        }
    free(Jt); free(Kt); free(Dn); free(pl);
    if (counts) { counts[0] += count; counts[1] += pcount; }
    return count;
}

long or_jk_direct_range(const or_basis* b, const double* D, const double* Q, double precision,
                        double multiplier, double* J, double* K, int nthreads, int s1lo, int s1hi,
                        long* counts) {
    return or_jk_direct_range_s(b, D, NULL, Q, precision, multiplier, J, K, nthreads, s1lo, s1hi, counts);
}

long or_jk_direct(const or_basis* b, const double* D, const double* Q, double precision,
                  double multiplier, double* J, double* K, int nthreads) {
    return or_jk_direct_range(b, D, Q, precision, multiplier, J, K, nthreads, 0, b->nsh, NULL);
}

/* ------------------------------------------------------------------------ */
/* In-core path: full ERI tensor (src/eris.cpp:18-87) + Fock::formJK        */
/* (src/fockbuild.cpp:36-51).                                                */
/* ------------------------------------------------------------------------ */

/* all (ij|kl) into out[n^4], filled from unique shell quartets by symmetry */
void or_eri_full(const or_basis* b, double* out) {
    const int nsh = b->nsh;
    const size_t n = b->nbf;
#pragma omp parallel
    {
        or_ws w = {0};
        double* buf = (double*)malloc(sizeof(double) * 50625 * 4);
        double* tmp = (double*)malloc(sizeof(double) * 50625 * 4);
#pragma omp for schedule(dynamic)
        for (int s1 = 0; s1 < nsh; ++s1)
            for (int s2 = 0; s2 <= s1; ++s2)
                for (int s3 = 0; s3 <= s1; ++s3) {
                    int s4max = (s1 == s3) ? s2 : s3;
                    for (int s4 = 0; s4 <= s4max; ++s4) {
                        const or_shell *A = &b->sh[s1], *B = &b->sh[s2], *C = &b->sh[s3], *D = &b->sh[s4];
                        eri_quartet(A, B, C, D, &w, buf, tmp);
                        for (int f1 = 0, f = 0; f1 < A->nbf; ++f1)
                            for (int f2 = 0; f2 < B->nbf; ++f2)
                                for (int f3 = 0; f3 < C->nbf; ++f3)
                                    for (int f4 = 0; f4 < D->nbf; ++f4, ++f) {
                                        size_t i = A->bf + f1, j = B->bf + f2, k = C->bf + f3, l = D->bf + f4;
                                        double v = buf[f];
                                        out[((i * n + j) * n + k) * n + l] = v;
                                        out[((j * n + i) * n + k) * n + l] = v;
                                        out[((i * n + j) * n + l) * n + k] = v;
                                        out[((j * n + i) * n + l) * n + k] = v;
                                        out[((k * n + l) * n + i) * n + j] = v;
                                        out[((l * n + k) * n + i) * n + j] = v;
                                        out[((k * n + l) * n + j) * n + i] = v;
                                        out[((l * n + k) * n + j) * n + i] = v;
                                    }
                    }
                }
        free(buf); free(tmp); ws_free(&w);
    }
}

/* Fock::formJK: J(u,v) += m P(s,l)(uv|sl); K(u,v) += m P(s,l)(ul|sv). */
void or_jk_incore(int n, const double* eri, const double* P, double multiplier, double* J, double* K) {
    const size_t N = n;
#pragma omp parallel for
    for (int u = 0; u < n; ++u)
        for (int v = 0; v < n; ++v) {
            double j = 0.0, k = 0.0;
            for (int s = 0; s < n; ++s)
                for (int l = 0; l < n; ++l) {
                    j += multiplier * P[s * N + l] * eri[((u * N + v) * N + s) * N + l];
                    k += multiplier * P[s * N + l] * eri[((u * N + l) * N + s) * N + v];
                }
            J[u * N + v] = j;
            K[u * N + v] = k;
        }
}

/* ------------------------------------------------------------------------ */
/* One-electron integrals (src/integrals.cpp:394-449)                        */
/* ------------------------------------------------------------------------ */

/* 1D overlap table S[i][j], i <= imax, j <= jmax (OS recursion) */
static void os1d(double PA, double PB, double p, int imax, int jmax, double S[][OR_LMAX + 4]) {
    S[0][0] = 1.0;
    for (int i = 0; i <= imax; ++i)
        for (int j = 0; j <= jmax; ++j) {
            if (i == 0 && j == 0) continue;
            double v;
            if (i > 0) {
                v = PA * S[i - 1][j];
                if (i > 1) v += (i - 1) / (2 * p) * S[i - 2][j];
                if (j > 0) v += j / (2 * p) * S[i - 1][j - 1];
            } else {
                v = PB * S[i][j - 1];
                if (j > 1) v += (j - 1) / (2 * p) * S[i][j - 2];
            }
            S[i][j] = v;
        }
}

/* overlap + kinetic for one shell pair, Cartesian, out[na][nb] */
static void st_cart(const or_shell* A, const or_shell* B, const double* ca, const double* cb,
                    double* Sout, double* Tout) {
    int na = ncart_l(A->l), nb = ncart_l(B->l);
    if (!ca) ca = A->c;
    if (!cb) cb = B->c;
    for (int i = 0; i < na * nb; ++i) { if (Sout) Sout[i] = 0; if (Tout) Tout[i] = 0; }
    double AB2 = 0;
    for (int k = 0; k < 3; ++k) AB2 += (A->O[k] - B->O[k]) * (A->O[k] - B->O[k]);
    for (int i = 0; i < A->nprim; ++i)
        for (int j = 0; j < B->nprim; ++j) {
            double a = A->a[i], b = B->a[j], p = a + b;
            double pref = pow(M_PI / p, 1.5) * exp(-a * b / p * AB2) * ca[i] * cb[j];
            double S1[3][OR_LMAX + 4][OR_LMAX + 4];
            for (int k = 0; k < 3; ++k) {
                double P = (a * A->O[k] + b * B->O[k]) / p;
                os1d(P - A->O[k], P - B->O[k], p, A->l + 1, B->l + 2, S1[k]);
            }
            for (int ia = 0; ia < na; ++ia) {
                const int* ea = g_comp[cum_l(A->l) + ia];
                for (int ib = 0; ib < nb; ++ib) {
                    const int* eb = g_comp[cum_l(B->l) + ib];
                    double sx = S1[0][ea[0]][eb[0]], sy = S1[1][ea[1]][eb[1]], sz = S1[2][ea[2]][eb[2]];
                    if (Sout) Sout[ia * nb + ib] += pref * sx * sy * sz;
                    if (Tout) {
                        double t[3];
                        for (int k = 0; k < 3; ++k) {
                            int ii = ea[k], jj = eb[k];
                            double v = -2.0 * b * (2 * jj + 1) * S1[k][ii][jj] + 4.0 * b * b * S1[k][ii][jj + 2];
                            if (jj >= 2) v += jj * (jj - 1) * S1[k][ii][jj - 2];
                            t[k] = -0.5 * v;
                        }
                        Tout[ia * nb + ib] += pref * (t[0] * sy * sz + sx * t[1] * sz + sx * sy * t[2]);
                    }
                }
            }
        }
}

/* nuclear attraction of one pair to one unit point charge at Cpt (sign: +1/r
 * integral, caller scales by -Z).  Cartesian out[na][nb] accumulated. */
static void v_cart(const or_shell* A, const or_shell* B, const double* ca, const double* cb,
                   const double* Cpt, double scale, double* out) {
    int la = A->l, lb = B->l, Lab = la + lb;
    int NE = cum_l(Lab + 1);
    int na = ncart_l(la), nb = ncart_l(lb);
    if (!ca) ca = A->c;
    if (!cb) cb = B->c;
    double AB[3], AB2 = 0;
    for (int k = 0; k < 3; ++k) { AB[k] = A->O[k] - B->O[k]; AB2 += AB[k] * AB[k]; }
    double* V = (double*)calloc((size_t)NE * (Lab + 1), sizeof(double));
    double* cont = (double*)calloc((size_t)NE, sizeof(double));
    double F[OR_LTOT + 8];
    for (int i = 0; i < A->nprim; ++i)
        for (int j = 0; j < B->nprim; ++j) {
            double a = A->a[i], b = B->a[j], p = a + b;
            double P[3], PA[3], PC[3], PC2 = 0;
            for (int k = 0; k < 3; ++k) {
                P[k] = (a * A->O[k] + b * B->O[k]) / p;
                PA[k] = P[k] - A->O[k];
                PC[k] = P[k] - Cpt[k];
                PC2 += PC[k] * PC[k];
            }
            double pref = 2.0 * M_PI / p * exp(-a * b / p * AB2) * ca[i] * cb[j];
            or_boys(Lab, p * PC2, F);
            for (int m = 0; m <= Lab; ++m) V[m] = pref * F[m];
            for (int le = 1; le <= Lab; ++le)
                for (int ei = cum_l(le); ei < cum_l(le + 1); ++ei) {
                    const int* e = g_comp[ei];
                    int ax = e[0] > 0 ? 0 : (e[1] > 0 ? 1 : 2);
                    int em[3] = {e[0], e[1], e[2]};
                    em[ax] -= 1;
                    int e1 = cpos(em[0], em[1], em[2]);
                    int e2 = -1;
                    if (em[ax] > 0) { int t[3] = {em[0], em[1], em[2]}; t[ax] -= 1; e2 = cpos(t[0], t[1], t[2]); }
                    for (int m = 0; m <= Lab - le; ++m) {
                        double v = PA[ax] * V[e1 * (Lab + 1) + m] - PC[ax] * V[e1 * (Lab + 1) + m + 1];
                        if (e2 >= 0) v += em[ax] / (2 * p) * (V[e2 * (Lab + 1) + m] - V[e2 * (Lab + 1) + m + 1]);
                        V[ei * (Lab + 1) + m] = v;
                    }
                }
            for (int ei = cum_l(la); ei < NE; ++ei) cont[ei] += V[ei * (Lab + 1)];
        }
    double* res = (double*)malloc(sizeof(double) * (size_t)na * nb + 8);
    double* t1 = (double*)malloc(sizeof(double) * (size_t)NE * 64 + 64);
    double* t2 = (double*)malloc(sizeof(double) * (size_t)NE * 64 + 64);
    hrr(la, lb, AB, cont, NE, 1, res, t1, t2);
    for (int i = 0; i < na * nb; ++i) out[i] += scale * res[i];
    free(res); free(t1); free(t2); free(V); free(cont);
}

static void pair_to_pure(const or_shell* A, const or_shell* B, double* buf) {
    const or_shell* S[4] = {A, B, NULL, NULL};
    int dims[2] = {A->ncart, B->ncart};
    double tmp[4096];
    for (int i = 0; i < 2; ++i) {
        const or_shell* X = S[i];
        if (!X->pure || X->l < 2) continue;
        int pre = (i == 0) ? 1 : dims[0];
        int post = (i == 0) ? dims[1] : 1;
        cart2pure_dim(buf, pre, dims[i], post, X->l, tmp);
        dims[i] = 2 * X->l + 1;
    }
}

/* S, T, V (n x n row-major); V uses point charges Z at xyz (natom). */
void or_one_body(const or_basis* b, int natom, const double* Z, const double* xyz,
                 double* S, double* T, double* V) {
    const int nsh = b->nsh, n = b->nbf;
    for (int s1 = 0; s1 < nsh; ++s1)
        for (int s2 = 0; s2 <= s1; ++s2) {
            const or_shell *A = &b->sh[s1], *B = &b->sh[s2];
            double bs[1024], bt[1024], bv[1024];
            st_cart(A, B, NULL, NULL, bs, bt);
            memset(bv, 0, sizeof(bv));
            for (int c = 0; c < natom; ++c) v_cart(A, B, NULL, NULL, xyz + 3 * c, -Z[c], bv);
            pair_to_pure(A, B, bs);
            pair_to_pure(A, B, bt);
            pair_to_pure(A, B, bv);
            for (int i = 0; i < A->nbf; ++i)
                for (int j = 0; j < B->nbf; ++j) {
                    int I = A->bf + i, Jj = B->bf + j;
                    double vs = bs[i * B->nbf + j], vt = bt[i * B->nbf + j], vv = bv[i * B->nbf + j];
                    S[I * n + Jj] = S[Jj * n + I] = vs;
                    T[I * n + Jj] = T[Jj * n + I] = vt;
                    V[I * n + Jj] = V[Jj * n + I] = vv;
                }
        }
}

/* ------------------------------------------------------------------------ */
/* DF: 3-centre (K|xy) packed and 2-centre (K|L) (src/eris.cpp:89-156,      */
/* 304-356).  The unit s shell (exponent 0, coefficient 1) plays libint's   */
/* BraKet::xs_xx dummy.                                                      */
/* ------------------------------------------------------------------------ */

static or_shell unit_s(void) {
    or_shell u;
    memset(&u, 0, sizeof(u));
    u.l = 0; u.nprim = 1; u.a[0] = 0.0; u.c[0] = 1.0; u.ncart = 1; u.nbf = 1;
    return u;
}

/* out: (n(n+1)/2) x naux row-major; row x(x+1)/2 + y for x >= y */
void or_eri3_packed(const or_basis* obs, const or_basis* aux, double* out) {
    const int n = obs->nbf, naux = aux->nbf;
    const size_t npair = (size_t)n * (n + 1) / 2;
    (void)npair;
#pragma omp parallel
    {
        or_ws w = {0};
        or_shell u = unit_s();
        double* buf = (double*)malloc(sizeof(double) * 50625 * 4);
        double* tmp = (double*)malloc(sizeof(double) * 50625 * 4);
#pragma omp for schedule(dynamic)
        for (int s1 = 0; s1 < aux->nsh; ++s1) {
            const or_shell* P = &aux->sh[s1];
            u.O[0] = P->O[0]; u.O[1] = P->O[1]; u.O[2] = P->O[2];
            for (int s2 = 0; s2 < obs->nsh; ++s2)
                for (int s3 = 0; s3 <= s2; ++s3) {
                    const or_shell *B = &obs->sh[s2], *C = &obs->sh[s3];
                    eri_quartet(P, &u, B, C, &w, buf, tmp);
                    for (int f1 = 0, f = 0; f1 < P->nbf; ++f1)
                        for (int f2 = 0; f2 < B->nbf; ++f2) {
                            int b2 = B->bf + f2;
                            for (int f3 = 0; f3 < C->nbf; ++f3, ++f) {
                                int b3 = C->bf + f3;
                                if (b2 >= b3) out[((size_t)b2 * (b2 + 1) / 2 + b3) * naux + P->bf + f1] = buf[f];
                            }
                        }
                }
        }
        free(buf); free(tmp); ws_free(&w);
    }
}

void or_eri2(const or_basis* aux, double* out) {
    const int naux = aux->nbf;
    or_ws w = {0};
    or_shell u1 = unit_s(), u2 = unit_s();
    double* buf = (double*)malloc(sizeof(double) * 50625);
    double* tmp = (double*)malloc(sizeof(double) * 50625);
    for (int s1 = 0; s1 < aux->nsh; ++s1)
        for (int s2 = 0; s2 <= s1; ++s2) {
            const or_shell *P = &aux->sh[s1], *Q = &aux->sh[s2];
            memcpy(u1.O, P->O, sizeof(u1.O));
            memcpy(u2.O, Q->O, sizeof(u2.O));
            eri_quartet(P, &u1, Q, &u2, &w, buf, tmp);
            for (int f1 = 0, f = 0; f1 < P->nbf; ++f1)
                for (int f2 = 0; f2 < Q->nbf; ++f2, ++f) {
                    out[(size_t)(P->bf + f1) * naux + Q->bf + f2] = buf[f];
                    out[(size_t)(Q->bf + f2) * naux + P->bf + f1] = buf[f];
                }
        }
    free(buf); free(tmp); ws_free(&w);
}

/* ------------------------------------------------------------------------ */
/* Gradient (src/rhfgrad.cpp:15-89, 418-598)                                 */
/* ------------------------------------------------------------------------ */

/* shell with l shifted by dl and coefficients multiplied by (2a)^pw */
static or_shell shifted(const or_shell* S, int dl, int pw) {
    or_shell r = *S;
    r.l = S->l + dl;
    r.ncart = ncart_l(r.l);
    for (int k = 0; k < S->nprim; ++k) r.c[k] = S->c[k] * (pw ? 2.0 * S->a[k] : 1.0);
    return r;
}

/* Cartesian derivative of a quartet with respect to centre `which` (0..3):
 * d[3][na*nb*nc*nd].  dX (ab|cd) = (2x)(a+1|..) - a_x (a-1|..). */
static void eri_cart_deriv(const or_shell* S0[4], int which, or_ws* w, double* d, double* bufp, double* bufm) {
    or_shell Sp = shifted(S0[which], +1, 1);
    const or_shell* Sp4[4] = {S0[0], S0[1], S0[2], S0[3]};
    Sp4[which] = &Sp;
    eri_cart(Sp4[0], Sp4[1], Sp4[2], Sp4[3], NULL, NULL, NULL, NULL, w, bufp);
    int haveM = S0[which]->l > 0;
    or_shell Sm;
    if (haveM) {
        Sm = shifted(S0[which], -1, 0);
        const or_shell* Sm4[4] = {S0[0], S0[1], S0[2], S0[3]};
        Sm4[which] = &Sm;
        eri_cart(Sm4[0], Sm4[1], Sm4[2], Sm4[3], NULL, NULL, NULL, NULL, w, bufm);
    }
    int nc[4], ncp[4], ncm[4];
    for (int i = 0; i < 4; ++i) { nc[i] = ncart_l(S0[i]->l); ncp[i] = nc[i]; ncm[i] = nc[i]; }
    ncp[which] = ncart_l(S0[which]->l + 1);
    ncm[which] = haveM ? ncart_l(S0[which]->l - 1) : 0;
    int lw = S0[which]->l;
    for (int i0 = 0; i0 < nc[0]; ++i0)
        for (int i1 = 0; i1 < nc[1]; ++i1)
            for (int i2 = 0; i2 < nc[2]; ++i2)
                for (int i3 = 0; i3 < nc[3]; ++i3) {
                    int idx[4] = {i0, i1, i2, i3};
                    const int* e = g_comp[cum_l(lw) + idx[which]];
                    size_t out = ((size_t)(i0 * nc[1] + i1) * nc[2] + i2) * nc[3] + i3;
                    for (int ax = 0; ax < 3; ++ax) {
                        int ep[3] = {e[0], e[1], e[2]};
                        ep[ax] += 1;
                        int ip[4] = {i0, i1, i2, i3};
                        ip[which] = cidx(ep[0], ep[1], ep[2]);
                        double v = bufp[((size_t)(ip[0] * ncp[1] + ip[1]) * ncp[2] + ip[2]) * ncp[3] + ip[3]];
                        if (e[ax] > 0) {
                            int em[3] = {e[0], e[1], e[2]};
                            em[ax] -= 1;
                            int im[4] = {i0, i1, i2, i3};
                            im[which] = cidx(em[0], em[1], em[2]);
                            v -= e[ax] * bufm[((size_t)(im[0] * ncm[1] + im[1]) * ncm[2] + im[2]) * ncm[3] + im[3]];
                        }
                        d[(size_t)ax * nc[0] * nc[1] * nc[2] * nc[3] + out] = v;
                    }
                }
}

/* two-electron gradient: g[natom*3] += 1/4 G1:D exactly as
 * compute_forces + compute_2body_fock_deriv<1> define it.  Quartet loop
 * follows src/rhfgrad.cpp:454-590 with the shell-pair list of
 * src/integrals.cpp:515-566 (overlap norm >= 1e-12 or same centre). */
void or_grad_2e(const or_basis* b, const double* D, const double* Q, double precision,
                int natom, double* grad, int nthreads) {
    const int nsh = b->nsh, n = b->nbf;
    double* Dn = (double*)malloc(sizeof(double) * nsh * nsh);
    or_shellblock_norm(b, D, Dn);
    /* shell pair list */
    char* sig = (char*)calloc((size_t)nsh * nsh, 1);
    for (int s1 = 0; s1 < nsh; ++s1)
        for (int s2 = 0; s2 <= s1; ++s2) {
            const or_shell *A = &b->sh[s1], *B = &b->sh[s2];
            int same = A->O[0] == B->O[0] && A->O[1] == B->O[1] && A->O[2] == B->O[2];
            int s = same;
            if (!same) {
                double bs[1024];
                st_cart(A, B, NULL, NULL, bs, NULL);
                pair_to_pure(A, B, bs);
                double nrm = 0;
                for (int i = 0; i < A->nbf * B->nbf; ++i) nrm += bs[i] * bs[i];
                s = sqrt(nrm) >= 1e-12;
            }
            sig[s1 * nsh + s2] = (char)s;
        }
#ifdef _OPENMP
    if (nthreads <= 0) nthreads = omp_get_max_threads();
#else
    nthreads = 1;
#endif
    double* gt = (double*)calloc((size_t)nthreads * natom * 3, sizeof(double));
#pragma omp parallel num_threads(nthreads)
    {
        int tid = 0;
#ifdef _OPENMP
        tid = omp_get_thread_num();
#endif
        double* gl = gt + (size_t)tid * natom * 3;
        or_ws w = {0};
        size_t big = 200000;
        double* d = (double*)malloc(sizeof(double) * big * 3);
        double* bp = (double*)malloc(sizeof(double) * big);
        double* bm = (double*)malloc(sizeof(double) * big);
        double* tmp = (double*)malloc(sizeof(double) * big);
#pragma omp for schedule(dynamic)
        for (int s1 = nsh - 1; s1 >= 0; --s1)
            for (int s2 = 0; s2 <= s1; ++s2) {
                if (!sig[s1 * nsh + s2]) continue;
                double Dn12 = Dn[s1 * nsh + s2];
                for (int s3 = 0; s3 <= s1; ++s3) {
                    double Dn123 = fmax(Dn[s1 * nsh + s3], fmax(Dn[s2 * nsh + s3], Dn12));
                    int s4max = (s1 == s3) ? s2 : s3;
                    for (int s4 = 0; s4 <= s4max; ++s4) {
                        if (!sig[s3 * nsh + s4]) continue;
                        double Dn1234 = fmax(Dn[s1 * nsh + s4], fmax(Dn[s2 * nsh + s4], fmax(Dn[s3 * nsh + s4], Dn123)));
                        if (Q && Dn1234 * Q[s1 * nsh + s2] * Q[s3 * nsh + s4] < precision) continue;
                        const or_shell* S[4] = {&b->sh[s1], &b->sh[s2], &b->sh[s3], &b->sh[s4]};
                        double deg = (s1 == s2 ? 1.0 : 2.0) * (s3 == s4 ? 1.0 : 2.0) *
                                     ((s1 == s3) ? (s2 == s4 ? 1.0 : 2.0) : 2.0);
                        int nb4 = S[0]->nbf * S[1]->nbf * S[2]->nbf * S[3]->nbf;
                        int nc4 = S[0]->ncart * S[1]->ncart * S[2]->ncart * S[3]->ncart;
                        for (int c = 0; c < 4; ++c) {
                            eri_cart_deriv(S, c, &w, d, bp, bm);
                            for (int ax = 0; ax < 3; ++ax) {
                                double* blk = d + (size_t)ax * nc4;
                                to_pure4(S, blk, tmp);
                                double acc = 0.0;
                                for (int f1 = 0, f = 0; f1 < S[0]->nbf; ++f1) {
                                    int b1 = S[0]->bf + f1;
                                    for (int f2 = 0; f2 < S[1]->nbf; ++f2) {
                                        int b2 = S[1]->bf + f2;
                                        for (int f3 = 0; f3 < S[2]->nbf; ++f3) {
                                            int b3 = S[2]->bf + f3;
                                            for (int f4 = 0; f4 < S[3]->nbf; ++f4, ++f) {
                                                int b4 = S[3]->bf + f4;
                                                double gam = 0.5 * D[b1 * n + b2] * D[b3 * n + b4] -
                                                             0.125 * (D[b1 * n + b3] * D[b2 * n + b4] + D[b1 * n + b4] * D[b2 * n + b3]);
                                                acc += blk[f] * gam;
                                            }
                                        }
                                    }
                                }
                                (void)nb4;
                                gl[S[c]->atom * 3 + ax] += deg * acc;
                            }
                        }
                    }
                }
            }
        free(d); free(bp); free(bm); free(tmp); ws_free(&w);
    }
    for (int t = 0; t < nthreads; ++t)
        for (int i = 0; i < natom * 3; ++i) grad[i] += gt[(size_t)t * natom * 3 + i];
    free(gt); free(Dn); free(sig);
}

/* one-electron gradient terms: sum D:(dT + dV) - 2 W:dS (src/rhfgrad.cpp:
 * 22-46), derivatives of the basis centres and of the nuclear charges. */
void or_grad_1e(const or_basis* b, int natom, const double* Z, const double* xyz,
                const double* D, const double* W, double* grad) {
    const int nsh = b->nsh, n = b->nbf;
    for (int s1 = 0; s1 < nsh; ++s1)
        for (int s2 = 0; s2 < nsh; ++s2) {
            const or_shell *A = &b->sh[s1], *B = &b->sh[s2];
            int na = A->ncart, nb = B->ncart;
            /* derivative w.r.t. centre A (bra) only; the (s2,s1) pass covers B */
            or_shell Ap = shifted(A, +1, 1);
            or_shell Am;
            int haveM = A->l > 0;
            if (haveM) Am = shifted(A, -1, 0);
            double sp[1024], tp[1024], sm[1024], tm[1024];
            st_cart(&Ap, B, NULL, NULL, sp, tp);
            if (haveM) st_cart(&Am, B, NULL, NULL, sm, tm);
            /* V derivative w.r.t. A for each nucleus, and w.r.t. the nucleus
             * (Hellmann-Feynman) via translational invariance of <a|1/r_C|b>
             * in A, B, C: dC = -(dA + dB); we add the dA part to atom(A) and
             * -dA to atom C here, the (s2,s1) pass adds the dB parts. */
            for (int c = -1; c < natom; ++c) {
                double vp[1024], vm[1024];
                memset(vp, 0, sizeof(vp));
                memset(vm, 0, sizeof(vm));
                if (c >= 0) {
                    v_cart(&Ap, B, NULL, NULL, xyz + 3 * c, -Z[c], vp);
                    if (haveM) v_cart(&Am, B, NULL, NULL, xyz + 3 * c, -Z[c], vm);
                }
                for (int ax = 0; ax < 3; ++ax) {
                    double ds[1024], dt[1024], dv[1024];
                    for (int ia = 0; ia < na; ++ia) {
                        const int* e = g_comp[cum_l(A->l) + ia];
                        int ep[3] = {e[0], e[1], e[2]};
                        ep[ax]++;
                        int ipn = cidx(ep[0], ep[1], ep[2]);
                        int imn = -1;
                        if (e[ax] > 0) { int em[3] = {e[0], e[1], e[2]}; em[ax]--; imn = cidx(em[0], em[1], em[2]); }
                        for (int ib = 0; ib < nb; ++ib) {
                            double s = sp[ipn * nb + ib], t = tp[ipn * nb + ib], v = vp[ipn * nb + ib];
                            if (imn >= 0) {
                                s -= e[ax] * sm[imn * nb + ib];
                                t -= e[ax] * tm[imn * nb + ib];
                                v -= e[ax] * vm[imn * nb + ib];
                            }
                            ds[ia * nb + ib] = s;
                            dt[ia * nb + ib] = t;
                            dv[ia * nb + ib] = v;
                        }
                    }
                    pair_to_pure(A, B, ds);
                    pair_to_pure(A, B, dt);
                    pair_to_pure(A, B, dv);
                    double accS = 0, accT = 0, accV = 0;
                    for (int i = 0; i < A->nbf; ++i)
                        for (int j = 0; j < B->nbf; ++j) {
                            int I = A->bf + i, Jj = B->bf + j;
                            accS += W[I * n + Jj] * ds[i * B->nbf + j];
                            accT += D[I * n + Jj] * dt[i * B->nbf + j];
                            accV += D[I * n + Jj] * dv[i * B->nbf + j];
                        }
                    /* ordered pairs with the bra-centre derivative only: the
                     * ket-centre derivative of the (s1,s2) block is the bra
                     * derivative of the (s2,s1) block (symmetric D, W), hence 2x */
                    if (c < 0) {
                        grad[A->atom * 3 + ax] += 2.0 * (accT - 2.0 * accS);
                    } else {
                        grad[A->atom * 3 + ax] += 2.0 * accV;
                        grad[c * 3 + ax] -= 2.0 * accV;
                    }
                }
            }
        }
}

/* nuclear repulsion gradient (src/rhfgrad.cpp:60-86) */
void or_grad_nuc(int natom, const double* Z, const double* xyz, double* grad) {
    for (int a1 = 1; a1 < natom; ++a1)
        for (int a2 = 0; a2 < a1; ++a2) {
            double d[3], r2 = 0;
            for (int k = 0; k < 3; ++k) { d[k] = xyz[3 * a1 + k] - xyz[3 * a2 + k]; r2 += d[k] * d[k]; }
            double r = sqrt(r2);
            double f = Z[a1] * Z[a2] / (r * r2);
            for (int k = 0; k < 3; ++k) {
                grad[3 * a1 + k] += -d[k] * f;
                grad[3 * a2 + k] -= -d[k] * f;
            }
        }
}

int or_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
