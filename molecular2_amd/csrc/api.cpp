// C ABI of the engine (include/m2.h): context, basis and shell-pair set-up,
// Schwarz bounds, one-electron integrals and the direct J/K driver.
#include "ctx.hpp"
#include "jk_lists.hpp"
#include "one_body.hpp"
#include "generated/sph_tables.inc"
#include "generated/jk_dispatch.inc"
#include "generated/deriv_classes.inc"

#include <hip/hip_runtime.h>
#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <numeric>
#include <string>
#include <tuple>
#include <vector>

namespace m2 {
void launch_one_body(const OneBodyArgs&, hipStream_t);
void launch_blocknorm(const double* A, int n, int nsh, const int* sh_bf, const int* sh_nbf, double* out, hipStream_t);
void launch_symmetrize(double* J, double* K, double* Jo, double* Ko, int n, hipStream_t);
void launch_zero(double* p, size_t n, hipStream_t);
hipError_t heavy_upload(int device);
void la_release(m2_ctx* ctx);
bool heavy_has(int la, int lb, int lc, int ld);
hipError_t launch_jk_heavy(int device, int la, int lb, int lc, int ld, int pure, const JKArgs& a,
                           const BoysTables& tab, int nitems, int mode, hipStream_t st);
hipError_t launch_schwarz_heavy(int device, int la, int lb, int pure, const JKArgs& a, const BoysTables& tab,
                                int p0, int np, double* Q, hipStream_t st);
}  // namespace m2

using namespace m2;

namespace {

int pair_class(int la, int lb) {
    for (int i = 0; i < NPAIRCLASS; ++i)
        if (PAIR_L[i][0] == la && PAIR_L[i][1] == lb) return i;
    return -1;
}

double dfact(int n) {
    double r = 1.0;
    for (int k = n; k > 1; k -= 2) r *= k;
    return r;
}

// libint2::Shell normalisation: primitive factor of x^l exp(-a r^2), then the
// contracted x^l function scaled to unit norm.
void normalise(HostShell& s) {
    const double pi32 = std::pow(M_PI, 1.5);
    const int l = s.l;
    for (size_t k = 0; k < s.a.size(); ++k) {
        const double ta = 2.0 * s.a[k];
        s.c[k] *= std::sqrt(std::pow(2.0, l) * std::pow(ta, l + 1.5) / (pi32 * dfact(2 * l - 1)));
    }
    double nrm = 0.0;
    for (size_t p = 0; p < s.a.size(); ++p)
        for (size_t q = 0; q < s.a.size(); ++q)
            nrm += s.c[p] * s.c[q] * dfact(2 * l - 1) * pi32 / (std::pow(2.0, l) * std::pow(s.a[p] + s.a[q], l + 1.5));
    nrm = 1.0 / std::sqrt(nrm);
    for (auto& x : s.c) x *= nrm;
}

// Boys F_m(T) on the host, for the device tables: series
// F_M(T) = e^-T sum_k (2T)^k / ((2M+1)(2M+3)...(2M+2k+1)) then downward recursion.
void boys_host(int mmax, double T, double* F) {
    const double et = std::exp(-T);
    double term = 1.0 / (2 * mmax + 1), sum = term;
    for (int k = 1; k < 2000; ++k) {
        term *= 2.0 * T / (2 * mmax + 2 * k + 1);
        sum += term;
        if (term < 1e-20 * sum) break;
    }
    F[mmax] = et * sum;
    for (int m = mmax - 1; m >= 0; --m) F[m] = (2.0 * T * F[m + 1] + et) / (2 * m + 1);
}

}  // namespace


namespace {

BoysTables tables(m2_ctx* ctx) { return boys_tables(ctx); }

m2_status upload_basis(m2_ctx* ctx, BasisData& b) {
    b.h_bf.resize(b.nsh);
    b.h_nbf.resize(b.nsh);
    b.h_l.resize(b.nsh);
    b.h_O.resize(3 * b.nsh);
    for (int s = 0; s < b.nsh; ++s) {
        b.h_bf[s] = b.sh[s].bf;
        b.h_nbf[s] = b.sh[s].nbf;
        b.h_l[s] = b.sh[s].l;
        for (int k = 0; k < 3; ++k) b.h_O[3 * s + k] = b.sh[s].O[k];
    }
    HIPCHK(b.d_O.reserve(3 * b.nsh));
    HIPCHK(b.d_bf.reserve(b.nsh));
    HIPCHK(b.d_nbf.reserve(b.nsh));
    HIPCHK(hipMemcpy(b.d_O.p, b.h_O.data(), sizeof(double) * 3 * b.nsh, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(b.d_bf.p, b.h_bf.data(), sizeof(int) * b.nsh, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(b.d_nbf.p, b.h_nbf.data(), sizeof(int) * b.nsh, hipMemcpyHostToDevice));
    return M2_OK;
}

int np_bucket(int np) {
    // primitive-pair count buckets: lanes of one wavefront differ by < 2x
    static const int edges[] = {1, 2, 4, 8, 16, 32, 64, 128, 1 << 30};
    for (int i = 0; i < 9; ++i)
        if (np <= edges[i]) return i;
    return 9;
}

JKArgs jk_args(m2_ctx* ctx) {
    JKArgs a{};
    a.prims = ctx->d_prims.p;
    a.prims_f = ctx->d_prims_f.p;
    a.kets_s = ctx->d_kets_s.p;
    a.pair_rec = ctx->d_pair_rec.p;
    a.kets_sf = ctx->d_kets_sf.p;
    a.pair_off = ctx->d_pair_off.p;
    a.pair_np = ctx->d_pair_np.p;
    a.pair_sh = ctx->d_pair_sh.p;
    a.pair_Q = ctx->d_pair_Q.p;
    a.sh_O = ctx->obs.d_O.p;
    a.sh_bf = ctx->obs.d_bf.p;
    a.Dn = ctx->d_Dn.p;
    a.prim_q = ctx->d_prim_q.p;
    a.fam_q = ctx->d_fam_q.p;
    a.prim_tau = 0.0;
    a.d_lds_max = 2;
    a.boys_lds_max = 1;
    a.n = ctx->obs.nbf;
    a.nsh = ctx->obs.nsh;
    if (ctx->fam_on) {
        a.fam_prims = ctx->d_fam_prims.p;
        a.fam_s = ctx->d_fam_s.p;
        a.fam_sf = ctx->d_fam_sf.p;
        a.fam_prims_f = ctx->d_fam_prims_f.p;
        a.fam_km = ctx->d_fam_km.p;
        a.fam_km_f = ctx->d_fam_km_f.p;
        a.fam_off = ctx->d_fam_off.p;
        a.fam_np = ctx->d_fam_np.p;
        a.fam_mem = ctx->d_fam_mem.p;
        a.pair_fam = ctx->d_pair_fam.p;
        a.pair_mem = ctx->d_pair_mem.p;
    }
    return a;
}

// s-s ket families: s shells with the same centre and identical exponents
// (a general contraction written as segmented shells, e.g. the two contracted
// s shells of O in cc-pVDZ) form a family of at most two shells; the class-0
// shell pairs whose shells come from the same two families form a family
// pair (<= FAM_MAXMEM members) whose primitive pairs are computed once; the
// record's K carries no contraction coefficients, and beside it every member
// keeps its full factor K c_c[i] c_d[j] (what K would be in that member's own
// shell-pair record), which the kernel applies per member.  Families are used when they cut the class-0 ket
// count by at least 15%.  Family pairs are ordered like pairs (np bucket, then
// the largest member Q descending); pair_fam / pair_mem give every class-0
// pair its family pair and slot.  Which class combinations use them is
// fam_combo (m2_common.hpp).
static m2_status build_families(m2_ctx* ctx) {
    BasisData& b = ctx->obs;
    ctx->fam_on = false;
    ctx->fam_wide = false;
    ctx->fam_seg = PairClassSeg{};
    const PairClassSeg& c0 = ctx->cls[0];
    if (c0.end <= c0.begin) return M2_OK;
    std::vector<int> fid(b.nsh, -1), fsize;
    {
        std::map<std::vector<double>, int> last;   // centre + exponents -> newest family of that key
        for (int s = 0; s < b.nsh; ++s) {
            const HostShell& h = b.sh[s];
            if (h.l != 0) continue;
            std::vector<double> key = {h.O[0], h.O[1], h.O[2]};
            key.insert(key.end(), h.a.begin(), h.a.end());
            auto it = last.find(key);
            if (it != last.end() && fsize[it->second] < 2) {
                fid[s] = it->second;
                ++fsize[fid[s]];
            } else {
                fid[s] = (int)fsize.size();
                fsize.push_back(1);
                last[key] = fid[s];
            }
        }
    }
    struct FP {
        int f1, f2, rep1, rep2;   // families (f1 >= f2) and a shell of each
        std::vector<int> mem;     // member pair indices
        double Q = 0.0;
        std::vector<PrimPair> pp;
        std::vector<double> km;
        std::vector<float> q;     // primitive bounds
    };
    std::map<std::pair<int, int>, int> index;
    std::vector<FP> fps;
    for (int i = c0.begin; i < c0.end; ++i) {
        const int sa = ctx->pair_sh[i].x, sb = ctx->pair_sh[i].y;
        const int f1 = std::max(fid[sa], fid[sb]), f2 = std::min(fid[sa], fid[sb]);
        auto it = index.find({f1, f2});
        int k;
        if (it == index.end()) {
            k = (int)fps.size();
            index[{f1, f2}] = k;
            FP fp;
            fp.f1 = f1;
            fp.f2 = f2;
            fp.rep1 = fid[sa] == f1 ? sa : sb;
            fp.rep2 = fid[sa] == f1 ? sb : sa;
            fps.push_back(fp);
        } else {
            k = it->second;
        }
        ARGCHK((int)fps[k].mem.size() < FAM_MAXMEM, "m2_basis_set: s-shell family pair with too many members");
        fps[k].mem.push_back(i);
        fps[k].Q = std::max(fps[k].Q, ctx->pair_Q[i]);
    }
    const int nfp = (int)fps.size();
    if (ctx->fam_mode == 0) return M2_OK;
    if (ctx->fam_mode == 1 && nfp * 1.15 > c0.end - c0.begin) return M2_OK;   // too little sharing
    // member activity: over the family pairs any quartet can reach
    // (Q_F max Q >= 1e-10), the mean number of members within 1e-4 of the
    // family's largest Q
    {
        double qmax = 0.0, nact = 0.0, nused = 0.0;
        for (double q : ctx->pair_Q) qmax = std::max(qmax, q);
        for (const FP& fp : fps) {
            if (fp.Q * qmax < 1e-10) continue;
            nused += 1.0;
            for (int j : fp.mem) nact += ctx->pair_Q[j] >= 1e-4 * fp.Q ? 1.0 : 0.0;
        }
        ctx->fam_wide = ctx->fam_mode == 2 || (nused > 0.0 && nact >= 1.6 * nused);
    }
    const double sqrt2pi52 = std::sqrt(2.0 * std::pow(M_PI, 2.5));
    for (FP& fp : fps) {
        const HostShell &A = b.sh[fp.rep1], &B = b.sh[fp.rep2];
        double AB2 = 0;
        for (int k = 0; k < 3; ++k) AB2 += (A.O[k] - B.O[k]) * (A.O[k] - B.O[k]);
        for (size_t i = 0; i < A.a.size(); ++i)
            for (size_t j = 0; j < B.a.size(); ++j) {
                const double a = A.a[i], bb = B.a[j], p = a + bb;
                const double ex = std::exp(-a * bb / p * AB2);
                if (ex < 1e-20) continue;   // the same primitive-pair screen as the shell pairs
                PrimPair q{};
                q.p = p;
                q.inv2p = 0.5 / p;
                q.K = sqrt2pi52 * ex / p;
                for (int k = 0; k < 3; ++k) {
                    q.P[k] = (a * A.O[k] + bb * B.O[k]) / p;
                    q.PA[k] = q.P[k] - A.O[k];
                }
                q.a = a;
                fp.pp.push_back(q);
                for (int m = 0; m < FAM_MAXMEM; ++m) {
                    double c = 0.0;
                    if (m < (int)fp.mem.size()) {
                        const int sc = ctx->pair_sh[fp.mem[m]].x, sd = ctx->pair_sh[fp.mem[m]].y;
                        // (sc sd| = sum_ij c_sc[i] c_sd[j] G_ij with i over the f1 family's exponents
                        c = fid[sc] == fp.f1 ? b.sh[sc].c[i] * b.sh[sd].c[j] : b.sh[sd].c[i] * b.sh[sc].c[j];
                    }
                    c *= q.K;   // the member's full ket factor K_cd of this primitive pair
                    fp.km.push_back(c);
                }
            }
        // primitive bounds: an s-s primitive pair's self-repulsion is
        // K^2 / sqrt(2p) (T = 0), so q = max_m |K_m| / (2p)^(1/4) over the
        // members' factors; primitives ordered by q descending
        const size_t np = fp.pp.size();
        std::vector<float> q(np);
        for (size_t k = 0; k < np; ++k) {
            double mx = 0.0;
            for (int m = 0; m < FAM_MAXMEM; ++m) mx = std::max(mx, std::fabs(fp.km[FAM_MAXMEM * k + m]));
            q[k] = (float)(mx / std::pow(2.0 * fp.pp[k].p, 0.25));
        }
        std::vector<int> idx(np);
        std::iota(idx.begin(), idx.end(), 0);
        std::stable_sort(idx.begin(), idx.end(), [&](int x, int y) { return q[x] > q[y]; });
        std::vector<PrimPair> pp(np);
        std::vector<double> km(fp.km.size());
        fp.q.resize(np);
        for (size_t k = 0; k < np; ++k) {
            pp[k] = fp.pp[idx[k]];
            fp.q[k] = q[idx[k]];
            for (int m = 0; m < FAM_MAXMEM; ++m) km[FAM_MAXMEM * k + m] = fp.km[FAM_MAXMEM * idx[k] + m];
        }
        fp.pp.swap(pp);
        fp.km.swap(km);
    }
    std::vector<int> order(nfp);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
        const int bx = np_bucket((int)fps[x].pp.size()), by = np_bucket((int)fps[y].pp.size());
        if (bx != by) return bx > by;
        return fps[x].Q > fps[y].Q;
    });
    const int npair = (int)ctx->pair_sh.size();
    std::vector<PrimPair> prims;
    std::vector<BraPair<float>> prims_f;
    std::vector<double> km, fQ(nfp);
    std::vector<float> km_f, fq;
    std::vector<int> off(nfp), np(nfp), pfam(npair, -1), pmem(npair, -1);
    std::vector<int4> mem(nfp);
    for (int r = 0; r < nfp; ++r) {
        const FP& fp = fps[order[r]];
        off[r] = (int)prims.size();
        np[r] = (int)fp.pp.size();
        fQ[r] = fp.Q;
        int m4[FAM_MAXMEM] = {-1, -1, -1, -1};
        for (size_t m = 0; m < fp.mem.size(); ++m) {
            m4[m] = fp.mem[m];
            pfam[fp.mem[m]] = r;
            pmem[fp.mem[m]] = (int)m;
        }
        mem[r] = make_int4(m4[0], m4[1], m4[2], m4[3]);
        for (const PrimPair& q : fp.pp) {
            prims.push_back(q);
            BraPair<float> f{};
            f.p = (float)q.p;
            f.inv2p = (float)q.inv2p;
            f.K = (float)q.K;
            for (int k = 0; k < 3; ++k) {
                f.P[k] = (float)q.P[k];
                f.PA[k] = (float)q.PA[k];
            }
            prims_f.push_back(f);
        }
        km.insert(km.end(), fp.km.begin(), fp.km.end());
        fq.insert(fq.end(), fp.q.begin(), fp.q.end());
    }
    km_f.assign(km.begin(), km.end());
    HIPCHK(ctx->d_fam_q.reserve(std::max<size_t>(fq.size(), 1)));
    if (!fq.empty()) HIPCHK(hipMemcpy(ctx->d_fam_q.p, fq.data(), sizeof(float) * fq.size(), hipMemcpyHostToDevice));
    HIPCHK(ctx->d_fam_prims.reserve(std::max<size_t>(prims.size(), 1)));
    HIPCHK(ctx->d_fam_prims_f.reserve(std::max<size_t>(prims.size(), 1)));
    HIPCHK(ctx->d_fam_km.reserve(std::max<size_t>(km.size(), 1)));
    HIPCHK(ctx->d_fam_km_f.reserve(std::max<size_t>(km.size(), 1)));
    HIPCHK(ctx->d_fam_Q.reserve(nfp));
    HIPCHK(ctx->d_fam_off.reserve(nfp));
    HIPCHK(ctx->d_fam_np.reserve(nfp));
    HIPCHK(ctx->d_fam_mem.reserve(nfp));
    HIPCHK(ctx->d_pair_fam.reserve(npair));
    HIPCHK(ctx->d_pair_mem.reserve(npair));
    std::vector<FamS> fs(prims.size());
    for (size_t i = 0; i < prims.size(); ++i) {
        fs[i] = FamS{};
        fs[i].p = prims[i].p;
        for (int k = 0; k < 3; ++k) fs[i].P[k] = prims[i].P[k];
        for (int m = 0; m < FAM_MAXMEM; ++m) fs[i].km[m] = km[FAM_MAXMEM * i + m];
        fs[i].q = fq[i];
    }
    std::vector<FamSf> fsf(prims.size());
    for (size_t i = 0; i < prims.size(); ++i) {
        fsf[i] = FamSf{};
        fsf[i].p = (float)prims[i].p;
        for (int k = 0; k < 3; ++k) fsf[i].P[k] = (float)prims[i].P[k];
        for (int m = 0; m < FAM_MAXMEM; ++m) fsf[i].km[m] = (float)km[FAM_MAXMEM * i + m];
        fsf[i].q = fq[i];
    }
    HIPCHK(ctx->d_fam_s.reserve(std::max<size_t>(fs.size(), 1)));
    HIPCHK(ctx->d_fam_sf.reserve(std::max<size_t>(fsf.size(), 1)));
    if (!prims.empty()) {
        HIPCHK(hipMemcpy(ctx->d_fam_s.p, fs.data(), sizeof(FamS) * fs.size(), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(ctx->d_fam_sf.p, fsf.data(), sizeof(FamSf) * fsf.size(), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(ctx->d_fam_prims.p, prims.data(), sizeof(PrimPair) * prims.size(), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(ctx->d_fam_prims_f.p, prims_f.data(), sizeof(BraPair<float>) * prims.size(),
                         hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(ctx->d_fam_km.p, km.data(), sizeof(double) * km.size(), hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(ctx->d_fam_km_f.p, km_f.data(), sizeof(float) * km.size(), hipMemcpyHostToDevice));
    }
    HIPCHK(hipMemcpy(ctx->d_fam_Q.p, fQ.data(), sizeof(double) * nfp, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->d_fam_off.p, off.data(), sizeof(int) * nfp, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->d_fam_np.p, np.data(), sizeof(int) * nfp, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->d_fam_mem.p, mem.data(), sizeof(int4) * nfp, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(ctx->d_pair_fam.p, pfam.data(), sizeof(int) * npair, hipMemcpyHostToDevice));
    ctx->h_pair_fam = pfam;
    HIPCHK(hipMemcpy(ctx->d_pair_mem.p, pmem.data(), sizeof(int) * npair, hipMemcpyHostToDevice));
    PairClassSeg& seg = ctx->fam_seg;
    seg.begin = 0;
    seg.end = nfp;
    for (int r = 0; r < nfp;) {
        const int bk = np_bucket(np[r]), g0 = r;
        while (r < nfp && np_bucket(np[r]) == bk) ++r;
        seg.groups.push_back(PairGroup{g0, r});
    }
    ctx->fam_on = true;
    return M2_OK;
}

// Build canonical shell pairs (la >= lb; equal l: a >= b), their primitive
// pair records, Schwarz bounds (device), then order each class by
// (primitive-count bucket, Q descending).
// Drops every captured J/K graph (their kernel arguments name this basis's
// buffers and launch shape) and moves to a new key generation.
void drop_graphs(m2_ctx* ctx) {
    for (JKGraph& g : ctx->graphs) (void)hipGraphExecDestroy(g.exec);
    ctx->graphs.clear();
    ++ctx->graph_gen;
}

// Grid bound of list ci at its current chunks per item: every bra with a
// nonzero Schwarz bound with all the kets it can pair with (kets up to its
// own position when bra and ket class agree), so no density gives more
// items (an unscreened build gives exactly this many).
void list_bound(m2_ctx* ctx, int ci) {
    JKList& L = ctx->lists[ci];
    long long items = 0;
    for (int i = 0; i < L.nbra; ++i) {
        const int bra = L.bra0 + i;
        if (!(ctx->pair_Q[bra] > 0.0)) continue;   // the list kernels give such bras no items
        const int lim = L.same ? (L.bra_fam ? ctx->h_pair_fam[bra] : bra) + 1 : INT_MAX;
        long long ch = 0;
        for (int g = 0; g < L.ngrp; ++g) {
            const int len = std::min(L.gend[g], lim) - L.gbeg[g];
            if (len > 0) ch += (len + L.kc - 1) / L.kc;
        }
        items += (ch + L.bt - 1) / L.bt;
    }
    L.bound = (int)std::min<long long>(items, L.cap);
}

// the list descriptors to the device (read by the list kernels); after a
// basis change or a new chunks-per-item choice
m2_status upload_jklists(m2_ctx* ctx) {
    for (int ci = 0; ci < NCOMBO; ++ci)
        if (ctx->lists[ci].nbra > 0) list_bound(ctx, ci);
    HIPCHK(ctx->d_jklists.reserve(NCOMBO));
    HIPCHK(hipMemcpy(ctx->d_jklists.p, ctx->lists, sizeof(JKList) * NCOMBO, hipMemcpyHostToDevice));
    return M2_OK;
}

m2_status build_pairs(m2_ctx* ctx) {
    BasisData& b = ctx->obs;
    struct P {
        int a, b, cls;
        std::vector<PrimPair> pp;
        std::vector<float> q;   // primitive-pair bounds (after the Schwarz pass)
    };
    std::vector<P> pairs;
    const double sqrt2pi52 = std::sqrt(2.0 * std::pow(M_PI, 2.5));
    for (int s1 = 0; s1 < b.nsh; ++s1)
        for (int s2 = 0; s2 < b.nsh; ++s2) {
            const HostShell &A = b.sh[s1], &B = b.sh[s2];
            if (A.l < B.l || (A.l == B.l && s1 < s2)) continue;
            P pr;
            pr.a = s1;
            pr.b = s2;
            pr.cls = pair_class(A.l, B.l);
            if (pr.cls < 0) {
                ctx->err = "m2_basis_set: orbital shells must have l <= 2";
                return M2_ERR_ARG;
            }
            double AB2 = 0;
            for (int k = 0; k < 3; ++k) AB2 += (A.O[k] - B.O[k]) * (A.O[k] - B.O[k]);
            for (size_t i = 0; i < A.a.size(); ++i)
                for (size_t j = 0; j < B.a.size(); ++j) {
                    const double a = A.a[i], bb = B.a[j], p = a + bb;
                    const double ex = std::exp(-a * bb / p * AB2);
                    // primitive-pair screen far below any FP64 J/K tolerance
                    if (ex < 1e-20) continue;
                    PrimPair q{};
                    q.p = p;
                    q.inv2p = 0.5 / p;
                    q.K = sqrt2pi52 * ex * A.c[i] * B.c[j] / p;
                    for (int k = 0; k < 3; ++k) {
                        q.P[k] = (a * A.O[k] + bb * B.O[k]) / p;
                        q.PA[k] = q.P[k] - A.O[k];
                    }
                    q.a = a;
                    pr.pp.push_back(q);
                }
            pairs.push_back(std::move(pr));
        }
    // first layout: by class, for the Schwarz pass
    std::stable_sort(pairs.begin(), pairs.end(), [](const P& x, const P& y) { return x.cls < y.cls; });
    auto upload = [&](void) -> m2_status {
        const int npair = (int)pairs.size();
        ctx->pair_sh.resize(npair);
        ctx->pair_off.resize(npair);
        ctx->pair_np.resize(npair);
        std::vector<PrimPair> prims;
        for (int i = 0; i < npair; ++i) {
            ctx->pair_sh[i] = make_int2(pairs[i].a, pairs[i].b);
            ctx->pair_off[i] = (int)prims.size();
            ctx->pair_np[i] = (int)pairs[i].pp.size();
            prims.insert(prims.end(), pairs[i].pp.begin(), pairs[i].pp.end());
        }
        HIPCHK(ctx->d_prims.reserve(prims.size()));
        HIPCHK(ctx->d_pair_off.reserve(npair));
        HIPCHK(ctx->d_pair_np.reserve(npair));
        HIPCHK(ctx->d_pair_sh.reserve(npair));
        HIPCHK(ctx->d_pair_Q.reserve(npair));
        HIPCHK(ctx->d_prims_f.reserve(prims.size()));
        if (!prims.empty()) {
            HIPCHK(hipMemcpy(ctx->d_prims.p, prims.data(), sizeof(PrimPair) * prims.size(), hipMemcpyHostToDevice));
            std::vector<BraPair<float>> pf(prims.size());
            for (size_t i = 0; i < prims.size(); ++i) {
                const PrimPair& q = prims[i];
                BraPair<float>& f = pf[i];
                f.p = (float)q.p;
                f.inv2p = (float)q.inv2p;
                f.K = (float)q.K;
                for (int k = 0; k < 3; ++k) {
                    f.P[k] = (float)q.P[k];
                    f.PA[k] = (float)q.PA[k];
                }
            }
            HIPCHK(hipMemcpy(ctx->d_prims_f.p, pf.data(), sizeof(BraPair<float>) * pf.size(), hipMemcpyHostToDevice));
        }
        HIPCHK(hipMemcpy(ctx->d_pair_off.p, ctx->pair_off.data(), sizeof(int) * npair, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(ctx->d_pair_np.p, ctx->pair_np.data(), sizeof(int) * npair, hipMemcpyHostToDevice));
        HIPCHK(hipMemcpy(ctx->d_pair_sh.p, ctx->pair_sh.data(), sizeof(int2) * npair, hipMemcpyHostToDevice));
        // primitive bounds once they exist (the first upload precedes them)
        HIPCHK(ctx->d_prim_q.reserve(std::max<size_t>(prims.size(), 1)));
        std::vector<float> q;
        for (int i = 0; i < npair; ++i) q.insert(q.end(), pairs[i].q.begin(), pairs[i].q.end());
        if (q.size() == prims.size() && !q.empty()) {
            HIPCHK(hipMemcpy(ctx->d_prim_q.p, q.data(), sizeof(float) * q.size(), hipMemcpyHostToDevice));
            std::vector<KetS> ks(prims.size());
            for (size_t i = 0; i < prims.size(); ++i) {
                ks[i].p = prims[i].p;
                ks[i].K = prims[i].K;
                for (int k = 0; k < 3; ++k) ks[i].P[k] = prims[i].P[k];
                ks[i].q = q[i];
                ks[i].pad = 0.0f;
            }
            HIPCHK(ctx->d_kets_s.reserve(ks.size()));
            HIPCHK(hipMemcpy(ctx->d_kets_s.p, ks.data(), sizeof(KetS) * ks.size(), hipMemcpyHostToDevice));
            std::vector<KetSf> ksf(prims.size());
            for (size_t i = 0; i < prims.size(); ++i) {
                ksf[i].p = (float)prims[i].p;
                ksf[i].K = (float)prims[i].K;
                for (int k = 0; k < 3; ++k) ksf[i].P[k] = (float)prims[i].P[k];
                ksf[i].q = q[i];
            }
            HIPCHK(ctx->d_kets_sf.reserve(ksf.size()));
            HIPCHK(hipMemcpy(ctx->d_kets_sf.p, ksf.data(), sizeof(KetSf) * ksf.size(), hipMemcpyHostToDevice));
        }
        return M2_OK;
    };
    m2_status st = upload();
    if (st) return st;
    // Schwarz bounds on the device, all pairs, FP64
    const int npair = (int)pairs.size();
    std::vector<double> Q(npair, 0.0);
    {
        JKArgs a = jk_args(ctx);
        int i0 = 0;
        while (i0 < npair) {
            int c = pairs[i0].cls, i1 = i0;
            while (i1 < npair && pairs[i1].cls == c) ++i1;
            if (SW_LAUNCH[c]) {
                SW_LAUNCH[c](b.pure, a, tables(ctx), i0, i1 - i0, ctx->d_pair_Q.p, ctx->stream);
                HIPCHK(hipGetLastError());
            } else {
                HIPCHK(launch_schwarz_heavy(ctx->device, PAIR_L[c][0], PAIR_L[c][1], b.pure, a, tables(ctx), i0,
                                            i1 - i0, ctx->d_pair_Q.p, ctx->stream));
            }
            i0 = i1;
        }
        HIPCHK(hipMemcpyAsync(Q.data(), ctx->d_pair_Q.p, sizeof(double) * npair, hipMemcpyDeviceToHost, ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
    }
    ctx->Qsh.assign((size_t)b.nsh * b.nsh, 0.0);
    for (int i = 0; i < npair; ++i) {
        ctx->Qsh[(size_t)pairs[i].a * b.nsh + pairs[i].b] = Q[i];
        ctx->Qsh[(size_t)pairs[i].b * b.nsh + pairs[i].a] = Q[i];
    }
    // Primitive-pair bounds q = sqrt(max |(ab|ab)|) of each primitive pair on
    // its own (contraction coefficients included), by the same Schwarz
    // kernels run over one-primitive "pairs"; by Cauchy-Schwarz every
    // primitive quartet satisfies |[ab|cd]| <= q_ab q_cd.  Each pair's
    // primitives are then ordered by q descending, which lets the J/K loops
    // stop at the first primitive below the screening threshold.
    std::vector<double> qprim;
    {
        const int nprim = npair ? ctx->pair_off[npair - 1] + ctx->pair_np[npair - 1] : 0;
        std::vector<int2> tsh(nprim);
        std::vector<int> toff(nprim), tnp(nprim, 1);
        for (int i = 0; i < npair; ++i)
            for (int k = 0; k < ctx->pair_np[i]; ++k) {
                tsh[ctx->pair_off[i] + k] = ctx->pair_sh[i];
                toff[ctx->pair_off[i] + k] = ctx->pair_off[i] + k;
            }
        DevBuf<int2> d_tsh;
        DevBuf<int> d_toff, d_tnp;
        DevBuf<double> d_tq;
        HIPCHK(d_tsh.reserve(nprim));
        HIPCHK(d_toff.reserve(nprim));
        HIPCHK(d_tnp.reserve(nprim));
        HIPCHK(d_tq.reserve(nprim));
        qprim.assign(nprim, 0.0);
        if (nprim > 0) {
            HIPCHK(hipMemcpy(d_tsh.p, tsh.data(), sizeof(int2) * nprim, hipMemcpyHostToDevice));
            HIPCHK(hipMemcpy(d_toff.p, toff.data(), sizeof(int) * nprim, hipMemcpyHostToDevice));
            HIPCHK(hipMemcpy(d_tnp.p, tnp.data(), sizeof(int) * nprim, hipMemcpyHostToDevice));
            JKArgs a = jk_args(ctx);
            a.pair_sh = d_tsh.p;
            a.pair_off = d_toff.p;
            a.pair_np = d_tnp.p;
            int i0 = 0;
            while (i0 < npair) {
                int c = pairs[i0].cls, i1 = i0;
                while (i1 < npair && pairs[i1].cls == c) ++i1;
                const int p0 = ctx->pair_off[i0], p1 = ctx->pair_off[i1 - 1] + ctx->pair_np[i1 - 1];
                if (p1 > p0) {
                    if (SW_LAUNCH[c]) {
                        SW_LAUNCH[c](b.pure, a, tables(ctx), p0, p1 - p0, d_tq.p, ctx->stream);
                        HIPCHK(hipGetLastError());
                    } else {
                        HIPCHK(launch_schwarz_heavy(ctx->device, PAIR_L[c][0], PAIR_L[c][1], b.pure, a, tables(ctx),
                                                    p0, p1 - p0, d_tq.p, ctx->stream));
                    }
                }
                i0 = i1;
            }
            HIPCHK(hipMemcpyAsync(qprim.data(), d_tq.p, sizeof(double) * nprim, hipMemcpyDeviceToHost, ctx->stream));
            HIPCHK(hipStreamSynchronize(ctx->stream));
        }
        for (int i = 0; i < npair; ++i) {
            P& pr = pairs[i];
            const int o = ctx->pair_off[i];
            std::vector<int> idx(pr.pp.size());
            std::iota(idx.begin(), idx.end(), 0);
            std::stable_sort(idx.begin(), idx.end(), [&](int x, int y) { return qprim[o + x] > qprim[o + y]; });
            std::vector<PrimPair> pp(pr.pp.size());
            pr.q.resize(pr.pp.size());
            for (size_t k = 0; k < idx.size(); ++k) {
                pp[k] = pr.pp[idx[k]];
                pr.q[k] = (float)qprim[o + idx[k]];
            }
            pr.pp.swap(pp);
        }
    }
    // final layout: class, np bucket, Q descending
    std::vector<int> order(npair);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) {
        if (pairs[x].cls != pairs[y].cls) return pairs[x].cls < pairs[y].cls;
        int bx = np_bucket((int)pairs[x].pp.size()), by = np_bucket((int)pairs[y].pp.size());
        if (bx != by) return bx > by;
        return Q[x] > Q[y];
    });
    std::vector<P> sorted;
    std::vector<double> Qs;
    sorted.reserve(npair);
    for (int i : order) {
        sorted.push_back(std::move(pairs[i]));
        Qs.push_back(Q[i]);
    }
    pairs.swap(sorted);
    st = upload();
    if (st) return st;
    ctx->pair_Q = Qs;
    HIPCHK(hipMemcpy(ctx->d_pair_Q.p, Qs.data(), sizeof(double) * npair, hipMemcpyHostToDevice));
    {
        std::vector<PairRec> rec(npair);
        for (int i = 0; i < npair; ++i) {
            PairRec& r = rec[i];
            r = PairRec{};
            r.sc = ctx->pair_sh[i].x;
            r.sd = ctx->pair_sh[i].y;
            r.off = ctx->pair_off[i];
            r.np = ctx->pair_np[i];
            r.Q = Qs[i];
            for (int k = 0; k < 3; ++k) r.CD[k] = b.sh[r.sc].O[k] - b.sh[r.sd].O[k];
            r.fc = b.sh[r.sc].bf;
            r.fd = b.sh[r.sd].bf;
        }
        HIPCHK(ctx->d_pair_rec.reserve(npair));
        HIPCHK(hipMemcpy(ctx->d_pair_rec.p, rec.data(), sizeof(PairRec) * npair, hipMemcpyHostToDevice));
    }
    for (int c = 0; c < NPAIRCLASS; ++c) ctx->cls[c] = PairClassSeg{};
    int i0 = 0;
    while (i0 < npair) {
        int c = pairs[i0].cls;
        PairClassSeg& seg = ctx->cls[c];
        seg.begin = i0;
        int i1 = i0;
        while (i1 < npair && pairs[i1].cls == c) {
            int bk = np_bucket((int)pairs[i1].pp.size()), g0 = i1;
            while (i1 < npair && pairs[i1].cls == c && np_bucket((int)pairs[i1].pp.size()) == bk) ++i1;
            seg.groups.push_back(PairGroup{g0, i1});
        }
        seg.end = i1;
        i0 = i1;
    }
    st = build_families(ctx);
    if (st) return st;
    // device work-list buffers of every class combination (sizes depend on
    // the basis only)
    ctx->max_nbra = 0;
    for (int X = 0, ci = 0; X < NPAIRCLASS; ++X)
        for (int Y = 0; Y <= X; ++Y, ++ci) {
            JKList& L = ctx->lists[ci];
            L = JKList{};
            const bool reg = JK_COMBO[X][Y] != nullptr;
            const bool fam = reg && ctx->fam_on && fam_combo(X, Y, ctx->fam_wide);
            const PairClassSeg &sx = ctx->cls[X], &sy = fam ? ctx->fam_seg : ctx->cls[Y];
            if (sx.end <= sx.begin || sy.end <= sy.begin) continue;
            ARGCHK((int)sy.groups.size() <= JK_MAXG, "m2_basis_set: too many primitive-count groups");
            L.bra0 = sx.begin;
            L.nbra = sx.end - sx.begin;
            L.same = X == Y;
            L.Qk = fam ? ctx->d_fam_Q.p : ctx->d_pair_Q.p;
            L.bra_fam = fam && X == Y ? ctx->d_pair_fam.p : nullptr;
            L.ngrp = (int)sy.groups.size();
            L.kc = reg ? 64 : JK_HEAVY_KETS;
            // the split and table-driven classes (l_a + l_b + l_c + l_d >= 6)
            // stay in FP64 in mixed builds: their few quartets gain less from
            // FP32 than a second launch costs
            L.fp32 = PAIR_L[X][0] + PAIR_L[X][1] + PAIR_L[Y][0] + PAIR_L[Y][1] < 6 ? 1 : 0;
            long long maxch = 0;   // chunks of the bra with the most kets
            for (int g = 0; g < L.ngrp; ++g) {
                L.gbeg[g] = sy.groups[g].begin;
                L.gend[g] = sy.groups[g].end;
                maxch += (sy.groups[g].end - sy.groups[g].begin + L.kc - 1) / L.kc;
            }
            // capacity: every bra at its largest chunk count, in items of
            // bt_min chunks (16 for the register classes, 1 for the heavy
            // ones), more when the list would otherwise pass
            // JK_LIST_MAX_ITEMS items
            L.bt_min = (int)std::max<long long>(reg ? JK_BLOCK_TASKS_MIN : 1,
                                                ((long long)L.nbra * maxch + JK_LIST_MAX_ITEMS - 1) / JK_LIST_MAX_ITEMS);
            const long long cap = (long long)L.nbra * ((maxch + L.bt_min - 1) / L.bt_min);
            ARGCHK(cap <= 2LL * JK_LIST_MAX_ITEMS, "m2_basis_set: work list too large");
            L.cap = (int)cap;
            L.bt = std::max(L.bt_min, ctx->tune_bt[ci] > 0 ? ctx->tune_bt[ci] : reg ? JK_BLOCK_TASKS : L.bt_min);
            DevBuf<int2>& cb = ctx->list_cut[ci];
            DevBuf<int>& ob = ctx->list_off[ci];
            DevBuf<int>& bb = ctx->list_bra[ci];
            HIPCHK(cb.reserve((size_t)L.nbra * JK_MAXG));
            HIPCHK(ob.reserve(2 * (size_t)(L.nbra + 1)));
            HIPCHK(bb.reserve(2 * (size_t)std::max(L.cap, 1)));
            L.cut = cb.p;
            L.off = ob.p;
            L.bra_of = bb.p;
            ctx->max_nbra = std::max(ctx->max_nbra, L.nbra);
        }
    return upload_jklists(ctx);
}

// kets per task of the gradient's program classes (split over workgroups)
constexpr int GRAD_PROG_CHUNK = GRAD_PROG_SPLIT * GRAD_PROG_KETS;

// gradient classes with generated register code (one quartet per lane)
static bool deriv_register(int la, int lb, int lc, int ld) {
    for (int k = 0; k < DERIV_REG_N; ++k)
        if (DERIV_REG_L[k][0] == la && DERIV_REG_L[k][1] == lb && DERIV_REG_L[k][2] == lc && DERIV_REG_L[k][3] == ld)
            return true;
    return false;
}

// Device memory one direct J/K build works in: the shell-block norms, the
// unsymmetrised J/K accumulators and the work-list totals (then max|D| as a
// double).  Either the context's own buffers (grown on demand) or, for the
// allocation-free entry point, fixed slices of the static scratch array
// (nothing is allocated).
struct JKWork {
    double* Dn = nullptr;
    double* J = nullptr;
    double* K = nullptr;
    int* lists = nullptr;
    bool fixed = false;
};

// The arguments of one build.  dScreen (optional): the density whose
// shell-block norms screen the quartets (UnrestrictedFock::formJKdirect
// screens both spin builds with Da + Db).  dD2 (optional): a second density
// digested from the same integrals into dJ2, dK2 (the unrestricted build).
struct JKCall {
    const double* dD;
    double multiplier, precision;
    unsigned flags;
    double* dJ;
    double* dK;
    const double* dScreen;
    const double* dD2;
    double* dJ2;
    double* dK2;
};

// class combinations launched by one build, in class order
struct JKLaunch {
    int X, Y, ci;
};

// On the device: shell-block norms and max|D|, then every class
// combination's work list (three launches for all of them), then one class
// kernel per combination and precision, then the symmetrisation.  By
// default the host reads nothing back: each class kernel's grid is the
// list's bound (a function of the basis), and workgroups past the total the
// list kernels wrote exit at once, so the whole build can also be captured
// into a graph.  With host_totals the totals are read back first (a stream
// synchronisation) and each grid is the exact item count.
m2_status jk_enqueue(m2_ctx* ctx, const JKCall& c, const JKWork& w, bool prof, std::vector<JKLaunch>& plan) {
    BasisData& b = ctx->obs;
    const int n = b.nbf, nsh = b.nsh;
    const bool mixed = c.flags == M2_MIXED;
    const int nden = c.dD2 ? 2 : 1;
    const bool bounded = !ctx->host_totals;
    launch_blocknorm(c.dScreen ? c.dScreen : c.dD, n, nsh, b.d_bf.p, b.d_nbf.p, w.Dn, ctx->stream);
    double* d_dmax = reinterpret_cast<double*>(w.lists + 2 * NCOMBO);
    launch_dmax(w.Dn, nsh * nsh, d_dmax, ctx->stream);
    launch_task_lists(ctx->d_jklists.p, NCOMBO, ctx->max_nbra, ctx->d_pair_Q.p, d_dmax, c.precision,
                      ctx->fp32_below, mixed ? 1 : 0, w.lists, ctx->stream);
    HIPCHK(hipGetLastError());
    if (!bounded) {
        HIPCHK(hipMemcpyAsync(ctx->h_lists, w.lists, sizeof(int) * (2 * NCOMBO + 2), hipMemcpyDeviceToHost,
                              ctx->stream));
        HIPCHK(hipStreamSynchronize(ctx->stream));
    }
    // launch plan: every combination with bra pairs (with host-read totals:
    // those with work items)
    plan.clear();
    long long units = 0;
    for (int X = 0, ci = 0; X < NPAIRCLASS; ++X)
        for (int Y = 0; Y <= X; ++Y, ++ci) {
            const JKList& L = ctx->lists[ci];
            if (L.nbra <= 0 || L.bound <= 0) continue;
            if (!bounded) {
                const int t64 = ctx->h_lists[2 * ci], t32 = ctx->h_lists[2 * ci + 1];
                units += t64 + t32;
                if (t64 + t32 == 0) continue;
            }
            plan.push_back(JKLaunch{X, Y, ci});
        }
    ctx->last_tasks = bounded ? -1 : units;
    ctx->last_tot = bounded ? w.lists : nullptr;
    launch_zero(w.J, (size_t)nden * n * n, ctx->stream);
    launch_zero(w.K, (size_t)nden * n * n, ctx->stream);
    JKArgs a = jk_args(ctx);
    a.Dn = w.Dn;
    a.counters = nullptr;
    if (prof) {
        HIPCHK(ctx->d_counters.reserve(108));
        launch_zero((double*)ctx->d_counters.p, 108, ctx->stream);
        if (ctx->prof_counters) a.counters = ctx->d_counters.p;
        while (ctx->ev.size() < 2 * plan.size() + 2) {
            hipEvent_t e;
            HIPCHK(hipEventCreate(&e));
            ctx->ev.push_back(e);
        }
    }
    a.D = c.dD;
    a.J = w.J;
    a.K = w.K;
    a.nden = nden;
    a.D2 = c.dD2;
    a.J2 = nden == 2 ? w.J + (size_t)n * n : nullptr;
    a.K2 = nden == 2 ? w.K + (size_t)n * n : nullptr;
    a.precision = c.precision;
    a.fp32_below = ctx->fp32_below;
    a.scale = c.multiplier;
    a.prim_tau = c.precision * ctx->prim_screen;
    const double prim_tau64 = a.prim_tau, prim_tau32 = c.precision * ctx->prim_screen32;
    for (size_t pi = 0; pi < plan.size(); ++pi) {
        const JKLaunch& pl = plan[pi];
        const JKList& L = ctx->lists[pl.ci];
        a.combo = 6 * pl.X + pl.Y;
        if (prof) HIPCHK(hipEventRecord(ctx->ev[2 * pi], ctx->stream));
        a.fam_kets = ctx->fam_on && fam_combo(pl.X, pl.Y, ctx->fam_wide) ? 1 : 0;
        a.cut = L.cut;
        a.bra0 = L.bra0;
        a.ngrp = L.ngrp;
        for (int g = 0; g < JK_MAXG; ++g) a.gbeg[g] = L.gbeg[g];
        a.block_tasks = L.bt;
        // FP64 heads, then FP32 tails (only mixed builds have tails)
        for (int tail = 0; tail < (mixed ? 2 : 1); ++tail) {
            int nblk;
            if (bounded) {
                nblk = L.bound;
                a.blk_total = w.lists + 2 * pl.ci + tail;
            } else {
                nblk = ctx->h_lists[2 * pl.ci + tail];
                if (nblk <= 0) continue;
                a.blk_total = nullptr;
            }
            a.tail = tail;
            a.prim_tau = tail ? prim_tau32 : prim_tau64;
            a.blk_off = L.off + tail * (L.nbra + 1);
            a.blk_bra = L.bra_of + (size_t)tail * L.cap;
            if (JK_COMBO[pl.X][pl.Y]) {
                JK_COMBO[pl.X][pl.Y](b.pure, a, tables(ctx), nblk, tail ? 2 : 0, ctx->stream);
                HIPCHK(hipGetLastError());
            } else {
                HIPCHK(launch_jk_heavy(ctx->device, PAIR_L[pl.X][0], PAIR_L[pl.X][1], PAIR_L[pl.Y][0],
                                       PAIR_L[pl.Y][1], b.pure, a, tables(ctx), nblk, tail ? 2 : 0, ctx->stream));
            }
        }
        if (prof) HIPCHK(hipEventRecord(ctx->ev[2 * pi + 1], ctx->stream));
    }
    launch_symmetrize(w.J, w.K, c.dJ, c.dK, n, ctx->stream);
    if (nden == 2) launch_symmetrize(w.J + (size_t)n * n, w.K + (size_t)n * n, c.dJ2, c.dK2, n, ctx->stream);
    HIPCHK(hipGetLastError());
    return M2_OK;
}

// per-combination kernel times and counters of a profiled build (after its stream synchronisation)
m2_status jk_profile_readout(m2_ctx* ctx, const std::vector<JKLaunch>& plan) {
    unsigned long long cnt[108];
    HIPCHK(hipMemcpy(cnt, ctx->d_counters.p, sizeof(cnt), hipMemcpyDeviceToHost));
    for (size_t pi = 0; pi < plan.size(); ++pi) {
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, ctx->ev[2 * pi], ctx->ev[2 * pi + 1]));
        const int c = 6 * plan[pi].X + plan[pi].Y;
        ctx->prof_ms[c] += ms;
        ctx->prof_launch[c] += 1;
    }
    for (int c = 0; c < 36; ++c) {
        ctx->prof_q[c] += (long long)cnt[3 * c];
        ctx->prof_pq[c] += (long long)cnt[3 * c + 1];
        ctx->prof_cpq[c] += (long long)cnt[3 * c + 2];
    }
    return M2_OK;
}

// Graph replay (m2_set_jk_graphs): a build is captured once per key -- the
// call's pointers and scalars, the context's settings that enter the kernel
// arguments, its work buffers, and a generation counter bumped by every
// change of basis or launch shape -- and later builds with the same key
// launch the instantiated graph.  The capture holds only kernel launches
// (the accumulators are cleared by a fill kernel, not a memset node).
bool jk_graph_key_eq(const JKGraphKey& x, const JKGraphKey& y) { return std::memcmp(&x, &y, sizeof(x)) == 0; }

m2_status jk_graph_run(m2_ctx* ctx, const JKCall& c, const JKWork& w) {
    JKGraphKey k;
    std::memset(&k, 0, sizeof(k));   // padding too: keys compare bytewise
    k.dD = c.dD;
    k.dJ = c.dJ;
    k.dK = c.dK;
    k.dScreen = c.dScreen;
    k.dD2 = c.dD2;
    k.dJ2 = c.dJ2;
    k.dK2 = c.dK2;
    k.multiplier = c.multiplier;
    k.precision = c.precision;
    k.flags = c.flags;
    k.fp32_below = ctx->fp32_below;
    k.prim_screen = ctx->prim_screen;
    k.prim_screen32 = ctx->prim_screen32;
    k.Dn = w.Dn;
    k.J = w.J;
    k.K = w.K;
    k.lists = w.lists;
    k.gen = ctx->graph_gen;
    for (size_t i = 0; i < ctx->graphs.size(); ++i) {
        if (!jk_graph_key_eq(ctx->graphs[i].key, k)) continue;
        std::rotate(ctx->graphs.begin(), ctx->graphs.begin() + i, ctx->graphs.begin() + i + 1);   // most recent first
        HIPCHK(hipGraphLaunch(ctx->graphs[0].exec, ctx->stream));
        ctx->last_tasks = -1;
        ctx->last_tot = w.lists;
        ++ctx->graph_replays;
        return M2_OK;
    }
    // capture (nothing runs), instantiate, then launch
    std::vector<JKLaunch> plan;
    HIPCHK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    const m2_status st = jk_enqueue(ctx, c, w, false, plan);
    hipGraph_t g = nullptr;
    const hipError_t ec = hipStreamEndCapture(ctx->stream, &g);
    if (st != M2_OK) {
        if (g) (void)hipGraphDestroy(g);
        return st;
    }
    if (ec != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        ctx->err = std::string("m2_jk_direct: graph capture: ") + hipGetErrorString(ec);
        return M2_ERR_HIP;
    }
    hipGraphExec_t ex = nullptr;
    const hipError_t ei = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ei != hipSuccess) {
        ctx->err = std::string("m2_jk_direct: graph instantiation: ") + hipGetErrorString(ei);
        return M2_ERR_HIP;
    }
    if (ctx->graphs.size() >= JK_GRAPH_CACHE) {
        (void)hipGraphExecDestroy(ctx->graphs.back().exec);
        ctx->graphs.pop_back();
    }
    ctx->graphs.insert(ctx->graphs.begin(), JKGraph{k, ex});
    ++ctx->graph_captures;
    HIPCHK(hipGraphLaunch(ex, ctx->stream));
    return M2_OK;
}

// One direct J/K build, enqueued on the context's stream and waited for.
m2_status jk_device_core(m2_ctx* ctx, const double* dD, double multiplier, double precision, unsigned flags,
                         double* dJ, double* dK, const double* dScreen = nullptr, const double* dD2 = nullptr,
                         double* dJ2 = nullptr, double* dK2 = nullptr, const JKWork* fixed = nullptr) {
    BasisData& b = ctx->obs;
    const int n = b.nbf, nsh = b.nsh;
    ARGCHK(n > 0, "m2_jk_direct: no orbital basis set");
    ARGCHK(flags <= M2_MIXED, "m2_jk_direct: unknown flags");
    const int nden = dD2 ? 2 : 1;
    JKWork w;
    if (fixed) {
        w = *fixed;
    } else {
        HIPCHK(ctx->d_Dn.reserve((size_t)nsh * nsh));
        HIPCHK(ctx->d_J.reserve((size_t)nden * n * n));
        HIPCHK(ctx->d_K.reserve((size_t)nden * n * n));
        HIPCHK(ctx->d_lists.reserve(2 * NCOMBO + 2));
        w.Dn = ctx->d_Dn.p;
        w.J = ctx->d_J.p;
        w.K = ctx->d_K.p;
        w.lists = ctx->d_lists.p;
    }
    const JKCall c{dD, multiplier, precision, flags, dJ, dK, dScreen, dD2, dJ2, dK2};
    // profiling creates events and a counter buffer: not in the fixed-memory build
    const bool prof = ctx->prof && !w.fixed;
    // (the allocation-free build never instantiates a graph: that allocates)
    if (ctx->use_graphs && !prof && !ctx->host_totals && !w.fixed) {
        const m2_status st = jk_graph_run(ctx, c, w);
        if (st) return st;
        HIPCHK(hipStreamSynchronize(ctx->stream));
        return M2_OK;
    }
    std::vector<JKLaunch> plan;
    const m2_status st = jk_enqueue(ctx, c, w, prof, plan);
    if (st) return st;
    HIPCHK(hipStreamSynchronize(ctx->stream));
    if (prof) return jk_profile_readout(ctx, plan);
    return M2_OK;
}

}  // namespace

namespace m2 {
struct ComboRangeG {
    int X, Y, t0, tmid, t1;
};
// The gradient's task list, built on the device into ctx->d_tasks (same
// tasks and order as build_tasks with gradient = true): per combination the
// tasks per bra are counted, one exclusive scan over all combinations gives
// the offsets, and each bra writes its tasks.  Only the scan (one int per bra
// and combination) comes back to the host, for the launch ranges.
hipError_t grad_tasks_device(m2_ctx* ctx, double precision, double Dmax, bool mixed, std::vector<ComboRangeG>& out) {
    out.clear();
    std::vector<GradTaskArgs> args;   // per combination: the FP64 head (part 0), then the FP32 tail (part 1)
    std::vector<int> start;           // first counter of each entry of args
    int n = 0;
    for (int X = 0; X < NPAIRCLASS; ++X) {
        const PairClassSeg& sx = ctx->cls[X];
        for (int Y = 0; Y <= X; ++Y) {
            const PairClassSeg& sy = ctx->cls[Y];
            if (sx.end <= sx.begin || sy.groups.empty()) continue;
            if ((int)sy.groups.size() > JK_MAXG) return hipErrorInvalidValue;
            const bool reg = deriv_register(PAIR_L[X][0], PAIR_L[X][1], PAIR_L[Y][0], PAIR_L[Y][1]);
            GradTaskArgs g{};
            g.Q = ctx->d_pair_Q.p;
            g.x0 = sx.begin;
            g.nbra = sx.end - sx.begin;
            g.same = X == Y;
            g.ngrp = (int)sy.groups.size();
            for (int k = 0; k < g.ngrp; ++k) {
                g.gbeg[k] = sy.groups[k].begin;
                g.gend[k] = sy.groups[k].end;
            }
            g.chunk = reg ? 64 : GRAD_PROG_CHUNK;
            g.precision = precision;
            g.dmax = Dmax;
            g.split = mixed ? std::max(precision, ctx->fp32_below) : precision;
            for (int part = 0; part < 2; ++part) {
                g.part = part;
                args.push_back(g);
                start.push_back(n);
                n += g.nbra;
            }
            out.push_back(ComboRangeG{X, Y, 0, 0, 0});
        }
    }
    if (n == 0) return hipSuccess;
    DevBuf<int> counts, offs;
    hipError_t e;
    if ((e = counts.reserve(n)) != hipSuccess || (e = offs.reserve((size_t)n + 1)) != hipSuccess) return e;
    for (size_t c = 0; c < args.size(); ++c) launch_grad_task_count(args[c], counts.p + start[c], ctx->stream);
    launch_scan_excl(counts.p, offs.p, n, ctx->stream);
    std::vector<int> h((size_t)n + 1);
    if ((e = hipMemcpyAsync(h.data(), offs.p, sizeof(int) * ((size_t)n + 1), hipMemcpyDeviceToHost, ctx->stream)) !=
            hipSuccess ||
        (e = hipStreamSynchronize(ctx->stream)) != hipSuccess)
        return e;
    const int total = h[n];
    for (size_t c = 0; c < out.size(); ++c) {
        out[c].t0 = h[start[2 * c]];
        out[c].tmid = h[start[2 * c + 1]];
        out[c].t1 = h[start[2 * c + 1] + args[2 * c + 1].nbra];
    }
    if (total > 0) {
        if ((e = ctx->d_tasks.reserve(total)) != hipSuccess) return e;
        for (size_t c = 0; c < args.size(); ++c)
            launch_grad_task_fill(args[c], offs.p + start[c], ctx->d_tasks.p, ctx->stream);
        if ((e = hipGetLastError()) != hipSuccess) return e;
    }
    // the buffers are freed once the fill kernels have run
    if ((e = hipStreamSynchronize(ctx->stream)) != hipSuccess) return e;
    counts.release();
    offs.release();
    std::vector<ComboRangeG> kept;
    for (const auto& r : out)
        if (r.t1 > r.t0) kept.push_back(r);
    out.swap(kept);
    return hipSuccess;
}

}  // namespace m2

namespace {

// Device memory of m2_jk_direct_device_noalloc: a file-scope array, one copy
// per GPU the library runs on, held until the process ends.  Host code never
// uses its name as an address (that is a host-side stand-in); each device's
// copy is reached through the address hipGetSymbolAddress returns with that
// device current, kept per device in g_scratch.
static __device__ __align__(256) unsigned char scratch[M2_NOALLOC_SCRATCH_BYTES];

std::mutex g_scratch_mu;
std::map<int, unsigned char*> g_scratch;

// once per device, with `device` current
hipError_t scratch_address(int device, unsigned char** out) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    auto it = g_scratch.find(device);
    if (it == g_scratch.end()) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        if (!p) return hipErrorInvalidSymbol;
        it = g_scratch.emplace(device, static_cast<unsigned char*>(p)).first;
    }
    *out = it->second;
    return hipSuccess;
}

size_t align256(size_t x) { return (x + 255) & ~size_t(255); }

}  // namespace

extern "C" {

m2_status m2_ctx_create(m2_ctx** out, int device) {
    if (!out) return M2_ERR_ARG;
    *out = nullptr;
    m2_ctx* ctx = new m2_ctx();
    ctx->device = device;
    hipError_t e = hipSetDevice(device);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        std::fprintf(stderr, "m2_ctx_create: %s\n", hipGetErrorString(e));
        delete ctx;
        return M2_ERR_HIP;
    }
    if (hipHostMalloc((void**)&ctx->h_lists, sizeof(int) * (2 * NCOMBO + 2)) != hipSuccess) {
        std::fprintf(stderr, "m2_ctx_create: pinned allocation failed\n");
        m2_ctx_destroy(ctx);
        return M2_ERR_HIP;
    }
    if ((e = scratch_address(device, &ctx->scratch)) != hipSuccess) {
        std::fprintf(stderr, "m2_ctx_create: hipGetSymbolAddress of the scratch array: %s\n", hipGetErrorString(e));
        m2_ctx_destroy(ctx);
        return M2_ERR_HIP;
    }
    if (heavy_upload(device) != hipSuccess) {
        std::fprintf(stderr, "m2_ctx_create: program table upload failed\n");
        m2_ctx_destroy(ctx);
        return M2_ERR_HIP;
    }
    // Boys tables, F_m(k * 0.1) for m < BOYS_MTAB
    std::vector<double> tab((size_t)BOYS_NGRID * BOYS_MTAB);
    std::vector<float> tabf(tab.size());
    for (int k = 0; k < BOYS_NGRID; ++k) boys_host(BOYS_MTAB - 1, k * BOYS_STEP, &tab[(size_t)k * BOYS_MTAB]);
    for (size_t i = 0; i < tab.size(); ++i) tabf[i] = (float)tab[i];
    // compact rows {F_(M+7), e^-T} of the low-m classes (same grid)
    std::vector<double> tabw(BOYS_WSIZE);
    for (int M = 0; M <= BOYS_DIRECT_MAX; ++M)
        for (int k = 0; k <= BOYS_WK[M]; ++k) {
            tabw[boys_woff(M) + (size_t)k * 2] = tab[(size_t)k * BOYS_MTAB + M + 7];
            tabw[boys_woff(M) + (size_t)k * 2 + 1] = std::exp(-(k * BOYS_STEP));
        }
    std::vector<float> tabfw(BOYS_FWSIZE);
    for (int M = 0; M <= BOYS_FW_MAX; ++M)
        for (int k = 0; k < BOYS_NGRID; ++k) {
            tabfw[((size_t)M * BOYS_NGRID + k) * 2] = tabf[(size_t)k * BOYS_MTAB + M + 4];
            tabfw[((size_t)M * BOYS_NGRID + k) * 2 + 1] = (float)std::exp(-(k * BOYS_STEP));
        }
    if (ctx->d_boys.reserve(tab.size()) != hipSuccess || ctx->d_boys_f.reserve(tab.size()) != hipSuccess ||
        ctx->d_boys_w.reserve(tabw.size()) != hipSuccess || ctx->d_boys_fw.reserve(tabfw.size()) != hipSuccess ||
        hipMemcpy(ctx->d_boys_fw.p, tabfw.data(), sizeof(float) * tabfw.size(), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(ctx->d_boys.p, tab.data(), sizeof(double) * tab.size(), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(ctx->d_boys_f.p, tabf.data(), sizeof(float) * tab.size(), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(ctx->d_boys_w.p, tabw.data(), sizeof(double) * tabw.size(), hipMemcpyHostToDevice) != hipSuccess) {
        std::fprintf(stderr, "m2_ctx_create: device allocation failed\n");
        m2_ctx_destroy(ctx);
        return M2_ERR_HIP;
    }
    *out = ctx;
    return M2_OK;
}

void m2_ctx_destroy(m2_ctx* ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    drop_graphs(ctx);
    local_df_release(ctx);
    for (BasisData* b : {&ctx->obs, &ctx->aux}) {
        b->d_O.release();
        b->d_bf.release();
        b->d_nbf.release();
    }
    ctx->d_prims.release();
    ctx->d_prims_f.release();
    ctx->d_kets_s.release();
    ctx->d_pair_rec.release();
    ctx->d_kets_sf.release();
    ctx->d_pair_off.release();
    ctx->d_pair_np.release();
    ctx->d_pair_sh.release();
    ctx->d_pair_Q.release();
    ctx->d_boys.release();
    ctx->d_boys_f.release();
    ctx->d_boys_w.release();
    ctx->d_boys_fw.release();
    ctx->d_D.release();
    ctx->d_J.release();
    ctx->d_K.release();
    ctx->d_Jo.release();
    ctx->d_Ko.release();
    ctx->d_Dn.release();
    ctx->d_tasks.release();
    ctx->d_jklists.release();
    for (int ci = 0; ci < NCOMBO; ++ci) {
        ctx->list_cut[ci].release();
        ctx->list_off[ci].release();
        ctx->list_bra[ci].release();
        ctx->lists[ci] = JKList{};
    }
    ctx->d_lists.release();
    if (ctx->h_lists) (void)hipHostFree(ctx->h_lists);
    ctx->h_lists = nullptr;
    ctx->d_counters.release();
    for (auto* x : {&ctx->d_B64, &ctx->d_dfV, &ctx->d_dfC, &ctx->d_dfX, &ctx->d_dfG, &ctx->d_dfJ, &ctx->d_dfK, &ctx->d_dfKp})
        x->release();
    ctx->d_df_jblk.release();
    for (auto* x : {&ctx->d_B32, &ctx->d_dfC32, &ctx->d_dfX32, &ctx->d_dfKp32})
        x->release();
    for (auto* x : {&ctx->d_aux_off, &ctx->d_aux_np, &ctx->d_aux_lists}) x->release();
    ctx->d_aux_prims.release();
    la_release(ctx);
    for (hipEvent_t e : ctx->ev) (void)hipEventDestroy(e);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

const char* m2_last_error(const m2_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

void* m2_ctx_stream(m2_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

int m2_ctx_device(const m2_ctx* ctx) { return ctx ? ctx->device : -1; }

// Empty the basis `which` (and for the orbital basis its shell pairs), so
// that every later call needing it returns M2_ERR_ARG.
static void clear_basis(m2_ctx* ctx, int which) {
    BasisData& b = (which == M2_ORBITAL) ? ctx->obs : ctx->aux;
    b.nsh = 0;
    b.nbf = 0;
    b.sh.clear();
    ctx->df_ready = false;
    if (which != M2_ORBITAL) return;
    ctx->pair_sh.clear();
    ctx->pair_off.clear();
    ctx->pair_np.clear();
    ctx->pair_Q.clear();
    ctx->Qsh.clear();
    for (auto& c : ctx->cls) c = PairClassSeg{};
    for (auto& L : ctx->lists) L = JKList{};
    ctx->fam_on = false;
    ctx->fam_seg = PairClassSeg{};
}

// Every input is checked before the context changes.  A call that fails
// (bad input, device upload, pair construction) leaves the basis empty
// (nbf = nsh = 0): later calls that need it return M2_ERR_ARG instead of
// mixing new shell arrays with old pair data or old sizes.
m2_status m2_basis_set(m2_ctx* ctx, int which, int nsh, const int* l, const int* nprim, const double* exps,
                       const double* coefs, const double* centers, const int* atoms, int pure) {
    if (!ctx) return M2_ERR_ARG;
    ARGCHK(which == M2_ORBITAL || which == M2_AUXILIARY, "m2_basis_set: which must be M2_ORBITAL or M2_AUXILIARY");
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    local_df_release(ctx);
    if (which == M2_ORBITAL) drop_graphs(ctx);
    std::vector<HostShell> shells;
    auto fail = [&](const char* msg) {
        clear_basis(ctx, which);
        ctx->err = msg;
        return M2_ERR_ARG;
    };
    if (!(nsh > 0 && l && nprim && exps && coefs && centers)) return fail("m2_basis_set: null or empty input");
    const int lmax = which == M2_ORBITAL ? 2 : 3;
    shells.resize(nsh);
    int off = 0, bf = 0;
    for (int s = 0; s < nsh; ++s) {
        HostShell& S = shells[s];
        if (nprim[s] <= 0) return fail("m2_basis_set: every shell needs nprim > 0");
        if (l[s] < 0 || l[s] > lmax)
            return fail(which == M2_ORBITAL ? "m2_basis_set: orbital shells must have l <= 2"
                                            : "m2_basis_set: auxiliary shells must have l <= 3");
        S.l = l[s];
        S.atom = atoms ? atoms[s] : -1;
        for (int k = 0; k < 3; ++k) S.O[k] = centers[3 * s + k];
        S.a.assign(exps + off, exps + off + nprim[s]);
        S.c.assign(coefs + off, coefs + off + nprim[s]);
        for (double x : S.a)
            if (!(x > 0.0)) return fail("m2_basis_set: exponents must be positive");
        off += nprim[s];
        normalise(S);
        S.nbf = pure ? 2 * S.l + 1 : (S.l + 1) * (S.l + 2) / 2;
        S.bf = bf;
        bf += S.nbf;
    }
    clear_basis(ctx, which);
    HIPCHK(hipSetDevice(ctx->device));
    BasisData& b = (which == M2_ORBITAL) ? ctx->obs : ctx->aux;
    b.pure = pure ? 1 : 0;
    b.sh.swap(shells);
    b.nsh = nsh;
    b.nbf = bf;
    m2_status st = upload_basis(ctx, b);
    if (st == M2_OK && which == M2_ORBITAL) st = build_pairs(ctx);
    if (st != M2_OK) {
        const std::string why = ctx->err;
        clear_basis(ctx, which);
        ctx->err = why;
    }
    return st;
}

m2_status m2_basis_info(m2_ctx* ctx, int which, int* nbf, int* nsh) {
    if (!ctx) return M2_ERR_ARG;
    BasisData& b = (which == M2_ORBITAL) ? ctx->obs : ctx->aux;
    if (nbf) *nbf = b.nbf;
    if (nsh) *nsh = b.nsh;
    return M2_OK;
}

m2_status m2_shell_coefs(m2_ctx* ctx, int which, int shell, double* out) {
    if (!ctx) return M2_ERR_ARG;
    BasisData& b = (which == M2_ORBITAL) ? ctx->obs : ctx->aux;
    ARGCHK(shell >= 0 && shell < b.nsh && out, "m2_shell_coefs: bad shell index");
    std::copy(b.sh[shell].c.begin(), b.sh[shell].c.end(), out);
    return M2_OK;
}

m2_status m2_schwarz(m2_ctx* ctx, double* Q) {
    if (!ctx) return M2_ERR_ARG;
    ARGCHK(Q && ctx->obs.nsh > 0, "m2_schwarz: no orbital basis or null output");
    std::copy(ctx->Qsh.begin(), ctx->Qsh.end(), Q);
    return M2_OK;
}

m2_status m2_shellblock_norm(m2_ctx* ctx, const double* A, double* out) {
    if (!ctx) return M2_ERR_ARG;
    BasisData& b = ctx->obs;
    ARGCHK(A && out && b.nbf > 0, "m2_shellblock_norm: bad arguments");
    HIPCHK(hipSetDevice(ctx->device));
    const size_t n2 = (size_t)b.nbf * b.nbf;
    HIPCHK(ctx->d_D.reserve(n2));
    HIPCHK(ctx->d_Dn.reserve((size_t)b.nsh * b.nsh));
    HIPCHK(hipMemcpyAsync(ctx->d_D.p, A, sizeof(double) * n2, hipMemcpyHostToDevice, ctx->stream));
    launch_blocknorm(ctx->d_D.p, b.nbf, b.nsh, b.d_bf.p, b.d_nbf.p, ctx->d_Dn.p, ctx->stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(out, ctx->d_Dn.p, sizeof(double) * b.nsh * b.nsh, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return M2_OK;
}

m2_status m2_one_body(m2_ctx* ctx, int natom, const double* Z, const double* xyz, double* S, double* T, double* V) {
    if (!ctx) return M2_ERR_ARG;
    BasisData& b = ctx->obs;
    ARGCHK(b.nbf > 0, "m2_one_body: no orbital basis set");
    ARGCHK(natom >= 0 && (natom == 0 || (Z && xyz)), "m2_one_body: bad charges");
    HIPCHK(hipSetDevice(ctx->device));
    const int n = b.nbf;
    std::vector<int> np(b.nsh), poff(b.nsh);
    std::vector<double> ex, cf;
    for (int s = 0; s < b.nsh; ++s) {
        np[s] = (int)b.sh[s].a.size();
        poff[s] = (int)ex.size();
        ex.insert(ex.end(), b.sh[s].a.begin(), b.sh[s].a.end());
        cf.insert(cf.end(), b.sh[s].c.begin(), b.sh[s].c.end());
    }
    DevBuf<int> d_l, d_np, d_poff;
    DevBuf<double> d_ex, d_cf, d_Z, d_xyz, d_S, d_T, d_V, d_sph;
    m2_status st = M2_OK;
    auto fail = [&](hipError_t e, const char* what) {
        ctx->err = std::string(what) + ": " + hipGetErrorString(e);
        st = M2_ERR_HIP;
    };
    hipError_t e = hipSuccess;
    if ((e = d_l.reserve(b.nsh)) || (e = d_np.reserve(b.nsh)) || (e = d_poff.reserve(b.nsh)) ||
        (e = d_ex.reserve(ex.size())) || (e = d_cf.reserve(cf.size())) || (e = d_Z.reserve(natom + 1)) ||
        (e = d_xyz.reserve(3 * natom + 3)) || (e = d_S.reserve((size_t)n * n)) || (e = d_T.reserve((size_t)n * n)) ||
        (e = d_V.reserve((size_t)n * n)) || (e = d_sph.reserve(sizeof(SPH_TAB) / sizeof(double)))) {
        fail(e, "m2_one_body: hipMalloc");
    }
    if (!st) {
        (void)hipMemcpy(d_l.p, b.h_l.data(), sizeof(int) * b.nsh, hipMemcpyHostToDevice);
        (void)hipMemcpy(d_np.p, np.data(), sizeof(int) * b.nsh, hipMemcpyHostToDevice);
        (void)hipMemcpy(d_poff.p, poff.data(), sizeof(int) * b.nsh, hipMemcpyHostToDevice);
        (void)hipMemcpy(d_ex.p, ex.data(), sizeof(double) * ex.size(), hipMemcpyHostToDevice);
        (void)hipMemcpy(d_cf.p, cf.data(), sizeof(double) * cf.size(), hipMemcpyHostToDevice);
        if (natom > 0) {
            (void)hipMemcpy(d_Z.p, Z, sizeof(double) * natom, hipMemcpyHostToDevice);
            (void)hipMemcpy(d_xyz.p, xyz, sizeof(double) * 3 * natom, hipMemcpyHostToDevice);
        }
        e = hipMemcpy(d_sph.p, SPH_TAB, sizeof(SPH_TAB), hipMemcpyHostToDevice);
        if (e) fail(e, "m2_one_body: upload");
    }
    if (!st) {
        OneBodyArgs a{};
        a.sh_O = b.d_O.p;
        a.sh_l = d_l.p;
        a.sh_bf = b.d_bf.p;
        a.sh_np = d_np.p;
        a.sh_poff = d_poff.p;
        a.exps = d_ex.p;
        a.coefs = d_cf.p;
        a.nsh = b.nsh;
        a.n = n;
        a.pure = b.pure;
        a.natom = natom;
        a.Z = d_Z.p;
        a.xyz = d_xyz.p;
        a.sph = d_sph.p;
        a.tab = tables(ctx);
        a.S = d_S.p;
        a.T = d_T.p;
        a.V = d_V.p;
        a.want_v = V ? 1 : 0;
        launch_one_body(a, ctx->stream);
        e = hipGetLastError();
        if (!e) e = hipStreamSynchronize(ctx->stream);
        if (!e && S) e = hipMemcpy(S, d_S.p, sizeof(double) * n * n, hipMemcpyDeviceToHost);
        if (!e && T) e = hipMemcpy(T, d_T.p, sizeof(double) * n * n, hipMemcpyDeviceToHost);
        if (!e && V) e = hipMemcpy(V, d_V.p, sizeof(double) * n * n, hipMemcpyDeviceToHost);
        if (e) fail(e, "m2_one_body: kernel");
    }
    for (auto* x : {&d_l, &d_np, &d_poff}) x->release();
    for (auto* x : {&d_ex, &d_cf, &d_Z, &d_xyz, &d_S, &d_T, &d_V, &d_sph}) x->release();
    return st;
}

m2_status m2_jk_direct_device(m2_ctx* ctx, const double* dD, double multiplier, double precision, unsigned flags,
                              double* dJ, double* dK) {
    if (!ctx) return M2_ERR_ARG;
    ARGCHK(dD && dJ && dK, "m2_jk_direct_device: null buffer");
    HIPCHK(hipSetDevice(ctx->device));
    return jk_device_core(ctx, dD, multiplier, precision, flags, dJ, dK);
}

m2_status m2_jk_direct_device_noalloc(m2_ctx* ctx, const double* dD, double multiplier, double precision,
                                      unsigned flags, double* dJ, double* dK) {
    if (!ctx) return M2_ERR_ARG;
    ARGCHK(dD && dJ && dK, "m2_jk_direct_device_noalloc: null buffer");
    ARGCHK(ctx->obs.nbf > 0, "m2_jk_direct_device_noalloc: no orbital basis set");
    if (!ctx->scratch) {
        ctx->err = "m2_jk_direct_device_noalloc: the scratch array has no device address";
        return M2_ERR_STATE;
    }
    const size_t n = ctx->obs.nbf, nsh = ctx->obs.nsh;
    // slices of the scratch array: norms | list totals + max|D| | J | K
    const size_t o_lists = align256(sizeof(double) * nsh * nsh);
    const size_t o_J = o_lists + align256(sizeof(int) * (2 * NCOMBO + 2));
    const size_t o_K = o_J + align256(sizeof(double) * n * n);
    const size_t o_end = o_K + align256(sizeof(double) * n * n);
    if (o_end > M2_NOALLOC_SCRATCH_BYTES) {
        ctx->err = "m2_jk_direct_device_noalloc: " + std::to_string(n) + " basis functions need " +
                   std::to_string(o_end) + " bytes of scratch; the static buffer holds " +
                   std::to_string((size_t)M2_NOALLOC_SCRATCH_BYTES);
        return M2_ERR_ARG;
    }
    HIPCHK(hipSetDevice(ctx->device));
    JKWork w;
    w.fixed = true;
    w.Dn = reinterpret_cast<double*>(ctx->scratch);
    w.lists = reinterpret_cast<int*>(ctx->scratch + o_lists);
    w.J = reinterpret_cast<double*>(ctx->scratch + o_J);
    w.K = reinterpret_cast<double*>(ctx->scratch + o_K);
    return jk_device_core(ctx, dD, multiplier, precision, flags, dJ, dK, nullptr, nullptr, nullptr, nullptr, &w);
}

m2_status m2_eri_packed(m2_ctx* ctx, double* out) {
    if (!ctx) return M2_ERR_ARG;
    ARGCHK(out, "m2_eri_packed: null output");
    const int n = ctx->obs.nbf;
    ARGCHK(n > 0, "m2_eri_packed: no orbital basis set");
    ARGCHK(n <= 256, "m2_eri_packed: the packed tensor is for bases of at most 256 functions");
    HIPCHK(hipSetDevice(ctx->device));
    const size_t n2 = (size_t)n * n, np = (size_t)n * (n + 1) / 2;
    HIPCHK(ctx->d_D.reserve(n2));
    HIPCHK(ctx->d_Jo.reserve(n2));
    HIPCHK(ctx->d_Ko.reserve(n2));
    std::vector<double> D(n2, 0.0), J(n2);
    // column (.. | k l) of the tensor = J of the symmetric unit density with
    // D_kl = D_lk = 1/2 (1 on the diagonal), every quartet computed (no
    // screening, FP64): J_ab = sum_cd (ab|cd) D_cd = (ab|kl)
    for (int k = 0; k < n; ++k)
        for (int l = 0; l <= k; ++l) {
            const double v = k == l ? 1.0 : 0.5;
            D[(size_t)k * n + l] = v;
            D[(size_t)l * n + k] = v;
            HIPCHK(hipMemcpyAsync(ctx->d_D.p, D.data(), sizeof(double) * n2, hipMemcpyHostToDevice, ctx->stream));
            m2_status st = jk_device_core(ctx, ctx->d_D.p, 1.0, 0.0, M2_FP64, ctx->d_Jo.p, ctx->d_Ko.p);
            if (st) return st;
            HIPCHK(hipMemcpyAsync(J.data(), ctx->d_Jo.p, sizeof(double) * n2, hipMemcpyDeviceToHost, ctx->stream));
            HIPCHK(hipStreamSynchronize(ctx->stream));
            D[(size_t)k * n + l] = 0.0;
            D[(size_t)l * n + k] = 0.0;
            const size_t col = (size_t)k * (k + 1) / 2 + l;
            for (int a = 0; a < n; ++a)
                for (int b = 0; b <= a; ++b) out[((size_t)a * (a + 1) / 2 + b) * np + col] = J[(size_t)a * n + b];
        }
    return M2_OK;
}

m2_status m2_jk_direct(m2_ctx* ctx, const double* D, double multiplier, double precision, unsigned flags, double* J,
                       double* K) {
    if (!ctx) return M2_ERR_ARG;
    ARGCHK(D && J && K, "m2_jk_direct: null buffer");
    ARGCHK(ctx->obs.nbf > 0, "m2_jk_direct: no orbital basis set");
    HIPCHK(hipSetDevice(ctx->device));
    const size_t n2 = (size_t)ctx->obs.nbf * ctx->obs.nbf;
    HIPCHK(ctx->d_D.reserve(n2));
    HIPCHK(ctx->d_Jo.reserve(n2));
    HIPCHK(ctx->d_Ko.reserve(n2));
    HIPCHK(hipMemcpyAsync(ctx->d_D.p, D, sizeof(double) * n2, hipMemcpyHostToDevice, ctx->stream));
    m2_status st = jk_device_core(ctx, ctx->d_D.p, multiplier, precision, flags, ctx->d_Jo.p, ctx->d_Ko.p);
    if (st) return st;
    HIPCHK(hipMemcpyAsync(J, ctx->d_Jo.p, sizeof(double) * n2, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipMemcpyAsync(K, ctx->d_Ko.p, sizeof(double) * n2, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return M2_OK;
}

m2_status m2_jk_direct_uhf(m2_ctx* ctx, const double* Da, const double* Db, double multiplier, double precision,
                           unsigned flags, double* Ja, double* Ka, double* Jb, double* Kb) {
    if (!ctx) return M2_ERR_ARG;
    ARGCHK(Da && Db && Ja && Ka && Jb && Kb, "m2_jk_direct_uhf: null buffer");
    ARGCHK(ctx->obs.nbf > 0, "m2_jk_direct_uhf: no orbital basis set");
    HIPCHK(hipSetDevice(ctx->device));
    const size_t n2 = (size_t)ctx->obs.nbf * ctx->obs.nbf;
    std::vector<double> Dt(n2);
    for (size_t i = 0; i < n2; ++i) Dt[i] = Da[i] + Db[i];
    HIPCHK(ctx->d_D.reserve(3 * n2));
    HIPCHK(ctx->d_Jo.reserve(2 * n2));
    HIPCHK(ctx->d_Ko.reserve(2 * n2));
    double *dDa = ctx->d_D.p, *dDb = dDa + n2, *dDt = dDb + n2;
    HIPCHK(hipMemcpyAsync(dDa, Da, sizeof(double) * n2, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(dDb, Db, sizeof(double) * n2, hipMemcpyHostToDevice, ctx->stream));
    HIPCHK(hipMemcpyAsync(dDt, Dt.data(), sizeof(double) * n2, hipMemcpyHostToDevice, ctx->stream));
    // one pass over the quartets, both spin densities digested from the same integrals
    m2_status st = jk_device_core(ctx, dDa, multiplier, precision, flags, ctx->d_Jo.p, ctx->d_Ko.p, dDt, dDb,
                                  ctx->d_Jo.p + n2, ctx->d_Ko.p + n2);
    if (st) return st;
    HIPCHK(hipMemcpyAsync(Ja, ctx->d_Jo.p, sizeof(double) * n2, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipMemcpyAsync(Ka, ctx->d_Ko.p, sizeof(double) * n2, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipMemcpyAsync(Jb, ctx->d_Jo.p + n2, sizeof(double) * n2, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipMemcpyAsync(Kb, ctx->d_Ko.p + n2, sizeof(double) * n2, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    return M2_OK;
}

m2_status m2_set_mixed_threshold(m2_ctx* ctx, double fp32_below) {
    if (!ctx) return M2_ERR_ARG;
    ARGCHK(fp32_below >= 0.0, "m2_set_mixed_threshold: threshold must be >= 0");
    ctx->fp32_below = fp32_below;
    return M2_OK;
}

m2_status m2_set_jk_families(m2_ctx* ctx, int mode) {
    if (!ctx) return M2_ERR_ARG;
    ARGCHK(mode >= 0 && mode <= 2, "m2_set_jk_families: mode must be 0, 1 or 2");
    ctx->fam_mode = mode;
    return M2_OK;
}

m2_status m2_set_profiling(m2_ctx* ctx, int on) {
    if (!ctx) return M2_ERR_ARG;
    ctx->prof = on != 0;
    ctx->prof_counters = on >= 2;
    return M2_OK;
}

m2_status m2_get_profile(m2_ctx* ctx, double* ms, long long* launches, long long* quartets, long long* primq) {
    if (!ctx) return M2_ERR_ARG;
    for (int c = 0; c < 36; ++c) {
        if (ms) ms[c] = ctx->prof_ms[c];
        if (launches) launches[c] = ctx->prof_launch[c];
        if (quartets) quartets[c] = ctx->prof_q[c];
        if (primq) primq[c] = ctx->prof_pq[c];
    }
    return M2_OK;
}

m2_status m2_get_profile_computed(m2_ctx* ctx, long long* computed) {
    if (!ctx || !computed) return M2_ERR_ARG;
    for (int c = 0; c < 36; ++c) computed[c] = ctx->prof_cpq[c];
    return M2_OK;
}

m2_status m2_set_prim_screening(m2_ctx* ctx, double relative) {
    if (!ctx) return M2_ERR_ARG;
    ARGCHK(relative >= 0.0 && relative <= 1.0, "m2_set_prim_screening: relative threshold must lie in [0, 1]");
    ctx->prim_screen = relative;
    return M2_OK;
}

m2_status m2_set_prim_screening_fp32(m2_ctx* ctx, double relative) {
    if (!ctx) return M2_ERR_ARG;
    ARGCHK(relative >= 0.0 && relative <= 1.0, "m2_set_prim_screening_fp32: relative threshold must lie in [0, 1]");
    ctx->prim_screen32 = relative;
    return M2_OK;
}

m2_status m2_set_jk_launch(m2_ctx* ctx, int host_totals) {
    if (!ctx) return M2_ERR_ARG;
    ctx->host_totals = host_totals != 0;
    return M2_OK;
}

m2_status m2_set_jk_graphs(m2_ctx* ctx, int on) {
    if (!ctx) return M2_ERR_ARG;
    HIPCHK(hipSetDevice(ctx->device));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    ctx->use_graphs = on != 0;
    if (!ctx->use_graphs) drop_graphs(ctx);
    return M2_OK;
}

m2_status m2_jk_graph_stats(m2_ctx* ctx, long long* captures, long long* replays, int* cached) {
    if (!ctx) return M2_ERR_ARG;
    if (captures) *captures = ctx->graph_captures;
    if (replays) *replays = ctx->graph_replays;
    if (cached) *cached = (int)ctx->graphs.size();
    return M2_OK;
}

m2_status m2_set_jk_block_tasks(m2_ctx* ctx, int X, int Y, int bt) {
    if (!ctx) return M2_ERR_ARG;
    ARGCHK(X >= 0 && X < NPAIRCLASS && Y >= 0 && Y <= X, "m2_set_jk_block_tasks: need 0 <= Y <= X < 6");
    ARGCHK(bt >= 0 && bt <= 4096, "m2_set_jk_block_tasks: chunks per work item must lie in [0, 4096]");
    const int ci = X * (X + 1) / 2 + Y;
    ctx->tune_bt[ci] = bt;
    JKList& L = ctx->lists[ci];
    if (L.nbra > 0) {
        HIPCHK(hipSetDevice(ctx->device));
        HIPCHK(hipStreamSynchronize(ctx->stream));   // no build reads the old descriptors any more
        drop_graphs(ctx);
        L.bt = std::max(L.bt_min, bt > 0 ? bt : JK_COMBO[X][Y] ? JK_BLOCK_TASKS : L.bt_min);
        return upload_jklists(ctx);
    }
    return M2_OK;
}

m2_status m2_jk_autotune(m2_ctx* ctx, const double* dD, double precision, unsigned flags, int repeats,
                         double* ms_before, double* ms_after) {
    if (!ctx) return M2_ERR_ARG;
    ARGCHK(dD, "m2_jk_autotune: null density");
    ARGCHK(ctx->obs.nbf > 0, "m2_jk_autotune: no orbital basis set");
    ARGCHK(repeats >= 1 && repeats <= 100, "m2_jk_autotune: repeats must lie in [1, 100]");
    HIPCHK(hipSetDevice(ctx->device));
    const size_t n2 = (size_t)ctx->obs.nbf * ctx->obs.nbf;
    HIPCHK(ctx->d_Jo.reserve(n2));
    HIPCHK(ctx->d_Ko.reserve(n2));
    double* dJ = ctx->d_Jo.p;
    double* dK = ctx->d_Ko.p;
    const bool prof0 = ctx->prof, cnt0 = ctx->prof_counters, graphs0 = ctx->use_graphs;
    ctx->use_graphs = false;
    auto restore = [&]() {
        ctx->prof = prof0;
        ctx->prof_counters = cnt0;
        ctx->use_graphs = graphs0;
        (void)m2_reset_profile(ctx);
    };
    // wall time of `repeats` builds between two events on the stream
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    auto timed = [&](double* ms) -> m2_status {
        ctx->prof = false;
        m2_status st = jk_device_core(ctx, dD, 1.0, precision, flags, dJ, dK);   // warm-up
        if (st) return st;
        HIPCHK(hipEventRecord(e0, ctx->stream));
        for (int r = 0; r < repeats; ++r)
            if ((st = jk_device_core(ctx, dD, 1.0, precision, flags, dJ, dK))) return st;
        HIPCHK(hipEventRecord(e1, ctx->stream));
        HIPCHK(hipEventSynchronize(e1));
        float t = 0.f;
        HIPCHK(hipEventElapsedTime(&t, e0, e1));
        *ms = t / repeats;
        return M2_OK;
    };
    const int REG[] = {16, 32, 64, 128, 256}, HEAVY[] = {1, 2, 4, 8, 16};
    constexpr int NCAND = 5;
    double best_ms[NCOMBO];
    int best_bt[NCOMBO];
    for (int ci = 0; ci < NCOMBO; ++ci) {
        best_ms[ci] = 1e300;
        best_bt[ci] = ctx->lists[ci].bt;
    }
    double before = 0.0, after = 0.0;
    m2_status st = timed(&before);
    for (int k = 0; k < NCAND && st == M2_OK; ++k) {
        for (int X = 0, ci = 0; X < NPAIRCLASS; ++X)
            for (int Y = 0; Y <= X; ++Y, ++ci) {
                JKList& L = ctx->lists[ci];
                if (L.nbra > 0) L.bt = std::max(L.bt_min, JK_COMBO[X][Y] ? REG[k] : HEAVY[k]);
            }
        drop_graphs(ctx);
        if ((st = upload_jklists(ctx))) break;
        ctx->prof = true;
        ctx->prof_counters = false;
        if ((st = jk_device_core(ctx, dD, 1.0, precision, flags, dJ, dK))) break;   // warm-up
        (void)m2_reset_profile(ctx);
        for (int r = 0; r < repeats && st == M2_OK; ++r) st = jk_device_core(ctx, dD, 1.0, precision, flags, dJ, dK);
        for (int X = 0, ci = 0; X < NPAIRCLASS; ++X)
            for (int Y = 0; Y <= X; ++Y, ++ci) {
                const double ms = ctx->prof_ms[6 * X + Y] / repeats;
                if (ctx->lists[ci].nbra > 0 && ms < best_ms[ci]) {
                    best_ms[ci] = ms;
                    best_bt[ci] = ctx->lists[ci].bt;
                }
            }
    }
    // keep the fastest per combination (also for later bases)
    for (int ci = 0; ci < NCOMBO; ++ci) {
        JKList& L = ctx->lists[ci];
        if (L.nbra <= 0) continue;
        L.bt = best_bt[ci];
        ctx->tune_bt[ci] = best_bt[ci];
    }
    drop_graphs(ctx);
    m2_status su = upload_jklists(ctx);
    if (st == M2_OK) st = su;
    if (st == M2_OK) st = timed(&after);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    restore();
    if (st) return st;
    if (ms_before) *ms_before = before;
    if (ms_after) *ms_after = after;
    return M2_OK;
}

m2_status m2_get_jk_launch(m2_ctx* ctx, int* host_totals, int* block_tasks, int* grid_bound) {
    if (!ctx) return M2_ERR_ARG;
    if (host_totals) *host_totals = ctx->host_totals ? 1 : 0;
    for (int X = 0, ci = 0; X < NPAIRCLASS; ++X)
        for (int Y = 0; Y < NPAIRCLASS; ++Y) {
            const bool on = Y <= X && ctx->lists[ci].nbra > 0;
            if (block_tasks) block_tasks[6 * X + Y] = on ? ctx->lists[ci].bt : 0;
            if (grid_bound) grid_bound[6 * X + Y] = on ? ctx->lists[ci].bound : 0;
            if (Y <= X) ++ci;
        }
    return M2_OK;
}

m2_status m2_reset_profile(m2_ctx* ctx) {
    if (!ctx) return M2_ERR_ARG;
    for (int c = 0; c < 36; ++c) {
        ctx->prof_ms[c] = 0.0;
        ctx->prof_launch[c] = ctx->prof_q[c] = ctx->prof_pq[c] = ctx->prof_cpq[c] = 0;
    }
    return M2_OK;
}

m2_status m2_last_stats(m2_ctx* ctx, long long* q64, long long* q32, long long* ntasks) {
    if (!ctx) return M2_ERR_ARG;
    if (q64) *q64 = -1;
    if (q32) *q32 = -1;
    if (ntasks) {
        // a bounded-grid build leaves its work-item totals on the device
        if (ctx->last_tot) {
            int tot[2 * NCOMBO];
            HIPCHK(hipSetDevice(ctx->device));
            HIPCHK(hipStreamSynchronize(ctx->stream));
            HIPCHK(hipMemcpy(tot, ctx->last_tot, sizeof(tot), hipMemcpyDeviceToHost));
            long long units = 0;
            for (int ci = 0; ci < NCOMBO; ++ci)
                if (ctx->lists[ci].nbra > 0) units += tot[2 * ci] + tot[2 * ci + 1];
            ctx->last_tasks = units;
            ctx->last_tot = nullptr;
        }
        *ntasks = ctx->last_tasks;
    }
    return M2_OK;
}

}  // extern "C"
