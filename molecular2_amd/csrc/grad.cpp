// C ABI of the RHF gradient (the reference's Fock::compute_forces,
// src/rhfgrad.cpp:15-89): one-electron and Pulay terms, the two-electron
// adjoint, and the nuclear-repulsion term.
#include "ctx.hpp"
#include "grad_reg.hpp"
#include "generated/sph_tables.inc"

#include <cmath>
#include <cstring>

namespace m2 {
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
hipError_t deriv_upload(int device);
hipError_t launch_grad_2e(int device, int la, int lb, int lc, int ld, const GradArgs& a, const BoysTables& tab,
                          const Task* t, int nt, int natom, int fp32, hipStream_t st);
hipError_t launch_grad_1e(const Grad1Args& a, const int2* pairs, const int* count, hipStream_t st);
void launch_blocknorm(const double* A, int n, int nsh, const int* sh_bf, const int* sh_nbf, double* out, hipStream_t);
void launch_zero(double* p, size_t n, hipStream_t);
// task list of the J/K driver (api.cpp)
struct ComboRangeG {
    int X, Y, t0, tmid, t1;
};
hipError_t grad_tasks_device(m2_ctx* ctx, double precision, double Dmax, bool mixed, std::vector<ComboRangeG>& out);
}  // namespace m2

using namespace m2;

extern "C" {

m2_status m2_grad(m2_ctx* ctx, int natom, const double* Z, const double* xyz, const double* D, const double* W,
                  double precision, double* g) {
    return m2_grad_ex(ctx, natom, Z, xyz, D, W, precision, 0, g);
}

m2_status m2_grad_ex(m2_ctx* ctx, int natom, const double* Z, const double* xyz, const double* D, const double* W,
                     double precision, unsigned flags, double* g) {
    if (!ctx) return M2_ERR_ARG;
    ARGCHK(flags <= M2_MIXED, "m2_grad_ex: unknown flags");
    BasisData& b = ctx->obs;
    ARGCHK(b.nbf > 0, "m2_grad: no orbital basis set");
    ARGCHK(natom > 0 && Z && xyz && D && W && g, "m2_grad: bad arguments");
    HIPCHK(hipSetDevice(ctx->device));
    HIPCHK(deriv_upload(ctx->device));
    const int n = b.nbf, nsh = b.nsh;
    // shell -> atom (by index when given, else by centre, src/integrals.cpp:383-392)
    std::vector<int> atom(nsh);
    for (int s = 0; s < nsh; ++s) {
        int a = b.sh[s].atom;
        if (a < 0 || a >= natom) {
            a = -1;
            for (int k = 0; k < natom; ++k)
                if (b.sh[s].O[0] == xyz[3 * k] && b.sh[s].O[1] == xyz[3 * k + 1] && b.sh[s].O[2] == xyz[3 * k + 2]) {
                    a = k;
                    break;
                }
        }
        ARGCHK(a >= 0, "m2_grad: a shell is not centred on any atom");
        atom[s] = a;
    }
    // overlap-significant shell pairs (compute_shellpair_list, threshold 1e-12)
    const int npair = (int)ctx->pair_sh.size();
    std::vector<double> S((size_t)n * n), T((size_t)n * n), V((size_t)n * n);
    {
        m2_status st = m2_one_body(ctx, 0, nullptr, nullptr, S.data(), nullptr, nullptr);
        if (st) return st;
    }
    std::vector<unsigned char> sig(npair);
    for (int p = 0; p < npair; ++p) {
        const int s1 = ctx->pair_sh[p].x, s2 = ctx->pair_sh[p].y;
        const HostShell &A = b.sh[s1], &B = b.sh[s2];
        bool same = A.O[0] == B.O[0] && A.O[1] == B.O[1] && A.O[2] == B.O[2];
        if (same) {
            sig[p] = 1;
            continue;
        }
        double nrm = 0.0;
        for (int i = 0; i < A.nbf; ++i)
            for (int j = 0; j < B.nbf; ++j) {
                const double v = S[(size_t)(A.bf + i) * n + B.bf + j];
                nrm += v * v;
            }
        sig[p] = std::sqrt(nrm) >= 1e-12 ? 1 : 0;
    }
    // the call's device buffers, freed on every return path
    struct Bufs {
        DevBuf<double> dD, dW, dG, dZ, dX, dsph, dex, dcf, dDc;
        DevBuf<int> datom, dl, dnp, dpoff, dcfo;
        DevBuf<unsigned char> dsig;
        DevBuf<int2> dp1;
        ~Bufs() {
            for (auto* x : {&dD, &dW, &dG, &dZ, &dX, &dsph, &dex, &dcf, &dDc}) x->release();
            for (auto* x : {&datom, &dl, &dnp, &dpoff, &dcfo}) x->release();
            dsig.release();
            dp1.release();
        }
    } bufs;
    auto &dD = bufs.dD, &dW = bufs.dW, &dG = bufs.dG, &dZ = bufs.dZ, &dX = bufs.dX, &dsph = bufs.dsph,
         &dex = bufs.dex, &dcf = bufs.dcf, &dDc = bufs.dDc;
    auto &datom = bufs.datom, &dl = bufs.dl, &dnp = bufs.dnp, &dpoff = bufs.dpoff, &dcfo = bufs.dcfo;
    auto& dsig = bufs.dsig;
    auto& dp1 = bufs.dp1;
    HIPCHK(dD.reserve((size_t)n * n));
    HIPCHK(dW.reserve((size_t)n * n));
    HIPCHK(dG.reserve((size_t)natom * 3));
    HIPCHK(dZ.reserve(natom));
    HIPCHK(dX.reserve(3 * (size_t)natom));
    HIPCHK(dsph.reserve(30));
    HIPCHK(datom.reserve(nsh));
    HIPCHK(dsig.reserve(npair));
    HIPCHK(ctx->d_Dn.reserve((size_t)nsh * nsh));
    HIPCHK(hipMemcpy(dD.p, D, sizeof(double) * n * n, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(dW.p, W, sizeof(double) * n * n, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(dZ.p, Z, sizeof(double) * natom, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(dX.p, xyz, sizeof(double) * 3 * natom, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(dsph.p, SPH_TAB + SPH_OFF[2], sizeof(double) * 30, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(datom.p, atom.data(), sizeof(int) * nsh, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(dsig.p, sig.data(), npair, hipMemcpyHostToDevice));
    launch_zero(dG.p, (size_t)natom * 3, ctx->stream);
    launch_blocknorm(dD.p, n, nsh, b.d_bf.p, b.d_nbf.p, ctx->d_Dn.p, ctx->stream);
    HIPCHK(hipGetLastError());
    std::vector<double> Dn((size_t)nsh * nsh);
    HIPCHK(hipMemcpyAsync(Dn.data(), ctx->d_Dn.p, sizeof(double) * nsh * nsh, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    double Dmax = 0.0;
    for (double x : Dn) Dmax = std::max(Dmax, x);
    std::vector<ComboRangeG> ranges;
    HIPCHK(grad_tasks_device(ctx, precision, Dmax, flags == M2_MIXED, ranges));
    GradArgs a{};
    a.prims = ctx->d_prims.p;
    a.pair_off = ctx->d_pair_off.p;
    a.pair_np = ctx->d_pair_np.p;
    a.pair_sh = ctx->d_pair_sh.p;
    a.pair_Q = ctx->d_pair_Q.p;
    a.pair_sig = dsig.p;
    a.sh_O = b.d_O.p;
    a.sh_bf = b.d_bf.p;
    a.sh_atom = datom.p;
    a.Dn = ctx->d_Dn.p;
    a.D = dD.p;
    a.sph2 = dsph.p;
    a.n = n;
    a.nsh = nsh;
    a.pure = b.pure;
    a.precision = precision;
    a.prim_q = ctx->d_prim_q.p;
    a.prim_tau = precision * ctx->prim_screen;
    a.grad = dG.p;
    // D in the Cartesian basis (Dc = T D T^T; T maps each pure d shell's five
    // functions to its six Cartesian components, identity elsewhere), so that
    // the kernels read each quartet's density blocks directly
    std::vector<int> cfo(nsh);
    int nc = 0;
    for (int s = 0; s < nsh; ++s) {
        cfo[s] = nc;
        nc += (b.sh[s].l + 1) * (b.sh[s].l + 2) / 2;
    }
    const double* sph2 = SPH_TAB + SPH_OFF[2];
    auto to_cart = [&](int s, const double* src, size_t stride, double* dst, size_t dstride, int ncol) {
        // rows of shell s: dst[i] = sum_u T[i][u] src[u] (for every column)
        const HostShell& sh = b.sh[s];
        const int ncart = (sh.l + 1) * (sh.l + 2) / 2;
        if (!(b.pure && sh.l >= 2)) {
            for (int i = 0; i < ncart; ++i)
                for (int c = 0; c < ncol; ++c) dst[i * dstride + c] = src[(size_t)(sh.bf + i) * stride + c];
            return;
        }
        for (int i = 0; i < ncart; ++i)
            for (int c = 0; c < ncol; ++c) {
                double v = 0.0;
                for (int u = 0; u < 5; ++u) v += sph2[u * 6 + i] * src[(size_t)(sh.bf + u) * stride + c];
                dst[i * dstride + c] = v;
            }
    };
    std::vector<double> Y((size_t)nc * n), YT((size_t)n * nc), Dct((size_t)nc * nc);
    for (int s = 0; s < nsh; ++s) to_cart(s, D, n, Y.data() + (size_t)cfo[s] * n, n, n);   // T D (nc x n)
    for (int r = 0; r < nc; ++r)
        for (int c = 0; c < n; ++c) YT[(size_t)c * nc + r] = Y[(size_t)r * n + c];
    for (int s = 0; s < nsh; ++s) to_cart(s, YT.data(), nc, Dct.data() + (size_t)cfo[s] * nc, nc, nc);   // (T D T^T)^T
    HIPCHK(dDc.reserve((size_t)nc * nc));
    HIPCHK(dcfo.reserve(nsh));
    HIPCHK(hipMemcpy(dDc.p, Dct.data(), sizeof(double) * nc * nc, hipMemcpyHostToDevice));   // symmetric
    HIPCHK(hipMemcpy(dcfo.p, cfo.data(), sizeof(int) * nsh, hipMemcpyHostToDevice));
    a.Dc = dDc.p;
    a.sh_cf = dcfo.p;
    a.nc = nc;
    const BoysTables tab = boys_tables(ctx);
    for (const ComboRangeG& r : ranges) {
        // FP64 tasks [t0, tmid), FP32 tasks [tmid, t1) (mixed precision only)
        HIPCHK(launch_grad_2e(ctx->device, PAIR_L[r.X][0], PAIR_L[r.X][1], PAIR_L[r.Y][0], PAIR_L[r.Y][1], a, tab,
                              ctx->d_tasks.p + r.t0, r.tmid - r.t0, natom, 0, ctx->stream));
        HIPCHK(launch_grad_2e(ctx->device, PAIR_L[r.X][0], PAIR_L[r.X][1], PAIR_L[r.Y][0], PAIR_L[r.Y][1], a, tab,
                              ctx->d_tasks.p + r.tmid, r.t1 - r.tmid, natom, 1, ctx->stream));
    }
    // one-electron part
    std::vector<int> np(nsh), poff(nsh);
    std::vector<double> ex, cf;
    for (int s = 0; s < nsh; ++s) {
        np[s] = (int)b.sh[s].a.size();
        poff[s] = (int)ex.size();
        ex.insert(ex.end(), b.sh[s].a.begin(), b.sh[s].a.end());
        cf.insert(cf.end(), b.sh[s].c.begin(), b.sh[s].c.end());
    }
    HIPCHK(dl.reserve(nsh));
    HIPCHK(dnp.reserve(nsh));
    HIPCHK(dpoff.reserve(nsh));
    HIPCHK(dex.reserve(ex.size()));
    HIPCHK(dcf.reserve(cf.size()));
    HIPCHK(hipMemcpy(dl.p, b.h_l.data(), sizeof(int) * nsh, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(dnp.p, np.data(), sizeof(int) * nsh, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(dpoff.p, poff.data(), sizeof(int) * nsh, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(dex.p, ex.data(), sizeof(double) * ex.size(), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(dcf.p, cf.data(), sizeof(double) * cf.size(), hipMemcpyHostToDevice));
    Grad1Args g1{};
    g1.sh_O = b.d_O.p;
    g1.sh_l = dl.p;
    g1.sh_bf = b.d_bf.p;
    g1.sh_np = dnp.p;
    g1.sh_poff = dpoff.p;
    g1.sh_atom = datom.p;
    g1.exps = dex.p;
    g1.coefs = dcf.p;
    g1.D = dD.p;
    g1.W = dW.p;
    g1.sph2 = dsph.p;
    g1.Z = dZ.p;
    g1.xyz = dX.p;
    g1.nsh = nsh;
    g1.n = n;
    g1.pure = b.pure;
    g1.natom = natom;
    g1.tab = tab;
    g1.grad = dG.p;
    // ordered shell pairs grouped by (l1, l2) for the class-templated kernel
    std::vector<int2> p1;
    int cnt1[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int c = 0; c < 9; ++c)
        for (int s1 = 0; s1 < nsh; ++s1)
            for (int s2 = 0; s2 < nsh; ++s2)
                if (b.sh[s1].l * 3 + b.sh[s2].l == c) {
                    p1.push_back(make_int2(s1, s2));
                    ++cnt1[c];
                }
    HIPCHK(dp1.reserve(p1.size()));
    HIPCHK(hipMemcpy(dp1.p, p1.data(), sizeof(int2) * p1.size(), hipMemcpyHostToDevice));
    HIPCHK(launch_grad_1e(g1, dp1.p, cnt1, ctx->stream));
    HIPCHK(hipMemcpyAsync(g, dG.p, sizeof(double) * natom * 3, hipMemcpyDeviceToHost, ctx->stream));
    HIPCHK(hipStreamSynchronize(ctx->stream));
    // nuclear repulsion (src/rhfgrad.cpp:60-86)
    for (int a1 = 1; a1 < natom; ++a1)
        for (int a2 = 0; a2 < a1; ++a2) {
            double d[3], r2 = 0.0;
            for (int k = 0; k < 3; ++k) {
                d[k] = xyz[3 * a1 + k] - xyz[3 * a2 + k];
                r2 += d[k] * d[k];
            }
            const double f = Z[a1] * Z[a2] / (std::sqrt(r2) * r2);
            for (int k = 0; k < 3; ++k) {
                g[3 * a1 + k] -= d[k] * f;
                g[3 * a2 + k] += d[k] * f;
            }
        }
    return M2_OK;
}

}  // extern "C"
