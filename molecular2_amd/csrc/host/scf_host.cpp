// Host C++ SCF over the MI355X engine: restatement of the reference's
// SCF::rhf loop (src/scf.cpp:90-185), Fock incremental build logic
// (src/fockbuild.cpp:176-222), Fock set-up / transform / diagonalise / makeDens
// (src/fock.cpp:32-73,114-244) and DIISEngine (src/diis.cpp:15-116).
// All O(n^3) work runs on the device (m2_gemm / m2_eigh), all two-electron
// work through m2_jk_direct / m2_jk_df.
#include "scf_host.hpp"

#include <algorithm>
#include <cstring>

namespace molecular {

// ---------------------------------------------------------------- Matrix
Matrix Matrix::transpose() const {
    Matrix t(c_, r_);
    for (int j = 0; j < c_; ++j)
        for (int i = 0; i < r_; ++i) t(j, i) = (*this)(i, j);
    return t;
}
Matrix Matrix::operator+(const Matrix& o) const {
    Matrix t(*this);
    t += o;
    return t;
}
Matrix Matrix::operator-(const Matrix& o) const {
    Matrix t(*this);
    t -= o;
    return t;
}
Matrix& Matrix::operator+=(const Matrix& o) {
    if (o.r_ != r_ || o.c_ != c_) throw Error("MATRIX", "size mismatch in +=");
    for (size_t k = 0; k < a_.size(); ++k) a_[k] += o.a_[k];
    return *this;
}
Matrix& Matrix::operator-=(const Matrix& o) {
    if (o.r_ != r_ || o.c_ != c_) throw Error("MATRIX", "size mismatch in -=");
    for (size_t k = 0; k < a_.size(); ++k) a_[k] -= o.a_[k];
    return *this;
}
Matrix Matrix::operator*(double s) const {
    Matrix t(*this);
    for (auto& x : t.a_) x *= s;
    return t;
}
double Matrix::norm() const {
    double s = 0;
    for (double x : a_) s += x * x;
    return std::sqrt(s);
}
double Matrix::dot(const Matrix& o) const {
    double s = 0;
    for (size_t k = 0; k < a_.size(); ++k) s += a_[k] * o.a_[k];
    return s;
}
Matrix Matrix::block(int r0, int c0, int nr, int nc) const {
    Matrix b(nr, nc);
    for (int j = 0; j < nc; ++j)
        for (int i = 0; i < nr; ++i) b(i, j) = (*this)(r0 + i, c0 + j);
    return b;
}

static void check(m2_ctx* ctx, m2_status st, const char* what) {
    if (st != M2_OK) throw Error("M2", std::string(what) + ": " + m2_last_error(ctx));
}

Matrix matmul(m2_ctx* ctx, const Matrix& A, const Matrix& B, bool tA, bool tB) {
    const int m = tA ? A.cols() : A.rows();
    const int k = tA ? A.rows() : A.cols();
    const int n = tB ? B.rows() : B.cols();
    if ((tB ? B.cols() : B.rows()) != k) throw Error("MATRIX", "size mismatch in matmul");
    Matrix C(m, n);
    check(ctx, m2_gemm(ctx, tA ? 'T' : 'N', tB ? 'T' : 'N', m, n, k, 1.0, A.data(), A.rows(), B.data(), B.rows(), 0.0,
                       C.data(), m),
          "m2_gemm");
    return C;
}

void eigh(m2_ctx* ctx, const Matrix& A, Vector& w, Matrix& V) {
    const int n = A.rows();
    w.assign(n, 0.0);
    V = Matrix(n, n);
    check(ctx, m2_eigh(ctx, n, A.data(), w.data(), V.data()), "m2_eigh");
}

// -------------------------------------------------------------- Molecule
int Molecule::getNel() const {
    int z = 0;
    for (const auto& a : atoms) z += a.Z;
    return z - charge;
}

double Molecule::getEnuc() const {
    double e = 0.0;
    for (size_t i = 0; i < atoms.size(); ++i)
        for (size_t j = i + 1; j < atoms.size(); ++j) {
            const double dx = atoms[i].x - atoms[j].x, dy = atoms[i].y - atoms[j].y, dz = atoms[i].z - atoms[j].z;
            e += atoms[i].Z * atoms[j].Z / std::sqrt(dx * dx + dy * dy + dz * dz);
        }
    return e;
}

// -------------------------------------------------------- IntegralEngine
static void set_basis(m2_ctx* ctx, int which, const std::vector<RawShell>& sh, bool pure) {
    std::vector<int> l, np, at;
    std::vector<double> ex, cf, cen;
    for (const auto& s : sh) {
        l.push_back(s.l);
        np.push_back((int)s.exps.size());
        at.push_back(s.atom);
        ex.insert(ex.end(), s.exps.begin(), s.exps.end());
        cf.insert(cf.end(), s.coefs.begin(), s.coefs.end());
        cen.insert(cen.end(), s.O, s.O + 3);
    }
    check(ctx, m2_basis_set(ctx, which, (int)sh.size(), l.data(), np.data(), ex.data(), cf.data(), cen.data(),
                            at.data(), pure ? 1 : 0),
          "m2_basis_set");
}

IntegralEngine::IntegralEngine(m2_ctx* c, const Molecule& m) : ctx(c) {
    set_basis(ctx, M2_ORBITAL, m.shells, m.pure);
    check(ctx, m2_basis_info(ctx, M2_ORBITAL, &nbf, &nsh), "m2_basis_info");
    if (!m.jkshells.empty()) {
        set_basis(ctx, M2_AUXILIARY, m.jkshells, m.pure);
        check(ctx, m2_basis_info(ctx, M2_AUXILIARY, &naux, nullptr), "m2_basis_info");
    }
    sints = tints = naints = Matrix(nbf, nbf);
    std::vector<double> Z, xyz;
    for (const auto& a : m.atoms) {
        Z.push_back(a.Z);
        xyz.push_back(a.x);
        xyz.push_back(a.y);
        xyz.push_back(a.z);
    }
    check(ctx, m2_one_body(ctx, (int)Z.size(), Z.data(), xyz.data(), sints.data(), tints.data(), naints.data()),
          "m2_one_body");
    prescreen = Matrix(nsh, nsh);
    check(ctx, m2_schwarz(ctx, prescreen.data()), "m2_schwarz");
}

Matrix IntegralEngine::compute_shellblock_norm(const Matrix& A) {
    Matrix out(nsh, nsh);
    check(ctx, m2_shellblock_norm(ctx, A.data(), out.data()), "m2_shellblock_norm");
    return out;
}

// ------------------------------------------------------------ DIISEngine
void DIISEngine::init(int maxd, bool use, double damp) {
    maxDiis = maxd > 0 ? maxd : 8;
    useDiis = use;
    damping_factor = damp;
    lastB = Matrix::Zero(1, 1);
    errs.clear();
}

static double vdot(const Vector& a, const Vector& b) {
    double s = 0;
    for (size_t i = 0; i < a.size(); ++i) s += a[i] * b[i];
    return s;
}

// Full-pivot LU of a small dense system; returns the rank like Eigen's
// FullPivLU (pivots above eps * size * max pivot).
static Vector lu_solve(Matrix A, Vector b, int& rank) {
    const int n = A.rows();
    std::vector<int> rp(n), cp(n);
    for (int i = 0; i < n; ++i) rp[i] = cp[i] = i;
    double maxpiv = 0.0;
    rank = 0;
    for (int k = 0; k < n; ++k) {
        int pi = k, pj = k;
        double best = -1.0;
        for (int i = k; i < n; ++i)
            for (int j = k; j < n; ++j)
                if (std::fabs(A(i, j)) > best) {
                    best = std::fabs(A(i, j));
                    pi = i;
                    pj = j;
                }
        if (k == 0) maxpiv = best;
        if (best <= maxpiv * n * 2.220446049250313e-16 || best == 0.0) break;
        ++rank;
        if (pi != k) {
            for (int j = 0; j < n; ++j) std::swap(A(pi, j), A(k, j));
            std::swap(b[pi], b[k]);
        }
        if (pj != k) {
            for (int i = 0; i < n; ++i) std::swap(A(i, pj), A(i, k));
            std::swap(cp[pj], cp[k]);
        }
        for (int i = k + 1; i < n; ++i) {
            const double f = A(i, k) / A(k, k);
            for (int j = k; j < n; ++j) A(i, j) -= f * A(k, j);
            b[i] -= f * b[k];
        }
    }
    Vector y(n, 0.0);
    for (int i = rank - 1; i >= 0; --i) {
        double s = b[i];
        for (int j = i + 1; j < rank; ++j) s -= A(i, j) * y[j];
        y[i] = s / A(i, i);
    }
    Vector x(n, 0.0);
    for (int i = 0; i < n; ++i) x[cp[i]] = y[i];
    return x;
}

Vector DIISEngine::solve(const Matrix& B, int start) {
    const int size = B.rows() - start;
    double norm = 1.0;
    const double scale = 1.0 + damping_factor;
    if (B(start, start) > 1e-10) norm = 1.0 / B(start, start);
    Matrix nB(size, size);
    for (int i = 0; i < size - 1; ++i) {
        nB(i, i) = B(i + start, i + start) * scale;
        for (int j = i; j < size - 1; ++j) {   // j == i replaces the damped diagonal, as the reference does
            nB(i, j) = norm * B(i + start, j + start);
            nB(j, i) = nB(i, j);
        }
        nB(i, size - 1) = nB(size - 1, i) = -1.0;
    }
    nB(size - 1, size - 1) = 0.0;
    Vector rhs(size, 0.0);
    rhs[size - 1] = -1.0;
    int rank = 0;
    Vector w = lu_solve(nB, rhs, rank);
    Vector out(w);
    out.push_back((double)rank);
    return out;
}

Vector DIISEngine::compute(std::vector<Vector>& errors) {
    Vector weights;
    const int it = (int)errs.size() + 1;
    if (useDiis) {
        if (it > maxDiis) errs.erase(errs.begin());
        Vector ne;
        for (const auto& e : errors) ne.insert(ne.end(), e.begin(), e.end());
        errs.push_back(ne);
        if (it == 1) lastB(0, 0) = vdot(ne, ne);
        if (it > 1) {
            const int lim = it < maxDiis ? it : maxDiis;
            Matrix B(lim + 1, lim + 1);
            B(lim, lim) = 0.0;
            int start = it > lim ? 1 : 0;
            for (int i = 0; i < lim - 1; ++i) {
                for (int j = i; j < lim - 1; ++j) {
                    B(i, j) = lastB(i + start, j + start);
                    B(j, i) = B(i, j);
                }
                B(lim, i) = -1.0;
                B(i, lim) = -1.0;
            }
            for (int i = 0; i < lim; ++i) {
                B(lim - 1, i) = vdot(errs[i], errs[lim - 1]);
                B(i, lim - 1) = B(lim - 1, i);
            }
            B(lim, lim - 1) = -1.0;
            B(lim - 1, lim) = -1.0;
            lastB = B;
            weights = solve(B, 0);
            double rank = weights.back();
            start = 1;
            while (rank < (lim - start + 1) && start < lim) {
                weights = solve(B, start++);
                rank = weights.back();
            }
            weights.resize(weights.size() - 2);
        }
    }
    return weights;
}

// ------------------------------------------------------------------ Fock
Fock::Fock(const Options& o, IntegralEngine& ints, const Molecule& m) : opt(o), integrals(ints), molecule(m) {
    formHCore();
    formOrthog();
    // the in-core path stores n(n+1)/2 squared integrals; larger bases run
    // direct, as the reference does when the estimate exceeds its memory
    // (src/integrals.cpp:86-93)
    direct = o.direct || nbfs > 256;
    density_fitted = o.df;
    diis = o.diis;
    iter = 0;
    nocc = molecule.getNel() / 2;
    MAX = o.maxdiis;
    precision = o.precision;
    if (density_fitted && integrals.nauxiliary() == 0) throw Error("DF", "no auxiliary (JKFIT) basis set");
}

void Fock::formHCore() {
    hcore = integrals.getKinetic() + integrals.getNucAttract();
    nbfs = hcore.rows();
}

void Fock::formOrthog() {
    Vector lambda;
    Matrix U;
    eigh(integrals.context(), integrals.getOverlap(), lambda, U);
    Matrix Us(U);
    for (int j = 0; j < nbfs; ++j) {
        if (lambda[j] <= 0.0) throw Error("ORTHOG", "overlap matrix is not positive definite");
        const double s = 1.0 / std::sqrt(lambda[j]);
        for (int i = 0; i < nbfs; ++i) Us(i, j) *= s;
    }
    orthog = matmul(integrals.context(), Us, U, false, true);   // U lambda^-1/2 U^T
}

void Fock::makeJK(Matrix& P, double multiplier) {
    if (density_fitted) formJKdf(P, multiplier);
    else if (!direct) formJK(P, multiplier);
    else formJKdirect(integrals.getPrescreen(), P, multiplier);
}

// src/fockbuild.cpp:36-51 over the packed tensor: J(u,v) = m sum_sl P(s,l)
// (uv|sl), K(u,v) = m sum_sl P(s,l) (ul|sv).  The tensor is computed once
// (m2_eri_packed: every quartet in FP64 on the device, like the reference's
// compute_eris, src/eris.cpp:18-87).
void Fock::formJK(Matrix& P, double multiplier) {
    m2_ctx* ctx = integrals.context();
    const int n = nbfs;
    const size_t np = (size_t)n * (n + 1) / 2;
    auto t0 = std::chrono::steady_clock::now();
    if (twoints.empty()) {
        twoints.resize(np * np);
        check(ctx, m2_eri_packed(ctx, twoints.data()), "m2_eri_packed");
    }
    auto pidx = [](int i, int j) { return i >= j ? (size_t)i * (i + 1) / 2 + j : (size_t)j * (j + 1) / 2 + i; };
    jints = Matrix(n, n);
    kints = Matrix(n, n);
    for (int u = 0; u < n; ++u)
        for (int v = 0; v < n; ++v) {
            double j = 0.0, k = 0.0;
            const double* row_uv = twoints.data() + pidx(u, v) * np;
            for (int s = 0; s < n; ++s)
                for (int l = 0; l < n; ++l) {
                    const double p = P(s, l);
                    j += p * row_uv[pidx(s, l)];
                    k += p * twoints[pidx(u, l) * np + pidx(s, v)];
                }
            jints(u, v) = multiplier * j;
            kints(u, v) = multiplier * k;
        }
    jk_seconds += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    jkints = jints - 0.5 * kints;
}

void Fock::formJKdirect(const Matrix& Schwarz, Matrix& D, double multiplier) {
    // The engine holds the device copy of the same Schwarz bounds; an empty
    // Schwarz matrix switches the quartet screening off, as in the reference
    // (src/fockbuild.cpp:74,114): every unique quartet is computed.
    const double screen = Schwarz.rows() != 0 && Schwarz.cols() != 0 ? precision : 0.0;
    m2_ctx* ctx = integrals.context();
    unsigned flags = M2_FP64;
    if (opt.mixed && !fp64_forced) {
        if (last_build_incremental) {
            check(ctx, m2_set_mixed_threshold(ctx, 1e300), "m2_set_mixed_threshold");   // all FP32
        } else {
            check(ctx, m2_set_mixed_threshold(ctx, opt.mixed_threshold), "m2_set_mixed_threshold");
        }
        flags = M2_MIXED;
        ++n_fp32_builds;
    }
    jints = Matrix(nbfs, nbfs);
    kints = Matrix(nbfs, nbfs);
    auto t0 = std::chrono::steady_clock::now();
    check(ctx, m2_jk_direct(ctx, D.data(), multiplier, screen, flags, jints.data(), kints.data()), "m2_jk_direct");
    jk_seconds += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    // leave the context's split where the caller set it (an all-FP32 Delta-D build changes it)
    if (flags == M2_MIXED) check(ctx, m2_set_mixed_threshold(ctx, opt.mixed_threshold), "m2_set_mixed_threshold");
    jkints = jints - 0.5 * kints;
}

void Fock::formJKdf(Matrix& Cocc, double multiplier) {
    m2_ctx* ctx = integrals.context();
    jints = Matrix(nbfs, nbfs);
    kints = Matrix(nbfs, nbfs);
    auto t0 = std::chrono::steady_clock::now();
    const int fp32 = opt.mixed && !fp64_forced ? 1 : 0;
    n_fp32_builds += fp32;
    check(ctx, m2_jk_df(ctx, Cocc.data(), Cocc.cols(), fp32, jints.data(), kints.data()), "m2_jk_df");
    jk_seconds += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    jkints = multiplier * (2.0 * jints - kints);
}

// src/fockbuild.cpp:176-203
void Fock::makeFock(Matrix& P, double multiplier) {
    if (!density_fitted) {
        focka = fockinc;
        if (!started_incremental && rms_error < incremental_threshold) {
            started_incremental = true;
            reset_incremental = false;
            next_reset = rms_error / 1e1;
            last_reset = iter - 1;
        }
        bool full = reset_incremental || !started_incremental;
        if (full) {
            focka = hcore;
            dens_diff = dens;
        }
        if (reset_incremental && started_incremental) {
            reset_incremental = false;
            last_reset = iter;
            next_reset = rms_error / 1e1;
        }
        last_build_incremental = !full;
        if (full) ++n_full;
        else ++n_incremental;
        // dens_diff may have been reassigned: build from it
        makeJK(full ? dens_diff : P, multiplier);
    } else {
        focka = hcore;
        makeJK(P, multiplier);
    }
    focka += jkints;
    if (!density_fitted) fockinc = focka;
    addDiis();
}

void Fock::addDiis() {
    if (diis) {
        if (iter >= MAX) focks.erase(focks.begin());
        focks.push_back(focka);
        iter++;
    }
}

void Fock::makeFock() {
    if (!density_fitted) makeFock(dens_diff);
    else {
        Matrix cocc = CP.block(0, 0, nbfs, nocc);
        makeFock(cocc);
    }
}

void Fock::average(const Vector& w) {
    if (diis && iter > 2) {
        focka = Matrix::Zero(nbfs, nbfs);
        const int offset = (int)focks.size() - (int)w.size();
        for (int i = offset; i < (int)focks.size(); ++i) focka += w[i - offset] * focks[i];
    }
}

void Fock::transform(bool first) {
    m2_ctx* ctx = integrals.context();
    if (first) {
        if (opt.guess == "hcore") {
            fockm = matmul(ctx, orthog, matmul(ctx, hcore, orthog), true, false);
            fockinc = hcore;
        } else if (opt.guess == "soad") {
            compute_soad_guess();
            fockm = matmul(ctx, orthog, matmul(ctx, focka, orthog), true, false);
        } else {
            throw Error("GUESS", "unknown guess '" + opt.guess + "' (hcore or soad)");
        }
    } else {
        fockm = matmul(ctx, orthog, matmul(ctx, focka, orthog), true, false);
    }
}

Vector sto3g_occupations(int Z) {
    if (Z < 1 || Z > 53) throw Error("SOAD", "STO-3G occupations are defined for 1 <= Z <= 53");
    int nao;
    if (Z <= 2) nao = 1;
    else if (Z <= 10) nao = 5;
    else if (Z <= 18) nao = 9;
    else if (Z <= 20) nao = 13;
    else if (Z <= 36) nao = 18;
    else if (Z <= 38) nao = 22;
    else nao = 27;
    Vector occ(nao, 0.0);
    size_t pos = 0;
    // put min(ne, 2 size) electrons evenly into the next `size` AOs; AOs
    // beyond the STO-3G set (3d of K and Ca) are dropped
    auto fill = [&](int size, int& ne) {
        const int n = std::min(ne, 2 * size);
        ne -= n;
        for (int f = 0; f < size; ++f)
            if (pos + f < occ.size()) occ[pos + f] = double(n) / size;
        pos += size;
    };
    int ne = Z;
    fill(1, ne);                                        // 1s
    if (Z > 2) { fill(1, ne); fill(3, ne); }            // 2s 2p
    if (Z > 10) { fill(1, ne); fill(3, ne); }           // 3s 3p
    if (Z > 18 && Z <= 36) {
        int n4s = (Z == 19 || Z == 24 || Z == 29) ? 1 : 2;
        ne -= n4s;
        fill(1, n4s);                                   // 4s
        int n4p = std::min(6, Z > 30 ? Z - 30 : 0);
        ne -= n4p;
        fill(3, n4p);                                   // 4p
        fill(5, ne);                                    // 3d
    }
    if (Z > 36) {
        fill(1, ne);                                    // 4s
        fill(3, ne);                                    // 4p
        int n5s = (Z == 37 || Z == 41 || Z == 42 || Z == 44 || Z == 45 || Z == 47) ? 1 : 2;
        ne -= n5s;
        fill(1, n5s);                                   // 5s
        int n5p = std::min(6, Z > 48 ? Z - 48 : 0);
        ne -= n5p;
        fill(3, n5p);                                   // 5p
        fill(5, ne);                                    // 3d
        fill(5, ne);                                    // 4d
    }
    return occ;
}

Matrix soad_two_electron(int device, const Molecule& mol, double precision) {
    if (mol.minshells.empty()) throw Error("SOAD", "no minimal (STO-3G) basis given for the SOAD guess");
    // combined basis: orbital shells, then the minimal ones
    std::vector<int> l, np, at;
    std::vector<double> ex, cf, cen;
    int nmin = 0;
    for (int pass = 0; pass < 2; ++pass)
        for (const RawShell& r : pass == 0 ? mol.shells : mol.minshells) {
            l.push_back(r.l);
            np.push_back((int)r.exps.size());
            at.push_back(r.atom);
            ex.insert(ex.end(), r.exps.begin(), r.exps.end());
            cf.insert(cf.end(), r.coefs.begin(), r.coefs.end());
            for (int k = 0; k < 3; ++k) cen.push_back(r.O[k]);
            if (pass == 1) {
                if (r.l > 1) throw Error("SOAD", "minimal basis shells must have l <= 1");
                nmin += 2 * r.l + 1;
            }
        }
    // minimal-basis density: per atom, the occupations in shell order, halved
    Vector dmin;
    for (int a = 0; a < mol.getNAtoms(); ++a) {
        int nao = 0;
        for (const RawShell& r : mol.minshells)
            if (r.atom == a) nao += 2 * r.l + 1;
        const Vector occ = sto3g_occupations(mol.atoms[a].Z);
        if ((int)occ.size() != nao)
            throw Error("SOAD", "minimal basis of atom " + std::to_string(a) + " does not match STO-3G");
        for (double o : occ) dmin.push_back(0.5 * o);
    }
    if ((int)dmin.size() != nmin) throw Error("SOAD", "minimal basis shells must be grouped by atom");
    m2_ctx* tmp = nullptr;
    if (m2_ctx_create(&tmp, device) != M2_OK) throw Error("SOAD", "m2_ctx_create failed");
    struct Guard {
        m2_ctx* c;
        ~Guard() { m2_ctx_destroy(c); }
    } guard{tmp};
    if (m2_basis_set(tmp, M2_ORBITAL, (int)l.size(), l.data(), np.data(), ex.data(), cf.data(), cen.data(), at.data(),
                     mol.pure ? 1 : 0) != M2_OK)
        throw Error("SOAD", m2_last_error(tmp));
    int ntot = 0;
    m2_basis_info(tmp, M2_ORBITAL, &ntot, nullptr);
    const int n = ntot - nmin;
    Matrix D(ntot, ntot), J(ntot, ntot), K(ntot, ntot);
    for (int i = 0; i < nmin; ++i) D(n + i, n + i) = dmin[i];
    if (m2_jk_direct(tmp, D.data(), 1.0, precision, M2_FP64, J.data(), K.data()) != M2_OK)
        throw Error("SOAD", m2_last_error(tmp));
    Matrix G(n, n);
    for (int j = 0; j < n; ++j)
        for (int i = 0; i < n; ++i) G(i, j) = 2.0 * J(i, j) - K(i, j);
    return G;
}

void Fock::compute_soad_guess() {
    // the minimal-basis J/K carries no screening worth the name in the
    // reference (engine precision = machine epsilon)
    Matrix G = soad_two_electron(m2_ctx_device(integrals.context()), molecule, 1e-15);
    focka = hcore + G;
    fockinc = focka;
}

void Fock::diagonalise() {
    Matrix U;
    eigh(integrals.context(), fockm, eps, U);
    CP = matmul(integrals.context(), orthog, U);
}

void Fock::makeDens() {
    if (dens.rows() != 0) dens_diff = -dens;
    Matrix Co = CP.block(0, 0, nbfs, nocc);
    dens = 2.0 * matmul(integrals.context(), Co, Co, false, true);
    if (dens_diff.rows() == 0) dens_diff = dens;
    else dens_diff += dens;
}

// W = C_occ eps_occ C_occ^T, then the device gradient (one-electron, Pulay,
// two-electron adjoint and nuclear terms)
void Fock::compute_forces(const std::vector<Atom>& atoms, int nocc_) {
    m2_ctx* ctx = integrals.context();
    auto t0 = std::chrono::steady_clock::now();
    Matrix Co = CP.block(0, 0, nbfs, nocc_);
    Matrix Ce(Co);
    for (int j = 0; j < nocc_; ++j)
        for (int i = 0; i < nbfs; ++i) Ce(i, j) *= eps[j];
    Matrix W = matmul(ctx, Ce, Co, false, true);
    std::vector<double> Z, xyz;
    for (const auto& a : atoms) {
        Z.push_back(a.Z);
        xyz.push_back(a.x);
        xyz.push_back(a.y);
        xyz.push_back(a.z);
    }
    std::vector<double> g(3 * atoms.size());
    // a mixed-precision SCF also takes the quartets below its FP32 threshold
    // of the gradient in FP32
    unsigned flags = 0;
    if (opt.mixed) {
        check(ctx, m2_set_mixed_threshold(ctx, opt.mixed_threshold), "m2_set_mixed_threshold");
        flags = M2_MIXED;
    }
    check(ctx,
          m2_grad_ex(ctx, (int)atoms.size(), Z.data(), xyz.data(), dens.data(), W.data(), precision, flags, g.data()),
          "m2_grad_ex");
    forces = Matrix((int)atoms.size(), 3);
    for (size_t a = 0; a < atoms.size(); ++a)
        for (int k = 0; k < 3; ++k) forces((int)a, k) = g[3 * a + k];
    grad_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

// ------------------------------------------------------------------- SCF
SCF::SCF(const Options& o, const Molecule& m, Fock& f) : opt(o), molecule(m), focker(f) {
    diis.init(o.maxdiis, o.diis);
}

double SCF::calcE(const Matrix& hcore, const Matrix& dens, const Matrix& fock) {
    // 0.5 tr(D H), 0.5 tr(D F): D, H, F symmetric so tr(D X) = sum D_ij X_ij
    one_E = 0.5 * dens.dot(hcore);
    two_E = 0.5 * dens.dot(fock) - one_E;
    one_E *= 2.0;
    return one_E + two_E;
}

void SCF::calcE() {
    last_energy = energy;
    energy = calcE(focker.getHCore(), focker.getDens(), focker.getFockAO()) + molecule.getEnuc();
}

// e = X^T (F D S - S D F) X, flattened column-major (src/scf.cpp:52-59)
Vector SCF::calcErr(const Matrix& F, const Matrix& D, const Matrix& S, const Matrix& orthog) {
    m2_ctx* ctx = focker.context();
    Matrix DS = matmul(ctx, D, S);
    Matrix FDS = matmul(ctx, F, DS);
    // S D F = (F D S)^T for symmetric F, D, S
    Matrix e = FDS - FDS.transpose();
    e = matmul(ctx, orthog, matmul(ctx, e, orthog), true, false);
    error = e.norm();
    return Vector(e.data(), e.data() + e.size());
}

Vector SCF::calcErr() {
    return calcErr(focker.getFockAO(), focker.getDens(), focker.getS(), focker.getOrthog());
}

bool SCF::testConvergence(double val) {
    const bool r = val < opt.densconverge;
    last_err = error;
    return r;
}

// src/scf.cpp:90-185, plus the precision policy of the engine: with
// Options::mixed the SCF first converges with mixed / FP32 builds, then
// forces a full FP64 rebuild and keeps FP64 for at least fp64_final more
// iterations before it may stop.
void SCF::rhf(bool print) {
    (void)print;
    const int nel = molecule.getNel();
    if (nel % 2 != 0) throw Error("RHF", "Molecule has an odd number of electrons.");
    if (molecule.getMultiplicity() != 1) throw Error("RHF", "Molecule is not a singlet state.");
    auto tic = std::chrono::steady_clock::now();
    auto lap = [&]() {
        auto t = std::chrono::steady_clock::now();
        double s = std::chrono::duration<double>(t - tic).count();
        tic = t;
        return s;
    };
    converged = false;
    log.clear();
    focker.fp64_forced = !opt.mixed;
    focker.transform(true);
    focker.diagonalise();
    focker.makeDens();
    Matrix old_dens = focker.getDens();
    focker.makeFock();
    std::vector<Vector> errs;
    errs.push_back(calcErr());
    Vector weights = diis.compute(errs);
    errs.clear();
    calcE();
    log.push_back({0, energy, 0.0, 0.0, lap(), 0, opt.mixed ? 1 : 0});
    focker.average(weights);
    focker.transform(false);
    int iter = 1;
    int fp64_iters = opt.mixed ? 0 : 1 << 30;
    while (!converged && iter < opt.maxiter) {
        focker.diagonalise();
        focker.makeDens();
        const double dd = (focker.getDens() - old_dens).norm();
        old_dens = focker.getDens();
        const int inc_before = focker.n_incremental;
        focker.makeFock();
        const int was_incremental = focker.n_incremental > inc_before;
        errs.push_back(calcErr());
        focker.rms_error = [&]() {
            double s = 0;
            for (double x : errs[0]) s += x * x;
            return std::sqrt(s);
        }() / ((double)old_dens.rows() * old_dens.cols());
        if (focker.rms_error < focker.next_reset || iter - focker.last_reset > 8) focker.reset_incremental = true;
        if (iter > opt.diis_start) weights = diis.compute(errs);
        errs.clear();
        calcE();
        const double delta = std::fabs(energy - last_energy);
        log.push_back({iter, energy, delta, dd, lap(), was_incremental, focker.fp64_forced ? 0 : 1});
        if (iter > opt.diis_start) focker.average(weights);
        focker.transform(false);
        bool conv = testConvergence(dd) && delta < opt.enconverge;
        if (focker.fp64_forced) ++fp64_iters;
        // the mixed phase cannot resolve changes below FP32 noise: it hands
        // over to FP64 at max(criterion, FP32 floor)
        const bool mixed_done = (dd < std::max(opt.densconverge, opt.handover_dens) &&
                                 delta < std::max(opt.enconverge, opt.handover_energy)) ||
                                iter >= opt.maxiter / 2;
        if (!focker.fp64_forced && mixed_done) conv = true;
        if (conv && !focker.fp64_forced) {
            // converged under mixed precision: switch to FP64, full rebuild next
            focker.fp64_forced = true;
            focker.reset_incremental = true;
            fp64_iters = 0;
            conv = false;
        }
        if (conv && fp64_iters < opt.fp64_final) conv = false;
        converged = conv;
        iter++;
    }
    focker.diagonalise();
    iterations = iter;
    // like the reference (src/scf.cpp:163-166) non-convergence is reported,
    // not thrown: the caller reads `converged`
}

// ---------------------------------------------------------------- UHF
UnrestrictedFock::UnrestrictedFock(const Options& o, IntegralEngine& ints, const Molecule& m) : Fock(o, ints, m) {
    nalpha = m.nalpha();
    nbeta = m.nbeta();
    if (nalpha < 0 || nbeta < 0 || nalpha > nbfs || (m.getNel() - m.multiplicity + 1) % 2 != 0)
        throw Error("UHF", "electron count and multiplicity do not fit");
    // mixed precision (FP32 quartets / fits with an FP64 hand-over) is built
    // for the restricted SCF only; the reference's UHF is FP64 throughout
    if (o.mixed) throw Error("UHF", "mixed precision is available for RHF only");
}

// src/fock.cpp:305-323
void UnrestrictedFock::transform(bool first) {
    m2_ctx* ctx = integrals.context();
    if (first) {
        if (opt.guess == "hcore") {
            fock_alpha_mo = matmul(ctx, orthog, matmul(ctx, hcore, orthog), true, false);
        } else if (opt.guess == "soad") {
            compute_soad_guess();
            fock_alpha_mo = matmul(ctx, orthog, matmul(ctx, focka, orthog), true, false);
        } else {
            throw Error("GUESS", "unknown guess '" + opt.guess + "' (hcore or soad)");
        }
        fock_beta_mo = fock_alpha_mo;
    } else {
        fock_alpha_mo = matmul(ctx, orthog, matmul(ctx, fock_alpha_ao, orthog), true, false);
        fock_beta_mo = matmul(ctx, orthog, matmul(ctx, fock_beta_ao, orthog), true, false);
    }
}

// src/fock.cpp:325-335
void UnrestrictedFock::diagonalise() {
    m2_ctx* ctx = integrals.context();
    Matrix Ua, Ub;
    eigh(ctx, fock_alpha_mo, eps_alpha, Ua);
    eigh(ctx, fock_beta_mo, eps_beta, Ub);
    CP_alpha = matmul(ctx, orthog, Ua);
    CP_beta = matmul(ctx, orthog, Ub);
}

// src/fock.cpp:337-342: D_s = C_s,occ C_s,occ^T (one electron per orbital)
void UnrestrictedFock::makeDens() {
    m2_ctx* ctx = integrals.context();
    Matrix Ca = CP_alpha.block(0, 0, nbfs, nalpha), Cb = CP_beta.block(0, 0, nbfs, nbeta);
    dens_alpha = nalpha > 0 ? matmul(ctx, Ca, Ca, false, true) : Matrix(nbfs, nbfs);
    dens_beta = nbeta > 0 ? matmul(ctx, Cb, Cb, false, true) : Matrix(nbfs, nbfs);
}

// src/fockbuild.cpp:412-429
void UnrestrictedFock::makeJK() {
    if (density_fitted) {
        Matrix ca = CP_alpha.block(0, 0, nbfs, nalpha), cb = CP_beta.block(0, 0, nbfs, nbeta);
        formJKdf(ca, cb);
    } else {
        formJKdirect(integrals.getPrescreen(), dens_alpha, dens_beta);
    }
}

void UnrestrictedFock::formJKdirect(const Matrix& Schwarz, Matrix& Da, Matrix& Db, double multiplier) {
    // empty Schwarz matrix: no quartet screening (src/fockbuild.cpp:74,114)
    const double screen = Schwarz.rows() != 0 && Schwarz.cols() != 0 ? precision : 0.0;
    m2_ctx* ctx = integrals.context();
    jints_alpha = Matrix(nbfs, nbfs);
    kints_alpha = Matrix(nbfs, nbfs);
    jints_beta = Matrix(nbfs, nbfs);
    kints_beta = Matrix(nbfs, nbfs);
    auto t0 = std::chrono::steady_clock::now();
    check(ctx, m2_jk_direct_uhf(ctx, Da.data(), Db.data(), multiplier, screen, M2_FP64, jints_alpha.data(),
                                kints_alpha.data(), jints_beta.data(), kints_beta.data()),
          "m2_jk_direct_uhf");
    jk_seconds += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    ++n_full;
}

// src/fockbuild.cpp:366-373 (J and K over C C^T of each spin)
void UnrestrictedFock::formJKdf(Matrix& ca_occ, Matrix& cb_occ, double multiplier) {
    m2_ctx* ctx = integrals.context();
    auto one = [&](Matrix& c, Matrix& J, Matrix& K) {
        J = Matrix(nbfs, nbfs);
        K = Matrix(nbfs, nbfs);
        if (c.cols() == 0) return;
        check(ctx, m2_jk_df(ctx, c.data(), c.cols(), 0, J.data(), K.data()), "m2_jk_df");
        J = multiplier * J;
        K = multiplier * K;
    };
    auto t0 = std::chrono::steady_clock::now();
    one(ca_occ, jints_alpha, kints_alpha);
    one(cb_occ, jints_beta, kints_beta);
    jk_seconds += std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    ++n_full;
}

// src/fockbuild.cpp:431-445
void UnrestrictedFock::makeFock() {
    fock_alpha_ao = hcore + jints_alpha + jints_beta;
    fock_beta_ao = fock_alpha_ao - kints_beta;
    fock_alpha_ao = fock_alpha_ao - kints_alpha;
    if (diis) {
        if (iter >= MAX) {
            alpha_focks.erase(alpha_focks.begin());
            beta_focks.erase(beta_focks.begin());
        }
        alpha_focks.push_back(fock_alpha_ao);
        beta_focks.push_back(fock_beta_ao);
        iter++;
    }
}

// src/fock.cpp:344-355
void UnrestrictedFock::average(const Vector& w) {
    if (diis && iter > 2) {
        fock_alpha_ao = Matrix(nbfs, nbfs);
        fock_beta_ao = Matrix(nbfs, nbfs);
        const int offset = (int)alpha_focks.size() - (int)w.size();
        for (int i = offset; i < (int)alpha_focks.size(); ++i) {
            fock_alpha_ao += w[i - offset] * alpha_focks[i];
            fock_beta_ao += w[i - offset] * beta_focks[i];
        }
    }
}

// src/scf.cpp:193-272: DIIS from the first iteration over both spins' error
// vectors; the energy uses the averaged Fock matrices.
void SCF::uhf(bool print) {
    (void)print;
    auto* uf = dynamic_cast<UnrestrictedFock*>(&focker);
    if (!uf) throw Error("UHF", "SCF::uhf needs an UnrestrictedFock");
    UnrestrictedFock& U = *uf;
    auto tic = std::chrono::steady_clock::now();
    auto lap = [&]() {
        auto t = std::chrono::steady_clock::now();
        double sec = std::chrono::duration<double>(t - tic).count();
        tic = t;
        return sec;
    };
    converged = false;
    log.clear();
    U.transform(true);
    U.diagonalise();
    const int n = U.getHCore().rows();
    Matrix old_a(n, n), old_b(n, n);
    int iter = 1;
    energy = 0.0;
    while (!converged && iter < opt.maxiter) {
        U.makeDens();
        U.makeJK();
        U.makeFock();
        std::vector<Vector> errs;
        errs.push_back(calcErr(U.getFockAlphaAO(), U.getDensAlpha(), U.getS(), U.getOrthog()));
        errs.push_back(calcErr(U.getFockBetaAO(), U.getDensBeta(), U.getS(), U.getOrthog()));
        Vector weights = diis.compute(errs);
        U.average(weights);
        const double ea = calcE(U.getHCore(), U.getDensAlpha(), U.getFockAlphaAO());
        const double eb = calcE(U.getHCore(), U.getDensBeta(), U.getFockBetaAO());
        U.transform(false);
        U.diagonalise();
        last_energy = energy;
        energy = ea + eb + molecule.getEnuc();
        const double delta = std::fabs(energy - last_energy);
        const double dist = (U.getDensAlpha() + U.getDensBeta() - old_a - old_b).norm();
        old_a = U.getDensAlpha();
        old_b = U.getDensBeta();
        log.push_back({iter, energy, delta, dist, lap(), 0, 0});
        if (delta < opt.enconverge && dist < opt.densconverge) converged = true;
        iter++;
    }
    U.diagonalise();
    iterations = iter;
}

// ------------------------------------------------------------ Hessian
// Fock::compute_hessian (src/rhfgrad.cpp:91-359), the analytic RHF Hessian
// H = H_nr + 2 H_r with the coupled-perturbed response of the orbitals, in
// the reference's steps and with its CPHF solver (cphf_group_solver,
// src/cphf.cpp:4-34: u <- Q + V u until ||du|| < tol, at most maxiter
// steps).  Two things differ in how the pieces are obtained:
//   * derivative integrals: the reference asks libint for first and second
//     derivative integrals.  Here the first-derivative matrices S1, T1+V1
//     and G1 = d(2J - K)[D]/dR (D = C_occ C_occ^T fixed) are central
//     differences (step bohr) of the engine's analytic integrals at
//     displaced geometries, and H_nr -- the second-derivative integrals
//     contracted with D, D x D and W plus the nuclear term -- is the central
//     difference of the engine's analytic gradient (m2_grad) taken at fixed
//     D and W, which is exactly its definition;
//   * MO integrals: the reference transforms the full ERI tensor (MP2
//     class).  Every MO-integral sum it forms is a J/K build of an AO
//     matrix contracted with MO coefficients: V u = [4 Cv^T J(X) Co - 2
//     Cv^T K(X) Co] / (e_i - e_a) with X = sym(Cv u Co^T), and the S1 terms
//     use J, K of Y = Co S1_oo Co^T; so each CPHF step is one direct J/K
//     build on the device instead of an O(n^5) transformation.
static Matrix sym_part(const Matrix& X) {
    Matrix Y = X + X.transpose();
    return Y * 0.5;
}

void Fock::compute_hessian(const std::vector<Atom>& atoms, int nocc_, double step, double cphf_tol, int cphf_maxiter) {
    m2_ctx* ctx = integrals.context();
    const int n = nbfs, N = (int)atoms.size(), nc = 3 * N, no = nocc_, nv = nbfs - nocc_;
    if (no <= 0 || nv <= 0) throw Error("HESSIAN", "needs occupied and virtual orbitals");
    const double hprec = std::min(precision, 1e-14);   // screening noise far below the difference quotients
    Matrix Co = CP.block(0, 0, n, no), Cv = CP.block(0, no, n, nv);
    Matrix Dh = matmul(ctx, Co, Co, false, true);       // dens / 2 (src/rhfgrad.cpp:93)
    Matrix Ce(Co);
    for (int j = 0; j < no; ++j)
        for (int i = 0; i < n; ++i) Ce(i, j) *= eps[j];
    Matrix W = matmul(ctx, Ce, Co, false, true);
    Matrix Dfull = Dh * 2.0;
    auto jk = [&](const Matrix& P, Matrix& J, Matrix& K) {
        J = Matrix(n, n);
        K = Matrix(n, n);
        check(ctx, m2_jk_direct(ctx, P.data(), 1.0, hprec, M2_FP64, J.data(), K.data()), "m2_jk_direct");
    };
    // first-derivative matrices and H_nr by central differences
    std::vector<Matrix> S1(nc, Matrix(n, n)), H1(nc, Matrix(n, n)), G1(nc, Matrix(n, n));
    Matrix Hnr(nc, nc);
    const double inv2h = 1.0 / (2.0 * step);
    for (int c = 0; c < nc; ++c)
        for (int sgn : {1, -1}) {
            Molecule m(molecule);
            const int a = c / 3, ax = c % 3;
            double* pa = ax == 0 ? &m.atoms[a].x : ax == 1 ? &m.atoms[a].y : &m.atoms[a].z;
            *pa += sgn * step;
            for (auto& sh : m.shells)
                if (sh.atom == a) sh.O[ax] += sgn * step;
            m.jkshells.clear();
            IntegralEngine ie(ctx, m);
            const double f = sgn * inv2h;
            S1[c] += ie.getOverlap() * f;
            H1[c] += (ie.getKinetic() + ie.getNucAttract()) * f;
            Matrix J, K;
            jk(Dh, J, K);
            G1[c] += (J * 2.0 - K) * f;
            std::vector<double> Z, xyz, g(nc);
            for (const auto& at : m.atoms) {
                Z.push_back(at.Z);
                xyz.push_back(at.x);
                xyz.push_back(at.y);
                xyz.push_back(at.z);
            }
            check(ctx, m2_grad(ctx, N, Z.data(), xyz.data(), Dfull.data(), W.data(), hprec, g.data()), "m2_grad");
            for (int r = 0; r < nc; ++r) Hnr(r, c) += g[r] * f;
        }
    {
        Molecule m(molecule);
        m.jkshells.clear();
        IntegralEngine restore(ctx, m);   // the context's basis back at the reference geometry
    }
    // CPHF per coordinate, then the perturbed density and energy-weighted density
    std::vector<Matrix> P1(nc), W1(nc);
    cphf_unconverged = 0;
    cphf_iterations = 0;
    for (int c = 0; c < nc; ++c) {
        Matrix S1mo = matmul(ctx, matmul(ctx, CP, S1[c], true, false), CP);
        Matrix H1mo = matmul(ctx, matmul(ctx, CP, H1[c], true, false), CP);
        Matrix G1mo = matmul(ctx, matmul(ctx, CP, G1[c], true, false), CP);
        Matrix Y = matmul(ctx, matmul(ctx, Co, S1mo.block(0, 0, no, no)), Co, false, true);
        Matrix JY, KY;
        jk(Y, JY, KY);
        Matrix GYvo = matmul(ctx, matmul(ctx, Cv, JY * 2.0 - KY, true, false), Co);   // 2(ai|lk)S1 - (ak|li)S1
        Matrix Q(nv, no);
        for (int i = 0; i < no; ++i)
            for (int a = 0; a < nv; ++a)
                Q(a, i) = (H1mo(a + no, i) - S1mo(a + no, i) * eps[i] - GYvo(a, i) + G1mo(a + no, i)) /
                          (eps[i] - eps[a + no]);
        // u <- Q + V u (cphf_single_solver)
        auto Vu = [&](const Matrix& U, Matrix& Jx, Matrix& Kx) {
            Matrix X = sym_part(matmul(ctx, matmul(ctx, Cv, U), Co, false, true));
            jk(X, Jx, Kx);
            Matrix T = matmul(ctx, matmul(ctx, Cv, Jx * 4.0 - Kx * 2.0, true, false), Co);
            for (int i = 0; i < no; ++i)
                for (int a = 0; a < nv; ++a) T(a, i) /= eps[i] - eps[a + no];
            return T;
        };
        Matrix U = Q, Jx, Kx;
        bool conv = false;
        int it = 0;
        while (!conv && it < cphf_maxiter) {
            Matrix U1 = Q + Vu(U, Jx, Kx);
            const double delta = (U1 - U).norm();
            U = U1;
            conv = delta < cphf_tol;
            ++it;
        }
        cphf_iterations += it;
        if (!conv) ++cphf_unconverged;   // the reference prints "CPHF failed to converge" and goes on
        // ftilde = H1 + G1 - C^T G(Y) C + C^T (4 J(X_u) - 2 K(X_u)) C over all MOs
        Matrix Xu = sym_part(matmul(ctx, matmul(ctx, Cv, U), Co, false, true));
        jk(Xu, Jx, Kx);
        Matrix Ft = H1mo + G1mo - matmul(ctx, matmul(ctx, CP, JY * 2.0 - KY, true, false), CP) +
                    matmul(ctx, matmul(ctx, CP, Jx * 4.0 - Kx * 2.0, true, false), CP);
        Matrix pt(n, n), wt(n, n);
        for (int i = 0; i < no; ++i) {
            for (int j = 0; j < no; ++j) {
                pt(i, j) = -S1mo(j, i);
                wt(i, j) = Ft(j, i) - (eps[i] + eps[j]) * S1mo(j, i);
            }
            for (int a = no; a < n; ++a) {
                pt(i, a) = 2.0 * U(a - no, i);
                wt(i, a) = eps[i] * pt(i, a);
            }
        }
        P1[c] = matmul(ctx, matmul(ctx, CP, pt), CP, false, true);
        W1[c] = matmul(ctx, matmul(ctx, CP, wt), CP, false, true);
    }
    // H = H_nr + 2 H_r over the upper triangle, mirrored (src/rhfgrad.cpp:335-358)
    hessian = Matrix(nc, nc);
    for (int i = 0; i < nc; ++i)
        for (int j = i; j < nc; ++j) {
            const double hr = P1[j].dot(H1[i]) + P1[j].dot(G1[i]) - W1[j].dot(S1[i]);
            hessian(i, j) = Hnr(i, j) + 2.0 * hr;
            hessian(j, i) = hessian(i, j);
        }
}

}  // namespace molecular
