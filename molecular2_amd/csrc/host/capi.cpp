// extern "C" driver of the host C++ SCF (molecular::SCF::rhf), for callers
// that cannot link C++ (the Python tests and bench use it through ctypes).
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstring>
#include <exception>
#include <fstream>
#include <iomanip>
#include <iterator>
#include <mutex>
#include <numeric>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

#include "inputs.hpp"
#include "m2_input.h"
#include "m2_scf.h"
#include "scf_host.hpp"

static thread_local std::string g_rhf_error;

static void load_shells(int ns, const int* ll, const int* np, const double* ex, const double* cf, const double* cen,
                        const int* at, std::vector<molecular::RawShell>& out) {
    int off = 0;
    for (int s = 0; s < ns; ++s) {
        molecular::RawShell r;
        r.l = ll[s];
        r.atom = at ? at[s] : -1;
        for (int k = 0; k < 3; ++k) r.O[k] = cen[3 * s + k];
        r.exps.assign(ex + off, ex + off + np[s]);
        r.coefs.assign(cf + off, cf + off + np[s]);
        off += np[s];
        out.push_back(r);
    }
}

// IntegralEngine::printERI (src/integrals.cpp:202-233): c1 >= c2, c1 >= c3 >=
// c4 (1-based), each line the weight (1/8 times the number of distinct
// index permutations of the quartet) times (c1 c2|c3 c4), zero (and counted
// as insignificant) below thrint; iostream default formatting as in the
// reference.
static void write_eri_file(std::ostream& out, int n, const std::vector<double>& packed, double thrint,
                           long long* nints, long long* ninsig) {
    const size_t np = (size_t)n * (n + 1) / 2;
    auto pidx = [](int i, int j) { return i >= j ? (size_t)i * (i + 1) / 2 + j : (size_t)j * (j + 1) / 2 + i; };
    out << " TWO ELECTRON INTEGRALS: \n";
    long long icount = 0, scount = 0;
    for (int c1 = 0; c1 < n; ++c1)
        for (int c2 = 0; c2 <= c1; ++c2)
            for (int c3 = 0; c3 <= c1; ++c3)
                for (int c4 = 0; c4 <= c3; ++c4) {
                    ++icount;
                    double w = 0.125;
                    if (c1 != c2) w *= 2.0;
                    if (c3 != c4) w *= 2.0;
                    if ((c1 + c2) != (c3 + c4)) w *= 2.0;
                    else if (c1 != c3 && c1 != c4) w *= 2.0;
                    else if (c2 != c3 && c2 != c4) w *= 2.0;
                    const double v = packed[pidx(c1, c2) * np + pidx(c3, c4)];
                    if (std::fabs(v) < thrint) {
                        ++scount;
                        w = 0.0;
                    }
                    out << std::setw(6) << c1 + 1 << std::setw(6) << c2 + 1 << std::setw(6) << c3 + 1
                        << std::setw(6) << c4 + 1 << std::setw(20) << w * v << "\n";
                }
    out << "N 2e Ints: " << icount << "\n";
    out << "N insig. Ints: " << scount << "\n";
    if (nints) *nints = icount;
    if (ninsig) *ninsig = scount;
}

static void print_eri_ctx(m2_ctx* ctx, const std::string& path, double thrint, long long* nints,
                          long long* ninsig) {
    int n = 0, nsh = 0;
    if (m2_basis_info(ctx, M2_ORBITAL, &n, &nsh) != 0 || n <= 0)
        throw std::runtime_error("m2_print_eri: no orbital basis set");
    const size_t np = (size_t)n * (n + 1) / 2;
    std::vector<double> packed(np * np);
    if (m2_eri_packed(ctx, packed.data()) != 0) throw std::runtime_error(std::string("m2_print_eri: ") + m2_last_error(ctx));
    std::ofstream out(path);
    if (!out) throw std::runtime_error("m2_print_eri: cannot open " + path);
    write_eri_file(out, n, packed, thrint, nints, ninsig);
    if (!out) throw std::runtime_error("m2_print_eri: write to " + path + " failed");
}

extern "C" {

const char* m2_rhf_last_error(void) { return g_rhf_error.c_str(); }

int m2_print_eri(m2_ctx* ctx, const char* path, double thrint, long long* nints, long long* ninsig) {
    g_rhf_error.clear();
    try {
        if (!ctx || !path) throw std::runtime_error("m2_print_eri: null argument");
        print_eri_ctx(ctx, path, thrint, nints, ninsig);
        return 0;
    } catch (const std::exception& e) {
        g_rhf_error = e.what();
        return 1;
    }
}

int m2_basis_file_shells(const char* path, int natom, const int* Z, const double* xyz, int* nsh, int* nprim,
                         m2_shells_out* out) {
    using namespace molecular;
    g_rhf_error.clear();
    try {
        if (!path || !Z || !xyz || !nsh || !nprim) throw std::runtime_error("m2_basis_file_shells: null argument");
        std::ifstream f(path);
        if (!f) throw std::runtime_error(std::string("BASISLIB: Could not open basis file ") + path);
        const std::string text((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
        std::vector<RawShell> all;
        for (int a = 0; a < natom; ++a) {
            std::vector<RawShell> sh = molpro_element_shells(text, element_symbol(Z[a]));
            if (sh.empty())
                throw std::runtime_error(std::string("BASISLIB: Could not read basis ") + path + " for atom " +
                                         element_symbol(Z[a]));
            for (RawShell& s : sh) {
                s.atom = a;
                for (int k = 0; k < 3; ++k) s.O[k] = xyz[3 * a + k];
                all.push_back(s);
            }
        }
        int np = 0;
        for (const RawShell& s : all) np += (int)s.exps.size();
        const bool fill = out && *nsh >= (int)all.size() && *nprim >= np;
        *nsh = (int)all.size();
        *nprim = np;
        if (!fill) return 0;
        int off = 0;
        for (size_t i = 0; i < all.size(); ++i) {
            const RawShell& s = all[i];
            if (out->l) out->l[i] = s.l;
            if (out->nprim) out->nprim[i] = (int)s.exps.size();
            if (out->atoms) out->atoms[i] = s.atom;
            if (out->centers)
                for (int k = 0; k < 3; ++k) out->centers[3 * i + k] = s.O[k];
            for (size_t k = 0; k < s.exps.size(); ++k, ++off) {
                if (out->exps) out->exps[off] = s.exps[k];
                if (out->coefs) out->coefs[off] = s.coefs[k];
            }
        }
        return 0;
    } catch (const std::exception& e) {
        g_rhf_error = e.what();
        return 1;
    }
}

int m2_soad_fock(int device, int natom, const double* Z, const double* xyz, int pure, int nsh, const int* l,
                 const int* nprim, const double* exps, const double* coefs, const double* centers, const int* atoms,
                 int nsh_min, const int* l_min, const int* nprim_min, const double* exps_min, const double* coefs_min,
                 const double* centers_min, const int* atoms_min, double* G) {
    using namespace molecular;
    g_rhf_error.clear();
    try {
        Molecule mol;
        mol.pure = pure != 0;
        for (int a = 0; a < natom; ++a) mol.atoms.push_back(Atom{(int)Z[a], xyz[3 * a], xyz[3 * a + 1], xyz[3 * a + 2]});
        load_shells(nsh, l, nprim, exps, coefs, centers, atoms, mol.shells);
        load_shells(nsh_min, l_min, nprim_min, exps_min, coefs_min, centers_min, atoms_min, mol.minshells);
        const Matrix g = soad_two_electron(device, mol, 1e-15);
        std::memcpy(G, g.data(), sizeof(double) * g.rows() * g.cols());
        return 0;
    } catch (const std::exception& e) {
        g_rhf_error = e.what();
        return 1;
    }
}

static molecular::Options options_from(const m2_scf_options_c* o) {
    molecular::Options opt;
    opt.diis = o->diis != 0;
    opt.maxdiis = o->maxdiis;
    opt.maxiter = o->maxiter;
    opt.enconverge = o->enconverge;
    opt.densconverge = o->densconverge;
    opt.precision = o->precision;
    opt.df = o->df != 0;
    opt.mixed = o->mixed != 0;
    opt.mixed_threshold = o->mixed_threshold;
    opt.fp64_final = o->fp64_final;
    opt.guess = o->guess == 1 ? "soad" : "hcore";
    opt.diis_start = o->diis_start;
    opt.direct = o->incore == 0;
    return opt;
}

// shells: orbital (nsh, l, nprim, exps, coefs, centers, atoms), optional
// auxiliary (nsh_aux may be 0).  Outputs D (n x n), C (n x n), eps (n) may
// be NULL.  Returns 0 on success; non-zero with m2_rhf_last_error() set.
int m2_rhf(m2_ctx* ctx, int natom, const double* Z, const double* xyz, int charge, int multiplicity, int pure,
           int nsh, const int* l, const int* nprim, const double* exps, const double* coefs, const double* centers,
           const int* atoms, int nsh_aux, const int* l_aux, const int* nprim_aux, const double* exps_aux,
           const double* coefs_aux, const double* centers_aux, const int* atoms_aux, int nsh_min, const int* l_min,
           const int* nprim_min, const double* exps_min, const double* coefs_min, const double* centers_min,
           const int* atoms_min, const m2_scf_options_c* o, m2_scf_result_c* res, double* D, double* C, double* eps,
           double* log5 /* maxiter x 5 or NULL */,
           double* grad /* natom x 3 or NULL: also compute Fock::compute_forces */) {
    using namespace molecular;
    g_rhf_error.clear();
    try {
        auto t0 = std::chrono::steady_clock::now();
        Molecule mol;
        mol.pure = pure != 0;
        mol.charge = charge;
        mol.multiplicity = multiplicity;
        for (int a = 0; a < natom; ++a) mol.atoms.push_back(Atom{(int)Z[a], xyz[3 * a], xyz[3 * a + 1], xyz[3 * a + 2]});
        load_shells(nsh, l, nprim, exps, coefs, centers, atoms, mol.shells);
        if (nsh_aux > 0)
            load_shells(nsh_aux, l_aux, nprim_aux, exps_aux, coefs_aux, centers_aux, atoms_aux, mol.jkshells);
        if (nsh_min > 0)
            load_shells(nsh_min, l_min, nprim_min, exps_min, coefs_min, centers_min, atoms_min, mol.minshells);
        const Options opt = options_from(o);
        IntegralEngine ints(ctx, mol);
        Fock fock(opt, ints, mol);
        SCF scf(opt, mol, fock);
        scf.rhf(false);
        const int n = ints.nbasis();
        if (grad) {
            fock.compute_forces(mol.atoms, fock.getNocc());
            for (int a = 0; a < natom; ++a)
                for (int k = 0; k < 3; ++k) grad[3 * a + k] = fock.getForces()(a, k);
        }
        if (res) {
            res->energy = scf.getEnergy();
            res->one_e = scf.getOneE();
            res->two_e = scf.getTwoE();
            res->enuc = mol.getEnuc();
            res->iterations = scf.iterations;
            res->converged = scf.converged ? 1 : 0;
            res->n_full = fock.n_full;
            res->n_incremental = fock.n_incremental;
            res->n_fp32 = fock.n_fp32_builds;
            res->jk_seconds = fock.jk_seconds;
            res->total_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            res->grad_seconds = fock.grad_seconds;
        }
        if (D) std::memcpy(D, fock.getDens().data(), sizeof(double) * n * n);
        if (C) std::memcpy(C, fock.getCP().data(), sizeof(double) * n * n);
        if (eps) std::memcpy(eps, fock.getEps().data(), sizeof(double) * n);
        if (log5) {
            int k = 0;
            for (const auto& r : scf.log) {
                if (k >= o->maxiter) break;
                log5[5 * k + 0] = r.energy;
                log5[5 * k + 1] = r.delta_e;
                log5[5 * k + 2] = r.delta_d;
                log5[5 * k + 3] = r.incremental;
                log5[5 * k + 4] = r.fp32;
                ++k;
            }
        }
        return 0;
    } catch (const std::exception& e) {
        g_rhf_error = e.what();
        return 1;
    }
}

int m2_uhf(m2_ctx* ctx, int natom, const double* Z, const double* xyz, int charge, int multiplicity, int pure,
           int nsh, const int* l, const int* nprim, const double* exps, const double* coefs, const double* centers,
           const int* atoms, int nsh_aux, const int* l_aux, const int* nprim_aux, const double* exps_aux,
           const double* coefs_aux, const double* centers_aux, const int* atoms_aux, int nsh_min, const int* l_min,
           const int* nprim_min, const double* exps_min, const double* coefs_min, const double* centers_min,
           const int* atoms_min, const m2_scf_options_c* o, m2_scf_result_c* res, double* Da, double* Db,
           double* eps_a, double* eps_b, double* log5) {
    using namespace molecular;
    g_rhf_error.clear();
    try {
        auto t0 = std::chrono::steady_clock::now();
        Molecule mol;
        mol.pure = pure != 0;
        mol.charge = charge;
        mol.multiplicity = multiplicity;
        for (int a = 0; a < natom; ++a) mol.atoms.push_back(Atom{(int)Z[a], xyz[3 * a], xyz[3 * a + 1], xyz[3 * a + 2]});
        load_shells(nsh, l, nprim, exps, coefs, centers, atoms, mol.shells);
        if (nsh_aux > 0)
            load_shells(nsh_aux, l_aux, nprim_aux, exps_aux, coefs_aux, centers_aux, atoms_aux, mol.jkshells);
        if (nsh_min > 0)
            load_shells(nsh_min, l_min, nprim_min, exps_min, coefs_min, centers_min, atoms_min, mol.minshells);
        const Options opt = options_from(o);
        IntegralEngine ints(ctx, mol);
        UnrestrictedFock fock(opt, ints, mol);
        SCF scf(opt, mol, fock);
        scf.uhf(false);
        const int n = ints.nbasis();
        if (res) {
            res->energy = scf.getEnergy();
            res->one_e = res->two_e = 0.0;
            res->enuc = mol.getEnuc();
            res->iterations = scf.iterations;
            res->converged = scf.converged ? 1 : 0;
            res->n_full = fock.n_full;
            res->n_incremental = 0;
            res->n_fp32 = 0;
            res->jk_seconds = fock.jk_seconds;
            res->total_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            res->grad_seconds = 0.0;
        }
        if (Da) std::memcpy(Da, fock.getDensAlpha().data(), sizeof(double) * n * n);
        if (Db) std::memcpy(Db, fock.getDensBeta().data(), sizeof(double) * n * n);
        if (eps_a) std::memcpy(eps_a, fock.getEpsAlpha().data(), sizeof(double) * n);
        if (eps_b) std::memcpy(eps_b, fock.getEpsBeta().data(), sizeof(double) * n);
        if (log5) {
            int k = 0;
            for (const auto& r : scf.log) {
                if (k >= o->maxiter) break;
                log5[5 * k + 0] = r.energy;
                log5[5 * k + 1] = r.delta_e;
                log5[5 * k + 2] = r.delta_d;
                log5[5 * k + 3] = 0;
                log5[5 * k + 4] = 0;
                ++k;
            }
        }
        return 0;
    } catch (const std::exception& e) {
        g_rhf_error = e.what();
        return 1;
    }
}

// Fock::compute_hessian after SCF::rhf (the reference's "hessian" path,
// src/optimiser.cpp:410): shells as for m2_rhf (the auxiliary basis is not
// used); hess: 3N x 3N; stats (optional): CPHF solves that did not converge
// and CPHF steps in total.
int m2_rhf_hessian_cphf(m2_ctx* ctx, int natom, const double* Z, const double* xyz, int charge, int multiplicity,
                        int pure, int nsh, const int* l, const int* nprim, const double* exps, const double* coefs,
                        const double* centers, const int* atoms, int nsh_min, const int* l_min, const int* nprim_min,
                        const double* exps_min, const double* coefs_min, const double* centers_min,
                        const int* atoms_min, const m2_scf_options_c* o, double step, double cphf_tol,
                        int cphf_maxiter, double* hess, m2_scf_result_c* res, int* stats) {
    using namespace molecular;
    g_rhf_error.clear();
    try {
        if (!(step > 0.0) || !(cphf_tol > 0.0) || cphf_maxiter < 1 || !hess)
            throw Error("HESSIAN", "step and cphf_tol must be positive, cphf_maxiter at least 1, hess non-null");
        auto t0 = std::chrono::steady_clock::now();
        Molecule mol;
        mol.pure = pure != 0;
        mol.charge = charge;
        mol.multiplicity = multiplicity;
        for (int a = 0; a < natom; ++a) mol.atoms.push_back(Atom{(int)Z[a], xyz[3 * a], xyz[3 * a + 1], xyz[3 * a + 2]});
        load_shells(nsh, l, nprim, exps, coefs, centers, atoms, mol.shells);
        if (nsh_min > 0)
            load_shells(nsh_min, l_min, nprim_min, exps_min, coefs_min, centers_min, atoms_min, mol.minshells);
        Options opt = options_from(o);
        if (opt.df) throw Error("HESSIAN", "the analytic Hessian needs the direct (non-DF) SCF");
        IntegralEngine ints(ctx, mol);
        Fock fock(opt, ints, mol);
        SCF scf(opt, mol, fock);
        scf.rhf(false);
        fock.compute_hessian(mol.atoms, fock.getNocc(), step, cphf_tol, cphf_maxiter);
        const int nc = 3 * natom;
        for (int j = 0; j < nc; ++j)
            for (int i = 0; i < nc; ++i) hess[(size_t)j * nc + i] = fock.getHessian()(i, j);
        if (stats) {
            stats[0] = fock.cphf_unconverged;
            stats[1] = fock.cphf_iterations;
        }
        if (res) {
            *res = m2_scf_result_c{};
            res->energy = scf.getEnergy();
            res->one_e = scf.getOneE();
            res->two_e = scf.getTwoE();
            res->enuc = mol.getEnuc();
            res->iterations = scf.iterations;
            res->converged = scf.converged ? 1 : 0;
            res->n_full = fock.n_full;
            res->n_incremental = fock.n_incremental;
            res->jk_seconds = fock.jk_seconds;
            res->total_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        }
        return 0;
    } catch (const std::exception& e) {
        g_rhf_error = e.what();
        return 1;
    }
}

// Batched RHF: molecule m runs m2_rhf on the context of whichever worker
// takes it; workers take molecules from one queue ordered by basis size,
// largest first, so the long runs start early and the short ones fill in.
int m2_rhf_batch(int device, int nworkers, int nmol, const int* natom, const int* charge, const double* Z,
                 const double* xyz, int pure, const int* nsh, const int* l, const int* nprim, const double* exps,
                 const double* coefs, const double* centers, const int* atoms, const m2_scf_options_c* o,
                 m2_scf_result_c* res) {
    if (o && o->guess != 0) {
        g_rhf_error = "m2_rhf_batch: the SOAD guess needs the minimal shells of m2_rhf_batch_min";
        return 1;
    }
    return m2_rhf_batch_min(device, nworkers, nmol, natom, charge, Z, xyz, pure, nsh, l, nprim, exps, coefs, centers,
                            atoms, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, o, res);
}

int m2_rhf_batch_min(int device, int nworkers, int nmol, const int* natom, const int* charge, const double* Z,
                     const double* xyz, int pure, const int* nsh, const int* l, const int* nprim, const double* exps,
                     const double* coefs, const double* centers, const int* atoms, const int* nsh_min,
                     const int* l_min, const int* nprim_min, const double* exps_min, const double* coefs_min,
                     const double* centers_min, const int* atoms_min, const m2_scf_options_c* o,
                     m2_scf_result_c* res) {
    g_rhf_error.clear();
    if (nmol < 0 || (nmol > 0 && (!natom || !Z || !xyz || !nsh || !l || !nprim || !exps || !coefs || !centers ||
                                  !atoms || !o || !res))) {
        g_rhf_error = "m2_rhf_batch: null argument";
        return 1;
    }
    const bool minimal = nsh_min != nullptr;
    if (minimal && (!l_min || !nprim_min || !exps_min || !coefs_min || !centers_min || !atoms_min)) {
        g_rhf_error = "m2_rhf_batch_min: null minimal-shell argument";
        return 1;
    }
    if (o->guess != 0 && !minimal) {
        g_rhf_error = "m2_rhf_batch_min: the SOAD guess (guess = 1) needs the minimal shells";
        return 1;
    }
    if (nmol == 0) return 0;
    // offsets of every molecule into the concatenated arrays
    std::vector<long> a_off(nmol + 1, 0), s_off(nmol + 1, 0), p_off(nmol + 1, 0);
    std::vector<long> ms_off(nmol + 1, 0), mp_off(nmol + 1, 0);
    std::vector<long> nbf(nmol, 0);
    for (int m = 0; m < nmol; ++m) {
        if (natom[m] <= 0 || nsh[m] <= 0) {
            g_rhf_error = "m2_rhf_batch: molecule " + std::to_string(m) + " has no atoms or no shells";
            return 1;
        }
        a_off[m + 1] = a_off[m] + natom[m];
        s_off[m + 1] = s_off[m] + nsh[m];
        long np = 0;
        for (long s = s_off[m]; s < s_off[m + 1]; ++s) {
            np += nprim[s];
            nbf[m] += pure ? 2 * l[s] + 1 : (l[s] + 1) * (l[s] + 2) / 2;
        }
        p_off[m + 1] = p_off[m] + np;
        if (minimal) {
            ms_off[m + 1] = ms_off[m] + nsh_min[m];
            long mp = 0;
            for (long s = ms_off[m]; s < ms_off[m + 1]; ++s) mp += nprim_min[s];
            mp_off[m + 1] = mp_off[m] + mp;
        }
    }
    std::vector<int> order(nmol);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return nbf[a] > nbf[b]; });
    if (nworkers <= 0) nworkers = 8;
    nworkers = std::min(nworkers, nmol);
    std::atomic<int> next{0};
    std::mutex mu;
    std::string first_error;
    auto worker = [&]() {
        m2_ctx* ctx = nullptr;
        if (m2_ctx_create(&ctx, device) != 0 || !ctx) {
            std::lock_guard<std::mutex> lk(mu);
            if (first_error.empty())
                first_error = std::string("m2_rhf_batch: context creation failed: ") +
                              (ctx ? m2_last_error(ctx) : "no context");
            if (ctx) m2_ctx_destroy(ctx);
            return;
        }
        for (int k; (k = next.fetch_add(1)) < nmol;) {
            const int m = order[k];
            const long a0 = a_off[m], s0 = s_off[m], p0 = p_off[m];
            const long ms0 = ms_off[m], mp0 = mp_off[m];
            const int st = m2_rhf(ctx, natom[m], Z + a0, xyz + 3 * a0, charge ? charge[m] : 0, 1, pure, nsh[m], l + s0,
                                  nprim + s0, exps + p0, coefs + p0, centers + 3 * s0, atoms + s0, 0, nullptr, nullptr,
                                  nullptr, nullptr, nullptr, nullptr, minimal ? nsh_min[m] : 0,
                                  minimal ? l_min + ms0 : nullptr, minimal ? nprim_min + ms0 : nullptr,
                                  minimal ? exps_min + mp0 : nullptr, minimal ? coefs_min + mp0 : nullptr,
                                  minimal ? centers_min + 3 * ms0 : nullptr, minimal ? atoms_min + ms0 : nullptr, o,
                                  &res[m], nullptr, nullptr, nullptr, nullptr, nullptr);
            if (st != 0) {
                std::lock_guard<std::mutex> lk(mu);
                if (first_error.empty()) first_error = "m2_rhf_batch: molecule " + std::to_string(m) + ": " + g_rhf_error;
                res[m].converged = 0;
                res[m].iterations = -1;
            }
        }
        m2_ctx_destroy(ctx);
    };
    std::vector<std::thread> pool;
    for (int w = 0; w < nworkers; ++w) pool.emplace_back(worker);
    for (auto& t : pool) t.join();
    if (!first_error.empty()) {
        g_rhf_error = first_error;
        return 1;
    }
    return 0;
}

// The shell arrays of one basis (orbital, auxiliary or minimal) as m2_rhf
// takes them, with the centres of one atom's shells movable.
struct ShellArgs {
    int nsh = 0;
    const int *l = nullptr, *nprim = nullptr, *atoms = nullptr;
    const double *exps = nullptr, *coefs = nullptr, *centers = nullptr;
    std::vector<double> moved;   // centres after a displacement
    const double* centres_for(int atom, int axis, double d) {
        if (nsh <= 0) return centers;
        moved.assign(centers, centers + 3 * nsh);
        for (int s = 0; s < nsh; ++s)
            if (atoms[s] == atom) moved[3 * s + axis] += d;
        return moved.data();
    }
};

int m2_rhf_hessian(int device, int nworkers, int natom, const double* Z, const double* xyz, int charge,
                   int multiplicity, int pure, int nsh, const int* l, const int* nprim, const double* exps,
                   const double* coefs, const double* centers, const int* atoms, int nsh_aux, const int* l_aux,
                   const int* nprim_aux, const double* exps_aux, const double* coefs_aux, const double* centers_aux,
                   const int* atoms_aux, int nsh_min, const int* l_min, const int* nprim_min, const double* exps_min,
                   const double* coefs_min, const double* centers_min, const int* atoms_min,
                   const m2_scf_options_c* o, double step, double* hess, m2_scf_result_c* res) {
    g_rhf_error.clear();
    if (natom <= 0 || !Z || !xyz || nsh <= 0 || !l || !nprim || !exps || !coefs || !centers || !atoms || !o ||
        !hess || (nsh_aux > 0 && (!l_aux || !nprim_aux || !exps_aux || !coefs_aux || !centers_aux || !atoms_aux)) ||
        (nsh_min > 0 && (!l_min || !nprim_min || !exps_min || !coefs_min || !centers_min || !atoms_min))) {
        g_rhf_error = "m2_rhf_hessian: null or empty argument";
        return 1;
    }
    const double h = step > 0 ? step : 0.01;
    const int ncoord = 3 * natom, njob = 2 * ncoord;
    // job k: atom k / 6, axis (k / 2) % 3, +h for even k, -h for odd k
    std::vector<double> g((size_t)njob * ncoord, 0.0);
    if (nworkers <= 0) nworkers = 8;
    nworkers = std::min(nworkers, njob);
    std::atomic<int> next{0};
    std::mutex mu;
    std::string first_error;
    auto worker = [&]() {
        m2_ctx* ctx = nullptr;
        if (m2_ctx_create(&ctx, device) != 0 || !ctx) {
            std::lock_guard<std::mutex> lk(mu);
            if (first_error.empty())
                first_error = std::string("m2_rhf_hessian: context creation failed: ") +
                              (ctx ? m2_last_error(ctx) : "no context");
            if (ctx) m2_ctx_destroy(ctx);
            return;
        }
        ShellArgs orb{nsh, l, nprim, atoms, exps, coefs, centers, {}};
        ShellArgs aux{nsh_aux, l_aux, nprim_aux, atoms_aux, exps_aux, coefs_aux, centers_aux, {}};
        ShellArgs mn{nsh_min, l_min, nprim_min, atoms_min, exps_min, coefs_min, centers_min, {}};
        std::vector<double> X(xyz, xyz + ncoord);
        for (int k; (k = next.fetch_add(1)) < njob;) {
            {
                std::lock_guard<std::mutex> lk(mu);
                if (!first_error.empty()) break;
            }
            const int atom = k / 6, axis = (k / 2) % 3;
            const double d = (k & 1) ? -h : h;
            X.assign(xyz, xyz + ncoord);
            X[3 * atom + axis] += d;
            m2_scf_result_c r{};
            const int st = m2_rhf(ctx, natom, Z, X.data(), charge, multiplicity, pure, orb.nsh, orb.l, orb.nprim,
                                  orb.exps, orb.coefs, orb.centres_for(atom, axis, d), orb.atoms, aux.nsh, aux.l,
                                  aux.nprim, aux.exps, aux.coefs, aux.centres_for(atom, axis, d), aux.atoms, mn.nsh,
                                  mn.l, mn.nprim, mn.exps, mn.coefs, mn.centres_for(atom, axis, d), mn.atoms, o, &r,
                                  nullptr, nullptr, nullptr, nullptr, g.data() + (size_t)k * ncoord);
            if (res) res[k] = r;
            if (st != 0) {
                std::lock_guard<std::mutex> lk(mu);
                if (first_error.empty())
                    first_error = "m2_rhf_hessian: displacement " + std::to_string(k) + ": " + g_rhf_error;
            }
        }
        m2_ctx_destroy(ctx);
    };
    std::vector<std::thread> pool;
    for (int w = 0; w < nworkers; ++w) pool.emplace_back(worker);
    for (auto& t : pool) t.join();
    if (!first_error.empty()) {
        g_rhf_error = first_error;
        return 1;
    }
    // row 3i+x: central difference of the whole gradient in the displacement
    // of atom i along x; entries within 1e-12 of zero are set to zero and the
    // lower triangle is then taken from the upper one
    for (int r = 0; r < ncoord; ++r)
        for (int c = 0; c < ncoord; ++c) {
            const double v = (g[(size_t)(2 * r) * ncoord + c] - g[(size_t)(2 * r + 1) * ncoord + c]) / (2.0 * h);
            hess[(size_t)r * ncoord + c] = std::fabs(v) > 1e-12 ? v : 0.0;
        }
    for (int r = 0; r < ncoord; ++r)
        for (int c = 0; c < r; ++c) hess[(size_t)r * ncoord + c] = hess[(size_t)c * ncoord + r];
    return 0;
}

static void copy_shells(const std::vector<molecular::RawShell>& sh, m2_shells_out* o) {
    if (!o || !o->l) return;
    long p = 0;
    for (size_t s = 0; s < sh.size(); ++s) {
        o->l[s] = sh[s].l;
        if (o->nprim) o->nprim[s] = (int)sh[s].exps.size();
        if (o->atoms) o->atoms[s] = sh[s].atom;
        for (int k = 0; k < 3; ++k)
            if (o->centers) o->centers[3 * s + k] = sh[s].O[k];
        for (size_t i = 0; i < sh[s].exps.size(); ++i, ++p) {
            if (o->exps) o->exps[p] = sh[s].exps[i];
            if (o->coefs) o->coefs[p] = sh[s].coefs[i];
        }
    }
}

static int nprim_of(const std::vector<molecular::RawShell>& sh) {
    int n = 0;
    for (const auto& s : sh) n += (int)s.exps.size();
    return n;
}

// The molecule construct number `imol` of a deck, with its orbital, jkbasis
// and minimal (STO-3G, for the SOAD guess) shells read from the library.
struct DeckData {
    molecular::DeckMolecule mol;
    std::vector<molecular::RawShell> orb, aux, min;
};

static DeckData load_deck_molecule(const std::string& text, int imol, const std::string& lib) {
    using namespace molecular;
    const Deck deck = parse_deck(text);
    if (imol < 0 || imol >= (int)deck.constructs.size())
        throw std::runtime_error("INPUT: the deck has no molecule construct " + std::to_string(imol));
    DeckData d;
    d.mol = deck_molecule(deck.constructs[imol]);
    d.orb = library_shells(lib, "basis", d.mol.atoms, d.mol.basis);
    if (!d.mol.jkbasis.empty()) d.aux = library_shells(lib, "jkfit", d.mol.atoms, d.mol.jkbasis);
    d.min = library_shells(lib, "basis", d.mol.atoms, std::map<int, std::string>{{0, "sto-3g"}});
    return d;
}

int m2_deck_molecule(const char* deck_text, int imol, const char* basislib, m2_deck_info* info, double* Z,
                     double* xyz, m2_shells_out* orb, m2_shells_out* aux, m2_shells_out* minimal) {
    g_rhf_error.clear();
    try {
        if (!deck_text || !basislib || !info) throw std::runtime_error("m2_deck_molecule: null argument");
        const DeckData d = load_deck_molecule(deck_text, imol, basislib);
        info->natom = (int)d.mol.atoms.size();
        info->charge = d.mol.charge;
        info->multiplicity = d.mol.multiplicity;
        info->nsh = (int)d.orb.size();
        info->nprim = nprim_of(d.orb);
        info->nsh_aux = (int)d.aux.size();
        info->nprim_aux = nprim_of(d.aux);
        info->nsh_min = (int)d.min.size();
        info->nprim_min = nprim_of(d.min);
        for (size_t a = 0; a < d.mol.atoms.size(); ++a) {
            if (Z) Z[a] = d.mol.atoms[a].Z;
            if (xyz) {
                xyz[3 * a] = d.mol.atoms[a].x;
                xyz[3 * a + 1] = d.mol.atoms[a].y;
                xyz[3 * a + 2] = d.mol.atoms[a].z;
            }
        }
        copy_shells(d.orb, orb);
        copy_shells(d.aux, aux);
        copy_shells(d.min, minimal);
        return 0;
    } catch (const std::exception& e) {
        g_rhf_error = e.what();
        return 1;
    }
}

int m2_deck_run(const char* deck_text, const char* basislib, int device, int pure, m2_deck_result* res, int maxres,
                int* nres) {
    using namespace molecular;
    g_rhf_error.clear();
    m2_ctx* ctx = nullptr;
    try {
        if (!deck_text || !basislib || !nres) throw std::runtime_error("m2_deck_run: null argument");
        *nres = 0;
        const Deck deck = parse_deck(deck_text);
        if (m2_ctx_create(&ctx, device) != 0 || !ctx)
            throw std::runtime_error(std::string("m2_deck_run: ") + (ctx ? m2_last_error(ctx) : "no context"));
        for (size_t ic = 0; ic < deck.commands.size(); ++ic) {
            const DeckCommand& cmd = deck.commands[ic];
            m2_deck_result r{};
            r.command = (int)ic;
            r.molecule = cmd.molecule;
            r.method = -1;
            if ((cmd.name == "hf" || cmd.name == "rhf" || cmd.name == "uhf") && cmd.molecule >= 0) {
                const DeckData d = load_deck_molecule(deck_text, cmd.molecule, basislib);
                // printeris / intfile / thrint (src/ProgramController.cpp:114-118):
                // the molecule's two-electron integrals in the printERI format
                const std::string* pe = deck.global("printeris");
                if (pe && (*pe == "true" || *pe == "1")) {
                    const std::string* fname = deck.global("intfile");
                    const std::string* thr = deck.global("thrint");
                    std::vector<int> l, np, at;
                    std::vector<double> ex, cf, cen;
                    for (const RawShell& s : d.orb) {
                        l.push_back(s.l);
                        np.push_back((int)s.exps.size());
                        at.push_back(s.atom);
                        ex.insert(ex.end(), s.exps.begin(), s.exps.end());
                        cf.insert(cf.end(), s.coefs.begin(), s.coefs.end());
                        cen.insert(cen.end(), {s.O[0], s.O[1], s.O[2]});
                    }
                    if (m2_basis_set(ctx, M2_ORBITAL, (int)l.size(), l.data(), np.data(), ex.data(), cf.data(),
                                     cen.data(), at.data(), pure) != 0)
                        throw std::runtime_error(std::string("m2_deck_run: ") + m2_last_error(ctx));
                    print_eri_ctx(ctx, fname ? *fname : "eris.ints", thr ? std::stod(*thr) : 1e-12, nullptr, nullptr);
                }
                int nel = -d.mol.charge;
                for (const Atom& a : d.mol.atoms) nel += a.Z;
                // ProgramController::call_hf (src/ProgramController.cpp:288-293)
                const bool uhf = cmd.name == "uhf" || (cmd.name == "hf" && (d.mol.multiplicity > 1 || nel % 2 != 0));
                // call_rhf / call_uhf defaults (src/ProgramController.cpp:295-326)
                m2_scf_options_c o{};
                o.diis = 1;
                o.maxdiis = 8;
                o.maxiter = 40;
                o.enconverge = 1e-7;
                o.densconverge = 1e-4;
                o.precision = 1e-12;
                o.df = 0;
                o.mixed = 0;
                o.mixed_threshold = 1e-4;
                o.fp64_final = 2;
                o.guess = 1;   // soad
                o.diis_start = 2;
                if (const std::string* v = cmd.find("diis")) o.diis = std::stoi(*v) != 0;
                if (const std::string* v = cmd.find("maxdiis")) o.maxdiis = std::stoi(*v);
                if (const std::string* v = cmd.find("maxiter")) o.maxiter = std::stoi(*v);
                if (const std::string* v = cmd.find("enconverge")) o.enconverge = std::stod(*v);
                if (const std::string* v = cmd.find("densconverge")) o.densconverge = std::stod(*v);
                if (const std::string* v = cmd.find("precision")) o.precision = std::stod(*v);
                if (const std::string* v = cmd.find("df")) o.df = std::stoi(*v) != 0;
                if (const std::string* v = cmd.find("guess")) o.guess = *v == "soad" ? 1 : 0;
                // "direct, false" (the reference's default, src/ProgramController.cpp:116)
                // selects the in-core path; this driver runs direct unless the deck says so
                if (const std::string* v = deck.global("direct")) o.incore = *v == "false" || *v == "0";
                if (o.df && d.aux.empty()) throw std::runtime_error("m2_deck_run: df needs a jkbasis construct");
                std::vector<double> Z, xyz;
                for (const Atom& a : d.mol.atoms) {
                    Z.push_back(a.Z);
                    xyz.insert(xyz.end(), {a.x, a.y, a.z});
                }
                struct Flat {
                    std::vector<int> l, np, at;
                    std::vector<double> ex, cf, cen;
                };
                auto flat = [](const std::vector<RawShell>& sh) {
                    Flat f;
                    for (const RawShell& s : sh) {
                        f.l.push_back(s.l);
                        f.np.push_back((int)s.exps.size());
                        f.at.push_back(s.atom);
                        f.ex.insert(f.ex.end(), s.exps.begin(), s.exps.end());
                        f.cf.insert(f.cf.end(), s.coefs.begin(), s.coefs.end());
                        f.cen.insert(f.cen.end(), {s.O[0], s.O[1], s.O[2]});
                    }
                    return f;
                };
                const Flat fo = flat(d.orb), fa = flat(d.aux), fm = flat(d.min);
                m2_scf_result_c sr{};
                const int na = o.df ? (int)d.aux.size() : 0, nm = o.guess == 1 ? (int)d.min.size() : 0;
                const int st =
                    uhf ? m2_uhf(ctx, (int)Z.size(), Z.data(), xyz.data(), d.mol.charge, d.mol.multiplicity, pure,
                                 (int)d.orb.size(), fo.l.data(), fo.np.data(), fo.ex.data(), fo.cf.data(),
                                 fo.cen.data(), fo.at.data(), na, fa.l.data(), fa.np.data(), fa.ex.data(),
                                 fa.cf.data(), fa.cen.data(), fa.at.data(), nm, fm.l.data(), fm.np.data(),
                                 fm.ex.data(), fm.cf.data(), fm.cen.data(), fm.at.data(), &o, &sr, nullptr, nullptr,
                                 nullptr, nullptr, nullptr)
                        : m2_rhf(ctx, (int)Z.size(), Z.data(), xyz.data(), d.mol.charge, d.mol.multiplicity, pure,
                                 (int)d.orb.size(), fo.l.data(), fo.np.data(), fo.ex.data(), fo.cf.data(),
                                 fo.cen.data(), fo.at.data(), na, fa.l.data(), fa.np.data(), fa.ex.data(),
                                 fa.cf.data(), fa.cen.data(), fa.at.data(), nm, fm.l.data(), fm.np.data(),
                                 fm.ex.data(), fm.cf.data(), fm.cen.data(), fm.at.data(), &o, &sr, nullptr, nullptr,
                                 nullptr, nullptr, nullptr);
                if (st != 0) throw std::runtime_error("m2_deck_run: command " + cmd.name + ": " + g_rhf_error);
                r.method = uhf ? 1 : 0;
                r.energy = sr.energy;
                r.iterations = sr.iterations;
                r.converged = sr.converged;
                r.seconds = sr.total_seconds;
            }
            if (res && *nres < maxres) res[*nres] = r;
            ++*nres;
        }
        m2_ctx_destroy(ctx);
        return 0;
    } catch (const std::exception& e) {
        if (ctx) m2_ctx_destroy(ctx);
        g_rhf_error = e.what();
        return 1;
    }
}

}  // extern "C"
