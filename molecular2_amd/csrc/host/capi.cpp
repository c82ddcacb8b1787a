// extern "C" driver of the host C++ SCF (molecular::SCF::rhf), for callers
// that cannot link C++ (the Python tests and bench use it through ctypes).
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstring>
#include <exception>
#include <mutex>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

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

extern "C" {

const char* m2_rhf_last_error(void) { return g_rhf_error.c_str(); }

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

// Batched RHF: molecule m runs m2_rhf on the context of whichever worker
// takes it; workers take molecules from one queue ordered by basis size,
// largest first, so the long runs start early and the short ones fill in.
int m2_rhf_batch(int device, int nworkers, int nmol, const int* natom, const int* charge, const double* Z,
                 const double* xyz, int pure, const int* nsh, const int* l, const int* nprim, const double* exps,
                 const double* coefs, const double* centers, const int* atoms, const m2_scf_options_c* o,
                 m2_scf_result_c* res) {
    g_rhf_error.clear();
    if (nmol < 0 || (nmol > 0 && (!natom || !Z || !xyz || !nsh || !l || !nprim || !exps || !coefs || !centers ||
                                  !atoms || !o || !res))) {
        g_rhf_error = "m2_rhf_batch: null argument";
        return 1;
    }
    if (o->guess != 0) {
        g_rhf_error = "m2_rhf_batch: only the core-Hamiltonian guess (guess = 0) is available in batch mode";
        return 1;
    }
    if (nmol == 0) return 0;
    // offsets of every molecule into the concatenated arrays
    std::vector<long> a_off(nmol + 1, 0), s_off(nmol + 1, 0), p_off(nmol + 1, 0);
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
            const int st = m2_rhf(ctx, natom[m], Z + a0, xyz + 3 * a0, charge ? charge[m] : 0, 1, pure, nsh[m], l + s0,
                                  nprim + s0, exps + p0, coefs + p0, centers + 3 * s0, atoms + s0, 0, nullptr, nullptr,
                                  nullptr, nullptr, nullptr, nullptr, 0, nullptr, nullptr, nullptr, nullptr, nullptr,
                                  nullptr, o, &res[m], nullptr, nullptr, nullptr, nullptr, nullptr);
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

}  // extern "C"
