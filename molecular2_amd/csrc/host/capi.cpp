// extern "C" driver of the host C++ SCF (molecular::SCF::rhf), for callers
// that cannot link C++ (the Python tests and bench use it through ctypes).
#include <chrono>
#include <cstring>
#include <exception>
#include <string>

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

}  // extern "C"
