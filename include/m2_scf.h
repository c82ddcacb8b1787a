/*
 * m2_scf.h -- extern "C" driver of the host C++ RHF (molecular::SCF::rhf over
 * the HIP engine), for callers that cannot link C++ (ctypes, other FFIs).
 *
 * C++ callers use the classes in molecular2_amd/csrc/host/scf_host.hpp
 * directly (IntegralEngine, Fock, SCF), which keep the reference's call
 * surface (inc/integrals.hpp:108, inc/fock.hpp:131-188, inc/scf.hpp:36-57).
 */
#ifndef M2_SCF_H
#define M2_SCF_H

#include "m2.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Options of SCF::rhf and Fock (defaults of the reference in
 * src/ProgramController.cpp:295-307 and src/fock.cpp:32-73):
 * diis 1, maxdiis 8, maxiter 40, enconverge 1e-7, densconverge 1e-4,
 * precision 1e-12 (formJKdirect screening), df 0 (1: density-fitted J/K with
 * the auxiliary shells), mixed 0 (1: FP32 quartets below mixed_threshold and
 * FP32 incremental builds, FP64 for the last fp64_final iterations). */
typedef struct {
    int diis, maxdiis, maxiter;
    double enconverge, densconverge, precision;
    int df, mixed;
    double mixed_threshold;
    int fp64_final;
    int guess;           /* 0: core Hamiltonian, 1: SOAD (needs the minimal basis) */
    int diis_start;      /* DIIS from iteration diis_start + 1 on (reference source: 2) */
    int incore;          /* 1: the reference's in-core path ("direct, false"): the
                            unique ERIs are stored once (m2_eri_packed) and J/K
                            formed from them (Fock::formJK, src/fockbuild.cpp:36-51)
                            for bases of at most 256 functions, else direct */
} m2_scf_options_c;

/* energy = one_e + two_e + enuc (SCF::calcE, src/scf.cpp:29-48); counts of
 * full / incremental / FP32 J/K builds; wall seconds of the J/K builds, the
 * whole call, and the gradient. */
typedef struct {
    double energy, one_e, two_e, enuc;
    int iterations, converged, n_full, n_incremental, n_fp32;
    double jk_seconds, total_seconds, grad_seconds;
} m2_scf_result_c;

/* SCF::rhf (src/scf.cpp:90-185) on a molecule given as natom nuclear charges
 * Z and positions xyz (bohr, 3*natom), charge and multiplicity, and its
 * orbital shells (nsh, l, nprim, exps, raw coefs, centers 3*nsh, atoms) plus
 * optional auxiliary shells (nsh_aux may be 0) and optional minimal-basis
 * (STO-3G) shells for the SOAD guess, grouped by atom in atom order (nsh_min
 * may be 0 unless o->guess == 1).  Outputs may be NULL: D
 * (n x n density), C (n x n MO coefficients, column-major), eps (n orbital
 * energies), log5 (maxiter x 5: energy, dE, dD, incremental, fp32 per
 * iteration), grad (natom x 3: Fock::compute_forces, src/rhfgrad.cpp:15-89).
 * Returns 0 on success; otherwise m2_rhf_last_error() says why.  A run that
 * does not converge returns 0 with res->converged = 0, as the reference logs
 * and continues. */
int m2_rhf(m2_ctx* ctx, int natom, const double* Z, const double* xyz, int charge, int multiplicity, int pure,
           int nsh, const int* l, const int* nprim, const double* exps, const double* coefs, const double* centers,
           const int* atoms, int nsh_aux, const int* l_aux, const int* nprim_aux, const double* exps_aux,
           const double* coefs_aux, const double* centers_aux, const int* atoms_aux, int nsh_min, const int* l_min,
           const int* nprim_min, const double* exps_min, const double* coefs_min, const double* centers_min,
           const int* atoms_min, const m2_scf_options_c* o, m2_scf_result_c* res, double* D, double* C, double* eps,
           double* log5, double* grad);
const char* m2_rhf_last_error(void);   /* of the calling thread */

/* SCF::uhf (src/scf.cpp:187-272) with UnrestrictedFock: arguments as for
 * m2_rhf (multiplicity = 2S + 1; n_alpha = (N - mult + 1) / 2 + mult - 1,
 * src/molecule.cpp:425-439).  Outputs may be NULL: Da, Db (n x n spin
 * densities C_s,occ C_s,occ^T), eps_a, eps_b (n), log5 (maxiter x 5:
 * energy, dE, dD per iteration from iteration 1).  J/K are direct (FP64) or,
 * with o->df, density fitted per spin. */
int m2_uhf(m2_ctx* ctx, int natom, const double* Z, const double* xyz, int charge, int multiplicity, int pure,
           int nsh, const int* l, const int* nprim, const double* exps, const double* coefs, const double* centers,
           const int* atoms, int nsh_aux, const int* l_aux, const int* nprim_aux, const double* exps_aux,
           const double* coefs_aux, const double* centers_aux, const int* atoms_aux, int nsh_min, const int* l_min,
           const int* nprim_min, const double* exps_min, const double* coefs_min, const double* centers_min,
           const int* atoms_min, const m2_scf_options_c* o, m2_scf_result_c* res, double* Da, double* Db,
           double* eps_a, double* eps_b, double* log5);

/* Batched SCF::rhf over nmol closed-shell molecules on one device.  No
 * reference counterpart: the reference runs one molecule per process
 * (src/ProgramController.cpp:112-120, one SCF::rhf per geometry).  Molecule m
 * has natom[m] atoms (Z, xyz concatenated over molecules) and nsh[m] shells (l,
 * nprim, atoms — indices local to the molecule — and centers concatenated over
 * molecules; exps/coefs over their primitives); charge may be NULL (all 0).
 * nworkers host threads (0: 8) each own an engine context with its own HIP
 * stream, so the SCFs of different molecules overlap on the GPU; molecules
 * are taken largest basis first.  res[m] is m2_rhf's result for molecule m
 * (iterations = -1 if its run failed).  Core-Hamiltonian guess only
 * (o->guess must be 0; m2_rhf_batch_min below takes the SOAD guess).
 * Returns 0 when every molecule ran (converged or
 * not); otherwise m2_rhf_last_error() names the first failure. */
int m2_rhf_batch(int device, int nworkers, int nmol, const int* natom, const int* charge, const double* Z,
                 const double* xyz, int pure, const int* nsh, const int* l, const int* nprim, const double* exps,
                 const double* coefs, const double* centers, const int* atoms, const m2_scf_options_c* o,
                 m2_scf_result_c* res);

/* m2_rhf_batch with each molecule's minimal (STO-3G) shells for the SOAD
 * guess (o->guess = 1, the reference's default, src/ProgramController.cpp:307):
 * nsh_min[m] shells per molecule, concatenated like the orbital shells and
 * grouped by atom in atom order (atoms_min local to the molecule).  With
 * nsh_min = NULL it is m2_rhf_batch. */
int m2_rhf_batch_min(int device, int nworkers, int nmol, const int* natom, const int* charge, const double* Z,
                     const double* xyz, int pure, const int* nsh, const int* l, const int* nprim, const double* exps,
                     const double* coefs, const double* centers, const int* atoms, const int* nsh_min,
                     const int* l_min, const int* nprim_min, const double* exps_min, const double* coefs_min,
                     const double* centers_min, const int* atoms_min, const m2_scf_options_c* o,
                     m2_scf_result_c* res);

/* Fock::compute_hessian_numeric (src/rhfgrad.cpp:362-416): the RHF nuclear
 * Hessian (3*natom x 3*natom, Eh/bohr^2) as central differences of the
 * analytic gradient, H(3i+x, :) = [g(R + h e_ix) - g(R - h e_ix)] / 2h, with
 * entries of magnitude <= 1e-12 set to zero and the lower triangle copied
 * from the upper one.  step = h in bohr (<= 0: the reference's 0.01).  Each
 * of the 6*natom displaced geometries runs a fresh m2_rhf (options o, same
 * shells with the displaced atom's centres moved) plus its gradient; nworkers
 * host threads (0: 8) each own an engine context and HIP stream on `device`,
 * so the displaced SCFs overlap on the GPU.  res (6*natom entries, may be
 * NULL) receives each displaced run's result, in the order atom, axis, +h
 * then -h.  Returns 0 when every run finished (converged or not, as the
 * reference logs and continues); otherwise m2_rhf_last_error() says why. */
int m2_rhf_hessian(int device, int nworkers, int natom, const double* Z, const double* xyz, int charge,
                   int multiplicity, int pure, int nsh, const int* l, const int* nprim, const double* exps,
                   const double* coefs, const double* centers, const int* atoms, int nsh_aux, const int* l_aux,
                   const int* nprim_aux, const double* exps_aux, const double* coefs_aux, const double* centers_aux,
                   const int* atoms_aux, int nsh_min, const int* l_min, const int* nprim_min, const double* exps_min,
                   const double* coefs_min, const double* centers_min, const int* atoms_min,
                   const m2_scf_options_c* o, double step, double* hess, m2_scf_result_c* res);

/* Fock::compute_hessian (src/rhfgrad.cpp:91-359) after SCF::rhf: the
 * analytic RHF nuclear Hessian H = H_nr + 2 H_r (3N x 3N, Eh/bohr^2, into
 * hess) with the coupled-perturbed Hartree-Fock response solved by the
 * reference's fixed-point iteration u <- Q + V u (cphf_group_solver,
 * src/cphf.cpp:4-34; the reference passes cphf_tol 1e-4, cphf_maxiter 30).
 * The derivative integrals are central differences (step bohr, e.g. 1e-4) of
 * the engine's analytic one-body and J/K integrals, and H_nr the central
 * difference of the analytic gradient at fixed density and energy-weighted
 * density; the MO-integral sums of the CPHF equations are direct J/K builds
 * of the perturbed densities.  Shells as for m2_rhf (no auxiliary basis:
 * the reference's Hessian needs the in-core SCF).  stats (optional, 2 ints):
 * CPHF solves that stopped at cphf_maxiter, CPHF steps in total. */
int m2_rhf_hessian_cphf(m2_ctx* ctx, int natom, const double* Z, const double* xyz, int charge, int multiplicity,
                        int pure, int nsh, const int* l, const int* nprim, const double* exps, const double* coefs,
                        const double* centers, const int* atoms, int nsh_min, const int* l_min, const int* nprim_min,
                        const double* exps_min, const double* coefs_min, const double* centers_min,
                        const int* atoms_min, const m2_scf_options_c* o, double step, double cphf_tol,
                        int cphf_maxiter, double* hess, m2_scf_result_c* res, int* stats);

/* The two-electron part of the SOAD guess Fock matrix (Fock::compute_soad_guess,
 * src/fock.cpp:174-225, with compute_2body_fock_general, src/fockbuild.cpp:
 * 477-583): G (n x n, orbital basis) = sum_ls D(ls) [2 (uv|ls) - (ul|vs)]
 * with D = diag(STO-3G ground-state occupations) / 2 over the minimal basis.
 * Shells as for m2_rhf.  Returns 0 on success (m2_rhf_last_error otherwise). */
int m2_soad_fock(int device, int natom, const double* Z, const double* xyz, int pure, int nsh, const int* l,
                 const int* nprim, const double* exps, const double* coefs, const double* centers, const int* atoms,
                 int nsh_min, const int* l_min, const int* nprim_min, const double* exps_min, const double* coefs_min,
                 const double* centers_min, const int* atoms_min, double* G);

#ifdef __cplusplus
}
#endif
#endif
