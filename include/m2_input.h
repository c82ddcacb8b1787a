/*
 * m2_input.h -- the reference's own input files for callers of this engine:
 * `.mol` input decks (ProgramController::parse, src/ProgramController.cpp:
 * 137-236; Molecule::init / parseGeomLine / parseBasis, src/molecule.cpp:
 * 20-240) and its XML basis library directory (`basislibrary/`: <kind>.list
 * and bases/build/<kind>N.xml, BasisReader, src/basisreader.cpp:27-132).
 * Host code only: m2_deck_molecule needs no GPU.
 */
#ifndef M2_INPUT_H
#define M2_INPUT_H

#include "m2_scf.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Sizes of one deck molecule: atoms, charge, multiplicity and the shell /
 * primitive counts of its orbital basis, its jkbasis (0 when the deck names
 * none) and the STO-3G minimal basis the SOAD guess uses. */
typedef struct {
    int natom, charge, multiplicity;
    int nsh, nprim;
    int nsh_aux, nprim_aux;
    int nsh_min, nprim_min;
} m2_deck_info;

/* Caller-owned shell arrays in the layout m2_basis_set / m2_rhf take (l,
 * nprim, atoms: nsh; exps, coefs: nprim total, raw coefficients; centers:
 * 3 nsh, bohr).  Any pointer may be NULL. */
typedef struct {
    int* l;
    int* nprim;
    double* exps;
    double* coefs;
    double* centers;
    int* atoms;
} m2_shells_out;

/* Molecule construct `imol` (0-based, in deck order) of the deck text:
 * nuclear charges Z and positions xyz (bohr; "angstrom" converts with
 * 1.889726124565) and its shells read from the basis library at `basislib`
 * (per element the basis the deck names, else its default, else sto-3g;
 * shells in atom order, library file order).  Call once with NULL arrays to
 * get the sizes in *info, then again with arrays of those sizes.  Returns 0,
 * or non-zero with m2_rhf_last_error() set (INPUT / INIT / BASISLIB errors,
 * as the reference names them). */
int m2_deck_molecule(const char* deck_text, int imol, const char* basislib, m2_deck_info* info, double* Z,
                     double* xyz, m2_shells_out* orb, m2_shells_out* aux, m2_shells_out* minimal);

/* One command of a deck run.  method: 0 RHF, 1 UHF, -1 not run here (a
 * post-HF or ALMO command of the reference). */
typedef struct {
    int command, molecule, method;
    double energy;
    int iterations, converged;
    double seconds;
} m2_deck_result;

/* Runs the deck's hf / rhf / uhf commands on `device` (hf chooses UHF for
 * open shells, ProgramController::call_hf), each with the options of its
 * command and the defaults of call_rhf / call_uhf (SOAD guess, DIIS 8,
 * maxiter 40, enconverge 1e-7, densconverge 1e-4, precision 1e-12, df with
 * the jkbasis).  pure: 1 spherical, 0 Cartesian functions (the reference's
 * recorded outputs use Cartesian d shells).  One result per deck command in
 * res (up to maxres); *nres = number of commands.  Returns 0, or non-zero
 * with m2_rhf_last_error() set. */
int m2_deck_run(const char* deck_text, const char* basislib, int device, int pure, m2_deck_result* res, int maxres,
                int* nres);

/* IntegralEngine::printERI (src/integrals.cpp:202-233): writes the
 * two-electron integrals of the context's orbital basis to `path` in the
 * reference's text format (header, one line "c1 c2 c3 c4 w*(c1c2|c3c4)" per
 * c1 >= c2, c1 >= c3 >= c4, 1-based, w the permutational weight, zero below
 * thrint; then the integral and insignificant counts).  The integrals come
 * from m2_eri_packed (so n <= 256).  nints / ninsig may be NULL.  A deck
 * with "printeris, true" writes this file (name "intfile", default
 * eris.ints; threshold "thrint", default 1e-12) for each molecule an hf /
 * rhf / uhf command runs.  Returns 0, or non-zero with m2_rhf_last_error(). */
int m2_print_eri(m2_ctx* ctx, const char* path, double thrint, long long* nints, long long* ninsig);

/* Shells of the atoms Z (positions xyz, bohr) from a Molpro basis text file
 * such as the reference's basissets/CC-PVDZfirst.basis: per atom in order,
 * the element's "c, i.j, ..." contraction lines in file order, one segmented
 * shell each, over the exponents of the preceding "<l>, <element>, ..."
 * line (the shells the reference's basis library holds for that file).
 * Call with out = NULL (or too small *nsh / *nprim) for the sizes; then with
 * arrays of those sizes.  Deck basis names ending in ".basis" are read the
 * same way.  Returns 0, or non-zero with m2_rhf_last_error() (BASISLIB). */
int m2_basis_file_shells(const char* path, int natom, const int* Z, const double* xyz, int* nsh, int* nprim,
                         m2_shells_out* out);

#ifdef __cplusplus
}
#endif
#endif
