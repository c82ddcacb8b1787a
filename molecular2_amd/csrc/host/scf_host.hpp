// Host C++ mirror of the reference's SCF call surface, calling the MI355X
// engine through the C ABI (include/m2.h):
//   IntegralEngine  (inc/integrals.hpp:96-352)   -> m2_basis_set / m2_one_body / m2_schwarz
//   Fock            (inc/fock.hpp:97-213)         -> m2_jk_direct / m2_jk_df / m2_eigh / m2_gemm
//   DIISEngine      (inc/diis.hpp, src/diis.cpp)
//   SCF             (inc/scf.hpp:36-57, src/scf.cpp)
// Matrices are column-major FP64 like the reference's Eigen Matrix
// (inc/eigen_wrapper.hpp:8).  Errors are thrown as Error(code, msg)
// (inc/error.hpp:18-35).
#pragma once
#include <chrono>
#include <cmath>
#include <stdexcept>
#include <string>
#include <vector>

#include "m2.h"

namespace molecular {

class Error : public std::runtime_error {
public:
    Error(const std::string& code, const std::string& msg) : std::runtime_error(code + ": " + msg), code(code), msg(msg) {}
    std::string code, msg;
};

// column-major dense matrix
class Matrix {
public:
    Matrix() = default;
    Matrix(int r, int c, double v = 0.0) : r_(r), c_(c), a_((size_t)r * c, v) {}
    static Matrix Zero(int r, int c) { return Matrix(r, c, 0.0); }
    static Matrix Identity(int n) {
        Matrix m(n, n);
        for (int i = 0; i < n; ++i) m(i, i) = 1.0;
        return m;
    }
    int rows() const { return r_; }
    int cols() const { return c_; }
    size_t size() const { return a_.size(); }
    double* data() { return a_.data(); }
    const double* data() const { return a_.data(); }
    double& operator()(int i, int j) { return a_[(size_t)j * r_ + i]; }
    double operator()(int i, int j) const { return a_[(size_t)j * r_ + i]; }
    double& operator[](size_t k) { return a_[k]; }
    double operator[](size_t k) const { return a_[k]; }
    Matrix transpose() const;
    Matrix operator+(const Matrix& o) const;
    Matrix operator-(const Matrix& o) const;
    Matrix& operator+=(const Matrix& o);
    Matrix& operator-=(const Matrix& o);
    Matrix operator*(double s) const;
    Matrix operator-() const { return (*this) * -1.0; }
    double norm() const;                      // Frobenius
    double dot(const Matrix& o) const;        // sum_ij A_ij B_ij
    Matrix block(int r0, int c0, int nr, int nc) const;

private:
    int r_ = 0, c_ = 0;
    std::vector<double> a_;
};
inline Matrix operator*(double s, const Matrix& m) { return m * s; }
using Vector = std::vector<double>;

struct Atom {
    int Z = 0;
    double x = 0, y = 0, z = 0;   // bohr
};

struct RawShell {
    int l = 0, atom = -1;
    double O[3] = {0, 0, 0};
    std::vector<double> exps, coefs;   // raw contraction coefficients
};

class Molecule {
public:
    std::vector<Atom> atoms;
    std::vector<RawShell> shells;      // orbital basis, reference shell order
    std::vector<RawShell> jkshells;    // DF auxiliary basis (may be empty)
    std::vector<RawShell> minshells;   // STO-3G basis of the SOAD guess (may be empty)
    bool pure = true;
    int charge = 0, multiplicity = 1;
    int getNel() const;
    // Molecule::nalpha / nbeta (src/molecule.cpp:425-439)
    int nalpha() const { return (getNel() - multiplicity + 1) / 2 + multiplicity - 1; }
    int nbeta() const { return (getNel() - multiplicity + 1) / 2; }
    int getNAtoms() const { return (int)atoms.size(); }
    int getMultiplicity() const { return multiplicity; }
    double getEnuc() const;            // Molecule::calcEnuc, src/molecule.cpp:443-457
};

// Command options with the reference's RHF defaults
// (src/ProgramController.cpp:295-307) plus the engine's precision policy.
struct Options {
    bool diis = true;
    int maxdiis = 8;
    int maxiter = 40;
    double enconverge = 1e-7;
    double densconverge = 1e-4;
    double precision = 1e-12;
    bool df = false;
    bool direct = true;
    // "soad" (superposition of atomic densities, the reference's default for
    // RHF and UHF, src/ProgramController.cpp:307,322) or "hcore"
    std::string guess = "soad";
    // DIIS extrapolation from SCF iteration diis_start + 1 on (2: the current
    // reference source, src/scf.cpp:148,154; 0 reproduces the build that
    // wrote the reference's recorded logs)
    int diis_start = 2;
    // mixed precision: full builds with FP32 quartets below mixed_threshold,
    // incremental Delta-D builds entirely FP32, the last fp64_final
    // iterations full FP64 builds.
    bool mixed = false;
    double mixed_threshold = 1e-4;
    int fp64_final = 2;
    // the mixed phase hands over to FP64 once |dE| and ||dD|| fall below
    // these (FP32 noise floors), or after maxiter/2 iterations
    double handover_energy = 1e-5;
    double handover_dens = 1e-3;
};

class IntegralEngine {
public:
    IntegralEngine(m2_ctx* ctx, const Molecule& m);
    Matrix& getOverlap() { return sints; }
    Matrix& getKinetic() { return tints; }
    Matrix& getNucAttract() { return naints; }
    Matrix& getPrescreen() { return prescreen; }
    Matrix compute_shellblock_norm(const Matrix& A);
    int nbasis() const { return nbf; }
    int nauxiliary() const { return naux; }
    m2_ctx* context() const { return ctx; }

private:
    m2_ctx* ctx;
    int nbf = 0, nsh = 0, naux = 0;
    Matrix sints, tints, naints, prescreen;
};

class DIISEngine {
public:
    void init(int maxDiis, bool useDiis, double damp = 0.02);
    Vector compute(std::vector<Vector>& errors);

private:
    Vector solve(const Matrix& B, int start);
    int maxDiis = 8;
    bool useDiis = true;
    double damping_factor = 0.02;
    std::vector<Vector> errs;
    Matrix lastB;
};

// Average STO-3G AO occupations of the neutral ground-state atom Z (NIST
// configurations, electrons smeared evenly over a subshell), in STO-3G shell
// order: the restatement of sto3g_ao_occupation_vector
// (inc/sto3g_atomic_density.hpp:56-110).  Z <= 53.
Vector sto3g_occupations(int Z);

// Two-electron SOAD Fock contribution in the orbital basis,
// G(uv) = sum_ls D(ls) [2 (uv|ls) - (ul|vs)] with D = diag(occupations) / 2
// over the minimal basis (compute_2body_fock_general, src/fockbuild.cpp:477-583,
// with a shell-diagonal D).  Runs on a temporary context whose basis is the
// orbital basis followed by the minimal one.
Matrix soad_two_electron(int device, const Molecule& mol, double precision);

// device linear algebra helpers
Matrix matmul(m2_ctx* ctx, const Matrix& A, const Matrix& B, bool tA = false, bool tB = false);
void eigh(m2_ctx* ctx, const Matrix& A, Vector& w, Matrix& V);

class Fock {
public:
    Fock(const Options& opt, IntegralEngine& ints, const Molecule& m);
    virtual ~Fock() = default;
    void formHCore();
    void formOrthog();
    void makeJK(Matrix& P, double multiplier = 1.0);
    void formJKdirect(const Matrix& Schwarz, Matrix& P, double multiplier = 1.0);
    void formJKdf(Matrix& Cocc, double multiplier = 1.0);
    // Fock::formJK (src/fockbuild.cpp:36-51): J and K from the stored unique
    // integrals (the in-core path, Options::direct false)
    void formJK(Matrix& P, double multiplier = 1.0);
    void makeFock();
    void makeFock(Matrix& P, double multiplier = 1.0);
    void addDiis();
    void average(const Vector& w);
    void transform(bool first = false);
    // Fock::compute_soad_guess (src/fock.cpp:174-225): focka = H + G[D_SOAD]
    void compute_soad_guess();
    void diagonalise();
    void makeDens();
    Matrix& getHCore() { return hcore; }
    Matrix& getFockAO() { return focka; }
    Matrix& getJK() { return jkints; }
    Matrix& getJ() { return jints; }
    Matrix& getK() { return kints; }
    Matrix& getDens() { return dens; }
    Matrix& getS() { return integrals.getOverlap(); }
    Matrix& getOrthog() { return orthog; }
    Matrix& getCP() { return CP; }
    Vector& getEps() { return eps; }
    int getNocc() const { return nocc; }
    m2_ctx* context() const { return integrals.context(); }
    // Fock::compute_forces (src/rhfgrad.cpp:15-89): dE/dR, natom x 3
    void compute_forces(const std::vector<Atom>& atoms, int nocc);
    Matrix& getForces() { return forces; }
    // Fock::compute_hessian (src/rhfgrad.cpp:91-359): the analytic RHF
    // Hessian with CPHF response; derivative integrals by central differences
    // of the engine's integrals (step bohr); the CPHF fixed point stops at
    // ||du|| < cphf_tol or after cphf_maxiter steps (src/cphf.cpp, the
    // reference passes 1e-4 and 30).  Leaves the context's basis at the
    // molecule's geometry.
    void compute_hessian(const std::vector<Atom>& atoms, int nocc, double step, double cphf_tol, int cphf_maxiter);
    Matrix& getHessian() { return hessian; }
    int cphf_unconverged = 0, cphf_iterations = 0;
    double grad_seconds = 0.0;

    // incremental Fock state (inc/fock.hpp:127-129)
    bool reset_incremental = false, started_incremental = false;
    double next_reset = 0.0, rms_error = 1.0, incremental_threshold = 1e-5;
    int last_reset = 0;
    // precision policy
    bool fp64_forced = false;       // set by SCF for the final iterations
    int n_full = 0, n_incremental = 0, n_fp32_builds = 0;
    double jk_seconds = 0.0;

protected:
    const Options opt;
    IntegralEngine& integrals;
    const Molecule& molecule;
    Matrix hcore, orthog, fockm, focka, fockinc, CP, dens, dens_diff, jints, kints, jkints, forces, hessian;
    Vector eps;
    std::vector<Matrix> focks;
    int nbfs = 0, nocc = 0, iter = 0, MAX = 8;
    bool diis, density_fitted, direct;
    std::vector<double> twoints;   // packed unique (ij|kl) of the in-core path
    double precision;
    bool last_build_incremental = false;
};

// UnrestrictedFock (inc/fock.hpp:215-273): alpha and beta Fock matrices,
// F_a = H + Ja + Jb - Ka, F_b = H + Ja + Jb - Kb (src/fockbuild.cpp:431-445),
// no incremental builds.
class UnrestrictedFock : public Fock {
public:
    UnrestrictedFock(const Options& opt, IntegralEngine& ints, const Molecule& m);
    void transform(bool first = false);
    void diagonalise();
    void makeDens();
    void makeJK();
    void formJKdirect(const Matrix& Schwarz, Matrix& Da, Matrix& Db, double multiplier = 1.0);
    void formJKdf(Matrix& ca_occ, Matrix& cb_occ, double multiplier = 1.0);
    void makeFock();
    void average(const Vector& w);
    Matrix& getFockAlphaAO() { return fock_alpha_ao; }
    Matrix& getFockBetaAO() { return fock_beta_ao; }
    Matrix& getDensAlpha() { return dens_alpha; }
    Matrix& getDensBeta() { return dens_beta; }
    Matrix& getJAlpha() { return jints_alpha; }
    Matrix& getJBeta() { return jints_beta; }
    Matrix& getKAlpha() { return kints_alpha; }
    Matrix& getKBeta() { return kints_beta; }
    Matrix& getCPAlpha() { return CP_alpha; }
    Matrix& getCPBeta() { return CP_beta; }
    Vector& getEpsAlpha() { return eps_alpha; }
    Vector& getEpsBeta() { return eps_beta; }
    int getNalpha() const { return nalpha; }
    int getNbeta() const { return nbeta; }

protected:
    Matrix fock_alpha_ao, fock_beta_ao, fock_alpha_mo, fock_beta_mo, dens_alpha, dens_beta;
    Matrix kints_alpha, kints_beta, jints_alpha, jints_beta, CP_alpha, CP_beta;
    Vector eps_alpha, eps_beta;
    std::vector<Matrix> alpha_focks, beta_focks;
    int nalpha = 0, nbeta = 0;
};

struct IterationRecord {
    int iter;
    double energy, delta_e, delta_d, seconds;
    int incremental, fp32;
};

class SCF {
public:
    SCF(const Options& opt, const Molecule& m, Fock& f);
    void calcE();
    double calcE(const Matrix& hcore, const Matrix& dens, const Matrix& fock);
    Vector calcErr();
    Vector calcErr(const Matrix& F, const Matrix& D, const Matrix& S, const Matrix& orthog);
    bool testConvergence(double val);
    void rhf(bool print = false);
    // SCF::uhf / uhf_internal (src/scf.cpp:187-272); the Fock object must be
    // an UnrestrictedFock
    void uhf(bool print = false);
    double getEnergy() const { return energy; }
    double getOneE() const { return one_E; }
    double getTwoE() const { return two_E; }
    bool converged = false;
    int iterations = 0;
    std::vector<IterationRecord> log;

private:
    const Options opt;
    const Molecule& molecule;
    Fock& focker;
    DIISEngine diis;
    double energy = 0, last_energy = 0, one_E = 0, two_E = 0, error = 0, last_err = 0;
};

}  // namespace molecular
