/*
 * m2.h -- C ABI of the MI355X RHF Fock-build engine (libm2.so).
 *
 * This is the thin extern "C" layer between the host C++ Fock /
 * IntegralEngine / SCF classes (molecular2_amd/csrc/host) and the HIP kernels.
 * Plain pointers and sizes only.
 *
 * Conventions (the reference's, src/fockbuild.cpp and inc/eigen_wrapper.hpp:8):
 *   - matrices are n x n FP64; all of them are symmetric, so row- and
 *     column-major storage coincide (the unsymmetrised intermediates never
 *     leave the library);
 *   - D is the full closed-shell density 2 C_occ C_occ^T, or a difference
 *     Delta-D; J = sum (uv|ls) D_ls and K = sum (ul|vs) D_ls, so the Fock
 *     contribution is J - K/2 (src/fockbuild.cpp:161-163);
 *   - the DF path takes unscaled C_occ (n x nocc, column-major) and returns J
 *     and K over C C^T (src/fockbuild.cpp:171-174: jkints = 2J - K);
 *   - basis functions: shells in the caller's order, Cartesian components in
 *     the standard order (xx, xy, xz, yy, yz, zz, ...), pure shells m=-l..l
 *     for l >= 2, p shells always x, y, z; contraction coefficients are passed
 *     raw and normalised inside as libint2::Shell does (src/basis.cpp:33-66).
 *
 * Host buffers belong to the caller, device buffers to the context.  Calls
 * are synchronous on return unless named *_device (those are ordered on the
 * context's stream and also synchronous on return).  One host thread per
 * context.  A non-zero status means failure; m2_last_error() says why (the
 * C++ wrapper turns it into Error(code, msg), inc/error.hpp:18-35).
 */
#ifndef M2_H
#define M2_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct m2_ctx m2_ctx;
typedef int m2_status;

enum { M2_OK = 0, M2_ERR_ARG = 1, M2_ERR_HIP = 2, M2_ERR_STATE = 3, M2_ERR_NUMERIC = 4 };

/* J/K precision flags */
enum {
    M2_FP64 = 0,       /* every quartet in FP64 */
    M2_MIXED = 1,      /* quartets with bound < fp32 threshold in FP32 */
};

/* which basis */
enum { M2_ORBITAL = 0, M2_AUXILIARY = 1 };

/* Context: device selection, stream, device buffers.  No reference
 * counterpart (the reference is single-process, single-threaded CPU code);
 * it owns what IntegralEngine and Fock own in src/integrals.cpp:42-124 and
 * src/fock.cpp:32-73. */
m2_status m2_ctx_create(m2_ctx** ctx, int device);
void m2_ctx_destroy(m2_ctx* ctx);
const char* m2_last_error(const m2_ctx* ctx);
/* the HIP stream (hipStream_t) every call of this context runs on */
void* m2_ctx_stream(m2_ctx* ctx);
/* the device the context was created on */
int m2_ctx_device(const m2_ctx* ctx);

/* Basis.  Replaces Basis::addShell + libint2::Shell construction
 * (src/basis.cpp:33-66) and IntegralEngine::map_shell_to_basis_function /
 * nbasis (src/integrals.cpp:329-371).  exps/coefs hold sum(nprim) values;
 * centers 3*nsh (bohr); atoms nsh.  For M2_ORBITAL this also builds the
 * shell-pair data and the Schwarz bounds on the device. */
m2_status m2_basis_set(m2_ctx* ctx, int which, int nsh, const int* l, const int* nprim,
                       const double* exps, const double* coefs, const double* centers,
                       const int* atoms, int pure);
m2_status m2_basis_info(m2_ctx* ctx, int which, int* nbf, int* nsh);
/* normalised coefficients of one shell (libint2 convention), for tests */
m2_status m2_shell_coefs(m2_ctx* ctx, int which, int shell, double* out);

/* IntegralEngine::compute_schwarz_ints (src/integrals.cpp:471-513):
 * Q[nsh*nsh], Q(s1,s2) = sqrt(max |(ab|ab)|). */
m2_status m2_schwarz(m2_ctx* ctx, double* Q);

/* IntegralEngine::compute_shellblock_norm (src/integrals.cpp:451-469). */
m2_status m2_shellblock_norm(m2_ctx* ctx, const double* A, double* out);

/* IntegralEngine::compute_1body_ints for overlap, kinetic and nuclear
 * attraction (src/integrals.cpp:394-449); point charges Z at xyz (bohr).
 * Any of S, T, V may be NULL. */
m2_status m2_one_body(m2_ctx* ctx, int natom, const double* Z, const double* xyz,
                      double* S, double* T, double* V);

/* Fock::formJKdirect (src/fockbuild.cpp:64-164): J and K of D, both
 * multiplied by `multiplier`, quartets screened by
 * max_6blocks(|D|) Q12 Q34 < precision.  flags: M2_FP64 or M2_MIXED. */
m2_status m2_jk_direct(m2_ctx* ctx, const double* D, double multiplier, double precision,
                       unsigned flags, double* J, double* K);
/* same on device buffers (n*n doubles each) */
m2_status m2_jk_direct_device(m2_ctx* ctx, const double* dD, double multiplier, double precision,
                              unsigned flags, double* dJ, double* dK);
/* m2_jk_direct_device without any allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device or pinned memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments, results and screening as m2_jk_direct_device;
 * the context's basis must already be set (m2_basis_set allocates).  The
 * per-build work memory (shell-block norms, unsymmetrised J/K accumulators,
 * work-list totals, heavy-class tasks) lives in one static device array of
 * M2_NOALLOC_SCRATCH_BYTES per GPU (held until the process ends), so the
 * basis may have at most ~2000 functions (2 n^2 doubles plus the tasks must
 * fit); larger inputs return M2_ERR_ARG with the size needed.  Every call on
 * one GPU shares that array: calls that overlap on one GPU (two contexts on
 * two host threads) race and must be serialised by the caller.  Profiling
 * (m2_set_profiling) is ignored on this path. */
#define M2_NOALLOC_SCRATCH_BYTES (64u << 20)
m2_status m2_jk_direct_device_noalloc(m2_ctx* ctx, const double* dD, double multiplier, double precision,
                                      unsigned flags, double* dJ, double* dK);
/* The two-electron integral tensor in the unique packed layout the
 * reference's in-core path stores (IntegralEngine::compute_eris,
 * src/eris.cpp:18-87, read by Fock::formJK, src/fockbuild.cpp:36-51, and
 * printed by IntegralEngine::printERI, src/integrals.cpp:202-233):
 * out[p(ij) * np + p(kl)] = (ij|kl) with p(ij) = i(i+1)/2 + j for i >= j
 * and np = n(n+1)/2, so out holds np^2 doubles.  Every quartet is computed
 * in FP64 without screening, one column (..|kl) per direct J build of a unit
 * density on the device.  For small bases (n <= 256). */
m2_status m2_eri_packed(m2_ctx* ctx, double* out);
/* UnrestrictedFock::formJKdirect (src/fockbuild.cpp:267-364): J and K of the
 * alpha and beta densities (Da = Ca_occ Ca_occ^T, Db likewise), every quartet
 * screened with the block norms of Da + Db; the unrestricted Fock matrices
 * are F_a = H + Ja + Jb - Ka and F_b = H + Ja + Jb - Kb
 * (UnrestrictedFock::makeFock, src/fockbuild.cpp:431-445). */
m2_status m2_jk_direct_uhf(m2_ctx* ctx, const double* Da, const double* Db, double multiplier, double precision,
                           unsigned flags, double* Ja, double* Ka, double* Jb, double* Kb);
/* quartets with Schwarz bound x block norm below this run in FP32 under
 * M2_MIXED (default 1e-4) */
m2_status m2_set_mixed_threshold(m2_ctx* ctx, double fp32_below);
/* s-s ket families for the direct J/K build, applied at the next
 * m2_basis_set.  No reference counterpart (libint computes every shell
 * quartet on its own): s shells of one centre with identical exponents share
 * their primitive pair integrals across the quartets they enter, with the
 * reference's screening and unique-quartet rule kept per contracted quartet.
 * mode 0: off; 1 (default): on when they save work, for (ss|ss) and (pp|ss)
 * always and for (ps|ss) and (ds|ss) when most family members are
 * significant together; 2: on for all four. */
m2_status m2_set_jk_families(m2_ctx* ctx, int mode);
/* Primitive screening of the direct J/K build.  Inside a shell quartet that
 * survives the screening of src/fockbuild.cpp:98-115, a primitive quartet is
 * skipped when q_ab q_cd max|D_blocks| < relative * precision, where q is the
 * primitive pair's own bound sqrt(max |(ab|ab)|) (so |[ab|cd]| <= q_ab q_cd).
 * The reference drops primitive quartets inside libint2 the same way, below
 * its engine precision min(precision/max|D|, eps)/K_max^4
 * (src/fockbuild.cpp:77-83).  Default relative = 1e-6 (each skipped
 * primitive quartet moves a J/K element by < 1e-18 at precision 1e-12);
 * 0 computes every primitive quartet. */
m2_status m2_set_prim_screening(m2_ctx* ctx, double relative);
/* The same threshold for the FP32 quartets of an M2_MIXED build (no
 * counterpart in the reference, which has no FP32 path).  Those quartets
 * are computed with ~6e-8 relative rounding error, so dropping primitive
 * quartets below relative * precision costs no accuracy they have.
 * Default 0.1; 0 <= relative <= 1. */
m2_status m2_set_prim_screening_fp32(m2_ctx* ctx, double relative);
/* Launch shape of the direct J/K class kernels (no counterpart in the
 * reference, whose Fock::formJKdirect walks the shell quartets on the host,
 * src/fockbuild.cpp:64-164).  host_totals = 0 (default): each class kernel's
 * grid is a bound on its work-item count that depends on the basis only,
 * and workgroups past the count the device list kernels wrote exit at once,
 * so a build reads nothing back before its end (and can be captured into a
 * graph).  host_totals = 1: the counts are read back first and every grid
 * is exact.  The results agree to rounding (the FP64 atomics sum in
 * another order). */
m2_status m2_set_jk_launch(m2_ctx* ctx, int host_totals);
/* Graph replay of the direct J/K builds (off by default; no counterpart in
 * the reference).  When on, a build with the bounded-grid launch and
 * profiling off is captured once into a HIP graph per key -- its device
 * pointers, multiplier, precision, flags, mixed threshold and primitive
 * screening -- and later builds with the same key launch that graph (one
 * launch instead of ~30).  Up to 8 graphs are kept, the least recently used
 * dropped first; a new orbital basis, a chunks-per-item change or turning
 * replay off drops them all.  The allocation-free entry point never uses
 * graphs.  Results equal those of the eager build to rounding. */
m2_status m2_set_jk_graphs(m2_ctx* ctx, int on);
/* graph captures and replays so far, and graphs held */
m2_status m2_jk_graph_stats(m2_ctx* ctx, long long* captures, long long* replays, int* cached);
/* Ket chunks per work item of class combination (X | Y), pair classes
 * ss=0 ps=1 pp=2 ds=3 dp=4 dd=5, 0 <= Y <= X; 0 restores the default (64
 * chunks of 64 kets for the register classes, 1 chunk of 4 kets for the
 * table-driven ones); values below the smallest the list capacity holds
 * (16 for the register classes) are raised to it.  Kept across basis sets. */
m2_status m2_set_jk_block_tasks(m2_ctx* ctx, int X, int Y, int bt);
/* Autotuner of the chunks per work item (no counterpart in the reference):
 * times `repeats` builds of the device density dD at precision/flags for
 * each of 16, 32, 64, 128, 256 chunks per item (register classes; 1, 2, 4,
 * 8, 16 for the table-driven ones), keeps the fastest per class
 * combination (as m2_set_jk_block_tasks would) and reports the mean build
 * time before and after (ms).  Uses the context's J/K output buffers;
 * profiling and graph settings are restored, the profile is cleared. */
m2_status m2_jk_autotune(m2_ctx* ctx, const double* dD, double precision, unsigned flags, int repeats,
                         double* ms_before, double* ms_after);
/* Current launch shape; block_tasks and grid_bound (optional, 36 entries,
 * index 6 X + Y): chunks per item and grid bound of every combination with
 * bra pairs in the current basis, else 0. */
m2_status m2_get_jk_launch(m2_ctx* ctx, int* host_totals, int* block_tasks, int* grid_bound);
/* statistics of the last J/K build: computed quartets (FP64, FP32; -1: not
 * counted), work items of the class kernels */
m2_status m2_last_stats(m2_ctx* ctx, long long* q64, long long* q32, long long* ntasks);

/* Density-fitted J/K: Fock::formJKdf / compute_2body_fock_df
 * (src/fockbuild.cpp:171-174,585-660) with the 3- and 2-centre integrals of
 * IntegralEngine::compute_eris_3index / compute_eris_2index
 * (src/eris.cpp:89-156,304-356).  Needs the M2_AUXILIARY basis (l <= 3).
 * m2_df_setup builds the fitted tensor (3c/2c integrals, FP64 Cholesky of the
 * metric, B = (xy|K) L^-T) on the device; m2_jk_df calls it on first use.
 * Cocc is n x nocc column-major (unscaled occupied orbitals); J and K are
 * over Cocc Cocc^T.  fp32 != 0: FP32 GEMMs with FP32 accumulation (MFMA);
 * fp32 == 0: FP64 GEMMs. */
m2_status m2_df_setup(m2_ctx* ctx);
m2_status m2_jk_df(m2_ctx* ctx, const double* Cocc, int nocc, int fp32, double* J, double* K);
m2_status m2_jk_df_device(m2_ctx* ctx, const double* dCocc, int nocc, int fp32, double* dJ, double* dK);
/* timing of the last DF call when profiling is on: ms of X = C^T B, of
 * K = X^T X, of the J part, and their sum */
m2_status m2_df_timing(m2_ctx* ctx, double* ms4);

/* One fragment of the fragment-local DF-JK (FragmentInfo, inc/fock.hpp:74-95):
 * its orbital functions [start, start + nbfs) and shells (nshells, contiguous
 * in basis order), auxiliary functions [auxstart, auxstart + naux) and shells
 * (ndfshells), occupied orbitals occ (its LMOs are consecutive columns of
 * Cocc, fragments in order), radius and centre of mass (bohr). */
typedef struct {
    int occ, nbfs, naux, start, auxstart, nshells, ndfshells;
    double radius;
    double com[3];
} m2_fragment;
/* Fock::compute_2body_fock_df_local (src/fockbuild.cpp:794-984), the ALMO
 * SCF's Fock build: Coulomb J (= 4 sum_K (xy|K) [V^-1 Pk]_K, Pk the kept
 * blocks of (xy|K) Pt) and exchange K (per LMO, fitted in its own fit
 * domain and added over its AO domain).  The first call after
 * m2_df_local_reset or a basis change builds the screened fragment blocks
 * of (xy|K) from Pt and the Schwarz bounds (build_blocked_eris, :985-1047;
 * block kept when 2 max|Q| max|Pt| sqrt(max|V_kk|) / r > thrint, or r < 1)
 * and the LMO, AO and fit domains of Cocc sigmainv^(1/2) (build_domains,
 * :662-792, with the reference's mo_thresh, fit_thresh and r_thresh); later
 * calls reuse them, as the reference does once its xyK is set.  Fragments
 * tile both bases in order.  Cocc: n x nocc column-major; sigmainv:
 * nocc x nocc; Pt, J, K: n x n (host).  Unlike the reference, Cocc is not
 * overwritten by Cocc sigmainv^(1/2). */
m2_status m2_jk_df_local(m2_ctx* ctx, int nfrag, const m2_fragment* frags, double mo_thresh, double fit_thresh,
                         double r_thresh, double thrint, const double* Cocc, int nocc, const double* sigmainv,
                         const double* Pt, double* J, double* K);
/* drops the fragment-local blocks and domains (the next call rebuilds them) */
m2_status m2_df_local_reset(m2_ctx* ctx);
/* the domains of the last set-up: which = 0 (LMO), 1 (AO), 2 (fit); per LMO
 * i, counts[i] fragment centres in centres[i * nfrag ...] (rest -1) */
m2_status m2_df_local_domains(m2_ctx* ctx, int which, int* centres, int* counts);

/* RHF nuclear gradient dE/dR (natom x 3, Eh/bohr): Fock::compute_forces
 * (src/rhfgrad.cpp:15-89) = D:(dT + dV) - 2 W:dS + 1/4 G1[D]:D + dE_nuc,
 * with the two-electron part of compute_2body_fock_deriv<1>
 * (src/rhfgrad.cpp:418-598) computed as a derivative-integral adjoint on the
 * device (integrals recomputed, never stored).  D = 2 C_occ C_occ^T and
 * W = C_occ diag(eps_occ) C_occ^T (n x n); quartets screened like the
 * reference's gradient loop (overlap shell-pair list, Schwarz x |D| against
 * `precision`).  Z / xyz are the nuclear charges and positions (bohr). */
m2_status m2_grad(m2_ctx* ctx, int natom, const double* Z, const double* xyz, const double* D,
                  const double* W, double precision, double* g);
/* m2_grad with flags: M2_MIXED computes the derivative quartets whose
 * bound Q_ab Q_cd max|D| lies below the context's FP32
 * threshold (m2_set_mixed_threshold) in FP32 arithmetic, summed in FP64 (the
 * mixed-precision counterpart of the reference's FP64 compute_forces,
 * src/rhfgrad.cpp:15-89); 0 is m2_grad. */
m2_status m2_grad_ex(m2_ctx* ctx, int natom, const double* Z, const double* xyz, const double* D,
                     const double* W, double precision, unsigned flags, double* g);

/* Dense FP64 linear algebra on the device for the SCF driver (host buffers,
 * column-major).  m2_eigh replaces Eigen::SelfAdjointEigenSolver in
 * Fock::formOrthog / Fock::diagonalise (src/fock.cpp:124-142,228-233):
 * ascending eigenvalues w[n], eigenvectors V (n x n, columns).  m2_gemm is
 * C = alpha op(A) op(B) + beta C with op = 'N' or 'T' (the matrix products of
 * SCF::calcErr and Fock::transform, src/scf.cpp:52-69, src/fock.cpp:156-172). */
m2_status m2_eigh(m2_ctx* ctx, int n, const double* A, double* w, double* V);
m2_status m2_gemm(m2_ctx* ctx, char ta, char tb, int m, int n, int k, double alpha, const double* A,
                  int lda, const double* B, int ldb, double beta, double* C, int ldc);

/* Kernel profiling (no reference counterpart; the reference only times whole
 * SCF iterations, src/logger.cpp:451-461).  When on, every J/K launch of a
 * class combination (bra pair class X, ket pair class Y, index 6X+Y with
 * classes ss, ps, pp, ds, dp, dd) is bracketed by HIP events on the context
 * stream and its computed quartets / primitive quartets are counted on the
 * device.  m2_get_profile fills 36-entry arrays accumulated since the last
 * reset.  on = 1: events only (timings undisturbed); on = 2: events and the
 * device quartet counters (these add atomics and distort the timings). */
m2_status m2_set_profiling(m2_ctx* ctx, int on);
m2_status m2_get_profile(m2_ctx* ctx, double* ms, long long* launches, long long* quartets,
                         long long* prim_quartets);
m2_status m2_reset_profile(m2_ctx* ctx);
/* primitive quartets actually computed per class combination (after
 * primitive screening; m2_get_profile's prim_quartets counts all of them),
 * accumulated with the device counters (on = 2) since the last reset */
m2_status m2_get_profile_computed(m2_ctx* ctx, long long* computed);

#ifdef __cplusplus
}
#endif
#endif
