// Shared device/host definitions of the MI355X Fock-build engine.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace m2 {

// Primitive shell-pair record, one per surviving primitive pair of a shell
// pair (a,b) with la >= lb.  K folds the contraction coefficients, the
// Gaussian-product prefactor exp(-ab/p |AB|^2), sqrt(2 pi^{5/2}) and 1/p, so
// that [00|00]^(m) = K_ab K_cd / sqrt(p+q) F_m(T).
struct alignas(16) PrimPair {
    double p;       // a + b
    double inv2p;   // 1 / (2p)
    double K;
    double P[3];    // product centre
    double PA[3];   // P - A (A = centre of the pair's first shell)
    double a;       // exponent of the first shell's primitive (derivatives)
};
static_assert(sizeof(PrimPair) == 80, "PrimPair layout");

// The fields of a PrimPair that the J/K kernels read, in the kernel's
// arithmetic type (the FP32 kernels read an FP32 copy of the pair records).
template <typename R>
struct alignas(16) BraPair {
    R p, inv2p, K;
    R P[3];
    R PA[3];
};
static_assert(sizeof(BraPair<double>) == 80 && sizeof(BraPair<float>) == 48, "BraPair layout");

// The per-chunk data of a shell pair read as a ket by the J/K kernels, in
// one 64-byte record (instead of ~10 scattered loads): its shells, first
// primitive pair and count, Schwarz bound, C - D and first basis functions
struct alignas(16) PairRec {
    int sc, sd, off, np;
    double Q;
    double CD[3];
    int fc, fd;
    int pad[2];
};
static_assert(sizeof(PairRec) == 64, "PairRec layout");

// The FP64 record a primitive pair is read through when it is the ket of an
// s-s pair class: the fields such kets use and the pair's primitive bound,
// 48 bytes instead of 80 (the lanes' ket records are the kernels' main
// L2 reads).
struct alignas(16) KetS {
    double p, K;
    double P[3];
    float q;   // bound of the primitive pair (its prim_q entry)
    float pad;
};
static_assert(sizeof(KetS) == 48, "KetS layout");

// ket primitive ik when ik < nket (a zero record otherwise), for loading the
// next ket primitive while the current one is computed
template <typename KP>
__device__ __forceinline__ KP load_ket(const KP* __restrict__ ket, int ik, int nket) {
    return ik < nket ? ket[ik] : KP{};
}
// bound of a loaded ket primitive: from the bound array, or from its KetS record
template <typename KP>
__device__ __forceinline__ float ket_q(const KP&, const float* __restrict__ kq, int ik) {
    return kq[ik];
}
struct KetS;
__device__ __forceinline__ float ket_q(const KetS& k, const float* __restrict__, int ik);

// bound of ket primitive ik: from the bound array, or from a KetS record
template <typename KP>
__device__ __forceinline__ float ket_bound(const KP* __restrict__, const float* __restrict__ kq, int ik) {
    return kq[ik];
}
__device__ __forceinline__ float ket_bound(const KetS* __restrict__ ket, const float* __restrict__, int ik) {
    return ket[ik].q;
}
__device__ __forceinline__ float ket_q(const KetS& k, const float* __restrict__, int) { return k.q; }

// The same for the FP32 kernels (24 bytes)
struct alignas(8) KetSf {
    float p, K;
    float P[3];
    float q;
};
static_assert(sizeof(KetSf) == 24, "KetSf layout");
__device__ __forceinline__ float ket_bound(const KetSf* __restrict__ ket, const float* __restrict__, int ik) {
    return ket[ik].q;
}
__device__ __forceinline__ float ket_q(const KetSf& k, const float* __restrict__, int) { return k.q; }

// Boys table: F_m(T_k) at T_k = k * BOYS_STEP, k = 0..BOYS_NGRID-1,
// m = 0..BOYS_MTAB-1.  Beyond BOYS_TMAX the asymptotic branch is used.
constexpr int BOYS_MTAB = 26;
constexpr int BOYS_NGRID = 601;
constexpr double BOYS_STEP = 0.1;
constexpr double BOYS_TMAX = 60.0;

// ket chunks (of <= 64 kets) per work item of the direct J/K register-class
// kernels (default, and the smallest value the work-list capacity allows; the
// autotuner picks per class combination); kets per chunk of the table-driven
// heavy classes (one wavefront walks them in turn; one chunk per work item by
// default); primitive-count groups per pair class, and the LDS budget of the
// K-row accumulators
constexpr int JK_BLOCK_TASKS = 64;
constexpr int JK_BLOCK_TASKS_MIN = 16;
constexpr int JK_HEAVY_KETS = 4;
constexpr long long JK_LIST_MAX_ITEMS = 1LL << 24;   // work items per list and precision (64 MiB)
// gradient program classes: a host task holds up to GRAD_PROG_SPLIT *
// GRAD_PROG_KETS kets; GRAD_PROG_SPLIT workgroups take GRAD_PROG_KETS each
constexpr int GRAD_PROG_KETS = 4, GRAD_PROG_SPLIT = 8;
constexpr int JK_MAXG = 10;
constexpr size_t JK_KROW_LDS_MAX = 96 * 1024;
constexpr size_t JK_DROW_LDS_MAX = 64 * 1024;   // K rows + staged density rows

struct JKArgs {
    const PrimPair* prims;
    const BraPair<float>* prims_f;   // the same records in FP32
    const KetS* kets_s;              // the same records as s-s kets (FP64)
    const PairRec* pair_rec;         // shell pairs as kets (per-chunk data)
    const struct KetSf* kets_sf;     // and as s-s kets of the FP32 kernels
    const int* pair_off;     // first PrimPair of each shell pair
    const int* pair_np;      // number of PrimPairs of each shell pair
    const int2* pair_sh;     // (a, b) shell indices, la >= lb
    const double* pair_Q;    // Schwarz bound of each pair
    const double* sh_O;      // shell centres, 3 per shell
    const int* sh_bf;        // first basis function of each shell
    const double* Dn;        // shell-block inf-norms of D, nsh x nsh
    const double* D;         // density (or Delta-D), n x n
    double* J;               // unsymmetrised accumulators, n x n
    double* K;
    int n;
    int nsh;
    double precision;        // quartet screening threshold
    double fp32_below;       // quartets with bound < this run in FP32 (mixed mode)
    double scale;            // multiplier
    unsigned long long* counters;  // profiling only: per class combo {quartets, primitive quartets, computed ones}
    int combo;               // 6 * X + Y of the launch
    int k_lds;               // set by the launcher: K rows of the bra shells accumulated in LDS
    int d_lds;               // set by the launcher: the bra shells' rows of the density norms (1) and also
                             // of D (2) staged in LDS after the K rows
    int d_lds_max;           // the most staging the launcher may choose (0, 1, 2)
    int boys_lds;            // set by the launcher: the class's compact Boys rows staged in LDS at byte boys_off
    int boys_off;
    int boys_lds_max;        // 0: never stage the Boys rows
    // device work list of the launch (tasks.hip): workgroup -> bra, and the
    // bra's ket segments per primitive-count group of the ket class
    const int2* cut;         // [bra][JK_MAXG] (mid, end)
    const int* blk_off;      // [bra] first workgroup of the bra
    const int* blk_bra;      // [workgroup] bra (index local to the class)
    const int* blk_total;    // device count of work items (grid sized from a bound), or null (grid = count)
    int block_tasks;         // ket chunks per work item
    int bra0;                // first pair of the bra class
    int ngrp;                // ket groups
    int tail;                // 0: FP64 head [gbeg, mid) of each group; 1: FP32 tail [mid, end)
    int gbeg[JK_MAXG];
    // unrestricted builds: a second density digested from the same integrals
    int nden;                // 1, or 2 with D2 / J2 / K2
    const double* D2;
    double* J2;
    double* K2;
    // s-s ket families (k_jk_fam): family pairs of s shells sharing their
    // primitive pairs among up to FAM_MAXMEM contracted member pairs
    int fam_kets;            // set by the host: ket indices of this launch are family pairs
    const PrimPair* fam_prims;        // primitive pairs, K without contraction coefficients
    const struct FamS* fam_s;         // the same with their factors and bounds (FP64 kernels)
    const struct FamSf* fam_sf;       // (FP32 kernels)
    const BraPair<float>* fam_prims_f;
    const double* fam_km;    // [prim][FAM_MAXMEM] member factors K c_c c_d (0 for absent members)
    const float* fam_km_f;
    const int* fam_off;      // first primitive pair of each family pair
    const int* fam_np;
    const int4* fam_mem;     // member shell-pair indices (-1 pads)
    const int* pair_fam;     // class-0 shell pair -> its family pair
    const int* pair_mem;     // class-0 shell pair -> its member slot
    // primitive screening: bounds sqrt(max |(ab|ab)|) of every primitive pair
    // (sorted descending within each shell pair) and of every family
    // primitive pair (largest member factor); a primitive quartet is dropped
    // when bound_ab bound_cd max|D_blocks| < prim_tau (0: none dropped)
    const float* prim_q;
    const float* fam_q;
    double prim_tau;
};

constexpr int FAM_MAXMEM = 4;

// The FP64 record a family primitive pair is read through by the family
// kernels: its centre, exponent sum, member factors and bound in one
// 80-byte record (instead of an 80-byte pair record of which 32 bytes are
// read, the factors and the bound in two more arrays)
struct alignas(16) FamS {
    double p;
    double P[3];
    double km[FAM_MAXMEM];   // member factors K c_c c_d (0 for absent members)
    float q;                 // bound of the family primitive pair
    float pad[3];
};
static_assert(sizeof(FamS) == 80, "FamS layout");
__device__ __forceinline__ float ket_bound(const FamS* __restrict__ ket, const float* __restrict__, int ik) {
    return ket[ik].q;
}
// member factor m of ket primitive ik: from the factor array, or from a FamS record
template <typename KP, typename KM>
__device__ __forceinline__ KM fam_factor(const KM* __restrict__ km, const KP* __restrict__, int ik, int m) {
    return km[FAM_MAXMEM * ik + m];
}
template <typename KM>
__device__ __forceinline__ double fam_factor(const KM* __restrict__, const FamS* __restrict__ ket, int ik, int m) {
    return ket[ik].km[m];
}
// member factor m of a loaded family primitive: from the factor array, or
// from its FamS / FamSf record
template <typename KP, typename KM>
__device__ __forceinline__ KM fam_factor_of(const KP&, const KM* __restrict__ km, int ik, int m) {
    return km[FAM_MAXMEM * ik + m];
}
template <typename KM>
__device__ __forceinline__ double fam_factor_of(const FamS& k, const KM* __restrict__, int, int m) {
    return k.km[m];
}
__device__ __forceinline__ float ket_q(const FamS& k, const float* __restrict__, int) { return k.q; }

// The same for the FP32 family kernels (48 bytes)
struct alignas(16) FamSf {
    float p;
    float P[3];
    float km[FAM_MAXMEM];
    float q;
    float pad[3];
};
static_assert(sizeof(FamSf) == 48, "FamSf layout");
__device__ __forceinline__ float ket_bound(const FamSf* __restrict__ ket, const float* __restrict__, int ik) {
    return ket[ik].q;
}
template <typename KM>
__device__ __forceinline__ float fam_factor(const KM* __restrict__, const FamSf* __restrict__ ket, int ik, int m) {
    return ket[ik].km[m];
}
template <typename KM>
__device__ __forceinline__ float fam_factor_of(const FamSf& k, const KM* __restrict__, int, int m) {
    return k.km[m];
}
__device__ __forceinline__ float ket_q(const FamSf& k, const float* __restrict__, int) { return k.q; }
// class combinations (bra X | ket Y) run with s-s family kets: Y = (ss) and
// X = (ss) or (pp) whenever the basis has families; X = (ps) or (ds) only when
// the family members are mostly significant together (`wide`), since those
// bra classes lose to the per-member bookkeeping when screening leaves about
// one member per family.  All four have run_fam (csrc/gen_eri.py FAM_BRA).
constexpr bool fam_combo(int X, int Y, bool wide) {
    return Y == 0 && (X == 0 || X == 2 || (wide && (X == 1 || X == 3)));
}



struct Task {                // one wavefront of quartets: bra pair x 1..64 ket pairs
    int bra;
    int ket0;
    int nket;
    int pad;
};

}  // namespace m2
