// Direct J/K shell-quartet kernels (templates; instantiated per class in the
// generated translation units).
//
// Work unit: one chunk = one bra shell pair x up to 64 consecutive ket shell
// pairs of one class and one primitive-count group (so every lane runs the
// same primitive loop); one lane per ket pair.  A workgroup owns one work
// item: a bra pair and up to a.block_tasks of its chunks, which its waves
// take in turn; the items come from the device work list of tasks.hip.  The
// grid is either the list's item count (read back by the host) or a bound on
// it that depends on the basis only (no host round trip): workgroups past
// the count the list kernels wrote exit at once.
// Each lane applies the exact screening of Fock::formJKdirect
// (src/fockbuild.cpp:98-115): max of the six shell-block norms of D times
// Q_ab Q_cd against `precision`, computes its contracted quartet with the
// generated OS/HGP code, and digests it with the degeneracy factors of
// src/fockbuild.cpp:121-149 into unsymmetrised J and K:
//   J(ab)  per-lane register accumulators over every Task of the wave,
//          summed across the wave once at the end;
//   K(a.), K(b.)  (every K contribution has a bra shell as its row) LDS rows
//          of the workgroup (ds_add_f64), flushed once to HBM with global
//          FP64 atomics; when the rows do not fit in LDS, direct atomics;
//   J(cd)  per-lane global FP64 atomics (hardware global_atomic_add_f64).
#pragma once
#include "m2_common.hpp"
#include "boys.hpp"
#include <type_traits>

namespace m2 {

template <int LA, int LB, int LC, int LD>
struct EriClass;

template <int L, bool PURE>
struct NFun {
    static constexpr int value = PURE ? 2 * L + 1 : (L + 1) * (L + 2) / 2;
};

__device__ __forceinline__ void atomic_add_f64(double* p, double v) {
    atomicAdd(p, v);   // -munsafe-fp-atomics: global_atomic_add_f64
}

// wavefront sum (all 64 lanes participate; inactive lanes pass 0)
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// K-row accumulation: LDS row (row < NA: shell a, else shell b) or HBM
template <int LM>
__device__ __forceinline__ void krow_add(double* __restrict__ sK, double* __restrict__ K, int n, int row, int grow,
                                         int col, double v) {
    if constexpr (LM >= 1)
        __hip_atomic_fetch_add((__attribute__((address_space(3))) double*)(sK + row * n + col), v, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_WORKGROUP);   // ds_add_f64
    else
        atomic_add_f64(&K[grow * n + col], v);
}

// D(row, col) for a row of the bra shells: rows r < NA are shell a's
// (global row fa + r), the rest shell b's; from the LDS copy when staged.
// LM, the LDS mode of a workgroup: 2 = K rows and D rows in LDS, 1 = K rows
// only, 0 = neither (K accumulates with global atomics)
template <int NA, int LM>
__device__ __forceinline__ double drow(const double* __restrict__ sD, const double* __restrict__ D, int n, int fa,
                                       int fb, int r, int col) {
    if constexpr (LM == 2) return sD[r * n + col];
    else return D[(r < NA ? fa + r : fb + r - NA) * n + col];
}

template <int NA, int NB, int NC, int ND, int LM, typename R>
__device__ __forceinline__ void digest(const R* __restrict__ v, int fa, int fb, int fc, int fd, double f,
                                       const double* __restrict__ Dab, double* __restrict__ jab,
                                       const double* __restrict__ D, int n, double* __restrict__ J,
                                       double* __restrict__ K, double* __restrict__ sK, const double* __restrict__ sD) {
    const double h = 0.5 * f;
    // J(ab) += f sum_cd D(cd) v  (lane registers)
#pragma unroll
    for (int c = 0; c < NC; ++c)
#pragma unroll
        for (int d = 0; d < ND; ++d) {
            const double dcd = f * D[(fc + c) * n + fd + d];
#pragma unroll
            for (int a = 0; a < NA; ++a)
#pragma unroll
                for (int b = 0; b < NB; ++b) jab[a * NB + b] += dcd * double(v[((a * NB + b) * NC + c) * ND + d]);
        }
    // K(ac) += h sum_bd D(bd) v
#pragma unroll
    for (int a = 0; a < NA; ++a)
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            double s = 0.0;
#pragma unroll
            for (int b = 0; b < NB; ++b)
#pragma unroll
                for (int d = 0; d < ND; ++d) s += drow<NA, LM>(sD, D, n, fa, fb, NA + b, fd + d) * double(v[((a * NB + b) * NC + c) * ND + d]);
            krow_add<LM>(sK, K, n, a, fa + a, fc + c, h * s);
        }
    // K(bd) += h sum_ac D(ac) v
#pragma unroll
    for (int b = 0; b < NB; ++b)
#pragma unroll
        for (int d = 0; d < ND; ++d) {
            double s = 0.0;
#pragma unroll
            for (int a = 0; a < NA; ++a)
#pragma unroll
                for (int c = 0; c < NC; ++c) s += drow<NA, LM>(sD, D, n, fa, fb, a, fc + c) * double(v[((a * NB + b) * NC + c) * ND + d]);
            krow_add<LM>(sK, K, n, NA + b, fb + b, fd + d, h * s);
        }
    // K(ad) += h sum_bc D(bc) v
#pragma unroll
    for (int a = 0; a < NA; ++a)
#pragma unroll
        for (int d = 0; d < ND; ++d) {
            double s = 0.0;
#pragma unroll
            for (int b = 0; b < NB; ++b)
#pragma unroll
                for (int c = 0; c < NC; ++c) s += drow<NA, LM>(sD, D, n, fa, fb, NA + b, fc + c) * double(v[((a * NB + b) * NC + c) * ND + d]);
            krow_add<LM>(sK, K, n, a, fa + a, fd + d, h * s);
        }
    // K(bc) += h sum_ad D(ad) v
#pragma unroll
    for (int b = 0; b < NB; ++b)
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            double s = 0.0;
#pragma unroll
            for (int a = 0; a < NA; ++a)
#pragma unroll
                for (int d = 0; d < ND; ++d) s += drow<NA, LM>(sD, D, n, fa, fb, a, fd + d) * double(v[((a * NB + b) * NC + c) * ND + d]);
            krow_add<LM>(sK, K, n, NA + b, fb + b, fc + c, h * s);
        }
    // J(cd) += f sum_ab D(ab) v (global atomics last: later loads of this
    // digestion need not wait for them)
#pragma unroll
    for (int c = 0; c < NC; ++c)
#pragma unroll
        for (int d = 0; d < ND; ++d) {
            double s = 0.0;
#pragma unroll
            for (int a = 0; a < NA; ++a)
#pragma unroll
                for (int b = 0; b < NB; ++b) s += Dab[a * NB + b] * double(v[((a * NB + b) * NC + c) * ND + d]);
            atomic_add_f64(&J[(fc + c) * n + fd + d], f * s);
        }
}

// f(std::integral_constant<int, LM>) for a workgroup's LDS mode (uniform)
template <typename F>
__device__ __forceinline__ void with_lds_mode(int lm, F&& f) {
    if (lm == 2) f(std::integral_constant<int, 2>{});
    else if (lm == 1) f(std::integral_constant<int, 1>{});
    else f(std::integral_constant<int, 0>{});
}

// Primitive quartets whose bound passes the lane's cut (profiling only; the
// generated loops compute these and, where lanes of a wavefront differ, a
// few more).
template <typename QB>
__device__ __forceinline__ int screened_prims(QB bq, int nbra, const float* kq, int nket, float cut) {
    int nb = nbra, tot = 0;
    for (int ik = 0; ik < nket; ++ik) {
        const float qk = kq[ik];
        while (nb > 0 && bq[nb - 1] * qk < cut) --nb;
        if (nb == 0) break;
        tot += nb;
    }
    return tot;
}

template <typename R>
struct PrimPairOf;
template <>
struct PrimPairOf<double> {
    using type = PrimPair;
    static __device__ __forceinline__ const PrimPair* get(const JKArgs& a) { return a.prims; }
};
template <>
struct PrimPairOf<float> {
    using type = BraPair<float>;
    static __device__ __forceinline__ const BraPair<float>* get(const JKArgs& a) { return a.prims_f; }
};

// Waves per workgroup: 8 (two per SIMD, <= 256 registers per wave) for
// classes whose VRR accumulators fit comfortably; 4 (one per SIMD, so a wave
// may use the 256 AGPRs as well) for the register-heavy ones.
template <int LA, int LB, int LC, int LD, typename R>
struct JKWaves {
    static constexpr int value = EriClass<LA, LB, LC, LD>::NACC * (int)sizeof(R) <= 280 ? 8 : 4;
};
// (ps|ss) in FP64 runs 4-wave workgroups held to 4 waves per SIMD (<= 128
// VGPRs: room for its next ket primitive in flight; its LDS, with the staged
// D and Boys rows, allows 4 workgroups per CU at n = 480); (ps|ps) in FP64 is
// held to 4 waves per SIMD too (it would take 129 VGPRs, 3 waves); elsewhere
// no minimum
template <int LA, int LB, int LC, int LD, typename R>
struct JKOccupancy {
    static constexpr bool ps = LA == 1 && LB == 0 && LC == 0 && LD == 0 && sizeof(R) == 8;
    static constexpr bool psps = LA == 1 && LB == 0 && LC == 1 && LD == 0 && sizeof(R) == 8;
    static constexpr int waves = ps ? 4 : JKWaves<LA, LB, LC, LD, R>::value;
    static constexpr int min_per_eu = ps || psps ? 4 : 1;
};

// Work item -> ket chunk `ch` of the bra's list: first ket and count (<= 64)
// of the chunk; the chunks run over the ket groups' head segments [gbeg,
// mid) (FP64) or tail segments [mid, end) (FP32).
__device__ __forceinline__ void chunk_kets(const JKArgs& a, const __attribute__((address_space(4))) int* cut, int ch,
                                           int& ket0, int& nket, int& grp) {
    ket0 = 0;
    nket = 0;
    grp = -1;
    int rem = ch;
    for (int g = 0; g < a.ngrp; ++g) {
        const int mid = cut[2 * g], end = cut[2 * g + 1];
        const int s0 = a.tail ? mid : a.gbeg[g], s1 = a.tail ? end : mid;
        const int nc = (s1 - s0 + 63) >> 6;
        if (rem < nc) {
            ket0 = s0 + 64 * rem;
            nket = min(64, s1 - ket0);
            grp = g;
            return;
        }
        rem -= nc;
    }
}

// Workgroup start: clear the K rows of the bra shells in LDS and stage the
// bra shells' rows of the density norms (d_lds >= 1) and of D (and D2;
// d_lds == 2) after them.  LDS layout: [K rows: NDEN (NA + NB) n]
// [D rows: NDEN (NA + NB) n when d_lds == 2] [norm rows of a, b: 2 nsh].
template <int NA, int NB, int NDEN>
__device__ __forceinline__ void stage_bra_rows(const JKArgs& a, int sa, int sb, int fa, int fb, double* sK,
                                               const double*& sD, const double*& sDn) {
    const int n = a.n, nsh = a.nsh;
    constexpr int NR = NA + NB;
    const int nk = a.k_lds ? NDEN * NR * n : 0;
    const int nd = a.d_lds == 2 ? NDEN * NR * n : 0;
    double* d = sK + nk;
    double* dn = d + nd;
    for (int i = threadIdx.x; i < nk; i += blockDim.x) sK[i] = 0.0;
    for (int i = threadIdx.x; i < nd; i += blockDim.x) {
        const int set = i / (NR * n), r = i - set * NR * n;
        const int row = r / n, col = r - row * n;
        d[i] = (set ? a.D2 : a.D)[(row < NA ? fa + row : fb + row - NA) * n + col];
    }
    if (a.d_lds)
        for (int i = threadIdx.x; i < 2 * nsh; i += blockDim.x) dn[i] = a.Dn[(i < nsh ? sa : sb) * nsh + i % nsh];
    sD = d;
    sDn = dn;
    __syncthreads();
}

// Quartets are computed when lo <= bound < hi (the launchers pass
// [precision, inf); precision is chosen per task list).
// Dynamic LDS (when a.k_lds): the (NA + NB) x n FP64 K rows of shells a and b.
// NDEN = 2 (unrestricted build): every quartet is digested with D and with D2
// into (J, K) and (J2, K2); the LDS then holds the K rows of both.
// One work item of k_jk: bra pair and up to a.block_tasks of its ket chunks.
template <int LA, int LB, int LC, int LD, typename R, bool PURE, int NDEN, bool BL>
__device__ __forceinline__ void jk_item(const JKArgs& a, const BoysTables& tab0, double lo, double hi, int witem) {
    constexpr int NA = NFun<LA, PURE>::value, NB = NFun<LB, PURE>::value;
    constexpr int NC = NFun<LC, PURE>::value, ND = NFun<LD, PURE>::value;
    constexpr int W = JKOccupancy<LA, LB, LC, LD, R>::waves;
    typedef const __attribute__((address_space(4))) int cint;
    // work item -> bra pair and its ket segments (written by the list kernels)
    const int bl = ((cint*)a.blk_bra)[witem];
    const int lb = witem - ((cint*)a.blk_off)[bl];
    const int bra = a.bra0 + bl;
    cint* cut = (cint*)a.cut + 2 * bl * JK_MAXG;   // (mid, end) per group
    int nch = 0;
    for (int g = 0; g < a.ngrp; ++g) {
        const int mid = cut[2 * g], end = cut[2 * g + 1];
        nch += ((a.tail ? end - mid : mid - a.gbeg[g]) + 63) >> 6;
    }
    const int c0 = lb * a.block_tasks, c1 = min(c0 + a.block_tasks, nch);
    const int lane = __lane_id();
    const int wave = threadIdx.x >> 6;
    const int2 sab = a.pair_sh[bra];
    const int sa = __builtin_amdgcn_readfirstlane(sab.x), sb = __builtin_amdgcn_readfirstlane(sab.y);
    const int n = a.n, nsh = a.nsh;
    const int fa = a.sh_bf[sa], fb = a.sh_bf[sb];
    const bool lds = a.k_lds != 0;
    // primitive pairs in the kernel's arithmetic type; the bra's are read
    // through the constant address space (never written by the kernel), so
    // the uniform bra loop loads them with scalar loads into SGPRs
    using PP = typename PrimPairOf<R>::type;
    const PP* gp = PrimPairOf<R>::get(a);
    typedef const __attribute__((address_space(4))) PP cPP;
    const int nbp = a.pair_np[bra];
    cPP* bprims = (cPP*)(gp + a.pair_off[bra]);
    typedef const __attribute__((address_space(4))) float cfl;
    cfl* bq = (cfl*)(a.prim_q + a.pair_off[bra]);
    extern __shared__ double sK[];
    // BL: the class's compact Boys rows read from LDS (copied here; the
    // bra-row staging below ends with the workgroup barrier)
    BoysTables tab = tab0;
    if constexpr (BL) tab = stage_boys_rows<R, LA + LB + LC + LD>(tab0, sK + a.boys_off / sizeof(double));
    const double *sD, *sDn;
    stage_bra_rows<NA, NB, NDEN>(a, sa, sb, fa, fb, sK, sD, sDn);
    const bool dnl = a.d_lds != 0;
    const int lm = lds ? (a.d_lds == 2 ? 2 : 1) : 0;   // LDS mode of the digestion
    const double* D = a.D;
    const double* D2 = a.D2;
    double Dab[NA * NB], jab[NA * NB];
    double Dab2[NDEN > 1 ? NA * NB : 1], jab2[NDEN > 1 ? NA * NB : 1];
#pragma unroll
    for (int x = 0; x < NA; ++x)
#pragma unroll
        for (int y = 0; y < NB; ++y) {
            Dab[x * NB + y] = D[(fa + x) * n + fb + y];
            jab[x * NB + y] = 0.0;
            if constexpr (NDEN > 1) {
                Dab2[x * NB + y] = D2[(fa + x) * n + fb + y];
                jab2[x * NB + y] = 0.0;
            }
        }
    const double* O = a.sh_O;
    const R ABx = R(O[3 * sa] - O[3 * sb]), ABy = R(O[3 * sa + 1] - O[3 * sb + 1]), ABz = R(O[3 * sa + 2] - O[3 * sb + 2]);
    const double* Dn = a.Dn;
    const double dnab = Dn[sa * nsh + sb];
    const double qab = a.pair_Q[bra];
    const double fab = (sa == sb ? 1.0 : 2.0) * a.scale;
    unsigned long long cnt_q = 0, cnt_pq = 0, cnt_cpq = 0;   // profiling counters of this wave
    for (int ch = c0 + wave; ch < c1; ch += W) {
        int ket0, nket, g;
        chunk_kets(a, cut, ch, ket0, nket, g);
        const int ket = ket0 + (lane < nket ? lane : 0);
        const PairRec pr = a.pair_rec[ket];   // one 64-byte record per ket
        const int sc = pr.sc, sd = pr.sd;
        // exact screening of src/fockbuild.cpp:98-115
        const double dn = fmax(fmax(dnab, Dn[sc * nsh + sd]),
                               dnl ? fmax(fmax(sDn[sc], sDn[sd]), fmax(sDn[nsh + sc], sDn[nsh + sd]))
                                   : fmax(fmax(Dn[sa * nsh + sc], Dn[sa * nsh + sd]),
                                          fmax(Dn[sb * nsh + sc], Dn[sb * nsh + sd])));
        const double bound = dn * qab * pr.Q;
        const bool active = lane < nket && bound >= lo && bound < hi;
        if (!__any(active)) continue;
        const float* kq = a.prim_q + pr.off;
        const float cut = float(a.prim_tau / dn);
        const int nkp = active ? pr.np : 0;
        if (a.counters) {
            cnt_q += __popcll(__ballot(active));
            cnt_pq += (unsigned long long)(wave_sum(double(nkp)) * nbp + 0.5);
            cnt_cpq += (unsigned long long)(wave_sum(double(screened_prims(bq, nbp, kq, nkp, cut))) + 0.5);
        }
        const R CDx = R(pr.CD[0]), CDy = R(pr.CD[1]), CDz = R(pr.CD[2]);
        R v[NA * NB * NC * ND > EriClass<LA, LB, LC, LD>::NCART ? NA * NB * NC * ND : EriClass<LA, LB, LC, LD>::NCART];
        // s-s kets through their compact records
        const auto kprims = [&]() {
            if constexpr (LC + LD == 0 && std::is_same<R, double>::value) return a.kets_s + pr.off;
            else if constexpr (LC + LD == 0) return a.kets_sf + pr.off;
            else return gp + pr.off;
        }();
        EriClass<LA, LB, LC, LD>::template run<R, PURE, true>(bprims, nbp, kprims, nkp,
                                                              ABx, ABy, ABz, CDx, CDy, CDz, tab, v, bq, kq, cut);
        if (active) {
            const double f = fab * (sc == sd ? 1.0 : 2.0) * (bra == ket ? 1.0 : 2.0);
            const int fc = a.pair_rec[ket].fc, fd = a.pair_rec[ket].fd;   // read late: not live in the loop
            with_lds_mode(lm, [&](auto M) {
                constexpr int LM = decltype(M)::value;
                digest<NA, NB, NC, ND, LM, R>(v, fa, fb, fc, fd, f, Dab, jab, D, n, a.J, a.K, sK, sD);
                if constexpr (NDEN > 1)
                    digest<NA, NB, NC, ND, LM, R>(v, fa, fb, fc, fd, f, Dab2, jab2, D2, n, a.J2, a.K2,
                                                  sK + (NA + NB) * n, sD + (NA + NB) * n);
            });
        }
    }
    if (a.counters && lane == 0 && cnt_q) {
        atomicAdd(&a.counters[3 * a.combo], cnt_q);
        atomicAdd(&a.counters[3 * a.combo + 1], cnt_pq);
        atomicAdd(&a.counters[3 * a.combo + 2], cnt_cpq);
    }
    // J(ab): one wave sum per function
#pragma unroll
    for (int x = 0; x < NA * NB; ++x) {
        const double s = wave_sum(jab[x]);
        if (lane == 0 && s != 0.0) atomic_add_f64(&a.J[(fa + x / NB) * n + fb + x % NB], s);
        if constexpr (NDEN > 1) {
            const double s2 = wave_sum(jab2[x]);
            if (lane == 0 && s2 != 0.0) atomic_add_f64(&a.J2[(fa + x / NB) * n + fb + x % NB], s2);
        }
    }
    if (!lds) return;
    __syncthreads();
    for (int i = threadIdx.x; i < NDEN * (NA + NB) * n; i += blockDim.x) {
        const double s = sK[i];
        if (s != 0.0) {
            const int set = i / ((NA + NB) * n), r = i - set * (NA + NB) * n;
            const int row = r / n, col = r - row * n;
            atomic_add_f64(&(set ? a.K2 : a.K)[(row < NA ? fa + row : fb + row - NA) * n + col], s);
        }
    }
}

template <int LA, int LB, int LC, int LD, typename R, bool PURE, int NDEN>
__global__ __launch_bounds__((JKOccupancy<LA, LB, LC, LD, R>::waves * 64),
                             (JKOccupancy<LA, LB, LC, LD, R>::min_per_eu)) void k_jk(JKArgs a, BoysTables tab, double lo,
                                                                                 double hi) {
    if (a.blk_total && (int)blockIdx.x >= *a.blk_total) return;   // grid sized from a bound: past the total
    if constexpr (boys_rows_bytes<R, LA + LB + LC + LD>() > 0) {
        if (a.boys_lds) jk_item<LA, LB, LC, LD, R, PURE, NDEN, true>(a, tab, lo, hi, blockIdx.x);
        else jk_item<LA, LB, LC, LD, R, PURE, NDEN, false>(a, tab, lo, hi, blockIdx.x);
    } else {
        jk_item<LA, LB, LC, LD, R, PURE, NDEN, false>(a, tab, lo, hi, blockIdx.x);
    }
}

// (ab|ss) with s-s family kets: one lane per ket family pair.  Its primitive
// pairs are shared by up to FAM_MAXMEM contracted member pairs (cd); each
// member keeps the exact screening of src/fockbuild.cpp:98-115 and, for the
// same class (X == Y == 0), the unique-pair rule: member rank (family,
// slot) <= the bra's rank.  Digestion per member as in k_jk.
template <typename R>
struct FamPrimOf;
template <>
struct FamPrimOf<double> {
    static __device__ __forceinline__ const PrimPair* prims(const JKArgs& a) { return a.fam_prims; }
    static __device__ __forceinline__ const double* km(const JKArgs& a) { return a.fam_km; }
};
template <>
struct FamPrimOf<float> {
    static __device__ __forceinline__ const BraPair<float>* prims(const JKArgs& a) { return a.fam_prims_f; }
    static __device__ __forceinline__ const float* km(const JKArgs& a) { return a.fam_km_f; }
};

template <int LA, int LB, typename R>
struct JKWavesFam {
    // (ss|ss) in FP64 runs 4-wave workgroups held to 4 waves per SIMD (<= 128
    // VGPRs: room for its next family primitive in flight)
    static constexpr bool ss = LA == 0 && LB == 0 && sizeof(R) == 8;
    static constexpr int value = ss ? 4 : EriClass<LA, LB, 0, 0>::NACC_FAM * (int)sizeof(R) <= 280 ? 8 : 4;
    static constexpr int min_per_eu = ss ? 4 : 1;
};

template <int LA, int LB, typename R, bool PURE, int NDEN, bool BL>
__device__ __forceinline__ void jk_fam_item(const JKArgs& a, const BoysTables& tab0, double lo, double hi,
                                            int witem) {
    constexpr int NA = NFun<LA, PURE>::value, NB = NFun<LB, PURE>::value;
    constexpr int NCART = EriClass<LA, LB, 0, 0>::NCART;
    constexpr int W = JKWavesFam<LA, LB, R>::value;
    typedef const __attribute__((address_space(4))) int cint;
    const int bl = ((cint*)a.blk_bra)[witem];
    const int lb = witem - ((cint*)a.blk_off)[bl];
    const int bra = a.bra0 + bl;
    cint* cut = (cint*)a.cut + 2 * bl * JK_MAXG;
    int nch = 0;
    for (int g = 0; g < a.ngrp; ++g) {
        const int mid = cut[2 * g], end = cut[2 * g + 1];
        nch += ((a.tail ? end - mid : mid - a.gbeg[g]) + 63) >> 6;
    }
    const int c0 = lb * a.block_tasks, c1 = min(c0 + a.block_tasks, nch);
    const int lane = __lane_id();
    const int wave = threadIdx.x >> 6;
    const int2 sab = a.pair_sh[bra];
    const int sa = __builtin_amdgcn_readfirstlane(sab.x), sb = __builtin_amdgcn_readfirstlane(sab.y);
    const int n = a.n, nsh = a.nsh;
    const int fa = a.sh_bf[sa], fb = a.sh_bf[sb];
    const bool lds = a.k_lds != 0;
    using PP = typename PrimPairOf<R>::type;
    typedef const __attribute__((address_space(4))) PP cPP;
    const int nbp = a.pair_np[bra];
    cPP* bprims = (cPP*)(PrimPairOf<R>::get(a) + a.pair_off[bra]);
    typedef const __attribute__((address_space(4))) float cfl;
    cfl* bq = (cfl*)(a.prim_q + a.pair_off[bra]);
    const auto* fprims = FamPrimOf<R>::prims(a);
    const auto* fkm = FamPrimOf<R>::km(a);
    // unique pairs of pairs inside class 0: the bra's own family and slot
    const bool same = a.pair_fam != nullptr && a.combo == 0;
    const int bfam = same ? a.pair_fam[bra] : -1, bmem = same ? a.pair_mem[bra] : 0;
    extern __shared__ double sK[];
    BoysTables tab = tab0;   // BL: compact Boys rows from LDS, as in jk_item
    if constexpr (BL) tab = stage_boys_rows<R, LA + LB>(tab0, sK + a.boys_off / sizeof(double));
    const double *sD, *sDn;
    stage_bra_rows<NA, NB, NDEN>(a, sa, sb, fa, fb, sK, sD, sDn);
    const bool dnl = a.d_lds != 0;
    const int lm = lds ? (a.d_lds == 2 ? 2 : 1) : 0;   // LDS mode of the digestion
    const double* D = a.D;
    const double* D2 = a.D2;
    double Dab[NA * NB], jab[NA * NB];
    double Dab2[NDEN > 1 ? NA * NB : 1], jab2[NDEN > 1 ? NA * NB : 1];
#pragma unroll
    for (int x = 0; x < NA; ++x)
#pragma unroll
        for (int y = 0; y < NB; ++y) {
            Dab[x * NB + y] = D[(fa + x) * n + fb + y];
            jab[x * NB + y] = 0.0;
            if constexpr (NDEN > 1) {
                Dab2[x * NB + y] = D2[(fa + x) * n + fb + y];
                jab2[x * NB + y] = 0.0;
            }
        }
    const double* O = a.sh_O;
    const R ABx = R(O[3 * sa] - O[3 * sb]), ABy = R(O[3 * sa + 1] - O[3 * sb + 1]), ABz = R(O[3 * sa + 2] - O[3 * sb + 2]);
    const double* Dn = a.Dn;
    const double dnab = Dn[sa * nsh + sb];
    const double qab = a.pair_Q[bra];
    const double fab = (sa == sb ? 1.0 : 2.0) * a.scale;
    unsigned long long cnt_q = 0, cnt_pq = 0, cnt_cpq = 0;
    for (int ch = c0 + wave; ch < c1; ch += W) {
        int ket0, nket, g;
        chunk_kets(a, cut, ch, ket0, nket, g);
        const int ket = ket0 + (lane < nket ? lane : 0);
        const int4 mem4 = a.fam_mem[ket];
        const int mem[FAM_MAXMEM] = {mem4.x, mem4.y, mem4.z, mem4.w};
        bool act[FAM_MAXMEM];
        bool active = false;
        int nact = 0;
        double dnmax = 0.0;
#pragma unroll
        for (int m = 0; m < FAM_MAXMEM; ++m) {
            act[m] = false;
            const int j = mem[m];
            if (lane < nket && j >= 0 && !(same && ket == bfam && m > bmem)) {
                const int2 scd = a.pair_sh[j];
                const int sc = scd.x, sd = scd.y;
                const double dn = fmax(fmax(dnab, Dn[sc * nsh + sd]),
                                       dnl ? fmax(fmax(sDn[sc], sDn[sd]), fmax(sDn[nsh + sc], sDn[nsh + sd]))
                                           : fmax(fmax(Dn[sa * nsh + sc], Dn[sa * nsh + sd]),
                                                  fmax(Dn[sb * nsh + sc], Dn[sb * nsh + sd])));
                const double bound = dn * qab * a.pair_Q[j];
                act[m] = bound >= lo && bound < hi;
                if (act[m]) dnmax = fmax(dnmax, dn);
            }
            active = active || act[m];
            nact += act[m] ? 1 : 0;
        }
        if (!__any(active)) continue;
        const int nkp = active ? a.fam_np[ket] : 0;
        const int off = a.fam_off[ket];
        const float* kq = a.fam_q + off;
        const float cut = float(a.prim_tau / dnmax);
        if (a.counters) {   // algorithmic counts: every member quartet with its own primitive quartets
            cnt_q += (unsigned long long)(wave_sum(double(nact)) + 0.5);
            cnt_pq += (unsigned long long)(wave_sum(double(nact) * a.fam_np[ket]) * nbp + 0.5);
            cnt_cpq += (unsigned long long)(wave_sum(double(screened_prims(bq, nbp, kq, nkp, cut))) + 0.5);
        }
        R v[FAM_MAXMEM * NCART];
        // the family primitive pairs through their FamS / FamSf records
        const auto kprims = [&]() {
            if constexpr (std::is_same<R, double>::value) return a.fam_s + off;
            else return a.fam_sf + off;
        }();
        EriClass<LA, LB, 0, 0>::template run_fam<R, PURE, true>(bprims, nbp, kprims, fkm + FAM_MAXMEM * off,
                                                                nkp, ABx, ABy, ABz, tab, v, bq, kq, cut);
#pragma unroll
        for (int m = 0; m < FAM_MAXMEM; ++m)
            if (act[m]) {
                const int2 scd = a.pair_sh[mem[m]];
                const int sc = scd.x, sd = scd.y;
                const double f = fab * (sc == sd ? 1.0 : 2.0) * (bra == mem[m] ? 1.0 : 2.0);
                const int fc = a.sh_bf[sc], fd = a.sh_bf[sd];
                with_lds_mode(lm, [&](auto M) {
                    constexpr int LM = decltype(M)::value;
                    digest<NA, NB, 1, 1, LM, R>(v + m * NCART, fa, fb, fc, fd, f, Dab, jab, D, n, a.J, a.K, sK, sD);
                    if constexpr (NDEN > 1)
                        digest<NA, NB, 1, 1, LM, R>(v + m * NCART, fa, fb, fc, fd, f, Dab2, jab2, D2, n, a.J2, a.K2,
                                                    sK + (NA + NB) * n, sD + (NA + NB) * n);
                });
            }
    }
    if (a.counters && lane == 0 && cnt_q) {
        atomicAdd(&a.counters[3 * a.combo], cnt_q);
        atomicAdd(&a.counters[3 * a.combo + 1], cnt_pq);
        atomicAdd(&a.counters[3 * a.combo + 2], cnt_cpq);
    }
#pragma unroll
    for (int x = 0; x < NA * NB; ++x) {
        const double s = wave_sum(jab[x]);
        if (lane == 0 && s != 0.0) atomic_add_f64(&a.J[(fa + x / NB) * n + fb + x % NB], s);
        if constexpr (NDEN > 1) {
            const double s2 = wave_sum(jab2[x]);
            if (lane == 0 && s2 != 0.0) atomic_add_f64(&a.J2[(fa + x / NB) * n + fb + x % NB], s2);
        }
    }
    if (!lds) return;
    __syncthreads();
    for (int i = threadIdx.x; i < NDEN * (NA + NB) * n; i += blockDim.x) {
        const double s = sK[i];
        if (s != 0.0) {
            const int set = i / ((NA + NB) * n), r = i - set * (NA + NB) * n;
            const int row = r / n, col = r - row * n;
            atomic_add_f64(&(set ? a.K2 : a.K)[(row < NA ? fa + row : fb + row - NA) * n + col], s);
        }
    }
}

template <int LA, int LB, typename R, bool PURE, int NDEN>
__global__ __launch_bounds__((JKWavesFam<LA, LB, R>::value * 64),
                             (JKWavesFam<LA, LB, R>::min_per_eu)) void k_jk_fam(JKArgs a, BoysTables tab, double lo,
                                                                               double hi) {
    if (a.blk_total && (int)blockIdx.x >= *a.blk_total) return;
    if constexpr (boys_rows_bytes<R, LA + LB>() > 0) {
        if (a.boys_lds) jk_fam_item<LA, LB, R, PURE, NDEN, true>(a, tab, lo, hi, blockIdx.x);
        else jk_fam_item<LA, LB, R, PURE, NDEN, false>(a, tab, lo, hi, blockIdx.x);
    } else {
        jk_fam_item<LA, LB, R, PURE, NDEN, false>(a, tab, lo, hi, blockIdx.x);
    }
}

// Split register classes (csrc/gen_eri.py SPLIT): the quartet runs in one
// pass per Cartesian component P of shell c, each returning the Cartesian
// (a b | c_P d) for all a, b, d (a, b then made spherical).  The pass is
// digested at once: shell c's spherical components m take c_P with the
// factors T[m][P] (EriSplit::CM / CC), so every contraction of
// src/fockbuild.cpp:144-149 with a c index runs over the passes:
//   J(ab) += f sum_d [sum_m T D(c_m, d)] (ab|c_P d)
//   J(c_m d) += f T sum_ab D(ab) (ab|c_P d)
//   K(a c_m) += h T sum_bd D(bd) (ab|c_P d),  K(b c_m) likewise with D(ad)
//   K(bd) += h sum_a [sum_m T D(a c_m)] (ab|c_P d),  K(ad) likewise.
template <int LA, int LB, int LC, int LD>
struct EriSplit;

template <int NA, int NB, int NDEN, int BB = 0>
size_t jk_lds_plan(JKArgs& a, int waves = 4);

template <int P, int N>
struct PassLoop {
    template <typename F>
    static __device__ __forceinline__ void run(F& f) {
        f(std::integral_constant<int, P>{});
        PassLoop<P + 1, N>::run(f);
    }
};
template <int N>
struct PassLoop<N, N> {
    template <typename F>
    static __device__ __forceinline__ void run(F&) {}
};

// pass P of a split class: its d components [d0, d0 + nd) and the
// spherical components of shell c its Cartesian component feeds
template <typename S, bool TC, int P>
struct PassTerms {
    static constexpr int c = S::PC[P];
    static constexpr int d0 = S::PD0[P];
    static constexpr int nd = S::PND[P];
    static constexpr int n = TC ? S::CN[c] : 1;
    static constexpr __device__ int m(int k) { return TC ? S::CM[c][k] : c; }
    static constexpr __device__ double t(int k) { return TC ? S::CC[c][k] : 1.0; }
};

template <typename PT, int NA, int NB, int LM, typename R>
__device__ __forceinline__ void split_digest(const R* __restrict__ v, int fa, int fb, int fc, int fd, double f,
                                             const double* __restrict__ Dab, double* __restrict__ jab,
                                             const double* __restrict__ D, int n, double* __restrict__ J,
                                             double* __restrict__ K, double* __restrict__ sK,
                                             const double* __restrict__ sD) {
    const double h = 0.5 * f;
    constexpr int ND = PT::nd;
    fd += PT::d0;
    // J(ab)
#pragma unroll
    for (int d = 0; d < ND; ++d) {
        double dcd = 0.0;
#pragma unroll
        for (int k = 0; k < PT::n; ++k) dcd += PT::t(k) * D[(fc + PT::m(k)) * n + fd + d];
        dcd *= f;
#pragma unroll
        for (int x = 0; x < NA * NB; ++x) jab[x] += dcd * double(v[x * ND + d]);
    }
    // K(a c_m) and K(a d)
#pragma unroll
    for (int a = 0; a < NA; ++a) {
        double s = 0.0;
#pragma unroll
        for (int b = 0; b < NB; ++b)
#pragma unroll
            for (int d = 0; d < ND; ++d) s += drow<NA, LM>(sD, D, n, fa, fb, NA + b, fd + d) * double(v[(a * NB + b) * ND + d]);
#pragma unroll
        for (int k = 0; k < PT::n; ++k) krow_add<LM>(sK, K, n, a, fa + a, fc + PT::m(k), h * PT::t(k) * s);
    }
#pragma unroll
    for (int a = 0; a < NA; ++a)
#pragma unroll
        for (int d = 0; d < ND; ++d) {
            double s = 0.0;
#pragma unroll
            for (int b = 0; b < NB; ++b) {
                double dbc = 0.0;
#pragma unroll
                for (int k = 0; k < PT::n; ++k) dbc += PT::t(k) * drow<NA, LM>(sD, D, n, fa, fb, NA + b, fc + PT::m(k));
                s += dbc * double(v[(a * NB + b) * ND + d]);
            }
            krow_add<LM>(sK, K, n, a, fa + a, fd + d, h * s);
        }
    // K(b c_m) and K(b d)
#pragma unroll
    for (int b = 0; b < NB; ++b) {
        double s = 0.0;
#pragma unroll
        for (int a = 0; a < NA; ++a)
#pragma unroll
            for (int d = 0; d < ND; ++d) s += drow<NA, LM>(sD, D, n, fa, fb, a, fd + d) * double(v[(a * NB + b) * ND + d]);
#pragma unroll
        for (int k = 0; k < PT::n; ++k) krow_add<LM>(sK, K, n, NA + b, fb + b, fc + PT::m(k), h * PT::t(k) * s);
    }
#pragma unroll
    for (int b = 0; b < NB; ++b)
#pragma unroll
        for (int d = 0; d < ND; ++d) {
            double s = 0.0;
#pragma unroll
            for (int a = 0; a < NA; ++a) {
                double dac = 0.0;
#pragma unroll
                for (int k = 0; k < PT::n; ++k) dac += PT::t(k) * drow<NA, LM>(sD, D, n, fa, fb, a, fc + PT::m(k));
                s += dac * double(v[(a * NB + b) * ND + d]);
            }
            krow_add<LM>(sK, K, n, NA + b, fb + b, fd + d, h * s);
        }
    // J(c_m d), global atomics last
#pragma unroll
    for (int d = 0; d < ND; ++d) {
        double s = 0.0;
#pragma unroll
        for (int x = 0; x < NA * NB; ++x) s += Dab[x] * double(v[x * ND + d]);
#pragma unroll
        for (int k = 0; k < PT::n; ++k) atomic_add_f64(&J[(fc + PT::m(k)) * n + fd + d], f * PT::t(k) * s);
    }
}

// One work item of k_jk_split: as jk_item, with the quartet evaluated and
// digested pass by pass.  4-wave workgroups, one wavefront per SIMD, so a
// pass may use the whole register file.
constexpr int JK_SPLIT_WAVES = 4;
template <int LA, int LB, int LC, int LD, typename R, bool PURE, int NDEN>
__device__ __forceinline__ void jk_split_item(const JKArgs& a, const BoysTables& tab, double lo, double hi,
                                              int witem) {
    using S = EriSplit<LA, LB, LC, LD>;
    constexpr int NA = NFun<LA, PURE>::value, NB = NFun<LB, PURE>::value;
    static_assert(LD <= 1, "the d index of a split class is not transformed");
    constexpr bool TC = PURE && LC >= 2;
    constexpr int W = JK_SPLIT_WAVES;
    typedef const __attribute__((address_space(4))) int cint;
    const int bl = ((cint*)a.blk_bra)[witem];
    const int lb = witem - ((cint*)a.blk_off)[bl];
    const int bra = a.bra0 + bl;
    cint* cut = (cint*)a.cut + 2 * bl * JK_MAXG;
    int nch = 0;
    for (int g = 0; g < a.ngrp; ++g) {
        const int mid = cut[2 * g], end = cut[2 * g + 1];
        nch += ((a.tail ? end - mid : mid - a.gbeg[g]) + 63) >> 6;
    }
    const int c0 = lb * a.block_tasks, c1 = min(c0 + a.block_tasks, nch);
    const int lane = __lane_id();
    const int wave = threadIdx.x >> 6;
    const int2 sab = a.pair_sh[bra];
    const int sa = __builtin_amdgcn_readfirstlane(sab.x), sb = __builtin_amdgcn_readfirstlane(sab.y);
    const int n = a.n, nsh = a.nsh;
    const int fa = a.sh_bf[sa], fb = a.sh_bf[sb];
    const bool lds = a.k_lds != 0;
    using PP = typename PrimPairOf<R>::type;
    const PP* gp = PrimPairOf<R>::get(a);
    typedef const __attribute__((address_space(4))) PP cPP;
    const int nbp = a.pair_np[bra];
    cPP* bprims = (cPP*)(gp + a.pair_off[bra]);
    typedef const __attribute__((address_space(4))) float cfl;
    cfl* bq = (cfl*)(a.prim_q + a.pair_off[bra]);
    extern __shared__ double sK[];
    const double *sD, *sDn;
    stage_bra_rows<NA, NB, NDEN>(a, sa, sb, fa, fb, sK, sD, sDn);
    const bool dnl = a.d_lds != 0;
    const int lm = lds ? (a.d_lds == 2 ? 2 : 1) : 0;   // LDS mode of the digestion
    const double* D = a.D;
    const double* D2 = a.D2;
    double Dab[NA * NB], jab[NA * NB];
    double Dab2[NDEN > 1 ? NA * NB : 1], jab2[NDEN > 1 ? NA * NB : 1];
#pragma unroll
    for (int x = 0; x < NA; ++x)
#pragma unroll
        for (int y = 0; y < NB; ++y) {
            Dab[x * NB + y] = D[(fa + x) * n + fb + y];
            jab[x * NB + y] = 0.0;
            if constexpr (NDEN > 1) {
                Dab2[x * NB + y] = D2[(fa + x) * n + fb + y];
                jab2[x * NB + y] = 0.0;
            }
        }
    const double* O = a.sh_O;
    const R ABx = R(O[3 * sa] - O[3 * sb]), ABy = R(O[3 * sa + 1] - O[3 * sb + 1]), ABz = R(O[3 * sa + 2] - O[3 * sb + 2]);
    const double* Dn = a.Dn;
    const double dnab = Dn[sa * nsh + sb];
    const double qab = a.pair_Q[bra];
    const double fab = (sa == sb ? 1.0 : 2.0) * a.scale;
    unsigned long long cnt_q = 0, cnt_pq = 0, cnt_cpq = 0;
    for (int ch = c0 + wave; ch < c1; ch += W) {
        int ket0, nket, g;
        chunk_kets(a, cut, ch, ket0, nket, g);
        const int ket = ket0 + (lane < nket ? lane : 0);
        const PairRec pr = a.pair_rec[ket];   // one 64-byte record per ket
        const int sc = pr.sc, sd = pr.sd;
        // exact screening of src/fockbuild.cpp:98-115
        const double dn = fmax(fmax(dnab, Dn[sc * nsh + sd]),
                               dnl ? fmax(fmax(sDn[sc], sDn[sd]), fmax(sDn[nsh + sc], sDn[nsh + sd]))
                                   : fmax(fmax(Dn[sa * nsh + sc], Dn[sa * nsh + sd]),
                                          fmax(Dn[sb * nsh + sc], Dn[sb * nsh + sd])));
        const double bound = dn * qab * pr.Q;
        const bool active = lane < nket && bound >= lo && bound < hi;
        if (!__any(active)) continue;
        const float* kq = a.prim_q + pr.off;
        const float cut = float(a.prim_tau / dn);
        const int nkp = active ? pr.np : 0;
        if (a.counters) {
            cnt_q += __popcll(__ballot(active));
            cnt_pq += (unsigned long long)(wave_sum(double(nkp)) * nbp + 0.5);
            cnt_cpq += (unsigned long long)(wave_sum(double(screened_prims(bq, nbp, kq, nkp, cut))) + 0.5);
        }
        const R CDx = R(pr.CD[0]), CDy = R(pr.CD[1]), CDz = R(pr.CD[2]);
        const int fc = pr.fc, fd = pr.fd;
        const double f = fab * (sc == sd ? 1.0 : 2.0) * (bra == ket ? 1.0 : 2.0);
        const PP* kprims = gp + pr.off;
        auto pass = [&](auto PC) {
            constexpr int P = decltype(PC)::value;
            using PT = PassTerms<S, TC, P>;
            R v[NA * NB * PT::nd];
            S::template run<P, R, PURE, true>(bprims, nbp, kprims, nkp, ABx, ABy, ABz, CDx, CDy, CDz, tab, v, bq, kq,
                                              cut);
            if (active)
                with_lds_mode(lm, [&](auto M) {
                    constexpr int LM = decltype(M)::value;
                    split_digest<PT, NA, NB, LM, R>(v, fa, fb, fc, fd, f, Dab, jab, D, n, a.J, a.K, sK, sD);
                    if constexpr (NDEN > 1)
                        split_digest<PT, NA, NB, LM, R>(v, fa, fb, fc, fd, f, Dab2, jab2, D2, n, a.J2, a.K2,
                                                        sK + (NA + NB) * n, sD + (NA + NB) * n);
                });
        };
        PassLoop<0, S::NPASS>::run(pass);
    }
    if (a.counters && lane == 0 && cnt_q) {
        atomicAdd(&a.counters[3 * a.combo], cnt_q);
        atomicAdd(&a.counters[3 * a.combo + 1], cnt_pq);
        atomicAdd(&a.counters[3 * a.combo + 2], cnt_cpq);
    }
#pragma unroll
    for (int x = 0; x < NA * NB; ++x) {
        const double s = wave_sum(jab[x]);
        if (lane == 0 && s != 0.0) atomic_add_f64(&a.J[(fa + x / NB) * n + fb + x % NB], s);
        if constexpr (NDEN > 1) {
            const double s2 = wave_sum(jab2[x]);
            if (lane == 0 && s2 != 0.0) atomic_add_f64(&a.J2[(fa + x / NB) * n + fb + x % NB], s2);
        }
    }
    if (!lds) return;
    __syncthreads();
    for (int i = threadIdx.x; i < NDEN * (NA + NB) * n; i += blockDim.x) {
        const double s = sK[i];
        if (s != 0.0) {
            const int set = i / ((NA + NB) * n), r = i - set * (NA + NB) * n;
            const int row = r / n, col = r - row * n;
            atomic_add_f64(&(set ? a.K2 : a.K)[(row < NA ? fa + row : fb + row - NA) * n + col], s);
        }
    }
}

template <int LA, int LB, int LC, int LD, typename R, bool PURE, int NDEN>
__global__ __launch_bounds__(JK_SPLIT_WAVES * 64, 1) void k_jk_split(JKArgs a, BoysTables tab, double lo, double hi) {
    if (a.blk_total && (int)blockIdx.x >= *a.blk_total) return;
    jk_split_item<LA, LB, LC, LD, R, PURE, NDEN>(a, tab, lo, hi, blockIdx.x);
}

template <int LA, int LB, int LC, int LD, bool PURE, int NDEN>
void launch_jk_split_n(const JKArgs& a0, const BoysTables& tab, int nblocks, int mode, hipStream_t st) {
    constexpr int NA = NFun<LA, PURE>::value, NB = NFun<LB, PURE>::value;
    JKArgs a = a0;
    const size_t lds = jk_lds_plan<NA, NB, NDEN>(a);
    const double inf = __builtin_huge_val();
    if (mode == 2)
        hipLaunchKernelGGL((k_jk_split<LA, LB, LC, LD, float, PURE, NDEN>), dim3(nblocks), dim3(64 * JK_SPLIT_WAVES), lds,
                           st, a, tab, a.precision, inf);
    else
        hipLaunchKernelGGL((k_jk_split<LA, LB, LC, LD, double, PURE, NDEN>), dim3(nblocks), dim3(64 * JK_SPLIT_WAVES),
                           lds, st, a, tab, a.precision, inf);
}

template <int LA, int LB, int LC, int LD, bool PURE>
void launch_jk_split(const JKArgs& a0, const BoysTables& tab, int nblocks, int mode, hipStream_t st) {
    if (nblocks <= 0) return;
    if (a0.nden > 1) launch_jk_split_n<LA, LB, LC, LD, PURE, 2>(a0, tab, nblocks, mode, st);
    else launch_jk_split_n<LA, LB, LC, LD, PURE, 1>(a0, tab, nblocks, mode, st);
}

// Schwarz: Q(pair) = sqrt(max_ab |(ab|ab)|), one lane per pair
// (IntegralEngine::compute_schwarz_ints, src/integrals.cpp:471-513).
template <int LA, int LB, bool PURE>
__global__ __launch_bounds__(256) void k_schwarz(JKArgs a, BoysTables tab, int p0, int np, double* __restrict__ Qout) {
    constexpr int NA = NFun<LA, PURE>::value, NB = NFun<LB, PURE>::value;
    constexpr int N2 = NA * NB;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= np) return;
    const int pr = p0 + t;
    const int2 s = a.pair_sh[pr];
    const double* O = a.sh_O;
    const double ABx = O[3 * s.x] - O[3 * s.y], ABy = O[3 * s.x + 1] - O[3 * s.y + 1], ABz = O[3 * s.x + 2] - O[3 * s.y + 2];
    double v[N2 * N2 > EriClass<LA, LB, LA, LB>::NCART ? N2 * N2 : EriClass<LA, LB, LA, LB>::NCART];
    const PrimPair* pp = a.prims + a.pair_off[pr];
    const int npp = a.pair_np[pr];
    EriClass<LA, LB, LA, LB>::template run<double, PURE>(pp, npp, pp, npp, ABx, ABy, ABz, ABx, ABy, ABz, tab, v);
    double mx = 0.0;
#pragma unroll
    for (int i = 0; i < N2; ++i) mx = fmax(mx, fabs(v[i * N2 + i]));
    Qout[pr] = sqrt(mx);
}

// LDS of one J/K workgroup: the K rows when they fit JK_KROW_LDS_MAX, then
// the staged density rows as far as the whole stays within JK_DROW_LDS_MAX
// (norm rows first, D rows too when there is room).  Sets a.k_lds, a.d_lds.
// BB > 0: the kernel can also read its BB bytes of compact Boys rows from
// LDS.  They are added (a.boys_lds, a.boys_off) when that keeps the
// workgroups per CU the LDS allowed (160 KB per CU, 512-byte granules) or
// leaves room for at least 5 waves per SIMD, if need be in place of the
// staged D rows (the Boys rows are read once per primitive quartet, the D
// rows once per quartet).
template <int NA, int NB, int NDEN, int BB>
size_t jk_lds_plan(JKArgs& a, int waves) {
    const size_t rows = sizeof(double) * (size_t)NDEN * (NA + NB) * a.n;
    const size_t norms = sizeof(double) * 2 * (size_t)a.nsh;
    a.k_lds = rows <= JK_KROW_LDS_MAX ? 1 : 0;
    a.d_lds = 0;
    a.boys_lds = 0;
    a.boys_off = 0;
    auto base = [&](int dl) { return a.k_lds ? rows + (dl == 2 ? rows : 0) + (dl ? norms : 0) : 0; };
    if (a.k_lds) {
        if (2 * rows + norms <= JK_DROW_LDS_MAX && a.d_lds_max >= 2) a.d_lds = 2;
        else if (rows + norms <= JK_DROW_LDS_MAX && a.d_lds_max >= 1) a.d_lds = 1;
    }
    if (BB > 0 && a.boys_lds_max) {
        constexpr size_t CU = 160 * 1024, G = 512;
        auto per_cu = [&](size_t b) { return b ? CU / ((b + G - 1) / G * G) : CU; };
        const size_t occ = per_cu(base(a.d_lds));
        for (int dl = a.d_lds; dl >= (a.d_lds ? 1 : 0); --dl) {
            const size_t off = (base(dl) + 15) & ~size_t(15), with = off + BB;
            if (with <= CU && (per_cu(with) >= occ || per_cu(with) * (size_t)waves >= 20)) {
                a.d_lds = dl;
                a.boys_lds = 1;
                a.boys_off = (int)off;
                return with;
            }
        }
    }
    return base(a.d_lds);
}

// mode 0: FP64 kernel over the FP64 heads of the ket groups; mode 2: FP32
// kernel over the FP32 tails (a.tail set by the caller).  `nblocks`
// workgroups of the work list in a.blk_bra.
template <int LA, int LB, int LC, int LD, bool PURE, int NDEN>
void launch_jk_n(const JKArgs& a0, const BoysTables& tab, int nblocks, int mode, hipStream_t st) {
    constexpr int NA = NFun<LA, PURE>::value, NB = NFun<LB, PURE>::value;
    JKArgs a = a0;
    const double inf = __builtin_huge_val();
    constexpr int M = LA + LB + LC + LD;
    if (mode == 2) {
        auto k = k_jk<LA, LB, LC, LD, float, PURE, NDEN>;
        constexpr int W = JKOccupancy<LA, LB, LC, LD, float>::waves;
        const size_t lds = jk_lds_plan<NA, NB, NDEN, boys_rows_bytes<float, M>()>(a, W);
        hipLaunchKernelGGL(k, dim3(nblocks), dim3(64 * W), lds, st, a, tab, a.precision, inf);
    } else {
        auto k = k_jk<LA, LB, LC, LD, double, PURE, NDEN>;
        constexpr int W = JKOccupancy<LA, LB, LC, LD, double>::waves;
        const size_t lds = jk_lds_plan<NA, NB, NDEN, boys_rows_bytes<double, M>()>(a, W);
        hipLaunchKernelGGL(k, dim3(nblocks), dim3(64 * W), lds, st, a, tab, a.precision, inf);
    }
}

template <int LA, int LB, int LC, int LD, bool PURE>
void launch_jk(const JKArgs& a0, const BoysTables& tab, int nblocks, int mode, hipStream_t st) {
    if (nblocks <= 0) return;
    if (a0.nden > 1) launch_jk_n<LA, LB, LC, LD, PURE, 2>(a0, tab, nblocks, mode, st);
    else launch_jk_n<LA, LB, LC, LD, PURE, 1>(a0, tab, nblocks, mode, st);
}

template <int LA, int LB, bool PURE, int NDEN>
void launch_jk_fam_n(const JKArgs& a0, const BoysTables& tab, int nblocks, int mode, hipStream_t st) {
    constexpr int NA = NFun<LA, PURE>::value, NB = NFun<LB, PURE>::value;
    JKArgs a = a0;
    const double inf = __builtin_huge_val();
    if (mode == 2) {
        auto k = k_jk_fam<LA, LB, float, PURE, NDEN>;
        constexpr int W = JKWavesFam<LA, LB, float>::value;
        const size_t lds = jk_lds_plan<NA, NB, NDEN, boys_rows_bytes<float, LA + LB>()>(a, W);
        hipLaunchKernelGGL(k, dim3(nblocks), dim3(64 * W), lds, st, a, tab, a.precision, inf);
    } else {
        auto k = k_jk_fam<LA, LB, double, PURE, NDEN>;
        constexpr int W = JKWavesFam<LA, LB, double>::value;
        const size_t lds = jk_lds_plan<NA, NB, NDEN, boys_rows_bytes<double, LA + LB>()>(a, W);
        hipLaunchKernelGGL(k, dim3(nblocks), dim3(64 * W), lds, st, a, tab, a.precision, inf);
    }
}

template <int LA, int LB, bool PURE>
void launch_jk_fam(const JKArgs& a0, const BoysTables& tab, int nblocks, int mode, hipStream_t st) {
    if (nblocks <= 0) return;
    if (a0.nden > 1) launch_jk_fam_n<LA, LB, PURE, 2>(a0, tab, nblocks, mode, st);
    else launch_jk_fam_n<LA, LB, PURE, 1>(a0, tab, nblocks, mode, st);
}

template <int LA, int LB, bool PURE>
void launch_schwarz(const JKArgs& a, const BoysTables& tab, int p0, int np, double* Q, hipStream_t st) {
    if (np <= 0) return;
    hipLaunchKernelGGL((k_schwarz<LA, LB, PURE>), dim3((np + 255) / 256), dim3(256), 0, st, a, tab, p0, np, Q);
}

}  // namespace m2
