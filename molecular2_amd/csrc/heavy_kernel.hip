// Heavy shell-quartet classes ((dp|dp), (dd|pp), (dd|ds), (dd|dp), (dd|dd)):
// one wavefront per quartet executing the class's generated recurrence
// program out of LDS.
//
// The program (csrc/gen_eri.py: gen_program) is the same Obara-Saika VRR +
// HGP HRR DAG the register-resident classes use, flattened into levels of
// independent instructions dst = sum_t coef_t * S[sid_t] * X[src_t].  The 64
// lanes execute one level at a time (barrier between levels): per primitive
// quartet lane m <= L evaluates F_m(T) straight from the Boys table, lane 0
// writes the per-primitive scalars (PA, WP, QC, WQ, 1/2p, ...), the VRR
// levels run, the [e0|f0] are accumulated; after the primitive loops the
// HRR / solid-harmonic levels run and the quartet is digested into J and K
// with the same screening, degeneracy and factors as the register kernels.
// Slots are reused by liveness, so (dd|dd) needs 3179 slots (25 KB FP64).
#include <mutex>
#include "m2_common.hpp"
#include "boys.hpp"
#include "generated/programs.inc"

namespace m2 {

constexpr int NSCALAR = 24;
// LDS scalar slots: the NSCALAR per-primitive scalars, then up to 8 constant
// coefficients of the pass-ordered post program (POSTW_COEF)
constexpr int NSCALAR_LDS = NSCALAR + 8;

struct HeavyProg {
    int L, nslots, base, nacc;
    const int* acc;      // nacc pairs (acc slot, vrr slot)
    const int* vlev;     // n_vrr_lev + 1 instruction offsets
    int nvlev;
    const int* plev;     // n_post_lev + 1 offsets
    int nplev;
    const int* out;      // nout slots, (a,b,c,d) order
    int na, nb, nc, nd;
    const int* ins;      // 14 ints per instruction
    const double* coef;  // PROG_MAXTERMS per instruction
    const int* vrrw;     // VRR instruction words in pass order (VRRW_WORDS of the class)
    int nvpass;          // VRR passes
    unsigned long long levend;   // bit p: pass p ends a VRR level
    const int* postw;    // post (HRR, transform) words in pass order (POSTW_WORDS)
    const int* pend;     // per post pass: 1 when it ends a level
    int npp;             // post passes
    const double* pcoef; // the program's non-unit coefficients (S[NSCALAR + i])
    int ncoef;
};

// The HRR / solid-harmonic levels from their pass-ordered words, next pass
// prefetched as in vrr_passes; a term is S[sid] * X[src] (coefficients live
// in S[NSCALAR + i]).
template <typename R>
__device__ __forceinline__ void post_passes(const HeavyProg& P, R* __restrict__ X, const R* __restrict__ S) {
    constexpr int NW = 1 + PROG_MAXTERMS;
    const int lane = threadIdx.x;
    const int* wp = P.postw + lane;
    int cur[NW], nxt[NW];
#pragma unroll
    for (int k = 0; k < NW; ++k) cur[k] = wp[k * 64];
    for (int p = 0; p < P.npp; ++p) {
        if (p + 1 < P.npp) {
#pragma unroll
            for (int k = 0; k < NW; ++k) nxt[k] = wp[((p + 1) * NW + k) * 64];
        }
        const int h = cur[0];
        if (h >= 0) {
            const int nt = h >> 16;
            R v = R(0);
#pragma unroll
            for (int t = 0; t < PROG_MAXTERMS; ++t)
                if (t < nt) v += S[cur[1 + t] >> 13] * X[cur[1 + t] & 8191];
            X[h & 65535] = v;
        }
        if (P.pend[p]) __syncthreads();
#pragma unroll
        for (int k = 0; k < NW; ++k) cur[k] = nxt[k];
    }
}

// the post program's coefficients into S[NSCALAR ..] (once per workgroup)
template <typename R>
__device__ __forceinline__ void post_coefs(const HeavyProg& P, R* __restrict__ S) {
    if ((int)threadIdx.x < P.ncoef) S[NSCALAR + threadIdx.x] = R(P.pcoef[threadIdx.x]);
}

// One primitive quartet's VRR from the pass-ordered instruction words: in
// pass p lane j computes one [e0|f0]^(m) slot from <= 6 terms coef * S[sid]
// * X[src].  The next pass's words load while this pass computes, so the
// table fetch overlaps the LDS work; a barrier ends every level.
template <typename R>
__device__ __forceinline__ void vrr_passes(const HeavyProg& P, R* __restrict__ X, const R* __restrict__ S) {
    constexpr int NW = 1 + PROG_MAXTERMS;
    const int lane = threadIdx.x;
    const int* wp = P.vrrw + lane;
    int cur[NW], nxt[NW];
#pragma unroll
    for (int k = 0; k < NW; ++k) cur[k] = wp[k * 64];
    for (int p = 0; p < P.nvpass; ++p) {
        if (p + 1 < P.nvpass) {
#pragma unroll
            for (int k = 0; k < NW; ++k) nxt[k] = wp[((p + 1) * NW + k) * 64];
        }
        const int h = cur[0];
        if (h >= 0) {
            const int nt = h >> 16;
            R v = R(0);
#pragma unroll
            for (int t = 0; t < PROG_MAXTERMS; ++t)
                if (t < nt) {
                    const int tw = cur[1 + t];
                    const int c6 = (tw >> 18) & 63;
                    v += R(c6 >= 32 ? c6 - 64 : c6) * S[(tw >> 13) & 31] * X[tw & 8191];
                }
            X[h & 65535] = v;
        }
        if ((P.levend >> p) & 1ull) __syncthreads();
#pragma unroll
        for (int k = 0; k < NW; ++k) cur[k] = nxt[k];
    }
}

template <typename R>
__device__ __forceinline__ R boys_one(double T, int m, const BoysTables& tab);

template <>
__device__ __forceinline__ double boys_one<double>(double T, int m, const BoysTables& tab) {
    if (T < BOYS_TMAX) {
        const int k = __double2int_rn(T * (1.0 / BOYS_STEP));
        const double d = k * BOYS_STEP - T;
        const double* r = tab.d + k * BOYS_MTAB + m;
        double s = r[7];
        s = r[6] + d * (1.0 / 7.0) * s;
        s = r[5] + d * (1.0 / 6.0) * s;
        s = r[4] + d * (1.0 / 5.0) * s;
        s = r[3] + d * (1.0 / 4.0) * s;
        s = r[2] + d * (1.0 / 3.0) * s;
        s = r[1] + d * (1.0 / 2.0) * s;
        return r[0] + d * s;
    }
    const double rs = rsqrt(T);
    double F = 0.88622692545275801365 * rs;
    const double h = 0.5 * rs * rs;
    for (int i = 0; i < m; ++i) F = (2 * i + 1) * F * h;
    return F;
}

template <>
__device__ __forceinline__ float boys_one<float>(double T, int m, const BoysTables& tab) {
    if (T < BOYS_TMAX) {
        const float Tf = float(T);
        const int k = __float2int_rn(Tf * float(1.0 / BOYS_STEP));
        const float d = k * float(BOYS_STEP) - Tf;
        const float* r = tab.f + k * BOYS_MTAB + m;
        float s = r[4];
        s = r[3] + d * (1.0f / 4.0f) * s;
        s = r[2] + d * (1.0f / 3.0f) * s;
        s = r[1] + d * (1.0f / 2.0f) * s;
        return r[0] + d * s;
    }
    return float(boys_one<double>(T, m, tab));
}

template <typename R>
__device__ __forceinline__ void run_level(const HeavyProg& P, int i0, int i1, R* __restrict__ X, const R* __restrict__ S) {
    for (int j = i0 + (int)threadIdx.x; j < i1; j += 64) {
        const int* in = P.ins + 14 * j;
        const double* cf = P.coef + PROG_MAXTERMS * j;
        const int nt = in[1];
        R v = R(0);
        for (int t = 0; t < nt; ++t) v += R(cf[t]) * S[in[8 + t]] * X[in[2 + t]];
        X[in[0]] = v;
    }
}

// computes the quartet (bra pair | ket pair) into X; returns with all lanes
// synchronised and the outputs at X[P.out[o]]
// NAP > 0: the lane's (accumulator, VRR slot) pairs come in accw (-1: none),
// loaded once per workgroup; NAP == 0: read from P.acc.
template <typename R, int NAP>
__device__ void heavy_quartet(const HeavyProg& P, const JKArgs& a, const BoysTables& tab, int bra, int ket,
                              double ABx, double ABy, double ABz, double CDx, double CDy, double CDz,
                              R* __restrict__ X, R* __restrict__ S, const int (&accw)[NAP > 0 ? NAP : 1][2]) {
    const int lane = threadIdx.x;
    if constexpr (NAP > 0) {
#pragma unroll
        for (int q = 0; q < NAP; ++q)
            if (accw[q][0] >= 0) X[accw[q][0]] = R(0);
    } else {
        for (int t = lane; t < P.nacc; t += 64) X[P.acc[2 * t]] = R(0);
    }
    const PrimPair* bp = a.prims + a.pair_off[bra];
    const PrimPair* kp = a.prims + a.pair_off[ket];
    const int nbp = a.pair_np[bra], nkp = a.pair_np[ket];
    for (int ib = 0; ib < nbp; ++ib) {
        const PrimPair B = bp[ib];
        for (int ik = 0; ik < nkp; ++ik) {
            const PrimPair K = kp[ik];
            const double p = B.p, q = K.p;
            const double dx = B.P[0] - K.P[0], dy = B.P[1] - K.P[1], dz = B.P[2] - K.P[2];
            const double ipq = 1.0 / (p + q);
            const double T = p * q * ipq * (dx * dx + dy * dy + dz * dz);
            __syncthreads();   // previous primitive's readers are done
            if (lane <= P.L) {
                const double pref = B.K * K.K * rsqrt(p + q);
                X[P.base + lane] = R(pref) * boys_one<R>(T, lane, tab);
            }
            if (lane == 0) {
                const double rp = q * ipq, rq = p * ipq;
                S[0] = R(1);
                S[1] = R(B.PA[0]); S[2] = R(B.PA[1]); S[3] = R(B.PA[2]);
                S[4] = R(-rp * dx); S[5] = R(-rp * dy); S[6] = R(-rp * dz);
                S[7] = R(K.PA[0]); S[8] = R(K.PA[1]); S[9] = R(K.PA[2]);
                S[10] = R(rq * dx); S[11] = R(rq * dy); S[12] = R(rq * dz);
                S[13] = R(B.inv2p); S[14] = R(K.inv2p); S[15] = R(0.5 * ipq);
                S[16] = R(B.inv2p * rp); S[17] = R(K.inv2p * rq);
                S[18] = R(ABx); S[19] = R(ABy); S[20] = R(ABz);
                S[21] = R(CDx); S[22] = R(CDy); S[23] = R(CDz);
            }
            __syncthreads();
            if (P.vrrw) {
                vrr_passes<R>(P, X, S);
            } else {
                for (int l = 0; l < P.nvlev; ++l) {
                    run_level<R>(P, P.vlev[l], P.vlev[l + 1], X, S);
                    __syncthreads();
                }
            }
            if constexpr (NAP > 0) {
#pragma unroll
                for (int q = 0; q < NAP; ++q)
                    if (accw[q][0] >= 0) X[accw[q][0]] += X[accw[q][1]];
            } else {
                for (int t = lane; t < P.nacc; t += 64) X[P.acc[2 * t]] += X[P.acc[2 * t + 1]];
            }
        }
    }
    __syncthreads();
    if (P.postw) {
        post_passes<R>(P, X, S);
    } else {
        for (int l = 0; l < P.nplev; ++l) {
            run_level<R>(P, P.plev[l], P.plev[l + 1], X, S);
            __syncthreads();
        }
    }
}

template <typename R>
__device__ __forceinline__ R outv(const HeavyProg& P, const R* X, int a_, int b_, int c_, int d_) {
    return X[P.out[((a_ * P.nb + b_) * P.nc + c_) * P.nd + d_]];
}

// LDS layout of k_jk_heavy: X[nslots] and S[NSCALAR_LDS] (type R), then the
// quartet's outputs Y[NA NB NC ND] in (a,b,c,d) order and the six density
// blocks it is digested with (FP64), 8-byte aligned.
template <typename R>
__host__ __device__ constexpr size_t heavy_y_offset(int nslots) {
    return ((sizeof(R) * (nslots + NSCALAR_LDS) + 7) / 8) * 8;
}

// The six density blocks a quartet is digested with, in LDS: D(c,d), D(a,b),
// D(b,d), D(a,c), D(b,c), D(a,d); one set per density.
template <int NA, int NB, int NC, int ND>
struct HeavyBlocks {
    static constexpr int CD = 0, AB = NC * ND, BD = AB + NA * NB, AC = BD + NB * ND, BC = AC + NA * NC,
                         AD = BC + NB * NC, SIZE = AD + NA * ND;
};

template <int NA, int NB, int NC, int ND>
__device__ __forceinline__ void heavy_stage_ket(const double* __restrict__ D, double* __restrict__ blk, int fa, int fb,
                                                int fc, int fd, int n, int lane) {
    using B = HeavyBlocks<NA, NB, NC, ND>;
    for (int i = lane; i < NC * ND; i += 64) blk[B::CD + i] = D[(fc + i / ND) * n + fd + i % ND];
    for (int i = lane; i < NB * ND; i += 64) blk[B::BD + i] = D[(fb + i / ND) * n + fd + i % ND];
    for (int i = lane; i < NA * NC; i += 64) blk[B::AC + i] = D[(fa + i / NC) * n + fc + i % NC];
    for (int i = lane; i < NB * NC; i += 64) blk[B::BC + i] = D[(fb + i / NC) * n + fc + i % NC];
    for (int i = lane; i < NA * ND; i += 64) blk[B::AD + i] = D[(fa + i / ND) * n + fd + i % ND];
}

// J and K contributions of one quartet (outputs Y in (a,b,c,d) order) with
// one density's blocks, degeneracy factor f (src/fockbuild.cpp:121-149)
template <int NA, int NB, int NC, int ND>
__device__ __forceinline__ void heavy_digest(const double* __restrict__ Y, const double* __restrict__ blk,
                                             double* __restrict__ J, double* __restrict__ K, double f, int fa, int fb,
                                             int fc, int fd, int n, int lane) {
    using B = HeavyBlocks<NA, NB, NC, ND>;
    const double h = 0.5 * f;
    for (int i = lane; i < NA * NB; i += 64) {   // J(ab) += D(cd) (ab|cd)
        const double* y = Y + i * (NC * ND);
        double s = 0.0;
#pragma unroll
        for (int t = 0; t < NC * ND; ++t) s += blk[B::CD + t] * y[t];
        atomicAdd(&J[(fa + i / NB) * n + fb + i % NB], f * s);
    }
    for (int i = lane; i < NC * ND; i += 64) {   // J(cd) += D(ab) (ab|cd)
        double s = 0.0;
#pragma unroll
        for (int t = 0; t < NA * NB; ++t) s += blk[B::AB + t] * Y[t * (NC * ND) + i];
        atomicAdd(&J[(fc + i / ND) * n + fd + i % ND], f * s);
    }
    for (int i = lane; i < NA * NC; i += 64) {   // K(ac) += D(bd) (ab|cd)
        const int x = i / NC, c = i % NC;
        double s = 0.0;
#pragma unroll
        for (int y = 0; y < NB; ++y)
#pragma unroll
            for (int d = 0; d < ND; ++d) s += blk[B::BD + y * ND + d] * Y[((x * NB + y) * NC + c) * ND + d];
        atomicAdd(&K[(fa + x) * n + fc + c], h * s);
    }
    for (int i = lane; i < NB * ND; i += 64) {   // K(bd) += D(ac) (ab|cd)
        const int y = i / ND, d = i % ND;
        double s = 0.0;
#pragma unroll
        for (int x = 0; x < NA; ++x)
#pragma unroll
            for (int c = 0; c < NC; ++c) s += blk[B::AC + x * NC + c] * Y[((x * NB + y) * NC + c) * ND + d];
        atomicAdd(&K[(fb + y) * n + fd + d], h * s);
    }
    for (int i = lane; i < NA * ND; i += 64) {   // K(ad) += D(bc) (ab|cd)
        const int x = i / ND, d = i % ND;
        double s = 0.0;
#pragma unroll
        for (int y = 0; y < NB; ++y)
#pragma unroll
            for (int c = 0; c < NC; ++c) s += blk[B::BC + y * NC + c] * Y[((x * NB + y) * NC + c) * ND + d];
        atomicAdd(&K[(fa + x) * n + fd + d], h * s);
    }
    for (int i = lane; i < NB * NC; i += 64) {   // K(bc) += D(ad) (ab|cd)
        const int y = i / NC, c = i % NC;
        double s = 0.0;
#pragma unroll
        for (int x = 0; x < NA; ++x)
#pragma unroll
            for (int d = 0; d < ND; ++d) s += blk[B::AD + x * ND + d] * Y[((x * NB + y) * NC + c) * ND + d];
        atomicAdd(&K[(fb + y) * n + fc + c], h * s);
    }
}

// One task: bra pair `bra` with the kets [ket0, ket0 + nket), one quartet at a
// time.  NDEN = 2 (unrestricted build): each quartet is digested with D into
// (J, K) and with D2 into (J2, K2).  Ends with a barrier: the LDS it uses is
// free for the next task.
template <typename R, int NA, int NB, int NC, int ND, int NAP, int NDEN>
__device__ __forceinline__ void heavy_task(const JKArgs& a, const BoysTables& tab, const HeavyProg& P, int bra,
                                           int ket0, int nket, double lo, double hi, R* X, R* S, double* Y,
                                           double* blk) {
    using B = HeavyBlocks<NA, NB, NC, ND>;
    constexpr int NOUT = NA * NB * NC * ND;
    const int lane = threadIdx.x;
    const int nsh = a.nsh, n = a.n;
    const int2 sab = a.pair_sh[bra];
    const int sa = sab.x, sb = sab.y;
    const double* O = a.sh_O;
    const double ABx = O[3 * sa] - O[3 * sb], ABy = O[3 * sa + 1] - O[3 * sb + 1], ABz = O[3 * sa + 2] - O[3 * sb + 2];
    const int fa = a.sh_bf[sa], fb = a.sh_bf[sb];
#pragma unroll
    for (int d = 0; d < NDEN; ++d) {
        const double* D = d ? a.D2 : a.D;
        for (int i = lane; i < NA * NB; i += 64) blk[d * B::SIZE + B::AB + i] = D[(fa + i / NB) * n + fb + i % NB];
    }
    post_coefs<R>(P, S);
    int accw[NAP][2];   // this lane's accumulation pairs, the same for every primitive quartet
#pragma unroll
    for (int q = 0; q < NAP; ++q) {
        const int t = 64 * q + lane;
        accw[q][0] = t < P.nacc ? P.acc[2 * t] : -1;
        accw[q][1] = t < P.nacc ? P.acc[2 * t + 1] : 0;
    }
    for (int kk = 0; kk < nket; ++kk) {
        const int ket = ket0 + kk;
        const int2 scd = a.pair_sh[ket];
        const int sc = scd.x, sd = scd.y;
        const double* Dn = a.Dn;
        const double dn = fmax(fmax(Dn[sa * nsh + sb], Dn[sc * nsh + sd]),
                               fmax(fmax(Dn[sa * nsh + sc], Dn[sa * nsh + sd]), fmax(Dn[sb * nsh + sc], Dn[sb * nsh + sd])));
        const double bound = dn * a.pair_Q[bra] * a.pair_Q[ket];
        if (!(bound >= lo && bound < hi)) continue;
        if (a.counters && lane == 0) {
            const unsigned long long npq = (unsigned long long)a.pair_np[bra] * a.pair_np[ket];
            atomicAdd(&a.counters[3 * a.combo], 1ull);
            atomicAdd(&a.counters[3 * a.combo + 1], npq);
            atomicAdd(&a.counters[3 * a.combo + 2], npq);   // no primitive screening here
        }
        const double CDx = O[3 * sc] - O[3 * sd], CDy = O[3 * sc + 1] - O[3 * sd + 1], CDz = O[3 * sc + 2] - O[3 * sd + 2];
        const int fc = a.sh_bf[sc], fd = a.sh_bf[sd];
        // density blocks of this quartet (read by the digestion below, after
        // the barriers inside heavy_quartet)
#pragma unroll
        for (int d = 0; d < NDEN; ++d)
            heavy_stage_ket<NA, NB, NC, ND>(d ? a.D2 : a.D, blk + d * B::SIZE, fa, fb, fc, fd, n, lane);
        heavy_quartet<R, NAP>(P, a, tab, bra, ket, ABx, ABy, ABz, CDx, CDy, CDz, X, S, accw);
        for (int o = lane; o < NOUT; o += 64) Y[o] = double(X[P.out[o]]);
        __syncthreads();
        const double deg = (sa == sb ? 1.0 : 2.0) * (sc == sd ? 1.0 : 2.0) * (bra == ket ? 1.0 : 2.0);
        const double f = deg * a.scale;
#pragma unroll
        for (int d = 0; d < NDEN; ++d)
            heavy_digest<NA, NB, NC, ND>(Y, blk + d * B::SIZE, d ? a.J2 : a.J, d ? a.K2 : a.K, f, fa, fb, fc, fd, n,
                                         lane);
        __syncthreads();   // Y and the density blocks consumed before the next quartet
    }
    __syncthreads();       // the bra's density block and coefficients consumed before the next task
}

// Heavy-class J/K over the device work list: one work item per workgroup, a
// bra and a.block_tasks chunks of JK_HEAVY_KETS kets (FP64 head or FP32
// tail by a.tail).  With a.blk_total set the grid is a bound on the item
// count and workgroups past the count exit at once.  The chunks of an item
// that fall in one primitive-count group form one contiguous ket range.
template <typename R, int NA, int NB, int NC, int ND, int NAP, int NDEN>
__global__ __launch_bounds__(64) void k_jk_heavy(JKArgs a, BoysTables tab, double lo, double hi, HeavyProg P) {
    extern __shared__ __align__(16) unsigned char smem[];
    R* X = reinterpret_cast<R*>(smem);
    R* S = X + P.nslots;
    double* Y = reinterpret_cast<double*>(smem + heavy_y_offset<R>(P.nslots));
    double* blk = Y + NA * NB * NC * ND;   // NDEN sets of HeavyBlocks
    const int w = blockIdx.x;
    if (a.blk_total && w >= *a.blk_total) return;
    const int bl = a.blk_bra[w];
    const int c0 = (w - a.blk_off[bl]) * a.block_tasks, c1 = c0 + a.block_tasks;
    const int bra = a.bra0 + bl;
    const int2* cut = a.cut + bl * JK_MAXG;   // (mid, end) per group
    int cb = 0;                               // chunks of the groups before g
    for (int g = 0; g < a.ngrp && cb < c1; ++g) {
        const int2 me = cut[g];
        const int k0 = a.tail ? me.x : a.gbeg[g], k1 = a.tail ? me.y : me.x;
        const int nch = (k1 - k0 + JK_HEAVY_KETS - 1) / JK_HEAVY_KETS;
        const int s0 = max(c0, cb), s1 = min(c1, cb + nch);
        if (s0 < s1) {
            const int j0 = k0 + (s0 - cb) * JK_HEAVY_KETS, j1 = min(k1, k0 + (s1 - cb) * JK_HEAVY_KETS);
            heavy_task<R, NA, NB, NC, ND, NAP, NDEN>(a, tab, P, bra, j0, j1 - j0, lo, hi, X, S, Y, blk);
        }
        cb += nch;
    }
}

// Schwarz of heavy pair classes: one wavefront per pair, (ab|ab) in FP64
__global__ __launch_bounds__(64) void k_schwarz_heavy(JKArgs a, BoysTables tab, int p0, int np, HeavyProg P,
                                                      double* __restrict__ Qout) {
    extern __shared__ __align__(16) unsigned char smem[];
    double* X = reinterpret_cast<double*>(smem);
    double* S = X + P.nslots;
    const int w = blockIdx.x;
    if (w >= np) return;
    const int pr = p0 + w;
    const int2 s = a.pair_sh[pr];
    const double* O = a.sh_O;
    const double ABx = O[3 * s.x] - O[3 * s.y], ABy = O[3 * s.x + 1] - O[3 * s.y + 1], ABz = O[3 * s.x + 2] - O[3 * s.y + 2];
    post_coefs<double>(P, S);
    const int acc0[1][2] = {{-1, 0}};
    heavy_quartet<double, 0>(P, a, tab, pr, pr, ABx, ABy, ABz, ABx, ABy, ABz, X, S, acc0);
    if (threadIdx.x == 0) {
        double mx = 0.0;
        for (int x = 0; x < P.na; ++x)
            for (int y = 0; y < P.nb; ++y) mx = fmax(mx, fabs(outv(P, X, x, y, x, y)));
        Qout[pr] = sqrt(mx);
    }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------

struct HeavyTables {
    int* vrrw = nullptr;
    int* postw = nullptr;
    int* pend = nullptr;
    double* pcoef = nullptr;
    int* ins = nullptr;
    double* coef = nullptr;
    int* lev = nullptr;
    int* acc = nullptr;
    int* out = nullptr;
};

static HeavyTables g_heavy[64];   // per device

hipError_t heavy_upload(int device) {
    if (device < 0 || device >= 64) return hipErrorInvalidDevice;
    static std::mutex mu;   // contexts may be created from several host threads
    std::lock_guard<std::mutex> lk(mu);
    HeavyTables& t = g_heavy[device];
    if (t.ins) return hipSuccess;
    hipError_t e;
    if ((e = hipMalloc((void**)&t.ins, sizeof(PROG_INS)))) return e;
    if ((e = hipMalloc((void**)&t.coef, sizeof(PROG_COEF)))) return e;
    if ((e = hipMalloc((void**)&t.lev, sizeof(PROG_LEV)))) return e;
    if ((e = hipMalloc((void**)&t.acc, sizeof(PROG_ACC)))) return e;
    if ((e = hipMalloc((void**)&t.out, sizeof(PROG_OUT)))) return e;
    if ((e = hipMalloc((void**)&t.vrrw, sizeof(VRRW_WORDS)))) return e;
    if ((e = hipMalloc((void**)&t.postw, sizeof(POSTW_WORDS)))) return e;
    if ((e = hipMalloc((void**)&t.pend, sizeof(POSTW_END)))) return e;
    if ((e = hipMalloc((void**)&t.pcoef, sizeof(POSTW_COEF)))) return e;
    if ((e = hipMemcpy(t.postw, POSTW_WORDS, sizeof(POSTW_WORDS), hipMemcpyHostToDevice))) return e;
    if ((e = hipMemcpy(t.pend, POSTW_END, sizeof(POSTW_END), hipMemcpyHostToDevice))) return e;
    if ((e = hipMemcpy(t.pcoef, POSTW_COEF, sizeof(POSTW_COEF), hipMemcpyHostToDevice))) return e;
    if ((e = hipMemcpy(t.vrrw, VRRW_WORDS, sizeof(VRRW_WORDS), hipMemcpyHostToDevice))) return e;
    if ((e = hipMemcpy(t.ins, PROG_INS, sizeof(PROG_INS), hipMemcpyHostToDevice))) return e;
    if ((e = hipMemcpy(t.coef, PROG_COEF, sizeof(PROG_COEF), hipMemcpyHostToDevice))) return e;
    if ((e = hipMemcpy(t.lev, PROG_LEV, sizeof(PROG_LEV), hipMemcpyHostToDevice))) return e;
    if ((e = hipMemcpy(t.acc, PROG_ACC, sizeof(PROG_ACC), hipMemcpyHostToDevice))) return e;
    return hipMemcpy(t.out, PROG_OUT, sizeof(PROG_OUT), hipMemcpyHostToDevice);
}

static bool find_prog(int device, int la, int lb, int lc, int ld, int pure, HeavyProg& P) {
    const HeavyTables& t = g_heavy[device];
    for (int k = 0; k < PROG_NCLASS; ++k) {
        const int* d = PROG_DESC + k * PROG_DESC_FIELDS;
        if (d[0] == la && d[1] == lb && d[2] == lc && d[3] == ld && d[4] == pure) {
            // fields: la lb lc ld pure L nslots base nacc acc_off vrr_lev_off n_vrr_lev post_lev_off
            //         n_post_lev out_off nout na nb nc nd
            P.L = d[5];
            P.nslots = d[6];
            P.base = d[7];
            P.nacc = d[8];
            P.acc = t.acc + 2 * d[9];
            P.vlev = t.lev + d[10];
            P.nvlev = d[11];
            P.plev = t.lev + d[12];
            P.nplev = d[13];
            P.out = t.out + d[14];
            P.na = d[16];
            P.nb = d[17];
            P.nc = d[18];
            P.nd = d[19];
            P.ins = t.ins;
            P.coef = t.coef;
            P.vrrw = nullptr;
            P.nvpass = 0;
            P.levend = 0;
            P.postw = nullptr;
            P.pend = nullptr;
            P.npp = 0;
            P.pcoef = nullptr;
            P.ncoef = 0;
            for (int v = 0; v < POSTW_NPROG; ++v) {
                const int* q = POSTW_DESC + 10 * v;
                if (q[0] == la && q[1] == lb && q[2] == lc && q[3] == ld && q[4] == pure) {
                    P.npp = q[5];
                    P.postw = t.postw + q[6];
                    P.pcoef = t.pcoef + q[7];
                    P.ncoef = q[8];
                    P.pend = t.pend + q[9];
                }
            }
            for (int v = 0; v < VRRW_NCLASS; ++v) {
                const int* q = VRRW_DESC + 6 * v;
                if (q[0] == la && q[1] == lb && q[2] == lc && q[3] == ld) {
                    P.vrrw = t.vrrw + q[5];
                    P.nvpass = q[4];
                    P.levend = VRRW_LEVEND[v];
                }
            }
            return true;
        }
    }
    return false;
}

bool heavy_has(int la, int lb, int lc, int ld) {
    for (int k = 0; k < PROG_NCLASS; ++k) {
        const int* d = PROG_DESC + k * PROG_DESC_FIELDS;
        if (d[0] == la && d[1] == lb && d[2] == lc && d[3] == ld) return true;
    }
    return false;
}

// nitems: the grid (the item count, or a bound on it when a.blk_total is set)
template <int NA, int NB, int NC, int ND, int NAP, int NDEN>
static void launch_heavy_dims(const HeavyProg& P, const JKArgs& a, const BoysTables& tab, int nitems, int mode,
                              hipStream_t st) {
    const double inf = __builtin_huge_val();
    const size_t extra = sizeof(double) * (NA * NB * NC * ND + NDEN * HeavyBlocks<NA, NB, NC, ND>::SIZE);
    if (mode == 2) {
        const size_t sm = heavy_y_offset<float>(P.nslots) + extra;
        hipLaunchKernelGGL((k_jk_heavy<float, NA, NB, NC, ND, NAP, NDEN>), dim3(nitems), dim3(64), sm, st, a, tab,
                           a.precision, inf, P);
    } else {
        const size_t sm = heavy_y_offset<double>(P.nslots) + extra;
        hipLaunchKernelGGL((k_jk_heavy<double, NA, NB, NC, ND, NAP, NDEN>), dim3(nitems), dim3(64), sm, st, a, tab,
                           a.precision, inf, P);
    }
}

template <int LA, int LB, int LC, int LD>
static void launch_heavy_class(int pure, const HeavyProg& P, const JKArgs& a, const BoysTables& tab, int nitems,
                               int mode, hipStream_t st) {
    constexpr int PA = 2 * LA + 1, PB = 2 * LB + 1, PC = 2 * LC + 1, PD = 2 * LD + 1;
    constexpr int CA = (LA + 1) * (LA + 2) / 2, CB = (LB + 1) * (LB + 2) / 2, CC = (LC + 1) * (LC + 2) / 2,
                  CD = (LD + 1) * (LD + 2) / 2;
    // accumulators: [e0|f0] with e over l = LA..LA+LB and f over l = LC..LC+LD, 64 per pass
    constexpr auto ncart_sum = [](int l0, int l1) {
        int n = 0;
        for (int l = l0; l <= l1; ++l) n += (l + 1) * (l + 2) / 2;
        return n;
    };
    constexpr int NAP = (ncart_sum(LA, LA + LB) * ncart_sum(LC, LC + LD) + 63) / 64;
    if (a.nden > 1) {
        if (pure) launch_heavy_dims<PA, PB, PC, PD, NAP, 2>(P, a, tab, nitems, mode, st);
        else launch_heavy_dims<CA, CB, CC, CD, NAP, 2>(P, a, tab, nitems, mode, st);
    } else {
        if (pure) launch_heavy_dims<PA, PB, PC, PD, NAP, 1>(P, a, tab, nitems, mode, st);
        else launch_heavy_dims<CA, CB, CC, CD, NAP, 1>(P, a, tab, nitems, mode, st);
    }
}

hipError_t launch_jk_heavy(int device, int la, int lb, int lc, int ld, int pure, const JKArgs& a,
                           const BoysTables& tab, int nitems, int mode, hipStream_t st) {
    if (nitems <= 0) return hipSuccess;
    HeavyProg P;
    if (!find_prog(device, la, lb, lc, ld, pure, P)) return hipErrorInvalidValue;
    const int key = ((la * 3 + lb) * 3 + lc) * 3 + ld;
    switch (key) {
        case ((2 * 3 + 1) * 3 + 2) * 3 + 1: launch_heavy_class<2, 1, 2, 1>(pure, P, a, tab, nitems, mode, st); break;
        case ((2 * 3 + 2) * 3 + 1) * 3 + 1: launch_heavy_class<2, 2, 1, 1>(pure, P, a, tab, nitems, mode, st); break;
        case ((2 * 3 + 2) * 3 + 2) * 3 + 0: launch_heavy_class<2, 2, 2, 0>(pure, P, a, tab, nitems, mode, st); break;
        case ((2 * 3 + 2) * 3 + 2) * 3 + 1: launch_heavy_class<2, 2, 2, 1>(pure, P, a, tab, nitems, mode, st); break;
        case ((2 * 3 + 2) * 3 + 2) * 3 + 2: launch_heavy_class<2, 2, 2, 2>(pure, P, a, tab, nitems, mode, st); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_schwarz_heavy(int device, int la, int lb, int pure, const JKArgs& a, const BoysTables& tab, int p0,
                                int np, double* Q, hipStream_t st) {
    if (np <= 0) return hipSuccess;
    HeavyProg P;
    if (!find_prog(device, la, lb, la, lb, pure, P)) return hipErrorInvalidValue;
    const size_t sm = sizeof(double) * (P.nslots + NSCALAR_LDS);
    hipLaunchKernelGGL(k_schwarz_heavy, dim3(np), dim3(64), sm, st, a, tab, p0, np, P, Q);
    return hipGetLastError();
}

}  // namespace m2
