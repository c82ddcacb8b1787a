// Direct J/K shell-quartet kernels (templates; instantiated per class in the
// generated translation units).
//
// Work unit: one Task = one wavefront = one bra shell pair (uniform across the
// wave, read into SGPRs) x up to 64 consecutive ket shell pairs of one class
// and one primitive-count group (so every lane runs the same primitive loop).
// Each lane applies the exact screening of Fock::formJKdirect
// (src/fockbuild.cpp:98-115): max of the six shell-block norms of D times
// Q_ab Q_cd against `precision`, computes its contracted quartet with the
// generated OS/HGP code, and digests it into unsymmetrised J and K with the
// degeneracy factors of src/fockbuild.cpp:121-149.  The bra-pair J block is
// summed across the wavefront before its atomics; the other five blocks are
// per-lane FP64 atomics (hardware global_atomic_add_f64).
#pragma once
#include "m2_common.hpp"
#include "boys.hpp"

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

template <int NA, int NB, int NC, int ND, typename R>
__device__ __forceinline__ void digest(const R* __restrict__ v, bool active, int fa, int fb, int fc, int fd,
                                       double f, const double* __restrict__ D, int n,
                                       double* __restrict__ J, double* __restrict__ K) {
    const double h = 0.5 * f;
    // J(ab) += f sum_cd D(cd) v : bra block shared by the whole wave
#pragma unroll
    for (int a = 0; a < NA; ++a) {
#pragma unroll
        for (int b = 0; b < NB; ++b) {
            double s = 0.0;
            if (active) {
#pragma unroll
                for (int c = 0; c < NC; ++c)
#pragma unroll
                    for (int d = 0; d < ND; ++d)
                        s += D[(fc + c) * n + fd + d] * double(v[((a * NB + b) * NC + c) * ND + d]);
                s *= f;
            }
            s = wave_sum(s);
            if (__lane_id() == 0 && s != 0.0) atomic_add_f64(&J[(fa + a) * n + fb + b], s);
        }
    }
    if (!active) return;
    // J(cd) += f sum_ab D(ab) v
#pragma unroll
    for (int c = 0; c < NC; ++c)
#pragma unroll
        for (int d = 0; d < ND; ++d) {
            double s = 0.0;
#pragma unroll
            for (int a = 0; a < NA; ++a)
#pragma unroll
                for (int b = 0; b < NB; ++b) s += D[(fa + a) * n + fb + b] * double(v[((a * NB + b) * NC + c) * ND + d]);
            atomic_add_f64(&J[(fc + c) * n + fd + d], f * s);
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
                for (int d = 0; d < ND; ++d) s += D[(fb + b) * n + fd + d] * double(v[((a * NB + b) * NC + c) * ND + d]);
            atomic_add_f64(&K[(fa + a) * n + fc + c], h * s);
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
                for (int c = 0; c < NC; ++c) s += D[(fa + a) * n + fc + c] * double(v[((a * NB + b) * NC + c) * ND + d]);
            atomic_add_f64(&K[(fb + b) * n + fd + d], h * s);
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
                for (int c = 0; c < NC; ++c) s += D[(fb + b) * n + fc + c] * double(v[((a * NB + b) * NC + c) * ND + d]);
            atomic_add_f64(&K[(fa + a) * n + fd + d], h * s);
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
                for (int d = 0; d < ND; ++d) s += D[(fa + a) * n + fd + d] * double(v[((a * NB + b) * NC + c) * ND + d]);
            atomic_add_f64(&K[(fb + b) * n + fc + c], h * s);
        }
}

// Quartets are computed when lo <= bound < hi (the launchers pass
// [precision, inf); precision is chosen per task list).
template <int LA, int LB, int LC, int LD, typename R, bool PURE>
__global__ __launch_bounds__(256) void k_jk(JKArgs a, BoysTables tab, const Task* __restrict__ tasks, int ntasks,
                                            double lo, double hi) {
    constexpr int NA = NFun<LA, PURE>::value, NB = NFun<LB, PURE>::value;
    constexpr int NC = NFun<LC, PURE>::value, ND = NFun<LD, PURE>::value;
    const int w = __builtin_amdgcn_readfirstlane(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6));
    if (w >= ntasks) return;
    const int lane = __lane_id();
    const int bra = __builtin_amdgcn_readfirstlane(tasks[w].bra);
    const int ket0 = __builtin_amdgcn_readfirstlane(tasks[w].ket0);
    const int nket = __builtin_amdgcn_readfirstlane(tasks[w].nket);
    const int ket = ket0 + (lane < nket ? lane : 0);
    const int2 sab = a.pair_sh[bra];
    const int2 scd = a.pair_sh[ket];
    const int sa = sab.x, sb = sab.y, sc = scd.x, sd = scd.y;
    const int nsh = a.nsh;
    // exact screening of src/fockbuild.cpp:98-115
    const double* Dn = a.Dn;
    double dn = fmax(fmax(Dn[sa * nsh + sb], Dn[sc * nsh + sd]),
                     fmax(fmax(Dn[sa * nsh + sc], Dn[sa * nsh + sd]), fmax(Dn[sb * nsh + sc], Dn[sb * nsh + sd])));
    const double bound = dn * a.pair_Q[bra] * a.pair_Q[ket];
    const bool active = lane < nket && bound >= lo && bound < hi;
    if (!__any(active)) return;
    if (a.counters) {
        const int na = __popcll(__ballot(active));
        const double nk = wave_sum(active ? double(a.pair_np[ket]) : 0.0);
        if (lane == 0) {
            atomicAdd(&a.counters[2 * a.combo], (unsigned long long)na);
            atomicAdd(&a.counters[2 * a.combo + 1], (unsigned long long)(nk * a.pair_np[bra] + 0.5));
        }
    }
    const double deg = (sa == sb ? 1.0 : 2.0) * (sc == sd ? 1.0 : 2.0) * (bra == ket ? 1.0 : 2.0);
    // the bra's primitive pairs are the same for the whole wavefront: stage
    // them in LDS once (converted to R), so the inner primitive loop reads
    // LDS broadcasts
    const int nbp = a.pair_np[bra];
    const PrimPair* gb = a.prims + a.pair_off[bra];
    extern __shared__ float4 s_raw[];
    BraPair<R>* bprims = reinterpret_cast<BraPair<R>*>(s_raw) + (threadIdx.x >> 6) * a.bra_lds;
    for (int i = lane; i < nbp; i += 64) {
        const PrimPair g = gb[i];
        BraPair<R> b;
        b.p = R(g.p);
        b.inv2p = R(g.inv2p);
        b.K = R(g.K);
        for (int k = 0; k < 3; ++k) {
            b.P[k] = R(g.P[k]);
            b.PA[k] = R(g.PA[k]);
        }
        bprims[i] = b;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const double* O = a.sh_O;
    const R ABx = R(O[3 * sa] - O[3 * sb]), ABy = R(O[3 * sa + 1] - O[3 * sb + 1]), ABz = R(O[3 * sa + 2] - O[3 * sb + 2]);
    const R CDx = R(O[3 * sc] - O[3 * sd]), CDy = R(O[3 * sc + 1] - O[3 * sd + 1]), CDz = R(O[3 * sc + 2] - O[3 * sd + 2]);
    R v[NA * NB * NC * ND > EriClass<LA, LB, LC, LD>::NCART ? NA * NB * NC * ND : EriClass<LA, LB, LC, LD>::NCART];
    const int nkp = active ? a.pair_np[ket] : 0;
    EriClass<LA, LB, LC, LD>::template run<R, PURE>(bprims, nbp, a.prims + a.pair_off[ket], nkp,
                                                    ABx, ABy, ABz, CDx, CDy, CDz, tab, v);
    digest<NA, NB, NC, ND, R>(v, active, a.sh_bf[sa], a.sh_bf[sb], a.sh_bf[sc], a.sh_bf[sd],
                              deg * a.scale, a.D, a.n, a.J, a.K);
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

// mode 0: FP64 kernel; mode 2: FP32 kernel (the caller hands each the tasks
// of its precision); both compute every quartet with bound >= precision.
template <int LA, int LB, int LC, int LD, bool PURE>
void launch_jk(const JKArgs& a, const BoysTables& tab, const Task* tasks, int ntasks, int mode, hipStream_t st) {
    if (ntasks <= 0) return;
    const int waves_per_block = 4;
    dim3 grid((ntasks + waves_per_block - 1) / waves_per_block), block(64 * waves_per_block);
    const double inf = __builtin_huge_val();
    if (mode == 2)
        hipLaunchKernelGGL((k_jk<LA, LB, LC, LD, float, PURE>), grid, block,
                           sizeof(BraPair<float>) * (size_t)a.bra_lds * waves_per_block, st, a, tab, tasks, ntasks,
                           a.precision, inf);
    else
        hipLaunchKernelGGL((k_jk<LA, LB, LC, LD, double, PURE>), grid, block,
                           sizeof(BraPair<double>) * (size_t)a.bra_lds * waves_per_block, st, a, tab, tasks, ntasks,
                           a.precision, inf);
}

template <int LA, int LB, bool PURE>
void launch_schwarz(const JKArgs& a, const BoysTables& tab, int p0, int np, double* Q, hipStream_t st) {
    if (np <= 0) return;
    hipLaunchKernelGGL((k_schwarz<LA, LB, PURE>), dim3((np + 255) / 256), dim3(256), 0, st, a, tab, p0, np, Q);
}

}  // namespace m2
