// Density-fitting integral kernels: 3-centre (P|xy) and 2-centre (P|Q)
// (the reference's IntegralEngine::compute_eris_3index / compute_eris_2index,
// src/eris.cpp:89-156,304-356), from the same generated OS/HGP class code as
// the 4-centre kernels with a unit s function (exponent 0) in the second
// position, as libint2's BraKet::xs_xx / xs_xs do.
#pragma once
#include "m2_common.hpp"
#include "boys.hpp"

namespace m2 {

template <int LA, int LB, int LC, int LD>
struct EriClass;

template <int L, bool PURE>
struct NFunDF {
    static constexpr int value = PURE ? 2 * L + 1 : (L + 1) * (L + 2) / 2;
};

struct DF3Args {
    const PrimPair* aux_prims;   // per aux shell: its primitives paired with the unit s
    const int* aux_off;
    const int* aux_np;
    const int* aux_bf;
    const PrimPair* prims;       // orbital shell pairs
    const int* pair_off;
    const int* pair_np;
    const int2* pair_sh;
    const double* sh_O;
    const int* sh_bf;
    double* M;                   // [K][x][y], naux x n x n
    int n, naux;
};

struct DF2Args {
    const PrimPair* aux_prims;
    const int* aux_off;
    const int* aux_np;
    const int* aux_bf;
    const int* P_list;           // aux shells of l = LP
    int nP;
    const int* Q_list;           // aux shells of l = LQ
    int nQ;
    double* V;                   // naux x naux
    int naux;
};

// Task.bra = aux shell, ket0..ket0+nket = orbital pairs of one class.
template <int LP, int LC, int LD, bool PURE>
__global__ __launch_bounds__(256) void k_df3(DF3Args a, BoysTables tab, const Task* __restrict__ tasks, int ntasks) {
    constexpr int NP = NFunDF<LP, PURE>::value, NC = NFunDF<LC, PURE>::value, ND = NFunDF<LD, PURE>::value;
    constexpr int NOUT = NP * NC * ND > EriClass<LP, 0, LC, LD>::NCART ? NP * NC * ND : EriClass<LP, 0, LC, LD>::NCART;
    const int w = __builtin_amdgcn_readfirstlane(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6));
    if (w >= ntasks) return;
    const int lane = __lane_id();
    const int P = __builtin_amdgcn_readfirstlane(tasks[w].bra);
    const int ket0 = __builtin_amdgcn_readfirstlane(tasks[w].ket0);
    const int nket = __builtin_amdgcn_readfirstlane(tasks[w].nket);
    if (lane >= nket) return;
    const int pr = ket0 + lane;
    const int2 s = a.pair_sh[pr];
    const double* O = a.sh_O;
    const double CDx = O[3 * s.x] - O[3 * s.y], CDy = O[3 * s.x + 1] - O[3 * s.y + 1], CDz = O[3 * s.x + 2] - O[3 * s.y + 2];
    double v[NOUT];
    EriClass<LP, 0, LC, LD>::template run<double, PURE>(a.aux_prims + a.aux_off[P], a.aux_np[P],
                                                        a.prims + a.pair_off[pr], a.pair_np[pr],
                                                        0.0, 0.0, 0.0, CDx, CDy, CDz, tab, v);
    const size_t n = a.n, n2 = n * n;
    const int kf = a.aux_bf[P], xf = a.sh_bf[s.x], yf = a.sh_bf[s.y];
#pragma unroll
    for (int p = 0; p < NP; ++p)
#pragma unroll
        for (int c = 0; c < NC; ++c)
#pragma unroll
            for (int d = 0; d < ND; ++d) {
                const double x = v[(p * NC + c) * ND + d];
                double* blk = a.M + (size_t)(kf + p) * n2;
                blk[(size_t)(xf + c) * n + yf + d] = x;
                blk[(size_t)(yf + d) * n + xf + c] = x;
            }
}

template <int LP, int LC, int LD, bool PURE>
void launch_df3(const DF3Args& a, const BoysTables& tab, const Task* tasks, int ntasks, hipStream_t st) {
    if (ntasks <= 0) return;
    hipLaunchKernelGGL((k_df3<LP, LC, LD, PURE>), dim3((ntasks + 3) / 4), dim3(256), 0, st, a, tab, tasks, ntasks);
}

// one lane per (P, Q) shell pair of the two l lists
template <int LP, int LQ, bool PURE>
__global__ __launch_bounds__(64) void k_df2(DF2Args a, BoysTables tab) {
    constexpr int NP = NFunDF<LP, PURE>::value, NQ = NFunDF<LQ, PURE>::value;
    constexpr int NOUT = NP * NQ > EriClass<LP, 0, LQ, 0>::NCART ? NP * NQ : EriClass<LP, 0, LQ, 0>::NCART;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= a.nP * a.nQ) return;
    const int P = a.P_list[t / a.nQ], Q = a.Q_list[t % a.nQ];
    double v[NOUT];
    EriClass<LP, 0, LQ, 0>::template run<double, PURE>(a.aux_prims + a.aux_off[P], a.aux_np[P],
                                                       a.aux_prims + a.aux_off[Q], a.aux_np[Q],
                                                       0.0, 0.0, 0.0, 0.0, 0.0, 0.0, tab, v);
    const int pf = a.aux_bf[P], qf = a.aux_bf[Q];
#pragma unroll
    for (int p = 0; p < NP; ++p)
#pragma unroll
        for (int q = 0; q < NQ; ++q) {
            a.V[(size_t)(pf + p) * a.naux + qf + q] = v[p * NQ + q];
            a.V[(size_t)(qf + q) * a.naux + pf + p] = v[p * NQ + q];
        }
}

template <int LP, int LQ, bool PURE>
void launch_df2_t(const DF2Args& a, const BoysTables& tab, hipStream_t st) {
    const int nt = a.nP * a.nQ;
    if (nt <= 0) return;
    hipLaunchKernelGGL((k_df2<LP, LQ, PURE>), dim3((nt + 63) / 64), dim3(64), 0, st, a, tab);
}

}  // namespace m2
