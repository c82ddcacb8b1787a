// Register-code derivative adjoint of the RHF gradient: GradArgs, the
// Cartesian density blocks, the screening of the reference's gradient loop
// and k_grad_reg, instantiated per class in the generated grad_reg_*.hip
// translation units (one per class, compiled in parallel).
#pragma once
#include "m2_common.hpp"
#include "boys.hpp"

namespace m2 {

template <int LA, int LB, int LC, int LD, int MASK>
struct EriDeriv;

constexpr int NSCALAR_D = 27;

struct GradArgs {
    const PrimPair* prims;
    const int* pair_off;
    const int* pair_np;
    const int2* pair_sh;
    const double* pair_Q;
    const unsigned char* pair_sig;   // overlap-significant pairs
    const double* sh_O;
    const int* sh_bf;
    const int* sh_atom;
    const double* Dn;
    const double* D;
    const double* sph2;              // 5 x 6 solid harmonics of l = 2
    int n, nsh, pure;
    double precision;
    double* grad;                    // natom x 3
    const double* Dc;                // D in the Cartesian basis, nc x nc
    const int* sh_cf;                // first Cartesian function of each shell
    int nc;
    // primitive screening (the J/K kernels' rule): bounds sqrt(max |(ab|ab)|)
    // of every primitive pair, aligned with prims and sorted descending within
    // a pair; a primitive quartet is skipped when its bound product times the
    // quartet's largest density-block norm is below prim_tau (0: none)
    const float* prim_q;
    double prim_tau;
};

__device__ __forceinline__ int ncart_d(int l) { return (l + 1) * (l + 2) / 2; }

// Cartesian density block Dt (N1 x N2) of two shells whose first Cartesian
// functions are c1 and c2
template <int L1, int L2>
__device__ __forceinline__ void cart_block(const GradArgs& a, int c1, int c2, double* Dt) {
    constexpr int N1 = (L1 + 1) * (L1 + 2) / 2, N2 = (L2 + 1) * (L2 + 2) / 2;
#pragma unroll
    for (int i = 0; i < N1; ++i)
#pragma unroll
        for (int j = 0; j < N2; ++j) Dt[i * N2 + j] = a.Dc[(size_t)(c1 + i) * a.nc + c2 + j];
}

// max of the six shell-block norms of D over a quartet
__device__ __forceinline__ double grad_dn(const GradArgs& a, int sa, int sb, int sc, int sd) {
    const int nsh = a.nsh;
    const double* Dn = a.Dn;
    return fmax(fmax(Dn[sa * nsh + sb], Dn[sc * nsh + sd]),
                fmax(fmax(Dn[sa * nsh + sc], Dn[sa * nsh + sd]), fmax(Dn[sb * nsh + sc], Dn[sb * nsh + sd])));
}

__device__ __forceinline__ bool grad_screen(const GradArgs& a, int bra, int ket, int sa, int sb, int sc, int sd) {
    if (!a.pair_sig[bra] || !a.pair_sig[ket]) return false;
    return grad_dn(a, sa, sb, sc, sd) * a.pair_Q[bra] * a.pair_Q[ket] >= a.precision;
}

__device__ __forceinline__ double wave_sum_g(double v) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// doubles of the per-block atom bins, rounded up to keep what follows them
// 16-byte aligned
__host__ __device__ constexpr int grad_bins_len(int natom) { return (natom * 3 + 1) & ~1; }

// register classes: one quartet per lane, per-block LDS bins for the atoms
// R: arithmetic of the integrals and their contraction (float for the
// quartets a mixed-precision gradient leaves below its FP32 threshold); the
// per-quartet sums are added up in double.
// FP64 (ss|ss) and (ps|ss) are held one wave per SIMD above the occupancy
// their register counts (100, 134) would otherwise give; elsewhere no minimum
template <int LA, int LB, int LC, int LD, int MASK, typename R>
struct GradOccupancy {
    static constexpr bool d = sizeof(R) == 8 && LB == 0 && LC == 0 && LD == 0 && MASK == 7;
    static constexpr int min_per_eu = d && LA == 0 ? 5 : d && LA == 1 ? 4 : 1;
};

template <int LA, int LB, int LC, int LD, int MASK, typename R>
__global__ __launch_bounds__(256, (GradOccupancy<LA, LB, LC, LD, MASK, R>::min_per_eu)) void k_grad_reg(
    GradArgs a, BoysTables tab, const Task* __restrict__ tasks, int ntasks, int natom) {
    constexpr int NA = (LA + 1) * (LA + 2) / 2, NB = (LB + 1) * (LB + 2) / 2;
    constexpr int NC = (LC + 1) * (LC + 2) / 2, ND = (LD + 1) * (LD + 2) / 2;
    extern __shared__ double bins[];   // natom * 3, then the compact Boys rows of the class (if any)
    constexpr int LP = LA + LB + LC + LD + 1;
    BoysTables tabx = tab;
    if constexpr (boys_rows_bytes<R, LP>() > 0) tabx = stage_boys_rows<R, LP>(tab, bins + grad_bins_len(natom));
    for (int i = threadIdx.x; i < natom * 3; i += blockDim.x) bins[i] = 0.0;
    __syncthreads();
    const int w = __builtin_amdgcn_readfirstlane(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6));
    if (w < ntasks) {
        const int lane = __lane_id();
        const int bra = __builtin_amdgcn_readfirstlane(tasks[w].bra);
        const int ket0 = __builtin_amdgcn_readfirstlane(tasks[w].ket0);
        const int nket = __builtin_amdgcn_readfirstlane(tasks[w].nket);
        const int ket = ket0 + (lane < nket ? lane : 0);
        const int2 sab = a.pair_sh[bra], scd = a.pair_sh[ket];
        const int sa = sab.x, sb = sab.y, sc = scd.x, sd = scd.y;
        const bool active = lane < nket && grad_screen(a, bra, ket, sa, sb, sc, sd);
        if (__any(active)) {
            R g[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
            if (active) {
                const int fa = a.sh_cf[sa], fb = a.sh_cf[sb], fc = a.sh_cf[sc], fd = a.sh_cf[sd];
                const double deg = (sa == sb ? 1.0 : 2.0) * (sc == sd ? 1.0 : 2.0) * (bra == ket ? 1.0 : 2.0);
                // Gamma (Cartesian) is formed after the primitive loop
                auto gam_fill = [&](R* Gam) {
                    double Dab[NA * NB], Dcd[NC * ND], Dac[NA * NC], Dbd[NB * ND], Dad[NA * ND], Dbc[NB * NC];
                    cart_block<LA, LB>(a, fa, fb, Dab);
                    cart_block<LC, LD>(a, fc, fd, Dcd);
                    cart_block<LA, LC>(a, fa, fc, Dac);
                    cart_block<LB, LD>(a, fb, fd, Dbd);
                    cart_block<LA, LD>(a, fa, fd, Dad);
                    cart_block<LB, LC>(a, fb, fc, Dbc);
#pragma unroll
                    for (int i = 0; i < NA; ++i)
#pragma unroll
                        for (int j = 0; j < NB; ++j)
#pragma unroll
                            for (int k = 0; k < NC; ++k)
#pragma unroll
                                for (int l = 0; l < ND; ++l)
                                    Gam[((i * NB + j) * NC + k) * ND + l] = R(
                                        deg * (0.5 * Dab[i * NB + j] * Dcd[k * ND + l] -
                                               0.125 * (Dac[i * NC + k] * Dbd[j * ND + l] + Dad[i * ND + l] * Dbc[j * NC + k])));
                };
                const double* O = a.sh_O;
                const float cut = float(a.prim_tau / grad_dn(a, sa, sb, sc, sd));
                EriDeriv<LA, LB, LC, LD, MASK>::template run<R>(
                    a.prims + a.pair_off[bra], a.pair_np[bra], a.prims + a.pair_off[ket], a.pair_np[ket],
                    R(O[3 * sa] - O[3 * sb]), R(O[3 * sa + 1] - O[3 * sb + 1]), R(O[3 * sa + 2] - O[3 * sb + 2]),
                    R(O[3 * sc] - O[3 * sd]), R(O[3 * sc + 1] - O[3 * sd + 1]), R(O[3 * sc + 2] - O[3 * sd + 2]), tabx,
                    gam_fill, g, a.prim_q + a.pair_off[bra], a.prim_q + a.pair_off[ket], cut);
            }
            const int at_a = a.sh_atom[sa], at_b = a.sh_atom[sb];
            // A and B are the same for the whole wavefront
#pragma unroll
            for (int i = 0; i < 3; ++i) {
                const double ga = wave_sum_g(double(g[i])), gb = wave_sum_g(double(g[3 + i]));
                if (lane == 0) {
                    atomicAdd(&bins[at_a * 3 + i], ga);
                    atomicAdd(&bins[at_b * 3 + i], gb);
                }
            }
            if (active) {
                const int at_c = a.sh_atom[sc], at_d = a.sh_atom[sd];
#pragma unroll
                for (int i = 0; i < 3; ++i) {
                    atomicAdd(&bins[at_c * 3 + i], double(g[6 + i]));
                    atomicAdd(&bins[at_d * 3 + i], -(double(g[i]) + double(g[3 + i]) + double(g[6 + i])));
                }
            }
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < natom * 3; i += blockDim.x)
        if (bins[i] != 0.0) atomicAdd(&a.grad[i], bins[i]);
}


template <int LA, int LB, int LC, int LD, int MASK, typename R>
void launch_reg(const GradArgs& a, const BoysTables& tab, const Task* t, int nt, int natom, hipStream_t st) {
    const size_t lds = sizeof(double) * grad_bins_len(natom) + boys_rows_bytes<R, LA + LB + LC + LD + 1>();
    hipLaunchKernelGGL((k_grad_reg<LA, LB, LC, LD, MASK, R>), dim3((nt + 3) / 4), dim3(256), lds, st, a, tab, t, nt,
                       natom);
}

}  // namespace m2
