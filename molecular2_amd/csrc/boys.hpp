// Device Boys function F_m(T), m = 0..M, for the ERI kernels.
//
// FP64: 8-term Taylor expansion about the nearest grid point (|dT| <= 0.05,
// truncation < 1e-15 relative), using F_m^(k) = (-1)^k F_{m+k}.  For
// M <= BOYS_DIRECT_MAX every F_m comes from its own expansion; the values
// F_0 .. F_(M+7) at the grid point are generated from a 16-byte table row,
// F_(M+7)(T0) and e^-T0, by the downward recursion F_m = (2 T0 F_(m+1) +
// e^-T0) / (2m+1) (stable: every step adds positive terms), so a lookup is
// one 16-byte load.  Above BOYS_DIRECT_MAX F_M is expanded from a full table
// row and the rest follow by the same recursion at T with e^-T.  For
// T >= BOYS_TMAX (60) the asymptotic F_0 = sqrt(pi/T)/2 (erfc(sqrt(60)) <
// 1e-27) and F_{m+1} = (2m+1) F_m / (2T) (the e^-T term is below 1e-26).
// FP32: 5-term Taylor expansion of F_M (truncation < 3e-9 relative) from an
// 8-byte row {F_(M+4)(T0), e^-T0} and the same recursion at T0, the lower F_m
// by the downward recursion at T; the large-T branch is evaluated in FP64.
#pragma once
#include "m2_common.hpp"

namespace m2 {

struct BoysTables {
    const double* d;   // [BOYS_NGRID][BOYS_MTAB]
    const float* f;
    const double* w;   // the compact window tables of BoysEval<double, M <= 3>
    const float* fw;   // the compact window tables of BoysEval<float, M <= BOYS_FW_MAX>
};

// Compact tables of the low-m classes: for M = 0..3, rows k = 0..BOYS_WK[M]
// of {F_(M+7)(T_k), e^-T_k} at T_k = k BOYS_STEP (two doubles a row), back
// to back from BoysTables::w.  Beyond BOYS_WK[M] BOYS_STEP the asymptotic
// form F_m = (2m-1)!! sqrt(pi) / (2^(m+1) T^(m+1/2)) is already within 1e-15
// relative (the neglected part is ~e^-T T^(m-1/2) / Gamma(m+1/2) of it), so
// the tables stop there: 5-7 KB each, which stay in the vector L1 cache of
// the kernels that read them.
constexpr int BOYS_WK[4] = {330, 380, 400, 440};
constexpr int boys_woff(int M) { return M == 0 ? 0 : boys_woff(M - 1) + (BOYS_WK[M - 1] + 1) * 2; }
constexpr int BOYS_WSIZE = boys_woff(4);
// FP32 windows {F_(M+4)(T_k), e^-T_k} (two floats a row, all BOYS_NGRID rows)
// for M = 0..BOYS_FW_MAX, back to back from BoysTables::fw: 4.8 KB each
constexpr int BOYS_FW_MAX = 8;
constexpr int BOYS_FWSIZE = (BOYS_FW_MAX + 1) * BOYS_NGRID * 2;

// 1/sqrt(x) for finite x > 0: the hardware estimate (v_rsq_f64, ~2^-23
// relative) refined by one third-order step of the iteration for x^(-1/2),
// y <- y (1 + e/2 + 3e^2/8) with e = 1 - x y^2, which leaves ~1 ulp.  The
// arguments here (p + q, T >= 60) are positive and normal, so no special
// cases are handled.
__device__ __forceinline__ double rsqrt_pos(double x) {
    const double y = __builtin_amdgcn_rsq(x);
    const double e = fma(-x * y, y, 1.0);
    return fma(y * e, fma(e, 0.375, 0.5), y);
}

template <typename R>
__device__ __forceinline__ R sqrt_inv(R x);
template <>
__device__ __forceinline__ double sqrt_inv<double>(double x) { return rsqrt_pos(x); }
template <>
__device__ __forceinline__ float sqrt_inv<float>(float x) { return rsqrtf(x); }

// Byte offset k * RB of row k (k < 2^24): one v_mul_u32_u24 whose result
// the compiler cannot fold into 64-bit address arithmetic, so that the row
// loads take the table base from SGPRs and this 32-bit offset from a VGPR.
template <unsigned RB>
__device__ __forceinline__ unsigned row_offset(unsigned k) {
    unsigned off;
    asm("v_mul_u32_u24 %0, %1, %2" : "=v"(off) : "i"(RB), "v"(k));
    return off;
}
template <typename T>
__device__ __forceinline__ const T* at_offset(const T* base, unsigned off) {
    return reinterpret_cast<const T*>(reinterpret_cast<const char*>(base) + off);
}

template <typename R, int M>
struct BoysEval;

// highest M for which every F_m is expanded directly (FP64)
constexpr int BOYS_DIRECT_MAX = 3;

template <int M>
struct BoysEval<double, M> {
    // the table branch covers T < TTAB; the asymptotic form the rest
    static constexpr double TTAB = M <= BOYS_DIRECT_MAX ? BOYS_WK[M <= BOYS_DIRECT_MAX ? M : 0] * BOYS_STEP : BOYS_TMAX;
    static constexpr int KMAX = M <= BOYS_DIRECT_MAX ? BOYS_WK[M <= BOYS_DIRECT_MAX ? M : 0] : BOYS_NGRID - 1;

    static __device__ __forceinline__ void run(double T, double* F, const BoysTables& tab) {
        if (T < TTAB) table(T, F, tab);
        else asym(T, F);
    }

    // T >= TTAB: F_0 = sqrt(pi/T)/2, F_(m+1) = (2m+1) F_m / (2T)
    static __device__ __forceinline__ void asym(double T, double* F) {
        const double rs = rsqrt_pos(T);
        F[0] = 0.88622692545275801365 * rs;   // sqrt(pi)/2 / sqrt(T)
        if constexpr (M > 0) {
            const double h = 0.5 * rs * rs;
#pragma unroll
            for (int m = 0; m < M; ++m) F[m + 1] = (2 * m + 1) * F[m] * h;
        }
    }

    // The table branch in two phases: load() gathers the row (and the offset
    // d from the grid point), eval() sums the Taylor series.
    static constexpr bool DIRECT = M <= BOYS_DIRECT_MAX;
    static constexpr int NROW = DIRECT ? 2 : 8;
    struct Row {
        double r[NROW];   // {F_(M+7)(T0), e^-T0} (DIRECT) or F_M .. F_(M+7) at T0
        double d, T, t2;  // T0 - T, T, 2 T0
    };
    // 0 <= T < TTAB.  The nearest grid index k = rint(T / BOYS_STEP) <= KMAX
    // comes from the low word of T / BOYS_STEP + 1.5 * 2^52 (whose last
    // mantissa bit is worth 1, so the addition rounds to an integer), and the
    // row is addressed by a 32-bit byte offset from the table base.
    static __device__ __forceinline__ void load(double T, Row& w, const BoysTables& tab) {
        constexpr double RND = 6755399441055744.0;   // 1.5 * 2^52
        const double x = fma(T, 1.0 / BOYS_STEP, RND);
        const unsigned k = unsigned(__double2loint(x));
        const double kd = x - RND;
        w.d = fma(kd, BOYS_STEP, -T);
        w.T = T;
        w.t2 = kd * (2.0 * BOYS_STEP);
        if constexpr (DIRECT) {
            const double2 v = *at_offset(reinterpret_cast<const double2*>(tab.w + boys_woff(M)), row_offset<16>(k));
            w.r[0] = v.x;
            w.r[1] = v.y;
        } else {
            const double* r = at_offset(tab.d + M, row_offset<BOYS_MTAB * 8>(k));
#pragma unroll
            for (int i = 0; i < NROW; ++i) w.r[i] = r[i];
        }
    }
    static __device__ __forceinline__ void eval(const Row& w, double* F) {
        const double d = w.d;
        if constexpr (DIRECT) {
            // F_m(T0), m = M+7 .. 0, by the downward recursion
            double r[M + 8];
            r[M + 7] = w.r[0];
#pragma unroll
            for (int m = M + 6; m >= 0; --m) r[m] = fma(w.t2, r[m + 1], w.r[1]) * (1.0 / (2 * m + 1));
            const double d2 = d * 0.5, d3 = d * (1.0 / 3.0), d4 = d * 0.25, d5 = d * 0.2, d6 = d * (1.0 / 6.0),
                         d7 = d * (1.0 / 7.0);
#pragma unroll
            for (int m = 0; m <= M; ++m) {
                double s = r[m + 6] + d7 * r[m + 7];
                s = r[m + 5] + d6 * s;
                s = r[m + 4] + d5 * s;
                s = r[m + 3] + d4 * s;
                s = r[m + 2] + d3 * s;
                s = r[m + 1] + d2 * s;
                F[m] = r[m] + d * s;
            }
        } else {
            const double* r = w.r;
            double s = r[7];
            s = r[6] + d * (1.0 / 7.0) * s;
            s = r[5] + d * (1.0 / 6.0) * s;
            s = r[4] + d * (1.0 / 5.0) * s;
            s = r[3] + d * (1.0 / 4.0) * s;
            s = r[2] + d * (1.0 / 3.0) * s;
            s = r[1] + d * (1.0 / 2.0) * s;
            s = r[0] + d * s;
            F[M] = s;
            if constexpr (M > 0) {
                const double e = exp(-w.T);
                const double t2 = 2.0 * w.T;
#pragma unroll
                for (int m = M - 1; m >= 0; --m) F[m] = (t2 * F[m + 1] + e) * (1.0 / (2 * m + 1));
            }
        }
    }
    static __device__ __forceinline__ void table(double T, double* F, const BoysTables& tab) {
        Row w;
        load(T, w, tab);
        eval(w, F);
    }
};

template <int M>
struct BoysEval<float, M> {
    static constexpr float TTAB = float(BOYS_TMAX);

    static __device__ __forceinline__ void run(float T, float* F, const BoysTables& tab) {
        if (T < TTAB) table(T, F, tab);
        else asym(T, F);
    }

    // T >= BOYS_TMAX: the asymptotic branch in FP64
    static __device__ __forceinline__ void asym(float T, float* F) {
        double Fd[M + 1];
        BoysEval<double, M>::asym(double(T), Fd);
#pragma unroll
        for (int m = 0; m <= M; ++m) F[m] = float(Fd[m]);
    }

    static constexpr bool FW = M <= BOYS_FW_MAX;
    static constexpr int NROW = FW ? 2 : 5;
    struct Row {
        float r[NROW];   // {F_(M+4)(T0), e^-T0} (FW) or F_M .. F_(M+4) at T0
        float d, T, t2;  // T0 - T, T, 2 T0
    };
    // 0 <= T < BOYS_TMAX: F_M by a 5-term Taylor series about the grid point
    // k = rint(T / BOYS_STEP) <= BOYS_NGRID - 1 (taken from the mantissa of
    // T / BOYS_STEP + 1.5 * 2^23, as in the FP64 case), the rest by the
    // downward recursion
    static __device__ __forceinline__ void load(float T, Row& w, const BoysTables& tab) {
        constexpr float RND = 12582912.0f;   // 1.5 * 2^23
        const float x = fmaf(T, float(1.0 / BOYS_STEP), RND);
        const unsigned k = unsigned(__float_as_int(x) - __float_as_int(RND));
        const float kd = x - RND;
        w.d = fmaf(kd, float(BOYS_STEP), -T);
        w.T = T;
        w.t2 = kd * float(2.0 * BOYS_STEP);
        if constexpr (FW) {
            const float2 v = *at_offset(reinterpret_cast<const float2*>(tab.fw + M * BOYS_NGRID * 2), row_offset<8>(k));
            w.r[0] = v.x;
            w.r[1] = v.y;
        } else {
            const float* r = at_offset(tab.f + M, row_offset<BOYS_MTAB * 4>(k));
#pragma unroll
            for (int i = 0; i < 5; ++i) w.r[i] = r[i];
        }
    }
    static __device__ __forceinline__ void eval(const Row& w, float* F) {
        const float d = w.d;
        float r[5];
        if constexpr (FW) {
            // F_(M+3) .. F_M at T0 by the downward recursion from F_(M+4)(T0)
            r[4] = w.r[0];
#pragma unroll
            for (int j = 3; j >= 0; --j) r[j] = fmaf(w.t2, r[j + 1], w.r[1]) * (1.0f / (2 * (M + j) + 1));
        } else {
#pragma unroll
            for (int j = 0; j < 5; ++j) r[j] = w.r[j];
        }
        float s = r[4];
        s = r[3] + d * (1.0f / 4.0f) * s;
        s = r[2] + d * (1.0f / 3.0f) * s;
        s = r[1] + d * (1.0f / 2.0f) * s;
        s = r[0] + d * s;
        F[M] = s;
        if constexpr (M > 0) {
            const float e = __expf(-w.T);
            const float t2 = 2.0f * w.T;
#pragma unroll
            for (int m = M - 1; m >= 0; --m) F[m] = (t2 * F[m + 1] + e) * (1.0f / (2 * m + 1));
        }
    }
    static __device__ __forceinline__ void table(float T, float* F, const BoysTables& tab) {
        Row w;
        load(T, w, tab);
        eval(w, F);
    }
};

// Bytes of the compact rows of BoysEval<R, M> (0: it reads the full tables)
template <typename R, int M>
constexpr int boys_rows_bytes() {
    return sizeof(R) == 8 ? (M <= BOYS_DIRECT_MAX ? (BOYS_WK[M <= BOYS_DIRECT_MAX ? M : 0] + 1) * 16 : 0)
                          : (M <= BOYS_FW_MAX ? BOYS_NGRID * 8 : 0);
}

// The workgroup copies the compact rows of BoysEval<R, M> to LDS at dst
// (16-byte aligned, boys_rows_bytes<R, M>() bytes) and gets tables whose
// compact rows are read from there; the caller synchronises before use.
template <typename R, int M>
__device__ __forceinline__ BoysTables stage_boys_rows(const BoysTables& tab, double* dst) {
    static_assert(boys_rows_bytes<R, M>() > 0, "no compact rows");
    BoysTables t = tab;
    if constexpr (sizeof(R) == 8) {
        const double2* src = reinterpret_cast<const double2*>(tab.w + boys_woff(M));
        double2* d = reinterpret_cast<double2*>(dst);
        for (int i = threadIdx.x; i <= BOYS_WK[M]; i += blockDim.x) d[i] = src[i];
        t.w = dst - boys_woff(M);
    } else {
        const float2* src = reinterpret_cast<const float2*>(tab.fw + M * BOYS_NGRID * 2);
        float2* d = reinterpret_cast<float2*>(dst);
        for (int i = threadIdx.x; i < BOYS_NGRID; i += blockDim.x) d[i] = src[i];
        t.fw = reinterpret_cast<const float*>(dst) - M * BOYS_NGRID * 2;
    }
    return t;
}

}  // namespace m2
