// Device-side work lists of the direct J/K class kernels.
//
// For one class combination (bra pair class X, ket pair class Y <= X) and
// every bra pair i of class X, the kets of class Y are kept in
// primitive-count groups, each sorted by Schwarz Q descending.  The
// quartets that can pass the screening of Fock::formJKdirect
// (src/fockbuild.cpp:98-115), max|D_blocks| Q_i Q_j >= precision, form a
// prefix of each group: Q_j >= precision / (Q_i max|D|).  Under mixed
// precision the prefix splits again at max(precision, fp32_below) / (Q_i
// max|D|) into an FP64 head and an FP32 tail (combinations with fp32 unset
// keep the whole prefix in FP64).  With X == Y only kets j <= i
// are unique quartets (src/fockbuild.cpp:88-109).  With s-s family kets the
// groups hold family pairs, sorted by the largest member Q; the triangle is
// cut at the bra's own family and finished per member by the kernel.
//
// Every class combination's list is built by the same three launches (the
// combination is a grid dimension):
//   k_task_count: per bra, the (mid, end) cut of every group (two binary
//                 searches) and the number of work items per precision;
//   k_task_scan:  exclusive prefix sums of the work-item counts and their
//                 totals;
//   k_task_fill:  work item -> bra map.
// A work item is a bra and up to bt of its ket chunks (64 kets per chunk for
// the register classes, JK_HEAVY_KETS for the table-driven heavy classes);
// the class kernels take their ket ranges straight from the cuts.
#include "jk_lists.hpp"

namespace m2 {

__device__ __forceinline__ int count_ge_dev(const double* __restrict__ Q, int b, int e, double thr) {
    int lo = b, hi = e;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (Q[mid] >= thr) lo = mid + 1;
        else hi = mid;
    }
    return lo - b;
}

// max over the shell-block norms of D (one workgroup)
__global__ __launch_bounds__(1024) void k_dmax(const double* __restrict__ Dn, int m, double* __restrict__ out) {
    __shared__ double s[16];
    double v = 0.0;
    for (int i = threadIdx.x; i < m; i += blockDim.x) v = fmax(v, Dn[i]);
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double r = 0.0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) r = fmax(r, s[w]);
        out[0] = r;
    }
}

// blockIdx.y: the class combination (its list in `lists`)
__global__ __launch_bounds__(256) void k_task_count(const JKList* __restrict__ lists, const double* __restrict__ Q,
                                                    const double* __restrict__ dmax, double precision,
                                                    double fp32_below, int mixed) {
    const JKList& c = lists[blockIdx.y];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= c.nbra) return;
    const int bra = c.bra0 + i;
    const double qi = Q[bra], Dmax = dmax[0];
    int ch64 = 0, ch32 = 0;
    const bool any = qi > 0.0 && Dmax > 0.0;
    const double thr = any ? precision / (qi * Dmax) : 0.0;
    const double thr64 = any ? (mixed && c.fp32 ? fmax(precision, fp32_below) / (qi * Dmax) : thr) : 0.0;
    for (int g = 0; g < c.ngrp; ++g) {
        const int gb = c.gbeg[g], ge = c.gend[g];
        int end = gb, mid = gb;
        if (any) {
            end = gb + count_ge_dev(c.Qk, gb, ge, thr);
            mid = gb + count_ge_dev(c.Qk, gb, ge, thr64);
            if (c.same) {   // unique pairs of pairs: kets up to the bra's own position
                const int lim = (c.bra_fam ? c.bra_fam[bra] : bra) + 1;
                end = min(end, lim);
                mid = min(mid, lim);
            }
            end = max(end, gb);
            mid = max(min(mid, end), gb);
        }
        c.cut[i * JK_MAXG + g] = make_int2(mid, end);
        ch64 += (mid - gb + c.kc - 1) / c.kc;
        ch32 += (end - mid + c.kc - 1) / c.kc;
    }
    c.off[i] = (ch64 + c.bt - 1) / c.bt;
    c.off[c.nbra + 1 + i] = (ch32 + c.bt - 1) / c.bt;
}

// exclusive scan in place of off[mode * (nbra + 1) + 0 .. nbra - 1]; the total
// goes to off[mode * (nbra + 1) + nbra] and tot[2 * combo + mode].
// blockIdx.x = mode, blockIdx.y = combo.  The list capacity holds every
// count the basis allows (nbra x the largest chunk count of a bra / bt_min),
// so the clamp to it never binds; it keeps the class kernels in bounds by
// construction.
__global__ __launch_bounds__(1024) void k_task_scan(const JKList* __restrict__ lists, int* __restrict__ tot) {
    const JKList& c = lists[blockIdx.y];
    if (c.nbra <= 0) return;
    __shared__ int s_w[16];
    __shared__ int s_carry;
    int* a = c.off + blockIdx.x * (c.nbra + 1);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    if (threadIdx.x == 0) s_carry = 0;
    __syncthreads();
    for (int base = 0; base < c.nbra; base += blockDim.x) {
        const int i = base + threadIdx.x;
        const int v = i < c.nbra ? a[i] : 0;
        int x = v;   // inclusive scan within the wave
        for (int off = 1; off < 64; off <<= 1) {
            const int y = __shfl_up(x, off, 64);
            if (lane >= off) x += y;
        }
        if (lane == 63) s_w[w] = x;
        __syncthreads();
        if (threadIdx.x == 0) {
            int run = 0;
            for (int k = 0; k < nw; ++k) {
                const int t = s_w[k];
                s_w[k] = run;
                run += t;
            }
        }
        __syncthreads();
        const int carry = s_carry;
        if (i < c.nbra) a[i] = carry + s_w[w] + x - v;
        __syncthreads();
        if (threadIdx.x == blockDim.x - 1) s_carry = carry + s_w[w] + x;
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        a[c.nbra] = s_carry;
        tot[2 * blockIdx.y + blockIdx.x] = min(s_carry, c.cap);
    }
}

// work item -> bra (local index); blockIdx.y = mode, blockIdx.z = combo.
// Entries beyond the capacity are not written.
__global__ __launch_bounds__(256) void k_task_fill(const JKList* __restrict__ lists) {
    const JKList& c = lists[blockIdx.z];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= c.nbra) return;
    const int mode = blockIdx.y;
    const int* off = c.off + mode * (c.nbra + 1);
    int* dst = c.bra_of + (size_t)mode * c.cap;
    const int e = min(off[i + 1], c.cap);
    for (int b = off[i]; b < e; ++b) dst[b] = i;
}

// ---------------------------------------------------------------------------
// Gradient task lists, built on the device: for every bra pair i of class X
// (in pair order) and every ket group of class Y, the screened prefix Q_j >=
// precision / (Q_i max|D|) (cut at j <= i when X == Y) in chunks of `chunk`
// kets, one Task per chunk; under mixed precision the prefix splits at
// split / (Q_i max|D|) into an FP64 head (part 0) and an FP32 tail (part 1).
//   k_grad_task_count: tasks per bra;  k_scan_excl: exclusive prefix sums;
//   k_grad_task_fill: the Task entries at each bra's offset.
// the ket range [lo, hi) of group k that bra i takes in g.part; qd = Q_i max|D|
__device__ __forceinline__ int2 grad_group_range(const GradTaskArgs& g, int i, double qd, int k) {
    int end = g.gbeg[k] + count_ge_dev(g.Q, g.gbeg[k], g.gend[k], g.precision / qd);
    int mid = g.gbeg[k] + count_ge_dev(g.Q, g.gbeg[k], g.gend[k], g.split / qd);
    if (g.same) {
        end = min(end, i + 1);
        mid = min(mid, i + 1);
    }
    mid = min(mid, end);
    return g.part ? make_int2(mid, end) : make_int2(g.gbeg[k], mid);
}

__global__ __launch_bounds__(256) void k_grad_task_count(GradTaskArgs g, int* __restrict__ counts) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.nbra) return;
    const int i = g.x0 + t;
    const double qi = g.Q[i];
    const double dmax = g.dmax_dev ? *g.dmax_dev : g.dmax;
    int n = 0;
    if (qi > 0.0 && dmax > 0.0) {
        for (int k = 0; k < g.ngrp; ++k) {
            const int2 r = grad_group_range(g, i, qi * dmax, k);
            if (r.y > r.x) n += (r.y - r.x + g.chunk - 1) / g.chunk;
        }
    }
    counts[t] = n;
}

__global__ __launch_bounds__(256) void k_grad_task_fill(GradTaskArgs g, const int* __restrict__ offs,
                                                        Task* __restrict__ out) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= g.nbra) return;
    const int i = g.x0 + t;
    const double qi = g.Q[i];
    const double dmax = g.dmax_dev ? *g.dmax_dev : g.dmax;
    if (!(qi > 0.0 && dmax > 0.0)) return;
    int w = offs[t];
    for (int k = 0; k < g.ngrp; ++k) {
        const int2 r = grad_group_range(g, i, qi * dmax, k);
        for (int j = r.x; j < r.y; j += g.chunk) out[w++] = Task{i, j, min(g.chunk, r.y - j), 0};
    }
}

// out[0..n-1]: exclusive prefix sums of in[0..n-1]; out[n]: the total
__global__ __launch_bounds__(1024) void k_scan_excl(const int* __restrict__ in, int* __restrict__ out, int n) {
    __shared__ int part[1024];
    const int per = (n + 1023) / 1024;
    const int b = min(n, (int)threadIdx.x * per), e = min(n, b + per);
    int s = 0;
    for (int k = b; k < e; ++k) s += in[k];
    part[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        int acc = 0;
        for (int k = 0; k < 1024; ++k) {
            const int v = part[k];
            part[k] = acc;
            acc += v;
        }
        out[n] = acc;
    }
    __syncthreads();
    int acc = part[threadIdx.x];
    for (int k = b; k < e; ++k) {
        out[k] = acc;
        acc += in[k];
    }
}

void launch_grad_task_count(const GradTaskArgs& g, int* counts, hipStream_t st) {
    if (g.nbra > 0) hipLaunchKernelGGL(k_grad_task_count, dim3((g.nbra + 255) / 256), dim3(256), 0, st, g, counts);
}
void launch_scan_excl(const int* in, int* out, int n, hipStream_t st) {
    hipLaunchKernelGGL(k_scan_excl, dim3(1), dim3(1024), 0, st, in, out, n);
}
void launch_grad_task_fill(const GradTaskArgs& g, const int* offs, Task* out, hipStream_t st) {
    if (g.nbra > 0) hipLaunchKernelGGL(k_grad_task_fill, dim3((g.nbra + 255) / 256), dim3(256), 0, st, g, offs, out);
}

void launch_dmax(const double* Dn, int m, double* out, hipStream_t st) {
    hipLaunchKernelGGL(k_dmax, dim3(1), dim3(1024), 0, st, Dn, m, out);
}

void launch_task_lists(const JKList* lists, int ncombo, int max_nbra, const double* Q, const double* dmax,
                       double precision, double fp32_below, int mixed, int* tot, hipStream_t st) {
    if (ncombo <= 0 || max_nbra <= 0) return;
    const int nb = (max_nbra + 255) / 256;
    hipLaunchKernelGGL(k_task_count, dim3(nb, ncombo), dim3(256), 0, st, lists, Q, dmax, precision, fp32_below,
                       mixed);
    hipLaunchKernelGGL(k_task_scan, dim3(2, ncombo), dim3(1024), 0, st, lists, tot);
    hipLaunchKernelGGL(k_task_fill, dim3(nb, 2, ncombo), dim3(256), 0, st, lists);
}

}  // namespace m2
