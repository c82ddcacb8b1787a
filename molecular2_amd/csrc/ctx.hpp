// Private: the context object behind the C ABI (shared by the engine's
// translation units).
#pragma once
#include "m2.h"
#include "m2_common.hpp"
#include "jk_lists.hpp"
#include "boys.hpp"
#include <hip/hip_runtime.h>
#include <algorithm>
#include <string>
#include <vector>

struct rocblas_handle_opaque;   // rocBLAS handle (typedef'd pointer in rocblas headers)

namespace m2 {

constexpr int NPAIRCLASS = 6;
constexpr int NCOMBO = NPAIRCLASS * (NPAIRCLASS + 1) / 2;
const int PAIR_L[NPAIRCLASS][2] = {{0, 0}, {1, 0}, {1, 1}, {2, 0}, {2, 1}, {2, 2}};
struct HostShell {
    int l = 0, atom = -1, bf = 0, nbf = 0;
    double O[3] = {0, 0, 0};
    std::vector<double> a, c;
};

// Device buffer owned by its holder: grown by reserve, freed by release or
// when the holder goes out of scope (so early error returns free it too).
template <typename T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    hipError_t reserve(size_t cnt) {
        if (cnt <= n && p) return hipSuccess;
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
        hipError_t e = hipMalloc((void**)&p, std::max<size_t>(cnt, 1) * sizeof(T));
        if (e == hipSuccess) n = cnt;
        return e;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
};

struct PairGroup {
    int begin, end;   // global pair index range, Q descending inside
};

struct PairClassSeg {
    int begin = 0, end = 0;
    std::vector<PairGroup> groups;
};

struct BasisData {
    int nsh = 0, nbf = 0, pure = 1;
    std::vector<HostShell> sh;
    std::vector<int> h_bf, h_nbf, h_l;
    std::vector<double> h_O;
    m2::DevBuf<double> d_O;
    m2::DevBuf<int> d_bf, d_nbf;
};


struct LocalDF;

// a captured direct J/K build: its key and executable graph
struct JKGraphKey {
    const void* dD;
    const void* dJ;
    const void* dK;
    const void* dScreen;
    const void* dD2;
    const void* dJ2;
    const void* dK2;
    double multiplier, precision, fp32_below, prim_screen, prim_screen32;
    const void* Dn;
    const void* J;
    const void* K;
    const void* lists;
    unsigned long long gen;
    unsigned flags;
};
struct JKGraph {
    JKGraphKey key;
    hipGraphExec_t exec;
};
constexpr size_t JK_GRAPH_CACHE = 8;

}  // namespace m2

struct m2_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    m2::BasisData obs, aux;
    // shell pairs of the orbital basis
    std::vector<int2> pair_sh;
    std::vector<int> pair_off, pair_np;
    std::vector<double> pair_Q;
    std::vector<double> Qsh;   // nsh x nsh
    m2::PairClassSeg cls[m2::NPAIRCLASS];
    m2::DevBuf<m2::PrimPair> d_prims;
    m2::DevBuf<m2::BraPair<float>> d_prims_f;
    m2::DevBuf<m2::KetS> d_kets_s;   // the pair records as s-s kets (with their bounds)
    m2::DevBuf<m2::PairRec> d_pair_rec;   // shell pairs as kets: shells, primitives, Q, C - D, first bfs
    m2::DevBuf<m2::KetSf> d_kets_sf;
    m2::DevBuf<m2::FamS> d_fam_s;    // the family primitive pairs with their factors and bounds
    m2::DevBuf<m2::FamSf> d_fam_sf;
    m2::DevBuf<int> d_pair_off, d_pair_np;
    m2::DevBuf<int2> d_pair_sh;
    m2::DevBuf<double> d_pair_Q;
    // primitive-pair bounds sqrt(max |(ab|ab)|), aligned with d_prims / d_fam_prims,
    // and the primitive screening threshold relative to the quartet precision
    m2::DevBuf<float> d_prim_q, d_fam_q;
    double prim_screen = 1e-6;
    // the same for the FP32 quartets of a mixed build (their arithmetic
    // carries ~6e-8 relative error, so skipped primitive quartets below
    // precision * prim_screen32 stay within it)
    double prim_screen32 = 0.1;
    // s-s ket family pairs (shells of one atom with identical exponents share
    // their primitive pairs), ordered like pairs: np bucket, Q descending
    int fam_mode = 1;                  // m2_set_jk_families
    bool fam_on = false;
    bool fam_wide = false;             // also for the (ps|ss) and (ds|ss) combinations
    m2::PairClassSeg fam_seg;          // groups over family-pair indices
    m2::DevBuf<m2::PrimPair> d_fam_prims;
    m2::DevBuf<m2::BraPair<float>> d_fam_prims_f;
    m2::DevBuf<double> d_fam_km, d_fam_Q;
    m2::DevBuf<float> d_fam_km_f;
    m2::DevBuf<int> d_fam_off, d_fam_np, d_pair_fam, d_pair_mem;
    std::vector<int> h_pair_fam;
    m2::DevBuf<int4> d_fam_mem;
    // Boys tables: full, FP32 copy, compact windows of the low-m classes
    m2::DevBuf<double> d_boys;
    m2::DevBuf<float> d_boys_f;
    m2::DevBuf<double> d_boys_w;
    m2::DevBuf<float> d_boys_fw;
    // J/K work buffers
    m2::DevBuf<double> d_D, d_J, d_K, d_Jo, d_Ko, d_Dn;
    m2::DevBuf<m2::Task> d_tasks;    // the gradient's task list
    // device work lists of the class combinations (index over X >= Y), their
    // device copy (read by the list kernels) and the largest bra count
    m2::JKList lists[m2::NCOMBO];
    m2::DevBuf<int2> list_cut[m2::NCOMBO];
    m2::DevBuf<int> list_off[m2::NCOMBO], list_bra[m2::NCOMBO];
    m2::DevBuf<m2::JKList> d_jklists;
    int max_nbra = 0;
    m2::DevBuf<int> d_lists;          // per combination work-item totals (FP64, FP32), then max|D|
    int* h_lists = nullptr;           // pinned copy of d_lists
    // launch shape (m2_set_jk_launch, m2_set_jk_block_tasks): grids sized
    // from the lists' bounds with no host round trip (default), or from the
    // totals read back; chunks per work item per combination (0: default)
    bool host_totals = false;
    int tune_bt[m2::NCOMBO] = {0};
    // totals of the last bounded build, read back only when asked for
    // (m2_last_stats)
    const int* last_tot = nullptr;
    // graph replay of the direct J/K build (m2_set_jk_graphs): the most
    // recently used graphs first; graph_gen changes with the basis and the
    // launch shape
    bool use_graphs = false;
    std::vector<m2::JKGraph> graphs;
    unsigned long long graph_gen = 0;
    long long graph_captures = 0, graph_replays = 0;
    std::vector<double> h_Dn;
    double fp32_below = 1e-4;
    long long last_tasks = 0;
    // profiling (m2_set_profiling): per class combo kernel time and counters
    bool prof = false;            // HIP events around each class launch
    bool prof_counters = false;   // also count computed quartets on the device (slows the kernels)
    std::vector<hipEvent_t> ev;
    double prof_ms[36] = {0};
    long long prof_launch[36] = {0};
    long long prof_q[36] = {0}, prof_pq[36] = {0}, prof_cpq[36] = {0};
    m2::DevBuf<unsigned long long> d_counters;
    // density fitting
    std::vector<m2::PrimPair> h_aux_prims;
    m2::DevBuf<m2::PrimPair> d_aux_prims;
    m2::DevBuf<int> d_aux_off, d_aux_np, d_aux_lists;
    m2::DevBuf<double> d_B64, d_dfV, d_dfC, d_dfX, d_dfG, d_dfJ, d_dfK, d_dfKp;
    m2::DevBuf<float> d_B32, d_dfC32, d_dfX32, d_dfKp32;   // Kp: split-K partial slices
    m2::DevBuf<int> d_df_jblk;        // (x, y0) per workgroup of the Coulomb contraction
    int df_jblocks = 0;
    bool df_ready = false;
    double df_last_ms[4] = {0, 0, 0, 0};   // X = C^T B, K = X^T X, J, total (profiling)
    // fragment-local DF-JK state (m2_jk_df_local), set up on first use
    m2::LocalDF* ldf = nullptr;
    // this device's copy of the static scratch array of the allocation-free
    // J/K entry point (hipGetSymbolAddress at context creation)
    unsigned char* scratch = nullptr;
    // dense linear algebra (rocBLAS / rocSOLVER)
    void* blas = nullptr;
    m2::DevBuf<double> la_a, la_b, la_c, la_w;
    m2::DevBuf<int> la_info;
};

namespace m2 {
inline BoysTables boys_tables(const m2_ctx* ctx) {
    return BoysTables{ctx->d_boys.p, ctx->d_boys_f.p, ctx->d_boys_w.p, ctx->d_boys_fw.p};
}
m2_status la_handle(m2_ctx* ctx, void** h);
void local_df_release(m2_ctx* ctx);   // fragment-local DF state (df_local.cpp)
}

#define HIPCHK(expr)                                                                   \
    do {                                                                               \
        hipError_t _e = (expr);                                                        \
        if (_e != hipSuccess) {                                                        \
            ctx->err = std::string(#expr) + ": " + hipGetErrorString(_e);              \
            return M2_ERR_HIP;                                                         \
        }                                                                              \
    } while (0)

#define ARGCHK(cond, msg)             \
    do {                              \
        if (!(cond)) {                \
            ctx->err = (msg);         \
            return M2_ERR_ARG;        \
        }                             \
    } while (0)
