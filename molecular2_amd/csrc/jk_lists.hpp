// Device work lists of the direct J/K class combinations (tasks.hip).
#pragma once
#include "m2_common.hpp"

namespace m2 {

struct JKList {
    int bra0, nbra;           // bra pairs [bra0, bra0 + nbra) of class X
    int same;                 // X == Y: only kets j <= i
    int ngrp;                 // primitive-count groups of the ket class Y
    int gbeg[JK_MAXG], gend[JK_MAXG];
    int2* cut;                // [nbra][JK_MAXG] (mid, end) of every group
    int* off;                 // [2][nbra + 1] workgroups per bra -> exclusive offsets (FP64 head, FP32 tail)
    int* bra_of;              // [2][cap] workgroup -> bra (local index)
    int cap;                  // capacity of each half of bra_of
    int kc;                   // kets per chunk: 64 (register classes) or JK_HEAVY_KETS (heavy classes)
    int bt;                   // ket chunks per work item
    int bt_min;               // the smallest bt the capacity holds
    int bound;                // most work items any density gives at this bt (per precision)
    const double* Qk;         // Schwarz bounds the ket ranges are sorted by (pairs or family pairs)
    const int* bra_fam;       // same class with family kets: bra pair -> its family pair (else null)
    int fp32;                 // mixed builds may run this combination's small quartets in FP32
};

// one class combination of the gradient's task list (tasks.hip)
struct GradTaskArgs {
    const double* Q;          // device Schwarz bounds of the pairs
    int x0, nbra;             // bra pairs [x0, x0 + nbra) of class X
    int same;                 // X == Y: only kets j <= i
    int ngrp;
    int gbeg[JK_MAXG], gend[JK_MAXG];
    int chunk;                // kets per task
    double precision, dmax;
    int part;                 // 0: kets with Q_i Q_j max|D| >= split (FP64), 1: the rest (FP32)
    double split;             // = precision when nothing runs in FP32
    const double* dmax_dev;   // device max|D| (read instead of dmax when set)
};
void launch_grad_task_count(const GradTaskArgs& g, int* counts, hipStream_t st);
void launch_scan_excl(const int* in, int* out, int n, hipStream_t st);
void launch_grad_task_fill(const GradTaskArgs& g, const int* offs, Task* out, hipStream_t st);

void launch_dmax(const double* Dn, int m, double* out, hipStream_t st);
// the lists of all `ncombo` combinations (device array `lists`; max_nbra: the
// largest nbra); totals to tot[2 * combo + mode]
void launch_task_lists(const JKList* lists, int ncombo, int max_nbra, const double* Q, const double* dmax,
                       double precision, double fp32_below, int mixed, int* tot, hipStream_t st);

}  // namespace m2
