// Arguments of the one-electron integral kernel (misc_kernels.hip), shared
// with the host code that fills them (api.cpp).
#pragma once
#include "m2_common.hpp"
#include "boys.hpp"

namespace m2 {

struct OneBodyArgs {
    const double* sh_O;
    const int* sh_l;
    const int* sh_bf;
    const int* sh_np;
    const int* sh_poff;
    const double* exps;
    const double* coefs;
    int nsh, n, pure, natom;
    int want_v;              // 0: overlap and kinetic only
    const double* Z;
    const double* xyz;
    const double* sph;
    double *S, *T, *V;
    BoysTables tab;
};

}  // namespace m2
