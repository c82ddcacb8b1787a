// Generated by molecular2_amd/csrc/gen_eri.py -- do not edit.
#include "../jk_kernel.hpp"
namespace m2 {
#include "eri_2010.inc"
void launch_jk_3_1(int pure, const JKArgs& a, const BoysTables& tab, const Task* t, int nt, int mode, hipStream_t st) {
  if (pure) launch_jk<2, 0, 1, 0, true>(a, tab, t, nt, mode, st);
  else launch_jk<2, 0, 1, 0, false>(a, tab, t, nt, mode, st);
}
}  // namespace m2
