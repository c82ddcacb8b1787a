// Generated by molecular2_amd/csrc/gen_eri.py -- do not edit.
#include "../jk_kernel.hpp"
namespace m2 {
#include "eri_2020.inc"
void launch_jk_3_3(int pure, const JKArgs& a, const BoysTables& tab, const Task* t, int nt, int mode, hipStream_t st) {
  if (pure) launch_jk<2, 0, 2, 0, true>(a, tab, t, nt, mode, st);
  else launch_jk<2, 0, 2, 0, false>(a, tab, t, nt, mode, st);
}
void launch_schwarz3(int pure, const JKArgs& a, const BoysTables& tab, int p0, int np, double* Q, hipStream_t st) {
  if (pure) launch_schwarz<2, 0, true>(a, tab, p0, np, Q, st); else launch_schwarz<2, 0, false>(a, tab, p0, np, Q, st);
}
}  // namespace m2
