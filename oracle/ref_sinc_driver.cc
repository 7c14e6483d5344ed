// Driver for the reference's base/sinc_bob.cc, compiled where that source
// lies (the directory is REF in oracle/Makefile).  TEST INFRASTRUCTURE: it exposes the
// reference's table generator (base/sinc_bob.cc:4-16) through a C symbol so
// the tests can pin the oracle's restatement bit for bit.
#include "sinc_bob.h"

extern "C" int ref_sinc_table(int jt, int nsinc, float *out) {
  sinc_bob s(jt, nsinc);
  if ((int)s.table.size() != jt) return -1;
  for (int it = 0; it < jt; it++)
    for (int k = 0; k < nsinc; k++) out[it * nsinc + k] = s.table[it][k];
  return 0;
}
