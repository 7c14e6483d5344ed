// Wavelet (util/Wavelet.cc): a Ricker wavelet of peak frequency fund,
// centred at 0.064 s, nt samples at dt.
//   Wavelet out=w.json [nt=128 dt=.004 fund=10]
// The expression keeps the reference's mix of float and double arithmetic.
#include <cmath>
#include <cstdio>
#include <exception>

#include "born_io.h"

int main(int argc, char **argv) {
  try {
    born_io::Params par(argc, argv);
    const int nt = par.i("nt", 128);
    const float dt = par.f("dt", .004f);
    const float fund = par.f("fund", 10.f);
    born_io::Cube c;
    born_io::Axis a;
    a.n = nt;
    a.o = 0.f;
    a.d = dt;
    c.axes.push_back(a);
    c.vals.assign((size_t)nt, 0.f);
    const float pi = std::atan(1.f) * 4;
    for (int it = 0; it < nt; it++) {
      const float tm = -.064 + it * dt;
      const float arg = -pi * pi * fund * fund * tm * tm;
      c.vals[it] = (float)((1. - 2. * (pi * pi) * fund * fund * tm * tm) * expf(arg));
    }
    born_io::write_cube(par.str("out"), c);
  } catch (const std::exception &e) {
    std::fprintf(stderr, "Wavelet: %s\n", e.what());
    return 1;
  }
  return 0;
}
