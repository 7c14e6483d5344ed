// Vel (util/Vel.cc): a constant velocity model on an nx x ny x nz grid
// with origin 0.
//   Vel out=v.json [nx=400 dx=.01 ny=400 dy=.01 nz=400 dz=.01 vconst=2.0]
#include <cstdio>
#include <exception>

#include "born_io.h"

int main(int argc, char **argv) {
  try {
    born_io::Params par(argc, argv);
    born_io::Cube c;
    for (const char *x : {"x", "y", "z"}) {
      born_io::Axis a;
      a.n = par.i(std::string("n") + x, 400);
      a.d = par.f(std::string("d") + x, .01f);
      c.axes.push_back(a);
    }
    c.vals.assign((size_t)c.n123(), par.f("vconst", 2.f));
    born_io::write_cube(par.str("out"), c);
  } catch (const std::exception &e) {
    std::fprintf(stderr, "Vel: %s\n", e.what());
    return 1;
  }
  return 0;
}
