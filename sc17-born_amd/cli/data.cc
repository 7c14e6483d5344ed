// Data (util/Data.cc): a one-shot data cube (time x receiver x x receiver
// y x shot x x shot y) with unit spikes every 50 samples from sample 30 on
// every 20th receiver from 20 on in both directions, plus one at the centre
// sample of the centre receiver.  The shot sits at the centre of the
// receiver grid.
//   Data out=d.json [nt=400 dt=.008 nx=300 dx=.01 ox=.5 ny=300 dy=.01 oy=.5]
#include <cstdio>
#include <exception>

#include "born_io.h"

int main(int argc, char **argv) {
  try {
    born_io::Params par(argc, argv);
    const int nt = par.i("nt", 400), nx = par.i("nx", 300), ny = par.i("ny", 300);
    const float dt = par.f("dt", .008f), dx = par.f("dx", .01f), dy = par.f("dy", .01f);
    const float ox = par.f("ox", .5f), oy = par.f("oy", .5f);
    born_io::Cube c;
    auto ax = [](int n, float o, float d) {
      born_io::Axis a;
      a.n = n;
      a.o = o;
      a.d = d;
      return a;
    };
    c.axes = {ax(nt, 0.f, dt), ax(nx, ox, dx), ax(ny, oy, dy), ax(1, ox + dx * nx / 2, dx),
              ax(1, oy + dy * ny / 2, dy)};
    c.vals.assign((size_t)c.n123(), 0.f);
    for (int i3 = 20; i3 < ny; i3 += 20)
      for (int i2 = 20; i2 < nx; i2 += 20)
        for (int i1 = 30; i1 < nt; i1 += 50) c.vals[i1 + (size_t)i2 * nt + (size_t)i3 * nx * nt] = 1.f;
    c.vals[nt / 2 + (size_t)(nx / 2) * nt + (size_t)(ny / 2) * nx * nt] = 1.f;
    born_io::write_cube(par.str("out"), c);
  } catch (const std::exception &e) {
    std::fprintf(stderr, "Data: %s\n", e.what());
    return 1;
  }
  return 0;
}
