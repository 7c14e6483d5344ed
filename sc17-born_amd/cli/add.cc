// Add (util/Add.cc): out = in1 + in2, element by element, with the axes of
// in1; the two must have the same number of samples on every axis.  out
// may be in2 (both inputs are read before out is written), as the
// reference's job script accumulates shot images (scripts/run.py:62-73).
//   Add in1=a.json in2=b.json out=c.json
#include <cstdio>
#include <exception>
#include <stdexcept>

#include "born_io.h"

int main(int argc, char **argv) {
  try {
    born_io::Params par(argc, argv);
    born_io::Cube a = born_io::read_cube(par.str("in1"));
    const born_io::Cube b = born_io::read_cube(par.str("in2"));
    bool same = a.axes.size() == b.axes.size();
    for (size_t k = 0; same && k < a.axes.size(); k++) same = a.axes[k].n == b.axes[k].n;
    if (!same) throw std::runtime_error("in1 and in2 differ in size");
    for (size_t i = 0; i < a.vals.size(); i++) a.vals[i] += b.vals[i];
    born_io::write_cube(par.str("out"), a);
  } catch (const std::exception &e) {
    std::fprintf(stderr, "Add: %s\n", e.what());
    return 1;
  }
  return 0;
}
