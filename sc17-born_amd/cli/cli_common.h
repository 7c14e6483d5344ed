// Shared set-up of the Born programs RTM3D and Model3D (rtm.cc:10-41,
// model.cc:10-51): the velocity model, the wavelet and the data geometry
// from hypercube files, the operator from libborn_host.so.
#pragma once

#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "born_host.h"
#include "born_io.h"

namespace born_cli {

struct Inputs {
  born_io::Cube vel, wavelet;
  std::vector<born_io::Axis> data_axes;  // t, receiver x, receiver y, shot x, shot y
  std::vector<float> src_x, src_y;
};

// Data axes: 1 time, 2 and 3 the receiver grid, 4 and 5 the shot grid
// (util/Data.cc:18-28); shots are numbered x fastest
// (source_func_3d.cc:129-138).
inline Inputs read_inputs(const born_io::Params &par) {
  Inputs in;
  in.vel = born_io::read_cube(par.str("velocity"));
  if (in.vel.axes.size() < 3) throw std::runtime_error("velocity must have 3 axes");
  in.wavelet = born_io::read_cube(par.str("wavelet"));
  in.data_axes = born_io::read_axes(par.str("data"));
  while (in.data_axes.size() < 5) in.data_axes.push_back(born_io::Axis());
  const born_io::Axis &sx = in.data_axes[3], &sy = in.data_axes[4];
  for (int i2 = 0; i2 < sy.n; i2++)
    for (int i1 = 0; i1 < sx.n; i1++) {
      in.src_x.push_back(sx.o + sx.d * i1);
      in.src_y.push_back(sy.o + sy.d * i2);
    }
  return in;
}

// Parameters as the reference reads them: src_depth (rtm.cc:19), rec_depth
// (data_rtm_3d.cc), nbound defaulting to bc_a = 50 (fd_prop_3d.cc:46-52).
// mode (0 the reference RTM, 1 the exact transpose), scatter, storage and
// device have no reference counterpart.
inline born_op *create_op(const born_io::Params &par, const Inputs &in) {
  born_op_desc d{};
  for (int a = 0; a < 3; a++) {
    d.vn[a] = in.vel.axes[a].n;
    d.vo[a] = in.vel.axes[a].o;
    d.vd[a] = in.vel.axes[a].d;
  }
  d.vel = in.vel.vals.data();
  d.nw = in.wavelet.axes[0].n;
  d.wdt = in.wavelet.axes[0].d;
  d.wavelet = in.wavelet.vals.data();
  d.nshots = (int)in.src_x.size();
  d.src_x = in.src_x.data();
  d.src_y = in.src_y.data();
  d.src_depth = par.f("src_depth", 0.f);
  const born_io::Axis &t = in.data_axes[0], &rx = in.data_axes[1], &ry = in.data_axes[2];
  d.nt = t.n;
  d.t0 = t.o;
  d.dt = t.d;
  d.rec_nx = rx.n;
  d.rec_ny = ry.n;
  d.rec_ox = rx.o;
  d.rec_oy = ry.o;
  d.rec_dx = rx.d;
  d.rec_dy = ry.d;
  d.rec_depth = par.f("rec_depth", 0.f);
  d.nbound = par.i("nbound", (int)par.f("bc_a", 50.f));
  d.mode = par.i("mode", 0);
  d.scatter = par.i("scatter", 0);
  d.device = par.i("device", 0);
  d.storage = par.str("storage", "f32") == "f16" ? 1 : 0;
  born_op *op = nullptr;
  if (born_op_create(&op, &d) != 0)
    throw std::runtime_error(std::string("born_op_create: ") + born_op_last_error(op));
  return op;
}

inline void check(born_op *op, int rc, const char *what) {
  if (rc != 0) throw std::runtime_error(std::string(what) + ": " + born_op_last_error(op));
}

}  // namespace born_cli
