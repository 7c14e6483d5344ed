// Launchers of the reference-mode loops: sourceProp and rtmAdjoint steps.
#include "born_ops.h"

namespace born {

hipError_t launch_src_step(const Geometry &geo, const Tune &tu, const SrcStep &a,
                           hipStream_t st) {
  SrcOp op{};
  op.s1 = a.s1; op.s0 = a.s0; op.vel = a.vel;
  op.pts = to_points(a.src); op.amp = a.amp;
  return launch(geo, tu, op, st);
}

hipError_t launch_radj_step(const Geometry &geo, const Tune &tu, const RadjStep &a,
                            hipStream_t st) {
  RadjOp op{};
  op.s1 = a.s1; op.r1 = a.r1; op.s0 = a.s0; op.r0 = a.r0;
  op.vel1 = a.vel1; op.vel2 = a.vel2; op.img = a.img;
  op.s_prop = a.s_prop; op.r_prop = a.r_prop;
  op.pts = to_points(a.src); op.amp_pre = a.amp_pre; op.pre = a.pre;
  op.rm = a.rec; op.tr = a.tr; op.tab = a.tab; op.sc = a.sc;
  op.id = a.id; op.inject = a.inject; op.nx = geo.nx; op.ny = geo.ny;
  return launch(geo, tu, op, st);
}

}  // namespace born
