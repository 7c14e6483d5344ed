// Launchers of the two-step (two time steps per launch) kernels.
#include "born_ops.h"
#include "born_pair.h"
#include "born_pair_adj.h"

namespace born {

const PairShape kPairShapes[kNumPairShapes] = {{2, 8}, {1, 8}, {2, 4}, {4, 4}};

template <int SCATTER, int RY, int WY>
static hipError_t launch_pair_shape(const Geometry &geo, int zc, const PairOp<SCATTER> &op,
                                    hipStream_t st) {
  const Grid g = pair_grid(geo.nx, geo.ny, geo.nz, RY * WY, zc > 0 ? zc : 64);
  const int nb = g.ntx * g.nty * g.ntz;
  hipLaunchKernelGGL((pair_kernel<SCATTER, RY, WY>), dim3(nb), dim3(kTX * WY), 0, st, g,
                     geo.taps, geo.sponge, op);
  return hipGetLastError();
}

template <int SCATTER>
static hipError_t launch_pair(const Geometry &geo, const Tune &tu, const PairOp<SCATTER> &op,
                              hipStream_t st) {
  if (geo.nx < 2 * kH + 1 || geo.ny < 2 * kH + 1 || geo.nz < 2 * kH + 1)
    return hipErrorInvalidValue;
  switch (tu.shape) {
    case 0: return launch_pair_shape<SCATTER, 2, 8>(geo, tu.zchunk, op, st);
    case 1: return launch_pair_shape<SCATTER, 1, 8>(geo, tu.zchunk, op, st);
    case 2: return launch_pair_shape<SCATTER, 2, 4>(geo, tu.zchunk, op, st);
    case 3: return launch_pair_shape<SCATTER, 4, 4>(geo, tu.zchunk, op, st);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_pair_step(const Geometry &geo, const Tune &tu, const PairStep &a,
                            hipStream_t st) {
  auto fill = [&](auto &op) {
    op.s_old = a.s_old; op.s_cur = a.s_cur; op.r_old = a.r_old; op.r_cur = a.r_cur;
    op.s_a = a.s_a; op.r_a = a.r_a; op.s_b = a.s_b; op.r_b = a.r_b;
    op.h_a = a.h_a; op.h_b = a.h_b; op.vel = a.vel; op.model = a.model;
    op.pts = to_points(a.src); op.amp_a = a.amp_a; op.amp_b = a.amp_b;
    op.rm = a.rec; op.tr = a.tr; op.tab_a = a.tab_a; op.tab_b = a.tab_b;
    op.id_a = a.id_a; op.id_b = a.id_b; op.ext_a = a.ext_a; op.ext_b = a.ext_b;
    return launch_pair(geo, tu, op, st);
  };
  if (a.scatter == 0) {
    PairOp<0> op{};
    return fill(op);
  }
  PairOp<1> op{};
  return fill(op);
}

template <int RY, int WY>
static hipError_t launch_lam_pair_shape(const Geometry &geo, int zc, const LamPairOp &op,
                                        hipStream_t st) {
  const Grid g = pair_grid(geo.nx, geo.ny, geo.nz, RY * WY, zc > 0 ? zc : 64);
  const int nb = g.ntx * g.nty * g.ntz;
  hipLaunchKernelGGL((lam_pair_kernel<RY, WY>), dim3(nb), dim3(kTX * WY), 0, st, g, geo.taps,
                     geo.sponge, op);
  return hipGetLastError();
}

hipError_t launch_lam_pair_step(const Geometry &geo, const Tune &tu, const LamPairStep &a,
                                hipStream_t st) {
  if (geo.nx < 2 * kH + 1 || geo.ny < 2 * kH + 1 || geo.nz < 2 * kH + 1)
    return hipErrorInvalidValue;
  LamPairOp op{};
  op.u_old = a.u_old; op.u_cur = a.u_cur; op.u_a = a.u_a; op.u_b = a.u_b;
  op.vel = a.vel; op.sig_a = a.sig_a; op.sig_b = a.sig_b; op.acc = a.acc;
  op.rm = a.rec; op.tr = a.tr; op.tab_a = a.tab_a; op.tab_b = a.tab_b;
  op.id_a = a.id_a; op.id_b = a.id_b; op.inj_a = a.inj_a; op.inj_b = a.inj_b;
  switch (tu.shape) {
    case 0: return launch_lam_pair_shape<2, 8>(geo, tu.zchunk, op, st);
    case 1: return launch_lam_pair_shape<1, 8>(geo, tu.zchunk, op, st);
    case 2: return launch_lam_pair_shape<2, 4>(geo, tu.zchunk, op, st);
    case 3: return launch_lam_pair_shape<4, 4>(geo, tu.zchunk, op, st);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace born
