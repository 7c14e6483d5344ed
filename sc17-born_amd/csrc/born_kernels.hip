// Fused time-step kernels of the Born propagator (gfx950).
//
// Each op below is one time step of one of the reference's loops
// (base/cpu_prop.cc), with every per-cell pass of that step fused into the
// single z-marching sweep of march_kernel:
//   FwdOp     rtmForward step (cpu_prop.cc:31-53): prop s, prop r, damp both,
//             inject source, Born scatter, extract receivers.
//   SrcOp     sourceProp step (cpu_prop.cc:112-122): undamped, inject source.
//   RadjOp    rtmAdjoint step (cpu_prop.cc:71-95): reverse s (vel1), r (vel2),
//             damp r, inject receivers, imaging condition.
//   SrecOp    background step of the exact adjoint (the s half of FwdOp) that
//             also stores the scatter operand of the step.
//   LamOp     adjoint-state step of the exact adjoint (see born_abi.hip).
//
// Storage convention: a field buffer holds each level exactly as the
// reference holds it right after that level was produced (damped, injected).
// The reference additionally damps the current level in place after the
// update (cpu_prop.cc:141-165 on p1); that value is only ever read back as the
// "old" level of the next step, so it is recomputed there as fl(b * old)
// instead of being written: same arithmetic, one HBM write less per field.
#include "born_device.h"
#include "born_launch.h"

namespace born {

struct NoBlk {};
struct RecPln {
  int slot;
};
struct SrcPln {
  uint32_t pm;
  int slot;
};

__device__ __forceinline__ int rec_first(const RecMap &rm, int slot, int x, int y,
                                         int nx, int ny) {
  return rm.rmap[((long long)slot * ny + y) * nx + x];
}

// ---------------------------------------------------------------- forward
template <int SCATTER>
struct FwdOp {
  static constexpr int NF = 2, NPRE = 4;
  static constexpr bool DAMP = true;
  const float *__restrict__ s1;
  const float *__restrict__ r1;
  float *__restrict__ s0;
  float *__restrict__ r0;
  const float *__restrict__ vel;
  const float *__restrict__ model;
  float *__restrict__ hist;  // optional scatter-operand history
  Points pts;
  const float *__restrict__ amp;  // source amplitudes of this step [npts]
  RecMap rm;
  float *__restrict__ tr;  // traces, time-major [ntb][nrec]
  const float *__restrict__ tab;
  int id, extract, nx, ny;

  __device__ const float *cur(int f) const { return f == 0 ? s1 : r1; }
  __device__ uint32_t block(int x0, int x1, int y0, int y1, int z0, int z1) const {
    return points_in_block(pts, x0, x1, y0, y1, z0, z1);
  }
  __device__ SrcPln plane(uint32_t blk, int z) const {
    return {points_on_plane(pts, blk, z), rm.plane_slot[z]};
  }
  __device__ void prefetch(float (&p)[NPRE], uint32_t o) const {
    p[0] = at(s0, o);
    p[1] = at(r0, o);
    p[2] = at(vel, o);
    p[3] = at(model, o);
  }
  __device__ void cell(const SrcPln &pl, const CellArgs &c, const float *lap,
                       const float *ctr, const float (&p)[NPRE]) const {
    const float v = p[2], b = c.b;
    float sn = b * (v * lap[0] + ctr[0] + ctr[0] - b * p[0]);
    if (pl.pm) sn = add_points(pts, pl.pm, c.x, c.y, sn, amp);
    float rn = b * (v * lap[1] + ctr[1] + ctr[1] - b * p[1]);
    float sig;
    if (SCATTER == 0) {
      // the reference's double literal: evaluated in double, rounded once
      // (cpu_prop.cc:220)
      rn = (float)((double)rn + .000001 * (double)sn * (double)p[3]);
      sig = sn;
    } else {
      sig = b * lap[0];
      rn = rn + p[3] * sig;
    }
    at(s0, c.o) = sn;
    at(r0, c.o) = rn;
    if (hist) at(hist, c.o) = sig;
    if (pl.slot >= 0 && extract) {
      for (int i = rec_first(rm, pl.slot, c.x, c.y, nx, ny); i >= 0; i = rm.rnext[i]) {
        float *q = tr + (long long)id * rm.nrec + i;
#pragma unroll
        for (int k = 0; k < 8; k++)
          q[(long long)k * rm.nrec] =
              __fadd_rn(q[(long long)k * rm.nrec], __fmul_rn(rn, tab[k]));
      }
    }
  }
};

// ------------------------------------------------------------- sourceProp
struct SrcOp {
  static constexpr int NF = 1, NPRE = 2;
  static constexpr bool DAMP = false;
  const float *__restrict__ s1;
  float *__restrict__ s0;
  const float *__restrict__ vel;
  Points pts;
  const float *__restrict__ amp;
  __device__ const float *cur(int) const { return s1; }
  __device__ uint32_t block(int x0, int x1, int y0, int y1, int z0, int z1) const {
    return points_in_block(pts, x0, x1, y0, y1, z0, z1);
  }
  __device__ uint32_t plane(uint32_t blk, int z) const {
    return points_on_plane(pts, blk, z);
  }
  __device__ void prefetch(float (&p)[NPRE], uint32_t o) const {
    p[0] = at(s0, o);
    p[1] = at(vel, o);
  }
  __device__ void cell(uint32_t pm, const CellArgs &c, const float *lap, const float *ctr,
                       const float (&p)[NPRE]) const {
    float sn = p[1] * lap[0] + ctr[0] + ctr[0] - p[0];
    if (pm) sn = add_points(pts, pm, c.x, c.y, sn, amp);
    at(s0, c.o) = sn;
  }
};

// ------------------------------------------------------ reference adjoint
struct RadjOp {
  static constexpr int NF = 2, NPRE = 5;
  static constexpr bool DAMP = true;
  const float *__restrict__ s1;
  const float *__restrict__ r1;
  float *__restrict__ s0;
  float *__restrict__ r0;
  const float *__restrict__ vel1;
  const float *__restrict__ vel2;
  float *__restrict__ img;
  int s_prop, r_prop;
  // source amplitude injected into the current level by the previous
  // iteration (cpu_prop.cc:89-92), applied here to the old level it became
  Points pts;
  const float *__restrict__ amp_pre;
  int pre;
  RecMap rm;
  const float *__restrict__ tr;
  const float *__restrict__ tab;
  float sc;  // (float)dir / (float)jt
  int id, inject, nx, ny;

  __device__ const float *cur(int f) const { return f == 0 ? s1 : r1; }
  __device__ uint32_t block(int x0, int x1, int y0, int y1, int z0, int z1) const {
    return pre ? points_in_block(pts, x0, x1, y0, y1, z0, z1) : 0u;
  }
  __device__ SrcPln plane(uint32_t blk, int z) const {
    return {points_on_plane(pts, blk, z), rm.plane_slot[z]};
  }
  __device__ void prefetch(float (&p)[NPRE], uint32_t o) const {
    p[0] = at(s0, o);
    p[1] = at(r0, o);
    p[2] = at(vel1, o);
    p[3] = at(vel2, o);
    p[4] = at(img, o);
  }
  __device__ void cell(const SrcPln &pl, const CellArgs &c, const float *lap,
                       const float *ctr, const float (&p)[NPRE]) const {
    float sn;
    if (s_prop) {
      float so = p[0];
      if (pl.pm) so = add_points(pts, pl.pm, c.x, c.y, so, amp_pre);
      sn = p[2] * lap[0] + ctr[0] + ctr[0] - so;
      at(s0, c.o) = sn;
    } else {
      sn = p[0];
    }
    const float b = c.b;
    float rn = r_prop ? b * (p[3] * lap[1] + ctr[1] + ctr[1] - b * p[1]) : b * (b * p[1]);
    if (pl.slot >= 0 && inject) {
      for (int i = rec_first(rm, pl.slot, c.x, c.y, nx, ny); i >= 0; i = rm.rnext[i])
        rn = __fadd_rn(rn, __fmul_rn(sc, sinc8(tab, tr + (long long)id * rm.nrec + i,
                                               rm.nrec)));
    }
    at(r0, c.o) = rn;
    at(img, c.o) = p[4] + sn * rn;
  }
};

// ------------------------------------- exact adjoint: background recompute
template <int SCATTER>
struct SrecOp {
  static constexpr int NF = 1, NPRE = 2;
  static constexpr bool DAMP = true;
  const float *__restrict__ s1;
  float *__restrict__ s0;
  const float *__restrict__ vel;
  float *__restrict__ sig;  // scatter operand of this step (nullptr: none)
  Points pts;
  const float *__restrict__ amp;
  __device__ const float *cur(int) const { return s1; }
  __device__ uint32_t block(int x0, int x1, int y0, int y1, int z0, int z1) const {
    return points_in_block(pts, x0, x1, y0, y1, z0, z1);
  }
  __device__ uint32_t plane(uint32_t blk, int z) const {
    return points_on_plane(pts, blk, z);
  }
  __device__ void prefetch(float (&p)[NPRE], uint32_t o) const {
    p[0] = at(s0, o);
    p[1] = at(vel, o);
  }
  __device__ void cell(uint32_t pm, const CellArgs &c, const float *lap, const float *ctr,
                       const float (&p)[NPRE]) const {
    const float b = c.b;
    float sn = b * (p[1] * lap[0] + ctr[0] + ctr[0] - b * p[0]);
    if (pm) sn = add_points(pts, pm, c.x, c.y, sn, amp);
    at(s0, c.o) = sn;
    if (sig) at(sig, c.o) = SCATTER == 0 ? sn : b * lap[0];
  }
};

// ---------------------------------------- exact adjoint: adjoint-state step
// nu = V B lambda obeys the forward recursion with V B-weighted receiver
// injection; the image accumulates sigma * nu and is divided by V B at the end.
struct LamOp {
  static constexpr int NF = 1, NPRE = 4;
  static constexpr bool DAMP = true;
  const float *__restrict__ u1;
  float *__restrict__ u0;
  const float *__restrict__ vel;
  const float *__restrict__ sig;
  float *__restrict__ acc;
  RecMap rm;
  const float *__restrict__ tr;
  const float *__restrict__ tab;
  int id, inject, nx, ny;
  __device__ const float *cur(int) const { return u1; }
  __device__ int block(int, int, int, int, int, int) const { return 0; }
  __device__ int plane(int, int z) const { return rm.plane_slot[z]; }
  __device__ void prefetch(float (&p)[NPRE], uint32_t o) const {
    p[0] = at(u0, o);
    p[1] = at(vel, o);
    p[2] = at(sig, o);
    p[3] = at(acc, o);
  }
  __device__ void cell(int slot, const CellArgs &c, const float *lap, const float *ctr,
                       const float (&p)[NPRE]) const {
    const float b = c.b, v = p[1];
    float un = b * (v * lap[0] + ctr[0] + ctr[0] - b * p[0]);
    if (slot >= 0 && inject) {
      float g = 0.f;
      for (int i = rec_first(rm, slot, c.x, c.y, nx, ny); i >= 0; i = rm.rnext[i])
        g += sinc8(tab, tr + (long long)id * rm.nrec + i, rm.nrec);
      un += v * b * g;
    }
    at(u0, c.o) = un;
    at(acc, c.o) = p[3] + p[2] * un;
  }
};

// ------------------------------------------------------------ small kernels

__global__ void fill_kernel(float *p, long long n, float v) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    p[i] = v;
}

// amp[s*npts + i] = scale * sinc8(tab[ii], src[i][id..id+7]) with
// tt = s + t_off, id = tt / jt, ii = tt - id*jt; zero when id + 7 >= ntsrc
// (the guard of cpu_prop.cc:167).
__global__ void src_amp_kernel(float *amp, int nsteps, int npts, const float *src,
                               int ntsrc, const float *tab, int nsinc, int jt, int ntab,
                               int t_off, float scale) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nsteps * npts) return;
  const int s = t / npts, i = t % npts;
  const int tt = s + t_off;
  const int id = tt / jt;
  const int ii = tt - id * jt;
  float a = 0.f;
  if (id + 7 < ntsrc && ii < ntab)
    a = __fmul_rn(scale, sinc8(tab + (long long)ii * nsinc, src + (long long)i * ntsrc + id, 1));
  amp[t] = a;
}

// out[i] += scale * acc[i] / (vel[i] * b[i]) on the interior.
__global__ void adj_finish_kernel(float *out, const float *acc, const float *vel,
                                  const Sponge sp, int nx, int ny, int nz, float scale) {
  const long long n12 = (long long)nx * ny;
  const long long n = n12 * nz;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x) {
    const int x = (int)(i % nx), y = (int)((i / nx) % ny), z = (int)(i / n12);
    if (x < kH || x >= nx - kH || y < kH || y >= ny - kH || z < kH || z >= nz - kH)
      continue;
    const float b = fminf(fminf(sp.bx[x], sp.by[y]), sp.bz[z]);
    out[i] += scale * acc[i] / (vel[i] * b);
  }
}

// [n0][n1] -> [n1][n0]
__global__ void transpose_kernel(float *dst, const float *src, int n0, int n1) {
  __shared__ float t[32][33];
  const int b0 = blockIdx.y * 32, b1 = blockIdx.x * 32;
  for (int k = threadIdx.y; k < 32; k += blockDim.y) {
    const int i0 = b0 + k, i1 = b1 + threadIdx.x;
    if (i0 < n0 && i1 < n1) t[k][threadIdx.x] = src[(long long)i0 * n1 + i1];
  }
  __syncthreads();
  for (int k = threadIdx.y; k < 32; k += blockDim.y) {
    const int i1 = b1 + k, i0 = b0 + threadIdx.x;
    if (i0 < n0 && i1 < n1) dst[(long long)i1 * n0 + i0] = t[threadIdx.x][k];
  }
}

// ------------------------------------------------------------- launchers

const TileShape kShapes[kNumShapes] = {{2, 4}, {4, 4}, {2, 8}, {2, 16}, {4, 8}, {3, 8}};

static Grid make_grid(const Geometry &geo, const TileShape &ts, int zc) {
  Grid g;
  g.nx = geo.nx;
  g.ny = geo.ny;
  g.nz = geo.nz;
  g.plane = (uint32_t)geo.nx * geo.ny;
  const int ty = ts.ry * ts.wy;
  g.ntx = (geo.nx - kH + kTX - 1) / kTX;
  g.nty = (geo.ny - 2 * kH + ty - 1) / ty;
  g.zc = zc > 0 ? zc : 32;
  g.ntz = (geo.nz - 2 * kH + g.zc - 1) / g.zc;
  return g;
}

template <class Op, int RY, int WY>
static hipError_t launch_shape(const Geometry &geo, int zc, const Op &op, hipStream_t st) {
  const Grid g = make_grid(geo, {RY, WY}, zc);
  const int nb = g.ntx * g.nty * g.ntz;
  hipLaunchKernelGGL((march_kernel<Op, RY, WY>), dim3(nb), dim3(kTX * WY), 0, st, g,
                     geo.taps, geo.sponge, op);
  return hipGetLastError();
}

template <class Op>
static hipError_t launch(const Geometry &geo, const Tune &tu, const Op &op, hipStream_t st) {
  switch (tu.shape) {
    case 0: return launch_shape<Op, 2, 4>(geo, tu.zchunk, op, st);
    case 1: return launch_shape<Op, 4, 4>(geo, tu.zchunk, op, st);
    case 2: return launch_shape<Op, 2, 8>(geo, tu.zchunk, op, st);
    case 3: return launch_shape<Op, 2, 16>(geo, tu.zchunk, op, st);
    case 4: return launch_shape<Op, 4, 8>(geo, tu.zchunk, op, st);
    case 5: return launch_shape<Op, 3, 8>(geo, tu.zchunk, op, st);
    default: return hipErrorInvalidValue;
  }
}

static Points to_points(const SrcPts &s) { return {(const int4 *)s.p, s.n}; }

hipError_t launch_fwd_step(const Geometry &geo, const Tune &tu, const FwdStep &a,
                           hipStream_t st) {
  auto fill = [&](auto &op) {
    op.s1 = a.s1; op.r1 = a.r1; op.s0 = a.s0; op.r0 = a.r0;
    op.vel = a.vel; op.model = a.model; op.hist = a.hist;
    op.pts = to_points(a.src); op.amp = a.amp;
    op.rm = a.rec; op.tr = a.tr; op.tab = a.tab; op.id = a.id;
    op.extract = a.extract; op.nx = geo.nx; op.ny = geo.ny;
  };
  if (a.scatter == 0) {
    FwdOp<0> op{};
    fill(op);
    return launch(geo, tu, op, st);
  }
  FwdOp<1> op{};
  fill(op);
  return launch(geo, tu, op, st);
}

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

hipError_t launch_srec_step(const Geometry &geo, const Tune &tu, const SrecStep &a,
                            hipStream_t st) {
  auto fill = [&](auto &op) {
    op.s1 = a.s1; op.s0 = a.s0; op.vel = a.vel; op.sig = a.sig;
    op.pts = to_points(a.src); op.amp = a.amp;
  };
  if (a.scatter == 0) {
    SrecOp<0> op{};
    fill(op);
    return launch(geo, tu, op, st);
  }
  SrecOp<1> op{};
  fill(op);
  return launch(geo, tu, op, st);
}

hipError_t launch_lam_step(const Geometry &geo, const Tune &tu, const LamStep &a,
                           hipStream_t st) {
  LamOp op{};
  op.u1 = a.u1; op.u0 = a.u0; op.vel = a.vel; op.sig = a.sig; op.acc = a.acc;
  op.rm = a.rec; op.tr = a.tr; op.tab = a.tab; op.id = a.id;
  op.inject = a.inject; op.nx = geo.nx; op.ny = geo.ny;
  return launch(geo, tu, op, st);
}

hipError_t launch_fill(float *p, long long n, float v, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  long long nb = (n + 255) / 256;
  if (nb > 65536) nb = 65536;
  hipLaunchKernelGGL(fill_kernel, dim3((unsigned)nb), dim3(256), 0, st, p, n, v);
  return hipGetLastError();
}

hipError_t launch_src_amp(float *amp, int nsteps, int npts, const float *src, int ntsrc,
                          const float *tab, int nsinc, int jt, int ntab, int t_off,
                          float scale, hipStream_t st) {
  const int n = nsteps * npts;
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(src_amp_kernel, dim3((n + 255) / 256), dim3(256), 0, st, amp, nsteps,
                     npts, src, ntsrc, tab, nsinc, jt, ntab, t_off, scale);
  return hipGetLastError();
}

hipError_t launch_adj_finish(const Geometry &geo, float *out, const float *acc,
                             const float *vel, float scale, hipStream_t st) {
  const long long n = (long long)geo.nx * geo.ny * geo.nz;
  long long nb = (n + 255) / 256;
  if (nb > 65536) nb = 65536;
  hipLaunchKernelGGL(adj_finish_kernel, dim3((unsigned)nb), dim3(256), 0, st, out, acc, vel,
                     geo.sponge, geo.nx, geo.ny, geo.nz, scale);
  return hipGetLastError();
}

hipError_t launch_transpose(float *dst, const float *src, int n0, int n1, hipStream_t st) {
  if (n0 <= 0 || n1 <= 0) return hipSuccess;
  hipLaunchKernelGGL(transpose_kernel, dim3((n1 + 31) / 32, (n0 + 31) / 32), dim3(32, 8), 0,
                     st, dst, src, n0, n1);
  return hipGetLastError();
}

}  // namespace born
