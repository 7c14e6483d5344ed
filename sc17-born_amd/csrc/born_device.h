// Device-side building blocks of the Born propagator on CDNA4 (gfx950).
//
// march_kernel advances one or two pressure fields by one leapfrog step of
// the 8th-order constant-density acoustic operator.  A block of WY 64-lane
// wavefronts owns a 64 x (RY*WY) cell tile of the xy plane (one wavefront per
// x row group, RY rows per thread) and a chunk of z planes.  z neighbours
// live in a 9-deep register queue per row; x/y neighbours of the current
// plane come from an LDS tile with a 4-cell halo.  The loop is software
// pipelined: the loads of plane z+1 (the queue's new leading plane, the LDS
// halo and the op's pointwise inputs) are issued before plane z is computed,
// so HBM latency overlaps the stencil arithmetic.
//
// What the step does with the Laplacian is supplied by an immutable "op"
// functor inlined into the kernel: sponge, source injection, Born scatter,
// receiver extraction/injection and imaging are fused into the same pass.
// Ops are passed by value as kernel arguments and never written, so their
// pointers stay in the global address space (global_load/store, not flat).
#pragma once
#include <hip/hip_fp16.h>
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace born {

constexpr int kTX = 64;  // x cells per tile (one wavefront)
constexpr int kH = 4;    // stencil half-width

// Stencil taps in the reference's order and axis convention
// (base/cpu_prop.cc:6-18, 233-248, 298-326): cx = fast axis, cy = middle,
// cz = slow.
struct Taps {
  float c0;
  float cx[4], cy[4], cz[4];
};

struct Grid {
  int nx, ny, nz;
  int ntx, nty, ntz;  // tiles along x, y and z chunks
  int zc;             // planes per z chunk
  uint32_t plane;     // nx*ny
};

// Per-axis sponge factors: bound[min(edge_x, edge_y, edge_z)] equals
// min(bx[x], by[y], bz[z]) because the table grows with the edge distance.
struct Sponge {
  const float *bx, *by, *bz;
};

// Point sources: (x, y, z, 0) per point in device memory.
struct Points {
  const int4 *p;
  int n;
};

// Receivers: every receiver plane z has a slot; rmap[slot][y][x] is the first
// receiver at that cell (or -1), rnext[] chains receivers sharing a cell.
struct RecMap {
  const int *plane_slot;  // [nz], -1 when the plane has no receiver
  const int *rmap;        // [nslots][ny][nx]
  const int *rnext;       // [nrec]
  int nrec;
  int ntb;                // trace length (nt_big)
};

// element at 32-bit element offset o (fields are < 2^30 cells)
template <class T>
__device__ __forceinline__ T &at(T *p, uint32_t o) {
  return *(T *)((char *)p + (uint64_t)(o * (uint32_t)sizeof(T)));
}
template <class T>
__device__ __forceinline__ const T &at(const T *p, uint32_t o) {
  return *(const T *)((const char *)p + (uint64_t)(o * (uint32_t)sizeof(T)));
}

// ---- storage of wavefield levels: float, or FP16 with a power-of-two scale
// per stored level (value = half * inv_scale).  Conversions saturate at the
// FP16 range instead of overflowing.
__device__ __forceinline__ float ld_level(const float *p, uint32_t o, float) {
  return at(p, o);
}
__device__ __forceinline__ float ld_level(const __half *p, uint32_t o, float inv) {
  return __half2float(at(p, o)) * inv;
}
__device__ __forceinline__ void st_level(float *p, uint32_t o, float v, float) {
  at(p, o) = v;
}
__device__ __forceinline__ void st_level(__half *p, uint32_t o, float v, float scale) {
  at(p, o) = __float2half_rn(fminf(fmaxf(v * scale, -65504.f), 65504.f));
}

// Scale for a level whose magnitude is at most bound: a power of two that
// maps bound to [8192, 16384), 4x below the FP16 maximum.  Relative precision
// of FP16 does not depend on the scale; the scale only keeps the level inside
// the normal range (a value 2^28 below the bound is still normal).
__device__ __forceinline__ float level_scale(float bound) {
  if (!(bound > 0.f)) return 1.f;
  int e;
  frexpf(bound, &e);  // bound < 2^e
  e = min(max(14 - e, -120), 120);
  return ldexpf(1.f, e);
}

// Scales of the stored levels of one FP16 field.  Level k lives in slot k % 4:
// scale[s] (stored = value * scale) and maxb[s], the largest |value| of the
// level as float bits (non-negative floats order like their bits).  The step
// that reads levels k (cur) and k-1 (old) and writes k+1 (nxt) derives the
// scale of k+1 from a bound on |level k+1| (gain * max k + max k-1 +
// injections), accumulates max |level k+1| with atomics, and clears slot
// k+2, which the following step fills.
struct Ring {
  float *scale;
  unsigned *maxb;
};
struct Slots {
  int cur, old, nxt, clr;
};
__device__ __forceinline__ float ring_max(const Ring &r, int s) {
  return __uint_as_float(r.maxb[s]);
}

// The hardware hands block h to XCD h % 8.  Give every XCD a contiguous range
// of logical tiles so that tiles sharing halo rows share an L2.
__device__ __forceinline__ int xcd_remap(int h, int nb) {
  const int q = nb / 8, r = nb % 8;
  const int xcd = h % 8, k = h / 8;
  return xcd * q + (xcd < r ? xcd : r) + k;
}

// Eight-tap sinc interpolation with the reference's rounding: products and
// sums rounded separately, left to right (base/cpu_prop.cc:169-176).
__device__ __forceinline__ float sinc8(const float *t, const float *v,
                                       long long stride) {
  float s = __fmul_rn(t[0], v[0]);
#pragma unroll
  for (int k = 1; k < 8; k++) s = __fadd_rn(s, __fmul_rn(t[k], v[k * stride]));
  return s;
}

// Uniform per-block mask of the source points inside the block's cells.
__device__ __forceinline__ uint32_t points_in_block(const Points &pts, int x0,
                                                   int x1, int y0, int y1,
                                                   int z0, int z1) {
  uint32_t m = 0;
  for (int i = 0; i < pts.n; i++) {
    const int4 q = pts.p[i];
    if (q.x >= x0 && q.x < x1 && q.y >= y0 && q.y < y1 && q.z >= z0 && q.z < z1)
      m |= 1u << i;
  }
  return m;
}

__device__ __forceinline__ uint32_t points_on_plane(const Points &pts,
                                                    uint32_t mask, int z) {
  uint32_t m = 0;
  for (uint32_t b = mask; b; b &= b - 1) {
    const int i = __builtin_ctz(b);
    if (pts.p[i].z == z) m |= 1u << i;
  }
  return m;
}

// v plus, in list order, amp[i] of the points at (x, y) among plane mask m
__device__ __forceinline__ float add_points(const Points &pts, uint32_t m, int x,
                                            int y, float v, const float *amp) {
  for (uint32_t b = m; b; b &= b - 1) {
    const int i = __builtin_ctz(b);
    const int4 q = pts.p[i];
    if (q.x == x && q.y == y) v = __fadd_rn(v, amp[i]);
  }
  return v;
}

struct CellArgs {
  int x, y, z;
  uint32_t o;  // element offset of the cell
  float b;     // sponge factor (1 outside the sponge)
};

// Generic z-marching leapfrog kernel.  Op supplies (all const):
//   static constexpr int NF, NPRE, NMAX; static constexpr bool DAMP;
//   Blk block(x0, x1, y0, y1, z0, z1);       uniform per-block state
//   float ld(const Blk&, int f, uint32_t o);  current level of field f
//   Pln plane(const Blk&, int z);            uniform per-plane state
//   void prefetch(const Blk&, float (&pre)[NPRE], uint32_t o);
//   void cell(const Blk&, const Pln&, const CellArgs&, const float *lap,
//             const float *ctr, const float (&pre)[NPRE], float (&tmax)[NF]);
//   void first(const Blk&);                  (NMAX > 0) once per launch
//   void publish_max(int f, float m);        (NMAX > 0) block max of |new|
template <class Op, int RY, int WY>
__device__ __forceinline__ void march_tile(const Grid &g, const Taps &t, const Sponge &sp,
                                           const Op &op, const int L) {
  constexpr int NF = Op::NF;
  constexpr int NP = Op::NPRE > 0 ? Op::NPRE : 1;
  constexpr int TY = RY * WY;
  constexpr int LW = kTX + 2 * kH;
  constexpr int LH = TY + 2 * kH;
  constexpr int NXH = 2 * kH * TY;  // x-halo cells of a plane
  __shared__ float tile[NF][LH][LW];

  const int bx = L % g.ntx;
  const int rest = L / g.ntx;
  const int by = rest % g.nty;
  const int bz = rest / g.nty;
  const int tx = (int)threadIdx.x % kTX;
  const int ty = (int)threadIdx.x / kTX;
  // x tiles start at the row start so that a wavefront's row maps onto whole
  // cache lines; the 4-cell shell at either end is masked
  const int x0 = bx * kTX, y0 = kH + by * TY;
  const int z0 = kH + bz * g.zc;
  const int z1 = min(z0 + g.zc, g.nz - kH);
  const int x = x0 + tx;
  const bool xin = x >= kH && x < g.nx - kH;
  const int xc = min(x, g.nx - 1);

  const auto blk = op.block(max(x0, kH), min(x0 + kTX, g.nx - kH), y0,
                            min(y0 + TY, g.ny - kH), z0, z1);
  if constexpr (Op::NMAX > 0)
    if (L == 0 && threadIdx.x == 0) op.first(blk);

  int yy[RY];
  bool in[RY];
  uint32_t row[RY];
  float bxy[RY];
#pragma unroll
  for (int j = 0; j < RY; j++) {
    yy[j] = y0 + ty + WY * j;
    in[j] = xin && yy[j] < g.ny - kH;
    const int yc = min(yy[j], g.ny - 1);
    row[j] = (uint32_t)yc * g.nx + xc;
    bxy[j] = Op::DAMP ? fminf(sp.bx[xc], sp.by[yc]) : 1.f;
  }

  // LDS halo: 8 y-halo rows of 64 cells and 8 x-halo columns of TY rows,
  // spread over the block's threads
  constexpr int NYH = 2 * kH * kTX;          // y-halo cells of a plane
  constexpr int NHALO = NYH + NXH;
  constexpr int HPT = (NHALO + kTX * WY - 1) / (kTX * WY);  // per thread
  uint32_t hoff[HPT];
  short hr[HPT], hc[HPT];
  bool hv[HPT];
#pragma unroll
  for (int k = 0; k < HPT; k++) {
    const int h = (int)threadIdx.x + k * kTX * WY;
    hv[k] = h < NHALO;
    int lr, lc;
    if (h < NYH) {  // rows 0..3 and TY+4..TY+7, columns 4..67
      const int r = h / kTX;
      lr = r < kH ? r : TY + r;
      lc = kH + h % kTX;
    } else {  // columns 0..3 and 68..71 of rows 4..TY+3
      const int hh = min(h - NYH, NXH - 1);
      const int c = hh % (2 * kH);
      lr = kH + hh / (2 * kH);
      lc = c < kH ? c : kTX + c;
    }
    hr[k] = (short)lr;
    hc[k] = (short)lc;
    const int gy = min(max(y0 - kH + lr, 0), g.ny - 1);
    const int gx = min(max(x0 - kH + lc, 0), g.nx - 1);
    hoff[k] = (uint32_t)gy * g.nx + gx;
  }

  float q[NF][RY][2 * kH + 1];
  float lead[NF][RY];
  float halo[NF][HPT];
  float pre[RY][NP], pre_n[RY][NP];
  float tmax[NF];
#pragma unroll
  for (int f = 0; f < NF; f++) {
    tmax[f] = 0.f;
#pragma unroll
    for (int j = 0; j < RY; j++) {
#pragma unroll
      for (int k = 0; k < 2 * kH; k++)
        q[f][j][k] = op.ld(blk, f, (uint32_t)(z0 - kH + k) * g.plane + row[j]);
      lead[f][j] = op.ld(blk, f, (uint32_t)(z0 + kH) * g.plane + row[j]);
    }
#pragma unroll
    for (int k = 0; k < HPT; k++)
      halo[f][k] = hv[k] ? op.ld(blk, f, (uint32_t)z0 * g.plane + hoff[k]) : 0.f;
  }
#pragma unroll
  for (int j = 0; j < RY; j++) op.prefetch(blk, pre[j], (uint32_t)z0 * g.plane + row[j]);

  for (int z = z0; z < z1; ++z) {
    const uint32_t zo = (uint32_t)z * g.plane;
#pragma unroll
    for (int f = 0; f < NF; f++) {
#pragma unroll
      for (int j = 0; j < RY; j++) {
        q[f][j][2 * kH] = lead[f][j];
        tile[f][kH + ty + WY * j][kH + tx] = q[f][j][kH];
      }
#pragma unroll
      for (int k = 0; k < HPT; k++)
        if (hv[k]) tile[f][hr[k]][hc[k]] = halo[f][k];
    }
    // next plane's loads, in flight during this plane's compute
    if (z + 1 < z1) {
      const uint32_t zn = zo + g.plane;
#pragma unroll
      for (int f = 0; f < NF; f++) {
#pragma unroll
        for (int j = 0; j < RY; j++) lead[f][j] = op.ld(blk, f, zn + kH * g.plane + row[j]);
#pragma unroll
        for (int k = 0; k < HPT; k++)
          if (hv[k]) halo[f][k] = op.ld(blk, f, zn + hoff[k]);
      }
#pragma unroll
      for (int j = 0; j < RY; j++) op.prefetch(blk, pre_n[j], zn + row[j]);
    }
    __syncthreads();
    const auto pl = op.plane(blk, z);
    const float bzv = Op::DAMP ? sp.bz[z] : 1.f;
#pragma unroll
    for (int j = 0; j < RY; j++) {
      if (!in[j]) continue;
      const int lr = kH + ty + WY * j, lc = kH + tx;
      float lap[NF], ctr[NF];
#pragma unroll
      for (int f = 0; f < NF; f++) {
        float a = t.c0 * q[f][j][kH];
#pragma unroll
        for (int k = 1; k <= kH; k++)
          a += t.cx[k - 1] * (tile[f][lr][lc - k] + tile[f][lr][lc + k]);
#pragma unroll
        for (int k = 1; k <= kH; k++)
          a += t.cy[k - 1] * (tile[f][lr - k][lc] + tile[f][lr + k][lc]);
#pragma unroll
        for (int k = 1; k <= kH; k++)
          a += t.cz[k - 1] * (q[f][j][kH - k] + q[f][j][kH + k]);
        lap[f] = a;
        ctr[f] = q[f][j][kH];
      }
      CellArgs c;
      c.x = x;
      c.y = yy[j];
      c.z = z;
      c.o = zo + row[j];
      c.b = fminf(bxy[j], bzv);
      op.cell(blk, pl, c, lap, ctr, pre[j], tmax);
    }
    __syncthreads();
#pragma unroll
    for (int f = 0; f < NF; f++)
#pragma unroll
      for (int j = 0; j < RY; j++)
#pragma unroll
        for (int k = 0; k < 2 * kH; k++) q[f][j][k] = q[f][j][k + 1];
#pragma unroll
    for (int j = 0; j < RY; j++)
#pragma unroll
      for (int k = 0; k < NP; k++) pre[j][k] = pre_n[j][k];
  }
  // block maxima of |new level| for the next step's FP16 scales
  if constexpr (Op::NMAX > 0) {
    __shared__ float red[NF][WY];
#pragma unroll
    for (int f = 0; f < NF; f++) {
      float m = tmax[f];
      for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
      if (tx == 0) red[f][ty] = m;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
      for (int f = 0; f < NF; f++) {
        float m = 0.f;
        for (int w = 0; w < WY; w++) m = fmaxf(m, red[f][w]);
        op.publish_max(f, m);
      }
    }
  }
}

template <class Op, int RY, int WY>
__global__ void __launch_bounds__(kTX *WY)
    march_kernel(const Grid g, const Taps t, const Sponge sp, const Op op) {
  march_tile<Op, RY, WY>(g, t, sp, op, xcd_remap((int)blockIdx.x, (int)gridDim.x));
}

// Tiles of a batch of shots with their own grids: shot s owns the logical
// tiles [tile0[s], tile0[s + 1]).
constexpr int kMaxBatch = 64;
struct BatchTiles {
  int n;
  int zc;
  int tile0[kMaxBatch + 1];
};

// Tile grid of a domain for tile shape (RY, WY) and z chunk zc.
__host__ __device__ inline Grid tile_grid(int nx, int ny, int nz, int ty, int zc) {
  Grid g;
  g.nx = nx;
  g.ny = ny;
  g.nz = nz;
  g.plane = (uint32_t)nx * ny;
  g.ntx = (nx - kH + kTX - 1) / kTX;
  g.nty = (ny - 2 * kH + ty - 1) / ty;
  g.zc = zc;
  g.ntz = (nz - 2 * kH + zc - 1) / zc;
  return g;
}

// One launch advances every shot of a batch by one step.  BOp::shot(s, ty,
// zc) returns {Grid, Sponge, Op} of shot s with the op's pointers moved to
// the shot's part of the batch arrays.
template <class BOp, int RY, int WY>
__global__ void __launch_bounds__(kTX *WY)
    march_batch_kernel(const BatchTiles bt, const Taps t, const BOp bop) {
  const int Lg = xcd_remap((int)blockIdx.x, (int)gridDim.x);
  int s = 0;
  while (s + 1 < bt.n && Lg >= bt.tile0[s + 1]) s++;
  const auto sh = bop.shot(s, RY * WY, bt.zc);
  march_tile<typename BOp::Op, RY, WY>(sh.g, t, sh.sp, sh.op, Lg - bt.tile0[s]);
}

}  // namespace born
