// Device-side building blocks of the Born propagator on CDNA4 (gfx950).
//
// One z-marching kernel template (march_kernel) advances one or two pressure
// fields by one leapfrog step of the 8th-order constant-density acoustic
// operator.  Each 256-thread block owns a 64 x 16 cell tile of the xy plane
// (one 64-lane wavefront per x row, 4 rows per thread) and a chunk of z
// planes.  The z neighbours live in a 9-deep register queue per row; the x/y
// neighbours of the current plane come from an LDS tile with a 4-cell halo.
// What the step does with the Laplacian is supplied by an "op" functor that
// is inlined into the kernel: sponge, source injection, Born scatter,
// receiver extraction/injection and imaging are all fused into the same pass
// over HBM.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace born {

constexpr int kTX = 64;              // x cells per tile (one wavefront)
constexpr int kWY = 4;               // wavefronts per block
constexpr int kRY = 4;               // y rows per thread
constexpr int kTY = kWY * kRY;       // y rows per tile
constexpr int kH = 4;                // stencil half-width
constexpr int kLW = kTX + 2 * kH;    // LDS tile width
constexpr int kLH = kTY + 2 * kH;    // LDS tile height
constexpr int kThreads = kTX * kWY;  // threads per block
constexpr int kMaxPts = 16;          // source points handled inside the step

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
  long long n12;
};

// Up to kMaxPts point sources injected into the new level inside the step.
struct PtList {
  int n;
  int x[kMaxPts], y[kMaxPts], z[kMaxPts];
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

// Per-axis sponge factors: bound[min(edge_x, edge_y, edge_z)] equals
// min(bx[x], by[y], bz[z]) because the table grows with the edge distance.
struct Sponge {
  const float *bx, *by, *bz;
};

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

struct CellArgs {
  int x, y, z;
  long long idx;
  float b;  // sponge factor (1 outside the sponge)
};

// Generic z-marching leapfrog kernel.  Op supplies:
//   static constexpr int NF;               stencil fields (1 or 2)
//   static constexpr bool DAMP;            whether the sponge factor is used
//   const float *cur(int f) const;         current level of field f
//   void block_init(x0, x1, y0, y1, z0, z1);  uniform per-block setup
//   void plane(int z);                     uniform per-plane setup
//   void cell(const CellArgs&, const float* lap, const float* ctr);
template <class Op>
__global__ void __launch_bounds__(kThreads)
    march_kernel(const Grid g, const Taps t, const Sponge sp, Op op) {
  constexpr int NF = Op::NF;
  __shared__ float tile[NF][kLH][kLW];

  const int L = xcd_remap((int)blockIdx.x, (int)gridDim.x);
  const int bx = L % g.ntx;
  const int rest = L / g.ntx;
  const int by = rest % g.nty;
  const int bz = rest / g.nty;
  const int tx = (int)threadIdx.x % kTX;
  const int ty = (int)threadIdx.x / kTX;
  const int x0 = kH + bx * kTX, y0 = kH + by * kTY;
  const int z0 = kH + bz * g.zc;
  const int z1 = min(z0 + g.zc, g.nz - kH);
  const int x = x0 + tx;
  const bool xin = x < g.nx - kH;
  const int xc = min(x, g.nx - 1);

  op.block_init(x0, min(x0 + kTX, g.nx - kH), y0, min(y0 + kTY, g.ny - kH),
                z0, z1);

  int yy[kRY];
  bool in[kRY];
  long long row[kRY];
  float bxy[kRY];
#pragma unroll
  for (int j = 0; j < kRY; j++) {
    yy[j] = y0 + ty + kWY * j;
    in[j] = xin && yy[j] < g.ny - kH;
    const int yc = min(yy[j], g.ny - 1);
    row[j] = (long long)yc * g.nx + xc;
    bxy[j] = Op::DAMP ? fminf(sp.bx[xc], sp.by[yc]) : 1.f;
  }

  // halo bookkeeping: 8 y-halo rows (two per thread) and 8 x-halo columns
  // of the 16 tile rows (one cell for each of the first 128 threads)
  const long long yha = (long long)(y0 - kH + ty) * g.nx + xc;
  const long long yhb = (long long)min(y0 + kTY + ty, g.ny - 1) * g.nx + xc;
  const int tid = (int)threadIdx.x;
  const bool xh = tid < 2 * kH * kTY;
  const int xh_r = (tid / (2 * kH)) % kTY;
  const int xh_c = tid % (2 * kH);
  const int xh_l = xh_c < kH ? xh_c : kTX + xh_c;
  const long long xho = (long long)min(y0 + xh_r, g.ny - 1) * g.nx +
                        min(x0 - kH + xh_l, g.nx - 1);

  float q[NF][kRY][2 * kH + 1];
#pragma unroll
  for (int f = 0; f < NF; f++) {
    const float *p = op.cur(f);
#pragma unroll
    for (int j = 0; j < kRY; j++)
#pragma unroll
      for (int k = 0; k < 2 * kH; k++)
        q[f][j][k] = p[(long long)(z0 - kH + k) * g.n12 + row[j]];
  }

  for (int z = z0; z < z1; ++z) {
    const long long zo = (long long)z * g.n12;
#pragma unroll
    for (int f = 0; f < NF; f++) {
      const float *p = op.cur(f);
#pragma unroll
      for (int j = 0; j < kRY; j++) q[f][j][2 * kH] = p[zo + kH * g.n12 + row[j]];
#pragma unroll
      for (int j = 0; j < kRY; j++) tile[f][kH + ty + kWY * j][kH + tx] = q[f][j][kH];
      tile[f][ty][kH + tx] = p[zo + yha];
      tile[f][kTY + kH + ty][kH + tx] = p[zo + yhb];
      if (xh) tile[f][kH + xh_r][xh_l] = p[zo + xho];
    }
    __syncthreads();
    op.plane(z);
    const float bzv = Op::DAMP ? sp.bz[z] : 1.f;
#pragma unroll
    for (int j = 0; j < kRY; j++) {
      if (!in[j]) continue;
      const int lr = kH + ty + kWY * j, lc = kH + tx;
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
      c.idx = zo + row[j];
      c.b = fminf(bxy[j], bzv);
      op.cell(c, lap, ctr);
    }
    __syncthreads();
#pragma unroll
    for (int f = 0; f < NF; f++)
#pragma unroll
      for (int j = 0; j < kRY; j++)
#pragma unroll
        for (int k = 0; k < 2 * kH; k++) q[f][j][k] = q[f][j][k + 1];
  }
}

// Uniform per-block source-point mask: bit i set when point i lies in the
// block's cells.  pz_mask (per plane) narrows it to points on plane z.
struct SrcInBlock {
  uint32_t mask;
  __device__ void init(const PtList &pts, int x0, int x1, int y0, int y1,
                       int z0, int z1) {
    mask = 0;
    for (int i = 0; i < pts.n; i++)
      if (pts.x[i] >= x0 && pts.x[i] < x1 && pts.y[i] >= y0 &&
          pts.y[i] < y1 && pts.z[i] >= z0 && pts.z[i] < z1)
        mask |= 1u << i;
  }
  __device__ uint32_t on_plane(const PtList &pts, int z) const {
    uint32_t m = 0;
    for (uint32_t b = mask; b; b &= b - 1) {
      const int i = __builtin_ctz(b);
      if (pts.z[i] == z) m |= 1u << i;
    }
    return m;
  }
  // sum, in list order, of amp[i] for the points at (x, y) among plane mask m
  __device__ static float add(const PtList &pts, uint32_t m, int x, int y,
                              float v, const float *amp) {
    for (uint32_t b = m; b; b &= b - 1) {
      const int i = __builtin_ctz(b);
      if (pts.x[i] == x && pts.y[i] == y) v = __fadd_rn(v, amp[i]);
    }
    return v;
  }
};

}  // namespace born
