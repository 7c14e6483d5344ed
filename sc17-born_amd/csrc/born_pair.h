// Two leapfrog steps per launch for the FP32 Born forward (temporal blocking).
//
// A single-step sweep reads the current and old levels of s and r, vel and
// model and writes the new levels: 24 B read and 8-12 B written per cell and
// step.  This kernel advances s and r by two steps in one z sweep: levels k-1
// and k (and vel, model) are read once, levels k+1 and k+2 are both written,
// so a cell costs about 24 + 16 B per two steps.
//
// Geometry of a block.  A wavefront row of 64 lanes covers 64 "extended"
// columns; the tile proper is the middle 56 (x0+4 .. x0+59) and the 4 columns
// on either side are the halo that step k+2 needs of level k+1.  In y the
// tile has TY = RY*WY rows and the extended region 4 more on either side.
//   step 1 (level k+1) is computed at plane z on the whole extended region,
//     from level k staged in LDS with a further 4-cell halo (Lk);
//   step 2 (level k+2) is computed at plane z-4 on the tile, from level k+1
//     staged in LDS (L1) and held in a 9-deep register queue per tile cell.
// Every thread owns RY tile-row cells (rows 4+w+WY*j of the extended region,
// w = wavefront) and HR = 8/WY cells of the 8 halo rows, all in its own lane
// column, so the z neighbours of every cell it updates are its own registers.
// Level k+1 outside the interior is zero, as in memory.  A z chunk [z0, z1)
// computes level k+1 on planes z0-4 .. z1+3 and stores only its own planes.
//
// Per cell the arithmetic is that of FwdOp (born_ops.h) in the same
// order, so the two steps give the same values as two single-step launches.
#pragma once
#include "born_device.h"

namespace born {

constexpr int kPairTW = kTX - 2 * kH;  // tile columns per block

template <int SCATTER>
struct PairOp {
  const float *__restrict__ s_old;  // level k-1
  const float *__restrict__ s_cur;  // level k
  const float *__restrict__ r_old;
  const float *__restrict__ r_cur;
  float *__restrict__ s_a;  // level k+1
  float *__restrict__ r_a;
  float *__restrict__ s_b;  // level k+2
  float *__restrict__ r_b;
  float *__restrict__ h_a;  // optional scatter-operand history of both steps
  float *__restrict__ h_b;
  const float *__restrict__ vel;
  const float *__restrict__ model;
  Points pts;
  const float *__restrict__ amp_a;  // source amplitudes of the two steps
  const float *__restrict__ amp_b;
  RecMap rm;
  float *__restrict__ tr;
  const float *__restrict__ tab_a;  // time-sinc rows of the two steps
  const float *__restrict__ tab_b;
  int id_a, id_b, ext_a, ext_b;
};

// sum of the 25 taps in the order of march_tile: centre, x, y, z pairs
__device__ __forceinline__ float lap25(const Taps &t, float c, const float (&xs)[2 * kH + 1],
                                       const float (&up)[kH], const float (&dn)[kH],
                                       const float (&zq)[2 * kH + 1]) {
  float a = mul_r(t.c0, c);
#pragma unroll
  for (int k = 1; k <= kH; k++) a = fma_r(t.cx[k - 1], add_r(xs[kH - k], xs[kH + k]), a);
#pragma unroll
  for (int k = 1; k <= kH; k++) a = fma_r(t.cy[k - 1], add_r(up[k - 1], dn[k - 1]), a);
#pragma unroll
  for (int k = 1; k <= kH; k++) a = fma_r(t.cz[k - 1], add_r(zq[kH - k], zq[kH + k]), a);
  return a;
}

// One forward cell update (FwdOp in born_ops.h): new s, r and the
// scatter operand (cpu_prop.cc:31-53 for the propagation, :220 for scatter 0).
template <int SCATTER>
__device__ __forceinline__ void pair_cell(const PairOp<SCATTER> &op, float bb, float v,
                                          float m, float lap_s, float lap_r, float cs,
                                          float cr, float os, float orr, uint32_t pm, int x,
                                          int y, const float *amp, float &sn, float &rn,
                                          float &sig) {
  sn = leap(bb, v, lap_s, cs, os);
  if (pm) sn = add_points(op.pts, pm, x, y, sn, amp);
  rn = leap(bb, v, lap_r, cr, orr);
  born_scatter<SCATTER>(bb, lap_s, sn, m, rn, sig);
}

// receiver extraction of one cell: traces[id + k][i] += r * tab[k]
template <int SCATTER>
__device__ __forceinline__ void pair_extract(const PairOp<SCATTER> &op, int slot, int x, int y,
                                             int nx, int ny, int id, const float *tab,
                                             float rn) {
  const RecMap &rm = op.rm;
  for (int i = rm.rmap[((long long)slot * ny + y) * nx + x]; i >= 0; i = rm.rnext[i]) {
    float *q = op.tr + (long long)id * rm.nrec + i;
    float v[8];
#pragma unroll
    for (int k = 0; k < 8; k++) v[k] = q[(long long)k * rm.nrec];
#pragma unroll
    for (int k = 0; k < 8; k++) q[(long long)k * rm.nrec] = add_r(v[k], mul_r(rn, ldc(tab, k)));
  }
}

// Tile grid of the pair kernel: tiles of 56 x TY cells, chunks of zc planes.
__host__ __device__ inline Grid pair_grid(int nx, int ny, int nz, int ty, int zc) {
  Grid g;
  g.nx = nx;
  g.ny = ny;
  g.nz = nz;
  g.plane = (uint32_t)nx * ny;
  g.ntx = (nx - 2 * kH + kPairTW - 1) / kPairTW;
  g.nty = (ny - 2 * kH + ty - 1) / ty;
  g.zc = zc;
  g.ntz = (nz - 2 * kH + zc - 1) / zc;
  return g;
}

template <int SCATTER, int RY, int WY>
__global__ void __launch_bounds__(kTX *WY)
    pair_kernel(const Grid g, const Taps t, const Sponge sp, const PairOp<SCATTER> op) {
  constexpr int NF = 2;
  constexpr int NT = kTX * WY;
  constexpr int TY = RY * WY;       // tile rows
  static_assert(2 * kH % WY == 0, "the 8 halo rows are spread over the wavefronts");
  constexpr int HR = 2 * kH / WY;   // halo-row cells per thread
  constexpr int C1 = RY + HR;       // step-1 cells per thread
  constexpr int ER = TY + 2 * kH;   // extended rows
  constexpr int KR = ER + 2 * kH;   // rows of the level-k LDS plane
  constexpr int KW = kTX + 2 * kH;  // columns of the level-k LDS plane
  constexpr int NQ = 2 * kH + 1;
  constexpr int NOH = 2 * kH * KW + ER * 2 * kH;  // outer halo cells of a plane
  constexpr int HPT = (NOH + NT - 1) / NT;
  __shared__ __align__(16) float Lk[NF][KR][KW];
  __shared__ __align__(16) float L1[NF][ER][kTX];

  const int L = xcd_remap((int)blockIdx.x, (int)gridDim.x);
  const int bx = L % g.ntx;
  const int rest = L / g.ntx;
  const int by = rest % g.nty;
  const int bz = rest / g.nty;
  const int tx = (int)threadIdx.x % kTX;
  const int w = (int)threadIdx.x / kTX;
  const int e0 = bx * kPairTW;  // x of extended column 0
  const int ey0 = by * TY;      // y of extended row 0
  const int z0 = kH + bz * g.zc;
  const int z1 = min(z0 + g.zc, g.nz - kH);
  const int zs = z0 - kH, ze = z1 + kH;  // planes of level k+1 computed here
  const int gx = e0 + tx;
  const bool xin = gx >= kH && gx < g.nx - kH;
  const bool xtile = xin && tx >= kH && tx < kTX - kH;
  const int gxc = min(gx, g.nx - 1);

  int er[C1];
  bool v1[C1];
  uint32_t off[C1];
  float bxy[C1];
#pragma unroll
  for (int c = 0; c < C1; c++) {
    if (c < RY) {
      er[c] = kH + w + WY * c;
    } else {
      const int hrow = w + WY * (c - RY);
      er[c] = hrow < kH ? hrow : TY + hrow;
    }
    const int gy = ey0 + er[c];
    v1[c] = xin && gy >= kH && gy < g.ny - kH;
    const int gyc = min(gy, g.ny - 1);
    off[c] = (uint32_t)gyc * g.nx + gxc;
    bxy[c] = fminf(sp.bx[gxc], sp.by[gyc]);
  }
  bool in2[RY];
#pragma unroll
  for (int j = 0; j < RY; j++) in2[j] = xtile && v1[j];

  // outer halo of level k: rows 0..3 and ER+4..ER+7 (all KW columns), then
  // columns 0..3 and 68..71 of rows 4..ER+3
  uint32_t hoff[HPT];
  short hr[HPT], hc[HPT];
  bool hv[HPT];
#pragma unroll
  for (int k = 0; k < HPT; k++) {
    const int h = (int)threadIdx.x + k * NT;
    hv[k] = h < NOH;
    const int hh = min(h, NOH - 1);
    int lr, lc;
    if (hh < 2 * kH * KW) {
      const int r = hh / KW;
      lr = r < kH ? r : ER + r;
      lc = hh % KW;
    } else {
      const int q = hh - 2 * kH * KW;
      const int c = q % (2 * kH);
      lr = kH + q / (2 * kH);
      lc = c < kH ? c : kTX + c;
    }
    hr[k] = (short)lr;
    hc[k] = (short)lc;
    const int gy = min(max(ey0 - kH + lr, 0), g.ny - 1);
    const int gxh = min(max(e0 - kH + lc, 0), g.nx - 1);
    hoff[k] = (uint32_t)gy * g.nx + gxh;
  }

  const uint32_t pm_blk =
      points_in_block(op.pts, max(e0, kH), min(e0 + kTX, g.nx - kH), max(ey0, kH),
                      min(ey0 + ER, g.ny - kH), max(zs, kH), min(ze, g.nz - kH));
  auto zp = [&](int z) { return (uint32_t)min(max(z, 0), g.nz - 1) * g.plane; };

  // Loads run two planes ahead in two register sets that alternate between
  // even and odd planes (the loop is unrolled by two), so no register is
  // copied while its load is in flight: a set is consumed (after a wait that
  // leaves the other set's loads outstanding) and refilled in the same plane.
  // Every load is unconditional (addresses are clamped), so the number of
  // memory operations issued between a set's loads and their use does not
  // depend on the path taken, and the wait counts stay exact.
  struct Set {
    float lead[NF][C1];  // level k at plane p+4 (queue lead)
    float halo[NF][HPT]; // level k at plane p, outer LDS halo
    float pre[C1][4];    // s_old, r_old, vel, model at plane p
  };
  const float *cur[NF] = {op.s_cur, op.r_cur};
  auto issue = [&](Set &S, int p) {
    const uint32_t zl = zp(p + kH), zc = zp(p);
#pragma unroll
    for (int f = 0; f < NF; f++) {
#pragma unroll
      for (int c = 0; c < C1; c++) S.lead[f][c] = at(cur[f], zl + off[c]);
#pragma unroll
      for (int k = 0; k < HPT; k++) S.halo[f][k] = at(cur[f], zc + hoff[k]);
    }
#pragma unroll
    for (int c = 0; c < C1; c++) {
      S.pre[c][0] = at(op.s_old, zc + off[c]);
      S.pre[c][1] = at(op.r_old, zc + off[c]);
      S.pre[c][2] = at(op.vel, zc + off[c]);
      S.pre[c][3] = at(op.model, zc + off[c]);
    }
  };

  float qk[NF][C1][NQ];      // level k, planes z-4 .. z+4
  float q1[NF][RY][NQ];      // level k+1 of the tile cells, planes z-8 .. z
  float d1[NF][HR][kH + 1];  // level k+1 of the halo-row cells, planes z-4 .. z
  float vm[RY][2][kH + 1];   // vel, model of the tile cells, planes z-4 .. z
  float pre[C1][4];          // the pointwise inputs of plane z

#pragma unroll
  for (int f = 0; f < NF; f++) {
#pragma unroll
    for (int c = 0; c < C1; c++)
#pragma unroll
      for (int i = 0; i < 2 * kH; i++) qk[f][c][i] = at(cur[f], zp(zs - kH + i) + off[c]);
#pragma unroll
    for (int j = 0; j < RY; j++)
#pragma unroll
      for (int i = 0; i < NQ; i++) q1[f][j][i] = 0.f;
#pragma unroll
    for (int h = 0; h < HR; h++)
#pragma unroll
      for (int i = 0; i <= kH; i++) d1[f][h][i] = 0.f;
  }
#pragma unroll
  for (int j = 0; j < RY; j++)
#pragma unroll
    for (int i = 0; i <= kH; i++) vm[j][0][i] = vm[j][1][i] = 0.f;
  Set A, B;
  issue(A, zs);
  issue(B, zs + 1);

  // stores of a plane are issued in the next plane, after its LDS staging
  // and before its loads
  float pa[RY][3], pb[RY][3];
  int za = -1, zb = -1;
  auto flush = [&]() {
    if (za >= 0) {
      const uint32_t o = (uint32_t)za * g.plane;
#pragma unroll
      for (int j = 0; j < RY; j++)
        if (in2[j]) {
          at(op.s_a, o + off[j]) = pa[j][0];
          at(op.r_a, o + off[j]) = pa[j][1];
          if (op.h_a) at(op.h_a, o + off[j]) = pa[j][2];
        }
    }
    if (zb >= 0) {
      const uint32_t o = (uint32_t)zb * g.plane;
#pragma unroll
      for (int j = 0; j < RY; j++)
        if (in2[j]) {
          at(op.s_b, o + off[j]) = pb[j][0];
          at(op.r_b, o + off[j]) = pb[j][1];
          if (op.h_b) at(op.h_b, o + off[j]) = pb[j][2];
        }
    }
  };

  // One plane: level k+1 at plane z, level k+2 at plane z-4.  Planes past
  // the chunk's last (an odd count is padded to even) only keep the loads
  // and barriers uniform.
  auto plane = [&](Set &S, int z) {
    // stage level k at plane z and level k+1 at plane z-4
#pragma unroll
    for (int f = 0; f < NF; f++) {
#pragma unroll
      for (int c = 0; c < C1; c++) {
        qk[f][c][2 * kH] = S.lead[f][c];
        Lk[f][er[c] + kH][tx + kH] = qk[f][c][kH];
      }
#pragma unroll
      for (int k = 0; k < HPT; k++)
        if (hv[k]) Lk[f][hr[k]][hc[k]] = S.halo[f][k];
#pragma unroll
      for (int j = 0; j < RY; j++) L1[f][er[j]][tx] = q1[f][j][kH];
#pragma unroll
      for (int h = 0; h < HR; h++) L1[f][er[RY + h]][tx] = d1[f][h][0];
    }
#pragma unroll
    for (int c = 0; c < C1; c++)
#pragma unroll
      for (int k = 0; k < 4; k++) pre[c][k] = S.pre[c][k];
    flush();
    // the set's old values are dead before its new loads are issued: keeps
    // the scheduler from hoisting the loads above their last use, so the
    // loads land in the set's own registers (no copy, and no wait, at the
    // loop's back edge)
    __builtin_amdgcn_sched_barrier(0);
    issue(S, z + 2);
    __syncthreads();

    // step 1: level k+1 at plane z on the extended region
    float n1[NF][C1];
#pragma unroll
    for (int f = 0; f < NF; f++)
#pragma unroll
      for (int c = 0; c < C1; c++) n1[f][c] = 0.f;
    za = -1;
    if (z >= kH && z < g.nz - kH && z < ze) {
      const uint32_t pm = points_on_plane(op.pts, pm_blk, z);
      const int slot = ldc(op.rm.plane_slot, z);
      const float bzv = ldc(sp.bz, z);
      const bool own = z >= z0 && z < z1;
#pragma unroll
      for (int c = 0; c < C1; c++) {
        if (!v1[c]) continue;
        const int lr = er[c] + kH, lc = tx + kH;
        float lap[NF];
#pragma unroll
        for (int f = 0; f < NF; f++) {
          float xs[NQ], up[kH], dn[kH], zq[NQ];
#pragma unroll
          for (int i = 0; i < NQ; i++) xs[i] = Lk[f][lr][lc - kH + i];
#pragma unroll
          for (int k = 1; k <= kH; k++) {
            up[k - 1] = Lk[f][lr - k][lc];
            dn[k - 1] = Lk[f][lr + k][lc];
          }
#pragma unroll
          for (int i = 0; i < NQ; i++) zq[i] = qk[f][c][i];
          lap[f] = lap25(t, qk[f][c][kH], xs, up, dn, zq);
        }
        float sn, rn, sig;
        pair_cell<SCATTER>(op, fminf(bxy[c], bzv), pre[c][2], pre[c][3], lap[0], lap[1],
                           qk[0][c][kH], qk[1][c][kH], pre[c][0], pre[c][1], pm, gx,
                           ey0 + er[c], op.amp_a, sn, rn, sig);
        n1[0][c] = sn;
        n1[1][c] = rn;
        if (c < RY) {
          pa[c][0] = sn;
          pa[c][1] = rn;
          pa[c][2] = sig;
          if (own && slot >= 0 && op.ext_a && in2[c])
            pair_extract(op, slot, gx, ey0 + er[c], g.nx, g.ny, op.id_a, op.tab_a, rn);
        }
      }
      if (own) za = z;
    }
#pragma unroll
    for (int f = 0; f < NF; f++) {
#pragma unroll
      for (int j = 0; j < RY; j++) q1[f][j][2 * kH] = n1[f][j];
#pragma unroll
      for (int h = 0; h < HR; h++) d1[f][h][kH] = n1[f][RY + h];
    }
#pragma unroll
    for (int j = 0; j < RY; j++) {
      vm[j][0][kH] = pre[j][2];
      vm[j][1][kH] = pre[j][3];
    }

    // step 2: level k+2 at plane z-4 on the tile
    const int zz = z - kH;
    zb = -1;
    if (zz >= z0 && zz < z1) {
      const uint32_t pm = points_on_plane(op.pts, pm_blk, zz);
      const int slot = ldc(op.rm.plane_slot, zz);
      const float bzv = ldc(sp.bz, zz);
#pragma unroll
      for (int j = 0; j < RY; j++) {
        if (!in2[j]) continue;
        const int lr = er[j], lc = tx;
        float lap[NF];
#pragma unroll
        for (int f = 0; f < NF; f++) {
          float xs[NQ], up[kH], dn[kH], zq[NQ];
#pragma unroll
          for (int i = 0; i < NQ; i++) xs[i] = L1[f][lr][lc - kH + i];
#pragma unroll
          for (int k = 1; k <= kH; k++) {
            up[k - 1] = L1[f][lr - k][lc];
            dn[k - 1] = L1[f][lr + k][lc];
          }
#pragma unroll
          for (int i = 0; i < NQ; i++) zq[i] = q1[f][j][i];
          lap[f] = lap25(t, q1[f][j][kH], xs, up, dn, zq);
        }
        float sn, rn, sig;
        pair_cell<SCATTER>(op, fminf(bxy[j], bzv), vm[j][0][0], vm[j][1][0], lap[0], lap[1],
                           q1[0][j][kH], q1[1][j][kH], qk[0][j][0], qk[1][j][0], pm, gx,
                           ey0 + er[j], op.amp_b, sn, rn, sig);
        pb[j][0] = sn;
        pb[j][1] = rn;
        pb[j][2] = sig;
        if (slot >= 0 && op.ext_b)
          pair_extract(op, slot, gx, ey0 + er[j], g.nx, g.ny, op.id_b, op.tab_b, rn);
      }
      zb = zz;
    }
    __syncthreads();

#pragma unroll
    for (int f = 0; f < NF; f++) {
#pragma unroll
      for (int c = 0; c < C1; c++)
#pragma unroll
        for (int i = 0; i < 2 * kH; i++) qk[f][c][i] = qk[f][c][i + 1];
#pragma unroll
      for (int j = 0; j < RY; j++)
#pragma unroll
        for (int i = 0; i < 2 * kH; i++) q1[f][j][i] = q1[f][j][i + 1];
#pragma unroll
      for (int h = 0; h < HR; h++)
#pragma unroll
        for (int i = 0; i < kH; i++) d1[f][h][i] = d1[f][h][i + 1];
    }
#pragma unroll
    for (int j = 0; j < RY; j++)
#pragma unroll
      for (int i = 0; i < kH; i++) {
        vm[j][0][i] = vm[j][0][i + 1];
        vm[j][1][i] = vm[j][1][i + 1];
      }
  };

  for (int z = zs; z < ze; z += 2) {
    plane(A, z);
    plane(B, z + 1);
  }
  flush();
}

}  // namespace born
