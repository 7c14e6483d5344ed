// Two adjoint-state steps per launch for the FP32 exact adjoint.
//
// The exact adjoint advances nu = V B lambda backwards with the forward's
// leapfrog form, injects the receiver residuals and accumulates the image
// acc += sigma_k nu (LamOp in born_ops.h).  One step reads nu (cur,
// old), vel, sigma and the image and writes nu and the image: 28 B per
// cell.  This kernel takes steps k and k-1 in one z sweep with the geometry
// of pair_kernel (born_pair.h): level j+1 of nu on the tile widened by the
// stencil radius at plane z, level j+2 on the tile at plane z-4.  The image
// is read once and written once for both steps, so a cell costs about
// 4 + 4 + 4 + 8 + 4 (read nu cur/old, vel, two sigma, image) + 12 (write two
// nu levels and the image) B per two steps.
//
// Per cell the arithmetic is LamOp's in the same order (leap(), lam_inject(),
// fma_r for the image), so the two steps give the bits of two single steps.
#pragma once
#include "born_pair.h"

namespace born {

struct LamPairOp {
  const float *__restrict__ u_old;  // level j-1
  const float *__restrict__ u_cur;  // level j
  float *__restrict__ u_a;          // level j+1 (step k)
  float *__restrict__ u_b;          // level j+2 (step k-1)
  const float *__restrict__ vel;
  const float *__restrict__ sig_a;  // sigma_k
  const float *__restrict__ sig_b;  // sigma_{k-1}
  float *__restrict__ acc;          // image accumulator
  RecMap rm;
  const float *__restrict__ tr;     // residual traces, time-major
  const float *__restrict__ tab_a;  // time-sinc rows of the two steps
  const float *__restrict__ tab_b;
  int id_a, id_b, inj_a, inj_b;
};

template <int RY, int WY>
__global__ void __launch_bounds__(kTX *WY)
    lam_pair_kernel(const Grid g, const Taps t, const Sponge sp, const LamPairOp op) {
  constexpr int NT = kTX * WY;
  constexpr int TY = RY * WY;
  static_assert(2 * kH % WY == 0, "the 8 halo rows are spread over the wavefronts");
  constexpr int HR = 2 * kH / WY;
  constexpr int C1 = RY + HR;
  constexpr int ER = TY + 2 * kH;
  constexpr int KR = ER + 2 * kH;
  constexpr int KW = kTX + 2 * kH;
  constexpr int NQ = 2 * kH + 1;
  constexpr int NOH = 2 * kH * KW + ER * 2 * kH;
  constexpr int HPT = (NOH + NT - 1) / NT;
  __shared__ __align__(16) float Lk[KR][KW];
  __shared__ __align__(16) float L1[ER][kTX];

  const int L = xcd_remap((int)blockIdx.x, (int)gridDim.x);
  const int bx = L % g.ntx;
  const int rest = L / g.ntx;
  const int by = rest % g.nty;
  const int bz = rest / g.nty;
  const int tx = (int)threadIdx.x % kTX;
  const int w = (int)threadIdx.x / kTX;
  const int e0 = bx * kPairTW;
  const int ey0 = by * TY;
  const int z0 = kH + bz * g.zc;
  const int z1 = min(z0 + g.zc, g.nz - kH);
  const int zs = z0 - kH, ze = z1 + kH;
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
  auto zp = [&](int z) { return (uint32_t)min(max(z, 0), g.nz - 1) * g.plane; };

  // loads two planes ahead in two alternating register sets (pair_kernel);
  // the tile cells' step-2 inputs of plane p-4 ride in the set of plane p
  struct Set {
    float lead[C1];     // nu level j at plane p+4
    float halo[HPT];    // nu level j at plane p, outer LDS halo
    float pre[C1][2];   // nu level j-1, vel at plane p
    float sa[RY], ac[RY];  // sigma_k and the image at plane p (tile cells)
    float sb[RY], vb[RY];  // sigma_{k-1} and vel at plane p-4 (tile cells)
  };
  auto issue = [&](Set &S, int p) {
    const uint32_t zl = zp(p + kH), zc = zp(p), z4 = zp(p - kH);
#pragma unroll
    for (int c = 0; c < C1; c++) S.lead[c] = at(op.u_cur, zl + off[c]);
#pragma unroll
    for (int k = 0; k < HPT; k++) S.halo[k] = at(op.u_cur, zc + hoff[k]);
#pragma unroll
    for (int c = 0; c < C1; c++) {
      S.pre[c][0] = at(op.u_old, zc + off[c]);
      S.pre[c][1] = at(op.vel, zc + off[c]);
    }
#pragma unroll
    for (int j = 0; j < RY; j++) {
      S.sa[j] = at(op.sig_a, zc + off[j]);
      S.ac[j] = at(op.acc, zc + off[j]);
      S.sb[j] = at(op.sig_b, z4 + off[j]);
      S.vb[j] = at(op.vel, z4 + off[j]);
    }
  };

  float qk[C1][NQ];      // level j, planes z-4 .. z+4
  float q1[RY][NQ];      // level j+1 of the tile cells, planes z-8 .. z
  float d1[HR][kH + 1];  // level j+1 of the halo-row cells, planes z-4 .. z
  float aq[RY][kH + 1];  // image after step k of the tile cells, planes z-4 .. z
  float pre[C1][2], sa[RY], ac[RY], sb[RY], vb[RY];

#pragma unroll
  for (int c = 0; c < C1; c++)
#pragma unroll
    for (int i = 0; i < 2 * kH; i++) qk[c][i] = at(op.u_cur, zp(zs - kH + i) + off[c]);
#pragma unroll
  for (int j = 0; j < RY; j++)
#pragma unroll
    for (int i = 0; i < NQ; i++) q1[j][i] = 0.f;
#pragma unroll
  for (int h = 0; h < HR; h++)
#pragma unroll
    for (int i = 0; i <= kH; i++) d1[h][i] = 0.f;
#pragma unroll
  for (int j = 0; j < RY; j++)
#pragma unroll
    for (int i = 0; i <= kH; i++) aq[j][i] = 0.f;
  Set A, B;
  issue(A, zs);
  issue(B, zs + 1);

  float pa[RY], pb[RY][2];  // stores: nu level j+1; nu level j+2 and image
  int za = -1, zb = -1;
  auto flush = [&]() {
    if (za >= 0) {
      const uint32_t o = (uint32_t)za * g.plane;
#pragma unroll
      for (int j = 0; j < RY; j++)
        if (in2[j]) at(op.u_a, o + off[j]) = pa[j];
    }
    if (zb >= 0) {
      const uint32_t o = (uint32_t)zb * g.plane;
#pragma unroll
      for (int j = 0; j < RY; j++)
        if (in2[j]) {
          at(op.u_b, o + off[j]) = pb[j][0];
          at(op.acc, o + off[j]) = pb[j][1];
        }
    }
  };

  auto plane = [&](Set &S, int z) {
#pragma unroll
    for (int c = 0; c < C1; c++) {
      qk[c][2 * kH] = S.lead[c];
      Lk[er[c] + kH][tx + kH] = qk[c][kH];
    }
#pragma unroll
    for (int k = 0; k < HPT; k++)
      if (hv[k]) Lk[hr[k]][hc[k]] = S.halo[k];
#pragma unroll
    for (int j = 0; j < RY; j++) L1[er[j]][tx] = q1[j][kH];
#pragma unroll
    for (int h = 0; h < HR; h++) L1[er[RY + h]][tx] = d1[h][0];
#pragma unroll
    for (int c = 0; c < C1; c++) {
      pre[c][0] = S.pre[c][0];
      pre[c][1] = S.pre[c][1];
    }
#pragma unroll
    for (int j = 0; j < RY; j++) {
      sa[j] = S.sa[j];
      ac[j] = S.ac[j];
      sb[j] = S.sb[j];
      vb[j] = S.vb[j];
    }
    flush();
    // the set's old values are dead before its new loads are issued
    __builtin_amdgcn_sched_barrier(0);
    issue(S, z + 2);
    __syncthreads();

    // step k: level j+1 at plane z on the extended region
    float n1[C1];
#pragma unroll
    for (int c = 0; c < C1; c++) n1[c] = 0.f;
    za = -1;
    if (z >= kH && z < g.nz - kH && z < ze) {
      const int slot = ldc(op.rm.plane_slot, z);
      const float bzv = ldc(sp.bz, z);
      const bool own = z >= z0 && z < z1;
#pragma unroll
      for (int c = 0; c < C1; c++) {
        if (!v1[c]) continue;
        const int lr = er[c] + kH, lc = tx + kH;
        float xs[NQ], up[kH], dn[kH], zq[NQ];
#pragma unroll
        for (int i = 0; i < NQ; i++) xs[i] = Lk[lr][lc - kH + i];
#pragma unroll
        for (int k = 1; k <= kH; k++) {
          up[k - 1] = Lk[lr - k][lc];
          dn[k - 1] = Lk[lr + k][lc];
        }
#pragma unroll
        for (int i = 0; i < NQ; i++) zq[i] = qk[c][i];
        const float lap = lap25(t, qk[c][kH], xs, up, dn, zq);
        const float bb = fminf(bxy[c], bzv), v = pre[c][1];
        float un = leap(bb, v, lap, qk[c][kH], pre[c][0]);
        if (slot >= 0 && op.inj_a)
          un = lam_inject(op.rm, slot, gx, ey0 + er[c], g.nx, g.ny, op.tr, op.tab_a, op.id_a, v,
                          bb, un);
        n1[c] = un;
        if (c < RY) {
          pa[c] = un;
          aq[c][kH] = fma_r(sa[c], un, ac[c]);
        }
      }
      if (own) za = z;
    }
#pragma unroll
    for (int j = 0; j < RY; j++) q1[j][2 * kH] = n1[j];
#pragma unroll
    for (int h = 0; h < HR; h++) d1[h][kH] = n1[RY + h];

    // step k-1: level j+2 at plane z-4 on the tile
    const int zz = z - kH;
    zb = -1;
    if (zz >= z0 && zz < z1) {
      const int slot = ldc(op.rm.plane_slot, zz);
      const float bzv = ldc(sp.bz, zz);
#pragma unroll
      for (int j = 0; j < RY; j++) {
        if (!in2[j]) continue;
        const int lr = er[j], lc = tx;
        float xs[NQ], up[kH], dn[kH], zq[NQ];
#pragma unroll
        for (int i = 0; i < NQ; i++) xs[i] = L1[lr][lc - kH + i];
#pragma unroll
        for (int k = 1; k <= kH; k++) {
          up[k - 1] = L1[lr - k][lc];
          dn[k - 1] = L1[lr + k][lc];
        }
#pragma unroll
        for (int i = 0; i < NQ; i++) zq[i] = q1[j][i];
        const float lap = lap25(t, q1[j][kH], xs, up, dn, zq);
        const float bb = fminf(bxy[j], bzv), v = vb[j];
        float un = leap(bb, v, lap, q1[j][kH], qk[j][0]);
        if (slot >= 0 && op.inj_b)
          un = lam_inject(op.rm, slot, gx, ey0 + er[j], g.nx, g.ny, op.tr, op.tab_b, op.id_b, v,
                          bb, un);
        pb[j][0] = un;
        pb[j][1] = fma_r(sb[j], un, aq[j][0]);
      }
      zb = zz;
    }
    __syncthreads();

#pragma unroll
    for (int c = 0; c < C1; c++)
#pragma unroll
      for (int i = 0; i < 2 * kH; i++) qk[c][i] = qk[c][i + 1];
#pragma unroll
    for (int j = 0; j < RY; j++) {
#pragma unroll
      for (int i = 0; i < 2 * kH; i++) q1[j][i] = q1[j][i + 1];
#pragma unroll
      for (int i = 0; i < kH; i++) aq[j][i] = aq[j][i + 1];
    }
#pragma unroll
    for (int h = 0; h < HR; h++)
#pragma unroll
      for (int i = 0; i < kH; i++) d1[h][i] = d1[h][i + 1];
  };

  for (int z = zs; z < ze; z += 2) {
    plane(A, z);
    plane(B, z + 1);
  }
  flush();
}

}  // namespace born
