// Fused time-step ops of the Born propagator (gfx950) and the templates
// that launch them; the launchers are instantiated in the born_k_*.hip
// translation units (compiled in parallel) and dispatched in born_kernels.hip.
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
#pragma once
#include <type_traits>

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

// Scaled storage (FP16) shares the state of the ops (unused with float and
// 24-bit storage): each stored level has a power-of-two scale derived from a
// bound on |new level|, gain * max|cur| + max|old| + injections, with
// gain = 2 + max(V) * tapsum and tapsum >= |L u| / max|u|.
template <class T>
constexpr bool kScaled = std::is_same<T, __half>::value;
// 24-bit storage (FP32 rounded to 24 bits)
template <class T>
constexpr bool kSplit = std::is_same<T, Split>::value;
// levels narrower than FP32 (FP16 or 24-bit): the forward stores s and r
// together, the steps run in pairs, the streams take the FP16 cache hints
template <class T>
constexpr bool kNarrow = !std::is_same<T, float>::value;
// two fields stored together: FP16 as __half2, 24-bit as SplitPair
template <class T>
using PairOf = typename std::conditional<kSplit<T>, SplitPair, __half2>::type;

struct Blk1 {  // one stored field
  PtsBlk pb;
  float ic, io, sn, sh;  // 1/scale of cur and old, scale of new and of sigma
};
struct Blk2 {  // two stored fields
  PtsBlk pb;
  float ic[2], io[2], sn[2], sh;
};

__device__ __forceinline__ float amp_sum(const Points &pts, const float *amp) {
  float a = 0.f;
  for (int i = 0; i < pts.n; i++) a += fabsf(ldc(amp, i));
  return a;
}

// ---------------------------------------------------------------- forward
// NT: non-temporal hint on the stores and FP32 pointwise loads (the shot
// batch runs without it, see BatchFwd)
template <int SCATTER, class T, bool NT = true>
struct FwdOp {
  static constexpr bool H = kScaled<T>;
  static constexpr int NF = 2, NPRE = 4, NOUT = 3, NMAX = H ? 2 : 0;
  // FP16 and 24-bit levels hold s and r together per cell (sr1: level k,
  // sr0: level k-1, srn: level k+1, the same buffer as sr0 in place)
  static constexpr bool PAIRED = kNarrow<T>;
  static constexpr bool Q2 = kNarrow<T>;  // steps in pairs (march_tile)
  static constexpr bool DAMP = true;
  const T *__restrict__ s1;
  const T *__restrict__ r1;
  // old s level (read) and new s level (written): the same buffer (in
  // place), or with scatter 0 and a kept history two levels of the history,
  // whose level k is s of level k+1
  const T *s0;
  T *s_new;
  T *__restrict__ r0;
  const PairOf<T> *sr1;
  const PairOf<T> *sr0;
  PairOf<T> *srn;
  const float *__restrict__ vel;
  const float *__restrict__ model;
  // NT (the single-shot forward): vel and model interleaved per cell, one
  // load and one stream for both
  const float2 *__restrict__ vm;
  T *__restrict__ hist;  // optional scatter-operand history
  Points pts;
  const float *__restrict__ amp;  // source amplitudes of this step [npts]
  RecMap rm;
  float *__restrict__ tr;  // traces, time-major [ntb][nrec]
  const float *__restrict__ tab;
  int id, extract, nx, ny;
  uint32_t nlev;  // cells of a level (24-bit storage: the lo bytes follow the hi values)
  // FP16 scale state in as few scalar registers as possible (the loop needs
  // none of it, but kernel arguments stay live in scalar registers, and
  // too many of them spill into vector lanes that every use reloads): the
  // ring block rg holds the s ring at 0, the r ring at 8 and max |model| at
  // 16 (float bits); slots packs cur, old, nxt, clr in its four bytes.
  float *rg;
  int slots;
  float gain, tapsum;
  float *hscale;  // FP16: scale of this step's history level

  __device__ Ring ring(int f) const { return {rg + 8 * f, (unsigned *)(rg + 8 * f + 4)}; }
  __device__ int slot(int i) const { return (slots >> (8 * i)) & 0xff; }

  __device__ Blk2 block(int x0, int x1, int y0, int y1, int z0, int z1) const {
    Blk2 b;
    b.pb = points_of_block(pts, x0, x1, y0, y1, z0, z1);
    for (int f = 0; f < 2; f++) b.ic[f] = b.io[f] = b.sn[f] = 1.f;
    b.sh = 1.f;
    if constexpr (H) {
      const int cur = slot(0), old = slot(1);
      for (int f = 0; f < 2; f++) {
        b.ic[f] = 1.f / ring(f).scale[cur];
        b.io[f] = 1.f / ring(f).scale[old];
      }
      const float ms = ring_max(ring(0), cur);
      const float bs = gain * ms + ring_max(ring(0), old) + amp_sum(pts, amp);
      const float mm = __uint_as_float(*(const unsigned *)(rg + 16));
      const float sig = SCATTER == 0 ? bs : tapsum * ms;
      const float br = gain * ring_max(ring(1), cur) + ring_max(ring(1), old) +
                       (SCATTER == 0 ? 1e-6f * bs * mm : mm * sig);
      b.sn[0] = level_scale(bs);
      b.sn[1] = level_scale(br);
      b.sh = SCATTER == 0 ? b.sn[0] : level_scale(sig);
    }
    return b;
  }
  __device__ void first(const Blk2 &b) const {
    for (int f = 0; f < 2; f++) {
      ring(f).scale[slot(2)] = b.sn[f];
      ring(f).maxb[slot(3)] = 0u;
    }
    if (hscale) *hscale = b.sh;
  }
  __device__ void publish_max(int f, float m) const {
    atomicMax(&ring(f).maxb[slot(2)], __float_as_uint(m));
  }
  template <int VX>
  __device__ void ldv(const Blk2 &b, int f, uint32_t o, float (&v)[VX]) const {
    ld_vec<VX>(f == 0 ? s1 : r1, o, b.ic[f], v, nlev);
  }
  template <int VX>
  __device__ void ldv2(const Blk2 &b, uint32_t o, float (&v)[2][VX]) const {
    ld_pair<VX>(sr1, o, b.ic, v, nlev);
  }
  __device__ SrcPln plane(const Blk2 &b, int z) const {
    return {points_on_plane(pts, b.pb, z), ldc(rm.plane_slot, z)};
  }
  template <int VX>
  __device__ void prefetchv(const Blk2 &b, float (&p)[NPRE][VX], uint32_t o) const {
    if constexpr (PAIRED) {
      float sr[2][VX];
      ld_pair<VX>(sr0, o, b.io, sr, nlev);
#pragma unroll
      for (int v = 0; v < VX; v++) {
        p[0][v] = sr[0][v];
        p[1][v] = sr[1][v];
      }
    } else if constexpr (NT) {
      ld_stream<VX>(s0, o, b.io[0], p[0], nlev);
      ld_stream<VX>(r0, o, b.io[1], p[1], nlev);
    } else {
      ld_vec<VX>(s0, o, b.io[0], p[0], nlev);
      ld_vec<VX>(r0, o, b.io[1], p[1], nlev);
    }
    if constexpr (NT) {
      const char *q = (const char *)vm + (uint64_t)o * 8u;
      if constexpr (VX == 1) {
        typedef float f2 __attribute__((ext_vector_type(2)));
        const f2 w = __builtin_nontemporal_load((const f2 *)q);
        p[2][0] = w.x;
        p[3][0] = w.y;
      } else {
        typedef float f4 __attribute__((ext_vector_type(4)));
        const f4 w = __builtin_nontemporal_load((const f4 *)q);
        p[2][0] = w.x;
        p[3][0] = w.y;
        p[2][1] = w.z;
        p[3][1] = w.w;
      }
    } else {
      ld_vec<VX>(vel, o, 1.f, p[2]);
      ld_vec<VX>(model, o, 1.f, p[3]);
    }
  }
  template <int VX>
  __device__ void storev(const Blk2 &b, uint32_t o, const float (&out)[NOUT][VX]) const {
    if constexpr (PAIRED) {
      st_pair<VX, NT>(srn, o, out[0], out[1], b.sn, nlev);
    } else {
      st_vec<VX, NT>(s_new, o, out[0], b.sn[0], nlev);
      st_vec<VX, NT>(r0, o, out[1], b.sn[1], nlev);
    }
    if (hist) st_vec<VX, NT>(hist, o, out[2], b.sh, nlev);
  }
  __device__ void cell(const Blk2 &b, const SrcPln &pl, const CellArgs &c,
                       const float *lap, const float *ctr, const float *p,
                       float (&out)[NOUT], float (&tmax)[NF]) const {
    const float v = p[2], bb = c.b;
    float sn = leap(bb, v, lap[0], ctr[0], p[0]);
    if (pl.pm) sn = add_points(pts, pl.pm, c.x, c.y, sn, amp);
    float rn = leap(bb, v, lap[1], ctr[1], p[1]);
    float sig;
    if constexpr (kNarrow<T> && SCATTER == 0) {
      // FP16 and 24-bit levels: the 1e-6 s m term in FP32.  FP32 levels
      // round it in FP64 as the reference does (bitwise parity); the
      // narrow levels' own rounding (2^-11, 2^-16) is far above FP32's.
      rn = fma_r(mul_r(1e-6f, sn), p[3], rn);
      sig = sn;
    } else {
      born_scatter<SCATTER>(bb, lap[0], sn, p[3], rn, sig);
    }
    out[0] = sn;
    out[1] = rn;
    out[2] = sig;
    if constexpr (H) {
      tmax[0] = fmaxf(tmax[0], fabsf(sn));
      tmax[1] = fmaxf(tmax[1], fabsf(rn));
    }
    if (pl.slot >= 0 && extract) {
      // the 8 samples of a receiver are distinct, but the compiler cannot
      // tell; all loads first, then all stores, one wait instead of eight
      for (int i = rec_first(rm, pl.slot, c.x, c.y, nx, ny); i >= 0; i = rm.rnext[i]) {
        float *q = tr + (long long)id * rm.nrec + i;
        float v[8];
#pragma unroll
        for (int k = 0; k < 8; k++) v[k] = q[(long long)k * rm.nrec];
#pragma unroll
        for (int k = 0; k < 8; k++)
          q[(long long)k * rm.nrec] = add_r(v[k], mul_r(rn, ldc(tab, k)));
      }
    }
  }
};

// ------------------------------------------------------------- sourceProp
struct SrcOp {
  static constexpr int NF = 1, NPRE = 2, NOUT = 1, NMAX = 0;
  static constexpr bool PAIRED = false;
  static constexpr bool Q2 = false;
  static constexpr bool DAMP = false;
  const float *__restrict__ s1;
  float *__restrict__ s0;
  const float *__restrict__ vel;
  Points pts;
  const float *__restrict__ amp;
  __device__ uint32_t block(int x0, int x1, int y0, int y1, int z0, int z1) const {
    return points_in_block(pts, x0, x1, y0, y1, z0, z1);
  }
  template <int VX>
  __device__ void ldv(uint32_t, int, uint32_t o, float (&v)[VX]) const {
    ld_vec<VX>(s1, o, 1.f, v);
  }
  __device__ uint32_t plane(uint32_t blk, int z) const {
    return points_on_plane(pts, blk, z);
  }
  template <int VX>
  __device__ void prefetchv(uint32_t, float (&p)[NPRE][VX], uint32_t o) const {
    ld_vec<VX>(s0, o, 1.f, p[0]);
    ld_vec<VX>(vel, o, 1.f, p[1]);
  }
  template <int VX>
  __device__ void storev(uint32_t, uint32_t o, const float (&out)[NOUT][VX]) const {
    st_vec<VX>(s0, o, out[0], 1.f);
  }
  __device__ void cell(uint32_t, uint32_t pm, const CellArgs &c, const float *lap,
                       const float *ctr, const float *p, float (&out)[NOUT],
                       float (&)[NF]) const {
    float sn = p[1] * lap[0] + ctr[0] + ctr[0] - p[0];
    if (pm) sn = add_points(pts, pm, c.x, c.y, sn, amp);
    out[0] = sn;
  }
};

// ------------------------------------------------------ reference adjoint
struct RadjOp {
  static constexpr int NF = 2, NPRE = 5, NOUT = 3, NMAX = 0;
  static constexpr bool PAIRED = false;
  static constexpr bool Q2 = false;
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

  __device__ uint32_t block(int x0, int x1, int y0, int y1, int z0, int z1) const {
    return pre ? points_in_block(pts, x0, x1, y0, y1, z0, z1) : 0u;
  }
  template <int VX>
  __device__ void ldv(uint32_t, int f, uint32_t o, float (&v)[VX]) const {
    ld_vec<VX>(f == 0 ? s1 : r1, o, 1.f, v);
  }
  __device__ SrcPln plane(uint32_t blk, int z) const {
    return {points_on_plane(pts, blk, z), ldc(rm.plane_slot, z)};
  }
  template <int VX>
  __device__ void prefetchv(uint32_t, float (&p)[NPRE][VX], uint32_t o) const {
    ld_vec<VX>(s0, o, 1.f, p[0]);
    ld_vec<VX>(r0, o, 1.f, p[1]);
    ld_vec<VX>(vel1, o, 1.f, p[2]);
    ld_vec<VX>(vel2, o, 1.f, p[3]);
    ld_vec<VX>(img, o, 1.f, p[4]);
  }
  template <int VX>
  __device__ void storev(uint32_t, uint32_t o, const float (&out)[NOUT][VX]) const {
    if (s_prop) st_vec<VX>(s0, o, out[0], 1.f);
    st_vec<VX>(r0, o, out[1], 1.f);
    st_vec<VX>(img, o, out[2], 1.f);
  }
  __device__ void cell(uint32_t, const SrcPln &pl, const CellArgs &c, const float *lap,
                       const float *ctr, const float *p, float (&out)[NOUT],
                       float (&)[NF]) const {
    float sn;
    if (s_prop) {
      float so = p[0];
      if (pl.pm) so = add_points(pts, pl.pm, c.x, c.y, so, amp_pre);
      sn = p[2] * lap[0] + ctr[0] + ctr[0] - so;
    } else {
      sn = p[0];
    }
    const float b = c.b;
    float rn = r_prop ? leap(b, p[3], lap[1], ctr[1], p[1]) : mul_r(b, mul_r(b, p[1]));
    if (pl.slot >= 0 && inject) {
      for (int i = rec_first(rm, pl.slot, c.x, c.y, nx, ny); i >= 0; i = rm.rnext[i])
        rn = add_r(rn, mul_r(sc, sinc8(tab, tr + (long long)id * rm.nrec + i, rm.nrec)));
    }
    out[0] = sn;
    out[1] = rn;
    out[2] = p[4] + sn * rn;
  }
};

// ------------------------------------- exact adjoint: background recompute
template <int SCATTER, class T>
struct SrecOp {
  static constexpr bool H = kScaled<T>;
  static constexpr int NF = 1, NPRE = 2, NOUT = 2, NMAX = H ? 1 : 0;
  static constexpr bool PAIRED = false;
  static constexpr bool Q2 = kNarrow<T>;  // steps in pairs (march_tile)
  static constexpr bool DAMP = true;
  const T *__restrict__ s1;
  // old s level (read) and new s level (written): the same buffer, or with
  // scatter 0 two levels of the segment history (level k of s is sigma of
  // step k-1)
  const T *s0;
  T *s_new;
  const float *__restrict__ vel;
  T *__restrict__ sig;  // scatter operand of this step (nullptr: none)
  Points pts;
  const float *__restrict__ amp;
  Ring ring;
  Slots sl;
  float gain, tapsum;
  float *hscale;  // FP16: scale of this step's sigma
  uint32_t nlev;  // cells of a level

  __device__ Blk1 block(int x0, int x1, int y0, int y1, int z0, int z1) const {
    Blk1 b;
    b.pb = points_of_block(pts, x0, x1, y0, y1, z0, z1);
    b.ic = b.io = b.sn = b.sh = 1.f;
    if constexpr (H) {
      b.ic = 1.f / ring.scale[sl.cur];
      b.io = 1.f / ring.scale[sl.old];
      const float ms = ring_max(ring, sl.cur);
      const float bs = gain * ms + ring_max(ring, sl.old) + amp_sum(pts, amp);
      b.sn = level_scale(bs);
      b.sh = SCATTER == 0 ? b.sn : level_scale(tapsum * ms);
    }
    return b;
  }
  __device__ void first(const Blk1 &b) const {
    ring.scale[sl.nxt] = b.sn;
    ring.maxb[sl.clr] = 0u;
    if (hscale) *hscale = b.sh;
  }
  __device__ void publish_max(int, float m) const {
    atomicMax(&ring.maxb[sl.nxt], __float_as_uint(m));
  }
  template <int VX>
  __device__ void ldv(const Blk1 &b, int, uint32_t o, float (&v)[VX]) const {
    ld_vec<VX>(s1, o, b.ic, v, nlev);
  }
  __device__ uint32_t plane(const Blk1 &b, int z) const {
    return points_on_plane(pts, b.pb, z);
  }
  template <int VX>
  __device__ void prefetchv(const Blk1 &b, float (&p)[NPRE][VX], uint32_t o) const {
    ld_stream<VX>(s0, o, b.io, p[0], nlev);
    // FP32: the velocity without the hint (114 -> 103 us per step at 304^3)
    if constexpr (kNarrow<T>)
      ld_stream<VX>(vel, o, 1.f, p[1]);
    else
      ld_vec<VX>(vel, o, 1.f, p[1]);
  }
  template <int VX>
  __device__ void storev(const Blk1 &b, uint32_t o, const float (&out)[NOUT][VX]) const {
    st_vec<VX>(s_new, o, out[0], b.sn, nlev);
    if (sig) st_vec<VX>(sig, o, out[1], b.sh, nlev);
  }
  __device__ void cell(const Blk1 &, uint32_t pm, const CellArgs &c, const float *lap,
                       const float *ctr, const float *p, float (&out)[NOUT],
                       float (&tmax)[NF]) const {
    const float bb = c.b;
    float sn = leap(bb, p[1], lap[0], ctr[0], p[0]);
    if (pm) sn = add_points(pts, pm, c.x, c.y, sn, amp);
    if constexpr (H) tmax[0] = fmaxf(tmax[0], fabsf(sn));
    out[0] = sn;
    out[1] = SCATTER == 0 ? sn : mul_r(bb, lap[0]);
  }
};

// ---------------------------------------- exact adjoint: adjoint-state step
// nu = V B lambda obeys the forward recursion with V B-weighted receiver
// injection; the image accumulates sigma * nu and is divided by V B at the end.
template <class T>
struct LamOp {
  static constexpr bool H = kScaled<T>;
  static constexpr int NF = 1, NPRE = 4, NOUT = 2, NMAX = H ? 1 : 0;
  static constexpr bool PAIRED = false;
  static constexpr bool Q2 = kNarrow<T>;  // steps in pairs (march_tile)
  static constexpr bool DAMP = true;
  const T *__restrict__ u1;
  T *__restrict__ u0;
  const float *__restrict__ vel;
  const T *__restrict__ sig;
  float *__restrict__ acc;
  RecMap rm;
  const float *__restrict__ tr;
  const float *__restrict__ tab;
  int id, inject, nx, ny;
  Ring ring;
  Slots sl;
  float gain;
  float inj;                // bound of |V B P^T w| per unit of max |trace|
  const unsigned *trmax;    // max |trace| as float bits
  const float *hscale;      // FP16: scale of sigma
  uint32_t nlev;            // cells of a level

  __device__ Blk1 block(int, int, int, int, int, int) const {
    Blk1 b;
    b.pb = PtsBlk{0u, 1, 0};
    b.ic = b.io = b.sn = b.sh = 1.f;
    if constexpr (H) {
      b.ic = 1.f / ring.scale[sl.cur];
      b.io = 1.f / ring.scale[sl.old];
      const float bu = gain * ring_max(ring, sl.cur) + ring_max(ring, sl.old) +
                       (inject ? inj * __uint_as_float(*trmax) : 0.f);
      b.sn = level_scale(bu);
      b.sh = 1.f / *hscale;
    }
    return b;
  }
  __device__ void first(const Blk1 &b) const {
    ring.scale[sl.nxt] = b.sn;
    ring.maxb[sl.clr] = 0u;
  }
  __device__ void publish_max(int, float m) const {
    atomicMax(&ring.maxb[sl.nxt], __float_as_uint(m));
  }
  template <int VX>
  __device__ void ldv(const Blk1 &b, int, uint32_t o, float (&v)[VX]) const {
    ld_vec<VX>(u1, o, b.ic, v, nlev);
  }
  __device__ int plane(const Blk1 &, int z) const { return ldc(rm.plane_slot, z); }
  template <int VX>
  __device__ void prefetchv(const Blk1 &b, float (&p)[NPRE][VX], uint32_t o) const {
    ld_stream<VX>(u0, o, b.io, p[0], nlev);
    ld_stream<VX>(sig, o, b.sh, p[2], nlev);
    // FP32: velocity and image without the non-temporal hint (read, and the
    // image written, every step: 143 against 155 us per step at 304^3);
    // the FP16 step is faster with it
    if constexpr (kNarrow<T>) {
      ld_stream<VX>(vel, o, 1.f, p[1]);
      ld_stream<VX>(acc, o, 1.f, p[3]);
    } else {
      ld_vec<VX>(vel, o, 1.f, p[1]);
      ld_vec<VX>(acc, o, 1.f, p[3]);
    }
  }
  template <int VX>
  __device__ void storev(const Blk1 &b, uint32_t o, const float (&out)[NOUT][VX]) const {
    st_vec<VX>(u0, o, out[0], b.sn, nlev);
    st_vec<VX, kNarrow<T>>(acc, o, out[1], 1.f);
  }
  __device__ void cell(const Blk1 &, int slot, const CellArgs &c, const float *lap,
                       const float *ctr, const float *p, float (&out)[NOUT],
                       float (&tmax)[NF]) const {
    const float bb = c.b, v = p[1];
    float un = leap(bb, v, lap[0], ctr[0], p[0]);
    if (slot >= 0 && inject) un = lam_inject(rm, slot, c.x, c.y, nx, ny, tr, tab, id, v, bb, un);
    if constexpr (H) tmax[0] = fmaxf(tmax[0], fabsf(un));
    out[0] = un;
    out[1] = fma_r(p[2], un, p[3]);
  }
};

// ------------------------------------------------------------ shot batches
// The forward op of every shot of a batch: the op's arrays start at the
// batch's first shot and are moved by the shot's descriptor.
template <int SCATTER, class T>
struct BatchFwd {
  // without the non-temporal hint: it slowed the batch (32 windowed shots)
  // by about 5 % where it sped up single shots
  using Op = FwdOp<SCATTER, T, false>;
  Op base;
  const float *sponge;
  float *rings;
  const ShotDesc *__restrict__ desc;
  int step, npts_total;
  struct Shot {
    Grid g;
    Sponge sp;
    Op op;
  };
  __device__ Shot shot(int s, int tw, int ty, int zc) const {
    const ShotDesc d = desc[s];
    Shot r;
    r.g = tile_grid(d.nx, d.ny, d.nz, tw, ty, zc);
    r.sp.bx = sponge + d.sp0;
    r.sp.by = sponge + d.sp0 + d.nx;
    r.sp.bz = sponge + d.sp0 + d.nx + d.ny;
    Op &o = r.op;
    o = base;
    o.s1 = base.s1 + d.cell0;
    o.r1 = base.r1 + d.cell0;
    o.s0 = base.s0 + d.cell0;
    o.s_new = base.s_new + d.cell0;
    o.r0 = base.r0 + d.cell0;
    o.sr1 = base.sr1 + d.cell0;
    o.sr0 = base.sr0 + d.cell0;
    o.srn = base.srn + d.cell0;
    o.vel = base.vel + d.cell0;
    o.model = base.model + d.cell0;
    o.hist = nullptr;
    o.pts.p = base.pts.p + d.pts0;
    o.pts.n = d.npts;
    o.amp = base.amp + (long long)step * npts_total + d.pts0;
    o.rm.plane_slot = base.rm.plane_slot + d.slot0;
    o.rm.rmap = base.rm.rmap + d.rmap0;
    o.rm.rnext = base.rm.rnext + d.rnext0;
    o.rm.nrec = d.nrec;
    o.tr = base.tr + d.tr0;
    o.nx = d.nx;
    o.ny = d.ny;
    o.rg = rings + d.ring0;
    o.gain = d.gain;
    o.nlev = (uint32_t)d.nx * d.ny * d.nz;
    return r;
  }
};

// The exact adjoint's ops for every shot of a batch, moved per shot as
// BatchFwd: the background recompute and the adjoint-state step, with each
// shot's scale rings, bound gain, receiver tables and traces.
template <int SCATTER, class T>
struct BatchSrec {
  using Op = SrecOp<SCATTER, T>;
  Op base;
  const float *sponge;
  float *rings;
  const ShotDesc *__restrict__ desc;
  int step, npts_total;
  struct Shot {
    Grid g;
    Sponge sp;
    Op op;
  };
  __device__ Shot shot(int s, int tw, int ty, int zc) const {
    const ShotDesc d = desc[s];
    Shot r;
    r.g = tile_grid(d.nx, d.ny, d.nz, tw, ty, zc);
    r.sp.bx = sponge + d.sp0;
    r.sp.by = sponge + d.sp0 + d.nx;
    r.sp.bz = sponge + d.sp0 + d.nx + d.ny;
    Op &o = r.op;
    o = base;
    o.s1 = base.s1 + d.cell0;
    o.s0 = base.s0 + d.cell0;
    o.s_new = base.s_new + d.cell0;
    o.vel = base.vel + d.cell0;
    o.sig = base.sig ? base.sig + d.cell0 : nullptr;
    o.pts.p = base.pts.p + d.pts0;
    o.pts.n = d.npts;
    o.amp = base.amp + (long long)step * npts_total + d.pts0;
    o.ring = {rings + d.ring0 + BR_S, (unsigned *)(rings + d.ring0 + BR_S + 4)};
    o.gain = d.gain;
    o.hscale = base.hscale ? base.hscale + s : nullptr;
    o.nlev = (uint32_t)d.nx * d.ny * d.nz;
    return r;
  }
};

template <class T>
struct BatchLam {
  using Op = LamOp<T>;
  Op base;
  const float *sponge;
  float *rings;
  const ShotDesc *__restrict__ desc;
  int step, npts_total;
  struct Shot {
    Grid g;
    Sponge sp;
    Op op;
  };
  __device__ Shot shot(int s, int tw, int ty, int zc) const {
    const ShotDesc d = desc[s];
    Shot r;
    r.g = tile_grid(d.nx, d.ny, d.nz, tw, ty, zc);
    r.sp.bx = sponge + d.sp0;
    r.sp.by = sponge + d.sp0 + d.nx;
    r.sp.bz = sponge + d.sp0 + d.nx + d.ny;
    Op &o = r.op;
    o = base;
    o.u1 = base.u1 + d.cell0;
    o.u0 = base.u0 + d.cell0;
    o.vel = base.vel + d.cell0;
    o.sig = base.sig + d.cell0;
    o.acc = base.acc + d.cell0;
    o.rm.plane_slot = base.rm.plane_slot + d.slot0;
    o.rm.rmap = base.rm.rmap + d.rmap0;
    o.rm.rnext = base.rm.rnext + d.rnext0;
    o.rm.nrec = d.nrec;
    o.tr = base.tr + d.tr0;
    o.nx = d.nx;
    o.ny = d.ny;
    o.ring = {rings + d.ring0 + BR_U, (unsigned *)(rings + d.ring0 + BR_U + 4)};
    o.gain = d.gain;
    // the single-shot bound's factors in its order: (max(V) |tab|) per cell
    o.inj = d.vmax * base.inj * (float)d.per_cell;
    o.trmax = (const unsigned *)(rings + d.ring0 + BR_MAXTR);
    o.hscale = base.hscale ? base.hscale + s : nullptr;
    o.nlev = (uint32_t)d.nx * d.ny * d.nz;
    return r;
  }
};

// ------------------------------------------------------------- launchers

// switch (shape) { case i: return F<..., RY, WY, VX, LA, LX>(ARGS); } over kShapes
// switch (shape) { case i: return F<..., RY, WY, VX, LA, LX>(ARGS); } over kShapes
#define BORN_SHAPE_CASES(F, OP, ...)                         \
  case 0: return F<OP, 2, 4, 1, 1, 64>(__VA_ARGS__);         \
  case 1: return F<OP, 4, 4, 1, 1, 64>(__VA_ARGS__);         \
  case 2: return F<OP, 2, 8, 1, 1, 64>(__VA_ARGS__);         \
  case 3: return F<OP, 2, 16, 1, 1, 64>(__VA_ARGS__);        \
  case 4: return F<OP, 1, 8, 1, 1, 64>(__VA_ARGS__);         \
  case 5: return F<OP, 1, 16, 1, 1, 64>(__VA_ARGS__);        \
  case 6: return F<OP, 1, 4, 2, 1, 32>(__VA_ARGS__);         \
  case 7: return F<OP, 2, 4, 2, 1, 32>(__VA_ARGS__);         \
  case 8: return F<OP, 1, 8, 2, 1, 32>(__VA_ARGS__);         \
  case 9: return F<OP, 2, 8, 2, 1, 32>(__VA_ARGS__);         \
  case 10: return F<OP, 1, 12, 2, 1, 32>(__VA_ARGS__);       \
  case 11: return F<OP, 2, 2, 2, 1, 32>(__VA_ARGS__);        \
  case 12: return F<OP, 2, 4, 1, 2, 64>(__VA_ARGS__);        \
  case 13: return F<OP, 4, 4, 1, 2, 64>(__VA_ARGS__);        \
  case 14: return F<OP, 2, 8, 1, 2, 64>(__VA_ARGS__);        \
  case 15: return F<OP, 2, 16, 1, 2, 64>(__VA_ARGS__);       \
  case 16: return F<OP, 1, 8, 1, 2, 64>(__VA_ARGS__);        \
  case 17: return F<OP, 1, 16, 1, 2, 64>(__VA_ARGS__);       \
  case 18: return F<OP, 1, 4, 2, 2, 32>(__VA_ARGS__);        \
  case 19: return F<OP, 2, 4, 2, 2, 32>(__VA_ARGS__);        \
  case 20: return F<OP, 1, 8, 2, 2, 32>(__VA_ARGS__);        \
  case 21: return F<OP, 2, 8, 2, 2, 32>(__VA_ARGS__);        \
  case 22: return F<OP, 1, 12, 2, 2, 32>(__VA_ARGS__);       \
  case 23: return F<OP, 2, 2, 2, 2, 32>(__VA_ARGS__);


inline Grid make_grid(const Geometry &geo, const TileShape &ts, int zc) {
  return tile_grid(geo.nx, geo.ny, geo.nz, ts.lx * ts.vx, ts.ry * (kTX * ts.wy / ts.lx),
                   zc > 0 ? zc : 32);
}

template <class Op, int RY, int WY, int VX, int LA, int LX>
hipError_t launch_shape(const Geometry &geo, int zc, const Op &op, hipStream_t st) {
  if (VX == 2 && geo.nx % 2) return hipErrorInvalidValue;
  const Grid g = make_grid(geo, {RY, WY, VX, LA, LX}, zc);
  const int nb = g.ntx * g.nty * g.ntz;
  hipLaunchKernelGGL((march_kernel<Op, RY, WY, VX, LA, LX>), dim3(nb), dim3(kTX * WY), 0, st,
                     g, geo.taps, geo.sponge, op);
  return hipGetLastError();
}

template <class Op>
hipError_t launch(const Geometry &geo, const Tune &tu, const Op &op, hipStream_t st) {
  switch (tu.shape) {
    BORN_SHAPE_CASES(launch_shape, Op, geo, tu.zchunk, op, st)
    default: return hipErrorInvalidValue;
  }
}

inline Points to_points(const SrcPts &s) { return {(const int4 *)s.p, s.n}; }
inline int pack_slots(const Slots &s) { return s.cur | s.old << 8 | s.nxt << 16 | s.clr << 24; }

template <class T>
T *as(const void *p) {
  return (T *)p;
}

template <class BOp, int RY, int WY, int VX, int LA, int LX>
hipError_t launch_batch_shape(const Taps &taps, int zc, const BatchCommon &b,
                                     const BOp &op, hipStream_t st) {
  if (b.nshots < 1 || b.nshots > kMaxBatch) return hipErrorInvalidValue;
  BatchTiles bt{};
  bt.n = b.nshots;
  bt.zc = zc > 0 ? zc : 32;
  int tiles = 0;
  for (int s = 0; s < b.nshots; s++) {
    const ShotDesc &d = b.host_desc[s];
    if (VX == 2 && d.nx % 2) return hipErrorInvalidValue;
    const Grid g = tile_grid(d.nx, d.ny, d.nz, LX * VX, RY * (kTX * WY / LX), bt.zc);
    bt.tile0[s] = tiles;
    tiles += g.ntx * g.nty * g.ntz;
  }
  bt.tile0[b.nshots] = tiles;
  hipLaunchKernelGGL((march_batch_kernel<BOp, RY, WY, VX, LA, LX>), dim3(tiles),
                     dim3(kTX * WY), 0, st, bt, taps, op);
  return hipGetLastError();
}

template <class BOp>
hipError_t launch_batch(const Taps &taps, const Tune &tu, const BatchCommon &b,
                               const BOp &op, hipStream_t st) {
  switch (tu.shape) {
    BORN_SHAPE_CASES(launch_batch_shape, BOp, taps, tu.zchunk, b, op, st)
    default: return hipErrorInvalidValue;
  }
}

// ---- the launchers of each op type.  They are templates here and are
// instantiated explicitly, each in one translation unit (born_k_*.hip, built
// in parallel); born_kernels.hip declares them extern and dispatches.

template <int SCATTER, class T>
hipError_t fwd_step(const Geometry &geo, const Tune &tu, const FwdStep &a, hipStream_t st) {
  FwdOp<SCATTER, T> op{};
  op.s1 = as<T>(a.s1); op.r1 = as<T>(a.r1); op.s0 = as<T>(a.s0); op.r0 = as<T>(a.r0);
  op.s_new = as<T>(a.s_new ? a.s_new : a.s0);
  op.sr1 = (const PairOf<T> *)a.s1; op.sr0 = (const PairOf<T> *)a.s0;
  op.srn = (PairOf<T> *)(a.s_new ? a.s_new : a.s0);
  op.vel = a.vel; op.model = a.model; op.vm = (const float2 *)a.vm; op.hist = as<T>(a.hist);
  if (!a.vm) return hipErrorInvalidValue;
  op.pts = to_points(a.src); op.amp = a.amp;
  op.rm = a.rec; op.tr = a.tr; op.tab = a.tab; op.id = a.id;
  op.extract = a.extract; op.nx = geo.nx; op.ny = geo.ny;
  op.rg = a.rg; op.slots = pack_slots(a.h.sl);
  op.gain = a.h.gain; op.tapsum = a.h.tapsum; op.hscale = a.hscale;
  op.nlev = (uint32_t)geo.nx * geo.ny * geo.nz;
  return launch(geo, tu, op, st);
}

template <int SCATTER, class T>
hipError_t srec_step(const Geometry &geo, const Tune &tu, const SrecStep &a, hipStream_t st) {
  SrecOp<SCATTER, T> op{};
  op.s1 = as<T>(a.s1); op.s0 = as<T>(a.s0); op.vel = a.vel; op.sig = as<T>(a.sig);
  op.s_new = as<T>(a.s_new ? a.s_new : a.s0);
  op.pts = to_points(a.src); op.amp = a.amp;
  op.ring = a.ring; op.sl = a.h.sl; op.gain = a.h.gain; op.tapsum = a.h.tapsum;
  op.hscale = a.hscale;
  op.nlev = (uint32_t)geo.nx * geo.ny * geo.nz;
  return launch(geo, tu, op, st);
}

template <class T>
hipError_t lam_step(const Geometry &geo, const Tune &tu, const LamStep &a, hipStream_t st) {
  LamOp<T> op{};
  op.u1 = as<T>(a.u1); op.u0 = as<T>(a.u0); op.vel = a.vel; op.sig = as<T>(a.sig);
  op.acc = a.acc; op.rm = a.rec; op.tr = a.tr; op.tab = a.tab; op.id = a.id;
  op.inject = a.inject; op.nx = geo.nx; op.ny = geo.ny;
  op.ring = a.ring; op.sl = a.h.sl; op.gain = a.h.gain; op.inj = a.inj;
  op.trmax = a.trmax; op.hscale = a.hscale;
  op.nlev = (uint32_t)geo.nx * geo.ny * geo.nz;
  return launch(geo, tu, op, st);
}

template <int SCATTER, class T>
hipError_t batch_fwd_step(const Taps &taps, const Tune &tu, const BatchFwdStep &b,
                          hipStream_t st) {
  const FwdStep &a = b.a;
  BatchFwd<SCATTER, T> bop{};
  auto &op = bop.base;
  op.s1 = as<T>(a.s1); op.r1 = as<T>(a.r1); op.s0 = as<T>(a.s0); op.r0 = as<T>(a.r0);
  op.s_new = as<T>(a.s0);
  op.sr1 = (const PairOf<T> *)a.s1; op.sr0 = (const PairOf<T> *)a.s0;
  op.srn = (PairOf<T> *)a.s0;
  op.vel = a.vel; op.model = a.model; op.hist = nullptr;
  op.pts = to_points(a.src); op.amp = a.amp;
  op.rm = a.rec; op.tr = a.tr; op.tab = a.tab; op.id = a.id;
  op.extract = a.extract;
  op.slots = pack_slots(a.h.sl); op.tapsum = a.h.tapsum; op.hscale = nullptr;
  bop.sponge = b.sponge; bop.rings = b.rings; bop.desc = b.desc; bop.step = b.step;
  bop.npts_total = b.npts_total;
  return launch_batch(taps, tu, b, bop, st);
}

template <int SCATTER, class T>
hipError_t batch_srec_step(const Taps &taps, const Tune &tu, const BatchSrecStep &b,
                           hipStream_t st) {
  const SrecStep &a = b.a;
  BatchSrec<SCATTER, T> bop{};
  auto &op = bop.base;
  op.s1 = as<T>(a.s1); op.s0 = as<T>(a.s0); op.vel = a.vel; op.sig = as<T>(a.sig);
  op.s_new = as<T>(a.s_new ? a.s_new : a.s0);
  op.pts = to_points(a.src); op.amp = a.amp;
  op.sl = a.h.sl; op.tapsum = a.h.tapsum; op.hscale = a.hscale;
  bop.sponge = b.sponge; bop.rings = b.rings; bop.desc = b.desc; bop.step = b.step;
  bop.npts_total = b.npts_total;
  return launch_batch(taps, tu, b, bop, st);
}

template <class T>
hipError_t batch_lam_step(const Taps &taps, const Tune &tu, const BatchLamStep &b,
                          hipStream_t st) {
  const LamStep &a = b.a;
  BatchLam<T> bop{};
  auto &op = bop.base;
  op.u1 = as<T>(a.u1); op.u0 = as<T>(a.u0); op.vel = a.vel; op.sig = as<T>(a.sig);
  op.acc = a.acc; op.rm = a.rec; op.tr = a.tr; op.tab = a.tab; op.id = a.id;
  op.inject = a.inject; op.sl = a.h.sl; op.inj = a.inj; op.hscale = a.hscale;
  bop.sponge = b.sponge; bop.rings = b.rings; bop.desc = b.desc; bop.step = b.step;
  bop.npts_total = b.npts_total;
  return launch_batch(taps, tu, b, bop, st);
}

// the instantiations, one translation unit each (F: template or extern template)
#define BORN_FWD_STEP(F, S, T) \
  F hipError_t fwd_step<S, T>(const Geometry &, const Tune &, const FwdStep &, hipStream_t)
#define BORN_SREC_STEP(F, S, T) \
  F hipError_t srec_step<S, T>(const Geometry &, const Tune &, const SrecStep &, hipStream_t)
#define BORN_LAM_STEP(F, T) \
  F hipError_t lam_step<T>(const Geometry &, const Tune &, const LamStep &, hipStream_t)
#define BORN_BATCH_SREC_STEP(F, S, T)                                                   \
  F hipError_t batch_srec_step<S, T>(const Taps &, const Tune &, const BatchSrecStep &, \
                                     hipStream_t)
#define BORN_BATCH_LAM_STEP(F, T) \
  F hipError_t batch_lam_step<T>(const Taps &, const Tune &, const BatchLamStep &, hipStream_t)
#define BORN_BATCH_FWD_STEP(F, S, T)                                                     \
  F hipError_t batch_fwd_step<S, T>(const Taps &, const Tune &, const BatchFwdStep &, \
                                    hipStream_t)

}  // namespace born
