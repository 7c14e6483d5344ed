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

// ---- Stores the compiler's wait-counter model does not see.  Loads and
// stores share one counter (vmcnt) on this architecture and the compiler
// takes a counter with both pending as unordered, so any wait for a load
// would drain every load in flight.  As inline asm the step kernels' stores
// are invisible to it: a wait for a load issued two steps earlier can leave
// the previous step's loads in flight.  That wait stays exact because a
// wavefront's loads return in order, so a vmcnt that counts only the loads
// issued after the one waited for covers it (pending stores only make the
// wait longer).  No step kernel reads what it stores before the launch ends,
// and a store of at most 8 bytes reads its data registers at issue.  The
// base is a wave-uniform pointer (scalar registers), off a byte offset.
// A vector memory instruction that reads a scalar register written by a
// vector instruction (v_readlane, as in a reload of a spilled scalar) needs
// 5 wait states before it, which the compiler does not insert for inline
// asm: each store starts with s_nop 4.  NT adds the non-temporal hint: no
// step reads a new level back before the next launch, and the hint took 6 %
// off the FP32 forward step at 304^3.
template <bool NT>
__device__ __forceinline__ void st_b8(void *p, uint32_t off, unsigned char v) {
  if constexpr (NT)
    asm volatile("s_nop 4\n\tglobal_store_byte %0, %1, %2 nt"
                 : : "v"(off), "v"((uint32_t)v), "s"(p) : "memory");
  else
    asm volatile("s_nop 4\n\tglobal_store_byte %0, %1, %2"
                 : : "v"(off), "v"((uint32_t)v), "s"(p) : "memory");
}
template <bool NT>
__device__ __forceinline__ void st_b16(void *p, uint32_t off, unsigned short v) {
  if constexpr (NT)
    asm volatile("s_nop 4\n\tglobal_store_short %0, %1, %2 nt"
                 : : "v"(off), "v"((uint32_t)v), "s"(p) : "memory");
  else
    asm volatile("s_nop 4\n\tglobal_store_short %0, %1, %2"
                 : : "v"(off), "v"((uint32_t)v), "s"(p) : "memory");
}
template <bool NT>
__device__ __forceinline__ void st_b32(void *p, uint32_t off, uint32_t v) {
  if constexpr (NT)
    asm volatile("s_nop 4\n\tglobal_store_dword %0, %1, %2 nt"
                 : : "v"(off), "v"(v), "s"(p) : "memory");
  else
    asm volatile("s_nop 4\n\tglobal_store_dword %0, %1, %2"
                 : : "v"(off), "v"(v), "s"(p) : "memory");
}
template <bool NT>
__device__ __forceinline__ void st_b64(void *p, uint32_t off, uint64_t v) {
  if constexpr (NT)
    asm volatile("s_nop 4\n\tglobal_store_dwordx2 %0, %1, %2 nt"
                 : : "v"(off), "v"(v), "s"(p) : "memory");
  else
    asm volatile("s_nop 4\n\tglobal_store_dwordx2 %0, %1, %2"
                 : : "v"(off), "v"(v), "s"(p) : "memory");
}

// Block barrier for LDS: the block's LDS accesses complete, then s_barrier.
// __syncthreads also fences global memory, which waits for every load in
// flight; the step kernels share nothing through global memory in a launch.
__device__ __forceinline__ void lds_barrier() {
  asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" : : : "memory");
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

// Read-only tables indexed by a uniform value (sponge factor of a plane, the
// receiver slot of a plane, source points): read through the constant
// address space, so they become scalar loads, tracked apart from the vector
// loads of the wavefields.  A vector load issued after the next plane's
// prefetch would otherwise make the wait for it a wait for the prefetch too.
template <class T>
__device__ __forceinline__ T ldc(const T *p, long long i) {
  static_assert(sizeof(T) % 4 == 0, "32-bit words");
  typedef const __attribute__((address_space(4))) uint32_t *cwords;
  const cwords w = (cwords)(p + i);
  uint32_t u[sizeof(T) / 4];
#pragma unroll
  for (int k = 0; k < (int)(sizeof(T) / 4); k++) u[k] = w[k];
  T v;
  __builtin_memcpy(&v, u, sizeof(T));
  return v;
}

// The hardware hands block h to XCD h % 8.  Give every XCD a contiguous range
// of logical tiles so that tiles sharing halo rows share an L2.
__device__ __forceinline__ int xcd_remap(int h, int nb) {
  const int q = nb / 8, r = nb % 8;
  const int xcd = h % 8, k = h / 8;
  return xcd * q + (xcd < r ? xcd : r) + k;
}

// Arithmetic with its rounding fixed.  The compiler fuses a product into a
// following sum where it can; these products and sums are never fused, and
// fma_r always is, so every kernel that computes a cell in the same terms
// gets the same bits.
__device__ __forceinline__ float add_r(float a, float b) {
#pragma clang fp contract(off)
  return a + b;
}
__device__ __forceinline__ float mul_r(float a, float b) {
#pragma clang fp contract(off)
  return a * b;
}
__device__ __forceinline__ double add_r(double a, double b) {
#pragma clang fp contract(off)
  return a + b;
}
__device__ __forceinline__ double mul_r(double a, double b) {
#pragma clang fp contract(off)
  return a * b;
}
__device__ __forceinline__ float fma_r(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

// One leapfrog update b * (V lap + 2 cur - b old) (cpu_prop.cc:31-53 with
// the damping of cpu_prop.cc:141-165), as fma(V, lap, cur) + cur - b old.
__device__ __forceinline__ float leap(float b, float v, float lap, float cur, float old) {
  return mul_r(b, fma_r(-b, old, add_r(fma_r(v, lap, cur), cur)));
}

// The Born source term of a forward step and the scatter operand kept for
// the exact adjoint.  Scatter 0 (image): r += 1e-6 s m in double, the
// reference's double literal (cpu_prop.cc:220); scatter 1 (dV): sigma =
// b L(s), r += m sigma.
template <int SCATTER>
__device__ __forceinline__ void born_scatter(float bb, float lap_s, float sn, float m,
                                             float &rn, float &sig) {
  if (SCATTER == 0) {
    rn = (float)add_r((double)rn, mul_r(mul_r(.000001, (double)sn), (double)m));
    sig = sn;
  } else {
    sig = mul_r(bb, lap_s);
    rn = fma_r(m, sig, rn);
  }
}

// Eight-tap sinc interpolation with the reference's rounding: products and
// sums rounded separately, left to right (base/cpu_prop.cc:169-176).
__device__ __forceinline__ float sinc8(const float *t, const float *v,
                                       long long stride) {
  float s = mul_r(t[0], v[0]);
#pragma unroll
  for (int k = 1; k < 8; k++) s = add_r(s, mul_r(t[k], v[k * stride]));
  return s;
}

// Uniform per-block mask of the source points inside the block's cells.
__device__ __forceinline__ uint32_t points_in_block(const Points &pts, int x0,
                                                   int x1, int y0, int y1,
                                                   int z0, int z1) {
  uint32_t m = 0;
  for (int i = 0; i < pts.n; i++) {
    const int4 q = ldc(pts.p, i);
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
    if (ldc(pts.p, i).z == z) m |= 1u << i;
  }
  return m;
}

// The source points of a block: the mask of those inside and the range of
// their z planes, so that the test of a plane off that range loads nothing
// (a loop of dependent scalar loads per plane cost the block that holds a
// 9-point source about 8 us of a 160 us step at 304^3).
struct PtsBlk {
  uint32_t m;
  int zlo, zhi;
};
__device__ __forceinline__ PtsBlk points_of_block(const Points &pts, int x0, int x1, int y0,
                                                  int y1, int z0, int z1) {
  PtsBlk b{0u, 1, 0};
  for (int i = 0; i < pts.n; i++) {
    const int4 q = ldc(pts.p, i);
    if (q.x >= x0 && q.x < x1 && q.y >= y0 && q.y < y1 && q.z >= z0 && q.z < z1) {
      b.zlo = b.m ? min(b.zlo, q.z) : q.z;
      b.zhi = b.m ? max(b.zhi, q.z) : q.z;
      b.m |= 1u << i;
    }
  }
  return b;
}
__device__ __forceinline__ uint32_t points_on_plane(const Points &pts, const PtsBlk &b,
                                                    int z) {
  return z < b.zlo || z > b.zhi ? 0u : points_on_plane(pts, b.m, z);
}

// v plus, in list order, amp[i] of the points at (x, y) among plane mask m
__device__ __forceinline__ float add_points(const Points &pts, uint32_t m, int x,
                                            int y, float v, const float *amp) {
  for (uint32_t b = m; b; b &= b - 1) {
    const int i = __builtin_ctz(b);
    const int4 q = ldc(pts.p, i);
    if (q.x == x && q.y == y) v = add_r(v, ldc(amp, i));
  }
  return v;
}

// Receiver injection of the adjoint-state step: u + V b g, g the sum over
// the receivers at (x, y) of their time-sinc interpolated residual
// (born_abi.hip, exact adjoint).  Rounding fixed as in leap().
__device__ __forceinline__ float lam_inject(const RecMap &rm, int slot, int x, int y, int nx,
                                            int ny, const float *tr, const float *tab, int id,
                                            float v, float b, float u) {
  float g = 0.f;
  for (int i = rm.rmap[((long long)slot * ny + y) * nx + x]; i >= 0; i = rm.rnext[i])
    g = add_r(g, sinc8(tab, tr + (long long)id * rm.nrec + i, rm.nrec));
  return fma_r(mul_r(v, b), g, u);
}

struct CellArgs {
  int x, y, z;
  uint32_t o;  // element offset of the cell
  float b;     // sponge factor (1 outside the sponge)
};

__device__ __forceinline__ float sat_half(float v) { return fminf(fmaxf(v, -65504.f), 65504.f); }

// ---- Two FP16 fields stored as one __half2 per cell (first field in the
// low half): one 4-byte load brings both, half the load instructions of two
// separate FP16 arrays.  inv = 1/scale of each field.
template <int VX>
__device__ __forceinline__ void ld_pair(const __half2 *p, uint32_t o, const float (&inv)[2],
                                        float (&v)[2][VX], uint32_t = 0) {
  if constexpr (VX == 1) {
    const __half2 w = at(p, o);
    v[0][0] = __low2float(w) * inv[0];
    v[1][0] = __high2float(w) * inv[1];
  } else {
    const uint2 u = at((const uint2 *)p, o >> 1);
    __half2 a, b;
    __builtin_memcpy(&a, &u.x, 4);
    __builtin_memcpy(&b, &u.y, 4);
    v[0][0] = __low2float(a) * inv[0];
    v[1][0] = __high2float(a) * inv[1];
    v[0][1] = __low2float(b) * inv[0];
    v[1][1] = __high2float(b) * inv[1];
  }
}
template <int VX, bool NT = true>
__device__ __forceinline__ void st_pair(__half2 *p, uint32_t o, const float (&a)[VX],
                                        const float (&b)[VX], const float (&scale)[2],
                                        uint32_t = 0) {
  const __half2 c0 = __floats2half2_rn(sat_half(a[0] * scale[0]), sat_half(b[0] * scale[1]));
  uint32_t u0;
  __builtin_memcpy(&u0, &c0, 4);
  if constexpr (VX == 1) {
    st_b32<NT>(p, o * 4u, u0);
  } else {
    const __half2 c1 = __floats2half2_rn(sat_half(a[1] * scale[0]), sat_half(b[1] * scale[1]));
    uint32_t u1;
    __builtin_memcpy(&u1, &c1, 4);
    st_b64<NT>(p, o * 4u, (uint64_t)u1 << 32 | u0);
  }
}

// ---- 24-bit storage: an FP32 value rounded to 24 bits per cell (sign, the
// 8 exponent bits and 15 significand bits; relative precision 2^-16, the
// whole FP32 range, no scale).  hi holds the top 16 bits of the rounded
// FP32 word and lo the 8 bits below them; the value's FP32 bits are
// hi << 16 | lo << 8.  Rounding adds 0x80 to the word and truncates: half
// way cases go away from zero, a finite value past the largest 24-bit one
// becomes an infinity as in FP32 rounding, infinities stay infinities.
//
// Layout of a level of n cells: the n hi halves, then the n lo bytes; for
// two fields stored together (s, r), n hi pairs then n lo byte pairs.  So
// a cell's hi and lo addresses are a shift of its offset, lo's plus the
// uniform 2n (4n) bytes.  A pointer to Split (SplitPair) advances by 3 (6)
// bytes per cell: the levels of a history and the shots of a batch are
// whole regions of 3n (6n) bytes.  The scale arguments of the loads and
// stores below are those of the scaled FP16 storage and are not used.
struct Split {
  unsigned char b[3];
};
struct SplitPair {
  unsigned char b[6];
};
static_assert(sizeof(Split) == 3 && sizeof(SplitPair) == 6, "packed split types");

// the rounded FP32 word of v
__device__ __forceinline__ uint32_t split_word(float v) { return __float_as_uint(v) + 0x80u; }
// value of a stored cell from hi (in the low 16 bits) and lo (low 8 bits)
__device__ __forceinline__ float split_val(uint32_t hi, uint32_t lo) {
  return __uint_as_float(hi << 16 | (lo & 0xffu) << 8);
}

// VX = 2: two consecutive cells per lane (o even, n even: a pair of hi
// values is 4-byte aligned, its lo bytes 2-byte aligned)
template <int VX>
__device__ __forceinline__ void ld_vec(const Split *p, uint32_t o, float, float (&v)[VX],
                                       uint32_t n) {
  const char *b = (const char *)p;
  if constexpr (VX == 1) {
    const uint32_t hi = *(const unsigned short *)(b + o * 2u);
    const uint32_t lo = *(const unsigned char *)(b + (2u * n + o));
    v[0] = split_val(hi, lo);
  } else {
    const uint32_t hi = *(const uint32_t *)(b + o * 2u);
    const uint32_t lo = *(const unsigned short *)(b + (2u * n + o));
    v[0] = __uint_as_float(hi << 16 | (lo & 0xffu) << 8);
    v[1] = __uint_as_float((hi & 0xffff0000u) | (lo & 0xff00u));
  }
}
template <int VX>
__device__ __forceinline__ void ld_stream(const Split *p, uint32_t o, float inv,
                                          float (&v)[VX], uint32_t n) {
  ld_vec<VX>(p, o, inv, v, n);
}
template <int VX, bool NT = true>
__device__ __forceinline__ void st_vec(Split *p, uint32_t o, const float (&v)[VX], float,
                                       uint32_t n) {
  const uint32_t q0 = split_word(v[0]);
  if constexpr (VX == 1) {
    st_b16<NT>(p, o * 2u, (unsigned short)(q0 >> 16));
    st_b8<NT>(p, 2u * n + o, (unsigned char)(q0 >> 8));
  } else {
    const uint32_t q1 = split_word(v[1]);
    st_b32<NT>(p, o * 2u, q0 >> 16 | (q1 & 0xffff0000u));
    st_b16<NT>(p, 2u * n + o, (unsigned short)((q0 >> 8 & 0xffu) | (q1 & 0xff00u)));
  }
}
// two fields stored together: first field in the low half of each pair
template <int VX>
__device__ __forceinline__ void ld_pair(const SplitPair *p, uint32_t o, const float (&)[2],
                                        float (&v)[2][VX], uint32_t n) {
  const char *b = (const char *)p;
  if constexpr (VX == 1) {
    const uint32_t hi = *(const uint32_t *)(b + o * 4u);
    const uint32_t lo = *(const unsigned short *)(b + (4u * n + o * 2u));
    v[0][0] = __uint_as_float(hi << 16 | (lo & 0xffu) << 8);
    v[1][0] = __uint_as_float((hi & 0xffff0000u) | (lo & 0xff00u));
  } else {
    const uint2 hi = *(const uint2 *)(b + o * 4u);
    const uint32_t lo = *(const uint32_t *)(b + (4u * n + o * 2u));
    v[0][0] = __uint_as_float(hi.x << 16 | (lo & 0xffu) << 8);
    v[1][0] = __uint_as_float((hi.x & 0xffff0000u) | (lo & 0xff00u));
    v[0][1] = __uint_as_float(hi.y << 16 | (lo >> 8 & 0xff00u));
    v[1][1] = __uint_as_float((hi.y & 0xffff0000u) | (lo >> 16 & 0xff00u));
  }
}
template <int VX, bool NT = true>
__device__ __forceinline__ void st_pair(SplitPair *p, uint32_t o, const float (&a)[VX],
                                        const float (&b)[VX], const float (&)[2],
                                        uint32_t n) {
  const uint32_t qa0 = split_word(a[0]), qb0 = split_word(b[0]);
  const uint32_t h0 = qa0 >> 16 | (qb0 & 0xffff0000u);
  const uint32_t l0 = (qa0 >> 8 & 0xffu) | (qb0 & 0xff00u);
  if constexpr (VX == 1) {
    st_b32<NT>(p, o * 4u, h0);
    st_b16<NT>(p, 4u * n + o * 2u, (unsigned short)l0);
  } else {
    const uint32_t qa1 = split_word(a[1]), qb1 = split_word(b[1]);
    const uint32_t h1 = qa1 >> 16 | (qb1 & 0xffff0000u);
    const uint32_t l1 = (qa1 >> 8 & 0xffu) | (qb1 & 0xff00u);
    st_b64<NT>(p, o * 4u, (uint64_t)h1 << 32 | h0);
    st_b32<NT>(p, 4u * n + o * 2u, l0 | l1 << 16);
  }
}

// ---- VX consecutive cells of a row (VX = 1 or 2).  A pair starts at an
// even element offset (x tiles start at even x and nx is even when VX = 2),
// so float2 / __half2 accesses are naturally aligned.
template <int VX>
__device__ __forceinline__ void ld_vec(const float *p, uint32_t o, float, float (&v)[VX],
                                       uint32_t = 0) {
  if constexpr (VX == 1) {
    v[0] = at(p, o);
  } else {
    const float2 w = at((const float2 *)p, o >> 1);
    v[0] = w.x;
    v[1] = w.y;
  }
}
template <int VX>
__device__ __forceinline__ void ld_vec(const __half *p, uint32_t o, float inv,
                                       float (&v)[VX], uint32_t = 0) {
  if constexpr (VX == 1) {
    v[0] = __half2float(at(p, o)) * inv;
  } else {
    const __half2 w = at((const __half2 *)p, o >> 1);
    v[0] = __low2float(w) * inv;
    v[1] = __high2float(w) * inv;
  }
}
// The same loads with the non-temporal hint, for inputs a launch reads once
// (the pointwise operands of a cell).
template <int VX>
__device__ __forceinline__ void ld_stream(const float *p, uint32_t o, float, float (&v)[VX],
                                          uint32_t = 0) {
  const float *q = (const float *)((const char *)p + (uint64_t)(o * 4u));
  if constexpr (VX == 1) {
    v[0] = __builtin_nontemporal_load(q);
  } else {
    typedef float f2 __attribute__((ext_vector_type(2)));
    const f2 w = __builtin_nontemporal_load((const f2 *)q);
    v[0] = w.x;
    v[1] = w.y;
  }
}
// FP16 levels keep the plain loads: with the hint the FP16 adjoint-state
// and background steps ran 15-20 % slower (304^3, 2- and 4-byte lanes).
template <int VX>
__device__ __forceinline__ void ld_stream(const __half *p, uint32_t o, float inv,
                                          float (&v)[VX], uint32_t = 0) {
  ld_vec<VX>(p, o, inv, v);
}

template <int VX, bool NT = true>
__device__ __forceinline__ void st_vec(float *p, uint32_t o, const float (&v)[VX], float,
                                       uint32_t = 0) {
  if constexpr (VX == 1)
    st_b32<NT>(p, o * 4u, __float_as_uint(v[0]));
  else
    st_b64<NT>(p, o * 4u, (uint64_t)__float_as_uint(v[1]) << 32 | __float_as_uint(v[0]));
}
template <int VX, bool NT = true>
__device__ __forceinline__ void st_vec(__half *p, uint32_t o, const float (&v)[VX],
                                       float scale, uint32_t = 0) {
  if constexpr (VX == 1) {
    st_b16<NT>(p, o * 2u, __half_as_ushort(__float2half_rn(sat_half(v[0] * scale))));
  } else {
    const __half2 c = __floats2half2_rn(sat_half(v[0] * scale), sat_half(v[1] * scale));
    uint32_t u;
    __builtin_memcpy(&u, &c, 4);
    st_b32<NT>(p, o * 2u, u);
  }
}

// Generic z-marching leapfrog tile.  VX cells per lane along x (a tile is
// 64*VX x RY*WY cells).  Op supplies (all const):
//   static constexpr int NF, NPRE, NOUT, NMAX; static constexpr bool DAMP;
//   Blk block(x0, x1, y0, y1, z0, z1);          uniform per-block state
//   void first(const Blk&);                     (NMAX > 0) once per launch
//   ldv<VX>(const Blk&, int f, uint32_t o, float (&v)[VX])   current level
//   (PAIRED ops instead: ldv2<VX>(const Blk&, uint32_t o, float (&v)[NF][VX])
//   loads every field of the current level at once)
//   Pln plane(const Blk&, int z);               uniform per-plane state
//   prefetchv<VX>(const Blk&, float (&p)[NPRE][VX], uint32_t o)
//   void cell(const Blk&, const Pln&, const CellArgs&, const float *lap,
//             const float *ctr, const float *pre, float (&out)[NOUT],
//             float (&tmax)[NF]);                the new values of one cell
//   storev<VX>(const Blk&, uint32_t o, const float (&out)[NOUT][VX])
//   void publish_max(int f, float m);           (NMAX > 0) block max of |new|
template <int B>
struct StepSet {
  static constexpr int value = B;
};

template <class Op, int RY, int WY, int VX, int PZ, int LX, int D>
__device__ __forceinline__ void march_tile(const Grid &g, const Taps &t, const Sponge &sp,
                                           const Op &op, const int L) {
  constexpr int NF = Op::NF;
  constexpr int NP = Op::NPRE > 0 ? Op::NPRE : 1;
  constexpr int NO = Op::NOUT;
  // LX lanes per row (64: a wavefront per row; 32: two rows per wavefront),
  // NG row groups per block
  constexpr int NG = kTX * WY / LX;
  constexpr int TW = LX * VX;
  constexpr int TY = RY * NG;
  constexpr int LW = TW + 2 * kH;
  constexpr int LH = TY + 2 * kH;
  // Ops that ask for it (Op::Q2, the FP16 ops: their steps ran 7-8 %
  // faster, the FP32 forward 2.5 % slower) with one step of lookahead unroll
  // the loop over pairs of steps: the z queue holds one plane more and moves
  // by two planes every second step (half the register moves), and the
  // pointwise inputs of the two steps sit in two buffers that the steps
  // refill alternately (no copy).
  constexpr bool Q2 = Op::Q2 && D == 1 && PZ == 1;
  constexpr int NB = Q2 ? 2 : D;  // buffer sets of loads in flight
  // queue: planes z - kH .. z + kH + PZ - 1 (Q2: one more)
  constexpr int NQ = 2 * kH + PZ + (Q2 ? 1 : 0);
  // LDS planes form a ring of NR = kH + PZ slots (plane q in slot q % NR):
  // a plane's halo is written when the plane enters the z queue as a lead,
  // kH planes before its centre rows and its use
  constexpr int NR = kH + PZ;
  __shared__ __align__(16) float tile[NR][NF][LH][LW];

  const int bx = L % g.ntx;
  const int rest = L / g.ntx;
  const int by = rest % g.nty;
  const int bz = rest / g.nty;
  const int tx = (int)threadIdx.x % LX;
  const int ty = (int)threadIdx.x / LX;
  // x tiles start at the row start so that a wavefront's row maps onto whole
  // cache lines; the 4-cell shell at either end is masked
  const int x0 = bx * TW, y0 = kH + by * TY;
  const int z0 = kH + bz * g.zc;
  const int z1 = min(z0 + g.zc, g.nz - kH);
  const int x = x0 + VX * tx;
  const bool xin = x >= kH && x < g.nx - kH;  // with VX = 2 a pair is in or out
  const int xc = min(x, g.nx - VX);

  const auto blk = op.block(max(x0, kH), min(x0 + TW, g.nx - kH), y0,
                            min(y0 + TY, g.ny - kH), z0, z1);
  if constexpr (Op::NMAX > 0)
    if (L == 0 && threadIdx.x == 0) op.first(blk);

  int yy[RY];
  bool in[RY];
  uint32_t row[RY];
  float bxy[RY][VX];
#pragma unroll
  for (int j = 0; j < RY; j++) {
    yy[j] = y0 + RY * ty + j;  // a thread's rows are consecutive
    in[j] = xin && yy[j] < g.ny - kH;
    const int yc = min(yy[j], g.ny - 1);
    row[j] = (uint32_t)yc * g.nx + xc;
#pragma unroll
    for (int v = 0; v < VX; v++) bxy[j][v] = Op::DAMP ? fminf(sp.bx[xc + v], sp.by[yc]) : 1.f;
  }

  // LDS halo in units of VX cells: 8 y-halo rows of 64 units and, per row,
  // kH/VX units on either side, spread over the block's threads
  constexpr int XU = kH / VX;
  constexpr int NYH = 2 * kH * LX;
  constexpr int NXH = 2 * XU * TY;
  constexpr int NHALO = NYH + NXH;
  constexpr int HPT = (NHALO + kTX * WY - 1) / (kTX * WY);
  uint32_t hoff[HPT];
  short hr[HPT], hc[HPT];
  bool hv[HPT];
#pragma unroll
  for (int k = 0; k < HPT; k++) {
    const int h = (int)threadIdx.x + k * kTX * WY;
    hv[k] = h < NHALO;
    int lr, lc;
    if (h < NYH) {  // rows 0..3 and TY+4..TY+7
      const int r = h / LX;
      lr = r < kH ? r : TY + r;
      lc = kH + VX * (h % LX);
    } else {  // columns 0..3 and TW+4..TW+7 of rows 4..TY+3
      const int hh = min(h - NYH, NXH - 1);
      const int c = hh % (2 * XU);
      lr = kH + hh / (2 * XU);
      lc = c < XU ? VX * c : kH + TW + VX * (c - XU);
    }
    hr[k] = (short)lr;
    hc[k] = (short)lc;
    const int gy = min(max(y0 - kH + lr, 0), g.ny - 1);
    const int gx = min(max(x0 - kH + lc, 0), g.nx - VX);
    hoff[k] = (uint32_t)gy * g.nx + gx;
  }

  // The tile marches PZ planes per step.  The loads of step z + D*PZ (queue
  // leads, LDS halos, pointwise inputs) are issued before step z is computed:
  // with D = 2 a load has two steps' compute to arrive, and the wait at a
  // step's start leaves the previous step's loads in flight (the stores are
  // invisible to the compiler's wait model and the barriers do not fence
  // global memory, see st_b32 and lds_barrier).  Loads past the chunk's last
  // step are clamped to planes the chunk reads anyway.
  //
  // The LDS halo of a plane is loaded together with the plane's centre rows
  // (when the plane enters the z queue as a lead), not kH planes later: the
  // neighbouring tiles read those rows at about the same time, so they come
  // from L2.  It waits in its LDS ring slot, not in registers; hl holds the
  // halos of the leads in flight.
  static_assert(D == 1 || PZ == 1, "a two-step lookahead takes one plane per step");
  float q[NF][RY][VX][NQ];
  float lead[NB][PZ][NF][RY][VX];
  float hl[NB][PZ][NF][HPT][VX];
  auto put_halo = [&](int slot, int f, int k, const float (&h)[VX]) {
    float *d = &tile[slot][f][hr[k]][hc[k]];
    if constexpr (VX == 1)
      d[0] = h[0];
    else
      *(float2 *)d = make_float2(h[0], h[1]);
  };
  // pointwise inputs of the steps in flight, and of the step computed
  float pre[NB][PZ][RY][NP][VX], cur[PZ][RY][NP][VX];
  float tmax[NF];
  // every field of the current level at element offset o
  auto ld_all = [&](uint32_t o, float (&v)[NF][VX]) {
    if constexpr (Op::PAIRED) {
      op.template ldv2<VX>(blk, o, v);
    } else {
#pragma unroll
      for (int f = 0; f < NF; f++) op.template ldv<VX>(blk, f, o, v[f]);
    }
  };
  // the loads of the step at plane zs into buffer set b
  auto issue = [&](int zs, int b) {
#pragma unroll
    for (int p = 0; p < PZ; p++) {
      const uint32_t zl = (uint32_t)min(zs + kH + p, z1 - 1 + kH) * g.plane;
      const uint32_t zn = (uint32_t)min(zs + p, z1 - 1) * g.plane;
#pragma unroll
      for (int j = 0; j < RY; j++) {
        float w[NF][VX];
        ld_all(zl + row[j], w);
#pragma unroll
        for (int f = 0; f < NF; f++)
#pragma unroll
          for (int v = 0; v < VX; v++) lead[b][p][f][j][v] = w[f][v];
      }
#pragma unroll
      for (int k = 0; k < HPT; k++) {
        float w[NF][VX];
        ld_all(zl + hoff[k], w);
#pragma unroll
        for (int f = 0; f < NF; f++)
#pragma unroll
          for (int v = 0; v < VX; v++) hl[b][p][f][k][v] = w[f][v];
      }
#pragma unroll
      for (int j = 0; j < RY; j++) op.template prefetchv<VX>(blk, pre[b][p][j], zn + row[j]);
    }
  };
#pragma unroll
  for (int f = 0; f < NF; f++) tmax[f] = 0.f;
#pragma unroll
  for (int j = 0; j < RY; j++) {
#pragma unroll
    for (int k = 0; k < 2 * kH; k++) {
      float w[NF][VX];
      ld_all((uint32_t)(z0 - kH + k) * g.plane + row[j], w);
#pragma unroll
      for (int f = 0; f < NF; f++)
#pragma unroll
        for (int v = 0; v < VX; v++) q[f][j][v][k] = w[f][v];
    }
  }
#pragma unroll
  for (int p = 0; p < kH; p++) {
#pragma unroll
    for (int k = 0; k < HPT; k++) {
      float h[NF][VX];
      ld_all((uint32_t)(z0 + p) * g.plane + hoff[k], h);
#pragma unroll
      for (int f = 0; f < NF; f++)
        if (hv[k]) put_halo((z0 + p) % NR, f, k, h[f]);
    }
  }
#pragma unroll
  for (int b = 0; b < D; b++) issue(z0 + b * PZ, b);

  // The new values of a step are stored at the start of the next step,
  // after its wait and before its loads: the wait for a step's inputs then
  // finds stores and loads that were issued a whole step earlier.
  float pend[PZ][RY][NO][VX];
  int pend_z = -1;
  auto flush = [&]() {
    if (pend_z >= 0) {
#pragma unroll
      for (int p = 0; p < PZ; p++)
        if (pend_z + p < z1) {
          const uint32_t po = (uint32_t)(pend_z + p) * g.plane;
#pragma unroll
          for (int j = 0; j < RY; j++)
            if (in[j]) op.template storev<VX>(blk, po + row[j], pend[p][j]);
        }
    }
  };

  // One step at plane z from buffer set b; its loads go to set bn.  R = 1
  // for the second step of a Q2 pair, whose queue entries are one register
  // further on.  b and R are compile-time constants.
  auto step = [&](const int z, auto bc, auto rc) {
    constexpr int b = decltype(bc)::value;
    constexpr int R = decltype(rc)::value;
    constexpr int bn = Q2 ? 1 - b : b;
    // the pointwise inputs are copied out of a buffer that this step's loads
    // refill
    constexpr bool COPY = bn == b;
    float(&pw)[PZ][RY][NP][VX] = COPY ? cur : pre[b];
    if constexpr (COPY) {
#pragma unroll
      for (int p = 0; p < PZ; p++)
#pragma unroll
        for (int j = 0; j < RY; j++)
#pragma unroll
          for (int k = 0; k < NP; k++)
#pragma unroll
            for (int v = 0; v < VX; v++) cur[p][j][k][v] = pre[b][p][j][k][v];
    }
#pragma unroll
    for (int f = 0; f < NF; f++) {
#pragma unroll
      for (int j = 0; j < RY; j++) {
#pragma unroll
        for (int p = 0; p < PZ; p++)
#pragma unroll
          for (int v = 0; v < VX; v++) q[f][j][v][2 * kH + p + R] = lead[b][p][f][j][v];
#pragma unroll
        for (int p = 0; p < PZ; p++) {
          float *d = &tile[(z + p) % NR][f][kH + RY * ty + j][kH + VX * tx];
          if constexpr (VX == 1)
            d[0] = q[f][j][0][kH + p + R];
          else
            *(float2 *)d = make_float2(q[f][j][0][kH + p + R], q[f][j][1][kH + p + R]);
        }
      }
#pragma unroll
      for (int p = 0; p < PZ; p++)
#pragma unroll
        for (int k = 0; k < HPT; k++)
          if (hv[k]) put_halo((z + kH + p) % NR, f, k, hl[b][p][f][k]);
    }
    flush();
    issue(z + D * PZ, bn);
    lds_barrier();
#pragma unroll
    for (int p = 0; p < PZ; p++) {
      const int zz = z + p;
      if (PZ > 1 && zz >= z1) break;
      const int sl = zz % NR;
      const uint32_t zo = (uint32_t)zz * g.plane;
      const auto pl = op.plane(blk, zz);
      const float bzv = Op::DAMP ? ldc(sp.bz, zz) : 1.f;
      // Laplacians field by field.  The column above and below this
      // thread's RY consecutive rows is read from LDS once per field (kH
      // rows either side; the thread's own rows come from its z queues) and
      // shared by the RY cells.
      const int lr0 = kH + RY * ty, lc = kH + VX * tx;
      float lap[RY][VX][NF];
#pragma unroll
      for (int f = 0; f < NF; f++) {
        float col[RY + 2 * kH][VX];
#pragma unroll
        for (int i = 0; i < kH; i++) {
          const int ra = lr0 - kH + i, rb = lr0 + RY + i;
          if constexpr (VX == 1) {
            col[i][0] = tile[sl][f][ra][lc];
            col[kH + RY + i][0] = tile[sl][f][rb][lc];
          } else {
            const float2 a2 = *(const float2 *)&tile[sl][f][ra][lc];
            const float2 b2 = *(const float2 *)&tile[sl][f][rb][lc];
            col[i][0] = a2.x;
            col[i][1] = a2.y;
            col[kH + RY + i][0] = b2.x;
            col[kH + RY + i][1] = b2.y;
          }
        }
#pragma unroll
        for (int j = 0; j < RY; j++)
#pragma unroll
          for (int v = 0; v < VX; v++) col[kH + j][v] = q[f][j][v][kH + p + R];
#pragma unroll
        for (int j = 0; j < RY; j++) {
          if (!in[j]) continue;
          // the row from lc - 4 to lc + VX + 3
          const int lr = lr0 + j;
          float xs[2 * kH + VX];
          if constexpr (VX == 1) {
#pragma unroll
            for (int i = 0; i < 2 * kH + 1; i++) xs[i] = tile[sl][f][lr][lc - kH + i];
          } else {
#pragma unroll
            for (int i = 0; i < kH + 1; i++) {
              const float2 w = *(const float2 *)&tile[sl][f][lr][lc - kH + 2 * i];
              xs[2 * i] = w.x;
              xs[2 * i + 1] = w.y;
            }
          }
#pragma unroll
          for (int v = 0; v < VX; v++) {
            // centre, then the x, y and z pairs, each pair summed first
            float a = mul_r(t.c0, q[f][j][v][kH + p + R]);
#pragma unroll
            for (int k = 1; k <= kH; k++)
              a = fma_r(t.cx[k - 1], add_r(xs[kH + v - k], xs[kH + v + k]), a);
#pragma unroll
            for (int k = 1; k <= kH; k++)
              a = fma_r(t.cy[k - 1], add_r(col[kH + j - k][v], col[kH + j + k][v]), a);
#pragma unroll
            for (int k = 1; k <= kH; k++)
              a = fma_r(t.cz[k - 1],
                        add_r(q[f][j][v][kH + p - k + R], q[f][j][v][kH + p + k + R]), a);
            lap[j][v][f] = a;
          }
        }
      }
#pragma unroll
      for (int j = 0; j < RY; j++) {
        if (!in[j]) continue;
        float ctr[VX][NF];
#pragma unroll
        for (int f = 0; f < NF; f++)
#pragma unroll
          for (int v = 0; v < VX; v++) ctr[v][f] = q[f][j][v][kH + p + R];
#pragma unroll
        for (int v = 0; v < VX; v++) {
          CellArgs c;
          c.x = x + v;
          c.y = yy[j];
          c.z = zz;
          c.o = zo + row[j] + v;
          c.b = fminf(bxy[j][v], bzv);
          float pv[NP], o[NO];
#pragma unroll
          for (int k = 0; k < NP; k++) pv[k] = pw[p][j][k][v];
          op.cell(blk, pl, c, lap[j][v], ctr[v], pv, o, tmax);
#pragma unroll
          for (int k = 0; k < NO; k++) pend[p][j][k][v] = o[k];
        }
      }
    }
    pend_z = z;
    lds_barrier();
    // the queue moves by PZ planes a step, or by two after a Q2 pair
    constexpr int MV = Q2 ? (R == 1 ? 2 : 0) : PZ;
    if constexpr (MV > 0) {
#pragma unroll
      for (int f = 0; f < NF; f++)
#pragma unroll
        for (int j = 0; j < RY; j++)
#pragma unroll
          for (int v = 0; v < VX; v++)
#pragma unroll
            for (int k = 0; k < 2 * kH; k++) q[f][j][v][k] = q[f][j][v][k + MV];
    }
  };

  if constexpr (Q2) {
    for (int z = z0; z < z1; z += 2) {
      step(z, StepSet<0>{}, StepSet<0>{});
      if (z + 1 >= z1) break;
      step(z + 1, StepSet<1>{}, StepSet<1>{});
    }
  } else {
    for (int z = z0; z < z1; z += D * PZ) {
      step(z, StepSet<0>{}, StepSet<0>{});
      if constexpr (D == 2) {
        if (z + PZ >= z1) break;
        step(z + PZ, StepSet<1>{}, StepSet<0>{});
      }
    }
  }
  flush();
  // block maxima of |new level| for the next step's FP16 scales
  if constexpr (Op::NMAX > 0) {
    __shared__ float red[NF][WY];
#pragma unroll
    for (int f = 0; f < NF; f++) {
      float m = tmax[f];
      for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
      if (threadIdx.x % kTX == 0) red[f][threadIdx.x / kTX] = m;
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

// The step kernels take one plane per step; LA steps of load lookahead.
template <class Op, int RY, int WY, int VX, int LA, int LX>
__global__ void __launch_bounds__(kTX *WY)
    march_kernel(const Grid g, const Taps t, const Sponge sp, const Op op) {
  march_tile<Op, RY, WY, VX, 1, LX, LA>(g, t, sp, op,
                                        xcd_remap((int)blockIdx.x, (int)gridDim.x));
}

// Tiles of a batch of shots with their own grids: shot s owns the logical
// tiles [tile0[s], tile0[s + 1]).
constexpr int kMaxBatch = 64;
struct BatchTiles {
  int n;
  int zc;
  int tile0[kMaxBatch + 1];
};

// Tile grid of a domain for tiles of tw x ty cells and z chunk zc.
__host__ __device__ inline Grid tile_grid(int nx, int ny, int nz, int tw, int ty, int zc) {
  Grid g;
  g.nx = nx;
  g.ny = ny;
  g.nz = nz;
  g.plane = (uint32_t)nx * ny;
  g.ntx = (nx - kH + tw - 1) / tw;
  g.nty = (ny - 2 * kH + ty - 1) / ty;
  g.zc = zc;
  g.ntz = (nz - 2 * kH + zc - 1) / zc;
  return g;
}

// One launch advances every shot of a batch by one step.  BOp::shot(s, tw,
// ty, zc) returns {Grid, Sponge, Op} of shot s with the op's pointers moved to
// the shot's part of the batch arrays.
template <class BOp, int RY, int WY, int VX, int LA, int LX>
__global__ void __launch_bounds__(kTX *WY)
    march_batch_kernel(const BatchTiles bt, const Taps t, const BOp bop) {
  const int Lg = xcd_remap((int)blockIdx.x, (int)gridDim.x);
  int s = 0;
  while (s + 1 < bt.n && Lg >= bt.tile0[s + 1]) s++;
  const auto sh = bop.shot(s, LX * VX, RY * (kTX * WY / LX), bt.zc);
  march_tile<typename BOp::Op, RY, WY, VX, 1, LX, LA>(sh.g, t, sh.sp, sh.op, Lg - bt.tile0[s]);
}

}  // namespace born
