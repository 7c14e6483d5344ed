// Host-side launch interface of the fused step kernels (ops in born_ops.h, launchers in
// born_kernels.hip and the born_k_*.hip units).
#pragma once
#include <hip/hip_runtime.h>

#include "born_device.h"

namespace born {

constexpr int kMaxPts = 32;  // source points handled inside the step

struct Geometry {
  int nx, ny, nz;
  Taps taps;
  Sponge sponge;
};

// Tile shape: RY rows per thread, WY wavefronts per block, VX cells per lane
// along x (VX = 2 needs an even nx), LA steps of load lookahead (1 or 2),
// LX lanes per row (64, or 32: a wavefront covers two rows).  The tile is
// LX*VX x RY*(64*WY/LX) cells.
struct TileShape {
  int ry, wy, vx, la, lx;
};
constexpr int kNumShapes = 24;
extern const TileShape kShapes[kNumShapes];

// Launch configuration of one step kernel: tile shape and z planes per block.
struct Tune {
  int shape;
  int zchunk;
};

// Source points: n entries of int4 (x, y, z, 0) in device memory.
struct SrcPts {
  const void *p;
  int n;
};

// FP16 storage of the step's wavefield levels (and of the history): the field
// pointers then address __half arrays; see Ring in born_device.h.
struct Half {
  int on;  // 0: float levels; 1: FP16; 2: FP16 with a refinement byte (Split)
  Slots sl;
  float gain, tapsum;
};

struct FwdStep {
  const float *s1, *r1;
  float *s0, *r0;
  float *s_new;  // where the new s level goes (null: s0, in place)
  const float *vel, *model;
  // vel and model interleaved cell by cell (2n floats): the single-shot
  // forward reads them as one stream
  const float *vm;
  float *hist;
  SrcPts src;
  const float *amp;
  RecMap rec;
  float *tr;
  const float *tab;
  int id, extract, scatter;
  Half h;
  float *rg;  // FP16 ring block: s ring at 0, r ring at 8, max |model| at 16
  float *hscale;
};

// Two forward steps in one launch (FP32 levels): reads levels k-1 and k of
// s and r, writes levels k+1 (s_a, r_a) and k+2 (s_b, r_b); four distinct
// buffers per field.  *_a belong to step it, *_b to step it+1.
struct PairStep {
  const float *s_old, *s_cur, *r_old, *r_cur;
  float *s_a, *r_a, *s_b, *r_b;
  float *h_a, *h_b;  // optional history levels it and it+1
  const float *vel, *model;
  SrcPts src;
  const float *amp_a, *amp_b;
  RecMap rec;
  float *tr;
  const float *tab_a, *tab_b;
  int id_a, id_b, ext_a, ext_b;
  int scatter;
};
// Two adjoint-state steps in one launch (FP32): steps k and k-1 of the exact
// adjoint; nu levels j-1, j read, j+1 (u_a) and j+2 (u_b) written; four
// distinct buffers.  *_a belong to step k, *_b to step k-1.
struct LamPairStep {
  const float *u_old, *u_cur;
  float *u_a, *u_b;
  const float *vel, *sig_a, *sig_b;
  float *acc;
  RecMap rec;
  const float *tr;
  const float *tab_a, *tab_b;
  int id_a, id_b, inj_a, inj_b;
};
// Pair kernel tile: RY tile rows per thread, WY wavefronts (56 x RY*WY cells).
struct PairShape {
  int ry, wy;
};
constexpr int kNumPairShapes = 4;
extern const PairShape kPairShapes[kNumPairShapes];

struct SrcStep {
  const float *s1;
  float *s0;
  const float *vel;
  SrcPts src;
  const float *amp;
};

struct RadjStep {
  const float *s1, *r1;
  float *s0, *r0;
  const float *vel1, *vel2;
  float *img;
  int s_prop, r_prop;
  SrcPts src;
  const float *amp_pre;
  int pre;
  RecMap rec;
  const float *tr;
  const float *tab;
  float sc;
  int id, inject;
};

struct SrecStep {
  const float *s1;
  float *s0;
  float *s_new;  // where the new s level goes (null: s0, in place)
  const float *vel;
  float *sig;
  SrcPts src;
  const float *amp;
  int scatter;
  Half h;
  Ring ring;
  float *hscale;
};

struct LamStep {
  const float *u1;
  float *u0;
  const float *vel, *sig;
  float *acc;
  RecMap rec;
  const float *tr;
  const float *tab;
  int id, inject;
  Half h;
  Ring ring;
  float inj;
  const unsigned *trmax;
  const float *hscale;
};

// Placement of one shot of a batch inside the batch's arrays (elements).
struct ShotDesc {
  int nx, ny, nz;
  int npts;
  long long cell0;  // fields, velocity and model
  long long rmap0;  // receiver map
  long long tr0;    // traces, time-major [ntb][nrec]
  int sp0;          // sponge tables: x at sp0, y at sp0 + nx, z at sp0 + nx + ny
  int pts0;         // source points (and amplitude columns)
  int slot0;        // receiver plane slots (nz entries)
  int rnext0;       // receiver chains
  int nrec;
  int ring0;        // scaled storage: the shot's kRingWords words of rings
  float gain;       // bound gain of scaled storage: 2 + max(V) * tapsum
  float vmax;       // max(V) of the shot
  int per_cell;     // most receivers sharing one cell
};
// Ring words of a shot (scaled storage), from ShotDesc::ring0: the s ring
// (scales, maxima), the r ring, max |model|, the adjoint state's ring and
// max |trace|.
constexpr int kRingWords = 32;
enum { BR_S = 0, BR_R = 8, BR_MAXM = 16, BR_U = 20, BR_MAXTR = 28 };

// What every batched step shares: the shot placement and per-shot tables.
// The step's own arrays hold the whole batch (shot 0 at offset 0) and are
// moved per shot by the descriptors.
struct BatchCommon {
  int npts_total;
  const float *sponge;
  float *rings;
  const ShotDesc *desc;      // device copy
  const ShotDesc *host_desc; // host copy (tile counts)
  int nshots;
  int step;                  // amplitude row of the step
};
struct BatchFwdStep : BatchCommon {
  FwdStep a;  // a.amp: amplitudes [step][npts_total], shot s at column pts0
};
// background recompute of the batched exact adjoint; a.hscale (scaled
// storage): this level's scales, one per shot
struct BatchSrecStep : BatchCommon {
  SrecStep a;
};
// adjoint-state step of the batched exact adjoint; a.inj: sum |tab_d| of
// the receiver table row set (times each shot's max(V) and receivers per
// cell on the device); a.hscale: the sigma level's scales, one per shot
struct BatchLamStep : BatchCommon {
  LamStep a;
};

hipError_t launch_fwd_step(const Geometry &geo, const Tune &tu, const FwdStep &a,
                           hipStream_t st);
hipError_t launch_pair_step(const Geometry &geo, const Tune &tu, const PairStep &a,
                            hipStream_t st);
hipError_t launch_lam_pair_step(const Geometry &geo, const Tune &tu, const LamPairStep &a,
                                hipStream_t st);
hipError_t launch_batch_fwd_step(const Taps &taps, const Tune &tu, const BatchFwdStep &b,
                                 hipStream_t st);
hipError_t launch_batch_ring_init(float *rings, int nshots, hipStream_t st);
hipError_t launch_batch_srec_step(const Taps &taps, const Tune &tu, const BatchSrecStep &b,
                                  hipStream_t st);
hipError_t launch_batch_lam_step(const Taps &taps, const Tune &tu, const BatchLamStep &b,
                                 hipStream_t st);
// ring words [at, at + 4) = 1 (scales) and [at + 4, at + 8) = 0 (maxima) of
// every shot; with maxima_only, one word at `at` = 0
hipError_t launch_batch_ring_reset(float *rings, int nshots, int at, int maxima_only,
                                   hipStream_t st);
hipError_t launch_src_step(const Geometry &geo, const Tune &tu, const SrcStep &a,
                           hipStream_t st);
hipError_t launch_radj_step(const Geometry &geo, const Tune &tu, const RadjStep &a,
                            hipStream_t st);
hipError_t launch_srec_step(const Geometry &geo, const Tune &tu, const SrecStep &a,
                            hipStream_t st);
hipError_t launch_lam_step(const Geometry &geo, const Tune &tu, const LamStep &a,
                           hipStream_t st);
hipError_t launch_fill(float *p, long long n, float v, hipStream_t st);
// dst (2n floats) = a and b interleaved cell by cell
hipError_t launch_interleave(float *dst, const float *a, const float *b, long long n,
                             hipStream_t st);
hipError_t launch_src_amp(float *amp, int nsteps, int npts, const float *src, int ntsrc,
                          const float *tab, int nsinc, int jt, int ntab, int t_off,
                          float scale, hipStream_t st, int stride = -1);
hipError_t launch_adj_finish(const Geometry &geo, float *out, const float *acc,
                             const float *vel, float scale, hipStream_t st);
hipError_t launch_transpose(float *dst, const float *src, int n0, int n1, hipStream_t st);
hipError_t launch_ring_init(float *scale, unsigned *maxb, int n, hipStream_t st);
hipError_t launch_absmax(const float *p, long long n, unsigned *out, hipStream_t st);

}  // namespace born
