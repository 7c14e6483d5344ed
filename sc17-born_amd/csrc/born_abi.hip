// C ABI of the Born propagator (include/born_hip.h): context, transfers and
// the time loops of the reference backend (base/cpu_prop.cc), each step one
// launch of a fused kernel of born_ops.h.
//
// Exact adjoint.  The Born forward's scattered field obeys
//   r_{k+1} = A r_k - B^2 r_{k-1} + f_k,   A = B (V L + 2),  d_k = P_k r_{k+1}
// (B sponge, V = v^2 dt^2, L the symmetric interior stencil).  Its adjoint
// state lambda obeys lambda_{k+1} = A^T lambda_{k+2} - B^2 lambda_{k+3} +
// P_k^T w_k with A^T = (L V + 2) B.  The change of variables nu = V B lambda
// turns that into nu_{k+1} = B (V L nu_{k+2} + 2 nu_{k+2} - B nu_{k+3}) +
// V B P_k^T w_k: the forward kernel again, with V B-weighted receiver
// injection.  The model gradient sum_k sigma_k lambda_{k+1} is accumulated as
// sum_k sigma_k nu_{k+1} and divided by V B once at the end.  sigma_k (the
// scatter operand of forward step k) is needed in reverse order; it is
// rebuilt segment by segment from checkpoints of the background, with the
// number of segments the smallest that fits the HBM budget (one segment:
// the whole history is kept and nothing is recomputed).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "born_hip.h"
#include "born_launch.h"
#include "host_par.h"

using born::Geometry;

struct born_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;

  bool have_space = false;
  int nx = 0, ny = 0, nz = 0;
  long long n = 0;
  float coeffs[13] = {};
  int nbound = 40;
  float bcB = .0005f;
  float *d_bx = nullptr, *d_by = nullptr, *d_bz = nullptr;
  int zchunk = 0;  // override of the tuned z chunk (0: tuned)
  int autotune = 1;
  born::Tune tune[24];
  float tune_ms[24];
  int pairs = 0;  // two-step forward launches for FP32 levels: 0 off, 1 auto, 2 on

  int nsinc_s = 0, jts = 0, nsinc_d = 0, jtd = 0;
  float *d_tab_s = nullptr, *d_tab_d = nullptr;
  // bumped by every sinc-table transfer: a captured batch time loop holds
  // the table's device address and is recaptured when this changes
  long long tab_version = 0;

  int npts_s = 0, nt_src = 0;
  float *d_src = nullptr;
  int *d_pts = nullptr;  // int4 (x, y, z, 0) per source point

  float *d_vel1 = nullptr, *d_vel2 = nullptr;

  int rec_nx = 0, rec_ny = 0, nrec = 0, nt_rec = 0;
  float *h_rec = nullptr;
  int *d_plane_slot = nullptr, *d_rmap = nullptr, *d_rnext = nullptr;
  int nslots = 0;

  int jt_state = 1;

  // wavefield storage of the Born forward / exact adjoint (BORN_STORAGE_*)
  int storage = BORN_STORAGE_F32;
  float vmax2 = 0.f;      // max of velocity 2 (V = v^2 dt^2)
  float vmin2 = 0.f;      // min of velocity 2 over the stencil interior
  float bmin = 1.f;       // smallest sponge factor
  float tabd_abs = 0.f;   // largest sum |tab_d[ii][k]| over the rows
  int rec_per_cell = 0;   // most receivers sharing one cell
  float *d_ring = nullptr;    // FP16 level scales, see ring()
  float *d_hscale = nullptr;  // FP16 scale per history level
  size_t hscale_cap = 0;
  float *d_ckring = nullptr;  // ring snapshot per checkpoint (8 words each)
  size_t ckring_cap = 0;

  float *d_w[4] = {nullptr, nullptr, nullptr, nullptr};
  size_t w_cap = 0;  // floats per work buffer
  float *d_w2[4] = {nullptr, nullptr, nullptr, nullptr};  // levels k+1, k+2 of pairs
  float *d_img = nullptr, *d_acc = nullptr, *d_p0 = nullptr, *d_p1 = nullptr;
  float *d_vm = nullptr;  // velocity 2 and the forward's model interleaved per cell
  size_t vm_cap = 0;
  float *d_tr = nullptr, *d_tr2 = nullptr;
  size_t tr_cap = 0;
  float *d_amp = nullptr;
  size_t amp_cap = 0;

  size_t budget = 0;
  float *d_seg = nullptr;
  size_t seg_cap = 0;  // fields
  // The step kernels write interior cells only, so every level they read
  // must hold zeros in its 4-cell shell.  seg_clean bytes from the start of
  // d_seg are zero outside the interiors of levels of seg_nx x seg_ny x
  // seg_nz cells of seg_esz bytes (cleared when that layout changes).
  size_t seg_clean = 0, seg_esz = 0;
  int seg_nx = -1, seg_ny = -1, seg_nz = -1;
  float *d_ckpt = nullptr;
  size_t ckpt_cap = 0;  // fields
  long long bg_version = 1;
  long long hist_version = -1;
  int hist_scatter = -1, hist_jt = -1, hist_nt = -1, hist_storage = -1;
  // a checkpoint plan of S segments of L levels left by the forward (FP32
  // levels): checkpoints of segments 1..S-1 in d_ckpt, the last segment's
  // scatter operands in d_seg; valid for ck_version == bg_version
  long long ck_version = -1;
  int ck_S = 0, ck_L = 0, ck_scatter = -1, ck_jt = -1, ck_nt = -1;
  int last_segments = 0, last_recomputed = 0, last_swept = 0;
  int last_pairs = 0;
  int last_lam_pairs = 0;

  // Placement of the work pair: the two levels of r of the FP32 forward (of
  // s and r together with narrow storage) and of nu of the adjoint state.
  // A step's launch time depends on where its streams sit in HBM relative
  // to one another (at 304^3 the FP32 forward 157-178 us and the 24-bit
  // forward 162-194 us for the same kernel and shape, by 32 GiB windows of
  // history offset), so a history longer than one window is laid out in
  // windows of seg_win bytes at fixed offsets of d_seg, each a gap for two
  // levels of the pair (and, for the adjoint state, the image accumulator)
  // followed by seg_lpw levels, and the pair sits in the work buffers or
  // the gap timed fastest for the window of the history levels a step
  // touches.  place_mode 0 keeps the plain layout (no gaps),
  // 1 times the places, 2 cycles through them (tests).
  int place_mode = 1;
  size_t place_window = (size_t)32 << 30;  // history bytes per window
  size_t seg_gap = 0;     // bytes of a gap of d_seg (0: no gaps)
  size_t seg_win = 0;     // bytes per window
  long long seg_lpw = 0;  // levels per window
  int seg_ngaps = 0;
  // the place timed fastest per window (-2: not timed) and what it was timed
  // for: the history buffer and window size, the grid, the other arrays the
  // step streams (their places matter too) and the launch shape
  std::vector<int> place_fwd, place_lam;
  std::vector<uintptr_t> place_fwd_key, place_lam_key;
  int last_fwd_moves = 0, last_lam_moves = 0;

  // this device's scratch array of born_dev_source_prop_noalloc (resolved
  // at born_create), or the reason it could not be resolved
  float *scratch = nullptr;
  std::string scratch_err;
};

namespace {

int fail(born_ctx *c, int code, const std::string &m) {
  if (c) c->err = m;
  return code;
}

#define HIPCHK(expr)                                                       \
  do {                                                                     \
    hipError_t e_ = (expr);                                                \
    if (e_ != hipSuccess)                                                  \
      return fail(ctx, BORN_ERR_HIP,                                       \
                  std::string(#expr) + ": " + hipGetErrorString(e_));      \
  } while (0)

#define CHK(expr)                   \
  do {                              \
    int r_ = (expr);                \
    if (r_ != BORN_OK) return r_;   \
  } while (0)

template <class T>
int dev_alloc(born_ctx *ctx, T **p, size_t count) {
  if (*p) {
    // work queued on the stream may still read the old buffer
    (void)hipStreamSynchronize(ctx->stream);
    (void)hipFree(*p);
    *p = nullptr;
  }
  if (count == 0) return BORN_OK;
  hipError_t e = hipMalloc((void **)p, count * sizeof(T));
  if (e != hipSuccess) {
    *p = nullptr;
    (void)hipGetLastError();
    return fail(ctx, BORN_ERR_NOMEM,
                "hipMalloc of " + std::to_string(count * sizeof(T)) +
                    " bytes failed: " + hipGetErrorString(e));
  }
  return BORN_OK;
}

// grow-only device buffer of floats
int ensure(born_ctx *ctx, float **p, size_t *cap, size_t need) {
  if (*cap >= need && *p) return BORN_OK;
  *cap = 0;
  CHK(dev_alloc(ctx, p, need));
  *cap = need;
  return BORN_OK;
}

int set_device(born_ctx *ctx) {
  HIPCHK(hipSetDevice(ctx->device));
  return BORN_OK;
}

Geometry geometry(const born_ctx *c) {
  Geometry g;
  g.nx = c->nx;
  g.ny = c->ny;
  g.nz = c->nz;
  // reference tap indices: C0=0, CZk=1+3(k-1), CXk=2+3(k-1), CYk=3+3(k-1)
  // (base/cpu_prop.cc:6-18)
  g.taps.c0 = c->coeffs[0];
  for (int k = 0; k < 4; k++) {
    g.taps.cz[k] = c->coeffs[1 + 3 * k];
    g.taps.cx[k] = c->coeffs[2 + 3 * k];
    g.taps.cy[k] = c->coeffs[3 + 3 * k];
  }
  g.sponge.bx = c->d_bx;
  g.sponge.by = c->d_by;
  g.sponge.bz = c->d_bz;
  return g;
}

born::RecMap recmap(const born_ctx *c) {
  born::RecMap m;
  m.plane_slot = c->d_plane_slot;
  m.rmap = c->d_rmap;
  m.rnext = c->d_rnext;
  m.nrec = c->nrec;
  m.ntb = c->nt_rec;
  return m;
}

// step kernels that are tuned separately
enum Kind {
  K_FWD0, K_FWD1, K_SRC, K_RADJ, K_SREC0, K_SREC1, K_LAM,
  K_FWD0H, K_FWD1H, K_SREC0H, K_SREC1H, K_LAMH, K_FWD0P, K_FWD1P, K_LAMP,
  K_FWD0S, K_FWD1S, K_SREC0S, K_SREC1S, K_LAMS, K_NUM
};
static_assert(K_NUM <= 24, "tuning slots");
// the step kernels of a storage form (BORN_STORAGE_*)
int fwd_kind(int storage, int scatter) {
  static const int k[3][2] = {{K_FWD0, K_FWD1}, {K_FWD0H, K_FWD1H}, {K_FWD0S, K_FWD1S}};
  return k[storage][scatter];
}
int srec_kind(int storage, int scatter) {
  static const int k[3][2] = {{K_SREC0, K_SREC1}, {K_SREC0H, K_SREC1H}, {K_SREC0S, K_SREC1S}};
  return k[storage][scatter];
}
int lam_kind(int storage) { return storage == 2 ? K_LAMS : storage == 1 ? K_LAMH : K_LAM; }
bool is_pair(int kind) { return kind == K_FWD0P || kind == K_FWD1P || kind == K_LAMP; }

born::SrcPts points(const born_ctx *c) { return {c->d_pts, c->npts_s}; }
born::SrcPts no_points() { return {nullptr, 0}; }

void reset_tuning(born_ctx *c) {
  for (int k = 0; k < K_NUM; k++) {
    // 64x32 tiles of 2 rows per thread, 64 planes; pairs: 56x16, 128 planes
    c->tune[k] = is_pair(k) ? born::Tune{0, 128} : born::Tune{3, 64};
    c->tune_ms[k] = -1.f;
  }
  c->place_fwd_key.clear();
  c->place_lam_key.clear();
}

// A shape of two cells per lane needs an even nx.  For an odd nx (a shape
// fixed by born_set_tuning or tuned on another grid) take the one-cell shape
// of the same tile and lookahead, else 2 x 16 rows; every shape computes
// the same bits, so only the speed changes.
born::Tune fit_nx(born::Tune t, bool even_nx) {
  const born::TileShape &ts = born::kShapes[t.shape];
  if (ts.vx == 1 || even_nx) return t;
  const int rows = ts.ry * (born::kTX * ts.wy / ts.lx);
  int pick = 3;
  for (int sh = born::kNumShapes - 1; sh >= 0; sh--) {
    const born::TileShape &c = born::kShapes[sh];
    if (c.vx == 1 && c.la == ts.la && c.ry * c.wy == rows) pick = sh;
  }
  t.shape = pick;
  return t;
}

born::Tune tuning(const born_ctx *c, int kind) {
  born::Tune t = c->tune[kind];
  if (c->zchunk > 0) t.zchunk = c->zchunk;
  return fit_nx(t, c->nx % 2 == 0);
}

// Autotuner: time every tile shape x z chunk for one step kernel and keep
// the fastest in *tune / *tune_ms (ms per launch; the four fastest of a
// first pass are timed again, round-robin, over more launches).  run(tune) must launch the kernel on
// scratch buffers only.  Pair shapes (2 cells per lane) need even_nx.
//
// z chunks: 32, 64 and 128 planes and, for nzi > 0 interior planes, every
// length that splits nzi evenly into k = 1..16 chunks (ceil(nzi / k), at
// least 16 planes), so that no block marches more planes than the others
// (at 296 planes, chunks of 60 end with one of 56 where chunks of 64 end
// with one of 40) and the number of blocks can match what the CUs hold at
// once.
constexpr int kMaxZChunks = 19;
static int zchunk_candidates(int nzi, int (&zcs)[kMaxZChunks]) {
  int n = 0;
  auto add = [&](int zc) {
    for (int i = 0; i < n; i++)
      if (zcs[i] == zc) return;
    if (n < kMaxZChunks) zcs[n++] = zc;
  };
  for (int target : {32, 64, 128}) add(target);
  if (nzi > 0)
    for (int k = 1; k <= 16; k++) {
      const int zc = (nzi + k - 1) / k;
      if (zc >= 16) add(zc);
    }
  return n;
}

template <class F>
int autotune_into(born_ctx *ctx, born::Tune *tune, float *tune_ms, bool even_nx, int nzi,
                  F run) {
  if (!ctx->autotune || *tune_ms >= 0.f) return BORN_OK;
  int zcs[kMaxZChunks];
  const int nzc = zchunk_candidates(nzi, zcs);
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  // ms per launch of `reps` launches after one untimed launch
  auto time = [&](const born::Tune &t, int reps, float *ms) -> int {
    HIPCHK(run(t));
    HIPCHK(hipEventRecord(e0, ctx->stream));
    for (int r = 0; r < reps; r++) HIPCHK(run(t));
    HIPCHK(hipEventRecord(e1, ctx->stream));
    HIPCHK(hipEventSynchronize(e1));
    HIPCHK(hipEventElapsedTime(ms, e0, e1));
    *ms /= reps;
    return BORN_OK;
  };
  // every candidate over 3 launches
  std::vector<std::pair<float, born::Tune>> cand;
  for (int sh = 0; sh < born::kNumShapes; sh++)
    for (int zi = 0; zi < nzc; zi++) {
      if (born::kShapes[sh].vx == 2 && !even_nx) continue;
      const born::Tune t{sh, ctx->zchunk > 0 ? ctx->zchunk : zcs[zi]};
      if (ctx->zchunk > 0 && zi > 0) continue;
      float ms = 0.f;
      CHK(time(t, 3, &ms));
      cand.push_back({ms, t});
    }
  std::sort(cand.begin(), cand.end(),
            [](const auto &a, const auto &b) { return a.first < b.first; });
  // the four fastest again, round-robin over three rounds of 30 launches
  // (a drift of the clocks during the search then favours none of them)
  const size_t top = std::min<size_t>(cand.size(), 4);
  float sum[4] = {0.f, 0.f, 0.f, 0.f};
  for (int round = 0; round < 3; round++)
    for (size_t i = 0; i < top; i++) {
      float ms = 0.f;
      CHK(time(cand[i].second, 30, &ms));
      sum[i] += ms / 3.f;
    }
  float best = 1e30f;
  born::Tune bt = *tune;
  for (size_t i = 0; i < top; i++)
    if (sum[i] < best) {
      best = sum[i];
      bt = cand[i].second;
    }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *tune = bt;
  *tune_ms = best;
  return BORN_OK;
}

template <class F>
int autotune(born_ctx *ctx, int kind, F run) {
  return autotune_into(ctx, &ctx->tune[kind], &ctx->tune_ms[kind], ctx->nx % 2 == 0,
                       ctx->nz - 2 * born::kH, run);
}

// The same for a two-step kernel: its shapes and z chunks that split the
// interior evenly (a chunk recomputes 8 planes of its neighbours, so chunks
// are long).  tune_ms is per launch (2 steps).
template <class F>
int autotune_pair(born_ctx *ctx, int kind, F run) {
  float *tune_ms = &ctx->tune_ms[kind];
  if (!ctx->autotune || *tune_ms >= 0.f) return BORN_OK;
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  float best = 1e30f;
  born::Tune bt = ctx->tune[kind];
  // chunks of equal length: every block marches the same number of planes
  int zcs[6], nzc = 0;
  const int nzi = std::max(ctx->nz - 8, 1);
  for (int k = 1; k <= 6 && nzc < 6; k++) {
    const int zc = (nzi + k - 1) / k;
    if (zc >= 16 && (nzc == 0 || zcs[nzc - 1] != zc)) zcs[nzc++] = zc;
  }
  if (nzc == 0) zcs[nzc++] = nzi;
  for (int sh = 0; sh < born::kNumPairShapes; sh++)
    for (int zi = 0; zi < nzc; zi++) {
      const born::Tune t{sh, ctx->zchunk > 0 ? ctx->zchunk : zcs[zi]};
      if (ctx->zchunk > 0 && zi > 0) continue;
      HIPCHK(run(t));  // warm
      HIPCHK(hipEventRecord(e0, ctx->stream));
      for (int r = 0; r < 3; r++) HIPCHK(run(t));
      HIPCHK(hipEventRecord(e1, ctx->stream));
      HIPCHK(hipEventSynchronize(e1));
      float ms = 0.f;
      HIPCHK(hipEventElapsedTime(&ms, e0, e1));
      if (ms < best) {
        best = ms;
        bt = t;
      }
    }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  ctx->tune[kind] = bt;
  *tune_ms = best / 3.f;
  return BORN_OK;
}

// Sponge factor per axis coordinate (cpu_prop.cc:141-165, 329-332).
std::vector<float> axis_sponge(int n, int width, const std::vector<float> &bound) {
  std::vector<float> b(n, 1.f);
  for (int i = 4; i < n - 4; i++) {
    int e = std::min(i - 4, n - 4 - i);
    if (e >= 0 && e < width) b[i] = bound[e];
  }
  return b;
}

int upload_sponge(born_ctx *ctx) {
  std::vector<float> bound(ctx->nbound);
  const float bcA = (float)ctx->nbound;
  for (int i = 0; i < ctx->nbound; i++) bound[i] = expf(-ctx->bcB * (bcA - i));
  ctx->bmin = 1.f;
  for (float b : bound) ctx->bmin = std::min(ctx->bmin, b);
  const int dims[3] = {ctx->nx, ctx->ny, ctx->nz};
  float **dst[3] = {&ctx->d_bx, &ctx->d_by, &ctx->d_bz};
  for (int a = 0; a < 3; a++) {
    std::vector<float> b = axis_sponge(dims[a], ctx->nbound, bound);
    CHK(dev_alloc(ctx, dst[a], b.size()));
    HIPCHK(hipMemcpy(*dst[a], b.data(), b.size() * sizeof(float),
                     hipMemcpyHostToDevice));
  }
  return BORN_OK;
}

bool interior(const born_ctx *c, long long loc, int *x, int *y, int *z) {
  if (loc < 0 || loc >= c->n) return false;
  const long long n12 = (long long)c->nx * c->ny;
  *z = (int)(loc / n12);
  *y = (int)((loc % n12) / c->nx);
  *x = (int)(loc % c->nx);
  return *x >= 4 && *x < c->nx - 4 && *y >= 4 && *y < c->ny - 4 && *z >= 4 &&
         *z < c->nz - 4;
}

int need_space(born_ctx *ctx) {
  if (!ctx) return BORN_ERR_ARG;
  if (!ctx->have_space)
    return fail(ctx, BORN_ERR_STATE, "born_create_space has not been called");
  return BORN_OK;
}

int check_dims(born_ctx *ctx, int n1, int n2, int n3) {
  if (n1 != ctx->nx || n2 != ctx->ny || n3 != ctx->nz)
    return fail(ctx, BORN_ERR_ARG,
                "grid " + std::to_string(n1) + "x" + std::to_string(n2) + "x" +
                    std::to_string(n3) + " differs from the created space");
  return BORN_OK;
}

// floats per work buffer: one float level, or a 24-bit level of two fields
// stored together (6 bytes per cell)
size_t work_floats(const born_ctx *c) {
  return c->storage == BORN_STORAGE_F24 ? ((size_t)c->n * 3 + 1) / 2 : (size_t)c->n;
}

int ensure_work(born_ctx *ctx) {
  const size_t need = work_floats(ctx);
  if (ctx->w_cap < need) {
    ctx->w_cap = 0;
    for (int i = 0; i < 4; i++) CHK(dev_alloc(ctx, &ctx->d_w[i], need));
    ctx->w_cap = need;
  }
  // the forward's interleaved velocity and model, allocated with the work
  // buffers so that it exists before any history is sized to the free memory
  CHK(ensure(ctx, &ctx->d_vm, &ctx->vm_cap, 2 * (size_t)ctx->n));
  return BORN_OK;
}

// 24-bit levels of two fields are addressed with 32-bit byte offsets up to
// 6 bytes per cell
int check_storage(born_ctx *ctx) {
  if (ctx->storage == BORN_STORAGE_F24 && (unsigned long long)ctx->n * 6 >= (1ull << 32))
    return fail(ctx, BORN_ERR_UNSUPPORTED,
                "24-bit storage needs fewer than 715,827,882 cells (32-bit byte offsets)");
  return BORN_OK;
}

int ensure_pair_work(born_ctx *ctx) {
  for (int i = 0; i < 4; i++)
    if (!ctx->d_w2[i]) CHK(dev_alloc(ctx, &ctx->d_w2[i], (size_t)ctx->n));
  return BORN_OK;
}

int zero(born_ctx *ctx, float *p, long long count) {
  HIPCHK(born::launch_fill(p, count, 0.f, ctx->stream));
  return BORN_OK;
}

// injection amplitudes of the source for nsteps steps (see src_amp_kernel)
int source_amps(born_ctx *ctx, int nsteps, int jt_idx, int t_off, float scale) {
  const size_t need = (size_t)std::max(nsteps, 1) * std::max(ctx->npts_s, 1);
  CHK(ensure(ctx, &ctx->d_amp, &ctx->amp_cap, need));
  HIPCHK(born::launch_src_amp(ctx->d_amp, nsteps, ctx->npts_s, ctx->d_src,
                              ctx->nt_src, ctx->d_tab_s, ctx->nsinc_s, jt_idx,
                              ctx->jts, t_off, scale, ctx->stream));
  return BORN_OK;
}

int need_born_inputs(born_ctx *ctx, bool vel1, bool vel2, bool rec) {
  CHK(need_space(ctx));
  if (!ctx->d_tab_s || !ctx->d_tab_d)
    return fail(ctx, BORN_ERR_STATE, "sinc tables have not been transferred");
  if (!ctx->d_src)
    return fail(ctx, BORN_ERR_STATE, "source function has not been transferred");
  if (vel1 && !ctx->d_vel1)
    return fail(ctx, BORN_ERR_STATE, "velocity 1 has not been transferred");
  if (vel2 && !ctx->d_vel2)
    return fail(ctx, BORN_ERR_STATE, "velocity 2 has not been transferred");
  if (rec && !ctx->d_rmap)
    return fail(ctx, BORN_ERR_STATE, "receivers have not been transferred");
  return BORN_OK;
}

size_t free_bytes(born_ctx *ctx) {
  size_t fr = 0, tot = 0;
  if (hipMemGetInfo(&fr, &tot) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return fr;
}

// Bytes available for background history: the set budget, or the free
// memory plus what the history buffers already hold, less a reserve.
size_t history_budget(born_ctx *ctx) {
  if (ctx->budget) return ctx->budget;
  const size_t held = (ctx->seg_cap + ctx->ckpt_cap) * sizeof(float);
  const size_t reserve = (size_t)2 << 30;
  const size_t fr = free_bytes(ctx) + held;
  return fr > reserve ? fr - reserve : 0;
}

// Layout of a segment history of L levels of esz bytes: a history longer
// than one window is laid out in windows of W bytes at offsets 0, W, 2W,
// ...: a gap (a place of the work pair, see born_ctx), then lpw levels.
struct SegLayout {
  size_t gap;     // bytes of a gap, 0: no gaps
  size_t win;     // bytes per window
  long long lpw;  // levels per window
  int ngaps;
  size_t bytes;   // the whole history
};
SegLayout seg_layout(const born_ctx *ctx, size_t L, size_t esz) {
  const size_t F = (size_t)ctx->n * esz;
  SegLayout s{0, 0, 0, 0, L * F};
  const size_t W = ctx->place_window;
  // a gap holds two levels of the forward's work pair (FP32 r levels, or
  // the s and r levels that narrow storage keeps together) or the adjoint
  // state's two levels and the FP32 image accumulator
  const size_t G = esz == 4 ? 3 * F : 4 * F;
  if (ctx->place_mode == 0 || F == 0 || W < G + F) return s;
  const long long lpw = (long long)((W - G) / F);
  if ((long long)L <= lpw) return s;
  s.gap = G;
  s.win = W;
  s.lpw = lpw;
  s.ngaps = (int)(((long long)L - 1) / lpw + 1);
  s.bytes = (size_t)(s.ngaps - 1) * W + s.gap + ((long long)L - (s.ngaps - 1) * lpw) * F;
  return s;
}

// level k of the segment history, in the layout of the last history_alloc
float *seg_level(const born_ctx *ctx, long long k) {
  const size_t F = (size_t)ctx->n * ctx->seg_esz;
  const size_t off = ctx->seg_gap ? (size_t)(k / ctx->seg_lpw) * ctx->seg_win + ctx->seg_gap +
                                        (size_t)(k % ctx->seg_lpw) * F
                                  : (size_t)k * F;
  return (float *)((char *)ctx->d_seg + off);
}

// the i-th field of `step` bytes in gap g of the segment history
float *seg_gap_field(const born_ctx *ctx, int g, int i, size_t step) {
  return (float *)((char *)ctx->d_seg + (size_t)g * ctx->seg_win + (size_t)i * step);
}

// history buffers: L levels of scatter operand + nck checkpoint levels, each
// n elements of esz bytes (capacities are counted in floats)
int history_alloc(born_ctx *ctx, size_t L, size_t nck, size_t esz) {
  const SegLayout lay = seg_layout(ctx, L, esz);
  const size_t bytes = lay.bytes;
  const size_t seg = (bytes + 3) / 4;
  const size_t ck = (nck * (size_t)ctx->n * esz + 3) / 4;
  if (ctx->seg_cap < seg) {
    ctx->seg_cap = 0;
    ctx->seg_clean = 0;
    ctx->hist_version = -1;
    CHK(dev_alloc(ctx, &ctx->d_seg, seg));
    ctx->seg_cap = seg;
  }
  // shells of the levels about to be used: zero unless cleared for this
  // layout (levels of another grid, element size or gap layout leave
  // values there)
  if (ctx->seg_nx != ctx->nx || ctx->seg_ny != ctx->ny || ctx->seg_nz != ctx->nz ||
      ctx->seg_esz != esz || ctx->seg_gap != lay.gap || ctx->seg_win != lay.win) {
    ctx->seg_clean = 0;
    ctx->seg_nx = ctx->nx;
    ctx->seg_ny = ctx->ny;
    ctx->seg_nz = ctx->nz;
    ctx->seg_esz = esz;
    ctx->seg_gap = lay.gap;
    ctx->seg_win = lay.win;
    ctx->seg_lpw = lay.lpw;
  }
  ctx->seg_ngaps = lay.ngaps;
  if (ctx->seg_clean < bytes) {
    const size_t from = ctx->seg_clean / 4;
    CHK(zero(ctx, ctx->d_seg + from, (long long)(seg - from)));
    ctx->seg_clean = seg * 4;
  }
  if (ctx->ckpt_cap < ck) {
    ctx->ckpt_cap = 0;
    CHK(dev_alloc(ctx, &ctx->d_ckpt, ck));
    ctx->ckpt_cap = ck;
  }
  return BORN_OK;
}

// The checkpoint plan of an exact adjoint over nt steps: the fewest
// segments S whose L = ceil(nt / S) levels of scatter operand (with the
// layout's gaps) plus 2 (S - 1) checkpoint levels fit the history budget.
int checkpoint_plan(born_ctx *ctx, int nt, size_t esz, int *S, int *L) {
  const size_t F = (size_t)ctx->n * esz;
  const size_t avail = history_budget(ctx);
  int segs = 0;
  for (int sg = 1; sg <= nt; sg++) {
    const size_t l = (size_t)((nt + sg - 1) / sg);
    if (seg_layout(ctx, l, esz).bytes + 2 * (size_t)(sg - 1) * F <= avail) {
      segs = sg;
      break;
    }
    if (l <= 2) break;
  }
  if (segs == 0) return fail(ctx, BORN_ERR_NOMEM, "no checkpoint plan fits the HBM budget");
  *L = (nt + segs - 1) / segs;
  *S = (nt + *L - 1) / *L;
  return BORN_OK;
}

int tr_buffers(born_ctx *ctx) {
  const size_t need = (size_t)ctx->nrec * ctx->nt_rec;
  if (ctx->tr_cap < need) {
    ctx->tr_cap = 0;
    CHK(dev_alloc(ctx, &ctx->d_tr, need));
    CHK(dev_alloc(ctx, &ctx->d_tr2, need));
    ctx->tr_cap = need;
  }
  return BORN_OK;
}

// host [nrec][ntb] -> device time-major [ntb][nrec] in ctx->d_tr
int traces_in(born_ctx *ctx, const float *h) {
  CHK(tr_buffers(ctx));
  const size_t bytes = (size_t)ctx->nrec * ctx->nt_rec * sizeof(float);
  HIPCHK(hipMemcpyAsync(ctx->d_tr2, h, bytes, hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(born::launch_transpose(ctx->d_tr, ctx->d_tr2, ctx->nrec, ctx->nt_rec,
                                ctx->stream));
  return BORN_OK;
}

int traces_out(born_ctx *ctx, float *h) {
  const size_t bytes = (size_t)ctx->nrec * ctx->nt_rec * sizeof(float);
  HIPCHK(born::launch_transpose(ctx->d_tr2, ctx->d_tr, ctx->nt_rec, ctx->nrec,
                                ctx->stream));
  HIPCHK(hipMemcpyAsync(h, ctx->d_tr2, bytes, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return BORN_OK;
}

// bytes per cell of one stored level
size_t elem(const born_ctx *c) {
  return c->storage == BORN_STORAGE_F16 ? 2 : c->storage == BORN_STORAGE_F24 ? 3 : 4;
}

// k-th field of element size esz in an array of fields of n cells
float *nth(float *base, long long k, long long n, size_t esz) {
  return (float *)((char *)base + (size_t)k * n * esz);
}

// Moves the pair (a, b) of levels of `bytes` each to place `to` (-1: the
// work buffers w0, w1; g >= 0: gap g of the segment history): two copies on
// the stream, none when the pair is there already.
int move_pair(born_ctx *ctx, int to, float *w0, float *w1, float *&a, float *&b, size_t bytes) {
  float *d0 = to < 0 ? w0 : seg_gap_field(ctx, to, 0, bytes);
  float *d1 = to < 0 ? w1 : seg_gap_field(ctx, to, 1, bytes);
  if ((a == d0 && b == d1) || (a == d1 && b == d0)) return BORN_OK;
  HIPCHK(hipMemcpyAsync(d0, a, bytes, hipMemcpyDeviceToDevice, ctx->stream));
  HIPCHK(hipMemcpyAsync(d1, b, bytes, hipMemcpyDeviceToDevice, ctx->stream));
  a = d0;
  b = d1;
  return BORN_OK;
}

// The place of the work pair in each of the nwin windows of the current
// layout, into use[w] (-1: the work buffers, g >= 0: gap g).  place_mode 1:
// the place timed fastest over 2 launches of run(window, place, i), i = 1,
// 2, ... the launch's position in a march (after an untimed i = 0); a choice
// is kept in `cache` and timed again only when `key` changes or the gap it
// names is past this layout's; without autotuning, the work buffers.
// The differences it finds are 5-20 % and come in classes: the gaps that
// time within 2 % of each other on the first window timed give the same
// time on every window (they sit in the same kind of HBM window), so that
// window times every place and the later ones the work buffers and one gap
// of each class, the fastest of the class on the first.
// place_mode 2: window w takes place w mod (gaps + 1) - 1, every place in
// turn.
template <class F>
int place_plan(born_ctx *ctx, std::vector<int> &cache, std::vector<uintptr_t> &held,
               const std::vector<uintptr_t> &key, int nwin, std::vector<int> &use, F run,
               int ncand = -1) {
  // places: -1 the work buffers, 0 .. ng - 1 the gaps (and, where the
  // caller has variants of a gap, ng .. ncand - 1)
  const int ng = ncand < 0 ? ctx->seg_ngaps : ncand;
  use.assign(nwin, -1);
  if (ctx->place_mode == 2) {
    for (int w = 0; w < nwin; w++) use[w] = w % (ng + 1) - 1;
    return BORN_OK;
  }
  if (!ctx->autotune) return BORN_OK;
  if (held != key) {
    cache.clear();
    held = key;
  }
  if ((int)cache.size() < nwin) cache.resize(nwin, -2);  // -2: not timed
  constexpr int reps = 2;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  auto time_gap = [&](int w, int g, float *ms) -> int {
    if (!e0) {
      HIPCHK(hipEventCreate(&e0));
      HIPCHK(hipEventCreate(&e1));
    }
    HIPCHK(run(w, g, 0));
    HIPCHK(hipEventRecord(e0, ctx->stream));
    for (int i = 1; i <= reps; i++) HIPCHK(run(w, g, i));
    HIPCHK(hipEventRecord(e1, ctx->stream));
    HIPCHK(hipEventSynchronize(e1));
    HIPCHK(hipEventElapsedTime(ms, e0, e1));
    return BORN_OK;
  };
  // the work buffers, then one gap per class (after the first timed window)
  std::vector<int> classes;
  int rc = BORN_OK;
  for (int w = 0; w < nwin && rc == BORN_OK; w++) {
    if (cache[w] < -1 || cache[w] >= ng) {
      if (classes.empty()) {
        float tw = 0.f;
        rc = time_gap(w, -1, &tw);
        std::vector<std::pair<float, int>> t(ng);
        for (int g = 0; g < ng && rc == BORN_OK; g++) {
          t[g].second = g;
          rc = time_gap(w, g, &t[g].first);
        }
        if (rc != BORN_OK) break;
        std::sort(t.begin(), t.end());
        classes.push_back(-1);
        classes.push_back(t[0].second);
        for (int k = 1; k < ng; k++)
          if (t[k].first > 1.02f * t[k - 1].first) classes.push_back(t[k].second);
        cache[w] = tw <= t[0].first ? -1 : t[0].second;
      } else {
        float best = 1e30f;
        for (int g : classes) {
          float ms = 0.f;
          rc = time_gap(w, g, &ms);
          if (rc != BORN_OK) break;
          if (ms < best) {
            best = ms;
            cache[w] = g;
          }
        }
      }
    }
    use[w] = cache[w];
  }
  if (e0) {
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
  }
  if (rc != BORN_OK) {
    cache.clear();
    use.assign(nwin, -1);
  }
  return rc;
}

// ctx->d_ring (64 words): rings s, r at 0, 8 (scale[4], maxb[4]) and max
// |model| at 16 (the forward's ring block); ring u at 24, max |trace| at 32;
// the autotuner's scratch ring block at 40 (rings at 40, 48, max at 56) and
// scratch history scale at 57.
enum { RING_S = 0, RING_R = 8, MAX_M = 16, RING_U = 24, MAX_TR = 32, RING_T0 = 40,
       RING_T1 = 48, MAX_T = 56, HS_T = 57 };
static_assert(RING_R == RING_S + 8 && MAX_M == RING_S + 16 && RING_T1 == RING_T0 + 8 &&
                  MAX_T == RING_T0 + 16 && HS_T < 64,
              "ring blocks");
born::Ring ring(born_ctx *c, int at) {
  return {c->d_ring + at, (unsigned *)(c->d_ring + at + 4)};
}
const unsigned *maxp(born_ctx *c, int at) { return (const unsigned *)(c->d_ring + at); }

int ring_buffers(born_ctx *ctx) {
  if (!ctx->d_ring) {
    CHK(dev_alloc(ctx, &ctx->d_ring, 64));
    HIPCHK(born::launch_fill(ctx->d_ring, 64, 0.f, ctx->stream));
  }
  return BORN_OK;
}

int ring_reset(born_ctx *ctx, int at) {
  born::Ring r = ring(ctx, at);
  HIPCHK(born::launch_ring_init(r.scale, r.maxb, 4, ctx->stream));
  return BORN_OK;
}

// slots of forward step k (reads levels k, k-1; writes k+1)
born::Slots slots(int k) { return {k % 4, (k + 3) % 4, (k + 1) % 4, (k + 2) % 4}; }

float tap_sum(const born_ctx *c) {
  float t = fabsf(c->coeffs[0]);
  for (int i = 1; i < 13; i++) t += 2.f * fabsf(c->coeffs[i]);
  return t;
}

born::Half half_state(const born_ctx *c, int k) {
  born::Half h;
  h.on = c->storage;  // BORN_STORAGE_F32 / F16 / F24 = 0 / 1 / 2
  h.sl = slots(k);
  h.tapsum = tap_sum(c);
  h.gain = 2.f + c->vmax2 * h.tapsum;
  return h;
}

int field_buf(born_ctx *ctx, float **p) {
  if (!*p) CHK(dev_alloc(ctx, p, (size_t)ctx->n));
  return BORN_OK;
}

// Receiver lookup of a domain: every plane holding receivers gets a slot;
// rmap[slot][y][x] is the first receiver at that cell (or -1) and rnext[]
// chains the receivers sharing a cell.
struct RecTables {
  std::vector<int> slot_of_z, rmap, rnext;
  int nslots = 0, per_cell = 0;
};

bool interior_of(int nx, int ny, int nz, long long loc, int *x, int *y, int *z) {
  const long long n12 = (long long)nx * ny;
  if (loc < 0 || loc >= n12 * nz) return false;
  *z = (int)(loc / n12);
  *y = (int)((loc % n12) / nx);
  *x = (int)(loc % nx);
  return *x >= 4 && *x < nx - 4 && *y >= 4 && *y < ny - 4 && *z >= 4 && *z < nz - 4;
}

int receiver_tables(born_ctx *ctx, int nx, int ny, int nz, int nrec, const int *locs,
                    RecTables *rt) {
  std::vector<int> px(nrec), py(nrec), pz(nrec);
  rt->slot_of_z.assign(nz, -1);
  rt->nslots = 0;
  for (int i = 0; i < nrec; i++) {
    if (!interior_of(nx, ny, nz, locs[i], &px[i], &py[i], &pz[i]))
      return fail(ctx, BORN_ERR_UNSUPPORTED,
                  "receiver " + std::to_string(i) + " outside the stencil interior");
    if (rt->slot_of_z[pz[i]] < 0) rt->slot_of_z[pz[i]] = rt->nslots++;
  }
  const long long n12 = (long long)nx * ny;
  rt->rmap.assign((size_t)rt->nslots * n12, -1);
  rt->rnext.assign(nrec, -1);
  std::vector<int> last((size_t)rt->nslots * n12, -1), count((size_t)rt->nslots * n12, 0);
  rt->per_cell = 0;
  for (int i = 0; i < nrec; i++) {
    const long long c = rt->slot_of_z[pz[i]] * n12 + (long long)py[i] * nx + px[i];
    if (rt->rmap[c] < 0)
      rt->rmap[c] = i;
    else
      rt->rnext[last[c]] = i;
    last[c] = i;
    rt->per_cell = std::max(rt->per_cell, ++count[c]);
  }
  return BORN_OK;
}

// cpu_prop.cc:296-326: the 13 stencil coefficients, double expressions
// rounded to float
void stencil_coeffs(float d1, float d2, float d3, float *c) {
  auto C00 = [](double d) { return 8.0 / (5.0 * d * d); };
  c[0] = (float)(-1025.0 / 576.0 * (C00(d1) + C00(d2) + C00(d3)));
  const double dd[3] = {d1, d2, d3};
  const double f[4] = {1.0, -8.0, 63.0, -896.0};
  for (int k = 0; k < 4; k++)
    for (int a = 0; a < 3; a++) c[1 + 3 * k + a] = (float)(k == 0 ? C00(dd[a]) : C00(dd[a]) / f[k]);
}

born::Taps taps_of(const float *coeffs) {
  born::Taps t;
  // reference tap indices: C0=0, CZk=1+3(k-1), CXk=2+3(k-1), CYk=3+3(k-1)
  // (base/cpu_prop.cc:6-18)
  t.c0 = coeffs[0];
  for (int k = 0; k < 4; k++) {
    t.cz[k] = coeffs[1 + 3 * k];
    t.cx[k] = coeffs[2 + 3 * k];
    t.cy[k] = coeffs[3 + 3 * k];
  }
  return t;
}

int check_jt(born_ctx *ctx, int jt) {
  // the forward indexes both sinc tables with the row it - (it/jt)*jt
  // (cpu_prop.cc:31-41)
  if (jt < 1 || jt > ctx->jts || jt > ctx->jtd)
    return fail(ctx, BORN_ERR_UNSUPPORTED,
                "jt=" + std::to_string(jt) + " exceeds the sinc tables (jts=" +
                    std::to_string(ctx->jts) + ", jtd=" + std::to_string(ctx->jtd) +
                    ")");
  return BORN_OK;
}

}  // namespace

// ---- Scratch of the allocation-free entry point born_dev_source_prop_noalloc:
// the per-step source amplitudes, at most kScratchFloats of them, in a
// file-scope device array that each GPU holds for the life of the process.
// Host code never uses the array's name as an address: its device address
// on each GPU is resolved once (hipGetSymbolAddress, with that GPU current)
// and kept per device.  Every call on one GPU shares that GPU's array.
constexpr size_t kScratchFloats = size_t(1) << 20;  // 4 MiB per GPU
static __device__ __align__(256) float scratch[kScratchFloats];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
float *g_scratch[kMaxDevices] = {};

// the scratch address of ctx's device (the device is current), once per device
void scratch_setup(born_ctx *ctx) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (ctx->device >= kMaxDevices) {
    ctx->scratch_err = "device index beyond " + std::to_string(kMaxDevices);
    return;
  }
  if (!g_scratch[ctx->device]) {
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      (void)hipGetLastError();
      ctx->scratch_err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
      return;
    }
    g_scratch[ctx->device] = (float *)p;
  }
  ctx->scratch = g_scratch[ctx->device];
}

}  // namespace

// ======================================================================
extern "C" {

int born_create(born_ctx **out, int device) {
  if (!out) return BORN_ERR_ARG;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) {
    (void)hipGetLastError();
    return BORN_ERR_HIP;
  }
  if (device < 0 || device >= ndev) return BORN_ERR_ARG;
  born_ctx *ctx = new born_ctx();
  ctx->device = device;
  reset_tuning(ctx);
  if (hipSetDevice(device) != hipSuccess ||
      hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
    (void)hipGetLastError();
    delete ctx;
    return BORN_ERR_HIP;
  }
  scratch_setup(ctx);
  *out = ctx;
  return BORN_OK;
}

int born_destroy(born_ctx *ctx) {
  if (!ctx) return BORN_ERR_ARG;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  void *bufs[] = {ctx->d_pts,   ctx->d_bx,    ctx->d_by,    ctx->d_bz,   ctx->d_tab_s,
                  ctx->d_tab_d, ctx->d_src,   ctx->d_vel1, ctx->d_vel2,
                  ctx->d_plane_slot, ctx->d_rmap, ctx->d_rnext, ctx->d_w[0],
                  ctx->d_w[1],  ctx->d_w[2],  ctx->d_w[3], ctx->d_w2[0],
                  ctx->d_w2[1], ctx->d_w2[2], ctx->d_w2[3], ctx->d_img,
                  ctx->d_acc,   ctx->d_p0,    ctx->d_p1,   ctx->d_tr,
                  ctx->d_tr2,   ctx->d_amp,   ctx->d_seg,  ctx->d_ckpt,
                  ctx->d_ring,  ctx->d_hscale, ctx->d_ckring, ctx->d_vm};
  for (void *b : bufs)
    if (b) (void)hipFree(b);
  (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return BORN_OK;
}

const char *born_last_error(const born_ctx *ctx) {
  return ctx ? ctx->err.c_str() : "null context";
}

int born_create_space(born_ctx *ctx, float d1, float d2, float d3, float bc_a,
                      float bc_b, float bc_y, int nx, int ny, int nz) {
  if (!ctx) return BORN_ERR_ARG;
  (void)bc_a;
  (void)bc_b;
  (void)bc_y;
  if (nx < 9 || ny < 9 || nz < 9)
    return fail(ctx, BORN_ERR_ARG, "each axis needs at least 9 samples");
  if (!(d1 > 0) || !(d2 > 0) || !(d3 > 0))
    return fail(ctx, BORN_ERR_ARG, "sampling must be positive");
  if ((long long)nx * ny * nz > (1ll << 30))
    return fail(ctx, BORN_ERR_UNSUPPORTED,
                "more than 2^30 cells (fields are addressed with 32-bit byte offsets)");
  CHK(set_device(ctx));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  const bool resized = nx != ctx->nx || ny != ctx->ny || nz != ctx->nz;
  ctx->nx = nx;
  ctx->ny = ny;
  ctx->nz = nz;
  ctx->n = (long long)nx * ny * nz;
  stencil_coeffs(d1, d2, d3, ctx->coeffs);
  ctx->nbound = 40;
  ctx->bcB = .0005f;
  if (resized) {
    // every buffer of the old cell count goes, the two-step levels included
    // (ensure_work / ensure_pair_work only allocate a missing buffer)
    for (int i = 0; i < 4; i++) CHK(dev_alloc(ctx, &ctx->d_w[i], 0));
    ctx->w_cap = 0;
    for (int i = 0; i < 4; i++) CHK(dev_alloc(ctx, &ctx->d_w2[i], 0));
    CHK(dev_alloc(ctx, &ctx->d_img, 0));
    CHK(dev_alloc(ctx, &ctx->d_acc, 0));
    CHK(dev_alloc(ctx, &ctx->d_vm, 0));
    ctx->vm_cap = 0;
    CHK(dev_alloc(ctx, &ctx->d_p0, 0));
    CHK(dev_alloc(ctx, &ctx->d_p1, 0));
    // The history and checkpoint buffers stay: freeing a buffer of hundreds
    // of GB and allocating it again took 5-7 s (the driver clears released
    // memory), a whole shot's work.  history_alloc clears the level shells
    // for the new layout and grows the buffers when the new grid needs more.
    CHK(dev_alloc(ctx, &ctx->d_vel1, 0));
    CHK(dev_alloc(ctx, &ctx->d_vel2, 0));
    CHK(dev_alloc(ctx, &ctx->d_rmap, 0));
    CHK(dev_alloc(ctx, &ctx->d_src, 0));
  }
  CHK(upload_sponge(ctx));
  if (resized) reset_tuning(ctx);
  ctx->have_space = true;
  ctx->bg_version++;
  return BORN_OK;
}

int born_set_sponge(born_ctx *ctx, int width, float bcB) {
  CHK(need_space(ctx));
  if (width < 0 || width > 4096) return fail(ctx, BORN_ERR_ARG, "bad sponge width");
  CHK(set_device(ctx));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  ctx->nbound = width;
  ctx->bcB = bcB;
  CHK(upload_sponge(ctx));
  ctx->bg_version++;
  return BORN_OK;
}

int born_set_ntblock(born_ctx *ctx, int nb) {
  if (!ctx) return BORN_ERR_ARG;
  (void)nb;
  return BORN_OK;
}

static int sinc_table(born_ctx *ctx, int nsinc, int jt, const float *table,
                      float **dst, int *ns_out, int *jt_out) {
  if (!ctx) return BORN_ERR_ARG;
  if (nsinc != 8) return fail(ctx, BORN_ERR_UNSUPPORTED, "nsinc must be 8");
  if (jt < 1 || !table) return fail(ctx, BORN_ERR_ARG, "bad sinc table");
  CHK(set_device(ctx));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  CHK(dev_alloc(ctx, dst, (size_t)nsinc * jt));
  HIPCHK(hipMemcpy(*dst, table, (size_t)nsinc * jt * sizeof(float),
                   hipMemcpyHostToDevice));
  *ns_out = nsinc;
  *jt_out = jt;
  ctx->bg_version++;
  ctx->tab_version++;
  return BORN_OK;
}

int born_transfer_sinc_table_s(born_ctx *ctx, int nsinc, int jts, const float *table) {
  return sinc_table(ctx, nsinc, jts, table, ctx ? &ctx->d_tab_s : nullptr,
                    ctx ? &ctx->nsinc_s : nullptr, ctx ? &ctx->jts : nullptr);
}

int born_transfer_sinc_table_d(born_ctx *ctx, int nsinc, int jtd, const float *table) {
  CHK(sinc_table(ctx, nsinc, jtd, table, ctx ? &ctx->d_tab_d : nullptr,
                 ctx ? &ctx->nsinc_d : nullptr, ctx ? &ctx->jtd : nullptr));
  float m = 0.f;
  for (int r = 0; r < jtd; r++) {
    float a = 0.f;
    for (int k = 0; k < nsinc; k++) a += fabsf(table[(size_t)r * nsinc + k]);
    m = std::max(m, a);
  }
  ctx->tabd_abs = m;
  return BORN_OK;
}

int born_transfer_source_func(born_ctx *ctx, int npts, int nt_big, const int *locs,
                              const float *vals) {
  CHK(need_space(ctx));
  if (npts < 0 || nt_big < 1 || (npts && (!locs || !vals)))
    return fail(ctx, BORN_ERR_ARG, "bad source function");
  if (npts > born::kMaxPts)
    return fail(ctx, BORN_ERR_UNSUPPORTED,
                "at most " + std::to_string(born::kMaxPts) + " source points");
  std::vector<int> pts(4 * (size_t)std::max(npts, 1), 0);
  for (int i = 0; i < npts; i++)
    if (!interior(ctx, locs[i], &pts[4 * i], &pts[4 * i + 1], &pts[4 * i + 2]))
      return fail(ctx, BORN_ERR_UNSUPPORTED,
                  "source point " + std::to_string(i) + " outside the stencil interior");
  CHK(set_device(ctx));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  CHK(dev_alloc(ctx, &ctx->d_src, (size_t)std::max(npts, 1) * nt_big));
  CHK(dev_alloc(ctx, &ctx->d_pts, pts.size()));
  HIPCHK(hipMemcpy(ctx->d_pts, pts.data(), pts.size() * sizeof(int), hipMemcpyHostToDevice));
  if (npts)
    HIPCHK(hipMemcpy(ctx->d_src, vals, (size_t)npts * nt_big * sizeof(float),
                     hipMemcpyHostToDevice));
  ctx->npts_s = npts;
  ctx->nt_src = nt_big;
  ctx->bg_version++;
  return BORN_OK;
}

static int vel_func(born_ctx *ctx, int nx, int ny, int nz, const float *v, float **dst) {
  CHK(need_space(ctx));
  CHK(check_dims(ctx, nx, ny, nz));
  if (!v) return fail(ctx, BORN_ERR_ARG, "null velocity");
  CHK(set_device(ctx));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  CHK(field_buf(ctx, dst));
  HIPCHK(hipMemcpy(*dst, v, (size_t)ctx->n * sizeof(float), hipMemcpyHostToDevice));
  ctx->bg_version++;
  return BORN_OK;
}

int born_transfer_vel_func1(born_ctx *ctx, int nx, int ny, int nz, const float *v) {
  return vel_func(ctx, nx, ny, nz, v, ctx ? &ctx->d_vel1 : nullptr);
}

int born_transfer_vel_func2(born_ctx *ctx, int nx, int ny, int nz, const float *v) {
  if (ctx && v && nx == ctx->nx && ny == ctx->ny && nz == ctx->nz) {
    // max |v| over the domain and min v over the stencil interior, per
    // thread block: neither depends on the order
    std::vector<float> part(born_par::threads(), 0.f), pmin(part.size(), INFINITY);
    const long nb = (long)part.size();
    const long long n = ctx->n;
    const int nx_ = ctx->nx, ny_ = ctx->ny, nz_ = ctx->nz;
    float *pp = part.data(), *pm = pmin.data();
    born_par::blocks(nb, 1, [=](long lo, long hi) {
      for (long b = lo; b < hi; b++) {
        float m = 0.f, mn = INFINITY;
        for (long long i = n * b / nb; i < n * (b + 1) / nb; i++) {
          m = std::max(m, fabsf(v[i]));
          const int x = (int)(i % nx_), y = (int)((i / nx_) % ny_), z = (int)(i / ((long long)nx_ * ny_));
          if (x >= 4 && x < nx_ - 4 && y >= 4 && y < ny_ - 4 && z >= 4 && z < nz_ - 4)
            mn = std::min(mn, v[i]);
        }
        pp[b] = m;
        pm[b] = mn;
      }
    });
    ctx->vmax2 = *std::max_element(part.begin(), part.end());
    ctx->vmin2 = *std::min_element(pmin.begin(), pmin.end());
  }
  return vel_func(ctx, nx, ny, nz, v, ctx ? &ctx->d_vel2 : nullptr);
}

int born_transfer_receiver_func(born_ctx *ctx, int rec_nx, int rec_ny, int nt_big,
                                const int *locs, float *traces) {
  CHK(need_space(ctx));
  if (rec_nx < 1 || rec_ny < 1 || nt_big < 1 || !locs)
    return fail(ctx, BORN_ERR_ARG, "bad receiver geometry");
  const int nrec = rec_nx * rec_ny;
  RecTables rt;
  CHK(receiver_tables(ctx, ctx->nx, ctx->ny, ctx->nz, nrec, locs, &rt));
  CHK(set_device(ctx));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  CHK(dev_alloc(ctx, &ctx->d_plane_slot, (size_t)ctx->nz));
  CHK(dev_alloc(ctx, &ctx->d_rmap, rt.rmap.size()));
  CHK(dev_alloc(ctx, &ctx->d_rnext, rt.rnext.size()));
  HIPCHK(hipMemcpy(ctx->d_plane_slot, rt.slot_of_z.data(), ctx->nz * sizeof(int),
                   hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(ctx->d_rmap, rt.rmap.data(), rt.rmap.size() * sizeof(int),
                   hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(ctx->d_rnext, rt.rnext.data(), rt.rnext.size() * sizeof(int),
                   hipMemcpyHostToDevice));
  ctx->rec_nx = rec_nx;
  ctx->rec_ny = rec_ny;
  ctx->nrec = nrec;
  ctx->nt_rec = nt_big;
  ctx->nslots = rt.nslots;
  ctx->rec_per_cell = rt.per_cell;
  ctx->h_rec = traces;
  return BORN_OK;
}

// ------------------------------------------------------------- device loops

int born_dev_forward(born_ctx *ctx, int jt, int nt, int scatter, const float *model,
                     float *traces, int keep_background) {
  CHK(need_born_inputs(ctx, false, true, true));
  CHK(check_jt(ctx, jt));
  if (nt < 0 || !model || !traces || (scatter != 0 && scatter != 1))
    return fail(ctx, BORN_ERR_ARG, "bad forward arguments");
  CHK(set_device(ctx));
  CHK(check_storage(ctx));
  CHK(ensure_work(ctx));
  CHK(ring_buffers(ctx));
  const long long n = ctx->n;
  const bool half = ctx->storage != BORN_STORAGE_F32;  // scaled 16-bit levels
  const size_t esz = elem(ctx);
  const bool try_pairs = ctx->pairs && !half && nt >= 2;
  if (try_pairs) CHK(ensure_pair_work(ctx));
  float *hist = nullptr;
  // With FP32 levels a history that does not fit becomes the adjoint's
  // checkpoint plan: checkpoints at segment starts and the last segment's
  // scatter operands, so that the adjoint need not sweep the background
  // first.  (Scaled levels keep s and r together; their adjoint sweeps.)
  int ckS = 0, ckL = 0;
  if (keep_background && nt > 0) {
    if (seg_layout(ctx, (size_t)nt, esz).bytes <= history_budget(ctx)) {
      CHK(history_alloc(ctx, (size_t)nt, 0, esz));
      hist = ctx->d_seg;
      if (half) CHK(ensure(ctx, &ctx->d_hscale, &ctx->hscale_cap, (size_t)nt));
    } else if (!half && !try_pairs) {
      CHK(checkpoint_plan(ctx, nt, esz, &ckS, &ckL));
      CHK(history_alloc(ctx, (size_t)ckL, 2 * (size_t)(ckS - 1), esz));
    }
  }
  ctx->hist_version = -1;
  ctx->ck_version = -1;
  ctx->jt_state = jt;
  const Geometry geo = geometry(ctx);
  const int kind = fwd_kind(ctx->storage, scatter);
  // velocity 2 and the model as one stream (8 bytes per cell; the buffer
  // comes with the work buffers)
  HIPCHK(born::launch_interleave(ctx->d_vm, ctx->d_vel2, model, n, ctx->stream));
  {  // tune on the work fields (zeroed below) and scratch scales, no traces written
    CHK(ring_reset(ctx, RING_T0));
    CHK(ring_reset(ctx, RING_T1));
    HIPCHK(born::launch_fill(ctx->d_ring + MAX_T, 2, 0.f, ctx->stream));
    born::FwdStep t{};
    t.s1 = ctx->d_w[1]; t.s0 = ctx->d_w[0]; t.r1 = ctx->d_w[3]; t.r0 = ctx->d_w[2];
    t.vel = ctx->d_vel2; t.model = model; t.vm = ctx->d_vm;
    t.src = no_points(); t.rec = recmap(ctx);
    t.tr = traces; t.tab = ctx->d_tab_d; t.extract = 0; t.scatter = scatter;
    t.h = half_state(ctx, 0);
    t.rg = ctx->d_ring + RING_T0; t.hscale = ctx->d_ring + HS_T;
    // The buffers the steps below use.  With the s levels in the kept
    // history each timed launch takes the next three history levels, as
    // the steps do (levels that the steps overwrite); the same three levels
    // launch after launch time up to 20 % faster than the steps run, and
    // pick another tile.  Else the history level written beside the
    // in-place s.
    const bool rotate = hist && scatter == 0 && !half && nt >= 3;
    if (!rotate) t.hist = hist;
    int lev = 0;
    CHK(autotune(ctx, kind, [&](const born::Tune &tu) {
      if (rotate) {
        t.s0 = seg_level(ctx, lev);
        t.s1 = seg_level(ctx, lev + 1);
        t.s_new = seg_level(ctx, lev + 2);
        lev = (lev + 1) % (nt - 2);
      }
      return born::launch_fwd_step(geo, tu, t, ctx->stream); }));
  }
  const born::Tune tu = tuning(ctx, kind);
  const int pkind = scatter == 0 ? K_FWD0P : K_FWD1P;
  bool use_pairs = false;
  if (try_pairs) {
    born::PairStep t{};
    t.s_old = ctx->d_w[0]; t.s_cur = ctx->d_w[1]; t.r_old = ctx->d_w[2]; t.r_cur = ctx->d_w[3];
    t.s_a = ctx->d_w2[0]; t.s_b = ctx->d_w2[1]; t.r_a = ctx->d_w2[2]; t.r_b = ctx->d_w2[3];
    t.vel = ctx->d_vel2; t.model = model; t.src = no_points(); t.rec = recmap(ctx);
    t.tr = traces; t.tab_a = t.tab_b = ctx->d_tab_d; t.scatter = scatter;
    CHK(autotune_pair(ctx, pkind, [&](const born::Tune &tp) {
      return born::launch_pair_step(geo, tp, t, ctx->stream); }));
    // pairs unless both kernels were timed and two single steps won
    const float mp = ctx->tune_ms[pkind], ms = ctx->tune_ms[kind];
    use_pairs = ctx->pairs == 2 || !(mp > 0.f && ms > 0.f && mp > 2.f * ms);
    for (int i = 0; i < 4; i++) CHK(zero(ctx, ctx->d_w2[i], n));
  }
  ctx->last_pairs = use_pairs;
  // Where the work pair sits (r, or with narrow storage s and r together),
  // by the window of the history level a step writes: level it of a kept
  // history, level it - klast in the last segment of a checkpoint plan;
  // before klast (s in the work buffers) a window of its own, index ngaps.
  const int klast = ckS > 1 ? (ckS - 1) * ckL : nt;
  const bool place = ctx->place_mode && ctx->seg_gap && !use_pairs && (hist || ckS > 1);
  const int wwork = hist ? 0 : 1;  // the window of steps that write no history level
  auto window_of = [&](int it) {
    const long long lv = hist ? it : it >= klast ? it - klast : -1;
    return lv < 0 ? ctx->seg_ngaps : (int)(lv / ctx->seg_lpw);
  };
  std::vector<int> rplaces;
  if (place) {
    const long long nlev = hist ? nt : ckL;
    born::FwdStep t{};  // a step without sources or traces
    t.vel = ctx->d_vel2; t.model = model; t.vm = ctx->d_vm; t.src = no_points();
    t.rec = recmap(ctx);
    t.tr = traces; t.tab = ctx->d_tab_d; t.extract = 0; t.scatter = scatter;
    t.h = half_state(ctx, 0); t.rg = ctx->d_ring + RING_T0; t.hscale = ctx->d_ring + HS_T;
    const std::vector<uintptr_t> key = {
        (uintptr_t)ctx->d_seg, ctx->seg_win, (uintptr_t)scatter, (uintptr_t)n,
        (uintptr_t)ctx->d_vel2, (uintptr_t)model, (uintptr_t)tu.shape, (uintptr_t)tu.zchunk,
        (uintptr_t)wwork, (uintptr_t)ctx->d_w[0]};
    CHK(place_plan(ctx, ctx->place_fwd, ctx->place_fwd_key, key, ctx->seg_ngaps + wwork,
                   rplaces, [&](int w, int g, int i) {
      // the pair's two levels at place g
      const size_t pb = (half ? 2 : 1) * (size_t)n * esz;  // a level of the pair
      float *p0 = g < 0 ? ctx->d_w[half ? 0 : 2] : seg_gap_field(ctx, g, 0, pb);
      float *p1 = g < 0 ? ctx->d_w[half ? 1 : 3] : seg_gap_field(ctx, g, 1, pb);
      if (i % 2) std::swap(p0, p1);
      if (w == ctx->seg_ngaps) {  // s in the work buffers, no history level
        t.s1 = ctx->d_w[1]; t.s0 = ctx->d_w[0]; t.s_new = nullptr; t.hist = nullptr;
        t.r0 = p0;
        t.r1 = p1;
        return born::launch_fwd_step(geo, tu, t, ctx->stream);
      }
      // three levels of the window, one on per launch, as in a march
      const long long a = (long long)w * ctx->seg_lpw;
      const long long b = std::min(nlev, a + ctx->seg_lpw);
      const long long k = a + i % std::max(1LL, b - a - 2);
      if (half) {  // s and r together in the pair, the scatter operand apart
        t.s0 = p0;
        t.s1 = p1;
        t.s_new = nullptr;
        t.hist = seg_level(ctx, k);
        return born::launch_fwd_step(geo, tu, t, ctx->stream);
      }
      if (scatter == 0) {
        t.s0 = seg_level(ctx, k);
        t.s1 = seg_level(ctx, std::min(k + 1, b - 1));
        t.s_new = seg_level(ctx, std::min(k + 2, b - 1));
        t.hist = nullptr;
      } else {
        t.s1 = ctx->d_w[1]; t.s0 = ctx->d_w[0]; t.s_new = nullptr;
        t.hist = seg_level(ctx, k);
      }
      t.r0 = p0;
      t.r1 = p1;
      return born::launch_fwd_step(geo, tu, t, ctx->stream); }));
  }
  for (int i = 0; i < 4; i++) CHK(zero(ctx, ctx->d_w[i], (long long)ctx->w_cap));
  CHK(source_amps(ctx, nt, jt, 0, 1 / (float)jt));
  if (half) {
    CHK(ring_reset(ctx, RING_S));
    CHK(ring_reset(ctx, RING_R));
    HIPCHK(born::launch_fill(ctx->d_ring + MAX_M, 1, 0.f, ctx->stream));
    HIPCHK(born::launch_absmax(model, n, (unsigned *)(ctx->d_ring + MAX_M), ctx->stream));
  }
  int it = 0;
  // levels k-1, k, k+1, k+2 of s and r rotate through four buffers each
  float *sb[4] = {ctx->d_w[0], ctx->d_w[1], ctx->d_w2[0], ctx->d_w2[1]};
  float *rb[4] = {ctx->d_w[2], ctx->d_w[3], ctx->d_w2[2], ctx->d_w2[3]};
  if (use_pairs) {
    const born::Tune tp = tuning(ctx, pkind);
    born::PairStep p{};
    p.vel = ctx->d_vel2;
    p.model = model;
    p.src = points(ctx);
    p.rec = recmap(ctx);
    p.tr = traces;
    p.scatter = scatter;
    for (; it + 1 < nt; it += 2) {
      const int ida = it / jt, idb = (it + 1) / jt;
      p.s_old = sb[0]; p.s_cur = sb[1]; p.s_a = sb[2]; p.s_b = sb[3];
      p.r_old = rb[0]; p.r_cur = rb[1]; p.r_a = rb[2]; p.r_b = rb[3];
      p.h_a = hist ? seg_level(ctx, it) : nullptr;
      p.h_b = hist ? seg_level(ctx, it + 1) : nullptr;
      p.amp_a = ctx->d_amp + (long long)it * ctx->npts_s;
      p.amp_b = ctx->d_amp + (long long)(it + 1) * ctx->npts_s;
      p.tab_a = ctx->d_tab_d + (long long)(it - ida * jt) * 8;
      p.tab_b = ctx->d_tab_d + (long long)(it + 1 - idb * jt) * 8;
      p.id_a = ida;
      p.id_b = idb;
      p.ext_a = ida + 7 < ctx->nt_rec;
      p.ext_b = idb + 7 < ctx->nt_rec;
      HIPCHK(born::launch_pair_step(geo, tp, p, ctx->stream));
      std::swap(sb[0], sb[2]);
      std::swap(sb[1], sb[3]);
      std::swap(rb[0], rb[2]);
      std::swap(rb[1], rb[3]);
    }
  }
  float *s_old = sb[0], *s_cur = sb[1];
  float *r_old = rb[0], *r_cur = rb[1];
  born::FwdStep a{};
  a.vel = ctx->d_vel2;
  a.model = model;
  a.vm = ctx->d_vm;
  a.src = points(ctx);
  a.rec = recmap(ctx);
  a.tr = traces;
  a.scatter = scatter;
  a.rg = ctx->d_ring + RING_S;
  // With scatter 0 the scatter operand of step it is s of level it+1, so a
  // kept history holds the s levels: step it reads levels it (history it-1)
  // and it-1 (history it-2) and writes level it+1 into history it, one
  // write per cell less than a separate s buffer and history.  (FP16 levels
  // keep s and r together as one __half2 per cell in sb[], and the history
  // is written apart.)
  const bool s_in_hist = hist && scatter == 0 && !half;
  // checkpoint plan: levels it and it-1 saved before step it at segment
  // starts it = c L (c = 1..S-1); from klast on the scatter operands go to
  // the segment history (scatter 0: the s levels themselves, level j > klast
  // in segment level j-1-klast, levels klast and klast-1 staying in s_cur
  // and s_old)
  const size_t Fb = (size_t)n * esz;
  int rplace = -1, moves = 0;  // where r is: the work buffers
  for (; it < nt; it++) {
    const int id = it / jt, ii = it - id * jt;
    if (place) {
      const int want = rplaces[window_of(it)];
      if (want != rplace) {
        if (half)
          CHK(move_pair(ctx, want, ctx->d_w[0], ctx->d_w[1], s_old, s_cur, 2 * Fb));
        else
          CHK(move_pair(ctx, want, ctx->d_w[2], ctx->d_w[3], r_old, r_cur, Fb));
        rplace = want;
        moves++;
      }
    }
    if (ckS > 1 && it > 0 && it % ckL == 0 && it / ckL <= ckS - 1) {
      const long long c = it / ckL - 1;
      HIPCHK(hipMemcpyAsync(nth(ctx->d_ckpt, 2 * c, n, esz), s_cur, Fb,
                            hipMemcpyDeviceToDevice, ctx->stream));
      HIPCHK(hipMemcpyAsync(nth(ctx->d_ckpt, 2 * c + 1, n, esz), s_old, Fb,
                            hipMemcpyDeviceToDevice, ctx->stream));
    }
    const bool into_seg = it >= klast;
    if (s_in_hist) {
      a.s1 = it >= 1 ? seg_level(ctx, it - 1) : s_cur;  // level 0: zero
      a.s0 = it >= 2 ? seg_level(ctx, it - 2) : s_old;
      a.s_new = seg_level(ctx, it);
      a.hist = nullptr;
    } else if (into_seg && scatter == 0) {
      a.s1 = it > klast ? seg_level(ctx, it - 1 - klast) : s_cur;
      a.s0 = it > klast + 1 ? seg_level(ctx, it - 2 - klast) : (it > klast ? s_cur : s_old);
      a.s_new = seg_level(ctx, it - klast);
      a.hist = nullptr;
    } else if (into_seg) {
      a.s1 = s_cur;
      a.s0 = s_old;
      a.s_new = nullptr;
      a.hist = seg_level(ctx, it - klast);
    } else {
      a.s1 = s_cur;
      a.s0 = s_old;
      a.s_new = nullptr;
      a.hist = hist ? seg_level(ctx, it) : nullptr;
    }
    a.r1 = r_cur;
    a.r0 = r_old;
    a.hscale = half && hist ? ctx->d_hscale + it : nullptr;
    a.amp = ctx->d_amp + (long long)it * ctx->npts_s;
    a.tab = ctx->d_tab_d + (long long)ii * 8;
    a.id = id;
    a.extract = id + 7 < ctx->nt_rec;
    a.h = half_state(ctx, it);
    HIPCHK(born::launch_fwd_step(geo, tu, a, ctx->stream));
    if (!(into_seg && scatter == 0)) std::swap(s_old, s_cur);
    std::swap(r_old, r_cur);
  }
  ctx->last_fwd_moves = moves;
  if (hist) {
    ctx->hist_version = ctx->bg_version;
    ctx->hist_scatter = scatter;
    ctx->hist_jt = jt;
    ctx->hist_nt = nt;
    ctx->hist_storage = ctx->storage;
  }
  if (ckS > 1) {
    ctx->ck_version = ctx->bg_version;
    ctx->ck_S = ckS;
    ctx->ck_L = ckL;
    ctx->ck_scatter = scatter;
    ctx->ck_jt = jt;
    ctx->ck_nt = nt;
  }
  return BORN_OK;
}

int born_dev_source_prop(born_ctx *ctx, int jts, int nt, float *p0, float *p1) {
  CHK(need_born_inputs(ctx, true, false, false));
  if (nt < 0 || jts < 1 || !p0 || !p1) return fail(ctx, BORN_ERR_ARG, "bad source_prop arguments");
  CHK(set_device(ctx));
  // cpu_prop.cc:108-122: _jt = jts scales, the table's jts indexes
  ctx->jt_state = jts;
  const Geometry geo = geometry(ctx);
  CHK(ensure_work(ctx));
  {  // tune on two work fields (unused by this loop)
    born::SrcStep t{};
    t.s1 = ctx->d_w[1]; t.s0 = ctx->d_w[0]; t.vel = ctx->d_vel1; t.src = no_points();
    CHK(autotune(ctx, K_SRC, [&](const born::Tune &tu) {
      return born::launch_src_step(geo, tu, t, ctx->stream); }));
  }
  const born::Tune tu = tuning(ctx, K_SRC);
  CHK(source_amps(ctx, nt + 1, ctx->jts, 0, 1 / (float)jts));
  float *a0 = p0, *a1 = p1;
  born::SrcStep s{};
  s.vel = ctx->d_vel1;
  s.src = points(ctx);
  for (int it = 0; it <= nt; it++) {
    s.s1 = a1;
    s.s0 = a0;
    s.amp = ctx->d_amp + (long long)it * ctx->npts_s;
    HIPCHK(born::launch_src_step(geo, tu, s, ctx->stream));
    std::swap(a0, a1);
  }
  if (nt % 2 == 1)  // cpu_prop.cc:123-128
    HIPCHK(hipMemcpyAsync(p1, p0, (size_t)ctx->n * sizeof(float),
                          hipMemcpyDeviceToDevice, ctx->stream));
  return BORN_OK;
}

// born_dev_source_prop without allocating: no autotuning (the launch shape
// tuned on this context so far, else the default one; every shape gives the
// same values) and the amplitudes in this device's scratch array.
int born_dev_source_prop_noalloc(born_ctx *ctx, int jts, int nt, float *p0, float *p1) {
  CHK(need_born_inputs(ctx, true, false, false));
  if (nt < 0 || jts < 1 || !p0 || !p1) return fail(ctx, BORN_ERR_ARG, "bad source_prop arguments");
  if (!ctx->scratch)
    return fail(ctx, BORN_ERR_HIP, "no scratch array on this device: " + ctx->scratch_err);
  const size_t need = (size_t)(nt + 1) * (size_t)ctx->npts_s;
  if (need > kScratchFloats)
    return fail(ctx, BORN_ERR_ARG,
                "(nt + 1) x source points = " + std::to_string(need) +
                    " amplitudes exceed the scratch array (" + std::to_string(kScratchFloats) +
                    ")");
  CHK(set_device(ctx));
  ctx->jt_state = jts;
  const Geometry geo = geometry(ctx);
  const born::Tune tu = tuning(ctx, K_SRC);
  // every amplitude the loop reads is written here first
  HIPCHK(born::launch_src_amp(ctx->scratch, nt + 1, ctx->npts_s, ctx->d_src, ctx->nt_src,
                              ctx->d_tab_s, ctx->nsinc_s, ctx->jts, ctx->jts, 0,
                              1 / (float)jts, ctx->stream));
  float *a0 = p0, *a1 = p1;
  born::SrcStep s{};
  s.vel = ctx->d_vel1;
  s.src = points(ctx);
  for (int it = 0; it <= nt; it++) {
    s.s1 = a1;
    s.s0 = a0;
    s.amp = ctx->scratch + (long long)it * ctx->npts_s;
    HIPCHK(born::launch_src_step(geo, tu, s, ctx->stream));
    std::swap(a0, a1);
  }
  if (nt % 2 == 1)  // cpu_prop.cc:123-128
    HIPCHK(hipMemcpyAsync(p1, p0, (size_t)ctx->n * sizeof(float), hipMemcpyDeviceToDevice,
                          ctx->stream));
  return BORN_OK;
}

int born_dev_rtm_adjoint(born_ctx *ctx, int nt, float *p0, float *p1, float *img,
                         const float *traces) {
  CHK(need_born_inputs(ctx, true, true, true));
  if (nt < 0 || !p0 || !p1 || !img || !traces)
    return fail(ctx, BORN_ERR_ARG, "bad rtm_adjoint arguments");
  CHK(set_device(ctx));
  CHK(ensure_work(ctx));
  CHK(field_buf(ctx, &ctx->d_acc));
  const long long n = ctx->n;
  const Geometry geo = geometry(ctx);
  {  // tune on work fields and the accumulator (no trace read, image untouched)
    born::RadjStep t{};
    t.s1 = ctx->d_w[3]; t.s0 = ctx->d_w[2]; t.r1 = ctx->d_w[1]; t.r0 = ctx->d_w[0];
    t.vel1 = ctx->d_vel1; t.vel2 = ctx->d_vel2; t.img = ctx->d_acc;
    t.s_prop = 1; t.r_prop = 1; t.src = no_points(); t.amp_pre = ctx->d_vel1; t.pre = 0;
    t.rec = recmap(ctx); t.tr = traces; t.tab = ctx->d_tab_d; t.inject = 0;
    CHK(autotune(ctx, K_RADJ, [&](const born::Tune &tu) {
      return born::launch_radj_step(geo, tu, t, ctx->stream); }));
  }
  const born::Tune tu = tuning(ctx, K_RADJ);
  CHK(zero(ctx, ctx->d_w[0], n));
  CHK(zero(ctx, ctx->d_w[1], n));
  // cpu_prop.cc:71-95 with _dir = -1 and _jt from the last sourceProp
  const float scale = -1 / (float)ctx->jt_state;
  if (nt > 0) CHK(source_amps(ctx, nt, ctx->jts, 1, scale));
  float *s_old = p0, *s_cur = p1;
  float *r_old = ctx->d_w[0], *r_cur = ctx->d_w[1];
  born::RadjStep a{};
  a.vel1 = ctx->d_vel1;
  a.vel2 = ctx->d_vel2;
  a.img = img;
  a.src = points(ctx);
  a.rec = recmap(ctx);
  a.tr = traces;
  a.sc = (float)-1 / (float)ctx->jt_state;
  for (int it = nt - 1; it >= 0; it--) {
    const int id = it / ctx->jtd, ii = it - id * ctx->jtd;
    a.s1 = s_cur;
    a.s0 = s_old;
    a.r1 = r_cur;
    a.r0 = r_old;
    a.s_prop = it < nt - 1;
    a.r_prop = it > 0;
    a.pre = it + 1 < nt - 1;
    a.amp_pre = a.pre ? ctx->d_amp + (long long)(it + 1) * ctx->npts_s : ctx->d_amp;
    a.tab = ctx->d_tab_d + (long long)ii * 8;
    a.id = id;
    a.inject = id + 7 < ctx->nt_rec;
    HIPCHK(born::launch_radj_step(geo, tu, a, ctx->stream));
    std::swap(s_old, s_cur);
    std::swap(r_old, r_cur);
  }
  return BORN_OK;
}

int born_dev_adjoint_exact(born_ctx *ctx, int jt, int nt, int scatter,
                           const float *traces, float *img) {
  CHK(need_born_inputs(ctx, false, true, true));
  CHK(check_jt(ctx, jt));
  if (nt < 0 || !traces || !img || (scatter != 0 && scatter != 1))
    return fail(ctx, BORN_ERR_ARG, "bad adjoint arguments");
  // the adjoint state runs in nu = V B lambda and the image is divided by
  // V B at the end: V B must be positive in the whole interior
  if (!(ctx->vmin2 * ctx->bmin > 0.f) || !std::isfinite(ctx->vmax2))
    return fail(ctx, BORN_ERR_UNSUPPORTED,
                "the exact adjoint needs velocity 2 and the sponge factors > 0 in the "
                "interior (min V = " + std::to_string(ctx->vmin2) + ", min sponge = " +
                    std::to_string(ctx->bmin) + ")");
  CHK(set_device(ctx));
  CHK(check_storage(ctx));
  CHK(ensure_work(ctx));
  CHK(ring_buffers(ctx));
  CHK(field_buf(ctx, &ctx->d_acc));
  const long long n = ctx->n;
  if (nt == 0) return BORN_OK;
  const bool half = ctx->storage != BORN_STORAGE_F32;  // scaled 16-bit levels
  const bool try_pairs = ctx->pairs && !half && nt >= 2;
  if (try_pairs) CHK(ensure_pair_work(ctx));
  const size_t esz = elem(ctx);
  const Geometry geo = geometry(ctx);
  const int skind = srec_kind(ctx->storage, scatter);
  const int lkind = lam_kind(ctx->storage);
  // bound of |V B P^T w| per unit of max |trace| (FP16 scales)
  const float inj = ctx->vmax2 * ctx->tabd_abs * (float)ctx->rec_per_cell;
  {  // tune both kernels on work fields, the accumulator and scratch scales
    CHK(ring_reset(ctx, RING_T0));
    HIPCHK(born::launch_fill(ctx->d_ring + MAX_T, 1, 0.f, ctx->stream));
    HIPCHK(born::launch_fill(ctx->d_ring + HS_T, 1, 1.f, ctx->stream));
    born::SrecStep t{};
    t.s1 = ctx->d_w[3]; t.s0 = ctx->d_w[2]; t.vel = ctx->d_vel2; t.sig = ctx->d_acc;
    t.src = no_points(); t.scatter = scatter;
    t.h = half_state(ctx, 0); t.ring = ring(ctx, RING_T0); t.hscale = ctx->d_ring + HS_T;
    CHK(autotune(ctx, skind, [&](const born::Tune &tu) {
      return born::launch_srec_step(geo, tu, t, ctx->stream); }));
    CHK(ring_reset(ctx, RING_T0));
    HIPCHK(born::launch_fill(ctx->d_ring + HS_T, 1, 1.f, ctx->stream));
    born::LamStep l{};
    l.u1 = ctx->d_w[1]; l.u0 = ctx->d_w[0]; l.vel = ctx->d_vel2; l.sig = ctx->d_w[2];
    l.acc = ctx->d_acc; l.rec = recmap(ctx); l.tr = traces; l.tab = ctx->d_tab_d;
    l.inject = 0;
    l.h = half_state(ctx, 0); l.ring = ring(ctx, RING_T0); l.inj = inj;
    l.trmax = maxp(ctx, MAX_T); l.hscale = ctx->d_ring + HS_T;
    CHK(autotune(ctx, lkind, [&](const born::Tune &tu) {
      return born::launch_lam_step(geo, tu, l, ctx->stream); }));
    if (try_pairs) {
      born::LamPairStep lp{};
      lp.u_old = ctx->d_w[0]; lp.u_cur = ctx->d_w[1]; lp.u_a = ctx->d_w2[0];
      lp.u_b = ctx->d_w2[1]; lp.vel = ctx->d_vel2; lp.sig_a = lp.sig_b = ctx->d_w[2];
      lp.acc = ctx->d_acc; lp.rec = recmap(ctx); lp.tr = traces;
      lp.tab_a = lp.tab_b = ctx->d_tab_d;
      CHK(autotune_pair(ctx, K_LAMP, [&](const born::Tune &tp) {
        return born::launch_lam_pair_step(geo, tp, lp, ctx->stream); }));
    }
  }
  const born::Tune stu = tuning(ctx, skind), ltu = tuning(ctx, lkind);
  bool lam_pairs = false;
  if (try_pairs) {
    // pairs unless both kernels were timed and two single steps won
    const float mp = ctx->tune_ms[K_LAMP], ms = ctx->tune_ms[lkind];
    lam_pairs = ctx->pairs == 2 || !(mp > 0.f && ms > 0.f && mp > 2.f * ms);
    CHK(zero(ctx, ctx->d_w2[0], n));
    CHK(zero(ctx, ctx->d_w2[1], n));
  }
  ctx->last_lam_pairs = lam_pairs;
  const bool cached = ctx->hist_version == ctx->bg_version &&
                      ctx->hist_scatter == scatter && ctx->hist_jt == jt &&
                      ctx->hist_nt == nt && ctx->hist_storage == ctx->storage &&
                      ctx->seg_cap * sizeof(float) >= (size_t)nt * n * esz;
  // the checkpoints and last segment that the forward left (FP32 levels)
  const bool planned = !cached && !half && ctx->ck_version == ctx->bg_version &&
                       ctx->ck_scatter == scatter && ctx->ck_jt == jt && ctx->ck_nt == nt;
  int S = 1, L = nt;
  if (planned) {
    S = ctx->ck_S;
    L = ctx->ck_L;
  }
  // this adjoint overwrites the segment history
  ctx->ck_version = -1;
  if (!cached && !planned) {
    ctx->hist_version = -1;  // the segment buffer is about to be reused
    CHK(checkpoint_plan(ctx, nt, esz, &S, &L));
    CHK(history_alloc(ctx, (size_t)L, 2 * (size_t)(S - 1), esz));
    if (half) {
      CHK(ensure(ctx, &ctx->d_hscale, &ctx->hscale_cap, (size_t)L));
      CHK(ensure(ctx, &ctx->d_ckring, &ctx->ckring_cap, 8 * (size_t)std::max(S - 1, 1)));
    }
  }
  CHK(source_amps(ctx, nt, jt, 0, 1 / (float)jt));
  float *s_old = ctx->d_w[2], *s_cur = ctx->d_w[3];
  // scatter 0: sigma of step k is s of level k+1, so the background steps
  // that fill the segment history write their s levels straight into it
  const bool s_in_seg = scatter == 0;
  born::SrecStep sr{};
  sr.vel = ctx->d_vel2;
  sr.src = points(ctx);
  sr.scatter = scatter;
  sr.ring = ring(ctx, RING_S);
  const size_t Fb = (size_t)n * esz;
  int recomputed = 0;
  if (!cached && !planned) {
    // pass 1: background sweep; checkpoints at segment starts (with the
    // FP16 scale ring), the last segment's scatter operands stored directly
    CHK(zero(ctx, s_old, n));
    CHK(zero(ctx, s_cur, n));
    if (half) CHK(ring_reset(ctx, RING_S));
    const int klast = (S - 1) * L;
    for (int k = 0; k < nt; k++) {
      if (k > 0 && k % L == 0 && k / L <= S - 1) {
        const long long c = k / L - 1;
        float *ck = nth(ctx->d_ckpt, 2 * c, n, esz);
        HIPCHK(hipMemcpyAsync(ck, s_cur, Fb, hipMemcpyDeviceToDevice, ctx->stream));
        HIPCHK(hipMemcpyAsync(nth(ctx->d_ckpt, 2 * c + 1, n, esz), s_old, Fb,
                              hipMemcpyDeviceToDevice, ctx->stream));
        if (half)
          HIPCHK(hipMemcpyAsync(ctx->d_ckring + 8 * c, ctx->d_ring + RING_S, 8 * sizeof(float),
                                hipMemcpyDeviceToDevice, ctx->stream));
      }
      float *sig = k >= klast ? seg_level(ctx, k - klast) : nullptr;
      const bool into_seg = s_in_seg && k >= klast;
      if (into_seg) {
        // levels klast and klast-1 stay in s_cur and s_old; level j > klast
        // is in segment level j-1-klast
        sr.s1 = k > klast ? seg_level(ctx, k - 1 - klast) : s_cur;
        sr.s0 = k > klast + 1 ? seg_level(ctx, k - 2 - klast)
                              : (k > klast ? s_cur : s_old);
        sr.s_new = sig;
        sr.sig = nullptr;
      } else {
        sr.s1 = s_cur;
        sr.s0 = s_old;
        sr.s_new = nullptr;
        sr.sig = sig;
      }
      sr.hscale = half && sig ? ctx->d_hscale + (k - klast) : nullptr;
      sr.amp = ctx->d_amp + (long long)k * ctx->npts_s;
      sr.h = half_state(ctx, k);
      HIPCHK(born::launch_srec_step(geo, stu, sr, ctx->stream));
      if (!into_seg) std::swap(s_old, s_cur);
    }
  }
  // where nu sits, by the window of the sigma level a step reads
  const bool place = ctx->place_mode && ctx->seg_gap && !lam_pairs;
  std::vector<int> uplaces;
  if (place) {
    born::LamStep t{};  // a step without receivers
    t.vel = ctx->d_vel2; t.acc = ctx->d_acc; t.rec = recmap(ctx); t.tr = traces;
    t.tab = ctx->d_tab_d; t.inject = 0;
    t.h = half_state(ctx, 0); t.ring = ring(ctx, RING_T0); t.inj = inj;
    t.trmax = maxp(ctx, MAX_T); t.hscale = ctx->d_ring + HS_T;
    const std::vector<uintptr_t> key = {
        (uintptr_t)ctx->d_seg, ctx->seg_win, (uintptr_t)n, (uintptr_t)ctx->d_vel2,
        (uintptr_t)ctx->d_acc, (uintptr_t)ltu.shape, (uintptr_t)ltu.zchunk};
    // places: gap g < ngaps holds nu, the accumulator stays in its buffer;
    // gap g - ngaps holds both for g >= ngaps
    CHK(place_plan(ctx, ctx->place_lam, ctx->place_lam_key, key, ctx->seg_ngaps, uplaces,
                   [&](int w, int g, int i) {
      // the sigma levels of the window, one on per launch
      const long long a = (long long)w * ctx->seg_lpw;
      const long long b = std::min<long long>(L, a + ctx->seg_lpw);
      t.sig = seg_level(ctx, a + i % std::max(1LL, b - a));
      const int gu = g < ctx->seg_ngaps ? g : g - ctx->seg_ngaps;
      float *u0 = g < 0 ? ctx->d_w[0] : seg_gap_field(ctx, gu, 0, Fb);
      float *u1 = g < 0 ? ctx->d_w[1] : seg_gap_field(ctx, gu, 1, Fb);
      t.acc = g < ctx->seg_ngaps ? ctx->d_acc : seg_gap_field(ctx, gu, 2, Fb);
      t.u0 = i % 2 ? u1 : u0;
      t.u1 = i % 2 ? u0 : u1;
      return born::launch_lam_step(geo, ltu, t, ctx->stream); }, 2 * ctx->seg_ngaps));
  }
  // pass 2: adjoint state, segment by segment in reverse
  float *u_old = ctx->d_w[0], *u_cur = ctx->d_w[1];
  CHK(zero(ctx, u_old, n));
  CHK(zero(ctx, u_cur, n));
  CHK(zero(ctx, ctx->d_acc, n));
  if (half) {
    CHK(ring_reset(ctx, RING_U));
    HIPCHK(born::launch_fill(ctx->d_ring + MAX_TR, 1, 0.f, ctx->stream));
    HIPCHK(born::launch_absmax(traces, (long long)ctx->nrec * ctx->nt_rec,
                               (unsigned *)(ctx->d_ring + MAX_TR), ctx->stream));
  }
  born::LamStep lm{};
  lm.vel = ctx->d_vel2;
  float *acc = ctx->d_acc;  // the image accumulator, moved with nu
  lm.rec = recmap(ctx);
  lm.tr = traces;
  lm.ring = ring(ctx, RING_U);
  lm.inj = inj;
  lm.trmax = maxp(ctx, MAX_TR);
  int step = 0;  // adjoint steps taken (indexes the ring of nu)
  int uplace = -1, umoves = 0;  // where nu is: the work buffers
  // nu levels j-1, j, j+1, j+2 rotate through four buffers in two-step launches
  float *ub[4] = {u_old, u_cur, ctx->d_w2[0], ctx->d_w2[1]};
  born::LamPairStep lp{};
  lp.vel = ctx->d_vel2;
  lp.acc = ctx->d_acc;
  lp.rec = recmap(ctx);
  lp.tr = traces;
  const born::Tune ptu = tuning(ctx, K_LAMP);
  for (int sg = S - 1; sg >= 0; sg--) {
    const int k0 = sg * L, k1 = std::min(nt, k0 + L);
    if (sg < S - 1) {
      if (sg == 0) {
        CHK(zero(ctx, s_old, n));
        CHK(zero(ctx, s_cur, n));
        if (half) CHK(ring_reset(ctx, RING_S));
      } else {
        const long long c = sg - 1;
        HIPCHK(hipMemcpyAsync(s_cur, nth(ctx->d_ckpt, 2 * c, n, esz), Fb,
                              hipMemcpyDeviceToDevice, ctx->stream));
        HIPCHK(hipMemcpyAsync(s_old, nth(ctx->d_ckpt, 2 * c + 1, n, esz), Fb,
                              hipMemcpyDeviceToDevice, ctx->stream));
        if (half)
          HIPCHK(hipMemcpyAsync(ctx->d_ring + RING_S, ctx->d_ckring + 8 * c, 8 * sizeof(float),
                                hipMemcpyDeviceToDevice, ctx->stream));
      }
      for (int k = k0; k < k1; k++) {
        float *sig = seg_level(ctx, k - k0);
        if (s_in_seg) {
          sr.s1 = k > k0 ? seg_level(ctx, k - 1 - k0) : s_cur;
          sr.s0 = k > k0 + 1 ? seg_level(ctx, k - 2 - k0) : (k > k0 ? s_cur : s_old);
          sr.s_new = sig;
          sr.sig = nullptr;
        } else {
          sr.s1 = s_cur;
          sr.s0 = s_old;
          sr.s_new = nullptr;
          sr.sig = sig;
        }
        sr.hscale = half ? ctx->d_hscale + (k - k0) : nullptr;
        sr.amp = ctx->d_amp + (long long)k * ctx->npts_s;
        sr.h = half_state(ctx, k);
        HIPCHK(born::launch_srec_step(geo, stu, sr, ctx->stream));
        if (!s_in_seg) std::swap(s_old, s_cur);
        recomputed++;
      }
    }
    for (int k = k1 - 1; k >= k0;) {
      const int id = k / jt, ii = k - id * jt;
      if (lam_pairs && k - 1 >= k0) {
        const int idb = (k - 1) / jt, iib = k - 1 - idb * jt;
        lp.u_old = ub[0]; lp.u_cur = ub[1]; lp.u_a = ub[2]; lp.u_b = ub[3];
        lp.sig_a = seg_level(ctx, k - k0);
        lp.sig_b = seg_level(ctx, k - 1 - k0);
        lp.tab_a = ctx->d_tab_d + (long long)ii * 8;
        lp.tab_b = ctx->d_tab_d + (long long)iib * 8;
        lp.id_a = id;
        lp.id_b = idb;
        lp.inj_a = id + 7 < ctx->nt_rec;
        lp.inj_b = idb + 7 < ctx->nt_rec;
        HIPCHK(born::launch_lam_pair_step(geo, ptu, lp, ctx->stream));
        std::swap(ub[0], ub[2]);
        std::swap(ub[1], ub[3]);
        step += 2;
        k -= 2;
        continue;
      }
      if (place) {
        const int want = uplaces[(int)((k - k0) / ctx->seg_lpw)];
        if (want != uplace) {
          const int gu = want < ctx->seg_ngaps ? want : want - ctx->seg_ngaps;
          CHK(move_pair(ctx, gu, ctx->d_w[0], ctx->d_w[1], ub[0], ub[1], Fb));
          float *to = want < ctx->seg_ngaps ? ctx->d_acc : seg_gap_field(ctx, gu, 2, Fb);
          if (to != acc)
            HIPCHK(hipMemcpyAsync(to, acc, (size_t)n * sizeof(float),
                                  hipMemcpyDeviceToDevice, ctx->stream));
          acc = to;
          uplace = want;
          umoves++;
        }
      }
      lm.u1 = ub[1];
      lm.u0 = ub[0];
      lm.acc = acc;
      lm.sig = seg_level(ctx, k - k0);
      lm.hscale = half ? ctx->d_hscale + (k - k0) : nullptr;
      lm.tab = ctx->d_tab_d + (long long)ii * 8;
      lm.id = id;
      lm.inject = id + 7 < ctx->nt_rec;
      lm.h = half_state(ctx, step++);
      HIPCHK(born::launch_lam_step(geo, ltu, lm, ctx->stream));
      std::swap(ub[0], ub[1]);
      k--;
    }
  }
  HIPCHK(born::launch_adj_finish(geo, img, acc, ctx->d_vel2,
                                 scatter == 0 ? 1e-6f : 1.f, ctx->stream));
  ctx->last_segments = S;
  ctx->last_recomputed = recomputed;
  ctx->last_swept = !cached && !planned ? nt : 0;
  ctx->last_lam_moves = umoves;
  return BORN_OK;
}

// ----------------------------------------------------- host-pointer entries

int born_rtm_forward(born_ctx *ctx, int n1, int n2, int n3, int jt, const float *img,
                     float *rec, int npts, int nt, int nt_big, int rec_nx, int rec_ny) {
  (void)rec;
  (void)npts;
  CHK(need_space(ctx));
  CHK(check_dims(ctx, n1, n2, n3));
  if (!img) return fail(ctx, BORN_ERR_ARG, "null image");
  if (!ctx->d_rmap || !ctx->h_rec)
    return fail(ctx, BORN_ERR_STATE, "receivers have not been transferred");
  if (nt_big != ctx->nt_rec || rec_nx * rec_ny != ctx->nrec)
    return fail(ctx, BORN_ERR_ARG, "trace geometry differs from the transferred one");
  CHK(set_device(ctx));
  CHK(field_buf(ctx, &ctx->d_img));
  HIPCHK(hipMemcpyAsync(ctx->d_img, img, (size_t)ctx->n * sizeof(float),
                        hipMemcpyHostToDevice, ctx->stream));
  CHK(traces_in(ctx, ctx->h_rec));
  CHK(born_dev_forward(ctx, jt, nt, 0, ctx->d_img, ctx->d_tr, 0));
  CHK(traces_out(ctx, ctx->h_rec));
  return BORN_OK;
}

int born_forward_scatter(born_ctx *ctx, int n1, int n2, int n3, int jt,
                         const float *img, float *rec, int nt, int scatter) {
  CHK(need_space(ctx));
  CHK(check_dims(ctx, n1, n2, n3));
  if (!img || !rec) return fail(ctx, BORN_ERR_ARG, "null argument");
  if (!ctx->d_rmap) return fail(ctx, BORN_ERR_STATE, "receivers have not been transferred");
  CHK(set_device(ctx));
  CHK(field_buf(ctx, &ctx->d_img));
  HIPCHK(hipMemcpyAsync(ctx->d_img, img, (size_t)ctx->n * sizeof(float),
                        hipMemcpyHostToDevice, ctx->stream));
  CHK(traces_in(ctx, rec));
  CHK(born_dev_forward(ctx, jt, nt, scatter, ctx->d_img, ctx->d_tr, 0));
  CHK(traces_out(ctx, rec));
  return BORN_OK;
}

int born_source_prop(born_ctx *ctx, int nx, int ny, int nz, int damp, int get_last,
                     float *p0, float *p1, int jts, int npts, int nt) {
  (void)damp;
  (void)get_last;
  (void)npts;
  CHK(need_space(ctx));
  CHK(check_dims(ctx, nx, ny, nz));
  if (!p0 || !p1) return fail(ctx, BORN_ERR_ARG, "null state");
  CHK(set_device(ctx));
  CHK(field_buf(ctx, &ctx->d_p0));
  CHK(field_buf(ctx, &ctx->d_p1));
  const size_t B = (size_t)ctx->n * sizeof(float);
  HIPCHK(hipMemcpyAsync(ctx->d_p0, p0, B, hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(hipMemcpyAsync(ctx->d_p1, p1, B, hipMemcpyHostToDevice, ctx->stream));
  CHK(born_dev_source_prop(ctx, jts, nt, ctx->d_p0, ctx->d_p1));
  HIPCHK(hipMemcpyAsync(p0, ctx->d_p0, B, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipMemcpyAsync(p1, ctx->d_p1, B, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return BORN_OK;
}

int born_rtm_adjoint(born_ctx *ctx, int n1, int n2, int n3, int jtd, float *p0,
                     float *p1, float *img, int npts_s, int nt) {
  (void)jtd;
  (void)npts_s;
  CHK(need_space(ctx));
  CHK(check_dims(ctx, n1, n2, n3));
  if (!p0 || !p1 || !img) return fail(ctx, BORN_ERR_ARG, "null argument");
  if (!ctx->d_rmap || !ctx->h_rec)
    return fail(ctx, BORN_ERR_STATE, "receivers have not been transferred");
  CHK(set_device(ctx));
  CHK(field_buf(ctx, &ctx->d_p0));
  CHK(field_buf(ctx, &ctx->d_p1));
  CHK(field_buf(ctx, &ctx->d_img));
  const size_t B = (size_t)ctx->n * sizeof(float);
  HIPCHK(hipMemcpyAsync(ctx->d_p0, p0, B, hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(hipMemcpyAsync(ctx->d_p1, p1, B, hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(hipMemcpyAsync(ctx->d_img, img, B, hipMemcpyHostToDevice, ctx->stream));
  CHK(traces_in(ctx, ctx->h_rec));
  CHK(born_dev_rtm_adjoint(ctx, nt, ctx->d_p0, ctx->d_p1, ctx->d_img, ctx->d_tr));
  HIPCHK(hipMemcpyAsync(img, ctx->d_img, B, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipMemcpyAsync(p0, ctx->d_p0, B, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipMemcpyAsync(p1, ctx->d_p1, B, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return BORN_OK;
}

int born_adjoint_exact(born_ctx *ctx, int n1, int n2, int n3, int jt, const float *rec,
                       float *img, int nt, int scatter) {
  CHK(need_space(ctx));
  CHK(check_dims(ctx, n1, n2, n3));
  if (!rec || !img) return fail(ctx, BORN_ERR_ARG, "null argument");
  if (!ctx->d_rmap) return fail(ctx, BORN_ERR_STATE, "receivers have not been transferred");
  CHK(set_device(ctx));
  CHK(field_buf(ctx, &ctx->d_img));
  const size_t B = (size_t)ctx->n * sizeof(float);
  HIPCHK(hipMemcpyAsync(ctx->d_img, img, B, hipMemcpyHostToDevice, ctx->stream));
  CHK(traces_in(ctx, rec));
  CHK(born_dev_adjoint_exact(ctx, jt, nt, scatter, ctx->d_tr, ctx->d_img));
  HIPCHK(hipMemcpyAsync(img, ctx->d_img, B, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return BORN_OK;
}

int born_set_checkpoint_budget(born_ctx *ctx, size_t bytes) {
  if (!ctx) return BORN_ERR_ARG;
  ctx->budget = bytes;
  return BORN_OK;
}

int born_last_checkpoint_plan(const born_ctx *ctx, int *segments, int *recomputed) {
  if (!ctx) return BORN_ERR_ARG;
  if (segments) *segments = ctx->last_segments;
  if (recomputed) *recomputed = ctx->last_recomputed;
  return BORN_OK;
}

int born_last_background_sweep(const born_ctx *ctx) {
  if (!ctx) return BORN_ERR_ARG;
  return ctx->last_swept;
}

int born_set_zchunk(born_ctx *ctx, int zchunk) {
  if (!ctx || zchunk < 0) return BORN_ERR_ARG;
  ctx->zchunk = zchunk;
  return BORN_OK;
}

int born_set_placement(born_ctx *ctx, int mode, long long window_bytes) {
  if (!ctx || mode < 0 || mode > 2 || window_bytes < 0) return BORN_ERR_ARG;
  ctx->place_mode = mode;
  ctx->place_window = window_bytes ? (size_t)window_bytes : (size_t)32 << 30;
  ctx->place_fwd_key.clear();
  ctx->place_lam_key.clear();
  return BORN_OK;
}

int born_last_placement(const born_ctx *ctx, int *forward_moves, int *adjoint_moves) {
  if (!ctx) return BORN_ERR_ARG;
  if (forward_moves) *forward_moves = ctx->last_fwd_moves;
  if (adjoint_moves) *adjoint_moves = ctx->last_lam_moves;
  return BORN_OK;
}

int born_set_step_pairs(born_ctx *ctx, int mode) {
  if (!ctx || mode < 0 || mode > 2) return BORN_ERR_ARG;
  ctx->pairs = mode;
  return BORN_OK;
}

int born_last_step_pairs(const born_ctx *ctx) {
  return ctx ? ctx->last_pairs | ctx->last_lam_pairs << 1 : 0;
}

int born_set_storage(born_ctx *ctx, int storage) {
  if (!ctx) return BORN_ERR_ARG;
  if (storage != BORN_STORAGE_F32 && storage != BORN_STORAGE_F16 &&
      storage != BORN_STORAGE_F24)
    return fail(ctx, BORN_ERR_ARG,
                "storage must be BORN_STORAGE_F32, BORN_STORAGE_F16 or BORN_STORAGE_F24");
  if (storage != ctx->storage) ctx->hist_version = -1;
  ctx->storage = storage;
  return BORN_OK;
}

int born_set_autotune(born_ctx *ctx, int on) {
  if (!ctx) return BORN_ERR_ARG;
  ctx->autotune = on != 0;
  if (on) reset_tuning(ctx);
  return BORN_OK;
}

int born_get_tuning(const born_ctx *ctx, int kernel, int *rows_per_thread,
                    int *waves_per_block, int *zchunk, float *ms) {
  if (!ctx || kernel < 0 || kernel >= K_NUM) return BORN_ERR_ARG;
  const born::Tune t = tuning(ctx, kernel);
  if (rows_per_thread)
    *rows_per_thread = is_pair(kernel) ? born::kPairShapes[t.shape].ry : born::kShapes[t.shape].ry;
  if (waves_per_block)
    *waves_per_block = is_pair(kernel) ? born::kPairShapes[t.shape].wy : born::kShapes[t.shape].wy;
  if (zchunk) *zchunk = t.zchunk;
  if (ms) *ms = ctx->tune_ms[kernel];
  return BORN_OK;
}

int born_get_tile(const born_ctx *ctx, int kernel, int *tile_x, int *tile_y,
                  int *cells_per_lane, int *lookahead) {
  if (!ctx || kernel < 0 || kernel >= K_NUM) return BORN_ERR_ARG;
  if (is_pair(kernel)) {
    const born::PairShape &ps = born::kPairShapes[tuning(ctx, kernel).shape];
    if (tile_x) *tile_x = born::kTX - 2 * born::kH;
    if (tile_y) *tile_y = ps.ry * ps.wy;
    if (cells_per_lane) *cells_per_lane = 1;
    if (lookahead) *lookahead = 1;
    return BORN_OK;
  }
  const born::TileShape &ts = born::kShapes[tuning(ctx, kernel).shape];
  if (tile_x) *tile_x = ts.lx * ts.vx;
  if (tile_y) *tile_y = ts.ry * (born::kTX * ts.wy / ts.lx);
  if (cells_per_lane) *cells_per_lane = ts.vx;
  if (lookahead) *lookahead = ts.la;
  return BORN_OK;
}

int born_set_tuning(born_ctx *ctx, int kernel, int rows_per_thread, int waves_per_block,
                    int cells_per_lane, int lookahead, int zchunk) {
  if (!ctx || kernel < 0 || kernel >= K_NUM) return BORN_ERR_ARG;
  if (zchunk < 1) return fail(ctx, BORN_ERR_ARG, "zchunk must be positive");
  if (is_pair(kernel)) {
    for (int sh = 0; sh < born::kNumPairShapes; sh++) {
      const born::PairShape &ps = born::kPairShapes[sh];
      if (ps.ry == rows_per_thread && ps.wy == waves_per_block && cells_per_lane == 1 &&
          lookahead == 1) {
        ctx->tune[kernel] = {sh, zchunk};
        ctx->tune_ms[kernel] = 0.f;
        return BORN_OK;
      }
    }
    return fail(ctx, BORN_ERR_ARG, "no such launch shape");
  }
  for (int sh = 0; sh < born::kNumShapes; sh++) {
    const born::TileShape &ts = born::kShapes[sh];
    if (ts.ry == rows_per_thread && ts.wy == waves_per_block && ts.vx == cells_per_lane &&
        ts.la == lookahead) {
      ctx->tune[kernel] = {sh, zchunk};
      ctx->tune_ms[kernel] = 0.f;  // fixed: the autotuner leaves it
      return BORN_OK;
    }
  }
  return fail(ctx, BORN_ERR_ARG, "no such launch shape");
}

void *born_get_stream(const born_ctx *ctx) { return ctx ? (void *)ctx->stream : nullptr; }

int born_synchronize(born_ctx *ctx) {
  if (!ctx) return BORN_ERR_ARG;
  CHK(set_device(ctx));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return BORN_OK;
}

// ============================================================ shot batches
// Many shots with their own windowed domains advance together: one launch
// per time step covers the tiles of every shot (march_batch_kernel), and
// the time loop is captured once in a HIP graph and replayed.

}  // extern "C"

struct born_batch {
  born_ctx *ctx = nullptr;
  int nshots = 0;
  std::vector<born::ShotDesc> desc;
  born::ShotDesc *d_desc = nullptr;
  float coeffs[13] = {};
  float tapsum = 0.f;
  long long cells = 0;
  int ntb = 0;
  bool even_nx = true;
  std::vector<long long> src0;  // source samples of shot s in d_src ([npts][ntb])
  int npts_total = 0;
  float *d_f[4] = {nullptr, nullptr, nullptr, nullptr};
  float *d_vel = nullptr, *d_model = nullptr, *d_sponge = nullptr, *d_src = nullptr;
  float *d_amp = nullptr;
  size_t amp_cap = 0;
  float *d_tr = nullptr, *d_tmp = nullptr;
  long long tr_total = 0;
  int max_nrec = 0;
  float *d_rings = nullptr;
  int *d_pts = nullptr, *d_slot = nullptr, *d_rmap = nullptr, *d_rnext = nullptr;
  // per storage form (F32, F16, F24): forward (scatter 0, 1), background
  // recompute (scatter 0, 1), adjoint state
  born::Tune tune[18];
  float tune_ms[18];
  size_t f_cap = 0;    // floats per field buffer (24-bit levels: 6 bytes per cell)
  long long max_shot_cells = 0;
  std::vector<float> vmin2;  // per shot: min of V over the stencil interior
  float bmin = 1.f;          // smallest sponge factor (the sponge of the batch)
  size_t acc_cap = 0, img_cap = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  int g_jt = -1, g_nt = -1, g_scatter = -1, g_storage = -1, g_shape = -1, g_zc = -1;
  long long g_tab = -1;  // ctx->tab_version the graph was captured with

  // exact adjoint: accumulator and image (cells floats), background history
  // and checkpoints (bytes), per-level scales [L][nshots] and checkpointed
  // s rings [S - 1][nshots][8] of scaled storage
  float *d_acc = nullptr, *d_img = nullptr;
  float *d_seg = nullptr, *d_ckpt = nullptr;
  size_t seg_bytes = 0, ckpt_bytes = 0;
  // bytes of d_seg whose level shells are zero for levels of seg_esz bytes
  // per cell (the step kernels never write a shell)
  size_t seg_clean = 0, seg_esz = 0;
  float *d_hscale = nullptr, *d_ckring = nullptr;
  size_t hscale_cap = 0, ckring_cap = 0;
  int last_segments = 0, last_recomputed = 0;
};

namespace {

void batch_free(born_batch *b) {
  born_ctx *ctx = b->ctx;
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
  for (int i = 0; i < 4; i++) (void)dev_alloc(ctx, &b->d_f[i], 0);
  (void)dev_alloc(ctx, &b->d_vel, 0);
  (void)dev_alloc(ctx, &b->d_model, 0);
  (void)dev_alloc(ctx, &b->d_sponge, 0);
  (void)dev_alloc(ctx, &b->d_src, 0);
  (void)dev_alloc(ctx, &b->d_amp, 0);
  (void)dev_alloc(ctx, &b->d_tr, 0);
  (void)dev_alloc(ctx, &b->d_tmp, 0);
  (void)dev_alloc(ctx, &b->d_rings, 0);
  (void)dev_alloc(ctx, &b->d_pts, 0);
  (void)dev_alloc(ctx, &b->d_slot, 0);
  (void)dev_alloc(ctx, &b->d_rmap, 0);
  (void)dev_alloc(ctx, &b->d_rnext, 0);
  (void)dev_alloc(ctx, &b->d_desc, 0);
  (void)dev_alloc(ctx, &b->d_acc, 0);
  (void)dev_alloc(ctx, &b->d_img, 0);
  (void)dev_alloc(ctx, &b->d_seg, 0);
  (void)dev_alloc(ctx, &b->d_ckpt, 0);
  (void)dev_alloc(ctx, &b->d_hscale, 0);
  (void)dev_alloc(ctx, &b->d_ckring, 0);
}

template <class T>
int upload(born_ctx *ctx, T **dst, const std::vector<T> &v) {
  CHK(dev_alloc(ctx, dst, std::max<size_t>(v.size(), 1)));
  if (!v.empty())
    HIPCHK(hipMemcpy(*dst, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  return BORN_OK;
}

int batch_build(born_batch *b, int nshots, const born_shot *shots) {
  born_ctx *ctx = b->ctx;
  b->nshots = nshots;
  b->desc.assign(nshots, born::ShotDesc{});
  b->src0.assign(nshots, 0);
  b->ntb = shots[0].nt_big;
  std::vector<float> sponge, src;
  std::vector<int> pts, slots, rmap, rnext;
  std::vector<float> bound(ctx->nbound);
  for (int i = 0; i < ctx->nbound; i++)
    bound[i] = expf(-ctx->bcB * ((float)ctx->nbound - i));
  b->bmin = 1.f;
  for (float v : bound) b->bmin = std::min(b->bmin, v);
  b->vmin2.clear();
  long long cells = 0, tr = 0;
  for (int s = 0; s < nshots; s++) {
    const born_shot &sh = shots[s];
    born::ShotDesc &d = b->desc[s];
    const std::string tag = "shot " + std::to_string(s) + ": ";
    if (sh.nx < 9 || sh.ny < 9 || sh.nz < 9)
      return fail(ctx, BORN_ERR_ARG, tag + "each axis needs at least 9 samples");
    const long long n = (long long)sh.nx * sh.ny * sh.nz;
    if (n > (1ll << 30)) return fail(ctx, BORN_ERR_UNSUPPORTED, tag + "more than 2^30 cells");
    if (!sh.vel || !sh.model) return fail(ctx, BORN_ERR_ARG, tag + "null velocity or model");
    if (sh.nt_big != b->ntb)
      return fail(ctx, BORN_ERR_ARG, tag + "all shots need the same trace length");
    if (sh.npts < 0 || sh.npts > born::kMaxPts || (sh.npts && (!sh.src_locs || !sh.src_vals)))
      return fail(ctx, BORN_ERR_UNSUPPORTED, tag + "bad source (at most 32 points)");
    if (sh.rec_nx < 1 || sh.rec_ny < 1 || !sh.rec_locs)
      return fail(ctx, BORN_ERR_ARG, tag + "bad receiver geometry");
    if (sh.nx % 2) b->even_nx = false;
    d.nx = sh.nx;
    d.ny = sh.ny;
    d.nz = sh.nz;
    d.npts = sh.npts;
    d.cell0 = cells;
    // every shot's first cell at a multiple of 64: whole blocks of 24-bit
    // levels, and an even offset for the two-cell-per-lane tiles
    cells += (n + 63) / 64 * 64;
    b->max_shot_cells = std::max(b->max_shot_cells, n);
    d.sp0 = (int)sponge.size();
    for (int ax = 0; ax < 3; ax++) {
      const int na = ax == 0 ? sh.nx : ax == 1 ? sh.ny : sh.nz;
      std::vector<float> t = axis_sponge(na, ctx->nbound, bound);
      sponge.insert(sponge.end(), t.begin(), t.end());
    }
    d.pts0 = (int)(pts.size() / 4);
    for (int i = 0; i < sh.npts; i++) {
      int x, y, z;
      if (!interior_of(sh.nx, sh.ny, sh.nz, sh.src_locs[i], &x, &y, &z))
        return fail(ctx, BORN_ERR_UNSUPPORTED,
                    tag + "source point " + std::to_string(i) + " outside the interior");
      pts.insert(pts.end(), {x, y, z, 0});
    }
    b->src0[s] = (long long)src.size();
    if (sh.npts) src.insert(src.end(), sh.src_vals, sh.src_vals + (size_t)sh.npts * b->ntb);
    const int nrec = sh.rec_nx * sh.rec_ny;
    RecTables rt;
    CHK(receiver_tables(ctx, sh.nx, sh.ny, sh.nz, nrec, sh.rec_locs, &rt));
    d.slot0 = (int)slots.size();
    slots.insert(slots.end(), rt.slot_of_z.begin(), rt.slot_of_z.end());
    d.rmap0 = (long long)rmap.size();
    rmap.insert(rmap.end(), rt.rmap.begin(), rt.rmap.end());
    d.rnext0 = (int)rnext.size();
    rnext.insert(rnext.end(), rt.rnext.begin(), rt.rnext.end());
    d.nrec = nrec;
    b->max_nrec = std::max(b->max_nrec, nrec);
    d.tr0 = tr;
    tr += (long long)nrec * b->ntb;
    d.ring0 = s * born::kRingWords;
    float vmax = 0.f, vmin = INFINITY;
    for (long long i = 0; i < n; i++) {
      vmax = std::max(vmax, fabsf(sh.vel[i]));
      const int x = (int)(i % sh.nx), y = (int)((i / sh.nx) % sh.ny);
      const int z = (int)(i / ((long long)sh.nx * sh.ny));
      if (x >= 4 && x < sh.nx - 4 && y >= 4 && y < sh.ny - 4 && z >= 4 && z < sh.nz - 4)
        vmin = std::min(vmin, sh.vel[i]);
    }
    d.gain = 2.f + vmax * b->tapsum;
    d.vmax = vmax;
    d.per_cell = rt.per_cell;
    b->vmin2.push_back(vmin);
  }
  b->cells = cells;
  b->tr_total = tr;
  b->npts_total = (int)(pts.size() / 4);
  CHK(set_device(ctx));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  for (int i = 0; i < 4; i++) CHK(dev_alloc(ctx, &b->d_f[i], (size_t)cells));
  b->f_cap = (size_t)cells;
  CHK(dev_alloc(ctx, &b->d_vel, (size_t)cells));
  CHK(dev_alloc(ctx, &b->d_model, (size_t)cells));
  for (int s = 0; s < nshots; s++) {
    const born::ShotDesc &d = b->desc[s];
    const size_t bytes = (size_t)d.nx * d.ny * d.nz * sizeof(float);
    HIPCHK(hipMemcpy(b->d_vel + d.cell0, shots[s].vel, bytes, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(b->d_model + d.cell0, shots[s].model, bytes, hipMemcpyHostToDevice));
  }
  CHK(upload(ctx, &b->d_sponge, sponge));
  CHK(upload(ctx, &b->d_src, src));
  CHK(upload(ctx, &b->d_pts, pts));
  CHK(upload(ctx, &b->d_slot, slots));
  CHK(upload(ctx, &b->d_rmap, rmap));
  CHK(upload(ctx, &b->d_rnext, rnext));
  CHK(upload(ctx, &b->d_desc, b->desc));
  CHK(dev_alloc(ctx, &b->d_tr, (size_t)std::max(tr, 1ll)));
  CHK(dev_alloc(ctx, &b->d_tmp, (size_t)b->max_nrec * b->ntb));
  CHK(dev_alloc(ctx, &b->d_rings, (size_t)nshots * born::kRingWords));
  for (int k = 0; k < 18; k++) {
    b->tune[k] = {3, 64};
    b->tune_ms[k] = -1.f;
  }
  return BORN_OK;
}

// the step launch of the batch at time step it
born::BatchFwdStep batch_step(born_batch *b, int it, int jt, int scatter, int storage) {
  born_ctx *ctx = b->ctx;
  born::BatchFwdStep st{};
  born::FwdStep &a = st.a;
  const bool odd = it % 2;
  a.s1 = b->d_f[odd ? 0 : 1];
  a.s0 = b->d_f[odd ? 1 : 0];
  a.r1 = b->d_f[odd ? 2 : 3];
  a.r0 = b->d_f[odd ? 3 : 2];
  a.vel = b->d_vel;
  a.model = b->d_model;
  a.src = {b->d_pts, 0};
  a.amp = b->d_amp;
  a.rec.plane_slot = b->d_slot;
  a.rec.rmap = b->d_rmap;
  a.rec.rnext = b->d_rnext;
  a.rec.nrec = 0;
  a.rec.ntb = b->ntb;
  a.tr = b->d_tr;
  const int id = it / jt, ii = it - id * jt;
  a.tab = ctx->d_tab_d + (long long)ii * 8;
  a.id = id;
  a.extract = id + 7 < b->ntb;
  a.scatter = scatter;
  a.h.on = storage;
  a.h.sl = slots(it);
  a.h.tapsum = b->tapsum;
  st.sponge = b->d_sponge;
  st.rings = b->d_rings;
  st.desc = b->d_desc;
  st.host_desc = b->desc.data();
  st.nshots = b->nshots;
  st.step = it;
  st.npts_total = b->npts_total;
  return st;
}

}  // namespace

extern "C" {

int born_batch_create(born_ctx *ctx, int nshots, const born_shot *shots, float d1, float d2,
                      float d3, born_batch **out) {
  if (!ctx || !out) return BORN_ERR_ARG;
  *out = nullptr;
  if (nshots < 1 || nshots > born::kMaxBatch || !shots)
    return fail(ctx, BORN_ERR_ARG,
                "a batch holds 1 to " + std::to_string(born::kMaxBatch) + " shots");
  if (!(d1 > 0) || !(d2 > 0) || !(d3 > 0))
    return fail(ctx, BORN_ERR_ARG, "sampling must be positive");
  born_batch *b = new born_batch;
  b->ctx = ctx;
  stencil_coeffs(d1, d2, d3, b->coeffs);
  b->tapsum = fabsf(b->coeffs[0]);
  for (int i = 1; i < 13; i++) b->tapsum += 2.f * fabsf(b->coeffs[i]);
  const int rc = batch_build(b, nshots, shots);
  if (rc != BORN_OK) {
    batch_free(b);
    delete b;
    return rc;
  }
  *out = b;
  return BORN_OK;
}

int born_batch_destroy(born_batch *b) {
  if (!b) return BORN_ERR_ARG;
  (void)set_device(b->ctx);
  (void)hipStreamSynchronize(b->ctx->stream);
  batch_free(b);
  delete b;
  return BORN_OK;
}

int born_batch_forward(born_batch *b, int jt, int nt, int scatter) {
  if (!b) return BORN_ERR_ARG;
  born_ctx *ctx = b->ctx;
  if (!ctx->d_tab_s || !ctx->d_tab_d)
    return fail(ctx, BORN_ERR_STATE, "sinc tables have not been transferred");
  CHK(check_jt(ctx, jt));
  if (nt < 0 || (scatter != 0 && scatter != 1))
    return fail(ctx, BORN_ERR_ARG, "bad forward arguments");
  CHK(set_device(ctx));
  const int storage = ctx->storage;
  const bool half = storage != BORN_STORAGE_F32;  // scaled 16-bit levels
  const bool split = storage == BORN_STORAGE_F24;
  const int kind = 6 * storage + scatter;
  if (split && (unsigned long long)b->max_shot_cells * 6 >= (1ull << 32))
    return fail(ctx, BORN_ERR_UNSUPPORTED,
                "24-bit storage needs fewer than 715,827,882 cells per shot");
  // 24-bit levels of two fields need 6 bytes per cell
  const size_t fneed = split ? ((size_t)b->cells * 3 + 1) / 2 : (size_t)b->cells;
  if (b->f_cap < fneed) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);  // the graph holds the old buffers
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    b->f_cap = 0;
    for (int i = 0; i < 4; i++) CHK(dev_alloc(ctx, &b->d_f[i], fneed));
    b->f_cap = fneed;
  }
  const born::Taps taps = taps_of(b->coeffs);
  CHK(ensure(ctx, &b->d_amp, &b->amp_cap,
             (size_t)std::max(1, b->npts_total) * std::max(nt, 1)));
  {  // tune on the batch arrays (cleared below), no traces written
    HIPCHK(born::launch_batch_ring_init(b->d_rings, b->nshots, ctx->stream));
    born::BatchFwdStep t = batch_step(b, 0, jt, scatter, storage);
    t.a.extract = 0;
    CHK(autotune_into(ctx, &b->tune[kind], &b->tune_ms[kind], b->even_nx, 0,
                      [&](const born::Tune &tu) {
                        return born::launch_batch_fwd_step(taps, tu, t, ctx->stream);
                      }));
  }
  born::Tune tu = b->tune[kind];
  if (ctx->zchunk > 0) tu.zchunk = ctx->zchunk;
  tu = fit_nx(tu, b->even_nx);
  // initial state: fields and traces zero, amplitudes of every shot, scales
  for (int i = 0; i < 4; i++) CHK(zero(ctx, b->d_f[i], (long long)b->f_cap));
  CHK(zero(ctx, b->d_tr, b->tr_total));
  for (int s = 0; s < b->nshots; s++) {
    const born::ShotDesc &d = b->desc[s];
    if (d.npts)
      HIPCHK(born::launch_src_amp(b->d_amp + d.pts0, nt, d.npts, b->d_src + b->src0[s], b->ntb,
                                  ctx->d_tab_s, ctx->nsinc_s, jt, ctx->jts, 0, 1 / (float)jt,
                                  ctx->stream, b->npts_total));
  }
  if (half) {
    HIPCHK(born::launch_batch_ring_init(b->d_rings, b->nshots, ctx->stream));
    for (int s = 0; s < b->nshots; s++) {
      const born::ShotDesc &d = b->desc[s];
      HIPCHK(born::launch_absmax(b->d_model + d.cell0, (long long)d.nx * d.ny * d.nz,
                                 (unsigned *)(b->d_rings + d.ring0 + 16), ctx->stream));
    }
  }
  if (nt == 0) return BORN_OK;
  // the time loop: captured once per (jt, nt, scatter, storage, shape), replayed
  const bool same = b->exec && b->g_jt == jt && b->g_nt == nt && b->g_scatter == scatter &&
                    b->g_storage == ctx->storage && b->g_shape == tu.shape &&
                    b->g_zc == tu.zchunk && b->g_tab == ctx->tab_version;
  if (!same) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    HIPCHK(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
    hipError_t le = hipSuccess;
    for (int it = 0; it < nt && le == hipSuccess; it++)
      le = born::launch_batch_fwd_step(taps, tu, batch_step(b, it, jt, scatter, storage),
                                       ctx->stream);
    hipGraph_t g = nullptr;
    const hipError_t ce = hipStreamEndCapture(ctx->stream, &g);
    if (le != hipSuccess || ce != hipSuccess) {
      if (g) (void)hipGraphDestroy(g);
      return fail(ctx, BORN_ERR_HIP,
                  std::string("capture of the batch time loop: ") +
                      hipGetErrorString(le != hipSuccess ? le : ce));
    }
    b->graph = g;
    HIPCHK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    b->g_jt = jt;
    b->g_nt = nt;
    b->g_scatter = scatter;
    b->g_storage = ctx->storage;
    b->g_shape = tu.shape;
    b->g_zc = tu.zchunk;
    b->g_tab = ctx->tab_version;
  }
  HIPCHK(hipGraphLaunch(b->exec, ctx->stream));
  return BORN_OK;
}

int born_batch_traces(born_batch *b, int shot, float *rec) {
  if (!b || !rec) return BORN_ERR_ARG;
  born_ctx *ctx = b->ctx;
  if (shot < 0 || shot >= b->nshots) return fail(ctx, BORN_ERR_ARG, "no such shot");
  CHK(set_device(ctx));
  const born::ShotDesc &d = b->desc[shot];
  HIPCHK(born::launch_transpose(b->d_tmp, b->d_tr + d.tr0, b->ntb, d.nrec, ctx->stream));
  HIPCHK(hipMemcpyAsync(rec, b->d_tmp, (size_t)d.nrec * b->ntb * sizeof(float),
                        hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return BORN_OK;
}

int born_batch_set_traces(born_batch *b, int shot, const float *rec) {
  if (!b || !rec) return BORN_ERR_ARG;
  born_ctx *ctx = b->ctx;
  if (shot < 0 || shot >= b->nshots) return fail(ctx, BORN_ERR_ARG, "no such shot");
  CHK(set_device(ctx));
  const born::ShotDesc &d = b->desc[shot];
  HIPCHK(hipMemcpyAsync(b->d_tmp, rec, (size_t)d.nrec * b->ntb * sizeof(float),
                        hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(born::launch_transpose(b->d_tr + d.tr0, b->d_tmp, d.nrec, b->ntb, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return BORN_OK;
}

int born_batch_adjoint_exact(born_batch *b, int jt, int nt, int scatter) {
  if (!b) return BORN_ERR_ARG;
  born_ctx *ctx = b->ctx;
  if (!ctx->d_tab_s || !ctx->d_tab_d)
    return fail(ctx, BORN_ERR_STATE, "sinc tables have not been transferred");
  CHK(check_jt(ctx, jt));
  if (nt < 0 || (scatter != 0 && scatter != 1))
    return fail(ctx, BORN_ERR_ARG, "bad adjoint arguments");
  for (int s = 0; s < b->nshots; s++)
    if (!(b->vmin2[s] * b->bmin > 0.f))
      return fail(ctx, BORN_ERR_UNSUPPORTED,
                  "shot " + std::to_string(s) +
                      ": the exact adjoint needs the velocity and the sponge factors > 0 "
                      "in the interior");
  CHK(set_device(ctx));
  const int storage = ctx->storage;
  const bool half = storage != BORN_STORAGE_F32;  // scaled 16-bit levels
  const bool split = storage == BORN_STORAGE_F24;
  if (split && (unsigned long long)b->max_shot_cells * 6 >= (1ull << 32))
    return fail(ctx, BORN_ERR_UNSUPPORTED,
                "24-bit storage needs fewer than 715,827,882 cells per shot");
  const size_t esz = elem(ctx);
  const long long cells = b->cells;
  const size_t fl = ((size_t)cells * esz + 3) / 4;  // floats of one single-field level
  CHK(ensure(ctx, &b->d_acc, &b->acc_cap, (size_t)cells));
  CHK(ensure(ctx, &b->d_img, &b->img_cap, (size_t)cells));
  CHK(zero(ctx, b->d_img, cells));
  if (nt == 0) return BORN_OK;
  CHK(ensure(ctx, &b->d_amp, &b->amp_cap,
             (size_t)std::max(1, b->npts_total) * std::max(nt, 1)));
  const born::Taps taps = taps_of(b->coeffs);
  const int skind = 6 * storage + 2 + scatter, lkind = 6 * storage + 4;
  auto common = [&](born::BatchCommon &c, int step) {
    c.npts_total = b->npts_total;
    c.sponge = b->d_sponge;
    c.rings = b->d_rings;
    c.desc = b->d_desc;
    c.host_desc = b->desc.data();
    c.nshots = b->nshots;
    c.step = step;
  };
  auto hstate = [&](int k) {
    born::Half h;
    h.on = storage;
    h.sl = slots(k);
    h.tapsum = b->tapsum;
    h.gain = 0.f;  // per shot (ShotDesc::gain)
    return h;
  };
  auto tune_of = [&](int kind) {
    born::Tune t = b->tune[kind];
    if (ctx->zchunk > 0) t.zchunk = ctx->zchunk;
    t = fit_nx(t, b->even_nx);
    return t;
  };
  float *s_old = b->d_f[2], *s_cur = b->d_f[3];
  float *u_old = b->d_f[0], *u_cur = b->d_f[1];
  {  // tune both kernels on the batch fields and scratch scales (rings reset below)
    HIPCHK(born::launch_batch_ring_init(b->d_rings, b->nshots, ctx->stream));
    HIPCHK(born::launch_batch_ring_reset(b->d_rings, b->nshots, born::BR_U, 0, ctx->stream));
    CHK(ensure(ctx, &b->d_hscale, &b->hscale_cap, (size_t)b->nshots));
    HIPCHK(born::launch_fill(b->d_hscale, b->nshots, 1.f, ctx->stream));
    born::BatchSrecStep t{};
    common(t, 0);
    t.a.s1 = s_cur; t.a.s0 = s_old; t.a.vel = b->d_vel; t.a.sig = b->d_acc;
    t.a.src = {b->d_pts, 0}; t.a.amp = b->d_amp; t.a.scatter = scatter;
    t.a.h = hstate(0); t.a.hscale = b->d_hscale;
    CHK(autotune_into(ctx, &b->tune[skind], &b->tune_ms[skind], b->even_nx, 0,
                      [&](const born::Tune &tu) {
                        return born::launch_batch_srec_step(taps, tu, t, ctx->stream);
                      }));
    born::BatchLamStep l{};
    common(l, 0);
    l.a.u1 = u_cur; l.a.u0 = u_old; l.a.vel = b->d_vel; l.a.sig = s_old; l.a.acc = b->d_acc;
    l.a.rec.plane_slot = b->d_slot; l.a.rec.rmap = b->d_rmap; l.a.rec.rnext = b->d_rnext;
    l.a.rec.ntb = b->ntb; l.a.tr = b->d_tr; l.a.tab = ctx->d_tab_d; l.a.inject = 0;
    l.a.h = hstate(0); l.a.inj = ctx->tabd_abs; l.a.hscale = b->d_hscale;
    CHK(autotune_into(ctx, &b->tune[lkind], &b->tune_ms[lkind], b->even_nx, 0,
                      [&](const born::Tune &tu) {
                        return born::launch_batch_lam_step(taps, tu, l, ctx->stream);
                      }));
  }
  const born::Tune stu = tune_of(skind), ltu = tune_of(lkind);
  // checkpoint plan: the fewest segments S with (L + 2 (S - 1)) levels in the
  // budget, L = ceil(nt / S) (every shot shares the plan)
  const size_t F = (size_t)cells * esz;
  size_t avail = ctx->budget;
  if (!avail) {
    const size_t fr = free_bytes(ctx) + b->seg_bytes + b->ckpt_bytes;
    const size_t reserve = (size_t)2 << 30;
    avail = fr > reserve ? fr - reserve : 0;
  }
  int S = 0;
  for (int sg = 1; sg <= nt; sg++) {
    const size_t l = (size_t)((nt + sg - 1) / sg);
    if ((l + 2 * (size_t)(sg - 1)) * F <= avail) {
      S = sg;
      break;
    }
    if (l <= 2) break;
  }
  if (S == 0) return fail(ctx, BORN_ERR_NOMEM, "no checkpoint plan fits the HBM budget");
  const int L = (nt + S - 1) / S;
  S = (nt + L - 1) / L;
  if (b->seg_bytes < (size_t)L * F) {
    b->seg_bytes = 0;
    b->seg_clean = 0;
    CHK(dev_alloc(ctx, &b->d_seg, ((size_t)L * F + 3) / 4));
    b->seg_bytes = (size_t)L * F;
  }
  if (b->seg_esz != esz) {
    b->seg_clean = 0;
    b->seg_esz = esz;
  }
  if (b->seg_clean < (size_t)L * F) {
    const size_t from = b->seg_clean / 4, to = ((size_t)L * F + 3) / 4;
    CHK(zero(ctx, b->d_seg + from, (long long)(to - from)));
    b->seg_clean = to * 4;
  }
  if (b->ckpt_bytes < 2 * (size_t)(S - 1) * F) {
    b->ckpt_bytes = 0;
    CHK(dev_alloc(ctx, &b->d_ckpt, (2 * (size_t)(S - 1) * F + 3) / 4));
    b->ckpt_bytes = 2 * (size_t)(S - 1) * F;
  }
  if (half) {
    CHK(ensure(ctx, &b->d_hscale, &b->hscale_cap, (size_t)L * b->nshots));
    CHK(ensure(ctx, &b->d_ckring, &b->ckring_cap, 8 * (size_t)b->nshots * std::max(S - 1, 1)));
  }
  // amplitudes of every shot, as the forward (src_amp per shot)
  for (int s = 0; s < b->nshots; s++) {
    const born::ShotDesc &d = b->desc[s];
    if (d.npts)
      HIPCHK(born::launch_src_amp(b->d_amp + d.pts0, nt, d.npts, b->d_src + b->src0[s], b->ntb,
                                  ctx->d_tab_s, ctx->nsinc_s, jt, ctx->jts, 0, 1 / (float)jt,
                                  ctx->stream, b->npts_total));
  }
  auto lev = [&](float *base, long long k) { return nth(base, k, cells, esz); };
  // the 8 s-ring words of every shot to / from checkpoint c
  auto ring_copy = [&](long long c, bool save) -> int {
    float *ck = b->d_ckring + 8 * (size_t)b->nshots * c;
    float *rg = b->d_rings + born::BR_S;
    HIPCHK(hipMemcpy2DAsync(save ? ck : rg, (save ? 8 : born::kRingWords) * sizeof(float),
                            save ? rg : ck, (save ? born::kRingWords : 8) * sizeof(float),
                            8 * sizeof(float), b->nshots, hipMemcpyDeviceToDevice,
                            ctx->stream));
    return BORN_OK;
  };
  const bool s_in_seg = scatter == 0;
  born::BatchSrecStep sr{};
  common(sr, 0);
  sr.a.vel = b->d_vel;
  sr.a.src = {b->d_pts, 0};
  sr.a.amp = b->d_amp;
  sr.a.scatter = scatter;
  auto srec = [&](int k, float *seg_level, float *sig_level, float *s1, float *s0,
                  float *s_new) -> int {
    sr.step = k;
    sr.a.s1 = s1;
    sr.a.s0 = s0;
    sr.a.s_new = s_new;
    sr.a.sig = sig_level;
    sr.a.hscale = half && seg_level ? b->d_hscale + (size_t)b->nshots * (k % L) : nullptr;
    sr.a.h = hstate(k);
    HIPCHK(born::launch_batch_srec_step(taps, stu, sr, ctx->stream));
    return BORN_OK;
  };
  // pass 1: background sweep; checkpoints at segment starts (with the s
  // rings), the last segment's scatter operands into the segment history
  CHK(zero(ctx, s_old, (long long)fl));
  CHK(zero(ctx, s_cur, (long long)fl));
  if (half) HIPCHK(born::launch_batch_ring_reset(b->d_rings, b->nshots, born::BR_S, 0, ctx->stream));
  const int klast = (S - 1) * L;
  for (int k = 0; k < nt; k++) {
    if (k > 0 && k % L == 0 && k / L <= S - 1) {
      const long long c = k / L - 1;
      HIPCHK(hipMemcpyAsync(lev(b->d_ckpt, 2 * c), s_cur, F, hipMemcpyDeviceToDevice, ctx->stream));
      HIPCHK(hipMemcpyAsync(lev(b->d_ckpt, 2 * c + 1), s_old, F, hipMemcpyDeviceToDevice,
                            ctx->stream));
      if (half) CHK(ring_copy(c, true));
    }
    float *sig = k >= klast ? lev(b->d_seg, k - klast) : nullptr;
    if (s_in_seg && k >= klast) {
      // levels klast and klast-1 stay in s_cur and s_old; level j > klast is
      // in segment level j-1-klast
      float *s1 = k > klast ? lev(b->d_seg, k - 1 - klast) : s_cur;
      float *s0 = k > klast + 1 ? lev(b->d_seg, k - 2 - klast) : (k > klast ? s_cur : s_old);
      CHK(srec(k, sig, nullptr, s1, s0, sig));
    } else {
      CHK(srec(k, sig, sig, s_cur, s_old, nullptr));
      std::swap(s_old, s_cur);
    }
  }
  // pass 2: adjoint state, segment by segment in reverse
  CHK(zero(ctx, u_old, (long long)fl));
  CHK(zero(ctx, u_cur, (long long)fl));
  CHK(zero(ctx, b->d_acc, cells));
  if (half) {
    HIPCHK(born::launch_batch_ring_reset(b->d_rings, b->nshots, born::BR_U, 0, ctx->stream));
    HIPCHK(born::launch_batch_ring_reset(b->d_rings, b->nshots, born::BR_MAXTR, 1, ctx->stream));
    for (int s = 0; s < b->nshots; s++) {
      const born::ShotDesc &d = b->desc[s];
      HIPCHK(born::launch_absmax(b->d_tr + d.tr0, (long long)d.nrec * b->ntb,
                                 (unsigned *)(b->d_rings + d.ring0 + born::BR_MAXTR),
                                 ctx->stream));
    }
  }
  born::BatchLamStep lm{};
  common(lm, 0);
  lm.a.vel = b->d_vel;
  lm.a.acc = b->d_acc;
  lm.a.rec.plane_slot = b->d_slot;
  lm.a.rec.rmap = b->d_rmap;
  lm.a.rec.rnext = b->d_rnext;
  lm.a.rec.ntb = b->ntb;
  lm.a.tr = b->d_tr;
  lm.a.inj = ctx->tabd_abs;
  int step = 0, recomputed = 0;
  for (int sg = S - 1; sg >= 0; sg--) {
    const int k0 = sg * L, k1 = std::min(nt, k0 + L);
    if (sg < S - 1) {
      if (sg == 0) {
        CHK(zero(ctx, s_old, (long long)fl));
        CHK(zero(ctx, s_cur, (long long)fl));
        if (half)
          HIPCHK(born::launch_batch_ring_reset(b->d_rings, b->nshots, born::BR_S, 0, ctx->stream));
      } else {
        const long long c = sg - 1;
        HIPCHK(hipMemcpyAsync(s_cur, lev(b->d_ckpt, 2 * c), F, hipMemcpyDeviceToDevice,
                              ctx->stream));
        HIPCHK(hipMemcpyAsync(s_old, lev(b->d_ckpt, 2 * c + 1), F, hipMemcpyDeviceToDevice,
                              ctx->stream));
        if (half) CHK(ring_copy(c, false));
      }
      for (int k = k0; k < k1; k++) {
        float *sig = lev(b->d_seg, k - k0);
        if (s_in_seg) {
          float *s1 = k > k0 ? lev(b->d_seg, k - 1 - k0) : s_cur;
          float *s0 = k > k0 + 1 ? lev(b->d_seg, k - 2 - k0) : (k > k0 ? s_cur : s_old);
          CHK(srec(k, sig, nullptr, s1, s0, sig));
        } else {
          CHK(srec(k, sig, sig, s_cur, s_old, nullptr));
          std::swap(s_old, s_cur);
        }
        recomputed++;
      }
    }
    for (int k = k1 - 1; k >= k0; k--) {
      const int id = k / jt, ii = k - id * jt;
      lm.step = k;
      lm.a.u1 = u_cur;
      lm.a.u0 = u_old;
      lm.a.sig = lev(b->d_seg, k - k0);
      lm.a.hscale = half ? b->d_hscale + (size_t)b->nshots * (k % L) : nullptr;
      lm.a.tab = ctx->d_tab_d + (long long)ii * 8;
      lm.a.id = id;
      lm.a.inject = id + 7 < b->ntb;
      lm.a.h = hstate(step++);
      HIPCHK(born::launch_batch_lam_step(taps, ltu, lm, ctx->stream));
      std::swap(u_old, u_cur);
    }
  }
  // image of every shot: acc / (V B), times 1e-6 for the reference's scatter
  for (int s = 0; s < b->nshots; s++) {
    const born::ShotDesc &d = b->desc[s];
    Geometry g{};
    g.nx = d.nx;
    g.ny = d.ny;
    g.nz = d.nz;
    g.sponge.bx = b->d_sponge + d.sp0;
    g.sponge.by = b->d_sponge + d.sp0 + d.nx;
    g.sponge.bz = b->d_sponge + d.sp0 + d.nx + d.ny;
    HIPCHK(born::launch_adj_finish(g, b->d_img + d.cell0, b->d_acc + d.cell0,
                                   b->d_vel + d.cell0, scatter == 0 ? 1e-6f : 1.f,
                                   ctx->stream));
  }
  b->last_segments = S;
  b->last_recomputed = recomputed;
  return BORN_OK;
}

int born_batch_image(born_batch *b, int shot, float *img) {
  if (!b || !img) return BORN_ERR_ARG;
  born_ctx *ctx = b->ctx;
  if (shot < 0 || shot >= b->nshots) return fail(ctx, BORN_ERR_ARG, "no such shot");
  if (!b->d_img) return fail(ctx, BORN_ERR_STATE, "no batch adjoint has run");
  CHK(set_device(ctx));
  const born::ShotDesc &d = b->desc[shot];
  HIPCHK(hipMemcpyAsync(img, b->d_img + d.cell0, (size_t)d.nx * d.ny * d.nz * sizeof(float),
                        hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return BORN_OK;
}

int born_batch_last_checkpoint_plan(const born_batch *b, int *segments, int *recomputed_steps) {
  if (!b) return BORN_ERR_ARG;
  if (segments) *segments = b->last_segments;
  if (recomputed_steps) *recomputed_steps = b->last_recomputed;
  return BORN_OK;
}

int born_batch_get_tile(const born_batch *b, int scatter, int *tile_x, int *tile_y,
                        int *zchunk, float *ms) {
  if (!b || (scatter != 0 && scatter != 1)) return BORN_ERR_ARG;
  const int kind = 6 * b->ctx->storage + scatter;
  const born::Tune t = b->tune[kind];
  const born::TileShape &ts = born::kShapes[t.shape];
  if (tile_x) *tile_x = ts.lx * ts.vx;
  if (tile_y) *tile_y = ts.ry * (born::kTX * ts.wy / ts.lx);
  if (zchunk) *zchunk = t.zchunk;
  if (ms) *ms = b->tune_ms[kind];
  return BORN_OK;
}

}  // extern "C"
