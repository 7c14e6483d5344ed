/*
 * born_hip.h -- C ABI of the MI355X (gfx950) Born propagator library
 * (sc17-born_amd/libborn_hip.so).
 *
 * This is the drop-in replacement for the reference's CUDA backend interface
 * base/gpu/gpu_funcs_3d.h:4-37 and for the baseProp virtuals it serves
 * (base/base_prop.h:10-68).  Every entry point of the first group names the
 * reference function it replaces; arguments keep the reference's meaning and
 * order.  Differences, all deliberate:
 *   - a context handle replaces the reference's file-scope device globals
 *     (gpu_finite_3d.h:37-99), so several propagators may coexist;
 *   - every call returns a status (BORN_OK or a negative BORN_ERR_*) and
 *     born_last_error() explains it, where the reference calls exit(-1)
 *     (gpu_funcs_3d.cu:58-70);
 *   - host-pointer calls follow the CPU backend's semantics (base/cpu_prop.cc),
 *     which is the parity target: the forward writes the traces into the buffer
 *     registered by born_transfer_receiver_func, source propagation returns the
 *     final two levels in p0/p1, and the adjoint accumulates into img.
 *
 * Layout: every 3-D array is [nz][ny][nx] float32 with x fastest
 * (base/cpu_prop.cc:229-231).  Host traces are [nrec][nt_big]
 * (base/rtm_zero_op_3d.cc:134); device-resident traces are time-major
 * [nt_big][nrec].
 *
 * Threading: one context per host thread; calls on one context are
 * serialised on the context's stream.
 */
#ifndef BORN_HIP_H
#define BORN_HIP_H
#include <stddef.h>
#ifdef __cplusplus
extern "C" {
#endif

#define BORN_OK 0
#define BORN_ERR_ARG (-1)         /* invalid argument or size mismatch */
#define BORN_ERR_STATE (-2)       /* call order violated (missing transfer) */
#define BORN_ERR_HIP (-3)         /* HIP runtime error */
#define BORN_ERR_NOMEM (-4)       /* device allocation failed */
#define BORN_ERR_UNSUPPORTED (-5) /* geometry the kernels do not handle */

typedef struct born_ctx born_ctx;

/* ---- reference interface (gpu_funcs_3d.h / baseProp) -------------------- */

/* setup_cuda(int) (gpu_funcs_3d.h:7, gpuProp ctor gpu_prop.cc:6-12):
 * binds the context to one HIP device and creates its stream. */
int born_create(born_ctx **ctx, int device);
int born_destroy(born_ctx *ctx);
const char *born_last_error(const born_ctx *ctx);

/* create_gpu_space (gpu_funcs_3d.h:35) / baseProp::createSpace
 * (base_prop.h:68; cpu_prop.cc:298-332).  Like the CPU backend, bc_a/bc_b/bc_y
 * are accepted and ignored: the sponge is 40 cells with bcB = .0005. */
int born_create_space(born_ctx *ctx, float d1, float d2, float d3, float bc_a,
                      float bc_b, float bc_y, int nx, int ny, int nz);

/* set_ntblock (gpu_funcs_3d.h:14) / baseProp::setNtblock (base_prop.h:13).
 * Accepted and ignored: trace blocking is not needed (traces stay in HBM). */
int born_set_ntblock(born_ctx *ctx, int nb);

/* transfer_sinc_table_s / _d (gpu_funcs_3d.h:18,21) / baseProp
 * transferSincTableS / D (base_prop.h:55,17).  table is [jt][nsinc]
 * row-major (the flattening of gpu_prop.cc:15-24); nsinc must be 8. */
int born_transfer_sinc_table_s(born_ctx *ctx, int nsinc, int jts,
                               const float *table);
int born_transfer_sinc_table_d(born_ctx *ctx, int nsinc, int jtd,
                               const float *table);

/* transfer_source_func (gpu_funcs_3d.h:23) / transferSourceFunc
 * (base_prop.h:22).  vals is [npts][nt_big]; at most 32 points, all inside the
 * stencil interior. */
int born_transfer_source_func(born_ctx *ctx, int npts, int nt_big,
                              const int *locs, const float *vals);

/* transfer_vel_func1 / 2 (gpu_funcs_3d.h:31,33) / transferVelFunc1/2
 * (base_prop.h:27-34): v^2 dt^2 over the domain.  vel1 drives source
 * propagation and its reversal, vel2 the Born fields. */
int born_transfer_vel_func1(born_ctx *ctx, int nx, int ny, int nz,
                            const float *v2dt2);
int born_transfer_vel_func2(born_ctx *ctx, int nx, int ny, int nz,
                            const float *v2dt2);

/* transfer_receiver_func (gpu_funcs_3d.h:26) / transferReceiverFunc
 * (base_prop.h:42).  Registers the receiver cells and, like the CPU backend
 * (cpu_prop.cc:281-288), the caller's trace buffer [rec_nx*rec_ny][nt_big]:
 * born_rtm_forward writes into it and born_rtm_adjoint reads it.  Receivers
 * must lie inside the stencil interior. */
int born_transfer_receiver_func(born_ctx *ctx, int rec_nx, int rec_ny,
                                int nt_big, const int *locs, float *traces);

/* rtm_forward (gpu_funcs_3d.h:15) / rtmForward (base_prop.h:48;
 * cpu_prop.cc:22-54): Born modelling of the domain image img (n1*n2*n3),
 * accumulated into the registered trace buffer.  As in the CPU backend the
 * rec argument is not used. */
int born_rtm_forward(born_ctx *ctx, int n1, int n2, int n3, int jt,
                     const float *img, float *rec, int npts, int nt,
                     int nt_big, int rec_nx, int rec_ny);

/* source_prop (gpu_funcs_3d.h:9) / sourceProp (base_prop.h:35;
 * cpu_prop.cc:106-129): nt+1 undamped steps with vel1 from the state in
 * p0/p1; the final two levels are returned in p0/p1 exactly as the CPU backend
 * leaves them.  damp, get_last and npts are ignored, as there. */
int born_source_prop(born_ctx *ctx, int nx, int ny, int nz, int damp,
                     int get_last, float *p0, float *p1, int jts, int npts,
                     int nt);

/* rtm_adjoint (gpu_funcs_3d.h:12) / rtmAdjoint (base_prop.h:60;
 * cpu_prop.cc:56-96): reverse-time migration of the registered traces with
 * the background held in p0/p1; accumulates into img (n1*n2*n3).  p0 and p1
 * are used as scratch (modified), as in the CPU backend. */
int born_rtm_adjoint(born_ctx *ctx, int n1, int n2, int n3, int jtd, float *p0,
                     float *p1, float *img, int npts_s, int nt);

/* ---- extensions (no reference counterpart) ------------------------------ */

/* Sponge width and decay (the reference hard-codes 40 and .0005). */
int born_set_sponge(born_ctx *ctx, int width, float bcB);

/* Exact adjoint of born_rtm_forward (scatter 0) or of the linearised damped
 * wave equation in V = v^2 dt^2 (scatter 1): img (n1*n2*n3) += F^T rec, with
 * rec [nrec][nt_big].  The background is rebuilt from HBM checkpoints.
 * The adjoint state is carried as V B lambda and the image divided by V B,
 * so velocity 2 and every sponge factor must be > 0 in the stencil interior;
 * otherwise the call returns BORN_ERR_UNSUPPORTED (the device form too). */
int born_adjoint_exact(born_ctx *ctx, int n1, int n2, int n3, int jt,
                       const float *rec, float *img, int nt, int scatter);

/* Born forward with a choice of scatter term (0: the reference's 1e-6 s m;
 * 1: dV * B * L(s)); traces accumulate into rec [nrec][nt_big]. */
int born_forward_scatter(born_ctx *ctx, int n1, int n2, int n3, int jt,
                         const float *img, float *rec, int nt, int scatter);

/* Device-resident entry points: all pointers are device memory on the
 * context's device; traces are time-major [nt_big][nrec].  Work is queued on
 * the context stream (born_get_stream) and is not synchronised. */
int born_dev_forward(born_ctx *ctx, int jt, int nt, int scatter,
                     const float *model, float *traces, int keep_background);
int born_dev_adjoint_exact(born_ctx *ctx, int jt, int nt, int scatter,
                           const float *traces, float *img);
int born_dev_source_prop(born_ctx *ctx, int jts, int nt, float *p0, float *p1);
int born_dev_rtm_adjoint(born_ctx *ctx, int nt, float *p0, float *p1,
                         float *img, const float *traces);

/* born_dev_source_prop without allocating anything, for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid allocation).  No counterpart in the reference.
 * The context must be set up by the transfer calls first (they allocate);
 * this call then runs no autotuning (it uses the launch shape tuned on this
 * context so far, else the default shape: every shape gives the same
 * values) and keeps its only scratch, the per-step source amplitudes, in a
 * fixed device array of 2^20 floats (4 MiB) that each GPU holds for the life
 * of the process.  (nt + 1) x source points above 2^20 returns BORN_ERR_ARG.
 * Every call on one GPU shares that GPU's array: calls that overlap on one
 * GPU (several contexts or host threads) race; calls queued on one stream
 * do not.  Results equal born_dev_source_prop's. */
int born_dev_source_prop_noalloc(born_ctx *ctx, int jts, int nt, float *p0, float *p1);

/* HBM budget (bytes) for the background history of the exact adjoint;
 * 0 = all free device memory less a reserve.  The number of checkpoints is
 * the smallest that fits (1 segment = the whole history kept). */
int born_set_checkpoint_budget(born_ctx *ctx, size_t bytes);
/* segments and recomputed steps of the last exact adjoint.  A Born forward
 * asked to keep the background (FP32 storage) whose history does not fit
 * keeps the checkpoint plan instead: checkpoints at segment starts and the
 * last segment, so the next exact adjoint of the same background, scatter,
 * jt and nt recomputes only the earlier segments. */
int born_last_checkpoint_plan(const born_ctx *ctx, int *segments,
                              int *recomputed_steps);
/* background steps the last exact adjoint swept before its segments (nt when
 * it found neither the history nor the checkpoint plan of a forward, else
 * 0); no counterpart in the reference, which keeps no checkpoints */
int born_last_background_sweep(const born_ctx *ctx);
/* Storage of the wavefield levels and background history of the Born
 * forward and exact adjoint (born_dev_forward, born_dev_adjoint_exact and
 * their host-pointer forms).  BORN_STORAGE_F16 keeps every level as FP16
 * with a power-of-two scale per level, chosen on the device from a bound on
 * the level's magnitude; arithmetic stays FP32, and traces, model and image
 * stay FP32.  No reference counterpart: the reference stores FP32 only.
 * The reference-mode loops (born_source_prop, born_rtm_adjoint) always use
 * FP32. */
#define BORN_STORAGE_F32 0
#define BORN_STORAGE_F16 1
/* BORN_STORAGE_F24: every level as its FP32 value rounded to 24 bits (3
 * bytes per cell: sign, the 8 exponent bits and 15 significand bits, the top
 * 16 bits and the byte below them stored apart); relative precision 2^-16
 * over the whole FP32 range, no scale, so the rounding noise injected at
 * every step stays below 1e-3 of the result over thousands of steps.
 * Fields of 24-bit levels hold fewer than 715,827,882 cells. */
#define BORN_STORAGE_F24 2
int born_set_storage(born_ctx *ctx, int storage);

/* Two time steps per launch with FP32 storage, in the Born forward and in
 * the adjoint-state sweep of the exact adjoint: one sweep reads levels k-1
 * and k once and writes levels k+1 and k+2; the values are those of single
 * steps.  mode 0 (default): single steps; 1: pairs unless the autotuner
 * timed two single steps faster; 2: pairs.  No reference counterpart: the
 * reference advances one step per pass (base/cpu_prop.cc:31-53, 71-95). */
int born_set_step_pairs(born_ctx *ctx, int mode);
/* bit 0: the last Born forward advanced in two-step launches; bit 1: the
 * last exact adjoint did */
int born_last_step_pairs(const born_ctx *ctx);

/* Placement in HBM of the two levels of the scattered field (Born forward;
 * of s and r together with FP16 or 24-bit storage) and of the adjoint
 * state (exact adjoint), for histories longer than one window.  A step
 * kernel's speed depends on where the arrays it streams sit relative to
 * one another (up to 20 % at 304^3, by 32 GiB windows of history offset),
 * so such a history is laid out in windows at fixed offsets of its buffer,
 * each a gap for two levels of the pair (the adjoint state's with the image
 * accumulator) and then history levels, and the pair moves (two or three
 * copies) between the work buffers and the gaps as the history levels a
 * step touches change window.  mode 1 (default): each
 * window takes the place timed fastest for it (timed once per grid,
 * history buffer and the buffers of velocity and model; the work buffers
 * when the autotuner is off); 2: the work buffers and the gaps in turn,
 * window after window; 0: no gaps and no moves.  window_bytes: bytes per
 * window (0: 32 GiB; at least a gap and one level).  Results do not depend
 * on either.  No reference counterpart: the reference holds no history on
 * a GPU. */
int born_set_placement(born_ctx *ctx, int mode, long long window_bytes);
/* pair moves of the last Born forward and of the last exact adjoint (the
 * first move, from the work buffers, counts) */
int born_last_placement(const born_ctx *ctx, int *forward_moves, int *adjoint_moves);

/* z planes per block of the step kernels (0 = chosen by the autotuner) */
int born_set_zchunk(born_ctx *ctx, int zchunk);

/* Autotuner of the step kernels' launch shape (on by default): before the
 * first launch of a step kernel on a new grid, every tile shape (rows per
 * thread x wavefronts per block) and z chunk is timed on scratch fields and
 * the fastest kept.  on = 0 keeps the default shape; on = 1 re-tunes. */
int born_set_autotune(born_ctx *ctx, int on);
/* The chosen launch shape of a step kernel: 0 Born forward, 1 Born forward
 * (dV scatter), 2 source propagation, 3 reference adjoint, 4/5 background
 * recompute, 6 adjoint state; 7-11 the FP16-storage forms of 0, 1, 4, 5, 6;
 * 12, 13, 14 the two-step forms of 0, 1, 6 (tile 56 x rows*waves); 15-19
 * the 24-bit-storage forms of 0, 1, 4, 5, 6.
 * ms = timed duration of one launch (-1 when not tuned). */
int born_get_tuning(const born_ctx *ctx, int kernel, int *rows_per_thread,
                    int *waves_per_block, int *zchunk, float *ms);

/* The chosen tile of a step kernel: tile_x x tile_y cells of the xy plane,
 * cells_per_lane consecutive x cells per thread (2: one 8-byte / 4-byte load
 * of a pair; needs an even nx), lookahead = how many z planes ahead of the
 * plane being computed the loads are issued (1 or 2; 1 for the two-step
 * kernels). */
int born_get_tile(const born_ctx *ctx, int kernel, int *tile_x, int *tile_y,
                  int *cells_per_lane, int *lookahead);
/* Fix the launch shape of a step kernel (rows per thread x waves per block
 * x cells per lane x lookahead, one of the autotuner's shapes) and its z
 * chunk; the autotuner then leaves that kernel. */
int born_set_tuning(born_ctx *ctx, int kernel, int rows_per_thread, int waves_per_block,
                    int cells_per_lane, int lookahead, int zchunk);

void *born_get_stream(const born_ctx *ctx);

/* ---- Shot batches (no reference counterpart: the reference runs one shot
 * per process, scripts/run.py).  Many shots, each with its own windowed
 * domain, velocity, model, source and receivers, advance together: one
 * launch per time step covers every shot's tiles, and the time loop is
 * captured once in a HIP graph and replayed.  The spacing is shared; the
 * sinc tables, storage (born_set_storage) and stream come from ctx at each
 * call (a new sinc table makes the next call capture the loop again), and
 * ctx must outlive the batch.  The sponge (born_set_sponge) is copied when
 * the batch is created: a later born_set_sponge does not change the batch.
 * Every shot's nx * ny * nz < 2^30. */
typedef struct born_batch born_batch;
typedef struct {
  int nx, ny, nz;        /* the shot's domain, x fastest */
  const float *vel;      /* host [nz][ny][nx], V = v^2 dt^2 (as transfer_vel_func2) */
  const float *model;    /* host [nz][ny][nx], the Born scatter operand */
  int npts, nt_big;      /* source points (<= 32) and trace length (same for all) */
  const int *src_locs;   /* linear cell index per source point */
  const float *src_vals; /* [npts][nt_big] */
  int rec_nx, rec_ny;    /* receivers */
  const int *rec_locs;   /* linear cell index per receiver, rec_nx * rec_ny */
} born_shot;
int born_batch_create(born_ctx *ctx, int nshots, const born_shot *shots, float d1, float d2,
                      float d3, born_batch **out);
int born_batch_destroy(born_batch *b);
/* Born forward of every shot (born_rtm_forward per shot: scatter 0 is the
 * reference's term, 1 the dV term).  Queued on the context stream. */
int born_batch_forward(born_batch *b, int jt, int nt, int scatter);
/* Traces of one shot after born_batch_forward, [nrec][nt_big] host memory
 * (synchronises the stream). */
int born_batch_traces(born_batch *b, int shot, float *rec);
/* Exact adjoint of born_batch_forward for every shot of the batch (the
 * transpose of born_rtm_forward per shot, scatter 0, or of the dV scatter,
 * born_adjoint_exact per shot): image_s = F_s^T traces_s, into the batch's
 * per-shot images (born_batch_image).  The traces are those of the last
 * born_batch_forward, or those set with born_batch_set_traces.  One launch
 * per time step covers every shot; the background is rebuilt from HBM
 * checkpoints of the whole batch, the fewest segments that fit the budget
 * (born_set_checkpoint_budget on ctx; 0: free memory).  Queued on the
 * context stream.  Each shot's image equals born_adjoint_exact of that shot
 * alone.  The reference migrates one shot per process and sums the images
 * (base/rtm_zero_op_3d.cc:240-299, util/Add.cc); there is no batched
 * counterpart. */
int born_batch_adjoint_exact(born_batch *b, int jt, int nt, int scatter);
/* host traces [nrec][nt_big] of one shot as the adjoint's input */
int born_batch_set_traces(born_batch *b, int shot, const float *rec);
/* image of one shot after born_batch_adjoint_exact: host [nz][ny][nx] of the
 * shot's domain (synchronises the stream) */
int born_batch_image(born_batch *b, int shot, float *img);
/* segments and recomputed steps of the last born_batch_adjoint_exact */
int born_batch_last_checkpoint_plan(const born_batch *b, int *segments,
                                    int *recomputed_steps);
/* Tile (cells) and z chunk the autotuner chose for the batch kernel. */
int born_batch_get_tile(const born_batch *b, int scatter, int *tile_x, int *tile_y,
                        int *zchunk, float *ms);
int born_synchronize(born_ctx *ctx);

#ifdef __cplusplus
}
#endif
#endif
