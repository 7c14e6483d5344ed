// Small kernels of the Born propagator (fills, source amplitudes, the exact
// adjoint's final scaling, trace transposes, FP16 scale rings), the table of
// step-kernel tile shapes, and the dispatch of the step launchers declared in
// born_launch.h to the op instantiations of the born_k_*.hip units.
#include "born_ops.h"

namespace born {

// per shot: both rings at scale 1, max 0, and max |model| 0
__global__ void batch_ring_init_kernel(float *rings, int nshots) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nshots * kRingWords) return;
  const int w = i % kRingWords;
  rings[i] = (w < 4 || (w >= 8 && w < 12)) ? 1.f : 0.f;
}

// ring words [at, at + 4) = 1 and [at + 4, at + 8) = 0 of every shot; or,
// with maxima_only, word at = 0
__global__ void batch_ring_reset_kernel(float *rings, int nshots, int at, int maxima_only) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int per = maxima_only ? 1 : 8;
  if (i >= nshots * per) return;
  const int s = i / per, w = i % per;
  rings[s * kRingWords + at + w] = (!maxima_only && w < 4) ? 1.f : 0.f;
}

// ring slots: scale 1, max 0
__global__ void ring_init_kernel(float *scale, unsigned *maxb, int n) {
  const int i = threadIdx.x;
  if (i < n) {
    scale[i] = 1.f;
    maxb[i] = 0u;
  }
}

// *out = max(*out, max |p[i]|) as float bits (*out starts at 0)
__global__ void absmax_kernel(const float *p, long long n, unsigned *out) {
  float m = 0.f;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    m = fmaxf(m, fabsf(p[i]));
  for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_xor(m, off));
  __shared__ float red[4];
  if (threadIdx.x % 64 == 0) red[threadIdx.x / 64] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < (int)blockDim.x / 64; w++) m = fmaxf(m, red[w]);
    atomicMax(out, __float_as_uint(m));
  }
}

// ------------------------------------------------------------ small kernels

__global__ void fill_kernel(float *p, long long n, float v) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    p[i] = v;
}

// amp[s*stride + i] = scale * sinc8(tab[ii], src[i][id..id+7]) with
// tt = s + t_off, id = tt / jt, ii = tt - id*jt; zero when id + 7 >= ntsrc
// (the guard of cpu_prop.cc:167).
__global__ void src_amp_kernel(float *amp, int nsteps, int npts, const float *src,
                               int ntsrc, const float *tab, int nsinc, int jt, int ntab,
                               int t_off, float scale, int stride) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nsteps * npts) return;
  const int s = t / npts, i = t % npts;
  const int tt = s + t_off;
  const int id = tt / jt;
  const int ii = tt - id * jt;
  float a = 0.f;
  if (id + 7 < ntsrc && ii < ntab)
    a = mul_r(scale, sinc8(tab + (long long)ii * nsinc, src + (long long)i * ntsrc + id, 1));
  amp[(long long)s * stride + i] = a;
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


// the twelve tiles with one step of load lookahead, then with two
#define BORN_TILES(LA)                                                                     \
  {2, 4, 1, LA, 64}, {4, 4, 1, LA, 64}, {2, 8, 1, LA, 64}, {2, 16, 1, LA, 64},             \
      {1, 8, 1, LA, 64}, {1, 16, 1, LA, 64}, {1, 4, 2, LA, 32}, {2, 4, 2, LA, 32},         \
      {1, 8, 2, LA, 32}, {2, 8, 2, LA, 32}, {1, 12, 2, LA, 32}, {2, 2, 2, LA, 32}
const TileShape kShapes[kNumShapes] = {BORN_TILES(1), BORN_TILES(2)};
#undef BORN_TILES

// the launchers of born_launch.h: dispatch on storage and scatter to the
// instantiations of the born_k_*.hip units
#define BORN_ALL_TYPES(M, ...) M(__VA_ARGS__, float); M(__VA_ARGS__, __half); M(__VA_ARGS__, Split)
BORN_ALL_TYPES(BORN_FWD_STEP, extern template, 0);
BORN_ALL_TYPES(BORN_FWD_STEP, extern template, 1);
BORN_ALL_TYPES(BORN_SREC_STEP, extern template, 0);
BORN_ALL_TYPES(BORN_SREC_STEP, extern template, 1);
BORN_ALL_TYPES(BORN_LAM_STEP, extern template);
BORN_ALL_TYPES(BORN_BATCH_FWD_STEP, extern template, 0);
BORN_ALL_TYPES(BORN_BATCH_FWD_STEP, extern template, 1);
BORN_ALL_TYPES(BORN_BATCH_SREC_STEP, extern template, 0);
BORN_ALL_TYPES(BORN_BATCH_SREC_STEP, extern template, 1);
BORN_ALL_TYPES(BORN_BATCH_LAM_STEP, extern template);
#undef BORN_ALL_TYPES

// F<T>() for the storage of Half::on
template <class F>
static hipError_t by_storage(int on, F f) {
  if (on == 2) return f(Split{});
  if (on == 1) return f(__half{});
  return f(0.f);
}

hipError_t launch_fwd_step(const Geometry &geo, const Tune &tu, const FwdStep &a,
                           hipStream_t st) {
  return by_storage(a.h.on, [&](auto t) {
    using T = decltype(t);
    return a.scatter == 0 ? fwd_step<0, T>(geo, tu, a, st) : fwd_step<1, T>(geo, tu, a, st);
  });
}

hipError_t launch_srec_step(const Geometry &geo, const Tune &tu, const SrecStep &a,
                            hipStream_t st) {
  return by_storage(a.h.on, [&](auto t) {
    using T = decltype(t);
    return a.scatter == 0 ? srec_step<0, T>(geo, tu, a, st) : srec_step<1, T>(geo, tu, a, st);
  });
}

hipError_t launch_lam_step(const Geometry &geo, const Tune &tu, const LamStep &a,
                           hipStream_t st) {
  return by_storage(a.h.on, [&](auto t) { return lam_step<decltype(t)>(geo, tu, a, st); });
}

hipError_t launch_batch_fwd_step(const Taps &taps, const Tune &tu, const BatchFwdStep &b,
                                 hipStream_t st) {
  return by_storage(b.a.h.on, [&](auto t) {
    using T = decltype(t);
    return b.a.scatter == 0 ? batch_fwd_step<0, T>(taps, tu, b, st)
                            : batch_fwd_step<1, T>(taps, tu, b, st);
  });
}

hipError_t launch_batch_srec_step(const Taps &taps, const Tune &tu, const BatchSrecStep &b,
                                  hipStream_t st) {
  return by_storage(b.a.h.on, [&](auto t) {
    using T = decltype(t);
    return b.a.scatter == 0 ? batch_srec_step<0, T>(taps, tu, b, st)
                            : batch_srec_step<1, T>(taps, tu, b, st);
  });
}

hipError_t launch_batch_lam_step(const Taps &taps, const Tune &tu, const BatchLamStep &b,
                                 hipStream_t st) {
  return by_storage(b.a.h.on,
                    [&](auto t) { return batch_lam_step<decltype(t)>(taps, tu, b, st); });
}

// dst[2i] = a[i], dst[2i+1] = b[i]
__global__ void interleave_kernel(float2 *dst, const float *a, const float *b, long long n) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    dst[i] = make_float2(a[i], b[i]);
}

hipError_t launch_interleave(float *dst, const float *a, const float *b, long long n,
                             hipStream_t st) {
  if (n <= 0) return hipSuccess;
  long long nb = (n + 255) / 256;
  if (nb > 65536) nb = 65536;
  hipLaunchKernelGGL(interleave_kernel, dim3((unsigned)nb), dim3(256), 0, st, (float2 *)dst, a,
                     b, n);
  return hipGetLastError();
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
                          float scale, hipStream_t st, int stride) {
  const int n = nsteps * npts;
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(src_amp_kernel, dim3((n + 255) / 256), dim3(256), 0, st, amp, nsteps,
                     npts, src, ntsrc, tab, nsinc, jt, ntab, t_off, scale,
                     stride < 0 ? npts : stride);
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

hipError_t launch_batch_ring_init(float *rings, int nshots, hipStream_t st) {
  const int n = nshots * kRingWords;
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(batch_ring_init_kernel, dim3((n + 255) / 256), dim3(256), 0, st, rings,
                     nshots);
  return hipGetLastError();
}

hipError_t launch_batch_ring_reset(float *rings, int nshots, int at, int maxima_only,
                                   hipStream_t st) {
  const int n = nshots * (maxima_only ? 1 : 8);
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(batch_ring_reset_kernel, dim3((n + 255) / 256), dim3(256), 0, st, rings,
                     nshots, at, maxima_only);
  return hipGetLastError();
}

hipError_t launch_ring_init(float *scale, unsigned *maxb, int n, hipStream_t st) {
  if (n > 64) return hipErrorInvalidValue;
  hipLaunchKernelGGL(ring_init_kernel, dim3(1), dim3(64), 0, st, scale, maxb, n);
  return hipGetLastError();
}

hipError_t launch_absmax(const float *p, long long n, unsigned *out, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  long long nb = (n + 255) / 256;
  if (nb > 2048) nb = 2048;
  hipLaunchKernelGGL(absmax_kernel, dim3((unsigned)nb), dim3(256), 0, st, p, n, out);
  return hipGetLastError();
}

hipError_t launch_transpose(float *dst, const float *src, int n0, int n1, hipStream_t st) {
  if (n0 <= 0 || n1 <= 0) return hipSuccess;
  hipLaunchKernelGGL(transpose_kernel, dim3((n1 + 31) / 32, (n0 + 31) / 32), dim3(32, 8), 0,
                     st, dst, src, n0, n1);
  return hipGetLastError();
}

}  // namespace born
