// Host-side preparation of the Born operator's inputs (see born_setup.h).
#include "born_host.h"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <vector>

#include "host_par.h"

namespace {

// Levinson solution of the symmetric Toeplitz system r f = g, with the
// reference's double accumulators over float storage (sinc_bob.cc:73-100).
void levinson(int m, const float *r, float *f, const float *g, float *a) {
  double v = r[0];
  a[0] = 1.f;
  f[0] = g[0] / r[0];
  for (int j = 1; j < m; j++) {
    a[j] = f[j] = 0.f;
    double e = 0.;
    for (int i = 0; i < j; i++) e += a[i] * r[j - i];
    double c = e / v;
    v -= e * c;
    for (int i = 0; i <= j / 2; i++) {
      const double lo = a[i], hi = a[j - i];
      const double hi_new = hi - c * lo;
      a[i] = (float)(lo - c * hi);
      a[j - i] = (float)hi_new;
    }
    double w = 0.;
    for (int i = 0; i < j; i++) w += f[i] * r[j - i];
    c = (g[j] - w) / v;
    for (int i = 0; i <= j; i++) f[i] = (float)(f[i] + c * a[j - i]);
  }
}

// Least-squares tapered sinc of even length n for fractional shift d
// (Hale's mksinc, sinc_bob.cc:36-64).  Single-precision spectral sums: in
// C++ cos(float) is the float overload.
void tapered_sinc(float *coef, int n, float d) {
  const float pi = 3.141592654;
  const float two_pi = pi * 2;
  const float nyq = 0.5f;
  const float s_nat = (float)(nyq * (0.066 + 0.265 * std::log((double)n)));
  const float ds = (s_nat - 0.f) / (n * 2 - 1);
  const float eta = (float)(n / 2 - 1.0 + d);
  std::vector<float> b(n), c(n), work(n);
  for (int j = 0; j < n; j++) {
    b[j] = c[j] = 0.f;
    for (float s = 0.f; s <= s_nat; s += ds) {
      b[j] += std::cos(two_pi * s * (float)j);
      c[j] += std::cos(two_pi * s * (eta - (float)j));
    }
  }
  levinson(n, b.data(), coef, c.data(), work.data());
}

int clamp_index(int v, int hi) { return std::max(0, std::min(hi, v)); }

// squared normalised distance into a random band of width w at index i of n
// (vel_fd_3d.cc:60-99)
float band_dist2(int i, int n, int w) {
  if (i < w) return (float)((w - i) * (w - i)) / (float)(w * w);
  if (i >= n - w) {
    const int k = i - (n - w - 1);
    return (float)(k * k) / (float)(w * w);
  }
  return 0.f;
}

}  // namespace

extern "C" {

void born_host_sinc_table(int jt, int nsinc, float *table) {
  const float step = (float)(1. / (float)jt);
  float t = 0.f;
  for (int it = 0; it < jt; it++, t += step) {
    float *row = table + (long)it * nsinc;
    std::fill(row, row + nsinc, 0.f);
    tapered_sinc(row, nsinc, t);
  }
}

void born_host_calc_stability(float vmax, float dmin, float ds0, int n, float *dt,
                              int *nt) {
  const float dstab = (float)(.47 * dmin / vmax);
  const float step = ds0 / std::ceil(ds0 / dstab);
  *dt = step;
  *nt = (n - 1) * (int)(ds0 / step) + 1;
}

int born_host_padded_size(int n, int nb, int fat) {
  const int m = n + 2 * (nb + fat);
  return (m + 15) / 16 * 16;
}

void born_host_wavelet_deriv(int n, float dt, const float *w, float *out) {
  for (int it = 0; it + 1 < n; it++) out[it] = (w[it + 1] - w[it]) / dt;
  if (n > 0) out[n - 1] = 0.f;
}

void born_host_source_func(int n1, int n2, float o1, float o2, float o3, float d1,
                           float d2, float d3, float sx, float sy, float sz,
                           const float *wdev, int nw, int nts, int *locs,
                           float *vals) {
  // +.5 then truncation: nearest cell; the stamp sits one plane above it
  const float fx = (float)((sx - o1) / d1 + .5);
  const float fy = (float)((sy - o2) / d2 + .5);
  const float fz = (float)((sz - o3) / d3 + .5);
  const int ix = (int)fx, iy = (int)fy, iz = (int)fz - 1;
  std::fill(vals, vals + (long)nts * 9, 0.f);
  // the reference copies all nw samples (source_func_3d.cc:103-105) and runs
  // past the trace when the wavelet is longer than nts - 4; truncate instead
  const int nw_fit = std::max(0, std::min(nw, nts - 4));
  const float zn = fz - iz;
  int k = 0;
  for (int y = iy - 1; y <= iy + 1; y++) {
    const float yn = fy - y;
    for (int x = ix - 1; x <= ix + 1; x++, k++) {
      const float xn = fx - x;
      locs[k] = x + y * n1 + iz * n1 * n2;
      const float w = std::exp(-zn * zn - xn * xn - yn * yn);
      for (int it = 0; it < nw_fit; it++) vals[(long)nts * k + it + 4] = w * wdev[it];
    }
  }
}

float born_host_rand_subspace(int irand, int nrand1, int nrand2, int nrand3, float dt,
                              int vn1, int vn2, int vn3, float vo1, float vo2,
                              float vo3, float vd1, float vd2, float vd3,
                              const float *vel, int n1, int n2, int n3, float o1,
                              float o2, float o3, float d1, float d2, float d3,
                              float *out) {
  double vm = vel[0];
  const long nv = (long)vn1 * vn2 * vn3;
  for (long i = 1; i < nv; i++) vm = std::max(vm, (double)vel[i]);
  const float vmax = (float)vm;
  std::srand(irand);
  long i = 0;
  for (int iz = 0; iz < n3; iz++) {
    const int zv = clamp_index((int)((o3 + d3 * iz - vo3) / vd3), vn3 - 1);
    const float dz2 = band_dist2(iz, n3, nrand3);
    for (int iy = 0; iy < n2; iy++) {
      const int yv = clamp_index((int)((o2 + d2 * iy - vo2) / vd2), vn2 - 1);
      const float dy2 = band_dist2(iy, n2, nrand2);
      for (int ix = 0; ix < n1; ix++, i++) {
        const int xv = clamp_index((int)((o1 + d1 * ix - vo1) / vd1), vn1 - 1);
        const float v = vel[xv + (long)yv * vn1 + (long)zv * vn1 * vn2];
        float dx2 = band_dist2(ix, n1, nrand1);
        if (dx2 < 0.f || dx2 > 1.f) dx2 = 1.f;
        const float dist = std::sqrt(dx2 + dy2 + dz2);
        float dev = 0.f;
        // rejection sampling of the boundary perturbation
        bool accepted = dist < .0001;
        while (!accepted && nrand2 != 0) {
          const float rnd = (float)((float)std::rand() / RAND_MAX - .5 - .49 * dist);
          dev = (float)(rnd * 1.3 * dist * v);
          accepted = std::fabs(dev + v) < vmax * 1.03 * (1. - std::min(1.f, dist) * .4) &&
                     dev + v > 0.0001;
        }
        out[i] = (v + dev) * (v + dev) * dt * dt;
      }
    }
  }
  return vmax;
}

void born_host_receiver_locs(int rec_nx, int rec_ny, float ro1, float ro2, float rd1,
                             float rd2, float rec_depth, int n1, int n2, float o1,
                             float o2, float o3, float d1, float d2, float d3,
                             int *locs) {
  const int iz = (int)((rec_depth - o3) / d3);
  for (int j = 0; j < rec_ny; j++)
    for (int i = 0; i < rec_nx; i++) {
      const int ix = (int)((ro1 + rd1 * i - o1) / d1);
      const int iy = (int)((ro2 + rd2 * j - o2) / d2);
      locs[i + j * rec_nx] = ix + iy * n1 + iz * n1 * n2;
    }
}

// Host loops below run over threads in blocks of whole traces, planes or
// elements; each output element sees the serial loop's operations in the
// same order.
void born_host_map_forward(int nt, int n2, int n3, int ntb, const float *m, float *d) {
  born_par::blocks((long)n2 * n3, 64, [=](long lo, long hi) {
    for (long tr = lo; tr < hi; tr++)
      for (int it = 0; it < nt; it++) d[tr * nt + it] += m[tr * ntb + it + 4];
  });
}

void born_host_map_adjoint(int nt, int n2, int n3, int ntb, float *m, const float *d) {
  born_par::blocks((long)n2 * n3, 64, [=](long lo, long hi) {
    for (long tr = lo; tr < hi; tr++)
      for (int it = 0; it < nt; it++) m[tr * ntb + it + 4] += d[tr * nt + it];
  });
}

// The reference's time derivative adds the first-sample difference three
// times to trace 0 of each plane and twice to every other trace
// (deriv_3d.cc:40-55); kept as is since it is self-adjoint-consistent.
void born_host_deriv_forward(int n1, int n2, int n3, float d1, const float *m, float *d) {
  born_par::blocks(n3, 1, [=](long p0, long p1) {
    for (long i3 = p0; i3 < p1; i3++) {
      float *dp = d + i3 * n1 * n2;
      const float *mp = m + i3 * n1 * n2;
      dp[1] += mp[0] / d1;
      dp[0] -= mp[0] / d1;
      for (int i2 = 0; i2 < n2; i2++) {
        float *dt = dp + (long)i2 * n1;
        const float *mt = mp + (long)i2 * n1;
        dt[1] += mt[0] / d1;
        dt[0] -= mt[0] / d1;
        for (int i1 = 0; i1 + 1 < n1; i1++) {
          dt[i1] -= mt[i1] / d1;
          dt[i1 + 1] += mt[i1] / d1;
        }
      }
    }
  });
}

void born_host_deriv_adjoint(int n1, int n2, int n3, float d1, float *m, const float *d) {
  born_par::blocks(n3, 1, [=](long p0, long p1) {
    for (long i3 = p0; i3 < p1; i3++) {
      float *mp = m + i3 * n1 * n2;
      const float *dp = d + i3 * n1 * n2;
      mp[0] += (dp[1] - dp[0]) / d1;
      for (int i2 = 0; i2 < n2; i2++) {
        float *mt = mp + (long)i2 * n1;
        const float *dt = dp + (long)i2 * n1;
        mt[0] += (dt[1] - dt[0]) / d1;
        for (int i1 = 0; i1 + 1 < n1; i1++) mt[i1] += (dt[i1 + 1] - dt[i1]) / d1;
      }
    }
  });
}

// The serial form is d[i + plane] += m[i]; d[i] -= m[i] for i below
// plane (n3 - 1): element j gets + m[j - plane] first, then - m[j].
void born_host_laplac_forward(int n1, int n2, int n3, const float *m, float *d) {
  const long plane = (long)n1 * n2, last = plane * (n3 - 1);
  if (n3 < 2) return;
  born_par::blocks(plane * n3, 1 << 16, [=](long lo, long hi) {
    for (long j = lo; j < hi; j++) {
      float v = d[j];
      if (j >= plane) v += m[j - plane];
      if (j < last) v -= m[j];
      d[j] = v;
    }
  });
}

void born_host_laplac_adjoint(int n1, int n2, int n3, float *m, const float *d) {
  const long plane = (long)n1 * n2;
  if (n3 < 2) return;
  born_par::blocks(plane * (n3 - 1), 1 << 16, [=](long lo, long hi) {
    for (long i = lo; i < hi; i++) m[i] += d[i + plane] - d[i];
  });
}

static void window(int n1o, int n2o, int n3o, float o1o, float o2o, float o3o,
                   float d1o, float d2o, float d3o, int n1i, int n2i, int n3i,
                   float o1i, float o2i, float o3i, float *image, float *img,
                   bool accumulate_into_image) {
  const int f1 = (int)((o1i - o1o) / d1o), f2 = (int)((o2i - o2o) / d2o),
            f3 = (int)((o3i - o3o) / d3o);
  const int z0 = std::max(0, -f3), z1 = std::min(n3i, n3o - f3);
  const int y0 = std::max(0, -f2), y1 = std::min(n2i, n2o - f2);
  const int x0 = std::max(0, -f1), x1 = std::min(n1i, n1o - f1);
  born_par::blocks(z1 - z0, 1, [=](long lo, long hi) {
    for (long i3 = z0 + lo; i3 < z0 + hi; i3++)
      for (int i2 = y0; i2 < y1; i2++)
        for (int i1 = x0; i1 < x1; i1++) {
          float &o = image[i1 + f1 + (long)(i2 + f2) * n1o + (i3 + f3) * n1o * n2o];
          float &in = img[i1 + (long)i2 * n1i + i3 * n1i * n2i];
          if (accumulate_into_image)
            o += in;
          else
            in = o;
        }
  });
}

void born_host_extract_sub(int n1o, int n2o, int n3o, float o1o, float o2o, float o3o,
                           float d1o, float d2o, float d3o, const float *image, int n1i,
                           int n2i, int n3i, float o1i, float o2i, float o3i, float *img) {
  window(n1o, n2o, n3o, o1o, o2o, o3o, d1o, d2o, d3o, n1i, n2i, n3i, o1i, o2i, o3i,
         const_cast<float *>(image), img, false);
}

void born_host_add_image(int n1o, int n2o, int n3o, float o1o, float o2o, float o3o,
                         float d1o, float d2o, float d3o, float *image, int n1i, int n2i,
                         int n3i, float o1i, float o2i, float o3i, const float *img) {
  window(n1o, n2o, n3o, o1o, o2o, o3o, d1o, d2o, d3o, n1i, n2i, n3i, o1i, o2i, o3i, image,
         const_cast<float *>(img), true);
}

}  // extern "C"
