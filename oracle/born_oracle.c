/*
 * CPU oracle for the Born forward / RTM adjoint path -- TEST INFRASTRUCTURE.
 *
 * A plain-C restatement of the reference's CPU backend (base/cpu_prop.cc) and
 * of the preparation code around it.  It is loaded only by tests/, by the
 * smoke check and by the CPU-baseline leg of bench.py, always as the checker.
 * Build: oracle/Makefile (gcc -O3 -fopenmp -ffp-contract=off).
 *
 * How each part is pinned (tests/test_oracle_*.py):
 *   coefficients   closed form: the taps equal the textbook 8th-order
 *                  second-derivative weights [-205/72, 8/5, -1/5, 8/315,
 *                  -1/560]/d^2, and the stencil differentiates polynomials of
 *                  degree <= 9 exactly (with the reference's axis permutation:
 *                  fast-axis taps use d2, middle d3, slow d1).
 *   prop           polynomial fields (exact second derivatives), the test.cu
 *                  input recipe against a brute-force numpy stencil, symmetry
 *                  of the interior operator, shell never written.
 *   damp           hand-derived special cells (edge 0 at the near face, edge 1
 *                  at the far face, untouched centre) and cell counts per edge.
 *   inject/extract brute force; jt=1 reduces to a single tap.
 *   sinc table     bit-exact against the reference sinc_bob.cc compiled from
 *                  its own source into oracle/_ref (jt = 1..8).
 *   loops          impulse responses derived by hand from the reference's
 *                  step lists (tests/test_oracle_kat.py: rtmForward data,
 *                  sourceProp levels and rtmAdjoint image for nt <= 3, one
 *                  source and receiver, exact one-tap tables); linearity in
 *                  the model, time reversibility of the undamped background,
 *                  and the exact adjoint's dot test against the forward (FP64).
 *   rand_subspace  nrand = 0 against the closed form (v^2 dt^2 of the nearest
 *                  lower model cell); the band against the reference's
 *                  per-cell acceptance envelope and the range of its draw.
 *                  The drawn values are parity unpinned: the reference calls
 *                  rand() inside a parallel loop, so its band is not
 *                  reproducible; the serial draw order is restated.
 * The CG solver (cgstep / lin_solver) is restated in numpy in oracle/cg.py.
 */
#include "born_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define C0 0
#define CZ1 1
#define CX1 2
#define CY1 3
#define CZ2 4
#define CX2 5
#define CY2 6
#define CZ3 7
#define CX3 8
#define CY3 9
#define CZ4 10
#define CX4 11
#define CY4 12

static int nthr(const orc_prop *p) { return p->nthreads > 0 ? p->nthreads : 1; }

/* cpu_prop.cc:296-332 (createSpace).  The bc arguments are ignored there: the
 * sponge is hard-coded to 40 cells with bcB = .0005 (cpu_prop.cc:329-332). */
#define C_C00(d) (8.0 / (5.0 * (d) * (d)))
void orc_create_space(orc_prop *p, float d1, float d2, float d3, float bc_a,
                      float bc_b, float bc_y, int nx, int ny, int nz) {
  (void)bc_a;
  (void)bc_b;
  (void)bc_y;
  p->nx = nx;
  p->ny = ny;
  p->nz = nz;
  p->coeffs[0] = -1025.0 / 576.0 * (C_C00(d1) + C_C00(d2) + C_C00(d3));
  p->coeffs[1] = C_C00(d1);
  p->coeffs[2] = C_C00(d2);
  p->coeffs[3] = C_C00(d3);
  p->coeffs[4] = -C_C00(d1) / 8.0;
  p->coeffs[5] = -C_C00(d2) / 8.0;
  p->coeffs[6] = -C_C00(d3) / 8.0;
  p->coeffs[7] = C_C00(d1) / 63.0;
  p->coeffs[8] = C_C00(d2) / 63.0;
  p->coeffs[9] = C_C00(d3) / 63.0;
  p->coeffs[10] = -C_C00(d1) / 896.0;
  p->coeffs[11] = -C_C00(d2) / 896.0;
  p->coeffs[12] = -C_C00(d3) / 896.0;
  orc_set_sponge(p, 40, .0005f);
}

/* Sponge table bound[i] = expf(-bcB*(bcA - i)) (cpu_prop.cc:329-332), with the
 * width as a parameter (the reference fixes it at 40). */
void orc_set_sponge(orc_prop *p, int width, float bcB) {
  float bcA = (float)width;
  if (width > ORC_MAX_BOUND) width = ORC_MAX_BOUND;
  p->nbound = width;
  for (int i = 0; i < width; i++) p->bound[i] = expf(-bcB * (bcA - i));
}

/* cpu_prop.cc:224-253 (prop): interior update, shell never written. */
void orc_prop_step(const orc_prop *p, float *p0, const float *p1,
                   const float *vel) {
  const int nx = p->nx, ny = p->ny, nz = p->nz;
  const long n12 = (long)nx * ny;
  const float *c = p->coeffs;
#pragma omp parallel for num_threads(nthr(p)) schedule(static)
  for (int i3 = 4; i3 < nz - 4; i3++) {
    for (int i2 = 4; i2 < ny - 4; i2++) {
      long ii = (long)i2 * nx + 4 + n12 * i3;
      for (int i1 = 4; i1 < nx - 4; i1++, ii++) {
        p0[ii] = vel[ii] * (c[C0] * p1[ii] + c[CX1] * (p1[ii - 1] + p1[ii + 1]) +
                            +c[CX2] * (p1[ii - 2] + p1[ii + 2]) +
                            +c[CX3] * (p1[ii - 3] + p1[ii + 3]) +
                            +c[CX4] * (p1[ii - 4] + p1[ii + 4]) +
                            +c[CY1] * (p1[ii - nx] + p1[ii + nx]) +
                            +c[CY2] * (p1[ii - 2 * nx] + p1[ii + 2 * nx]) +
                            +c[CY3] * (p1[ii - 3 * nx] + p1[ii + 3 * nx]) +
                            +c[CY4] * (p1[ii - 4 * nx] + p1[ii + 4 * nx]) +
                            +c[CZ1] * (p1[ii - 1 * n12] + p1[ii + 1 * n12]) +
                            +c[CZ2] * (p1[ii - 2 * n12] + p1[ii + 2 * n12]) +
                            +c[CZ3] * (p1[ii - 3 * n12] + p1[ii + 3 * n12]) +
                            +c[CZ4] * (p1[ii - 4 * n12] + p1[ii + 4 * n12])) +
                 p1[ii] + p1[ii] - p0[ii];
      }
    }
  }
}

static inline int imin(int a, int b) { return a < b ? a : b; }

/* cpu_prop.cc:141-165 (damp): both levels, edge = distance to the interior
 * face, far side offset by one (n-4-i). */
void orc_damp(const orc_prop *p, float *p0, float *p1) {
  const int nx = p->nx, ny = p->ny, nz = p->nz;
  const long n12 = (long)nx * ny;
#pragma omp parallel for num_threads(nthr(p)) schedule(static)
  for (int i3 = 4; i3 < nz - 4; i3++) {
    int edge1 = imin(i3 - 4, nz - 4 - i3);
    for (int i2 = 4; i2 < ny - 4; i2++) {
      int edge2 = imin(edge1, imin(i2 - 4, ny - 4 - i2));
      long ii = (long)i2 * nx + 4 + n12 * i3;
      for (int i1 = 4; i1 < nx - 4; i1++, ii++) {
        int edge = imin(edge2, imin(i1 - 4, nx - 4 - i1));
        if (edge >= 0 && edge < p->nbound) {
          p0[ii] *= p->bound[edge];
          p1[ii] *= p->bound[edge];
        }
      }
    }
  }
}

/* cpu_prop.cc:166-179 (injectSource): 8-tap sinc, scale dir/(float)jt. */
void orc_inject_source(const orc_prop *p, int id, int ii, float *field) {
  if (id + 7 >= p->nt_src) return;
  const float *t = p->tab_s + (long)ii * p->nsinc_s;
  for (int i = 0; i < p->npts_s; i++) {
    const float *v = p->src + (long)p->nt_src * i + id;
    field[p->locs_s[i]] +=
        p->dir / (float)p->jt *
        (t[0] * v[0] + t[1] * v[1] + t[2] * v[2] + t[3] * v[3] + t[4] * v[4] +
         t[5] * v[5] + t[6] * v[6] + t[7] * v[7]);
  }
}

/* cpu_prop.cc:180-197 (injectReceivers).  Run serially, which equals the
 * reference's parallel loop whenever no two receivers share a cell. */
void orc_inject_receivers(const orc_prop *p, int id, int ii, float *field) {
  if (id + 7 >= p->nt_rec) return;
  float sc = (float)p->dir / (float)p->jt;
  const float *t = p->tab_d + (long)ii * p->nsinc_d;
  for (int i = 0; i < p->nrec; i++) {
    const float *v = p->rec + (long)p->nt_rec * i + id;
    field[p->locs_r[i]] +=
        sc * (t[0] * v[0] + t[1] * v[1] + t[2] * v[2] + t[3] * v[3] +
              t[4] * v[4] + t[5] * v[5] + t[6] * v[6] + t[7] * v[7]);
  }
}

/* cpu_prop.cc:198-215 (dataExtract): no 1/jt. */
void orc_data_extract(orc_prop *p, int id, int ii, const float *field) {
  if (id + 7 >= p->nt_rec) return;
  const float *t = p->tab_d + (long)ii * p->nsinc_d;
  for (int i = 0; i < p->nrec; i++) {
    float *r = p->rec + (long)p->nt_rec * i + id;
    float v = field[p->locs_r[i]];
    for (int k = 0; k < 8; k++) r[k] += v * t[k];
  }
}

/* cpu_prop.cc:216-223 (imageAdd): the literal .000001 is a double, so the
 * product and the sum are evaluated in double and rounded once. */
void orc_image_add(const orc_prop *p, const float *img, float *rec_field,
                   const float *src_field) {
  const long n = (long)p->nx * p->ny * p->nz;
#pragma omp parallel for num_threads(nthr(p)) schedule(static)
  for (long i = 0; i < n; i++)
    rec_field[i] += .000001 * src_field[i] * img[i];
}

/* cpu_prop.cc:97-105 (imageCondition). */
void orc_image_condition(const orc_prop *p, const float *rec, const float *src,
                         float *img) {
  const long n = (long)p->nx * p->ny * p->nz;
#pragma omp parallel for num_threads(nthr(p)) schedule(static)
  for (long i = 0; i < n; i++) img[i] += src[i] * rec[i];
}

static void swapp(float **a, float **b) {
  float *t = *a;
  *a = *b;
  *b = t;
}

/* cpu_prop.cc:22-54 (rtmForward): writes the traces into p->rec (+=). */
void orc_rtm_forward(orc_prop *p, int jt, const float *img, int nt) {
  const long n = (long)p->nx * p->ny * p->nz;
  p->jt = jt;
  float *buf = (float *)calloc((size_t)(4 * n), sizeof(float));
  float *s_p0 = buf, *s_p1 = buf + n, *r_p0 = buf + 2 * n, *r_p1 = buf + 3 * n;
  p->dir = 1;
  for (int it = 0; it < nt; it++) {
    int id = it / jt;
    int ii = it - id * jt;
    orc_prop_step(p, s_p0, s_p1, p->vel2);
    orc_prop_step(p, r_p0, r_p1, p->vel2);
    orc_damp(p, s_p0, s_p1);
    orc_damp(p, r_p0, r_p1);
    orc_inject_source(p, id, ii, s_p0);
    orc_image_add(p, img, r_p0, s_p0);
    orc_data_extract(p, id, ii, r_p0);
    swapp(&s_p0, &s_p1);
    swapp(&r_p0, &r_p1);
  }
  free(buf);
}

/* cpu_prop.cc:106-129 (sourceProp): nt+1 undamped steps with vel1; the final
 * two levels stay in the caller's arrays; odd nt copies one over the other
 * (cpu_prop.cc:123-128). */
void orc_source_prop(orc_prop *p, float *p0, float *p1, int jts, int nt) {
  const long n = (long)p->nx * p->ny * p->nz;
  float *a = p0, *b = p1;
  p->jt = jts;
  p->dir = 1;
  for (int it = 0; it <= nt; it++) {
    int id = it / p->jts;
    int ii = it - id * p->jts;
    orc_prop_step(p, a, b, p->vel1);
    orc_inject_source(p, id, ii, a);
    swapp(&a, &b);
  }
  if (nt % 2 == 1) {
    /* after the loop a is the caller's p0 */
    memcpy(p1, p0, sizeof(float) * (size_t)n);
  }
  (void)a;
  (void)b;
}

/* cpu_prop.cc:56-96 (rtmAdjoint): reverse loop; accumulates into img. */
void orc_rtm_adjoint(orc_prop *p, float *src_p0, float *src_p1, float *img,
                     int nt) {
  const long n = (long)p->nx * p->ny * p->nz;
  float *buf = (float *)calloc((size_t)(2 * n), sizeof(float));
  float *r_p0 = buf, *r_p1 = buf + n;
  p->dir = -1;
  for (int it = nt - 1; it >= 0; it--) {
    int id_s = (it + 1) / p->jts;
    int i_s = it + 1 - id_s * p->jts;
    int id = it / p->jtd;
    int ii = it - id * p->jtd;
    if (it > 0) orc_prop_step(p, r_p0, r_p1, p->vel2);
    if (it < nt - 1) {
      orc_prop_step(p, src_p0, src_p1, p->vel1);
      orc_inject_source(p, id_s, i_s, src_p1);
    }
    orc_damp(p, r_p0, r_p1);
    orc_inject_receivers(p, id, ii, r_p0);
    orc_image_condition(p, r_p0, src_p0, img);
    swapp(&src_p0, &src_p1);
    swapp(&r_p0, &r_p1);
  }
  free(buf);
}

/* ------------------------------------------------------------------------ */
/* Born forward / exact adjoint / nonlinear modelling, in a generic real type */

#define R double
#define SFX(name) name##_d
#include "born_oracle_exact.inc"
#undef R
#undef SFX

#define R float
#define SFX(name) name##_f
#define ORC_NO_FORWARD
#include "born_oracle_exact.inc"
#undef ORC_NO_FORWARD
#undef R
#undef SFX

/* ------------------------------------------------------------------------ */
/* Host-side preparation                                                      */

/* sinc_bob.cc:36-64 (mksinc) and :73-100 (toep), float storage with the
 * reference's double internals. */
static void orc_toep(int m, float *r, float *f, float *g, float *a) {
  int i, j, jh;
  double c, e, v, w, bot;
  a[0] = 1.;
  v = r[0];
  f[0] = g[0] / r[0];
  for (j = 1; j < m; j++) {
    e = a[j] = f[j] = 0.;
    for (i = 0; i < j; i++) e += a[i] * r[j - i];
    c = e / v;
    v -= e * c;
    jh = j / 2;
    for (i = 0; i <= jh; i++) {
      bot = a[j - i] - c * a[i];
      a[i] -= c * a[j - i];
      a[j - i] = bot;
    }
    w = 0;
    for (i = 0; i < j; i++) w += f[i] * r[j - i];
    c = (g[j] - w) / v;
    for (i = 0; i <= j; i++) f[i] += c * a[j - i];
  }
}

static void orc_mksinc(float *sinc, int lsinc, float d, float *space) {
  int j;
  float *b, *c, *work, pi, pi2, snyq, snat, s0, ds, eta, s;
  pi = 3.141592654;
  pi2 = pi * 2;
  snyq = 0.5;
  snat = snyq * (0.066 + 0.265 * log((double)lsinc));
  s0 = 0.0;
  ds = (snat - s0) / (lsinc * 2 - 1);
  eta = lsinc / 2 - 1.0 + d;
  b = space;
  c = b + lsinc;
  work = c + lsinc;
  for (j = 0; j < lsinc; j++) {
    for (s = s0, b[j] = c[j] = 0.0; s <= snat; s += ds) {
      /* C++ resolves cos(float) to the float overload (sinc_bob.cc:57-58) */
      b[j] += cosf(pi2 * s * j);
      c[j] += cosf(pi2 * s * (eta - j));
    }
  }
  orc_toep(lsinc, b, sinc, c, work);
}

/* sinc_bob.cc:4-16: table[it][nsinc] for fractions t = it/jt. */
void orc_sinc_table(int jt, int nsinc, float *table) {
  float dtab = 1. / (float)(jt);
  float t = 0.;
  float *junk = (float *)calloc((size_t)nsinc * 3, sizeof(float));
  for (int it = 0; it < jt; it++, t += dtab) {
    for (int k = 0; k < nsinc; k++) table[it * nsinc + k] = 0.f;
    orc_mksinc(table + it * nsinc, nsinc, t, junk);
  }
  free(junk);
}

/* fd_prop_3d.cc:6-33 (calc_stability); dmin is the value get_min_samp returns
 * (vel_fd_3d.h:43-51). */
void orc_calc_stability(float vmax, float dmin, const float *ds, int nds, int n,
                        float *dt_out, int *nt_out) {
  (void)nds;
  float d = .47 * dmin / vmax;
  float dt = ds[0] / ceilf(ds[0] / d);
  *dt_out = dt;
  *nt_out = (n - 1) * (int)(ds[0] / dt) + 1;
}

/* source_func_3d.cc:4-35 (set_compute_size): pad by nb+fat on both sides,
 * round up to a multiple of 16. */
int orc_padded_size(int n, int nb, int fat) {
  int m = n + 2 * (nb + fat);
  int rem = m - ((int)(m / 16)) * 16;
  if (rem != 0) rem = 16 - rem;
  return m + rem;
}

/* source_func_3d.cc:112-127: forward difference of the wavelet, last = 0. */
void orc_wavelet_deriv(int n, float dt, const float *w, float *out) {
  out[n - 1] = 0;
  for (int it = 0; it < n - 1; it++) out[it] = (w[it + 1] - w[it]) / dt;
}

/* source_func_3d.cc:76-109 (wavelet_source_func::get_source_func): 3x3 stamp
 * in the plane iz-1, samples at offset +4.  vals is [9][nts]. */
void orc_source_func(int n1, int n2, float o1, float o2, float o3, float d1,
                     float d2, float d3, float sx, float sy, float sz,
                     const float *wdev, int nw, int nts, int *locs,
                     float *vals) {
  float fz = (sz - o3) / d3 + .5;
  float fy = (sy - o2) / d2 + .5;
  float fx = (sx - o1) / d1 + .5;
  int iy = (int)fy;
  int ix = (int)fx;
  int iz = (int)fz;
  int i = 0;
  for (int it = 0; it < nts * 9; it++) vals[it] = 0;
  for (int i3 = iz - 1; i3 <= iz - 1; i3++) {
    float zn = fz - i3;
    for (int i2 = iy - 1; i2 <= iy + 1; i2++) {
      float yn = fy - i2;
      for (int i1 = ix - 1; i1 <= ix + 1; i1++, i++) {
        float xn = fx - i1;
        locs[i] = i1 + i2 * n1 + i3 * n1 * n2;
        float scale = expf(-zn * zn - xn * xn - yn * yn);
        /* truncated to the trace (the reference writes past it when
           nw > nts - 4, source_func_3d.cc:103-105) */
        for (int it = 0; it < nw && it + 4 < nts; it++)
          vals[nts * i + it + 4] = scale * wdev[it];
      }
    }
  }
}

static inline int iclamp(int v, int lo, int hi) {
  return v < lo ? lo : (v > hi ? hi : v);
}

/* vel_fd_3d.cc:29-134 (rand_subspace), run serially in (iz, iy, ix) order.
 * Returns vmax; out = (v+dev)^2 dt^2 over the domain. */
float orc_rand_subspace(int irand, int nrand1, int nrand2, int nrand3, float dt,
                        int vn1, int vn2, int vn3, float vo1, float vo2,
                        float vo3, float vd1, float vd2, float vd3,
                        const float *vel, int n1, int n2, int n3, float o1,
                        float o2, float o3, float d1, float d2, float d3,
                        float *out) {
  double vm = vel[0];
  for (long i = 1; i < (long)vn1 * vn2 * vn3; i++)
    if (vel[i] > vm) vm = vel[i];
  float vmax = (float)vm;
  srand(irand);
  for (int iz = 0; iz < n3; iz++) {
    float dy2, dx2, dz2, rnd;
    long i = (long)iz * n2 * n1;
    int iz_v = iclamp((int)((o3 + d3 * iz - vo3) / vd3), 0, vn3 - 1);
    if (iz < nrand3)
      dz2 = (float)((nrand3 - iz) * (nrand3 - iz)) / (float)(nrand3 * nrand3);
    else if (iz >= n3 - nrand3)
      dz2 = (float)((iz - (n3 - nrand3 - 1)) * (iz - (n3 - nrand3 - 1))) /
            (float)(nrand3 * nrand3);
    else
      dz2 = 0;
    for (int iy = 0; iy < n2; iy++) {
      int iy_v = iclamp((int)((o2 + d2 * iy - vo2) / vd2), 0, vn2 - 1);
      if (iy < nrand2)
        dy2 = (float)((nrand2 - iy) * (nrand2 - iy)) / (float)(nrand2 * nrand2);
      else if (iy >= n2 - nrand2)
        dy2 = (float)((iy - (n2 - nrand2 - 1)) * (iy - (n2 - nrand2 - 1))) /
              (float)(nrand2 * nrand2);
      else
        dy2 = 0;
      for (int ix = 0; ix < n1; ix++, i++) {
        int ix_v = iclamp((int)((o1 + d1 * ix - vo1) / vd1), 0, vn1 - 1);
        float val = vel[ix_v + (long)iy_v * vn1 + (long)iz_v * vn1 * vn2];
        if (ix < nrand1)
          dx2 = (float)((nrand1 - ix) * (nrand1 - ix)) / (float)(nrand1 * nrand1);
        else if (ix >= n1 - nrand1)
          dx2 = (float)((ix - (n1 - nrand1 - 1)) * (ix - (n1 - nrand1 - 1))) /
                (float)(nrand1 * nrand1);
        else
          dx2 = 0;
        if (dx2 < 0. || dx2 > 1) dx2 = 1;
        float dist = sqrtf(dx2 + dy2 + dz2);
        int found = 0;
        float dev;
        if (dist < .0001) found = 1;
        dev = 0.;
        while (!found && nrand2 != 0) {
          rnd = (float)rand() / RAND_MAX - .5 - .49 * dist;
          dev = rnd * 1.3 * dist * val;
          float mind = dist < 1.f ? dist : 1.f;
          if (fabsf(dev + val) < vmax * 1.03 * (1. - mind * .4) &&
              dev + val > 0.0001)
            found = 1;
        }
        out[i] = (val + dev) * (val + dev) * dt * dt;
      }
    }
  }
  return vmax;
}

/* data_rtm_3d.cc:59-90 (fractional receiver coordinates) then
 * map_data_3d.cc:17-31 (floor to a grid index). */
void orc_receiver_locs(int rec_nx, int rec_ny, float ro1, float ro2, float rd1,
                       float rd2, float rec_depth, int n1, int n2, float o1,
                       float o2, float o3, float d1, float d2, float d3,
                       int *locs) {
  int i = 0;
  for (int i3 = 0; i3 < rec_ny; i3++) {
    for (int i2 = 0; i2 < rec_nx; i2++, i++) {
      float s_z = (rec_depth - o3) / d3;
      float s_x = (ro1 + rd1 * i2 - o1) / d1;
      float s_y = (ro2 + rd2 * i3 - o2) / d2;
      int ix = (int)s_x;
      int iz = (int)s_z;
      int iy = (int)s_y;
      locs[i] = ix + iy * n1 + iz * n1 * n2;
    }
  }
}

/* map_data_3d.cc:57-75: d[t, i2, i3] += m[i2, i3][t+4]. */
void orc_map_data_forward(int nt, int n2, int n3, int ntb, const float *m,
                          float *d) {
  for (int i3 = 0; i3 < n3; i3++)
    for (int i2 = 0; i2 < n2; i2++)
      for (int it = 0; it < nt; it++)
        d[it + (long)i2 * nt + (long)i3 * nt * n2] +=
            m[(long)ntb * i2 + it + 4 + (long)i3 * ntb * n2];
}

/* map_data_3d.cc:37-55. */
void orc_map_data_adjoint(int nt, int n2, int n3, int ntb, float *m,
                          const float *d) {
  for (int i3 = 0; i3 < n3; i3++)
    for (int i2 = 0; i2 < n2; i2++)
      for (int it = 0; it < nt; it++)
        m[(long)ntb * i2 + it + 4 + (long)i3 * ntb * n2] +=
            d[it + (long)i2 * nt + (long)i3 * nt * n2];
}

/* deriv_3d.cc:31-56 (forward), including the duplicated first-sample terms
 * (once per i3 plane, once per trace). */
void orc_deriv_forward(int n1, int n2, int n3, float d1, const float *m,
                       float *d) {
  const long n12 = (long)n1 * n2;
  for (int i3 = 0; i3 < n3; i3++) {
    d[1 + i3 * n12] += m[i3 * n12] / d1;
    d[i3 * n12] -= m[i3 * n12] / d1;
    for (int i2 = 0; i2 < n2; i2++) {
      d[1 + i2 * n1 + i3 * n12] += m[i2 * n1 + i3 * n12] / d1;
      d[i2 * n1 + i3 * n12] -= m[i2 * n1 + i3 * n12] / d1;
      for (int i1 = 0; i1 < n1 - 1; i1++) {
        d[i1 + i2 * n1 + i3 * n12] -= m[i3 * n12 + i2 * n1 + i1] / d1;
        d[i1 + i2 * n1 + i3 * n12 + 1] += m[i3 * n12 + i2 * n1 + i1] / d1;
      }
    }
  }
}

/* deriv_3d.cc:4-29 (adjoint). */
void orc_deriv_adjoint(int n1, int n2, int n3, float d1, float *m,
                       const float *d) {
  const long n12 = (long)n1 * n2;
  for (int i3 = 0; i3 < n3; i3++) {
    m[i3 * n12] += (d[i3 * n12 + 1] - d[i3 * n12]) / d1;
    for (int i2 = 0; i2 < n2; i2++) {
      m[i2 * n1 + i3 * n12] +=
          (d[i2 * n1 + i3 * n12 + 1] - d[i2 * n1 + i3 * n12]) / d1;
      for (int i1 = 0; i1 < n1 - 1; i1++)
        m[i3 * n12 + i2 * n1 + i1] +=
            (d[i1 + i2 * n1 + i3 * n12 + 1] - d[i1 + i2 * n1 + i3 * n12]) / d1;
    }
  }
}

/* laplac_3d.cc:51-75 (forward): vertical first difference of the image. */
void orc_laplac_forward(int n1, int n2, int n3, const float *m, float *d) {
  const long n12 = (long)n1 * n2;
  for (int i3 = 0; i3 < n3 - 1; i3++)
    for (int i2 = 0; i2 < n2; i2++)
      for (int i1 = 0; i1 < n1; i1++) {
        long i = i3 * n12 + (long)i2 * n1 + i1;
        d[i + n12] += m[i];
        d[i] -= m[i];
      }
}

/* laplac_3d.cc:4-24 (adjoint). */
void orc_laplac_adjoint(int n1, int n2, int n3, float *m, const float *d) {
  const long n12 = (long)n1 * n2;
  for (int i3 = 0; i3 < n3 - 1; i3++)
    for (int i2 = 0; i2 < n2; i2++)
      for (int i1 = 0; i1 < n1; i1++) {
        long i = i3 * n12 + (long)i2 * n1 + i1;
        m[i] += d[i + n12] - d[i];
      }
}

/* image_rtm_3d.cc:53-95 (extract_sub): integer window from the model-size
 * image into the domain; img must be zeroed by the caller. */
void orc_extract_sub(int n1o, int n2o, int n3o, float o1o, float o2o,
                     float o3o, float d1o, float d2o, float d3o,
                     const float *image, int n1i, int n2i, int n3i, float o1i,
                     float o2i, float o3i, float *img) {
  int f3 = (int)((o3i - o3o) / d3o);
  int f2 = (int)((o2i - o2o) / d2o);
  int f1 = (int)((o1i - o1o) / d1o);
  for (int i3 = 0; i3 < n3i; i3++) {
    if (i3 + f3 < 0 || i3 + f3 >= n3o) continue;
    for (int i2 = 0; i2 < n2i; i2++) {
      if (i2 + f2 < 0 || i2 + f2 >= n2o) continue;
      for (int i1 = 0; i1 < n1i; i1++) {
        if (i1 + f1 < 0 || i1 + f1 >= n1o) continue;
        img[i1 + (long)i2 * n1i + (long)i3 * n1i * n2i] =
            image[i1 + f1 + (long)(i2 + f2) * n1o + (long)(i3 + f3) * n1o * n2o];
      }
    }
  }
}

/* image_rtm_3d.cc:24-51 (add_image). */
void orc_add_image(int n1o, int n2o, int n3o, float o1o, float o2o, float o3o,
                   float d1o, float d2o, float d3o, float *image, int n1i,
                   int n2i, int n3i, float o1i, float o2i, float o3i,
                   const float *img) {
  int f3 = (int)((o3i - o3o) / d3o);
  int f2 = (int)((o2i - o2o) / d2o);
  int f1 = (int)((o1i - o1o) / d1o);
  for (int i3 = 0; i3 < n3i; i3++) {
    if (i3 + f3 < 0 || i3 + f3 >= n3o) continue;
    for (int i2 = 0; i2 < n2i; i2++) {
      if (i2 + f2 < 0 || i2 + f2 >= n2o) continue;
      for (int i1 = 0; i1 < n1i; i1++) {
        if (i1 + f1 < 0 || i1 + f1 >= n1o) continue;
        image[i1 + f1 + (long)(i2 + f2) * n1o + (long)(i3 + f3) * n1o * n2o] +=
            img[i1 + (long)i2 * n1i + (long)i3 * n1i * n2i];
      }
    }
  }
}
