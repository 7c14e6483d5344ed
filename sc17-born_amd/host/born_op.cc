// Host-side operator layer of the Born path (see born_op.h).
#include "born_op.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <ctime>

#include "born_hip.h"
#include "born_host.h"
#include "host_par.h"

namespace born_host {

// One progress line per shot on stderr when BORN_VERBOSE is set: a
// full-size pass takes seconds per shot, minutes per CG iteration.
static void shot_done(const char *pass, int ishot, int nshots) {
  static const bool on = getenv("BORN_VERBOSE") != nullptr;
  if (!on) return;
  static const auto t0 = std::chrono::steady_clock::now();
  const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  fprintf(stderr, "[born] %s shot %d/%d done, %.1f s since the first line\n", pass, ishot + 1,
          nshots, s);
}

// ------------------------------------------------------------ hypercube_float

static long long count(const std::vector<axis> &axes) {
  long long n = 1;
  for (const axis &a : axes) n *= a.n;
  return n;
}

hypercube_float::hypercube_float(const std::vector<axis> &axes)
    : vals((size_t)count(axes), 0.f), axes_(axes) {
  name = "hypercube_float";
}

hypercube_float::hypercube_float(const std::vector<axis> &axes, const float *v)
    : vals(v, v + count(axes)), axes_(axes) {
  name = "hypercube_float";
}

std::shared_ptr<my_vector> hypercube_float::clone_vec() {
  return std::make_shared<hypercube_float>(axes_, vals.data());
}

std::shared_ptr<my_vector> hypercube_float::clone_space() {
  return std::make_shared<hypercube_float>(axes_);
}

static hypercube_float &as_hf(const std::shared_ptr<my_vector> &v) {
  auto *h = dynamic_cast<hypercube_float *>(v.get());
  if (!h) throw std::runtime_error("vector is not a hypercube_float");
  return *h;
}

// Element-wise vector arithmetic runs over threads in contiguous blocks:
// each element gets the serial loop's operations.
template <class F>
static void each(size_t n, F &&f) {
  born_par::blocks((long)n, 1 << 16, [&](long lo, long hi) {
    for (long i = lo; i < hi; i++) f((size_t)i);
  });
}

static void fill0(std::vector<float> &v) {
  float *p = v.data();
  each(v.size(), [=](size_t i) { p[i] = 0.f; });
}

void hypercube_float::set_val(double v) {
  float *p = vals.data();
  const float f = (float)v;
  each(vals.size(), [=](size_t i) { p[i] = f; });
}

void hypercube_float::scale(double s) {
  float *p = vals.data();
  each(vals.size(), [=](size_t i) { p[i] = (float)(p[i] * s); });
}

// hypercube_float.cc:213-222
void hypercube_float::scale_add(double sc_me, std::shared_ptr<my_vector> other, double sc_o) {
  float *p = vals.data();
  const float *o = as_hf(other).vals.data();
  each(vals.size(), [=](size_t i) { p[i] = (float)(p[i] * sc_me + sc_o * o[i]); });
}

void hypercube_float::add(std::shared_ptr<my_vector> other) {
  float *p = vals.data();
  const float *o = as_hf(other).vals.data();
  each(vals.size(), [=](size_t i) { p[i] += o[i]; });
}

// hypercube_float.cc:120-130: FP64 accumulation.  Partial sums over fixed
// chunks of 2^18 elements, added in chunk order: the result does not depend
// on the number of threads.
double hypercube_float::dot(std::shared_ptr<my_vector> other) {
  const float *a = vals.data(), *b = as_hf(other).vals.data();
  const size_t n = vals.size(), chunk = size_t(1) << 18;
  std::vector<double> part((n + chunk - 1) / chunk, 0.);
  double *pp = part.data();
  born_par::blocks((long)part.size(), 1, [=](long lo, long hi) {
    for (long c = lo; c < hi; c++) {
      double r = 0.;
      const size_t e = std::min(n, (size_t)(c + 1) * chunk);
      for (size_t i = (size_t)c * chunk; i < e; i++) r += (double)a[i] * (double)b[i];
      pp[c] = r;
    }
  });
  double r = 0.;
  for (double x : part) r += x;
  return r;
}

// hypercube_float.cc:115-119: U(-.5, .5) from rand()
void hypercube_float::random() {
  for (float &x : vals) x = (float)((float)rand() / (float)RAND_MAX - .5);
}

bool hypercube_float::check_match(std::shared_ptr<my_vector> other) {
  auto *h = dynamic_cast<hypercube_float *>(other.get());
  if (!h || h->axes_.size() != axes_.size()) return false;
  for (size_t i = 0; i < axes_.size(); i++) {
    if (axes_[i].n != h->axes_[i].n) return false;
    if (std::fabs((axes_[i].o - h->axes_[i].o) / axes_[i].d) > .01) return false;
  }
  return true;
}

// ---------------------------------------------------------------- my_operator

// my_operator.cc:13-62
bool my_operator::dot_test(bool verb, double tol, double *dots) {
  auto mod1 = domain->clone_vec(), mod2 = domain->clone_vec();
  auto dat1 = range->clone_vec(), dat2 = range->clone_vec();
  mod1->random();
  dat2->random();
  forward(false, mod1, dat1);
  const double dot1 = dat1->dot(dat2);
  adjoint(false, mod2, dat2);
  const double dot2 = mod2->dot(mod1);
  if (verb) fprintf(stderr, "Dot(add=false) %g %g\n", dot1, dot2);
  forward(true, mod1, dat1);
  adjoint(true, mod2, dat2);
  const double dot3 = mod1->dot(mod2);
  const double dot4 = dat1->dot(dat2);
  if (verb) fprintf(stderr, "Dot(add=true) %g %g\n", dot4, dot3);
  if (dots) {
    dots[0] = dot1;
    dots[1] = dot2;
    dots[2] = dot3;
    dots[3] = dot4;
  }
  if (std::fabs(dot1) < 1e-12 || std::fabs(dot3) < 1e-12) return false;
  if (std::fabs((dot1 - dot2) / dot1) > tol) return false;
  if (std::fabs((dot3 - dot4) / dot3) > tol) return false;
  return true;
}

// my_operator.cc:63-68
void my_operator::hessian(std::shared_ptr<my_vector> in, std::shared_ptr<my_vector> out) {
  auto d = range->clone_space();
  forward(false, in, d);
  adjoint(false, out, d);
}

// --------------------------------------------------------------------- cgstep

void cgstep::alloc_step(std::shared_ptr<my_vector> mod, std::shared_ptr<my_vector> dat) {
  s = mod->clone_space();
  ss = dat->clone_space();
  forget_it();
}

void cgstep::forget_it() {
  forget = true;
  if (s) s->zero();
  if (ss) ss->zero();
  small = 1e-20;
}

// cgstep.cc:11-54: Claerbout's conjugate-gradient step, 2x2 system for the
// step sizes in FP64
bool cgstep::step_it(int iter, std::shared_ptr<my_vector> m, std::shared_ptr<my_vector> g,
                     std::shared_ptr<my_vector> rr, std::shared_ptr<my_vector> gg, double *vv) {
  (void)iter;
  double alfa, beta;
  if (forget) {
    beta = 0.;
    const double gg_gg = gg->dot(gg);
    if (std::fabs(gg_gg) < small) return false;
    alfa = -gg->dot(rr) / gg_gg;
  } else {
    const double gg_gg = gg->dot(gg), ss_ss = ss->dot(ss), gg_ss = gg->dot(ss);
    if (std::fabs(gg_gg) < small || std::fabs(ss_ss) < small) return false;
    const double determ = gg_gg * ss_ss * (1.0 - gg_ss / gg_gg * gg_ss / ss_ss);
    const double gg_rr = -gg->dot(rr), ss_rr = -ss->dot(rr);
    alfa = (ss_ss * gg_rr - gg_ss * ss_rr) / determ;
    beta = (-gg_ss * gg_rr + gg_gg * ss_rr) / determ;
  }
  s->scale_add(beta, g, alfa);
  ss->scale_add(beta, gg, alfa);
  m->add(s);
  rr->add(ss);
  *vv = rr->dot(rr);
  forget = false;
  return true;
}

// ----------------------------------------------------------------- lin_solver

// lin_solver.cc:6-15 and 73-98: rr = F m0 - d
simple_solver::simple_solver(step *s, std::shared_ptr<my_vector> data,
                             std::shared_ptr<my_operator> op, std::shared_ptr<my_vector> m0) {
  st = s;
  oper = op;
  rr = data->clone_vec();
  rr->scale(-1.);
  if (m0) op->forward(true, m0, rr);
  gg = rr->clone_space();
  m = op->domain->clone_space();
  g = op->domain->clone_space();
  st->alloc_step(m, rr);
}

// lin_solver.cc:104-131
bool lin_solver::solve(int niter) {
  residuals.clear();
  for (int iter = 0; iter < niter; iter++) {
    const time_t t0 = time(nullptr);
    oper->adjoint(false, g, rr, iter);
    oper->forward(false, g, gg, iter);
    double val = 0.;
    const bool valid = st->step_it(iter, m, g, rr, gg, &val);
    residuals.push_back(val);
    if (verb)
      fprintf(stderr, "Iteration %d, residual %g, wall clock time = %f\n", iter, val,
              difftime(time(nullptr), t0));
    if (!valid) return false;
  }
  return true;
}

// -------------------------------------------------------------------- hipProp

hipProp::hipProp(int device) {
  const int rc = born_create(&ctx_, device);
  if (rc != BORN_OK) throw std::runtime_error("born_create failed (no HIP device?)");
}

hipProp::~hipProp() {
  if (ctx_) born_destroy(ctx_);
}

void hipProp::check(int rc, const char *what) {
  if (rc != BORN_OK)
    throw std::runtime_error(std::string(what) + ": " + born_last_error(ctx_));
}

static std::vector<float> flatten(const std::vector<std::vector<float>> &t) {
  std::vector<float> f;
  for (const auto &row : t) f.insert(f.end(), row.begin(), row.end());
  return f;
}

void hipProp::setNtblock(int nb) { check(born_set_ntblock(ctx_, nb), "setNtblock"); }

void hipProp::transferSincTableD(int nsinc, int jtd, std::vector<std::vector<float>> &table) {
  const auto f = flatten(table);
  check(born_transfer_sinc_table_d(ctx_, nsinc, jtd, f.data()), "transferSincTableD");
}

void hipProp::transferSincTableS(int nsinc, int jts, std::vector<std::vector<float>> &table) {
  const auto f = flatten(table);
  check(born_transfer_sinc_table_s(ctx_, nsinc, jts, f.data()), "transferSincTableS");
}

void hipProp::transferSourceFunc(int npts, int nt_big, std::vector<int> &locs, float *vals) {
  check(born_transfer_source_func(ctx_, npts, nt_big, locs.data(), vals), "transferSourceFunc");
}

void hipProp::transferVelFunc1(int nx, int ny, int nz, float *v) {
  check(born_transfer_vel_func1(ctx_, nx, ny, nz, v), "transferVelFunc1");
}

void hipProp::transferVelFunc2(int nx, int ny, int nz, float *v) {
  check(born_transfer_vel_func2(ctx_, nx, ny, nz, v), "transferVelFunc2");
}

void hipProp::sourceProp(int nx, int ny, int nz, bool damp, bool getLast, float *p0, float *p1,
                         int jts, int npts, int nt) {
  check(born_source_prop(ctx_, nx, ny, nz, damp, getLast, p0, p1, jts, npts, nt), "sourceProp");
}

void hipProp::transferReceiverFunc(int nx, int ny, int nt, std::vector<int> &locs, float *rec) {
  check(born_transfer_receiver_func(ctx_, nx, ny, nt, locs.data(), rec), "transferReceiverFunc");
}

void hipProp::rtmForward(int n1, int n2, int n3, int jt, float *img, float *rec, int npts,
                         int nt, int nt_big, int rec_nx, int rec_ny) {
  check(born_rtm_forward(ctx_, n1, n2, n3, jt, img, rec, npts, nt, nt_big, rec_nx, rec_ny),
        "rtmForward");
}

void hipProp::rtmAdjoint(int n1, int n2, int n3, int jtd, float *p0, float *p1, float *img,
                         int npts_s, int nt) {
  check(born_rtm_adjoint(ctx_, n1, n2, n3, jtd, p0, p1, img, npts_s, nt), "rtmAdjoint");
}

void hipProp::createSpace(float d1, float d2, float d3, float bc_a, float bc_b, float bc_y,
                          int nx, int ny, int nz) {
  check(born_create_space(ctx_, d1, d2, d3, bc_a, bc_b, bc_y, nx, ny, nz), "createSpace");
}

void hipProp::bornAdjointExact(int n1, int n2, int n3, int jt, float *rec, float *img, int nt,
                               int scatter) {
  check(born_adjoint_exact(ctx_, n1, n2, n3, jt, rec, img, nt, scatter), "bornAdjointExact");
}

void hipProp::bornForwardScatter(int n1, int n2, int n3, int jt, float *img, float *rec, int nt,
                                 int scatter) {
  check(born_forward_scatter(ctx_, n1, n2, n3, jt, img, rec, nt, scatter), "bornForwardScatter");
}

// ---------------------------------------------------------------- rtm_zero_op

static std::vector<std::vector<float>> sinc_rows(int jt) {
  std::vector<float> t((size_t)jt * 8);
  born_host_sinc_table(jt, 8, t.data());
  std::vector<std::vector<float>> rows(jt);
  for (int i = 0; i < jt; i++) rows[i].assign(t.begin() + 8 * i, t.begin() + 8 * (i + 1));
  return rows;
}

// rtm_zero_op_3d.cc:9-44 with fd_prop_3d.cc:6-66
rtm_zero_op::rtm_zero_op(std::shared_ptr<baseProp> prop, const BornSetup &setup)
    : prop_(std::move(prop)), s_(setup) {
  if (s_.vel_axes.size() != 3 || (long long)s_.vel.size() != count(s_.vel_axes))
    throw std::runtime_error("velocity model must be 3-D and match its axes");
  if (s_.src_x.size() != s_.src_y.size() || s_.src_x.empty())
    throw std::runtime_error("one (x, y) source position per shot is required");
  // fd_prop::calc_stability with the reference's min/max sampling rules
  // (vel_fd_3d.h:43-59)
  const float dx = s_.vel_axes[0].d, dy = s_.vel_axes[1].d, dz = s_.vel_axes[2].d;
  float dmin = dx;
  if (dz < dx && dz < dy)
    dmin = dz;
  else if (dy < dx)
    dmin = dz;
  double vmx = s_.vel[0];
  for (float v : s_.vel) vmx = std::max(vmx, (double)v);
  int nt_unused = 0;
  born_host_calc_stability((float)vmx, dmin, s_.t_axis.d, s_.t_axis.n, &dt_, &nt_unused);
  jts_ = (int)(s_.wavelet_axis.d / dt_);
  jtd_ = (int)(s_.t_axis.d / dt_);
  if (jts_ < 1 || jtd_ < 1) throw std::runtime_error("data/wavelet sampling finer than dt");
  nt_ = (s_.t_axis.n - 1) * jtd_ + 1;
  nt_big_ = 8 + s_.t_axis.n;
  // source_func::set_compute_size (source_func_3d.cc:4-35) + the origin of
  // create_domain (:64-67), computed once: the reference shifts the origin
  // again on every call, which this does not repeat
  const int pad = s_.nbound + 4;
  for (int a = 0; a < 3; a++) {
    dom_[a].n = born_host_padded_size(s_.vel_axes[a].n, s_.nbound, 4);
    dom_[a].d = s_.vel_axes[a].d;
    dom_[a].o = s_.vel_axes[a].o - s_.vel_axes[a].d * pad;
  }
  auto tab_s = sinc_rows(jts_);
  prop_->transferSincTableS(8, jts_, tab_s);
  prop_->createSpace(dx, dy, dz, (float)s_.nbound, .0005f, .0005f, dom_[0].n, dom_[1].n,
                     dom_[2].n);
  prop_->setNtblock(50);
  auto tab_d = sinc_rows(jtd_);
  prop_->transferSincTableD(8, jtd_, tab_d);
  // wavelet_source_func ctor (source_func_3d.cc:112-127)
  wavelet_deriv_.resize(s_.wavelet.size());
  born_host_wavelet_deriv((int)s_.wavelet.size(), s_.wavelet_axis.d, s_.wavelet.data(),
                          wavelet_deriv_.data());
  set_domain(model_space());
  set_range(data_space());
}

std::shared_ptr<hypercube_float> rtm_zero_op::model_space() const {
  return std::make_shared<hypercube_float>(s_.vel_axes);
}

std::shared_ptr<hypercube_float> rtm_zero_op::data_space() const {
  return std::make_shared<hypercube_float>(std::vector<axis>{
      s_.t_axis, s_.rec_x, s_.rec_y, axis((int)s_.src_x.size())});
}

// Per-shot inputs (rtm_zero_op_3d.cc:139-174): source stamp, velocity
// subspaces with and without the random band, receiver cells.
rtm_zero_op::Shot &rtm_zero_op::prepare_shot(int ishot) {
  const long long n = (long long)dom_[0].n * dom_[1].n * dom_[2].n;
  const size_t nshots = s_.src_x.size();
  if (shots_.size() != nshots) shots_.resize(nshots);
  if (shots_[ishot]) return *shots_[ishot];
  const auto &va = s_.vel_axes;
  auto band = [&](int seed, int nr, float *out) {
    born_host_rand_subspace(seed, nr, nr, nr, dt_, va[0].n, va[1].n, va[2].n, va[0].o, va[1].o,
                            va[2].o, va[0].d, va[1].d, va[2].d, s_.vel.data(), dom_[0].n,
                            dom_[1].n, dom_[2].n, dom_[0].o, dom_[1].o, dom_[2].o, dom_[0].d,
                            dom_[1].d, dom_[2].d, out);
  };
  const bool keep = (double)n * sizeof(float) * nshots <= kShotCacheBytes;
  auto fresh = std::make_unique<Shot>();
  Shot &sh = keep ? *fresh : uncached_;
  sh.src_locs.resize(9);
  sh.src_vals.resize((size_t)9 * nt_big_);
  born_host_source_func(dom_[0].n, dom_[1].n, dom_[0].o, dom_[1].o, dom_[2].o, dom_[0].d,
                        dom_[1].d, dom_[2].d, s_.src_x[ishot], s_.src_y[ishot], s_.src_depth,
                        wavelet_deriv_.data(), (int)wavelet_deriv_.size(), nt_big_,
                        sh.src_locs.data(), sh.src_vals.data());
  const int seed = sh.src_locs[3] - ((int)sh.src_locs[3] / 10000) * 10000;
  sh.vrand.resize((size_t)n);
  band(seed, s_.nbound, sh.vrand.data());
  // with no band the draw never runs: the seed does not change the result
  if (vnone_.empty()) {
    vnone_.resize((size_t)n);
    band(seed, 0, vnone_.data());
  }
  sh.rec_locs.resize((size_t)s_.rec_x.n * s_.rec_y.n);
  born_host_receiver_locs(s_.rec_x.n, s_.rec_y.n, s_.rec_x.o, s_.rec_y.o, s_.rec_x.d,
                          s_.rec_y.d, s_.rec_depth, dom_[0].n, dom_[1].n, dom_[0].o, dom_[1].o,
                          dom_[2].o, dom_[0].d, dom_[1].d, dom_[2].d, sh.rec_locs.data());
  if (!keep) return uncached_;
  shots_[ishot] = std::move(fresh);
  return *shots_[ishot];
}

// rtm_zero_op_3d.cc:101-197, one shot at a time, reading the model and
// writing the data through the arguments.
bool rtm_zero_op::forward(bool add, std::shared_ptr<my_vector> mvec,
                          std::shared_ptr<my_vector> dvec, int iter) {
  (void)iter;
  auto &m = as_hf(mvec);
  auto &d = as_hf(dvec);
  if (!add) d.zero();
  const int n1 = dom_[0].n, n2 = dom_[1].n, n3 = dom_[2].n;
  const long long n = (long long)n1 * n2 * n3;
  const int rnx = s_.rec_x.n, rny = s_.rec_y.n, ntd = s_.t_axis.n;
  const auto &va = s_.vel_axes;
  std::vector<float> img((size_t)n), img_lap((size_t)n);
  std::vector<float> rec_func((size_t)rnx * rny * nt_big_), shot_deriv((size_t)rnx * rny * ntd),
      shot((size_t)rnx * rny * ntd);
  for (int ishot = 0; ishot < (int)s_.src_x.size(); ishot++) {
    Shot &sh = prepare_shot(ishot);
    prop_->transferSourceFunc(9, nt_big_, sh.src_locs, sh.src_vals.data());
    fill0(img);
    born_host_extract_sub(va[0].n, va[1].n, va[2].n, va[0].o, va[1].o, va[2].o, va[0].d,
                          va[1].d, va[2].d, m.vals.data(), n1, n2, n3, dom_[0].o, dom_[1].o,
                          dom_[2].o, img.data());
    prop_->transferVelFunc1(n1, n2, n3, sh.vrand.data());
    prop_->transferVelFunc2(n1, n2, n3, vnone_.data());
    fill0(rec_func);
    prop_->transferReceiverFunc(rnx, rny, nt_big_, sh.rec_locs, rec_func.data());
    fill0(img_lap);
    born_host_laplac_forward(n1, n2, n3, img.data(), img_lap.data());
    if (s_.scatter == 0)
      prop_->rtmForward(n1, n2, n3, jtd_, img_lap.data(), rec_func.data(), 9, nt_, nt_big_, rnx,
                        rny);
    else
      prop_->bornForwardScatter(n1, n2, n3, jtd_, img_lap.data(), rec_func.data(), nt_,
                                s_.scatter);
    fill0(shot_deriv);
    born_host_map_forward(ntd, rnx, rny, nt_big_, rec_func.data(), shot_deriv.data());
    fill0(shot);
    born_host_deriv_forward(ntd, rnx, rny, s_.t_axis.d, shot_deriv.data(), shot.data());
    float *dst = d.vals.data() + (size_t)ishot * rnx * rny * ntd;
    const float *src = shot.data();
    each(shot.size(), [=](size_t i) { dst[i] += src[i]; });
    shot_done("forward", ishot, (int)s_.src_x.size());
  }
  return true;
}

// rtm_zero_op_3d.cc:199-299: reference RTM (random-boundary background
// rebuilt by time reversal) or the exact transpose of forward().
bool rtm_zero_op::adjoint(bool add, std::shared_ptr<my_vector> mvec,
                          std::shared_ptr<my_vector> dvec, int iter) {
  (void)iter;
  auto &m = as_hf(mvec);
  auto &d = as_hf(dvec);
  if (!add) m.zero();
  const int n1 = dom_[0].n, n2 = dom_[1].n, n3 = dom_[2].n;
  const long long n = (long long)n1 * n2 * n3;
  const int rnx = s_.rec_x.n, rny = s_.rec_y.n, ntd = s_.t_axis.n;
  const auto &va = s_.vel_axes;
  std::vector<float> img((size_t)n), img_lap((size_t)n), p0((size_t)n), p1((size_t)n);
  std::vector<float> rec_func((size_t)rnx * rny * nt_big_), shot_deriv((size_t)rnx * rny * ntd);
  for (int ishot = 0; ishot < (int)s_.src_x.size(); ishot++) {
    Shot &sh = prepare_shot(ishot);
    prop_->transferSourceFunc(9, nt_big_, sh.src_locs, sh.src_vals.data());
    prop_->transferVelFunc1(n1, n2, n3, sh.vrand.data());
    prop_->transferVelFunc2(n1, n2, n3, vnone_.data());
    const float *shot = d.vals.data() + (size_t)ishot * rnx * rny * ntd;
    fill0(shot_deriv);
    born_host_deriv_adjoint(ntd, rnx, rny, s_.t_axis.d, shot_deriv.data(), shot);
    fill0(rec_func);
    born_host_map_adjoint(ntd, rnx, rny, nt_big_, rec_func.data(), shot_deriv.data());
    prop_->transferReceiverFunc(rnx, rny, nt_big_, sh.rec_locs, rec_func.data());
    fill0(img);
    if (s_.mode == BornSetup::REFERENCE) {
      fill0(p0);
      fill0(p1);
      prop_->sourceProp(n1, n2, n3, false, true, p0.data(), p1.data(), jts_, 9, nt_);
      prop_->rtmAdjoint(n1, n2, n3, jtd_, p1.data(), p0.data(), img.data(), 9, nt_);
    } else {
      prop_->bornAdjointExact(n1, n2, n3, jtd_, rec_func.data(), img.data(), nt_, s_.scatter);
    }
    fill0(img_lap);
    born_host_laplac_adjoint(n1, n2, n3, img_lap.data(), img.data());
    born_host_add_image(va[0].n, va[1].n, va[2].n, va[0].o, va[1].o, va[2].o, va[0].d, va[1].d,
                        va[2].d, m.vals.data(), n1, n2, n3, dom_[0].o, dom_[1].o, dom_[2].o,
                        img_lap.data());
    shot_done("adjoint", ishot, (int)s_.src_x.size());
  }
  return true;
}

}  // namespace born_host
