// Host-side operator layer of the Born path: the reference's vector /
// operator / solver interfaces (base/my_vector.h, base/my_operator.h,
// base/cgstep.h, base/lin_solver.h), its propagator seam (base/base_prop.h)
// with a HIP implementation, and the Born operator rtm_zero_op
// (base/rtm_zero_op_3d.h) built on them.  Names and signatures follow the
// reference so code written against it drops in; the SEP I/O types are
// replaced by a plain axis and an in-memory float vector.
#pragma once
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

struct born_ctx;

namespace born_host {

// SEP::axis (n, o, d)
struct axis {
  int n = 1;
  float o = 0.f, d = 1.f;
  axis() = default;
  axis(int n_, float o_ = 0.f, float d_ = 1.f) : n(n_), o(o_), d(d_) {}
};

// base/my_vector.h:7-91 (the part the operators and solvers use)
class my_vector {
 public:
  virtual ~my_vector() = default;
  virtual std::shared_ptr<my_vector> clone_vec() = 0;    // copy of the values
  virtual std::shared_ptr<my_vector> clone_space() = 0;  // same space, zeros
  std::shared_ptr<my_vector> clone_zero() { return clone_space(); }
  virtual void set_val(double v) = 0;
  void zero() { set_val(0.); }
  virtual void scale(double s) = 0;
  // this = this * sc_me + other * sc_other
  virtual void scale_add(double sc_me, std::shared_ptr<my_vector> other, double sc_other) = 0;
  virtual void add(std::shared_ptr<my_vector> other) = 0;
  virtual double dot(std::shared_ptr<my_vector> other) = 0;  // FP64 accumulation
  virtual void random() = 0;                                  // U(-.5, .5)
  virtual bool check_match(std::shared_ptr<my_vector> other) = 0;
  std::string name;
};

// base/hypercube_float.h:7-64: in-memory float vector with axes
class hypercube_float : public my_vector {
 public:
  explicit hypercube_float(const std::vector<axis> &axes);
  hypercube_float(const std::vector<axis> &axes, const float *vals);
  std::shared_ptr<my_vector> clone_vec() override;
  std::shared_ptr<my_vector> clone_space() override;
  void set_val(double v) override;
  void scale(double s) override;
  void scale_add(double sc_me, std::shared_ptr<my_vector> other, double sc_other) override;
  void add(std::shared_ptr<my_vector> other) override;
  double dot(std::shared_ptr<my_vector> other) override;
  void random() override;
  bool check_match(std::shared_ptr<my_vector> other) override;
  const axis &getAxis(int i) const { return axes_.at(i - 1); }  // 1-based
  int getNdim() const { return (int)axes_.size(); }
  long long getN123() const { return (long long)vals.size(); }
  std::vector<float> vals;

 private:
  std::vector<axis> axes_;
};

// base/my_operator.h:6-48 / my_operator.cc:13-68
class my_operator {
 public:
  virtual ~my_operator() = default;
  virtual bool forward(bool add, std::shared_ptr<my_vector> model,
                       std::shared_ptr<my_vector> data, int iter = 1) = 0;
  virtual bool adjoint(bool add, std::shared_ptr<my_vector> model,
                       std::shared_ptr<my_vector> data, int iter = 1) = 0;
  void set_domain(std::shared_ptr<my_vector> dom) { domain = dom->clone_space(); }
  void set_range(std::shared_ptr<my_vector> ran) { range = ran->clone_space(); }
  // <F m1, d2> against <m1, F^T d2> with add=false and add=true; relative
  // tolerance tol (the reference uses 1e-4).  dots = {dot1, dot2, dot3, dot4}
  bool dot_test(bool verb, double tol = 1e-4, double *dots = nullptr);
  virtual void hessian(std::shared_ptr<my_vector> in, std::shared_ptr<my_vector> out);
  std::shared_ptr<my_vector> domain, range;
};

// base/step.h / base/cgstep.h:6-25 / cgstep.cc:11-54 (FP64 scalars)
class step {
 public:
  virtual ~step() = default;
  virtual void alloc_step(std::shared_ptr<my_vector> mod, std::shared_ptr<my_vector> dat) = 0;
  virtual bool step_it(int iter, std::shared_ptr<my_vector> x, std::shared_ptr<my_vector> g,
                       std::shared_ptr<my_vector> rr, std::shared_ptr<my_vector> gg,
                       double *sc) = 0;
};

class cgstep : public step {
 public:
  void alloc_step(std::shared_ptr<my_vector> mod, std::shared_ptr<my_vector> dat) override;
  bool step_it(int iter, std::shared_ptr<my_vector> x, std::shared_ptr<my_vector> g,
               std::shared_ptr<my_vector> rr, std::shared_ptr<my_vector> gg,
               double *sc) override;
  void forget_it();

 private:
  std::shared_ptr<my_vector> s, ss;
  bool forget = true;
  double small = 1e-20;
};

// base/lin_solver.h:7-47 (simple_solver: min |F m - d|)
class lin_solver {
 public:
  virtual ~lin_solver() = default;
  bool solve(int niter);
  std::shared_ptr<my_vector> return_model() { return m->clone_vec(); }
  std::shared_ptr<my_vector> g, m, gg, rr;
  std::shared_ptr<my_operator> oper;
  step *st = nullptr;
  int verb = 0;
  std::vector<double> residuals;  // |rr|^2 after each iteration (FP64)
};

class simple_solver : public lin_solver {
 public:
  simple_solver(step *s, std::shared_ptr<my_vector> data, std::shared_ptr<my_operator> op,
                std::shared_ptr<my_vector> m0 = nullptr);
};

// base/base_prop.h:4-72 without the SEP I/O handle
class baseProp {
 public:
  virtual ~baseProp() = default;
  virtual void setNtblock(int nb) { (void)nb; }
  virtual void transferSincTableD(int nsinc, int jtd, std::vector<std::vector<float>> &table) = 0;
  virtual void transferSourceFunc(int npts, int nt_big, std::vector<int> &locs, float *vals) = 0;
  virtual void transferVelFunc1(int nx, int ny, int nz, float *vloc) = 0;
  virtual void transferVelFunc2(int nx, int ny, int nz, float *vloc) = 0;
  virtual void sourceProp(int nx, int ny, int nz, bool damp, bool getLast, float *p0, float *p1,
                          int jts, int npts, int nt) = 0;
  virtual void transferReceiverFunc(int nx, int ny, int nt, std::vector<int> &locs,
                                    float *rec) = 0;
  virtual void rtmForward(int n1, int n2, int n3, int jt, float *img, float *rec, int npts,
                          int nt, int nt_big, int rec_nx, int rec_ny) = 0;
  virtual void transferSincTableS(int nsinc, int jts, std::vector<std::vector<float>> &table) = 0;
  virtual void rtmAdjoint(int n1, int n2, int n3, int jtd, float *p0, float *p1, float *img,
                          int npts_s, int nt) = 0;
  virtual void createSpace(float d1, float d2, float d3, float bc_a, float bc_b, float bc_y,
                           int nx, int ny, int nz) = 0;
  // extension: exact adjoint of rtmForward (img += F^T rec)
  virtual void bornAdjointExact(int n1, int n2, int n3, int jt, float *rec, float *img, int nt,
                                int scatter) {
    (void)n1, (void)n2, (void)n3, (void)jt, (void)rec, (void)img, (void)nt, (void)scatter;
    throw std::runtime_error("this propagator has no exact adjoint");
  }
  // extension: Born forward with the dV * B * L(s) scatter term
  virtual void bornForwardScatter(int n1, int n2, int n3, int jt, float *img, float *rec, int nt,
                                  int scatter) {
    (void)n1, (void)n2, (void)n3, (void)jt, (void)img, (void)rec, (void)nt, (void)scatter;
    throw std::runtime_error("this propagator has no scatter choice");
  }
};

// The MI355X backend: every virtual forwards to the C ABI of libborn_hip.so
// (include/born_hip.h); errors become std::runtime_error.
class hipProp : public baseProp {
 public:
  explicit hipProp(int device = 0);
  ~hipProp() override;
  void setNtblock(int nb) override;
  void transferSincTableD(int nsinc, int jtd, std::vector<std::vector<float>> &table) override;
  void transferSourceFunc(int npts, int nt_big, std::vector<int> &locs, float *vals) override;
  void transferVelFunc1(int nx, int ny, int nz, float *vloc) override;
  void transferVelFunc2(int nx, int ny, int nz, float *vloc) override;
  void sourceProp(int nx, int ny, int nz, bool damp, bool getLast, float *p0, float *p1, int jts,
                  int npts, int nt) override;
  void transferReceiverFunc(int nx, int ny, int nt, std::vector<int> &locs, float *rec) override;
  void rtmForward(int n1, int n2, int n3, int jt, float *img, float *rec, int npts, int nt,
                  int nt_big, int rec_nx, int rec_ny) override;
  void transferSincTableS(int nsinc, int jts, std::vector<std::vector<float>> &table) override;
  void rtmAdjoint(int n1, int n2, int n3, int jtd, float *p0, float *p1, float *img, int npts_s,
                  int nt) override;
  void createSpace(float d1, float d2, float d3, float bc_a, float bc_b, float bc_y, int nx,
                   int ny, int nz) override;
  void bornAdjointExact(int n1, int n2, int n3, int jt, float *rec, float *img, int nt,
                        int scatter) override;
  void bornForwardScatter(int n1, int n2, int n3, int jt, float *img, float *rec, int nt,
                          int scatter) override;
  born_ctx *context() { return ctx_; }

 private:
  void check(int rc, const char *what);
  born_ctx *ctx_ = nullptr;
};

// Born operator configuration (what the reference reads from its SEP files
// and parameters: velocity, wavelet, data axes, src_depth, rec_depth, nbound).
struct BornSetup {
  std::vector<axis> vel_axes;  // x, y, z
  std::vector<float> vel;      // [nz][ny][nx]
  axis wavelet_axis;           // nw samples at the wavelet dt
  std::vector<float> wavelet;
  std::vector<float> src_x, src_y;  // one entry per shot
  float src_depth = 0.f;
  axis t_axis;              // data time axis (nt, o, dt)
  axis rec_x, rec_y;        // receiver grid (data axes 2 and 3)
  float rec_depth = 0.f;
  int nbound = 50;          // fd_prop_3d.cc:52 default (bc_a)
  enum Mode { REFERENCE = 0, EXACT = 1 };
  int mode = REFERENCE;     // adjoint: reference RTM or exact transpose
  int scatter = 0;          // 0: 1e-6 s m (reference), 1: dV B L(s)
};

// base/rtm_zero_op_3d.h:10-42: model = image on the velocity grid
// ([nz][ny][nx]), data = traces [shot][rec_y][rec_x][t].
class rtm_zero_op : public my_operator {
 public:
  rtm_zero_op(std::shared_ptr<baseProp> prop, const BornSetup &setup);
  bool forward(bool add, std::shared_ptr<my_vector> model, std::shared_ptr<my_vector> data,
               int iter = 1) override;
  bool adjoint(bool add, std::shared_ptr<my_vector> model, std::shared_ptr<my_vector> data,
               int iter = 1) override;
  std::shared_ptr<hypercube_float> model_space() const;
  std::shared_ptr<hypercube_float> data_space() const;
  int nt() const { return nt_; }
  float dt() const { return dt_; }
  int jts() const { return jts_; }
  int jtd() const { return jtd_; }
  const axis &domain_axis(int i) const { return dom_[i]; }

 private:
  struct Shot {
    std::vector<int> src_locs, rec_locs;
    std::vector<float> src_vals, vrand;
  };
  // Source stamp, receiver cells and random-boundary velocity of one shot.
  // Velocity and geometry are fixed for the operator's life, so shots are
  // prepared once and kept while all shots' velocities fit in
  // kShotCacheBytes of host memory (else prepared again on each call).
  Shot &prepare_shot(int ishot);
  static constexpr double kShotCacheBytes = 64.0 * (1 << 30);
  std::vector<std::unique_ptr<Shot>> shots_;
  Shot uncached_;
  std::vector<float> vnone_;  // velocity without the random band: one for all shots
  std::shared_ptr<baseProp> prop_;
  BornSetup s_;
  axis dom_[3];
  float dt_ = 0.f;
  int nt_ = 0, jts_ = 1, jtd_ = 1, nt_big_ = 0;
  std::vector<float> wavelet_deriv_;
};

}  // namespace born_host
