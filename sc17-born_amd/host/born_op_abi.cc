// C ABI of the Born operator layer (include/born_host.h, part 2).
#include <cstdlib>
#include <cstring>
#include <functional>
#include <exception>
#include <memory>
#include <string>

#include "born_hip.h"
#include "born_host.h"
#include "born_op.h"

using namespace born_host;

struct born_op {
  std::shared_ptr<hipProp> prop;
  std::shared_ptr<rtm_zero_op> op;
  std::string err;
};

namespace {

int guard(born_op *h, const std::function<void()> &f) {
  try {
    f();
    return 0;
  } catch (const std::exception &e) {
    if (h) h->err = e.what();
    return -1;
  }
}

std::shared_ptr<hypercube_float> wrap(const std::shared_ptr<hypercube_float> &space,
                                      const float *v) {
  auto h = std::dynamic_pointer_cast<hypercube_float>(space->clone_space());
  if (v) std::memcpy(h->vals.data(), v, h->vals.size() * sizeof(float));
  return h;
}

}  // namespace

extern "C" {

int born_op_create(born_op **out, const born_op_desc *d) {
  if (!out || !d) return -1;
  *out = nullptr;
  auto *h = new born_op();
  const int rc = guard(h, [&] {
    BornSetup s;
    for (int a = 0; a < 3; a++) s.vel_axes.emplace_back(d->vn[a], d->vo[a], d->vd[a]);
    s.vel.assign(d->vel, d->vel + (size_t)d->vn[0] * d->vn[1] * d->vn[2]);
    s.wavelet_axis = axis(d->nw, 0.f, d->wdt);
    s.wavelet.assign(d->wavelet, d->wavelet + d->nw);
    s.src_x.assign(d->src_x, d->src_x + d->nshots);
    s.src_y.assign(d->src_y, d->src_y + d->nshots);
    s.src_depth = d->src_depth;
    s.t_axis = axis(d->nt, d->t0, d->dt);
    s.rec_x = axis(d->rec_nx, d->rec_ox, d->rec_dx);
    s.rec_y = axis(d->rec_ny, d->rec_oy, d->rec_dy);
    s.rec_depth = d->rec_depth;
    s.nbound = d->nbound;
    s.mode = d->mode;
    s.scatter = d->scatter;
    h->prop = std::make_shared<hipProp>(d->device);
    if (born_set_storage(h->prop->context(), d->storage) != BORN_OK)
      throw std::runtime_error("storage must be 0 (FP32) or 1 (FP16)");
    h->op = std::make_shared<rtm_zero_op>(h->prop, s);
  });
  if (rc) {
    // keep the handle so the caller can read the message
    *out = h;
    return rc;
  }
  *out = h;
  return 0;
}

int born_op_destroy(born_op *h) {
  delete h;
  return 0;
}

const char *born_op_last_error(const born_op *h) { return h ? h->err.c_str() : "null"; }

int born_op_sizes(const born_op *h, long long *model_n, long long *data_n, int *nt, float *dt,
                  int *dom) {
  if (!h || !h->op) return -1;
  if (model_n) *model_n = h->op->model_space()->getN123();
  if (data_n) *data_n = h->op->data_space()->getN123();
  if (nt) *nt = h->op->nt();
  if (dt) *dt = h->op->dt();
  if (dom)
    for (int a = 0; a < 3; a++) dom[a] = h->op->domain_axis(a).n;
  return 0;
}

int born_op_forward(born_op *h, int add, const float *model, float *data) {
  if (!h || !h->op || !model || !data) return -1;
  return guard(h, [&] {
    auto m = wrap(h->op->model_space(), model);
    auto dv = wrap(h->op->data_space(), add ? data : nullptr);
    h->op->forward(add != 0, m, dv);
    std::memcpy(data, dv->vals.data(), dv->vals.size() * sizeof(float));
  });
}

int born_op_adjoint(born_op *h, int add, float *model, const float *data) {
  if (!h || !h->op || !model || !data) return -1;
  return guard(h, [&] {
    auto m = wrap(h->op->model_space(), add ? model : nullptr);
    auto dv = wrap(h->op->data_space(), data);
    h->op->adjoint(add != 0, m, dv);
    std::memcpy(model, m->vals.data(), m->vals.size() * sizeof(float));
  });
}

int born_op_dot_test(born_op *h, double tol, double *dots, int *pass) {
  if (!h || !h->op) return -1;
  return guard(h, [&] {
    const bool ok = h->op->dot_test(false, tol, dots);
    if (pass) *pass = ok ? 1 : 0;
  });
}

int born_op_lsrtm(born_op *h, const float *data, float *model, int niter, double *residuals) {
  if (!h || !h->op || !data || !model || niter < 0) return -1;
  return guard(h, [&] {
    auto dv = wrap(h->op->data_space(), data);
    cgstep cg;
    simple_solver solver(&cg, dv, h->op);
    solver.verb = getenv("BORN_VERBOSE") != nullptr;  // one line per iteration
    solver.solve(niter);
    auto m = std::dynamic_pointer_cast<hypercube_float>(solver.m);
    std::memcpy(model, m->vals.data(), m->vals.size() * sizeof(float));
    for (size_t i = 0; i < solver.residuals.size() && residuals; i++)
      residuals[i] = solver.residuals[i];
    for (size_t i = solver.residuals.size(); i < (size_t)niter && residuals; i++)
      residuals[i] = -1.;
  });
}

}  // extern "C"
