/*
 * born_host.h -- C ABI of the host library (sc17-born_amd/libborn_host.so).
 *
 * Part 1, preparation: the work the reference does on the CPU before handing
 * arrays to its propagator backend (base/fd_prop_3d.cc,
 * base/source_func_3d.cc, base/vel_fd_3d.cc, base/sinc_bob.cc,
 * base/data_rtm_3d.cc, base/map_data_3d.cc, base/deriv_3d.cc,
 * base/laplac_3d.cc, base/image_rtm_3d.cc).  Arrays are float32, x fastest.
 * Floating-point expressions keep the reference's order and float/double
 * promotions, so the prepared inputs are the reference's bit for bit.
 *
 * Part 2, the Born operator: the reference's rtm_zero_op forward()/adjoint()
 * (base/rtm_zero_op_3d.h:22-25) on the MI355X propagator, with its dot test
 * (base/my_operator.cc:13-62) and the conjugate-gradient solver
 * (base/lin_solver.cc:104-131, base/cgstep.cc:11-54), behind a handle.
 */
#ifndef BORN_HOST_H
#define BORN_HOST_H
#ifdef __cplusplus
extern "C" {
#endif

/* sinc_bob(jt, nsinc).table flattened [jt][nsinc] (sinc_bob.cc:4-100) */
void born_host_sinc_table(int jt, int nsinc, float *table);

/* fd_prop::calc_stability (fd_prop_3d.cc:6-33): dt and the number of
 * propagation steps for n samples of interval ds0 */
void born_host_calc_stability(float vmax, float dmin, float ds0, int n,
                              float *dt, int *nt);

/* source_func::set_compute_size (source_func_3d.cc:4-35): n + 2(nb+fat)
 * rounded up to a multiple of 16 */
int born_host_padded_size(int n, int nb, int fat);

/* wavelet_source_func ctor (source_func_3d.cc:112-127): forward difference */
void born_host_wavelet_deriv(int n, float dt, const float *w, float *out);

/* wavelet_source_func::get_source_func (source_func_3d.cc:76-109):
 * 9 locations and [9][nts] samples of the stamped wavelet derivative; a
 * wavelet longer than nts - 4 samples is truncated to the trace */
void born_host_source_func(int n1, int n2, float o1, float o2, float o3,
                           float d1, float d2, float d3, float sx, float sy,
                           float sz, const float *wdev, int nw, int nts,
                           int *locs, float *vals);

/* vel_fd_3d::rand_subspace (vel_fd_3d.cc:29-134), random band drawn
 * serially in (z, y, x) order; returns vmax */
float born_host_rand_subspace(int irand, int nrand1, int nrand2, int nrand3,
                              float dt, int vn1, int vn2, int vn3, float vo1,
                              float vo2, float vo3, float vd1, float vd2,
                              float vd3, const float *vel, int n1, int n2,
                              int n3, float o1, float o2, float o3, float d1,
                              float d2, float d3, float *out);

/* data_rtm_3d::get_source_func (data_rtm_3d.cc:59-90) + map_data ctor
 * (map_data_3d.cc:4-35): receiver cells of a regular surface grid */
void born_host_receiver_locs(int rec_nx, int rec_ny, float ro1, float ro2,
                             float rd1, float rd2, float rec_depth, int n1,
                             int n2, float o1, float o2, float o3, float d1,
                             float d2, float d3, int *locs);

/* map_data forward/adjoint (map_data_3d.cc:37-75), += */
void born_host_map_forward(int nt, int n2, int n3, int ntb, const float *m,
                           float *d);
void born_host_map_adjoint(int nt, int n2, int n3, int ntb, float *m,
                           const float *d);
/* deriv forward/adjoint (deriv_3d.cc:4-56), += */
void born_host_deriv_forward(int n1, int n2, int n3, float d1, const float *m,
                             float *d);
void born_host_deriv_adjoint(int n1, int n2, int n3, float d1, float *m,
                             const float *d);
/* laplac forward/adjoint (laplac_3d.cc:4-75), += */
void born_host_laplac_forward(int n1, int n2, int n3, const float *m, float *d);
void born_host_laplac_adjoint(int n1, int n2, int n3, float *m, const float *d);
/* image_rtm_3d::extract_sub / add_image (image_rtm_3d.cc:24-95) */
void born_host_extract_sub(int n1o, int n2o, int n3o, float o1o, float o2o,
                           float o3o, float d1o, float d2o, float d3o,
                           const float *image, int n1i, int n2i, int n3i,
                           float o1i, float o2i, float o3i, float *img);
void born_host_add_image(int n1o, int n2o, int n3o, float o1o, float o2o,
                         float o3o, float d1o, float d2o, float d3o,
                         float *image, int n1i, int n2i, int n3i, float o1i,
                         float o2i, float o3i, const float *img);

/* ---- Born operator (rtm_zero_op) ---------------------------------------- */

typedef struct born_op born_op;

/* What the reference reads from its SEP files and parameters
 * (rtm.cc:13-36, model.cc:13-46). */
typedef struct {
  int vn[3];           /* velocity model x, y, z sizes */
  float vo[3], vd[3];  /* origins and spacings */
  const float *vel;    /* [vn2][vn1][vn0] m/s */
  int nw;              /* wavelet samples */
  float wdt;           /* wavelet sampling */
  const float *wavelet;
  int nshots;
  const float *src_x, *src_y; /* shot positions (m) */
  float src_depth;
  int nt;              /* data samples */
  float t0, dt;
  int rec_nx, rec_ny;  /* receiver grid */
  float rec_ox, rec_oy, rec_dx, rec_dy, rec_depth;
  int nbound;          /* padding / random band width (fd_prop_3d.cc:52) */
  int mode;            /* adjoint: 0 reference RTM, 1 exact transpose */
  int scatter;         /* 0: 1e-6 s m (reference), 1: dV B L(s) */
  int device;
  int storage;         /* wavefield storage of the exact pair: 0 FP32, 1 FP16
                          (born_set_storage; no reference counterpart) */
} born_op_desc;

/* rtm_zero_op ctor (rtm_zero_op_3d.cc:9-44) on a HIP device */
int born_op_create(born_op **op, const born_op_desc *desc);
int born_op_destroy(born_op *op);
const char *born_op_last_error(const born_op *op);
/* model = image on the velocity grid; data = [shot][rec_y][rec_x][t];
 * nt/dt = propagation steps and interval; dom = padded domain (x, y, z) */
int born_op_sizes(const born_op *op, long long *model_n, long long *data_n, int *nt,
                  float *dt, int *dom);
/* rtm_zero_op::forward / adjoint (rtm_zero_op_3d.cc:101-299); add = 0
 * overwrites the output, add = 1 accumulates */
int born_op_forward(born_op *op, int add, const float *model, float *data);
int born_op_adjoint(born_op *op, int add, float *model, const float *data);
/* my_operator::dot_test (my_operator.cc:13-62): dots = {<Fm,d>, <m,F'd>,
 * add=true pair}; pass = mismatch within tol */
int born_op_dot_test(born_op *op, double tol, double *dots, int *pass);
/* simple_solver + cgstep (lin_solver.cc:6-15, 104-131): niter CG iterations
 * of min |F m - d| from m = 0; model out, residuals |r|^2 per iteration */
int born_op_lsrtm(born_op *op, const float *data, float *model, int niter,
                  double *residuals);

#ifdef __cplusplus
}
#endif
#endif
