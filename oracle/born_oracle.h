/*
 * CPU oracle for the Born forward / RTM adjoint path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library restates, in plain C, the
 * arithmetic of the reference's CPU backend (base/cpu_prop.cc) and of the
 * host-side preparation around it (sinc tables, source stamp, velocity
 * subspace, receiver mapping, trace/image linear operators).  Only the test
 * suite, the smoke check and the CPU-baseline leg of the benchmark may load
 * it, and only as the checker: the shipped library never calls it.
 *
 * Every function cites the reference file:line it follows.  Floating-point
 * expressions keep the reference's evaluation order and its float/double
 * promotions (for instance the double literal in the Born scatter), and the
 * library is compiled with -ffp-contract=off so no FMA changes a rounding.
 *
 * Pinning (how each function is checked independently of this text) is listed
 * in the header comment of oracle/born_oracle.c.
 */
#ifndef BORN_ORACLE_H
#define BORN_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAX_BOUND 256

/* State of the reference cpuProp (base/cpu_prop.h:36-53), with the aliases it
 * keeps to caller-owned buffers. */
typedef struct {
  int nx, ny, nz;            /* domain, x fastest (cpu_prop.cc:229-231) */
  float coeffs[13];          /* C0,CZ1,CX1,CY1,... (cpu_prop.cc:6-18, 298-326) */
  int nbound;                /* sponge width W (hard-coded 40, cpu_prop.cc:330) */
  float bound[ORC_MAX_BOUND];
  int npts_s, nt_src;        /* transferSourceFunc (cpu_prop.cc:261-267) */
  const int *locs_s;
  const float *src;          /* [npts_s][nt_src] */
  int nsinc_s, jts;          /* transferSincTableS (cpu_prop.cc:289-295) */
  const float *tab_s;        /* [jts][nsinc_s] */
  int nrec, nt_rec;          /* transferReceiverFunc (cpu_prop.cc:281-288) */
  const int *locs_r;
  float *rec;                /* [nrec][nt_rec], written by the forward */
  int nsinc_d, jtd;          /* transferSincTableD (cpu_prop.cc:254-260) */
  const float *tab_d;        /* [jtd][nsinc_d] */
  const float *vel1, *vel2;  /* transferVelFunc1/2 (cpu_prop.cc:268-279) */
  int jt;                    /* cpuProp::_jt, last set by rtmForward/sourceProp */
  int dir;                   /* cpuProp::_dir */
  int nthreads;              /* OpenMP threads for the cell loops (0 = default) */
} orc_prop;

/* ---- propagator primitives (reference semantics, float) ---- */
void orc_create_space(orc_prop *p, float d1, float d2, float d3, float bc_a,
                      float bc_b, float bc_y, int nx, int ny, int nz);
void orc_set_sponge(orc_prop *p, int width, float bcB);
void orc_prop_step(const orc_prop *p, float *p0, const float *p1,
                   const float *vel);
void orc_damp(const orc_prop *p, float *p0, float *p1);
void orc_inject_source(const orc_prop *p, int id, int ii, float *field);
void orc_inject_receivers(const orc_prop *p, int id, int ii, float *field);
void orc_data_extract(orc_prop *p, int id, int ii, const float *field);
void orc_image_add(const orc_prop *p, const float *img, float *rec_field,
                   const float *src_field);
void orc_image_condition(const orc_prop *p, const float *rec, const float *src,
                         float *img);

/* ---- propagator loops (reference semantics, float) ---- */
void orc_rtm_forward(orc_prop *p, int jt, const float *img, int nt);
void orc_source_prop(orc_prop *p, float *p0, float *p1, int jts, int nt);
void orc_rtm_adjoint(orc_prop *p, float *p0, float *p1, float *img, int nt);

/* ---- Born forward and its exact adjoint (plain transpose) ----
 * The forward is the recursion of orc_rtm_forward: a damped background s with
 * vel2 and the source, a damped scattered field r driven by
 *   scatter == 0 :  1e-6 * s_new * m        (the reference's Born term)
 *   scatter == 1 :  m * B * L(s_cur)        (m = dV, the linearisation of the
 *                                            damped wave equation in V=v^2dt^2)
 * and receiver extraction.  orc_exact_adjoint_* is its transpose, written in
 * the plain form lambda_j = A^T lambda_{j+1} - B^2 lambda_{j+2} + P^T w with
 * A^T = (L V + 2) B, keeping the whole background history in memory.
 * Suffix _d computes in double, _f in float.  rec_* is [nrec][nt_rec];
 * outputs accumulate (+=). */
void orc_born_forward_d(const orc_prop *p, int jt, int scatter,
                        const float *img, int nt, double *rec_out);
void orc_exact_adjoint_d(const orc_prop *p, int jt, int scatter,
                         const double *rec_in, int nt, double *img_out);
void orc_exact_adjoint_f(const orc_prop *p, int jt, int scatter,
                         const float *rec_in, int nt, float *img_out);

/* ---- nonlinear modelling (for the finite-difference check) ----
 * Damped constant-density acoustic modelling with velocity field vel2dt2,
 * recording every receiver exactly like orc_rtm_forward records the Born
 * field.  rec_out is [nrec][nt_rec] (+=). */
void orc_nonlinear_d(const orc_prop *p, int jt, const double *vel2dt2, int nt,
                     double *rec_out);

/* ---- host-side preparation (reference semantics) ---- */
void orc_sinc_table(int jt, int nsinc, float *table);
void orc_calc_stability(float vmax, float dmin, const float *ds, int nds,
                        int n, float *dt_out, int *nt_out);
int orc_padded_size(int n, int nb, int fat);
void orc_wavelet_deriv(int n, float dt, const float *w, float *out);
void orc_source_func(int n1, int n2, float o1, float o2, float o3, float d1,
                     float d2, float d3, float sx, float sy, float sz,
                     const float *wdev, int nw, int nts, int *locs,
                     float *vals);
float orc_rand_subspace(int irand, int nrand1, int nrand2, int nrand3, float dt,
                        int vn1, int vn2, int vn3, float vo1, float vo2,
                        float vo3, float vd1, float vd2, float vd3,
                        const float *vel, int n1, int n2, int n3, float o1,
                        float o2, float o3, float d1, float d2, float d3,
                        float *out);
void orc_receiver_locs(int rec_nx, int rec_ny, float ro1, float ro2, float rd1,
                       float rd2, float rec_depth, int n1, int n2, float o1,
                       float o2, float o3, float d1, float d2, float d3,
                       int *locs);
void orc_map_data_forward(int nt, int n2, int n3, int ntb, const float *m,
                          float *d);
void orc_map_data_adjoint(int nt, int n2, int n3, int ntb, float *m,
                          const float *d);
void orc_deriv_forward(int n1, int n2, int n3, float d1, const float *m,
                       float *d);
void orc_deriv_adjoint(int n1, int n2, int n3, float d1, float *m,
                       const float *d);
void orc_laplac_forward(int n1, int n2, int n3, const float *m, float *d);
void orc_laplac_adjoint(int n1, int n2, int n3, float *m, const float *d);
void orc_extract_sub(int n1o, int n2o, int n3o, float o1o, float o2o,
                     float o3o, float d1o, float d2o, float d3o,
                     const float *image, int n1i, int n2i, int n3i, float o1i,
                     float o2i, float o3i, float *img);
void orc_add_image(int n1o, int n2o, int n3o, float o1o, float o2o, float o3o,
                   float d1o, float d2o, float d3o, float *image, int n1i,
                   int n2i, int n3i, float o1i, float o2i, float o3i,
                   const float *img);

#ifdef __cplusplus
}
#endif
#endif
