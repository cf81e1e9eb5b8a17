/*
 * hft_oracle.h -- CPU restatement of HipFT's flux-transport hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker, never the product: only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
 *
 * Arrays use the reference layout (src/hipft.f90:2114): Fortran column-major
 * f(ntm,npm,nr) with theta fastest and the two-point phi overlap
 * (column 1 == column npm-1, column npm == column 2).
 */
#ifndef HFT_ORACLE_H
#define HFT_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc orc;

/* Mesh: t[ntm] theta nodes on [0,pi]; p[npm1] phi nodes on [0,2pi]. */
orc *orc_create(int ntm, int npm1, const double *t, const double *p, int nr);
void orc_destroy(orc *o);
void orc_dims(const orc *o, int *out /* ntm npm nt np nr */);
/* which: 0 t 1 p 2 th 3 ph 4 dt 5 dth 6 dp 7 dph 8 st 9 sth 10 ct 11 cth
 *        12 st_i 13 sth_i 14 dt_i 15 dth_i 16 dp_i 17 dph_i */
int orc_mesh(const orc *o, int which, double *out);
/* Operator data for other restatements of the step (the differentiable
 * PyTorch transcription): WENO3 weights (which 0..6 theta CP MC M_T CP_T P_T
 * MC_T eps, 7..13 phi; 1-based arrays shifted to 0-based, length nt or np),
 * Lax-Friedrichs alphas (ntm,npm,nr), diffusion coefficients
 * coef(2:ntm-1,2:npm-1,5,nr), and the pole factors (bc_flow n/s, bc_diff
 * n/s). */
int orc_weno(const orc *o, int which, double *out);
/* OpenMP threads of the loops over points (results do not depend on it). */
void orc_set_threads(int n);
int orc_get_threads(void);
void orc_alpha(const orc *o, double *alpha_t, double *alpha_p);
void orc_coef(const orc *o, double *coef);
void orc_pole_fac(const orc *o, double *out4);

void orc_set_f(orc *o, const double *f);
void orc_get_f(const orc *o, double *f);
/* Initial-map ingest: PT map (npm1,ntm,nr) -> f, seam average, pole m=0. */
void orc_ingest_pt(orc *o, const double *map_pt, double mult_fac);

/* Flows (reference units: coefficients in m/s, omega and const in km/s). */
void orc_flow_reset(orc *o);
void orc_flow_add_mf(orc *o, const double *p1, const double *p3, const double *p5);
void orc_flow_add_dr(orc *o, const double *p0, const double *p2, const double *p4);
void orc_flow_add_rigid(orc *o, double omega_kms);
void orc_flow_add_vt_const(orc *o, double vt_kms);
void orc_flow_set(orc *o, const double *vt, const double *vp); /* (nt,npm,nr), (ntm,np,nr) */
void orc_flow_get(const orc *o, double *vt, double *vp);
void orc_flow_attenuate(orc *o, const double *att_value); /* per realization value in G */
void orc_set_lf_alpha(orc *o);
double orc_flow_dtmax(orc *o, int strang);

void orc_advection_operator(orc *o, const double *x, double *y);
void orc_advection_step(orc *o, double dt);

/* Diffusion with constant per-realization coefficient eta[nr] (km^2/s). */
void orc_load_diffusion(orc *o, const double *eta, double factor);
double orc_dt_euler(const orc *o);
double orc_dt_euler_r(const orc *o, int r);
void orc_diffusion_operator(orc *o, const double *x, double *y);
double orc_ptl(orc *o, double dt_in);
long orc_sts_rkg2(double dt, double dt_euler, double *uj, double *vj, double *ubj,
                  double *gj, long smax);
long orc_diffusion_sts(orc *o, double dt);
/* Full diffusion_step with PTL subcycles.  Per-cycle records (time_stepped,
 * dt, s, diagnostic subcycles) and, when hist != NULL, per-cycle analysis
 * rows [cycle][r][14].  Returns the number of cycles. */
int orc_diffusion_step(orc *o, double dt, int subcycles, int max_cycles,
                       double *cyc_time, double *cyc_dt, long *cyc_s, long *cyc_sc,
                       double *hist, const double *fval_pt);

/* History diagnostics (14 per realization, reference column order). */
void orc_analysis(const orc *o, const double *fval_pt, double *out);

/* Data assimilation: data (npm1,ntm,3,nr) slices data/weight/mu. */
void orc_assimilate(orc *o, const double *data, int custom_from_mu, const double *mu_power,
                    const double *mu_limit, const double *lat_limit, double mult_fac,
                    int add_mode, int balance);
void orc_set_weno_eps(orc *o, double eps);
void orc_set_pole_flux_lat_limit(orc *o, double deg);
void orc_get_flux(const orc *o, const double *map /*(ntm,npm)*/, double *fp, double *fm);
void orc_balance_flux(const orc *o, double *map);

/* Sources: random flux emergence frame (rnd theta fastest over k = 2..npm-1,
 * out (ntm,npm,nr)) and source_step f += dt * src (ntm,npm,nr). */
void orc_rfe_frame(const orc *o, const double *rnd, int model, const double *uf_per_hour,
                   double *out);
void orc_source_step(orc *o, double dtl, const double *src);

/* Spherical-harmonic validation field Y(np,nt) (Holmes et al. recursion). */
void orc_tesseral(int m, int l, const double *pvec, const double *tvec, int np, int nt,
                  double *Y);

#ifdef __cplusplus
}
#endif
#endif
