/*
 * hft_oracle.c -- CPU restatement of HipFT's flux-transport hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP library:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it, and never as the thing measured or shipped.
 *
 * Each routine restates one routine of the reference (src/hipft.f90, cited by
 * line) in plain C, on the reference's own array layout: Fortran column-major
 * f(ntm,npm,nr), theta fastest, two-point phi overlap.  Expressions keep the
 * reference's operation order (left-to-right Fortran evaluation, no FMA
 * contraction: build with -ffp-contract=off) so that results are reproducible
 * to the last bit.
 *
 * The one deliberate choice the reference leaves open: the polar closure sums
 * of `do concurrent ... reduce(+)` (src/hipft.f90:7905-7908, 7972-7979) have
 * no defined order.  The oracle fixes it (pole_sum64 below: 64 strided
 * partial sums, then a halving tree), which is also the order the HIP kernels
 * use, so the two agree bit for bit.
 */
#include "hft_oracle.h"

#include <math.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include <stdlib.h>
#include <string.h>

/* ---- constants: src/hipft.f90:83-136 ---------------------------------- */
static const double zero = 0.0, one = 1.0, two = 2.0, three = 3.0, four = 4.0;
static const double six = 6.0, ten = 10.0, fifteen = 15.0;
static const double half = 0.5, quarter = 0.25;
static const double twentyfour = 24.0, twentyfive = 25.0, three_quarter = 0.75;
static const double sixth = 0.16666666666666666;
static const double two_third = 0.66666666666666666;
static const double third = 0.33333333333333333;
static const double pi = 3.1415926535897932;
static const double pi_two = 1.5707963267948966;
static const double pi_i = 0.3183098861837907;
static const double twopi = 6.2831853071795864;
static const double twopi_i = 0.15915494309189535;
static const double d2r = 0.017453292519943295;
static const double rsun_cm2 = 4.84416e21;
static const double km_s_to_rs_hr = 0.005172413793103448;
static const double m_s_to_rs_hr = 5.172413793103448e-06;
static const double safety = 0.95;

struct orc {
  int ntm, npm, nt, np, nr, npm1;
  /* mesh (1-based storage: element [0] unused) src/hipft.f90:242-259 */
  double *t, *p, *th, *ph, *dt, *dth, *dp, *dph;
  double *dt_i, *dth_i, *dp_i, *dph_i, *st, *sth, *ct, *cth, *st_i, *sth_i;
  double *f;               /* (ntm,npm,nr) */
  double *vt, *vp;         /* (nt,npm,nr), (ntm,np,nr) */
  double *vt0, *vp0;       /* unattenuated copies for flow_attenuate */
  double *alpha_t, *alpha_p; /* (ntm,npm,nr) */
  /* WENO grid weights src/hipft.f90:432-457 (1-based) */
  double *D_C_CPt, *D_C_MCt, *D_M_Tt, *D_CP_Tt, *D_P_Tt, *D_MC_Tt, *eps_t;
  double *D_C_CPp, *D_C_MCp, *D_M_Tp, *D_CP_Tp, *D_P_Tp, *D_MC_Tp, *eps_p;
  double bc_flow_npole_fac, bc_flow_spole_fac;
  /* diffusion */
  double *dcoef;           /* diffusion_coef (nt,np,nr) */
  double *coef;            /* coef(2:ntm-1,2:npm-1,5,nr) */
  double bc_diff_npole_fac, bc_diff_spole_fac;
  double dt_euler;
  double *dt_euler_r;
  int have_diffusion;
  double pole_flux_lat_limit; /* degrees, src/hipft.f90:566 (default 30) */
};

#define F3(a, n1, n2, j, k, i) (a)[((j) - 1) + (size_t)(n1) * (((k) - 1) + (size_t)(n2) * ((i) - 1))]
#define FF(o, a, j, k, i) F3(a, (o)->ntm, (o)->npm, j, k, i)
#define VT(o, j, k, i) F3((o)->vt, (o)->nt, (o)->npm, j, k, i)
#define VP(o, j, k, i) F3((o)->vp, (o)->ntm, (o)->np, j, k, i)
#define DC(o, j, k, i) F3((o)->dcoef, (o)->nt, (o)->np, j, k, i)
#define COEF(o, j, k, d, i)                                                            \
  (o)->coef[((j) - 2) + (size_t)((o)->ntm - 2) *                                      \
                            (((k) - 2) + (size_t)((o)->npm - 2) * (((d) - 1) + 5 * (size_t)((i) - 1)))]

static double *alloc1(int n) { return (double *)calloc((size_t)n + 1, sizeof(double)); }

/* Halving-tree polar sum over v[0..n-1] (see file header). */
static double pole_sum64(const double *v, int n) {
  double part[64];
  for (int l = 0; l < 64; l++) {
    double s = 0.0;
    for (int c = l; c < n; c += 64) s = s + v[c];
    part[l] = s;
  }
  for (int h = 32; h >= 1; h >>= 1)
    for (int l = 0; l < h; l++) part[l] = part[l] + part[l + h];
  return part[0];
}

static inline double dmax(double a, double b) { return a > b ? a : b; }
static inline double dmin(double a, double b) { return a < b ? a : b; }

/* x**n for integer n >= 0 by binary powering (square-and-multiply, the
 * result seeded with the first set bit's square), the expansion compilers use
 * for an integer exponent. */
static double ipow(double x, int n) {
  double res = 1.0, sq = x;
  int have = 0;
  while (n) {
    if (n & 1) {
      res = have ? res * sq : sq;
      have = 1;
    }
    n >>= 1;
    if (n) sq = sq * sq;
  }
  return res;
}

/* ---- set_mesh: src/hipft.f90:6495-6653 -------------------------------- */
orc *orc_create(int ntm, int npm1, const double *t, const double *p, int nr) {
  orc *o = (orc *)calloc(1, sizeof(orc));
  o->ntm = ntm;
  o->npm1 = npm1;
  o->nt = ntm + 1;
  o->np = npm1 + 1;
  o->npm = o->np;
  o->nr = nr;
  int nt = o->nt, np = o->np, npm = o->npm;
  o->t = alloc1(ntm);
  o->p = alloc1(np);
  for (int i = 1; i <= ntm; i++) o->t[i] = t[i - 1];
  for (int i = 1; i <= npm1; i++) o->p[i] = p[i - 1];
  o->p[npm1 + 1] = o->p[2] + twopi; /* src/hipft.f90:2132 */
  o->dt = alloc1(ntm); o->dt_i = alloc1(ntm); o->st = alloc1(ntm); o->st_i = alloc1(ntm);
  o->ct = alloc1(ntm); o->dp = alloc1(npm); o->dp_i = alloc1(npm);
  o->th = alloc1(nt); o->ph = alloc1(np); o->dth = alloc1(nt); o->dth_i = alloc1(nt);
  o->sth = alloc1(nt); o->sth_i = alloc1(nt); o->cth = alloc1(nt);
  o->dph = alloc1(np); o->dph_i = alloc1(np);
  int ntm1 = ntm;
  for (int i = 2; i <= ntm1; i++) {
    o->th[i] = half * (o->t[i] + o->t[i - 1]);
    o->dth[i] = o->t[i] - o->t[i - 1];
  }
  o->th[1] = o->th[2] - o->dth[2];
  o->th[nt] = o->th[ntm1] + o->dth[ntm1];
  o->dth[1] = o->dth[2];
  o->dth[nt] = o->dth[ntm1];
  for (int i = 1; i <= nt; i++) o->dth_i[i] = one / o->dth[i];
  for (int i = 1; i <= ntm; i++) {
    o->dt[i] = o->th[i + 1] - o->th[i];
    o->dt_i[i] = one / o->dt[i];
  }
  for (int i = 1; i <= ntm; i++) o->st[i] = sin(o->t[i]);
  o->st[1] = 0.0;
  o->st[ntm] = 0.0;
  for (int i = 1; i <= ntm; i++) o->ct[i] = cos(o->t[i]);
  o->ct[1] = one;
  o->ct[ntm] = one;
  for (int i = 2; i <= ntm - 1; i++) o->st_i[i] = one / o->st[i];
  o->st_i[1] = 0.0;
  o->st_i[ntm] = 0.0;
  for (int i = 1; i <= nt; i++) o->sth[i] = sin(fabs(o->th[i]));
  for (int i = 2; i <= ntm1; i++) o->sth_i[i] = one / o->sth[i];
  o->sth_i[1] = 0.0;
  o->sth_i[nt] = 0.0;
  for (int i = 1; i <= nt; i++) o->cth[i] = cos(o->th[i]);
  for (int i = 2; i <= np; i++) {
    o->ph[i] = half * (o->p[i] + o->p[i - 1]);
    o->dph[i] = o->p[i] - o->p[i - 1];
  }
  o->ph[1] = o->ph[npm1] - twopi;
  o->dph[1] = o->dph[npm1];
  for (int i = 1; i <= np; i++) o->dph_i[i] = one / o->dph[i];
  for (int i = 1; i <= npm - 1; i++) o->dp[i] = o->ph[i + 1] - o->ph[i];
  o->dp[npm] = o->dp[2];
  for (int i = 1; i <= npm; i++) o->dp_i[i] = one / o->dp[i];

  size_t nf = (size_t)ntm * npm * nr;
  o->f = (double *)calloc(nf, sizeof(double));
  o->vt = (double *)calloc((size_t)nt * npm * nr, sizeof(double));
  o->vp = (double *)calloc((size_t)ntm * np * nr, sizeof(double));
  o->vt0 = (double *)calloc((size_t)nt * npm * nr, sizeof(double));
  o->vp0 = (double *)calloc((size_t)ntm * np * nr, sizeof(double));
  o->alpha_t = (double *)calloc(nf, sizeof(double));
  o->alpha_p = (double *)calloc(nf, sizeof(double));

  /* load_weno: src/hipft.f90:6846-6938 (weno_eps = 0 -> cell widths) */
  o->D_C_CPt = alloc1(nt); o->D_C_MCt = alloc1(nt); o->D_M_Tt = alloc1(nt);
  o->D_CP_Tt = alloc1(nt); o->D_P_Tt = alloc1(nt); o->D_MC_Tt = alloc1(nt);
  o->eps_t = alloc1(ntm);
  o->D_C_CPp = alloc1(np); o->D_C_MCp = alloc1(np); o->D_M_Tp = alloc1(np);
  o->D_CP_Tp = alloc1(np); o->D_P_Tp = alloc1(np); o->D_MC_Tp = alloc1(np);
  o->eps_p = alloc1(npm);
  for (int i = 1; i <= ntm; i++) o->eps_t[i] = o->dt[i];
  for (int i = 1; i <= npm; i++) o->eps_p[i] = o->dp[i];
  o->pole_flux_lat_limit = 30.0;
  const double *d = o->dt;
  for (int j = 2; j <= nt - 2; j++) {
    double tot = d[j - 1] + d[j] + d[j + 1];
    o->D_C_CPt[j] = d[j] / (d[j] + d[j + 1]);
    o->D_C_MCt[j] = d[j] / (d[j - 1] + d[j]);
    o->D_M_Tt[j] = d[j - 1] / tot;
    o->D_CP_Tt[j] = (d[j] + d[j + 1]) / tot;
    o->D_P_Tt[j] = d[j + 1] / tot;
    o->D_MC_Tt[j] = (d[j - 1] + d[j]) / tot;
  }
  d = o->dp;
  for (int k = 2; k <= np - 1; k++) {
    double tot = d[k - 1] + d[k] + d[k + 1];
    o->D_C_CPp[k] = d[k] / (d[k] + d[k + 1]);
    o->D_C_MCp[k] = d[k] / (d[k - 1] + d[k]);
    o->D_M_Tp[k] = d[k - 1] / tot;
    o->D_CP_Tp[k] = (d[k] + d[k + 1]) / tot;
    o->D_P_Tp[k] = d[k + 1] / tot;
    o->D_MC_Tp[k] = (d[k - 1] + d[k]) / tot;
  }
  double *pw[6] = {o->D_C_CPp, o->D_C_MCp, o->D_M_Tp, o->D_CP_Tp, o->D_P_Tp, o->D_MC_Tp};
  for (int q = 0; q < 6; q++) { /* set_periodic_bc_1d src/hipft.f90:4841-4866 */
    pw[q][1] = pw[q][np - 1];
    pw[q][np] = pw[q][2];
  }
  o->bc_flow_npole_fac = two * o->dth_i[2] * pi_i; /* src/hipft.f90:2568-2569 */
  o->bc_flow_spole_fac = two * o->dth_i[nt - 1] * pi_i;
  return o;
}

void orc_destroy(orc *o) {
  if (!o) return;
  double **ptrs[] = {&o->t, &o->p, &o->th, &o->ph, &o->dt, &o->dth, &o->dp, &o->dph,
                     &o->dt_i, &o->dth_i, &o->dp_i, &o->dph_i, &o->st, &o->sth, &o->ct,
                     &o->cth, &o->st_i, &o->sth_i, &o->f, &o->vt, &o->vp, &o->vt0, &o->vp0,
                     &o->alpha_t, &o->alpha_p, &o->D_C_CPt, &o->D_C_MCt, &o->D_M_Tt,
                     &o->D_CP_Tt, &o->D_P_Tt, &o->D_MC_Tt, &o->eps_t, &o->D_C_CPp,
                     &o->D_C_MCp, &o->D_M_Tp, &o->D_CP_Tp, &o->D_P_Tp, &o->D_MC_Tp,
                     &o->eps_p, &o->dcoef, &o->coef, &o->dt_euler_r};
  for (size_t i = 0; i < sizeof(ptrs) / sizeof(ptrs[0]); i++) free(*ptrs[i]);
  free(o);
}

void orc_dims(const orc *o, int *out) {
  out[0] = o->ntm; out[1] = o->npm; out[2] = o->nt; out[3] = o->np; out[4] = o->nr;
}

int orc_mesh(const orc *o, int which, double *out) {
  const double *src[] = {o->t, o->p, o->th, o->ph, o->dt, o->dth, o->dp, o->dph, o->st,
                         o->sth, o->ct, o->cth, o->st_i, o->sth_i, o->dt_i, o->dth_i,
                         o->dp_i, o->dph_i};
  int len[] = {o->ntm, o->np, o->nt, o->np, o->ntm, o->nt, o->npm, o->np, o->ntm,
               o->nt, o->ntm, o->nt, o->ntm, o->nt, o->ntm, o->nt, o->npm, o->np};
  if (which < 0 || which > 17) return -1;
  memcpy(out, src[which] + 1, sizeof(double) * (size_t)len[which]);
  return len[which];
}

void orc_set_threads(int n) {
#ifdef _OPENMP
  omp_set_num_threads(n > 0 ? n : 1);
#else
  (void)n;
#endif
}

int orc_get_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

int orc_weno(const orc *o, int which, double *out) {
  const double *src[] = {o->D_C_CPt, o->D_C_MCt, o->D_M_Tt, o->D_CP_Tt, o->D_P_Tt, o->D_MC_Tt,
                         o->eps_t, o->D_C_CPp, o->D_C_MCp, o->D_M_Tp, o->D_CP_Tp, o->D_P_Tp,
                         o->D_MC_Tp, o->eps_p};
  if (which < 0 || which > 13) return -1;
  int n = which < 7 ? o->nt : o->np;
  memcpy(out, src[which] + 1, sizeof(double) * (size_t)n);
  return n;
}

void orc_alpha(const orc *o, double *alpha_t, double *alpha_p) {
  size_t n = (size_t)o->ntm * o->npm * o->nr;
  memcpy(alpha_t, o->alpha_t, sizeof(double) * n);
  memcpy(alpha_p, o->alpha_p, sizeof(double) * n);
}

void orc_coef(const orc *o, double *coef) {
  memcpy(coef, o->coef, sizeof(double) * (size_t)(o->ntm - 2) * (o->npm - 2) * 5 * o->nr);
}

void orc_pole_fac(const orc *o, double *out4) {
  out4[0] = o->bc_flow_npole_fac;
  out4[1] = o->bc_flow_spole_fac;
  out4[2] = o->bc_diff_npole_fac;
  out4[3] = o->bc_diff_spole_fac;
}

void orc_set_f(orc *o, const double *f) {
  memcpy(o->f, f, sizeof(double) * (size_t)o->ntm * o->npm * o->nr);
}
void orc_get_f(const orc *o, double *f) {
  memcpy(f, o->f, sizeof(double) * (size_t)o->ntm * o->npm * o->nr);
}

/* set_periodic_bc_3d: src/hipft.f90:4897-4925 */
static void periodic3(double *a, int n1, int n2, int n3) {
  for (int i = 1; i <= n3; i++)
    for (int j = 1; j <= n1; j++) {
      F3(a, n1, n2, j, 1, i) = F3(a, n1, n2, j, n2 - 1, i);
      F3(a, n1, n2, j, n2, i) = F3(a, n1, n2, j, 2, i);
    }
}

/* ---- IC ingest: load_initial_condition src/hipft.f90:2116-2173 --------- */
void orc_ingest_pt(orc *o, const double *map, double mult_fac) {
  int ntm = o->ntm, npm1 = o->npm1, npm = o->npm;
  for (int i = 1; i <= o->nr; i++) {
    for (int k = 1; k <= npm1; k++)
      for (int j = 1; j <= ntm; j++)
        FF(o, o->f, j, k, i) = F3(map, npm1, ntm, k, j, i); /* transpose 2118-2120 */
    for (int j = 1; j <= ntm; j++) {                       /* 2124-2126 */
      FF(o, o->f, j, 1, i) = half * (FF(o, o->f, j, 1, i) + FF(o, o->f, j, npm1, i));
      FF(o, o->f, j, npm1, i) = FF(o, o->f, j, 1, i);
      FF(o, o->f, j, npm1 + 1, i) = FF(o, o->f, j, 2, i);
    }
    double fn1 = 0.0, fs1 = 0.0; /* 2140-2157 */
    for (int k = 2; k <= npm - 1; k++) {
      fn1 = fn1 + FF(o, o->f, 1, k, i) * o->dp[k];
      fs1 = fs1 + FF(o, o->f, ntm, k, i) * o->dp[k];
    }
    for (int k = 1; k <= npm; k++) {
      FF(o, o->f, 1, k, i) = fn1 * twopi_i;
      FF(o, o->f, ntm, k, i) = fs1 * twopi_i;
    }
    for (int k = 1; k <= npm; k++) /* 2171-2173 */
      for (int j = 1; j <= ntm; j++) FF(o, o->f, j, k, i) = mult_fac * FF(o, o->f, j, k, i);
  }
}

/* ---- flows: update_flow src/hipft.f90:3411-3448, 7219-7285 ------------- */
void orc_flow_reset(orc *o) {
  memset(o->vt, 0, sizeof(double) * (size_t)o->nt * o->npm * o->nr);
  memset(o->vp, 0, sizeof(double) * (size_t)o->ntm * o->np * o->nr);
}

void orc_flow_add_mf(orc *o, const double *p1, const double *p3, const double *p5) {
  for (int i = 1; i <= o->nr; i++)
    for (int k = 1; k <= o->npm; k++)
      for (int j = 1; j <= o->nt; j++) {
        double c = o->cth[j];
        VT(o, j, k, i) = VT(o, j, k, i) - m_s_to_rs_hr * o->sth[j] *
                                              (p1[i - 1] * c + p3[i - 1] * ipow(c, 3) +
                                               p5[i - 1] * ipow(c, 5));
      }
}

void orc_flow_add_dr(orc *o, const double *p0, const double *p2, const double *p4) {
  for (int i = 1; i <= o->nr; i++)
    for (int k = 1; k <= o->np; k++)
      for (int j = 1; j <= o->ntm; j++) {
        double c = o->ct[j];
        VP(o, j, k, i) = VP(o, j, k, i) + m_s_to_rs_hr * o->st[j] *
                                              (p0[i - 1] + p2[i - 1] * ipow(c, 2) +
                                               p4[i - 1] * ipow(c, 4));
      }
}

void orc_flow_add_rigid(orc *o, double omega) {
  for (int i = 1; i <= o->nr; i++)
    for (int k = 1; k <= o->np; k++)
      for (int j = 1; j <= o->ntm; j++)
        VP(o, j, k, i) = VP(o, j, k, i) + omega * km_s_to_rs_hr * o->st[j];
}

void orc_flow_add_vt_const(orc *o, double vtc) {
  for (int i = 1; i <= o->nr; i++)
    for (int k = 1; k <= o->npm; k++)
      for (int j = 1; j <= o->nt; j++) VT(o, j, k, i) = VT(o, j, k, i) + vtc * km_s_to_rs_hr;
}

void orc_flow_set(orc *o, const double *vt, const double *vp) {
  memcpy(o->vt, vt, sizeof(double) * (size_t)o->nt * o->npm * o->nr);
  memcpy(o->vp, vp, sizeof(double) * (size_t)o->ntm * o->np * o->nr);
}

void orc_flow_get(const orc *o, double *vt, double *vp) {
  memcpy(vt, o->vt, sizeof(double) * (size_t)o->nt * o->npm * o->nr);
  memcpy(vp, o->vp, sizeof(double) * (size_t)o->ntm * o->np * o->nr);
}

/* Flow attenuation by |Br|: src/hipft.f90:3489-3507.  att_value[r] is the
 * realization's flow_attenuate_value (G); <=0 means HUGE inverse (2266-2271). */
void orc_flow_attenuate(orc *o, const double *att_value) {
  int nt = o->nt, np = o->np, npm = o->npm, ntm = o->ntm;
  for (int i = 1; i <= o->nr; i++) {
    double inv = att_value[i - 1] > zero ? one / att_value[i - 1] : 1.7976931348623157e308;
    for (int k = 2; k <= npm - 1; k++)
      for (int j = 2; j <= nt - 1; j++) {
        double a = fabs(FF(o, o->f, j - 1, k, i)), b = fabs(FF(o, o->f, j, k, i));
        double m = a > b ? a : b;
        VT(o, j, k, i) = VT(o, j, k, i) * (one - tanh(m * inv));
      }
    for (int k = 2; k <= np - 1; k++)
      for (int j = 1; j <= ntm; j++) {
        double a = fabs(FF(o, o->f, j, k - 1, i)), b = fabs(FF(o, o->f, j, k, i));
        double m = a > b ? a : b;
        VP(o, j, k, i) = VP(o, j, k, i) * (one - tanh(m * inv));
      }
  }
  periodic3(o->vt, nt, npm, o->nr);
  periodic3(o->vp, ntm, np, o->nr);
}

/* ---- set_lf_alpha: src/hipft.f90:6940-7048 ---------------------------- */
void orc_set_lf_alpha(orc *o) {
  int nt = o->nt, np = o->np, npm = o->npm, ntm = o->ntm;
  for (int i = 1; i <= o->nr; i++) {
    for (int k = 1; k <= npm; k++) {
      for (int j = 3; j <= nt - 3; j++) {
        double m = fabs(VT(o, j - 2, k, i));
        for (int q = -1; q <= 3; q++) m = dmax(m, fabs(VT(o, j + q, k, i)));
        FF(o, o->alpha_t, j, k, i) = m;
      }
      int j = 1;
      FF(o, o->alpha_t, j, k, i) = dmax(dmax(fabs(VT(o, j, k, i)), fabs(VT(o, j + 1, k, i))),
                                        dmax(fabs(VT(o, j + 2, k, i)), fabs(VT(o, j + 3, k, i))));
      j = 2;
      {
        double m = fabs(VT(o, j - 1, k, i));
        for (int q = 0; q <= 3; q++) m = dmax(m, fabs(VT(o, j + q, k, i)));
        FF(o, o->alpha_t, j, k, i) = m;
      }
      j = nt - 2;
      {
        double m = fabs(VT(o, j - 2, k, i));
        for (int q = -1; q <= 2; q++) m = dmax(m, fabs(VT(o, j + q, k, i)));
        FF(o, o->alpha_t, j, k, i) = m;
      }
      j = nt - 1;
      {
        double m = fabs(VT(o, j - 2, k, i));
        for (int q = -1; q <= 1; q++) m = dmax(m, fabs(VT(o, j + q, k, i)));
        FF(o, o->alpha_t, j, k, i) = m;
      }
    }
    for (int j = 2; j <= ntm - 1; j++) {
      for (int k = 3; k <= np - 3; k++) {
        double m = fabs(VP(o, j, k - 2, i));
        for (int q = -1; q <= 3; q++) m = dmax(m, fabs(VP(o, j, k + q, i)));
        FF(o, o->alpha_p, j, k, i) = m;
      }
      int k = 2;
      {
        double m = fabs(VP(o, j, np - 2, i));
        for (int q = -1; q <= 3; q++) m = dmax(m, fabs(VP(o, j, k + q, i)));
        FF(o, o->alpha_p, j, k, i) = m;
      }
      k = np - 2;
      {
        double m = fabs(VP(o, j, k - 2, i));
        for (int q = -1; q <= 2; q++) m = dmax(m, fabs(VP(o, j, k + q, i)));
        m = dmax(m, fabs(VP(o, j, 3, i)));
        FF(o, o->alpha_p, j, k, i) = m;
      }
      k = np - 1;
      {
        double m = fabs(VP(o, j, k - 2, i));
        for (int q = -1; q <= 1; q++) m = dmax(m, fabs(VP(o, j, k + q, i)));
        m = dmax(m, fabs(VP(o, j, 3, i)));
        m = dmax(m, fabs(VP(o, j, 4, i)));
        FF(o, o->alpha_p, j, k, i) = m;
      }
    }
  }
  periodic3(o->alpha_p, ntm, npm, o->nr);
}

/* ---- get_flow_dtmax: src/hipft.f90:7287-7350 (SSPRK43) ------------------ */
double orc_flow_dtmax(orc *o, int strang) {
  double dtmax = 1.7976931348623157e308;
  for (int i = 1; i <= o->nr; i++)
    for (int k = 2; k <= o->npm - 1; k++)
      for (int j = 2; j <= o->ntm - 1; j++) {
        double kdotv = dmax(fabs(VT(o, j, k, i)), fabs(VT(o, j + 1, k, i))) * o->dt_i[j] +
                       dmax(fabs(VP(o, j, k, i)), fabs(VP(o, j, k + 1, i))) * o->st_i[j] *
                           o->dp_i[k];
        double deltat = half / dmax(kdotv, 2.2250738585072014e-308);
        dtmax = dmin(dtmax, deltat);
      }
  double r = safety * dtmax;
  if (strang) r = two * r;
  r = two * r; /* SSPRK(4,3) CFL = 2, 7340-7342 */
  return r;
}

/* ---- advection_operator_weno3: src/hipft.f90:7711-7925 ----------------- */
void orc_advection_operator(orc *o, const double *x, double *aop) {
  int ntm = o->ntm, npm = o->npm, nt = o->nt, np = o->np, nr = o->nr;
  double *flux_t = (double *)calloc((size_t)nt * npm, sizeof(double));
  double *flux_p = (double *)calloc((size_t)ntm * np, sizeof(double));
  double *LP = (double *)calloc((size_t)ntm * (npm + 1), sizeof(double));
  double *LN = (double *)calloc((size_t)ntm * (npm + 1), sizeof(double));
  double *pv = (double *)malloc(sizeof(double) * (size_t)npm);
#define FT(j, k) flux_t[((j) - 1) + (size_t)nt * ((k) - 1)]
#define FPX(j, k) flux_p[((j) - 1) + (size_t)ntm * ((k) - 1)]
#define L2(a, j, k) (a)[((j) - 1) + (size_t)ntm * (k)] /* k = 0..npm */
  for (int i = 1; i <= nr; i++) {
    /* theta: LF split 7766-7769 */
    #pragma omp parallel for schedule(static)
    for (int k = 1; k <= npm; k++)
      for (int j = 1; j <= ntm; j++) {
        double f = FF(o, x, j, k, i), a = FF(o, o->alpha_t, j, k, i);
        L2(LP, j, k) = half * f * (VT(o, j + 1, k, i) - a);
        L2(LN, j, k) = half * f * (VT(o, j, k, i) + a);
      }
    const double *CP = o->D_C_CPt, *MC = o->D_C_MCt;
    #pragma omp parallel for schedule(static)
    for (int k = 1; k <= npm; k++)
      for (int j = 3; j <= nt - 2; j++) { /* 7773-7803 */
        double LNm2 = L2(LN, j - 2, k), LNm1 = L2(LN, j - 1, k), LN0 = L2(LN, j, k);
        double LPm1 = L2(LP, j - 1, k), LP0 = L2(LP, j, k), LPp1 = L2(LP, j + 1, k);
        double p0m = (one + MC[j - 1]) * LNm1 - MC[j - 1] * LNm2;
        double p1m = MC[j] * LNm1 + CP[j - 1] * LN0;
        double p0p = (one + CP[j]) * LP0 - CP[j] * LPp1;
        double p1p = CP[j - 1] * LP0 + MC[j] * LPm1;
        double t0 = MC[j - 1] * (LNm1 - LNm2), t1 = CP[j - 1] * (LN0 - LNm1);
        double t2 = CP[j] * (LPp1 - LP0), t3 = MC[j] * (LP0 - LPm1);
        double B0m = four * (t0 * t0), B1m = four * (t1 * t1);
        double B0p = four * (t2 * t2), B1p = four * (t3 * t3);
        double e0 = o->eps_t[j - 1] + B0m, e1 = o->eps_t[j - 1] + B1m;
        double e2 = o->eps_t[j] + B0p, e3 = o->eps_t[j] + B1p;
        double w0m = o->D_P_Tt[j - 1] / (e0 * e0);
        double w1m = o->D_MC_Tt[j - 1] / (e1 * e1);
        double w0p = o->D_M_Tt[j] / (e2 * e2);
        double w1p = o->D_CP_Tt[j] / (e3 * e3);
        double wm_sum = w0m + w1m, wp_sum = w0p + w1p;
        double OM0m = w0m / wm_sum, OM1m = w1m / wm_sum;
        double OM0p = w0p / wp_sum, OM1p = w1p / wp_sum;
        double um = OM0m * p0m + OM1m * p1m;
        double up = OM0p * p0p + OM1p * p1p;
        FT(j, k) = up + um;
      }
    for (int k = 1; k <= npm; k++) { /* pole-face upwind 7807-7817 */
      double v = VT(o, 2, k, i);
      double cct = v >= 0.0 ? one : -one; /* sign(upwind,v), upwind = 1 */
      if (v == 0.0 && signbit(v)) cct = -one;
      FT(2, k) = v * half * ((one - cct) * FF(o, x, 2, k, i) + (one + cct) * FF(o, x, 1, k, i));
      v = VT(o, nt - 1, k, i);
      cct = v >= 0.0 ? one : -one;
      if (v == 0.0 && signbit(v)) cct = -one;
      FT(nt - 1, k) = v * half * ((one - cct) * FF(o, x, ntm, k, i) +
                                  (one + cct) * FF(o, x, ntm - 1, k, i));
    }
    /* phi: LF split 7833-7845 */
    #pragma omp parallel for schedule(static)
    for (int k = 1; k <= npm - 1; k++)
      for (int j = 2; j <= ntm - 1; j++) {
        double f = FF(o, x, j, k, i), a = FF(o, o->alpha_p, j, k, i);
        L2(LP, j, k) = half * f * (VP(o, j, k + 1, i) - a);
        L2(LN, j, k) = half * f * (VP(o, j, k, i) + a);
      }
    for (int j = 2; j <= ntm - 1; j++) {
      L2(LP, j, 0) = L2(LP, j, npm - 2);
      L2(LN, j, 0) = L2(LN, j, npm - 2);
      L2(LP, j, npm) = L2(LP, j, 2);
      L2(LN, j, npm) = L2(LN, j, 2);
    }
    const double *CPp = o->D_C_CPp, *MCp = o->D_C_MCp;
    #pragma omp parallel for schedule(static)
    for (int k = 2; k <= o->npm1; k++)
      for (int j = 2; j <= ntm - 1; j++) { /* 7850-7880 */
        double LNm2 = L2(LN, j, k - 2), LNm1 = L2(LN, j, k - 1), LN0 = L2(LN, j, k);
        double LPm1 = L2(LP, j, k - 1), LP0 = L2(LP, j, k), LPp1 = L2(LP, j, k + 1);
        double p0m = (one + MCp[k - 1]) * LNm1 - MCp[k - 1] * LNm2;
        double p1m = MCp[k] * LNm1 + CPp[k - 1] * LN0;
        double p0p = (one + CPp[k]) * LP0 - CPp[k] * LPp1;
        double p1p = CPp[k - 1] * LP0 + MCp[k] * LPm1;
        double t0 = MCp[k - 1] * (LNm1 - LNm2), t1 = CPp[k - 1] * (LN0 - LNm1);
        double t2 = CPp[k] * (LPp1 - LP0), t3 = MCp[k] * (LP0 - LPm1);
        double B0m = four * (t0 * t0), B1m = four * (t1 * t1);
        double B0p = four * (t2 * t2), B1p = four * (t3 * t3);
        double e0 = o->eps_p[k - 1] + B0m, e1 = o->eps_p[k - 1] + B1m;
        double e2 = o->eps_p[k] + B0p, e3 = o->eps_p[k] + B1p;
        double w0m = o->D_P_Tp[k - 1] / (e0 * e0);
        double w1m = o->D_MC_Tp[k - 1] / (e1 * e1);
        double w0p = o->D_M_Tp[k] / (e2 * e2);
        double w1p = o->D_CP_Tp[k] / (e3 * e3);
        double wm_sum = w0m + w1m, wp_sum = w0p + w1p;
        double OM0m = w0m / wm_sum, OM1m = w1m / wm_sum;
        double OM0p = w0p / wp_sum, OM1p = w1p / wp_sum;
        double um = OM0m * p0m + OM1m * p1m;
        double up = OM0p * p0p + OM1p * p1p;
        FPX(j, k) = up + um;
      }
    for (int j = 1; j <= ntm; j++) { /* 7882 */
      FPX(j, 1) = FPX(j, np - 1);
      FPX(j, np) = FPX(j, 2);
    }
    #pragma omp parallel for schedule(static)
    for (int k = 2; k <= npm - 1; k++) /* 7890-7898 */
      for (int j = 2; j <= ntm - 1; j++) {
        double a = (o->sth[j + 1] * FT(j + 1, k) - o->sth[j] * FT(j, k)) * o->st_i[j] * o->dt_i[j];
        double b = (FPX(j, k + 1) - FPX(j, k)) * o->st_i[j] * o->dp_i[k];
        FF(o, aop, j, k, i) = a + b;
      }
    /* pole closure 7902-7915 */
    for (int k = 2; k <= npm - 1; k++) pv[k - 2] = FT(2, k) * o->dp[k];
    double fn = pole_sum64(pv, npm - 2);
    for (int k = 2; k <= npm - 1; k++) pv[k - 2] = FT(nt - 1, k) * o->dp[k];
    double fs = pole_sum64(pv, npm - 2);
    for (int k = 2; k <= npm - 1; k++) {
      FF(o, aop, 1, k, i) = fn * o->bc_flow_npole_fac;
      FF(o, aop, ntm, k, i) = -fs * o->bc_flow_spole_fac;
    }
  }
  periodic3(aop, ntm, npm, nr);
#undef FT
#undef FPX
#undef L2
  free(flux_t); free(flux_p); free(LP); free(LN); free(pv);
}

/* ---- advection_step_ssprk43: src/hipft.f90:7515-7582 -------------------- */
void orc_advection_step(orc *o, double dtl) {
  size_t n = (size_t)o->ntm * o->npm * o->nr;
  double *f1 = (double *)malloc(sizeof(double) * n);
  double *aop = (double *)malloc(sizeof(double) * n);
  double *f = o->f;
  orc_advection_operator(o, f, aop);
  #pragma omp parallel for schedule(static)
  for (size_t q = 0; q < n; q++) f1[q] = f[q] - half * dtl * aop[q];
  orc_advection_operator(o, f1, aop);
  #pragma omp parallel for schedule(static)
  for (size_t q = 0; q < n; q++) f1[q] = f1[q] - half * dtl * aop[q];
  orc_advection_operator(o, f1, aop);
  #pragma omp parallel for schedule(static)
  for (size_t q = 0; q < n; q++) f1[q] = two_third * f[q] + third * f1[q] - sixth * dtl * aop[q];
  orc_advection_operator(o, f1, aop);
  #pragma omp parallel for schedule(static)
  for (size_t q = 0; q < n; q++) f[q] = f1[q] - half * dtl * aop[q];
  free(f1);
  free(aop);
}

/* ---- load_diffusion (constant eta) + load_diffusion_matrix + Euler dt ----
 * src/hipft.f90:6692-6702, 6821-6843, 4927-4983, 4985-5040 */
void orc_load_diffusion(orc *o, const double *eta, double factor) {
  int nt = o->nt, np = o->np, ntm = o->ntm, npm = o->npm, nr = o->nr;
  free(o->dcoef);
  free(o->coef);
  free(o->dt_euler_r);
  o->dcoef = (double *)malloc(sizeof(double) * (size_t)nt * np * nr);
  o->coef = (double *)calloc((size_t)(ntm - 2) * (npm - 2) * 5 * nr, sizeof(double));
  o->dt_euler_r = (double *)calloc((size_t)nr, sizeof(double));
  for (int i = 1; i <= nr; i++)
    for (int k = 1; k <= np; k++)
      for (int j = 1; j <= nt; j++) DC(o, j, k, i) = factor * eta[i - 1];
  const double *dph_i = o->dph_i, *dp_i = o->dp_i, *st_i = o->st_i, *dth_i = o->dth_i;
  const double *dt_i = o->dt_i, *sth = o->sth;
  for (int i = 1; i <= nr; i++)
    for (int k = 2; k <= npm - 1; k++)
      for (int j = 2; j <= ntm - 1; j++) {
        double c1 = half * (DC(o, j, k, i) + DC(o, j + 1, k, i)) * dph_i[k] * dp_i[k] * st_i[j] * st_i[j];
        double c2 = half * (DC(o, j, k, i) + DC(o, j, k + 1, i)) * dth_i[j] * dt_i[j] * st_i[j] * sth[j];
        double c4 = half * (DC(o, j + 1, k, i) + DC(o, j + 1, k + 1, i)) * dth_i[j + 1] * dt_i[j] * st_i[j] * sth[j + 1];
        double c5 = half * (DC(o, j, k + 1, i) + DC(o, j + 1, k + 1, i)) * dph_i[k + 1] * dp_i[k] * st_i[j] * st_i[j];
        COEF(o, j, k, 1, i) = c1;
        COEF(o, j, k, 2, i) = c2;
        COEF(o, j, k, 4, i) = c4;
        COEF(o, j, k, 5, i) = c5;
        COEF(o, j, k, 3, i) = -(c1 + c2 + c4 + c5);
      }
  double max_eig_all = 0.0;
  for (int i = 1; i <= nr; i++) {
    double max_eig = 0.0;
    for (int k = 2; k <= npm - 1; k++)
      for (int j = 2; j <= ntm - 1; j++) {
        double g = zero;
        for (int d = 1; d <= 5; d++) g = g + fabs(COEF(o, j, k, d, i));
        max_eig = dmax(g, max_eig);
      }
    o->dt_euler_r[i - 1] = safety * (two / max_eig);
    max_eig_all = dmax(max_eig_all, max_eig);
  }
  o->dt_euler = safety * (two / max_eig_all);
  o->bc_diff_npole_fac = pi_i * o->dth_i[2] * o->dth_i[2];
  o->bc_diff_spole_fac = pi_i * o->dth_i[nt - 1] * o->dth_i[nt - 1];
  o->have_diffusion = 1;
}

double orc_dt_euler(const orc *o) { return o->dt_euler; }
double orc_dt_euler_r(const orc *o, int r) { return o->dt_euler_r[r]; }

/* ---- diffusion_operator_cd: src/hipft.f90:7927-7990 --------------------- */
void orc_diffusion_operator(orc *o, const double *x, double *y) {
  int ntm = o->ntm, npm = o->npm, nt = o->nt, nr = o->nr;
  double *pv = (double *)malloc(sizeof(double) * (size_t)npm);
  for (int i = 1; i <= nr; i++) {
    #pragma omp parallel for schedule(static)
    for (int k = 2; k <= npm - 1; k++)
      for (int j = 2; j <= ntm - 1; j++)
        FF(o, y, j, k, i) = COEF(o, j, k, 1, i) * FF(o, x, j, k - 1, i) +
                            COEF(o, j, k, 2, i) * FF(o, x, j - 1, k, i) +
                            COEF(o, j, k, 3, i) * FF(o, x, j, k, i) +
                            COEF(o, j, k, 4, i) * FF(o, x, j + 1, k, i) +
                            COEF(o, j, k, 5, i) * FF(o, x, j, k + 1, i);
    for (int k = 2; k <= npm - 1; k++)
      pv[k - 2] = (DC(o, 1, k, i) + DC(o, 2, k, i)) * (FF(o, x, 2, k, i) - FF(o, x, 1, k, i)) * o->dp[k];
    double fn = pole_sum64(pv, npm - 2);
    for (int k = 2; k <= npm - 1; k++)
      pv[k - 2] = (DC(o, nt - 1, k, i) + DC(o, nt, k, i)) *
                  (FF(o, x, ntm, k, i) - FF(o, x, ntm - 1, k, i)) * o->dp[k];
    double fs = pole_sum64(pv, npm - 2);
    for (int k = 1; k <= npm; k++) {
      FF(o, y, 1, k, i) = fn * o->bc_diff_npole_fac;
      FF(o, y, ntm, k, i) = -fs * o->bc_diff_spole_fac;
    }
  }
  periodic3(y, ntm, npm, nr);
  free(pv);
}

/* ---- get_dtime_diffusion_ptl: src/hipft.f90:5042-5161 ------------------- */
double orc_ptl(orc *o, double dtime_in) {
  int ntm = o->ntm, npm = o->npm, nr = o->nr;
  const double fmin = 1e-14, safe = 0.95;
  double dtime_ptl = 1.7976931348623157e308;
  size_t n = (size_t)ntm * npm * nr;
  double *Af = (double *)malloc(sizeof(double) * n);
  double *f = o->f;
  orc_diffusion_operator(o, f, Af);
  double axabsmax = -one;
  for (size_t q = 0; q < n; q++)
    if (fabs(f[q]) > fmin) axabsmax = dmax(fabs(Af[q]), axabsmax);
  if (axabsmax > zero) {
    for (int i = 1; i <= nr; i++)
      for (int k = 2; k <= npm - 1; k++)
        for (int j = 2; j <= ntm - 1; j++) {
          if (axabsmax == fabs(FF(o, Af, j, k, i)) && fabs(FF(o, f, j, k, i)) > fmin) {
            double c = FF(o, f, j, k, i), ca = FF(o, Af, j, k, i);
            int nj[4] = {j - 1, j + 1, j, j}, nk[4] = {k, k, k - 1, k + 1};
            for (int q = 0; q < 4; q++) {
              double deltau = c - FF(o, f, nj[q], nk[q], i);
              double deltaf = ca - FF(o, Af, nj[q], nk[q], i);
              if (deltau * deltaf < -fmin) dtime_ptl = dmin(-deltau / deltaf, dtime_ptl);
            }
          }
        }
    dtime_ptl = safe * dtime_ptl;
    if (dtime_ptl < dtime_in) dtime_ptl = dtime_in;
  }
  free(Af);
  return dtime_ptl;
}

/* ---- load_sts_rkg2: src/hipft.f90:5361-5475 ----------------------------- */
long orc_sts_rkg2(double dtl, double dt_euler, double *uj, double *vj, double *ubj,
                  double *gj, long smax) {
  double s_real = half * (sqrt(twentyfive + twentyfour * (dtl / dt_euler)) - three);
  long s = (long)ceil(s_real);
  if (s < 3) s = 3;
  if (s % 2 == 0) s = s + 1;
  if (!uj || s > smax) return s;
  double w = six / (double)((s + 4) * (s - 1));
  double *b = (double *)malloc(sizeof(double) * (size_t)(s + 1));
  b[1] = one / three;
  b[2] = one / fifteen;
  uj[1 - 1] = one;
  vj[1 - 1] = -9999.0;
  ubj[1 - 1] = w;
  gj[1 - 1] = -9999.0;
  uj[2 - 1] = one / two;
  vj[2 - 1] = -one / ten;
  ubj[2 - 1] = uj[2 - 1] * w;
  gj[2 - 1] = zero;
  for (long j = 3; j <= s; j++) {
    double dj = (double)j;
    b[j] = (four * (double)(j - 1) * (double)(j + 4)) / (three * dj * (double)(j + 1) * (double)(j + 2) * (double)(j + 3));
    double bj_bjm1 = b[j] / b[j - 1];
    double bj_bjm2 = b[j] / b[j - 2];
    uj[j - 1] = bj_bjm1 * (two + one / dj);
    vj[j - 1] = -bj_bjm2 * (one / dj + one);
    ubj[j - 1] = uj[j - 1] * w;
    gj[j - 1] = (half * dj * (dj + one) * b[j - 1] - one) * ubj[j - 1];
  }
  free(b);
  return s;
}

/* ---- diffusion_step_sts_cd: src/hipft.f90:5163-5242 --------------------- */
long orc_diffusion_sts(orc *o, double dtl) {
  long s = orc_sts_rkg2(dtl, o->dt_euler, NULL, NULL, NULL, NULL, 0);
  double *uj = malloc(sizeof(double) * s), *vj = malloc(sizeof(double) * s);
  double *ubj = malloc(sizeof(double) * s), *gj = malloc(sizeof(double) * s);
  orc_sts_rkg2(dtl, o->dt_euler, uj, vj, ubj, gj, s);
  size_t n = (size_t)o->ntm * o->npm * o->nr;
  double *u0 = malloc(sizeof(double) * n), *dty0 = malloc(sizeof(double) * n);
  double *ykm1 = malloc(sizeof(double) * n), *ukm1 = malloc(sizeof(double) * n);
  double *ukm2 = malloc(sizeof(double) * n);
  double *f = o->f;
  orc_diffusion_operator(o, f, dty0);
  #pragma omp parallel for schedule(static)
  for (size_t q = 0; q < n; q++) {
    u0[q] = f[q];
    ukm2[q] = f[q];
    dty0[q] = dtl * dty0[q];
    ukm1[q] = f[q] + ubj[0] * dty0[q];
  }
  for (long k = 2; k <= s; k++) {
    orc_diffusion_operator(o, ukm1, ykm1);
    double a = uj[k - 1], b = vj[k - 1], c = ubj[k - 1], g = gj[k - 1];
    #pragma omp parallel for schedule(static)
    for (size_t q = 0; q < n; q++) {
      f[q] = a * ukm1[q] + b * ukm2[q] + (one - a - b) * u0[q] + c * dtl * ykm1[q] + g * dty0[q];
      ukm2[q] = ukm1[q];
      ukm1[q] = f[q];
    }
  }
  free(uj); free(vj); free(ubj); free(gj);
  free(u0); free(dty0); free(ykm1); free(ukm1); free(ukm2);
  return s;
}

/* ---- diffusion_step: src/hipft.f90:4551-4695 (RKG2, optional PTL) ------- */
int orc_diffusion_step(orc *o, double dtl, int subcycles, int max_cycles, double *cyc_time,
                       double *cyc_dt, long *cyc_s, long *cyc_sc, double *hist,
                       const double *fval_pt) {
  const int subcycles_max = 60;
  int auto_sc = subcycles == 0;
  int i = 1, done = 0;
  double time_stepped = 0.0;
  double dtl2 = auto_sc ? zero : dtl / (double)subcycles;
  for (;;) {
    if (auto_sc) dtl2 = orc_ptl(o, dtl2);
    if (auto_sc && time_stepped + dtl2 >= dtl) {
      dtl2 = dtl - time_stepped;
      done = 1;
    }
    if (auto_sc && i == subcycles_max - 1) {
      dtl2 = dtl - time_stepped;
      done = 1;
    }
    long sc = subcycles;
    if (auto_sc) {
      double q = ceil((dtl - time_stepped) / dtl2);
      sc = q > 2e9 ? 2000000000L : (long)q;
    }
    long s = orc_diffusion_sts(o, dtl2);
    time_stepped = time_stepped + dtl2;
    if (i <= max_cycles) {
      if (cyc_time) cyc_time[i - 1] = time_stepped;
      if (cyc_dt) cyc_dt[i - 1] = dtl2;
      if (cyc_s) cyc_s[i - 1] = s;
      if (cyc_sc) cyc_sc[i - 1] = sc;
      if (hist) orc_analysis(o, fval_pt, hist + (size_t)(i - 1) * o->nr * 14);
    }
    if (!auto_sc && i == subcycles) done = 1;
    if (done) break;
    i = i + 1;
  }
  return i;
}

/* ---- analysis_step: src/hipft.f90:3145-3325 ------------------------------
 * out[r*14 + c], c: 0 BR_MIN 1 BR_MAX 2 BR_ABS_MIN 3 FLUX_POSITIVE
 * 4 FLUX_NEGATIVE 5 NPOLE_FLUX_POSITIVE 6 NPOLE_FLUX_NEGATIVE
 * 7 SPOLE_FLUX_POSITIVE 8 SPOLE_FLUX_NEGATIVE 9 NPOLE_AREA 10 SPOLE_AREA
 * 11 EQ_DIPOLE 12 AX_DIPOLE 13 VALIDATION_ERR_HHabs */
void orc_analysis(const orc *o, const double *fval_pt, double *out) {
  int ntm = o->ntm, npm = o->npm, npm1 = npm - 1;
  const double *t = o->t, *p = o->p;
  for (int k = 1; k <= o->nr; k++) {
    double mn = 1.7976931348623157e308, mx = 2.2250738585072014e-308;
    double mnabs = 1.7976931348623157e308, valerr = zero, sumfval2 = zero;
    for (int j = 1; j <= npm - 2; j++)
      for (int i = 1; i <= ntm; i++) {
        double f = FF(o, o->f, i, j, k);
        double fv_ref = fval_pt ? F3(fval_pt, npm1, ntm, j, i, k) : f;
        mn = dmin(f, mn);
        mx = dmax(f, mx);
        mnabs = dmin(fabs(f), mnabs);
        double d = f - fv_ref;
        valerr = valerr + d * d;
        sumfval2 = sumfval2 + fabs(f) * fabs(fv_ref);
      }
    double *r = out + (size_t)(k - 1) * 14;
    r[0] = mn;
    r[1] = mx;
    r[2] = mnabs;
    if (sumfval2 == zero)
      r[13] = sqrt(valerr / (double)(ntm * (npm - 1)));
    else
      r[13] = sqrt(valerr / sumfval2);
    double fp = 0, fm = 0, fppn = 0, fmpn = 0, fpps = 0, fmps = 0, apn = 0, aps = 0;
    double eqd1 = 0, eqd2 = 0, ax = 0;
    for (int i = 1; i <= npm - 1; i++)
      for (int j = 1; j <= ntm; j++) {
        double sn_t, cs_t, d_t, da_t, da_p;
        if (j == 1) {
          double tav = half * (t[1] + t[2]);
          sn_t = sin(tav);
          cs_t = cos(tav);
          d_t = o->dth[2];
          da_t = quarter * sn_t * d_t;
        } else if (j == ntm) {
          double tav = half * (t[ntm] + t[ntm - 1]);
          sn_t = sin(tav);
          cs_t = cos(tav);
          d_t = o->dth[ntm];
          da_t = quarter * sn_t * d_t;
        } else {
          sn_t = sin(t[j]);
          cs_t = cos(t[j]);
          d_t = o->dt[j];
          da_t = sn_t * d_t;
        }
        if (i == 1)
          da_p = half * o->dph[1];
        else if (i == npm - 1)
          da_p = half * o->dph[npm - 1];
        else
          da_p = o->dp[i];
        double cs_p = cos(p[i]), sn_p = sin(p[i]);
        double f = FF(o, o->f, j, i, k);
        double fv = f * da_t * da_p;
        if (fv > 0.)
          fp = fp + fv;
        else
          fm = fm + fv;
        if (t[j] <= o->pole_flux_lat_limit * d2r) {
          if (fv > 0.)
            fppn = fppn + fv;
          else
            fmpn = fmpn + fv;
          apn = apn + da_t * da_p;
        }
        if (t[j] >= pi - o->pole_flux_lat_limit * d2r) {
          if (fv > 0.)
            fpps = fpps + fv;
          else
            fmps = fmps + fv;
          aps = aps + da_t * da_p;
        }
        eqd1 = eqd1 + f * sn_t * cs_p * da_t * da_p;
        eqd2 = eqd2 + f * sn_t * sn_p * da_t * da_p;
        ax = ax + f * cs_t * da_t * da_p;
      }
    r[3] = rsun_cm2 * fp;
    r[4] = rsun_cm2 * fm;
    r[5] = rsun_cm2 * fppn;
    r[6] = rsun_cm2 * fmpn;
    r[7] = rsun_cm2 * fpps;
    r[8] = rsun_cm2 * fmps;
    r[9] = rsun_cm2 * apn;
    r[10] = rsun_cm2 * aps;
    r[12] = three_quarter * pi_i * ax;
    eqd1 = three_quarter * pi_i * eqd1;
    eqd2 = three_quarter * pi_i * eqd2;
    r[11] = sqrt(eqd1 * eqd1 + eqd2 * eqd2);
  }
}

/* ---- weno_eps (load_weno, src/hipft.f90:6895-6901): eps <= 0 gives the
 * cell widths dt(j), dp(k); otherwise every smoothness weight uses eps. */
void orc_set_weno_eps(orc *o, double eps) {
  for (int i = 1; i <= o->ntm; i++) o->eps_t[i] = eps <= 0. ? o->dt[i] : eps;
  for (int i = 1; i <= o->npm; i++) o->eps_p[i] = eps <= 0. ? o->dp[i] : eps;
}

/* pole_flux_lat_limit in degrees (analysis_step, src/hipft.f90:3276-3285). */
void orc_set_pole_flux_lat_limit(orc *o, double deg) { o->pole_flux_lat_limit = deg; }

/* ---- get_flux / balance_flux: src/hipft.f90:8672-8807 ------------------- */
void orc_get_flux(const orc *o, const double *map, double *flux_p, double *flux_m) {
  int ntm = o->ntm, npm = o->npm;
  double fp = zero, fm = zero;
  for (int k = 1; k <= npm - 1; k++)
    for (int j = 1; j <= ntm; j++) {
      double da_t, da_p;
      if (j == 1) {
        double tav = half * (o->t[1] + o->t[2]);
        da_t = quarter * sin(tav) * o->dth[2];
      } else if (j == ntm) {
        double tav = half * (o->t[ntm] + o->t[ntm - 1]);
        da_t = quarter * sin(tav) * o->dth[ntm];
      } else {
        da_t = o->st[j] * o->dt[j];
      }
      if (k == 1)
        da_p = half * o->dph[1];
      else if (k == npm - 1)
        da_p = half * o->dph[npm - 1];
      else
        da_p = o->dp[k];
      double fv = map[(j - 1) + (size_t)ntm * (k - 1)] * da_t * da_p;
      if (fv > zero)
        fp = fp + fv;
      else
        fm = fm + fv;
    }
  *flux_p = fp;
  *flux_m = fm;
}

void orc_balance_flux(const orc *o, double *map) {
  double fp, fm;
  orc_get_flux(o, map, &fp, &fm);
  if (fm + fp == zero) return;
  double ratio = sqrt(fabs(fp / fm));
  size_t n = (size_t)o->ntm * o->npm;
  for (size_t q = 0; q < n; q++) {
    double br = map[q];
    map[q] = br > 0 ? br / ratio : br * ratio;
  }
}

/* ---- sources: generate_rfe (src/hipft.f90:8809-8921) with the balance_flux
 * of load_source / update_source (7203-7208, 3932-3935), and source_step
 * (4697-4733).  rnd is rnd_nums(1:ntm, 2:npm-1), theta fastest (the random
 * numbers already times pi/sqrt(2 pi), 8842-8846); uf_per_hour[nr] in
 * 1e21 Mx/hr (7167-7194); out (ntm,npm,nr). */
void orc_rfe_frame(const orc *o, const double *rnd, int model, const double *uf_per_hour,
                   double *out) {
  const double rsun_cm2 = 4.84416e21, input_flux_fac = 1.0e21;
  int ntm = o->ntm, npm = o->npm, nr = o->nr;
  int num_cells = ntm * (npm - 2);
  for (int i = 1; i <= nr; i++) {
    double sc;
    if (model == 1)
      sc = uf_per_hour[i - 1] * input_flux_fac / (rsun_cm2 * (double)num_cells);
    else
      sc = uf_per_hour[i - 1] * input_flux_fac / (rsun_cm2 * four * pi);
    for (int k = 2; k <= npm - 1; k++)
      for (int j = 1; j <= ntm; j++) {
        double rn = rnd[(j - 1) + (size_t)ntm * (k - 2)];
        double v;
        if (model == 1) {
          double da_ti;
          if (j == 1) {
            double tav = half * (o->t[1] + o->t[2]);
            da_ti = one / (quarter * sin(tav) * o->dth[2]);
          } else if (j == ntm) {
            double tav = half * (o->t[ntm] + o->t[ntm - 1]);
            da_ti = one / (quarter * sin(tav) * o->dth[ntm]);
          } else {
            da_ti = o->st_i[j] * o->dt_i[j];
          }
          v = rn * sc * da_ti * o->dp_i[k];
        } else {
          v = rn * sc;
        }
        FF(o, out, j, k, i) = v;
      }
    for (int j = 1; j <= ntm; j++) {
      FF(o, out, j, 1, i) = FF(o, out, j, npm - 1, i);
      FF(o, out, j, npm, i) = FF(o, out, j, 2, i);
    }
    double fn1 = zero, fs1 = zero;
    for (int k = 1; k <= npm - 2; k++) {
      fn1 = fn1 + FF(o, out, 1, k, i) * o->dp[k];
      fs1 = fs1 + FF(o, out, ntm, k, i) * o->dp[k];
    }
    fn1 = fn1 * twopi_i;
    fs1 = fs1 * twopi_i;
    for (int k = 1; k <= npm; k++) {
      FF(o, out, 1, k, i) = fn1;
      FF(o, out, ntm, k, i) = fs1;
    }
    orc_balance_flux(o, &FF(o, out, 1, 1, i));
  }
}

/* f += dt * src, src (ntm,npm,nr). */
void orc_source_step(orc *o, double dtl, const double *src) {
  size_t n = (size_t)o->ntm * o->npm * o->nr;
  for (size_t q = 0; q < n; q++) o->f[q] = o->f[q] + dtl * src[q];
}

/* ---- apply_data_assimilation + assimilate_new_data -----------------------
 * src/hipft.f90:4066-4079, 4083-4146, 4242-4275 */
void orc_assimilate(orc *o, const double *data_in, int custom_from_mu, const double *mu_power,
                    const double *mu_limit, const double *lat_limit, double mult_fac,
                    int add_mode, int balance) {
  int ntm = o->ntm, npm = o->npm, npm1 = o->npm1, nr = o->nr;
  size_t nslice = (size_t)npm1 * ntm;
  double *nd = (double *)malloc(sizeof(double) * nslice * 3 * nr);
  memcpy(nd, data_in, sizeof(double) * nslice * 3 * nr);
#define ND(l, k, s, i) nd[((l) - 1) + (size_t)npm1 * (((k) - 1) + (size_t)ntm * (((s) - 1) + 3 * (size_t)((i) - 1)))]
  for (int i = 1; i <= nr; i++) {
    if (custom_from_mu)
      for (int k = 1; k <= ntm; k++)
        for (int l = 1; l <= npm1; l++)
          ND(l, k, 2, i) = ND(l, k, 3, i) > mu_limit[i - 1] ? pow(ND(l, k, 3, i), mu_power[i - 1]) : zero;
    if (lat_limit[i - 1] > 0) {
      int tidx0 = -1, tidx1 = -1;
      for (int j = 1; j <= ntm; j++)
        if (o->t[j] >= lat_limit[i - 1] * d2r && tidx0 == -1) tidx0 = j;
      for (int j = ntm; j >= 1; j--)
        if (o->t[j] <= (pi - lat_limit[i - 1] * d2r) && tidx1 == -1) tidx1 = j;
      for (int k = 1; k <= tidx0; k++)
        for (int l = 1; l <= npm1; l++) ND(l, k, 2, i) = zero;
      if (tidx1 >= 1)
        for (int k = tidx1; k <= ntm; k++)
          for (int l = 1; l <= npm1; l++) ND(l, k, 2, i) = zero;
    }
    for (int k = 1; k <= ntm; k++)
      for (int l = 1; l <= npm1; l++) ND(l, k, 1, i) = mult_fac * ND(l, k, 1, i);
  }
  double *deltaf = (double *)calloc((size_t)ntm * npm * nr, sizeof(double));
  for (int i = 1; i <= nr; i++)
    for (int k = 1; k <= npm1; k++)
      for (int j = 1; j <= ntm; j++) {
        if (add_mode)
          FF(o, deltaf, j, k, i) = ND(k, j, 2, i) * ND(k, j, 1, i);
        else
          FF(o, deltaf, j, k, i) = ND(k, j, 2, i) * (ND(k, j, 1, i) - FF(o, o->f, j, k, i));
      }
  periodic3(deltaf, ntm, npm, nr);
  if (balance)
    for (int i = 1; i <= nr; i++) orc_balance_flux(o, deltaf + (size_t)ntm * npm * (i - 1));
  size_t n = (size_t)ntm * npm * nr;
  for (size_t q = 0; q < n; q++) o->f[q] = o->f[q] + deltaf[q];
  periodic3(o->f, ntm, npm, nr);
#undef ND
  free(deltaf);
  free(nd);
}

/* ---- tesseral_spherical_harmonic: src/hipft.f90:2390-2533 --------------- */
void orc_tesseral(int m, int l, const double *pvec, const double *tvec, int np, int nt,
                  double *Y) {
  int abs_m = abs(m);
  const double lpi = 3.141592653589793;
  double *ct = malloc(sizeof(double) * nt), *Pold = malloc(sizeof(double) * nt);
  double *Poldold = malloc(sizeof(double) * nt);
  for (int i = 0; i < nt; i++) {
    ct[i] = cos(tvec[i]);
    Pold[i] = one;
    Poldold[i] = 0.0;
  }
  if (abs_m > 0) {
    for (int i = 0; i < nt; i++) Pold[i] = sqrt(three);
    for (int im = 2; im <= abs_m; im++)
      for (int i = 0; i < nt; i++) Pold[i] = sqrt((two * im + one) / (two * im)) * Pold[i];
  }
  for (int il = abs_m + 1; il <= l; il++) {
    double anm = sqrt(((two * il - one) * (two * il + one)) / (double)((il - abs_m) * (il + abs_m)));
    double bnm = sqrt(((two * il + one) * (double)(il + abs_m - 1) * (double)(il - abs_m - 1)) /
                      ((double)((il - abs_m) * (il + abs_m)) * (two * il - three)));
    for (int i = 0; i < nt; i++) {
      double Pl = anm * ct[i] * Pold[i] - bnm * Poldold[i];
      Poldold[i] = Pold[i];
      Pold[i] = Pl;
    }
  }
  double one_over_sqrt_4pi = one / sqrt(four * lpi);
  for (int i = 0; i < nt; i++)
    Pold[i] = ((abs_m % 2) ? -one : one) * ipow(sin(tvec[i]), abs_m) * one_over_sqrt_4pi * Pold[i];
  int pos = (m + 1) >= 0 ? 1 : 0; /* abs(max(0,sign(1,m+1))) */
  for (int j = 0; j < nt; j++)
    for (int i = 0; i < np; i++)
      Y[i + (size_t)np * j] = Pold[j] * ((double)pos * cos(m * pvec[i]) + (one - pos) * sin(abs_m * pvec[i]));
  free(ct);
  free(Pold);
  free(Poldold);
}
