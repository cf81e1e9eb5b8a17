// hft_lib.hip -- libhft: host controller and C ABI (include/hft.h) of the
// MI355X flux-transport core.  Kernels live in hft_kernels.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <functional>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/hft.h"
#include "hft_kernels.hip"
#include "hft_adjoint_kernels.hip"

using namespace hft;

namespace {

thread_local std::string g_err;

int fail(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define HIPCHK(x)                                                                  \
  do {                                                                             \
    hipError_t e_ = (x);                                                           \
    if (e_ != hipSuccess)                                                          \
      return fail(HFT_E_HIP, "%s failed: %s (%s:%d)", #x, hipGetErrorString(e_),  \
                  __FILE__, __LINE__);                                             \
  } while (0)

// Reference constants (src/hipft.f90:83-136).
const double zero = 0.0, one = 1.0, two = 2.0, three = 3.0, four = 4.0, six = 6.0;
const double ten = 10.0, fifteen = 15.0, half = 0.5, quarter = 0.25;
const double twentyfour = 24.0, twentyfive = 25.0;
const double sixth = 0.16666666666666666, two_third = 0.66666666666666666;
const double third = 0.33333333333333333;
const double pi = 3.1415926535897932, pi_i = 0.3183098861837907;
const double twopi = 6.2831853071795864, twopi_i = 0.15915494309189535;
const double d2r = 0.017453292519943295;
const double km_s_to_rs_hr = 0.005172413793103448, m_s_to_rs_hr = 5.172413793103448e-06;
const double safety = 0.95;
const double rsun_cm2 = 4.84416e21, input_flux_fac = 1.0e21;
const double kHuge = 1.7976931348623157e308, kTiny = 2.2250738585072014e-308;

// x**n, integer n >= 0, by square-and-multiply (the expansion compilers emit
// for an integer exponent).
double ipow(double x, int n) {
  double res = 1.0, sq = x;
  bool have = false;
  while (n) {
    if (n & 1) {
      res = have ? res * sq : sq;
      have = true;
    }
    n >>= 1;
    if (n) sq = sq * sq;
  }
  return res;
}

template <typename T>
int dalloc(T **p, size_t n) {
  if (n == 0) n = 1;
  hipError_t e = hipMalloc((void **)p, n * sizeof(T));
  if (e != hipSuccess) return fail(HFT_E_NOMEM, "hipMalloc(%zu bytes): %s", n * sizeof(T), hipGetErrorString(e));
  return 0;
}

struct TimedPair {
  hipEvent_t a, b;
  int group;
  double units;   // compulsory field passes x realizations of the launch
  double sunits;  // the same work done stage by stage (differs for fused launches)
};

// Kernel timing groups; launches timed under "rkg2_fused" also count in "rkg2".
// The phase_* groups time whole phases of the step (events at the phase
// boundaries on the context's stream), the reference's wtime_flux_transport*
// and wtime_source buckets.
const char *kGroups[] = {"rkg2", "weno", "ptl", "analysis", "assim", "flow", "rkg2_fused",
                         "adjoint", "phase_advection", "phase_diffusion", "phase_source",
                         "phase_flux_transport"};
constexpr int kNGroups = 12;
constexpr int kGroupFused = 6;
constexpr int kGroupAdjoint = 7;
constexpr int kPhaseAdvection = 8;
constexpr int kPhaseDiffusion = 9;
constexpr int kPhaseSource = 10;
constexpr int kPhaseStep = 11;

}  // namespace

struct hft_ctx {
  int ntm, npm1, npm, nt, np, nu, pitch, nr, mode, device;
  size_t rstride, nfield;
  // host mesh, reference indexing (element 0 unused)
  std::vector<double> t, p, th, ph, dt, dth, dp, dph, dt_i, dth_i, dp_i, dph_i, st, sth, ct,
      cth, st_i, sth_i;
  std::vector<double> wt[7], wp[7];  // WENO weights CP MC MT CPT PT MCT eps
  hipStream_t stream = nullptr;
  // device mesh storage
  std::vector<double *> dmesh;
  Mesh m;
  // fields
  double *f = nullptr, *buf0 = nullptr, *buf1 = nullptr, *dty0 = nullptr, *fval = nullptr;
  double *buf2 = nullptr, *buf3 = nullptr;  // extra stage buffers of the fused RKG2 path
  // tuning (hft_set_tuning): fused multi-stage RKG2 launches
  int fused = 1, fused_T = 4, band_rows = 32, fused_min = 8;
  int fused_target_blocks = 0;  // shrink bands while fewer blocks than this (0: off)
  // 1: each fused launch picks its band count against whole rounds of
  // co-resident blocks (fused_band_count); 0: bands of band_rows rows
  int fused_bands = 0;
  int fused_band_share = 32;  // percent: slowdown of a round from an idle to a full device
  int fused_band_min_rows = 8;  // thinnest band fused_band_count may choose
  int fused_cap[6] = {0, 0, 0, 0, 0, 0};  // co-resident k_rkg2_fused2<T> blocks on the device
  int fused_kernel = 3;  // 1: k_rkg2_fused, 2: k_rkg2_march, 3: k_rkg2_fused2 (2 columns/thread)
  int fused_body = 1;    // k_rkg2_fused2: interior march steps without row tests
  int weno_fastdiv = 1;  // WENO3 weights: short divides where the operands allow
  int adv_band = 1, adv_band_rows = 16;  // banded WENO3 stage kernel
  // persistent super-step (k_rkg2_persist): 0 = off, 1 = on; band rows of
  // the realizations with the most stages and of the others
  int persist = 0, persist_rows = 32, persist_rows_long = 32, persist_long = 0;
  unsigned *d_ptiles = nullptr, *h_ptiles = nullptr;
  size_t ptiles_cap = 0;
  int *d_pint = nullptr, *h_pint = nullptr;  // zH, znb, zoff (device) / staging (pinned)
  size_t pint_cap = 0;
  int *d_pctl = nullptr;  // counter, err, done[...]
  size_t pctl_cap = 0;
  hipEvent_t ptiles_ev = nullptr;
  int persist_cap = 0;
  int wall_khz = 0;  // wall_clock64() rate of the device
  // independent mode: per-realization CFL sub-steps of the advection step
  int adv_substeps = 1;
  int64_t flow_gen = 0, adv_lim_gen = -1;
  std::vector<double> adv_lim;
  int *d_adv_order = nullptr, *h_adv_order = nullptr;
  double *d_adv_dt = nullptr, *h_adv_dt = nullptr;
  hipEvent_t adv_ev = nullptr;
  // hipGraph replay of RKG2 super-step launch sequences, keyed by
  // (use_af, realization -> stage count in launch order)
  int graphs_on = 0;
  std::vector<std::pair<std::vector<int64_t>, hipGraphExec_t>> graphs;
  int64_t graph_hits = 0, graph_builds = 0;
  // whole flux-transport steps planned on the device and replayed as one
  // hipGraph (tuning "step_graph"): the captured sequence, its key, the key
  // of the previous step, and the largest stage count seen in each round of
  // sub-cycles (sizes the sequence)
  int step_graph = 1, step_graph_rounds = 1;
  hipGraphExec_t sg_exec = nullptr;
  std::vector<int64_t> sg_key, sg_prev, sg_smax;
  int64_t sg_hits = 0, sg_builds = 0, sg_overflows = 0;
  int64_t epoch = 0;  // bumped by every setter that changes what a captured step reads
  char *d_plan = nullptr, *h_plan = nullptr;  // controller state (device) / its pinned copy
  size_t plan_bytes = 0;
  int *d_slots = nullptr, *d_htab = nullptr;
  size_t slots_cap = 0;
  double *d_dte = nullptr;
  std::vector<int64_t> *obs_smax = nullptr;  // host controller: record each round's largest s
  const int *adv_skip = nullptr;             // advection launches skip when *adv_skip != 0
  bool have_fval = false;
  // flows
  bool flow_set = false, flow2d = false;
  std::vector<double> prof_vt, prof_vp;  // base profiles (code units) or empty
  std::vector<double> an_p0, an_p2, an_p4, an_p1, an_p3, an_p5;
  double an_omega = 0, an_vtc = 0;
  std::vector<double> vt1, vp1, at1, ap1;  // [nr][nt], [nr][ntm] x3
  double *d_vt1 = nullptr, *d_vp1 = nullptr, *d_at1 = nullptr, *d_ap1 = nullptr;
  std::vector<double> atten_inv;
  bool atten = false;
  double *d_atten_inv = nullptr, *d_vt2 = nullptr, *d_vp2 = nullptr, *d_at2 = nullptr,
         *d_ap2 = nullptr;
  size_t vt_rstride = 0;
  // 2-D base flows (flows from files): one shared or one per realization,
  // and the analytic terms update_flow adds to them
  bool ff = false;
  size_t ff_vt_rstride = 0, ff_vp_rstride = 0;
  double *d_vtf = nullptr, *d_vpf = nullptr, *d_mft = nullptr, *d_drp = nullptr,
         *d_omp = nullptr;
  bool ff_mf = false, ff_dr = false, ff_om = false;
  // sources: a constant map (source from file) and two random flux
  // emergence frames blended as (1 - alpha) rfe[first] + alpha rfe[1 - first]
  double *d_src = nullptr, *d_rfe[2] = {nullptr, nullptr}, *d_rnd = nullptr;
  double *d_rfe_dati = nullptr, *d_rfe_scale = nullptr;
  bool src_on = false;
  int rfe_first = -1;
  double rfe_alpha = 0.0;
  // diffusion
  bool diff_set = false;
  std::vector<double> D, dt_euler_r;
  double dt_euler = 0;
  double *d_D = nullptr;
  // RKG2 / PTL work arrays
  StageCoef *d_coef = nullptr;
  size_t coef_cap = 0;
  StageCoef *h_coef = nullptr;  // pinned
  size_t h_coef_cap = 0;
  hipEvent_t coef_ev = nullptr;
  int64_t *d_s = nullptr;
  int *d_order = nullptr, *d_order_id = nullptr;
  double *d_cdt = nullptr, *d_ub1 = nullptr;
  unsigned long long *d_absmax = nullptr, *d_dtmin = nullptr, *d_target = nullptr;
  uint4 *d_prec = nullptr;  // PTL tile records of k_laplacian (nr x ntm x column tiles)
  int *h_small = nullptr;  // pinned scratch for small uploads
  double *h_smalld = nullptr;
  std::vector<int64_t> last_s, last_cycles;
  // analysis
  double *d_part = nullptr, *d_hist = nullptr;
  double *h_hist = nullptr;  // pinned copy of d_hist (hft_analysis)
  double *d_da_t = nullptr, *d_sn_t = nullptr, *d_cs_t = nullptr, *d_tt = nullptr;
  double *d_da_p = nullptr, *d_cs_p = nullptr, *d_sn_p = nullptr;
  int *d_pcol = nullptr;
  int ana_nblk = 0, ana_rows = 0;
  // assimilation
  double *d_map = nullptr;
  size_t map_cap = 0;
  double *d_apar = nullptr;
  int *d_ipar = nullptr;
  // adjoint (hft_step_adjoint): field buffers allocated on first use
  std::vector<double *> adj_fix, adj_ck, adj_w, adj_st;
  double *d_apD = nullptr, *d_apA = nullptr, *d_gout = nullptr;
  size_t apA_cap = 0;
  int adj_band_rows = 16, adj_seg = 0;
  int adj_fused = 1, adj_fused_rows = 16;  // multi-stage RKG2 adjoint over stored stages
  int64_t adj_mem_mb = 32768;
  hipStream_t own_stream = nullptr;  // the context's stream when a caller's is in use
  // timing
  bool timing = false;        // events around every launch
  bool phase_timing = false;  // events at the boundaries of the step's phases
  double *scratch_dev = nullptr;  // this device's copy of `scratch` (hft_analysis_noalloc)
  double pole_lat = 30.0;         // pole_flux_lat_limit, degrees (src/hipft.f90:566)
  std::vector<TimedPair> pending;
  std::vector<hipEvent_t> free_ev;
  double t_ms[kNGroups] = {0};
  int64_t t_n[kNGroups] = {0};
  double t_units[kNGroups] = {0};
  double t_sunits[kNGroups] = {0};
};

namespace {

int ev_get(hft_ctx *c, hipEvent_t *e) {
  if (!c->free_ev.empty()) {
    *e = c->free_ev.back();
    c->free_ev.pop_back();
    return 0;
  }
  HIPCHK(hipEventCreate(e));
  return 0;
}

int drain_timing(hft_ctx *c) {
  if (c->pending.empty()) return 0;
  HIPCHK(hipStreamSynchronize(c->stream));
  for (auto &tp : c->pending) {
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, tp.a, tp.b));
    c->t_ms[tp.group] += ms;
    c->t_n[tp.group] += 1;
    c->t_units[tp.group] += tp.units;
    c->t_sunits[tp.group] += tp.sunits;
    if (tp.group == kGroupFused) {
      c->t_ms[0] += ms;
      c->t_n[0] += 1;
      c->t_units[0] += tp.units;
      c->t_sunits[0] += tp.sunits;
    }
    c->free_ev.push_back(tp.a);
    c->free_ev.push_back(tp.b);
  }
  c->pending.clear();
  return 0;
}

// Launch helper: records events around the launch when timing is enabled.
// units: compulsory field passes (reads + writes of a realization's field)
// the launch makes, times its realizations; used for algorithmic bytes.
// sunits: the passes the same work takes one stage per launch (a fused
// multi-stage launch keeps its intermediate stages on chip); < 0 = units.
template <typename K, typename A>
int launch_ex(hft_ctx *c, int group, K kern, int blocks, int threads, size_t shmem, const A &args,
              double units, double sunits = -1.0) {
  if (blocks <= 0) return 0;
  hipEvent_t a = nullptr, b = nullptr;
  if (c->timing) {
    if (c->pending.size() > 4096) {
      int rc = drain_timing(c);
      if (rc) return rc;
    }
    int rc = ev_get(c, &a);
    if (!rc) rc = ev_get(c, &b);
    if (rc) return rc;
    HIPCHK(hipEventRecord(a, c->stream));
  }
  hipLaunchKernelGGL(kern, dim3(blocks), dim3(threads), shmem, c->stream, args);
  HIPCHK(hipGetLastError());
  if (c->timing) {
    HIPCHK(hipEventRecord(b, c->stream));
    c->pending.push_back({a, b, group, units, sunits < 0 ? units : sunits});
  }
  return 0;
}

// Phase timing: phase_begin records the opening event (none when phase
// timing is off), phase_end the closing one and queues the pair; a failed
// phase returns its events to the pool.
int phase_begin(hft_ctx *c, hipEvent_t *a) {
  *a = nullptr;
  if (!c->phase_timing) return 0;
  if (c->pending.size() > 4096) {
    int rc = drain_timing(c);
    if (rc) return rc;
  }
  int rc = ev_get(c, a);
  if (rc) return rc;
  HIPCHK(hipEventRecord(*a, c->stream));
  return 0;
}

int phase_end(hft_ctx *c, hipEvent_t a, int group, int body_rc) {
  if (!a) return body_rc;
  if (body_rc) {
    c->free_ev.push_back(a);
    return body_rc;
  }
  hipEvent_t b = nullptr;
  int rc = ev_get(c, &b);
  if (rc) return rc;
  HIPCHK(hipEventRecord(b, c->stream));
  c->pending.push_back({a, b, group, 0.0, 0.0});
  return 0;
}

template <typename K, typename A>
int launch(hft_ctx *c, int group, K kern, int blocks, const A &args, double units) {
  return launch_ex(c, group, kern, blocks, kBlock, 0, args, units);
}

Tiles tiles_for(const hft_ctx *c, int nreal) {
  Tiles g;
  g.ctiles = (c->nu + kBlock - 1) / kBlock;
  g.ntm = c->ntm;
  g.nreal = nreal;
  return g;
}
int nblocks(const Tiles &g) { return g.ctiles * g.ntm * g.nreal; }
// k_laplacian: one block per column tile of kLapRows rows
int lap_blocks(const Tiles &g) { return g.ctiles * ((g.ntm + kLapRows - 1) / kLapRows) * g.nreal; }

// ---- set_mesh: src/hipft.f90:6495-6653 ------------------------------------
void build_mesh(hft_ctx *c, const double *tin, const double *pin) {
  int ntm = c->ntm, npm1 = c->npm1, nt = c->nt, np = c->np, npm = c->npm;
  auto V = [](int n) { return std::vector<double>(n + 1, 0.0); };
  c->t = V(ntm);
  c->p = V(np);
  for (int i = 1; i <= ntm; i++) c->t[i] = tin[i - 1];
  for (int i = 1; i <= npm1; i++) c->p[i] = pin[i - 1];
  c->p[npm1 + 1] = c->p[2] + twopi;
  c->dt = V(ntm); c->dt_i = V(ntm); c->st = V(ntm); c->st_i = V(ntm); c->ct = V(ntm);
  c->dp = V(npm); c->dp_i = V(npm);
  c->th = V(nt); c->ph = V(np); c->dth = V(nt); c->dth_i = V(nt); c->sth = V(nt);
  c->sth_i = V(nt); c->cth = V(nt); c->dph = V(np); c->dph_i = V(np);
  auto &t = c->t, &p = c->p;
  for (int i = 2; i <= ntm; i++) {
    c->th[i] = half * (t[i] + t[i - 1]);
    c->dth[i] = t[i] - t[i - 1];
  }
  c->th[1] = c->th[2] - c->dth[2];
  c->th[nt] = c->th[ntm] + c->dth[ntm];
  c->dth[1] = c->dth[2];
  c->dth[nt] = c->dth[ntm];
  for (int i = 1; i <= nt; i++) c->dth_i[i] = one / c->dth[i];
  for (int i = 1; i <= ntm; i++) {
    c->dt[i] = c->th[i + 1] - c->th[i];
    c->dt_i[i] = one / c->dt[i];
  }
  for (int i = 1; i <= ntm; i++) c->st[i] = std::sin(t[i]);
  c->st[1] = 0.0;
  c->st[ntm] = 0.0;
  for (int i = 1; i <= ntm; i++) c->ct[i] = std::cos(t[i]);
  c->ct[1] = one;
  c->ct[ntm] = one;
  for (int i = 2; i <= ntm - 1; i++) c->st_i[i] = one / c->st[i];
  c->st_i[1] = 0.0;
  c->st_i[ntm] = 0.0;
  for (int i = 1; i <= nt; i++) c->sth[i] = std::sin(std::fabs(c->th[i]));
  for (int i = 2; i <= ntm; i++) c->sth_i[i] = one / c->sth[i];
  c->sth_i[1] = 0.0;
  c->sth_i[nt] = 0.0;
  for (int i = 1; i <= nt; i++) c->cth[i] = std::cos(c->th[i]);
  for (int i = 2; i <= np; i++) {
    c->ph[i] = half * (p[i] + p[i - 1]);
    c->dph[i] = p[i] - p[i - 1];
  }
  c->ph[1] = c->ph[npm1] - twopi;
  c->dph[1] = c->dph[npm1];
  for (int i = 1; i <= np; i++) c->dph_i[i] = one / c->dph[i];
  for (int i = 1; i <= npm - 1; i++) c->dp[i] = c->ph[i + 1] - c->ph[i];
  c->dp[npm] = c->dp[2];
  for (int i = 1; i <= npm; i++) c->dp_i[i] = one / c->dp[i];
  // load_weno: src/hipft.f90:6895-6932 (weno_eps = 0: cell widths)
  for (int q = 0; q < 7; q++) {
    c->wt[q] = V(nt);
    c->wp[q] = V(np);
  }
  for (int j = 1; j <= ntm; j++) c->wt[6][j] = c->dt[j];
  for (int k = 1; k <= npm; k++) c->wp[6][k] = c->dp[k];
  auto weights = [](std::vector<double> *w, const std::vector<double> &d, int lo, int hi) {
    for (int j = lo; j <= hi; j++) {
      double tot = d[j - 1] + d[j] + d[j + 1];
      w[0][j] = d[j] / (d[j] + d[j + 1]);          // D_C_CP
      w[1][j] = d[j] / (d[j - 1] + d[j]);          // D_C_MC
      w[2][j] = d[j - 1] / tot;                    // D_M_T
      w[3][j] = (d[j] + d[j + 1]) / tot;           // D_CP_T
      w[4][j] = d[j + 1] / tot;                    // D_P_T
      w[5][j] = (d[j - 1] + d[j]) / tot;           // D_MC_T
    }
  };
  weights(c->wt, c->dt, 2, nt - 2);
  weights(c->wp, c->dp, 2, np - 1);
  for (int q = 0; q < 6; q++) {
    c->wp[q][1] = c->wp[q][np - 1];
    c->wp[q][np] = c->wp[q][2];
  }
}

// Mesh::fastdiv: every WENO3 weight that is a numerator of the weights'
// divides (MT, CPT, PT, MCT) and every eps, over the faces the kernels use
// (theta reference indices 2..nt-2, phi 1..np-1), lies in [2^-100, 2^100].
void set_fastdiv(hft_ctx *c) {
  auto in_range = [](double v) { return v >= 0x1p-100 && v <= 0x1p100; };
  bool ok = c->weno_fastdiv != 0;
  for (int q = 2; q <= 6 && ok; q++) {
    for (int j = 2; j <= c->nt - 2; j++) ok = ok && in_range(c->wt[q][j]);
    for (int k = 1; k <= c->np - 1; k++) ok = ok && in_range(c->wp[q][k]);
  }
  c->m.fastdiv = ok ? 1 : 0;
}

int upload_vec(hft_ctx *c, const std::vector<double> &v, const double **out) {
  double *d;
  int rc = dalloc(&d, v.size());
  if (rc) return rc;
  c->dmesh.push_back(d);
  // reference arrays are 1-based: drop element 0
  HIPCHK(hipMemcpy(d, v.data() + 1, (v.size() - 1) * sizeof(double), hipMemcpyHostToDevice));
  *out = d;
  return 0;
}

template <typename T>
int upload_raw(hft_ctx *c, const std::vector<T> &v, T **out) {
  int rc = dalloc(out, v.size());
  if (rc) return rc;
  HIPCHK(hipMemcpy(*out, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
  return 0;
}

int setup_device(hft_ctx *c) {
  Mesh &m = c->m;
  m.ntm = c->ntm; m.nt = c->nt; m.npm = c->npm; m.np = c->np; m.nu = c->nu; m.pitch = c->pitch;
  int rc = 0;
  rc |= upload_vec(c, c->dt_i, &m.dt_i);
  rc |= upload_vec(c, c->st_i, &m.st_i);
  rc |= upload_vec(c, c->sth, &m.sth);
  rc |= upload_vec(c, c->dth_i, &m.dth_i);
  rc |= upload_vec(c, c->dth, &m.dth);
  rc |= upload_vec(c, c->dp_i, &m.dp_i);
  rc |= upload_vec(c, c->dph_i, &m.dph_i);
  rc |= upload_vec(c, c->dp, &m.dp);
  const double **tw[7] = {&m.tCP, &m.tMC, &m.tMT, &m.tCPT, &m.tPT, &m.tMCT, &m.teps};
  const double **pw[7] = {&m.pCP, &m.pMC, &m.pMT, &m.pCPT, &m.pPT, &m.pMCT, &m.peps};
  for (int q = 0; q < 7; q++) {
    rc |= upload_vec(c, c->wt[q], tw[q]);
    rc |= upload_vec(c, c->wp[q], pw[q]);
  }
  if (rc) return rc;
  m.bc_diff_n = pi_i * c->dth_i[2] * c->dth_i[2];  // src/hipft.f90:6841-6842
  m.bc_diff_s = pi_i * c->dth_i[c->nt - 1] * c->dth_i[c->nt - 1];
  m.bc_flow_n = two * c->dth_i[2] * pi_i;  // src/hipft.f90:2568-2569
  m.bc_flow_s = two * c->dth_i[c->nt - 1] * pi_i;
  set_fastdiv(c);
  size_t nf = c->nfield;
  if ((rc = dalloc(&c->f, nf)) || (rc = dalloc(&c->buf0, nf)) || (rc = dalloc(&c->buf1, nf)) ||
      (rc = dalloc(&c->dty0, nf)))
    return rc;
  HIPCHK(hipMemset(c->f, 0, nf * sizeof(double)));
  HIPCHK(hipMemset(c->buf0, 0, nf * sizeof(double)));
  HIPCHK(hipMemset(c->buf1, 0, nf * sizeof(double)));
  HIPCHK(hipMemset(c->dty0, 0, nf * sizeof(double)));
  int nr = c->nr;
  std::vector<int> id(nr);
  for (int i = 0; i < nr; i++) id[i] = i;
  if ((rc = upload_raw(c, id, &c->d_order_id))) return rc;
  if ((rc = dalloc(&c->d_order, nr)) || (rc = dalloc(&c->d_s, nr)) ||
      (rc = dalloc(&c->d_cdt, nr)) || (rc = dalloc(&c->d_ub1, nr)) ||
      (rc = dalloc(&c->d_absmax, nr)) || (rc = dalloc(&c->d_dtmin, nr)) ||
      (rc = dalloc(&c->d_target, 1)) || (rc = dalloc(&c->d_D, nr)) ||
      (rc = dalloc(&c->d_prec, (size_t)nblocks(tiles_for(c, nr)))))
    return rc;
  HIPCHK(hipHostMalloc((void **)&c->h_small, sizeof(int) * (nr * 8 + 64)));
  HIPCHK(hipHostMalloc((void **)&c->h_smalld, sizeof(double) * (nr * 8 + 64)));
  HIPCHK(hipEventCreateWithFlags(&c->coef_ev, hipEventDisableTiming));
  // analysis weights (analysis_step, src/hipft.f90:3235-3262)
  int ntm = c->ntm, npm = c->npm;
  std::vector<double> da_t(ntm), sn_t(ntm), cs_t(ntm), tt(ntm);
  for (int j = 1; j <= ntm; j++) {
    double sn, cs, d_t, dat;
    if (j == 1) {
      double tav = half * (c->t[1] + c->t[2]);
      sn = std::sin(tav); cs = std::cos(tav); d_t = c->dth[2]; dat = quarter * sn * d_t;
    } else if (j == ntm) {
      double tav = half * (c->t[ntm] + c->t[ntm - 1]);
      sn = std::sin(tav); cs = std::cos(tav); d_t = c->dth[ntm]; dat = quarter * sn * d_t;
    } else {
      sn = std::sin(c->t[j]); cs = std::cos(c->t[j]); d_t = c->dt[j]; dat = sn * d_t;
    }
    da_t[j - 1] = dat; sn_t[j - 1] = sn; cs_t[j - 1] = cs; tt[j - 1] = c->t[j];
  }
  std::vector<double> da_p(npm - 1), cs_p(npm - 1), sn_p(npm - 1);
  std::vector<int> pcol(npm - 1);
  for (int i = 1; i <= npm - 1; i++) {
    double dap = (i == 1) ? half * c->dph[1] : (i == npm - 1 ? half * c->dph[npm - 1] : c->dp[i]);
    da_p[i - 1] = dap;
    cs_p[i - 1] = std::cos(c->p[i]);
    sn_p[i - 1] = std::sin(c->p[i]);
    pcol[i - 1] = (i == 1) ? c->nu - 1 : i - 2;
  }
  if ((rc = upload_raw(c, da_t, &c->d_da_t)) || (rc = upload_raw(c, sn_t, &c->d_sn_t)) ||
      (rc = upload_raw(c, cs_t, &c->d_cs_t)) || (rc = upload_raw(c, tt, &c->d_tt)) ||
      (rc = upload_raw(c, da_p, &c->d_da_p)) || (rc = upload_raw(c, cs_p, &c->d_cs_p)) ||
      (rc = upload_raw(c, sn_p, &c->d_sn_p)) || (rc = upload_raw(c, pcol, &c->d_pcol)))
    return rc;
  c->ana_rows = std::max(1, (ntm + 63) / 64);
  c->ana_nblk = (ntm + c->ana_rows - 1) / c->ana_rows;
  if ((rc = dalloc(&c->d_part, (size_t)nr * c->ana_nblk * kNPart)) ||
      (rc = dalloc(&c->d_hist, (size_t)nr * 14)) || (rc = dalloc(&c->d_apar, (size_t)nr * 4)) ||
      (rc = dalloc(&c->d_ipar, (size_t)nr * 4)))
    return rc;
  HIPCHK(hipHostMalloc((void **)&c->h_hist, (size_t)nr * 14 * sizeof(double)));
  return 0;
}

// ---- field layout conversions ---------------------------------------------
// reference f(ntm,npm,nr) <-> device [r][j][c] (c <-> column c+2)
void tp_to_dev(const hft_ctx *c, const double *tp, std::vector<double> &dev) {
  dev.assign(c->nfield, 0.0);
  for (int r = 0; r < c->nr; r++)
    for (int k = 2; k <= c->npm - 1; k++)
      for (int j = 0; j < c->ntm; j++)
        dev[fidx(r, j, k - 2, c->ntm, c->pitch)] =
            tp[j + (size_t)c->ntm * ((k - 1) + (size_t)c->npm * r)];
}

void dev_to_tp(const hft_ctx *c, const std::vector<double> &dev, double *tp) {
  for (int r = 0; r < c->nr; r++)
    for (int k = 1; k <= c->npm; k++) {
      int cc = (k == 1) ? c->nu - 1 : (k == c->npm ? 0 : k - 2);
      for (int j = 0; j < c->ntm; j++)
        tp[j + (size_t)c->ntm * ((k - 1) + (size_t)c->npm * r)] =
            dev[fidx(r, j, cc, c->ntm, c->pitch)];
    }
}

// device layout -> PT map (npm1, ntm, nr), one-point phi overlap
void dev_to_pt(const hft_ctx *c, const std::vector<double> &dev, double *f) {
  for (int r = 0; r < c->nr; r++)
    for (int j = 0; j < c->ntm; j++)
      for (int l = 1; l <= c->npm1; l++) {
        int cc = (l == 1) ? c->nu - 1 : l - 2;
        f[(l - 1) + (size_t)c->npm1 * (j + (size_t)c->ntm * r)] = dev[fidx(r, j, cc, c->ntm, c->pitch)];
      }
}

// Validation field PT (npm1,ntm,nr) -> device layout.  analysis_step pairs
// state columns 1..npm-2 with field columns 1..npm-2 (src/hipft.f90:3190-3197),
// so the seam column (device c = nu-1, reference column 1 == npm-1) takes
// the field's first column (phi = 0), not its last: an analytic field need
// not be periodic.
void pt_to_dev(const hft_ctx *c, const double *pt, std::vector<double> &dev) {
  dev.assign(c->nfield, 0.0);
  for (int r = 0; r < c->nr; r++)
    for (int j = 0; j < c->ntm; j++)
      for (int cc = 0; cc < c->nu; cc++) {
        int l = (cc == c->nu - 1) ? 1 : cc + 2;
        dev[fidx(r, j, cc, c->ntm, c->pitch)] =
            pt[(l - 1) + (size_t)c->npm1 * (j + (size_t)c->ntm * r)];
      }
}

int put_field(hft_ctx *c, double *d, const std::vector<double> &h) {
  HIPCHK(hipMemcpyAsync(d, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

int get_field(hft_ctx *c, const double *d, std::vector<double> &h) {
  h.resize(c->nfield);
  HIPCHK(hipMemcpyAsync(h.data(), d, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

// ---- flows ------------------------------------------------------------------
// update_flow (src/hipft.f90:3411-3448, 7243-7283) and set_lf_alpha
// (6968-7041) for phi-independent flows.
int rebuild_flow(hft_ctx *c) {
  int nr = c->nr, nt = c->nt, ntm = c->ntm;
  c->vt1.assign((size_t)nr * nt, 0.0);
  c->vp1.assign((size_t)nr * ntm, 0.0);
  c->at1.assign((size_t)nr * ntm, 0.0);
  c->ap1.assign((size_t)nr * ntm, 0.0);
  for (int r = 0; r < nr; r++) {
    double *vt = &c->vt1[(size_t)r * nt];
    double *vp = &c->vp1[(size_t)r * ntm];
    if (!c->prof_vt.empty())
      for (int j = 0; j < nt; j++) vt[j] = c->prof_vt[(size_t)r * nt + j];
    if (!c->prof_vp.empty())
      for (int j = 0; j < ntm; j++) vp[j] = c->prof_vp[(size_t)r * ntm + j];
    if (!c->an_p1.empty())
      for (int jr = 1; jr <= nt; jr++) {
        double cc = c->cth[jr];
        vt[jr - 1] = vt[jr - 1] - m_s_to_rs_hr * c->sth[jr] *
                                      (c->an_p1[r] * cc + c->an_p3[r] * ipow(cc, 3) +
                                       c->an_p5[r] * ipow(cc, 5));
      }
    if (!c->an_p0.empty())
      for (int jr = 1; jr <= ntm; jr++) {
        double cc = c->ct[jr];
        vp[jr - 1] = vp[jr - 1] + m_s_to_rs_hr * c->st[jr] *
                                      (c->an_p0[r] + c->an_p2[r] * ipow(cc, 2) +
                                       c->an_p4[r] * ipow(cc, 4));
      }
    if (c->an_omega != 0.0)
      for (int jr = 1; jr <= ntm; jr++)
        vp[jr - 1] = vp[jr - 1] + c->an_omega * km_s_to_rs_hr * c->st[jr];
    if (c->an_vtc != 0.0)
      for (int jr = 1; jr <= nt; jr++) vt[jr - 1] = vt[jr - 1] + c->an_vtc * km_s_to_rs_hr;
    // alphas
    double *at = &c->at1[(size_t)r * ntm];
    double *ap = &c->ap1[(size_t)r * ntm];
    auto av = [&](int jf) { return std::fabs(vt[jf - 1]); };
    for (int jr = 1; jr <= nt - 1; jr++) {
      int lo, hi;
      if (jr == 1) { lo = 1; hi = 4; }
      else if (jr == 2) { lo = 1; hi = 5; }
      else if (jr == nt - 2) { lo = jr - 2; hi = jr + 2; }
      else if (jr == nt - 1) { lo = jr - 2; hi = jr + 1; }
      else { lo = jr - 2; hi = jr + 3; }
      double mm = av(lo);
      for (int q = lo + 1; q <= hi; q++) mm = std::max(mm, av(q));
      at[jr - 1] = mm;
    }
    for (int jr = 2; jr <= ntm - 1; jr++) ap[jr - 1] = std::fabs(vp[jr - 1]);
  }
  auto up = [&](std::vector<double> &h, double **d) -> int {
    if (!*d) {
      int rc = dalloc(d, h.size());
      if (rc) return rc;
    }
    HIPCHK(hipMemcpy(*d, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
    return 0;
  };
  int rc;
  if ((rc = up(c->vt1, &c->d_vt1)) || (rc = up(c->vp1, &c->d_vp1)) || (rc = up(c->at1, &c->d_at1)) ||
      (rc = up(c->ap1, &c->d_ap1)))
    return rc;
  if (c->ff) {
    // the analytic terms alone, added to the 2-D base flows on the device
    c->ff_mf = !c->an_p1.empty();
    c->ff_dr = !c->an_p0.empty();
    c->ff_om = c->an_omega != 0.0;
    std::vector<double> mft, drp, omp;
    if (c->ff_mf) {
      mft.assign((size_t)nr * nt, 0.0);
      for (int r = 0; r < nr; r++)
        for (int jr = 1; jr <= nt; jr++) {
          double cc = c->cth[jr];
          mft[(size_t)r * nt + jr - 1] =
              m_s_to_rs_hr * c->sth[jr] *
              (c->an_p1[r] * cc + c->an_p3[r] * ipow(cc, 3) + c->an_p5[r] * ipow(cc, 5));
        }
      if ((rc = up(mft, &c->d_mft))) return rc;
    }
    if (c->ff_dr) {
      drp.assign((size_t)nr * ntm, 0.0);
      for (int r = 0; r < nr; r++)
        for (int jr = 1; jr <= ntm; jr++) {
          double cc = c->ct[jr];
          drp[(size_t)r * ntm + jr - 1] =
              m_s_to_rs_hr * c->st[jr] *
              (c->an_p0[r] + c->an_p2[r] * ipow(cc, 2) + c->an_p4[r] * ipow(cc, 4));
        }
      if ((rc = up(drp, &c->d_drp))) return rc;
    }
    if (c->ff_om) {
      omp.assign(ntm, 0.0);
      for (int jr = 1; jr <= ntm; jr++) omp[jr - 1] = c->an_omega * km_s_to_rs_hr * c->st[jr];
      if ((rc = up(omp, &c->d_omp))) return rc;
    }
  }
  c->flow_set = true;
  c->flow_gen++;
  return 0;
}

int ensure_flow2d(hft_ctx *c) {
  if (c->d_vt2) return 0;
  c->vt_rstride = (size_t)c->nt * c->pitch;
  int rc;
  if ((rc = dalloc(&c->d_vt2, c->vt_rstride * c->nr)) || (rc = dalloc(&c->d_vp2, c->nfield)) ||
      (rc = dalloc(&c->d_at2, c->nfield)) || (rc = dalloc(&c->d_ap2, c->nfield)))
    return rc;
  HIPCHK(hipMemset(c->d_vt2, 0, c->vt_rstride * c->nr * sizeof(double)));
  HIPCHK(hipMemset(c->d_vp2, 0, c->nfield * sizeof(double)));
  HIPCHK(hipMemset(c->d_at2, 0, c->nfield * sizeof(double)));
  HIPCHK(hipMemset(c->d_ap2, 0, c->nfield * sizeof(double)));
  return 0;
}

int update_flow_2d(hft_ctx *c) {
  int rc = ensure_flow2d(c);
  if (rc) return rc;
  AttenArgs a;
  a.f = c->f; a.vt1 = c->d_vt1; a.vp1 = c->d_vp1; a.vt = c->d_vt2; a.vp = c->d_vp2;
  a.vtf = c->ff ? c->d_vtf : nullptr; a.vpf = c->ff ? c->d_vpf : nullptr;
  a.vtf_rstride = c->ff_vt_rstride; a.vpf_rstride = c->ff_vp_rstride;
  a.mft = c->ff_mf ? c->d_mft : nullptr; a.drp = c->ff_dr ? c->d_drp : nullptr;
  a.omp = c->ff_om ? c->d_omp : nullptr;
  a.vtc = c->an_vtc * km_s_to_rs_hr; a.has_vtc = c->an_vtc != 0.0;
  a.inv = c->atten ? c->d_atten_inv : nullptr;
  a.ntm = c->ntm; a.nt = c->nt; a.nu = c->nu; a.pitch = c->pitch;
  a.rstride = c->rstride; a.vt_rstride = c->vt_rstride;
  dim3 g((c->nu + kBlock - 1) / kBlock, c->nt, c->nr);
  hipLaunchKernelGGL(k_attenuate, g, dim3(kBlock), 0, c->stream, a);
  HIPCHK(hipGetLastError());
  dim3 g2((c->nu + kBlock - 1) / kBlock, c->ntm, c->nr);
  hipLaunchKernelGGL(k_lf_alpha, g2, dim3(kBlock), 0, c->stream, (const double *)c->d_vt2,
                     (const double *)c->d_vp2, c->d_at2, c->d_ap2, c->ntm, c->nt, c->nu, c->pitch,
                     c->rstride, c->vt_rstride);
  HIPCHK(hipGetLastError());
  c->flow2d = true;
  c->flow_gen++;
  return 0;
}

// ---- RKG2 coefficients: load_sts_rkg2, src/hipft.f90:5399-5461 ----------------
int64_t sts_stages(double dtl, double dte) {
  double s_real = half * (std::sqrt(twentyfive + twentyfour * (dtl / dte)) - three);
  double sc = std::ceil(s_real);
  int64_t s = sc > 9.0e15 ? (int64_t)9.0e15 : (int64_t)sc;
  if (s < 3) s = 3;
  if (s % 2 == 0) s = s + 1;
  return s;
}

// uj, vj, ubj, gj for j = 1..s (vectors of size s+1, element 0 unused)
void sts_coefs(int64_t s, std::vector<double> &uj, std::vector<double> &vj,
               std::vector<double> &ubj, std::vector<double> &gj) {
  uj.assign(s + 1, 0.0); vj.assign(s + 1, 0.0); ubj.assign(s + 1, 0.0); gj.assign(s + 1, 0.0);
  std::vector<double> b(s + 1, 0.0);
  double w = six / (double)((s + 4) * (s - 1));
  b[1] = one / three;
  b[2] = one / fifteen;
  uj[1] = one; vj[1] = -9999.0; ubj[1] = w; gj[1] = -9999.0;
  uj[2] = one / two; vj[2] = -one / ten; ubj[2] = uj[2] * w; gj[2] = zero;
  for (int64_t j = 3; j <= s; j++) {
    double dj = (double)j;
    b[j] = (four * (double)(j - 1) * (double)(j + 4)) /
           (three * dj * (double)(j + 1) * (double)(j + 2) * (double)(j + 3));
    double bj_bjm1 = b[j] / b[j - 1];
    double bj_bjm2 = b[j] / b[j - 2];
    uj[j] = bj_bjm1 * (two + one / dj);
    vj[j] = -bj_bjm2 * (one / dj + one);
    ubj[j] = uj[j] * w;
    gj[j] = (half * dj * (dj + one) * b[j - 1] - one) * ubj[j];
  }
}

// Captured RKG2 super-steps (hipGraph execs) of the context.
void clear_graphs(hft_ctx *c) {
  for (auto &g : c->graphs) (void)hipGraphExecDestroy(g.second);
  c->graphs.clear();
  if (c->sg_exec) (void)hipGraphExecDestroy(c->sg_exec);
  c->sg_exec = nullptr;
  c->sg_key.clear();
}

struct CoefCache {
  int64_t s = -1;
  std::vector<double> u, v, ub, g;
};

// Stage tables of a batched RKG2 super-step: realizations sorted by stage
// count (descending) into ord / d_order, coefficients of stage k of
// realization r at d_coef[k*nr + r], s, cycle dt and ubj(1) per realization.
int upload_rkg2(hft_ctx *c, const std::vector<int> &reals, const std::vector<double> &dt2,
                const std::vector<int64_t> &s, std::vector<int> &ord,
                const std::vector<double *> *W = nullptr) {
  int n = (int)reals.size();
  int nr = c->nr;
  ord = reals;
  std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return s[a] > s[b]; });
  int64_t smax = s[ord[0]];
  size_t need = (size_t)(smax + 1) * nr;
  if (need > c->coef_cap) {
    // captured super-steps hold the old table's address: drop them
    clear_graphs(c);
    if (c->d_coef) HIPCHK(hipFree(c->d_coef));
    c->d_coef = nullptr;
    size_t cap = std::max(need, (size_t)1024);
    int rc = dalloc(&c->d_coef, cap);
    if (rc) return rc;
    c->coef_cap = cap;
  }
  // the previous upload from the pinned buffers must be complete
  HIPCHK(hipEventSynchronize(c->coef_ev));
  if (!W) {
    // the table is built on the device from the stage counts and cycle
    // steps (k_rkg2_coefs: the expressions of sts_coefs), so only those go
    // up, not the (smax + 1) x nr table of which the realizations still
    // cycling use a few columns
    int64_t *hs = (int64_t *)(c->h_small + 2 * nr + 16);
    for (int z = 0; z < n; z++) {
      int r = ord[z];
      c->h_small[z] = r;
      hs[r] = s[r];
      c->h_smalld[r] = dt2[r];                                        // cycle dt
      c->h_smalld[nr + r] = six / (double)((s[r] + 4) * (s[r] - 1));  // sts_ubj(1)
    }
    HIPCHK(hipMemcpyAsync(c->d_order, c->h_small, n * sizeof(int), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->d_s, hs, nr * sizeof(int64_t), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->d_cdt, c->h_smalld, nr * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->d_ub1, c->h_smalld + nr, nr * sizeof(double), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipEventRecord(c->coef_ev, c->stream));
    if (smax >= 2) {
      CoefArgs ca;
      ca.coef = c->d_coef; ca.order = c->d_order; ca.s = c->d_s; ca.cdt = c->d_cdt;
      ca.nr = nr; ca.n = n; ca.kspan = (int)(smax - 1);
      int64_t threads = (int64_t)n * ca.kspan;
      return launch_ex(c, 0, k_rkg2_coefs, (int)((threads + kBlock - 1) / kBlock), kBlock, 0, ca,
                       0.0);
    }
    return 0;
  }
  if (need > c->h_coef_cap) {
    if (c->h_coef) HIPCHK(hipHostFree(c->h_coef));
    size_t cap = std::max(need, (size_t)1024);
    HIPCHK(hipHostMalloc((void **)&c->h_coef, cap * sizeof(StageCoef)));
    c->h_coef_cap = cap;
  }
  CoefCache cache;
  for (int z = 0; z < n; z++) {
    int r = ord[z];
    if (cache.s != s[r]) {
      cache.s = s[r];
      sts_coefs(s[r], cache.u, cache.v, cache.ub, cache.g);
    }
    double dtl = dt2[r];
    for (int64_t k = 2; k <= s[r]; k++) {
      StageCoef &cf = c->h_coef[(size_t)k * nr + r];
      cf.a = cache.u[k];
      cf.b = cache.v[k];
      cf.c = one - cache.u[k] - cache.v[k];
      cf.d = cache.ub[k] * dtl;
      cf.g = cache.g[k];
      // stored stage u_{k-1} (adjoint sweeps over a kept forward super-step)
      cf.w = (W && (size_t)(k - 1) < W->size()) ? (*W)[k - 1] : nullptr;
    }
    c->h_smalld[r] = dtl;           // cycle dt
    c->h_smalld[nr + r] = cache.ub[1];
  }
  int64_t *hs = (int64_t *)(c->h_small + 2 * nr + 16);
  for (int z = 0; z < n; z++) {
    c->h_small[z] = ord[z];
    hs[ord[z]] = s[ord[z]];
  }
  HIPCHK(hipMemcpyAsync(c->d_coef, c->h_coef, need * sizeof(StageCoef), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->d_order, c->h_small, n * sizeof(int), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->d_s, hs, nr * sizeof(int64_t), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->d_cdt, c->h_smalld, nr * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->d_ub1, c->h_smalld + nr, nr * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipEventRecord(c->coef_ev, c->stream));
  return 0;
}

// Interior march steps of k_rkg2_fused2 (tuning fused_body): they address
// rows through 32-bit buffer offsets, so the realization's field must stay
// below 2 GiB.  Five stages per launch run the general steps only: their
// interior steps gave values that differ from the general steps on some
// band shapes (512x1024, 32- and 64-row bands), a cause not found.
int fused_body_level(const hft_ctx *c) {
  if (c->fused_T > 4 || c->rstride * sizeof(double) >= (size_t)0x7fff0000) return 0;
  return c->fused_body;
}

bool fused_usable(const hft_ctx *c) {
  return c->fused && c->nu <= 1024 && c->fused_T >= 2 &&
         (c->fused_T <= 4 || (c->fused_T == 5 && c->fused_kernel == 3)) &&
         c->band_rows + 2 * c->fused_T <= kFusedMaxRows;
}

int superstep_launches(hft_ctx *c, int n, const std::vector<int> &ord,
                       const std::vector<int64_t> &s, bool use_af);

// One RKG2 super-step for the listed realizations, each with its own cycle
// time step dt2[r] and stage count s[r] (diffusion_step_sts_cd,
// src/hipft.f90:5163-5242).  use_af: dty0 already holds L(f) from the PTL
// pass (same operator, so identical to recomputing it).
int rkg2_superstep(hft_ctx *c, const std::vector<int> &reals, const std::vector<double> &dt2,
                   const std::vector<int64_t> &s, bool use_af) {
  int n = (int)reals.size();
  if (n == 0) return 0;
  std::vector<int> ord;
  int rc = upload_rkg2(c, reals, dt2, s, ord);
  if (rc) return rc;
  if ((fused_usable(c) || c->persist) && !c->buf2) {
    if ((rc = dalloc(&c->buf2, c->nfield)) || (rc = dalloc(&c->buf3, c->nfield))) return rc;
  }
  // the launch sequence depends only on the stage counts in launch order
  // (the coefficient tables are device data uploaded above)
  if (!c->graphs_on || c->timing || c->stream == nullptr) return superstep_launches(c, n, ord, s, use_af);
  std::vector<int64_t> key;
  key.reserve(2 * n + 1);
  key.push_back(use_af ? 1 : 0);
  for (int z = 0; z < n; z++) {
    key.push_back(ord[z]);
    key.push_back(s[ord[z]]);
  }
  for (auto &g : c->graphs)
    if (g.first == key) {
      c->graph_hits++;
      HIPCHK(hipGraphLaunch(g.second, c->stream));
      return 0;
    }
  if (c->graphs.size() >= 64) clear_graphs(c);
  HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  int rcl = superstep_launches(c, n, ord, s, use_af);
  hipGraph_t graph = nullptr;
  hipError_t e = hipStreamEndCapture(c->stream, &graph);
  if (rcl) {
    if (graph) (void)hipGraphDestroy(graph);
    return rcl;
  }
  if (e != hipSuccess) return fail(HFT_E_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
  hipGraphExec_t exec = nullptr;
  e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (e != hipSuccess) return fail(HFT_E_HIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
  c->graphs.push_back({key, exec});
  c->graph_builds++;
  HIPCHK(hipGraphLaunch(exec, c->stream));
  return 0;
}

// Blocks of k_rkg2_fused2<T> that the device holds at once (blocks per CU
// from the occupancy calculator x CUs), cached per T.
template <int T>
int fused2_capacity(hft_ctx *c, int threads, size_t shmem, int *cap) {
  if (c->fused_cap[T] == 0) {
    int per = 0, cus = 0;
    HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_rkg2_fused2<T>, threads, shmem));
    HIPCHK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device));
    c->fused_cap[T] = std::max(1, per) * std::max(1, cus);
  }
  *cap = c->fused_cap[T];
  return 0;
}

// Band count of a fused launch over `active` realizations when the device
// holds `cap` blocks at once.  A block marches its band's rows plus about 3T
// rows of halo and pipeline fill, one block per slot, so a launch takes
// rounds of cap blocks; a round's time grows with the rows per band and,
// by kShare (about a third; tuning key "fused_band_share" in percent) from
// an idle to a full device, with the blocks sharing the memory system
// (measured on 512 x 1024 bands of 32-128 rows).  Of the band counts
// whose bands have hmin..(kFusedMaxRows - 2T) rows, the one with the least
// modelled time; a launch whose last round is nearly empty (active * nbands
// just above a multiple of cap) moves to fewer, taller bands or to more,
// shorter ones that fill the rounds.
int fused_band_count(int ntm, int T, int active, int cap, double kShare, int hmin) {
  int hmax = kFusedMaxRows - 2 * T;
  int best = (ntm + hmax - 1) / hmax;
  double best_t = 1e300;
  for (int nb = std::max(1, (ntm + hmax - 1) / hmax); nb <= std::max(1, ntm / hmin); nb++) {
    int H = (ntm + nb - 1) / nb;
    if (H > hmax) continue;
    int64_t blocks = (int64_t)active * nb;
    int64_t rounds = (blocks + cap - 1) / cap;
    int64_t last = blocks - (rounds - 1) * cap;
    double per = (double)(H + 3 * T);
    double t = per * ((double)(rounds - 1) * (1.0 + kShare) + 1.0 + kShare * last / cap);
    if (t < best_t) {
      best_t = t;
      best = nb;
    }
  }
  return best;
}

// Band rows of a fused launch over `active` realizations.
int fused_rows(hft_ctx *c, int active, int *H) {
  int T = c->fused_T;
  int h = c->band_rows;
  if (c->fused_target_blocks > 0 &&
      (int64_t)active * ((c->ntm + h - 1) / h) < c->fused_target_blocks) {
    // few realizations left: thinner bands so the launch still fills the GPU
    int64_t t = (int64_t)c->ntm * active / c->fused_target_blocks;
    h = (int)std::max<int64_t>(2, std::min<int64_t>(h, t));
  }
  if (c->fused_bands && c->fused_kernel == 3) {
    int ncol = ((c->nu + 63) / 64) * 64;
    int threads = ((ncol / 2 + 63) / 64) * 64, cap = 0, rc;
    size_t sh2 = (size_t)(T + 1) * 3 * 2 * threads * sizeof(double);
    switch (T) {
      case 2: rc = fused2_capacity<2>(c, threads, sh2, &cap); break;
      case 3: rc = fused2_capacity<3>(c, threads, sh2, &cap); break;
      case 5: rc = fused2_capacity<5>(c, threads, sh2, &cap); break;
      default: rc = fused2_capacity<4>(c, threads, sh2, &cap); break;
    }
    if (rc) return rc;
    int nb = fused_band_count(c->ntm, T, active, cap, 0.01 * c->fused_band_share,
                              c->fused_band_min_rows);
    h = (c->ntm + nb - 1) / nb;
  }
  *H = h;
  return 0;
}

// Launch of k_rkg2_fused2<T> (the LDS row length a compile-time constant
// when it is 1024 and interior steps are on).
int launch_fused2(hft_ctx *c, const FusedArgs &fa, int blocks, double units, double sunits) {
  int T = c->fused_T;
  int threads = ((fa.ncol / 2 + 63) / 64) * 64;
  size_t sh2 = (size_t)(T + 1) * 3 * 2 * threads * sizeof(double);
  if (fa.body && 2 * threads == 1024) {
    switch (T) {
      case 2: return launch_ex(c, kGroupFused, k_rkg2_fused2<2, 1024>, blocks, threads, sh2, fa, units, sunits);
      case 3: return launch_ex(c, kGroupFused, k_rkg2_fused2<3, 1024>, blocks, threads, sh2, fa, units, sunits);
      case 5: return launch_ex(c, kGroupFused, k_rkg2_fused2<5, 1024>, blocks, threads, sh2, fa, units, sunits);
      default: return launch_ex(c, kGroupFused, k_rkg2_fused2<4, 1024>, blocks, threads, sh2, fa, units, sunits);
    }
  }
  switch (T) {
    case 2: return launch_ex(c, kGroupFused, k_rkg2_fused2<2>, blocks, threads, sh2, fa, units, sunits);
    case 3: return launch_ex(c, kGroupFused, k_rkg2_fused2<3>, blocks, threads, sh2, fa, units, sunits);
    case 5: return launch_ex(c, kGroupFused, k_rkg2_fused2<5>, blocks, threads, sh2, fa, units, sunits);
    default: return launch_ex(c, kGroupFused, k_rkg2_fused2<4>, blocks, threads, sh2, fa, units, sunits);
  }
}

// One fused launch: stages k..k+T-1 of the first `active` realizations of
// ord, reading u_{k-1} (pA) and u_{k-2} (pB); the last stage goes to oA (or
// each realization's final stage, if earlier), the one before to oB, and
// stage k+t-1 also to W[t] where W[t] is not null.
int launch_fused(hft_ctx *c, int64_t k, int active, const std::vector<int> &ord,
                 const std::vector<int64_t> &s, const double *pA, const double *pB,
                 const double *dty0, double *oA, double *oB, double *const W[5]) {
  int nr = c->nr, T = c->fused_T;
  FusedArgs fa{};
  fa.m = c->m;
  fa.ncol = ((c->nu + 63) / 64) * 64;
  {
    int rc = fused_rows(c, active, &fa.H);
    if (rc) return rc;
  }
  fa.nbands = (c->ntm + fa.H - 1) / fa.H;
  fa.k0 = (int)k;
  fa.A = pA; fa.B = pB; fa.f = c->f; fa.dty0 = dty0; fa.outA = oA; fa.outB = oB;
  fa.D = c->d_D; fa.coef = c->d_coef; fa.nr_table = nr; fa.s = c->d_s; fa.order = c->d_order;
  fa.rstride = c->rstride;
  fa.cdt = c->d_cdt;
  for (int t = 0; t < 5; t++) fa.W[t] = W[t];
  size_t shmem = (size_t)T * 3 * fa.ncol * sizeof(double);
  // compulsory passes: read u_{k-1}, u_{k-2}, u0 (and dty0, except the
  // two-column kernel, which forms it from u0) and write two stage levels
  // per realization, plus each stored stage; stage by stage the same stages
  // would move 5 passes each
  int nstore = 0;
  for (int t = 0; t < 5; t++) nstore += W[t] ? 1 : 0;
  double units = ((c->fused_kernel == 3 ? 5.0 : 6.0) + nstore) * active, sunits = 0;
  for (int z = 0; z < active; z++) {
    int64_t st = std::min<int64_t>(T, s[ord[z]] - k + 1);
    sunits += 5.0 * st;
  }
  int blocks = active * fa.nbands;
  bool store = W[0] || W[1] || W[2] || W[3] || W[4];
  if (c->fused_kernel == 3) {
    // two phi columns per thread: a multiple of 64 threads, LDS rows of 2 *
    // threads (3-row rings of the T stage levels and of u0); interior steps
    // without row tests when the buffers' byte offsets fit the 32-bit
    // buffer-resource range
    fa.body = fused_body_level(c);
    return launch_fused2(c, fa, blocks, units, sunits);
  }
  if (c->fused_kernel == 2 && !store) {
    switch (T) {
      case 2: return launch_ex(c, kGroupFused, k_rkg2_march<2>, blocks, fa.ncol, shmem, fa, units, sunits);
      case 3: return launch_ex(c, kGroupFused, k_rkg2_march<3>, blocks, fa.ncol, shmem, fa, units, sunits);
      default: return launch_ex(c, kGroupFused, k_rkg2_march<4>, blocks, fa.ncol, shmem, fa, units, sunits);
    }
  }
  switch (T) {
    case 2: return launch_ex(c, kGroupFused, k_rkg2_fused<2>, blocks, fa.ncol, shmem, fa, units, sunits);
    case 3: return launch_ex(c, kGroupFused, k_rkg2_fused<3>, blocks, fa.ncol, shmem, fa, units, sunits);
    default: return launch_ex(c, kGroupFused, k_rkg2_fused<4>, blocks, fa.ncol, shmem, fa, units, sunits);
  }
}

int copy_finals(hft_ctx *c, const std::vector<double *> &final_loc);

bool persist_usable(const hft_ctx *c, int64_t smax) {
  int T = c->fused_T;
  int64_t groups = (smax - 1 + T - 1) / T;
  return c->persist && c->fused && c->fused_kernel == 3 && c->nu <= 1024 && T >= 2 && T <= 4 &&
         c->nr < (1 << kPersistZBits) && groups < (1 << (32 - kPersistZBits - kPersistBBits)) &&
         c->persist_rows >= 1 && c->persist_rows + 2 * T <= kFusedMaxRows &&
         c->persist_rows_long >= 1 && c->persist_rows_long + 2 * T <= kFusedMaxRows;
}

// Stages 2..s of a super-step in one persistent launch (k_rkg2_persist):
// the tile list (group-major, realizations in stage-count order, each
// realization's bands), its per-realization band shapes and cleared
// counters, then the launch; returns where each realization's last stage is.
int superstep_persistent(hft_ctx *c, int n, const std::vector<int> &ord,
                         const std::vector<int64_t> &s, std::vector<double *> &final_loc) {
  int T = c->fused_T, ntm = c->ntm;
  std::vector<int> G(n), H(n), nb(n), off(n + 1, 0);
  int gmax = 0;
  for (int z = 0; z < n; z++) {
    int64_t sz = s[ord[z]];
    G[z] = sz >= 2 ? (int)((sz - 1 + T - 1) / T) : 0;
    gmax = std::max(gmax, G[z]);
    H[z] = z < c->persist_long ? c->persist_rows_long : c->persist_rows;
    H[z] = std::min(H[z], ntm);
    nb[z] = (ntm + H[z] - 1) / H[z];
    if (nb[z] >= (1 << kPersistBBits)) return fail(HFT_E_ARG, "persistent super-step: %d bands", nb[z]);
    off[z + 1] = off[z] + nb[z];
  }
  size_t ntiles = 0;
  for (int z = 0; z < n; z++) ntiles += (size_t)G[z] * nb[z];
  if (ntiles == 0) return 0;
  if (ntiles > (size_t)INT32_MAX) return fail(HFT_E_ARG, "persistent super-step: too many tiles");
  if (!c->ptiles_ev) HIPCHK(hipEventCreateWithFlags(&c->ptiles_ev, hipEventDisableTiming));
  HIPCHK(hipEventSynchronize(c->ptiles_ev));  // the previous upload has been read
  if (ntiles > c->ptiles_cap) {
    if (c->d_ptiles) HIPCHK(hipFree(c->d_ptiles));
    if (c->h_ptiles) HIPCHK(hipHostFree(c->h_ptiles));
    c->d_ptiles = nullptr;
    c->h_ptiles = nullptr;
    size_t cap = std::max(ntiles, (size_t)4096);
    int rc = dalloc(&c->d_ptiles, cap);
    if (rc) return rc;
    HIPCHK(hipHostMalloc((void **)&c->h_ptiles, cap * sizeof(unsigned)));
    c->ptiles_cap = cap;
  }
  size_t nint = 3 * (size_t)n;
  if (nint > c->pint_cap) {
    if (c->d_pint) HIPCHK(hipFree(c->d_pint));
    if (c->h_pint) HIPCHK(hipHostFree(c->h_pint));
    c->d_pint = nullptr;
    c->h_pint = nullptr;
    int rc = dalloc(&c->d_pint, nint);
    if (rc) return rc;
    HIPCHK(hipHostMalloc((void **)&c->h_pint, nint * sizeof(int)));
    c->pint_cap = nint;
  }
  size_t nctl = 2 + (size_t)off[n];
  if (nctl > c->pctl_cap) {
    if (c->d_pctl) HIPCHK(hipFree(c->d_pctl));
    c->d_pctl = nullptr;
    int rc = dalloc(&c->d_pctl, nctl);
    if (rc) return rc;
    HIPCHK(hipMemset(c->d_pctl, 0, nctl * sizeof(int)));
    c->pctl_cap = nctl;
  }
  size_t q = 0;
  for (int g = 0; g < gmax; g++)
    for (int z = 0; z < n; z++)
      if (G[z] > g)
        for (int b = 0; b < nb[z]; b++)
          c->h_ptiles[q++] = (unsigned)z | ((unsigned)b << kPersistZBits) |
                             ((unsigned)g << (kPersistZBits + kPersistBBits));
  for (int z = 0; z < n; z++) {
    c->h_pint[z] = H[z];
    c->h_pint[n + z] = nb[z];
    c->h_pint[2 * n + z] = off[z];
  }
  HIPCHK(hipMemcpyAsync(c->d_ptiles, c->h_ptiles, ntiles * sizeof(unsigned), hipMemcpyHostToDevice,
                        c->stream));
  HIPCHK(hipMemcpyAsync(c->d_pint, c->h_pint, nint * sizeof(int), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipEventRecord(c->ptiles_ev, c->stream));
  // counter and done counters start at zero; the error flag is kept (read
  // and cleared by the diffusion step)
  int *counter = c->d_pctl, *err = c->d_pctl + 1, *done = c->d_pctl + 2;
  hipLaunchKernelGGL(k_zero_ints, dim3(std::max(1, std::min(64, (off[n] + 256) / 256))), dim3(256), 0,
                     c->stream, counter, 1);
  HIPCHK(hipGetLastError());
  hipLaunchKernelGGL(k_zero_ints, dim3(std::max(1, std::min(64, (off[n] + 255) / 256))), dim3(256), 0,
                     c->stream, done, off[n]);
  HIPCHK(hipGetLastError());
  PersistArgs pa{};
  FusedArgs &fa = pa.fa;
  fa.m = c->m;
  fa.ncol = ((c->nu + 63) / 64) * 64;
  fa.nbands = 0;
  fa.H = 0;
  fa.k0 = 2;
  fa.A = fa.B = nullptr;
  fa.outA = fa.outB = nullptr;
  fa.f = c->f; fa.dty0 = c->dty0; fa.D = c->d_D; fa.coef = c->d_coef; fa.nr_table = c->nr;
  fa.s = c->d_s; fa.order = c->d_order; fa.rstride = c->rstride; fa.cdt = c->d_cdt;
  for (int t = 0; t < 5; t++) fa.W[t] = nullptr;
  pa.P[0] = c->buf0; pa.P[1] = c->buf1; pa.P[2] = c->buf2; pa.P[3] = c->buf3;
  pa.tiles = c->d_ptiles;
  pa.zH = c->d_pint; pa.znb = c->d_pint + n; pa.zoff = c->d_pint + 2 * n;
  pa.ntiles = (int)ntiles;
  pa.counter = counter; pa.err = err; pa.done = done;
  if (c->wall_khz == 0) {
    int khz = 0;
    HIPCHK(hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, c->device));
    c->wall_khz = khz > 0 ? khz : 100000;
  }
  pa.wait_ticks = 2000LL * c->wall_khz;  // 2 s: far above any real wait
  int threads = ((fa.ncol / 2 + 63) / 64) * 64;
  size_t sh = (size_t)(T + 1) * 3 * 2 * threads * sizeof(double);
  if (c->persist_cap == 0) {
    int per = 0, cus = 0;
    switch (T) {
      case 2: HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_rkg2_persist<2>, threads, sh)); break;
      case 3: HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_rkg2_persist<3>, threads, sh)); break;
      default: HIPCHK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_rkg2_persist<4>, threads, sh)); break;
    }
    HIPCHK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, c->device));
    c->persist_cap = std::max(1, per) * std::max(1, cus);
  }
  int blocks = (int)std::min<size_t>(ntiles, (size_t)c->persist_cap);
  double units = 0, sunits = 0;
  for (int z = 0; z < n; z++) {
    units += 5.0 * G[z];
    sunits += 5.0 * (double)(s[ord[z]] - 1);
  }
  int rc;
  switch (T) {
    case 2: rc = launch_ex(c, kGroupFused, k_rkg2_persist<2>, blocks, threads, sh, pa, units, sunits); break;
    case 3: rc = launch_ex(c, kGroupFused, k_rkg2_persist<3>, blocks, threads, sh, pa, units, sunits); break;
    default: rc = launch_ex(c, kGroupFused, k_rkg2_persist<4>, blocks, threads, sh, pa, units, sunits); break;
  }
  if (rc) return rc;
  for (int z = 0; z < n; z++) {
    int gl = G[z] - 1;
    if (gl >= 0) final_loc[ord[z]] = (gl == 0 || !(gl & 1)) ? c->buf2 : c->buf0;
  }
  return 0;
}

// Error flag of the persistent super-steps since the last check (a wait
// that ran past its bound); clears it.
int persist_check(hft_ctx *c) {
  if (!c->d_pctl) return 0;
  int err = 0;
  HIPCHK(hipMemcpyAsync(&err, c->d_pctl + 1, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (err) {
    int zero = 0;
    HIPCHK(hipMemcpy(c->d_pctl + 1, &zero, sizeof(int), hipMemcpyHostToDevice));
    return fail(HFT_E_HIP, "persistent RKG2 super-step: a tile waited past its bound");
  }
  return 0;
}

// The launches of one RKG2 super-step (tables already uploaded).
int superstep_launches(hft_ctx *c, int n, const std::vector<int> &ord,
                       const std::vector<int64_t> &s, bool use_af) {
  int nr = c->nr;
  int64_t smax = s[ord[0]];
  Stage1Args a1;
  a1.m = c->m; a1.g = tiles_for(c, n); a1.f = c->f; a1.dty0 = c->dty0; a1.u1 = c->buf1;
  a1.D = c->d_D; a1.dt = c->d_cdt; a1.ub1 = c->d_ub1; a1.order = c->d_order;
  a1.Af = use_af ? c->dty0 : nullptr; a1.rstride = c->rstride;
  int rc = launch(c, 0, k_rkg2_stage1, nblocks(a1.g), a1, 3.0 * n);
  if (rc) return rc;
  if (persist_usable(c, smax) && c->buf2) {
    std::vector<double *> fl(nr, nullptr);
    if ((rc = superstep_persistent(c, n, ord, s, fl))) return rc;
    return copy_finals(c, fl);
  }
  // Stage buffers: pA holds u_{k-1}, pB u_{k-2}.  Single-stage launches
  // overwrite u_{k-2} in place (pointwise); fused launches write two fresh
  // buffers.  u_s of a realization that finishes in a fused launch is left
  // in that launch's outA and copied into f afterwards.
  bool fused_ok = fused_usable(c) && c->buf2;
  double *pA = c->buf1, *pB = c->f;
  std::vector<double *> final_loc(nr, nullptr);
  auto pick_free = [&](double *x, double *y) {
    double *cand[4] = {c->buf0, c->buf1, c->buf2, c->buf3};
    for (double *q : cand)
      if (q && q != x && q != y && q != pA && q != pB) return q;
    return (double *)nullptr;
  };
  StageArgs a;
  a.m = c->m; a.f = c->f; a.dty0 = c->dty0; a.D = c->d_D; a.s = c->d_s; a.order = c->d_order;
  a.rstride = c->rstride;
  int active = n;
  int64_t k = 2;
  while (k <= smax) {
    while (active > 0 && s[ord[active - 1]] < k) active--;
    int T = c->fused_T;
    if (fused_ok && active >= c->fused_min) {
      double *oA = pick_free(nullptr, nullptr);
      double *oB = pick_free(oA, nullptr);
      double *none[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
      rc = launch_fused(c, k, active, ord, s, pA, pB, c->dty0, oA, oB, none);
      if (rc) return rc;
      for (int z = 0; z < active; z++)
        if (s[ord[z]] <= k + T - 1) final_loc[ord[z]] = oA;
      pA = oA;
      pB = oB;
      k += T;
    } else {
      double *out = (pB == c->f) ? pick_free(nullptr, nullptr) : pB;
      a.g = tiles_for(c, active);
      a.k = (int)k;
      a.x = pA;
      a.ukm2 = pB;
      a.out = out;
      a.coef = c->d_coef + (size_t)k * nr;
      rc = launch(c, 0, k_rkg2_stage, nblocks(a.g), a, 5.0 * active);
      if (rc) return rc;
      pB = pA;
      pA = out;
      k += 1;
    }
  }
  return copy_finals(c, final_loc);
}

// Final stages of realizations that finished in fused launches into f
// (final_loc[r] = the buffer holding realization r's last stage, or null).
int copy_finals(hft_ctx *c, const std::vector<double *> &final_loc) {
  int nr = c->nr;
  if (nr <= kCopyMax) {
    CopyFinalsArgs ca;
    ca.dst = c->f;
    ca.rstride = c->rstride;
    ca.n = 0;
    double *bufs[4] = {c->buf0, c->buf1, c->buf2, c->buf3};
    for (int q = 0; q < 4; q++) ca.src[q] = bufs[q];
    for (int r = 0; r < nr; r++) {
      if (!final_loc[r]) continue;
      int q = 0;
      while (q < 4 && bufs[q] != final_loc[r]) q++;
      if (q == 4) return fail(HFT_E_STATE, "final RKG2 stage in an unknown buffer");
      ca.real[ca.n] = (uint16_t)r;
      ca.which[ca.n] = (uint8_t)q;
      ca.n++;
    }
    ca.per = (int)((c->rstride / 2 + kBlock - 1) / kBlock);
    return launch(c, 0, k_copy_finals, ca.n * ca.per, ca, 0.0);
  }
  for (int r = 0; r < nr; r++)
    if (final_loc[r])
      HIPCHK(hipMemcpyAsync(c->f + (size_t)r * c->rstride, final_loc[r] + (size_t)r * c->rstride,
                            c->rstride * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
  return 0;
}

// get_dtime_diffusion_ptl (src/hipft.f90:5077-5156) for a set of
// realizations.  Leaves L(f) in dty0.  res[r] = PTL time step.
int ptl_pass(hft_ctx *c, const std::vector<int> &reals, bool coupled,
             const std::vector<double> &dt_in, std::vector<double> &res) {
  int n = (int)reals.size();
  int nr = c->nr;
  // The previous super-step's upload_rkg2 copies out of the same pinned
  // buffers asynchronously: wait for those copies before overwriting them.
  HIPCHK(hipEventSynchronize(c->coef_ev));
  for (int z = 0; z < n; z++) c->h_small[z] = reals[z];
  unsigned long long *hmax = (unsigned long long *)(c->h_smalld);
  unsigned long long *hmin = hmax + nr;
  unsigned long long hugebits;
  double hv = kHuge;
  memcpy(&hugebits, &hv, 8);
  for (int r = 0; r < nr; r++) {
    hmax[r] = 0ull;
    hmin[r] = hugebits;
  }
  HIPCHK(hipMemcpyAsync(c->d_order, c->h_small, n * sizeof(int), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->d_absmax, hmax, nr * 8, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->d_dtmin, hmin, nr * 8, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipEventRecord(c->coef_ev, c->stream));
  LapArgs la;
  la.m = c->m; la.g = tiles_for(c, n); la.x = c->f; la.y = c->dty0; la.D = c->d_D;
  la.order = c->d_order; la.rstride = c->rstride; la.absmax = c->d_absmax; la.rec = c->d_prec;
  int rc = launch(c, 2, k_laplacian, lap_blocks(la.g), la, 2.0 * n);
  if (rc) return rc;
  PtlArgs pa;
  pa.m = c->m; pa.g = la.g; pa.f = c->f; pa.Af = c->dty0; pa.absmax = c->d_absmax;
  pa.dtmin = c->d_dtmin; pa.order = c->d_order; pa.coupled = coupled ? 1 : 0; pa.nlaunch = n;
  pa.rstride = c->rstride; pa.rec = c->d_prec;
  rc = launch_ex(c, 2, k_ptl_pick, n, 256, 0, pa, 0.0);
  if (rc) return rc;
  std::vector<unsigned long long> mx(nr), mn(nr);
  HIPCHK(hipMemcpyAsync(mx.data(), c->d_absmax, nr * 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipMemcpyAsync(mn.data(), c->d_dtmin, nr * 8, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  auto asd = [](unsigned long long b) {
    double d;
    memcpy(&d, &b, 8);
    return d;
  };
  const double safe = 0.95;
  res.assign(nr, kHuge);
  if (coupled) {
    double ax = 0.0, dmin = kHuge;
    for (int r : reals) ax = std::max(ax, asd(mx[r]));
    for (int r : reals) dmin = std::min(dmin, asd(mn[r]));
    double dtp = kHuge;
    if (ax > zero) {
      dtp = safe * dmin;
      if (dtp < dt_in[reals[0]]) dtp = dt_in[reals[0]];
    }
    for (int r : reals) res[r] = dtp;
  } else {
    for (int r : reals) {
      double dtp = kHuge;
      if (asd(mx[r]) > zero) {
        dtp = safe * asd(mn[r]);
        if (dtp < dt_in[r]) dtp = dt_in[r];
      }
      res[r] = dtp;
    }
  }
  return 0;
}

}  // namespace

// Partial sums of hft_analysis_noalloc: a file-scope device array, so that
// entry point allocates nothing.  Every GPU the library runs on holds its own
// copy (8 MiB) until the process ends; host code must use the per-device
// address from hipGetSymbolAddress, never the array's name.
constexpr size_t kScratchDoubles = (size_t)1 << 20;
static __device__ __align__(256) double scratch[kScratchDoubles];

namespace {
std::mutex g_scratch_mu;
std::vector<double *> g_scratch_ptr;  // per device id

// Address of this device's `scratch`, looked up once per device with that
// device current (called from hft_create).
int init_scratch(hft_ctx *c) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if ((int)g_scratch_ptr.size() <= c->device) g_scratch_ptr.resize(c->device + 1, nullptr);
  if (!g_scratch_ptr[c->device]) {
    void *ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      return fail(HFT_E_HIP, "hipGetSymbolAddress(scratch) on device %d: %s", c->device,
                  hipGetErrorString(e));
    g_scratch_ptr[c->device] = (double *)ptr;
  }
  c->scratch_dev = g_scratch_ptr[c->device];
  return 0;
}
}  // namespace

// ============================================================================
// C ABI
// ============================================================================
int set_adjoint_attributes();

extern "C" {

const char *hft_last_error(void) { return g_err.c_str(); }

int hft_create(int ntm, int npm1, const double *t, const double *p, int nr, int mode,
               hft_ctx **out) {
  if (!out || !t || !p) return fail(HFT_E_ARG, "hft_create: null argument");
  *out = nullptr;
  if (ntm < 5 || npm1 < 6 || nr < 1)
    return fail(HFT_E_ARG, "hft_create: grid too small (ntm=%d npm1=%d nr=%d)", ntm, npm1, nr);
  if (mode != HFT_MODE_COUPLED && mode != HFT_MODE_INDEPENDENT)
    return fail(HFT_E_ARG, "hft_create: bad mode %d", mode);
  // src/hipft.f90:6542-6564: the mesh must cover the sphere
  const double eps = 1.e-5;
  if (std::fabs(t[0]) > eps || std::fabs(t[ntm - 1] - pi) > eps || std::fabs(p[0]) > eps ||
      std::fabs(p[npm1 - 1] - twopi) > eps)
    return fail(HFT_E_ARG, "hft_create: mesh does not cover the sphere (t %g..%g, p %g..%g)",
                t[0], t[ntm - 1], p[0], p[npm1 - 1]);
  hft_ctx *c = new hft_ctx();
  c->ntm = ntm; c->npm1 = npm1; c->npm = npm1 + 1; c->np = npm1 + 1; c->nt = ntm + 1;
  c->nu = npm1 - 1; c->nr = nr; c->mode = mode;
  c->pitch = ((c->nu + 63) / 64) * 64;
  c->rstride = (size_t)ntm * c->pitch;
  c->nfield = c->rstride * nr;
  hipError_t e = hipGetDevice(&c->device);
  if (e != hipSuccess) {
    delete c;
    return fail(HFT_E_HIP, "hipGetDevice: %s", hipGetErrorString(e));
  }
  e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    delete c;
    return fail(HFT_E_HIP, "hipStreamCreate: %s", hipGetErrorString(e));
  }
  build_mesh(c, t, p);
  int rc = setup_device(c);
  if (!rc) rc = init_scratch(c);
  if (rc) {
    std::string msg = g_err;
    hft_destroy(c);
    g_err = msg;
    return rc;
  }
  c->last_s.assign(nr, 0);
  c->last_cycles.assign(nr, 0);
  {
    // the fused RKG2 kernels use T stages x 3 rows x up to 1024 doubles of LDS
    int row = 3 * 1024 * (int)sizeof(double);
    struct {
      const void *fn;
      int bytes;
    } lds[] = {{(const void *)k_rkg2_fused<2>, 2 * row},  {(const void *)k_rkg2_fused<3>, 3 * row},
               {(const void *)k_rkg2_fused<4>, 4 * row},  {(const void *)k_rkg2_march<2>, 2 * row},
               {(const void *)k_rkg2_march<3>, 3 * row},  {(const void *)k_rkg2_march<4>, 4 * row},
               {(const void *)k_rkg2_fused2<2>, 3 * row}, {(const void *)k_rkg2_fused2<3>, 4 * row},
               {(const void *)k_rkg2_fused2<4>, 5 * row}, {(const void *)k_rkg2_fused2<5>, 6 * row},
               {(const void *)k_rkg2_fused2<2, 1024>, 3 * row},
               {(const void *)k_rkg2_fused2<3, 1024>, 4 * row},
               {(const void *)k_rkg2_fused2<4, 1024>, 5 * row},
               {(const void *)k_rkg2_fused2<5, 1024>, 6 * row},
               {(const void *)k_rkg2_persist<2>, 3 * row}, {(const void *)k_rkg2_persist<3>, 4 * row},
               {(const void *)k_rkg2_persist<4>, 5 * row}};
    for (auto &q : lds) {
      hipError_t e2 = hipFuncSetAttribute(q.fn, hipFuncAttributeMaxDynamicSharedMemorySize, q.bytes);
      if (e2 != hipSuccess) rc = fail(HFT_E_HIP, "hipFuncSetAttribute(%d bytes of LDS): %s", q.bytes,
                                      hipGetErrorString(e2));
    }
    if (!rc) rc = set_adjoint_attributes();
    if (rc) {
      std::string msg = g_err;
      hft_destroy(c);
      g_err = msg;
      return rc;
    }
  }
  *out = c;
  return 0;
}

int hft_destroy(hft_ctx *c) {
  if (!c) return 0;
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  for (double *d : c->dmesh) (void)hipFree(d);
  void *ptrs[] = {c->f, c->buf0, c->buf1, c->buf2, c->buf3, c->dty0, c->fval, c->d_vt1, c->d_vp1, c->d_at1,
                  c->d_ap1, c->d_atten_inv, c->d_vt2, c->d_vp2, c->d_at2, c->d_ap2, c->d_D,
                  c->d_coef, c->d_s, c->d_order, c->d_order_id, c->d_cdt, c->d_ub1,
                  c->d_absmax, c->d_dtmin, c->d_target, c->d_part, c->d_hist, c->d_da_t,
                  c->d_sn_t, c->d_cs_t, c->d_tt, c->d_da_p, c->d_cs_p, c->d_sn_p, c->d_pcol,
                  c->d_map, c->d_apar, c->d_ipar, c->d_vtf, c->d_vpf, c->d_mft, c->d_drp,
                  c->d_omp, c->d_src, c->d_rfe[0], c->d_rfe[1], c->d_rnd, c->d_rfe_dati,
                  c->d_rfe_scale};
  for (void *q : ptrs)
    if (q) (void)hipFree(q);
  if (c->h_coef) (void)hipHostFree(c->h_coef);
  if (c->h_small) (void)hipHostFree(c->h_small);
  if (c->h_smalld) (void)hipHostFree(c->h_smalld);
  for (auto &tp : c->pending) {
    (void)hipEventDestroy(tp.a);
    (void)hipEventDestroy(tp.b);
  }
  for (auto e : c->free_ev) (void)hipEventDestroy(e);
  if (c->coef_ev) (void)hipEventDestroy(c->coef_ev);
  for (auto &g : c->graphs) (void)hipGraphExecDestroy(g.second);
  if (c->sg_exec) (void)hipGraphExecDestroy(c->sg_exec);
  for (void *q : {(void *)c->d_plan, (void *)c->d_slots, (void *)c->d_htab, (void *)c->d_dte,
                  (void *)c->d_prec})
    if (q) (void)hipFree(q);
  if (c->h_plan) (void)hipHostFree(c->h_plan);
  if (c->h_hist) (void)hipHostFree(c->h_hist);
  for (auto *vec : {&c->adj_fix, &c->adj_ck, &c->adj_w, &c->adj_st})
    for (double *q : *vec)
      if (q) (void)hipFree(q);
  for (double *q : {c->d_apD, c->d_apA, c->d_gout, c->d_adv_dt})
    if (q) (void)hipFree(q);
  if (c->d_adv_order) (void)hipFree(c->d_adv_order);
  if (c->h_adv_order) (void)hipHostFree(c->h_adv_order);
  if (c->h_adv_dt) (void)hipHostFree(c->h_adv_dt);
  if (c->adv_ev) (void)hipEventDestroy(c->adv_ev);
  if (c->ptiles_ev) (void)hipEventDestroy(c->ptiles_ev);
  for (void *q : {(void *)c->d_ptiles, (void *)c->d_pint, (void *)c->d_pctl})
    if (q) (void)hipFree(q);
  if (c->h_ptiles) (void)hipHostFree(c->h_ptiles);
  if (c->h_pint) (void)hipHostFree(c->h_pint);
  if (c->own_stream) {
    (void)hipStreamSynchronize(c->own_stream);
    (void)hipStreamDestroy(c->own_stream);  // c->stream belongs to the caller
  } else if (c->stream) {
    (void)hipStreamDestroy(c->stream);
  }
  delete c;
  return 0;
}

int hft_dims(const hft_ctx *c, int *ntm, int *npm, int *nr) {
  if (!c) return fail(HFT_E_ARG, "null context");
  if (ntm) *ntm = c->ntm;
  if (npm) *npm = c->npm;
  if (nr) *nr = c->nr;
  return 0;
}

int hft_set_state(hft_ctx *c, const double *f, int layout, double mult_fac) {
  if (!c || !f) return fail(HFT_E_ARG, "hft_set_state: null argument");
  std::vector<double> dev;
  if (layout == HFT_LAYOUT_TP) {
    tp_to_dev(c, f, dev);
  } else if (layout == HFT_LAYOUT_PT) {
    // load_initial_condition ingest, src/hipft.f90:2116-2157, 2171-2173
    int ntm = c->ntm, npm1 = c->npm1, npm = c->npm;
    std::vector<double> tp((size_t)ntm * npm * c->nr);
    auto F = [&](int j, int k, int r) -> double & {
      return tp[(j - 1) + (size_t)ntm * ((k - 1) + (size_t)npm * r)];
    };
    for (int r = 0; r < c->nr; r++) {
      for (int k = 1; k <= npm1; k++)
        for (int j = 1; j <= ntm; j++) F(j, k, r) = f[(k - 1) + (size_t)npm1 * ((j - 1) + (size_t)ntm * r)];
      for (int j = 1; j <= ntm; j++) {
        F(j, 1, r) = half * (F(j, 1, r) + F(j, npm1, r));
        F(j, npm1, r) = F(j, 1, r);
        F(j, npm1 + 1, r) = F(j, 2, r);
      }
      double fn1 = 0.0, fs1 = 0.0;
      for (int k = 2; k <= npm - 1; k++) {
        fn1 = fn1 + F(1, k, r) * c->dp[k];
        fs1 = fs1 + F(ntm, k, r) * c->dp[k];
      }
      for (int k = 1; k <= npm; k++) {
        F(1, k, r) = fn1 * twopi_i;
        F(ntm, k, r) = fs1 * twopi_i;
      }
      for (int k = 1; k <= npm; k++)
        for (int j = 1; j <= ntm; j++) F(j, k, r) = mult_fac * F(j, k, r);
    }
    tp_to_dev(c, tp.data(), dev);
  } else {
    return fail(HFT_E_ARG, "hft_set_state: bad layout %d", layout);
  }
  return put_field(c, c->f, dev);
}

int hft_get_state(hft_ctx *c, double *f, int layout) {
  if (!c || !f) return fail(HFT_E_ARG, "hft_get_state: null argument");
  std::vector<double> dev;
  int rc = get_field(c, c->f, dev);
  if (rc) return rc;
  if (layout == HFT_LAYOUT_TP) {
    dev_to_tp(c, dev, f);
  } else if (layout == HFT_LAYOUT_PT) {
    dev_to_pt(c, dev, f);
  } else {
    return fail(HFT_E_ARG, "hft_get_state: bad layout %d", layout);
  }
  return 0;
}

int hft_set_flow_analytic(hft_ctx *c, const double *p0, const double *p2, const double *p4,
                          const double *p1, const double *p3, const double *p5,
                          double vp_rigid_omega, double vt_const) {
  if (!c) return fail(HFT_E_ARG, "null context");
  int nr = c->nr;
  auto cp = [nr](const double *s, std::vector<double> &d) {
    if (s) d.assign(s, s + nr); else d.clear();
  };
  if ((p0 == nullptr) != (p2 == nullptr) || (p0 == nullptr) != (p4 == nullptr))
    return fail(HFT_E_ARG, "hft_set_flow_analytic: p0,p2,p4 must be all set or all NULL");
  if ((p1 == nullptr) != (p3 == nullptr) || (p1 == nullptr) != (p5 == nullptr))
    return fail(HFT_E_ARG, "hft_set_flow_analytic: p1,p3,p5 must be all set or all NULL");
  cp(p0, c->an_p0); cp(p2, c->an_p2); cp(p4, c->an_p4);
  cp(p1, c->an_p1); cp(p3, c->an_p3); cp(p5, c->an_p5);
  c->an_omega = vp_rigid_omega;
  c->an_vtc = vt_const;
  int rc = rebuild_flow(c);
  if (rc) return rc;
  if (c->atten || c->ff) return update_flow_2d(c);
  return 0;
}

int hft_set_flow_profiles(hft_ctx *c, const double *vt, const double *vp) {
  if (!c) return fail(HFT_E_ARG, "null context");
  if (vt) c->prof_vt.assign(vt, vt + (size_t)c->nr * c->nt); else c->prof_vt.clear();
  if (vp) c->prof_vp.assign(vp, vp + (size_t)c->nr * c->ntm); else c->prof_vp.clear();
  c->ff = false;  // profiles replace 2-D base flows
  c->flow2d = false;
  int rc = rebuild_flow(c);
  if (rc) return rc;
  if (c->atten) return update_flow_2d(c);
  return 0;
}

// Flows from files (add_flow_from_files, src/hipft.f90:3525-3660): 2-D base
// flows in the reference's storage, vt(nt,npm,nrf) and vp(ntm,np,nrf) in code
// units (rs/hr), theta fastest; nrf = nr when per_realization, else 1 (the
// reference's files are shared by all realizations).  The unique phi columns
// k = 2..npm-1 are taken; the seam columns follow the periodic rule
// (set_periodic_bc_3d).  Analytic flows set before or after are added in
// update_flow's order.  vt = vp = NULL removes the base flows.
int hft_set_flow_arrays(hft_ctx *c, const double *vt, const double *vp, int per_realization) {
  if (!c) return fail(HFT_E_ARG, "null context");
  if ((vt == nullptr) != (vp == nullptr))
    return fail(HFT_E_ARG, "hft_set_flow_arrays: vt and vp must be both set or both NULL");
  int rc;
  if (!vt) {
    c->ff = false;
    c->flow2d = false;
    if ((rc = rebuild_flow(c))) return rc;
    return c->atten ? update_flow_2d(c) : 0;
  }
  int nrf = per_realization ? c->nr : 1;
  int nt = c->nt, ntm = c->ntm, npm = c->npm, nu = c->nu, pitch = c->pitch;
  size_t rt = (size_t)nt * pitch, rp = (size_t)ntm * pitch;
  if (!c->d_vtf) {
    // room for per-realization flows, used shared or not
    if ((rc = dalloc(&c->d_vtf, rt * c->nr)) || (rc = dalloc(&c->d_vpf, rp * c->nr))) return rc;
  }
  std::vector<double> ht(rt * nrf, 0.0), hp(rp * nrf, 0.0);
  for (int r = 0; r < nrf; r++)
    for (int q = 0; q < nu; q++) {
      int k = q + 2;  // reference column
      for (int j = 0; j < nt; j++)
        ht[r * rt + (size_t)j * pitch + q] = vt[((size_t)r * npm + (k - 1)) * nt + j];
      for (int j = 0; j < ntm; j++)
        hp[r * rp + (size_t)j * pitch + q] = vp[((size_t)r * npm + (k - 1)) * ntm + j];
    }
  HIPCHK(hipMemcpy(c->d_vtf, ht.data(), ht.size() * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->d_vpf, hp.data(), hp.size() * sizeof(double), hipMemcpyHostToDevice));
  c->ff_vt_rstride = per_realization ? rt : 0;
  c->ff_vp_rstride = per_realization ? rp : 0;
  c->ff = true;
  c->prof_vt.clear();  // 2-D base flows replace profiles
  c->prof_vp.clear();
  if ((rc = rebuild_flow(c))) return rc;
  return update_flow_2d(c);
}

int hft_set_flow_attenuation(hft_ctx *c, const double *value) {
  if (!c) return fail(HFT_E_ARG, "null context");
  if (!value) {
    c->atten = false;
    c->flow2d = false;
    c->flow_gen++;
    return c->ff ? update_flow_2d(c) : 0;
  }
  c->atten_inv.resize(c->nr);
  for (int r = 0; r < c->nr; r++) c->atten_inv[r] = value[r] > zero ? one / value[r] : kHuge;
  if (!c->d_atten_inv) {
    int rc = dalloc(&c->d_atten_inv, c->nr);
    if (rc) return rc;
  }
  HIPCHK(hipMemcpy(c->d_atten_inv, c->atten_inv.data(), c->nr * sizeof(double), hipMemcpyHostToDevice));
  c->atten = true;
  if (!c->flow_set) {
    int rc = rebuild_flow(c);
    if (rc) return rc;
  }
  return update_flow_2d(c);
}

int hft_update_flow(hft_ctx *c) {
  if (!c) return fail(HFT_E_ARG, "null context");
  if (!c->atten) return 0;
  return update_flow_2d(c);
}

int hft_flow_dtmax(hft_ctx *c, int strang, double *dt, double *dt_r) {
  if (!c || !dt) return fail(HFT_E_ARG, "hft_flow_dtmax: null argument");
  if (!c->flow_set) return fail(HFT_E_STATE, "hft_flow_dtmax: no flow set");
  int nr = c->nr, ntm = c->ntm, npm = c->npm;
  std::vector<double> per(nr, kHuge);
  if (c->flow2d) {
    unsigned long long hb;
    double hv = kHuge;
    memcpy(&hb, &hv, 8);
    std::vector<unsigned long long> init(nr, hb);
    HIPCHK(hipMemcpyAsync(c->d_dtmin, init.data(), nr * 8, hipMemcpyHostToDevice, c->stream));
    dim3 g((c->nu + kBlock - 1) / kBlock, ntm - 2, nr);
    hipLaunchKernelGGL(k_flow_dtmax, g, dim3(kBlock), 0, c->stream, (const double *)c->d_vt2,
                       (const double *)c->d_vp2, c->m.dt_i, c->m.st_i, c->m.dp_i, ntm, c->nu,
                       c->pitch, c->rstride, c->vt_rstride, c->d_dtmin);
    HIPCHK(hipGetLastError());
    std::vector<unsigned long long> out(nr);
    HIPCHK(hipMemcpyAsync(out.data(), c->d_dtmin, nr * 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    for (int r = 0; r < nr; r++) memcpy(&per[r], &out[r], 8);
  } else {
    // phi-independent flows: kdotv grows with dp_i(k), so the minimum over k
    // is taken at the largest dp_i (rounding is monotone).
    double dpmax = 0.0;
    for (int k = 2; k <= npm - 1; k++) dpmax = std::max(dpmax, c->dp_i[k]);
    for (int r = 0; r < nr; r++) {
      const double *vt = &c->vt1[(size_t)r * c->nt];
      const double *vp = &c->vp1[(size_t)r * ntm];
      double dtmax = kHuge;
      for (int j = 2; j <= ntm - 1; j++) {
        double a = std::max(std::fabs(vt[j - 1]), std::fabs(vt[j]));
        double b = std::max(std::fabs(vp[j - 1]), std::fabs(vp[j - 1]));
        double kdotv = a * c->dt_i[j] + b * c->st_i[j] * dpmax;
        double deltat = half / std::max(kdotv, kTiny);
        dtmax = std::min(dtmax, deltat);
      }
      per[r] = dtmax;
    }
  }
  double mn = kHuge;
  for (int r = 0; r < nr; r++) {
    double v = safety * per[r];
    if (strang) v = two * v;
    v = two * v;  // SSPRK(4,3), src/hipft.f90:7340-7342
    per[r] = v;
    mn = std::min(mn, v);
  }
  *dt = mn;
  if (dt_r) memcpy(dt_r, per.data(), nr * sizeof(double));
  return 0;
}

int hft_set_diffusion(hft_ctx *c, const double *eta, double factor) {
  if (!c || !eta) return fail(HFT_E_ARG, "hft_set_diffusion: null argument");
  c->epoch++;
  int nr = c->nr, ntm = c->ntm, npm = c->npm;
  c->D.resize(nr);
  c->dt_euler_r.resize(nr);
  for (int r = 0; r < nr; r++) {
    if (!(eta[r] >= 0.0)) return fail(HFT_E_ARG, "hft_set_diffusion: eta[%d] = %g", r, eta[r]);
    c->D[r] = factor * eta[r];
  }
  // Gershgorin bound of the constant-coefficient matrix, evaluated with the
  // coefficient expressions of load_diffusion_matrix (4960-4981) and
  // get_dtime_diffusion_euler (5017-5033).
  double max_all = 0.0;
  for (int r = 0; r < nr; r++) {
    double D = c->D[r], max_eig = 0.0;
    for (int k = 2; k <= npm - 1; k++)
      for (int j = 2; j <= ntm - 1; j++) {
        double c1 = half * (D + D) * c->dph_i[k] * c->dp_i[k] * c->st_i[j] * c->st_i[j];
        double c2 = half * (D + D) * c->dth_i[j] * c->dt_i[j] * c->st_i[j] * c->sth[j];
        double c4 = half * (D + D) * c->dth_i[j + 1] * c->dt_i[j] * c->st_i[j] * c->sth[j + 1];
        double c5 = half * (D + D) * c->dph_i[k + 1] * c->dp_i[k] * c->st_i[j] * c->st_i[j];
        double c3 = -(c1 + c2 + c4 + c5);
        double g = zero;
        g = g + std::fabs(c1);
        g = g + std::fabs(c2);
        g = g + std::fabs(c3);
        g = g + std::fabs(c4);
        g = g + std::fabs(c5);
        max_eig = std::max(g, max_eig);
      }
    c->dt_euler_r[r] = safety * (two / max_eig);
    max_all = std::max(max_all, max_eig);
  }
  c->dt_euler = safety * (two / max_all);
  HIPCHK(hipMemcpy(c->d_D, c->D.data(), nr * sizeof(double), hipMemcpyHostToDevice));
  c->diff_set = true;
  return 0;
}

int hft_dt_euler(hft_ctx *c, double *dt, double *dt_r) {
  if (!c) return fail(HFT_E_ARG, "null context");
  if (!c->diff_set) return fail(HFT_E_STATE, "hft_dt_euler: diffusion not set");
  if (dt) *dt = c->dt_euler;
  if (dt_r) memcpy(dt_r, c->dt_euler_r.data(), c->nr * sizeof(double));
  return 0;
}

// Realizations of one advection launch: the first n of order, each with its
// own sub-step dtr[r] (cdt = cmul * dtr[r]).  Null plan: every realization,
// one dt.
struct AdvPlan {
  const int *order;
  const double *dtr;
  int n;
};

static int advection_stage(hft_ctx *c, const double *x, const double *y, double *out, double ca,
                           double cb, double cdt, bool aop_only, const AdvPlan *plan = nullptr,
                           double cmul = 0.0) {
  int n = plan ? plan->n : c->nr;
  const int *order = plan ? plan->order : c->d_order_id;
  const double *dtr = plan ? plan->dtr : nullptr;
  double u = (y ? 3.0 : 2.0) * n;
  if (c->adv_band && c->nu <= 1024) {
    AdvBandArgs a;
    a.m = c->m;
    a.f1.vt = c->d_vt1; a.f1.at = c->d_at1; a.f1.vp = c->d_vp1; a.f1.ap = c->d_ap1;
    a.f2.vt = c->d_vt2; a.f2.vp = c->d_vp2; a.f2.at = c->d_at2; a.f2.ap = c->d_ap2;
    a.f2.vt_rstride = c->vt_rstride;
    a.x = x; a.y = y; a.out = out; a.ca = ca; a.cb = cb; a.cdt = cdt;
    a.dtr = dtr; a.cmul = cmul;
    a.write_aop = aop_only ? 1 : 0;
    a.H = std::min(c->adv_band_rows, kAdvMaxBandRows);
    a.nbands = (c->ntm + a.H - 1) / a.H;
    a.ncol = ((c->nu + 63) / 64) * 64;
    a.order = order; a.rstride = c->rstride;
    a.skip = c->adv_skip;
    int blocks = n * a.nbands;
    size_t shmem = ((size_t)17 * a.ncol + (size_t)kAdvRowVals * (a.H + 6)) * sizeof(double);
    if (c->flow2d) {
      if (y) return launch_ex(c, 1, k_advection_band<true, true>, blocks, a.ncol, shmem, a, u);
      return launch_ex(c, 1, k_advection_band<true, false>, blocks, a.ncol, shmem, a, u);
    }
    if (y) return launch_ex(c, 1, k_advection_band<false, true>, blocks, a.ncol, shmem, a, u);
    return launch_ex(c, 1, k_advection_band<false, false>, blocks, a.ncol, shmem, a, u);
  }
  AdvArgs a;
  a.m = c->m;
  a.g = tiles_for(c, n);
  a.f1.vt = c->d_vt1; a.f1.at = c->d_at1; a.f1.vp = c->d_vp1; a.f1.ap = c->d_ap1;
  a.f2.vt = c->d_vt2; a.f2.vp = c->d_vp2; a.f2.at = c->d_at2; a.f2.ap = c->d_ap2;
  a.f2.vt_rstride = c->vt_rstride;
  a.x = x; a.y = y; a.out = out; a.ca = ca; a.cb = cb; a.cdt = cdt;
  a.dtr = dtr; a.cmul = cmul;
  a.order = order; a.rstride = c->rstride; a.write_aop = aop_only ? 1 : 0;
  a.skip = c->adv_skip;
  int nb = nblocks(a.g);
  if (c->flow2d) {
    if (y) return launch(c, 1, k_advection_stage<true, true>, nb, a, u);
    return launch(c, 1, k_advection_stage<true, false>, nb, a, u);
  }
  if (y) return launch(c, 1, k_advection_stage<false, true>, nb, a, u);
  return launch(c, 1, k_advection_stage<false, false>, nb, a, u);
}

// CFL sub-steps per realization of an advection step of length dtl
// (independent mode): n_r = ceil(dtl / limit_r), limit_r from the flow CFL
// rule of get_flow_dtmax (src/hipft.f90:7287-7350) for an unsplit step.
static int advection_substeps(hft_ctx *c, double dtl, std::vector<int> &nsub, int &nmax) {
  int nr = c->nr;
  nsub.assign(nr, 1);
  nmax = 1;
  if (c->mode != HFT_MODE_INDEPENDENT || !c->adv_substeps) return 0;
  if (c->adv_lim_gen != c->flow_gen) {
    double mn;
    c->adv_lim.assign(nr, kHuge);
    int rc = hft_flow_dtmax(c, 0, &mn, c->adv_lim.data());
    if (rc) return rc;
    c->adv_lim_gen = c->flow_gen;
  }
  for (int r = 0; r < nr; r++) {
    double q = dtl / c->adv_lim[r];
    int n = q > 1.0 ? (int)std::min(std::ceil(q), 1.0e6) : 1;
    nsub[r] = n;
    nmax = std::max(nmax, n);
  }
  return 0;
}

static int advection_step_impl(hft_ctx *c, double dtl);

int hft_advection_step(hft_ctx *c, double dtl) {
  if (!c) return fail(HFT_E_ARG, "null context");
  if (!c->flow_set) return fail(HFT_E_STATE, "hft_advection_step: no flow set");
  hipEvent_t ev;
  int rc = phase_begin(c, &ev);
  if (rc) return rc;
  return phase_end(c, ev, kPhaseAdvection, advection_step_impl(c, dtl));
}

static int advection_step_impl(hft_ctx *c, double dtl) {
  int rc, nmax;
  std::vector<int> nsub;
  if ((rc = advection_substeps(c, dtl, nsub, nmax))) return rc;
  if (nmax == 1) {
    // advection_step_ssprk43, src/hipft.f90:7551-7574
    if ((rc = advection_stage(c, c->f, nullptr, c->buf0, 0, 0, half * dtl, false))) return rc;
    if ((rc = advection_stage(c, c->buf0, nullptr, c->buf1, 0, 0, half * dtl, false))) return rc;
    if ((rc = advection_stage(c, c->buf1, c->f, c->buf0, two_third, third, sixth * dtl, false)))
      return rc;
    return advection_stage(c, c->buf0, nullptr, c->f, 0, 0, half * dtl, false);
  }
  // realization r takes nsub[r] SSPRK(4,3) steps of dtl / nsub[r]; sub-step
  // i advances the realizations with nsub > i (a prefix of the sorted order)
  int nr = c->nr;
  std::vector<int> ord(nr);
  for (int r = 0; r < nr; r++) ord[r] = r;
  std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return nsub[a] > nsub[b]; });
  if (!c->d_adv_order) {
    int rc2;
    if ((rc2 = dalloc(&c->d_adv_order, nr)) || (rc2 = dalloc(&c->d_adv_dt, nr))) return rc2;
    HIPCHK(hipHostMalloc((void **)&c->h_adv_order, nr * sizeof(int)));
    HIPCHK(hipHostMalloc((void **)&c->h_adv_dt, nr * sizeof(double)));
    HIPCHK(hipEventCreateWithFlags(&c->adv_ev, hipEventDisableTiming));
  } else {
    HIPCHK(hipEventSynchronize(c->adv_ev));  // the previous upload has been read
  }
  for (int z = 0; z < nr; z++) c->h_adv_order[z] = ord[z];
  for (int r = 0; r < nr; r++) c->h_adv_dt[r] = dtl / (double)nsub[r];
  HIPCHK(hipMemcpyAsync(c->d_adv_order, c->h_adv_order, nr * sizeof(int), hipMemcpyHostToDevice,
                        c->stream));
  HIPCHK(hipMemcpyAsync(c->d_adv_dt, c->h_adv_dt, nr * sizeof(double), hipMemcpyHostToDevice,
                        c->stream));
  HIPCHK(hipEventRecord(c->adv_ev, c->stream));
  int active = nr;
  for (int i = 0; i < nmax; i++) {
    while (active > 0 && nsub[ord[active - 1]] <= i) active--;
    AdvPlan plan{c->d_adv_order, c->d_adv_dt, active};
    if ((rc = advection_stage(c, c->f, nullptr, c->buf0, 0, 0, 0, false, &plan, half))) return rc;
    if ((rc = advection_stage(c, c->buf0, nullptr, c->buf1, 0, 0, 0, false, &plan, half))) return rc;
    if ((rc = advection_stage(c, c->buf1, c->f, c->buf0, two_third, third, 0, false, &plan, sixth)))
      return rc;
    if ((rc = advection_stage(c, c->buf0, nullptr, c->f, 0, 0, 0, false, &plan, half))) return rc;
  }
  return 0;
}

// Controller of hft_diffusion_step.  Hooks for the adjoint: pre(round) runs
// before each round of cycles; superstep, when given, replaces the batched
// RKG2 super-step (same arguments as rkg2_superstep).
typedef std::function<int(const std::vector<int> &, const std::vector<double> &,
                          const std::vector<int64_t> &, bool)>
    SuperstepFn;
// Controller state of a step resumed part-way (after the rounds a planned
// step ran on the device): round index to start at (0-based) and, per
// controller, cycle number, done flag, time stepped, last cycle dt and
// diagnostic sub-cycle count.
struct CtlResume {
  int round0 = 0;
  std::vector<int> cyc;
  std::vector<char> done;
  std::vector<double> stepped, dt2;
  std::vector<int64_t> sdiag;
};
static int diffusion_impl(hft_ctx *c, double dtl, int subcycles, hft_cycle_fn cb, void *user,
                          const std::function<int(int)> *pre, const SuperstepFn *superstep,
                          const CtlResume *resume = nullptr);

int hft_diffusion_step(hft_ctx *c, double dtl, int subcycles, hft_cycle_fn cb, void *user) {
  if (!c) return fail(HFT_E_ARG, "null context");
  hipEvent_t ev;
  int rc = phase_begin(c, &ev);
  if (rc) return rc;
  return phase_end(c, ev, kPhaseDiffusion,
                   diffusion_impl(c, dtl, subcycles, cb, user, nullptr, nullptr));
}

static int diffusion_impl(hft_ctx *c, double dtl, int subcycles, hft_cycle_fn cb, void *user,
                          const std::function<int(int)> *pre, const SuperstepFn *superstep,
                          const CtlResume *resume) {
  if (!c) return fail(HFT_E_ARG, "null context");
  if (!c->diff_set) return fail(HFT_E_STATE, "hft_diffusion_step: diffusion not set");
  if (subcycles < 0) return fail(HFT_E_ARG, "hft_diffusion_step: subcycles < 0");
  // diffusion_step, src/hipft.f90:4589-4690.  One controller for the whole
  // batch (coupled) or one per realization (independent); all controllers
  // that still cycle advance together in one batched RKG2 super-step.
  const int subcycles_max = 60;
  int nr = c->nr;
  bool coupled = c->mode == HFT_MODE_COUPLED;
  bool auto_sc = subcycles == 0;
  int ng = coupled ? 1 : nr;
  std::vector<int> cyc(ng, 1);
  std::vector<char> done(ng, 0);
  std::vector<double> stepped(ng, 0.0), dt2(ng, auto_sc ? zero : dtl / (double)subcycles);
  std::vector<int64_t> sdiag(ng, subcycles);
  std::vector<hft_cycle> recs(nr);
  std::vector<double> dt2r(nr), dtin(nr), ptl;
  std::vector<int64_t> sr(nr, 0);
  int round1 = 1;
  if (resume) {
    if ((int)resume->cyc.size() != ng) return fail(HFT_E_STATE, "diffusion resume: %d controllers", ng);
    cyc = resume->cyc;
    done = resume->done;
    stepped = resume->stepped;
    dt2 = resume->dt2;
    sdiag = resume->sdiag;
    round1 = resume->round0 + 1;
  }
  for (int round = round1;; round++) {
    std::vector<int> gs, reals;
    for (int g = 0; g < ng; g++)
      if (!done[g]) gs.push_back(g);
    if (gs.empty()) break;
    if (pre) {
      int rc = (*pre)(round);
      if (rc) return rc;
    }
    for (int g : gs) {
      if (coupled)
        for (int r = 0; r < nr; r++) reals.push_back(r);
      else
        reals.push_back(g);
    }
    if (auto_sc) {
      for (int g : gs) {
        if (coupled)
          for (int r = 0; r < nr; r++) dtin[r] = dt2[0];
        else
          dtin[g] = dt2[g];
      }
      int rc = ptl_pass(c, reals, coupled, dtin, ptl);
      if (rc) return rc;
      for (int g : gs) dt2[g] = ptl[coupled ? 0 : g];
    }
    for (int g : gs) {
      if (auto_sc && stepped[g] + dt2[g] >= dtl) {
        dt2[g] = dtl - stepped[g];
        done[g] = 1;
      }
      if (auto_sc && cyc[g] == subcycles_max - 1) {
        dt2[g] = dtl - stepped[g];
        done[g] = 1;
      }
      if (auto_sc) {
        double q = std::ceil((dtl - stepped[g]) / dt2[g]);
        sdiag[g] = q > 2e9 ? 2000000000LL : (int64_t)q;
      }
      double dte = coupled ? c->dt_euler : c->dt_euler_r[g];
      int64_t s = sts_stages(dt2[g], dte);
      if (coupled)
        for (int r = 0; r < nr; r++) { sr[r] = s; dt2r[r] = dt2[g]; }
      else {
        sr[g] = s;
        dt2r[g] = dt2[g];
      }
    }
    if (c->obs_smax) {
      int64_t sm = 0;
      for (int r : reals) sm = std::max(sm, sr[r]);
      c->obs_smax->push_back(sm);
    }
    int rc = superstep ? (*superstep)(reals, dt2r, sr, auto_sc)
                       : rkg2_superstep(c, reals, dt2r, sr, auto_sc);
    if (rc) return rc;
    for (auto &rr : recs) rr = hft_cycle{0, 0, 0, 0};
    for (int g : gs) {
      stepped[g] = stepped[g] + dt2[g];
      auto put = [&](int r) {
        recs[r].time_stepped = stepped[g];
        recs[r].dt = dt2[g];
        recs[r].sts_s = sr[r];
        recs[r].subcycles = sdiag[g];
        c->last_s[r] = sr[r];
        c->last_cycles[r] = cyc[g];
      };
      if (coupled)
        for (int r = 0; r < nr; r++) put(r);
      else
        put(g);
    }
    if (cb) cb(user, round, recs.data());
    for (int g : gs) {
      if (!auto_sc && cyc[g] == subcycles) done[g] = 1;
      if (!done[g]) cyc[g] += 1;
    }
  }
  return c->persist ? persist_check(c) : 0;
}

int hft_diffusion_stats(hft_ctx *c, int64_t *s, int64_t *cycles) {
  if (!c) return fail(HFT_E_ARG, "null context");
  if (s) memcpy(s, c->last_s.data(), c->nr * sizeof(int64_t));
  if (cycles) memcpy(cycles, c->last_cycles.data(), c->nr * sizeof(int64_t));
  return 0;
}

static bool sources_active(const hft_ctx *c);

static int flux_transport_impl(hft_ctx *c, double dtl, int strang, int subcycles,
                               hft_cycle_fn cb, void *user);

int hft_flux_transport_step(hft_ctx *c, double dtl, int strang, int subcycles, hft_cycle_fn cb,
                            void *user) {
  if (!c) return fail(HFT_E_ARG, "null context");
  hipEvent_t ev;
  int rc = phase_begin(c, &ev);
  if (rc) return rc;
  return phase_end(c, ev, kPhaseStep, flux_transport_impl(c, dtl, strang, subcycles, cb, user));
}

// ---- device-planned steps replayed as one hipGraph --------------------------
// A flux-transport step whose diffusion part is planned on the device
// (k_plan_* kernels) is a fixed launch sequence: the advection stages, then
// per round of sub-cycles the PTL pass, the controller, the coefficient
// table, stage 1, gcap[round] fused RKG2 launches and the final copies, then
// the second advection half-step.  The sequence is captured once per key
// (step length, splitting, sub-cycles, the epoch bumped by setters that
// change what the launches take by value, the device arrays and mesh they
// read, launch capacities) and replayed with one hipGraphLaunch; flow and
// diffusivity values rewritten in place between steps keep the capture
// valid.  The host reads back the controller
// state once per step.  gcap[0] holds every stage count a step can need (a
// cycle is never longer than the step, so s <= sts_stages(dt, dt_euler)).
// By default the sequence holds one round (tuning "step_graph_rounds" = 1):
// a step that needs more rounds is finished by the host controller from
// the device state (bit-identical: same kernels, same arithmetic).  Extra
// rounds usually cycle a few realizations with many stages, where empty
// launches held in reserve cost more than the host controller's launches.
// With step_graph_rounds = 0 the number of rounds and their launch counts
// are learned from the stage counts of earlier steps (fixed sub-cycles:
// every round alike, known in advance).
constexpr int kPlanRoundsMax = 64;
constexpr int kPlanSlotsMax = 4096;

struct PlanLayout {
  size_t stepped, dt2, sdiag, last_s, last_cycles, smax, cyc, done, status, bytes;
};

static PlanLayout plan_layout(int ng, int nr) {
  PlanLayout L;
  size_t o = 0;
  L.stepped = o; o += 8 * (size_t)ng;
  L.dt2 = o; o += 8 * (size_t)ng;
  L.sdiag = o; o += 8 * (size_t)ng;
  L.last_s = o; o += 8 * (size_t)nr;
  L.last_cycles = o; o += 8 * (size_t)nr;
  L.smax = o; o += 8 * (size_t)kPlanRoundsMax;
  L.cyc = o; o += 4 * (size_t)ng;
  L.done = o; o += 4 * (size_t)ng;
  L.status = o; o += 4 * 4;
  L.bytes = (o + 255) & ~(size_t)255;
  return L;
}

static bool sg_eligible(const hft_ctx *c, hft_cycle_fn cb, int subcycles) {
  return c->step_graph && !cb && c->stream && !c->timing && !c->phase_timing && c->diff_set &&
         !sources_active(c) && !c->persist && fused_usable(c) && c->fused_kernel == 3 &&
         c->nr <= 1024 && subcycles >= 0 && subcycles <= kPlanRoundsMax;
}

// Largest RKG2 stage count a cycle of this step can have.
static int64_t sg_stage_bound(const hft_ctx *c, double dtl, int subcycles) {
  double d = subcycles > 0 ? dtl / (double)subcycles : dtl;
  if (c->mode == HFT_MODE_COUPLED) return sts_stages(d, c->dt_euler);
  int64_t b = 0;
  for (int r = 0; r < c->nr; r++) b = std::max(b, sts_stages(d, c->dt_euler_r[r]));
  return b;
}

// Fused launches per round of the captured sequence (false: too many).
static bool sg_capacity(const hft_ctx *c, int subcycles, int64_t sbound, std::vector<int> &gcap) {
  int T = c->fused_T;
  auto G = [&](int64_t s) { return s >= 2 ? (int)((s - 1 + T - 1) / T) : 0; };
  int gb = G(sbound);
  gcap.clear();
  if (subcycles > 0) {
    gcap.assign(subcycles, gb);
  } else {
    int R = 1;
    for (size_t i = 0; i < c->sg_smax.size(); i++)
      if (c->sg_smax[i] > 0) R = std::max(R, (int)i + 1);
    if (c->step_graph_rounds > 0) R = std::min(R, c->step_graph_rounds);
    R = std::min(R, kPlanRoundsMax);
    gcap.push_back(gb);
    for (int i = 1; i < R; i++) {
      int g = i < (int)c->sg_smax.size() ? G(c->sg_smax[i]) : 0;
      gcap.push_back(std::min(gb, g + std::max(1, g / 8)));
    }
  }
  int tot = 0;
  for (int g : gcap) tot += g;
  return gb >= 1 && tot <= kPlanSlotsMax;
}

// Record the largest stage count of each round, starting at round `from`.
static void sg_merge(hft_ctx *c, const std::vector<int64_t> &sm, int from) {
  for (size_t i = 0; i < sm.size(); i++) {
    size_t q = from + i;
    if (q >= (size_t)kPlanRoundsMax) break;
    if (c->sg_smax.size() <= q) c->sg_smax.resize(q + 1, 0);
    c->sg_smax[q] = std::max(c->sg_smax[q], sm[i]);
  }
}

// Device buffers of a planned step, the band-row table and the Euler limits.
static int sg_prepare(hft_ctx *c, int64_t sbound, int gmax, int *maxtiles) {
  int nr = c->nr, ng = c->mode == HFT_MODE_COUPLED ? 1 : nr, rc;
  if (!c->buf2) {
    if ((rc = dalloc(&c->buf2, c->nfield)) || (rc = dalloc(&c->buf3, c->nfield))) return rc;
  }
  size_t need = (size_t)(sbound + 1) * nr;
  if (need > c->coef_cap) {
    clear_graphs(c);
    HIPCHK(hipStreamSynchronize(c->stream));
    if (c->d_coef) HIPCHK(hipFree(c->d_coef));
    c->d_coef = nullptr;
    size_t cap = std::max(need, (size_t)1024);
    if ((rc = dalloc(&c->d_coef, cap))) return rc;
    c->coef_cap = cap;
  }
  // the host controller that finishes an overflowing step fills the same
  // table from pinned memory: size that buffer now, so no later round of a
  // step allocates pinned memory (a stage count never exceeds sbound)
  if (need > c->h_coef_cap) {
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipEventSynchronize(c->coef_ev));
    if (c->h_coef) HIPCHK(hipHostFree(c->h_coef));
    c->h_coef = nullptr;
    size_t cap = std::max(need, (size_t)1024);
    HIPCHK(hipHostMalloc((void **)&c->h_coef, cap * sizeof(StageCoef)));
    c->h_coef_cap = cap;
  }
  PlanLayout L = plan_layout(ng, nr);
  if (!c->d_plan) {
    if ((rc = dalloc(&c->d_plan, L.bytes))) return rc;
    HIPCHK(hipMemset(c->d_plan, 0, L.bytes));
    HIPCHK(hipHostMalloc((void **)&c->h_plan, L.bytes));
    c->plan_bytes = L.bytes;
    if ((rc = dalloc(&c->d_htab, (size_t)nr + 1)) || (rc = dalloc(&c->d_dte, (size_t)ng))) return rc;
  }
  if ((size_t)gmax * 3 > c->slots_cap) {
    clear_graphs(c);
    HIPCHK(hipStreamSynchronize(c->stream));
    if (c->d_slots) HIPCHK(hipFree(c->d_slots));
    c->d_slots = nullptr;
    size_t cap = std::max((size_t)gmax * 3, (size_t)384);
    if ((rc = dalloc(&c->d_slots, cap))) return rc;
    c->slots_cap = cap;
  }
  std::vector<int> htab(nr + 1, 1);
  *maxtiles = 1;
  for (int a = 1; a <= nr; a++) {
    if ((rc = fused_rows(c, a, &htab[a]))) return rc;
    int nb = (c->ntm + htab[a] - 1) / htab[a];
    *maxtiles = std::max(*maxtiles, a * nb);
  }
  std::vector<double> dte(ng);
  for (int g = 0; g < ng; g++) dte[g] = ng == 1 && c->mode == HFT_MODE_COUPLED ? c->dt_euler : c->dt_euler_r[g];
  HIPCHK(hipStreamSynchronize(c->stream));
  HIPCHK(hipMemcpy(c->d_htab, htab.data(), (nr + 1) * sizeof(int), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->d_dte, dte.data(), ng * sizeof(double), hipMemcpyHostToDevice));
  return 0;
}

static PlanArgs sg_args(hft_ctx *c, double dtl, int subcycles, int64_t sbound) {
  int nr = c->nr;
  bool coupled = c->mode == HFT_MODE_COUPLED;
  int ng = coupled ? 1 : nr;
  PlanLayout L = plan_layout(ng, nr);
  char *b = c->d_plan;
  PlanArgs p{};
  p.stepped = (double *)(b + L.stepped);
  p.dt2 = (double *)(b + L.dt2);
  p.sdiag = (int64_t *)(b + L.sdiag);
  p.cyc = (int *)(b + L.cyc);
  p.done = (int *)(b + L.done);
  p.status = (int *)(b + L.status);
  p.smax = (int64_t *)(b + L.smax);
  p.last_s = (int64_t *)(b + L.last_s);
  p.last_cycles = (int64_t *)(b + L.last_cycles);
  p.order = c->d_order;
  p.s = c->d_s;
  p.cdt = c->d_cdt;
  p.ub1 = c->d_ub1;
  p.coef = c->d_coef;
  p.slots = c->d_slots;
  p.htab = c->d_htab;
  p.dte = c->d_dte;
  p.absmax = c->d_absmax;
  p.dtmin = c->d_dtmin;
  p.dtl = dtl;
  p.dt2_fixed = subcycles > 0 ? dtl / (double)subcycles : 0.0;
  p.subcycles = subcycles;
  p.coupled = coupled ? 1 : 0;
  p.nr = nr;
  p.ng = ng;
  p.T = c->fused_T;
  p.ntm = c->ntm;
  p.kspan = (int)(sbound - 1);
  return p;
}

// The launch sequence of a planned step (enqueued while the stream captures).
static int sg_enqueue(hft_ctx *c, double dtl, int strang, int subcycles, int64_t sbound,
                      const std::vector<int> &gcap, int maxtiles) {
  int nr = c->nr, T = c->fused_T, rc;
  bool adv = c->flow_set, auto_sc = subcycles == 0;
  double h = strang ? dtl * half : dtl;
  if (adv && (rc = advection_step_impl(c, h))) return rc;
  PlanArgs p = sg_args(c, dtl, subcycles, sbound);
  int *status = p.status;
  if ((rc = launch_ex(c, 2, k_plan_init, 1, 1024, 0, p, 0.0))) return rc;
  Tiles gall = tiles_for(c, nr);
  int ncol = ((c->nu + 63) / 64) * 64;
  int ncoef = nr * p.kspan;
  int per = (int)((c->rstride / 2 + kBlock - 1) / kBlock);
  int R = (int)gcap.size();
  for (int round = 0; round < R; round++) {
    p.round = round;
    p.gcap = gcap[round];
    if (auto_sc) {
      // ptl_pass over the realizations still cycling
      if ((rc = launch_ex(c, 2, k_plan_prep, 1, 1024, 0, p, 0.0))) return rc;
      LapArgs la;
      la.m = c->m; la.g = gall; la.x = c->f; la.y = c->dty0; la.D = c->d_D;
      la.order = c->d_order; la.rstride = c->rstride; la.absmax = c->d_absmax; la.dn = status + 2;
      la.rec = c->d_prec;
      if ((rc = launch(c, 2, k_laplacian, lap_blocks(gall), la, 0.0))) return rc;
      PtlArgs pa;
      pa.m = c->m; pa.g = gall; pa.f = c->f; pa.Af = c->dty0; pa.absmax = c->d_absmax;
      pa.dtmin = c->d_dtmin; pa.order = c->d_order; pa.coupled = p.coupled; pa.nlaunch = nr;
      pa.rstride = c->rstride; pa.dn = status + 2; pa.rec = c->d_prec;
      if ((rc = launch_ex(c, 2, k_ptl_pick, nr, 256, 0, pa, 0.0))) return rc;
    }
    if ((rc = launch_ex(c, 2, k_plan_round, 1, 1024, 0, p, 0.0))) return rc;
    if ((rc = launch_ex(c, 0, k_plan_coefs, (ncoef + kBlock - 1) / kBlock, kBlock, 0, p, 0.0))) return rc;
    Stage1Args a1;
    a1.m = c->m; a1.g = gall; a1.f = c->f; a1.dty0 = c->dty0; a1.u1 = c->buf1;
    a1.D = c->d_D; a1.dt = c->d_cdt; a1.ub1 = c->d_ub1; a1.order = c->d_order;
    a1.Af = auto_sc ? c->dty0 : nullptr; a1.rstride = c->rstride; a1.dn = status + 2;
    if ((rc = launch(c, 0, k_rkg2_stage1, nblocks(gall), a1, 0.0))) return rc;
    // launch q reads u_{k0-1}, u_{k0-2} from the previous launch's outputs
    // (launch 0: u_1 from stage 1 and u_0 = f) and writes buf0/buf2 (q
    // even) or buf1/buf3 (q odd)
    for (int q = 0; q < gcap[round]; q++) {
      FusedArgs fa{};
      fa.m = c->m;
      fa.ncol = ncol;
      fa.H = 1;
      fa.nbands = 1;
      fa.k0 = 2 + q * T;
      fa.A = q == 0 ? c->buf1 : ((q & 1) ? c->buf0 : c->buf1);
      fa.B = q == 0 ? c->f : ((q & 1) ? c->buf2 : c->buf3);
      fa.outA = (q & 1) ? c->buf1 : c->buf0;
      fa.outB = (q & 1) ? c->buf3 : c->buf2;
      fa.f = c->f; fa.dty0 = c->dty0; fa.D = c->d_D; fa.coef = c->d_coef; fa.nr_table = nr;
      fa.s = c->d_s; fa.order = c->d_order; fa.rstride = c->rstride; fa.cdt = c->d_cdt;
      fa.body = fused_body_level(c);
      fa.slot = c->d_slots + 3 * q;
      if ((rc = launch_fused2(c, fa, maxtiles, 0.0, 0.0))) return rc;
    }
    PlanCopyArgs ca;
    ca.dst = c->f; ca.src[0] = c->buf0; ca.src[1] = c->buf1; ca.order = c->d_order;
    ca.n = status + 2; ca.s = c->d_s; ca.rstride = c->rstride; ca.per = per; ca.T = T;
    if ((rc = launch(c, 0, k_copy_finals_plan, nr * per, ca, 0.0))) return rc;
  }
  p.round = R;
  if ((rc = launch_ex(c, 2, k_plan_end, 1, 64, 0, p, 0.0))) return rc;
  if (adv && strang) {
    // the second half-step does nothing when the diffusion step overflowed
    c->adv_skip = status;
    rc = advection_step_impl(c, h);
    c->adv_skip = nullptr;
    if (rc) return rc;
  }
  return 0;
}

// Run the step as a replayed planned sequence when possible (*ran = true).
static int sg_step(hft_ctx *c, double dtl, int strang, int subcycles, bool *ran) {
  *ran = false;
  int nr = c->nr, rc;
  bool coupled = c->mode == HFT_MODE_COUPLED, adv = c->flow_set;
  int ng = coupled ? 1 : nr;
  double h = strang ? dtl * half : dtl;
  if (adv) {
    // per-realization CFL sub-steps vary the advection launches: not planned
    std::vector<int> nsub;
    int nmax = 1;
    if ((rc = advection_substeps(c, h, nsub, nmax))) return rc;
    if (nmax != 1) return 0;
  }
  int64_t dbits;
  memcpy(&dbits, &dtl, sizeof(dbits));
  std::vector<int64_t> base = {dbits, strang, subcycles, c->epoch, adv ? 1 : 0, c->flow2d ? 1 : 0};
  // device arrays and mesh the captured launches read (flow and diffusivity
  // values are rewritten in place between steps, their arrays stay)
  for (const void *q : {(const void *)c->f, (const void *)c->buf0, (const void *)c->buf1,
                        (const void *)c->dty0, (const void *)c->d_D, (const void *)c->d_vt1,
                        (const void *)c->d_vp1, (const void *)c->d_at1, (const void *)c->d_ap1,
                        (const void *)c->d_vt2, (const void *)c->d_vp2, (const void *)c->d_at2,
                        (const void *)c->d_ap2, (const void *)c->d_order_id})
    base.push_back((int64_t)(intptr_t)q);
  {
    uint64_t hsh = 1469598103934665603ull;  // FNV-1a over the mesh descriptor
    const unsigned char *mb = reinterpret_cast<const unsigned char *>(&c->m);
    for (size_t q = 0; q < sizeof(Mesh); q++) hsh = (hsh ^ mb[q]) * 1099511628211ull;
    base.push_back((int64_t)hsh);
  }
  // with more than one round captured, the rounds' stage counts of a new
  // kind of step are learned from a host-controlled step first
  bool learn = subcycles == 0 && c->step_graph_rounds != 1;
  if (base != c->sg_prev) {
    c->sg_prev = base;
    c->sg_smax.clear();
    if (learn) return 0;
  }
  if (learn && c->sg_smax.empty()) return 0;
  int64_t sbound = sg_stage_bound(c, dtl, subcycles);
  std::vector<int> gcap;
  if (!sg_capacity(c, subcycles, sbound, gcap)) return 0;
  std::vector<int64_t> key = base;
  key.push_back(sbound);
  key.push_back(c->fused_T);
  for (int g : gcap) key.push_back(g);
  if (!c->sg_exec || key != c->sg_key) {
    int gmax = *std::max_element(gcap.begin(), gcap.end()), maxtiles = 1;
    if ((rc = sg_prepare(c, sbound, gmax, &maxtiles))) return rc;
    clear_graphs(c);
    HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    int rcl = sg_enqueue(c, dtl, strang, subcycles, sbound, gcap, maxtiles);
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamEndCapture(c->stream, &graph);
    c->adv_skip = nullptr;
    if (rcl) {
      if (graph) (void)hipGraphDestroy(graph);
      return rcl;
    }
    if (e != hipSuccess) return fail(HFT_E_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
    hipGraphExec_t exec = nullptr;
    e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) return fail(HFT_E_HIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
    c->sg_exec = exec;
    c->sg_key = key;
    c->sg_builds++;
  } else {
    c->sg_hits++;
  }
  HIPCHK(hipGraphLaunch(c->sg_exec, c->stream));
  *ran = true;
  PlanLayout L = plan_layout(ng, nr);
  HIPCHK(hipMemcpyAsync(c->h_plan, c->d_plan, L.bytes, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  const char *hb = c->h_plan;
  const int *st = (const int *)(hb + L.status);
  memcpy(c->last_s.data(), hb + L.last_s, nr * sizeof(int64_t));
  memcpy(c->last_cycles.data(), hb + L.last_cycles, nr * sizeof(int64_t));
  int R = (int)gcap.size(), ovf = st[0], c0 = st[1];
  int nread = ovf ? std::min(c0 + 1, R) : R;
  const int64_t *sm = (const int64_t *)(hb + L.smax);
  sg_merge(c, std::vector<int64_t>(sm, sm + nread), 0);
  if (!ovf) return 0;
  // finish the step on the host from the controller state the device kept
  c->sg_overflows++;
  if (c0 < 0 || c0 > R) return fail(HFT_E_STATE, "planned step: bad resume round %d", c0);
  CtlResume res;
  res.round0 = c0;
  const double *stp = (const double *)(hb + L.stepped), *d2 = (const double *)(hb + L.dt2);
  const int64_t *sd = (const int64_t *)(hb + L.sdiag);
  const int *cy = (const int *)(hb + L.cyc), *dn = (const int *)(hb + L.done);
  res.stepped.assign(stp, stp + ng);
  res.dt2.assign(d2, d2 + ng);
  res.sdiag.assign(sd, sd + ng);
  res.cyc.assign(cy, cy + ng);
  res.done.resize(ng);
  for (int g = 0; g < ng; g++) res.done[g] = dn[g] ? 1 : 0;
  std::vector<int64_t> obs;
  c->obs_smax = &obs;
  rc = diffusion_impl(c, dtl, subcycles, nullptr, nullptr, nullptr, nullptr, &res);
  c->obs_smax = nullptr;
  if (rc) return rc;
  sg_merge(c, obs, c0);
  if (adv && strang && (rc = advection_step_impl(c, h))) return rc;
  return 0;
}

static int flux_transport_host(hft_ctx *c, double dtl, int strang, int subcycles,
                               hft_cycle_fn cb, void *user);

static int flux_transport_impl(hft_ctx *c, double dtl, int strang, int subcycles,
                               hft_cycle_fn cb, void *user) {
  bool learn = false;
  if (c->diff_set && sg_eligible(c, cb, subcycles)) {
    bool ran = false;
    int rc = sg_step(c, dtl, strang, subcycles, &ran);
    if (rc || ran) return rc;
    learn = subcycles == 0 && c->step_graph_rounds != 1;
  }
  std::vector<int64_t> obs;
  if (learn) c->obs_smax = &obs;
  int rc = flux_transport_host(c, dtl, strang, subcycles, cb, user);
  c->obs_smax = nullptr;
  if (!rc && learn) sg_merge(c, obs, 0);
  return rc;
}

static int flux_transport_host(hft_ctx *c, double dtl, int strang, int subcycles,
                               hft_cycle_fn cb, void *user) {
  int rc;
  bool adv = c->flow_set, dif = c->diff_set, src = sources_active(c);
  if (strang) {
    double h = dtl * half;
    if (src && (rc = hft_source_step(c, h))) return rc;
    if (adv && (rc = hft_advection_step(c, h))) return rc;
    if (dif && (rc = hft_diffusion_step(c, dtl, subcycles, cb, user))) return rc;
    if (adv && (rc = hft_advection_step(c, h))) return rc;
    if (src && (rc = hft_source_step(c, h))) return rc;
  } else {
    if (adv && (rc = hft_advection_step(c, dtl))) return rc;
    if (dif && (rc = hft_diffusion_step(c, dtl, subcycles, cb, user))) return rc;
    if (src && (rc = hft_source_step(c, dtl))) return rc;
  }
  return 0;
}

int hft_apply_diffusion_operator(hft_ctx *c, const double *x, double *y) {
  if (!c || !x || !y) return fail(HFT_E_ARG, "null argument");
  if (!c->diff_set) return fail(HFT_E_STATE, "diffusion not set");
  std::vector<double> dev;
  tp_to_dev(c, x, dev);
  int rc = put_field(c, c->buf0, dev);
  if (rc) return rc;
  LapArgs la;
  la.m = c->m; la.g = tiles_for(c, c->nr); la.x = c->buf0; la.y = c->buf1; la.D = c->d_D;
  la.order = c->d_order_id; la.rstride = c->rstride; la.absmax = nullptr;
  if ((rc = launch(c, 2, k_laplacian, lap_blocks(la.g), la, 2.0 * c->nr))) return rc;
  if ((rc = get_field(c, c->buf1, dev))) return rc;
  dev_to_tp(c, dev, y);
  return 0;
}

int hft_apply_advection_operator(hft_ctx *c, const double *x, double *y) {
  if (!c || !x || !y) return fail(HFT_E_ARG, "null argument");
  if (!c->flow_set) return fail(HFT_E_STATE, "no flow set");
  std::vector<double> dev;
  tp_to_dev(c, x, dev);
  int rc = put_field(c, c->buf0, dev);
  if (rc) return rc;
  if ((rc = advection_stage(c, c->buf0, nullptr, c->buf1, 0, 0, 0, true))) return rc;
  if ((rc = get_field(c, c->buf1, dev))) return rc;
  dev_to_tp(c, dev, y);
  return 0;
}

static int assimilate_impl(hft_ctx *c, const double *map_dev, int per_real, int custom,
                           const double *mu_power, const double *mu_limit,
                           const double *lat_limit, double mult, int add_mode, int balance) {
  int nr = c->nr, ntm = c->ntm;
  // lat-limit indices (load_data_assimilation, src/hipft.f90:4066-4079)
  std::vector<double> ap(nr * 2);
  std::vector<int> ip(nr * 3);
  for (int r = 0; r < nr; r++) {
    ap[r] = mu_power ? mu_power[r] : 4.0;
    ap[nr + r] = mu_limit ? mu_limit[r] : 0.1;
    double ll = lat_limit ? lat_limit[r] : 0.0;
    int t0 = -1, t1 = -1;
    for (int j = 1; j <= ntm; j++)
      if (c->t[j] >= ll * d2r && t0 == -1) t0 = j;
    for (int j = ntm; j >= 1; j--)
      if (c->t[j] <= (pi - ll * d2r) && t1 == -1) t1 = j;
    ip[r] = t0;
    ip[nr + r] = t1;
    ip[2 * nr + r] = ll > 0 ? 1 : 0;
  }
  HIPCHK(hipMemcpyAsync(c->d_apar, ap.data(), ap.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->d_ipar, ip.data(), ip.size() * sizeof(int), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));  // ap and ip are pageable locals
  AssimArgs a;
  a.m = c->m; a.g = tiles_for(c, nr); a.f = c->f; a.delta = balance ? c->buf0 : nullptr;
  a.map = map_dev; a.map_rstride = per_real ? (size_t)3 * ntm * c->npm1 : 0; a.npm1 = c->npm1;
  a.custom = custom; a.mu_power = c->d_apar; a.mu_limit = c->d_apar + nr;
  a.tidx0 = c->d_ipar; a.tidx1 = c->d_ipar + nr; a.lat_on = c->d_ipar + 2 * nr;
  a.mult = mult; a.add_mode = add_mode; a.order = c->d_order_id; a.rstride = c->rstride;
  int rc = launch(c, 4, k_assimilate, nblocks(a.g), a, 4.0 * nr);
  if (rc) return rc;
  if (balance) {
    // balance_flux on the increment (src/hipft.f90:4129-4134, 8752-8801)
    AnaArgs fa{};
    fa.f = c->buf0; fa.ntm = ntm; fa.nu = c->nu; fa.pitch = c->pitch; fa.nr = nr;
    fa.rstride = c->rstride; fa.da_t = c->d_da_t; fa.da_p = c->d_da_p; fa.pcol = c->d_pcol;
    fa.rows_per_block = c->ana_rows; fa.part = c->d_part; fa.nblk = c->ana_nblk;
    hipLaunchKernelGGL(k_flux_part, dim3(c->ana_nblk, nr), dim3(kBlock), 0, c->stream, fa);
    HIPCHK(hipGetLastError());
    std::vector<double> part((size_t)nr * c->ana_nblk * 2);
    HIPCHK(hipMemcpyAsync(part.data(), c->d_part, part.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    std::vector<double> ratio(nr);
    for (int r = 0; r < nr; r++) {
      double fp = 0, fm = 0;
      for (int b = 0; b < c->ana_nblk; b++) {
        fp += part[((size_t)r * c->ana_nblk + b) * 2];
        fm += part[((size_t)r * c->ana_nblk + b) * 2 + 1];
      }
      ratio[r] = (fm + fp == zero) ? 0.0 : std::sqrt(std::fabs(fp / fm));
    }
    HIPCHK(hipMemcpyAsync(c->d_apar, ratio.data(), nr * sizeof(double), hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(k_apply_balanced, dim3(256, nr), dim3(kBlock), 0, c->stream, c->f,
                       (const double *)c->buf0, (const double *)c->d_apar, ntm, c->nu, c->pitch,
                       c->rstride);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->stream));
  }
  return 0;
}

int hft_assimilate(hft_ctx *c, const double *map, int per_real, int custom, const double *mu_power,
                   const double *mu_limit, const double *lat_limit, double mult, int add_mode,
                   int balance) {
  if (!c || !map) return fail(HFT_E_ARG, "hft_assimilate: null argument");
  size_t n = (size_t)3 * c->ntm * c->npm1 * (per_real ? c->nr : 1);
  if (n > c->map_cap) {
    if (c->d_map) HIPCHK(hipFree(c->d_map));
    c->d_map = nullptr;
    int rc = dalloc(&c->d_map, n);
    if (rc) return rc;
    c->map_cap = n;
  }
  HIPCHK(hipMemcpyAsync(c->d_map, map, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
  int rc = assimilate_impl(c, c->d_map, per_real, custom, mu_power, mu_limit, lat_limit, mult,
                           add_mode, balance);
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

// ---- sources: load_source (src/hipft.f90:7050-7217), update_source
// (3875-3947), generate_rfe (8809-8921), source_step (4697-4733) ----------
int hft_set_source(hft_ctx *c, const double *src) {
  if (!c) return fail(HFT_E_ARG, "null context");
  if (!src) {
    c->src_on = false;
    return 0;
  }
  int ntm = c->ntm, pitch = c->pitch, rc;
  if (!c->d_src && (rc = dalloc(&c->d_src, (size_t)ntm * pitch))) return rc;
  std::vector<double> h((size_t)ntm * pitch, 0.0);
  for (int k = 2; k <= c->npm - 1; k++)
    for (int j = 0; j < ntm; j++) h[(size_t)j * pitch + (k - 2)] = src[j + (size_t)ntm * (k - 1)];
  HIPCHK(hipMemcpy(c->d_src, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
  c->src_on = true;
  return 0;
}

// Multiplicative flux balance of every realization of a device field in
// place (balance_flux, src/hipft.f90:8752-8801).
static int balance_field(hft_ctx *c, double *fld) {
  int nr = c->nr;
  AnaArgs fa{};
  fa.f = fld; fa.ntm = c->ntm; fa.nu = c->nu; fa.pitch = c->pitch; fa.nr = nr;
  fa.rstride = c->rstride; fa.da_t = c->d_da_t; fa.da_p = c->d_da_p; fa.pcol = c->d_pcol;
  fa.rows_per_block = c->ana_rows; fa.part = c->d_part; fa.nblk = c->ana_nblk;
  hipLaunchKernelGGL(k_flux_part, dim3(c->ana_nblk, nr), dim3(kBlock), 0, c->stream, fa);
  HIPCHK(hipGetLastError());
  std::vector<double> part((size_t)nr * c->ana_nblk * 2);
  HIPCHK(hipMemcpyAsync(part.data(), c->d_part, part.size() * sizeof(double),
                        hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  std::vector<double> ratio(nr);
  for (int r = 0; r < nr; r++) {
    double fp = 0, fm = 0;
    for (int b = 0; b < c->ana_nblk; b++) {
      fp += part[((size_t)r * c->ana_nblk + b) * 2];
      fm += part[((size_t)r * c->ana_nblk + b) * 2 + 1];
    }
    ratio[r] = (fm + fp == zero) ? 0.0 : std::sqrt(std::fabs(fp / fm));
  }
  HIPCHK(hipMemcpy(c->d_apar, ratio.data(), nr * sizeof(double), hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_balance_inplace, dim3(256, nr), dim3(kBlock), 0, c->stream, fld,
                     (const double *)c->d_apar, c->ntm, c->nu, c->pitch, c->rstride);
  HIPCHK(hipGetLastError());
  return 0;
}

int hft_rfe_generate(hft_ctx *c, int slot, const double *rnd, int model,
                     const double *uf_per_hour) {
  if (!c || !rnd || !uf_per_hour) return fail(HFT_E_ARG, "hft_rfe_generate: null argument");
  if (slot != 0 && slot != 1) return fail(HFT_E_ARG, "hft_rfe_generate: slot must be 0 or 1");
  if (model != 1 && model != 2) return fail(HFT_E_ARG, "hft_rfe_generate: model must be 1 or 2");
  int nr = c->nr, ntm = c->ntm, npm = c->npm, nu = c->nu, pitch = c->pitch, rc;
  if (!c->d_rfe[slot] && (rc = dalloc(&c->d_rfe[slot], c->nfield))) return rc;
  if (!c->d_rnd) {
    if ((rc = dalloc(&c->d_rnd, (size_t)ntm * pitch)) || (rc = dalloc(&c->d_rfe_dati, ntm)) ||
        (rc = dalloc(&c->d_rfe_scale, nr)))
      return rc;
  }
  std::vector<double> h((size_t)ntm * pitch, 0.0);
  for (int q = 0; q < nu; q++)
    for (int j = 0; j < ntm; j++) h[(size_t)j * pitch + q] = rnd[(size_t)q * ntm + j];
  // per-realization scale (7173-7194) and the model-1 cell factor (8860-8874)
  std::vector<double> scale(nr), dati(ntm, 0.0);
  for (int r = 0; r < nr; r++) {
    double uf = uf_per_hour[r] * input_flux_fac;
    scale[r] = model == 1 ? uf / (rsun_cm2 * (double)(ntm * (npm - 2)))
                          : uf / (rsun_cm2 * four * pi);
  }
  for (int jr = 1; jr <= ntm; jr++) {
    if (jr == 1) {
      double tav = half * (c->t[1] + c->t[2]);
      dati[0] = one / (quarter * std::sin(tav) * c->dth[2]);
    } else if (jr == ntm) {
      double tav = half * (c->t[ntm] + c->t[ntm - 1]);
      dati[ntm - 1] = one / (quarter * std::sin(tav) * c->dth[ntm]);
    } else {
      dati[jr - 1] = c->st_i[jr] * c->dt_i[jr];
    }
  }
  HIPCHK(hipMemcpy(c->d_rnd, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->d_rfe_scale, scale.data(), nr * sizeof(double), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(c->d_rfe_dati, dati.data(), ntm * sizeof(double), hipMemcpyHostToDevice));
  double *fr = c->d_rfe[slot];
  hipLaunchKernelGGL(k_rfe_frame, dim3(256, nr), dim3(kBlock), 0, c->stream, fr,
                     (const double *)c->d_rnd, (const double *)c->d_rfe_scale,
                     (const double *)c->d_rfe_dati, c->m.dp_i, model, ntm, nu, pitch, c->rstride);
  HIPCHK(hipGetLastError());
  hipLaunchKernelGGL(k_rfe_poles, dim3(nr), dim3(64), 0, c->stream, fr, c->m.dp, ntm, nu, pitch,
                     c->rstride, twopi_i);
  HIPCHK(hipGetLastError());
  if ((rc = balance_field(c, fr))) return rc;
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

int hft_rfe_weights(hft_ctx *c, int first_slot, double alpha) {
  if (!c) return fail(HFT_E_ARG, "null context");
  if (first_slot < -1 || first_slot > 1) return fail(HFT_E_ARG, "hft_rfe_weights: slot");
  if (first_slot >= 0 && (!c->d_rfe[0] || !c->d_rfe[1]))
    return fail(HFT_E_STATE, "hft_rfe_weights: both frames must be generated first");
  c->rfe_first = first_slot;
  c->rfe_alpha = alpha;
  return 0;
}

int hft_get_rfe(hft_ctx *c, int slot, double *out_tp) {
  if (!c || !out_tp) return fail(HFT_E_ARG, "null argument");
  if ((slot != 0 && slot != 1) || !c->d_rfe[slot])
    return fail(HFT_E_STATE, "hft_get_rfe: frame %d not generated", slot);
  std::vector<double> dev;
  int rc = get_field(c, c->d_rfe[slot], dev);
  if (rc) return rc;
  dev_to_tp(c, dev, out_tp);
  return 0;
}

static bool sources_active(const hft_ctx *c) { return c->src_on || c->rfe_first >= 0; }

static int source_step_impl(hft_ctx *c, double dtl) {
  const double *r1 = c->rfe_first >= 0 ? c->d_rfe[c->rfe_first] : nullptr;
  const double *r2 = c->rfe_first >= 0 ? c->d_rfe[1 - c->rfe_first] : nullptr;
  hipLaunchKernelGGL(k_source_step, dim3(256, c->nr), dim3(kBlock), 0, c->stream, c->f,
                     (const double *)(c->src_on ? c->d_src : nullptr), r1, r2, c->rfe_alpha, dtl,
                     c->ntm, c->nu, c->pitch, c->rstride);
  HIPCHK(hipGetLastError());
  return 0;
}

int hft_source_step(hft_ctx *c, double dtl) {
  if (!c) return fail(HFT_E_ARG, "null context");
  hipEvent_t ev;
  int rc = phase_begin(c, &ev);
  if (rc) return rc;
  return phase_end(c, ev, kPhaseSource, source_step_impl(c, dtl));
}

int hft_assimilate_device(hft_ctx *c, const double *map_dev, int per_real, int custom,
                          const double *mu_power, const double *mu_limit, const double *lat_limit,
                          double mult, int add_mode, int balance) {
  if (!c || !map_dev) return fail(HFT_E_ARG, "hft_assimilate_device: null argument");
  return assimilate_impl(c, map_dev, per_real, custom, mu_power, mu_limit, lat_limit, mult,
                         add_mode, balance);
}

int hft_set_validation_field(hft_ctx *c, const double *fval_pt) {
  if (!c) return fail(HFT_E_ARG, "null context");
  if (!fval_pt) {
    c->have_fval = false;
    return 0;
  }
  if (!c->fval) {
    int rc = dalloc(&c->fval, c->nfield);
    if (rc) return rc;
  }
  std::vector<double> dev;
  pt_to_dev(c, fval_pt, dev);
  int rc = put_field(c, c->fval, dev);
  if (rc) return rc;
  c->have_fval = true;
  return 0;
}

int hft_analysis(hft_ctx *c, const double *fval_pt, const double *fval_scale, double *out) {
  if (!c || !out) return fail(HFT_E_ARG, "hft_analysis: null argument");
  int nr = c->nr;
  const double *fv = nullptr;
  const double *scale = nullptr;
  if (fval_pt) {
    int rc = hft_set_validation_field(c, fval_pt);
    if (rc) return rc;
    fv = c->fval;
  } else if (fval_scale) {
    if (!c->have_fval) return fail(HFT_E_STATE, "hft_analysis: no validation field set");
    HIPCHK(hipMemcpyAsync(c->d_apar, fval_scale, nr * sizeof(double), hipMemcpyHostToDevice, c->stream));
    fv = c->fval;
    scale = c->d_apar;
  }
  AnaArgs a;
  a.f = c->f; a.fval = fv; a.fval_scale = scale; a.ntm = c->ntm; a.nu = c->nu; a.pitch = c->pitch;
  a.nr = nr; a.rstride = c->rstride; a.da_t = c->d_da_t; a.sn_t = c->d_sn_t; a.cs_t = c->d_cs_t;
  a.tt = c->d_tt; a.da_p = c->d_da_p; a.cs_p = c->d_cs_p; a.sn_p = c->d_sn_p; a.pcol = c->d_pcol;
  a.lat_n = c->pole_lat * d2r;
  a.lat_s = pi - c->pole_lat * d2r;
  a.rows_per_block = c->ana_rows; a.part = c->d_part; a.nblk = c->ana_nblk;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (c->timing) {
    int rc = ev_get(c, &e0);
    if (!rc) rc = ev_get(c, &e1);
    if (rc) return rc;
    HIPCHK(hipEventRecord(e0, c->stream));
  }
  hipLaunchKernelGGL(k_analysis_part, dim3(c->ana_nblk, nr), dim3(kBlock), 0, c->stream, a);
  HIPCHK(hipGetLastError());
  hipLaunchKernelGGL(k_analysis_final, dim3((nr * kNPart + 63) / 64), dim3(64), 0, c->stream,
                     (const double *)c->d_part, c->ana_nblk, nr, c->ntm, c->npm, c->d_hist);
  HIPCHK(hipGetLastError());
  if (c->timing) {
    HIPCHK(hipEventRecord(e1, c->stream));
    c->pending.push_back({e0, e1, 3, 1.0 * nr, 1.0 * nr});
  }
  // through pinned memory: a copy into the caller's pageable array would be
  // staged by the runtime
  HIPCHK(hipMemcpyAsync(c->h_hist, c->d_hist, (size_t)nr * 14 * sizeof(double),
                        hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  memcpy(out, c->h_hist, (size_t)nr * 14 * sizeof(double));
  return 0;
}

int hft_analysis_noalloc(hft_ctx *c, const double *f_dev, int64_t ld, double *out_dev) {
  if (!c || !out_dev) return fail(HFT_E_ARG, "hft_analysis_noalloc: null argument");
  if (!c->scratch_dev) return fail(HFT_E_STATE, "hft_analysis_noalloc: no scratch address");
  int nr = c->nr;
  size_t need = (size_t)nr * c->ana_nblk * kNPart;
  if (need > kScratchDoubles)
    return fail(HFT_E_ARG,
                "hft_analysis_noalloc: %d realizations x %d row blocks x %d partials = %zu "
                "doubles exceed the %zu-double scratch array",
                nr, c->ana_nblk, kNPart, need, kScratchDoubles);
  if (f_dev && ld < c->nu)
    return fail(HFT_E_ARG, "hft_analysis_noalloc: ld %lld < %d unique phi columns", (long long)ld,
                c->nu);
  AnaArgs a;
  a.f = f_dev ? f_dev : c->f;
  a.fval = nullptr;
  a.fval_scale = nullptr;
  a.ntm = c->ntm; a.nu = c->nu; a.nr = nr;
  a.pitch = f_dev ? (int)ld : c->pitch;
  a.rstride = f_dev ? (size_t)c->ntm * (size_t)ld : c->rstride;
  a.da_t = c->d_da_t; a.sn_t = c->d_sn_t; a.cs_t = c->d_cs_t;
  a.tt = c->d_tt; a.da_p = c->d_da_p; a.cs_p = c->d_cs_p; a.sn_p = c->d_sn_p; a.pcol = c->d_pcol;
  a.lat_n = c->pole_lat * d2r;
  a.lat_s = pi - c->pole_lat * d2r;
  a.rows_per_block = c->ana_rows; a.part = c->scratch_dev; a.nblk = c->ana_nblk;
  // every partial is written before it is read: nothing to clear
  hipLaunchKernelGGL(k_analysis_part, dim3(c->ana_nblk, nr), dim3(kBlock), 0, c->stream, a);
  HIPCHK(hipGetLastError());
  hipLaunchKernelGGL(k_analysis_final, dim3((nr * kNPart + 63) / 64), dim3(64), 0, c->stream,
                     (const double *)c->scratch_dev, c->ana_nblk, nr, c->ntm, c->npm, out_dev);
  HIPCHK(hipGetLastError());
  return 0;
}

int hft_set_weno_eps(hft_ctx *c, double eps) {
  if (!c) return fail(HFT_E_ARG, "null context");
  c->epoch++;
  if (!(eps == eps)) return fail(HFT_E_ARG, "hft_set_weno_eps: eps is NaN");
  // load_weno, src/hipft.f90:6895-6901
  for (int j = 1; j <= c->ntm; j++) c->wt[6][j] = eps <= zero ? c->dt[j] : eps;
  for (int k = 1; k <= c->npm; k++) c->wp[6][k] = eps <= zero ? c->dp[k] : eps;
  // launches already queued may still read the device copies
  HIPCHK(hipStreamSynchronize(c->stream));
  HIPCHK(hipMemcpy((void *)c->m.teps, c->wt[6].data() + 1, (c->wt[6].size() - 1) * sizeof(double),
                   hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy((void *)c->m.peps, c->wp[6].data() + 1, (c->wp[6].size() - 1) * sizeof(double),
                   hipMemcpyHostToDevice));
  set_fastdiv(c);
  return 0;
}

int hft_set_pole_flux_lat_limit(hft_ctx *c, double deg) {
  if (!c) return fail(HFT_E_ARG, "null context");
  if (!(deg == deg)) return fail(HFT_E_ARG, "hft_set_pole_flux_lat_limit: NaN");
  c->pole_lat = deg;
  return 0;
}

int hft_balance_flux(hft_ctx *c, double *out_pt) {
  if (!c) return fail(HFT_E_ARG, "null context");
  int rc;
  if (!out_pt) {
    if ((rc = balance_field(c, c->f))) return rc;
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
  }
  HIPCHK(hipMemcpyAsync(c->buf0, c->f, c->nfield * sizeof(double), hipMemcpyDeviceToDevice,
                        c->stream));
  if ((rc = balance_field(c, c->buf0))) return rc;
  std::vector<double> dev;
  if ((rc = get_field(c, c->buf0, dev))) return rc;
  dev_to_pt(c, dev, out_pt);
  return 0;
}

int hft_state_device(hft_ctx *c, double **f_dev, int *pitch) {
  if (!c) return fail(HFT_E_ARG, "null context");
  if (f_dev) *f_dev = c->f;
  if (pitch) *pitch = c->pitch;
  return 0;
}

int hft_synchronize(hft_ctx *c) {
  if (!c) return fail(HFT_E_ARG, "null context");
  HIPCHK(hipStreamSynchronize(c->stream));
  return 0;
}

int hft_stream(hft_ctx *c, void **stream) {
  if (!c || !stream) return fail(HFT_E_ARG, "null argument");
  *stream = (void *)c->stream;
  return 0;
}

int hft_set_tuning(hft_ctx *c, const char *key, int value) {
  if (!c || !key) return fail(HFT_E_ARG, "hft_set_tuning: null argument");
  clear_graphs(c);  // captured launch sequences depend on the tuning
  c->epoch++;
  if (!strcmp(key, "fused")) c->fused = value != 0;
  else if (!strcmp(key, "fused_stages")) {
    if (value < 2 || value > 5) return fail(HFT_E_ARG, "fused_stages must be 2..5 (5: fused_kernel 3 only)");
    c->fused_T = value;
  } else if (!strcmp(key, "band_rows")) {
    if (value < 2 || value > 4096) return fail(HFT_E_ARG, "band_rows must be 2..4096");
    c->band_rows = value;
  } else if (!strcmp(key, "fused_min_realizations")) {
    if (value < 1) return fail(HFT_E_ARG, "fused_min_realizations must be >= 1");
    c->fused_min = value;
  } else if (!strcmp(key, "fused_target_blocks")) {
    if (value < 0) return fail(HFT_E_ARG, "fused_target_blocks must be >= 0");
    c->fused_target_blocks = value;
  } else if (!strcmp(key, "fused_kernel")) {
    if (value < 1 || value > 3) return fail(HFT_E_ARG, "fused_kernel must be 1, 2 or 3");
    c->fused_kernel = value;
  } else if (!strcmp(key, "weno_fastdiv")) {
    if (value < 0 || value > 1) return fail(HFT_E_ARG, "weno_fastdiv must be 0 or 1");
    c->weno_fastdiv = value;
    set_fastdiv(c);
  } else if (!strcmp(key, "fused_body")) {
    if (value < 0 || value > 1) return fail(HFT_E_ARG, "fused_body must be 0 or 1");
    c->fused_body = value;
  } else if (!strcmp(key, "fused_bands")) {
    if (value < 0 || value > 1) return fail(HFT_E_ARG, "fused_bands must be 0 or 1");
    c->fused_bands = value;
  } else if (!strcmp(key, "fused_band_min_rows")) {
    if (value < 2 || value > 64) return fail(HFT_E_ARG, "fused_band_min_rows must be 2..64");
    c->fused_band_min_rows = value;
  } else if (!strcmp(key, "fused_band_share")) {
    if (value < 0 || value > 200) return fail(HFT_E_ARG, "fused_band_share must be 0..200");
    c->fused_band_share = value;
  } else if (!strcmp(key, "persist")) {
    // The persistent super-step (k_rkg2_persist) did not finish on the
    // MI355X in its first runs (a 91 x 181 batch of six realizations), with
    // no cause found; it stays switched off.
    if (value != 0)
      return fail(HFT_E_ARG, "persist: the persistent RKG2 super-step is switched off (it did not "
                             "complete on the GPU; cause not found)");
    c->persist = 0;
  } else if (!strcmp(key, "persist_rows") || !strcmp(key, "persist_rows_long")) {
    if (value < 4 || value > kFusedMaxRows - 10)
      return fail(HFT_E_ARG, "%s must be 4..%d", key, kFusedMaxRows - 10);
    (key[12] == '_' ? c->persist_rows_long : c->persist_rows) = value;
  } else if (!strcmp(key, "persist_long")) {
    if (value < 0) return fail(HFT_E_ARG, "persist_long must be >= 0");
    c->persist_long = value;
  } else if (!strcmp(key, "graphs")) {
    c->graphs_on = value != 0;
  } else if (!strcmp(key, "step_graph")) {
    if (value < 0 || value > 1) return fail(HFT_E_ARG, "step_graph must be 0 or 1");
    c->step_graph = value;
  } else if (!strcmp(key, "step_graph_rounds")) {
    if (value < 0 || value > kPlanRoundsMax)
      return fail(HFT_E_ARG, "step_graph_rounds must be 0..%d", kPlanRoundsMax);
    c->step_graph_rounds = value;
  } else if (!strcmp(key, "adjoint_segment")) {
    if (value < 0) return fail(HFT_E_ARG, "adjoint_segment must be >= 0");
    c->adj_seg = value;
  } else if (!strcmp(key, "adjoint_mem_mb")) {
    if (value < 0) return fail(HFT_E_ARG, "adjoint_mem_mb must be >= 0");
    c->adj_mem_mb = value;
  } else if (!strcmp(key, "adjoint_band_rows")) {
    if (value < 1 || value > 4096) return fail(HFT_E_ARG, "adjoint_band_rows must be 1..4096");
    c->adj_band_rows = value;
  } else if (!strcmp(key, "adjoint_fused")) {
    c->adj_fused = value != 0;
  } else if (!strcmp(key, "adjoint_fused_rows")) {
    if (value < 4 || value > kFusedMaxRows - 8)
      return fail(HFT_E_ARG, "adjoint_fused_rows must be 4..%d", kFusedMaxRows - 8);
    c->adj_fused_rows = value;
  } else if (!strcmp(key, "advection_substeps")) {
    c->adv_substeps = value != 0;
  } else if (!strcmp(key, "advection_band")) {
    c->adv_band = value != 0;
  } else if (!strcmp(key, "advection_band_rows")) {
    if (value < 1 || value > 4096) return fail(HFT_E_ARG, "advection_band_rows must be 1..4096");
    c->adv_band_rows = value;
  } else {
    return fail(HFT_E_ARG, "unknown tuning key '%s'", key);
  }
  return 0;
}

int hft_get_tuning(hft_ctx *c, const char *key, int *value) {
  if (!c || !key || !value) return fail(HFT_E_ARG, "hft_get_tuning: null argument");
  if (!strcmp(key, "fused")) *value = c->fused;
  else if (!strcmp(key, "fused_stages")) *value = c->fused_T;
  else if (!strcmp(key, "band_rows")) *value = c->band_rows;
  else if (!strcmp(key, "fused_min_realizations")) *value = c->fused_min;
  else if (!strcmp(key, "fused_target_blocks")) *value = c->fused_target_blocks;
  else if (!strcmp(key, "fused_bands")) *value = c->fused_bands;
  else if (!strcmp(key, "fused_body")) *value = c->fused_body;
  else if (!strcmp(key, "weno_fastdiv")) *value = c->weno_fastdiv;
  else if (!strcmp(key, "fused_band_share")) *value = c->fused_band_share;
  else if (!strcmp(key, "fused_band_min_rows")) *value = c->fused_band_min_rows;
  else if (!strcmp(key, "fused_kernel")) *value = c->fused_kernel;
  else if (!strcmp(key, "persist")) *value = c->persist;
  else if (!strcmp(key, "persist_rows")) *value = c->persist_rows;
  else if (!strcmp(key, "persist_rows_long")) *value = c->persist_rows_long;
  else if (!strcmp(key, "persist_long")) *value = c->persist_long;
  else if (!strcmp(key, "graphs")) *value = c->graphs_on;
  else if (!strcmp(key, "graph_hits")) *value = (int)c->graph_hits;
  else if (!strcmp(key, "graph_builds")) *value = (int)c->graph_builds;
  else if (!strcmp(key, "step_graph")) *value = c->step_graph;
  else if (!strcmp(key, "step_graph_rounds")) *value = c->step_graph_rounds;
  else if (!strcmp(key, "step_graph_hits")) *value = (int)c->sg_hits;
  else if (!strcmp(key, "step_graph_builds")) *value = (int)c->sg_builds;
  else if (!strcmp(key, "step_graph_overflows")) *value = (int)c->sg_overflows;
  else if (!strcmp(key, "adjoint_segment")) *value = c->adj_seg;
  else if (!strcmp(key, "adjoint_mem_mb")) *value = (int)c->adj_mem_mb;
  else if (!strcmp(key, "adjoint_band_rows")) *value = c->adj_band_rows;
  else if (!strcmp(key, "adjoint_fused")) *value = c->adj_fused;
  else if (!strcmp(key, "adjoint_fused_rows")) *value = c->adj_fused_rows;
  else if (!strcmp(key, "advection_substeps")) *value = c->adv_substeps;
  else if (!strcmp(key, "advection_band")) *value = c->adv_band;
  else if (!strcmp(key, "advection_band_rows")) *value = c->adv_band_rows;
  else return fail(HFT_E_ARG, "unknown tuning key '%s'", key);
  return 0;
}

int hft_set_timing(hft_ctx *c, int enable) {
  if (!c) return fail(HFT_E_ARG, "null context");
  int rc = drain_timing(c);
  if (rc) return rc;
  for (int g = 0; g < kNGroups; g++) {
    c->t_ms[g] = 0;
    c->t_n[g] = 0;
    c->t_units[g] = 0;
    c->t_sunits[g] = 0;
  }
  c->timing = (enable & 1) != 0;
  c->phase_timing = (enable & 2) != 0;
  return 0;
}

int hft_kernel_time(hft_ctx *c, const char *group, double *total_ms, int64_t *launches,
                    double *alg_bytes) {
  if (!c || !group) return fail(HFT_E_ARG, "null argument");
  int rc = drain_timing(c);
  if (rc) return rc;
  for (int g = 0; g < kNGroups; g++)
    if (strcmp(group, kGroups[g]) == 0) {
      if (total_ms) *total_ms = c->t_ms[g];
      if (launches) *launches = c->t_n[g];
      if (alg_bytes) *alg_bytes = c->t_units[g] * 8.0 * (double)c->ntm * (double)c->nu;
      return 0;
    }
  return fail(HFT_E_ARG, "unknown kernel group '%s'", group);
}

int hft_kernel_stage_bytes(hft_ctx *c, const char *group, double *stage_bytes) {
  if (!c || !group || !stage_bytes) return fail(HFT_E_ARG, "null argument");
  int rc = drain_timing(c);
  if (rc) return rc;
  for (int g = 0; g < kNGroups; g++)
    if (strcmp(group, kGroups[g]) == 0) {
      *stage_bytes = c->t_sunits[g] * 8.0 * (double)c->ntm * (double)c->nu;
      return 0;
    }
  return fail(HFT_E_ARG, "unknown kernel group '%s'", group);
}

}  // extern "C"

#include "hft_adjoint.hip"
