// hft_adjoint.hip -- host controller of the discrete adjoint of one
// flux-transport step (part of the libhft translation unit: included by
// hft_lib.hip after the forward controller).
//
// hft_step_adjoint recomputes the step from the context's state, storing what
// the transposed sweeps need (the four SSPRK(4,3) stage inputs of each
// advection step, the input of every diffusion round and the RKG2 stages),
// then runs the sweeps backwards:
//   Strang:  lam <- A'^T lam, D'^T lam, A'^T lam
// The RKG2 recurrence is transposed stage by stage (k = s..1).  With fixed
// sub-cycles whose stages fit in "adjoint_mem_mb" the recomputation keeps
// every stage (one-stage launches); otherwise (PTL sub-cycling, or
// "adjoint_segment" > 0) it keeps each round's input and, in the backward
// sweep, pairs (u_{k0-1}, u_{k0-2}) at segment starts, recomputing each
// segment before its adjoint stages.

namespace {

double *adj_buf(hft_ctx *c, std::vector<double *> &v, size_t i, int *rc) {
  while (v.size() <= i) v.push_back(nullptr);
  if (!v[i]) {
    *rc = dalloc(&v[i], c->nfield);
    if (*rc) return nullptr;
  }
  return v[i];
}

// Fixed adjoint buffers
enum {
  AB_XA = 0,     // 4: stage inputs of the first advection step
  AB_XB = 4,     // 4: stage inputs of the second advection step (Strang)
  AB_T1 = 8,
  AB_T2 = 9,
  AB_LF = 10,
  AB_LAM = 11,   // adjoint state (pitched)
  AB_DTY = 12,   // dty0 of the cycle
  AB_LB = 13,    // 3: rolling RKG2 stage adjoints
  AB_L0 = 16,
  AB_R = 17,     // 3: rolling stage states of the checkpoint pass
  AB_U1 = 20,
  AB_Q3 = 21,    // fourth rotating buffer of the fused sweep
  AB_L1 = 22,    // lam_1 (fused sweep)
  AB_COUNT = 23
};

int fill_dev(hft_ctx *c, double *p, size_t n, double v) {
  if (n == 0) return 0;
  size_t nb = (n + kBlock - 1) / kBlock;
  hipLaunchKernelGGL(k_fill, dim3((unsigned)nb), dim3(kBlock), 0, c->stream, p, n, v);
  HIPCHK(hipGetLastError());
  return 0;
}

int copy_real(hft_ctx *c, double *dst, const double *src, int r) {
  size_t off = (size_t)r * c->rstride;
  HIPCHK(hipMemcpyAsync(dst + off, src + off, c->rstride * sizeof(double),
                        hipMemcpyDeviceToDevice, c->stream));
  return 0;
}

// Forward advection step from f, keeping the stage inputs X[0..3]
// (X[0] = step input); same launches as hft_advection_step.
int advection_store(hft_ctx *c, double dtl, double *X[4]) {
  int rc;
  HIPCHK(hipMemcpyAsync(X[0], c->f, c->nfield * sizeof(double), hipMemcpyDeviceToDevice,
                        c->stream));
  if ((rc = advection_stage(c, X[0], nullptr, X[1], 0, 0, half * dtl, false))) return rc;
  if ((rc = advection_stage(c, X[1], nullptr, X[2], 0, 0, half * dtl, false))) return rc;
  if ((rc = advection_stage(c, X[2], X[0], X[3], two_third, third, sixth * dtl, false))) return rc;
  return advection_stage(c, X[3], nullptr, c->f, 0, 0, half * dtl, false);
}

int vjp_stage(hft_ctx *c, const double *x, const double *lam, const double *P, double *out,
              double *acc, double ca, double cb, double cdt, bool frozen) {
  AdvVjpArgs a;
  a.m = c->m;
  a.f1.vt = c->d_vt1; a.f1.at = c->d_at1; a.f1.vp = c->d_vp1; a.f1.ap = c->d_ap1;
  a.x = x; a.lam = lam; a.P = P; a.out = out; a.acc = acc;
  a.ca = ca; a.cb = cb; a.cdt = cdt;
  a.H = c->adj_band_rows;
  a.nbands = (c->ntm + a.H - 1) / a.H;
  a.ncol = ((c->nu + 63) / 64) * 64;
  a.rstride = c->rstride;
  a.part = c->d_apA;
  int blocks = c->nr * a.nbands;
  size_t shmem = ((size_t)19 * a.ncol + 16) * sizeof(double);
  double u = 4.0 * c->nr;
  // blocks of up to 512 threads get the register budget of two waves per SIMD
  bool b512 = a.ncol <= 512;
#define HFT_VJP(FZ, HP)                                                                         \
  (b512 ? launch_ex(c, kGroupAdjoint, k_advection_vjp_band<FZ, HP, 512>, blocks, a.ncol, shmem, a, u) \
        : launch_ex(c, kGroupAdjoint, k_advection_vjp_band<FZ, HP, 1024>, blocks, a.ncol, shmem, a, u))
  if (frozen) return P ? HFT_VJP(true, true) : HFT_VJP(true, false);
  return P ? HFT_VJP(false, true) : HFT_VJP(false, false);
#undef HFT_VJP
}

// Adjoint of one SSPRK(4,3) step (forward: X1 = X0 - h W X0, X2 = X1 - h W X1,
// X3 = 2/3 X0 + 1/3 X2 - dtl/6 W X2, out = X3 - h W X3), lam in place.
int advection_adjoint(hft_ctx *c, double dtl, double *X[4], double *lam, bool frozen) {
  int rc;
  double h = half * dtl;
  double *T1 = adj_buf(c, c->adj_fix, AB_T1, &rc);
  if (!T1) return rc;
  double *T2 = adj_buf(c, c->adj_fix, AB_T2, &rc);
  if (!T2) return rc;
  double *LF = adj_buf(c, c->adj_fix, AB_LF, &rc);
  if (!LF) return rc;
  if ((rc = vjp_stage(c, X[3], lam, nullptr, T1, nullptr, 0, one, h, frozen))) return rc;
  if ((rc = vjp_stage(c, X[2], T1, nullptr, T2, LF, two_third, third, sixth * dtl, frozen))) return rc;
  if ((rc = vjp_stage(c, X[1], T2, nullptr, T1, nullptr, 0, one, h, frozen))) return rc;
  return vjp_stage(c, X[0], T1, LF, lam, nullptr, 0, one, h, frozen);
}

struct CycleRec {
  std::vector<int> reals;
  std::vector<double> dt;     // [nr]
  std::vector<int64_t> s;     // [nr]
  std::vector<double *> W;    // stored stages W[k] = u_k (k >= 1); empty: recompute
};

// Forward RKG2 super-step that keeps every stage u_1..u_{s-1} of the round
// in rec.W (the realizations' final stages go into the state as usual).
int superstep_store(hft_ctx *c, const std::vector<int> &reals, const std::vector<double> &dt2,
                    const std::vector<int64_t> &s, bool use_af, CycleRec &rec, size_t &next) {
  int nr = c->nr, rc;
  std::vector<int> ord;
  if ((rc = upload_rkg2(c, reals, dt2, s, ord))) return rc;
  int n = (int)ord.size();
  int64_t smax = s[ord[0]];
  double *DTY = adj_buf(c, c->adj_fix, AB_DTY, &rc);
  if (!DTY) return rc;
  rec.W.assign(smax, nullptr);
  for (int64_t k = 1; k < smax; k++) {
    rec.W[k] = adj_buf(c, c->adj_st, next++, &rc);
    if (!rec.W[k]) return rc;
  }
  Stage1Args a1;
  a1.m = c->m; a1.g = tiles_for(c, n); a1.f = c->f; a1.dty0 = DTY; a1.u1 = rec.W[1];
  a1.D = c->d_D; a1.dt = c->d_cdt; a1.ub1 = c->d_ub1; a1.order = c->d_order;
  a1.Af = use_af ? c->dty0 : nullptr; a1.rstride = c->rstride;
  if ((rc = launch(c, 0, k_rkg2_stage1, nblocks(a1.g), a1, 3.0 * n))) return rc;
  if (fused_usable(c) && !c->buf2) {
    if ((rc = dalloc(&c->buf2, c->nfield)) || (rc = dalloc(&c->buf3, c->nfield))) return rc;
  }
  bool fused_ok = fused_usable(c) && c->buf2;
  std::vector<double *> final_loc(nr, nullptr);
  StageArgs sa;
  sa.m = c->m; sa.f = c->f; sa.dty0 = DTY; sa.D = c->d_D; sa.s = c->d_s; sa.order = c->d_order;
  sa.rstride = c->rstride;
  int active = n;
  int64_t k = 2;
  while (k <= smax) {
    while (active > 0 && s[ord[active - 1]] < k) active--;
    const double *pA = rec.W[k - 1];
    const double *pB = k == 2 ? c->f : rec.W[k - 2];
    int T = c->fused_T;
    if (fused_ok && active >= c->fused_min) {
      // several stages per launch, each stage's rows also kept in rec.W
      double *Wt[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
      for (int t = 0; t < T; t++)
        if (k + t < smax) Wt[t] = rec.W[k + t];
      if ((rc = launch_fused(c, k, active, ord, s, pA, pB, DTY, c->buf0, c->buf2, Wt))) return rc;
      for (int z = 0; z < active; z++)
        if (s[ord[z]] <= k + T - 1) final_loc[ord[z]] = c->buf0;
      k += T;
    } else {
      sa.g = tiles_for(c, active);
      sa.k = (int)k;
      sa.x = pA;
      sa.ukm2 = pB;
      sa.out = k < smax ? rec.W[k] : c->buf1;  // stage smax finishes every active realization
      sa.coef = c->d_coef + (size_t)k * nr;
      if ((rc = launch(c, 0, k_rkg2_stage, nblocks(sa.g), sa, 5.0 * active))) return rc;
      k += 1;
    }
  }
  return copy_finals(c, final_loc);
}

// Adjoint of one round over its stored stages with k_rkg2_adj_fused: T
// transposed stages per launch, realizations in relative steps (all start
// at their own s_r together; the s-sorted prefix with stages left is
// active).  lam is the incoming adjoint (level 0 of the first launch) and
// receives the adjoint of the round's input.
int cycle_adjoint_fused(hft_ctx *c, const CycleRec &rec, const double *U0, double *lam) {
  constexpr int T = 4;
  int nr = c->nr, rc;
  std::vector<int> ord;
  if ((rc = upload_rkg2(c, rec.reals, rec.dt, rec.s, ord, &rec.W))) return rc;
  int n = (int)ord.size();
  const std::vector<int64_t> &s = rec.s;
  double *Q[4];
  for (int i = 0; i < 3; i++)
    if (!(Q[i] = adj_buf(c, c->adj_fix, AB_LB + i, &rc))) return rc;
  if (!(Q[3] = adj_buf(c, c->adj_fix, AB_Q3, &rc))) return rc;
  double *L1 = adj_buf(c, c->adj_fix, AB_L1, &rc);
  if (!L1) return rc;
  double *L0 = adj_buf(c, c->adj_fix, AB_L0, &rc);
  if (!L0) return rc;
  AdjFusedArgs fa;
  fa.m = c->m;
  fa.H = std::min(c->adj_fused_rows, c->ntm);
  fa.nbands = (c->ntm + fa.H - 1) / fa.H;
  fa.ncol = ((c->nu + 63) / 64) * 64;
  fa.ctiles = (c->nu + kBlock - 1) / kBlock;
  fa.L1 = L1; fa.L0 = L0; fa.u0 = U0; fa.D = c->d_D; fa.coef = c->d_coef; fa.nr_table = nr;
  fa.s = c->d_s; fa.dt = c->d_cdt; fa.order = c->d_order; fa.rstride = c->rstride;
  fa.part = c->d_apD;
  size_t shmem = (size_t)T * 3 * fa.ncol * sizeof(double);
  const double *A = lam, *B = lam;  // B is not read at m0 = 0
  for (int64_t m0 = 0;; m0 += T) {
    int active = 0;  // realizations with stages left: s_r - 1 - m0 >= 1
    while (active < n && s[ord[active]] - 1 - m0 >= 1) active++;
    if (active == 0) break;
    double *oA = nullptr, *oB = nullptr;
    for (double *x : Q) {
      if (x == A || x == B) continue;
      if (!oA) oA = x;
      else if (!oB) oB = x;
    }
    fa.m0 = m0;
    fa.A = A; fa.B = B; fa.outA = oA; fa.outB = oB;
    double units = 0;
    for (int z = 0; z < active; z++)
      units += 7.0 * (double)std::min<int64_t>(T, s[ord[z]] - 1 - m0);  // as stage by stage
    rc = fa.ncol <= 512 ? launch_ex(c, kGroupAdjoint, k_rkg2_adj_fused<T, 512>, active * fa.nbands,
                                    fa.ncol, shmem, fa, units)
                        : launch_ex(c, kGroupAdjoint, k_rkg2_adj_fused<T, 1024>, active * fa.nbands,
                                    fa.ncol, shmem, fa, units);
    if (rc) return rc;
    A = oA;
    B = oB;
  }
  AdjCloseArgs ca;
  ca.m = c->m; ca.g = tiles_for(c, n); ca.lam1 = L1; ca.L0 = L0; ca.u0 = U0; ca.D = c->d_D;
  ca.dt = c->d_cdt; ca.ub1 = c->d_ub1; ca.order = c->d_order; ca.rstride = c->rstride;
  ca.part = c->d_apD;
  if ((rc = launch(c, kGroupAdjoint, k_rkg2_adj_close, nblocks(ca.g), ca, 3.0 * n))) return rc;
  dim3 gi((unsigned)((c->rstride + kBlock - 1) / kBlock), (unsigned)n);
  hipLaunchKernelGGL(k_copy_reals, gi, dim3(kBlock), 0, c->stream, lam, (const double *)L0,
                     (const int *)c->d_order, c->rstride);
  HIPCHK(hipGetLastError());
  return 0;
}

// Adjoint of one round of RKG2 cycles from its input U0 (realizations of
// the round only; the others pass through unchanged).  Stage states come
// from rec.W when the forward recomputation kept them, else from segment
// recomputation.
int cycle_adjoint(hft_ctx *c, const CycleRec &rec, const double *U0, double *lam) {
  if (!rec.W.empty() && c->adj_fused && c->nu <= 1024 &&
      c->adj_fused_rows + 8 <= kFusedMaxRows)
    return cycle_adjoint_fused(c, rec, U0, lam);
  int nr = c->nr, rc;
  std::vector<int> ord;
  if ((rc = upload_rkg2(c, rec.reals, rec.dt, rec.s, ord))) return rc;
  int n = (int)ord.size();
  const std::vector<int64_t> &s = rec.s;
  int64_t smax = s[ord[0]];
  auto nactive = [&](int64_t kmin) {  // realizations with s_r >= kmin (a prefix of ord)
    int q = 0;
    while (q < n && s[ord[q]] >= kmin) q++;
    return q;
  };
  double *LB[3], *R[3];
  for (int i = 0; i < 3; i++)
    if (!(LB[i] = adj_buf(c, c->adj_fix, AB_LB + i, &rc))) return rc;
  double *L0 = adj_buf(c, c->adj_fix, AB_L0, &rc);
  if (!L0) return rc;
  bool stored = !rec.W.empty();
  // segment length (one segment covering every stage when they are stored)
  int64_t K = stored ? smax - 1 : c->adj_seg;
  if (K <= 0) {
    double field_mb = (double)c->nfield * 8.0 / 1048576.0;
    double all_mb = (double)(smax + 24) * field_mb;
    K = all_mb <= (double)c->adj_mem_mb ? smax - 1
                                        : (int64_t)std::ceil(std::sqrt((double)(smax - 1)));
  }
  if (K < 1) K = 1;
  int64_t nseg = (smax - 1 + K - 1) / K;
  std::vector<const double *> ck1(nseg), ck2(nseg);
  StageArgs sa;
  auto fwd_stage = [&](int64_t k, const double *x, const double *xm2, double *out) -> int {
    int na = nactive(k + 1);  // u_k is needed by realizations with s_r >= k+1
    if (na == 0) return 0;
    sa.g = tiles_for(c, na);
    sa.k = (int)k;
    sa.x = x;
    sa.ukm2 = xm2;
    sa.out = out;
    sa.coef = c->d_coef + (size_t)k * nr;
    return launch(c, 0, k_rkg2_stage, nblocks(sa.g), sa, 5.0 * na);
  };
  ck2[0] = U0;
  if (stored) {
    ck1[0] = rec.W[1];
  } else {
    double *DTY = adj_buf(c, c->adj_fix, AB_DTY, &rc);
    if (!DTY) return rc;
    double *U1 = adj_buf(c, c->adj_fix, AB_U1, &rc);
    if (!U1) return rc;
    // stage 1: dty0 = dt L u0, u1 = u0 + ub1 dty0
    Stage1Args a1;
    a1.m = c->m; a1.g = tiles_for(c, n); a1.f = U0; a1.dty0 = DTY; a1.u1 = U1;
    a1.D = c->d_D; a1.dt = c->d_cdt; a1.ub1 = c->d_ub1; a1.order = c->d_order;
    a1.Af = nullptr; a1.rstride = c->rstride;
    if ((rc = launch(c, 0, k_rkg2_stage1, nblocks(a1.g), a1, 3.0 * n))) return rc;
    sa.m = c->m; sa.f = const_cast<double *>(U0); sa.dty0 = DTY; sa.D = c->d_D; sa.s = c->d_s;
    sa.order = c->d_order; sa.rstride = c->rstride;
    // checkpoint pass: pairs (u_{k0-1}, u_{k0-2}) at segment starts k0 = 2 + i K
    ck1[0] = U1;
    if (nseg > 1) {
      for (int i = 0; i < 3; i++)
        if (!(R[i] = adj_buf(c, c->adj_fix, AB_R + i, &rc))) return rc;
      const double *pm2 = U0, *pm1 = U1;
      int ri = 0;
      int64_t klast = 2 + (nseg - 1) * K;
      for (int64_t k = 2; k <= klast - 1; k++) {
        double *out = R[ri];
        ri = (ri + 1) % 3;
        if ((rc = fwd_stage(k, pm1, pm2, out))) return rc;
        pm2 = pm1;
        pm1 = out;
        if ((k - 1) % K == 0) {  // k = k0 - 1 of segment i = (k-1)/K
          int64_t i = (k - 1) / K;
          double *b1 = adj_buf(c, c->adj_ck, 2 * i, &rc);
          if (!b1) return rc;
          double *b2 = adj_buf(c, c->adj_ck, 2 * i + 1, &rc);
          if (!b2) return rc;
          HIPCHK(hipMemcpyAsync(b1, pm1, c->nfield * sizeof(double), hipMemcpyDeviceToDevice,
                                c->stream));
          HIPCHK(hipMemcpyAsync(b2, pm2, c->nfield * sizeof(double), hipMemcpyDeviceToDevice,
                                c->stream));
          ck1[i] = b1;
          ck2[i] = b2;
        }
      }
    }
  }
  // adjoint initial values: lam_{s_r} = lam_r, adjoint of u0 = 0
  dim3 gi((unsigned)((c->rstride + kBlock - 1) / kBlock), (unsigned)n);
  hipLaunchKernelGGL(k_adj_cycle_init, gi, dim3(kBlock), 0, c->stream, LB[0], LB[1], LB[2], L0,
                     (const double *)lam, (const int *)c->d_order, (const int64_t *)c->d_s,
                     c->rstride);
  HIPCHK(hipGetLastError());
  AdjStageArgs ja;
  ja.m = c->m; ja.L0 = L0; ja.u0 = U0; ja.D = c->d_D; ja.s = c->d_s; ja.dt = c->d_cdt;
  ja.order = c->d_order; ja.rstride = c->rstride; ja.part = c->d_apD;
  for (int64_t i = nseg - 1; i >= 0; i--) {
    int64_t k0 = 2 + i * K, k1 = std::min<int64_t>(k0 + K - 1, smax);
    // W[t] = u_{k0-2+t}
    std::vector<const double *> W(k1 - k0 + 2);
    W[0] = ck2[i];
    W[1] = ck1[i];
    for (int64_t k = k0; k <= k1 - 1; k++) {
      if (stored) {
        W[k - k0 + 2] = rec.W[k];
        continue;
      }
      double *out = adj_buf(c, c->adj_w, (size_t)(k - k0), &rc);
      if (!out) return rc;
      if ((rc = fwd_stage(k, W[k - k0 + 1], W[k - k0], out))) return rc;
      W[k - k0 + 2] = out;
    }
    for (int64_t k = k1; k >= k0; k--) {
      int na = nactive(k);
      if (na == 0) continue;
      ja.g = tiles_for(c, na);
      ja.k = (int)k;
      ja.lam = LB[k % 3];
      ja.lamn = LB[(k + 1) % 3];
      ja.out = LB[(k - 1) % 3];
      ja.u = W[k - k0 + 1];
      ja.coef = c->d_coef + (size_t)k * nr;
      ja.coefn = c->d_coef + (size_t)(k + 1) * nr;
      if ((rc = launch(c, kGroupAdjoint, k_rkg2_adj_stage, nblocks(ja.g), ja, 7.0 * na))) return rc;
    }
  }
  AdjCloseArgs ca;
  ca.m = c->m; ca.g = tiles_for(c, n); ca.lam1 = LB[1]; ca.L0 = L0; ca.u0 = U0; ca.D = c->d_D;
  ca.dt = c->d_cdt; ca.ub1 = c->d_ub1; ca.order = c->d_order; ca.rstride = c->rstride;
  ca.part = c->d_apD;
  if ((rc = launch(c, kGroupAdjoint, k_rkg2_adj_close, nblocks(ca.g), ca, 3.0 * n))) return rc;
  hipLaunchKernelGGL(k_copy_reals, gi, dim3(kBlock), 0, c->stream, lam, (const double *)L0,
                     (const int *)c->d_order, c->rstride);
  HIPCHK(hipGetLastError());
  return 0;
}

// What one step's forward recomputation keeps for its adjoint.
struct StepStore {
  double *XA[4] = {nullptr, nullptr, nullptr, nullptr};
  double *XB[4] = {nullptr, nullptr, nullptr, nullptr};
  std::vector<CycleRec> recs;
  std::vector<double *> UC;  // input of every diffusion round
};

// Fixed sub-cycles: the largest RKG2 stage count of a cycle (0 for PTL).
int64_t fixed_smax(hft_ctx *c, double dtl, int subcycles) {
  if (subcycles <= 0 || !c->diff_set) return 0;
  int64_t smax = 0;
  for (int r = 0; r < c->nr; r++) {
    double dte = c->mode == HFT_MODE_COUPLED ? c->dt_euler : c->dt_euler_r[r];
    smax = std::max(smax, sts_stages(dtl / (double)subcycles, dte));
  }
  return smax;
}

// Fields one step keeps when its RKG2 stages are stored.
double step_fields(hft_ctx *c, double dtl, int strang, int subcycles) {
  double f = c->flow_set ? (strang ? 8.0 : 4.0) : 0.0;
  if (c->diff_set) f += (double)subcycles * (double)(fixed_smax(c, dtl, subcycles) + 2);
  return f;
}

// Forward step from the state, keeping in st (fields drawn from c->adj_st
// starting at next) what backward_step needs.  store_stages: keep the RKG2
// stages (fixed sub-cycles only), else only the round inputs.
int forward_store_step(hft_ctx *c, double dtl, int strang, int subcycles, StepStore &st,
                       size_t &next, bool store_stages) {
  int rc;
  int nr = c->nr;
  bool adv = c->flow_set, dif = c->diff_set;
  double adt = strang ? dtl * half : dtl;
  if (adv) {
    for (int i = 0; i < 4; i++)
      if (!(st.XA[i] = adj_buf(c, c->adj_st, next++, &rc))) return rc;
    if ((rc = advection_store(c, adt, st.XA))) return rc;
  }
  if (dif) {
    std::function<int(int)> pre = [&](int round) -> int {
      (void)round;
      int rc2;
      double *u = adj_buf(c, c->adj_st, next++, &rc2);
      if (!u) return rc2;
      st.UC.push_back(u);
      HIPCHK(hipMemcpyAsync(u, c->f, c->nfield * sizeof(double), hipMemcpyDeviceToDevice,
                            c->stream));
      return 0;
    };
    bool store = store_stages && subcycles > 0;
    SuperstepFn ss = [&](const std::vector<int> &reals, const std::vector<double> &dt2,
                         const std::vector<int64_t> &sv, bool use_af) -> int {
      CycleRec cr;
      cr.reals = reals;
      cr.dt.assign(nr, 0.0);
      cr.s.assign(nr, 0);
      for (int r : reals) {
        cr.dt[r] = dt2[r];
        cr.s[r] = sv[r];
      }
      int rc2 = store ? superstep_store(c, reals, dt2, sv, use_af, cr, next)
                      : rkg2_superstep(c, reals, dt2, sv, use_af);
      st.recs.push_back(std::move(cr));
      return rc2;
    };
    if ((rc = diffusion_impl(c, dtl, subcycles, nullptr, nullptr, &pre, &ss))) return rc;
  }
  if (strang && adv) {
    for (int i = 0; i < 4; i++)
      if (!(st.XB[i] = adj_buf(c, c->adj_st, next++, &rc))) return rc;
    if ((rc = advection_store(c, adt, st.XB))) return rc;
  }
  return 0;
}

// Transposed sweeps of one step over LAM (pitched layout).
int backward_step(hft_ctx *c, double dtl, int strang, const StepStore &st, double *LAM,
                  bool frozen) {
  int rc;
  bool adv = c->flow_set, dif = c->diff_set;
  double adt = strang ? dtl * half : dtl;
  double *XB[4] = {st.XB[0], st.XB[1], st.XB[2], st.XB[3]};
  double *XA[4] = {st.XA[0], st.XA[1], st.XA[2], st.XA[3]};
  if (strang && adv && (rc = advection_adjoint(c, adt, XB, LAM, frozen))) return rc;
  if (dif) {
    for (int q = (int)st.recs.size() - 1; q >= 0; q--) {
      if (st.recs[q].reals.empty()) continue;
      if ((rc = cycle_adjoint(c, st.recs[q], st.UC[q], LAM))) return rc;
    }
  }
  if (adv && (rc = advection_adjoint(c, adt, XA, LAM, frozen))) return rc;
  return 0;
}

// Common set-up of the adjoint entry points: partial sums zeroed, LAM
// filled from the caller's device array.
int adjoint_begin(hft_ctx *c, const double *lam_dev, int64_t ld, double **LAM, int *nbands_out) {
  int nr = c->nr, rc;
  int ctiles = (c->nu + kBlock - 1) / kBlock;
  size_t nD = (size_t)nr * c->ntm * ctiles;
  int nbands = (c->ntm + c->adj_band_rows - 1) / c->adj_band_rows;
  size_t nA = (size_t)nr * nbands * 2;
  if (!c->d_apD) {
    if ((rc = dalloc(&c->d_apD, nD)) || (rc = dalloc(&c->d_gout, (size_t)nr * 3))) return rc;
  }
  if (nA > c->apA_cap) {
    if (c->d_apA) HIPCHK(hipFree(c->d_apA));
    c->d_apA = nullptr;
    if ((rc = dalloc(&c->d_apA, nA))) return rc;
    c->apA_cap = nA;
  }
  if ((rc = fill_dev(c, c->d_apD, nD, 0.0)) || (rc = fill_dev(c, c->d_apA, nA, 0.0)) ||
      (rc = fill_dev(c, c->d_gout, (size_t)nr * 3, 0.0)))
    return rc;
  if (!(*LAM = adj_buf(c, c->adj_fix, AB_LAM, &rc))) return rc;
  HIPCHK(hipMemcpy2DAsync(*LAM, c->pitch * sizeof(double), lam_dev, (size_t)ld * sizeof(double),
                          c->nu * sizeof(double), (size_t)nr * c->ntm, hipMemcpyDeviceToDevice,
                          c->stream));
  *nbands_out = nbands;
  return 0;
}

// LAM back to the caller; parameter sensitivities added into grad (host).
int adjoint_end(hft_ctx *c, double *lam_dev, int64_t ld, const double *LAM, int nbands,
                double *grad) {
  int nr = c->nr;
  int ctiles = (c->nu + kBlock - 1) / kBlock;
  HIPCHK(hipMemcpy2DAsync(lam_dev, (size_t)ld * sizeof(double), LAM, c->pitch * sizeof(double),
                          c->nu * sizeof(double), (size_t)nr * c->ntm, hipMemcpyDeviceToDevice,
                          c->stream));
  hipLaunchKernelGGL(k_sum_parts, dim3(nr), dim3(256), 0, c->stream, (const double *)c->d_apD,
                     c->ntm * ctiles, c->ntm * ctiles, 0, 1, c->d_gout, 3, 0);
  hipLaunchKernelGGL(k_sum_parts, dim3(nr), dim3(256), 0, c->stream, (const double *)c->d_apA,
                     nbands, nbands * 2, 0, 2, c->d_gout, 3, 1);
  hipLaunchKernelGGL(k_sum_parts, dim3(nr), dim3(256), 0, c->stream, (const double *)c->d_apA,
                     nbands, nbands * 2, 1, 2, c->d_gout, 3, 2);
  HIPCHK(hipGetLastError());
  if (grad) {
    std::vector<double> g((size_t)nr * 3);
    HIPCHK(hipMemcpyAsync(g.data(), c->d_gout, g.size() * sizeof(double), hipMemcpyDeviceToHost,
                          c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    for (size_t i = 0; i < g.size(); i++) grad[i] += g[i];
  }
  return 0;
}

int adjoint_check(hft_ctx *c, const double *lam_dev, int64_t ld, int subcycles, double dtl,
                  int strang) {
  if (!c || !lam_dev) return fail(HFT_E_ARG, "null argument");
  if (ld < c->nu) return fail(HFT_E_ARG, "adjoint: ld < %d", c->nu);
  if (subcycles < 0) return fail(HFT_E_ARG, "adjoint: subcycles < 0");
  if (c->flow_set && (c->flow2d || c->atten))
    return fail(HFT_E_STATE, "adjoint: 2-D flows (from files or attenuated) have no adjoint");
  if (c->flow_set && c->nu > 1024)
    return fail(HFT_E_ARG, "adjoint: more than 1024 unique phi columns");
  if (sources_active(c)) return fail(HFT_E_STATE, "adjoint: source terms are not supported");
  if (c->flow_set) {
    std::vector<int> nsub;
    int nmax, rc = advection_substeps(c, strang ? half * dtl : dtl, nsub, nmax);
    if (rc) return rc;
    if (nmax > 1)
      return fail(HFT_E_ARG,
                  "adjoint: the advection step exceeds the CFL limit of a realization "
                  "(%d sub-steps); shorten dt", nmax);
  }
  return 0;
}

}  // namespace

int set_adjoint_attributes() {
  int lds = (19 * 1024 + 16) * (int)sizeof(double);
  struct {
    const void *fn;
    int bytes;
  } req[] = {{(const void *)k_rkg2_adj_fused<4, 1024>, 4 * 3 * 1024 * (int)sizeof(double)},
             {(const void *)k_advection_vjp_band<true, true, 1024>, lds},
             {(const void *)k_advection_vjp_band<true, false, 1024>, lds},
             {(const void *)k_advection_vjp_band<false, true, 1024>, lds},
             {(const void *)k_advection_vjp_band<false, false, 1024>, lds}};
  for (auto &q : req) {
    hipError_t e = hipFuncSetAttribute(q.fn, hipFuncAttributeMaxDynamicSharedMemorySize, q.bytes);
    if (e != hipSuccess)
      return fail(HFT_E_HIP, "hipFuncSetAttribute(%d bytes of LDS, adjoint): %s", q.bytes,
                  hipGetErrorString(e));
  }
  return 0;
}

extern "C" {

int hft_set_stream(hft_ctx *c, void *stream, int own) {
  if (!c) return fail(HFT_E_ARG, "null context");
  HIPCHK(hipStreamSynchronize(c->stream));
  c->epoch++;
  if (!own) {
    if (!c->own_stream) c->own_stream = c->stream;
    c->stream = (hipStream_t)stream;  // NULL: the null stream
  } else if (c->own_stream) {
    c->stream = c->own_stream;
    c->own_stream = nullptr;
  }
  return 0;
}

int hft_set_state_device(hft_ctx *c, const double *src, int64_t ld) {
  if (!c || !src) return fail(HFT_E_ARG, "null argument");
  if (ld < c->nu) return fail(HFT_E_ARG, "hft_set_state_device: ld < %d", c->nu);
  HIPCHK(hipMemcpy2DAsync(c->f, c->pitch * sizeof(double), src, (size_t)ld * sizeof(double),
                          c->nu * sizeof(double), (size_t)c->nr * c->ntm, hipMemcpyDeviceToDevice,
                          c->stream));
  return 0;
}

int hft_get_state_device(hft_ctx *c, double *dst, int64_t ld) {
  if (!c || !dst) return fail(HFT_E_ARG, "null argument");
  if (ld < c->nu) return fail(HFT_E_ARG, "hft_get_state_device: ld < %d", c->nu);
  HIPCHK(hipMemcpy2DAsync(dst, (size_t)ld * sizeof(double), c->f, c->pitch * sizeof(double),
                          c->nu * sizeof(double), (size_t)c->nr * c->ntm, hipMemcpyDeviceToDevice,
                          c->stream));
  return 0;
}

int hft_step_adjoint(hft_ctx *c, double dtl, int strang, int subcycles, double *lam_dev,
                     int64_t ld, double *grad, int flags) {
  int rc = adjoint_check(c, lam_dev, ld, subcycles, dtl, strang);
  if (rc) return rc;
  bool frozen = (flags & HFT_ADJ_FROZEN_WENO) != 0;
  double *LAM;
  int nbands;
  if ((rc = adjoint_begin(c, lam_dev, ld, &LAM, &nbands))) return rc;
  double field_mb = (double)c->nfield * 8.0 / 1048576.0;
  bool store = c->adj_seg == 0 &&
               (step_fields(c, dtl, strang, subcycles) + 24.0) * field_mb <= (double)c->adj_mem_mb;
  StepStore st;
  size_t next = 0;
  if ((rc = forward_store_step(c, dtl, strang, subcycles, st, next, store))) return rc;
  if ((rc = backward_step(c, dtl, strang, st, LAM, frozen))) return rc;
  return adjoint_end(c, lam_dev, ld, LAM, nbands, grad);
}

int hft_steps_adjoint(hft_ctx *c, int nsteps, double dtl, int strang, int subcycles,
                      double *lam_dev, int64_t ld, double *grad, int flags) {
  int rc = adjoint_check(c, lam_dev, ld, subcycles, dtl, strang);
  if (rc) return rc;
  if (nsteps < 1) return fail(HFT_E_ARG, "hft_steps_adjoint: nsteps < 1");
  bool frozen = (flags & HFT_ADJ_FROZEN_WENO) != 0;
  double *LAM;
  int nbands;
  if ((rc = adjoint_begin(c, lam_dev, ld, &LAM, &nbands))) return rc;
  double field_mb = (double)c->nfield * 8.0 / 1048576.0;
  double per_step = step_fields(c, dtl, strang, subcycles);
  bool all = c->adj_seg == 0 && subcycles > 0 &&
             ((double)nsteps * per_step + 24.0) * field_mb <= (double)c->adj_mem_mb;
  size_t next = 0;
  if (all) {
    // one forward pass keeping every step's data, then the reverse sweep
    std::vector<StepStore> st(nsteps);
    for (int n = 0; n < nsteps; n++)
      if ((rc = forward_store_step(c, dtl, strang, subcycles, st[n], next, true))) return rc;
    for (int n = nsteps - 1; n >= 0; n--)
      if ((rc = backward_step(c, dtl, strang, st[n], LAM, frozen))) return rc;
  } else {
    // keep the step inputs, recompute each step before its adjoint
    std::vector<double *> X(nsteps);
    for (int n = 0; n < nsteps; n++) {
      if (!(X[n] = adj_buf(c, c->adj_st, next++, &rc))) return rc;
      HIPCHK(hipMemcpyAsync(X[n], c->f, c->nfield * sizeof(double), hipMemcpyDeviceToDevice,
                            c->stream));
      if ((rc = hft_flux_transport_step(c, dtl, strang, subcycles, nullptr, nullptr))) return rc;
    }
    double *XN = nullptr;
    if (!(XN = adj_buf(c, c->adj_st, next++, &rc))) return rc;
    HIPCHK(hipMemcpyAsync(XN, c->f, c->nfield * sizeof(double), hipMemcpyDeviceToDevice,
                          c->stream));
    bool store = c->adj_seg == 0 && (per_step + (double)nsteps + 26.0) * field_mb <=
                                        (double)c->adj_mem_mb;
    size_t mark = next;
    for (int n = nsteps - 1; n >= 0; n--) {
      HIPCHK(hipMemcpyAsync(c->f, X[n], c->nfield * sizeof(double), hipMemcpyDeviceToDevice,
                            c->stream));
      StepStore st;
      next = mark;
      if ((rc = forward_store_step(c, dtl, strang, subcycles, st, next, store))) return rc;
      if ((rc = backward_step(c, dtl, strang, st, LAM, frozen))) return rc;
    }
    HIPCHK(hipMemcpyAsync(c->f, XN, c->nfield * sizeof(double), hipMemcpyDeviceToDevice,
                          c->stream));
  }
  return adjoint_end(c, lam_dev, ld, LAM, nbands, grad);
}

}  // extern "C"
