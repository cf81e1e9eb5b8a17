// hft_adjoint_kernels.hip -- discrete adjoint (vector-Jacobian products) of
// the flux-transport step, for gfx950 (FP64).
//
// The forward map of one step is f -> A(D(A(f))) (Strang) or D(A(f)) with
//   A = SSPRK(4,3) over the WENO3 operator W (src/hipft.f90:7551-7574,
//       7711-7925),
//   D = RKG2 super-steps over the diffusion operator L (src/hipft.f90:5163-5242,
//       7927-7990).
// These kernels apply the transposed Jacobians of W and of the RKG2 stage
// recurrence exactly as the forward kernels evaluate them: every entry of
// the ntm x nu state (the pole rows hold one value per column) is an
// independent variable, and the transposes follow the operators' own
// stencils, pole closures and periodic seam.  Parameter sensitivities are the
// log-derivatives with respect to a per-realization diffusivity scale
// (L = D_r * L1) and scales of the theta and phi velocities (the
// Lax-Friedrichs alpha scales with them).  Discrete choices of the forward
// run (stage counts, sub-cycle time steps, upwind signs) are held fixed.
//
// Unlike the forward kernels these need not round like the reference, so
// they let the compiler contract multiply-adds (HFT_ADJ_FAST opens each body)
// and use reciprocals.
#include "hft_device.h"

#define HFT_ADJ_FAST _Pragma("clang fp contract(fast)")

namespace hft {

// Deterministic block sum for blockDim.x a multiple of 64 (<= 1024).
__device__ inline double block_sum(double *sh, double v) {
  for (int h = 32; h >= 1; h >>= 1) v = v + __shfl_xor(v, h, 64);
  int w = threadIdx.x >> 6, nw = blockDim.x >> 6;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) sh[w] = v;
  __syncthreads();
  double s = 0.0;
  for (int i = 0; i < nw; i++) s = s + sh[i];
  __syncthreads();
  return s;
}

// ---------------------------------------------------------------------------
// Diffusion operator transpose.
// Forward (lap_interior / lap_pole): for interior rows
//   y(j,c) = c1 x(j,c-1) + c2 x(j-1,c) + c3 x(j,c) + c4 x(j+1,c) + c5 x(j,c+1)
// and pole rows y(0,:) = bc_n sum_c 2D (x(1,c)-x(0,c)) dp(c),
//                y(ntm-1,:) = -bc_s sum_c 2D (x(ntm-1,c)-x(ntm-2,c)) dp(c).
// ---------------------------------------------------------------------------
__device__ inline void lap_c15(const Mesh &m, double D, int j, int c, double &c1, double &c5) {
  HFT_ADJ_FAST
  double sti = m.st_i[j], dpi = m.dp_i[c + 1];
  c1 = D * m.dph_i[c + 1];
  c1 = c1 * dpi;
  c1 = c1 * sti;
  c1 = c1 * sti;
  c5 = D * m.dph_i[c + 2];
  c5 = c5 * dpi;
  c5 = c5 * sti;
  c5 = c5 * sti;
}
__device__ inline double lap_c2(const Mesh &m, double D, int j) {
  double sti = m.st_i[j];
  double c2 = D * m.dth_i[j];
  c2 = c2 * m.dt_i[j];
  c2 = c2 * sti;
  return c2 * m.sth[j];
}
__device__ inline double lap_c4(const Mesh &m, double D, int j) {
  double sti = m.st_i[j];
  double c4 = D * m.dth_i[j + 1];
  c4 = c4 * m.dt_i[j];
  c4 = c4 * sti;
  return c4 * m.sth[j + 1];
}

// (L^T lam)(j,c).  Ln, Ls: sums of lam over the north / south pole rows
// (only read for rows 0,1 / ntm-2,ntm-1).
__device__ inline double lap_T(const Mesh &m, double D, const double *lam, int j, int c,
                               double Ln, double Ls) {
  HFT_ADJ_FAST
  const int ntm = m.ntm, nu = m.nu, pitch = m.pitch;
  const double *row = lam + (size_t)j * pitch;
  double y = 0.0;
  if (j >= 1 && j <= ntm - 2) {
    double c1, c5;
    lap_c15(m, D, j, c, c1, c5);
    double c2 = lap_c2(m, D, j), c4 = lap_c4(m, D, j);
    double c3 = -(((c1 + c2) + c4) + c5);
    int cw = wrapc(c - 1, nu), ce = wrapc(c + 1, nu);
    double w1, w5, e1, e5;
    lap_c15(m, D, j, cw, w1, w5);
    lap_c15(m, D, j, ce, e1, e5);
    y = c3 * row[c];
    y = y + w5 * row[cw];
    y = y + e1 * row[ce];
  }
  if (j - 1 >= 1 && j - 1 <= ntm - 2) y = y + lap_c4(m, D, j - 1) * row[c - pitch];
  if (j + 1 >= 1 && j + 1 <= ntm - 2) y = y + lap_c2(m, D, j + 1) * row[c + pitch];
  double D2 = D + D;
  double pn = D2 * m.dp[c + 1] * m.bc_diff_n, ps = D2 * m.dp[c + 1] * m.bc_diff_s;
  if (j == 0) y = y - pn * Ln;
  if (j == 1) y = y + pn * Ln;
  if (j == ntm - 1) y = y - ps * Ls;
  if (j == ntm - 2) y = y + ps * Ls;
  return y;
}

// Row sums of the pole rows of lam for the tiles that need them; every lane
// of the wave takes part.
__device__ inline void pole_sums(const Mesh &m, const double *lam, int j, double &Ln, double &Ls) {
  Ln = 0.0;
  Ls = 0.0;
  const int nu = m.nu;
  if (j <= 1) Ln = wave_pole_sum(nu, [&](int k) { return lam[k]; });
  if (j >= m.ntm - 2) {
    const double *rs = lam + (size_t)(m.ntm - 1) * m.pitch;
    Ls = wave_pole_sum(nu, [&](int k) { return rs[k]; });
  }
}

// RKG2 adjoint of stage k >= 2.  Forward: u_k = a_k u_{k-1} + b_k u_{k-2} +
// c_k u0 + d_k L u_{k-1} + g_k dty0, u_1 = u0 + ub_1 dty0, dty0 = dt L u0.
// With lam_j the adjoint of u_j, stage k (descending) forms
//   lam_{k-1} = b_{k+1} lam_{k+1} + a_k lam_k + d_k L^T lam_k
// (b_{k+1} lam_{k+1} is absent at k = s), and adds to the adjoint of u0
//   c_k lam_k + g_k dt L^T lam_k
// (the dty0 adjoint sum_k g_k lam_k + ub_1 lam_1 goes through dt L^T at
// once, since L^T lam_k is formed here anyway).  k == 2 also adds b_2 lam_2
// and lam_1 (u_0 = u0 and u_1 = u0 + ...); k_rkg2_adj_close adds
// ub_1 dt L^T lam_1.  part[slot] += <L^T lam_k, d_k u_{k-1} + g_k dt u0>,
// the diffusivity log-derivative, one slot per tile (fixed order, no atomics).
struct AdjStageArgs {
  Mesh m;
  Tiles g;
  const double *lam;   // lam_k
  const double *lamn;  // lam_{k+1}
  double *out;         // lam_{k-1}
  double *L0;          // adjoint of u0
  const double *u;     // u_{k-1}
  const double *u0;
  const double *D;
  const StageCoef *coef;   // stage k, [nr]
  const StageCoef *coefn;  // stage k+1, [nr] (read where k < s_r)
  const int64_t *s;
  const double *dt;        // [nr] cycle time step
  const int *order;
  int k;
  size_t rstride;
  double *part;  // [nr][ntm*ctiles]
};

__global__ __launch_bounds__(kBlock) void k_rkg2_adj_stage(AdjStageArgs a) {
  HFT_ADJ_FAST
  __shared__ double sh[kBlock / 64];
  int t = xcd_remap(blockIdx.x, gridDim.x);
  int z, j, ct;
  tile_of(t, a.g, z, j, ct);
  int r = a.order[z];
  const Mesh &m = a.m;
  int c = ct * kBlock + threadIdx.x;
  bool col = c < m.nu;
  int cc = col ? c : 0;
  size_t base = (size_t)r * a.rstride;
  const double *lam = a.lam + base;
  double Ln, Ls;
  pole_sums(m, lam, j, Ln, Ls);
  double D = a.D[r];
  double lt = lap_T(m, D, lam, j, cc, Ln, Ls);
  double e = 0.0;
  if (col) {
    size_t p = base + (size_t)j * m.pitch + c;
    const StageCoef cf = a.coef[r];
    double lk = lam[(size_t)j * m.pitch + c];
    double lp = cf.a * lk + cf.d * lt;
    if (a.k < a.s[r]) lp = lp + a.coefn[r].b * a.lamn[p];
    a.out[p] = lp;
    double gdt = cf.g * a.dt[r];
    double l0 = a.L0[p] + cf.c * lk + gdt * lt;
    if (a.k == 2) l0 = l0 + cf.b * lk + lp;
    a.L0[p] = l0;
    e = lt * (cf.d * a.u[p] + gdt * a.u0[p]);
  }
  double s = block_sum(sh, e);
  if (threadIdx.x == 0) {
    size_t slot = (size_t)r * (m.ntm * a.g.ctiles) + (size_t)j * a.g.ctiles + ct;
    a.part[slot] = a.part[slot] + s;
  }
}

// Closing step: the dty0 adjoint's ub_1 lam_1 term, L0 += ub_1 dt L^T lam_1,
// part += ub_1 dt <L^T lam_1, u0>.
struct AdjCloseArgs {
  Mesh m;
  Tiles g;
  const double *lam1;
  double *L0;
  const double *u0;
  const double *D;
  const double *dt;   // [nr] cycle time step
  const double *ub1;  // [nr]
  const int *order;
  size_t rstride;
  double *part;
};

__global__ __launch_bounds__(kBlock) void k_rkg2_adj_close(AdjCloseArgs a) {
  HFT_ADJ_FAST
  __shared__ double sh[kBlock / 64];
  int t = xcd_remap(blockIdx.x, gridDim.x);
  int z, j, ct;
  tile_of(t, a.g, z, j, ct);
  int r = a.order[z];
  const Mesh &m = a.m;
  int c = ct * kBlock + threadIdx.x;
  bool col = c < m.nu;
  int cc = col ? c : 0;
  size_t base = (size_t)r * a.rstride;
  const double *l1 = a.lam1 + base;
  double Ln, Ls;
  pole_sums(m, l1, j, Ln, Ls);
  double lt = lap_T(m, a.D[r], l1, j, cc, Ln, Ls);
  double e = 0.0;
  if (col) {
    size_t p = base + (size_t)j * m.pitch + c;
    double w = a.ub1[r] * a.dt[r];
    a.L0[p] = a.L0[p] + w * lt;
    e = w * lt * a.u0[p];
  }
  double s = block_sum(sh, e);
  if (threadIdx.x == 0) {
    size_t slot = (size_t)r * (m.ntm * a.g.ctiles) + (size_t)j * a.g.ctiles + ct;
    a.part[slot] = a.part[slot] + s;
  }
}

// Per-realization sum of n partials in fixed order: out[r*ostride + oslot] += sum.
__global__ void k_sum_parts(const double *part, int n, int stride, int comp, int ncomp,
                            double *out, int ostride, int oslot) {
  __shared__ double sh[16];
  int r = blockIdx.x;
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x)
    s = s + part[(size_t)r * stride + (size_t)i * ncomp + comp];
  s = block_sum(sh, s);
  if (threadIdx.x == 0) out[r * ostride + oslot] = out[r * ostride + oslot] + s;
}

// ---------------------------------------------------------------------------
// WENO3 face flux gradient.  weno_face (forward) maps the split fluxes
// (LNm2, LNm1, LN0 | LPm1, LP0, LPp1) to the face flux F = um + up with
//   um = (w0 p0m + w1 p1m)/(w0 + w1),  w_i = C_i / (eps + 4 t_i^2)^2.
// dum = OM0 dp0m + OM1 dp1m + (p0m - um)/S dw0 + (p1m - um)/S dw1,
// dw_i/dt_i = -16 t_i w_i / e_i.  FROZEN drops the weight derivatives
// (the linearisation with frozen smoothness weights).
// ---------------------------------------------------------------------------
template <bool FROZEN>
__device__ inline void weno_face_grad(double LNm2, double LNm1, double LN0, double LPm1, double LP0,
                                      double LPp1, const WenoW &w, double g[6]) {
  HFT_ADJ_FAST
  const double one = 1.0, four = 4.0;
  double p0m = (one + w.mcm) * LNm1 - w.mcm * LNm2;
  double p1m = w.mc0 * LNm1 + w.cpm * LN0;
  double p0p = (one + w.cp0) * LP0 - w.cp0 * LPp1;
  double p1p = w.cpm * LP0 + w.mc0 * LPm1;
  double t0 = w.mcm * (LNm1 - LNm2), t1 = w.cpm * (LN0 - LNm1);
  double t2 = w.cp0 * (LPp1 - LP0), t3 = w.mc0 * (LP0 - LPm1);
  double i0 = one / (w.em + four * (t0 * t0)), i1 = one / (w.em + four * (t1 * t1));
  double i2 = one / (w.e0 + four * (t2 * t2)), i3 = one / (w.e0 + four * (t3 * t3));
  double w0m = w.ptm * (i0 * i0), w1m = w.mctm * (i1 * i1);
  double w0p = w.mt0 * (i2 * i2), w1p = w.cpt0 * (i3 * i3);
  double iSm = one / (w0m + w1m), iSp = one / (w0p + w1p);
  double O0m = w0m * iSm, O1m = w1m * iSm, O0p = w0p * iSp, O1p = w1p * iSp;
  double k0 = 0.0, k1 = 0.0, k2 = 0.0, k3 = 0.0;
  if (!FROZEN) {
    double um = O0m * p0m + O1m * p1m;
    double up = O0p * p0p + O1p * p1p;
    k0 = (p0m - um) * (-16.0 * t0 * O0m * i0);
    k1 = (p1m - um) * (-16.0 * t1 * O1m * i1);
    k2 = (p0p - up) * (-16.0 * t2 * O0p * i2);
    k3 = (p1p - up) * (-16.0 * t3 * O1p * i3);
  }
  // LN side: t0 = mcm (LNm1 - LNm2), t1 = cpm (LN0 - LNm1)
  g[0] = -w.mcm * (O0m + k0);
  g[1] = O0m * (one + w.mcm) + O1m * w.mc0 + k0 * w.mcm - k1 * w.cpm;
  g[2] = w.cpm * (O1m + k1);
  // LP side: t2 = cp0 (LPp1 - LP0), t3 = mc0 (LP0 - LPm1)
  g[3] = w.mc0 * (O1p - k3);
  g[4] = O0p * (one + w.cp0) + O1p * w.cpm - k2 * w.cp0 + k3 * w.mc0;
  g[5] = w.cp0 * (k2 - O0p);
}

// Vector-Jacobian product of one SSPRK(4,3) stage for 1-D (phi-independent)
// flows:  out = P + cb*lam - cdt * W'(x)^T lam,  acc = ca*lam (optional),
// part[r][band][0..1] += -cdt * (<lam, dW/dsigma_t>, <lam, dW/dsigma_p>).
// Layout as k_advection_band: one block = one realization x a band of theta
// rows over the full phi width.  Theta faces are visited once per block (3
// halo faces above the band); each face's adjoint psi (from the divergence
// and the pole closure) times its gradient is scattered into a 4-row register
// window of the column.  Phi faces of a row go through LDS.
struct AdvVjpArgs {
  Mesh m;
  Flow1D f1;
  const double *x, *lam, *P;
  double *out, *acc;
  double ca, cb, cdt;
  int nbands, H, ncol;
  size_t rstride;
  double *part;  // [nr][nbands][2]
};

template <bool FROZEN, bool HAS_P>
__global__ __launch_bounds__(1024) void k_advection_vjp_band(AdvVjpArgs a) {
  HFT_ADJ_FAST
  extern __shared__ double sh[];  // 9 x ncol + 16
  const Mesh &m = a.m;
  const int ncol = a.ncol, nu = m.nu, ntm = m.ntm, nt = m.nt, pitch = m.pitch;
  double *sLN = sh, *sLP = sh + ncol, *sLam = sh + 2 * ncol, *sG = sh + 3 * ncol;
  double *sRed = sh + 9 * ncol;
  int L = xcd_remap(blockIdx.x, gridDim.x);
  int r = L / a.nbands, band = L - r * a.nbands;
  int c = threadIdx.x;
  bool col = c < nu;
  int cc = col ? c : 0;
  size_t base = (size_t)r * a.rstride;
  const double *x = a.x + base;
  const double *lam = a.lam + base;
  const double *vt = a.f1.vt + (size_t)r * nt, *at = a.f1.at + (size_t)r * ntm;
  const double *vp = a.f1.vp + (size_t)r * ntm, *ap = a.f1.ap + (size_t)r * ntm;
  const double half = 0.5, one = 1.0;
  int j0 = band * a.H, j1 = min(ntm, j0 + a.H);
  double Ln = 0.0, Ls = 0.0;
  if (j0 <= 2) Ln = wave_pole_sum(nu, [&](int k) { return lam[k]; });
  if (j1 + 2 >= ntm) {
    const double *rs = lam + (size_t)(ntm - 1) * pitch;
    Ls = wave_pole_sum(nu, [&](int k) { return rs[k]; });
  }
  auto X = [&](int j) { return x[(size_t)j * pitch + cc]; };
  auto LAM = [&](int j) { return lam[(size_t)j * pitch + cc]; };
  auto dLN = [&](int j) { return half * (vt[j] + at[j]); };
  auto dLP = [&](int j) { return half * (vt[j + 1] - at[j]); };
  // theta row coefficients of the divergence (interior rows)
  auto adn = [&](int j) { return m.sth[j + 1] * m.st_i[j] * m.dt_i[j]; };
  auto aup = [&](int j) { return -m.sth[j] * m.st_i[j] * m.dt_i[j]; };
  double gt = 0.0, gp = 0.0;  // sensitivities of owned faces
  // Contribution of theta face q (between rows q and q+1) to rows q-1..q+2.
  auto tface = [&](int q, double cr[4], bool own) {
    cr[0] = cr[1] = cr[2] = cr[3] = 0.0;
    if (q < 0 || q > ntm - 2) return;
    double psi = 0.0;
    if (q >= 1) psi = psi + LAM(q) * adn(q);
    if (q + 1 <= ntm - 2) psi = psi + LAM(q + 1) * aup(q + 1);
    if (q == 0) psi = psi + Ln * m.dp[cc + 1] * m.bc_flow_n;
    if (q == ntm - 2) psi = psi - Ls * m.dp[cc + 1] * m.bc_flow_s;
    int jf = q + 2;  // reference face
    if (jf == 2 || jf == nt - 1) {
      double v = vt[jf - 1];
      double cct = fsign1(v);
      double ga = v * half * (one - cct), gb = v * half * (one + cct);  // x(q+1), x(q)
      cr[1] = psi * gb;
      cr[2] = psi * ga;
      if (own) gt = gt + psi * (ga * X(q + 1) + gb * X(q));
      return;
    }
    double ln[3], lp[3];
    for (int d = 0; d < 3; d++) {
      int jn = q - 1 + d, jp = q + d;
      ln[d] = half * X(jn) * (vt[jn] + at[jn]);
      lp[d] = half * X(jp) * (vt[jp + 1] - at[jp]);
    }
    WenoW w = weno_w(m.tCP, m.tMC, m.tMT, m.tCPT, m.tPT, m.tMCT, m.teps, jf - 2, jf - 1);
    double g[6];
    weno_face_grad<FROZEN>(ln[0], ln[1], ln[2], lp[0], lp[1], lp[2], w, g);
    cr[0] = psi * g[0] * dLN(q - 1);
    cr[1] = psi * (g[1] * dLN(q) + g[3] * dLP(q));
    cr[2] = psi * (g[2] * dLN(q + 1) + g[4] * dLP(q + 1));
    cr[3] = psi * g[5] * dLP(q + 2);
    if (own) {
      double s = g[0] * ln[0] + g[1] * ln[1] + g[2] * ln[2];
      s = s + (g[3] * lp[0] + g[4] * lp[1] + g[5] * lp[2]);
      gt = gt + psi * s;
    }
  };
  // window Y[i] = theta part of (W'^T lam) at row j + i
  double Y[4] = {0.0, 0.0, 0.0, 0.0};
  {
    double cr[4];
    tface(j0 - 2, cr, false);  // rows j0-3 .. j0
    Y[0] += cr[3];
    tface(j0 - 1, cr, false);  // rows j0-2 .. j0+1
    Y[0] += cr[2];
    Y[1] += cr[3];
    tface(j0, cr, true);  // rows j0-1 .. j0+2
    Y[0] += cr[1];
    Y[1] += cr[2];
    Y[2] += cr[3];
  }
  for (int j = j0; j < j1; j++) {
    double cr[4];
    tface(j + 1, cr, j + 1 < j1);  // rows j .. j+3
    Y[0] += cr[0];
    Y[1] += cr[1];
    Y[2] += cr[2];
    Y[3] += cr[3];
    double yv = Y[0];
    bool pole = (j == 0 || j == ntm - 1);
    if (!pole) {
      // phi split fluxes and adjoint of this row through LDS
      double xv = X(j);
      double vpj = vp[j], apj = ap[j];
      double lnc = half * xv * (vpj + apj), lpc = half * xv * (vpj - apj);
      sLN[c] = lnc;
      sLP[c] = lpc;
      sLam[c] = col ? LAM(j) : 0.0;
      lds_barrier();
      double sti = m.st_i[j];
      if (col) {
        int c2 = wrapc(cc - 2, nu), c1 = wrapc(cc - 1, nu), cp = wrapc(cc + 1, nu);
        double psi = sLam[c1] * sti * m.dp_i[c1 + 1] - sLam[cc] * sti * m.dp_i[cc + 1];
        WenoW w = weno_w(m.pCP, m.pMC, m.pMT, m.pCPT, m.pPT, m.pMCT, m.peps, cc, cc + 1);
        double g[6];
        double i0 = sLN[c2], i1 = sLN[c1], i2 = sLN[cc], i3 = sLP[c1], i4 = sLP[cc], i5 = sLP[cp];
        weno_face_grad<FROZEN>(i0, i1, i2, i3, i4, i5, w, g);
        for (int d = 0; d < 6; d++) sG[d * ncol + c] = psi * g[d];
        double s = g[0] * i0 + g[1] * i1 + g[2] * i2;
        s = s + (g[3] * i3 + g[4] * i4 + g[5] * i5);
        gp = gp + psi * s;
      } else {
        for (int d = 0; d < 6; d++) sG[d * ncol + c] = 0.0;
      }
      lds_barrier();
      if (col) {
        int cm = wrapc(cc - 1, nu), cp = wrapc(cc + 1, nu), cp2 = wrapc(cc + 2, nu);
        double sn = sG[2 * ncol + cc] + sG[1 * ncol + cp] + sG[0 * ncol + cp2];
        double sp = sG[5 * ncol + cm] + sG[4 * ncol + cc] + sG[3 * ncol + cp];
        yv = yv + (half * (vpj + apj) * sn + half * (vpj - apj) * sp);
      }
      lds_barrier();
    }
    if (col) {
      size_t off = (size_t)j * pitch + c;
      double lv = lam[off];
      double v = a.cb * lv - a.cdt * yv;
      if (HAS_P) v = a.P[base + off] + v;
      a.out[base + off] = v;
      if (a.acc) a.acc[base + off] = a.ca * lv;
    }
    Y[0] = Y[1];
    Y[1] = Y[2];
    Y[2] = Y[3];
    Y[3] = 0.0;
  }
  double st = block_sum(sRed, col ? gt : 0.0);
  double sp = block_sum(sRed, col ? gp : 0.0);
  if (threadIdx.x == 0) {
    double *pp = a.part + ((size_t)r * a.nbands + band) * 2;
    pp[0] = pp[0] - a.cdt * st;
    pp[1] = pp[1] - a.cdt * sp;
  }
}

// Start of a cycle adjoint for the realizations ord[0..n): lam_{s_r} = lam,
// adjoint of u0 = 0.
__global__ __launch_bounds__(kBlock) void k_adj_cycle_init(double *LB0, double *LB1, double *LB2,
                                                           double *L0, const double *lam,
                                                           const int *order, const int64_t *s,
                                                           size_t rstride) {
  size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= rstride) return;
  int r = order[blockIdx.y];
  size_t p = (size_t)r * rstride + i;
  double *LB[3] = {LB0, LB1, LB2};
  LB[s[r] % 3][p] = lam[p];
  L0[p] = 0.0;
}

// dst = src for the realizations ord[0..n).
__global__ __launch_bounds__(kBlock) void k_copy_reals(double *dst, const double *src,
                                                       const int *order, size_t rstride) {
  size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x;
  if (i >= rstride) return;
  size_t p = (size_t)order[blockIdx.y] * rstride + i;
  dst[p] = src[p];
}

// Pointwise helpers of the adjoint controller.
__global__ __launch_bounds__(kBlock) void k_fill(double *p, size_t n, double v) {
  size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x;
  if (i < n) p[i] = v;
}

}  // namespace hft
