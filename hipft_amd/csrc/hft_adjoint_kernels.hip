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

// ---------------------------------------------------------------------------
// Fused RKG2 adjoint: T consecutive transposed stages in one launch, over
// the stored stages of the forward recomputation.
//
// Realizations are indexed by the number of adjoint stages taken since they
// entered (relative step m = s_r - k): each realization's sweep starts at its
// own k = s_r with lam_{s_r} = the incoming adjoint, so in relative steps all
// start together and each finishes at m = s_r - 2 (stage k = 2), as the
// forward stages of a super-step start together and finish at s_r.  Level 0
// is lam at relative step m0 (buffer A); the pointwise term of level 1 reads
// relative step m0-1 (buffer B; absent at m0 = 0); level t = 1..T forms lam
// at relative step m0+t, i.e. lam_{k-1} of stage k = s_r - m0 - t + 1:
//   lam_{k-1} = a_k lam_k + d_k L^T lam_k + b_{k+1} lam_{k+1}   (k < s_r)
//   L0       += c_k lam_k + g_k dt L^T lam_k (+ b_2 lam_2 + lam_1 at k = 2)
//   e        += <L^T lam_k, d_k u_{k-1} + g_k dt u0>
// with u_{k-1} read from the stored stage (StageCoef::w).
//
// Block = (realization, band of H theta rows, full phi width), one thread
// per column, marching down the band and a T-row halo with a lag of one row
// per level, as k_rkg2_fused: L^T at row R of level t needs level t-1's rows
// R-1, R, R+1 of the own column (registers; R+1 formed earlier in the same
// step) and row R at the neighbouring columns (3-row LDS ring per level).
// Row 0 needs the sum of level t-1's row 0 (in LDS since the previous step);
// rows ntm-2 and ntm-1 need the sum of level t-1's row ntm-1, which for row
// ntm-2 is formed in the same step, so that step has one more barrier.  The
// L0 accumulators of the T rows in flight and u0 of those rows rotate in
// registers; L0 of an own row is read once (not at m0 = 0, where it starts
// at zero) and written once, after the last level.  e is summed per thread
// and per block in a fixed order into the slot of the band's first row.
// ---------------------------------------------------------------------------
struct AdjFusedArgs {
  Mesh m;
  int nbands, H, ncol, ctiles;
  int64_t m0;
  const double *A;    // lam at relative step m0
  const double *B;    // lam at relative step m0-1 (read where m0 > 0)
  double *outA;       // lam at relative step m0+T
  double *outB;       // lam at relative step m0+T-1
  double *L1;         // lam_1 of the realizations that finish in this launch
  double *L0;         // adjoint of u0
  const double *u0;
  const double *D;
  const StageCoef *coef;  // table [k][nr]; .w = stored stage u_{k-1}
  int nr_table;
  const int64_t *s;
  const double *dt;   // [nr] cycle time step
  const int *order;
  size_t rstride;
  double *part;       // [nr][ntm*ctiles]
};

template <int T>
struct AdjFusedState {
  double w[T][4];      // levels 0..T-1, own column, slot = age-rotated
  double hu[4];        // u0 of rows q-1-d
  double acc[4];       // L0 accumulators of rows q-1-d
  double pf[2][4 + T]; // prefetched inputs of the next two steps
};

template <int T, int U>
__device__ __forceinline__ void adj_fused_step(const AdjFusedArgs &a, AdjFusedState<T> &S,
                                               double *lds, const double *rowc,
                                               const double (*cf)[6], int i, int lo0, int j0,
                                               int j1, int hi0, int nst, bool finishing,
                                               bool col, int c, int cc, int cw, int ce,
                                               double P1, double P5, double W5, double E1,
                                               double pnc, double psc, const double *A,
                                               const double *B, const double *U0,
                                               const double *L0, const double *const *Ut,
                                               double *OA, double *OB, double *OL1, double *OL0,
                                               double &esum) {
  HFT_ADJ_FAST
  const Mesh &m = a.m;
  const int ncol = a.ncol, nu = m.nu, ntm = m.ntm, pitch = m.pitch;
  // slot of age d at this step: (U - d) & 3
  constexpr int s0 = U & 3, s1 = (U + 3) & 3, s2 = (U + 2) & 3, s3 = (U + 1) & 3;
  constexpr int ps = U & 1;
  const int q = lo0 + i;
  double pv[4 + T];
#pragma unroll
  for (int e = 0; e < 4 + T; e++) pv[e] = S.pf[ps][e];
  {
    // prefetch step i+2 into the slot just consumed; rows clamped into the
    // field (values of rows outside the band's range never reach an output)
    int qa = min(q + 2, ntm - 1), qb = min(max(q + 1, 0), ntm - 1);
    S.pf[ps][0] = A[(size_t)qa * pitch + cc];
    S.pf[ps][1] = B[(size_t)qb * pitch + cc];
    S.pf[ps][2] = U0[(size_t)qb * pitch + cc];
    S.pf[ps][3] = L0[(size_t)qb * pitch + cc];
#pragma unroll
    for (int t = 1; t <= T; t++) {
      int qt = min(max(q + 2 - t, 0), ntm - 1);
      S.pf[ps][3 + t] = Ut[t - 1][(size_t)qt * pitch + cc];
    }
  }
  // level 0 (row q); u0 and the L0 accumulator of row q-1 (age 0)
  S.w[0][s0] = pv[0];
  if (q < hi0) lds[(size_t)(q % 3) * ncol + c] = pv[0];
  S.hu[s0] = pv[2];
  S.acc[s0] = a.m0 == 0 ? 0.0 : pv[3];
#pragma unroll
  for (int t = 1; t <= T; t++) {
    const int R = q - t;
    const int rlo = max(0, j0 - T + t), rhi = min(ntm, j1 + T - t);
    const double *lp = lds + (size_t)((t - 1) * 3) * ncol;  // level t-1 ring
    constexpr int dummy = 0;
    (void)dummy;
    const int as = (t == 1) ? s0 : (t == 2 ? s1 : (t == 3 ? s2 : s3));  // age t-1
    // row ntm-2 needs the sum of level t-1's row ntm-1, written in this step
    if (R == ntm - 2 && t <= nst) lds_barrier();
    bool act = t <= nst && R >= rlo && R < rhi;
    double v = 0.0;
    if (act) {
      const double xc = S.w[t - 1][s1];  // row R
      const double xs = S.w[t - 1][s0];  // row R+1
      const double xn = S.w[t - 1][s2];  // row R-1
      const double *rr = rowc + 4 * (R - lo0 + 1);  // {sti^2, c2, c4, 0} of row R
      double lt;
      if (R == 0) {
        double Ln = wave_pole_sum(nu, [&](int k) { return lp[k]; });
        lt = rr[5] * xs - pnc * Ln;
      } else if (R == ntm - 1) {
        const double *rs = lp + (size_t)((ntm - 1) % 3) * ncol;
        double Ls = wave_pole_sum(nu, [&](int k) { return rs[k]; });
        lt = rr[-2] * xn - psc * Ls;
      } else {
        const double *rc = lp + (size_t)(R % 3) * ncol;
        const double s2v = rr[0];
        double c1 = P1 * s2v, c5 = P5 * s2v;
        double c3 = -(((c1 + rr[1]) + rr[2]) + c5);
        lt = c3 * xc;
        lt = lt + (W5 * s2v) * rc[cw];
        lt = lt + (E1 * s2v) * rc[ce];
        lt = lt + rr[-2] * xn;  // c4(R-1) (zero at R = 1)
        lt = lt + rr[5] * xs;   // c2(R+1) (zero at R = ntm-2)
        if (R == 1) lt = lt + pnc * wave_pole_sum(nu, [&](int k) { return lp[k]; });
        if (R == ntm - 2) {
          const double *rs = lp + (size_t)((ntm - 1) % 3) * ncol;
          lt = lt + psc * wave_pole_sum(nu, [&](int k) { return rs[k]; });
        }
      }
      const double *k = cf[t - 1];  // a, d, c (+ b at k = 2), g dt, b_{k+1}, [k == 2]
      const double bnv = (t == 1) ? pv[1] : S.w[t >= 2 ? t - 2 : 0][s2];  // lam_{k+1} at row R
      v = k[0] * xc + k[1] * lt;
      v = v + k[4] * bnv;
      double acc = S.acc[as] + k[2] * xc + k[3] * lt;
      if (k[5] != 0.0) acc = acc + v;
      S.acc[as] = acc;
      if (R >= j0 && R < j1) {
        esum = esum + lt * (k[1] * pv[3 + t] + k[3] * S.hu[as]);
        if (col) {
          size_t o = (size_t)R * pitch + c;
          if (t == nst) {
            if (finishing) OL1[o] = v;
            else OA[o] = v;
            OL0[o] = acc;
          } else if (t == T - 1 && !finishing) {
            OB[o] = v;
          }
        }
      }
    }
    if (t < T) {
      S.w[t < T ? t : 0][s0] = v;
      if (act) lds[(size_t)(t * 3 + (R % 3)) * ncol + c] = v;
    }
  }
  lds_barrier();
}

template <int T, int NT>
__global__ __launch_bounds__(NT) void k_rkg2_adj_fused(AdjFusedArgs a) {
  HFT_ADJ_FAST
  extern __shared__ double lds[];  // [T][3][ncol]
  __shared__ double cf[T][6];
  __shared__ double rowc[4 * (kFusedMaxRows + 2)];
  __shared__ double sred[16];
  const Mesh &m = a.m;
  const int nu = m.nu, ntm = m.ntm;
  int L = xcd_remap(blockIdx.x, gridDim.x);
  int z = L / a.nbands, band = L - z * a.nbands;
  int r = a.order[z];
  int c = threadIdx.x;
  bool col = c < nu;
  int cc = col ? c : 0;
  int cw = wrapc(cc - 1, nu), ce = wrapc(cc + 1, nu);
  int j0 = band * a.H, j1 = min(ntm, j0 + a.H);
  int lo0 = max(0, j0 - T), hi0 = min(ntm, j1 + T);
  int64_t sr = a.s[r];
  int64_t left = sr - 1 - a.m0;  // stages still to take (>= 1 in an active block)
  int nst = (int)(left < T ? left : T);
  bool finishing = left <= T;
  size_t base = (size_t)r * a.rstride;
  double D = a.D[r];
  double dtr = a.dt[r];
  // c1 and c5 of the own column and of the neighbours without sin^-2 theta
  double P1 = (D * m.dph_i[cc + 1]) * m.dp_i[cc + 1];
  double P5 = (D * m.dph_i[cc + 2]) * m.dp_i[cc + 1];
  double W5 = (D * m.dph_i[cw + 2]) * m.dp_i[cw + 1];
  double E1 = (D * m.dph_i[ce + 1]) * m.dp_i[ce + 1];
  double D2 = D + D;
  double pnc = D2 * m.dp[cc + 1] * m.bc_diff_n, psc = D2 * m.dp[cc + 1] * m.bc_diff_s;
  if (threadIdx.x < T) {
    int t = threadIdx.x + 1;
    double v[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (t <= nst) {
      int64_t k = sr - a.m0 - t + 1;
      const StageCoef &q = a.coef[(size_t)k * a.nr_table + r];
      v[0] = q.a;
      v[1] = q.d;
      v[2] = k == 2 ? q.c + q.b : q.c;
      v[3] = q.g * dtr;
      v[4] = (a.m0 + t - 1 > 0) ? a.coef[(size_t)(k + 1) * a.nr_table + r].b : 0.0;
      v[5] = k == 2 ? 1.0 : 0.0;
    }
#pragma unroll
    for (int e = 0; e < 6; e++) cf[t - 1][e] = v[e];
  }
  // rows lo0-1 .. hi0: sin^-2 theta and the theta coefficients c2, c4 (zero
  // outside the interior rows)
  for (int idx = threadIdx.x; idx < hi0 - lo0 + 2; idx += blockDim.x) {
    int R = lo0 - 1 + idx;
    double s2v = 0.0, c2 = 0.0, c4 = 0.0;
    if (R >= 1 && R <= ntm - 2) {
      double sti = m.st_i[R];
      s2v = sti * sti;
      c2 = D * m.dth_i[R] * m.dt_i[R] * sti * m.sth[R];
      c4 = D * m.dth_i[R + 1] * m.dt_i[R] * sti * m.sth[R + 1];
    }
    rowc[4 * idx] = s2v;
    rowc[4 * idx + 1] = c2;
    rowc[4 * idx + 2] = c4;
    rowc[4 * idx + 3] = 0.0;
  }
  const double *A = a.A + base, *B = a.B + base, *U0 = a.u0 + base;
  double *L0 = a.L0 + base;
  double *OA = a.outA + base, *OB = a.outB + base, *OL1 = a.L1 + base;
  const double *Ut[T];
#pragma unroll
  for (int t = 1; t <= T; t++) {
    const double *w = nullptr;
    if (t <= nst) w = a.coef[(size_t)(sr - a.m0 - t + 1) * a.nr_table + r].w;
    Ut[t - 1] = (w ? w : a.u0) + base;  // levels beyond nst read u0 (unused)
  }
  AdjFusedState<T> S;
#pragma unroll
  for (int d = 0; d < 4; d++) {
    S.hu[d] = 0.0;
    S.acc[d] = 0.0;
#pragma unroll
    for (int t = 0; t < T; t++) S.w[t][d] = 0.0;
  }
#pragma unroll
  for (int i = 0; i < 2; i++) {
    int q = lo0 + i;
    int qa = min(q, ntm - 1), qb = min(max(q - 1, 0), ntm - 1);
    S.pf[i][0] = A[(size_t)qa * m.pitch + cc];
    S.pf[i][1] = B[(size_t)qb * m.pitch + cc];
    S.pf[i][2] = U0[(size_t)qb * m.pitch + cc];
    S.pf[i][3] = L0[(size_t)qb * m.pitch + cc];
#pragma unroll
    for (int t = 1; t <= T; t++) {
      int qt = min(max(q - t, 0), ntm - 1);
      S.pf[i][3 + t] = Ut[t - 1][(size_t)qt * m.pitch + cc];
    }
  }
  lds_barrier();
  double esum = 0.0;
  int nsteps = (hi0 - lo0) + T;
  for (int i = 0; i < nsteps; i += 4) {
    adj_fused_step<T, 0>(a, S, lds, rowc, cf, i, lo0, j0, j1, hi0, nst, finishing, col, c, cc, cw,
                         ce, P1, P5, W5, E1, pnc, psc, A, B, U0, L0, Ut, OA, OB, OL1, L0, esum);
    if (i + 1 < nsteps)
      adj_fused_step<T, 1>(a, S, lds, rowc, cf, i + 1, lo0, j0, j1, hi0, nst, finishing, col, c,
                           cc, cw, ce, P1, P5, W5, E1, pnc, psc, A, B, U0, L0, Ut, OA, OB, OL1,
                           L0, esum);
    if (i + 2 < nsteps)
      adj_fused_step<T, 2>(a, S, lds, rowc, cf, i + 2, lo0, j0, j1, hi0, nst, finishing, col, c,
                           cc, cw, ce, P1, P5, W5, E1, pnc, psc, A, B, U0, L0, Ut, OA, OB, OL1,
                           L0, esum);
    if (i + 3 < nsteps)
      adj_fused_step<T, 3>(a, S, lds, rowc, cf, i + 3, lo0, j0, j1, hi0, nst, finishing, col, c,
                           cc, cw, ce, P1, P5, W5, E1, pnc, psc, A, B, U0, L0, Ut, OA, OB, OL1,
                           L0, esum);
  }
  double s = block_sum(sred, col ? esum : 0.0);
  if (threadIdx.x == 0) {
    size_t slot = (size_t)r * (ntm * a.ctiles) + (size_t)j0 * a.ctiles;
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
// 1/x for the adjoint's WENO3 weights, whose arguments are at least the
// positive smoothness epsilon (a cell width) and far from the subnormal
// range: the hardware reciprocal refined by two Newton steps (each doubles
// the correct bits), within about one ulp, in place of the correctly rounded
// divide sequence the forward kernels need.
__device__ inline double rcp_nr(double x) {
  double r = __builtin_amdgcn_rcp(x);
  double e = __builtin_fma(-x, r, 1.0);
  r = __builtin_fma(r, e, r);
  e = __builtin_fma(-x, r, 1.0);
  return __builtin_fma(r, e, r);
}

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
  double i0 = rcp_nr(w.em + four * (t0 * t0)), i1 = rcp_nr(w.em + four * (t1 * t1));
  double i2 = rcp_nr(w.e0 + four * (t2 * t2)), i3 = rcp_nr(w.e0 + four * (t3 * t3));
  double w0m = w.ptm * (i0 * i0), w1m = w.mctm * (i1 * i1);
  double w0p = w.mt0 * (i2 * i2), w1p = w.cpt0 * (i3 * i3);
  double iSm = rcp_nr(w0m + w1m), iSp = rcp_nr(w0p + w1p);
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

template <bool FROZEN, bool HAS_P, int NT>
__global__ __launch_bounds__(NT) void k_advection_vjp_band(AdvVjpArgs a) {
  HFT_ADJ_FAST
  // LN, LP, lam and the six face-gradient terms of a row (9 x ncol), the
  // reduction scratch (16), then the phi WENO weights of each column's west
  // face (5 x ncol pairs; each thread reads only its own)
  extern __shared__ double sh[];
  const Mesh &m = a.m;
  const int ncol = a.ncol, nu = m.nu, ntm = m.ntm, nt = m.nt, pitch = m.pitch;
  double *sLN = sh, *sLP = sh + ncol, *sLam = sh + 2 * ncol, *sG = sh + 3 * ncol;
  double *sRed = sh + 9 * ncol;
  double2 *sW = reinterpret_cast<double2 *>(sh + 9 * ncol + 16);
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
  {
    WenoW w = weno_w(m.pCP, m.pMC, m.pMT, m.pCPT, m.pPT, m.pMCT, m.peps, cc, cc + 1);
    sW[0 * ncol + c] = make_double2(w.mcm, w.cpm);
    sW[1 * ncol + c] = make_double2(w.mc0, w.cp0);
    sW[2 * ncol + c] = make_double2(w.ptm, w.mctm);
    sW[3 * ncol + c] = make_double2(w.mt0, w.cpt0);
    sW[4 * ncol + c] = make_double2(w.em, w.e0);
  }
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
        WenoW w;
        double2 q0 = sW[0 * ncol + c], q1 = sW[1 * ncol + c], q2 = sW[2 * ncol + c];
        double2 q3 = sW[3 * ncol + c], q4 = sW[4 * ncol + c];
        w.mcm = q0.x; w.cpm = q0.y; w.mc0 = q1.x; w.cp0 = q1.y; w.ptm = q2.x;
        w.mctm = q2.y; w.mt0 = q3.x; w.cpt0 = q3.y; w.em = q4.x; w.e0 = q4.y;
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
