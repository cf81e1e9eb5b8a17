// hft_kernels.hip -- HIP kernels of the flux-transport step for gfx950 (FP64).
//
// Every kernel works on whole batches of realizations.  Work is tiled as
// (realization, theta row, block of phi columns); the 1-D launch index is
// remapped so that each of the 8 XCDs receives a contiguous range of tiles
// (neighbouring rows of one realization), which keeps the stencil's row
// neighbours in that XCD's L2.
//
// Floating-point expressions follow the reference's operation order and the
// library is built with -ffp-contract=off, so results match the CPU
// restatement of src/hipft.f90 bit for bit (polar sums use the fixed order
// of wave_pole_sum).
#include "hft_device.h"

namespace hft {

static constexpr int kBlock = 256;

// Tile decomposition -------------------------------------------------------
struct Tiles {
  int ctiles;  // column tiles per row
  int ntm;
  int nreal;   // realizations in the launch
};

__device__ inline int xcd_remap(int L, int T) {
  // Blocks are dispatched round-robin over 8 XCDs (block L -> XCD L%8);
  // give XCD x the contiguous tile range [start_x, start_x + size_x).
  int per = T >> 3, rem = T & 7;
  int x = L & 7, slot = L >> 3;
  return x * per + (x < rem ? x : rem) + slot;
}

__device__ inline void tile_of(int t, const Tiles &g, int &z, int &j, int &ct) {
  ct = t % g.ctiles;
  int rj = t / g.ctiles;
  j = rj % g.ntm;
  z = rj / g.ntm;
}

__device__ inline int wrapc(int c, int nu) {
  return c < 0 ? c + nu : (c >= nu ? c - nu : c);
}

// Sum over v(c), c = 0..nu-1, in the fixed order: lane l of a wavefront
// accumulates c = l, l+64, ... in increasing c, then a butterfly combines
// partials (l, l+32), (l, l+16), ... (bitwise the halving tree of the CPU
// restatement).  Every lane returns the same value.
template <typename F>
__device__ inline double wave_pole_sum(int nu, F v) {
  int lane = threadIdx.x & 63;
  double s = 0.0;
  for (int c = lane; c < nu; c += 64) s = s + v(c);
  for (int h = 32; h >= 1; h >>= 1) s = s + __shfl_xor(s, h, 64);
  return s;
}

// ---------------------------------------------------------------------------
// Diffusion: 5-point surface Laplacian with polar closure
// (diffusion_operator_cd, src/hipft.f90:7927-7990) for a constant
// per-realization diffusivity D_r (coef from load_diffusion_matrix,
// src/hipft.f90:4960-4981, recomputed in registers in reference order).
// ---------------------------------------------------------------------------
__device__ inline double lap_interior(const Mesh &m, double D, const double *x, int j, int c,
                                      int nu, int pitch) {
  // reference indices: jj = j+1, k = c+2
  const double *row = x + (size_t)j * pitch;
  double xc = row[c];
  double xw = row[wrapc(c - 1, nu)];
  double xe = row[wrapc(c + 1, nu)];
  double xn = row[c - pitch];
  double xs = row[c + pitch];
  double sti = m.st_i[j];
  double dph_lo = m.dph_i[c + 1], dph_hi = m.dph_i[c + 2], dpi = m.dp_i[c + 1];
  double c1 = D * dph_lo;
  c1 = c1 * dpi;
  c1 = c1 * sti;
  c1 = c1 * sti;
  double c2 = D * m.dth_i[j];
  c2 = c2 * m.dt_i[j];
  c2 = c2 * sti;
  c2 = c2 * m.sth[j];
  double c4 = D * m.dth_i[j + 1];
  c4 = c4 * m.dt_i[j];
  c4 = c4 * sti;
  c4 = c4 * m.sth[j + 1];
  double c5 = D * dph_hi;
  c5 = c5 * dpi;
  c5 = c5 * sti;
  c5 = c5 * sti;
  double c3 = -(((c1 + c2) + c4) + c5);
  double y = c1 * xw;
  y = y + c2 * xn;
  y = y + c3 * xc;
  y = y + c4 * xs;
  y = y + c5 * xe;
  return y;
}

__device__ inline double lap_pole(const Mesh &m, double D, const double *x, bool north,
                                  int pitch, int ntm, int nu) {
  double D2 = D + D;
  if (north) {
    const double *r0 = x, *r1 = x + pitch;
    double fn = wave_pole_sum(nu, [&](int c) { return (D2 * (r1[c] - r0[c])) * m.dp[c + 1]; });
    return fn * m.bc_diff_n;
  }
  const double *r0 = x + (size_t)(ntm - 1) * pitch, *r1 = x + (size_t)(ntm - 2) * pitch;
  double fs = wave_pole_sum(nu, [&](int c) { return (D2 * (r0[c] - r1[c])) * m.dp[c + 1]; });
  return -fs * m.bc_diff_s;
}

__device__ inline double laplacian(const Mesh &m, double D, const double *x, int j, int c,
                                   int ntm, int nu, int pitch) {
  if (j == 0) return lap_pole(m, D, x, true, pitch, ntm, nu);
  if (j == ntm - 1) return lap_pole(m, D, x, false, pitch, ntm, nu);
  return lap_interior(m, D, x, j, c, nu, pitch);
}

// RKG2 first stage (src/hipft.f90:5211-5218):
//   dty0 = dt*L(f);  u1 = f + ubj(1)*dty0
struct Stage1Args {
  Mesh m;
  Tiles g;
  const double *f;
  double *dty0;
  double *u1;
  const double *D;       // [nr] diffusivity in code units
  const double *dt;      // [nr] cycle time step
  const double *ub1;     // [nr] sts_ubj(1)
  const int *order;      // launch index -> realization
  const double *Af;      // optional precomputed L(f) (PTL pass), may be null
  size_t rstride;
  // device count of the launch's realizations (g.nreal is then the most
  // there may be): blocks past that count return at once
  const int *dn = nullptr;
};

// First tile index of a block, with the XCD remap over the tiles that exist:
// n realizations x tiles per realization when a device count is given (the
// blocks beyond return), else the whole grid.  Returns -1 for such a block.
__device__ inline int tile_index(const Tiles &g, const int *dn) {
  int nt = (int)gridDim.x;
  if (dn) {
    nt = *dn * g.ntm * g.ctiles;
    if ((int)blockIdx.x >= nt) return -1;
  }
  return xcd_remap(blockIdx.x, nt);
}

__global__ __launch_bounds__(kBlock) void k_rkg2_stage1(Stage1Args a) {
  int t = tile_index(a.g, a.dn);
  if (t < 0) return;
  int z, j, ct;
  tile_of(t, a.g, z, j, ct);
  int r = a.order[z];
  const Mesh &m = a.m;
  int c = ct * kBlock + threadIdx.x;
  size_t base = (size_t)r * a.rstride;
  const double *x = a.f + base;
  double L;
  if (a.Af) {
    L = (c < m.nu) ? a.Af[base + (size_t)j * m.pitch + c] : 0.0;
  } else {
    // pole rows: every lane takes part in the wave sum
    L = (j == 0 || j == m.ntm - 1 || c < m.nu) ? laplacian(m, a.D[r], x, j, c < m.nu ? c : 0,
                                                           m.ntm, m.nu, m.pitch)
                                               : 0.0;
  }
  if (c >= m.nu) return;
  size_t p = base + (size_t)j * m.pitch + c;
  double d0 = a.dt[r] * L;
  a.dty0[p] = d0;
  a.u1[p] = x[(size_t)j * m.pitch + c] + a.ub1[r] * d0;
}

// RKG2 stage k >= 2 (src/hipft.f90:5222-5235).  u_{k-2} is overwritten in
// place (only the point itself reads it); the last stage of each
// realization writes straight into the state array (which also holds u0).
struct StageArgs {
  Mesh m;
  Tiles g;
  const double *x;      // u_{k-1}
  const double *ukm2;   // u_{k-2}
  double *out;          // u_k unless last stage
  double *f;            // state (u0), receives u_s
  const double *dty0;
  const double *D;
  const StageCoef *coef;  // coefficients of this stage, [nr]
  const int64_t *s;       // [nr] stage counts
  const int *order;
  int k;
  size_t rstride;
};

__global__ __launch_bounds__(kBlock) void k_rkg2_stage(StageArgs a) {
  int t = xcd_remap(blockIdx.x, gridDim.x);
  int z, j, ct;
  tile_of(t, a.g, z, j, ct);
  int r = a.order[z];
  const Mesh &m = a.m;
  int c = ct * kBlock + threadIdx.x;
  size_t base = (size_t)r * a.rstride;
  const double *x = a.x + base;
  double L = (j == 0 || j == m.ntm - 1 || c < m.nu)
                 ? laplacian(m, a.D[r], x, j, c < m.nu ? c : 0, m.ntm, m.nu, m.pitch)
                 : 0.0;
  if (c >= m.nu) return;
  size_t off = (size_t)j * m.pitch + c;
  size_t p = base + off;
  const StageCoef cf = a.coef[r];
  double v = cf.a * x[off];
  v = v + cf.b * a.ukm2[p];
  v = v + cf.c * a.f[p];
  v = v + cf.d * L;
  v = v + cf.g * a.dty0[p];
  double *o = (a.k == a.s[r]) ? a.f : a.out;
  o[p] = v;
}

// Final stages of a super-step into f: realization real[z] is copied from
// src[which[z]], z = 0..n-1, in 16-byte words (rstride is even).  One launch
// replaces a device-to-device copy per realization.
constexpr int kCopyMax = 256;
struct CopyFinalsArgs {
  double *dst;
  const double *src[4];
  size_t rstride;
  int n, per;  // realizations, blocks per realization
  uint16_t real[kCopyMax];
  uint8_t which[kCopyMax];
};

__global__ __launch_bounds__(kBlock) void k_copy_finals(CopyFinalsArgs a) {
  int z = blockIdx.x / a.per;
  size_t i = (size_t)(blockIdx.x - z * a.per) * kBlock + threadIdx.x;
  if (i >= a.rstride / 2) return;
  size_t p = (size_t)a.real[z] * a.rstride + 2 * i;
  *reinterpret_cast<double2 *>(a.dst + p) =
      *reinterpret_cast<const double2 *>(a.src[a.which[z]] + p);
}

// ---------------------------------------------------------------------------
// Fused RKG2 stages: T consecutive stages k0..k0+T-1 in one launch.
//
// A block owns one realization and a band of H theta rows over the full phi
// width (one thread per phi column, nu <= 1024).  It marches down the band
// plus a T-row halo above and below: at march step i it loads row q of
// u_{k0-1} (level 0) and stage t = 1..T computes row q - t of its level from
// level t-1.  Level t-1's rows R-1, R, R+1 at the thread's own column are in
// registers (R+1 was computed earlier in the same step); row R at the
// neighbouring columns c-1, c+1 and the full rows the polar sums need come
// from a 3-row LDS ring per level, written in an earlier step (one barrier
// per step).  The north-pole row of stage t needs level t-1's rows 0 and 1
// at every column, so it is formed one step late, together with row 1.
//
// Register windows and histories are circular buffers whose slots rotate
// with the step index; the march is unrolled by 4 so every slot index is a
// compile-time constant (no register moves).
//
// Only u_{k0-1}, u_{k0-2}, u0 and dty0 are read and the last two stages
// written, so a launch moves about 4 + 2 fields (plus halo) for T stages
// instead of 5T.  Expressions and the polar-sum order are those of
// k_rkg2_stage, so the results are bit-identical.
// ---------------------------------------------------------------------------
struct FusedArgs {
  Mesh m;
  int nbands, H, k0, ncol;
  const double *A;     // u_{k0-1}
  const double *B;     // u_{k0-2}
  const double *f;     // u0 (not written: halo rows are shared)
  const double *dty0;
  double *outA;        // u_{k0+T-1}, or u_s of a realization that finishes
  double *outB;        // u_{k0+T-2}
  const double *D;
  const StageCoef *coef;  // table [k][nr]
  int nr_table;
  const int64_t *s;
  const int *order;
  size_t rstride;
  double *W[5];        // optional: stage k0+t-1 rows into W[t-1] (null: not kept)
  const double *cdt;   // [nr] cycle time step (k_rkg2_fused2 forms dty0 = dt L(u0) itself)
  int body;            // k_rkg2_fused2: interior steps without row tests (fused2_body_step)
  // k_rkg2_fused2 launched from a device plan: {tiles, H, nbands} of this
  // launch in device memory (nbands, H and the grid size above are then
  // ignored; blocks past the tile count return at once)
  const int *slot = nullptr;
};

constexpr int kFusedMaxRows = 160;  // H + 2T rows of per-row coefficients

template <int T>
struct FusedState {
  double w[T][4];     // level t (0..T-1) own-column values, slot = age-rotated
  double p0[T];       // level t, north-pole row 0 (formed late, kept aside)
  double hu[4], hd[4];  // u0 and dty0 at rows q-1-d
  double pf[2][4];    // prefetched inputs of the next two steps
};

template <int T, int U>
__device__ __forceinline__ void fused_step(const FusedArgs &a, FusedState<T> &S, double *lds,
                                           const double *rowc, int i, int nsteps, int lo0,
                                           int hi0, int j0, int j1, int nst, bool final_here,
                                           bool col, int c, int cw, int ce, double D2,
                                           double P1, double P5, const double *A,
                                           const double *B, const double *U0,
                                           const double *DT, double *OA, double *OB,
                                           double u0_pole, double dt_pole,
                                           const double (*cfs)[5], double *const *Wst) {
  const Mesh &m = a.m;
  const int ncol = a.ncol, nu = m.nu, ntm = m.ntm, pitch = m.pitch;
  // slot of age d at this sub-step: (U - d) & 3
  constexpr int s0 = U & 3, s1 = (U + 3) & 3, s2 = (U + 2) & 3;
  constexpr int p0s = U & 1, p1s = (U + 1) & 1;
  int q = lo0 + i;
  double a0 = S.pf[p0s][0], b1 = S.pf[p0s][1], u1 = S.pf[p0s][2], d1 = S.pf[p0s][3];
  {
    // prefetch step i+2 into the slot just consumed
    int qn = q + 2, rn = qn - 1;
    double v0 = 0.0, v1 = 0.0, v2 = 0.0, v3 = 0.0;
    if (col && i + 2 < nsteps) {
      if (qn < hi0) v0 = A[(size_t)qn * pitch + c];
      if (rn >= 0 && rn < ntm) {
        size_t o = (size_t)rn * pitch + c;
        v1 = B[o];
        v2 = U0[o];
        v3 = DT[o];
      }
    }
    S.pf[p0s][0] = v0;
    S.pf[p0s][1] = v1;
    S.pf[p0s][2] = v2;
    S.pf[p0s][3] = v3;
  }
  (void)p1s;
  // u0/dty0 history: slot of row q-1-d is (U - d) & 3
  S.hu[s0] = u1;
  S.hd[s0] = d1;
  // level 0
  S.w[0][s0] = a0;
  if (q < hi0) {
    lds[(size_t)(0 * 3 + (q % 3)) * ncol + c] = a0;
    if (q == 0) S.p0[0] = a0;
  }
#pragma unroll
  for (int t = 1; t <= T; t++) {
    int R = q - t;
    int rlo = max(0, j0 - T + t), rhi = min(ntm, j1 + T - t);
    const double *lp = lds + (size_t)((t - 1) * 3) * ncol;  // level t-1 ring
    const double *k = cfs[t - 1];
    // u0/dty0 at row R = q - t: age t-1
    double u0v = (t == 1) ? S.hu[s0] : (t == 2 ? S.hu[s1] : (t == 3 ? S.hu[s2] : S.hu[(U + 1) & 3]));
    double dtv = (t == 1) ? S.hd[s0] : (t == 2 ? S.hd[s1] : (t == 3 ? S.hd[s2] : S.hd[(U + 1) & 3]));
    bool act = t <= nst && R >= rlo && R < rhi;
    double v = 0.0;
    if (act && R == 1 && rlo == 0) {
      // late north-pole row 0 of stage t, from level t-1 rows 0 and 1 in LDS
      const double *r0 = lp + 0 * ncol, *r1p = lp + 1 * ncol;
      double fn = wave_pole_sum(nu, [&](int kk) { return (D2 * (r1p[kk] - r0[kk])) * m.dp[kk + 1]; });
      double Lv = fn * m.bc_diff_n;
      double xc = (t == 1) ? S.p0[0] : S.p0[t - 1];
      double ukm2 = (t == 1) ? B[c < nu ? c : 0] : S.p0[t - 2];
      double vv = k[0] * xc;
      vv = vv + k[1] * ukm2;
      vv = vv + k[2] * u0_pole;
      vv = vv + k[3] * Lv;
      vv = vv + k[4] * dt_pole;
      if (t < T) S.p0[t] = vv;
      if (t < T) lds[(size_t)(t * 3 + 0) * ncol + c] = vv;
      if (col && 0 >= j0 && 0 < j1) {
        if (t == nst) OA[c] = vv;
        if (!final_here && t == T - 1) OB[c] = vv;
        if (Wst[t - 1]) Wst[t - 1][c] = vv;
      }
    }
    if (act && R > 0) {
      double Lv;
      if (R == ntm - 1) {
        const double *r0 = lp + ((ntm - 1) % 3) * ncol, *r1p = lp + ((ntm - 2) % 3) * ncol;
        double fs = wave_pole_sum(nu, [&](int kk) { return (D2 * (r0[kk] - r1p[kk])) * m.dp[kk + 1]; });
        Lv = -fs * m.bc_diff_s;
      } else {
        const double *rc = lp + (R % 3) * ncol;
        double xc = S.w[t - 1][s1];
        double xs = S.w[t - 1][s0];
        double xn = (R == 1) ? S.p0[t - 1] : S.w[t - 1][s2];
        double xw = rc[cw], xe = rc[ce];
        const double *rr = rowc + 4 * (R - lo0);
        double sti = rr[0];
        double c1 = (P1 * sti) * sti;
        double c2 = rr[1];
        double c4 = rr[2];
        double c5 = (P5 * sti) * sti;
        double c3 = -(((c1 + c2) + c4) + c5);
        Lv = c1 * xw;
        Lv = Lv + c2 * xn;
        Lv = Lv + c3 * xc;
        Lv = Lv + c4 * xs;
        Lv = Lv + c5 * xe;
      }
      double xc = S.w[t - 1][s1];
      double ukm2;
      if (t == 1) ukm2 = b1;
      else if (R == 1 && t >= 2) ukm2 = S.w[t - 2][s2];
      else ukm2 = S.w[t - 2][s2];
      v = k[0] * xc;
      v = v + k[1] * ukm2;
      v = v + k[2] * u0v;
      v = v + k[3] * Lv;
      v = v + k[4] * dtv;
      if (col && R >= j0 && R < j1) {
        size_t o = (size_t)R * pitch + c;
        if (t == nst) OA[o] = v;
        if (!final_here && t == T - 1) OB[o] = v;
        if (Wst[t - 1]) Wst[t - 1][o] = v;
      }
    }
    if (t < T) {
      S.w[t][s0] = v;
      if (act && R > 0) lds[(size_t)(t * 3 + (R % 3)) * ncol + c] = v;
    }
  }
  lds_barrier();
}

template <int T>
__global__ __launch_bounds__(1024) void k_rkg2_fused(FusedArgs a) {
  extern __shared__ double lds[];  // [T][3][ncol]
  __shared__ double cfs[4][5];
  __shared__ double rowc[4 * kFusedMaxRows];
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
  int nst = (int)((sr - a.k0 + 1) < T ? (sr - a.k0 + 1) : T);
  bool final_here = (a.k0 + nst - 1) == sr;
  size_t base = (size_t)r * a.rstride;
  const double *A = a.A + base, *B = a.B + base, *U0 = a.f + base, *DT = a.dty0 + base;
  double *OA = a.outA + base, *OB = a.outB + base;
  double *Wst[T];
#pragma unroll
  for (int t = 0; t < T; t++) Wst[t] = a.W[t] ? a.W[t] + base : nullptr;
  double D = a.D[r];
  double D2 = D + D;
  double P1 = (D * m.dph_i[cc + 1]) * m.dp_i[cc + 1];
  double P5 = (D * m.dph_i[cc + 2]) * m.dp_i[cc + 1];
  // stage coefficients (block-uniform) and per-row coefficient factors
  if (threadIdx.x < T * 5) {
    int t = threadIdx.x / 5, e = threadIdx.x - 5 * t;
    double v = 0.0;
    if (t < nst) {  // the table holds stages 2..s of this realization only
      const StageCoef &k = a.coef[(size_t)(a.k0 + t) * a.nr_table + r];
      v = e == 0 ? k.a : (e == 1 ? k.b : (e == 2 ? k.c : (e == 3 ? k.d : k.g)));
    }
    cfs[t][e] = v;
  }
  for (int R = lo0 + (int)threadIdx.x; R < hi0; R += blockDim.x) {
    double *rr = rowc + 4 * (R - lo0);
    if (R > 0 && R < ntm - 1) {
      double sti = m.st_i[R];
      double c2 = D * m.dth_i[R];
      c2 = c2 * m.dt_i[R];
      c2 = c2 * sti;
      c2 = c2 * m.sth[R];
      double c4 = D * m.dth_i[R + 1];
      c4 = c4 * m.dt_i[R];
      c4 = c4 * sti;
      c4 = c4 * m.sth[R + 1];
      rr[0] = sti;
      rr[1] = c2;
      rr[2] = c4;
    } else {
      rr[0] = rr[1] = rr[2] = 0.0;
    }
    rr[3] = 0.0;
  }
  FusedState<T> S;
#pragma unroll
  for (int t = 0; t < T; t++) {
    S.p0[t] = 0.0;
#pragma unroll
    for (int d = 0; d < 4; d++) S.w[t][d] = 0.0;
  }
#pragma unroll
  for (int d = 0; d < 4; d++) S.hu[d] = S.hd[d] = 0.0;
  double u0_pole = 0.0, dt_pole = 0.0;
  if (lo0 == 0 && col) {
    u0_pole = U0[c];
    dt_pole = DT[c];
  }
  int nsteps = (hi0 - lo0) + T;
  // steps 0 and 1 inputs: pf slot (i & 1)
#pragma unroll
  for (int i = 0; i < 2; i++) {
    int q = lo0 + i, rn = q - 1;
    double v0 = 0.0, v1 = 0.0, v2 = 0.0, v3 = 0.0;
    if (col && i < nsteps) {
      if (q < hi0) v0 = A[(size_t)q * m.pitch + c];
      if (rn >= 0 && rn < ntm) {
        size_t o = (size_t)rn * m.pitch + c;
        v1 = B[o];
        v2 = U0[o];
        v3 = DT[o];
      }
    }
    S.pf[i][0] = v0;
    S.pf[i][1] = v1;
    S.pf[i][2] = v2;
    S.pf[i][3] = v3;
  }
  lds_barrier();
  for (int i = 0; i < nsteps; i += 4) {
    fused_step<T, 0>(a, S, lds, rowc, i, nsteps, lo0, hi0, j0, j1, nst, final_here, col, c, cw,
                     ce, D2, P1, P5, A, B, U0, DT, OA, OB, u0_pole, dt_pole, cfs, Wst);
    if (i + 1 < nsteps)
      fused_step<T, 1>(a, S, lds, rowc, i + 1, nsteps, lo0, hi0, j0, j1, nst, final_here, col,
                       c, cw, ce, D2, P1, P5, A, B, U0, DT, OA, OB, u0_pole, dt_pole, cfs, Wst);
    if (i + 2 < nsteps)
      fused_step<T, 2>(a, S, lds, rowc, i + 2, nsteps, lo0, hi0, j0, j1, nst, final_here, col,
                       c, cw, ce, D2, P1, P5, A, B, U0, DT, OA, OB, u0_pole, dt_pole, cfs, Wst);
    if (i + 3 < nsteps)
      fused_step<T, 3>(a, S, lds, rowc, i + 3, nsteps, lo0, hi0, j0, j1, nst, final_here, col,
                       c, cw, ce, D2, P1, P5, A, B, U0, DT, OA, OB, u0_pole, dt_pole, cfs, Wst);
  }
}

// ---------------------------------------------------------------------------
// Fused RKG2 stages, lean march (k_rkg2_march<T>): the same schedule and
// arithmetic as k_rkg2_fused, with the per-step bookkeeping reduced to a few
// scalar instructions.  The march is unrolled by 6, so every ring slot is a
// compile-time constant (own-column windows: 3 slots by age; u0/dty0
// histories: shifted by one each step; LDS rows: slot (R - lo0) mod 3 =
// (i - t) mod 3); row
// pointers advance by one pitch per step; and every row condition is a
// comparison of the step index with a per-block constant.
// ---------------------------------------------------------------------------
template <int T>
struct MarchState {
  double w[T][3];       // level t own-column rows, slot (i - age) mod 3
  double p0[T];         // level t north-pole row (formed late)
  double hu[4], hd[4];  // u0 and dty0 at rows q-1-age, index age (shifted each step)
  double pf[2][4];      // prefetched inputs of the next two steps
};

template <int T>
struct MarchPlan {
  int nsteps, iA, iB;
  int ia[T + 1], ib[T + 1];  // level t active for ia <= i < ib
  int b_one;                 // level t computes row 1 at step b_one + t (R >= 1 from there)
  int north_tmax;            // levels t <= north_tmax form their north-pole row late
  int b_south;               // level t computes row ntm-1 at step b_south + t
  int b_o0, b_o1;            // level t rows inside the band: b_o0 + t <= i < b_o1 + t
  int sS1, sS2;              // LDS ring offsets of rows ntm-1, ntm-2
  int nst;                   // stages this launch advances for the realization
  bool final_here, pole0_out, lo0zero;
};

template <int T, int U>
__device__ __forceinline__ void march_step(MarchState<T> &S, double *lds, const double *rowc,
                                           const MarchPlan<T> &P, const Mesh &m, int i,
                                           bool col, int c, int cw, int ce, int ncol, double D2,
                                           double P1, double P5, const double *A,
                                           const double *B, const double *U0, const double *DT,
                                           size_t pitch, size_t &rowq, double *OA, double *OB,
                                           double u0_pole, double dt_pole,
                                           const double (*cfs)[5]) {
  const int nu = m.nu;
  constexpr int s0 = U % 3, s1 = (U + 2) % 3, s2 = (U + 1) % 3;  // ages 0, 1, 2
  constexpr int ps = U & 1;
  double a0 = S.pf[ps][0], b1 = S.pf[ps][1], u1 = S.pf[ps][2], d1 = S.pf[ps][3];
  {
    // prefetch for step i+2: A row q+2, B/u0/dty0 row q+1 (rowq = q*pitch)
    double v0 = 0.0, v1 = 0.0, v2 = 0.0, v3 = 0.0;
    if (col) {
      if (i + 2 < P.iA) v0 = A[rowq + 2 * pitch + c];
      if (i + 2 < P.nsteps && i + 1 < P.iB) {
        size_t o = rowq + pitch + c;
        v1 = B[o];
        v2 = U0[o];
        v3 = DT[o];
      }
    }
    S.pf[ps][0] = v0;
    S.pf[ps][1] = v1;
    S.pf[ps][2] = v2;
    S.pf[ps][3] = v3;
  }
#pragma unroll
  for (int d = 3; d > 0; d--) {
    S.hu[d] = S.hu[d - 1];
    S.hd[d] = S.hd[d - 1];
  }
  S.hu[0] = u1;
  S.hd[0] = d1;
  // level 0: row q of u_{k0-1}
  S.w[0][s0] = a0;
  if (i < P.iA) {
    lds[(size_t)(U % 3) * ncol + c] = a0;
    if (P.lo0zero && i == 0) S.p0[0] = a0;
  }
#pragma unroll
  for (int t = 1; t <= T; t++) {
    const int sl = ((U - t) % 3 + 3) % 3;  // LDS slot of row R = q - t
    const double *lp = lds + (size_t)((t - 1) * 3) * ncol;  // level t-1 ring
    const double *k = cfs[t - 1];
    double u0v = S.hu[t - 1], dtv = S.hd[t - 1];  // u0/dty0 of row R: age t-1
    const int ione = P.b_one + t;
    bool act = t <= P.nst && i >= P.ia[t] && i < P.ib[t];
    bool wantA = t == P.nst, wantB = !P.final_here && t == T - 1;
    double v = 0.0;
    if (act && t <= P.north_tmax && i == ione) {
      // late north-pole row 0 of stage t, from level t-1 rows 0 and 1 (slots 0, 1)
      const double *r0 = lp, *r1p = lp + ncol;
      double fn = wave_pole_sum(nu, [&](int kk) { return (D2 * (r1p[kk] - r0[kk])) * m.dp[kk + 1]; });
      double Lv = fn * m.bc_diff_n;
      double xc = S.p0[t - 1];
      double ukm2 = (t == 1) ? B[c < nu ? c : 0] : S.p0[t < 2 ? 0 : t - 2];
      double vv = k[0] * xc;
      vv = vv + k[1] * ukm2;
      vv = vv + k[2] * u0_pole;
      vv = vv + k[3] * Lv;
      vv = vv + k[4] * dt_pole;
      if (t < T) {
        S.p0[t] = vv;
        lds[(size_t)(t * 3 + 0) * ncol + c] = vv;
      }
      if (col && P.pole0_out) {
        if (wantA) OA[c] = vv;
        if (wantB) OB[c] = vv;
      }
    }
    if (act && i >= ione) {
      double Lv;
      if (i == P.b_south + t) {
        const double *r0 = lp + P.sS1, *r1p = lp + P.sS2;
        double fs = wave_pole_sum(nu, [&](int kk) { return (D2 * (r0[kk] - r1p[kk])) * m.dp[kk + 1]; });
        Lv = -fs * m.bc_diff_s;
      } else {
        const double *rc = lp + (size_t)sl * ncol;
        double xc = S.w[t - 1][s1];
        double xs = S.w[t - 1][s0];
        double xn = (i == ione) ? S.p0[t - 1] : S.w[t - 1][s2];
        double xw = rc[cw], xe = rc[ce];
        const double *rr = rowc + 4 * (i - t);
        double sti = rr[0];
        double c1 = (P1 * sti) * sti;
        double c2 = rr[1];
        double c4 = rr[2];
        double c5 = (P5 * sti) * sti;
        double c3 = -(((c1 + c2) + c4) + c5);
        Lv = c1 * xw;
        Lv = Lv + c2 * xn;
        Lv = Lv + c3 * xc;
        Lv = Lv + c4 * xs;
        Lv = Lv + c5 * xe;
      }
      double xc = S.w[t - 1][s1];
      double ukm2 = (t == 1) ? b1 : S.w[t < 2 ? 0 : t - 2][s2];
      v = k[0] * xc;
      v = v + k[1] * ukm2;
      v = v + k[2] * u0v;
      v = v + k[3] * Lv;
      v = v + k[4] * dtv;
      if (col && i >= P.b_o0 + t && i < P.b_o1 + t) {
        size_t o = rowq - (size_t)t * pitch + c;
        if (wantA) OA[o] = v;
        if (wantB) OB[o] = v;
      }
    }
    if (t < T) {
      S.w[t][s0] = v;
      if (act && i >= ione) lds[(size_t)(t * 3 + sl) * ncol + c] = v;
    }
  }
  rowq += pitch;
  lds_barrier();
}

template <int T>
__global__ __launch_bounds__(1024) void k_rkg2_march(FusedArgs a) {
  extern __shared__ double lds[];  // [T][3][ncol]
  __shared__ double cfs[4][5];
  __shared__ double rowc[4 * kFusedMaxRows];
  const Mesh &m = a.m;
  const int nu = m.nu, ntm = m.ntm, ncol = a.ncol;
  const size_t pitch = (size_t)m.pitch;
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
  int nst = (int)((sr - a.k0 + 1) < T ? (sr - a.k0 + 1) : T);
  size_t base = (size_t)r * a.rstride;
  const double *A = a.A + base, *B = a.B + base, *U0 = a.f + base, *DT = a.dty0 + base;
  double *OA = a.outA + base, *OB = a.outB + base;
  double D = a.D[r];
  double D2 = D + D;
  double P1 = (D * m.dph_i[cc + 1]) * m.dp_i[cc + 1];
  double P5 = (D * m.dph_i[cc + 2]) * m.dp_i[cc + 1];
  if (threadIdx.x < T * 5) {
    int t = threadIdx.x / 5, e = threadIdx.x - 5 * t;
    double v = 0.0;
    if (t < nst) {
      const StageCoef &k = a.coef[(size_t)(a.k0 + t) * a.nr_table + r];
      v = e == 0 ? k.a : (e == 1 ? k.b : (e == 2 ? k.c : (e == 3 ? k.d : k.g)));
    }
    cfs[t][e] = v;
  }
  for (int R = lo0 + (int)threadIdx.x; R < hi0; R += blockDim.x) {
    double *rr = rowc + 4 * (R - lo0);
    if (R > 0 && R < ntm - 1) {
      double sti = m.st_i[R];
      double c2 = D * m.dth_i[R];
      c2 = c2 * m.dt_i[R];
      c2 = c2 * sti;
      c2 = c2 * m.sth[R];
      double c4 = D * m.dth_i[R + 1];
      c4 = c4 * m.dt_i[R];
      c4 = c4 * sti;
      c4 = c4 * m.sth[R + 1];
      rr[0] = sti;
      rr[1] = c2;
      rr[2] = c4;
    } else {
      rr[0] = rr[1] = rr[2] = 0.0;
    }
    rr[3] = 0.0;
  }
  MarchPlan<T> P;
  P.nsteps = (hi0 - lo0) + T;
  P.iA = hi0 - lo0;
  P.iB = ntm - lo0;
#pragma unroll
  for (int t = 1; t <= T; t++) {
    int rlo = max(0, j0 - T + t), rhi = min(ntm, j1 + T - t);
    P.ia[t] = rlo - lo0 + t;
    P.ib[t] = rhi - lo0 + t;
  }
  P.ia[0] = P.ib[0] = 0;
  P.b_one = 1 - lo0;
  P.north_tmax = T - j0;  // rlo_t = max(0, j0 - T + t) == 0
  P.b_south = ntm - 1 - lo0;
  P.b_o0 = j0 - lo0;
  P.b_o1 = j1 - lo0;
  P.sS1 = (((ntm - 1 - lo0) % 3 + 3) % 3) * ncol;
  P.sS2 = (((ntm - 2 - lo0) % 3 + 3) % 3) * ncol;
  P.nst = nst;
  P.final_here = (a.k0 + nst - 1) == sr;
  P.pole0_out = j0 == 0;
  P.lo0zero = lo0 == 0;
  MarchState<T> S;
#pragma unroll
  for (int t = 0; t < T; t++) {
    S.p0[t] = 0.0;
#pragma unroll
    for (int d = 0; d < 3; d++) S.w[t][d] = 0.0;
  }
#pragma unroll
  for (int d = 0; d < 4; d++) S.hu[d] = S.hd[d] = 0.0;
  double u0_pole = 0.0, dt_pole = 0.0;
  if (lo0 == 0 && col) {
    u0_pole = U0[c];
    dt_pole = DT[c];
  }
  // inputs of steps 0 and 1: A rows lo0, lo0+1; B/u0/dty0 rows lo0-1, lo0
#pragma unroll
  for (int i = 0; i < 2; i++) {
    int q = lo0 + i, rn = q - 1;
    double v0 = 0.0, v1 = 0.0, v2 = 0.0, v3 = 0.0;
    if (col && i < P.nsteps) {
      if (q < hi0) v0 = A[(size_t)q * pitch + c];
      if (rn >= 0 && rn < ntm) {
        size_t o = (size_t)rn * pitch + c;
        v1 = B[o];
        v2 = U0[o];
        v3 = DT[o];
      }
    }
    S.pf[i][0] = v0;
    S.pf[i][1] = v1;
    S.pf[i][2] = v2;
    S.pf[i][3] = v3;
  }
  size_t rowq = (size_t)lo0 * pitch;
  lds_barrier();
  const int nsteps = P.nsteps;
  for (int i = 0; i < nsteps; i += 6) {
    march_step<T, 0>(S, lds, rowc, P, m, i, col, c, cw, ce, ncol, D2, P1, P5, A, B, U0, DT, pitch,
                     rowq, OA, OB, u0_pole, dt_pole, cfs);
    if (i + 1 >= nsteps) break;
    march_step<T, 1>(S, lds, rowc, P, m, i + 1, col, c, cw, ce, ncol, D2, P1, P5, A, B, U0, DT,
                     pitch, rowq, OA, OB, u0_pole, dt_pole, cfs);
    if (i + 2 >= nsteps) break;
    march_step<T, 2>(S, lds, rowc, P, m, i + 2, col, c, cw, ce, ncol, D2, P1, P5, A, B, U0, DT,
                     pitch, rowq, OA, OB, u0_pole, dt_pole, cfs);
    if (i + 3 >= nsteps) break;
    march_step<T, 3>(S, lds, rowc, P, m, i + 3, col, c, cw, ce, ncol, D2, P1, P5, A, B, U0, DT,
                     pitch, rowq, OA, OB, u0_pole, dt_pole, cfs);
    if (i + 4 >= nsteps) break;
    march_step<T, 4>(S, lds, rowc, P, m, i + 4, col, c, cw, ce, ncol, D2, P1, P5, A, B, U0, DT,
                     pitch, rowq, OA, OB, u0_pole, dt_pole, cfs);
    if (i + 5 >= nsteps) break;
    march_step<T, 5>(S, lds, rowc, P, m, i + 5, col, c, cw, ce, ncol, D2, P1, P5, A, B, U0, DT,
                     pitch, rowq, OA, OB, u0_pole, dt_pole, cfs);
  }
}

// ---------------------------------------------------------------------------
// Fused RKG2 stages, two phi columns per thread (k_rkg2_fused2<T>): the
// schedule of k_rkg2_fused<T> (bands, lag-one march, 3-row LDS rings, late
// north pole) with thread x owning columns 2x and 2x+1.  The block has half
// as many waves, so the block-uniform bookkeeping (row conditions, ring
// slots) is issued half as often per column, the east neighbour of the even
// column and the west neighbour of the odd column come from registers, and
// row loads, LDS stores and row stores move 16 bytes per lane.  Every point
// is computed with the expressions of k_rkg2_stage: bit-identical results.
// The block has a multiple of 64 threads (full wavefronts for the polar
// sums) and LDS rows of 2 * blockDim doubles.
// ---------------------------------------------------------------------------
// Prefetch distance of the row loads in steps; each load gets its own slot
// (2 * distance slots), so a slot is consumed before it is loaded again and
// the loads stay in flight across the unrolled loop's back edge.  Two steps
// for up to four stages: with dty0 formed in the kernel (three row loads per
// step instead of four) the slots fit the register budget of two wavefronts
// per SIMD; five stages keep one.
template <int T>
struct Fused2Cfg {
  static constexpr int PD = T >= 5 ? 1 : 2;
  static constexpr int NS = 2 * PD;
};

template <int T>
struct Fused2State {
  double w[T][4][2];
  double p0[T][2];
  double hu[4][2], hd[4][2];  // u0 and dty0 of rows q-1-d (d = 0..3)
  double u0n[2];              // u0 of row q-1 (loaded one step ahead of hu)
  double pf[Fused2Cfg<T>::NS][3][2];  // prefetched u_{k0-1} (row q), u_{k0-2} (row q-1), u0 (row q)
};

// Columns c0, c0+1 of a row (c0 even, c0 < nu <= pitch, pitch even): one
// 16-byte access.  When nu is odd the last thread's second column is the
// row's first padding column, which no kernel reads as data; it may receive
// that thread's (unused) value.
// The rows of u_{k0-1}, u_{k0-2} and the outputs stream through once per
// launch (the fields of a batch exceed the caches), so their loads and
// stores carry the non-temporal hint: measured 5-6 % faster fused launches
// at 64 realizations of 512x1024.  u0 is read again by every launch of the
// super-step and by the kernels around it, and one copy of it for the
// batch (268 MB at 64 realizations of 512x1024) is about the size of the
// 256 MB Infinity Cache: its rows are loaded with the default policy, so
// they may stay there (fused time per diffusion step 15.9 -> 15.4 ms, whole
// step 20.6 -> 19.5 ms on that batch).
typedef double fused_d2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ void load_pair(const double *row, int c0, double &x, double &y) {
  fused_d2 v = __builtin_nontemporal_load(reinterpret_cast<const fused_d2 *>(row + c0));
  x = v.x;
  y = v.y;
}

__device__ __forceinline__ void load_pair_u0(const double *row, int c0, double &x, double &y) {
  fused_d2 v = *reinterpret_cast<const fused_d2 *>(row + c0);
  x = v.x;
  y = v.y;
}

__device__ __forceinline__ void store_pair(double *row, int c0, int nu, double x, double y) {
  if (c0 < nu) __builtin_nontemporal_store(fused_d2{x, y}, reinterpret_cast<fused_d2 *>(row + c0));
}

template <int T, int U>
__device__ __forceinline__ void fused2_step(const FusedArgs &a, Fused2State<T> &S, double *lds,
                                            const double *rowc, int i, int nsteps, int lo0,
                                            int hi0, int j0, int j1, int nst, bool final_here,
                                            int c0, int cl, int cw0, int ce0, int ce1,
                                            bool col1, int nrow,
                                            double D2, const double (&P1)[2],
                                            const double (&P5)[2], const double *A,
                                            const double *B, const double *U0, double dtr,
                                            double *OA, double *OB, const double (&u0_pole)[2],
                                            const double (&dt_pole)[2], const double (*cfs)[5],
                                            double *const *Wst, const double *sdp) {
  const Mesh &m = a.m;
  const int nu = m.nu, ntm = m.ntm, pitch = m.pitch;
  constexpr int s0 = U & 3, s1 = (U + 3) & 3, s2 = (U + 2) & 3;
  constexpr int kPD = Fused2Cfg<T>::PD, kNS = Fused2Cfg<T>::NS;
  constexpr int p0s = U % kNS, pls = (U + kPD) % kNS;  // slot read now, slot loaded now
  int q = lo0 + i;
  double a0[2], b1[2], u0q[2];
#pragma unroll
  for (int e = 0; e < 2; e++) {
    a0[e] = S.pf[p0s][0][e];
    b1[e] = S.pf[p0s][1][e];
    u0q[e] = S.pf[p0s][2][e];
  }
  {
    // prefetch step i+kPD into its own slot.  The loads are
    // unconditional (rows and columns clamped into the field): values of rows
    // outside the band's range or of columns >= nu never reach a stored
    // output, and straight-line loads let each wait cover only the load it
    // needs instead of every load in flight.
    int qn = min(q + kPD, hi0 - 1), rn = min(max(q + kPD - 1, 0), ntm - 1);
    size_t oa = (size_t)qn * pitch + cl, o = (size_t)rn * pitch + cl;
    load_pair(A + oa, 0, S.pf[pls][0][0], S.pf[pls][0][1]);
    load_pair(B + o, 0, S.pf[pls][1][0], S.pf[pls][1][1]);
    load_pair_u0(U0 + oa, 0, S.pf[pls][2][0], S.pf[pls][2][1]);
  }
  // u0 and dty0 of row q-5 (age 4: the slot that rows q-1 take now), for a
  // fifth stage
  double hu4[2] = {S.hu[s0][0], S.hu[s0][1]}, hd4[2] = {S.hd[s0][0], S.hd[s0][1]};
#pragma unroll
  for (int e = 0; e < 2; e++) {
    S.hu[s0][e] = S.u0n[e];  // u0 of row q-1
    S.u0n[e] = u0q[e];
    S.w[0][s0][e] = a0[e];
  }
  // u0 ring (after the T stage levels' rings): rows of u0 for the phi
  // neighbours of dty0 = dt L(u0), formed by stage 1 one row behind
  double *lu0 = lds + (size_t)(T * 3) * nrow;
  if (q < hi0) {
    *reinterpret_cast<double2 *>(lu0 + (size_t)(q % 3) * nrow + c0) = make_double2(u0q[0], u0q[1]);
    *reinterpret_cast<double2 *>(lds + (size_t)(q % 3) * nrow + c0) = make_double2(a0[0], a0[1]);
    if (q == 0) {
      S.p0[0][0] = a0[0];
      S.p0[0][1] = a0[1];
    }
  }
#pragma unroll
  for (int t = 1; t <= T; t++) {
    int R = q - t;
    int rlo = max(0, j0 - T + t), rhi = min(ntm, j1 + T - t);
    const double *lp = lds + (size_t)((t - 1) * 3) * nrow;  // level t-1 ring
    const double *k = cfs[t - 1];
    const int ha = (t == 1) ? s0 : (t == 2 ? s1 : (t == 3 ? s2 : ((U + 1) & 3)));
    bool act = t <= nst && R >= rlo && R < rhi;
    double v[2] = {0.0, 0.0};
    if (act && R == 1 && rlo == 0) {
      // late north-pole row 0 of stage t, from level t-1 rows 0 and 1 in LDS
      const double *r0 = lp + 0 * nrow, *r1p = lp + 1 * nrow;
      double fn = wave_pole_sum(nu, [&](int kk) { return (D2 * (r1p[kk] - r0[kk])) * sdp[kk]; });
      double Lv = fn * m.bc_diff_n;
      double bp0 = 0.0, bp1 = 0.0;
      if (t == 1 && c0 < nu) load_pair(B, c0, bp0, bp1);
      double vv[2];
#pragma unroll
      for (int e = 0; e < 2; e++) {
        double xc = S.p0[t - 1][e];
        double ukm2 = (t == 1) ? (e == 0 ? bp0 : bp1) : S.p0[t >= 2 ? t - 2 : 0][e];
        double x = k[0] * xc;
        x = x + k[1] * ukm2;
        x = x + k[2] * u0_pole[e];
        x = x + k[3] * Lv;
        x = x + k[4] * dt_pole[e];
        vv[e] = x;
      }
      if (t < T) {
        S.p0[t < T ? t : 0][0] = vv[0];
        S.p0[t < T ? t : 0][1] = vv[1];
        *reinterpret_cast<double2 *>(lds + (size_t)(t * 3 + 0) * nrow + c0) = make_double2(vv[0], vv[1]);
      }
      if (0 >= j0 && 0 < j1) {
        if (t == nst) store_pair(OA, c0, nu, vv[0], vv[1]);
        if (!final_here && t == T - 1) store_pair(OB, c0, nu, vv[0], vv[1]);
        if (Wst[t - 1]) store_pair(Wst[t - 1], c0, nu, vv[0], vv[1]);
      }
    }
    if (act && R > 0 && t == 1) {
      // dty0 = dt L(u0) of row R (as k_rkg2_stage1 / the PTL pass form it:
      // same coefficient and summation order), kept for the later stages
      double Lu[2];
      if (R == ntm - 1) {
        // the south-pole row's dty0 as stage 1 wrote it (the same
        // expression, dt times the polar closure's sum), uniform along the row
        double d = a.dty0[(U0 - a.f) + (size_t)(ntm - 1) * pitch];
        S.hd[s0][0] = S.hd[s0][1] = d;
      } else {
        const double *rc = lu0 + (R % 3) * nrow;
        const double *rr = rowc + 4 * (R - lo0);
        double sti = rr[0];
        double c2 = rr[1];
        double c4 = rr[2];
        double xw[2], xe[2];
        xw[0] = rc[cw0];
        xe[0] = col1 ? S.hu[s0][1] : rc[ce0];
        xw[1] = S.hu[s0][0];
        xe[1] = rc[ce1];
#pragma unroll
        for (int e = 0; e < 2; e++) {
          double c1 = (P1[e] * sti) * sti;
          double c5 = (P5[e] * sti) * sti;
          double c3 = -(((c1 + c2) + c4) + c5);
          double L = c1 * xw[e];
          L = L + c2 * S.hu[s1][e];
          L = L + c3 * S.hu[s0][e];
          L = L + c4 * S.u0n[e];
          L = L + c5 * xe[e];
          Lu[e] = L;
        }
        S.hd[s0][0] = dtr * Lu[0];
        S.hd[s0][1] = dtr * Lu[1];
      }
    }
    if (act && R > 0) {
      double Lv[2];
      if (R == ntm - 1) {
        const double *r0 = lp + ((ntm - 1) % 3) * nrow, *r1p = lp + ((ntm - 2) % 3) * nrow;
        double fs = wave_pole_sum(nu, [&](int kk) { return (D2 * (r0[kk] - r1p[kk])) * sdp[kk]; });
        Lv[0] = Lv[1] = -fs * m.bc_diff_s;
      } else {
        const double *rc = lp + (R % 3) * nrow;
        const double *rr = rowc + 4 * (R - lo0);
        double sti = rr[0];
        double c2 = rr[1];
        double c4 = rr[2];
        double xc0 = S.w[t - 1][s1][0], xc1 = S.w[t - 1][s1][1];
        double xw[2], xe[2];
        xw[0] = rc[cw0];
        xe[0] = col1 ? xc1 : rc[ce0];
        xw[1] = xc0;
        xe[1] = rc[ce1];
#pragma unroll
        for (int e = 0; e < 2; e++) {
          double xc = S.w[t - 1][s1][e];
          double xs = S.w[t - 1][s0][e];
          double xn = (R == 1) ? S.p0[t - 1][e] : S.w[t - 1][s2][e];
          double c1 = (P1[e] * sti) * sti;
          double c5 = (P5[e] * sti) * sti;
          double c3 = -(((c1 + c2) + c4) + c5);
          double L = c1 * xw[e];
          L = L + c2 * xn;
          L = L + c3 * xc;
          L = L + c4 * xs;
          L = L + c5 * xe[e];
          Lv[e] = L;
        }
      }
#pragma unroll
      for (int e = 0; e < 2; e++) {
        double xc = S.w[t - 1][s1][e];
        double ukm2 = (t == 1) ? b1[e] : S.w[t >= 2 ? t - 2 : 0][s2][e];
        double x = k[0] * xc;
        x = x + k[1] * ukm2;
        x = x + k[2] * (t <= 4 ? S.hu[ha][e] : hu4[e]);
        x = x + k[3] * Lv[e];
        x = x + k[4] * (t <= 4 ? S.hd[ha][e] : hd4[e]);
        v[e] = x;
      }
      if (R >= j0 && R < j1) {
        size_t o = (size_t)R * pitch;
        if (t == nst) store_pair(OA + o, c0, nu, v[0], v[1]);
        if (!final_here && t == T - 1) store_pair(OB + o, c0, nu, v[0], v[1]);
        if (Wst[t - 1]) store_pair(Wst[t - 1] + o, c0, nu, v[0], v[1]);
      }
    }
    if (t < T) {
      S.w[t < T ? t : 0][s0][0] = v[0];
      S.w[t < T ? t : 0][s0][1] = v[1];
      if (act && R > 0)
        *reinterpret_cast<double2 *>(lds + (size_t)(t * 3 + (R % 3)) * nrow + c0) = make_double2(v[0], v[1]);
    }
  }
  lds_barrier();
}

// ---------------------------------------------------------------------------
// Interior march step of k_rkg2_fused2<T, NR> (T <= 4).  The rows of every
// level are interior rows of the band's own range (2 <= R <= ntm-2,
// j0 <= R < j1) and all T stages are active, so the step has no row tests:
// the values are those of fused2_step, computed in the same order.  U is
// the step index modulo 12 within a run of body steps that starts at a step
// index divisible by 4 and a row q divisible by 3, so register slots
// (mod 4), prefetch slots (mod 4) and LDS ring slots (mod 3) are
// compile-time constants; with NR (LDS row length) a constant too, LDS
// addresses are per-thread bases plus immediate offsets.  Row loads and
// stores use buffer resources over the realization's field: the row's byte
// offset is a scalar operand, the column offset a per-thread constant, and
// a store from a thread whose columns lie beyond the row (offset past the
// resource's size) writes nothing.  Stage coefficients are scalar
// registers.
// ---------------------------------------------------------------------------
typedef int fused_v4i __attribute__((ext_vector_type(4)));
constexpr int kBufNT = 2;    // cache-policy operand: non-temporal
constexpr int kBufKeep = 0;  // default policy (u0 rows, which are read again)

__device__ __forceinline__ __amdgpu_buffer_rsrc_t field_rsrc(const double *p, int bytes) {
  // raw buffer (stride 0), 32-bit data format word of the gfx9 resource
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<double *>(p), (short)0, bytes, 0x00020000);
}

template <int POL = kBufNT>
__device__ __forceinline__ void bload_pair(__amdgpu_buffer_rsrc_t r, int voff, int soff, double &x,
                                           double &y) {
  fused_v4i v = __builtin_amdgcn_raw_buffer_load_b128(r, voff, soff, POL);
  double2 d = __builtin_bit_cast(double2, v);
  x = d.x;
  y = d.y;
}

__device__ __forceinline__ void bstore_pair(__amdgpu_buffer_rsrc_t r, int voff, int soff, double x,
                                            double y) {
  __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(fused_v4i, make_double2(x, y)), r, voff,
                                         soff, kBufNT);
}

__device__ __forceinline__ double uniform_double(double x) {
  long long b = __double_as_longlong(x);
  int lo = __builtin_amdgcn_readfirstlane((int)(b & 0xffffffffLL));
  int hi = __builtin_amdgcn_readfirstlane((int)(b >> 32));
  return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}

template <int T>
struct BodyArgs {
  int hi0, ntm, pitch8;  // pitch8: row pitch in bytes
  int vld, vst;          // column byte offsets: loads (clamped into the row), stores (or past the end)
  bool storeB;           // stage k0+T-2 goes to outB (the realization does not finish here)
  double kc[T][5];       // stage coefficients (uniform)
};

template <int T, int U, int NR>
__device__ __forceinline__ void fused2_body_step(Fused2State<T> &S, double *lds, const double *rowb,
                                                 int nrow_rt, int q, const BodyArgs<T> &ba,
                                                 __amdgpu_buffer_rsrc_t rA,
                                                 __amdgpu_buffer_rsrc_t rB,
                                                 __amdgpu_buffer_rsrc_t rU,
                                                 __amdgpu_buffer_rsrc_t rOA,
                                                 __amdgpu_buffer_rsrc_t rOB, int c0, int cw0,
                                                 int ce0, int ce1, const double (&P1)[2],
                                                 const double (&P5)[2], double dtr) {
  static_assert(T >= 2 && T <= 5, "interior step: 2..5 stages");
  constexpr int s0 = U & 3, s1 = (U + 3) & 3, s2 = (U + 2) & 3;
  constexpr int kPD = Fused2Cfg<T>::PD, kNS = Fused2Cfg<T>::NS;
  constexpr int p0s = U % kNS, pls = (U + kPD) % kNS;
  constexpr int rq = U % 3;  // ring slot of row q
  const int nrow = NR ? NR : nrow_rt;
  double a0[2], b1[2], u0q[2];
#pragma unroll
  for (int e = 0; e < 2; e++) {
    a0[e] = S.pf[p0s][0][e];
    b1[e] = S.pf[p0s][1][e];
    u0q[e] = S.pf[p0s][2][e];
  }
  {
    int qn = min(q + kPD, ba.hi0 - 1), rn = min(q + kPD - 1, ba.ntm - 1);
    int oa = qn * ba.pitch8, ob = rn * ba.pitch8;
    bload_pair(rA, ba.vld, oa, S.pf[pls][0][0], S.pf[pls][0][1]);
    bload_pair(rB, ba.vld, ob, S.pf[pls][1][0], S.pf[pls][1][1]);
    bload_pair<kBufKeep>(rU, ba.vld, oa, S.pf[pls][2][0], S.pf[pls][2][1]);
  }
  // u0 and dty0 of row q-5 (age 4, the slot rows q-1 take now), for a fifth stage
  double hu4[2] = {S.hu[s0][0], S.hu[s0][1]}, hd4[2] = {S.hd[s0][0], S.hd[s0][1]};
#pragma unroll
  for (int e = 0; e < 2; e++) {
    S.hu[s0][e] = S.u0n[e];
    S.u0n[e] = u0q[e];
    S.w[0][s0][e] = a0[e];
  }
  double *lu0 = lds + (size_t)(T * 3) * nrow;
  // neighbour values of every level, read before any of the step's work
  // (index 0: the u0 ring for dty0; t: level t-1's ring at row q-t)
  double nbw[T + 1], nbe0[T + 1], nbe1[T + 1];
#pragma unroll
  for (int t = 0; t <= T; t++) {
    const int rs = ((U - (t ? t : 1)) % 3 + 3) % 3;
    const double *rc = (t == 0 ? lu0 : lds + (size_t)((t - 1) * 3) * nrow) + (size_t)rs * nrow;
    nbw[t] = rc[cw0];
    nbe0[t] = rc[ce0];
    nbe1[t] = rc[ce1];
  }
  // row R = q-t's (sti, c2, c4) of every level
  double rsti[T], rc2[T], rc4[T];
#pragma unroll
  for (int t = 1; t <= T; t++) {
    const double *rr = rowb + 4 * (T - t);
    rsti[t - 1] = rr[0];
    rc2[t - 1] = rr[1];
    rc4[t - 1] = rr[2];
  }
#pragma unroll
  for (int t = 1; t <= T; t++) {
    const int ha = (t == 1) ? s0 : (t == 2 ? s1 : (t == 3 ? s2 : ((U + 1) & 3)));
    double sti = rsti[t - 1];
    double c2 = rc2[t - 1];
    double c4 = rc4[t - 1];
    if (t == 1) {
      // dty0 = dt L(u0) of row R (the expressions of fused2_step)
      double xw[2], xe[2];
      xw[0] = nbw[0];
      xe[0] = nbe0[0];
      xw[1] = S.hu[s0][0];
      xe[1] = nbe1[0];
      double Lu[2];
#pragma unroll
      for (int e = 0; e < 2; e++) {
        double c1 = (P1[e] * sti) * sti;
        double c5 = (P5[e] * sti) * sti;
        double c3 = -(((c1 + c2) + c4) + c5);
        double L = c1 * xw[e];
        L = L + c2 * S.hu[s1][e];
        L = L + c3 * S.hu[s0][e];
        L = L + c4 * S.u0n[e];
        L = L + c5 * xe[e];
        Lu[e] = L;
      }
      S.hd[s0][0] = dtr * Lu[0];
      S.hd[s0][1] = dtr * Lu[1];
    }
    double xc0 = S.w[t - 1][s1][0], xc1 = S.w[t - 1][s1][1];
    double xw[2], xe[2];
    xw[0] = nbw[t];
    xe[0] = nbe0[t];
    xw[1] = xc0;
    xe[1] = nbe1[t];
    double v[2];
#pragma unroll
    for (int e = 0; e < 2; e++) {
      double xc = e == 0 ? xc0 : xc1;
      double xs = S.w[t - 1][s0][e];
      double xn = S.w[t - 1][s2][e];
      double c1 = (P1[e] * sti) * sti;
      double c5 = (P5[e] * sti) * sti;
      double c3 = -(((c1 + c2) + c4) + c5);
      double L = c1 * xw[e];
      L = L + c2 * xn;
      L = L + c3 * xc;
      L = L + c4 * xs;
      L = L + c5 * xe[e];
      double ukm2 = (t == 1) ? b1[e] : S.w[t >= 2 ? t - 2 : 0][s2][e];
      const double *k = ba.kc[t - 1];
      double x = k[0] * xc;
      x = x + k[1] * ukm2;
      x = x + k[2] * (t <= 4 ? S.hu[ha][e] : hu4[e]);
      x = x + k[3] * L;
      x = x + k[4] * (t <= 4 ? S.hd[ha][e] : hd4[e]);
      v[e] = x;
    }
    if (t == T) bstore_pair(rOA, ba.vst, (q - t) * ba.pitch8, v[0], v[1]);
    if (t == T - 1 && ba.storeB) bstore_pair(rOB, ba.vst, (q - t) * ba.pitch8, v[0], v[1]);
    if (t < T) {
      S.w[t < T ? t : 0][s0][0] = v[0];
      S.w[t < T ? t : 0][s0][1] = v[1];
    }
  }
  // The step's LDS writes come after all its LDS reads: each goes to a ring
  // slot no level reads in this step (row q of u0 and of level 0, row q-t of
  // level t).  With the reads hoisted above (same values, same arithmetic)
  // the step waits on LDS about 4 times instead of 8; measured launch times
  // are unchanged within 0.3 %, so LDS latency does not bound the step.
  *reinterpret_cast<double2 *>(lu0 + (size_t)rq * nrow + c0) = make_double2(u0q[0], u0q[1]);
  *reinterpret_cast<double2 *>(lds + (size_t)rq * nrow + c0) = make_double2(a0[0], a0[1]);
#pragma unroll
  for (int t = 1; t < T; t++) {
    const int rs = ((U - t) % 3 + 3) % 3;
    *reinterpret_cast<double2 *>(lds + (size_t)(t * 3 + rs) * nrow + c0) =
        make_double2(S.w[t][s0][0], S.w[t][s0][1]);
  }
  lds_barrier();
}

// One (realization, band) tile of k_rkg2_fused2<T>: stages k0 .. k0+T-1 of
// realization a.order[z] over theta rows [band*H, band*H + H), reading
// u_{k0-1} from Ain and u_{k0-2} from Bin, writing outA/outB (the other
// fields and tables from a).  lds, cfs and rowc are the block's shared
// arrays; the tile begins after a barrier that ends every earlier use.
// With body set (and T <= 4, all T stages active, no stored stages), the
// steps whose rows are all interior run as fused2_body_step.
template <int T, int NR = 0>
__device__ __forceinline__ void fused2_tile(const FusedArgs &a, int z, int band, int Hb, int k0,
                                            const double *Ain, const double *Bin, double *OAin,
                                            double *OBin, double *lds, double (*cfs)[5],
                                            double *rowc, double *sdp, bool body = false) {
  const Mesh &m = a.m;
  const int nu = m.nu, ntm = m.ntm;
  const int nrow = 2 * (int)blockDim.x;
  int r = a.order[z];
  int c0 = 2 * (int)threadIdx.x;
  int cl = min(c0, m.pitch - 2);  // column pair of the loads, inside the row
  bool col1 = c0 + 1 < nu;
  int e0 = c0 < nu ? c0 : 0, e1 = col1 ? c0 + 1 : 0;
  int cw0 = wrapc(e0 - 1, nu), ce0 = wrapc(e0 + 1, nu), ce1 = wrapc(e1 + 1, nu);
  int j0 = band * Hb, j1 = min(ntm, j0 + Hb);
  int lo0 = max(0, j0 - T), hi0 = min(ntm, j1 + T);
  int64_t sr = a.s[r];
  int nst = (int)((sr - k0 + 1) < T ? (sr - k0 + 1) : T);
  bool final_here = (k0 + nst - 1) == sr;
  size_t base = (size_t)r * a.rstride;
  const double *A = Ain + base, *B = Bin + base, *U0 = a.f + base;
  double *OA = OAin + base, *OB = OBin + base;
  double *Wst[T];
#pragma unroll
  for (int t = 0; t < T; t++) Wst[t] = a.W[t] ? a.W[t] + base : nullptr;
  double D = a.D[r];
  double D2 = D + D;
  double dtr = a.cdt[r];
  double P1[2], P5[2];
  P1[0] = (D * m.dph_i[e0 + 1]) * m.dp_i[e0 + 1];
  P5[0] = (D * m.dph_i[e0 + 2]) * m.dp_i[e0 + 1];
  P1[1] = (D * m.dph_i[e1 + 1]) * m.dp_i[e1 + 1];
  P5[1] = (D * m.dph_i[e1 + 2]) * m.dp_i[e1 + 1];
  if (threadIdx.x < T * 5) {
    int t = threadIdx.x / 5, e = threadIdx.x - 5 * t;
    double v = 0.0;
    if (t < nst) {
      const StageCoef &k = a.coef[(size_t)(k0 + t) * a.nr_table + r];
      v = e == 0 ? k.a : (e == 1 ? k.b : (e == 2 ? k.c : (e == 3 ? k.d : k.g)));
    }
    cfs[t][e] = v;
  }
  for (int R = lo0 + (int)threadIdx.x; R < hi0; R += blockDim.x) {
    double *rr = rowc + 4 * (R - lo0);
    if (R > 0 && R < ntm - 1) {
      double sti = m.st_i[R];
      double c2 = D * m.dth_i[R];
      c2 = c2 * m.dt_i[R];
      c2 = c2 * sti;
      c2 = c2 * m.sth[R];
      double c4 = D * m.dth_i[R + 1];
      c4 = c4 * m.dt_i[R];
      c4 = c4 * sti;
      c4 = c4 * m.sth[R + 1];
      rr[0] = sti;
      rr[1] = c2;
      rr[2] = c4;
    } else {
      rr[0] = rr[1] = rr[2] = 0.0;
    }
    rr[3] = 0.0;
  }
  Fused2State<T> S;
#pragma unroll
  for (int t = 0; t < T; t++)
#pragma unroll
    for (int e = 0; e < 2; e++) {
      S.p0[t][e] = 0.0;
#pragma unroll
      for (int d = 0; d < 4; d++) S.w[t][d][e] = 0.0;
    }
#pragma unroll
  for (int d = 0; d < 4; d++)
#pragma unroll
    for (int e = 0; e < 2; e++) S.hu[d][e] = S.hd[d][e] = 0.0;
  S.u0n[0] = S.u0n[1] = 0.0;
  double u0_pole[2] = {0.0, 0.0}, dt_pole[2] = {0.0, 0.0};
  if (lo0 == 0 || hi0 == ntm) {
    // the phi cell widths of the polar sums, read from LDS by the march
    for (int c = threadIdx.x; c < nu; c += blockDim.x) sdp[c] = m.dp[c + 1];
  }
  if (lo0 == 0) {
    // dty0 of the north-pole row, dt L(u0) with the polar closure's sum:
    // stage 1 wrote it (the same expression and summation order), uniform
    // along the row.  Reading it spares the pole band a wave sum over two
    // rows of u0 before its march, which made it the launch's last block.
    dt_pole[0] = dt_pole[1] = a.dty0[base];
    if (c0 < nu) load_pair_u0(U0, c0, u0_pole[0], u0_pole[1]);
  }
  int nsteps = (hi0 - lo0) + T;
#pragma unroll
  for (int i = 0; i < Fused2Cfg<T>::PD; i++) {
    int q = min(lo0 + i, hi0 - 1), rn = min(max(lo0 + i - 1, 0), ntm - 1);
    size_t oa = (size_t)q * m.pitch + cl, o = (size_t)rn * m.pitch + cl;
    load_pair(A + oa, 0, S.pf[i][0][0], S.pf[i][0][1]);
    load_pair(B + o, 0, S.pf[i][1][0], S.pf[i][1][1]);
    load_pair_u0(U0 + oa, 0, S.pf[i][2][0], S.pf[i][2][1]);
  }
  lds_barrier();
  int i = 0;
  {
    // body rows: q in [qlo, qhi] puts every level's row R = q - t in the
    // band's own range and away from the poles; the run starts at a step
    // index divisible by 4 whose row is divisible by 3 and covers whole
    // groups of 4 steps (register and prefetch slots are the general
    // steps' again after each group; ring slots follow the row)
    int qlo = max(j0, 2) + T, qhi = min(j1, ntm - 1);
    int ib = (qlo - lo0 + 3) & ~3;
    while ((lo0 + ib) % 3) ib += 4;
    int nquads = (qhi - (lo0 + ib) + 1) / 4;
    if (body && nst == T && nquads > 0) {
      for (; i < ib; i += 4) {
        fused2_step<T, 0>(a, S, lds, rowc, i, nsteps, lo0, hi0, j0, j1, nst, final_here, c0, cl,
                          cw0, ce0, ce1, col1, nrow, D2, P1, P5, A, B, U0, dtr, OA, OB, u0_pole,
                          dt_pole, cfs, Wst, sdp);
        fused2_step<T, 1>(a, S, lds, rowc, i + 1, nsteps, lo0, hi0, j0, j1, nst, final_here, c0,
                          cl, cw0, ce0, ce1, col1, nrow, D2, P1, P5, A, B, U0, dtr, OA, OB,
                          u0_pole, dt_pole, cfs, Wst, sdp);
        fused2_step<T, 2>(a, S, lds, rowc, i + 2, nsteps, lo0, hi0, j0, j1, nst, final_here, c0,
                          cl, cw0, ce0, ce1, col1, nrow, D2, P1, P5, A, B, U0, dtr, OA, OB,
                          u0_pole, dt_pole, cfs, Wst, sdp);
        fused2_step<T, 3>(a, S, lds, rowc, i + 3, nsteps, lo0, hi0, j0, j1, nst, final_here, c0,
                          cl, cw0, ce0, ce1, col1, nrow, D2, P1, P5, A, B, U0, dtr, OA, OB,
                          u0_pole, dt_pole, cfs, Wst, sdp);
      }
      BodyArgs<T> ba;
      ba.hi0 = hi0;
      ba.ntm = ntm;
      ba.pitch8 = m.pitch * 8;
      ba.vld = cl * 8;
      ba.vst = c0 < nu ? c0 * 8 : 0x7ffffff0;
      ba.storeB = !final_here;
#pragma unroll
      for (int t = 0; t < T; t++)
#pragma unroll
        for (int e = 0; e < 5; e++) ba.kc[t][e] = uniform_double(cfs[t][e]);
      int bytes = (int)(a.rstride * sizeof(double));
      __amdgpu_buffer_rsrc_t rA = field_rsrc(A, bytes), rB = field_rsrc(B, bytes),
                             rU = field_rsrc(U0, bytes), rOA = field_rsrc(OA, bytes),
                             rOB = field_rsrc(OB, bytes);
      // runs of 12 steps (the ring slots' period) that may end after 4 or 8
      // (exits elsewhere in a run cost register spills)
      const int iend = ib + 4 * nquads;
      while (i < iend) {
        int q = lo0 + i;
        const double *rb = rowc + 4 * (i - T);
#define HFT_BODY(u)                                                                           \
  fused2_body_step<T, u, NR>(S, lds, rb + 4 * (u), nrow, q + (u), ba, rA, rB, rU, rOA, rOB, c0, \
                             cw0, ce0, ce1, P1, P5, dtr)
        HFT_BODY(0); HFT_BODY(1); HFT_BODY(2); HFT_BODY(3);
        if (i + 4 >= iend) {
          i += 4;
          break;
        }
        HFT_BODY(4); HFT_BODY(5); HFT_BODY(6); HFT_BODY(7);
        if (i + 8 >= iend) {
          i += 8;
          break;
        }
        HFT_BODY(8); HFT_BODY(9); HFT_BODY(10); HFT_BODY(11);
#undef HFT_BODY
        i += 12;
      }
    }
  }
  for (; i < nsteps; i += 4) {
    fused2_step<T, 0>(a, S, lds, rowc, i, nsteps, lo0, hi0, j0, j1, nst, final_here, c0, cl, cw0,
                      ce0, ce1, col1, nrow, D2, P1, P5, A, B, U0, dtr, OA, OB, u0_pole, dt_pole, cfs, Wst, sdp);
    if (i + 1 < nsteps)
      fused2_step<T, 1>(a, S, lds, rowc, i + 1, nsteps, lo0, hi0, j0, j1, nst, final_here, c0,
                        cl, cw0, ce0, ce1, col1, nrow, D2, P1, P5, A, B, U0, dtr, OA, OB, u0_pole, dt_pole,
                        cfs, Wst, sdp);
    if (i + 2 < nsteps)
      fused2_step<T, 2>(a, S, lds, rowc, i + 2, nsteps, lo0, hi0, j0, j1, nst, final_here, c0,
                        cl, cw0, ce0, ce1, col1, nrow, D2, P1, P5, A, B, U0, dtr, OA, OB, u0_pole, dt_pole,
                        cfs, Wst, sdp);
    if (i + 3 < nsteps)
      fused2_step<T, 3>(a, S, lds, rowc, i + 3, nsteps, lo0, hi0, j0, j1, nst, final_here, c0,
                        cl, cw0, ce0, ce1, col1, nrow, D2, P1, P5, A, B, U0, dtr, OA, OB, u0_pole, dt_pole,
                        cfs, Wst, sdp);
  }
}

// NR: the LDS row length 2 * blockDim when it is a compile-time constant
// (0: taken from blockDim).
template <int T, int NR = 0>
__global__ __launch_bounds__(512) void k_rkg2_fused2(FusedArgs a) {
  extern __shared__ double lds[];  // [T + 1][3][2 * blockDim]: stage levels 0..T-1, then u0
  __shared__ double cfs[T][5];
  __shared__ __align__(16) double rowc[4 * kFusedMaxRows];
  __shared__ double sdp[1024];  // phi cell widths for the polar sums (pole bands)
  int nt = (int)gridDim.x, nb = a.nbands, H = a.H;
  if (a.slot) {
    nt = a.slot[0];
    H = a.slot[1];
    nb = a.slot[2];
    if ((int)blockIdx.x >= nt) return;
  }
  int L = xcd_remap(blockIdx.x, nt);
  int z = L / nb, band = L - z * nb;
  bool stored = false;
#pragma unroll
  for (int t = 0; t < T; t++) stored = stored || a.W[t] != nullptr;
  fused2_tile<T, NR>(a, z, band, H, a.k0, a.A, a.B, a.outA, a.outB, lds, cfs, rowc, sdp,
                     a.body != 0 && !stored);
}

// ---------------------------------------------------------------------------
// Persistent RKG2 super-step (k_rkg2_persist<T>): every fused tile of one
// super-step -- realization z (position in the stage-count order), band b,
// stage group g = stages 2+gT .. 2+gT+T-1 (clipped at s_z) -- in ONE launch.
// Blocks take tiles from a list ordered group by group (longest realizations
// first); a tile of group g waits until group g-1 of its band and of the
// bands either side is done (the rows it reads as halo and the buffer it
// overwrites).  Every tile a block waits for precedes its own in the list
// and was taken by a running block, so the wait always ends; a bound on the
// spins turns a fault of this argument into an error flag instead of a hang.
// Stage levels rotate through four buffers: group 0 reads u_1 (P1) and u0,
// writes P2/P3; odd groups read P2/P3 and write P0/P1; even groups read
// P0/P1 and write P2/P3.
// ---------------------------------------------------------------------------
struct PersistArgs {
  FusedArgs fa;           // mesh, u0 (f), D, coef, s, order, rstride, cdt; H per tile
  double *P[4];
  const unsigned *tiles;  // packed: z | band << 12 | g << 20
  const int *zH;          // [n] band height of realization z
  const int *znb;         // [n] band count of realization z
  const int *zoff;        // [n] offset of z's counters in done
  int ntiles;
  int *counter;           // next tile
  int *done;              // groups completed, per (z, band)
  int *err;               // set when a wait outlasts wait_ticks
  long long wait_ticks;   // bound of one wait, in wall_clock64() ticks
};

// Current value of a counter other blocks update with atomics: read with an
// atomic read-modify-write, which is performed where the atomics are (a
// plain or relaxed atomic load may keep hitting a stale copy in this XCD's
// L2 while another XCD's updates go elsewhere).
__device__ __forceinline__ int poll_int(int *p) {
  return __hip_atomic_fetch_add(p, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

constexpr int kPersistZBits = 12, kPersistBBits = 8;

__global__ void k_zero_ints(int *p, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) p[i] = 0;
}

template <int T>
__global__ __launch_bounds__(512) void k_rkg2_persist(PersistArgs p) {
  extern __shared__ double lds[];
  __shared__ double cfs[T][5];
  __shared__ double rowc[4 * kFusedMaxRows];
  __shared__ double sdp[1024];
  __shared__ int sh_tile, sh_stop;
  const FusedArgs &a = p.fa;
  for (;;) {
    if (threadIdx.x == 0) {
      int stop = 0, tau = 0;
      if (poll_int(p.err)) {
        stop = 1;
      } else {
        tau = atomicAdd(p.counter, 1);
        if (tau >= p.ntiles) {
          stop = 1;
        } else {
          unsigned tl = p.tiles[tau];
          int z = (int)(tl & ((1u << kPersistZBits) - 1));
          int b = (int)((tl >> kPersistZBits) & ((1u << kPersistBBits) - 1));
          int g = (int)(tl >> (kPersistZBits + kPersistBBits));
          if (g > 0) {
            int *dz = p.done + p.zoff[z];
            int nb = p.znb[z], blo = b > 0 ? b - 1 : 0, bhi = b + 1 < nb ? b + 1 : nb - 1;
            long long t0 = wall_clock64();
            for (;;) {
              bool ok = true;
              for (int bb = blo; bb <= bhi; bb++)
                if (poll_int(dz + bb) < g) ok = false;
              if (ok) break;
              if (wall_clock64() - t0 > p.wait_ticks || poll_int(p.err)) {
                __hip_atomic_fetch_or(p.err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                stop = 1;
                break;
              }
              __builtin_amdgcn_s_sleep(2);
            }
          }
        }
      }
      sh_tile = tau;
      sh_stop = stop;
      // acquire: the rows the tile reads were released by the tiles waited for
      __atomic_thread_fence(__ATOMIC_ACQUIRE);
    }
    __syncthreads();
    if (sh_stop) return;
    unsigned tl = p.tiles[sh_tile];
    int z = (int)(tl & ((1u << kPersistZBits) - 1));
    int b = (int)((tl >> kPersistZBits) & ((1u << kPersistBBits) - 1));
    int g = (int)(tl >> (kPersistZBits + kPersistBBits));
    // group 0 reads u_1 (P1) and u0; odd groups P2/P3 -> P0/P1; even P0/P1 -> P2/P3
    const double *Ain = g == 0 ? p.P[1] : ((g & 1) ? p.P[2] : p.P[0]);
    const double *Bin = g == 0 ? a.f : ((g & 1) ? p.P[3] : p.P[1]);
    double *OAo = (g & 1) ? p.P[0] : p.P[2];
    double *OBo = (g & 1) ? p.P[1] : p.P[3];
    fused2_tile<T>(a, z, b, p.zH[z], 2 + g * T, Ain, Bin, OAo, OBo, lds, cfs, rowc, sdp);
    // release: this thread's stores of the tile, then the block's count
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0)
      __hip_atomic_fetch_add(p.done + p.zoff[z] + b, 1, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// Plain operator application y = L x (tests, PTL pass).
struct LapArgs {
  Mesh m;
  Tiles g;
  const double *x;
  double *y;
  const double *D;
  const int *order;
  size_t rstride;
  // PTL: per-realization max |Lx| over points with |x| > 1e-14
  unsigned long long *absmax;  // may be null
  const int *dn = nullptr;     // device realization count (as Stage1Args)
  // PTL: per tile (logical tile index) the tile's max |Lx| over those points
  // (bits, x = low and y = high word), how many of its points reach it (z)
  // and the first such thread (w); may be null
  uint4 *rec = nullptr;
};

// Rows of one k_laplacian block: a column tile of kLapRows consecutive rows,
// whose interior points load their stencils before any arithmetic (four
// rows of loads in flight per thread instead of one).
static constexpr int kLapRows = 4;

__device__ inline int lap_groups(int ntm) { return (ntm + kLapRows - 1) / kLapRows; }

// Eight wavefronts per SIMD (at most 64 VGPRs): more rows in flight per CU
// than the five the unconstrained allocation (over 80 VGPRs) allows,
// 150 -> 132 us per PTL pass at 64 realizations of 512x1024.
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(8))) void k_laplacian(LapArgs a) {
  const Mesh &m = a.m;
  const int ng = lap_groups(m.ntm);
  int nt = (int)gridDim.x;
  if (a.dn) {
    nt = *a.dn * ng * a.g.ctiles;
    if ((int)blockIdx.x >= nt) return;
  }
  const int t = xcd_remap(blockIdx.x, nt);
  const int ct = t % a.g.ctiles, rg = t / a.g.ctiles;
  const int jg = rg % ng, z = rg / ng;
  const int r = a.order[z];
  const int c = ct * kBlock + threadIdx.x, j0 = jg * kLapRows;
  const bool inc = c < m.nu;
  const size_t base = (size_t)r * a.rstride;
  const double *x = a.x + base;
  const double D = a.D[r];
  double L[kLapRows];
  // interior rows: every load unconditional (row and column clamped into
  // the interior, which has rows 1 .. ntm-2 as ntm >= 5), the value kept
  // where the point is an interior one
  const int cc = inc ? c : 0;
#pragma unroll
  for (int q = 0; q < kLapRows; q++) {
    int j = min(max(j0 + q, 1), m.ntm - 2);
    L[q] = lap_interior(m, D, x, j, cc, m.nu, m.pitch);
  }
  // polar rows: every thread of the row takes part in the wavefront sum
  if (j0 == 0) L[0] = lap_pole(m, D, x, true, m.pitch, m.ntm, m.nu);
  const int qs = m.ntm - 1 - j0;
  if (qs < kLapRows) {
    double ls = lap_pole(m, D, x, false, m.pitch, m.ntm, m.nu);
#pragma unroll
    for (int q = 0; q < kLapRows; q++)
      if (q == qs) L[q] = ls;
  }
  double cand[kLapRows];
#pragma unroll
  for (int q = 0; q < kLapRows; q++) {
    cand[q] = 0.0;
    int j = j0 + q;
    if (inc && j < m.ntm) {
      size_t off = (size_t)j * m.pitch + c;
      a.y[base + off] = L[q];
      if (fabs(x[off]) > 1e-14) cand[q] = fabs(L[q]);
    }
  }
  if (a.absmax) {
    // per row: the tile's max, then one atomic per block for the rows'
    // max (|Af| >= 0: the integer order of the bit patterns is the numeric
    // order)
    constexpr int kW = kBlock / 64;
    __shared__ double wmax[kLapRows][kW];
    __shared__ int wcnt[kLapRows][kW], wfirst[kLapRows][kW];
    const int w = threadIdx.x >> 6;
#pragma unroll
    for (int q = 0; q < kLapRows; q++) {
      double wm = cand[q];
      for (int h = 32; h >= 1; h >>= 1) wm = fmax(wm, __shfl_xor(wm, h, 64));
      if ((threadIdx.x & 63) == 0) wmax[q][w] = wm;
    }
    __syncthreads();
    double mx[kLapRows], bmax = 0.0;
#pragma unroll
    for (int q = 0; q < kLapRows; q++) {
      mx[q] = wmax[q][0];
      for (int v = 1; v < kW; v++) mx[q] = fmax(mx[q], wmax[q][v]);
      bmax = fmax(bmax, mx[q]);
    }
    if (threadIdx.x == 0 && bmax > 0.0)
      atomicMax(&a.absmax[r], (unsigned long long)__double_as_longlong(bmax));
    if (a.rec) {
      // each row tile's record: its max, how many points reach it, the
      // first one
#pragma unroll
      for (int q = 0; q < kLapRows; q++) {
        unsigned long long hit = __ballot(mx[q] > 0.0 && cand[q] == mx[q]);
        if ((threadIdx.x & 63) == 0) {
          wcnt[q][w] = __popcll(hit);
          wfirst[q][w] = hit ? (threadIdx.x & ~63) + __ffsll((long long)hit) - 1 : -1;
        }
      }
      __syncthreads();
      const int q = threadIdx.x;
      if (q < kLapRows && j0 + q < m.ntm) {
        int cnt = 0, first = -1;
        for (int v = 0; v < kW; v++) {
          cnt += wcnt[q][v];
          if (first < 0) first = wfirst[q][v];
        }
        double mq = wmax[q][0];
        for (int v = 1; v < kW; v++) mq = fmax(mq, wmax[q][v]);
        unsigned long long b = (unsigned long long)__double_as_longlong(mq);
        size_t rt = ((size_t)z * m.ntm + (j0 + q)) * a.g.ctiles + ct;
        a.rec[rt] = make_uint4((unsigned)b, (unsigned)(b >> 32), (unsigned)cnt, (unsigned)first);
      }
    }
  }
}

// PTL step limit at the points where |Af| equals the maximum
// (get_dtime_diffusion_ptl, src/hipft.f90:5109-5136).
struct PtlArgs {
  Mesh m;
  Tiles g;
  const double *f, *Af;
  const unsigned long long *absmax;  // [nr]
  unsigned long long *dtmin;         // [nr] bits of min(-du/dF)
  const int *order;
  int coupled;      // target = max over the launch's realizations
  int nlaunch;
  size_t rstride;
  const int *dn = nullptr;  // device realization count (replaces nlaunch)
  const uint4 *rec = nullptr;  // k_ptl_pick: k_laplacian's tile records
};

// The PTL step limit from the tile records of k_laplacian: one block per
// realization of the launch visits the records of its tiles and evaluates
// the limit only at the points of tiles whose max equals the target (the
// recorded point when it is the tile's only one, else every point of the
// tile): at each point off the poles where |Af| equals the target and
// |f| > 1e-14, min over the four neighbours with du dF < -1e-14 of
// -du/dF.  The minimum does not depend on the order the points are visited
// in, and no pass over f and Af is needed to find them.
__global__ __launch_bounds__(256) void k_ptl_pick(PtlArgs a) {
  const uint4 *rec = a.rec;
  const Mesh &m = a.m;
  int z = blockIdx.x;
  int nl = a.dn ? *a.dn : a.nlaunch;
  if (z >= nl) return;
  int r = a.order[z];
  double target;
  if (a.coupled) {
    unsigned long long mx = 0;
    for (int q = 0; q < nl; q++) {
      unsigned long long v = a.absmax[a.order[q]];
      mx = v > mx ? v : mx;
    }
    target = __longlong_as_double((long long)mx);
  } else {
    target = __longlong_as_double((long long)a.absmax[r]);
  }
  double best = 1.7976931348623157e308;
  if (target > 0.0) {
    size_t base = (size_t)r * a.rstride;
    const double *f = a.f + base, *A = a.Af + base;
    auto eval = [&](int j, int c) {
      size_t off = (size_t)j * m.pitch + c;
      double fc = f[off], ac = A[off];
      if (target == fabs(ac) && fabs(fc) > 1e-14) {
        size_t nb[4] = {off - m.pitch, off + m.pitch, (size_t)j * m.pitch + wrapc(c - 1, m.nu),
                        (size_t)j * m.pitch + wrapc(c + 1, m.nu)};
        for (int q = 0; q < 4; q++) {
          double du = fc - f[nb[q]];
          double dF = ac - A[nb[q]];
          if (du * dF < -1e-14) best = fmin(-du / dF, best);
        }
      }
    };
    const int per = a.g.ntm * a.g.ctiles;
    for (int k = threadIdx.x; k < per; k += blockDim.x) {
      uint4 q = rec[(size_t)z * per + k];
      unsigned long long b = (unsigned long long)q.x | ((unsigned long long)q.y << 32);
      if (__longlong_as_double((long long)b) != target || q.z == 0) continue;
      int j = k / a.g.ctiles, ct = k - j * a.g.ctiles;
      if (j == 0 || j == m.ntm - 1) continue;
      if (q.z == 1) {
        int c = ct * kBlock + (int)q.w;
        if (c < m.nu) eval(j, c);
      } else {
        for (int c = ct * kBlock; c < min(m.nu, (ct + 1) * kBlock); c++) eval(j, c);
      }
    }
  }
  __shared__ double wb[4];
  for (int h = 32; h >= 1; h >>= 1) best = fmin(best, __shfl_xor(best, h, 64));
  if ((threadIdx.x & 63) == 0) wb[threadIdx.x >> 6] = best;
  __syncthreads();
  if (threadIdx.x == 0) {
    double b = wb[0];
    for (int q = 1; q < (int)(blockDim.x >> 6); q++) b = fmin(b, wb[q]);
    if (b < 1.7976931348623157e308)
      atomicMin(&a.dtmin[r], (unsigned long long)__double_as_longlong(b));
  }
}

// ---------------------------------------------------------------------------
// Advection: WENO3-CS(h) with local Lax-Friedrichs splitting on the
// nonuniform grid (advection_operator_weno3, src/hipft.f90:7711-7925) fused
// with one SSPRK(4,3) stage (src/hipft.f90:7551-7574).
// ---------------------------------------------------------------------------
struct WenoW {
  double mcm, cpm, mc0, cp0, ptm, mctm, mt0, cpt0, em, e0;
};

__device__ inline WenoW weno_w(const double *CP, const double *MC, const double *MT,
                               const double *CPT, const double *PT, const double *MCT,
                               const double *eps, int im1, int i0) {
  WenoW w;
  w.mcm = MC[im1];
  w.cpm = CP[im1];
  w.mc0 = MC[i0];
  w.cp0 = CP[i0];
  w.ptm = PT[im1];
  w.mctm = MCT[im1];
  w.mt0 = MT[i0];
  w.cpt0 = CPT[i0];
  w.em = eps[im1];
  w.e0 = eps[i0];
  return w;
}

// Correctly rounded FP64 quotients without the divide's scaling and fix-up
// steps.  The compiler's FP64 divide is div_scale (numerator, denominator),
// a hardware reciprocal refined by two Newton steps, the quotient, its
// residual and div_fmas (a fused multiply-add, scaled only when div_scale
// scaled), then div_fixup (special values).  When numerator and
// denominator are finite, normal, within 2^767 of each other's exponent
// and the quotient is normal, div_scale leaves both operands unchanged,
// div_fmas is a plain fma and div_fixup returns its input, so this sequence
// gives the same bits.  Two quotients by one denominator share its refined
// reciprocal.
struct RefinedRecip {
  double b, y;
};

__device__ __forceinline__ RefinedRecip refined_recip(double b) {
  double y = __builtin_amdgcn_rcp(b);
  double e = fma(-b, y, 1.0);
  y = fma(y, e, y);
  e = fma(-b, y, 1.0);
  y = fma(y, e, y);
  return {b, y};
}

__device__ __forceinline__ double quotient(double a, const RefinedRecip &r) {
  double q = a * r.y;
  double res = fma(-r.b, q, a);
  return fma(res, r.y, q);
}

// One face flux from the split fluxes at the six surrounding cells
// (src/hipft.f90:7775-7801, 7852-7878).  fastdiv (Mesh::fastdiv): the
// weights' numerators and eps lie in [2^-100, 2^100], so where the four
// smoothness terms e do too (e >= eps holds; e <= 2^100 is tested), every
// quotient below has operands in [2^-300, 2^301], no more than 2^601
// apart, and a normal result: the short divides give the divide's bits.
// Elsewhere (large or non-finite terms) the full divide runs.
__device__ inline double weno_face(double LNm2, double LNm1, double LN0, double LPm1, double LP0,
                                   double LPp1, const WenoW &w, int fastdiv = 0) {
  const double one = 1.0, four = 4.0;
  double p0m = (one + w.mcm) * LNm1 - w.mcm * LNm2;
  double p1m = w.mc0 * LNm1 + w.cpm * LN0;
  double p0p = (one + w.cp0) * LP0 - w.cp0 * LPp1;
  double p1p = w.cpm * LP0 + w.mc0 * LPm1;
  double t0 = w.mcm * (LNm1 - LNm2), t1 = w.cpm * (LN0 - LNm1);
  double t2 = w.cp0 * (LPp1 - LP0), t3 = w.mc0 * (LP0 - LPm1);
  double B0m = four * (t0 * t0), B1m = four * (t1 * t1);
  double B0p = four * (t2 * t2), B1p = four * (t3 * t3);
  double e0 = w.em + B0m, e1 = w.em + B1m, e2 = w.e0 + B0p, e3 = w.e0 + B1p;
  double OM0m, OM1m, OM0p, OM1p;
  const double kHi = 0x1p100;
  if (fastdiv && e0 <= kHi && e1 <= kHi && e2 <= kHi && e3 <= kHi) {
    double w0m = quotient(w.ptm, refined_recip(e0 * e0));
    double w1m = quotient(w.mctm, refined_recip(e1 * e1));
    double w0p = quotient(w.mt0, refined_recip(e2 * e2));
    double w1p = quotient(w.cpt0, refined_recip(e3 * e3));
    RefinedRecip rm = refined_recip(w0m + w1m), rp = refined_recip(w0p + w1p);
    OM0m = quotient(w0m, rm);
    OM1m = quotient(w1m, rm);
    OM0p = quotient(w0p, rp);
    OM1p = quotient(w1p, rp);
  } else {
    double w0m = w.ptm / (e0 * e0);
    double w1m = w.mctm / (e1 * e1);
    double w0p = w.mt0 / (e2 * e2);
    double w1p = w.cpt0 / (e3 * e3);
    double wm_sum = w0m + w1m, wp_sum = w0p + w1p;
    OM0m = w0m / wm_sum;
    OM1m = w1m / wm_sum;
    OM0p = w0p / wp_sum;
    OM1p = w1p / wp_sum;
  }
  double um = OM0m * p0m + OM1m * p1m;
  double up = OM0p * p0p + OM1p * p1p;
  return up + um;
}

// sign(1, v) of Fortran: -1 for negative values (and -0.0), else +1.
__device__ inline double fsign1(double v) { return signbit(v) ? -1.0 : 1.0; }

template <bool F2D>
struct FlowAcc {
  Flow1D f1;
  Flow2D f2;
  int nt, ntm, pitch, nu;
  size_t rstride;
  int r;
  // vt at theta face jf (0-based ref jf-1), column c
  __device__ double vt(int jf0, int c) const {
    if (F2D) return f2.vt[(size_t)r * f2.vt_rstride + (size_t)jf0 * pitch + c];
    return f1.vt[(size_t)r * nt + jf0];
  }
  __device__ double at(int j, int c) const {
    if (F2D) return f2.at[(size_t)r * rstride + (size_t)j * pitch + c];
    return f1.at[(size_t)r * ntm + j];
  }
  // vp at phi face q (q = reference k-2 of the face), row j
  __device__ double vp(int j, int q) const {
    if (F2D) return f2.vp[(size_t)r * rstride + (size_t)j * pitch + q];
    return f1.vp[(size_t)r * ntm + j];
  }
  __device__ double ap(int j, int c) const {
    if (F2D) return f2.ap[(size_t)r * rstride + (size_t)j * pitch + c];
    return f1.ap[(size_t)r * ntm + j];
  }
};

struct AdvArgs {
  Mesh m;
  Tiles g;
  Flow1D f1;
  Flow2D f2;
  const double *x;   // operator input
  const double *y;   // second input of stage 3 (may be null)
  double *out;
  double ca, cb, cdt;  // out = ca*y + cb*x - cdt*A(x)   (or x - cdt*A(x))
  const double *dtr;   // per-realization sub-step, when set: cdt = cmul * dtr[r]
  double cmul;
  const int *order;
  size_t rstride;
  int write_aop;       // operator-only mode: out = A(x)
  const int *skip = nullptr;  // when set and *skip != 0 the launch does nothing
};

template <bool F2D>
__device__ inline double theta_face_upwind(const FlowAcc<F2D> &fa, const double *x, int jf0,
                                           int c, int pitch, int ntm) {
  // jf0 = 1 (reference face 2) or nt-2 (reference face nt-1)
  const double one = 1.0, half = 0.5;
  double v = fa.vt(jf0, c);
  double cct = fsign1(v);
  double xa, xb;  // (one-cct) multiplies xa, (one+cct) multiplies xb
  if (jf0 == 1) {
    xa = x[(size_t)1 * pitch + c];
    xb = x[c];
  } else {
    xa = x[(size_t)(ntm - 1) * pitch + c];
    xb = x[(size_t)(ntm - 2) * pitch + c];
  }
  return v * half * ((one - cct) * xa + (one + cct) * xb);
}

template <bool F2D>
__device__ inline double theta_face_weno(const Mesh &m, const FlowAcc<F2D> &fa, const double *x,
                                         int jf, int c) {
  // reference face jf (3..nt-2): cells jf-2..jf+1 (reference) = rows jf-3..jf
  const double half = 0.5;
  double LN[3], LP[3];
  for (int q = 0; q < 3; q++) {
    int jr = jf - 2 + q;  // reference cell for LN: jf-2, jf-1, jf
    int j0 = jr - 1;
    double xv = x[(size_t)j0 * m.pitch + c];
    double a = fa.at(j0, c);
    LN[q] = half * xv * (fa.vt(jr - 1, c) + a);
  }
  for (int q = 0; q < 3; q++) {
    int jr = jf - 1 + q;  // reference cell for LP: jf-1, jf, jf+1
    int j0 = jr - 1;
    double xv = x[(size_t)j0 * m.pitch + c];
    double a = fa.at(j0, c);
    LP[q] = half * xv * (fa.vt(jr, c) - a);
  }
  WenoW w = weno_w(m.tCP, m.tMC, m.tMT, m.tCPT, m.tPT, m.tMCT, m.teps, jf - 2, jf - 1);
  return weno_face(LN[0], LN[1], LN[2], LP[0], LP[1], LP[2], w, m.fastdiv);
}

template <bool F2D>
__device__ inline double phi_face_weno(const Mesh &m, const FlowAcc<F2D> &fa, const double *xrow,
                                       int j, int q) {
  // face q <-> reference k = q+2; cells k-2..k+1 <-> columns q-2..q+1
  const double half = 0.5;
  int nu = m.nu;
  double LN[3], LP[3];
  for (int d = 0; d < 3; d++) {
    int cc = wrapc(q - 2 + d, nu);
    double xv = xrow[cc];
    double a = fa.ap(j, cc);
    LN[d] = half * xv * (fa.vp(j, cc) + a);  // vp at the cell's west face
  }
  for (int d = 0; d < 3; d++) {
    int cc = wrapc(q - 1 + d, nu);
    double xv = xrow[cc];
    double a = fa.ap(j, cc);
    LP[d] = half * xv * (fa.vp(j, wrapc(cc + 1, nu)) - a);  // east face
  }
  WenoW w = weno_w(m.pCP, m.pMC, m.pMT, m.pCPT, m.pPT, m.pMCT, m.peps, q, q + 1);
  return weno_face(LN[0], LN[1], LN[2], LP[0], LP[1], LP[2], w, m.fastdiv);
}

template <bool F2D>
__device__ inline double advection_op(const Mesh &m, const FlowAcc<F2D> &fa, const double *x,
                                      int j, int c) {
  int ntm = m.ntm, nt = m.nt, pitch = m.pitch, nu = m.nu;
  if (j == 0 || j == ntm - 1) {
    bool north = j == 0;
    int jf0 = north ? 1 : nt - 2;
    double s = wave_pole_sum(nu, [&](int cc) {
      return theta_face_upwind(fa, x, jf0, cc, pitch, ntm) * m.dp[cc + 1];
    });
    return north ? s * m.bc_flow_n : -s * m.bc_flow_s;
  }
  int jr = j + 1;  // reference cell
  double ft0 = (jr == 2) ? theta_face_upwind(fa, x, 1, c, pitch, ntm)
                         : theta_face_weno(m, fa, x, jr, c);
  double ft1 = (jr + 1 == nt - 1) ? theta_face_upwind(fa, x, nt - 2, c, pitch, ntm)
                                  : theta_face_weno(m, fa, x, jr + 1, c);
  const double *xrow = x + (size_t)j * pitch;
  double fp0 = phi_face_weno(m, fa, xrow, j, c);
  double fp1 = phi_face_weno(m, fa, xrow, j, wrapc(c + 1, nu));
  double sti = m.st_i[j];
  double at_ = (m.sth[j + 1] * ft1 - m.sth[j] * ft0) * sti;
  at_ = at_ * m.dt_i[j];
  double ap_ = (fp1 - fp0) * sti;
  ap_ = ap_ * m.dp_i[c + 1];
  return at_ + ap_;
}

template <bool F2D, bool HAS_Y>
__global__ __launch_bounds__(kBlock) void k_advection_stage(AdvArgs a) {
  if (a.skip && *a.skip) return;
  int t = xcd_remap(blockIdx.x, gridDim.x);
  int z, j, ct;
  tile_of(t, a.g, z, j, ct);
  int r = a.order[z];
  const Mesh &m = a.m;
  int c = ct * kBlock + threadIdx.x;
  FlowAcc<F2D> fa;
  fa.f1 = a.f1;
  fa.f2 = a.f2;
  fa.nt = m.nt;
  fa.ntm = m.ntm;
  fa.pitch = m.pitch;
  fa.nu = m.nu;
  fa.rstride = a.rstride;
  fa.r = r;
  size_t base = (size_t)r * a.rstride;
  const double *x = a.x + base;
  bool pole = (j == 0 || j == m.ntm - 1);
  if (!pole && c >= m.nu) return;
  double A = advection_op<F2D>(m, fa, x, j, c < m.nu ? c : 0);
  if (c >= m.nu) return;
  size_t off = (size_t)j * m.pitch + c;
  double cdt = a.dtr ? a.cmul * a.dtr[r] : a.cdt;
  double v;
  if (a.write_aop) {
    v = A;
  } else if (HAS_Y) {
    v = a.ca * a.y[base + off] + a.cb * x[off];
    v = v - cdt * A;
  } else {
    v = x[off] - cdt * A;
  }
  a.out[base + off] = v;
}

// ---------------------------------------------------------------------------
// Banded WENO3 stage: each face flux computed once.
//
// A block owns one realization and a band of theta rows over the full phi
// width (one thread per column) and marches down the band.  Theta: the
// thread keeps its column's Lax-Friedrichs split fluxes LP/LN of four rows in
// registers and computes the face below the current row once, carrying the
// face above from the previous row.  Phi: each row's split fluxes go to LDS;
// a thread forms its cell's west face, publishes it in LDS and takes the east
// face from its neighbour after the next row's barrier (one barrier per row).
// Pole rows sum the upwind pole-face fluxes over the full row (fixed wave
// order).  Expressions are those of advection_op.
// ---------------------------------------------------------------------------
struct AdvBandArgs {
  Mesh m;
  Flow1D f1;
  Flow2D f2;
  const double *x, *y;
  double *out;
  double ca, cb, cdt;
  const double *dtr;   // per-realization sub-step, when set: cdt = cmul * dtr[r]
  double cmul;
  int write_aop;
  int nbands, H, ncol;
  const int *order;
  size_t rstride;
  const int *skip = nullptr;  // when set and *skip != 0 the launch does nothing
};

// Row-uniform values per row of the band table, and the most rows a band
// may have (the host caps the band height at kAdvMaxBandRows).
constexpr int kAdvRowVals = 17;
constexpr int kAdvMaxBandRows = 128;

template <bool F2D, bool HAS_Y>
__global__ __launch_bounds__(1024) void k_advection_band(AdvBandArgs a) {
  // phi split fluxes LN, LP and west faces, two buffers each by row parity,
  // the pole-face terms (7 x ncol doubles), the own column's phi WENO
  // weights (5 x ncol pairs, each thread reads only its own), then the row
  // table (kAdvRowVals x (H + 6))
  extern __shared__ double sh[];
  const Mesh &m = a.m;
  const int ncol = a.ncol, nu = m.nu, ntm = m.ntm, nt = m.nt, pitch = m.pitch;
  double2 *sW = reinterpret_cast<double2 *>(sh + 7 * ncol);
  if (a.skip && *a.skip) return;
  int L = xcd_remap(blockIdx.x, gridDim.x);
  int z = L / a.nbands, band = L - z * a.nbands;
  int r = a.order[z];
  int c = threadIdx.x;
  bool col = c < nu;
  int cc = col ? c : 0;
  {
    // west face of column cc: weights at reference faces cc (m1) and cc+1 (0)
    WenoW w = weno_w(m.pCP, m.pMC, m.pMT, m.pCPT, m.pPT, m.pMCT, m.peps, cc, cc + 1);
    sW[0 * ncol + c] = make_double2(w.mcm, w.cpm);
    sW[1 * ncol + c] = make_double2(w.mc0, w.cp0);
    sW[2 * ncol + c] = make_double2(w.ptm, w.mctm);
    sW[3 * ncol + c] = make_double2(w.mt0, w.cpt0);
    sW[4 * ncol + c] = make_double2(w.em, w.e0);
  }
  const double dpi = m.dp_i[cc + 1];
  FlowAcc<F2D> fa;
  fa.f1 = a.f1;
  fa.f2 = a.f2;
  fa.nt = nt;
  fa.ntm = ntm;
  fa.pitch = pitch;
  fa.nu = nu;
  fa.rstride = a.rstride;
  fa.r = r;
  size_t base = (size_t)r * a.rstride;
  const double *x = a.x + base;
  const double half = 0.5;
  const double cdt = a.dtr ? a.cmul * a.dtr[r] : a.cdt;
  int j0 = band * a.H, j1 = min(ntm, j0 + a.H);
  // The band's row-uniform values (rows and faces j0-2 .. j1+3) in LDS, read
  // as broadcasts: waits on them do not wait on the row loads in flight.
  // k = 0 vt(face), 1 at, 2 vp, 3 ap (1-D flows), 4 st_i, 5 sth, 6 dt_i,
  // 7..16 the theta WENO weights of the face below row j.
  const int R0 = j0 - 2, NR = j1 + 4 - R0;
  double *tb = sh + 17 * ncol;
  for (int idx = threadIdx.x; idx < kAdvRowVals * NR; idx += blockDim.x) {
    int k = idx / NR, jj = R0 + (idx - k * NR);
    double v = 0.0;
    bool row = jj >= 0 && jj < ntm, face = jj >= 0 && jj < nt;
    bool wf = jj >= 0 && jj + 1 < nt;
    switch (k) {
      case 0: if (!F2D && face) v = a.f1.vt[(size_t)r * nt + jj]; break;
      case 1: if (!F2D && row) v = a.f1.at[(size_t)r * ntm + jj]; break;
      case 2: if (!F2D && row) v = a.f1.vp[(size_t)r * ntm + jj]; break;
      case 3: if (!F2D && row) v = a.f1.ap[(size_t)r * ntm + jj]; break;
      case 4: if (row) v = m.st_i[jj]; break;
      case 5: if (face) v = m.sth[jj]; break;
      case 6: if (row) v = m.dt_i[jj]; break;
      case 7: if (wf) v = m.tMC[jj]; break;
      case 8: if (wf) v = m.tCP[jj]; break;
      case 9: if (wf) v = m.tMC[jj + 1]; break;
      case 10: if (wf) v = m.tCP[jj + 1]; break;
      case 11: if (wf) v = m.tPT[jj]; break;
      case 12: if (wf) v = m.tMCT[jj]; break;
      case 13: if (wf) v = m.tMT[jj + 1]; break;
      case 14: if (wf) v = m.tCPT[jj + 1]; break;
      case 15: if (wf) v = m.teps[jj]; break;
      default: if (wf) v = m.teps[jj + 1]; break;
    }
    tb[idx] = v;
  }
  lds_barrier();
  auto T_ = [&](int k, int jj) { return tb[k * NR + (jj - R0)]; };
  auto VT = [&](int jf0) { return F2D ? fa.vt(jf0, cc) : T_(0, jf0); };
  auto AT = [&](int jj) { return F2D ? fa.at(jj, cc) : T_(1, jj); };
  // theta split fluxes of the own column: LN/LP of reference cell jr
  auto ln_t = [&](int j0r, double xv) { return half * xv * (VT(j0r) + AT(j0r)); };
  auto lp_t = [&](int j0r, double xv) { return half * xv * (VT(j0r + 1) - AT(j0r)); };
  auto xat = [&](int j) { return (j >= 0 && j < ntm) ? x[(size_t)j * pitch + cc] : 0.0; };
  // theta face flux below row j (between rows j and j+1, reference face j+2)
  auto tface = [&](int j, const double *LN4, const double *LP4, double xj, double xj1) -> double {
    int jf = j + 2;  // reference face
    if (jf == 2 || jf == nt - 1) {
      const double one = 1.0;
      double v = VT(jf - 1);
      double cct = fsign1(v);
      // face 2: (one-cct)*x(row1) + (one+cct)*x(row0); face nt-1: rows ntm-1, ntm-2
      return v * half * ((one - cct) * xj1 + (one + cct) * xj);
    }
    // weights of weno_w(theta arrays, jf-2, jf-1), from the row table
    WenoW w;
    w.mcm = T_(7, j); w.cpm = T_(8, j); w.mc0 = T_(9, j); w.cp0 = T_(10, j);
    w.ptm = T_(11, j); w.mctm = T_(12, j); w.mt0 = T_(13, j); w.cpt0 = T_(14, j);
    w.em = T_(15, j); w.e0 = T_(16, j);
    // LN at rows j-1, j, j+1; LP at rows j, j+1, j+2 (slots 0..3 = rows j-1..j+2)
    return weno_face(LN4[0], LN4[1], LN4[2], LP4[1], LP4[2], LP4[3], w, m.fastdiv);
  };
  // rows j-1..j+2 of the own column
  double X[4], LN4[4], LP4[4];
  for (int d = 0; d < 4; d++) {
    int jj = j0 - 1 + d;
    X[d] = xat(jj);
    LN4[d] = (jj >= 0 && jj < ntm) ? ln_t(jj, X[d]) : 0.0;
    LP4[d] = (jj >= 0 && jj < ntm) ? lp_t(jj, X[d]) : 0.0;
  }
  // face above the first row (between j0-1 and j0)
  double ft_up = 0.0;
  if (j0 >= 1) {
    double Xm[4], LNm[4], LPm[4];
    for (int d = 0; d < 4; d++) {
      int jj = j0 - 2 + d;
      Xm[d] = xat(jj);
      LNm[d] = (jj >= 0 && jj < ntm) ? ln_t(jj, Xm[d]) : 0.0;
      LPm[d] = (jj >= 0 && jj < ntm) ? lp_t(jj, Xm[d]) : 0.0;
    }
    ft_up = tface(j0 - 1, LNm, LPm, Xm[1], Xm[2]);
  }
  // Row j is finished one barrier later, when its east phi face (the west
  // face of column c+1) is in LDS: pj, pX, pfu, pfd, pfw hold its value, theta
  // faces and west face meanwhile, and its store waits for the end of the
  // next row.  LDS buffers alternate by row parity, so one barrier per row
  // separates every write from the reads of it and from the next write to
  // the same buffer.  Row j+4 of x is loaded during row j (clamped into the
  // field, straight-line) into one of two registers used in turn, so a wait
  // for it never waits for the load issued after it.
  const double dpc = m.dp[cc + 1];
  auto store = [&](int jr, double v) {
    if (col) a.out[base + (size_t)jr * pitch + c] = v;
  };
  auto value = [&](double xv, double yv, double A) {
    double v;
    if (a.write_aop) {
      v = A;
    } else if (HAS_Y) {
      v = a.ca * yv + a.cb * xv;
      v = v - cdt * A;
    } else {
      v = xv - cdt * A;
    }
    return v;
  };
  auto xload = [&](int jj) { return x[(size_t)min(max(jj, 0), ntm - 1) * pitch + cc]; };
  bool pend = false;
  int pj = 0;
  double pX = 0.0, pfu = 0.0, pfd = 0.0, pfw = 0.0;
  bool due = false;
  int dj = 0;
  double dX = 0.0, dY = 0.0, dA = 0.0;
  auto row = [&](int j, double xcur, double &xnext) {
    // xcur: row j+3 of x, loaded during row j-1 (any value when outside)
    // X[1] = row j, X[0] = j-1, X[2] = j+1, X[3] = j+2
    bool pole = (j == 0 || j == ntm - 1);
    const int b = j & 1;
    double *sLNb = sh + b * ncol, *sLPb = sh + (2 + b) * ncol, *sFPb = sh + (4 + b) * ncol;
    double *sPL = sh + 6 * ncol;
    double ft_dn = 0.0;
    if (j < ntm - 1) ft_dn = tface(j, LN4, LP4, X[1], X[2]);
    if (pole) {
      // upwind pole-face flux of the own column, summed over the row below
      double fface = (j == 0) ? ft_dn : ft_up;
      if (col) sPL[c] = fface * dpc;
    } else {
      // phi split fluxes of row j (own cell), shared through LDS
      double xv = X[1];
      double av = F2D ? fa.ap(j, cc) : T_(3, j);
      double vw = F2D ? fa.vp(j, cc) : T_(2, j);
      double ve = F2D ? fa.vp(j, wrapc(cc + 1, nu)) : T_(2, j);
      sLNb[c] = half * xv * (vw + av);
      sLPb[c] = half * xv * (ve - av);
    }
    lds_barrier();
    if (pend) {
      double fpe = sh[(4 + (pj & 1)) * ncol + wrapc(cc + 1, nu)];
      double sti = T_(4, pj);
      double at_ = (T_(5, pj + 1) * pfd - T_(5, pj) * pfu) * sti;
      at_ = at_ * T_(6, pj);
      double ap_ = (fpe - pfw) * sti;
      ap_ = ap_ * dpi;
      due = true;
      dj = pj;
      dX = pX;
      dA = at_ + ap_;
      if (HAS_Y) dY = a.y[base + (size_t)pj * pitch + cc];
      pend = false;
    }
    xnext = xload(j + 4);  // after the y load, so a wait for y does not wait for it
    if (pole) {
      double s = wave_pole_sum(nu, [&](int k) { return sPL[k]; });
      double yv = HAS_Y ? a.y[base + (size_t)j * pitch + cc] : 0.0;
      store(j, value(X[1], yv, (j == 0) ? s * m.bc_flow_n : -s * m.bc_flow_s));
    } else {
      // west face of the own cell: face q = c
      double fpw = 0.0;
      if (col) {
        int c2 = wrapc(cc - 2, nu), c1 = wrapc(cc - 1, nu), cp = wrapc(cc + 1, nu);
        WenoW w;
        double2 q0 = sW[0 * ncol + c], q1 = sW[1 * ncol + c], q2 = sW[2 * ncol + c];
        double2 q3 = sW[3 * ncol + c], q4 = sW[4 * ncol + c];
        w.mcm = q0.x; w.cpm = q0.y; w.mc0 = q1.x; w.cp0 = q1.y; w.ptm = q2.x;
        w.mctm = q2.y; w.mt0 = q3.x; w.cpt0 = q3.y; w.em = q4.x; w.e0 = q4.y;
        fpw = weno_face(sLNb[c2], sLNb[c1], sLNb[cc], sLPb[c1], sLPb[cc], sLPb[cp], w, m.fastdiv);
      }
      sFPb[c] = fpw;
      pend = true;
      pj = j;
      pX = X[1];
      pfu = ft_up;
      pfd = ft_dn;
      pfw = fpw;
    }
    // advance the window by one row
    ft_up = ft_dn;
    int jn = j + 3;
    double xn = (jn < ntm) ? xcur : 0.0;
#pragma unroll
    for (int d = 0; d < 3; d++) {
      X[d] = X[d + 1];
      LN4[d] = LN4[d + 1];
      LP4[d] = LP4[d + 1];
    }
    X[3] = xn;
    LN4[3] = (jn < ntm) ? ln_t(jn, xn) : 0.0;
    LP4[3] = (jn < ntm) ? lp_t(jn, xn) : 0.0;
    if (due) {
      store(dj, value(dX, dY, dA));
      due = false;
    }
  };
  double xa = xload(j0 + 3), xb = 0.0;
  for (int j = j0; j < j1; j += 2) {
    row(j, xa, xb);
    if (j + 1 < j1) row(j + 1, xb, xa);
  }
  if (pend) {
    lds_barrier();
    double fpe = sh[(4 + (pj & 1)) * ncol + wrapc(cc + 1, nu)];
    double sti = T_(4, pj);
    double at_ = (T_(5, pj + 1) * pfd - T_(5, pj) * pfu) * sti;
    at_ = at_ * T_(6, pj);
    double ap_ = (fpe - pfw) * sti;
    ap_ = ap_ * dpi;
    double yv = HAS_Y ? a.y[base + (size_t)pj * pitch + cc] : 0.0;
    store(pj, value(pX, yv, at_ + ap_));
  }
}

// ---------------------------------------------------------------------------
// Analysis (analysis_step, src/hipft.f90:3145-3325): per-block partials then
// a fixed-order final reduction.
// ---------------------------------------------------------------------------
static constexpr int kNPart = 16;
struct AnaArgs {
  const double *f;
  const double *fval;        // validation field in device layout or null
  const double *fval_scale;  // [nr] or null
  int ntm, nu, pitch, nr;
  size_t rstride;
  // 1-D weights: theta [ntm]: da_t, sn_t, cs_t, t; phi [npm1] over the
  // reference's i = 1..npm-1: da_p, cs_p, sn_p, column index
  const double *da_t, *sn_t, *cs_t, *tt;
  const double *da_p, *cs_p, *sn_p;
  const int *pcol;
  double lat_n, lat_s;   // pole_flux_lat_limit*d2r and pi - that
  int rows_per_block;
  double *part;          // [nr][nblk][kNPart]
  int nblk;
};

__device__ inline void block_reduce(double *sh, double v, int op, double *out) {
  // op 0 sum, 1 min, 2 max; fixed tree order
  int tid = threadIdx.x;
  sh[tid] = v;
  __syncthreads();
  for (int h = kBlock / 2; h >= 1; h >>= 1) {
    if (tid < h) {
      double a = sh[tid], b = sh[tid + h];
      sh[tid] = op == 0 ? a + b : (op == 1 ? fmin(a, b) : fmax(a, b));
    }
    __syncthreads();
  }
  if (tid == 0) *out = sh[0];
  __syncthreads();
}

// Columns per chunk of a k_analysis_part thread (1280 = 5 x 256 columns
// cover the 1024-column map in one chunk).
static constexpr int kAnaU = 5;

__global__ __launch_bounds__(kBlock) void k_analysis_part(AnaArgs a) {
  int r = blockIdx.y, b = blockIdx.x;
  const double huge = 1.7976931348623157e308, tiny = 2.2250738585072014e-308;
  double mn = huge, mx = tiny, mnabs = huge, verr = 0.0, sfv = 0.0;
  double fp = 0, fm = 0, fppn = 0, fmpn = 0, fpps = 0, fmps = 0, apn = 0, aps = 0;
  double e1 = 0, e2 = 0, ax = 0;
  int j0 = b * a.rows_per_block, j1 = min(a.ntm, j0 + a.rows_per_block);
  size_t base = (size_t)r * a.rstride;
  double scale = a.fval_scale ? a.fval_scale[r] : 1.0;
  for (int j = j0; j < j1; j++) {
    const double *row = a.f + base + (size_t)j * a.pitch;
    const double *vrow = a.fval ? a.fval + base + (size_t)j * a.pitch : nullptr;
    double dat = a.da_t[j], snt = a.sn_t[j], cst = a.cs_t[j], tj = a.tt[j];
    // one pass over the row: the reference's columns i = 1..npm-1 are the
    // unique columns nu-1, 0, 1, ..., nu-1 (pcol, computed here rather than
    // loaded, so a point's load does not wait on another); the
    // min/max/error terms take each unique column once (i >= 1, column i-1).
    // A thread takes its columns in chunks of kAnaU, every load of a chunk
    // issued before its arithmetic, which runs in the column order.
    for (int i0 = threadIdx.x; i0 < a.nu + 1; i0 += kAnaU * kBlock) {
      double fr[kAnaU];
#pragma unroll
      for (int u = 0; u < kAnaU; u++) {
        int i = i0 + u * kBlock;
        fr[u] = i < a.nu + 1 ? row[i == 0 ? a.nu - 1 : i - 1] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < kAnaU; u++) {
        int i = i0 + u * kBlock;
        if (i >= a.nu + 1) break;
        double f = fr[u];
        if (i >= 1) {
          int c = i - 1;
          double fv = vrow ? (a.fval_scale ? vrow[c] * scale : vrow[c]) : f;
          mn = fmin(f, mn);
          mx = fmax(f, mx);
          mnabs = fmin(fabs(f), mnabs);
          double d = f - fv;
          verr = verr + d * d;
          sfv = sfv + fabs(f) * fabs(fv);
        }
        double dap = a.da_p[i];
        double fv = f * dat;
        fv = fv * dap;
        if (fv > 0.0) fp = fp + fv; else fm = fm + fv;
        double area = dat * dap;
        if (tj <= a.lat_n) {
          if (fv > 0.0) fppn = fppn + fv; else fmpn = fmpn + fv;
          apn = apn + area;
        }
        if (tj >= a.lat_s) {
          if (fv > 0.0) fpps = fpps + fv; else fmps = fmps + fv;
          aps = aps + area;
        }
        double q = f * snt;
        e1 = e1 + ((q * a.cs_p[i]) * dat) * dap;
        e2 = e2 + ((q * a.sn_p[i]) * dat) * dap;
        ax = ax + ((f * cst) * dat) * dap;
      }
    }
  }
  double *o = a.part + ((size_t)r * a.nblk + b) * kNPart;
  double vals[kNPart] = {mn, mx, mnabs, verr, sfv, fp, fm, fppn, fmpn, fpps, fmps, apn, aps, e1, e2, ax};
  // every partial reduced across the wavefront by a fixed butterfly, then
  // the block's wavefronts combined in order by one thread per partial (one
  // barrier for all sixteen instead of a shared-memory tree each)
  __shared__ double wpart[kBlock / 64][kNPart];
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int q = 0; q < kNPart; q++) {
    double v = vals[q];
    for (int h = 32; h >= 1; h >>= 1) {
      double u = __shfl_xor(v, h, 64);
      v = q == 0 || q == 2 ? fmin(v, u) : (q == 1 ? fmax(v, u) : v + u);
    }
    if (lane == 0) wpart[w][q] = v;
  }
  __syncthreads();
  if (threadIdx.x < kNPart) {
    int q = threadIdx.x;
    double v = wpart[0][q];
    for (int k = 1; k < kBlock / 64; k++) {
      double u = wpart[k][q];
      v = q == 0 || q == 2 ? fmin(v, u) : (q == 1 ? fmax(v, u) : v + u);
    }
    o[q] = v;
  }
}

// kNPart lanes per realization (launched with kNPart * nr threads in blocks
// of 64): lane q combines partial q over the row blocks in block order, then
// the realization's first lane forms the 14 history values.
__global__ void k_analysis_final(const double *part, int nblk, int nr, int ntm, int npm,
                                 double *out) {
  static_assert(64 % kNPart == 0, "a realization's lanes lie in one wavefront");
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  const int r = g / kNPart, q = g % kNPart;
  const double huge = 1.7976931348623157e308, tiny = 2.2250738585072014e-308;
  const int op = q == 0 || q == 2 ? 0 : (q == 1 ? 1 : 2);  // min, max, sum
  double acc = op == 0 ? huge : (op == 1 ? tiny : 0.0);
  auto comb = [&](double v) { acc = op == 0 ? fmin(acc, v) : (op == 1 ? fmax(acc, v) : acc + v); };
  if (r < nr) {
    const double *pr = part + (size_t)r * nblk * kNPart + q;
    int b = 0;
    for (; b + 8 <= nblk; b += 8) {
      double v[8];
#pragma unroll
      for (int u = 0; u < 8; u++) v[u] = pr[(size_t)(b + u) * kNPart];
#pragma unroll
      for (int u = 0; u < 8; u++) comb(v[u]);
    }
    for (; b < nblk; b++) comb(pr[(size_t)b * kNPart]);
  }
  double accs[kNPart];
  const int lane0 = (threadIdx.x & 63) & ~(kNPart - 1);
#pragma unroll
  for (int k = 0; k < kNPart; k++) accs[k] = __shfl(acc, lane0 + k, 64);
  if (r >= nr || q != 0) return;
  const double rsun_cm2 = 4.84416e21, three_quarter = 0.75, pi_i = 0.3183098861837907;
  double *o = out + (size_t)r * 14;
  o[0] = accs[0];
  o[1] = accs[1];
  o[2] = accs[2];
  o[13] = (accs[4] == 0.0) ? sqrt(accs[3] / (double)(ntm * (npm - 1))) : sqrt(accs[3] / accs[4]);
  for (int k = 0; k < 8; k++) o[3 + k] = rsun_cm2 * accs[5 + k];
  o[12] = three_quarter * pi_i * accs[15];
  double e1 = three_quarter * pi_i * accs[13];
  double e2 = three_quarter * pi_i * accs[14];
  o[11] = sqrt(e1 * e1 + e2 * e2);
}

// ---------------------------------------------------------------------------
// Data assimilation blend (src/hipft.f90:4242-4271, 4115-4141).
// ---------------------------------------------------------------------------
struct AssimArgs {
  Mesh m;
  Tiles g;
  double *f;
  double *delta;            // optional: store delta instead of applying (balance)
  const double *map;        // [(nr)][3][ntm][npm1]
  size_t map_rstride;       // 0 when shared by all realizations
  int npm1;
  int custom;
  const double *mu_power, *mu_limit;
  const int *tidx0, *tidx1; // 1-based lat-limit indices, -1 = none; used if lat_on
  const int *lat_on;
  double mult;
  int add_mode;
  const int *order;
  size_t rstride;
};

__global__ __launch_bounds__(kBlock) void k_assimilate(AssimArgs a) {
  int t = xcd_remap(blockIdx.x, gridDim.x);
  int z, j, ct;
  tile_of(t, a.g, z, j, ct);
  int r = a.order[z];
  int c = ct * kBlock + threadIdx.x;
  if (c >= a.m.nu) return;
  // unique column c <-> reference k = c+2 <-> map column l = k (1-based)
  const double *mp = a.map + (size_t)r * a.map_rstride;
  size_t sl = (size_t)a.m.ntm * a.npm1;
  size_t q = (size_t)j * a.npm1 + (c + 1);
  double d = mp[q], w = mp[sl + q];
  if (a.custom) {
    double mu = mp[2 * sl + q];
    w = (mu > a.mu_limit[r]) ? pow(mu, a.mu_power[r]) : 0.0;
  }
  if (a.lat_on[r]) {
    int jr = j + 1;
    if (jr <= a.tidx0[r] || (a.tidx1[r] >= 1 && jr >= a.tidx1[r])) w = 0.0;
  }
  d = a.mult * d;
  size_t p = (size_t)r * a.rstride + (size_t)j * a.m.pitch + c;
  double fo = a.f[p];
  double dl = a.add_mode ? w * d : w * (d - fo);
  if (a.delta)
    a.delta[p] = dl;
  else
    a.f[p] = fo + dl;
}

// Flux of a field per realization (get_flux, src/hipft.f90:8672-8738) as
// per-block partials [nr][nblk][2]; reuses the analysis weights.
__global__ __launch_bounds__(kBlock) void k_flux_part(AnaArgs a) {
  __shared__ double sh[kBlock];
  int r = blockIdx.y, b = blockIdx.x;
  double fp = 0, fm = 0;
  int j0 = b * a.rows_per_block, j1 = min(a.ntm, j0 + a.rows_per_block);
  for (int j = j0; j < j1; j++) {
    const double *row = a.f + (size_t)r * a.rstride + (size_t)j * a.pitch;
    double dat = a.da_t[j];
    for (int i = threadIdx.x; i < a.nu + 1; i += kBlock) {
      double fv = row[a.pcol[i]] * dat;
      fv = fv * a.da_p[i];
      if (fv > 0.0) fp = fp + fv; else fm = fm + fv;
    }
  }
  double *o = a.part + ((size_t)r * a.nblk + b) * 2;
  block_reduce(sh, fp, 0, o);
  block_reduce(sh, fm, 0, o + 1);
}

// f += balanced delta (balance_flux, src/hipft.f90:8788-8801).
__global__ __launch_bounds__(kBlock) void k_apply_balanced(double *f, const double *delta,
                                                           const double *ratio, int ntm, int nu,
                                                           int pitch, size_t rstride) {
  int r = blockIdx.y;
  size_t n = (size_t)ntm * pitch;
  for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock) {
    int c = (int)(i % pitch);
    if (c >= nu) continue;
    size_t p = (size_t)r * rstride + i;
    double br = delta[p];
    double rt = ratio[r];
    double d = (rt == 0.0) ? br : (br > 0 ? br / rt : br * rt);
    f[p] = f[p] + d;
  }
}

// ---------------------------------------------------------------------------
// Sources (update_source, src/hipft.f90:3875-3947; generate_rfe 8809-8921;
// source_step 4697-4733).
// ---------------------------------------------------------------------------
// Random flux emergence frame: rfe = rnd * scale_r (model 2) or
// ((rnd * scale_r) * da_ti(j)) * dp_i(k) (model 1), rnd one shared map.
__global__ __launch_bounds__(kBlock) void k_rfe_frame(double *rfe, const double *rnd,
                                                      const double *scale, const double *da_ti,
                                                      const double *dp_i, int model, int ntm,
                                                      int nu, int pitch, size_t rstride) {
  int r = blockIdx.y;
  size_t n = (size_t)ntm * pitch;
  double sc = scale[r];
  for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (size_t)gridDim.x * kBlock) {
    int c = (int)(i % pitch);
    if (c >= nu) continue;
    int j = (int)(i / pitch);
    double v;
    if (model == 1) {
      v = rnd[i] * sc;
      v = v * da_ti[j];
      v = v * dp_i[c + 1];
    } else {
      v = rnd[i] * sc;
    }
    rfe[(size_t)r * rstride + i] = v;
  }
}

// Pole rows of a frame := sum_{k=1}^{npm-2} rfe(pole,k) dp(k) / (2 pi), in
// the reference's sequential k order (k = 1 is the seam column nu-1).
// Block r, thread 0 sums; the block writes the row.
__global__ __launch_bounds__(64) void k_rfe_poles(double *rfe, const double *dp, int ntm, int nu,
                                                  int pitch, size_t rstride, double twopi_i) {
  __shared__ double fpole[2];
  double *base = rfe + (size_t)blockIdx.x * rstride;
  if (threadIdx.x < 2) {
    const double *row = base + (size_t)(threadIdx.x == 0 ? 0 : ntm - 1) * pitch;
    double s = 0.0;
    s = s + row[nu - 1] * dp[0];
    for (int c = 0; c <= nu - 2; c++) s = s + row[c] * dp[c + 1];
    fpole[threadIdx.x] = s * twopi_i;
  }
  __syncthreads();
  for (int c = threadIdx.x; c < nu; c += blockDim.x) {
    base[c] = fpole[0];
    base[(size_t)(ntm - 1) * pitch + c] = fpole[1];
  }
}

// map := balanced map (balance_flux, src/hipft.f90:8788-8801); ratio 0 leaves
// the map unchanged (zero net flux).
__global__ __launch_bounds__(kBlock) void k_balance_inplace(double *map, const double *ratio,
                                                            int ntm, int nu, int pitch,
                                                            size_t rstride) {
  int r = blockIdx.y;
  size_t n = (size_t)ntm * pitch;
  double rt = ratio[r];
  if (rt == 0.0) return;
  for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (size_t)gridDim.x * kBlock) {
    int c = (int)(i % pitch);
    if (c >= nu) continue;
    size_t p = (size_t)r * rstride + i;
    double br = map[p];
    map[p] = br > 0 ? br / rt : br * rt;
  }
}

// f += dt * source, source = ((0 + file) + (1 - alpha) rfe1) + alpha rfe2
// with the absent terms left out (update_source's order).
__global__ __launch_bounds__(kBlock) void k_source_step(double *f, const double *file,
                                                        const double *r1, const double *r2,
                                                        double alpha, double dt, int ntm, int nu,
                                                        int pitch, size_t rstride) {
  int r = blockIdx.y;
  size_t n = (size_t)ntm * pitch;
  const double one = 1.0;
  for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < n;
       i += (size_t)gridDim.x * kBlock) {
    int c = (int)(i % pitch);
    if (c >= nu) continue;
    size_t p = (size_t)r * rstride + i;
    double s = 0.0;
    if (file) s = s + file[i];
    if (r1) {
      s = s + (one - alpha) * r1[p];
      s = s + alpha * r2[p];
    }
    f[p] = f[p] + dt * s;
  }
}

// 2-D flows: vt (nt rows) and vp (ntm rows, faces q) per realization.
struct AttenArgs {
  const double *f;
  const double *vt1, *vp1;   // [nr][nt], [nr][ntm] assembled 1-D flows
  // 2-D base flows (flows from files; null: the 1-D flows above are used):
  // vtf [rf][nt][pitch], vpf [rf][ntm][pitch], rf = r or 0 (shared, stride 0)
  const double *vtf, *vpf;
  size_t vtf_rstride, vpf_rstride;
  // analytic terms added to the base flows in update_flow's order
  // (src/hipft.f90:3427-3445): vt - mft[r][j] + vtc, vp + drp[r][j] + omp[j]
  const double *mft, *drp, *omp;
  double vtc;
  int has_vtc;
  double *vt, *vp;           // 2-D outputs
  const double *inv;         // [nr] 1/flow_attenuate_value, or null (no attenuation)
  int ntm, nt, nu, pitch;
  size_t rstride, vt_rstride;
};

// 2-D flows: update_flow (src/hipft.f90:3411-3510) from 2-D base flows and/or
// 1-D flows, attenuated by |Br| (3489-3507) when a.inv is set.
__global__ __launch_bounds__(kBlock) void k_attenuate(AttenArgs a) {
  int r = blockIdx.z, j = blockIdx.y;
  int c = blockIdx.x * kBlock + threadIdx.x;
  if (c >= a.nu) return;
  const double one = 1.0;
  const double *f = a.f + (size_t)r * a.rstride;
  // vt(jf) for jf = 1..nt (0-based j = jf-1 < nt): attenuated for jf = 2..nt-1
  if (j < a.nt) {
    double v;
    if (a.vtf) {
      v = a.vtf[(size_t)r * a.vtf_rstride + (size_t)j * a.pitch + c];
      if (a.mft) v = v - a.mft[(size_t)r * a.nt + j];
      if (a.has_vtc) v = v + a.vtc;
    } else {
      v = a.vt1[(size_t)r * a.nt + j];
    }
    if (a.inv && j >= 1 && j <= a.nt - 2) {
      double b0 = fabs(f[(size_t)(j - 1) * a.pitch + c]), b1 = fabs(f[(size_t)j * a.pitch + c]);
      double m = b0 > b1 ? b0 : b1;
      v = v * (one - tanh(m * a.inv[r]));
    }
    a.vt[(size_t)r * a.vt_rstride + (size_t)j * a.pitch + c] = v;
  }
  // vp at face q = c (reference k = c+2, between columns c-1 and c)
  if (j < a.ntm) {
    double v;
    if (a.vpf) {
      v = a.vpf[(size_t)r * a.vpf_rstride + (size_t)j * a.pitch + c];
      if (a.drp) v = v + a.drp[(size_t)r * a.ntm + j];
      if (a.omp) v = v + a.omp[j];
    } else {
      v = a.vp1[(size_t)r * a.ntm + j];
    }
    if (a.inv) {
      double b0 = fabs(f[(size_t)j * a.pitch + wrapc(c - 1, a.nu)]);
      double b1 = fabs(f[(size_t)j * a.pitch + c]);
      double m = b0 > b1 ? b0 : b1;
      v = v * (one - tanh(m * a.inv[r]));
    }
    a.vp[(size_t)r * a.rstride + (size_t)j * a.pitch + c] = v;
  }
}

// Lax-Friedrichs alphas for 2-D flows (set_lf_alpha, src/hipft.f90:6968-7046).
__global__ __launch_bounds__(kBlock) void k_lf_alpha(const double *vt, const double *vp,
                                                     double *at, double *ap, int ntm, int nt,
                                                     int nu, int pitch, size_t rstride,
                                                     size_t vt_rstride) {
  int r = blockIdx.z, j = blockIdx.y;
  int c = blockIdx.x * kBlock + threadIdx.x;
  if (c >= nu) return;
  const double *V = vt + (size_t)r * vt_rstride;
  // alpha_t(jr) for jr = 1..nt-1 (= ntm): window of |vt| faces
  int jr = j + 1;
  int lo, hi;
  if (jr == 1) { lo = 1; hi = 4; }
  else if (jr == 2) { lo = 1; hi = 5; }
  else if (jr == nt - 2) { lo = jr - 2; hi = jr + 2; }
  else if (jr == nt - 1) { lo = jr - 2; hi = jr + 1; }
  else { lo = jr - 2; hi = jr + 3; }
  double m = fabs(V[(size_t)(lo - 1) * pitch + c]);
  for (int q = lo + 1; q <= hi; q++) m = fmax(m, fabs(V[(size_t)(q - 1) * pitch + c]));
  at[(size_t)r * rstride + (size_t)j * pitch + c] = m;
  // alpha_p: rows 2..ntm-1, window k-2..k+3 over faces (periodic)
  double mp = 0.0;
  if (jr >= 2 && jr <= ntm - 1) {
    const double *P = vp + (size_t)r * rstride + (size_t)j * pitch;
    mp = fabs(P[wrapc(c - 2, nu)]);
    for (int d = -1; d <= 3; d++) mp = fmax(mp, fabs(P[wrapc(c + d, nu)]));
  }
  ap[(size_t)r * rstride + (size_t)j * pitch + c] = mp;
}

// CFL limit of 2-D flows (get_flow_dtmax, src/hipft.f90:7321-7327):
// per-realization min of 1/2/max(kdotv, tiny) as bits (positive doubles).
__global__ __launch_bounds__(kBlock) void k_flow_dtmax(const double *vt, const double *vp,
                                                       const double *dt_i, const double *st_i,
                                                       const double *dp_i, int ntm, int nu,
                                                       int pitch, size_t rstride,
                                                       size_t vt_rstride,
                                                       unsigned long long *out) {
  int r = blockIdx.z, j = blockIdx.y + 1;  // rows 1..ntm-2
  int c = blockIdx.x * kBlock + threadIdx.x;
  double best = 1.7976931348623157e308;
  if (c < nu && j <= ntm - 2) {
    const double *V = vt + (size_t)r * vt_rstride;
    const double *P = vp + (size_t)r * rstride + (size_t)j * pitch;
    double a = fmax(fabs(V[(size_t)j * pitch + c]), fabs(V[(size_t)(j + 1) * pitch + c]));
    double b = fmax(fabs(P[c]), fabs(P[wrapc(c + 1, nu)]));
    double kdotv = a * dt_i[j] + b * st_i[j] * dp_i[c + 1];
    best = 0.5 / fmax(kdotv, 2.2250738585072014e-308);
  }
  for (int h = 32; h >= 1; h >>= 1) best = fmin(best, __shfl_xor(best, h, 64));
  if ((threadIdx.x & 63) == 0) atomicMin(&out[r], (unsigned long long)__double_as_longlong(best));
}


// ---------------------------------------------------------------------------
// Device-planned diffusion step: the controller of diffusion_step
// (src/hipft.f90:4589-4690), the PTL step rule of get_dtime_diffusion_ptl
// (5077-5156) and the RKG2 tables of load_sts_rkg2 (5399-5461) evaluated on
// the device, so that a whole time step is one fixed launch sequence that
// can be replayed as a hipGraph without a host round trip:
//   k_plan_init;
//   per round:  [k_plan_prep, k_laplacian, k_ptl_pick]  (PTL sub-cycling)
//               k_plan_round, k_plan_coefs, k_rkg2_stage1,
//               gcap x k_rkg2_fused2 (launch q: stages 2+qT .. 2+qT+T-1),
//               k_copy_finals_plan;
//   k_plan_end.
// k_plan_round is one block: every controller (one per realization, or one
// for the batch in coupled mode) takes the host controller's steps in the
// same floating-point order, so step sizes, stage counts and coefficients
// are the host's bit for bit.  A round whose super-step would need more
// fused launches than the sequence holds (or a step that needs more rounds)
// sets the overflow flag instead: that round and every later kernel of the
// sequence do nothing, and the host resumes the step from the controller
// state in device memory.
// ---------------------------------------------------------------------------
struct PlanArgs {
  // controller state, one entry per controller g (ng = 1 coupled, nr independent)
  double *stepped, *dt2;
  int64_t *sdiag;
  int *cyc, *done;
  int *status;        // [0] overflow, [1] round to resume at, [2] realizations this round
  int64_t *smax;      // [rounds] largest stage count a round needed (0: none cycled)
  int64_t *last_s, *last_cycles;  // [nr] hft_diffusion_stats
  // super-step tables read by the round's kernels
  int *order;         // launch index -> realization (PTL pass: ascending; super-step: by s)
  int64_t *s;         // [nr]
  double *cdt, *ub1;  // [nr] cycle time step, ubj(1)
  StageCoef *coef;    // [k][nr]
  int *slots;         // [gcap][3]: tiles, H, nbands of fused launch q
  const int *htab;    // [nr + 1] band rows of a fused launch over a realizations
  const double *dte;  // [ng] explicit Euler limit
  unsigned long long *absmax, *dtmin;  // [nr] PTL results (bits of non-negative doubles)
  double dtl;         // step length
  double dt2_fixed;   // dtl / subcycles when subcycles > 0
  int subcycles, coupled, nr, ng, T, ntm, round, gcap, kspan;
};

// Exclusive count of set flags over the block (thread order), and the total.
__device__ inline int block_prefix(bool flag, int *wsum, int *total) {
  unsigned long long mk = __ballot(flag);
  int lane = threadIdx.x & 63, w = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
  int before = __popcll(mk & ((1ull << lane) - 1ull));
  if (lane == 0) wsum[w] = __popcll(mk);
  __syncthreads();
  int off = 0, tot = 0;
  for (int q = 0; q < nw; q++) {
    if (q < w) off += wsum[q];
    tot += wsum[q];
  }
  __syncthreads();
  *total = tot;
  return off + before;
}

// load_sts_rkg2 (src/hipft.f90:5399-5461): stage count of a cycle of length
// dtl for the explicit limit dte, and the coefficients of stage k >= 2.
__device__ inline int64_t plan_sts_stages(double dtl, double dte) {
  double s_real = 0.5 * (sqrt(25.0 + 24.0 * (dtl / dte)) - 3.0);
  double sc = ceil(s_real);
  int64_t s = sc > 9.0e15 ? (int64_t)9.0e15 : (int64_t)sc;
  if (s < 3) s = 3;
  if (s % 2 == 0) s = s + 1;
  return s;
}

__device__ inline double plan_sts_b(int64_t j) {
  if (j == 1) return 1.0 / 3.0;
  if (j == 2) return 1.0 / 15.0;
  double dj = (double)j;
  return (4.0 * (double)(j - 1) * (double)(j + 4)) /
         (3.0 * dj * (double)(j + 1) * (double)(j + 2) * (double)(j + 3));
}

__device__ inline double plan_sts_w(int64_t s) { return 6.0 / (double)((s + 4) * (s - 1)); }

__device__ inline StageCoef plan_sts_coef(int64_t s, int64_t k, double dtl) {
  double w = plan_sts_w(s);
  double u, v, ub, gm;
  if (k == 2) {
    u = 1.0 / 2.0;
    v = -1.0 / 10.0;
    ub = u * w;
    gm = 0.0;
  } else {
    double dj = (double)k;
    double bj = plan_sts_b(k), bjm1 = plan_sts_b(k - 1), bjm2 = plan_sts_b(k - 2);
    double bj_bjm1 = bj / bjm1;
    double bj_bjm2 = bj / bjm2;
    u = bj_bjm1 * (2.0 + 1.0 / dj);
    v = -bj_bjm2 * (1.0 / dj + 1.0);
    ub = u * w;
    gm = (0.5 * dj * (dj + 1.0) * bjm1 - 1.0) * ub;
  }
  StageCoef cf;
  cf.a = u;
  cf.b = v;
  cf.c = 1.0 - u - v;
  cf.d = ub * dtl;
  cf.g = gm;
  cf.w = nullptr;
  cf.pad[0] = cf.pad[1] = 0.0;
  return cf;
}

__global__ __launch_bounds__(1024) void k_plan_init(PlanArgs p) {
  for (int g = threadIdx.x; g < p.ng; g += blockDim.x) {
    p.stepped[g] = 0.0;
    p.dt2[g] = p.subcycles > 0 ? p.dt2_fixed : 0.0;
    p.sdiag[g] = p.subcycles;
    p.cyc[g] = 1;
    p.done[g] = 0;
  }
  if (threadIdx.x == 0) {
    p.status[0] = 0;
    p.status[1] = -1;
    p.status[2] = 0;
  }
}

// PTL pass set-up of a round (ptl_pass): the realizations whose controller
// still cycles, ascending, into order / status[2]; the max and min
// accumulators reset.  One block, nr <= blockDim.
__global__ __launch_bounds__(1024) void k_plan_prep(PlanArgs p) {
  __shared__ int wsum[16];
  int r = threadIdx.x, total;
  bool ovf = p.status[0] != 0;
  bool act = !ovf && r < p.nr && !p.done[p.coupled ? 0 : r];
  int idx = block_prefix(act, wsum, &total);
  if (act) p.order[idx] = r;
  if (r < p.nr) {
    p.absmax[r] = 0ull;
    p.dtmin[r] = (unsigned long long)__double_as_longlong(1.7976931348623157e308);
  }
  if (threadIdx.x == 0) p.status[2] = total;
}

// One round of the controller (diffusion_step's loop body, with ptl_pass's
// step rule) and the super-step's tables.  One block, nr <= blockDim.
__global__ __launch_bounds__(1024) void k_plan_round(PlanArgs p) {
  __shared__ int wsum[16];
  __shared__ long long s_of[1024];   // stage count per realization (-1: not in the round)
  __shared__ long long s_sorted[1024];
  __shared__ unsigned long long red_mx[16], red_mn[16];
  __shared__ long long red_s[16];
  __shared__ double b_d2, b_stp;
  __shared__ long long b_s, b_sd;
  __shared__ int b_act, b_cyc, b_ovf;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, nw = (blockDim.x + 63) >> 6;
  const bool auto_sc = p.subcycles == 0;
  const double kBig = 1.7976931348623157e308;
  if (p.status[0]) {
    for (int q = tid; q < p.gcap; q += blockDim.x) {
      p.slots[3 * q] = 0;
      p.slots[3 * q + 1] = 1;
      p.slots[3 * q + 2] = 1;
    }
    if (tid == 0) p.status[2] = 0;
    return;
  }
  // coupled PTL: one step for the batch, from the max of |L f| and the min
  // of the step limits over every realization
  unsigned long long mx = 0ull, mn = (unsigned long long)__double_as_longlong(kBig);
  if (p.coupled && auto_sc) {
    for (int r = tid; r < p.nr; r += blockDim.x) {
      mx = p.absmax[r] > mx ? p.absmax[r] : mx;
      mn = p.dtmin[r] < mn ? p.dtmin[r] : mn;
    }
    for (int h = 32; h >= 1; h >>= 1) {
      unsigned long long a = __shfl_xor(mx, h, 64), b = __shfl_xor(mn, h, 64);
      mx = a > mx ? a : mx;
      mn = b < mn ? b : mn;
    }
    if (lane == 0) {
      red_mx[w] = mx;
      red_mn[w] = mn;
    }
    __syncthreads();
    mx = 0ull;
    mn = (unsigned long long)__double_as_longlong(kBig);
    for (int q = 0; q < nw; q++) {
      mx = red_mx[q] > mx ? red_mx[q] : mx;
      mn = red_mn[q] < mn ? red_mn[q] : mn;
    }
  }
  // controller g = tid: this round's cycle time step, stage count and
  // diagnostic sub-cycle count
  int g = tid;
  bool act = g < p.ng && !p.done[g];
  double d2 = 0.0, stp = 0.0;
  int dn = 0, cyc = 0;
  long long sd = 0, sg = 0;
  if (act) {
    stp = p.stepped[g];
    d2 = p.dt2[g];
    cyc = p.cyc[g];
    if (auto_sc) {
      unsigned long long axb = p.coupled ? mx : p.absmax[g];
      unsigned long long dmb = p.coupled ? mn : p.dtmin[g];
      double ax = __longlong_as_double((long long)axb), dmin = __longlong_as_double((long long)dmb);
      double dtp = kBig;
      if (ax > 0.0) {
        dtp = 0.95 * dmin;
        if (dtp < d2) dtp = d2;
      }
      d2 = dtp;
    }
    if (auto_sc && stp + d2 >= p.dtl) {
      d2 = p.dtl - stp;
      dn = 1;
    }
    if (auto_sc && cyc == 60 - 1) {
      d2 = p.dtl - stp;
      dn = 1;
    }
    sd = p.sdiag[g];
    if (auto_sc) {
      double q = ceil((p.dtl - stp) / d2);
      sd = q > 2e9 ? 2000000000LL : (long long)q;
    }
    sg = plan_sts_stages(d2, p.dte[g]);
  }
  // largest stage count of the round -> fused launches it needs
  long long sm = act ? sg : 0;
  for (int h = 32; h >= 1; h >>= 1) {
    long long o = __shfl_xor(sm, h, 64);
    sm = o > sm ? o : sm;
  }
  if (lane == 0) red_s[w] = sm;
  if (p.coupled && tid == 0) {
    b_act = act ? 1 : 0;
    b_d2 = d2;
    b_stp = stp;
    b_s = sg;
    b_sd = sd;
    b_cyc = cyc;
  }
  __syncthreads();
  sm = 0;
  for (int q = 0; q < nw; q++) sm = red_s[q] > sm ? red_s[q] : sm;
  int need = sm > 0 ? (int)((sm - 1 + p.T - 1) / p.T) : 0;
  if (tid == 0) {
    p.smax[p.round] = sm;
    b_ovf = need > p.gcap ? 1 : 0;
    if (b_ovf) {
      p.status[0] = 1;
      p.status[1] = p.round;
    }
  }
  __syncthreads();
  if (b_ovf || sm == 0) {
    for (int q = tid; q < p.gcap; q += blockDim.x) {
      p.slots[3 * q] = 0;
      p.slots[3 * q + 1] = 1;
      p.slots[3 * q + 2] = 1;
    }
    if (tid == 0) p.status[2] = 0;
    return;
  }
  // per realization r = tid
  int r = tid;
  bool ract = false;
  double rd2 = 0.0, rstp = 0.0;
  long long rs = 0, rsd = 0;
  int rcyc = 0;
  if (r < p.nr) {
    if (p.coupled) {
      ract = b_act != 0;
      rd2 = b_d2; rstp = b_stp; rs = b_s; rsd = b_sd; rcyc = b_cyc;
    } else {
      ract = act;
      rd2 = d2; rstp = stp; rs = sg; rsd = sd; rcyc = cyc;
    }
    s_of[r] = ract ? rs : -1;
  }
  __syncthreads();
  if (ract) {
    // stable order by stage count, descending (ties in realization order)
    int rank = 0;
    for (int q = 0; q < p.nr; q++) {
      long long o = s_of[q];
      rank += (o > rs || (o == rs && q < r)) ? 1 : 0;
    }
    p.order[rank] = r;
    s_sorted[rank] = rs;
    p.s[r] = rs;
    p.cdt[r] = rd2;
    p.ub1[r] = plan_sts_w(rs);
    p.last_s[r] = rs;
    p.last_cycles[r] = rcyc;
  }
  int n;
  block_prefix(ract, wsum, &n);  // (also a barrier: s_sorted complete)
  // controller state after the round
  if (act) {
    p.dt2[g] = d2;
    p.sdiag[g] = sd;
    p.stepped[g] = stp + d2;
    int dnn = dn;
    if (!auto_sc && cyc == p.subcycles) dnn = 1;
    if (dnn) p.done[g] = 1;
    else p.cyc[g] = cyc + 1;
  }
  // fused launch q advances the realizations with s >= 2 + qT
  for (int q = tid; q < p.gcap; q += blockDim.x) {
    long long k0 = 2 + (long long)q * p.T;
    int lo = 0, hi = n;
    while (lo < hi) {
      int mid = (lo + hi) >> 1;
      if (s_sorted[mid] >= k0) lo = mid + 1;
      else hi = mid;
    }
    int a = lo;
    if (a == 0) {
      p.slots[3 * q] = 0;
      p.slots[3 * q + 1] = 1;
      p.slots[3 * q + 2] = 1;
    } else {
      int H = p.htab[a];
      int nb = (p.ntm + H - 1) / H;
      p.slots[3 * q] = a * nb;
      p.slots[3 * q + 1] = H;
      p.slots[3 * q + 2] = nb;
    }
  }
  if (tid == 0) p.status[2] = n;
}

// Coefficient table of the round's super-step: thread (z, k) writes stage
// k = 2..s of realization order[z].
__global__ __launch_bounds__(kBlock) void k_plan_coefs(PlanArgs p) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int z = i / p.kspan;
  int64_t k = 2 + (i - z * p.kspan);
  if (z >= p.status[2]) return;
  int r = p.order[z];
  int64_t s = p.s[r];
  if (k > s) return;
  p.coef[(size_t)k * p.nr + r] = plan_sts_coef(s, k, p.cdt[r]);
}

// The stage table of a host-controlled super-step on the device (the
// expressions of k_plan_coefs): realizations order[0..n) with stage counts
// s and cycle steps cdt, one thread per (realization, stage k >= 2).
struct CoefArgs {
  StageCoef *coef;  // [k][nr]
  const int *order;
  const int64_t *s;
  const double *cdt;
  int nr, n, kspan;  // kspan: stages 2..kspan+1 per realization
};

__global__ __launch_bounds__(kBlock) void k_rkg2_coefs(CoefArgs a) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int z = i / a.kspan;
  int64_t k = 2 + (i - z * a.kspan);
  if (z >= a.n) return;
  int r = a.order[z];
  int64_t s = a.s[r];
  if (k > s) return;
  a.coef[(size_t)k * a.nr + r] = plan_sts_coef(s, k, a.cdt[r]);
}

// Final stages of a planned super-step into f: fused launch q writes its
// last stage to src[q & 1], and realization r finishes in launch (s_r-2)/T.
struct PlanCopyArgs {
  double *dst;
  const double *src[2];
  const int *order;
  const int *n;
  const int64_t *s;
  size_t rstride;
  int per, T;
};

__global__ __launch_bounds__(kBlock) void k_copy_finals_plan(PlanCopyArgs a) {
  int z = blockIdx.x / a.per;
  if (z >= *a.n) return;
  size_t i = (size_t)(blockIdx.x - z * a.per) * kBlock + threadIdx.x;
  if (i >= a.rstride / 2) return;
  int r = a.order[z];
  int q = (int)((a.s[r] - 2) / a.T);
  size_t o = (size_t)r * a.rstride + 2 * i;
  *reinterpret_cast<double2 *>(a.dst + o) = *reinterpret_cast<const double2 *>(a.src[q & 1] + o);
}

// After the last round of the sequence (p.round = the number of rounds it
// holds): a controller that has not finished its step needs more rounds.
__global__ void k_plan_end(PlanArgs p) {
  if (threadIdx.x != 0 || p.status[0]) return;
  for (int g = 0; g < p.ng; g++)
    if (!p.done[g]) {
      p.status[0] = 1;
      p.status[1] = p.round;
      return;
    }
}

}  // namespace hft
