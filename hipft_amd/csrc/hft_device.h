// hft_device.h -- device-side data structures shared by the HIP kernels and
// the host controller of libhft.
//
// Storage of a field (state, RKG2 stage buffers, operator output):
//   a[(r*ntm + j)*pitch + c]
// r = realization, j = 0..ntm-1 theta row (reference row j+1, poles at 0 and
// ntm-1), c = 0..nu-1 unique phi column (reference column c+2; the reference's
// overlap columns 1 and npm are columns nu-1 and 0 here, so the phi direction
// is purely periodic with period nu = npm-2).  pitch >= nu is padded to a
// multiple of 64 doubles so every row starts on a 512-byte boundary.
//
// One-dimensional mesh arrays are the reference's 1-based arrays shifted to
// 0-based: m.X[i-1] == X(i) of src/hipft.f90 (module mesh, 242-259).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace hft {

struct Mesh {
  int ntm, nt, npm, np, nu, pitch;
  // theta arrays: lengths ntm (t, dt_i, st_i) or nt (sth, dth_i)
  const double *dt_i, *st_i, *sth, *dth_i, *dth;
  // phi arrays of length np (dp_i, dph_i, dp) indexed by reference k-1
  const double *dp_i, *dph_i, *dp;
  // WENO3 grid weights (load_weno, src/hipft.f90:6905-6932), reference index-1
  const double *tCP, *tMC, *tMT, *tCPT, *tPT, *tMCT, *teps;  // theta, length nt
  const double *pCP, *pMC, *pMT, *pCPT, *pPT, *pMCT, *peps;  // phi, length np
  double bc_diff_n, bc_diff_s;  // bc_diffusion_{n,s}pole_fac
  double bc_flow_n, bc_flow_s;  // bc_flow_{n,s}pole_fac
  // every WENO3 numerator weight (PT, MCT, MT, CPT) and eps lies in
  // [2^-100, 2^100]: weno_face may take its short divide where the
  // smoothness terms allow (set by the host)
  int fastdiv;
};

// u_k = a*u_{k-1} + b*u_{k-2} + c*u0 + d*L(u_{k-1}) + g*dty0
// (diffusion_step_sts_cd stage k, src/hipft.f90:5226-5231, with
//  c = one-uj-vj and d = ubj*dt evaluated on the host in reference order).
// w: the stored stage u_{k-1} of the forward recomputation (field base),
// set only for adjoint sweeps over stored stages.
struct StageCoef {
  double a, b, c, d, g;
  const double *w;
  double pad[2];
};

// Per-realization 1-D flow profiles for phi-independent flows.
struct Flow1D {
  const double *vt;  // [nr][nt]   vt(j) on theta faces
  const double *at;  // [nr][ntm]  alpha_t(j)
  const double *vp;  // [nr][ntm]  vp(j)
  const double *ap;  // [nr][ntm]  alpha_p(j)
};

// Per-realization 2-D flows (attenuated flows): same storage as a field but
// vt has nt rows; alpha_t, alpha_p have ntm rows.
struct Flow2D {
  const double *vt, *vp, *at, *ap;
  size_t vt_rstride;  // nt*pitch
};

// Workgroup barrier that orders LDS only.  The band kernels exchange data
// between waves through LDS alone; a full __syncthreads() would also wait
// for every outstanding global load (vmcnt(0)) and so stall the loads they
// prefetch rows ahead.
__device__ inline void lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

__host__ __device__ inline size_t fidx(int r, int j, int c, int ntm, int pitch) {
  return ((size_t)r * ntm + j) * (size_t)pitch + c;
}

}  // namespace hft
