/*
 * hft.h -- C ABI of the MI355X flux-transport core (libhft.so).
 *
 * The reference (HipFT, src/hipft.f90) has no plugin layer: its time loop
 * (src/hipft.f90:903-948) calls Fortran subroutines that act on module state.
 * This ABI replaces those call sites.  A Fortran driver binds it through
 * iso_c_binding (include/hft.f90); Python binds it through ctypes
 * (hipft_amd/core.py).
 *
 * Conventions
 *  - Every function returns int: 0 on success, a negative HFT_E_* code on
 *    error; hft_last_error() returns the message of the calling thread's last
 *    error.  Nothing in the library calls exit() or abort().
 *  - All arrays are double precision.  Host pointers are borrowed for the call
 *    only; the library copies what it needs.
 *  - Layout HFT_LAYOUT_TP is the reference's state array f(ntm,npm,nr)
 *    (Fortran column-major, theta fastest, two-point phi overlap:
 *    column 1 == column npm-1, column npm == column 2), src/hipft.f90:2114.
 *    HFT_LAYOUT_PT is a map as stored in files: (npm1,ntm,nr), phi fastest,
 *    one-point overlap (first and last phi columns are the same meridian).
 *  - Realizations are the batch axis.  In HFT_MODE_COUPLED the time-step
 *    controls (Euler limit, PTL sub-cycles, RKG2 stage count, flow CFL) are
 *    shared by all realizations exactly as in the reference
 *    (src/hipft.f90:5019-5037, 5094-5143, 7321-7346).  In
 *    HFT_MODE_INDEPENDENT every realization runs its own controller and gets
 *    the result a single-realization reference run would give it.
 *  - One context drives one GPU (the current HIP device at hft_create) from
 *    one host thread.  Calls are stream-ordered; functions that return host
 *    data synchronise.
 */
#ifndef HFT_H
#define HFT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HFT_OK 0
#define HFT_E_ARG -1      /* invalid argument */
#define HFT_E_HIP -2      /* HIP runtime error */
#define HFT_E_STATE -3    /* call out of order (e.g. diffusion not loaded) */
#define HFT_E_NOMEM -4    /* device allocation failed */

#define HFT_LAYOUT_TP 0
#define HFT_LAYOUT_PT 1

#define HFT_MODE_COUPLED 0
#define HFT_MODE_INDEPENDENT 1

typedef struct hft_ctx hft_ctx;

/* Number of history diagnostics per realization (analysis_step columns
 * BR_MIN .. VALIDATION_ERR_HHabs, src/hipft.f90:1630-1646). */
#define HFT_NHIST 14

/* Per-cycle record of a diffusion step (src/hipft.f90:4607-4683). */
typedef struct hft_cycle {
  double time_stepped; /* part of the step covered after this cycle */
  double dt;           /* cycle time step */
  int64_t sts_s;       /* RKG2 stage count of this cycle */
  int64_t subcycles;   /* diagnostic sub-cycle estimate printed in histories */
} hft_cycle;

/* Called after each round of diffusion sub-cycles with rec[nr]: the record
 * of every realization that took a cycle in this round (rec[r].dt == 0 for a
 * realization that had already finished its step).  In coupled mode all
 * realizations cycle together.  Used to write the extra history rows the
 * reference writes at verbose>1 (src/hipft.f90:4659-4669). */
typedef void (*hft_cycle_fn)(void *user, int cycle, const hft_cycle *rec);

/* ---- lifetime ----------------------------------------------------------
 * Replaces load_initial_condition's mesh set-up and set_mesh
 * (src/hipft.f90:2104-2136, 6495-6653).  t[ntm] are the theta nodes on
 * [0,pi] and p[npm1] the phi nodes on [0,2pi] (file scales, one-point
 * overlap).  mode: HFT_MODE_*. */
int hft_create(int ntm, int npm1, const double *t, const double *p, int nr, int mode,
               hft_ctx **out);
int hft_destroy(hft_ctx *ctx);
const char *hft_last_error(void);
int hft_dims(const hft_ctx *ctx, int *ntm, int *npm, int *nr);

/* ---- state ---------------------------------------------------------------
 * HFT_LAYOUT_TP: f(ntm,npm,nr) as the reference holds it.
 * HFT_LAYOUT_PT on set: the initial-map ingest of src/hipft.f90:2116-2157,
 * 2171-2173 (transpose, seam average, pole rows set to their phi average,
 * times mult_fac).  On get: PT map (npm1,ntm,nr) with the one-point overlap. */
int hft_set_state(hft_ctx *ctx, const double *f, int layout, double mult_fac);
int hft_get_state(hft_ctx *ctx, double *f, int layout);

/* ---- flows ---------------------------------------------------------------
 * update_flow, src/hipft.f90:3371-3523, for time-independent flows.
 * p0,p2,p4: differential-rotation coefficients (m/s) per realization
 * (add_flow_differential_rotation_analytic 7219-7251), NULL = none.
 * p1,p3,p5: meridional-flow coefficients (m/s) per realization
 * (add_flow_meridianal_analytic 7253-7285), NULL = none.
 * vp_rigid_omega (km/s) and vt_const (km/s): src/hipft.f90:3436-3448.
 * Also sets the Lax-Friedrichs alphas (set_lf_alpha 6940-7048). */
int hft_set_flow_analytic(hft_ctx *ctx, const double *p0, const double *p2, const double *p4,
                          const double *p1, const double *p3, const double *p5,
                          double vp_rigid_omega, double vt_const);
/* Base flow given as theta profiles (phi-independent), in code units
 * (R_sun/hr): vt[nr][nt] on the theta half mesh, vp[nr][ntm] on the main
 * theta mesh.  This is what flows from files reduce to when they do not
 * depend on phi (add_flow_from_files, src/hipft.f90:3525-3673); as there,
 * the analytic terms of hft_set_flow_analytic are added on top of it
 * (src/hipft.f90:3421-3448).  NULL vt or vp = zero. */
int hft_set_flow_profiles(hft_ctx *ctx, const double *vt, const double *vp);
/* Flows from files (add_flow_from_files, src/hipft.f90:3525-3660; the
 * driver reads the files and converts m/s to code units as 3596-3598,
 * 3625-3627 do): 2-D base flows in the reference's storage,
 * vt(nt,npm,nrf) and vp(ntm,np,nrf) with np = npm, theta fastest,
 * R_sun/hr; nrf = nr when per_realization != 0, else 1 (one flow for every
 * realization, as the reference's files are).  Only the unique phi columns
 * 2..npm-1 are read; the seam columns are periodic.  The analytic terms of
 * hft_set_flow_analytic are added in update_flow's order (3421-3448) and
 * attenuation applies on top.  Replaces profiles set by
 * hft_set_flow_profiles; vt = vp = NULL removes the base flows. */
int hft_set_flow_arrays(hft_ctx *ctx, const double *vt, const double *vp, int per_realization);
/* Flow attenuation by |Br| (src/hipft.f90:3489-3515): value[nr] in Gauss
 * (<= 0 disables for that realization).  When enabled, the flow becomes a
 * per-step 2-D field recomputed from the state at every hft_update_flow. */
int hft_set_flow_attenuation(hft_ctx *ctx, const double *value);
/* Recompute attenuated flows from the current state (update_flow for
 * flow_attenuate=.true.); a no-op for static flows. */
int hft_update_flow(hft_ctx *ctx);
/* get_flow_dtmax (src/hipft.f90:7287-7350, SSPRK(4,3)): per-realization
 * limits in dt_r[nr] (may be NULL); *dt = min over realizations. */
int hft_flow_dtmax(hft_ctx *ctx, int strang, double *dt, double *dt_r);

/* weno_eps (load_weno, src/hipft.f90:6895-6901): the smoothness-weight
 * epsilon of WENO3; eps <= 0 (the reference's default 0) uses the cell
 * widths dt(j) in theta and dp(k) in phi. */
int hft_set_weno_eps(hft_ctx *ctx, double eps);

/* ---- diffusion -----------------------------------------------------------
 * load_diffusion with constant per-realization coefficient eta[nr] (km^2/s),
 * factor = diffusion_coef_factor (7.43162901307967e-09 converts km^2/s to
 * R_sun^2/hr), src/hipft.f90:6692-6702, 6821-6843, 4927-4983.  Computes the
 * explicit Euler limit (get_dtime_diffusion_euler 4985-5040). */
int hft_set_diffusion(hft_ctx *ctx, const double *eta, double factor);
/* Euler limit: *dt = coupled (min over realizations), dt_r[nr] per
 * realization (may be NULL). */
int hft_dt_euler(hft_ctx *ctx, double *dt, double *dt_r);

/* ---- time step ------------------------------------------------------------
 * flux_transport_step, src/hipft.f90:4463-4507: Strang A(dt/2) D(dt) A(dt/2)
 * when strang != 0, else A(dt) D(dt); with a source term set (below),
 * S(dt/2) A(dt/2) D(dt) A(dt/2) S(dt/2), else A(dt) D(dt) S(dt).
 * subcycles = diffusion_subcycles (0 = automatic PTL sub-cycling).  cb/user
 * may be NULL. */
int hft_flux_transport_step(hft_ctx *ctx, double dt, int strang, int subcycles,
                            hft_cycle_fn cb, void *user);
/* advection_step (SSPRK(4,3) + WENO3, src/hipft.f90:4509-4549, 7515-7582).
 * Independent mode: a realization whose flow CFL limit (hft_flow_dtmax with
 * strang = 0) is below dt takes n_r = ceil(dt / limit_r) SSPRK(4,3) steps of
 * dt / n_r, the others one step of dt (tuning key "advection_substeps"). */
int hft_advection_step(hft_ctx *ctx, double dt);
/* diffusion_step (RKG2 with PTL sub-cycles, src/hipft.f90:4551-4695). */
int hft_diffusion_step(hft_ctx *ctx, double dt, int subcycles, hft_cycle_fn cb, void *user);
/* Last diffusion step summary: per realization stage count of the last
 * cycle (s[nr]) and number of cycles (cycles[nr]); either may be NULL. */
int hft_diffusion_stats(hft_ctx *ctx, int64_t *s, int64_t *cycles);

/* ---- sources ---------------------------------------------------------------
 * The source term of update_source (src/hipft.f90:3875-3947):
 *   source = ((0 + map) + (1 - alpha) rfe[first]) + alpha rfe[1 - first]
 * with absent terms left out; source_step (4697-4733) is f += dt * source.
 * hft_set_source: a constant map (source_from_file after its interpolation
 * onto the mesh, 7074-7124), reference storage (ntm, npm), theta fastest,
 * G/hr; shared by all realizations; NULL removes it. */
int hft_set_source(hft_ctx *ctx, const double *src);
/* Random flux emergence frame into slot 0 or 1 (generate_rfe 8809-8921 and
 * the balance_flux of load_source / update_source): rnd is the reference's
 * random map rnd_nums(1:ntm, 2:npm-1) (already times pi/sqrt(2 pi)),
 * theta fastest; model 1 scales by the unsigned flux per cell, model 2 per
 * area, uf_per_hour[nr] = source_rfe_total_unsigned_flux_per_hour(s) in
 * 1e21 Mx/hr (7167-7194); pole rows are phi averages; each realization's
 * frame is flux balanced. */
int hft_rfe_generate(hft_ctx *ctx, int slot, const double *rnd, int model,
                     const double *uf_per_hour);
/* Blend of the two frames in the source term: (1 - alpha) rfe[first_slot] +
 * alpha rfe[1 - first_slot]; first_slot = -1 leaves the frames out. */
int hft_rfe_weights(hft_ctx *ctx, int first_slot, double alpha);
/* Frame of a slot, TP layout (ntm, npm, nr). */
int hft_get_rfe(hft_ctx *ctx, int slot, double *out_tp);
/* source_step: f += dt * source. */
int hft_source_step(hft_ctx *ctx, double dt);

/* ---- operators (for tests and adjoint checks) -----------------------------
 * y = A x for the whole batch, TP layout, x and y host arrays. */
int hft_apply_diffusion_operator(hft_ctx *ctx, const double *x, double *y);
int hft_apply_advection_operator(hft_ctx *ctx, const double *x, double *y);

/* balance_flux (src/hipft.f90:8752-8801) of every realization: with
 * out_pt == NULL the state itself (initial_map_flux_balance, 2163-2167);
 * otherwise a balanced copy of the state is written to out_pt as a PT map
 * (npm1,ntm,nr) and the state is left as it is (output_map_flux_balance,
 * 2839-2843). */
int hft_balance_flux(hft_ctx *ctx, double *out_pt);

/* ---- data assimilation -----------------------------------------------------
 * apply_data_assimilation + assimilate_new_data, src/hipft.f90:4148-4306,
 * 4083-4146.  map: (npm1,ntm,3) PT slices data/weight/mu, one map shared by
 * all realizations (per_realization=0) or (npm1,ntm,3,nr).  Per-realization
 * mu_power, mu_limit, lat_limit (degrees); custom_from_mu builds the weight
 * slice from mu (4242-4252); add_mode is assimilate_data_add; balance is
 * assimilate_data_balance_flux. */
int hft_assimilate(hft_ctx *ctx, const double *map, int per_realization, int custom_from_mu,
                   const double *mu_power, const double *mu_limit, const double *lat_limit,
                   double mult_fac, int add_mode, int balance);
/* Same, with the map already in device memory (PT slices, same shapes). */
int hft_assimilate_device(hft_ctx *ctx, const double *map_dev, int per_realization,
                          int custom_from_mu, const double *mu_power, const double *mu_limit,
                          const double *lat_limit, double mult_fac, int add_mode, int balance);

/* ---- analysis ---------------------------------------------------------------
 * analysis_step, src/hipft.f90:3145-3325: out[nr][HFT_NHIST].
 * Validation field: fval_pt (npm1,ntm,nr) host array, NULL = none (error 0);
 * if fval_scale[nr] is non-NULL the field set by hft_set_validation_field is
 * used, multiplied per realization by fval_scale[r]. */
int hft_set_validation_field(hft_ctx *ctx, const double *fval_pt);
/* pole_flux_lat_limit (degrees, src/hipft.f90:566; default 30): the polar
 * caps of the analysis' polar flux and area columns (3276-3285). */
int hft_set_pole_flux_lat_limit(hft_ctx *ctx, double deg);
int hft_analysis(hft_ctx *ctx, const double *fval_pt, const double *fval_scale, double *out);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops): the history
 * diagnostics of hft_analysis without a validation field (its
 * VALIDATION_ERR column is 0), computed on the context's stream without
 * any allocation, copy or synchronisation.  f_dev: a device batch laid out
 * [r][j][c] with row length ld >= npm-2 (unique phi columns), or NULL for
 * the context's state; out_dev: device array [nr][HFT_NHIST].  The block
 * partial sums live in a fixed 8 MiB device array shared by every context
 * on a GPU (nr x row blocks x 16 doubles, at most 2^20; larger batches get
 * HFT_E_ARG), so calls that overlap on one GPU (other streams or host
 * threads) race; calls on one stream are ordered. */
int hft_analysis_noalloc(hft_ctx *ctx, const double *f_dev, int64_t ld, double *out_dev);

/* ---- device access (Python/PyTorch interop, benchmarks) -------------------
 * Device pointer of the state and its layout: f_dev[(r*ntm + j)*pitch + c],
 * c = 0..npm-3 the unique phi columns (c <-> reference column c+2). */
int hft_state_device(hft_ctx *ctx, double **f_dev, int *pitch);
int hft_synchronize(hft_ctx *ctx);
/* HIP stream the context launches on (hipStream_t as void*). */
int hft_stream(hft_ctx *ctx, void **stream);
/* Launch tuning of one context (chosen per grid and batch by
 * hipft_amd/tune.py; the defaults suit 512x1024 batches):
 *   "fused"                  1 = run RKG2 stages several per launch (default)
 *   "fused_stages"           stages per fused launch, 2..4 (default 4); 5 with
 *                            fused_kernel 3
 *   "band_rows"              theta rows per block of a fused launch (32; 2..4096)
 *   "fused_min_realizations" fused launches only while at least this many
 *                            realizations are still stepping (8)
 *   "fused_target_blocks"    when a fused launch would have fewer blocks
 *                            (realizations x bands) than this, its bands are
 *                            made thinner (>= 2 rows) to fill the GPU (0: off)
 *   "fused_kernel"           fused launch kernel: 1 = one phi column per
 *                            thread, 2 = the same with lean bookkeeping,
 *                            3 = two phi columns per thread (default)
 *   "fused_bands"            1 = each fused launch of kernel 3 picks its band
 *                            count (bands of fused_band_min_rows..(160 - 2 x
 *                            stages) rows) so
 *                            that its blocks fill whole rounds of the blocks
 *                            the device holds at once; band_rows and
 *                            fused_target_blocks are then unused (0)
 *   "fused_band_share"       that choice's model of how much longer a round
 *                            takes on a full device than on an idle one, in
 *                            percent (32)
 *   "fused_band_min_rows"    the thinnest band that choice may pick (8; 2..64):
 *                            a launch over few realizations is bound by the
 *                            latency of its march, so thinner bands finish
 *                            sooner
 *   "advection_band"         1 = WENO3 stages by theta bands computing each
 *                            face once (default); 0 = one thread per point
 *   "advection_band_rows"    theta rows per block of the banded WENO3 (16)
 *   "advection_substeps"     1 = per-realization CFL sub-steps of the
 *                            advection step in independent mode (default);
 *                            0 = one step of dt for every realization
 *   "graphs"                 1 = capture each RKG2 super-step's launches in a
 *                            hipGraph and replay it when the same
 *                            realizations have the same stage counts (0);
 *                            "graph_hits", "graph_builds" (read only) count
 *                            replays and captures
 *   "step_graph"             1 = plan hft_flux_transport_step on the device
 *                            (PTL step rule, stage counts and coefficient
 *                            tables of diffusion_step / load_sts_rkg2,
 *                            src/hipft.f90:4589-4690, 5399-5461) and replay the
 *                            whole step as one captured hipGraph (default);
 *                            not used with a cycle callback, timing, sources,
 *                            per-realization CFL sub-steps, the null stream
 *                            or more than 1024 realizations
 *   "step_graph_rounds"      sub-cycle rounds the captured step holds (1); a
 *                            step needing more is finished by the host
 *                            controller; 0 = learned from earlier steps.
 *                            "step_graph_builds", "step_graph_hits",
 *                            "step_graph_overflows" (read only) count
 *                            captures, replays and steps finished on the host
 *   "adjoint_segment"       RKG2 stages per recomputed segment of the
 *                            adjoint (0 = all stages of a cycle when they fit
 *                            in "adjoint_mem_mb", else sqrt(s))
 *   "adjoint_mem_mb"         device memory for stored RKG2 stages (32768)
 *   "adjoint_band_rows"      theta rows per block of the WENO3 adjoint (16)
 *   "adjoint_fused"          1 = when a cycle's RKG2 stages are stored, its
 *                            adjoint takes 4 transposed stages per launch
 *                            (default); 0 = one launch per stage
 *   "adjoint_fused_rows"     theta rows per block of those launches (16)
 * Forward results do not depend on these settings (bit-identical); adjoint
 * results agree to rounding (the adjoint kernels contract multiply-adds and
 * sum in launch-dependent order). */
int hft_set_tuning(hft_ctx *ctx, const char *key, int value);
int hft_get_tuning(hft_ctx *ctx, const char *key, int *value);

/* Kernel timing with HIP events on the context's stream.  hft_set_timing
 * resets the counters; enable is a bit set: 1 = events around every launch
 * (this also turns off hipGraph replay of RKG2 super-steps), 2 = events at the
 * boundaries of the step's phases only, 0 = off.
 * hft_kernel_time returns, for a kernel group ("rkg2", "weno", "ptl",
 * "analysis", "assim", "flow", "adjoint", or "rkg2_fused" = the multi-stage
 * RKG2 launches, which "rkg2" includes) since the reset: total device time (ms), number of
 * launches, and the algorithmic bytes those launches move (compulsory
 * reads + writes of FP64 fields over the ntm x (npm-2) unique points; a
 * multi-stage launch counts what it must move: its input levels, u0 and
 * its two output levels, not its intermediate stages).
 * With phase events on, the groups "phase_advection", "phase_diffusion",
 * "phase_source" and "phase_flux_transport" hold the stream time of every
 * hft_advection_step, hft_diffusion_step, hft_source_step and
 * hft_flux_transport_step call (the reference's wtime_flux_transport_advection,
 * wtime_flux_transport_diffusion, wtime_source and wtime_flux_transport,
 * src/hipft.f90:4505, 4546, 4692, 4730); their alg_bytes are 0. */
int hft_set_timing(hft_ctx *ctx, int enable);
int hft_kernel_time(hft_ctx *ctx, const char *group, double *total_ms, int64_t *launches,
                    double *alg_bytes);
/* The bytes the same group's work would move one RKG2 stage per launch (5
 * field passes per stage and realization): equal to alg_bytes except for the
 * multi-stage "rkg2_fused" launches (and "rkg2", which includes them), whose
 * intermediate stages stay on chip.  stage_bytes / alg_bytes is the traffic
 * that temporal blocking saves.  No counterpart in the reference (its
 * wtime_* timers, src/hipft.f90:478-486, count time only). */
int hft_kernel_stage_bytes(hft_ctx *ctx, const char *group, double *stage_bytes);

/* ---- differentiation (no counterpart in the reference) ---------------------
 * Device-resident state for callers that keep maps in HBM (PyTorch): src/dst
 * are device arrays [r][j][c] with row length ld >= npm-2 (unique phi
 * columns only), copied on the context's stream. */
int hft_set_state_device(hft_ctx *ctx, const double *src, int64_t ld);
int hft_get_state_device(hft_ctx *ctx, double *dst, int64_t ld);
/* own = 0: launch on the caller's HIP stream (hipStream_t as void*; NULL is
 * the null stream); own = 1: on the context's own stream again. */
int hft_set_stream(hft_ctx *ctx, void *stream, int own);

/* Discrete adjoint of one hft_flux_transport_step(ctx, dt, strang,
 * subcycles, ...) taken from the current state x_n.
 *   lam  (device, layout as hft_set_state_device) in: dJ/dx_{n+1};
 *        out: dJ/dx_n
 *   grad (host [nr][3], accumulated +=): dJ/dlog of the per-realization
 *        diffusivity (eta), of a common scale of the theta velocity vt and
 *        of a common scale of the phi velocity vp (alpha scales with them).
 * The step's discrete choices (sub-cycle time steps, RKG2 stage counts,
 * upwind signs) are those of the recomputed forward step and are held
 * fixed.  On return the state is x_{n+1}.  Flows must not be attenuated.
 * flags: HFT_ADJ_FROZEN_WENO linearises WENO3 with frozen nonlinear weights
 * (otherwise the exact Jacobian of the weights is used). */
#define HFT_ADJ_FROZEN_WENO 1
int hft_step_adjoint(hft_ctx *ctx, double dt, int strang, int subcycles, double *lam,
                     int64_t ld, double *grad, int flags);
/* Adjoint of nsteps consecutive steps from the current state x_n: lam goes
 * from dJ/dx_{n+nsteps} to dJ/dx_n, grad accumulates as above, and the state
 * is x_{n+nsteps} on return.  With fixed sub-cycles and everything within
 * "adjoint_mem_mb", one forward pass keeps every step's stage data and the
 * reverse sweep follows (no recomputation); otherwise the step inputs are
 * kept and each step is recomputed before its adjoint. */
int hft_steps_adjoint(hft_ctx *ctx, int nsteps, double dt, int strang, int subcycles,
                      double *lam, int64_t ld, double *grad, int flags);

#ifdef __cplusplus
}
#endif
#endif
