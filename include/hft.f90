!
! hft.f90 -- iso_c_binding interface to libhft (include/hft.h).
!
! A Fortran time loop (the reference's program HIPFT, src/hipft.f90:903-948)
! uses this module to drive the MI355X flux-transport core: the calls
! replace update_flow / update_timestep helpers, flux_transport_step,
! apply_data_assimilation and analysis_step.  Arrays are passed by reference
! in the reference layout f(ntm,npm,nr); scalars by value.  Every function
! returns 0 on success; hft_last_error() gives the message of a failure.
!
module hft
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: HFT_LAYOUT_TP = 0, HFT_LAYOUT_PT = 1
  integer(c_int), parameter :: HFT_MODE_COUPLED = 0, HFT_MODE_INDEPENDENT = 1
  integer(c_int), parameter :: HFT_NHIST = 14
  integer(c_int), parameter :: HFT_ADJ_FROZEN_WENO = 1

  type, bind(c) :: hft_cycle
    real(c_double) :: time_stepped, dt
    integer(c_int64_t) :: sts_s, subcycles
  end type

  interface
    integer(c_int) function hft_create(ntm, npm1, t, p, nr, mode, ctx) bind(c, name="hft_create")
      import :: c_int, c_double, c_ptr
      integer(c_int), value :: ntm, npm1, nr, mode
      real(c_double), intent(in) :: t(*), p(*)
      type(c_ptr), intent(out) :: ctx
    end function
    integer(c_int) function hft_destroy(ctx) bind(c, name="hft_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    type(c_ptr) function hft_last_error() bind(c, name="hft_last_error")
      import :: c_ptr
    end function
    integer(c_int) function hft_dims(ctx, ntm, npm, nr) bind(c, name="hft_dims")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), intent(out) :: ntm, npm, nr
    end function
    integer(c_int) function hft_set_state(ctx, f, layout, mult_fac) bind(c, name="hft_set_state")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: f(*)
      integer(c_int), value :: layout
      real(c_double), value :: mult_fac
    end function
    integer(c_int) function hft_get_state(ctx, f, layout) bind(c, name="hft_get_state")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: f(*)
      integer(c_int), value :: layout
    end function
    integer(c_int) function hft_set_flow_analytic(ctx, p0, p2, p4, p1, p3, p5, omega, vtc) &
        bind(c, name="hft_set_flow_analytic")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      type(c_ptr), value :: p0, p2, p4, p1, p3, p5   ! c_loc(array) or c_null_ptr
      real(c_double), value :: omega, vtc
    end function
    integer(c_int) function hft_set_flow_profiles(ctx, vt, vp) bind(c, name="hft_set_flow_profiles")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, vt, vp
    end function
    integer(c_int) function hft_set_flow_arrays(ctx, vt, vp, per_realization) &
        bind(c, name="hft_set_flow_arrays")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, vt, vp
      integer(c_int), value :: per_realization
    end function
    integer(c_int) function hft_set_source(ctx, src) bind(c, name="hft_set_source")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, src
    end function
    integer(c_int) function hft_rfe_generate(ctx, slot, rnd, model, uf_per_hour) &
        bind(c, name="hft_rfe_generate")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: slot, model
      real(c_double), intent(in) :: rnd(*), uf_per_hour(*)
    end function
    integer(c_int) function hft_rfe_weights(ctx, first_slot, alpha) bind(c, name="hft_rfe_weights")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: first_slot
      real(c_double), value :: alpha
    end function
    integer(c_int) function hft_get_rfe(ctx, slot, out) bind(c, name="hft_get_rfe")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: slot
      real(c_double), intent(out) :: out(*)
    end function
    integer(c_int) function hft_source_step(ctx, dt) bind(c, name="hft_source_step")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: dt
    end function
    integer(c_int) function hft_set_flow_attenuation(ctx, value) bind(c, name="hft_set_flow_attenuation")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, value
    end function
    integer(c_int) function hft_update_flow(ctx) bind(c, name="hft_update_flow")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function hft_flow_dtmax(ctx, strang, dt, dt_r) bind(c, name="hft_flow_dtmax")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: strang
      real(c_double), intent(out) :: dt
      type(c_ptr), value :: dt_r
    end function
    integer(c_int) function hft_set_diffusion(ctx, eta, factor) bind(c, name="hft_set_diffusion")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: eta(*)
      real(c_double), value :: factor
    end function
    integer(c_int) function hft_dt_euler(ctx, dt, dt_r) bind(c, name="hft_dt_euler")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: dt
      type(c_ptr), value :: dt_r
    end function
    integer(c_int) function hft_flux_transport_step(ctx, dt, strang, subcycles, cb, user) &
        bind(c, name="hft_flux_transport_step")
      import :: c_int, c_double, c_ptr, c_funptr
      type(c_ptr), value :: ctx
      real(c_double), value :: dt
      integer(c_int), value :: strang, subcycles
      type(c_funptr), value :: cb
      type(c_ptr), value :: user
    end function
    integer(c_int) function hft_advection_step(ctx, dt) bind(c, name="hft_advection_step")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double), value :: dt
    end function
    integer(c_int) function hft_diffusion_step(ctx, dt, subcycles, cb, user) &
        bind(c, name="hft_diffusion_step")
      import :: c_int, c_double, c_ptr, c_funptr
      type(c_ptr), value :: ctx
      real(c_double), value :: dt
      integer(c_int), value :: subcycles
      type(c_funptr), value :: cb
      type(c_ptr), value :: user
    end function
    integer(c_int) function hft_diffusion_stats(ctx, s, cycles) bind(c, name="hft_diffusion_stats")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, s, cycles
    end function
    integer(c_int) function hft_apply_diffusion_operator(ctx, x, y) &
        bind(c, name="hft_apply_diffusion_operator")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: x(*)
      real(c_double), intent(out) :: y(*)
    end function
    integer(c_int) function hft_apply_advection_operator(ctx, x, y) &
        bind(c, name="hft_apply_advection_operator")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: x(*)
      real(c_double), intent(out) :: y(*)
    end function
    integer(c_int) function hft_assimilate(ctx, map, per_real, custom, mu_power, mu_limit, &
        lat_limit, mult, add_mode, balance) bind(c, name="hft_assimilate")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: map(*)
      integer(c_int), value :: per_real, custom, add_mode, balance
      type(c_ptr), value :: mu_power, mu_limit, lat_limit
      real(c_double), value :: mult
    end function
    integer(c_int) function hft_assimilate_device(ctx, map_dev, per_real, custom, mu_power, &
        mu_limit, lat_limit, mult, add_mode, balance) bind(c, name="hft_assimilate_device")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx, map_dev
      integer(c_int), value :: per_real, custom, add_mode, balance
      type(c_ptr), value :: mu_power, mu_limit, lat_limit
      real(c_double), value :: mult
    end function
    integer(c_int) function hft_set_validation_field(ctx, fval) bind(c, name="hft_set_validation_field")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, fval
    end function
    integer(c_int) function hft_set_weno_eps(ctx, eps) bind(c, name="hft_set_weno_eps")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: eps
    end function
    integer(c_int) function hft_set_pole_flux_lat_limit(ctx, deg) &
        bind(c, name="hft_set_pole_flux_lat_limit")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: deg
    end function
    ! out_pt = c_null_ptr balances the state in place
    integer(c_int) function hft_balance_flux(ctx, out_pt) bind(c, name="hft_balance_flux")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, out_pt
    end function
    integer(c_int) function hft_analysis(ctx, fval, fval_scale, out) bind(c, name="hft_analysis")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx, fval, fval_scale
      real(c_double), intent(out) :: out(*)
    end function
    ! no counterpart in the reference: hft_analysis for callers that cannot
    ! allocate (device input and output, partial sums in a fixed device array)
    integer(c_int) function hft_analysis_noalloc(ctx, f_dev, ld, out_dev) &
        bind(c, name="hft_analysis_noalloc")
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: ctx, f_dev, out_dev
      integer(c_int64_t), value :: ld
    end function
    integer(c_int) function hft_state_device(ctx, f_dev, pitch) bind(c, name="hft_state_device")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      type(c_ptr), intent(out) :: f_dev
      integer(c_int), intent(out) :: pitch
    end function
    integer(c_int) function hft_synchronize(ctx) bind(c, name="hft_synchronize")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function hft_stream(ctx, stream) bind(c, name="hft_stream")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      type(c_ptr), intent(out) :: stream
    end function
    integer(c_int) function hft_set_tuning(ctx, key, value) bind(c, name="hft_set_tuning")
      import :: c_int, c_ptr, c_char
      type(c_ptr), value :: ctx
      character(kind=c_char), intent(in) :: key(*)
      integer(c_int), value :: value
    end function
    integer(c_int) function hft_get_tuning(ctx, key, value) bind(c, name="hft_get_tuning")
      import :: c_int, c_ptr, c_char
      type(c_ptr), value :: ctx
      character(kind=c_char), intent(in) :: key(*)
      integer(c_int), intent(out) :: value
    end function
    integer(c_int) function hft_set_timing(ctx, enable) bind(c, name="hft_set_timing")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: enable
    end function
    integer(c_int) function hft_kernel_time(ctx, group, total_ms, launches, alg_bytes) &
        bind(c, name="hft_kernel_time")
      import :: c_int, c_double, c_ptr, c_char, c_int64_t
      type(c_ptr), value :: ctx
      character(kind=c_char), intent(in) :: group(*)
      real(c_double), intent(out) :: total_ms, alg_bytes
      integer(c_int64_t), intent(out) :: launches
    end function
    integer(c_int) function hft_kernel_stage_bytes(ctx, group, stage_bytes) &
        bind(c, name="hft_kernel_stage_bytes")
      import :: c_int, c_double, c_ptr, c_char
      type(c_ptr), value :: ctx
      character(kind=c_char), intent(in) :: group(*)
      real(c_double), intent(out) :: stage_bytes
    end function
    ! device-resident state and the discrete adjoint of one step
    integer(c_int) function hft_set_state_device(ctx, src, ld) bind(c, name="hft_set_state_device")
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: ctx, src
      integer(c_int64_t), value :: ld
    end function
    integer(c_int) function hft_get_state_device(ctx, dst, ld) bind(c, name="hft_get_state_device")
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: ctx, dst
      integer(c_int64_t), value :: ld
    end function
    integer(c_int) function hft_set_stream(ctx, stream, own) bind(c, name="hft_set_stream")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, stream
      integer(c_int), value :: own
    end function
    integer(c_int) function hft_step_adjoint(ctx, dt, strang, subcycles, lam, ld, grad, flags) &
        bind(c, name="hft_step_adjoint")
      import :: c_int, c_double, c_ptr, c_int64_t
      type(c_ptr), value :: ctx, lam
      real(c_double), value :: dt
      integer(c_int), value :: strang, subcycles, flags
      integer(c_int64_t), value :: ld
      real(c_double), intent(inout) :: grad(*)
    end function
    integer(c_int) function hft_steps_adjoint(ctx, nsteps, dt, strang, subcycles, lam, ld, grad, &
        flags) bind(c, name="hft_steps_adjoint")
      import :: c_int, c_double, c_ptr, c_int64_t
      type(c_ptr), value :: ctx, lam
      integer(c_int), value :: nsteps, strang, subcycles, flags
      real(c_double), value :: dt
      integer(c_int64_t), value :: ld
      real(c_double), intent(inout) :: grad(*)
    end function
  end interface
end module hft
