!
! hft_hipft.f90 -- the reference's executable (program HIPFT, src/hipft.f90)
! as a Fortran driver over libhft: it reads the &hipft_input_parameters
! namelist of hipft.in (read_input_file, src/hipft.f90:8307-8456; every item
! of the group is accepted, with the reference's defaults, 493-726), fills
! the per-realization arrays from their scalars (process_input_parameters,
! 8542-8579), runs the time loop of 903-948 through the C ABI (include/hft.h,
! module hft) and writes hipft_history_sol_r######.out and
! hipft_history_num_r######.out in the reference's layout (headers
! 1616-1646, rows 2633-2634), with the extra rows of every diffusion
! sub-cycle at verbose > 1 (4659-4669).
!
! Supported: grids from res_nt/res_np with the validation fields
! (validation_run = 1 soccer ball, 2 Gaussian blobs; 1892-2101), analytic
! flows (differential rotation, meridional flow, rigid rotation, constant vt;
! 3371-3523, 7219-7285) with attenuation (3489-3515), constant per-realization
! diffusion with PTL or fixed sub-cycles, Strang or unsplit steps, the
! history cadence, weno_eps and pole_flux_lat_limit.  Items that need map
! files or features this driver does not read (initial_map_filename, flows
! from files, data assimilation, sources, map output) stop the run with a
! message naming them; the Python driver (hipft_amd/driver.py) covers those.
!
! Usage:  hft_hipft [hipft.in]
!
module hft_hipft_state
  use iso_c_binding
  implicit none
  integer, parameter :: r8 = c_double
  integer, parameter :: max_realizations = 2000  ! src/hipft.f90:236
  real(r8), parameter :: pi = 3.1415926535897932_r8, twopi = 6.2831853071795864_r8
  real(r8), parameter :: pi_two = 1.5707963267948966_r8, pi_four = 0.7853981633974483_r8
  real(r8), parameter :: threepi_two = 4.71238898038469_r8
  real(r8), parameter :: km_s_to_rs_hr = 0.005172413793103448_r8
  real(r8), parameter :: diff_km2_s_to_rs2_hr = 7.43162901307967e-09_r8

  ! run state shared with the diffusion sub-cycle callback
  type(c_ptr) :: ctx = c_null_ptr
  integer :: nr = 1, ntm = 0, npm1 = 0, verbose = 0, ntime = 0
  real(r8) :: time_now = 0, dtime = 0, dtmax_flow = 0, dt_adv_used = 0, dt_euler = 0
  real(r8) :: step_start = 0
  integer(c_int64_t), allocatable, target :: sts_s(:), subcyc(:)
  ! history rows [16 or 8, rows, nr]
  real(r8), allocatable :: sol(:,:,:), num(:,:,:)
  integer :: nrows = 0
  ! validation
  integer :: vkind = 0
  real(r8) :: vwidth = 0.03_r8, vt_const = 0, vp_omega = 0, time_start = 0, dfac = 0
  real(r8), allocatable, target :: eta(:), vscale(:), tmesh(:), pmesh(:), sti(:), fval(:)
  real(r8) :: hist_cadence = 0
  integer :: idx_hist = 0

contains

  subroutine check(rc, what)
    integer(c_int), intent(in) :: rc
    character(*), intent(in) :: what
    character(kind=c_char), pointer :: msg(:)
    integer :: i
    if (rc == 0) return
    call c_f_pointer(hft_last_error_ptr(), msg, [512])
    write (*,'(a,a,i0,a)', advance='no') what, ' failed (', rc, '): '
    do i = 1, 512
      if (msg(i) == c_null_char) exit
      write (*,'(a)', advance='no') msg(i)
    end do
    write (*,*)
    stop 1
  end subroutine

  type(c_ptr) function hft_last_error_ptr()
    use hft, only: hft_last_error
    hft_last_error_ptr = hft_last_error()
  end function

  ! exact validation_run = 2 field at time_now on the run mesh (PT layout)
  ! the field of every realization: out holds npm1*ntm*nr values, the slices alike
  subroutine gauss_field(elapsed, out)
    real(r8), intent(in) :: elapsed
    real(r8), intent(out) :: out(:)
    real(r8) :: vtt, vpt, p1, p2, tt, pp
    integer :: j, l, r, n
    vtt = vt_const*km_s_to_rs_hr*elapsed
    vpt = vp_omega*km_s_to_rs_hr*elapsed
    p1 = (pi_two + vpt) - floor((pi_two + vpt)/twopi)*twopi
    p2 = (threepi_two + vpt) - floor((threepi_two + vpt)/twopi)*twopi
    do j = 1, ntm
      tt = tmesh(j)
      do l = 1, npm1
        pp = pmesh(l)
        out(l + npm1*(j - 1)) = -sti(j)*exp(-(tt - pi_four - vtt)**2/vwidth - (pp - p1)**2/vwidth) &
                                + sti(j)*exp(-(tt - pi_four - vtt)**2/vwidth - (pp - p2)**2/vwidth)
      end do
    end do
    n = npm1*ntm
    do r = 2, size(out)/n
      out((r - 1)*n + 1:r*n) = out(1:n)
    end do
  end subroutine

  ! analysis_step + one history row of every realization (output_histories)
  subroutine emit(t_row)
    use hft
    real(r8), intent(in) :: t_row
    real(r8), target :: h(HFT_NHIST, nr)
    integer :: r
    if (hist_cadence > 0) then
      if (t_row < time_start + idx_hist*hist_cadence) return
    end if
    idx_hist = idx_hist + 1
    select case (vkind)
    case (1)
      do r = 1, nr
        vscale(r) = exp(-42.0_r8*eta(r)*dfac*(t_row - time_start))
      end do
      call check(hft_analysis(ctx, c_null_ptr, c_loc(vscale), h), 'hft_analysis')
    case (2)
      call gauss_field(t_row - time_start, fval)
      call check(hft_analysis(ctx, c_loc(fval), c_null_ptr, h), 'hft_analysis')
    case default
      call check(hft_analysis(ctx, c_null_ptr, c_null_ptr, h), 'hft_analysis')
    end select
    nrows = nrows + 1
    if (nrows > size(sol, 2)) call grow()
    do r = 1, nr
      sol(1, nrows, r) = real(ntime, r8)
      sol(2, nrows, r) = t_row
      sol(3:, nrows, r) = h(:, r)
      num(:, nrows, r) = [real(ntime, r8), t_row, dtime, dtmax_flow, dt_adv_used, dt_euler, &
                          real(sts_s(r), r8), real(subcyc(r), r8)]
    end do
  end subroutine

  subroutine grow()
    real(r8), allocatable :: a(:,:,:), b(:,:,:)
    integer :: n
    n = size(sol, 2)
    allocate (a(size(sol, 1), 2*n, nr), b(size(num, 1), 2*n, nr))
    a(:, 1:n, :) = sol
    b(:, 1:n, :) = num
    call move_alloc(a, sol)
    call move_alloc(b, num)
  end subroutine

  ! after each round of diffusion sub-cycles (diffusion_step, 4659-4669)
  subroutine on_cycle(user, cycle, rec) bind(c)
    use hft, only: hft_cycle
    type(c_ptr), value :: user
    integer(c_int), value :: cycle
    type(hft_cycle), intent(in) :: rec(*)
    integer :: r
    do r = 1, nr
      if (rec(r)%dt /= 0) then
        sts_s(r) = rec(r)%sts_s
        if (verbose > 0) subcyc(r) = rec(r)%subcycles
      end if
    end do
    if (verbose > 1) call emit(step_start + rec(1)%time_stepped)
  end subroutine

end module hft_hipft_state

program hft_hipft
  use iso_c_binding
  use hft
  use hft_hipft_state
  implicit none

  ! &hipft_input_parameters with the reference's defaults (src/hipft.f90:493-726)
  integer :: res_nt = 0, res_np = 0, n_realizations = 1, validation_run = 0
  real(r8) :: validation_run_width = 0.03_r8
  character(512) :: initial_map_filename = ''
  logical :: initial_map_flux_balance = .false.
  real(r8) :: initial_map_mult_fac = 1.0_r8
  character(512) :: output_map_root_filename = 'hipft_brmap', output_map_directory = 'output_maps'
  logical :: output_map_flux_balance = .false., output_flows = .false.
  integer(8) :: output_map_idx_start = 1, output_map_idx_cadence = 0
  character(512) :: output_flows_directory = 'output_flows'
  real(r8) :: output_map_time_cadence = 0, output_history_time_cadence = 0
  logical :: output_map_2d = .true., output_single_precision = .true.
  real(r8) :: time_end = 1.0_r8, dt_min = 1.0e-15_r8, dt_max = huge(1.0_r8)
  logical :: strang_splitting = .false.
  real(r8) :: pole_flux_lat_limit = 30.0_r8
  logical :: advance_flow = .false., flow_attenuate = .false.
  real(r8) :: flow_vp_rigid_omega = 0, flow_rigid_omega = 0, flow_vt_const = 0
  real(r8) :: flow_attenuate_value = 500.0_r8
  real(r8), target :: flow_attenuate_values(max_realizations) = -1.0_r8
  integer :: flow_dr_model = 0, flow_mf_model = 0
  real(r8) :: flow_dr_coef_p0_value = 46.0_r8, flow_dr_coef_p2_value = -262.0_r8
  real(r8) :: flow_dr_coef_p4_value = -379.0_r8
  real(r8), target :: flow_dr_coef_p0_values(max_realizations) = -1000000.0_r8
  real(r8), target :: flow_dr_coef_p2_values(max_realizations) = -1000000.0_r8
  real(r8), target :: flow_dr_coef_p4_values(max_realizations) = -1000000.0_r8
  real(r8) :: flow_mf_coef_p1_value = 22.0_r8, flow_mf_coef_p3_value = 11.0_r8
  real(r8) :: flow_mf_coef_p5_value = -28.0_r8
  real(r8), target :: flow_mf_coef_p1_values(max_realizations) = -1000000.0_r8
  real(r8), target :: flow_mf_coef_p3_values(max_realizations) = -1000000.0_r8
  real(r8), target :: flow_mf_coef_p5_values(max_realizations) = -1000000.0_r8
  logical :: use_flow_from_files = .false.
  character(512) :: flow_list_filename = ' ', flow_root_dir = '.'
  integer :: flow_num_method = 4
  real(r8) :: upwind = 1.0_r8, weno_eps = 0
  logical :: advance_diffusion = .false., diffusion_coef_grid = .false.
  real(r8) :: diffusion_coef_constant = 0, diffusion_coef_factor = diff_km2_s_to_rs2_hr
  real(r8) :: diffusion_coef_constants(max_realizations) = -1.0_r8
  character(512) :: diffusion_coef_filename = ' '
  integer :: diffusion_num_method = 3, diffusion_subcycles = 0
  logical :: assimilate_data = .false., assimilate_data_balance_flux = .false.
  logical :: assimilate_data_add = .false., assimilate_data_custom_from_mu = .false.
  real(r8) :: assimilate_data_mult_fac = 1.0_r8
  character(512) :: assimilate_data_map_list_filename = ' ', assimilate_data_map_root_dir = '.'
  real(r8) :: assimilate_data_lat_limit = 0, assimilate_data_mu_power = 4.0_r8
  real(r8) :: assimilate_data_mu_limit = 0.1_r8
  real(r8) :: assimilate_data_lat_limits(max_realizations) = -1.0_r8
  real(r8) :: assimilate_data_mu_powers(max_realizations) = -1.0_r8
  real(r8) :: assimilate_data_mu_limits(max_realizations) = -1.0_r8
  logical :: advance_source = .false., source_from_file = .false.
  character(512) :: source_filename = ' '
  integer :: source_rfe_model = 0, source_rfe_seed = -9999
  real(r8) :: source_rfe_total_unsigned_flux_per_hour = 0, source_rfe_lifetime = 0.3_r8
  real(r8) :: source_rfe_total_unsigned_flux_per_hours(max_realizations) = -1.0_r8

  namelist /hipft_input_parameters/ verbose, res_nt, res_np, n_realizations, &
    initial_map_filename, initial_map_flux_balance, initial_map_mult_fac, validation_run, &
    validation_run_width, time_start, time_end, output_map_idx_start, output_map_directory, &
    output_map_root_filename, output_map_idx_cadence, output_map_time_cadence, output_map_2d, &
    output_map_flux_balance, output_flows, output_flows_directory, &
    output_history_time_cadence, output_single_precision, dt_min, dt_max, strang_splitting, &
    pole_flux_lat_limit, advance_flow, flow_vp_rigid_omega, flow_rigid_omega, flow_vt_const, &
    flow_attenuate, flow_attenuate_value, flow_attenuate_values, flow_dr_model, &
    flow_dr_coef_p0_value, flow_dr_coef_p2_value, flow_dr_coef_p4_value, &
    flow_dr_coef_p0_values, flow_dr_coef_p2_values, flow_dr_coef_p4_values, flow_mf_model, &
    flow_mf_coef_p1_value, flow_mf_coef_p3_value, flow_mf_coef_p5_value, &
    flow_mf_coef_p1_values, flow_mf_coef_p3_values, flow_mf_coef_p5_values, &
    use_flow_from_files, flow_list_filename, flow_root_dir, flow_num_method, weno_eps, upwind, &
    advance_diffusion, diffusion_coef_filename, diffusion_coef_grid, diffusion_coef_factor, &
    diffusion_num_method, diffusion_subcycles, diffusion_coef_constant, &
    diffusion_coef_constants, assimilate_data, assimilate_data_balance_flux, &
    assimilate_data_mult_fac, assimilate_data_map_list_filename, assimilate_data_map_root_dir, &
    assimilate_data_custom_from_mu, assimilate_data_mu_power, assimilate_data_mu_powers, &
    assimilate_data_lat_limit, assimilate_data_lat_limits, assimilate_data_mu_limit, &
    assimilate_data_mu_limits, assimilate_data_add, advance_source, source_from_file, &
    source_filename, source_rfe_model, source_rfe_total_unsigned_flux_per_hour, &
    source_rfe_total_unsigned_flux_per_hours, source_rfe_lifetime, source_rfe_seed

  character(len=1024) :: infile
  integer :: u, ios, i, j, l, r
  real(r8), allocatable, target :: map0(:), p0(:), p2(:), p4(:), p1(:), p3(:), p5(:), att(:)
  real(r8), allocatable :: y0(:,:), y5(:,:)
  logical :: ts_needs, tsflow_needs, flow_needs, done
  real(r8) :: dt, d1, d2, dur
  type(c_funptr) :: cb

  infile = 'hipft.in'
  if (command_argument_count() >= 1) call get_command_argument(1, infile)
  open (newunit=u, file=trim(infile), status='old', action='read', iostat=ios)
  if (ios /= 0) then
    write (*,*) '### hft_hipft: could not open the input file ', trim(infile)
    stop 1
  end if
  read (u, nml=hipft_input_parameters)
  close (u)

  ! process_input_parameters (src/hipft.f90:8482-8579) and this driver's scope
  if (output_map_time_cadence > 0 .and. output_map_idx_cadence > 0) &
    call refuse('output_map_time_cadence and output_map_idx_cadence together')
  if (n_realizations > max_realizations) call refuse('n_realizations above 2000')
  if (len_trim(initial_map_filename) > 0) call refuse('initial_map_filename (map files)')
  if (res_nt < 3 .or. res_np < 3) call refuse('res_nt/res_np below 3 without a map file')
  if (use_flow_from_files) call refuse('use_flow_from_files')
  if (assimilate_data) call refuse('assimilate_data')
  if (advance_source) call refuse('advance_source')
  if (output_map_time_cadence > 0 .or. output_map_idx_cadence > 0) call refuse('map output')
  if (initial_map_flux_balance .or. output_map_flux_balance) call refuse('flux balancing of maps')
  if (flow_num_method /= 4) call refuse('flow_num_method other than 4 (SSPRK(4,3)+WENO3)')
  if (diffusion_num_method /= 3) call refuse('diffusion_num_method other than 3 (RKG2)')
  if (upwind /= 1.0_r8) call refuse('upwind other than 1')
  if (flow_rigid_omega /= 0) call refuse('flow_rigid_omega')
  if (len_trim(diffusion_coef_filename) > 0 .or. diffusion_coef_grid) &
    call refuse('diffusivity from a file or grid')
  if (output_flows) call refuse('output_flows')
  nr = n_realizations
  do r = 1, max_realizations
    if (flow_attenuate_values(r) < 0) flow_attenuate_values(r) = flow_attenuate_value
    if (flow_dr_coef_p0_values(r) < -100000.0_r8) flow_dr_coef_p0_values(r) = flow_dr_coef_p0_value
    if (flow_dr_coef_p2_values(r) < -100000.0_r8) flow_dr_coef_p2_values(r) = flow_dr_coef_p2_value
    if (flow_dr_coef_p4_values(r) < -100000.0_r8) flow_dr_coef_p4_values(r) = flow_dr_coef_p4_value
    if (flow_mf_coef_p1_values(r) < -100000.0_r8) flow_mf_coef_p1_values(r) = flow_mf_coef_p1_value
    if (flow_mf_coef_p3_values(r) < -100000.0_r8) flow_mf_coef_p3_values(r) = flow_mf_coef_p3_value
    if (flow_mf_coef_p5_values(r) < -100000.0_r8) flow_mf_coef_p5_values(r) = flow_mf_coef_p5_value
    if (diffusion_coef_constants(r) < 0) diffusion_coef_constants(r) = diffusion_coef_constant
  end do

  ! mesh of a resolution-only run: phi nodes i*2pi/(res_np-1), theta i*pi/(res_nt-1)
  ntm = res_nt
  npm1 = res_np
  allocate (tmesh(ntm), pmesh(npm1), sti(ntm), map0(npm1*ntm*nr), fval(npm1*ntm*nr))
  d1 = twopi/real(res_np - 1, r8)
  d2 = pi/real(res_nt - 1, r8)
  do l = 1, npm1
    pmesh(l) = real(l - 1, r8)*d1
  end do
  do j = 1, ntm
    tmesh(j) = real(j - 1, r8)*d2
    sti(j) = 0
    if (j > 1 .and. j < ntm) sti(j) = 1.0_r8/sin(tmesh(j))
  end do
  map0 = 0
  vkind = validation_run
  vwidth = validation_run_width
  dur = time_end - time_start
  select case (validation_run)
  case (1)  ! Y_6^0 + sqrt(14/11) Y_6^5
    allocate (y0(npm1, ntm), y5(npm1, ntm))
    call tesseral(0, 6, y0)
    call tesseral(5, 6, y5)
    do r = 1, nr
      map0((r - 1)*npm1*ntm + 1:r*npm1*ntm) = reshape(y0 + sqrt(14.0_r8/11.0_r8)*y5, [npm1*ntm])
    end do
  case (2)  ! two blobs; the flows take one revolution / a quarter turn
    if (abs(flow_vt_const) > 0) flow_vt_const = (pi_two/dur)/km_s_to_rs_hr
    if (abs(flow_vp_rigid_omega) > 0) flow_vp_rigid_omega = (twopi/dur)/km_s_to_rs_hr
    vt_const = flow_vt_const
    vp_omega = flow_vp_rigid_omega
    call gauss_field(0.0_r8, map0)
  case (0)
  case default
    call refuse('validation_run other than 0, 1, 2')
  end select

  ! set-up (set_mesh, load_initial_condition, load_flows, load_diffusion)
  call check(hft_create(ntm, npm1, tmesh, pmesh, nr, HFT_MODE_COUPLED, ctx), 'hft_create')
  call check(hft_set_state(ctx, map0, HFT_LAYOUT_PT, initial_map_mult_fac), 'hft_set_state')
  if (weno_eps /= 0) call check(hft_set_weno_eps(ctx, weno_eps), 'hft_set_weno_eps')
  if (pole_flux_lat_limit /= 30.0_r8) &
    call check(hft_set_pole_flux_lat_limit(ctx, pole_flux_lat_limit), 'hft_set_pole_flux_lat_limit')
  allocate (eta(nr), vscale(nr), sts_s(nr), subcyc(nr))
  allocate (sol(2 + HFT_NHIST, 64, nr), num(8, 64, nr))
  eta = diffusion_coef_constants(1:nr)
  dfac = diffusion_coef_factor
  if (vkind == 1) call check(hft_set_validation_field(ctx, c_loc(map0)), 'hft_set_validation_field')
  allocate (p0(nr), p2(nr), p4(nr), p1(nr), p3(nr), p5(nr), att(nr))
  p0 = flow_dr_coef_p0_values(1:nr); p2 = flow_dr_coef_p2_values(1:nr)
  p4 = flow_dr_coef_p4_values(1:nr); p1 = flow_mf_coef_p1_values(1:nr)
  p3 = flow_mf_coef_p3_values(1:nr); p5 = flow_mf_coef_p5_values(1:nr)
  att = flow_attenuate_values(1:nr)
  sts_s = 0
  subcyc = diffusion_subcycles
  dt_euler = 0
  if (advance_diffusion) then
    call check(hft_set_diffusion(ctx, eta, diffusion_coef_factor), 'hft_set_diffusion')
    call check(hft_dt_euler(ctx, dt_euler, c_null_ptr), 'hft_dt_euler')
  end if
  hist_cadence = output_history_time_cadence
  time_now = time_start
  ntime = 0
  dtime = 0
  dtmax_flow = 0
  dt_adv_used = 0
  call emit(time_now)
  ts_needs = .true.
  tsflow_needs = .true.
  flow_needs = .true.
  cb = c_null_funptr
  if (advance_diffusion) cb = c_funloc(on_cycle)

  ! main loop (src/hipft.f90:903-948)
  do
    ntime = ntime + 1
    ! update_step: flows (update_flow 3371-3523)
    if (advance_flow .and. flow_needs) then
      flow_needs = .false.
      tsflow_needs = .true.
      ts_needs = .true.
      call set_flows()
      if (flow_attenuate) then
        call check(hft_set_flow_attenuation(ctx, c_loc(att)), 'hft_set_flow_attenuation')
        flow_needs = .true.
      end if
    end if
    ! update_timestep (4308-4461): end of run, flow CFL, dt_max, dt_min
    if (ts_needs) then
      dtime = time_end - time_now
      if (advance_flow) then
        dtmax_flow = huge(1.0_r8)
        if (tsflow_needs) then
          call check(hft_flow_dtmax(ctx, merge(1, 0, strang_splitting), dtmax_flow, c_null_ptr), &
                     'hft_flow_dtmax')
          tsflow_needs = .false.
        end if
        if (dtmax_flow < dtime) dtime = dtmax_flow
      end if
      if (dt_max < dtime) dtime = dt_max
      if (dtime < dt_min) write (*,*) '   UPDATE_TIMESTEP: Warning! Time step is smaller than DTMIN!'
      ts_needs = .false.
    end if
    if (time_now + dtime >= time_end) then
      dtime = time_end - time_now
      ts_needs = .false.
    end if
    dt = dtime
    if (advance_flow) dt_adv_used = merge(0.5_r8*dt, dt, strang_splitting)
    step_start = time_now
    ! flux_transport_step (4463-4507)
    call check(hft_flux_transport_step(ctx, dt, merge(1, 0, strang_splitting), &
                                       diffusion_subcycles, cb, c_null_ptr), &
               'hft_flux_transport_step')
    if (advance_diffusion .and. diffusion_subcycles == 0) &
      call check(hft_diffusion_stats(ctx, c_null_ptr, c_loc(subcyc)), 'hft_diffusion_stats')
    time_now = time_now + dt
    ! analysis_step + output_histories
    call emit(time_now)
    done = time_now >= time_end
    if (done) exit
  end do

  call write_histories()
  call check(hft_destroy(ctx), 'hft_destroy')
  write (*,'(a,i0,a,i0,a)') 'hft_hipft: ', ntime, ' steps of ', nr, ' realizations done'

contains

  subroutine refuse(what)
    character(*), intent(in) :: what
    write (*,'(a,a,a)') '### hft_hipft: ', what, &
      ' is not supported by this driver (the Python driver hipft_amd/driver.py covers it)'
    stop 1
  end subroutine

  subroutine set_flows()
    type(c_ptr) :: a0, a2, a4, a1, a3, a5
    a0 = c_null_ptr; a2 = c_null_ptr; a4 = c_null_ptr
    a1 = c_null_ptr; a3 = c_null_ptr; a5 = c_null_ptr
    if (flow_dr_model == 1) then
      a0 = c_loc(p0); a2 = c_loc(p2); a4 = c_loc(p4)
    end if
    if (flow_mf_model == 1) then
      a1 = c_loc(p1); a3 = c_loc(p3); a5 = c_loc(p5)
    end if
    call check(hft_set_flow_analytic(ctx, a0, a2, a4, a1, a3, a5, flow_vp_rigid_omega, &
                                     flow_vt_const), 'hft_set_flow_analytic')
  end subroutine

  ! Real spherical harmonic Y(np, nt) of degree l, order m (fully
  ! normalised, Condon-Shortley phase), by the forward column recursion of
  ! Holmes & Featherstone (2002) from the sectoral seed (2m+1)!!-type
  ! product; cos(m phi) for m >= 0, sin(|m| phi) for m < 0.
  subroutine tesseral(m, ll, y)
    integer, intent(in) :: m, ll
    real(r8), intent(out) :: y(:,:)
    real(r8) :: pold(ntm), poldold(ntm), pl, anm, bnm, ct, stm, ang, pos
    integer :: am, im, il, jj, kk, q
    am = abs(m)
    pold = 1.0_r8
    poldold = 0
    if (am > 0) then
      pold = sqrt(3.0_r8)
      do im = 2, am
        pold = sqrt((2.0_r8*im + 1.0_r8)/(2.0_r8*im))*pold
      end do
    end if
    do il = am + 1, ll
      anm = sqrt(((2.0_r8*il - 1.0_r8)*(2.0_r8*il + 1.0_r8))/real((il - am)*(il + am), r8))
      bnm = sqrt(((2.0_r8*il + 1.0_r8)*real(il + am - 1, r8)*real(il - am - 1, r8)) &
                 /(real((il - am)*(il + am), r8)*(2.0_r8*il - 3.0_r8)))
      do jj = 1, ntm
        ct = cos(tmesh(jj))
        pl = anm*ct*pold(jj) - bnm*poldold(jj)
        poldold(jj) = pold(jj)
        pold(jj) = pl
      end do
    end do
    pos = 0
    if (m + 1 >= 0) pos = 1
    do jj = 1, ntm
      ! sin(t)^|m| by square-and-multiply, as a compiler expands an integer power
      stm = 1.0_r8
      if (am > 0) then
        stm = ipow(sin(tmesh(jj)), am)
      end if
      q = 1
      if (mod(am, 2) == 1) q = -1
      pl = real(q, r8)*stm*(1.0_r8/sqrt(4.0_r8*3.141592653589793_r8))*pold(jj)
      do kk = 1, npm1
        ang = pos*cos(real(m, r8)*pmesh(kk)) + (1.0_r8 - pos)*sin(real(am, r8)*pmesh(kk))
        y(kk, jj) = pl*ang
      end do
    end do
  end subroutine

  real(r8) function ipow(x, n)
    real(r8), intent(in) :: x
    integer, intent(in) :: n
    real(r8) :: sq
    integer :: k
    logical :: have
    ipow = 1.0_r8
    sq = x
    k = n
    have = .false.
    do while (k > 0)
      if (iand(k, 1) == 1) then
        if (have) then
          ipow = ipow*sq
        else
          ipow = sq
        end if
        have = .true.
      end if
      k = ishft(k, -1)
      if (k > 0) sq = sq*sq
    end do
  end function

  ! hipft_history_{sol,num}_r######.out (headers 1616-1646, rows 2633-2634)
  subroutine write_histories()
    character(len=22), parameter :: solh(15) = [character(len=22) :: 'TIME', 'BR_MIN', 'BR_MAX', &
      'BR_ABS_MIN', 'FLUX_POSITIVE', 'FLUX_NEGATIVE', 'NPOLE_FLUX_POSITIVE', &
      'NPOLE_FLUX_NEGATIVE', 'SPOLE_FLUX_POSITIVE', 'SPOLE_FLUX_NEGATIVE', 'NPOLE_AREA', &
      'SPOLE_AREA', 'EQ_DIPOLE', 'AX_DIPOLE', 'VALIDATION_ERR_HHabs']
    character(len=22), parameter :: numh(7) = [character(len=22) :: 'TIME', 'DTIME', &
      'DTIME_ADV_STB', 'DTIME_ADV_USED', 'DTIME_DIFF_STB', 'N_DIFF_PER_STEP', 'N_DIFF_CYCLES']
    character(len=64) :: fname
    integer :: uu, rr, k
    do rr = 1, nr
      write (fname, '(a,i6.6,a)') 'hipft_history_sol_r', rr, '.out'
      open (newunit=uu, file=trim(fname), status='replace', action='write')
      write (uu, '(a10,15(1x,a22),a)') 'STEP', (adjustr(solh(k)), k = 1, 15), ' '
      do k = 1, nrows
        write (uu, '(i10,1p,15(1x,e23.15e3),a)') nint(sol(1, k, rr)), sol(2:, k, rr), ' '
      end do
      close (uu)
      write (fname, '(a,i6.6,a)') 'hipft_history_num_r', rr, '.out'
      open (newunit=uu, file=trim(fname), status='replace', action='write')
      write (uu, '(a10,7(1x,a22),a)') 'STEP', (adjustr(numh(k)), k = 1, 7), ' '
      do k = 1, nrows
        write (uu, '(i10,1x,5(1pe23.15e3,1x),i15,1x,i15)') nint(num(1, k, rr)), num(2:6, k, rr), &
          nint(num(7, k, rr), 8), nint(num(8, k, rr), 8)
      end do
      close (uu)
    end do
  end subroutine

end program hft_hipft
