!
! hft_run.f90 -- a Fortran time loop over libhft through the iso_c_binding
! module hft (include/hft.f90): the shape of the reference's main loop
! (program HIPFT, src/hipft.f90:903-948) with the flux-transport step,
! the per-step analysis and the history rows handed to the C ABI.
!
! Usage:  hft_run <control file>
! The control file holds the namelist group &hft_run:
!   ntm, npm1, nr            grid and number of realizations
!   t_file, p_file           theta (ntm) and phi (npm1) scales, raw float64
!   map_file                 initial maps, raw float64 PT layout (npm1,ntm,nr)
!   mode                     0 coupled (reference semantics), 1 independent
!   eta(nr)                  diffusivity (km^2/s); 0 = no diffusion
!   dr_p0/p2/p4(nr)          differential-rotation coefficients (m/s)
!   mf_p1/p3/p5(nr)          meridional-flow coefficients (m/s)
!   advect                   .true. to advance the flows
!   dt, nsteps, strang, subcycles
!   hist_file, final_file    outputs: raw float64 history rows
!                            [nsteps][nr][2+14] (step, time, diagnostics)
!                            and the final state, TP layout (ntm,npm,nr)
!
program hft_run
  use iso_c_binding
  use hft
  implicit none
  integer, parameter :: maxr = 256
  real(c_double), parameter :: diff_km2_s_to_rs2_hr = 7.43162901307967e-09_c_double
  integer :: ntm = 0, npm1 = 0, nr = 1, mode = 1, nsteps = 1, strang = 1, subcycles = 0
  real(c_double) :: dt = 1.0_c_double
  logical :: advect = .false.
  real(c_double), target :: eta(maxr) = 0, dr_p0(maxr) = 0, dr_p2(maxr) = 0, dr_p4(maxr) = 0
  real(c_double), target :: mf_p1(maxr) = 0, mf_p3(maxr) = 0, mf_p5(maxr) = 0
  character(len=512) :: t_file = '', p_file = '', map_file = '', hist_file = 'hist.bin'
  character(len=512) :: final_file = ''
  namelist /hft_run/ ntm, npm1, nr, mode, nsteps, strang, subcycles, dt, advect, eta, &
                     dr_p0, dr_p2, dr_p4, mf_p1, mf_p3, mf_p5, t_file, p_file, map_file, &
                     hist_file, final_file
  character(len=512) :: ctl
  real(c_double), allocatable :: t(:), p(:), map(:), f(:), hist(:,:), rows(:,:,:)
  type(c_ptr) :: ctx
  integer(c_int) :: ierr
  integer :: u, n, r
  real(c_double) :: time

  if (command_argument_count() < 1) then
    write (*,*) 'usage: hft_run <control file>'
    stop 2
  end if
  call get_command_argument(1, ctl)
  open (newunit=u, file=trim(ctl), status='old', action='read')
  read (u, nml=hft_run)
  close (u)
  if (nr > maxr) stop 'hft_run: nr exceeds 256'

  allocate (t(ntm), p(npm1), map(npm1*ntm*nr), f(ntm*(npm1+1)*nr))
  allocate (hist(HFT_NHIST, nr), rows(2+HFT_NHIST, nr, nsteps))
  call read_raw(t_file, t)
  call read_raw(p_file, p)
  call read_raw(map_file, map)

  ! setup (set_mesh, load_initial_condition, update_flow, load_diffusion)
  call check(hft_create(ntm, npm1, t, p, nr, mode, ctx), 'hft_create')
  call check(hft_set_state(ctx, map, HFT_LAYOUT_PT, 1.0_c_double), 'hft_set_state')
  if (advect) then
    call check(hft_set_flow_analytic(ctx, c_loc(dr_p0), c_loc(dr_p2), c_loc(dr_p4), &
                                     c_loc(mf_p1), c_loc(mf_p3), c_loc(mf_p5), &
                                     0.0_c_double, 0.0_c_double), 'hft_set_flow_analytic')
  end if
  if (any(eta(1:nr) > 0)) then
    call check(hft_set_diffusion(ctx, eta, diff_km2_s_to_rs2_hr), 'hft_set_diffusion')
  end if

  ! main loop: flux_transport_step + analysis_step each step
  time = 0
  do n = 1, nsteps
    call check(hft_flux_transport_step(ctx, dt, strang, subcycles, c_null_funptr, c_null_ptr), &
               'hft_flux_transport_step')
    time = time + dt
    call check(hft_analysis(ctx, c_null_ptr, c_null_ptr, hist), 'hft_analysis')
    do r = 1, nr
      rows(1, r, n) = real(n, c_double)
      rows(2, r, n) = time
      rows(3:, r, n) = hist(:, r)
    end do
  end do

  call write_raw(hist_file, reshape(rows, [size(rows)]))
  if (len_trim(final_file) > 0) then
    call check(hft_get_state(ctx, f, HFT_LAYOUT_TP), 'hft_get_state')
    call write_raw(final_file, f)
  end if
  call check(hft_destroy(ctx), 'hft_destroy')
  write (*,'(a,i0,a,i0,a)') 'hft_run: ', nsteps, ' steps of ', nr, ' realizations done'

contains

  subroutine check(rc, what)
    integer(c_int), intent(in) :: rc
    character(*), intent(in) :: what
    character(kind=c_char), pointer :: msg(:)
    integer :: i
    if (rc == 0) return
    call c_f_pointer(hft_last_error(), msg, [512])
    write (*,'(a,a,i0,a)', advance='no') what, ' failed (', rc, '): '
    do i = 1, 512
      if (msg(i) == c_null_char) exit
      write (*,'(a)', advance='no') msg(i)
    end do
    write (*,*)
    stop 1
  end subroutine

  subroutine read_raw(name, a)
    character(*), intent(in) :: name
    real(c_double), intent(out) :: a(:)
    integer :: uu
    open (newunit=uu, file=trim(name), access='stream', form='unformatted', status='old', &
          action='read')
    read (uu) a
    close (uu)
  end subroutine

  subroutine write_raw(name, a)
    character(*), intent(in) :: name
    real(c_double), intent(in) :: a(:)
    integer :: uu
    open (newunit=uu, file=trim(name), access='stream', form='unformatted', status='replace', &
          action='write')
    write (uu) a
    close (uu)
  end subroutine

end program hft_run
