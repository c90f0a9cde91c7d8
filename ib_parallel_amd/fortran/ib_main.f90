! ib_main.f90 -- PETSc-free Fortran driver for the MI355X time step.
!
! Same program shape as the reference's main (src/main.F90:16-57) for one
! process: read the namelist input/ib.inp (src/parameters.F90:51-78) and the
! body decks input/body.NNN.inp (src/grid.F90:252-288), then
!   do while (it < istop): advance; save-step .var file; force, redistributed
!   force, theta and iteration files
! with advance(it) replaced by ib_advance through iso_c_binding.  A run with
! istart > 0 resumes from output/data00_<istart>.var as setup_variables does
! (src/variables.F90:237-245): read_variables, then the weights rebuilt from
! the saved body positions (ib_set_body).  The output
! files keep the reference's names and record layouts for nproc = 1, rank 0:
!   output/dataRR_IIIIIII.var   write_variables (src/variables.F90:709-803)
!   output/force01_00.dat       write_force     (src/operators_struct.F90:655-679)
!   output/forcerdst01_00.dat   write_force_rdst (src/operators_struct.F90:683-707)
!   output/theta01_BB.dat       write_theta     (src/operators_struct.F90:624-651)
!   output/iter01.dat           write_iterations (src/timestepper.F90:799-815)
!   output/pressure00_IIIIIII.var  write_pressure (src/operators_pressure.F90:857-880)
!                                  on save steps when compute_pressure = T
! Text outputs are opened with a long record length so a step never wraps
! onto two lines (the reference is built with ifort -no-wrap-margin,
! Makefile:58, for the same reason).
program ib_main
  use iso_c_binding
  use ib_gpu_mod
  implicit none

  ! namelist READ_PARAMETERS (src/parameters.F90:55-59)
  integer :: istart = 0, istop = 10, isave = 10, irestart = 0, m = 0, n = 0, mgridlev = 1, dimen = 2
  real(8) :: dt = 0.d0, re = 0.d0, fsi_tol = 1.d-7, atol = 1.d-7, rtol = 1.d-7, len = 0.d0
  real(8) :: offsetx = 0.d0, offsety = 0.d0, sdxi = 0.d0, sdxe = 0.d0, sdyi = 0.d0, sdye = 0.d0, delta_sd = 0.d0
  logical :: compute_pressure = .false., num_stat = .false., motion_prescribed = .false.
  logical :: sub_domain = .false., sub_domain_full_rectangular = .false., sub_domain_precomputed = .false.
  namelist /read_parameters/ istart, istop, isave, irestart, m, n, dt, re, &
                             fsi_tol, atol, rtol, len, offsetx, offsety, &
                             mgridlev, compute_pressure, &
                             dimen, num_stat, motion_prescribed, sub_domain, sub_domain_full_rectangular, &
                             sub_domain_precomputed, sdxi, sdxe, sdyi, sdye, delta_sd

  type :: body_deck
    integer :: npts, fam
    real(8) :: gamma0 = 0.d0, alpha0 = 0.d0, itheta = 0.d0, ktheta = 0.d0, ctheta = 0.d0
    real(c_double), allocatable :: x(:), y(:)
  end type body_deck

  type(body_deck), allocatable, target :: decks(:)
  type(ib_body), allocatable :: cbodies(:)
  type(ib_params) :: p
  type(ib_subdomain) :: sdp
  type(ib_step_info) :: info
  type(ib_sizes) :: sz
  type(c_ptr) :: h
  integer(c_int) :: rc
  integer(c_int32_t) :: it
  integer :: nbod, i
  logical :: exists
  real(8), parameter :: pi = 4.0d0 * atan(1.0d0)
  character(len=3) :: num

  inquire(file='input/ib.inp', exist=exists)
  if (.not. exists) stop 'cannot find input file'
  open(unit=3, file='input/ib.inp', form='formatted', status='old')
  read(unit=3, nml=read_parameters)
  close(3)
  write(*,*) 'read input file'

  ! body decks: input/body.001.inp, 002, ... until one is missing
  nbod = 0
  do
    write(num, '(I3.3)') nbod + 1
    inquire(file='input/body.'//num//'.inp', exist=exists)
    if (.not. exists) exit
    nbod = nbod + 1
  end do
  allocate(decks(max(nbod, 1)), cbodies(max(nbod, 1)))
  do i = 1, nbod
    call read_body(i, decks(i))
    write(*,*) 'body no.', i, 'has', decks(i)%npts, 'points. Family?', decks(i)%fam
    cbodies(i)%npts = decks(i)%npts
    cbodies(i)%fam = decks(i)%fam
    cbodies(i)%gamma0 = decks(i)%gamma0
    cbodies(i)%alpha0 = decks(i)%alpha0
    cbodies(i)%itheta = decks(i)%itheta
    cbodies(i)%ktheta = decks(i)%ktheta
    cbodies(i)%ctheta = decks(i)%ctheta
    cbodies(i)%x = c_loc(decks(i)%x)
    cbodies(i)%y = c_loc(decks(i)%y)
  end do

  p%m = m; p%n = n; p%mgridlev = mgridlev; p%istart = istart
  p%dt = dt; p%re = re; p%len = len; p%offsetx = offsetx; p%offsety = offsety
  p%fsi_tol = fsi_tol; p%atol = atol; p%rtol = rtol
  p%gmres_restart = 30; p%gmres_max_it = 10000; p%fsi_max_count = 1000
  p%deltafnc_f32_literals = 1
  p%num_stat = merge(1, 0, num_stat); p%sub_domain = merge(1, 0, sub_domain)
  p%launch_mode = 0
  p%schur_mode = 0
  p%face_margin = 0
  p%gmres_form = 0
  p%reserved = 0
  call getenv_int('IB_GMRES_RESTART', p%gmres_restart)

  if (sub_domain) then
    ! sub-domain approach (src/timestepper.F90:79-97); interp_points,
    ! droptol and the sparsity window take the reference's values (0)
    sdp%sdxi = sdxi; sdp%sdxe = sdxe; sdp%sdyi = sdyi; sdp%sdye = sdye; sdp%delta_sd = delta_sd
    sdp%full_rectangular = merge(1, 0, sub_domain_full_rectangular)
    sdp%precomputed = merge(1, 0, sub_domain_precomputed)
    sdp%droptol = 0.d0; sdp%interp_points = 0; sdp%support_nz = 0
    rc = ib_create_subdomain(p, sdp, int(nbod, c_int32_t), cbodies, h)
  else
    rc = ib_create(p, int(nbod, c_int32_t), cbodies, h)
  end if
  if (rc /= IB_OK) then
    write(*,*) 'ib_create failed (', rc, '): ', trim(ib_error_message(c_null_ptr))
    stop 1
  end if
  rc = ib_get_sizes(h, sz)

  call execute_command_line('mkdir -p output')
  if (istart > 0) call read_variables(int(istart, c_int32_t))
  it = istart
  do while (it < istop)
    rc = ib_advance(h, it, info)
    if (rc /= IB_OK .and. rc /= IB_E_NOCONV_GMRES .and. rc /= IB_E_NOCONV_FSI) then
      write(*,*) 'ib_advance failed (', rc, '): ', trim(ib_error_message(h))
      stop 1
    end if
    if (mod(it, 10) == 0) write(*,*) '...Advancing to itime =', it
    if (it >= irestart) then
      if ((mod(it - irestart, isave) == 0) .or. (it == istop)) then
        call write_variables(it)
        if (compute_pressure) call write_pressure(it)
      end if
    end if
    call write_force(it, 0, 'force')
    call write_force(it, 1, 'forcerdst')
    call write_theta(it)
    if (.not. num_stat) call write_iterations(it, info)
  end do
  rc = ib_destroy(h)

contains

  subroutine getenv_int(name, val)
    character(len=*), intent(in) :: name
    integer(c_int32_t), intent(inout) :: val
    character(len=32) :: s
    integer :: st, v
    call get_environment_variable(name, s, status=st)
    if (st == 0) then
      read(s, *, iostat=st) v
      if (st == 0) val = v
    end if
  end subroutine getenv_int

  ! setup_geometry deck format (src/grid.F90:256-270)
  subroutine read_body(ib, d)
    integer, intent(in) :: ib
    type(body_deck), intent(inout) :: d
    character(len=3) :: fn
    integer :: j
    write(fn, '(I3.3)') ib
    open(unit=8, file='input/body.'//fn//'.inp', form='formatted', status='old')
    read(8, *) d%npts
    read(8, *) d%fam
    ! family 3 (a beam; a stub in the reference, which reads no parameters
    ! for it) carries the same five lines as family 2: pitch, alpha0, mass
    ! per unit length, bending stiffness EI, damping
    if (d%fam == 2 .or. d%fam == 3) then
      read(8, *) d%gamma0; d%gamma0 = d%gamma0 * pi / 180
      read(8, *) d%alpha0; d%alpha0 = d%alpha0 * pi / 180
      read(8, *) d%itheta
      read(8, *) d%ktheta
      read(8, *) d%ctheta
    end if
    allocate(d%x(d%npts), d%y(d%npts))
    do j = 1, d%npts
      read(8, *) d%x(j), d%y(j)
    end do
    close(8)
  end subroutine read_body

  ! write_variables (src/variables.F90:709-803), nproc = 1
  subroutine write_variables(itn)
    integer(c_int32_t), intent(in) :: itn
    character(len=7) :: charit
    real(c_double), allocatable :: w(:), s(:), qx(:), qy(:), ro(:), rhs_old(:)
    real(c_double), allocatable :: f(:), fr(:), xb(:), yb(:), th(:), thd(:), thdd(:), q0(:)
    integer :: k, rcl
    integer(c_int32_t) :: pn
    allocate(w(sz%n_nodes), s(sz%n_nodes), ro(sz%n_nodes), qx(sz%n_xfaces), qy(sz%n_yfaces))
    allocate(rhs_old(int(sz%n_nodes) * mgridlev), q0(max(sz%n_xfaces, sz%n_yfaces)))
    allocate(f(sz%nf), fr(sz%nf), xb(sz%nf), yb(sz%nf))
    allocate(th(max(sz%ntors, 1)), thd(max(sz%ntors, 1)), thdd(max(sz%ntors, 1)))
    write(*,*) 'writing variables at it=', itn
    write(charit, '(I7.7)') itn
    open(unit=100, file='output/data00_'//charit//'.var', form='unformatted', status='unknown')
    write(100) int(m, c_int32_t), int(n, c_int32_t), int(mgridlev, c_int32_t), int(sz%nb, c_int32_t)
    write(100) re, dt, len, offsetx, offsety
    write(100) 2_c_int32_t, int(m, c_int32_t), int(m - 1, c_int32_t), 2_c_int32_t, int(n, c_int32_t), &
               int(n - 1, c_int32_t)
    write(100) 1_c_int32_t, int(m + 1, c_int32_t), int(m + 1, c_int32_t), 1_c_int32_t, int(n, c_int32_t), &
               int(n, c_int32_t)
    write(100) 1_c_int32_t, int(m, c_int32_t), int(m, c_int32_t), 1_c_int32_t, int(n + 1, c_int32_t), &
               int(n + 1, c_int32_t)
    do k = 1, mgridlev
      rcl = ib_get_level_fields(h, int(k, c_int32_t), w, s, qx, qy, ro)
      rhs_old((k - 1) * sz%n_nodes + 1:k * sz%n_nodes) = ro
      write(100) w
      write(100) s
      write(100) qx
      ! motion_potential / motion_rotation with motion_grid = (-1, 0, 0, 0, 0)
      q0(1:sz%n_xfaces) = len / m * 2.d0**(k - 1)
      write(100) q0(1:sz%n_xfaces)
      q0(1:sz%n_xfaces) = 0.d0
      write(100) q0(1:sz%n_xfaces)
      write(100) qy
      q0(1:sz%n_yfaces) = 0.d0
      write(100) q0(1:sz%n_yfaces)
      write(100) q0(1:sz%n_yfaces)
    end do
    write(100) rhs_old
    write(100) 0.d0, 0.d0, 0.d0
    pn = int(sz%nf, c_int32_t)
    write(100) pn, 1_c_int32_t, pn
    rcl = ib_get_body(h, f, fr, xb, yb, th, thd, thdd)
    write(100) f
    write(100) xb
    write(100) yb
    write(100) int(sz%ntors, c_int32_t)
    if (sz%ntors > 0) write(100) th(1:sz%ntors), thd(1:sz%ntors), thdd(1:sz%ntors)
    write(100) fr
    close(100)
  end subroutine write_variables

  ! read_variables (src/variables.F90:807-913), nproc = 1: the records that
  ! write_variables writes, in the same order
  subroutine read_variables(itn)
    integer(c_int32_t), intent(in) :: itn
    character(len=7) :: charit
    real(c_double), allocatable :: w(:, :), s(:, :), qx(:, :), qy(:, :), rhs_old(:), q0(:)
    real(c_double), allocatable :: f(:), xb(:), yb(:), th(:), thd(:), thdd(:)
    integer(c_int32_t) :: hm, hn, hl, hnb, i6(6), pn, pi0, pe, nt
    real(c_double) :: hre, hdt, hlen, hox, hoy, rot(3)
    integer :: k, rcl, np
    logical :: there
    write(charit, '(I7.7)') itn
    inquire(file='output/data00_'//charit//'.var', exist=there)
    if (.not. there) then
      write(*,*) 'restart file output/data00_'//charit//'.var not found'
      stop 1
    end if
    write(*,*) 'reading variables at it=', itn
    np = sz%n_nodes
    allocate(w(np, mgridlev), s(np, mgridlev), qx(sz%n_xfaces, mgridlev), qy(sz%n_yfaces, mgridlev))
    allocate(rhs_old(np * mgridlev), q0(max(sz%n_xfaces, sz%n_yfaces)))
    allocate(f(max(sz%nf, 1)), xb(max(sz%nf, 1)), yb(max(sz%nf, 1)))
    allocate(th(max(sz%ntors, 1)), thd(max(sz%ntors, 1)), thdd(max(sz%ntors, 1)))
    open(unit=100, file='output/data00_'//charit//'.var', form='unformatted', status='old')
    read(100) hm, hn, hl, hnb
    if (hm /= m .or. hn /= n .or. hl /= mgridlev .or. hnb /= sz%nb) then
      write(*,*) 'restart file does not match the deck (m, n, mgridlev, nb):', hm, hn, hl, hnb
      stop 1
    end if
    read(100) hre, hdt, hlen, hox, hoy
    read(100) i6
    read(100) i6
    read(100) i6
    ! per level: omega, s, qx, q0px, q0rx, qy, q0py, q0ry (the uniform
    ! freestream q0 is recomputed every step, so its records are skipped)
    do k = 1, mgridlev
      read(100) w(:, k)
      read(100) s(:, k)
      read(100) qx(:, k)
      read(100) q0(1:sz%n_xfaces)
      read(100) q0(1:sz%n_xfaces)
      read(100) qy(:, k)
      read(100) q0(1:sz%n_yfaces)
      read(100) q0(1:sz%n_yfaces)
    end do
    read(100) rhs_old
    do k = 1, mgridlev
      rcl = ib_set_level_fields(h, int(k, c_int32_t), w(:, k), s(:, k), qx(:, k), qy(:, k), &
                                rhs_old((k - 1) * np + 1:k * np))
    end do
    read(100) rot
    read(100) pn, pi0, pe
    read(100) f(1:sz%nf)
    read(100) xb(1:sz%nf)
    read(100) yb(1:sz%nf)
    read(100) nt
    if (nt > 0) read(100) th(1:nt), thd(1:nt), thdd(1:nt)
    close(100)
    rcl = ib_set_body(h, f, xb, yb, th, thd, thdd)
    if (rcl /= IB_OK) then
      write(*,*) 'ib_set_body failed (', rcl, '): ', trim(ib_error_message(h))
      stop 1
    end if
  end subroutine read_variables

  ! calculate_pressure + write_pressure (src/main.F90:40-43,
  ! src/operators_pressure.F90:857-880), nproc = 1
  subroutine write_pressure(itn)
    integer(c_int32_t), intent(in) :: itn
    character(len=7) :: charit
    real(c_double), allocatable :: pr(:)
    integer :: k, rcl, nc
    nc = m * n
    allocate(pr(nc * mgridlev))
    rcl = ib_pressure(h, pr)
    if (rcl /= IB_OK) then
      write(*,*) 'ib_pressure failed (', rcl, '): ', trim(ib_error_message(h))
      stop 1
    end if
    write(charit, '(I7.7)') itn
    open(unit=100, file='output/pressure00_'//charit//'.var', form='unformatted', status='unknown')
    write(100) 1_c_int32_t, int(m, c_int32_t), int(m, c_int32_t), 1_c_int32_t, int(n, c_int32_t), int(n, c_int32_t)
    do k = 1, mgridlev
      write(100) pr((k - 1) * nc + 1:k * nc)
    end do
    close(100)
  end subroutine write_pressure

  ! write_force / write_force_rdst with write_force2 (src/operators_struct.F90:655-731)
  subroutine write_force(itn, which, stem)
    integer(c_int32_t), intent(in) :: itn
    integer, intent(in) :: which
    character(len=*), intent(in) :: stem
    real(c_double) :: fx(max(nbod, 1)), fy(max(nbod, 1))
    integer :: rcl, k
    rcl = ib_get_forces(h, int(which, c_int32_t), fx, fy)
    if (itn == 1) then
      open(unit=103, file='output/'//stem//'01_00.dat', form='formatted', status='replace', recl=65536)
    else
      open(unit=103, file='output/'//stem//'01_00.dat', form='formatted', status='unknown', position='append', recl=65536)
    end if
    write(103, *) itn, (fx(k), k = 1, nbod), (fy(k), k = 1, nbod)
    close(103)
  end subroutine write_force

  ! write_theta (src/operators_struct.F90:624-651): one file per torsional body
  subroutine write_theta(itn)
    integer(c_int32_t), intent(in) :: itn
    real(c_double), allocatable :: f(:), th(:), thd(:), thdd(:)
    character(len=2) :: cb
    integer :: rcl, t, b
    if (sz%ntors == 0) return
    allocate(f(sz%nf), th(sz%ntors), thd(sz%ntors), thdd(sz%ntors))
    rcl = ib_get_body(h, f, f, f, f, th, thd, thdd)
    t = 0
    do b = 1, nbod
      if (decks(b)%fam /= 2) cycle
      t = t + 1
      write(cb, '(I2.2)') b
      if (itn == 1) then
        open(unit=104, file='output/theta01_'//cb//'.dat', form='formatted', status='replace', recl=65536)
      else
        open(unit=104, file='output/theta01_'//cb//'.dat', form='formatted', status='unknown', position='append', recl=65536)
      end if
      write(104, *) itn, th(t)
      close(104)
    end do
  end subroutine write_theta

  ! write_iterations (src/timestepper.F90:799-815)
  subroutine write_iterations(itn, inf)
    integer(c_int32_t), intent(in) :: itn
    type(ib_step_info), intent(in) :: inf
    if (itn == 1) then
      open(unit=105, file='output/iter01.dat', form='formatted', status='replace', recl=65536)
    else
      open(unit=105, file='output/iter01.dat', form='formatted', status='unknown', position='append', recl=65536)
    end if
    write(105, *) itn, inf%ksp_iter, inf%fsi_iter
    close(105)
  end subroutine write_iterations

end program ib_main
