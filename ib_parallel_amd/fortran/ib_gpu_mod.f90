! ib_gpu_mod.f90 -- iso_c_binding interface to libibgpu.so (include/ib_gpu.h).
!
! The derived types match the C structs field for field (bind(C)); every
! function returns the C status code (0 = IB_OK).  This is the binding a
! Fortran host such as the reference's main program (src/main.F90) uses to
! call the MI355X time step instead of advance() (src/timestepper.F90:18).
module ib_gpu_mod
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: IB_OK = 0, IB_E_INVAL = 1, IB_E_NOCONV_GMRES = 2, &
                               IB_E_NOCONV_FSI = 3, IB_E_HIP = 4, IB_E_UNSUPPORTED = 5

  type, bind(C) :: ib_params
    integer(c_int32_t) :: m, n, mgridlev, istart
    real(c_double) :: dt, re, len, offsetx, offsety
    real(c_double) :: fsi_tol, atol, rtol
    integer(c_int32_t) :: gmres_restart, gmres_max_it, fsi_max_count
    integer(c_int32_t) :: deltafnc_f32_literals, num_stat, sub_domain, launch_mode
    integer(c_int32_t) :: schur_mode, face_margin
    integer(c_int32_t) :: gmres_form
    integer(c_int32_t) :: reserved(2)
  end type ib_params

  type, bind(C) :: ib_subdomain
    real(c_double) :: sdxi, sdxe, sdyi, sdye, delta_sd
    integer(c_int32_t) :: full_rectangular, precomputed
    real(c_double) :: droptol
    integer(c_int32_t) :: interp_points, support_nz
  end type ib_subdomain

  type, bind(C) :: ib_body
    integer(c_int32_t) :: npts, fam
    real(c_double) :: gamma0, alpha0, itheta, ktheta, ctheta
    type(c_ptr) :: x, y
  end type ib_body

  type, bind(C) :: ib_step_info
    integer(c_int32_t) :: itime, ksp_iter, fsi_iter, status
    real(c_double) :: err_fsi
  end type ib_step_info

  type, bind(C) :: ib_sizes
    integer(c_int32_t) :: m, n, mgridlev, nbc
    integer(c_int32_t) :: n_nodes, n_xfaces, n_yfaces
    integer(c_int32_t) :: nb, nf, nbodies, ntors
  end type ib_sizes

  interface
    integer(c_int) function ib_create(p, nbodies, bodies, h) bind(C, name="ib_create")
      import :: c_int, c_int32_t, c_ptr, ib_params, ib_body
      type(ib_params), intent(in) :: p
      integer(c_int32_t), value :: nbodies
      type(ib_body), intent(in) :: bodies(*)
      type(c_ptr), intent(out) :: h
    end function ib_create

    integer(c_int) function ib_create_subdomain(p, sd, nbodies, bodies, h) bind(C, name="ib_create_subdomain")
      import :: c_int, c_int32_t, c_ptr, ib_params, ib_subdomain, ib_body
      type(ib_params), intent(in) :: p
      type(ib_subdomain), intent(in) :: sd
      integer(c_int32_t), value :: nbodies
      type(ib_body), intent(in) :: bodies(*)
      type(c_ptr), intent(out) :: h
    end function ib_create_subdomain

    integer(c_int) function ib_destroy(h) bind(C, name="ib_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function ib_destroy

    integer(c_int) function ib_get_sizes(h, s) bind(C, name="ib_get_sizes")
      import :: c_int, c_ptr, ib_sizes
      type(c_ptr), value :: h
      type(ib_sizes), intent(out) :: s
    end function ib_get_sizes

    integer(c_int) function ib_last_error(h, buf, blen) bind(C, name="ib_last_error")
      import :: c_int, c_int32_t, c_ptr, c_char
      type(c_ptr), value :: h
      character(kind=c_char), intent(out) :: buf(*)
      integer(c_int32_t), value :: blen
    end function ib_last_error

    integer(c_int) function ib_advance(h, itime, info) bind(C, name="ib_advance")
      import :: c_int, c_int32_t, c_ptr, ib_step_info
      type(c_ptr), value :: h
      integer(c_int32_t), intent(inout) :: itime
      type(ib_step_info), intent(out) :: info
    end function ib_advance

    integer(c_int) function ib_get_level_fields(h, k, omega, s, qx, qy, rhs_old) &
        bind(C, name="ib_get_level_fields")
      import :: c_int, c_int32_t, c_ptr, c_double
      type(c_ptr), value :: h
      integer(c_int32_t), value :: k
      real(c_double), intent(out) :: omega(*), s(*), qx(*), qy(*), rhs_old(*)
    end function ib_get_level_fields

    integer(c_int) function ib_set_level_fields(h, k, omega, s, qx, qy, rhs_old) &
        bind(C, name="ib_set_level_fields")
      import :: c_int, c_int32_t, c_ptr, c_double
      type(c_ptr), value :: h
      integer(c_int32_t), value :: k
      real(c_double), intent(in) :: omega(*), s(*), qx(*), qy(*), rhs_old(*)
    end function ib_set_level_fields

    integer(c_int) function ib_pressure(h, p) bind(C, name="ib_pressure")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(out) :: p(*)
    end function ib_pressure

    integer(c_int) function ib_set_body(h, f, xb, yb, theta, thetad, thetadd) bind(C, name="ib_set_body")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(in) :: f(*), xb(*), yb(*), theta(*), thetad(*), thetadd(*)
    end function ib_set_body

    integer(c_int) function ib_get_body(h, f, f_rdst, xb, yb, theta, thetad, thetadd) &
        bind(C, name="ib_get_body")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      real(c_double), intent(out) :: f(*), f_rdst(*), xb(*), yb(*), theta(*), thetad(*), thetadd(*)
    end function ib_get_body

    integer(c_int) function ib_get_forces(h, which, fx, fy) bind(C, name="ib_get_forces")
      import :: c_int, c_int32_t, c_ptr, c_double
      type(c_ptr), value :: h
      integer(c_int32_t), value :: which
      real(c_double), intent(out) :: fx(*), fy(*)
    end function ib_get_forces
  end interface

contains

  ! Message of the last failing call as a Fortran string.
  function ib_error_message(h) result(msg)
    type(c_ptr), intent(in) :: h
    character(len=512) :: msg
    character(kind=c_char) :: buf(512)
    integer(c_int) :: n, i
    n = ib_last_error(h, buf, 512_c_int32_t)
    msg = ' '
    do i = 1, min(int(n), 511)
      msg(i:i) = buf(i)
    end do
  end function ib_error_message

end module ib_gpu_mod
