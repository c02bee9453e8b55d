!> iso_c_binding interface to libmonc_tvd.so (include/monc_tvd.h).
!!
!! This is the binding MONC's Fortran host uses to reach the gfx950 TVD
!! advection kernels.  It follows the reference's own C-interop precedent
!! (io/src/writers/io_state_writer.F90:26-60, io/src/forthread/forthread_ciface.F90):
!! bind(C) interfaces with value arguments and type(c_ptr) handles.
module monc_tvd_interface_mod
  use iso_c_binding, only : c_int, c_double, c_ptr, c_char, c_null_char, c_f_pointer, c_associated
  implicit none

  integer(c_int), parameter :: MONC_TVD_OK = 0
  integer(c_int), parameter :: MONC_TVD_CENTRED_STEPPING = 0, MONC_TVD_FORWARD_STEPPING = 1
  integer(c_int), parameter :: MONC_TVD_FP64 = 0, MONC_TVD_FP32 = 1
  integer(c_int), parameter :: MONC_TVD_ADVECT_FLOW = 1, MONC_TVD_ADVECT_THETA = 2, MONC_TVD_ADVECT_Q = 4, &
       MONC_TVD_ADVECT_TRACER = 8
  integer(c_int), parameter :: MONC_TVD_U = 0, MONC_TVD_V = 1, MONC_TVD_W = 2, MONC_TVD_TH = 3
  integer(c_int), parameter :: MONC_TVD_PROG = 0, MONC_TVD_ZFIELD = 1, MONC_TVD_SOURCE = 2
  integer(c_int), parameter :: MONC_TVD_ERR_ARG = 1, MONC_TVD_ERR_HIP = 2, MONC_TVD_ERR_STATE = 3
  integer(c_int), parameter :: MONC_TVD_DIAG_TENDENCY = 1, MONC_TVD_DIAG_ZFACE = 2
  integer(c_int), parameter :: MONC_TVD_DIAG_3D = 0, MONC_TVD_DIAG_PROFILE = 1, MONC_TVD_DIAG_TABS_3D = 2, &
       MONC_TVD_DIAG_TABS_PROFILE = 3, MONC_TVD_DIAG_ZFACES = 4
  integer(c_int), parameter :: MONC_TVD_COMPONENT_TVD = 0, MONC_TVD_COMPONENT_PW = 1, MONC_TVD_COMPONENT_THADV = 2

  interface
    integer(c_int) function monc_tvd_create(nz, ny, nx, halo, n_q, n_tracers, advect_flags, cx, cy, rdz, rdzn, &
         dz, dzn, tzc1, tzc2, tzd1, tzd2, precision, ctx) bind(C, name="monc_tvd_create")
      import :: c_int, c_double, c_ptr
      integer(c_int), value :: nz, ny, nx, halo, n_q, n_tracers, advect_flags, precision
      real(c_double), value :: cx, cy
      real(c_double), dimension(*), intent(in) :: rdz, rdzn, dz, dzn, tzc1, tzc2, tzd1, tzd2
      type(c_ptr), intent(out) :: ctx
    end function monc_tvd_create

    integer(c_int) function monc_tvd_set_field_active(ctx, field, active) bind(C, name="monc_tvd_set_field_active")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: field, active
    end function monc_tvd_set_field_active

    integer(c_int) function monc_tvd_destroy(ctx) bind(C, name="monc_tvd_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function monc_tvd_destroy

    integer(c_int) function monc_tvd_set_field(ctx, field, variant, host) bind(C, name="monc_tvd_set_field")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: field, variant
      real(c_double), dimension(*), intent(in) :: host
    end function monc_tvd_set_field

    integer(c_int) function monc_tvd_get_field(ctx, field, variant, host) bind(C, name="monc_tvd_get_field")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: field, variant
      real(c_double), dimension(*), intent(out) :: host
    end function monc_tvd_get_field

    integer(c_int) function monc_tvd_halo_fill(ctx) bind(C, name="monc_tvd_halo_fill")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function monc_tvd_halo_fill

    integer(c_int) function monc_tvd_tendencies(ctx, dtm, field_stepping, momentum_stepping, scalar_stepping) &
         bind(C, name="monc_tvd_tendencies")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double), value :: dtm
      integer(c_int), value :: field_stepping, momentum_stepping, scalar_stepping
    end function monc_tvd_tendencies

    integer(c_int) function monc_tvd_forward_steps(ctx, dtm, nsteps) bind(C, name="monc_tvd_forward_steps")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double), value :: dtm
      integer(c_int), value :: nsteps
    end function monc_tvd_forward_steps

    integer(c_int) function monc_tvd_leapfrog_steps(ctx, dtm, tsmth, ugal, vgal, nsteps, first_forward) &
        bind(C, name="monc_tvd_leapfrog_steps")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double), value :: dtm, tsmth, ugal, vgal
      integer(c_int), value :: nsteps, first_forward
    end function monc_tvd_leapfrog_steps

    ! th_advection's timestep callback (thadvection.F90:124-211) on the same context
    integer(c_int) function monc_thadv_tendencies(ctx, anelastic, dthref, baroclinicity, thref, thref0, fcoriol, &
        g, rate_x, rate_y, ugal, vgal) bind(C, name="monc_thadv_tendencies")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: anelastic, baroclinicity
      real(c_double), intent(in) :: dthref(*), thref(*)
      real(c_double), value :: thref0, fcoriol, g, rate_x, rate_y, ugal, vgal
    end function monc_thadv_tendencies

    ! pw_advection's timestep callback (pwadvection.F90:258-610) on the same context
    integer(c_int) function monc_pw_tendencies(ctx, advect_flow, advect_th, advect_q, tcx, tcy) &
        bind(C, name="monc_pw_tendencies")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: advect_flow, advect_th, advect_q
      real(c_double), value :: tcx, tcy
    end function monc_pw_tendencies

    integer(c_int) function monc_tvd_bind_field(ctx, field, variant, device_ptr) bind(C, name="monc_tvd_bind_field")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: field, variant
      type(c_ptr), value :: device_ptr
    end function monc_tvd_bind_field

    !> Tendency diagnostics (the published tend_* fields): rprefrcp may be
    !! c_null_ptr, else the address of a length-nz vector.
    integer(c_int) function monc_tvd_set_diagnostics(ctx, flags, rprefrcp) bind(C, name="monc_tvd_set_diagnostics")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, rprefrcp
      integer(c_int), value :: flags
    end function monc_tvd_set_diagnostics

    integer(c_int) function monc_tvd_get_diagnostic(ctx, field, which, host) bind(C, name="monc_tvd_get_diagnostic")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: field, which
      real(c_double), dimension(*), intent(out) :: host
    end function monc_tvd_get_diagnostic

    integer(c_int) function monc_tvd_get_component_diagnostic(ctx, component, field, which, host) &
        bind(C, name="monc_tvd_get_component_diagnostic")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: component, field, which
      real(c_double), dimension(*), intent(out) :: host
    end function monc_tvd_get_component_diagnostic

    integer(c_int) function monc_tvd_synchronize(ctx) bind(C, name="monc_tvd_synchronize")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function monc_tvd_synchronize

    type(c_ptr) function monc_tvd_last_error_c() bind(C, name="monc_tvd_last_error")
      import :: c_ptr
    end function monc_tvd_last_error_c
  end interface

contains

  !> The last error message as a Fortran string (for log_log(LOG_ERROR, ...))
  function monc_tvd_last_error() result(msg)
    character(len=:), allocatable :: msg
    character(kind=c_char), dimension(:), pointer :: chars
    type(c_ptr) :: p
    integer :: n

    p = monc_tvd_last_error_c()
    msg = ""
    if (.not. c_associated(p)) return
    call c_f_pointer(p, chars, [4096])
    n = 0
    do while (n < 4096)
      if (chars(n+1) == c_null_char) exit
      n = n + 1
    end do
    allocate(character(len=n) :: msg)
    do n = 1, len(msg)
      msg(n:n) = chars(n)
    end do
  end function monc_tvd_last_error
end module monc_tvd_interface_mod
