!> The MONC-facing logic of the tvd_advection shim, written against MONC's own
!! names but free of MONC's modules, so that it compiles (and the drop-in
!! driver exercises it) without the model.  The component module of the shim
!! (the descriptor and callbacks, in the integration notes) only forwards
!! current_state's members to these routines:
!!
!!   tvd_shim_create(current_state%local_grid%size, current_state%number_q_fields,
!!        current_state%n_tracers, advect_flow, advect_th, advect_q,
!!        current_state%th%active, [(current_state%q(n)%active, n = 1, number_q_fields)],
!!        horizontal%cx, horizontal%cy, vertical%rdz, ..., vertical%tzd2, MONC_TVD_FP64, ctx)
!!
!! with advect_flow, advect_th and advect_q the selectors of
!! tvdadvection.F90:349-351 (determine_if_advection_here of the three
!! advection options).
module tvd_shim_core_mod
  use iso_c_binding, only : c_int, c_double, c_ptr
  use monc_tvd_interface_mod
  implicit none
  private

  !> Dimension indices of local_grid%size (model_core/src/grid/grids.F90:14).
  integer, parameter, public :: Z_INDEX = 1, Y_INDEX = 2, X_INDEX = 3

  public :: tvd_shim_selected, tvd_shim_flags, tvd_shim_create, tvd_shim_q_id, tvd_shim_tracer_id
contains

  !> Whether an advection option names this scheme: an empty value, "tvd" or
  !! "any" selects it (the rule of tvdadvection.F90:1218-1230).
  logical function tvd_shim_selected(option_value)
    character(len=*), intent(in) :: option_value

    tvd_shim_selected = len_trim(option_value) == 0 .or. trim(option_value) == "tvd" .or. &
         trim(option_value) == "any"
  end function tvd_shim_selected

  !> The library's advect flags for the three selectors.  The reference
  !! advects tracers when the *theta* option says so (advect_tracer is set
  !! from "advection_theta_field", tvdadvection.F90:352), so tracers follow
  !! advect_th.
  integer(c_int) function tvd_shim_flags(advect_flow, advect_th, advect_q)
    logical, intent(in) :: advect_flow, advect_th, advect_q

    tvd_shim_flags = 0
    if (advect_flow) tvd_shim_flags = tvd_shim_flags + MONC_TVD_ADVECT_FLOW
    if (advect_th) tvd_shim_flags = tvd_shim_flags + MONC_TVD_ADVECT_THETA + MONC_TVD_ADVECT_TRACER
    if (advect_q) tvd_shim_flags = tvd_shim_flags + MONC_TVD_ADVECT_Q
  end function tvd_shim_flags

  !> Library field id of q(n), n = 1..number_q_fields.
  integer(c_int) function tvd_shim_q_id(n)
    integer, intent(in) :: n
    tvd_shim_q_id = int(3 + n, c_int)
  end function tvd_shim_q_id

  !> Library field id of tracer(n), n = 1..n_tracers.
  integer(c_int) function tvd_shim_tracer_id(number_q_fields, n)
    integer, intent(in) :: number_q_fields, n
    tvd_shim_tracer_id = int(3 + number_q_fields + n, c_int)
  end function tvd_shim_tracer_id

  !> initialisation_callback (tvdadvection.F90:286-481) on the library: the
  !! context for the local grid, with the active flags of th and each q field
  !! (:661, :706).  local_size is local_grid%size, indexed by Z_INDEX,
  !! Y_INDEX and X_INDEX.  Returns the library status (0 = ok).
  integer(c_int) function tvd_shim_create(local_size, number_q_fields, n_tracers, advect_flow, advect_th, &
       advect_q, th_active, q_active, cx, cy, rdz, rdzn, dz, dzn, tzc1, tzc2, tzd1, tzd2, precision, ctx) &
       result(rc)
    integer, intent(in) :: local_size(3), number_q_fields, n_tracers
    logical, intent(in) :: advect_flow, advect_th, advect_q, th_active
    logical, intent(in) :: q_active(:)
    real(c_double), intent(in) :: cx, cy
    real(c_double), dimension(*), intent(in) :: rdz, rdzn, dz, dzn, tzc1, tzc2, tzd1, tzd2
    integer(c_int), intent(in) :: precision
    type(c_ptr), intent(out) :: ctx
    integer :: n

    rc = monc_tvd_create(int(local_size(Z_INDEX), c_int), int(local_size(Y_INDEX), c_int), &
         int(local_size(X_INDEX), c_int), 2_c_int, int(number_q_fields, c_int), int(n_tracers, c_int), &
         tvd_shim_flags(advect_flow, advect_th, advect_q), cx, cy, rdz, rdzn, dz, dzn, tzc1, tzc2, tzd1, tzd2, &
         precision, ctx)
    if (rc /= MONC_TVD_OK) return
    rc = monc_tvd_set_field_active(ctx, MONC_TVD_TH, merge(1_c_int, 0_c_int, th_active))
    if (rc /= MONC_TVD_OK) return
    do n = 1, min(number_q_fields, size(q_active))
      rc = monc_tvd_set_field_active(ctx, tvd_shim_q_id(n), merge(1_c_int, 0_c_int, q_active(n)))
      if (rc /= MONC_TVD_OK) return
    end do
  end function tvd_shim_create
end module tvd_shim_core_mod
