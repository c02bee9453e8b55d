!> Fortran host driver of the drop-in path: reads one MONC local state from a
!! raw file, creates the context through the shim core (tvd_shim_create, as
!! the component's initialisation callback does), runs the TVD timestep
!! through the C ABI as the component's timestep callback would (halo swap,
!! then tendencies into zeroed source fields), and writes the source fields
!! back out.
!!
!! Input file (stream, native endian): 13 int32 (local_grid%size in Z_INDEX,
!! Y_INDEX, X_INDEX order, number_q_fields, n_tracers, field_stepping,
!! momentum_stepping, scalar_stepping, precision, advect_flow, advect_th,
!! advect_q, th%active), number_q_fields int32 (q(n)%active), 3 float64 (cx,
!! cy, dtm), the 8 metric vectors (rdz, rdzn, dz, dzn, tzc1, tzc2, tzd1, tzd2;
!! nz each), then the prognostics and the z-fields of u, v, w, th, q(:),
!! tracer(:), each (nz, ny+4, nx+4) float64.
!! Output file: the source fields in the same layout and order, then for u,
!! v, w, th and each q the TVD tendency diagnostics as the component's
!! field_value_retrieval would hand them out (tend_*_tvdadvection_3d_local,
!! (nz, ny, nx), and _profile_total_local, nz), zeros for a field that was
!! not advected (no diagnostic recorded).
program tvd_dropin_driver
  use iso_c_binding, only : c_int, c_double, c_ptr, c_null_ptr
  use monc_tvd_interface_mod
  use tvd_shim_core_mod
  implicit none

  character(len=4096) :: in_path, out_path
  integer(c_int) :: hdr(13), fs, ms, ss, prec, n
  integer :: local_size(3), nq, ntr, nf
  integer(c_int), allocatable :: qact(:)
  integer(c_int) :: rc
  real(c_double) :: rhdr(3)
  real(c_double), allocatable :: met(:,:), fld(:,:,:), zero(:,:,:), t3d(:,:,:), prof(:)
  type(c_ptr) :: ctx
  integer :: u_in, u_out

  call get_command_argument(1, in_path)
  call get_command_argument(2, out_path)
  open(newunit=u_in, file=trim(in_path), access="stream", form="unformatted", status="old", action="read")
  read(u_in) hdr
  local_size = hdr(1:3)
  nq = hdr(4); ntr = hdr(5)
  fs = hdr(6); ms = hdr(7); ss = hdr(8); prec = hdr(9)
  allocate(qact(nq))
  if (nq > 0) read(u_in) qact
  read(u_in) rhdr
  nf = 4 + nq + ntr
  allocate(met(local_size(Z_INDEX), 8), fld(local_size(Z_INDEX), local_size(Y_INDEX)+4, local_size(X_INDEX)+4))
  allocate(zero(local_size(Z_INDEX), local_size(Y_INDEX)+4, local_size(X_INDEX)+4))
  read(u_in) met
  zero = 0.0_c_double

  call check(tvd_shim_create(local_size, nq, ntr, hdr(10) /= 0, hdr(11) /= 0, hdr(12) /= 0, hdr(13) /= 0, &
       qact /= 0, rhdr(1), rhdr(2), met(:,1), met(:,2), met(:,3), met(:,4), met(:,5), met(:,6), met(:,7), &
       met(:,8), prec, ctx), "tvd_shim_create")
  do n = 0, nf - 1
    read(u_in) fld
    call check(monc_tvd_set_field(ctx, n, MONC_TVD_PROG, fld), "set prognostic")
  end do
  do n = 0, nf - 1
    read(u_in) fld
    call check(monc_tvd_set_field(ctx, n, MONC_TVD_ZFIELD, fld), "set z-field")
    ! clearsourceterms zeroes every source field before the dynamics group
    call check(monc_tvd_set_field(ctx, n, MONC_TVD_SOURCE, zero), "clear source")
  end do
  close(u_in)

  call check(monc_tvd_halo_fill(ctx), "halo fill")
  call check(monc_tvd_set_diagnostics(ctx, MONC_TVD_DIAG_TENDENCY, c_null_ptr), "diagnostics")
  call check(monc_tvd_tendencies(ctx, rhdr(3), fs, ms, ss), "tendencies")

  open(newunit=u_out, file=trim(out_path), access="stream", form="unformatted", status="replace", action="write")
  do n = 0, nf - 1
    call check(monc_tvd_get_field(ctx, n, MONC_TVD_SOURCE, fld), "get source")
    write(u_out) fld
  end do
  allocate(t3d(local_size(Z_INDEX), local_size(Y_INDEX), local_size(X_INDEX)), prof(local_size(Z_INDEX)))
  do n = 0, 3 + nq
    rc = monc_tvd_get_diagnostic(ctx, n, MONC_TVD_DIAG_3D, t3d)
    if (rc == MONC_TVD_ERR_STATE) then
      t3d = 0.0_c_double
      prof = 0.0_c_double
    else
      call check(rc, "tend_3d")
      call check(monc_tvd_get_diagnostic(ctx, n, MONC_TVD_DIAG_PROFILE, prof), "profile total")
    end if
    write(u_out) t3d
    write(u_out) prof
  end do
  close(u_out)
  call check(monc_tvd_destroy(ctx), "destroy")
  print '(a)', "tvd_dropin_driver: ok"

contains

  subroutine check(status, what)
    integer(c_int), intent(in) :: status
    character(len=*), intent(in) :: what

    if (status /= MONC_TVD_OK) then
      print '(a,a,a,i0,a,a)', "tvd_dropin_driver: ", what, " failed with status ", status, ": ", &
           monc_tvd_last_error()
      error stop 1
    end if
  end subroutine check
end program tvd_dropin_driver
