!> Fortran host driver of the drop-in path: reads one MONC local state from a
!! raw file, runs the TVD timestep through the C ABI exactly as the component's
!! timestep callback would (halo swap, then tendencies into zeroed source
!! fields), and writes the source fields back out.
!!
!! Input file (stream, native endian): 8 int32 (nz, ny, nx, n_q, field_stepping,
!! momentum_stepping, scalar_stepping, precision), 3 float64 (cx, cy, dtm), the 8
!! metric vectors (rdz, rdzn, dz, dzn, tzc1, tzc2, tzd1, tzd2; nz each), then
!! the 4+n_q prognostics and the 4+n_q z-fields, each (nz, ny+4, nx+4) float64.
!! Output file: the 4+n_q source fields in the same layout.
program tvd_dropin_driver
  use iso_c_binding, only : c_int, c_double, c_ptr
  use monc_tvd_interface_mod
  implicit none

  character(len=4096) :: in_path, out_path
  integer(c_int) :: hdr(8), nz, ny, nx, n_q, fs, ms, ss, prec, rc, n
  real(c_double) :: rhdr(3)
  real(c_double), allocatable :: met(:,:), fld(:,:,:), zero(:,:,:)
  type(c_ptr) :: ctx
  integer :: u_in, u_out

  call get_command_argument(1, in_path)
  call get_command_argument(2, out_path)
  open(newunit=u_in, file=trim(in_path), access="stream", form="unformatted", status="old", action="read")
  read(u_in) hdr
  read(u_in) rhdr
  nz = hdr(1); ny = hdr(2); nx = hdr(3); n_q = hdr(4)
  fs = hdr(5); ms = hdr(6); ss = hdr(7); prec = hdr(8)
  allocate(met(nz, 8), fld(nz, ny+4, nx+4), zero(nz, ny+4, nx+4))
  read(u_in) met
  zero = 0.0_c_double

  rc = monc_tvd_create(nz, ny, nx, 2_c_int, n_q, MONC_TVD_ADVECT_FLOW + MONC_TVD_ADVECT_THETA + MONC_TVD_ADVECT_Q, &
       rhdr(1), rhdr(2), met(:,1), met(:,2), met(:,3), met(:,4), met(:,5), met(:,6), met(:,7), met(:,8), prec, ctx)
  call check(rc, "monc_tvd_create")
  do n = 0, 3 + n_q
    read(u_in) fld
    call check(monc_tvd_set_field(ctx, n, MONC_TVD_PROG, fld), "set prognostic")
  end do
  do n = 0, 3 + n_q
    read(u_in) fld
    call check(monc_tvd_set_field(ctx, n, MONC_TVD_ZFIELD, fld), "set z-field")
    ! clearsourceterms zeroes every source field before the dynamics group
    call check(monc_tvd_set_field(ctx, n, MONC_TVD_SOURCE, zero), "clear source")
  end do
  close(u_in)

  call check(monc_tvd_halo_fill(ctx), "halo fill")
  call check(monc_tvd_tendencies(ctx, rhdr(3), fs, ms, ss), "tendencies")

  open(newunit=u_out, file=trim(out_path), access="stream", form="unformatted", status="replace", action="write")
  do n = 0, 3 + n_q
    call check(monc_tvd_get_field(ctx, n, MONC_TVD_SOURCE, fld), "get source")
    write(u_out) fld
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
