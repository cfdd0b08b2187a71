!=======================================================================
! ss3d_capi -- ISO_C_BINDING interface of libspecsim3d.so (include/specsim3d.h).
!
! The Fortran driver (specsim3d.f90) reaches the HIP engine only through these interfaces:
! plain C types, handles as c_ptr, integer return codes (0 = success, message from
! ss3d_last_error).
!=======================================================================
module ss3d_capi
  use iso_c_binding
  implicit none
  private
  public :: ss3d_structure_t, ss3d_spec_diag_t, ss3d_outputs_t, ss3d_cond_columns_t, ss3d_cond_params_t, &
            ss3d_cond_info_t, SS3D_FP64, SS3D_FP32
  public :: ss3d_context_create, ss3d_context_destroy, ss3d_periodic_grid, ss3d_spectrum_create, &
            ss3d_spectrum_destroy, ss3d_realize, ss3d_draws_per_realization, ss3d_check, &
            ss3d_variogram, ss3d_npairs, ss3d_covariance, ss3d_writer_open, ss3d_writer_slot, &
            ss3d_writer_realize, ss3d_writer_close, ss3d_cond_prepare, ss3d_cond_create, ss3d_cond_info_get, &
            ss3d_cond_destroy, ss3d_writer_condition, ss3d_error_has

  integer(c_int), parameter :: SS3D_FP64 = 0, SS3D_FP32 = 1

  ! ss3d_structure: scalar members of the reference's `type structure`
  type, bind(C) :: ss3d_structure_t
    integer(c_int32_t) :: itype = 1
    integer(c_int32_t) :: pad_ = 0
    real(c_double) :: sill = 0, ahmax = 1, ahmin = 1, avert = 1, az = 0, dip = 0, plg = 0
  end type

  type, bind(C) :: ss3d_spec_diag_t
    real(c_double) :: max_imag = 0, max_real = 0, c0 = 0, spos = 0, sneg = 0
  end type

  ! ss3d_outputs: the files of one run (NUL-terminated paths, c_null_ptr = not written)
  type, bind(C) :: ss3d_outputs_t
    type(c_ptr) :: bin_path = c_null_ptr, vti_path = c_null_ptr, vtk_base = c_null_ptr
    integer(c_int32_t) :: vtk_count = 0, reserved = 0
    type(c_ptr) :: gslib_path = c_null_ptr, vtk_ascii_path = c_null_ptr
    real(c_double) :: origin(3) = 0, spacing(3) = 1
  end type

  ! conditioning (cond_mod, f90:548-1083): data columns, search and transform, what is printed
  type, bind(C) :: ss3d_cond_columns_t
    integer(c_int32_t) :: ixc = 1, iyc = 2, izc = 0, ivc = 3, iwc = 0, itrans = 0
    real(c_double) :: tmin = -1.0d21, tmax = 1.0d21, skmean = 0
  end type

  type, bind(C) :: ss3d_cond_params_t
    real(c_double) :: rmax = 1, rmed = 1, rvert = 1, sa1 = 0, sa2 = 0, sa3 = 0
    integer(c_int32_t) :: ndmax = 16, itrans = 0
    real(c_double) :: skmean = 0
    integer(c_int32_t) :: nt = 0, ltail = 1, utail = 1, reserved = 0
    type(c_ptr) :: vr = c_null_ptr, gt = c_null_ptr
    real(c_double) :: zlo = 0, zhi = 0, ltpar = 1, utpar = 1
  end type

  type, bind(C) :: ss3d_cond_info_t
    integer(c_int32_t) :: nsbx = 0, nsby = 0, nsbz = 0, nnone = 0, nexact = 0, nfail = 0
    integer(c_int64_t) :: nsum = 0
    real(c_double) :: weights_s = 0, apply_s = 0, backtr_s = 0
  end type

  interface
    integer(c_int) function ss3d_context_create(device, ctx) bind(C, name='ss3d_context_create')
      import :: c_int, c_ptr
      integer(c_int), value :: device
      type(c_ptr), intent(out) :: ctx
    end function

    integer(c_int) function ss3d_context_destroy(ctx) bind(C, name='ss3d_context_destroy')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function

    integer(c_int) function ss3d_periodic_grid(st, nst, dx, dy, dz, nx, ny, nz, ipad, npx, npy, npz) &
        bind(C, name='ss3d_periodic_grid')
      import :: c_int, c_double, ss3d_structure_t
      type(ss3d_structure_t), intent(in) :: st(*)
      integer(c_int), value :: nst, nx, ny, nz, ipad
      real(c_double), value :: dx, dy, dz
      integer(c_int), intent(out) :: npx, npy, npz
    end function

    integer(c_int) function ss3d_spectrum_create(ctx, st, nst, nugget, dx, dy, dz, npx, npy, npz, spec, diag) &
        bind(C, name='ss3d_spectrum_create')
      import :: c_int, c_double, c_ptr, ss3d_structure_t, ss3d_spec_diag_t
      type(c_ptr), value :: ctx
      type(ss3d_structure_t), intent(in) :: st(*)
      integer(c_int), value :: nst, npx, npy, npz
      real(c_double), value :: nugget, dx, dy, dz
      type(c_ptr), intent(out) :: spec
      type(ss3d_spec_diag_t), intent(out) :: diag
    end function

    integer(c_int) function ss3d_spectrum_destroy(spec) bind(C, name='ss3d_spectrum_destroy')
      import :: c_int, c_ptr
      type(c_ptr), value :: spec
    end function

    integer(c_int64_t) function ss3d_draws_per_realization(npx, npy, npz, igauss) &
        bind(C, name='ss3d_draws_per_realization')
      import :: c_int, c_int64_t
      integer(c_int), value :: npx, npy, npz, igauss
    end function

    integer(c_int) function ss3d_realize(ctx, spec, iseed, igauss, r0, nreal, nx, ny, nz, precision, host_out, &
        mean, var, maximag) bind(C, name='ss3d_realize')
      import :: c_int, c_int64_t, c_double, c_ptr
      type(c_ptr), value :: ctx, spec, host_out
      integer(c_int), value :: iseed, igauss, nreal, nx, ny, nz, precision
      integer(c_int64_t), value :: r0
      real(c_double), intent(out) :: mean(*), var(*), maximag(*)
    end function

    ! variograms (vario_lag f90:405-431): gam(nlagmax, ndir, nreal) in Fortran order
    integer(c_int) function ss3d_variogram(ctx, field, on_device, precision, nx, ny, nz, nreal, ndir, dirs, gam) &
        bind(C, name='ss3d_variogram')
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx, field
      integer(c_int), value :: on_device, precision, nx, ny, nz, nreal, ndir
      integer(c_int), intent(in) :: dirs(*)
      real(c_double), intent(out) :: gam(*)
    end function

    integer(c_int64_t) function ss3d_npairs(nx, ny, nz, lx, ly, lz) bind(C, name='ss3d_npairs')
      import :: c_int, c_int64_t
      integer(c_int), value :: nx, ny, nz, lx, ly, lz
    end function

    real(c_double) function ss3d_covariance(st, nst, nugget, hx, hy, hz) bind(C, name='ss3d_covariance')
      import :: c_int, c_double, ss3d_structure_t
      type(ss3d_structure_t), intent(in) :: st(*)
      integer(c_int), value :: nst
      real(c_double), value :: nugget, hx, hy, hz
    end function

    ! streaming output writer (the reference's writers, f90:307-391, 1357-1414)
    integer(c_int) function ss3d_writer_open(ctx, o, nx, ny, nz, nreal, precision, w) bind(C, name='ss3d_writer_open')
      import :: c_int, c_ptr, ss3d_outputs_t
      type(c_ptr), value :: ctx
      type(ss3d_outputs_t), intent(in) :: o
      integer(c_int), value :: nx, ny, nz, nreal, precision
      type(c_ptr), intent(out) :: w
    end function

    integer(c_int) function ss3d_writer_slot(w) bind(C, name='ss3d_writer_slot')
      import :: c_int, c_ptr
      type(c_ptr), value :: w
    end function

    integer(c_int) function ss3d_writer_realize(w, spec, iseed, igauss, nreal, mean, var, maximag, dev_field) &
        bind(C, name='ss3d_writer_realize')
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: w, spec
      integer(c_int), value :: iseed, igauss, nreal
      real(c_double), intent(out) :: mean(*), var(*), maximag(*)
      type(c_ptr), intent(out) :: dev_field
    end function

    integer(c_int) function ss3d_writer_close(w) bind(C, name='ss3d_writer_close')
      import :: c_int, c_ptr
      type(c_ptr), value :: w
    end function

    integer(c_int) function ss3d_cond_prepare(rows, nrow, ncol, cols, nx, ny, nz, xmin, ymin, zmin, dx, dy, dz, &
        nd, nmerged, nout, ijk, dval, dg, vr, gt) bind(C, name='ss3d_cond_prepare')
      import :: c_int, c_int32_t, c_double, ss3d_cond_columns_t
      real(c_double), intent(in) :: rows(*)
      integer(c_int), value :: nrow, ncol, nx, ny, nz
      type(ss3d_cond_columns_t), intent(in) :: cols
      real(c_double), value :: xmin, ymin, zmin, dx, dy, dz
      integer(c_int32_t), intent(out) :: nd, nmerged, nout
      integer(c_int32_t), intent(out) :: ijk(*)
      real(c_double), intent(out) :: dval(*), dg(*), vr(*), gt(*)
    end function

    integer(c_int) function ss3d_cond_create(ctx, st, nst, nugget, nx, ny, nz, xmin, ymin, zmin, dx, dy, dz, nd, &
        ijk, g, p, cond, info) bind(C, name='ss3d_cond_create')
      import :: c_int, c_int32_t, c_double, c_ptr, ss3d_structure_t, ss3d_cond_params_t, ss3d_cond_info_t
      type(c_ptr), value :: ctx
      type(ss3d_structure_t), intent(in) :: st(*)
      integer(c_int), value :: nst, nx, ny, nz
      integer(c_int32_t), value :: nd
      real(c_double), value :: nugget, xmin, ymin, zmin, dx, dy, dz
      integer(c_int32_t), intent(in) :: ijk(*)
      real(c_double), intent(in) :: g(*)
      type(ss3d_cond_params_t), intent(in) :: p
      type(c_ptr), intent(out) :: cond
      type(ss3d_cond_info_t), intent(out) :: info
    end function

    integer(c_int) function ss3d_cond_info_get(cond, info) bind(C, name='ss3d_cond_info_get')
      import :: c_int, c_ptr, ss3d_cond_info_t
      type(c_ptr), value :: cond
      type(ss3d_cond_info_t), intent(out) :: info
    end function

    integer(c_int) function ss3d_cond_destroy(cond) bind(C, name='ss3d_cond_destroy')
      import :: c_int, c_ptr
      type(c_ptr), value :: cond
    end function

    integer(c_int) function ss3d_writer_condition(w, cond) bind(C, name='ss3d_writer_condition')
      import :: c_int, c_ptr
      type(c_ptr), value :: w, cond
    end function

    integer(c_int) function ss3d_last_error_c(buf, length) bind(C, name='ss3d_last_error')
      import :: c_int, c_int64_t, c_char
      character(kind=c_char), intent(out) :: buf(*)
      integer(c_int64_t), value :: length
    end function
  end interface

contains

  ! True when the library's last error message on this thread contains s.
  logical function ss3d_error_has(s)
    character(len=*), intent(in) :: s
    character(kind=c_char) :: buf(512)
    character(len=512) :: msg
    integer :: i, rc
    buf = c_null_char
    rc = ss3d_last_error_c(buf, int(size(buf), c_int64_t))
    msg = ''
    do i = 1, size(buf)
      if (buf(i) == c_null_char) exit
      msg(i:i) = buf(i)
    end do
    ss3d_error_has = index(msg, s) > 0
  end function ss3d_error_has

  ! Stops the program with the library's message when rc is not 0.
  subroutine ss3d_check(rc, what)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: what
    character(kind=c_char) :: buf(512)
    character(len=512) :: msg
    integer :: i, rc2
    if (rc == 0) return
    buf = c_null_char
    rc2 = ss3d_last_error_c(buf, int(size(buf), c_int64_t))
    msg = ''
    do i = 1, size(buf)
      if (buf(i) == c_null_char) exit
      msg(i:i) = buf(i)
    end do
    write(*,'(a,a,a,i0,a,a)') ' ', what, ' failed (code ', rc, '): ', trim(msg)
    stop 1
  end subroutine ss3d_check

end module ss3d_capi
