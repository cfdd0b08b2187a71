!=======================================================================
! specsim3d -- drop-in driver of the MI355X engine for SPECSIM3D's unconditional path.
!
! Keeps the reference program's process contract (src/specsim3d.f90 of stebond/specsim3d):
!   command line   specsim3d [file.par] [--gauss] [--bin] [--vtk]            (f90:1111-1141)
!   parameter file optional header up to 'START OF PARAMETERS', grid, model,
!                  seed line with optional igauss, base name, optional output
!                  codes and icond                                           (f90:1146-1232)
!   stdout         grid line, spectrum diagnostics, timers, one line per
!                  realization, in the reference's formats                  (f90:1254,1289-1300,1343)
!   files          GSLIB .out, raw float32 .bin, one .vti with every
!                  realization, binary/ASCII legacy .vtk                     (f90:307-391,1357-1414)
! The numerical work runs in libspecsim3d.so (HIP, gfx950) through the ISO_C_BINDING
! interfaces of module ss3d_capi: grid sizing, the FP64 spectral density, and the
! realizations (FP64 by default; --fp32 selects the FP32 half-spectrum path).
!
! Options beyond the reference: --fp32 (FP32 realizations), --no-grid-cap (periodic grids
! above the reference's 2^27-node limit).  Not part of this engine: conditioning (--cond,
! icond=1), which stops with a message.  --vario computes the experimental variograms on the GPU.
!=======================================================================
module ss3d_writers
  use iso_c_binding, only: c_double
  implicit none
  private
  public :: write_bin, write_vtk_binary, write_vti_file, write_gslib, write_vtk_ascii
  integer, parameter :: dp = c_double
  character(len=1), parameter :: lf = achar(10)

contains

  ! raw float32, realizations one after the other, x fastest (f90:1362-1376)
  subroutine write_bin(fname, res, n, nreal)
    character(len=*), intent(in) :: fname
    integer, intent(in) :: n, nreal
    real(dp), intent(in) :: res(n, nreal)
    integer :: u, r
    open(newunit=u, file=trim(fname), status='replace', action='write', access='stream', form='unformatted')
    do r = 1, nreal
      write(u) real(res(:, r), 4)
    end do
    close(u)
  end subroutine write_bin

  ! legacy VTK STRUCTURED_POINTS, BINARY big-endian float32 (f90:362-391)
  subroutine write_vtk_binary(fname, f, nx, ny, nz, org, spc)
    character(len=*), intent(in) :: fname
    integer, intent(in) :: nx, ny, nz
    real(dp), intent(in) :: f(nx*ny*nz), org(3), spc(3)
    character(len=256) :: ln
    integer :: u
    open(newunit=u, file=trim(fname), status='replace', access='stream', form='unformatted', convert='big_endian')
    write(u) '# vtk DataFile Version 3.0'//lf//'SPECSIM3D realization'//lf//'BINARY'//lf// &
             'DATASET STRUCTURED_POINTS'//lf
    write(ln, '(a,3(1x,i0))') 'DIMENSIONS', nx, ny, nz
    write(u) trim(ln)//lf
    write(ln, '(a,3(1x,es16.8))') 'ORIGIN', org
    write(u) trim(ln)//lf
    write(ln, '(a,3(1x,es16.8))') 'SPACING', spc
    write(u) trim(ln)//lf
    write(ln, '(a,i0)') 'POINT_DATA ', int(nx, 8)*int(ny, 8)*int(nz, 8)
    write(u) trim(ln)//lf//'SCALARS sim float 1'//lf//'LOOKUP_TABLE default'//lf
    write(u) real(f, 4)
    write(u) lf
    close(u)
  end subroutine write_vtk_binary

  ! one VTK XML ImageData file with every realization, raw appended little-endian float32 with
  ! UInt64 block sizes (f90:307-359)
  subroutine write_vti_file(fname, res, nx, ny, nz, nreal, org, spc)
    character(len=*), intent(in) :: fname
    integer, intent(in) :: nx, ny, nz, nreal
    real(dp), intent(in) :: res(nx*ny*nz, nreal), org(3), spc(3)
    character(len=512) :: ln
    character(len=48) :: ext
    integer(8) :: nbytes
    integer :: u, r
    integer(1) :: probe(4)
    probe = transfer(1_4, probe)
    if (probe(1) /= 1) stop 'write_vti: big-endian machines are not supported'
    nbytes = 4_8*int(nx, 8)*int(ny, 8)*int(nz, 8)
    write(ext, '(a,i0,a,i0,a,i0)') '0 ', nx - 1, ' 0 ', ny - 1, ' 0 ', nz - 1
    open(newunit=u, file=trim(fname), status='replace', access='stream', form='unformatted')
    write(u) '<?xml version="1.0"?>'//lf
    write(u) '<VTKFile type="ImageData" version="1.0" byte_order="LittleEndian" header_type="UInt64">'//lf
    write(ln, '(a,a,a,3(1x,es16.8),a,3(1x,es16.8),a)') '  <ImageData WholeExtent="', trim(ext), '" Origin="', org, &
      '" Spacing="', spc, '">'
    write(u) trim(ln)//lf
    write(ln, '(a,a,a)') '    <Piece Extent="', trim(ext), '">'
    write(u) trim(ln)//lf
    write(u) '      <PointData Scalars="sim_1">'//lf
    do r = 1, nreal
      write(ln, '(a,i0,a,i0,a)') '        <DataArray type="Float32" Name="sim_', r, '" format="appended" offset="', &
        (nbytes + 8_8)*int(r - 1, 8), '"/>'
      write(u) trim(ln)//lf
    end do
    write(u) '      </PointData>'//lf//'      <CellData/>'//lf//'    </Piece>'//lf//'  </ImageData>'//lf
    write(u) '  <AppendedData encoding="raw">'//lf//'   _'
    do r = 1, nreal
      write(u) nbytes
      write(u) real(res(:, r), 4)
    end do
    write(u) lf//'  </AppendedData>'//lf//'</VTKFile>'//lf
    close(u)
  end subroutine write_vti_file

  ! GSLIB grid file: title, nreal, one name per column, one row per node x fastest (f90:1384-1396)
  subroutine write_gslib(fname, res, n, nreal)
    character(len=*), intent(in) :: fname
    integer, intent(in) :: n, nreal
    real(dp), intent(in) :: res(n, nreal)
    integer :: u, r, p
    open(newunit=u, file=trim(fname), status='replace', action='write')
    write(u, '(a)') 'SPECSIM3D realizations'
    write(u, '(i4)') nreal
    do r = 1, nreal
      write(u, '(a,i0)') 'sim', r
    end do
    do p = 1, n
      write(u, '(*(es15.7))') res(p, :)
    end do
    close(u)
  end subroutine write_gslib

  ! legacy default: ASCII VTK of realization 1 (f90:1397-1413)
  subroutine write_vtk_ascii(fname, f, nx, ny, nz, org, spc)
    character(len=*), intent(in) :: fname
    integer, intent(in) :: nx, ny, nz
    real(dp), intent(in) :: f(nx, ny*nz), org(3), spc(3)
    integer :: u, row
    open(newunit=u, file=trim(fname), status='replace', action='write')
    write(u, '(a)') '# vtk DataFile Version 3.0'
    write(u, '(a)') 'SPECSIM3D realization 1'
    write(u, '(a)') 'ASCII'
    write(u, '(a)') 'DATASET STRUCTURED_POINTS'
    write(u, '(a,3i6)') 'DIMENSIONS', nx, ny, nz
    write(u, '(a,3es15.7)') 'ORIGIN', org
    write(u, '(a,3es15.7)') 'SPACING', spc
    write(u, '(a,i0)') 'POINT_DATA ', nx*ny*nz
    write(u, '(a)') 'SCALARS sim float 1'
    write(u, '(a)') 'LOOKUP_TABLE default'
    do row = 1, ny*nz
      write(u, '(*(es13.5))') f(:, row)
    end do
    close(u)
  end subroutine write_vtk_ascii

end module ss3d_writers


program specsim3d
  use iso_c_binding
  use ss3d_capi
  use ss3d_writers
  implicit none
  integer, parameter :: dp = c_double, maxst = 10
  ! run description
  character(len=256) :: parfile, outbase, fname, line, arg, varfile
  real(dp) :: org(3), spc(3), nugget
  integer :: nx, ny, nz, nst, iseed, nreal, ipad, igauss, icode, ival, nout, icond
  type(ss3d_structure_t) :: st(maxst)
  logical :: f_bin, f_vtk, f_cond, f_vario, f_fp32, f_nocap, legacy, found, exists
  logical :: o_gslib, o_vtk_all, o_bin, o_vti, o_vtk1_ascii, o_vtk1_bin
  integer :: ncli, i, k, ios, ia, na, ihead
  ! engine
  type(c_ptr) :: ctx, spec
  type(ss3d_spec_diag_t) :: diag
  integer(c_int) :: npx, npy, npz, prec
  integer(8) :: ncell
  real(dp), allocatable, target :: res(:, :)
  real(4), allocatable, target :: buf32(:, :)
  real(dp), allocatable :: zmean(:), zvar(:), zimag(:)
  integer :: r0, cnt, chunk
  integer(8) :: t0, t_spec, t_real, t_vario, t_end, rate

  call system_clock(t0, rate)
  parfile = 'SPECSIM3D.PAR'
  if (command_argument_count() >= 1) call get_command_argument(1, parfile)
  igauss = 0
  f_bin = .false.; f_vtk = .false.; f_cond = .false.; f_vario = .false.; f_fp32 = .false.; f_nocap = .false.
  ncli = 0
  na = command_argument_count()
  ia = 2
  do while (ia <= na)
    call get_command_argument(ia, arg)
    select case (trim(arg))
    case ('--gauss')
      igauss = 1
    case ('--bin')
      f_bin = .true.; ncli = ncli + 1
    case ('--vtk')
      f_vtk = .true.; ncli = ncli + 1
    case ('--vario')
      ia = ia + 1
      if (ia > na) stop '--vario needs a variogram parameter file'
      call get_command_argument(ia, varfile)
      f_vario = .true.
    case ('--cond')
      ia = ia + 1
      if (ia > na) stop '--cond needs a conditioning parameter file'
      f_cond = .true.; ncli = ncli + 1
    case ('--fp32')
      f_fp32 = .true.
    case ('--no-grid-cap')
      f_nocap = .true.
    case default
      write(*,*) 'unknown option: ', trim(arg); stop 1
    end select
    ia = ia + 1
  end do

  inquire(file=trim(parfile), exist=exists)
  if (.not. exists) then
    write(*,*) 'Parameter file not found: ', trim(parfile); stop 1
  end if
  open(10, file=trim(parfile), status='old', action='read')
  ! optional free-text header ended by a line containing START OF PARAMETERS (any case)
  found = .false.
  do ihead = 1, 30
    read(10, '(a)', iostat=ios) line
    if (ios /= 0) exit
    if (index(upper(line), 'START OF PARAMETERS') > 0) then
      found = .true.; exit
    end if
  end do
  if (.not. found) rewind(10)
  read(10, *) org
  read(10, *) spc
  read(10, *) nx, ny, nz
  read(10, *) nugget, nst
  if (nst < 1 .or. nst > maxst) stop 'nstruct out of range'
  do k = 1, nst
    read(10, *) st(k)%itype, st(k)%sill, st(k)%ahmax, st(k)%ahmin, st(k)%avert, st(k)%az, st(k)%dip, st(k)%plg
    if (st(k)%ahmax <= 0 .or. st(k)%ahmin <= 0 .or. st(k)%avert <= 0) stop 'ranges must be positive'
  end do
  ! seed line: iseed nreal ipad [igauss]
  read(10, '(a)') line
  read(line, *, iostat=ios) iseed, nreal, ipad, ival
  if (ios == 0) then
    if (ival == 1) igauss = 1
    if (ival < 0 .or. ival > 1) then
      write(*,*) 'parameter file: igauss must be 0 or 1'; stop 1
    end if
  else
    read(line, *) iseed, nreal, ipad
  end if
  read(10, '(a)') outbase
  outbase = adjustl(outbase)

  o_gslib = .false.; o_vtk_all = .false.; o_bin = .false.; o_vti = .false.
  o_vtk1_ascii = .false.; o_vtk1_bin = .false.
  legacy = .false.
  nout = 0
  read(10, *, iostat=ios) nout
  if (ios < 0) then
    legacy = .true.
  else if (ios > 0) then
    write(*,*) 'parameter file: cannot read the number of outputs (nout)'; stop 1
  end if
  if (.not. legacy) then
    if (nout < 1 .or. nout > 9) then
      write(*,*) 'parameter file: nout must be in 1..9'; stop 1
    end if
    do i = 1, nout
      read(10, *, iostat=ios) icode
      if (ios /= 0) then
        write(*,*) 'parameter file: missing output code'; stop 1
      end if
      select case (icode)
      case (1); o_gslib = .true.
      case (2); o_vtk_all = .true.
      case (3); o_bin = .true.
      case (4); o_vti = .true.
      case default
        write(*,*) 'parameter file: output code must be 1 (GSLIB), 2 (VTK), 3 (binary) or 4 (VTI)'; stop 1
      end select
    end do
    read(10, *, iostat=ios) icond
    if (ios /= 0) then
      write(*,*) 'parameter file: missing conditioning flag (icond)'; stop 1
    end if
    if (icond < 0 .or. icond > 1) then
      write(*,*) 'parameter file: icond must be 0 or 1'; stop 1
    end if
    f_cond = icond == 1
    if (ncli > 0) write(*,*) 'Note: --bin, --vtk and --cond are ignored: outputs and conditioning are set in the parameter file'
  else
    o_vtk_all = f_vtk
    if (.not. f_vtk) then
      if (f_bin) then
        o_bin = .true.; o_vtk1_bin = .true.
      else
        o_gslib = .true.; o_vtk1_ascii = .true.
      end if
    end if
  end if
  close(10)
  if (f_cond) then
    write(*,*) 'conditional simulation is not part of this engine (unconditional path only)'; stop 1
  end if

  ! ---- periodic grid (library: sizing rule of the reference)
  call ss3d_check(ss3d_periodic_grid(st, int(nst, c_int), spc(1), spc(2), spc(3), int(nx, c_int), int(ny, c_int), &
                                     int(nz, c_int), int(ipad, c_int), npx, npy, npz), 'periodic grid')
  if (.not. f_nocap .and. real(npx, dp)*real(npy, dp)*real(npz, dp) > 2.0_dp**27) &
    stop 'periodic grid too large (>2^27 nodes)'
  write(*, '(a,3i6,a,3i6)') ' SPECSIM3D  grid:', nx, ny, nz, '   periodic grid:', npx, npy, npz

  ! ---- spectral density on the GPU
  call ss3d_check(ss3d_context_create(0_c_int, ctx), 'GPU context')
  call ss3d_check(ss3d_spectrum_create(ctx, st, int(nst, c_int), nugget, spc(1), spc(2), spc(3), npx, npy, npz, &
                                       spec, diag), 'spectral density')
  write(*, '(a,es10.3)') ' max |Im S| / max |Re S|        : ', diag%max_imag/diag%max_real
  write(*, '(a,f10.5)')  ' target variance C(0)           : ', diag%c0
  write(*, '(a,f10.5)')  ' sum of positive spectrum       : ', diag%spos
  write(*, '(a,es10.3)') ' negative spectral mass / total : ', abs(diag%sneg)/(diag%spos - diag%sneg)
  if (igauss == 1) then
    write(*,*) 'spectral coefficients : Gaussian (exactly Gaussian field)'
  else
    write(*,*) 'spectral coefficients : fixed amplitudes, random phases (default)'
  end if
  call system_clock(t_spec)
  write(*, '(a,f8.2,a)') ' time covariance+spectrum: ', real(t_spec - t0)/real(rate), ' s'

  ! ---- realizations
  ncell = int(nx, 8)*int(ny, 8)*int(nz, 8)
  allocate(res(ncell, max(nreal, 1)), zmean(max(nreal, 1)), zvar(max(nreal, 1)), zimag(max(nreal, 1)))
  if (nreal > 0) then
    if (f_fp32) then
      prec = SS3D_FP32
      chunk = int(max(1_8, min(int(nreal, 8), (256_8*1024_8*1024_8)/max(ncell, 1_8))))
      allocate(buf32(ncell, chunk))
      do r0 = 0, nreal - 1, chunk
        cnt = min(chunk, nreal - r0)
        call ss3d_check(ss3d_realize(ctx, spec, int(iseed, c_int), int(igauss, c_int), int(r0, c_int64_t), &
                                     int(cnt, c_int), int(nx, c_int), int(ny, c_int), int(nz, c_int), prec, &
                                     c_loc(buf32), zmean(r0 + 1:), zvar(r0 + 1:), zimag(r0 + 1:)), 'realizations')
        res(:, r0 + 1:r0 + cnt) = real(buf32(:, 1:cnt), dp)
      end do
      deallocate(buf32)
    else
      prec = SS3D_FP64
      call ss3d_check(ss3d_realize(ctx, spec, int(iseed, c_int), int(igauss, c_int), 0_c_int64_t, int(nreal, c_int), &
                                   int(nx, c_int), int(ny, c_int), int(nz, c_int), prec, c_loc(res), zmean, zvar, &
                                   zimag), 'realizations')
    end if
  end if
  do i = 1, nreal
    write(*, '(a,i3,a,f9.4,a,f9.4,a,es9.2)') ' realization', i, '  mean=', zmean(i), '  variance=', zvar(i), &
      '  max|Im|=', zimag(i)
  end do
  i = ss3d_spectrum_destroy(spec)
  call system_clock(t_real)
  write(*, '(a,f8.2,a)') ' time realizations         : ', real(t_real - t_spec)/real(rate), ' s'
  if (f_vario) then
    call vario_report(trim(varfile))
    call system_clock(t_vario)
    write(*, '(a,f8.2,a)') ' time variograms           : ', real(t_vario - t_real)/real(rate), ' s'
  end if
  i = ss3d_context_destroy(ctx)

  ! ---- outputs (same files and order as the reference)
  if (o_bin) call write_bin(trim(outbase)//'.bin', res, int(ncell), nreal)
  if (o_vtk_all .or. o_vtk1_bin) then
    do i = 1, nreal
      if (o_vtk_all .or. i == 1) then
        write(fname, '(a,a,i0,a)') trim(outbase), '_', i, '.vtk'
        call write_vtk_binary(fname, res(:, i), nx, ny, nz, org, spc)
      end if
    end do
  end if
  if (o_vti) call write_vti_file(trim(outbase)//'.vti', res, nx, ny, nz, nreal, org, spc)
  if (o_gslib) call write_gslib(trim(outbase)//'.out', res, int(ncell), nreal)
  if (o_vtk1_ascii) call write_vtk_ascii(trim(outbase)//'_1.vtk', res(:, 1), nx, ny, nz, org, spc)

  call system_clock(t_end)
  write(*, '(a,f8.2,a)') ' elapsed time: ', real(t_end - t0)/real(rate), ' s'

contains

  ! Experimental variograms of every realization along the directions of the variogram file,
  ! their ensemble mean and sd and the model variogram: the report and the <base>_vario.csv table
  ! of vario_report (f90:434-529), with the pair sums computed on the GPU (ss3d_variogram).
  ! File: ndir iperreal / ndir lines sx sy sz nlag (lag k = k*(sx,sy,sz) cells).
  subroutine vario_report(vfile)
    character(len=*), intent(in) :: vfile
    integer :: ndir, iper, id, il, r, nlagmax, nfit, u
    integer(c_int), allocatable :: dirs(:, :)
    real(dp), allocatable, target :: gam(:, :, :)
    real(dp) :: c0, gm, gsd, gth, h, rms, mx, e, hx, hy, hz
    integer(8) :: np
    character(len=512) :: vname
    open(newunit=u, file=vfile, status='old', action='read', iostat=ios)
    if (ios /= 0) then
      write(*,*) 'cannot open variogram file: ', vfile; return
    end if
    read(u, *) ndir, iper
    allocate(dirs(4, ndir))
    do id = 1, ndir
      read(u, *) dirs(1, id), dirs(2, id), dirs(3, id), dirs(4, id)
    end do
    close(u)
    nlagmax = max(1, maxval(dirs(4, :)))
    allocate(gam(nlagmax, ndir, max(nreal, 1)))
    gam = 0.0_dp
    if (nreal > 0) call ss3d_check(ss3d_variogram(ctx, c_loc(res), 0_c_int, SS3D_FP64, int(nx, c_int), &
                                   int(ny, c_int), int(nz, c_int), int(nreal, c_int), int(ndir, c_int), &
                                   dirs, gam), 'variograms')
    c0 = nugget
    do i = 1, nst
      c0 = c0 + st(i)%sill
    end do
    vname = trim(outbase)//'_vario.csv'
    open(newunit=u, file=trim(vname), status='replace', action='write')
    write(u, '(a)', advance='no') 'dir,sx,sy,sz,lag,h,npairs,gamma_model,gamma_mean,gamma_sd'
    if (iper == 1) then
      do r = 1, nreal
        write(u, '(a,i0)', advance='no') ',gamma_r', r
      end do
    end if
    write(u, '(a)') ''
    write(*, '(a)') ' variograms (Fortran): direction  lags  RMS(mean-model)  max|err|   [rising part, ' // &
                    'gamma_model < 0.95 sill]'
    do id = 1, ndir
      rms = 0.0_dp; mx = 0.0_dp; nfit = 0
      do il = 1, dirs(4, id)
        np = ss3d_npairs(int(nx, c_int), int(ny, c_int), int(nz, c_int), int(il*dirs(1, id), c_int), &
                         int(il*dirs(2, id), c_int), int(il*dirs(3, id), c_int))
        if (np == 0) cycle
        hx = il*dirs(1, id)*spc(1); hy = il*dirs(2, id)*spc(2); hz = il*dirs(3, id)*spc(3)
        h = sqrt(hx*hx + hy*hy + hz*hz)
        gth = c0 - ss3d_covariance(st, int(nst, c_int), nugget, hx, hy, hz)
        gm = sum(gam(il, id, 1:nreal))/real(nreal, dp)
        if (nreal > 1) then
          gsd = sqrt(sum((gam(il, id, 1:nreal) - gm)**2)/real(nreal - 1, dp))
        else
          gsd = 0.0_dp
        end if
        write(u, '(i0,3(",",i0),",",i0,",",es14.6,",",i0,3(",",es14.6))', advance='no') &
              id, dirs(1, id), dirs(2, id), dirs(3, id), il, h, np, gth, gm, gsd
        if (iper == 1) then
          do r = 1, nreal
            write(u, '(",",es14.6)', advance='no') gam(il, id, r)
          end do
        end if
        write(u, '(a)') ''
        if (gth < 0.95_dp*c0) then
          e = gm - gth
          rms = rms + e*e; mx = max(mx, abs(e)); nfit = nfit + 1
        end if
      end do
      if (nfit > 0) rms = sqrt(rms/real(nfit, dp))
      write(*, '(a,i3,3i4,i6,2f12.5)') '   dir', id, dirs(1, id), dirs(2, id), dirs(3, id), dirs(4, id), rms, mx
    end do
    close(u)
    write(*, '(a,a)') ' variogram table: ', trim(vname)
  end subroutine vario_report

  function upper(s) result(u)
    character(len=*), intent(in) :: s
    character(len=len(s)) :: u
    integer :: j
    u = s
    do j = 1, len_trim(u)
      if (u(j:j) >= 'a' .and. u(j:j) <= 'z') u(j:j) = achar(iachar(u(j:j)) - 32)
    end do
  end function upper

end program specsim3d
