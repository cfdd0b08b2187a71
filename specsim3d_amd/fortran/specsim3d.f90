!=======================================================================
! specsim3d -- drop-in driver of the MI355X engine for SPECSIM3D's unconditional path.
!
! Keeps the reference program's process contract (src/specsim3d.f90 of stebond/specsim3d):
!   command line   specsim3d [file.par] [--gauss] [--bin] [--vtk] [--vario f] (f90:1111-1141)
!   parameter file optional header up to 'START OF PARAMETERS', grid, model,
!                  seed line with optional igauss, base name, optional output
!                  codes and icond                                           (f90:1146-1232)
!   stdout         grid line, spectrum diagnostics, timers, one line per
!                  realization, in the reference's formats                  (f90:1254,1289-1300,1343)
!   files          GSLIB .out, raw float32 .bin, one .vti with every
!                  realization, binary/ASCII legacy .vtk                     (f90:307-391,1357-1414)
! The numerical work runs in libspecsim3d.so (HIP, gfx950) through the ISO_C_BINDING
! interfaces of module ss3d_capi: grid sizing, the FP64 spectral density, the realizations
! (FP64 by default; --fp32 selects the FP32 half-spectrum path) and the output files, which the
! library's writer streams to disk while later realizations are computed (ss3d_writer_*).
!
! Options beyond the reference: --fp32 (FP32 realizations), --no-grid-cap (periodic grids
! above the reference's 2^27-node limit).  --vario computes the experimental variograms on the
! GPU, chunk by chunk on the device copies of the realizations.  Conditioning (--cond file, or
! icond=1 with the block in the parameter file; f90:700-1004): the driver reads the block and the
! GSLIB data file, the library prepares the data (ss3d_cond_prepare), computes the kriging weights
! of every node once and conditions each chunk of realizations on the device before it is written.
!=======================================================================
program specsim3d
  use iso_c_binding
  use ss3d_capi
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
  real(dp), allocatable :: zmean(:), zvar(:), zimag(:)
  integer :: r0, cnt, chunk
  integer(8) :: t0, t_spec, t_real, t_vario, t_end, rate, tv0, tv1, tvario
  ! conditioning (block of f90:731-747 and the data file)
  character(len=256) :: condfile
  character(len=512) :: datafile, cline
  type(ss3d_cond_columns_t) :: ccols
  type(ss3d_cond_params_t) :: cpar
  type(ss3d_cond_info_t) :: cinfo
  type(c_ptr) :: cond
  integer :: ciu, nvar, nraw, cnd
  integer(c_int32_t) :: c_nd, c_nmerged, c_nout
  real(dp), allocatable :: crows(:, :), crow(:)
  real(dp), allocatable, target :: c_dval(:), c_dg(:), c_vr(:), c_gt(:)
  integer(c_int32_t), allocatable :: c_ijk(:, :)
  ! output files (written by the library's streaming writer)
  type(c_ptr) :: wr, dfield
  type(ss3d_outputs_t), target :: outs
  character(kind=c_char), target :: p_bin(512), p_vti(512), p_vtk(512), p_out(512), p_vtka(512)
  ! variograms: directions read before the realizations, pair sums taken per chunk on the device
  logical :: vario_ok
  integer :: ndir, iper
  integer(c_int), allocatable :: dirs(:, :)
  real(dp), allocatable :: gam(:, :, :)

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
      call get_command_argument(ia, condfile)
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
  if (f_cond) call cond_read()
  close(10)

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

  ! ---- output files: every header now, realizations streamed as they are computed
  ncell = int(nx, 8)*int(ny, 8)*int(nz, 8)
  allocate(zmean(max(nreal, 1)), zvar(max(nreal, 1)), zimag(max(nreal, 1)))
  prec = SS3D_FP64
  if (f_fp32) prec = SS3D_FP32
  outs%origin = org
  outs%spacing = spc
  if (o_bin) outs%bin_path = cstr(trim(outbase)//'.bin', p_bin)
  if (o_vti) outs%vti_path = cstr(trim(outbase)//'.vti', p_vti)
  if (o_vtk_all .or. o_vtk1_bin) then
    outs%vtk_base = cstr(trim(outbase), p_vtk)
    outs%vtk_count = int(nreal, c_int32_t)
    if (.not. o_vtk_all) outs%vtk_count = int(min(1, nreal), c_int32_t)
  end if
  if (o_gslib) outs%gslib_path = cstr(trim(outbase)//'.out', p_out)
  if (o_vtk1_ascii) outs%vtk_ascii_path = cstr(trim(outbase)//'_1.vtk', p_vtka)
  vario_ok = .false.
  if (f_vario) call vario_read(trim(varfile))
  call ss3d_check(ss3d_writer_open(ctx, outs, int(nx, c_int), int(ny, c_int), int(nz, c_int), int(nreal, c_int), &
                                   prec, wr), 'output files')
  if (f_cond) then
    call cond_setup()
    call ss3d_check(ss3d_writer_condition(wr, cond), 'conditioning')
  end if

  ! ---- realizations
  chunk = int(ss3d_writer_slot(wr))
  tvario = 0
  do r0 = 0, nreal - 1, chunk
    cnt = min(chunk, nreal - r0)
    call ss3d_check(ss3d_writer_realize(wr, spec, int(iseed, c_int), int(igauss, c_int), int(cnt, c_int), &
                                        zmean(r0 + 1:), zvar(r0 + 1:), zimag(r0 + 1:), dfield), 'realizations')
    if (vario_ok) then
      call system_clock(tv0)
      call ss3d_check(ss3d_variogram(ctx, dfield, 1_c_int, prec, int(nx, c_int), int(ny, c_int), int(nz, c_int), &
                                     int(cnt, c_int), int(ndir, c_int), dirs, gam(:, :, r0 + 1:)), 'variograms')
      call system_clock(tv1)
      tvario = tvario + (tv1 - tv0)
    end if
  end do
  do i = 1, nreal
    write(*, '(a,i3,a,f9.4,a,f9.4,a,es9.2)') ' realization', i, '  mean=', zmean(i), '  variance=', zvar(i), &
      '  max|Im|=', zimag(i)
  end do
  i = ss3d_spectrum_destroy(spec)
  if (f_cond) call cond_report()
  call system_clock(t_real)
  write(*, '(a,f8.2,a)') ' time realizations         : ', real(t_real - t_spec - tvario)/real(rate), ' s'
  if (f_vario) then
    call vario_report(trim(varfile))
    call system_clock(t_vario)
    write(*, '(a,f8.2,a)') ' time variograms           : ', real(t_vario - t_real + tvario)/real(rate), ' s'
  end if

  ! ---- outputs: wait for the streamed files, write the GSLIB/ASCII files (same files as the reference)
  call ss3d_check(ss3d_writer_close(wr), 'output files')
  if (f_cond) i = ss3d_cond_destroy(cond)
  i = ss3d_context_destroy(ctx)

  call system_clock(t_end)
  write(*, '(a,f8.2,a)') ' elapsed time: ', real(t_end - t0)/real(rate), ' s'

contains

  ! Reads the conditioning block (f90:731-749) from the parameter file (new format, unit 10, still
  ! open) or from the --cond file (legacy), and the GSLIB data file (f90:757-779).
  subroutine cond_read()
    integer :: j, ir
    if (legacy) then
      ciu = 15
      open(ciu, file=trim(condfile), status='old', action='read', iostat=ios)
      if (ios /= 0) then
        write(*,*) 'Conditioning parameter file not found: ', trim(condfile); stop 1
      end if
    else
      ciu = 10
    end if
    read(ciu, '(a)') cline
    cline = adjustl(cline)
    datafile = cline(1:index(cline//' ', ' ') - 1)
    read(ciu, *) ccols%ixc, ccols%iyc, ccols%izc, ccols%ivc, ccols%iwc
    read(ciu, *) ccols%tmin, ccols%tmax
    read(ciu, *) ccols%itrans, ccols%skmean
    read(ciu, *) cpar%zlo, cpar%zhi
    read(ciu, *) cpar%ltail, cpar%ltpar
    read(ciu, *) cpar%utail, cpar%utpar
    read(ciu, *) cpar%rmax, cpar%rmed, cpar%rvert
    read(ciu, *) cpar%sa1, cpar%sa2, cpar%sa3
    read(ciu, *) cpar%ndmax
    if (legacy) close(ciu)
    if (cpar%ndmax < 1 .or. cpar%ndmax > 100) stop 'cond: ndmax must be in 1..100'
    if (cpar%rmax <= 0 .or. cpar%rmed <= 0 .or. cpar%rvert <= 0) stop 'cond: search radii must be positive'
    cpar%itrans = ccols%itrans
    cpar%skmean = ccols%skmean
    open(16, file=trim(datafile), status='old', action='read', iostat=ios)
    if (ios /= 0) then
      write(*,*) 'Data file not found: ', trim(datafile); stop 1
    end if
    read(16, '(a)') cline
    read(16, *) nvar
    do j = 1, nvar
      read(16, '(a)') cline
    end do
    nraw = 0
    do
      read(16, '(a)', iostat=ios) cline
      if (ios /= 0) exit
      if (len_trim(cline) > 0) nraw = nraw + 1
    end do
    rewind(16)
    read(16, '(a)') cline
    read(16, *) nvar
    do j = 1, nvar
      read(16, '(a)') cline
    end do
    allocate(crow(nvar), crows(nvar, max(nraw, 1)))
    ir = 0
    do j = 1, nraw
      read(16, '(a)', iostat=ios) cline
      if (ios /= 0) exit
      if (len_trim(cline) == 0) cycle
      read(cline, *) crow
      ir = ir + 1
      crows(:, ir) = crow
    end do
    nraw = ir
    close(16)
  end subroutine cond_read

  ! Prepares the data (library, f90:760-846), writes <base>_nscore.trn and <base>_conddata.dat as
  ! the reference does (f90:826-855) and computes the kriging weights of every node on the GPU.
  subroutine cond_setup()
    integer :: u, j, rc
    allocate(c_ijk(3, max(nraw, 1)), c_dval(max(nraw, 1)), c_dg(max(nraw, 1)), c_vr(max(nraw, 1)), c_gt(max(nraw, 1)))
    rc = ss3d_cond_prepare(crows, int(nraw, c_int), int(nvar, c_int), ccols, int(nx, c_int), int(ny, c_int), &
                           int(nz, c_int), org(1), org(2), org(3), spc(1), spc(2), spc(3), c_nd, c_nmerged, c_nout, &
                           c_ijk, c_dval, c_dg, c_vr, c_gt)
    if (rc /= 0) then
      if (ss3d_error_has('no conditioning data')) stop 'cond: no conditioning data inside the grid'
    end if
    call ss3d_check(rc, 'conditioning data')
    cnd = int(c_nd)
    if (ccols%itrans == 1) then
      open(newunit=u, file=trim(outbase)//'_nscore.trn', status='replace', action='write')
      write(u, '(a)') 'SPECSIM3D normal score transform table (value, normal score)'
      write(u, '(i0)') 2
      write(u, '(a)') 'value'
      write(u, '(a)') 'nscore'
      do j = 1, cnd
        write(u, '(2es17.8)') c_vr(j), c_gt(j)
      end do
      close(u)
      cpar%nt = c_nd
      cpar%vr = c_loc(c_vr)
      cpar%gt = c_loc(c_gt)
    end if
    open(newunit=u, file=trim(outbase)//'_conddata.dat', status='replace', action='write')
    write(u, '(a)') 'SPECSIM3D conditioning data used (grid nodes)'
    write(u, '(i0)') 5
    write(u, '(a)') 'x'
    write(u, '(a)') 'y'
    write(u, '(a)') 'z'
    write(u, '(a)') 'value'
    write(u, '(a)') 'gauss'
    do j = 1, cnd
      write(u, '(5es17.8)') org(1) + (c_ijk(1, j) - 1)*spc(1), org(2) + (c_ijk(2, j) - 1)*spc(2), &
                            org(3) + (c_ijk(3, j) - 1)*spc(3), c_dval(j), c_dg(j)
    end do
    close(u)
    call ss3d_check(ss3d_cond_create(ctx, st, int(nst, c_int), nugget, int(nx, c_int), int(ny, c_int), &
                                     int(nz, c_int), org(1), org(2), org(3), spc(1), spc(2), spc(3), c_nd, c_ijk, &
                                     c_dg, cpar, cond, cinfo), 'kriging weights')
  end subroutine cond_setup

  ! The reference's conditioning lines (f90:818-819, 844-845, 890, 927-931).
  subroutine cond_report()
    character(len=24) :: tb
    call ss3d_check(ss3d_cond_info_get(cond, cinfo), 'conditioning')
    write(*, '(a,i8,a,i8,a,i8,a)') ' conditioning data: ', c_nd, ' nodes used (', c_nmerged, ' merged, ', c_nout, &
      ' outside the grid)'
    if (ccols%itrans == 1) write(*, '(a,f9.4,a,f9.4)') ' normal score transform: data mean=', &
      sum(c_dg(1:cnd))/cnd, '  variance=', sum((c_dg(1:cnd) - sum(c_dg(1:cnd))/cnd)**2)/cnd
    write(*, '(a,3i5,a,i0,a)') ' super blocks:', cinfo%nsbx, cinfo%nsby, cinfo%nsbz, '  (search ', cpar%ndmax, &
      ' nearest data)'
    write(*, '(a,f6.1,a,i0,a,i0,a)') ' mean neighbours per node: ', real(cinfo%nsum, dp)/(real(nx, dp)*ny*nz), &
      '  (nodes without data: ', cinfo%nnone, ', data nodes: ', cinfo%nexact, ')'
    if (cinfo%nfail > 0) write(*, '(a,i0)') ' WARNING: kriging systems regularized: ', cinfo%nfail
    write(tb, '(f10.2)') real(cinfo%backtr_s)
    write(*, '(a,f8.2,a)') ' time conditioning: ', real(cinfo%weights_s + cinfo%apply_s), &
      ' s (kriging) + '//trim(adjustl(tb))//' s (back transform)'
  end subroutine cond_report

  ! Reads the variogram file (f90:454-462): ndir iperreal / ndir lines sx sy sz nlag (lag k =
  ! k*(sx,sy,sz) cells).  A file that cannot be opened is reported where the reference reports it,
  ! after the realizations.
  subroutine vario_read(vfile)
    character(len=*), intent(in) :: vfile
    integer :: id, u, nlagmax
    open(newunit=u, file=vfile, status='old', action='read', iostat=ios)
    if (ios /= 0) return
    read(u, *) ndir, iper
    allocate(dirs(4, ndir))
    do id = 1, ndir
      read(u, *) dirs(1, id), dirs(2, id), dirs(3, id), dirs(4, id)
    end do
    close(u)
    nlagmax = max(1, maxval(dirs(4, :)))
    allocate(gam(nlagmax, ndir, max(nreal, 1)))
    gam = 0.0_dp
    vario_ok = .true.
  end subroutine vario_read

  ! Total sill of the model: the nugget plus every structure's sill, added in structure order
  ! (the value the reference's report takes as C(0)).
  real(dp) function model_sill()
    integer :: is
    model_sill = nugget
    do is = 1, nst
      model_sill = model_sill + st(is)%sill
    end do
  end function model_sill

  ! The rows of one direction of the variogram table: the lags 1..nlag whose shift keeps pairs
  ! inside the grid, each with its distance, pair count, model semivariogram (sill - C(h)) and the
  ! ensemble mean and sd of the realizations' values; nrow of them.
  subroutine vario_rows(id, sill, nrow, lagv, dist, npair, gmod, gmean, gsdev)
    integer, intent(in) :: id
    real(dp), intent(in) :: sill
    integer, intent(out) :: nrow
    integer, intent(out) :: lagv(:)
    integer(8), intent(out) :: npair(:)
    real(dp), intent(out) :: dist(:), gmod(:), gmean(:), gsdev(:)
    integer :: lag
    integer(8) :: np
    real(dp) :: shift(3)
    nrow = 0
    do lag = 1, dirs(4, id)
      np = ss3d_npairs(int(nx, c_int), int(ny, c_int), int(nz, c_int), int(lag*dirs(1, id), c_int), &
                       int(lag*dirs(2, id), c_int), int(lag*dirs(3, id), c_int))
      if (np <= 0) cycle
      nrow = nrow + 1
      shift = real(lag*dirs(1:3, id), dp)*spc
      lagv(nrow) = lag
      npair(nrow) = np
      dist(nrow) = sqrt(shift(1)*shift(1) + shift(2)*shift(2) + shift(3)*shift(3))
      gmod(nrow) = sill - ss3d_covariance(st, int(nst, c_int), nugget, shift(1), shift(2), shift(3))
      gmean(nrow) = sum(gam(lag, id, 1:nreal))/real(nreal, dp)
      gsdev(nrow) = 0.0_dp
      if (nreal > 1) gsdev(nrow) = sqrt(sum((gam(lag, id, 1:nreal) - gmean(nrow))**2)/real(nreal - 1, dp))
    end do
  end subroutine vario_rows

  ! The variogram report (f90:434-529): <base>_vario.csv with one row per direction and lag (and the
  ! realizations' values when the file asks for them), and per direction the RMS and largest
  ! deviation of the ensemble mean from the model over the rising part of the model
  ! (gamma_model < 0.95 sill).  The pair sums come from the GPU (ss3d_variogram, on each chunk).
  subroutine vario_report(vfile)
    character(len=*), intent(in) :: vfile
    integer :: id, k, nrow, nfit, u, nper, r
    integer, allocatable :: lagv(:)
    integer(8), allocatable :: npair(:)
    real(dp), allocatable :: dist(:), gmod(:), gmean(:), gsdev(:), dev(:)
    logical, allocatable :: rising(:)
    real(dp) :: sill, rmsd, maxd
    character(len=512) :: vname
    if (.not. vario_ok) then
      write(*,*) 'cannot open variogram file: ', vfile; return
    end if
    sill = model_sill()
    nper = 0
    if (iper == 1) nper = nreal
    vname = trim(outbase)//'_vario.csv'
    open(newunit=u, file=trim(vname), status='replace', action='write')
    write(u, '(a)', advance='no') 'dir,sx,sy,sz,lag,h,npairs,gamma_model,gamma_mean,gamma_sd'
    if (nper > 0) write(u, '(*(a,i0))', advance='no') (',gamma_r', r, r = 1, nper)
    write(u, '(a)') ''
    write(*, '(a)') ' variograms (Fortran): direction  lags  RMS(mean-model)  max|err|   [rising part, ' // &
                    'gamma_model < 0.95 sill]'
    do id = 1, ndir
      k = max(1, dirs(4, id))
      allocate(lagv(k), npair(k), dist(k), gmod(k), gmean(k), gsdev(k))
      call vario_rows(id, sill, nrow, lagv, dist, npair, gmod, gmean, gsdev)
      do k = 1, nrow
        write(u, '(i0,3(",",i0),",",i0,",",es14.6,",",i0,3(",",es14.6))', advance='no') id, dirs(1:3, id), &
          lagv(k), dist(k), npair(k), gmod(k), gmean(k), gsdev(k)
        if (nper > 0) write(u, '(*(",",es14.6))', advance='no') gam(lagv(k), id, 1:nper)
        write(u, '(a)') ''
      end do
      rising = gmod(1:nrow) < 0.95_dp*sill
      dev = gmean(1:nrow) - gmod(1:nrow)
      nfit = count(rising)
      rmsd = 0.0_dp
      maxd = 0.0_dp
      if (nfit > 0) then
        rmsd = sqrt(sum(dev*dev, mask=rising)/real(nfit, dp))
        maxd = maxval(abs(dev), mask=rising)
      end if
      write(*, '(a,i3,3i4,i6,2f12.5)') '   dir', id, dirs(1:3, id), dirs(4, id), rmsd, maxd
      deallocate(lagv, npair, dist, gmod, gmean, gsdev)
    end do
    close(u)
    write(*, '(a,a)') ' variogram table: ', trim(vname)
  end subroutine vario_report

  ! Copies s into buf as a NUL-terminated C string and returns its address.
  function cstr(s, buf) result(p)
    character(len=*), intent(in) :: s
    character(kind=c_char), target, intent(inout) :: buf(:)
    type(c_ptr) :: p
    integer :: j
    if (len(s) + 1 > size(buf)) stop 'output file name too long'
    do j = 1, len(s)
      buf(j) = s(j:j)
    end do
    buf(len(s) + 1) = c_null_char
    p = c_loc(buf)
  end function cstr

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
