!! particles_host -- drop-in host driver for the MI355X integrator.
!!
!! Does what the reference program does (gkowal/particles,
!! sources/particles.F90:29-387) with the integration on the GPU:
!!   * reads the same parameter file (`-i file`, default ./params.in) with
!!     the same keys and defaults (particles.F90:53-84,124-155; field files
!!     fitsio.F90:79-85);
!!   * the same unit algebra (vunit / c, thermal speed, q/m of protons and
!!     electrons, particles.F90:163-214), region clamp (:150-155) and
!!     snapshot time grid (:240-280);
!!   * initial positions uniform in the region and momenta normal(0, vth)
!!     with the intrinsic random_number in the reference's call order
!!     (random.F90:104-220), and the moment printout (random.F90:231-353);
!!     for an even number of particles the state equals, bit for bit, the
!!     one the reference's own random module draws in a fresh process;
!!   * the six field files (velx..magz[.fits|.fits.gz] in datadir) read as
!!     FP64 and scaled by vunit / bunit (fields.F90:63-74);
!!   * params(1:17) packed as particles.F90:317-333, the whole particle loop
!!     (:337-356) as one ptc_integrate call;
!!   * counter.fits, time.fits (TUNIT) and state.fits (TEMP) as
!!     particles.F90:365-371.
!! Extra command-line switches: --mixed (FP32 fields, mixed precision),
!! --kernels auto|lanes|group|N (which device kernels integrate: the
!! library's "tail" option, "lanes" = one lane per particle; default auto),
!! --layout aos|columns|pairs|soa|brick and --f32-layout nodes|columns|quads|
!! pairs (how the FP64 and the --mixed FP32 fields are stored on the device:
!! the library's "layout" / "f32_layout" options; default aos and nodes;
!! columns, four times the memory of aos, is the fastest form measured on
!! 256^3 to 1024^3 fields),
!! --dry-run (write the initial state and params(1:17) to initial_state.fits
!! and params.fits and stop before the integration; needs no GPU).
!! Extra keys (the reference's field is static): field_snapshots = N > 1
!! reads N field snapshots, snapshot s from <name>_<ssss>.fits[.gz]
!! (s = 0001 ..), at times (s - 1) * field_dt in the unit of tmax, blended
!! linearly in time at every stage (ptc_fields_series); default 1, the
!! reference's single static field.
!! Differences on purpose: "zmax" sets zmax (the reference stores it into
!! zmax_domain, particles.F90:149); an odd number of particles gets normal
!! momenta too (the reference's split of the array needs an even count);
!! unknown methods are an error instead of integrating nothing; particles
!! whose integration could not finish are reported (status of ptc_integrate).

module host_params
  implicit none
  private
  public :: read_params, get_param

  integer, parameter :: nlen = 64, vlen = 256
  character(len=nlen), allocatable :: pname(:)
  character(len=vlen), allocatable :: pvalue(:)
  integer :: nparams = 0

  interface get_param
    module procedure get_char, get_int, get_real
  end interface

contains

  ! "name = value  # comment" lines; blank lines and lines starting with '#'
  ! are skipped; a quoted value is the text between the first and the last
  ! quote.  A name given twice keeps its last value.
  subroutine read_params(fname, ok)
    character(len=*), intent(in) :: fname
    logical, intent(out)         :: ok
    character(len=1024) :: line, tmp
    integer :: u, ios, n, p, c, q1, q2
    ok = .false.
    open(newunit = u, file = fname, status = 'old', action = 'read', iostat = ios)
    if (ios /= 0) return
    n = 0
    do
      read(u, '(a)', iostat = ios) line
      if (ios /= 0) exit
      n = n + 1
    end do
    allocate(pname(n), pvalue(n))
    rewind(u)
    nparams = 0
    do
      read(u, '(a)', iostat = ios) line
      if (ios /= 0) exit
      tmp = adjustl(line)
      if (len_trim(tmp) == 0 .or. tmp(1:1) == '#') cycle
      p = index(line, '=')
      if (p <= 1) cycle
      q1 = scan(line, '"''')
      if (q1 > p) then
        q2 = index(line, line(q1:q1), back = .true.)
        tmp = line(q1 + 1:max(q1, q2 - 1))
      else
        c = index(line, '#')
        if (c > p) then
          tmp = line(p + 1:c - 1)
        else
          tmp = line(p + 1:)
        end if
      end if
      nparams = nparams + 1
      pname(nparams) = adjustl(line(1:p - 1))
      pvalue(nparams) = adjustl(tmp)
    end do
    close(u)
    ok = .true.
  end subroutine

  integer function find(name)
    character(len=*), intent(in) :: name
    integer :: i
    find = 0
    do i = 1, nparams
      if (trim(pname(i)) == trim(name) .and. len_trim(pvalue(i)) > 0) find = i
    end do
  end function

  subroutine get_char(name, value)
    character(len=*), intent(in)    :: name
    character(len=*), intent(inout) :: value
    integer :: i
    i = find(name)
    if (i > 0) value = trim(pvalue(i))
  end subroutine

  subroutine get_int(name, value)
    character(len=*), intent(in) :: name
    integer, intent(inout)       :: value
    integer :: i, ios, v
    i = find(name)
    if (i == 0) return
    read(pvalue(i), *, iostat = ios) v
    if (ios == 0) then
      value = v
    else
      write(*, '(a)') ' warning: cannot read integer parameter ' // trim(name)
    end if
  end subroutine

  subroutine get_real(name, value)
    character(len=*), intent(in)  :: name
    real(8), intent(inout)        :: value
    integer :: i, ios
    real(8) :: v
    i = find(name)
    if (i == 0) return
    read(pvalue(i), *, iostat = ios) v
    if (ios == 0) then
      value = v
    else
      write(*, '(a)') ' warning: cannot read real parameter ' // trim(name)
    end if
  end subroutine

end module host_params


program particles_host

  use iso_c_binding
  use ptc_binding
  use host_params
  implicit none

  ! physical constants (SI, the reference's values)
  real(8), parameter :: c  = 2.99792458d+08
  real(8), parameter :: kb = 1.38064852d-23
  real(8), parameter :: mp = 1.672621898d-27
  real(8), parameter :: me = 9.10938356d-31
  real(8), parameter :: e  = 1.6021766208d-19

  ! parameters and their defaults
  character(len=32)  :: simulation = 'newtonian', ptype = 'proton', tscale = 'linear'
  character(len=32)  :: method = 'rk5'
  integer            :: nparticles = 1000, nsteps = 100
  real(8) :: dtini = 1.0d-04, dtmax = 1.0d-03, tmin = 1.0d-03, tmax = 1.0d+00
  real(8) :: atol = 1.0d-06, rtol = 1.0d-06, facmin = 3.33d-01, facmax = 6.0d+00
  real(8) :: safe = 9.0d-01, beta = 0.0d+00, temp = 1.0d+03, vth = 1.0d-03
  real(8) :: bunit = 1.0d+00, vunit = 1.0d+00, tunit = 1.0d+00, lunit = 1.0d+00
  real(8) :: qom = 9.5788332241480675d+03
  real(8) :: xmin = 0.0d0, ymin = 0.0d0, zmin = 0.0d0, xmax = 1.0d0, ymax = 1.0d0, zmax = 1.0d0
  real(8) :: xmin_domain = 0.0d0, ymin_domain = 0.0d0, zmin_domain = 0.0d0
  real(8) :: xmax_domain = 1.0d0, ymax_domain = 1.0d0, zmax_domain = 1.0d0
  character(len=256) :: datadir = './'
  character(len=64)  :: fnames(6) = [character(len=64) :: 'velx', 'vely', 'velz', 'magx', &
                                     'magy', 'magz']
  character(len=16), parameter :: fkeys(6) = [character(len=16) :: 'x-velocity-name', &
      'y-velocity-name', 'z-velocity-name', 'x-magnetic-name', 'y-magnetic-name', &
      'z-magnetic-name']
  character(len=4), parameter :: label(6) = ['posx', 'posy', 'posz', 'momx', 'momy', 'momz']

  character(len=512) :: pfile = './params.in', arg, path
  logical  :: ok, dry_run = .false., mixed = .false., exists
  character(len=32) :: kernels = 'auto', layout = 'aos', f32_layout = 'nodes'
  integer  :: i, k, iret, nargs, naxis, bitpix, imethod, nbad, s
  integer  :: nfs = 1
  real(8)  :: fdt = 0.0d0
  character(len=64) :: fbase(6)
  type(c_ptr), allocatable :: snaps(:)
  real(8), allocatable :: stimes(:)
  integer(c_int64_t) :: dims(8), nn, n
  integer(8) :: t1, t2, rate
  integer(c_int) :: dm(3)
  real(8) :: ltmn, ltmx, moms(4), params(32)
  real(8), allocatable :: uu(:,:,:,:), bb(:,:,:,:), time(:), state(:,:,:), tmp(:,:)
  integer(c_int64_t), allocatable :: counter(:,:)
  integer(c_int), allocatable, target :: status(:)
  type(c_ptr) :: fh

  ! ---- command line ----
  nargs = command_argument_count()
  i = 1
  do while (i <= nargs)
    call get_command_argument(i, arg)
    select case (trim(arg))
    case ('-i', '--input')
      i = i + 1
      call get_command_argument(i, pfile)
    case ('--dry-run')
      dry_run = .true.
    case ('--mixed')
      mixed = .true.
    case ('--kernels')
      i = i + 1
      call get_command_argument(i, kernels)
      if (trim(kernels) == 'lanes') kernels = '0'
      if (ptc_set_option('tail' // c_null_char, trim(kernels) // c_null_char) /= PTC_OK) then
        write(*, '(a)') ' --kernels: ' // ptc_last_error()
        stop 1
      end if
    case ('--layout')
      i = i + 1
      call get_command_argument(i, layout)
      if (ptc_set_option('layout' // c_null_char, trim(layout) // c_null_char) /= PTC_OK) then
        write(*, '(a)') ' --layout: ' // ptc_last_error()
        stop 1
      end if
    case ('--f32-layout')
      i = i + 1
      call get_command_argument(i, f32_layout)
      if (ptc_set_option('f32_layout' // c_null_char, trim(f32_layout) // c_null_char) /= PTC_OK) then
        write(*, '(a)') ' --f32-layout: ' // ptc_last_error()
        stop 1
      end if
    case default
      write(*, '(a)') ' unknown option ' // trim(arg)
      stop 1
    end select
    i = i + 1
  end do

  write(*, "(80('-'))")
  write(*, "(a)") ' particles (MI355X integrator, ' // ptc_version_string() // ')'
  call read_params(trim(pfile), ok)
  if (.not. ok) then
    write(*, '(a)') ' cannot open the parameter file ' // trim(pfile)
    stop 1
  end if

  call get_param('simulation', simulation)
  call get_param('method', method)
  call get_param('particle_type', ptype)
  call get_param('time_scale', tscale)
  call get_param('nparticles', nparticles)
  call get_param('nsteps', nsteps)
  call get_param('bunit', bunit)
  call get_param('vunit', vunit)
  call get_param('tunit', tunit)
  call get_param('temperature', temp)
  call get_param('tmin', tmin)
  call get_param('tmax', tmax)
  call get_param('dtini', dtini)
  call get_param('dtmax', dtmax)
  call get_param('atol', atol)
  call get_param('rtol', rtol)
  call get_param('facmin', facmin)
  call get_param('facmax', facmax)
  call get_param('safe', safe)
  call get_param('beta', beta)
  call get_param('xmin', xmin)
  call get_param('ymin', ymin)
  call get_param('zmin', zmin)
  call get_param('xmax', xmax)
  call get_param('ymax', ymax)
  call get_param('zmax', zmax)
  call get_param('xmin_domain', xmin_domain)
  call get_param('ymin_domain', ymin_domain)
  call get_param('zmin_domain', zmin_domain)
  call get_param('xmax_domain', xmax_domain)
  call get_param('ymax_domain', ymax_domain)
  call get_param('zmax_domain', zmax_domain)
  call get_param('datadir', datadir)
  do k = 1, 6
    call get_param(trim(fkeys(k)), fnames(k))
  end do
  call get_param('field_snapshots', nfs)
  call get_param('field_dt', fdt)
  if (nfs < 1 .or. (nfs > 1 .and. .not. fdt > 0.0d0)) then
    write(*, '(a)') ' field_snapshots must be >= 1, and field_dt > 0 for more than one'
    stop 1
  end if

  ! the seeding region lies inside the domain
  xmin = max(xmin, xmin_domain)
  ymin = max(ymin, ymin_domain)
  zmin = max(zmin, zmin_domain)
  xmax = min(xmax, xmax_domain)
  ymax = min(ymax, ymax_domain)
  zmax = min(zmax, zmax_domain)

  ! units: velocities in c, thermal speed of protons at `temperature`
  vunit = vunit / c
  vth = sqrt(kb * temp / mp) / c
  select case (trim(simulation))
  case ('relativistic', 'rel', 'r', 'srmhd', 'srhd')
    lunit = tunit * c
  case default
    lunit = tunit * vunit * c
  end select
  select case (trim(ptype))
  case ('proton', 'p')
    qom = 1.0d-04 * e * tunit / mp
  case ('electron', 'e')
    qom = -1.0d-04 * e * tunit / me
  end select
  select case (trim(method))
  case ('rk5')
    imethod = PTC_RK5
  case ('dop853', 'dp853', 'rk8')
    imethod = PTC_DOP853
  case default
    write(*, '(a)') ' unknown integration method ' // trim(method)
    stop 1
  end select
  select case (trim(tscale))
  case ('lin', 'linear')
    nsteps = nsteps + 1
  case ('log', 'logarithmic', 'log10')
    nsteps = int(log10(tmax) - log10(tmin)) * nsteps + 2
  case default
    write(*, '(a)') ' unknown time scale ' // trim(tscale)
    stop 1
  end select

  write(*, '(a)') ' Fields:'
  write(*, "(3x,'simulation    = ',a)") trim(simulation)
  write(*, "(3x,'B unit        = ',es13.6,' G')") bunit
  write(*, "(3x,'V unit        = ',es13.6,' c')") vunit
  write(*, "(3x,'time unit     = ',es13.6,' s')") tunit
  write(*, "(3x,'length unit   = ',es13.6,' m')") lunit
  write(*, '(a)') ' Particles:'
  write(*, "(3x,'nparticles    = ',i0)") nparticles
  write(*, "(3x,'particle type = ',a)") trim(ptype)
  write(*, "(3x,'q/m           = ',es22.15,' [1/s]')") qom
  write(*, "(3x,'temperature   = ',es22.15,' [K]')") temp
  write(*, "(3x,'thermal speed = ',es22.15,' [c]')") vth
  write(*, '(a)') ' Methods:'
  write(*, "(3x,'integration   = ',a)") trim(method)
  write(*, "(3x,'dtini dtmax   = ',2es13.5)") dtini, dtmax
  write(*, "(3x,'atol rtol     = ',2es13.5)") atol, rtol
  write(*, "(3x,'facmin facmax = ',2es13.5)") facmin, facmax
  write(*, "(3x,'safe beta     = ',2es13.5)") safe, beta
  write(*, "(3x,'precision     = ',a)") merge('mixed (FP32 fields)', 'FP64               ', mixed)
  write(*, "(3x,'kernels       = ',a)") trim(kernels)
  if (mixed) then
    write(*, "(3x,'field storage = ',a)") trim(f32_layout)
  else
    write(*, "(3x,'field storage = ',a)") trim(layout)
  end if
  if (nfs > 1) write(*, "(3x,'field snaps   = ',i0,' every ',es13.5)") nfs, fdt
  write(*, '(a)') ' Output:'
  write(*, "(3x,'time scale    = ',a)") trim(tscale)
  write(*, "(3x,'nsteps        = ',i0)") nsteps
  write(*, "(3x,'tmin tmax     = ',2es13.5)") tmin, tmax

  ! ---- fields (velx .. magz in datadir, .fits or .fits.gz; snapshot s of
  !      a series in <name>_<ssss>.fits[.gz]) ----
  if (len_trim(datadir) > 0) then
    if (datadir(len_trim(datadir):len_trim(datadir)) /= '/') datadir = trim(datadir) // '/'
  end if
  fbase = fnames
  call field_paths(1)
  iret = ptc_fits_dims(trim(fnames(1)) // c_null_char, naxis, dims, bitpix)
  if (iret /= 0 .or. naxis /= 3) then
    write(*, '(a)') ' cannot read the field dimensions: ' // ptc_fits_last_error()
    stop 1
  end if
  dm = int(dims(1:3), c_int)
  allocate(uu(dm(1), dm(2), dm(3), 3), bb(dm(1), dm(2), dm(3), 3))
  nn = int(dm(1), c_int64_t) * dm(2) * dm(3)
  do s = 2, nfs
    ! every snapshot present and of the first one's shape
    call field_paths(s)
    do k = 1, 6
      iret = ptc_fits_dims(trim(fnames(k)) // c_null_char, naxis, dims, bitpix)
      if (iret /= 0 .or. naxis /= 3 .or. any(int(dims(1:3), c_int) /= dm)) then
        write(*, '(a)') ' field snapshot ' // trim(fnames(k)) // ' does not match the first'
        stop 1
      end if
    end do
  end do
  call read_snapshot(1)
  write(*, "(' Field data     = ',i0,' x ',i0,' x ',i0,', ',f10.3,' GB')") dm, &
       16.0d0 * 3 * real(nn, 8) * nfs / 1024.0d0**3

  ! ---- snapshot times ----
  allocate(time(nsteps), counter(3, nparticles), status(nparticles))
  allocate(state(8, nsteps, nparticles))
  time = 0.0d0
  counter = 0
  state = 0.0d0
  if (tscale(1:3) == 'lin') then
    do n = 2, nsteps
      time(n) = (tmax * (n - 1)) / (nsteps - 1)
    end do
  else
    ltmn = log10(tmin)
    ltmx = log10(tmax)
    do n = 2, nsteps
      time(n) = 1.0d+01**(((nsteps - n) * ltmn + (n - 2) * ltmx) / (nsteps - 2))
    end do
  end if

  ! ---- initial state ----
  call system_clock(t1, rate)
  allocate(tmp(3, nparticles))
  call random_number(tmp)
  state(1, 1, :) = (xmax - xmin) * tmp(1, :) + xmin
  state(2, 1, :) = (ymax - ymin) * tmp(2, :) + ymin
  state(3, 1, :) = (zmax - zmin) * tmp(3, :) + zmin
  call normal_polar(0.0d0, vth, tmp)
  state(4:6, 1, :) = tmp
  call system_clock(t2)
  write(*, "(' Initial state generated in ',f10.4,' seconds.')") real(t2 - t1, 8) / rate
  write(*, '(a)') ' Statistics of the initial state:'
  write(*, "(6x,4a24)") 'mean', 'standard deviation', 'skewness', 'kurtosis'
  do k = 1, 6
    call moments(state(k, 1, :), moms)
    write(*, "(3x,a,': ',4es24.15)") label(k), moms
  end do

  params = 0.0d0
  params(1) = qom
  params(2) = vunit
  params(3) = tmax
  params(4) = dtini
  params(5) = dtmax
  params(6) = atol
  params(7) = rtol
  params(8) = facmin
  params(9) = facmax
  params(10) = safe
  params(11) = beta
  params(12) = xmin_domain
  params(13) = ymin_domain
  params(14) = zmin_domain
  params(15) = xmax_domain - xmin_domain
  params(16) = ymax_domain - ymin_domain
  params(17) = zmax_domain - zmin_domain

  if (dry_run) then
    write(*, "(' params(1:17) = ',/,(3x,4es24.15))") params(1:17)
    write(*, "(' time(1:nsteps) = ',/,(3x,4es24.15))") time
    ! the inputs of the integration, exactly: state(1:6,1,:) and params(1:17)
    call write_f64('!initial_state.fits', 2, [6_c_int64_t, int(nparticles, c_int64_t)], &
                   state(1:6, 1, :), 'TEMP', temp, 'Initial thermal temperature')
    call write_f64('!params.fits', 1, [17_c_int64_t], params(1:17), 'TUNIT', tunit, &
                   'Time unit in seconds')
    write(*, '(a)') ' Dry run: initial_state.fits and params.fits written; stopping before' // &
                    ' the integration.'
  else
    call integrate_and_store()
  end if

contains

  subroutine integrate_and_store()
  ! ---- integration on the GPU ----
  write(*, '(a)') ' Integrating trajectories...'
  call system_clock(t1)
  if (nfs == 1) then
    iret = ptc_fields_create(dm, uu, bb, PTC_MODE_VB, merge(PTC_FP32, PTC_FP64, mixed), fh)
    if (iret /= PTC_OK) then
      write(*, '(a)') ' ptc_fields_create: ' // ptc_last_error()
      stop 2
    end if
  else
    ! one snapshot on the host at a time: read, upload, next; then the
    ! series copies them into one block and the single handles go
    allocate(snaps(nfs), stimes(nfs))
    do s = 1, nfs
      if (s > 1) call read_snapshot(s)
      stimes(s) = (s - 1) * fdt
      iret = ptc_fields_create(dm, uu, bb, PTC_MODE_VB, merge(PTC_FP32, PTC_FP64, mixed), snaps(s))
      if (iret /= PTC_OK) then
        write(*, '(a)') ' ptc_fields_create: ' // ptc_last_error()
        stop 2
      end if
    end do
    iret = ptc_fields_series(snaps, stimes, int(nfs, c_int), fh)
    if (iret /= PTC_OK) then
      write(*, '(a)') ' ptc_fields_series: ' // ptc_last_error()
      stop 2
    end if
    do s = 1, nfs
      iret = ptc_fields_destroy(snaps(s))
    end do
  end if
  iret = ptc_integrate(fh, imethod, int(nparticles, c_int64_t), nsteps, params, size(params), &
                       time, state, counter, c_loc(status))
  if (iret /= PTC_OK) then
    write(*, '(a)') ' ptc_integrate: ' // ptc_last_error()
    stop 2
  end if
  iret = ptc_fields_destroy(fh)
  call system_clock(t2)
  write(*, "(' Integration completed in ',f12.4,' seconds.')") real(t2 - t1, 8) / rate
  nbad = count(status /= 0)
  if (nbad > 0) write(*, "(' warning: ',i0,' particles did not reach tmax (status /= 0)')") nbad

  ! ---- output files ----
  call system_clock(t1)
  call write_i64('!counter.fits', 2, [3_c_int64_t, int(nparticles, c_int64_t)], counter)
  call write_f64('!time.fits', 1, [int(nsteps, c_int64_t)], time, 'TUNIT', tunit, &
                 'Time unit in seconds')
  call write_f64('!state.fits', 3, [8_c_int64_t, int(nsteps, c_int64_t), &
                 int(nparticles, c_int64_t)], state, 'TEMP', temp, 'Initial thermal temperature')
  call system_clock(t2)
  write(*, "(' Data stored in ',f10.4,' seconds.')") real(t2 - t1, 8) / rate
  end subroutine

  subroutine field_paths(snap)
  ! fnames = the six files of snapshot `snap` (the plain names for a static
  ! field), .fits or else .fits.gz
  integer, intent(in) :: snap
  character(len=16) :: suffix
  integer :: j
  logical :: there
  suffix = ''
  if (nfs > 1) write(suffix, "('_',i4.4)") snap
  do j = 1, 6
    path = trim(datadir) // trim(fbase(j)) // trim(suffix) // '.fits'
    inquire(file = trim(path), exist = there)
    if (.not. there) then
      path = trim(path) // '.gz'
      inquire(file = trim(path), exist = there)
    end if
    if (.not. there) then
      write(*, '(a)') ' missing field file ' // trim(datadir) // trim(fbase(j)) // trim(suffix) &
                      // '.fits[.gz]'
      stop 1
    end if
    fnames(j) = trim(path)
  end do
  end subroutine

  subroutine read_snapshot(snap)
  ! uu, bb = snapshot `snap` in code units (fields.F90:63-74)
  integer, intent(in) :: snap
  integer :: j
  call field_paths(snap)
  do j = 1, 3
    iret = ptc_fits_read(trim(fnames(j)) // c_null_char, nn, uu(:,:,:,j))
    if (iret == 0) iret = ptc_fits_read(trim(fnames(j + 3)) // c_null_char, nn, bb(:,:,:,j))
    if (iret /= 0) then
      write(*, '(a)') ' cannot read a field: ' // ptc_fits_last_error()
      stop 1
    end if
  end do
  uu = vunit * uu
  bb = bunit * bb
  end subroutine

  function ptc_version_string() result(v)
    character(len=:), allocatable :: v
    interface
      type(c_ptr) function ptc_version_c() bind(C, name="ptc_version")
        import :: c_ptr
      end function
    end interface
    v = c_string(ptc_version_c())
  end function

  ! normal(mu, sigma) by the Marsaglia polar method: pairs (u, v) uniform in
  ! (-1, 1)^2 redrawn until 0 < u^2 + v^2 <= 1, the first half of the
  ! columns from u, the second from v (random.F90:104-220 draws in this
  ! order: all u, all v, then per element re-draws, column-major)
  subroutine normal_polar(mu, sigma, arr)
    real(8), intent(in)    :: mu, sigma
    real(8), intent(inout) :: arr(:,:)
    real(8), allocatable :: u(:,:), v(:,:), s(:,:)
    integer :: i, j, m, h, nc
    m = size(arr, 1)
    nc = size(arr, 2)
    h = (nc + 1) / 2
    allocate(u(m, h), v(m, h), s(m, h))
    call random_number(u)
    u = 2.0d0 * u - 1.0d0
    call random_number(v)
    v = 2.0d0 * v - 1.0d0
    s = u * u + v * v
    do j = 1, h
      do i = 1, m
        do while (s(i, j) > 1.0d0 .or. s(i, j) == 0.0d0)
          call random_number(u(i, j))
          call random_number(v(i, j))
          u(i, j) = 2.0d0 * u(i, j) - 1.0d0
          v(i, j) = 2.0d0 * v(i, j) - 1.0d0
          s(i, j) = u(i, j) * u(i, j) + v(i, j) * v(i, j)
        end do
      end do
    end do
    s = sqrt(-2.0d0 * log(s) / s)
    arr(:, 1:h) = u * s * sigma + mu
    if (nc > h) arr(:, h + 1:nc) = v(:, 1:nc - h) * s(:, 1:nc - h) * sigma + mu
  end subroutine

  ! mean, standard deviation, skewness, excess kurtosis; means by
  ! compensated (Neumaier) summation
  subroutine moments(x, mom)
    real(8), intent(in)  :: x(:)
    real(8), intent(out) :: mom(4)
    real(8), allocatable :: d(:)
    mom(1) = csum(x) / size(x)
    allocate(d(size(x)))
    d = x - mom(1)
    mom(2) = sqrt(csum(d * d) / size(x))
    if (mom(2) /= 0.0d0) d = d / mom(2)
    mom(3) = csum(d**3) / size(x)
    mom(4) = csum(d**4) / size(x) - 3.0d0
  end subroutine

  real(8) function csum(x)
    real(8), intent(in) :: x(:)
    real(8) :: s, t, comp
    integer :: i
    s = x(1)
    comp = 0.0d0
    do i = 2, size(x)
      t = s + x(i)
      if (abs(s) >= abs(x(i))) then
        comp = comp + (s - t) + x(i)
      else
        comp = comp + (x(i) - t) + s
      end if
      s = t
    end do
    csum = s + comp
  end function

  subroutine write_f64(fname, naxis, dims, data, key, value, comment)
    character(len=*), intent(in)   :: fname, key, comment
    integer, intent(in)            :: naxis
    integer(c_int64_t), intent(in) :: dims(:)
    real(8), intent(in)            :: data(*)
    real(8), intent(in)            :: value
    character(kind=c_char, len=len(key) + 1), target     :: ckey
    character(kind=c_char, len=len(comment) + 1), target :: ccom
    type(c_ptr) :: kp(1), cp(1)
    ckey = key // c_null_char
    ccom = comment // c_null_char
    kp(1) = c_loc(ckey)
    cp(1) = c_loc(ccom)
    if (ptc_fits_write_f64(fname // c_null_char, naxis, dims, data, 1, kp, [value], cp) /= 0) then
      write(*, '(a)') ' cannot write ' // fname // ': ' // ptc_fits_last_error()
      stop 3
    end if
  end subroutine

  subroutine write_i64(fname, naxis, dims, data)
    character(len=*), intent(in)   :: fname
    integer, intent(in)            :: naxis
    integer(c_int64_t), intent(in) :: dims(:), data(*)
    if (ptc_fits_write_i64(fname // c_null_char, naxis, dims, data) /= 0) then
      write(*, '(a)') ' cannot write ' // fname // ': ' // ptc_fits_last_error()
      stop 3
    end if
  end subroutine

end program particles_host
