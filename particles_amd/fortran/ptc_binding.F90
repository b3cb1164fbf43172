!! ptc_binding -- Fortran 2003 interfaces (iso_c_binding) to the C-ABI of
!! libparticles_amd.so (include/particles_amd.h).
!!
!! With this module the reference driver's particle loop
!! (sources/particles.F90:337-356) becomes one call:
!!
!!   use ptc_binding
!!   type(c_ptr) :: fh
!!   iret = ptc_fields_create(dm, uu, bb, PTC_MODE_VB, PTC_FP64, fh)
!!   iret = ptc_integrate(fh, PTC_DOP853, int(nparticles, c_int64_t), nsteps, &
!!                        params, size(params), time, state, counter, c_null_ptr)
!!   iret = ptc_fields_destroy(fh)
!!
!! Arrays are passed as the reference declares them: uu/bb(dm1,dm2,dm3,3),
!! params(32), time(nsteps), state(8,nsteps,nparticles), counter(3,nparticles).
module ptc_binding

  use iso_c_binding
  implicit none

  integer(c_int), parameter :: PTC_OK = 0, PTC_EARG = 1, PTC_EHIP = 2, PTC_ENOMEM = 3, PTC_ESTATE = 4
  integer(c_int), parameter :: PTC_RK5 = 0, PTC_DOP853 = 1
  integer(c_int), parameter :: PTC_MODE_VB = 0, PTC_MODE_EB = 1
  integer(c_int), parameter :: PTC_FP64 = 0, PTC_FP32 = 1
  integer(c_int), parameter :: PTC_INIT_MAXWELL = 0, PTC_INIT_GAMMA = 1, PTC_INIT_GAMMA_LOG = 2

  interface

    integer(c_int) function ptc_fields_create(dm, f1, bb, mode, prec, out) bind(C, name="ptc_fields_create")
      import :: c_int, c_double, c_ptr
      integer(c_int), intent(in) :: dm(3)
      real(c_double), intent(in) :: f1(*), bb(*)
      integer(c_int), value      :: mode, prec
      type(c_ptr), intent(out)   :: out
    end function

    integer(c_int) function ptc_fields_destroy(f) bind(C, name="ptc_fields_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: f
    end function

    ! time-dependent fields: n snapshot handles at strictly increasing times
    integer(c_int) function ptc_fields_series(snaps, times, n, out) bind(C, name="ptc_fields_series")
      import :: c_int, c_double, c_ptr
      type(c_ptr), intent(in)    :: snaps(*)
      real(c_double), intent(in) :: times(*)
      integer(c_int), value      :: n
      type(c_ptr), intent(out)   :: out
    end function

    integer(c_int) function ptc_integrate(f, method, np, m, params, nparams, time, state, counter, &
                                          status) bind(C, name="ptc_integrate")
      import :: c_int, c_int64_t, c_double, c_ptr
      type(c_ptr), value             :: f
      integer(c_int), value          :: method, m, nparams
      integer(c_int64_t), value      :: np
      real(c_double), intent(in)     :: params(*), time(*)
      real(c_double), intent(inout)  :: state(*)
      integer(c_int64_t), intent(out):: counter(*)
      type(c_ptr), value             :: status
    end function

    integer(c_int) function ptc_particles_create(f, method, np, m, params, nparams, time, state, out) &
                                                 bind(C, name="ptc_particles_create")
      import :: c_int, c_int64_t, c_double, c_ptr
      type(c_ptr), value         :: f
      integer(c_int), value      :: method, m, nparams
      integer(c_int64_t), value  :: np
      real(c_double), intent(in) :: params(*), time(*), state(*)
      type(c_ptr), intent(out)   :: out
    end function

    integer(c_int) function ptc_particles_advance(p, t_stop) bind(C, name="ptc_particles_advance")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value    :: p
      real(c_double), value :: t_stop
    end function

    integer(c_int) function ptc_particles_download(p, state, counter, status) &
                                                   bind(C, name="ptc_particles_download")
      import :: c_int, c_ptr
      type(c_ptr), value :: p, state, counter, status
    end function

    integer(c_int) function ptc_particles_destroy(p) bind(C, name="ptc_particles_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: p
    end function

    integer(c_int) function ptc_particles_generate(p, kind, seed, lo, hi, a, b) &
                                                   bind(C, name="ptc_particles_generate")
      import :: c_int, c_int64_t, c_double, c_ptr
      type(c_ptr), value         :: p
      integer(c_int), value      :: kind
      integer(c_int64_t), value  :: seed
      real(c_double), intent(in) :: lo(3), hi(3)
      real(c_double), value      :: a, b
    end function

    integer(c_int) function ptc_particles_moments(p, moments) bind(C, name="ptc_particles_moments")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value          :: p
      real(c_double), intent(out) :: moments(24)
    end function

    integer(c_int) function ptc_particles_advance_graph(p, n, t_stop) &
                                                        bind(C, name="ptc_particles_advance_graph")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value         :: p
      integer(c_int), value      :: n
      real(c_double), intent(in) :: t_stop(*)
    end function

    integer(c_int) function ptc_particles_autotune(p, t_end, chosen) bind(C, name="ptc_particles_autotune")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value           :: p
      real(c_double), value        :: t_end
      integer(c_int), intent(out)  :: chosen(5)  ! block, sort, order, tail, group lanes
    end function

    integer(c_int) function ptc_particles_totals(p, totals) bind(C, name="ptc_particles_totals")
      import :: c_int, c_int64_t, c_ptr
      type(c_ptr), value              :: p
      integer(c_int64_t), intent(out) :: totals(5)
    end function

    integer(c_int) function ptc_set_option(key, value) bind(C, name="ptc_set_option")
      import :: c_int, c_char
      character(kind=c_char), intent(in) :: key(*), value(*)
    end function

    type(c_ptr) function ptc_last_error_c() bind(C, name="ptc_last_error")
      import :: c_ptr
    end function

    ! FITS primary-image I/O (include/particles_amd_fits.h); paths are
    ! NUL-terminated
    integer(c_int) function ptc_fits_dims(path, naxis, dims, bitpix) bind(C, name="ptc_fits_dims")
      import :: c_int, c_int64_t, c_char
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int), intent(out)        :: naxis, bitpix
      integer(c_int64_t), intent(out)    :: dims(8)
    end function

    integer(c_int) function ptc_fits_read(path, n, out) bind(C, name="ptc_fits_read")
      import :: c_int, c_int64_t, c_char, c_double
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int64_t), value          :: n
      real(c_double), intent(out)        :: out(*)
    end function

    integer(c_int) function ptc_fits_write_f64(path, naxis, dims, data, nkeys, keys, values, &
                                               comments) bind(C, name="ptc_fits_write_f64")
      import :: c_int, c_int64_t, c_char, c_double, c_ptr
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int), value              :: naxis, nkeys
      integer(c_int64_t), intent(in)     :: dims(*)
      real(c_double), intent(in)         :: data(*), values(*)
      type(c_ptr), intent(in)            :: keys(*), comments(*)
    end function

    integer(c_int) function ptc_fits_write_i64(path, naxis, dims, data) bind(C, name="ptc_fits_write_i64")
      import :: c_int, c_int64_t, c_char
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int), value              :: naxis
      integer(c_int64_t), intent(in)     :: dims(*), data(*)
    end function

    type(c_ptr) function ptc_fits_last_error_c() bind(C, name="ptc_fits_last_error")
      import :: c_ptr
    end function

  end interface

contains

  ! ptc_last_error() as a Fortran string
  function ptc_last_error() result(msg)
    character(len=:), allocatable :: msg
    msg = c_string(ptc_last_error_c())
  end function

  ! ptc_fits_last_error() as a Fortran string
  function ptc_fits_last_error() result(msg)
    character(len=:), allocatable :: msg
    msg = c_string(ptc_fits_last_error_c())
  end function

  ! a NUL-terminated C string as a Fortran string
  function c_string(cp) result(msg)
    type(c_ptr), intent(in) :: cp
    character(len=:), allocatable :: msg
    character(kind=c_char), pointer :: p(:)
    integer :: n
    call c_f_pointer(cp, p, [4096])
    n = 0
    do while (n < 4096)
      if (p(n + 1) == c_null_char) exit
      n = n + 1
    end do
    allocate(character(len=n) :: msg)
    msg = transfer(p(1:n), msg)
  end function

end module ptc_binding
