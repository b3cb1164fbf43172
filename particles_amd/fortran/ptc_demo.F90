!! Minimal caller of the C-ABI through ptc_binding: uniform B along z, ten
!! particles, DOP853, prints the step counters.  Build:
!!   flang -c ptc_binding.F90 && flang ptc_demo.F90 ptc_binding.o \
!!         -L.. -lparticles_amd -Wl,-rpath,<dir of libparticles_amd.so>
program ptc_demo
  use iso_c_binding
  use ptc_binding
  implicit none
  integer(c_int), parameter :: n = 8, np = 10, m = 3
  integer(c_int) :: dm(3), iret, i
  real(c_double), allocatable :: uu(:,:,:,:), bb(:,:,:,:), state(:,:,:)
  real(c_double) :: params(32), time(m)
  integer(c_int64_t) :: counter(3, np)
  type(c_ptr) :: fh

  dm = n
  allocate(uu(n, n, n, 3), bb(n, n, n, 3), state(8, m, np))
  uu = 0.0d0; bb = 0.0d0; bb(:,:,:,3) = 1.0d0
  params = 0.0d0
  params(1:11) = [1.0d0, 1.0d0, 6.283185307179586d0, 1.0d-4, 1.0d0, 1.0d-8, 1.0d-8, &
                  0.333d0, 6.0d0, 0.9d0, 0.0d0]
  params(15:17) = 1.0d0
  time = [0.0d0, 3.141592653589793d0, 6.283185307179586d0]
  state = 0.0d0
  do i = 1, np
    state(1:6, 1, i) = [0.1d0 * i, 0.5d0, 0.5d0, 0.1d0 * i, 0.0d0, 0.0d0]
  end do

  iret = ptc_fields_create(dm, uu, bb, PTC_MODE_VB, PTC_FP64, fh)
  if (iret /= PTC_OK) then
    print *, 'ptc_fields_create failed: ', ptc_last_error()
    stop 1
  end if
  iret = ptc_integrate(fh, PTC_DOP853, int(np, c_int64_t), m, params, 32, time, state, counter, &
                       c_null_ptr)
  if (iret /= PTC_OK) then
    print *, 'ptc_integrate failed: ', ptc_last_error()
    stop 1
  end if
  do i = 1, np
    print '(a,i3,a,3i8,a,2es14.6)', 'particle', i, ' steps', counter(:, i), '  x(T0)', state(1:2, m, i)
  end do
  iret = ptc_fields_destroy(fh)
end program ptc_demo
