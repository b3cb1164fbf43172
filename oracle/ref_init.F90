!! ref_init -- TEST INFRASTRUCTURE ONLY: the reference program's initial
!! particle state, drawn by the reference's own random module
!! (sources/random.F90, compiled from the reference sources into
!! oracle/_ref), in the reference's call order (sources/particles.F90:289-296)
!! and with its thermal speed (particles.F90:163-164, constants :41-45).
!! A fresh process starts the compiler's random_number stream at its default
!! seed, as the reference program does.
!!
!! usage: ref_init np xmin xmax ymin ymax zmin zmax temperature outfile
!! writes state(1:6, 1, 1:np) (FP64, column-major, raw stream) to outfile.
!! np must be even (the reference's normal() fills an even count).
program ref_init
  use random, only : uniform, normal
  implicit none
  real(8), parameter :: c  = 2.99792458d+08
  real(8), parameter :: kb = 1.38064852d-23
  real(8), parameter :: mp = 1.672621898d-27
  character(len=256) :: arg, fname
  integer :: np, i, u
  real(8) :: b(6), temp, vth
  real(8), allocatable :: state(:,:,:)

  call get_command_argument(1, arg)
  read(arg, *) np
  do i = 1, 6
    call get_command_argument(1 + i, arg)
    read(arg, *) b(i)
  end do
  call get_command_argument(8, arg)
  read(arg, *) temp
  call get_command_argument(9, fname)
  vth = sqrt(kb * temp / mp) / c
  allocate(state(8, 2, np))
  state = 0.0d0
  call uniform(real(1.0, kind=8), state(1:3,1,1:np))
  state(1,1,1:np) = (b(2) - b(1)) * state(1,1,1:np) + b(1)
  state(2,1,1:np) = (b(4) - b(3)) * state(2,1,1:np) + b(3)
  state(3,1,1:np) = (b(6) - b(5)) * state(3,1,1:np) + b(5)
  call normal(real(0.0, kind=8), vth, state(4:6,1,1:np))
  open(newunit=u, file=trim(fname), access='stream', form='unformatted', status='replace')
  write(u) state(1:6, 1, 1:np)
  close(u)
end program ref_init
