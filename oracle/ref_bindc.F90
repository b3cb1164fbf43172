!! Test infrastructure only: a C-callable harness around the reference's own
!! hot-path modules (interpolations, accelerations, integrations), compiled
!! from the reference sources where they lie.  It fills nothing by itself and
!! never reads files: fields, parameters and initial states arrive as plain
!! arrays, which sidesteps the reference's FITS reader (its FP64 read path
!! calls the float32 reader, sources/fitsio.F90:258-261).
!!
!! Entry points mirror the reference call sites:
!!   ref_interpolate_lin  -> interpolate_lin   (sources/interpolations.F90:104)
!!   ref_acceleration     -> acceleration      (sources/accelerations.F90:59)
!!   ref_dti_guess        -> dti_guess         (sources/integrations.F90:524)
!!   ref_integrate        -> the particle loop (sources/particles.F90:337-356),
!!                           OpenMP static schedule like particles.F90:339
!!
!! Only tests/, the fixture generator and the CPU-baseline leg of the benchmark
!! load the library built from this file; nothing in the product calls it.
module ref_bindc

  use iso_c_binding
  use interpolations, only : interpolate_lin
  use accelerations,  only : acceleration
  use integrations,   only : integrate_rk5, integrate_dp853, dti_guess

  implicit none

contains

  subroutine ref_interpolate_lin(dm, xl, xs, uu, bb, pos, u, b)               &
                                 bind(C, name="ref_interpolate_lin")
    integer(c_int), dimension(3), intent(in)  :: dm
    real(c_double), dimension(3), intent(in)  :: xl, xs, pos
    real(c_double), dimension(dm(1),dm(2),dm(3),3), intent(in) :: uu, bb
    real(c_double), dimension(3), intent(out) :: u, b
    call interpolate_lin(dm, xl, xs, uu, bb, pos, u, b)
  end subroutine ref_interpolate_lin

  subroutine ref_acceleration(qom, vun, dm, xl, xs, uu, bb, vec, acc)         &
                              bind(C, name="ref_acceleration")
    real(c_double), value                     :: qom, vun
    integer(c_int), dimension(3), intent(in)  :: dm
    real(c_double), dimension(3), intent(in)  :: xl, xs
    real(c_double), dimension(dm(1),dm(2),dm(3),3), intent(in) :: uu, bb
    real(c_double), dimension(6), intent(in)  :: vec
    real(c_double), dimension(6), intent(out) :: acc
    call acceleration(qom, vun, dm, xl, xs, uu, bb, vec, acc)
  end subroutine ref_acceleration

  function ref_dti_guess(iord, dm, params, uu, bb, si) result(dti)            &
                         bind(C, name="ref_dti_guess")
    integer(c_int), value                     :: iord
    integer(c_int), dimension(3), intent(in)  :: dm
    real(c_double), dimension(32), intent(in) :: params
    real(c_double), dimension(dm(1),dm(2),dm(3),3), intent(in) :: uu, bb
    real(c_double), dimension(6), intent(in)  :: si
    real(c_double)                            :: dti
    dti = dti_guess(iord, dm, params, uu, bb, si)
  end function ref_dti_guess

  ! method: 0 = 'rk5', 1 = 'dop853'/'dp853'/'rk8'
  subroutine ref_integrate(method, np, m, dm, params, time, uu, bb, state, cnt) &
                           bind(C, name="ref_integrate")
    integer(c_int), value                     :: method, m
    integer(c_int64_t), value                 :: np
    integer(c_int), dimension(3), intent(in)  :: dm
    real(c_double), dimension(32), intent(in) :: params
    real(c_double), dimension(m), intent(in)  :: time
    real(c_double), dimension(dm(1),dm(2),dm(3),3), intent(in) :: uu, bb
    real(c_double), dimension(8,m,np), intent(inout) :: state
    integer(c_int64_t), dimension(3,np), intent(out) :: cnt
    integer(c_int64_t) :: n

    if (method == 0) then
!$omp parallel do
      do n = 1, np
        call integrate_rk5(m, dm, params, time, uu, bb, state(:,:,n), cnt(:,n))
      end do
    else
!$omp parallel do
      do n = 1, np
        call integrate_dp853(m, dm, params, time, uu, bb, state(:,:,n), cnt(:,n))
      end do
    end if
  end subroutine ref_integrate

  integer(c_int) function ref_num_threads() bind(C, name="ref_num_threads")
!$  integer, external :: omp_get_max_threads
    ref_num_threads = 1
!$  ref_num_threads = omp_get_max_threads()
  end function ref_num_threads

end module ref_bindc
