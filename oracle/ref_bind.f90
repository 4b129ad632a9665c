! ref_bind.f90 -- C entry points onto the reference's own convolution code,
! compiled together with it into oracle/_ref/libref_conv.so.
!
! TEST INFRASTRUCTURE ONLY (the checker and the CPU baseline): nothing of the
! product library links it.  The compute is the reference's: circ3d_sub and
! myzfft3d (SD/toeplitz_subs.f90:264-294,561-589), myzfft1d with the DFFTPACK
! backend (CP/ftcore.f90:137-157) and the vendored DFFTPACK, all compiled
! from the reference sources where they lie.  What this file adds is the
! framing of op_tinv_notscaled around circ3d_sub (SD/coreproject.f90:449-450,
! 509-521): the density is placed in the first octant of the zero-filled
! (2n)^3 array, circ3d_sub convolves circsize of them at a time with the
! kernel array threed_two(-n:n-1)^3 (SD/sincproject.f90:110), and the second
! octant of the result is the pair potential.

subroutine ref_op_tinv(n, threed, den, red, allsize, circsize) bind(C, name="ref_op_tinv")
  use iso_c_binding
  use circ3dsubmod
  implicit none
  integer(c_int), value :: n, allsize, circsize
  complex(c_double_complex), intent(in) :: threed(2*n, 2*n, 2*n, 1, 1, 1)
  complex(c_double_complex), intent(in) :: den(n, n, n, allsize)
  complex(c_double_complex), intent(out) :: red(n, n, n, allsize)
  complex(c_double_complex), allocatable :: big(:, :, :, :), back(:, :, :, :)
  integer :: c0, lo, hi
  allocate(big(2*n, 2*n, 2*n, circsize), back(2*n, 2*n, 2*n, circsize))
  do c0 = 0, allsize / circsize - 1
     lo = c0 * circsize + 1
     hi = lo + circsize - 1
     big(:, :, :, :) = 0
     big(1:n, 1:n, 1:n, :) = den(:, :, :, lo:hi)
     call circ3d_sub(threed, big, back, n, circsize)
     red(:, :, :, lo:hi) = back(n+1:2*n, n+1:2*n, n+1:2*n, :)
  enddo
  deallocate(big, back)
end subroutine ref_op_tinv

subroutine ref_set_threads(nt) bind(C, name="ref_set_threads")
  use iso_c_binding
  use omp_lib
  implicit none
  integer(c_int), value :: nt
  call omp_set_num_threads(nt)
end subroutine ref_set_threads

function ref_num_threads() bind(C, name="ref_num_threads") result(nt)
  use iso_c_binding
  use omp_lib
  implicit none
  integer(c_int) :: nt
  nt = omp_get_max_threads()
end function ref_num_threads
