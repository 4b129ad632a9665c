! Writes and reads orbital files with the Fortran runtime's own unformatted
! sequential records, in the layout of S/loadstuff.f90 (header record
! dims(3), num, cflag; then the complex*16 orbital array), so the Python
! reader/writer can be checked against a real Fortran compiler.
!   spfio_check write FILE   -> dims 3 4 5, 2 orbitals, a(i,j) = (i + 100 j, -i)
!   spfio_check read FILE    -> expects dims 3 4 5, 2 orbitals, a(i,j) = (i + 1000 j, i / 2)
program spfio_check
  implicit none
  integer :: dims(3), num, cflag, i, j, iu
  complex*16, allocatable :: a(:,:)
  character(len=512) :: mode, path
  call get_command_argument(1, mode)
  call get_command_argument(2, path)
  if (trim(mode) == 'write') then
     dims = (/ 3, 4, 5 /)
     num = 2
     cflag = 1
     allocate(a(60, 2))
     do j = 1, 2
        do i = 1, 60
           a(i, j) = cmplx(dble(i + 100 * j), dble(-i), kind=8)
        end do
     end do
     open(newunit=iu, file=trim(path), form='unformatted', status='replace')
     write(iu) dims, num, cflag
     write(iu) a
     close(iu)
  else
     open(newunit=iu, file=trim(path), form='unformatted', status='old')
     read(iu) dims, num, cflag
     if (dims(1) /= 3 .or. dims(2) /= 4 .or. dims(3) /= 5 .or. num /= 2 .or. cflag /= 1) stop 2
     allocate(a(dims(1) * dims(2) * dims(3), num))
     read(iu) a
     close(iu)
     do j = 1, 2
        do i = 1, 60
           if (a(i, j) /= cmplx(dble(i + 1000 * j), dble(i) / 2d0, kind=8)) stop 3
        end do
     end do
  end if
  print '(a)', 'spfio ok'
end program spfio_check
