! Conformance of the Fortran drop-in (v1-1_amd/fortran/sincdvr_shim.f90):
! calls op_tinv, call_twoe_matel00, mult_ke and mult_reducedpot with the
! argument shapes of their reference call sites and checks the results
! against each other and against direct Fortran sums:
!   call_twoe_matel00(1, N, spfs, spfs, ...)   S/matel.f90:95   (same array twice)
!   call_twoe_matel00(2, 3, ...)               a lowspf:highspf sub-range
!   op_tinv(den, red, N, N, ...)               SD/coreproject.f90:1078
!   mult_ke(spfs, out, N, timingdir, notiming) S/matel.f90:264
!   mult_reducedpot(1, N, spfs, out, rp)       S/derivs.f90:763
! in the hermnorm build, the c-norm build (CONJUGATE empty) and with smooth
! exterior complex scaling, and the timing side effects: op_tinv's counters
! and twoematel.time.dat under notiming = 0 in the directory given as the
! first command-line argument.
program conformance
  use tinvsubmod
  implicit none
  integer, parameter :: n = 21, norb = 4, n3 = n**3
  complex*16, allocatable :: spfs(:,:), tm(:,:,:,:), tr(:,:,:), den(:,:), red(:,:), keo(:,:), mf(:,:), spf2(:,:)
  complex*16, allocatable :: tm2(:,:,:,:), tr2(:,:,:), tms(:,:,:,:), trs(:,:,:), keo2(:,:)
  complex*16 :: s, ref
  real*8 :: x, y, z, err, scl, pts(n)
  integer :: i, j, k, a, b, a2, b2, p, t1, t3, t4, t5, ft(10), nlines, ios
  character(len=512) :: tdir
  character(len=200) :: line

  call get_command_argument(1, tdir)
  if (len_trim(tdir) == 0) tdir = "."
  call sincdvr_gpu_init(n, 0.4d0, norb, 0, 0, 0d0, 1d0, 1d4, 5d0, 240, 40, 0, 0, 0)
  allocate(spfs(n3, norb), spf2(n3, norb), tm(norb, norb, norb, norb), tr(n3, norb, norb), den(n3, norb), &
       red(n3, norb), keo(n3, norb), mf(n3, norb), tm2(norb, norb, norb, norb), tr2(n3, norb, norb), &
       tms(norb, norb, norb, 2:3), trs(n3, norb, 2:3), keo2(n3, norb))
  do i = 1, n
     pts(i) = i * 0.4d0
  end do
  pts = pts - (pts(1) + pts(n)) / 2d0
  do a = 1, norb
     do k = 1, n
        do j = 1, n
           do i = 1, n
              x = pts(i) - 0.3d0 * a; y = pts(j) + 0.2d0 * a; z = pts(k) - 0.1d0
              spfs(i + n * (j - 1) + n * n * (k - 1), a) = exp(-(0.5d0 + 0.2d0 * a) * (x * x + y * y + z * z)) * &
                   cmplx(cos(0.3d0 * a * x), sin(0.2d0 * y + a), kind=8)
           end do
        end do
     end do
  end do

  ! two-electron matrix elements with the same array as bra and ket; timed
  ! (notiming = 0): the first call of the run writes the header of
  ! twoematel.time.dat, every timed call one line (SD/coreproject.f90:1112-1142)
  call call_twoe_matel00(1, norb, spfs, spfs, tm, tr, tdir, 0)

  ! op_tinv on the densities of each spf2b batch reproduces twoereduced; it
  ! adds its time to times4 only
  err = 0; scl = maxval(abs(tr))
  t1 = 11; t3 = 13; t4 = 0; t5 = 15; ft = 17
  do b2 = 1, norb
     do a2 = 1, norb
        den(:, a2) = spfs(:, a2) * conjg(spfs(:, b2))
     end do
     call op_tinv(den, red, norb, norb, t1, t3, t4, t5, ft)
     err = max(err, maxval(abs(red - tr(:, :, b2))) / scl)
  end do
  print '(A,ES10.3)', "op_tinv vs twoereduced        rel err ", err
  if (err > 1d-12) stop 2
  if (t1 /= 11 .or. t3 /= 13 .or. t5 /= 15 .or. any(ft /= 17) .or. t4 < 0) stop 20
  print '(A,I6)', "op_tinv times4 (ms)           ", t4

  ! a lowspf:highspf sub-range gives the matching slices of the full call
  call call_twoe_matel00(2, 3, spfs, spfs, tms, trs, "timing", 1)
  err = max(maxval(abs(tms - tm(:, :, :, 2:3))) / maxval(abs(tm)), maxval(abs(trs - tr(:, :, 2:3))) / scl)
  print '(A,ES10.3)', "sub-range 2:3 vs full         rel err ", err
  if (err > 1d-12) stop 21

  ! one more timed call: the header and two lines in twoematel.time.dat
  call call_twoe_matel00(1, norb, spfs, spfs, tm2, tr2, tdir, 0)
  if (maxval(abs(tm2 - tm)) > 0d0) stop 25
  open(31, file=trim(tdir)//"/twoematel.time.dat", status="old", iostat=ios)
  if (ios /= 0) stop 22
  nlines = 0
  do
     read(31, '(A)', iostat=ios) line
     if (ios /= 0) exit
     nlines = nlines + 1
     if (nlines == 1 .and. index(line, "F.T.") == 0) stop 23
  end do
  close(31)
  print '(A,I3)', "twoematel.time.dat lines      ", nlines
  if (nlines /= 3) stop 24

  ! twoematel(1a,1b,2a,2b) = sum_x spfs(x,1a) conj(spfs(x,1b)) twoereduced(x,2a,2b)
  err = 0; scl = maxval(abs(tm))
  do b2 = 1, norb
     do a2 = 1, norb
        do b = 1, norb
           do a = 1, norb
              s = sum(spfs(:, a) * conjg(spfs(:, b)) * tr(:, a2, b2))
              err = max(err, abs(s - tm(a, b, a2, b2)) / scl)
           end do
        end do
     end do
  end do
  print '(A,ES10.3)', "twoematel vs direct sum       rel err ", err
  if (err > 1d-12) stop 3

  ! the general (bra /= ket) path agrees with the symmetric one
  spf2 = spfs
  call call_twoe_matel00(1, norb, spfs, spf2, tm2, tr2, "timing", 1)
  err = maxval(abs(tm2 - tm)) / maxval(abs(tm))
  print '(A,ES10.3)', "general vs symmetric pairs    rel err ", err
  if (err > 1d-12) stop 4

  ! kinetic energy is Hermitian on the grid: <a|T b> = conj(<b|T a>)
  call mult_ke(spfs, keo, norb, "timing", 1)
  err = 0
  do b = 1, norb
     do a = 1, norb
        s = sum(conjg(spfs(:, a)) * keo(:, b))
        ref = conjg(sum(conjg(spfs(:, b)) * keo(:, a)))
        err = max(err, abs(s - ref) / abs(ref))
     end do
  end do
  print '(A,ES10.3)', "mult_ke hermiticity           rel err ", err
  if (err > 1d-12) stop 5

  ! mean-field apply with reducedpot = twoereduced: out(:,i) = sum_k rp(:,k,i) spfs(:,k)
  call mult_reducedpot(1, norb, spfs, mf, tr)
  err = 0
  do p = 1, norb
     err = max(err, maxval(abs(mf(:, p) - sum(tr(:, :, p) * spfs, dim=2))))
  end do
  err = err / maxval(abs(mf))
  print '(A,ES10.3)', "mult_reducedpot vs direct sum rel err ", err
  if (err > 1d-12) stop 6

  ! c-norm build: densities and matrix elements without conjugation
  call sincdvr_gpu_init(n, 0.4d0, norb, 1, 0, 0d0, 1d0, 1d4, 5d0, 240, 40, 0, 0, 0)
  call call_twoe_matel00(1, norb, spfs, spfs, tm, tr, "timing", 1)
  err = 0; scl = maxval(abs(tr))
  do b2 = 1, norb
     do a2 = 1, norb
        den(:, a2) = spfs(:, a2) * spfs(:, b2)
     end do
     call op_tinv(den, red, norb, norb, t1, t3, t4, t5, ft)
     err = max(err, maxval(abs(red - tr(:, :, b2))) / scl)
  end do
  scl = maxval(abs(tm))
  do b2 = 1, norb
     do a2 = 1, norb
        do b = 1, norb
           do a = 1, norb
              s = sum(spfs(:, a) * spfs(:, b) * tr(:, a2, b2))
              err = max(err, abs(s - tm(a, b, a2, b2)) / scl)
           end do
        end do
     end do
  end do
  print '(A,ES10.3)', "cnorm twoematel, twoereduced  rel err ", err
  if (err > 1d-12) stop 7

  ! exterior complex scaling: ramp beyond the grid leaves mult_ke unchanged;
  ! on the grid the operator J^-1 K J^-1 + scalediag is complex symmetric
  call sincdvr_gpu_init(n, 0.4d0, norb, 1, 0, 0d0, 1d0, 1d4, 5d0, 240, 40, 0, 0, 0)
  call mult_ke(spfs, keo, norb, "timing", 1)
  call sincdvr_gpu_init(n, 0.4d0, norb, 1, 1, 0.3d0, 1d0, 50d0, 5d0, 240, 40, 0, 0, 0)
  call mult_ke(spfs, keo2, norb, "timing", 1)
  err = maxval(abs(keo2 - keo)) / maxval(abs(keo))
  print '(A,ES10.3)', "ECS off the grid vs plain     rel err ", err
  if (err > 1d-14) stop 8
  call sincdvr_gpu_init(n, 0.4d0, norb, 1, 1, 0.3d0, 1d0, 2.5d0, 1.5d0, 240, 40, 0, 0, 0)
  call mult_ke(spfs, keo2, norb, "timing", 1)
  if (maxval(abs(keo2 - keo)) < 1d-6 * maxval(abs(keo))) stop 9      ! the ramp must act
  err = 0
  do b = 1, norb
     do a = 1, norb
        s = sum(spfs(:, a) * keo2(:, b))
        ref = sum(spfs(:, b) * keo2(:, a))
        err = max(err, abs(s - ref) / abs(ref))
     end do
  end do
  print '(A,ES10.3)', "ECS mult_ke complex symmetry  rel err ", err
  if (err > 1d-12) stop 10
  print '(A)', "CONFORMANCE OK"
end program conformance
