! sincdvr_shim.f90 -- the SINCDVR project subroutine interface of
! LBNL-AMO-MCTDHF, forwarded through iso_c_binding to the gfx950 library
! (include/sincdvr.h).  Linked in place of the project's own objects, it lets
! the unchanged MCTDHF driver call:
!
!   module tinvsubmod :: op_tinv            SD/coreproject.f90:250-258
!   subroutine mult_ke                      SD/coreproject.f90:2104-2116
!   subroutine call_twoe_matel00            SD/coreproject.f90:945-955
!   subroutine mult_reducedpot              SD/coreproject.f90:1200-1205
!
! with the reference's argument lists (DATATYPE = complex*16 build flavours).
! The project state the reference keeps in modules myparams/myprojectmod
! (grid, spacing, flags, kernel) lives in the library context created by
! sincdvr_gpu_init, which the project's init_project would call once.
! A nonzero library status prints the library message and stops, as the
! reference's CFLST does (S/Definitions.ALL:45).
!
! Timing side effects, as the reference's routines have them: op_tinv adds
! its elapsed milliseconds to the caller's times4 counter (the F.T. column;
! the padded-array fill and the MPI exchanges that times1/3/5 count have no
! counterpart here, and the non-MPI reference leaves fttimes alone);
! call_twoe_matel00 keeps the reference's per-call times(1:7) and, with
! notiming = 0, writes timingdir/twoematel.time.dat: the header line on the
! first call, then one line per call (SD/coreproject.f90:1103-1142) with the
! columns etc, 1den, 1mpi, F.T., 2mpi, dot, ALL and the seven fttimes (zero
! without MPI), in milliseconds.  The F.T. and dot columns come from the
! library's per-kernel event timing (sdv_profile, on only while notiming = 0);
! the density formation runs inside the x-forward kernel, so 1den is 0 and
! the densities count under F.T.; etc is the rest of the call (transfers).

module sincdvr_capi
  use, intrinsic :: iso_c_binding
  implicit none

  type, bind(c) :: sdv_config
     integer(c_int) :: numpoints
     real(c_double) :: spacing
     integer(c_int) :: toothnbig, toothnsmall, numspf, cnormflag, scalingflag
     real(c_double) :: scalingtheta, scalingstretch, scalingdistance, smoothness
     integer(c_int) :: notwoflag, debugflag, precision
     integer(c_int) :: maskflag, masknumpoints, velflag
  end type sdv_config

  ! maskflag / masknumpoints of sinc_params (SD/sinc_params.f90:73-74), set by
  ! sincdvr_gpu_set_mask before sincdvr_gpu_init
  integer :: sdv_maskflag = 0, sdv_masknumpoints = 0

  interface
     subroutine sdv_config_default(cfg) bind(c, name="sdv_config_default")
       import :: sdv_config
       type(sdv_config), intent(out) :: cfg
     end subroutine sdv_config_default
     integer(c_int) function sdv_init(cfg, ctx) bind(c, name="sdv_init")
       import :: sdv_config, c_int, c_ptr
       type(sdv_config), intent(in) :: cfg
       type(c_ptr), intent(out) :: ctx
     end function sdv_init
     integer(c_int) function sdv_free(ctx) bind(c, name="sdv_free")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function sdv_free
     type(c_ptr) function sdv_last_error() bind(c, name="sdv_last_error")
       import :: c_ptr
     end function sdv_last_error
     integer(c_int) function sdv_op_tinv(ctx, den, red, allsize, circsize) bind(c, name="sdv_op_tinv")
       import :: c_int, c_ptr, c_double_complex
       type(c_ptr), value :: ctx
       complex(c_double_complex), intent(in) :: den(*)
       complex(c_double_complex), intent(out) :: red(*)
       integer(c_int), value :: allsize, circsize
     end function sdv_op_tinv
     integer(c_int) function sdv_call_twoe_matel00(ctx, low, high, p10, p20, tm, tr) &
          bind(c, name="sdv_call_twoe_matel00")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx, p10, p20, tm, tr
       integer(c_int), value :: low, high
     end function sdv_call_twoe_matel00
     integer(c_int) function sdv_mult_ke(ctx, in, out, howmany) bind(c, name="sdv_mult_ke")
       import :: c_int, c_ptr, c_double_complex
       type(c_ptr), value :: ctx
       complex(c_double_complex), intent(in) :: in(*)
       complex(c_double_complex), intent(out) :: out(*)
       integer(c_int), value :: howmany
     end function sdv_mult_ke
     integer(c_int) function sdv_profile(ctx, enable) bind(c, name="sdv_profile")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: enable
     end function sdv_profile
     integer(c_int) function sdv_profile_read(ctx, ms, cnt) bind(c, name="sdv_profile_read")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: ms(*)
       integer(c_int), intent(out) :: cnt(*)
     end function sdv_profile_read
     integer(c_int) function sdv_mult_reducedpot(ctx, first, last, in, out, rp) bind(c, name="sdv_mult_reducedpot")
       import :: c_int, c_ptr, c_double_complex
       type(c_ptr), value :: ctx
       integer(c_int), value :: first, last
       complex(c_double_complex), intent(in) :: in(*), rp(*)
       complex(c_double_complex), intent(out) :: out(*)
     end function sdv_mult_reducedpot
  end interface

  type(c_ptr), save :: sdv_ctx_ptr = c_null_ptr
  integer, save :: sdv_totpoints = -1, sdv_numspf = -1

  integer, parameter :: sdv_nstages = 7     ! SDV_NSTAGES: xfwd, yz, xinv, matel assemble/gemm/reduce, ke

contains

  ! Wall clock in milliseconds of the day, the unit of the reference's timing
  ! counters (clockmod's myclock, S/parameters.f90:30-35).
  integer function sdv_clock_ms()
    integer :: v(8)
    call date_and_time(values=v)
    sdv_clock_ms = v(8) + 1000 * (v(7) + 60 * (v(6) + 60 * v(5)))
  end function sdv_clock_ms

  ! timingdir(1:getlen(timingdir)): up to the first blank (S/utils.f90:38-53)
  integer function sdv_dirlen(dir)
    character(*), intent(in) :: dir
    sdv_dirlen = index(dir, " ") - 1
    if (sdv_dirlen < 0) sdv_dirlen = len(dir)
  end function sdv_dirlen

  subroutine sdv_check(rc, where)
    integer(c_int), intent(in) :: rc
    character(*), intent(in) :: where
    character(kind=c_char), pointer :: msg(:)
    integer :: i
    if (rc == 0) return
    call c_f_pointer(sdv_last_error(), msg, [512])
    write(*, '(A,A,I4,A)', advance='no') where, " failed (", rc, "): "
    do i = 1, 512
       if (msg(i) == c_null_char) exit
       write(*, '(A)', advance='no') msg(i)
    end do
    write(*, *)
    stop 1
  end subroutine sdv_check

end module sincdvr_capi


! Project set-up: the grid, flags and kernel of myparams / myprojectmod.
subroutine sincdvr_gpu_init(numpoints, spacing, numspf, cnormflag, scalingflag, scalingtheta, &
     scalingstretch, scalingdistance, smoothness, toothnbig, toothnsmall, notwoflag, debugflag, precision)
  use sincdvr_capi
  implicit none
  integer, intent(in) :: numpoints, numspf, cnormflag, scalingflag, toothnbig, toothnsmall, &
       notwoflag, debugflag, precision
  real*8, intent(in) :: spacing, scalingtheta, scalingstretch, scalingdistance, smoothness
  type(sdv_config) :: cfg
  if (c_associated(sdv_ctx_ptr)) call sdv_check(sdv_free(sdv_ctx_ptr), "sdv_free")
  call sdv_config_default(cfg)
  cfg%numpoints = numpoints; cfg%spacing = spacing; cfg%numspf = numspf
  cfg%cnormflag = cnormflag; cfg%scalingflag = scalingflag; cfg%scalingtheta = scalingtheta
  cfg%scalingstretch = scalingstretch; cfg%scalingdistance = scalingdistance; cfg%smoothness = smoothness
  cfg%toothnbig = toothnbig; cfg%toothnsmall = toothnsmall; cfg%notwoflag = notwoflag
  cfg%debugflag = debugflag; cfg%precision = precision
  cfg%maskflag = sdv_maskflag; cfg%masknumpoints = sdv_masknumpoints
  call sdv_check(sdv_init(cfg, sdv_ctx_ptr), "sdv_init")
  sdv_totpoints = numpoints**3
  sdv_numspf = numspf
end subroutine sincdvr_gpu_init

! The kinetic apply's edge mask (maskflag, masknumpoints of sinc_params), for
! the next sincdvr_gpu_init; used when scalingflag is 0.
subroutine sincdvr_gpu_set_mask(maskflag, masknumpoints)
  use sincdvr_capi
  implicit none
  integer, intent(in) :: maskflag, masknumpoints
  sdv_maskflag = maskflag
  sdv_masknumpoints = masknumpoints
end subroutine sincdvr_gpu_set_mask


module tinvsubmod
contains
  ! SD/coreproject.f90:250-258.  The timing counters are the caller's; the
  ! library keeps its own (sdv_profile).
  subroutine op_tinv(twoeden03, twoereduced, allsize, circsize, times1, times3, times4, times5, fttimes)
    use sincdvr_capi
    implicit none
    integer, intent(in) :: allsize, circsize
    integer, intent(inout) :: times1, times3, times4, times5, fttimes(10)
    complex*16, intent(in) :: twoeden03(sdv_totpoints, allsize)
    complex*16, intent(out) :: twoereduced(sdv_totpoints, allsize)
    integer :: itime
    itime = sdv_clock_ms()
    call sdv_check(sdv_op_tinv(sdv_ctx_ptr, twoeden03, twoereduced, allsize, circsize), "op_tinv")
    times4 = times4 + sdv_clock_ms() - itime
  end subroutine op_tinv
end module tinvsubmod


subroutine mult_ke(in, out, howmany, timingdir, notiming)
  use sincdvr_capi
  implicit none
  integer, intent(in) :: howmany, notiming
  character, intent(in) :: timingdir*(*)
  complex*16, intent(in) :: in(sdv_totpoints, howmany)
  complex*16, intent(out) :: out(sdv_totpoints, howmany)
  call sdv_check(sdv_mult_ke(sdv_ctx_ptr, in, out, howmany), "mult_ke")
end subroutine mult_ke


subroutine call_twoe_matel00(lowspf, highspf, inspfs10, inspfs20, twoematel, twoereduced, timingdir, notiming)
  use sincdvr_capi
  implicit none
  integer, intent(in) :: notiming, lowspf, highspf
  complex*16, target, intent(in) :: inspfs10(sdv_totpoints, sdv_numspf), inspfs20(sdv_totpoints, sdv_numspf)
  complex*16, target, intent(out) :: twoematel(sdv_numspf, sdv_numspf, sdv_numspf, lowspf:highspf), &
       twoereduced(sdv_totpoints, sdv_numspf, lowspf:highspf)
  character, intent(in) :: timingdir*(*)
  type(c_ptr) :: p10, p20
  integer, save :: xcount = 0
  integer :: times(10), fttimes(10), firsttime, myiostat
  real(c_double) :: ms(sdv_nstages)
  integer(c_int) :: cnt(sdv_nstages)
  if (highspf < lowspf) return      ! SD/coreproject.f90:969-972
  times(:) = 0; fttimes(:) = 0      ! per call, as the reference zeroes them
  firsttime = sdv_clock_ms()
  if (notiming == 0) call sdv_check(sdv_profile(sdv_ctx_ptr, 1), "sdv_profile")
  p10 = c_loc(inspfs10)
  p20 = c_loc(inspfs20)     ! the same address when the caller passes one array twice
  call sdv_check(sdv_call_twoe_matel00(sdv_ctx_ptr, lowspf, highspf, p10, p20, c_loc(twoematel), &
       c_loc(twoereduced)), "call_twoe_matel00")
  times(7) = sdv_clock_ms() - firsttime
  if (notiming == 0) then
     call sdv_check(sdv_profile_read(sdv_ctx_ptr, ms, cnt), "sdv_profile_read")
     call sdv_check(sdv_profile(sdv_ctx_ptr, 0), "sdv_profile")
     times(4) = nint(ms(1) + ms(2) + ms(3))            ! F.T.: the three convolution kernels
     times(6) = nint(ms(4) + ms(5) + ms(6))            ! dot: the matrix-element reduction
     times(1) = max(0, times(7) - times(4) - times(6))  ! etc
  end if
  xcount = xcount + 1
  if (notiming == 0) then
     if (xcount == 1) then
        open(853, file=timingdir(1:sdv_dirlen(timingdir))//"/twoematel.time.dat", status="unknown", &
             iostat=myiostat)
        if (myiostat == 0) then
           write(853, '(100A11)') "etc", "1den", "1mpi", "F.T.", "2mpi", "dot", "ALL"
           close(853)
        end if
     end if
     open(853, file=timingdir(1:sdv_dirlen(timingdir))//"/twoematel.time.dat", status="unknown", &
          position="append", iostat=myiostat)
     if (myiostat /= 0) then
        write(*, *) "opening twoematel timing sincdvr failed: ", timingdir(1:sdv_dirlen(timingdir))
        stop 1
     end if
     write(853, '(100I11)') times(1:7), fttimes(1:7)
     close(853)
  end if
end subroutine call_twoe_matel00


subroutine mult_reducedpot(firstspf, lastspf, inspfs, outspfs, reducedpot)
  use sincdvr_capi
  implicit none
  integer, intent(in) :: firstspf, lastspf
  complex*16, intent(out) :: outspfs(sdv_totpoints, firstspf:lastspf)
  complex*16, intent(in) :: inspfs(sdv_totpoints, sdv_numspf), reducedpot(sdv_totpoints, sdv_numspf, firstspf:lastspf)
  call sdv_check(sdv_mult_reducedpot(sdv_ctx_ptr, firstspf, lastspf, inspfs, outspfs, reducedpot), "mult_reducedpot")
end subroutine mult_reducedpot
