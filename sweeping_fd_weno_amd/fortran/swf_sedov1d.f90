! swf_sedov1d.f90 - the 1D Sedov program on the MI355X library.
!
! Keeps the parameter surface, initial data, time loop, progress lines and output of
! the reference program ex2_sedov_1d/main.f90 (L1/Linf against its `exact`
! function, fort.7, den.dat / pre.dat / vel.dat, the sweep statistics) and replaces
! its time-stepping procedures (`rk3` with wenofd, scheme, checkpositivity and
! sweep_lim) by the swf1d_* entry points of libswf.so.
!
!   swf_sedov1d [cfl swp_num ambient max_iter precision]
!
! Defaults are the program as shipped: cfl 1.2, swp_num 50, ambient 0 (E = 1d-9/gamma1;
! 1 sets E = 1d-12), max_iter 10000, precision 0 (FP64-exact, bit-identical to the
! reference built with -fdefault-real-8; 1 = FP64, 2 = FP32-bulk).  The 1D benchmark
! case is `swf_sedov1d 0.6 2 1`.  Compiled like that build, with -fdefault-real-8:
! the d-exponent literals (loop test, progress percentage, 1d-9) are binary128 as there.
program swf_sedov1d
   use swf
   use iso_c_binding
   implicit none
   ! parameter surface (ex2_sedov_1d/main.f90:4-7; weno5.f90 for pi)
   integer, parameter :: n = 800, ierror = 0
   real(8), parameter :: gamma = 1.4, gamma1 = gamma - 1., eps = 1d-13, time_terminal = 0.001
   real(8), parameter :: pi = 3.14159265358979323846264338327950
   real(8) :: cfl, max_iter
   integer :: swp_num, ambient, prec
   character(len=64) :: arg
   real(8) :: u(3, -2:n + 3), x(-2:n + 3)
   real(8) :: dx, time, xleft, xright
   integer :: i, rc, nargs, kcount
   integer(c_int64_t) :: steps, want, nt
   integer(c_int64_t), allocatable :: tq(:), tn(:), tp(:)
   type(swf1d_config) :: cfg
   type(swf_sweep_stats) :: st
   type(c_ptr) :: ctx

   cfl = 1.2
   swp_num = 50
   ambient = 0
   max_iter = 10000
   prec = SWF_FP64_EXACT
   nargs = command_argument_count()
   if (nargs >= 1) then
      call get_command_argument(1, arg)
      read (arg, *) cfl
   end if
   if (nargs >= 2) then
      call get_command_argument(2, arg)
      read (arg, *) swp_num
   end if
   if (nargs >= 3) then
      call get_command_argument(3, arg)
      read (arg, *) ambient
   end if
   if (nargs >= 4) then
      call get_command_argument(4, arg)
      read (arg, *) max_iter
   end if
   if (nargs >= 5) then
      call get_command_argument(5, arg)
      read (arg, *) prec
   end if

   ! initdata (ex2_sedov_1d/main.f90:405-428)
   xleft = -2.
   xright = 2.
   dx = (xright - xleft)/n
   do i = -2, n + 3
      x(i) = xleft + i*dx - dx/2.
   end do
   u = 0d0
   do i = -2, n + 3
      u(1, i) = 1.
      if (ambient == 0) then
         u(3, i) = 1d-9/gamma1
      else
         u(3, i) = 1d-12
      end if
   end do
   u(3, n/2 + 1) = 3200000.0/dx

   cfg%n = n
   cfg%dx = dx
   cfg%gamma = gamma
   cfg%cfl = cfl
   cfg%eps = eps
   cfg%use_limiter = 2
   cfg%swp_num = swp_num
   cfg%precision = prec
   cfg%device = 0
   rc = swf1d_create(cfg, ctx)
   call check(rc)
   rc = swf1d_upload(ctx, u)
   call check(rc)

   ! time loop (ex2_sedov_1d/main.f90:38-54): progress every 50 steps, max_iter stop
   time = 0d0
   kcount = 0
   do while (time <= time_terminal - 1d-10)
      want = 50 - mod(kcount, 50)
      if (kcount + want > max_iter) want = int(max_iter) - kcount
      rc = swf1d_advance(ctx, time, time_terminal, want, steps)
      call check(rc)
      kcount = kcount + int(steps)
      if (kcount/50*50 .eq. kcount .and. steps > 0) write (6, *) kcount, time/time_terminal*100d0, '%'
      if (kcount .eq. max_iter) then
         rc = swf1d_download(ctx, u)
         call check(rc)
         print *, "min_rho:", minval(u(1, :))
         call output
         call swf1d_destroy(ctx)
         stop "Finished to max_iters"
      end if
      if (steps < want) exit
   end do
   rc = swf1d_download(ctx, u)
   call check(rc)
   call output
   call swf1d_destroy(ctx)

contains

   ! exact (ex2_sedov_1d/main.f90:430-439), operation for operation
   function exact(xx, t)
      real(8), intent(in) :: xx, t
      real(8), dimension(1:3) :: exact
      exact(1) = 1.+0.99999999*sin(2*pi*(xx - t))
      exact(2) = exact(1)
      exact(3) = 1./gamma1 + 0.5*exact(1)
   end function exact

   ! output (ex2_sedov_1d/main.f90:441-512)
   subroutine output
      real(8) :: er1, eri, eriold, er1old, o1, oi, uue(3), avg
      er1 = 0.
      eri = 0.
      do i = 1, n
         uue = exact(x(i), time)
         er1 = er1 + abs(uue(1) - u(1, i))
         eri = max(eri, abs(uue(1) - u(1, i)))
      end do
      er1 = er1/n
      eriold = 1.
      er1old = 1.
      if (ierror .eq. 1) then
         read (7, 11) er1old, eriold
      end if
      o1 = log(er1old/er1)/log(2D0)
      oi = log(eriold/eri)/log(2D0)
      print *, 'L1 = ', ' Linfinity ='
      print *, er1, eri
      print *, 'L1order=', ' Liorder='
      print *, o1, oi
11    format(6e13.5)
      rewind (7)
      write (7, 11) er1, eri
      open (73, file='den.dat')
      open (74, file='pre.dat')
      open (75, file='vel.dat')
      do i = 1, n
         write (73, *) x(i), u(1, i)
         write (74, *) x(i), gamma1*(u(3, i) - 0.5*u(2, i)**2/u(1, i))
         write (75, *) x(i), u(2, i)/u(1, i)
      end do
      rc = swf1d_stats(ctx, st)
      call check(rc)
      allocate (tq(max(st%sweep_count, 1_8)), tn(max(st%sweep_count, 1_8)), tp(max(st%sweep_count, 1_8)))
      nt = swf1d_track(ctx, tq, tn, tp, st%sweep_count)
      ! the reference's track arrays are default integers: maxval of none of them
      ! prints -huge(0) - 1, as there
      print *, 'total sweep = ', int(sum(tq(1:st%sweep_count)))
      avg = real(sum(tq(1:st%sweep_count)), 8)/real(st%sweep_count, 8)
      print *, 'avg sweep = ', avg
      print *, 'max sweep = ', maxval(int(tq(1:st%sweep_count)))
      print *, 'total negative pairs = ', int(sum(tn(1:st%sweep_count)))
      avg = real(sum(tn(1:st%sweep_count)), 8)/real(st%sweep_count, 8)
      print *, 'avg negative pairs = ', avg
      print *, 'max negative pairs = ', maxval(int(tn(1:st%sweep_count)))
      print *, 'total positive neighbors = ', int(sum(tp(1:st%sweep_count)))
      avg = real(sum(tp(1:st%sweep_count)), 8)/real(st%sweep_count, 8)
      print *, 'avg positive neighbors = ', avg
      print *, 'max positive neighbors = ', maxval(int(tp(1:st%sweep_count)))
      print *, '# of limiter', int(st%den_count), int(st%pre_count)
   end subroutine output

   subroutine check(code)
      integer, intent(in) :: code
      if (code /= SWF_OK) then
         print *, swf_message(int(code, c_int))
         stop 'Program terminated due to incorrect physical conditions.'
      end if
   end subroutine check

end program swf_sedov1d
