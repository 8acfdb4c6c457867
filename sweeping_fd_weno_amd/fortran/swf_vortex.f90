! swf_vortex.f90 - low-density isentropic vortex accuracy driver on the MI355X library.
!
! Keeps the parameter surface, initial data, time loop and output of the reference
! program ex3_vortex_accuracy_2d/main.f90 (exact solution at t = 0 on every cell,
! periodic bc; L1/Linf density errors against the exact solution, their orders
! against the previous grid read from fort.7, the err_table.dat row, the field
! files and the sweep statistics) and replaces its time-stepping procedures (`rk`
! and callees) by the C-ABI of libswf.so.
!
!   swf_vortex [nx precision new_table use_limiter]
!
! Runs in the working directory like the reference (fort.7, err_table.dat and the
! .dat files there).  precision: 0 = FP64-exact (bit-identical to the reference),
! 1 = FP64, 2 = FP32-bulk.  new_table = 1 starts a new err_table.dat (orders 0).
! use_limiter: the reference's compile-time switch (main.f90:5), 2 = sweeping
! (as shipped), 1 = Zhang-Shu flux limiter.
! Compiled like the reference, with -fdefault-real-8: the d-exponent literals
! (loop test, log(2D0)) are binary128 exactly as there.
program swf_vortex
   use swf
   use iso_c_binding
   implicit none
   ! parameter surface (ex3_vortex_accuracy_2d/main.f90:5-9; weno5.f90:78 for pi)
   integer, parameter :: ierror = 1, swp_num = 20000
   real(8), parameter :: time_terminal = 0.01, gamma = 1.4, gamma1 = gamma - 1., eps = 1d-13
   real(8), parameter :: pi = 3.14159265358979323846264338327950
   integer :: nx, ny, n, prec, itab, use_limiter
   logical :: new_table
   character(len=64) :: arg
   real(8), allocatable :: u(:, :, :), x(:), y(:)
   real(8) :: dx, dy, time, xleft, xright, yleft, yright
   real(8) :: er1, eri, er1old, eriold, o1, oi, uue(4), avg
   integer :: i, j, rc, nargs
   integer(c_int64_t) :: kcount, steps, nstep
   integer(c_int64_t), allocatable :: tq(:), tn(:), tp(:)
   type(swf_config) :: cfg
   type(swf_sweep_stats) :: st
   type(c_ptr) :: ctx

   nx = 180
   prec = SWF_FP64_EXACT
   itab = 0
   use_limiter = 2
   nargs = command_argument_count()
   if (nargs >= 1) then
      call get_command_argument(1, arg)
      read (arg, *) nx
   end if
   if (nargs >= 2) then
      call get_command_argument(2, arg)
      read (arg, *) prec
   end if
   if (nargs >= 3) then
      call get_command_argument(3, arg)
      read (arg, *) itab
   end if
   if (nargs >= 4) then
      call get_command_argument(4, arg)
      read (arg, *) use_limiter
   end if
   new_table = itab /= 0
   ny = nx
   n = nx

   ! initdata (ex3_vortex_accuracy_2d/main.f90:670-696)
   allocate (u(4, -2:nx + 3, -2:ny + 3), x(-2:nx + 3), y(-2:ny + 3))
   xleft = 0.
   xright = 10.
   dx = (xright - xleft)/nx
   yleft = 0.
   yright = 10.
   dy = (yright - yleft)/ny
   do i = -2, nx + 3
      x(i) = xleft + i*dx - dx/2.
   end do
   do i = -2, ny + 3
      y(i) = yleft + i*dy - dy/2.
   end do
   do j = -2, ny + 3
      do i = -2, nx + 3
         u(:, i, j) = exact(x(i), y(j), 0.)
      end do
   end do

   cfg%nx = nx
   cfg%ny = ny
   cfg%xlo = xleft
   cfg%ylo = yleft
   cfg%dx = dx
   cfg%dy = dy
   cfg%gamma = gamma
   cfg%cfl = 0.5
   cfg%eps = eps
   cfg%eps2 = 1d-8
   cfg%delta = 1d-12
   cfg%bc = SWF_BC_PERIODIC
   cfg%use_limiter = use_limiter
   cfg%swp_num = swp_num
   cfg%alpha_in_scheme = 1
   cfg%roe_eps_double = 1
   cfg%ydens_plain = 1
   cfg%eps2_quad_literal = 1
   cfg%findalpha_ex3 = 1
   cfg%dt_rule_dx53 = 1
   cfg%poscheck_le = 1
   cfg%success_ge = 0
   cfg%pbar_check = 0
   cfg%fail_stop = 0
   cfg%jet_rho = 0.
   cfg%jet_u = 0.
   cfg%amb_rho = 0.
   cfg%amb_p = 0.
   cfg%jet_halfwidth = 0.
   cfg%step_total = 0
   cfg%alpha_lines = 0
   cfg%weno_eps12 = 0
   cfg%psum_check = 0
   cfg%in_rho = 0.
   cfg%in_u = 0.
   cfg%in_p = 0.
   cfg%precision = prec
   cfg%device = 0
   rc = swf_create(cfg, ctx)
   call check(rc)
   rc = swf_upload(ctx, u)
   call check(rc)
   rc = swf_bc(ctx)  ! initdata ends with `call bc`
   call check(rc)

   ! time loop (ex3_vortex_accuracy_2d/main.f90:49-55), progress every 20 steps
   time = 0d0
   kcount = 0
   do while (time <= time_terminal - 1d-10)
      nstep = 20
      rc = swf_advance(ctx, time, time_terminal, nstep, steps)
      call check(rc)
      kcount = kcount + steps
      if (steps == 20) write (6, *) kcount, time/time_terminal*100d0, '%'
      if (steps < 20) exit
   end do

   ! output (ex3_vortex_accuracy_2d/main.f90:720-837)
   rc = swf_download(ctx, u)
   call check(rc)
   er1 = 0.
   eri = 0.
   do i = 1, nx
      do j = 1, ny
         uue = exact(x(i), y(j), time)
         er1 = er1 + abs(uue(1) - u(1, i, j))
         eri = max(eri, abs(uue(1) - u(1, i, j)))
      end do
   end do
   er1 = er1/nx/ny
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
   if (new_table) then
      open (unit=10, file='err_table.dat', status='replace')
   else
      open (unit=10, file='err_table.dat', position="append", status='old')
   end if
   if (new_table) then
      o1 = 0.0
      oi = 0.0
   end if
   write (10, *) n, er1, o1, eri, oi
11 format(6e13.5)
   rewind (7)
   write (7, 11) er1, eri

   open (73, file='den.dat')
   open (74, file='pre.dat')
   open (75, file='m1.dat')
   open (76, file='m2.dat')
   open (77, file='den2.dat')
   open (78, file='pre2.dat')
   do j = 1, ny
      do i = 1, nx
         uue = exact(x(i), y(j), time)
         write (73, *) x(i), y(j), u(1, i, j), uue(1), abs(uue(1) - u(1, i, j))
         write (74, *) x(i), y(j), gamma1*(u(4, i, j) - 0.5*u(2, i, j)**2/u(1, i, j) - 0.5*u(3, i, j)**2/u(1, i, j))
         write (75, *) x(i), y(j), u(2, i, j)
         write (76, *) x(i), y(j), u(3, i, j)
      end do
   end do
   do j = 1, ny
      do i = 1, nx
         write (77, *) y(j), u(1, i, j)
         write (78, *) y(j), gamma1*(u(4, i, j) - 0.5*u(2, i, j)**2/u(1, i, j) - 0.5*u(3, i, j)**2/u(1, i, j))
      end do
   end do

   rc = swf_stats(ctx, st)
   call check(rc)
   allocate (tq(max(st%sweep_count, 1_8)), tn(max(st%sweep_count, 1_8)), tp(max(st%sweep_count, 1_8)))
   nstep = swf_track(ctx, tq, tn, tp, st%sweep_count)
   ! the reference's track arrays are default integers: maxval of none of them
   ! prints -huge(0) - 1, as there
   print *, 'total sweep = ', sum(tq(1:st%sweep_count))
   avg = real(sum(tq(1:st%sweep_count)), 8)/real(st%sweep_count, 8)
   print *, 'avg sweep = ', avg
   print *, 'max sweep = ', maxval(int(tq(1:st%sweep_count)))
   print *, 'total negative pairs = ', sum(tn(1:st%sweep_count))
   avg = real(sum(tn(1:st%sweep_count)), 8)/real(st%sweep_count, 8)
   print *, 'avg negative pairs = ', avg
   print *, 'max negative pairs = ', maxval(int(tn(1:st%sweep_count)))
   print *, 'total positive neighbors = ', sum(tp(1:st%sweep_count))
   avg = real(sum(tp(1:st%sweep_count)), 8)/real(st%sweep_count, 8)
   print *, 'avg positive neighbors = ', avg
   print *, 'max positive neighbors = ', maxval(int(tp(1:st%sweep_count)))
   print *, '# of limiter', st%den_count, st%pre_count
   call swf_destroy(ctx)

contains

   ! exact solution (ex3_vortex_accuracy_2d/main.f90:699-718); the expression is the
   ! reference's, operation for operation, so that the errors are the same numbers
   function exact(xx, yy, t)
      real(8), intent(in) :: xx, yy, t
      real(8), dimension(1:4) :: exact
      real(8) :: mag, r
      r = sqrt((xx - 5.-t)**2 + (yy - 5.-t)**2)
      mag = 10.0828
      exact(1) = (1 - gamma1*mag**2/8./gamma/pi**2*exp(1.-r**2))**(1./gamma1)
      exact(2) = exact(1)*(1.-mag/2./pi*exp((1.-r**2)/2.)*(yy - 5.))
      exact(3) = exact(1)*(1.+mag/2./pi*exp((1.-r**2)/2.)*(xx - 5.))
      exact(4) = exact(1)**gamma/gamma1 + 0.5*exact(2)**2/exact(1) + 0.5*exact(3)**2/exact(1)
   end function exact

   subroutine check(code)
      integer, intent(in) :: code
      if (code /= SWF_OK) then
         print *, swf_message(int(code, c_int))
         stop 'Program terminated due to incorrect physical conditions.'
      end if
   end subroutine check

end program swf_vortex
