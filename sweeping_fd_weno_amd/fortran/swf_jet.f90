! swf_jet.f90 - the Mach 2000 jet program on the MI355X library.
!
! Keeps the parameter surface, initial data, time loop, progress lines and output of
! the reference program ex5_mach2000_jet/main.f90 (the half domain [0,1] x [0,0.25]
! with reflection at y = 0; log-scaled density, pressure and p/rho written mirrored
! to [-0.25, 0.25] in den.dat / pre.dat / m1.dat; the sweep statistics) and replaces
! its time-stepping procedures (`rk` and callees) by the C-ABI of libswf.so.
!
!   swf_jet [nx ny precision time_terminal]
!
! Defaults are the program's 800 x 200 and T = 0.001; precision 0 = FP64-exact
! (bit-identical to the reference), 1 = FP64, 2 = FP32-bulk.  Compiled like the reference, with
! -fdefault-real-8: the d-exponent literals (loop test, percentage) are binary128.
program swf_jet
   use swf
   use iso_c_binding
   implicit none
   ! parameter surface (ex5_mach2000_jet/main.f90:4-5, :597)
   integer, parameter :: swp_num = 20000
   real(8), parameter :: gamma = 5./3., gamma1 = gamma - 1., eps = 1d-13, eps2 = 1d-8
   real(8), parameter :: ujet = 800.
   integer :: nx, ny, total_x, total_y, prec
   real(8) :: time_terminal
   character(len=64) :: arg
   real(8), allocatable :: u(:, :, :)
   real(8) :: dx, dy, time, xleft, xright, yleft, yright, avg
   real(8) :: lrho, lm, ln, le, pres
   integer :: i, j, rc, nargs, kcount
   integer(c_int64_t) :: steps, nstep
   integer(c_int64_t), allocatable :: tq(:), tn(:), tp(:)
   type(swf_config) :: cfg
   type(swf_sweep_stats) :: st
   type(c_ptr) :: ctx

   nx = 800
   ny = 200
   prec = SWF_FP64_EXACT
   time_terminal = 0.001
   nargs = command_argument_count()
   if (nargs >= 1) then
      call get_command_argument(1, arg)
      read (arg, *) nx
   end if
   if (nargs >= 2) then
      call get_command_argument(2, arg)
      read (arg, *) ny
   end if
   if (nargs >= 3) then
      call get_command_argument(3, arg)
      read (arg, *) prec
   end if
   if (nargs >= 4) then
      call get_command_argument(4, arg)
      read (arg, *) time_terminal
   end if
   total_x = nx
   total_y = ny

   ! initdata (ex5_mach2000_jet/main.f90:744-768), ending with `call bc`
   allocate (u(4, -2:nx + 3, -2:ny + 3))
   xleft = 0.
   xright = 1.
   dx = (xright - xleft)/nx
   yleft = 0.
   yright = 0.25
   dy = (yright - yleft)/ny
   u = 0.
   u(1, :, :) = 0.5
   u(4, :, :) = 0.4127/gamma1

   cfg%nx = nx
   cfg%ny = ny
   cfg%xlo = xleft
   cfg%ylo = yleft
   cfg%dx = dx
   cfg%dy = dy
   cfg%gamma = gamma
   cfg%cfl = 1./4.
   cfg%eps = eps
   cfg%eps2 = eps2
   cfg%delta = 1d-12
   cfg%bc = SWF_BC_JET_HALF
   cfg%use_limiter = 2
   cfg%swp_num = swp_num
   cfg%alpha_in_scheme = 1
   cfg%roe_eps_double = 1
   cfg%ydens_plain = 1
   cfg%eps2_quad_literal = 1
   cfg%findalpha_ex3 = 0
   cfg%dt_rule_dx53 = 0
   cfg%poscheck_le = 0
   cfg%success_ge = 1
   cfg%pbar_check = 0
   cfg%fail_stop = 0
   cfg%jet_rho = 5.
   cfg%jet_u = ujet
   cfg%amb_rho = 0.5
   cfg%amb_p = 0.4127
   cfg%jet_halfwidth = 0.05
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
   rc = swf_bc(ctx)
   call check(rc)

   ! time loop (ex5_mach2000_jet/main.f90:47-55), progress every 20 steps
   time = 0d0
   kcount = 0
   do while (time <= time_terminal - 1d-10)
      nstep = 20
      rc = swf_advance(ctx, time, time_terminal, nstep, steps)
      call check(rc)
      kcount = kcount + int(steps)
      if (steps == 20) write (6, *) kcount, time/time_terminal*100d0, '%'
      if (steps < 20) exit
   end do

   ! output (ex5_mach2000_jet/main.f90:791-879)
   rc = swf_download(ctx, u)
   call check(rc)
   open (73, file='den.dat')
   open (74, file='pre.dat')
   open (75, file='m1.dat')
   write (73, *) 'VARIABLES = "X", "Y", "DENSITY"'
   write (73, *) 'zone i=', total_x, ', j=', total_y*2
   write (74, *) 'zone i=', total_x, ', j=', total_y*2
   write (75, *) 'zone i=', total_x, ', j=', total_y*2
   do j = 1, total_y
      do i = 1, total_x
         lrho = u(1, i, total_y - j + 1)
         lm = u(2, i, total_y - j + 1)
         ln = u(3, i, total_y - j + 1)
         le = u(4, i, total_y - j + 1)
         pres = gamma1*(le - 0.5*lm**2/lrho - 0.5*ln**2/lrho)
         write (73, *) dx*(i - 0.5), dy*(j - 0.5) - 0.25, log(abs(lrho))
         write (74, *) dx*(i - 0.5), dy*(j - 0.5) - 0.25, log(abs(pres))
         write (75, *) dx*(i - 0.5), dy*(j - 0.5) - 0.25, log(abs(pres/lrho))
      end do
   end do
   do j = 1, total_y
      do i = 1, total_x
         lrho = u(1, i, j)
         lm = u(2, i, j)
         ln = u(3, i, j)
         le = u(4, i, j)
         pres = gamma1*(le - 0.5*lm**2/lrho - 0.5*ln**2/lrho)
         write (73, *) dx*(i - 0.5), dy*(j - 0.5), log(abs(lrho))
         write (74, *) dx*(i - 0.5), dy*(j - 0.5), log(abs(pres))
         write (75, *) dx*(i - 0.5), dy*(j - 0.5), log(abs(pres/lrho))
      end do
   end do
   close (73)
   close (74)
   close (75)

   rc = swf_stats(ctx, st)
   call check(rc)
   allocate (tq(max(st%sweep_count, 1_8)), tn(max(st%sweep_count, 1_8)), tp(max(st%sweep_count, 1_8)))
   nstep = swf_track(ctx, tq, tn, tp, st%sweep_count)
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
   call swf_destroy(ctx)

contains

   subroutine check(code)
      integer, intent(in) :: code
      if (code /= SWF_OK) then
         print *, swf_message(int(code, c_int))
         stop 'Program terminated due to incorrect physical conditions.'
      end if
   end subroutine check

end program swf_jet
