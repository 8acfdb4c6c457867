! swf_driver.f90 - 2D Sedov driver on the MI355X library.
!
! Keeps the parameter surface, initial data, time loop and output of the reference
! program ex4_sedov_2d/main.f90 (run as one domain, cut_x = cut_y = 1) and replaces
! its time-stepping procedures (`rk` and callees) by the C-ABI of libswf.so.
!
!   swf_driver [total_x total_y time_terminal precision outdir]
!
! precision: 0 = FP64-exact (bit-identical to the reference), 1 = FP64, 2 = FP32-bulk.
! Compiled like the reference, with -fdefault-real-8, so the loop test
! `time <= time_terminal - 1d-10` is evaluated in binary128 exactly as there.
program swf_driver
   use swf
   use iso_c_binding
   implicit none
   ! parameter surface (ex4_sedov_2d/main.f90:15-20,50; cfl of `rk`, :687)
   integer, parameter :: total_x_default = 640, total_y_default = 640
   real(8), parameter :: time_terminal_default = 1, gamma = 1.4, eps = 1d-13, eps2 = 1d-8, xm = 1.3, ym = 1.3
   real(8), parameter :: cfl = 0.5, delta = 1d-12
   integer, parameter :: use_limiter = 2, swp_num = 20000
   integer :: total_x, total_y, prec
   real(8) :: time_terminal
   character(len=256) :: arg, output_dir
   real(8), allocatable :: u(:, :, :)
   real(8) :: dx, dy, time, gamma1, lrho, lm, ln, le, pres, avg
   integer :: i, j, rc, nargs
   integer(c_int64_t) :: kcount, steps, n
   integer(c_int64_t), allocatable :: tq(:), tn(:), tp(:)
   integer(8) :: c0, c1, crate
   type(swf_config) :: cfg
   type(swf_sweep_stats) :: st
   type(c_ptr) :: ctx

   total_x = total_x_default
   total_y = total_y_default
   time_terminal = time_terminal_default
   prec = SWF_FP64_EXACT
   output_dir = ''
   nargs = command_argument_count()
   if (nargs >= 2) then
      call get_command_argument(1, arg)
      read (arg, *) total_x
      call get_command_argument(2, arg)
      read (arg, *) total_y
   end if
   if (nargs >= 3) then
      call get_command_argument(3, arg)
      read (arg, *) time_terminal
   end if
   if (nargs >= 4) then
      call get_command_argument(4, arg)
      read (arg, *) prec
   end if
   if (nargs >= 5) call get_command_argument(5, output_dir)

   ! initdata (ex4_sedov_2d/main.f90:1121-1189, one domain)
   dx = (xm - 0.)/total_x
   dy = (ym - 0.)/total_y
   allocate (u(4, -2:total_x + 3, -2:total_y + 3))
   u = 0.
   u(1, :, :) = 1
   u(4, :, :) = 1d-12
   do j = 1, total_y
      do i = 1, total_x
         if (abs(dx*(i - 0.5) - 0.0) <= dx/2.0 .and. abs(dy*(j - 0.5) - 0.0) <= dy/2.0) then
            u(4, i, j) = 0.244816/(dx*dy)
         end if
      end do
   end do

   cfg%nx = total_x
   cfg%ny = total_y
   cfg%xlo = 0.
   cfg%ylo = 0.
   cfg%dx = dx
   cfg%dy = dy
   cfg%gamma = gamma
   cfg%cfl = cfl
   cfg%eps = eps
   cfg%eps2 = eps2
   cfg%delta = delta
   cfg%bc = SWF_BC_SEDOV
   cfg%use_limiter = use_limiter
   cfg%swp_num = swp_num
   cfg%alpha_in_scheme = 0
   cfg%roe_eps_double = 0
   cfg%ydens_plain = 0
   cfg%eps2_quad_literal = 0
   cfg%findalpha_ex3 = 0
   cfg%dt_rule_dx53 = 0
   cfg%poscheck_le = 0
   cfg%success_ge = 0
   cfg%pbar_check = 1
   cfg%fail_stop = 1
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

   ! time loop (ex4_sedov_2d/main.f90:109-119): the device evaluates the loop
   ! test and the CFL step; progress every 50 steps as the reference prints it
   time = 0d0
   kcount = 0
   call system_clock(c0, crate)
   do while (time <= time_terminal - 1d-10)
      n = 50
      rc = swf_advance(ctx, time, time_terminal, n, steps)
      call check(rc)
      kcount = kcount + steps
      if (steps == 50) write (6, *) kcount, time/time_terminal*100d0, '%'
      if (steps < 50) exit
   end do
   call system_clock(c1)
   print *, 'CPU Time', real(c1 - c0, 8)/real(crate, 8), kcount, time

   ! output (ex4_sedov_2d/main.f90:1193-1354)
   rc = swf_download(ctx, u)
   call check(rc)
   gamma1 = gamma - 1.
   if (len_trim(output_dir) > 0) then
      open (73, file=trim(output_dir)//'/den.dat')
      open (74, file=trim(output_dir)//'/pre.dat')
      open (75, file=trim(output_dir)//'/m1.dat')
      open (76, file=trim(output_dir)//'/m2.dat')
      write (73, *) 'VARIABLES = "X", "Y", "DENSITY"'
      write (73, *) 'zone i=', total_x, ', j=', total_y
      write (74, *) 'zone i=', total_x, ', j=', total_y
      write (75, *) 'zone i=', total_x, ', j=', total_y
      write (76, *) 'zone i=', total_x, ', j=', total_y
      do j = 1, total_y
         do i = 1, total_x
            lrho = u(1, i, j)
            lm = u(2, i, j)
            ln = u(3, i, j)
            le = u(4, i, j)
            pres = gamma1*(le - 0.5*lm**2/lrho - 0.5*ln**2/lrho)
            write (73, *) dx*(i - 0.5), dy*(j - 0.5), lrho
            write (74, *) dx*(i - 0.5), dy*(j - 0.5), pres
            write (75, *) dx*(i - 0.5), dy*(j - 0.5), lm
            write (76, *) dx*(i - 0.5), dy*(j - 0.5), ln
         end do
      end do
      close (73)
      close (74)
      close (75)
      close (76)
   end if

   rc = swf_stats(ctx, st)
   call check(rc)
   allocate (tq(max(st%sweep_count, 1_8)), tn(max(st%sweep_count, 1_8)), tp(max(st%sweep_count, 1_8)))
   n = swf_track(ctx, tq, tn, tp, st%sweep_count)
   print *, 'total sweep = ', sum(tq(1:st%sweep_count))
   avg = real(sum(tq(1:st%sweep_count)), 8)/real(st%sweep_count, 8)
   print *, 'avg sweep = ', avg
   print *, 'max sweep = ', maxval(tq(1:st%sweep_count))
   print *, 'total negative pairs = ', sum(tn(1:st%sweep_count))
   avg = real(sum(tn(1:st%sweep_count)), 8)/real(st%sweep_count, 8)
   print *, 'avg negative pairs = ', avg
   print *, 'max negative pairs = ', maxval(tn(1:st%sweep_count))
   print *, 'total positive neighbors = ', sum(tp(1:st%sweep_count))
   avg = real(sum(tp(1:st%sweep_count)), 8)/real(st%sweep_count, 8)
   print *, 'avg positive neighbors = ', avg
   print *, 'max positive neighbors = ', maxval(tp(1:st%sweep_count))
   print *, '# of limiter', st%den_count, st%pre_count
   call swf_destroy(ctx)

contains

   subroutine check(code)
      integer, intent(in) :: code
      if (code /= SWF_OK) then
         print *, swf_message(int(code, c_int))
         stop 'Program terminated due to incorrect physical conditions.'
      end if
   end subroutine check

end program swf_driver
