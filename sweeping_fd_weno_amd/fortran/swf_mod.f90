! swf_mod.f90 - iso_c_binding interface to libswf.so (include/swf.h).
!
! The Fortran side of the drop-in boundary: a driver keeps the reference's
! PARAMETER surface and time loop and calls these entry points in place of the
! internal procedures `rk`, `wenofd`, `findalpha`, `checkpositivity`, `sweep_lim`
! and `bc` (ex4_sedov_2d/main.f90).  The state array keeps the reference layout
! u(4, -2:nx+3, -2:ny+3); the 1D program's (swf1d_*, ex2_sedov_1d/main.f90) keeps
! u(3, -2:n+3), the detonation program's (swf7_*, ex7_detonation/main.f90) u(5, -2:nx+3, -2:ny+3).
module swf
   use iso_c_binding
   implicit none

   integer(c_int32_t), parameter :: SWF_FP64_EXACT = 0, SWF_FP64 = 1, SWF_FP32_BULK = 2
   integer(c_int32_t), parameter :: SWF_BC_SEDOV = 0, SWF_BC_PERIODIC = 1, SWF_BC_JET_HALF = 2, &
                                    SWF_BC_JET_FULL = 3, SWF_BC_STEP = 4
   integer(c_int), parameter :: SWF_OK = 0, SWF_E_PBAR = 1, SWF_E_SWEEP_FAIL = 2, SWF_E_PSUM = 3, &
                                SWF_E_ORDER2_EDGE = 4

   type, bind(C) :: swf_config
      integer(c_int32_t) :: nx, ny
      real(c_double) :: xlo, ylo, dx, dy
      real(c_double) :: gamma, cfl, eps, eps2, delta
      integer(c_int32_t) :: bc, use_limiter, swp_num
      integer(c_int32_t) :: alpha_in_scheme, roe_eps_double, ydens_plain, eps2_quad_literal, findalpha_ex3
      integer(c_int32_t) :: dt_rule_dx53, poscheck_le, success_ge, pbar_check, fail_stop
      real(c_double) :: jet_rho, jet_u, amb_rho, amb_p, jet_halfwidth
      integer(c_int32_t) :: step_total, alpha_lines, weno_eps12, psum_check
      real(c_double) :: in_rho, in_u, in_p
      integer(c_int32_t) :: precision, device
   end type swf_config

   type, bind(C) :: swf1d_config
      integer(c_int32_t) :: n
      real(c_double) :: dx, gamma, cfl, eps
      integer(c_int32_t) :: use_limiter, swp_num, precision, device
   end type swf1d_config

   ! the detonation program (swf7_*, ex7_detonation/main.f90), u(5, -2:nx+3, -2:ny+3)
   type, bind(C) :: swf7_config
      integer(c_int32_t) :: total, step
      real(c_double) :: dx, dy
      real(c_double) :: gamma, q_const, t_const, k_const
      real(c_double) :: eps, eps2, cfl
      integer(c_int32_t) :: swp_num, device
   end type swf7_config

   type, bind(C) :: swf_sweep_stats
      integer(c_int64_t) :: sweep_count, den_count, pre_count
      integer(c_int64_t) :: total_sweep, max_sweep, total_neg, max_neg, total_pos, max_pos
      integer(c_int64_t) :: stage_sweeps, steps
   end type swf_sweep_stats

   interface
      integer(c_int) function swf_create(cfg, ctx) bind(C, name='swf_create')
         import :: swf_config, c_ptr, c_int
         type(swf_config), intent(in) :: cfg
         type(c_ptr), intent(out) :: ctx
      end function swf_create

      subroutine swf_destroy(ctx) bind(C, name='swf_destroy')
         import :: c_ptr
         type(c_ptr), value :: ctx
      end subroutine swf_destroy

      integer(c_int) function swf_upload(ctx, u) bind(C, name='swf_upload')
         import :: c_ptr, c_int, c_double
         type(c_ptr), value :: ctx
         real(c_double), intent(in) :: u(*)
      end function swf_upload

      integer(c_int) function swf_download(ctx, u) bind(C, name='swf_download')
         import :: c_ptr, c_int, c_double
         type(c_ptr), value :: ctx
         real(c_double), intent(out) :: u(*)
      end function swf_download

      integer(c_int) function swf_step(ctx, time, tfinal, dt) bind(C, name='swf_step')
         import :: c_ptr, c_int, c_double
         type(c_ptr), value :: ctx
         real(c_double), value :: time, tfinal
         real(c_double), intent(out) :: dt
      end function swf_step

      integer(c_int) function swf_advance(ctx, time, tfinal, max_steps, steps) bind(C, name='swf_advance')
         import :: c_ptr, c_int, c_double, c_int64_t
         type(c_ptr), value :: ctx
         real(c_double), intent(inout) :: time
         real(c_double), value :: tfinal
         integer(c_int64_t), value :: max_steps
         integer(c_int64_t), intent(out) :: steps
      end function swf_advance

      integer(c_int) function swf_bc(ctx) bind(C, name='swf_bc')
         import :: c_ptr, c_int
         type(c_ptr), value :: ctx
      end function swf_bc

      integer(c_int) function swf_stats(ctx, s) bind(C, name='swf_stats')
         import :: c_ptr, c_int, swf_sweep_stats
         type(c_ptr), value :: ctx
         type(swf_sweep_stats), intent(out) :: s
      end function swf_stats

      integer(c_int64_t) function swf_track(ctx, q, neg, pos, cap) bind(C, name='swf_track')
         import :: c_ptr, c_int64_t
         type(c_ptr), value :: ctx
         integer(c_int64_t), intent(out) :: q(*), neg(*), pos(*)
         integer(c_int64_t), value :: cap
      end function swf_track

      integer(c_int) function swf1d_create(cfg, ctx) bind(C, name='swf1d_create')
         import :: swf1d_config, c_ptr, c_int
         type(swf1d_config), intent(in) :: cfg
         type(c_ptr), intent(out) :: ctx
      end function swf1d_create

      subroutine swf1d_destroy(ctx) bind(C, name='swf1d_destroy')
         import :: c_ptr
         type(c_ptr), value :: ctx
      end subroutine swf1d_destroy

      integer(c_int) function swf1d_upload(ctx, u) bind(C, name='swf1d_upload')
         import :: c_ptr, c_int, c_double
         type(c_ptr), value :: ctx
         real(c_double), intent(in) :: u(*)
      end function swf1d_upload

      integer(c_int) function swf1d_download(ctx, u) bind(C, name='swf1d_download')
         import :: c_ptr, c_int, c_double
         type(c_ptr), value :: ctx
         real(c_double), intent(out) :: u(*)
      end function swf1d_download

      integer(c_int) function swf1d_advance(ctx, time, tfinal, max_steps, steps) bind(C, name='swf1d_advance')
         import :: c_ptr, c_int, c_double, c_int64_t
         type(c_ptr), value :: ctx
         real(c_double), intent(inout) :: time
         real(c_double), value :: tfinal
         integer(c_int64_t), value :: max_steps
         integer(c_int64_t), intent(out) :: steps
      end function swf1d_advance

      integer(c_int) function swf1d_stats(ctx, s) bind(C, name='swf1d_stats')
         import :: c_ptr, c_int, swf_sweep_stats
         type(c_ptr), value :: ctx
         type(swf_sweep_stats), intent(out) :: s
      end function swf1d_stats

      integer(c_int64_t) function swf1d_track(ctx, q, neg, pos, cap) bind(C, name='swf1d_track')
         import :: c_ptr, c_int64_t
         type(c_ptr), value :: ctx
         integer(c_int64_t), intent(out) :: q(*), neg(*), pos(*)
         integer(c_int64_t), value :: cap
      end function swf1d_track

      integer(c_int) function swf7_create(cfg, ctx) bind(C, name='swf7_create')
         import :: swf7_config, c_ptr, c_int
         type(swf7_config), intent(in) :: cfg
         type(c_ptr), intent(out) :: ctx
      end function swf7_create

      subroutine swf7_destroy(ctx) bind(C, name='swf7_destroy')
         import :: c_ptr
         type(c_ptr), value :: ctx
      end subroutine swf7_destroy

      integer(c_int) function swf7_upload(ctx, u) bind(C, name='swf7_upload')
         import :: c_ptr, c_int, c_double
         type(c_ptr), value :: ctx
         real(c_double), intent(in) :: u(*)
      end function swf7_upload

      integer(c_int) function swf7_download(ctx, u) bind(C, name='swf7_download')
         import :: c_ptr, c_int, c_double
         type(c_ptr), value :: ctx
         real(c_double), intent(out) :: u(*)
      end function swf7_download

      integer(c_int) function swf7_bc(ctx) bind(C, name='swf7_bc')
         import :: c_ptr, c_int
         type(c_ptr), value :: ctx
      end function swf7_bc

      integer(c_int) function swf7_advance(ctx, time, tfinal, max_steps, steps) bind(C, name='swf7_advance')
         import :: c_ptr, c_int, c_double, c_int64_t
         type(c_ptr), value :: ctx
         real(c_double), intent(inout) :: time
         real(c_double), value :: tfinal
         integer(c_int64_t), value :: max_steps
         integer(c_int64_t), intent(out) :: steps
      end function swf7_advance

      integer(c_int) function swf7_sweep(ctx) bind(C, name='swf7_sweep')
         import :: c_int, c_ptr
         type(c_ptr), value :: ctx
      end function swf7_sweep
      integer(c_int) function swf7_stats(ctx, s) bind(C, name='swf7_stats')
         import :: c_ptr, c_int, swf_sweep_stats
         type(c_ptr), value :: ctx
         type(swf_sweep_stats), intent(out) :: s
      end function swf7_stats

      type(c_ptr) function swf_strerror(code) bind(C, name='swf_strerror')
         import :: c_ptr, c_int
         integer(c_int), value :: code
      end function swf_strerror
   end interface

contains

   ! C string returned by swf_strerror as a Fortran string
   function swf_message(code) result(msg)
      integer(c_int), intent(in) :: code
      character(len=:), allocatable :: msg
      character(kind=c_char), pointer :: p(:)
      integer :: n
      call c_f_pointer(swf_strerror(code), p, [256])
      n = 0
      do while (n < 256)
         if (p(n + 1) == c_null_char) exit
         n = n + 1
      end do
      allocate (character(len=n) :: msg)
      msg = transfer(p(1:n), msg)
   end function swf_message

end module swf
