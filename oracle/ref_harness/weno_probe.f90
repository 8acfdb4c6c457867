! Known-answer harness for the reference WENO-JS point reconstructions.
!
! Links the reference module `weno5` (ex4_sedov_2d/weno5.f90, compiled unmodified
! with the reference's own flags) and evaluates weno_right / weno_left
! (weno5.f90:120-146 / :86-113) on a stream of 5-point stencils.
!
! Input  file (argv 1): raw little-endian float64, n*5 values a(-2..2) per stencil.
! Output file (argv 2): raw float64, n*2 values (weno_right, weno_left) per stencil.
! This program is part of the oracle's test infrastructure only.
program weno_probe
   use weno5
   implicit none
   character(len=512) :: fin, fout
   integer(8) :: nbytes, n, k
   real(8), allocatable :: a(:, :), r(:, :)
   real(8) :: ur, ul

   call get_command_argument(1, fin)
   call get_command_argument(2, fout)
   open (10, file=trim(fin), access='stream', form='unformatted', status='old')
   inquire (unit=10, size=nbytes)
   n = nbytes/40
   allocate (a(-2:2, n), r(2, n))
   read (10) a
   close (10)
   do k = 1, n
      call weno_right(a(:, k), ur)
      call weno_left(a(:, k), ul)
      r(1, k) = ur
      r(2, k) = ul
   end do
   open (11, file=trim(fout), access='stream', form='unformatted', status='replace')
   write (11) r
   close (11)
end program weno_probe
