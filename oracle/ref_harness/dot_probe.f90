! Summation order of flang's dot_product for REAL(8) (built like ex2_sedov_1d with
! -O2 -fdefault-real-8).  ex2_sedov_1d/main.f90:577-579 forms the sweep's mean state
! with dot_product(rho_t, weights); this program prints, as hexadecimal bit
! patterns, dot_product and an explicit left-to-right loop on data where the
! summation order changes the result (a +-1d17 pair around ordinary values).
program dot_probe
  implicit none
  integer, parameter :: n = 806
  real :: x(n), w(n), s, t
  integer :: i
  do i = 1, n
     x(i) = 1.0 + mod(i, 7)*0.37
     w(i) = 1/(800. + 6)
  end do
  x(1) = 1d17
  x(400) = -1d17
  s = dot_product(x, w)
  t = 0
  do i = 1, n
     t = t + x(i)*w(i)
  end do
  write (*, '(z16.16,1x,z16.16)') s, t
end program dot_probe
