! smooth_setup_example -- host-only use of the smoothing set-up through m_osiris_amd:
! fills the nl_smooth parameters the way an input deck would (compensated along x,
! 2 passes; binomial along y, 3 passes) and prints the filter orders and coefficients.
program smooth_setup_example
  use iso_c_binding
  use m_osiris_amd
  implicit none
  type(osr_smooth_params) :: p
  integer(c_int) :: order(3), st, d, k
  real(c_float) :: coef(31, 3)
  real(c_double) :: dx(2) = [0.1_c_double, 0.1_c_double]

  p%type(1:2) = [OSR_SMOOTH_COMPENSATED, OSR_SMOOTH_BINOMIAL]
  p%order(1:2) = [2, 3]
  st = osr_smooth_setup(2_c_int, p, dx, 1.0_c_double, order, coef)
  if (st /= OSR_OK) then
    print *, "osr_smooth_setup failed: ", osr_error_message()
    stop 1
  end if
  do d = 1, 2
    write (*, '(a,i0,a,i0)') "dir ", d, " order ", order(d)
    do k = 16 - order(d), 16 + order(d)
      write (*, '(es16.8)') coef(k, d)
    end do
  end do
end program smooth_setup_example
