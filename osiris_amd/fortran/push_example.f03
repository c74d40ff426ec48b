! push_example -- a Fortran host driving the GPU hot path through m_osiris_amd.
!
! 1. The stateless drop-in: a 2D quadratic push of particles held in Fortran arrays
!    laid out as in t_species (x(2,N), ix(2,N), p(3,N), q(N)) with the guarded field
!    arrays e/b/jay(3, 1-gc:nx+gc, ...), exactly as spec/os-spec-vpush.f03:144-169 calls
!    the native kernel, then the periodic Yee advance of the same fields.
! 2. The resident path: a small two-stream run whose energy history is printed.
! Prints "push_example: OK" and exits 0 on success.
program push_example
  use, intrinsic :: iso_c_binding
  use m_osiris_amd
  implicit none

  integer, parameter :: nx1 = 32, nx2 = 32, gcl = 2, gcu = 3, ppc = 4
  integer, parameter :: np = nx1 * nx2 * ppc
  integer(c_int) :: ix(2, np), i0, i1, emf_size(2), emf_offset(2), st
  real(c_float) :: x(2, np), u(3, np), q(np), rqm
  real(c_float), allocatable :: e(:, :, :), b(:, :, :), jay(:, :, :)
  real(c_double) :: dx(2), dt, ene
  integer(c_int) :: nxa(2), ld(2), gca(2)
  integer :: k, i1c, i2c, kk
  real :: r(5)
  type(osr_sim_desc) :: desc
  type(c_ptr) :: sim
  integer(c_int) :: sid
  integer(c_int64_t) :: nsteps, cnt
  real(c_double) :: fld(6, 10), par(2, 10)

  if (osr_device_count() < 1) then
    print *, "push_example: no GPU"
    stop 1
  end if

  allocate (e(3, 1 - gcl:nx1 + gcu, 1 - gcl:nx2 + gcu), b(3, 1 - gcl:nx1 + gcu, 1 - gcl:nx2 + gcu))
  allocate (jay(3, 1 - gcl:nx1 + gcu, 1 - gcl:nx2 + gcu))
  call random_seed()
  call random_number(e)
  call random_number(b)
  e = 1.0e-2 * (e - 0.5)
  b = 1.0e-1 * (b - 0.5)
  jay = 0.0
  k = 0
  do i2c = 1, nx2
    do i1c = 1, nx1
      do kk = 1, ppc
        k = k + 1
        call random_number(r)
        ix(:, k) = [i1c, i2c]
        x(:, k) = real(r(1:2) - 0.5, c_float)
        x(:, k) = min(x(:, k), 0.49999)
        u(:, k) = real(0.1 * (r(3:5) - 0.5), c_float)
        q(k) = -1.0 / ppc
      end do
    end do
  end do
  i0 = 1
  i1 = np
  rqm = -1.0
  emf_size = [nx1 + gcl + gcu, nx2 + gcl + gcu]
  emf_offset = [gcl, gcl]
  dx = [0.1d0, 0.1d0]
  dt = 0.05d0
  ene = 0.0d0
  ! same call as the reference's p_simd branch, with the hip_ routine
  call hip_vadvance_deposit_2d_s2(ix, x, u, q, i0, i1, rqm, e, b, emf_size, emf_offset, &
                                  jay, emf_size, emf_offset, dx, dt, ene)
  ! ene accumulates q*u^2/(gamma+1) with q < 0 here (normalised by rqm*dV later)
  if (any(x < -0.5) .or. any(x >= 0.5) .or. ene >= 0.0d0) then
    print *, "push_example: drop-in push produced invalid state"
    stop 1
  end if
  ! the current is folded into the interior by the caller (jay%update_boundary); for the
  ! Yee call it is enough that jay is finite
  nxa = [nx1, nx2]
  ld = emf_size
  gca = [gcl, gcl]
  st = osr_yee_advance(2, nxa, ld, gca, dx, dt, e, b, jay)
  if (st /= OSR_OK) then
    print *, "push_example: osr_yee_advance failed: ", osr_error_message()
    stop 1
  end if

  ! ---- resident path: two counter-streaming beams
  desc%dims = 2
  desc%nx = [64, 64, 1]
  desc%dx = [0.05d0, 0.05d0, 0.0d0]
  desc%dt = 0.035d0
  desc%order = 2
  st = osr_sim_create(desc, sim)
  if (st /= OSR_OK) then
    print *, "push_example: osr_sim_create failed: ", osr_error_message()
    stop 1
  end if
  do k = 1, 2
    st = osr_sim_add_species(sim, -1.0_c_float, OSR_BORIS, sid)
    if (st == OSR_OK) st = osr_species_generate(sim, sid, [4_c_int, 4_c_int, 1_c_int], 0.5d0, &
                                                 [merge(0.2d0, -0.2d0, k == 1), 0d0, 0d0], &
                                                 [0.01d0, 0.01d0, 0.01d0], &
                                                 int(1000 + k, c_int64_t), 1_c_int)
    if (st /= OSR_OK) then
      print *, "push_example: species setup failed: ", osr_error_message()
      stop 1
    end if
  end do
  st = osr_sim_run(sim, 10)
  if (st == OSR_OK) st = osr_sim_steps(sim, nsteps)
  if (st == OSR_OK) st = osr_sim_history(sim, 0_c_int64_t, 10_c_int64_t, fld, par)
  if (st == OSR_OK) st = osr_species_count(sim, 0, cnt)
  if (st /= OSR_OK .or. nsteps /= 10 .or. cnt /= 64 * 64 * 16) then
    print *, "push_example: resident run failed: ", osr_error_message()
    stop 1
  end if
  do k = 1, 10, 3
    print '(a,i3,a,es23.16,a,es23.16)', " step", k - 1, "  fld E1 ", fld(4, k), "  par ", par(1, k)
  end do
  st = osr_sim_destroy(sim)
  print *, "push_example: OK"
end program push_example
