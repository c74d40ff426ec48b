! m_osiris_amd -- ISO_C_BINDING interfaces of the MI355X PIC hot path (include/osiris_amd.h)
!
! This is the thin layer a Fortran host calls.  It mirrors the seams of the reference:
!  * hip_vadvance_deposit_{2d,3d}_s{1,2,3} have the argument list of the native fused
!    kernels FNAME(vadvance_deposit_*) (simd/os-spec-push-avx2.c:471-475, 761-765) as
!    they are called from spec/os-spec-vpush.f03:144-169 (every argument by reference),
!    so a p_hip case beside p_simd in t_species%push (spec/os-species.f03:1499-1515)
!    only changes the name of the routine it calls;
!  * osr_yee_advance replaces t_emf_solver%advance for periodic boundaries
!    (emf/solver/os-emf-solver-define.f03:26-34);
!  * the osr_sim_* handle keeps particles and fields resident on the GPU and advances
!    whole time steps (t_particles%advance_deposit + current fold + field advance).
! Status-returning functions give 0 (OSR_OK) on success; osr_error_message() returns the
! text of the last error as a Fortran string.
module m_osiris_amd
  use, intrinsic :: iso_c_binding
  implicit none
  private

  integer(c_int), parameter, public :: OSR_OK = 0, OSR_ERR_ARG = 1, OSR_ERR_HIP = 2, &
                                       OSR_ERR_NOMEM = 3, OSR_ERR_STATE = 4, OSR_ERR_LAYOUT = 5
  integer(c_int), parameter, public :: OSR_BORIS = 0, OSR_VAY = 1
  integer(c_int), parameter, public :: OSR_PATH_FUSED = 0, OSR_PATH_PLAIN = 1
  ! x walls: periodic, or Lindman absorbing fields + absorbing particles (emf_bound
  ! "lindman", spec_bound "absorbing")
  integer(c_int), parameter, public :: OSR_BND_PERIODIC = 0, OSR_BND_OPEN = 1, &
                                       OSR_BND_LINDMAN = 1, OSR_BND_VPML = 2

  type, bind(c), public :: osr_sim_desc
    integer(c_int) :: dims = 2
    integer(c_int) :: nx(3) = 0
    real(c_double) :: dx(3) = 0.0_c_double
    real(c_double) :: dt = 0.0_c_double
    integer(c_int) :: order = 2
    integer(c_int) :: tile(3) = 0
    integer(c_int) :: path = OSR_PATH_FUSED
    integer(c_int) :: kernel = 0
    integer(c_int) :: xbnd = OSR_BND_PERIODIC
    integer(c_int) :: reorder = 0
    integer(c_int) :: vpml_size = 0
    integer(c_int) :: vpml_diffuse(2) = 0
    integer(c_int) :: push_waves = 0
  end type osr_sim_desc

  ! current smoothing (vdf/os-vdf-smooth.f90): the reference's filter type codes and the
  ! nl_smooth namelist; swfj has the reference's shape swfj(3, p_order_max, p_x_dim)
  integer(c_int), parameter, public :: OSR_SMOOTH_NONE = 0, OSR_SMOOTH_CUSTOM = 1, &
                                       OSR_SMOOTH_BINOMIAL = 2, OSR_SMOOTH_COMPENSATED = 3, &
                                       OSR_SMOOTH_5PASS = 4, OSR_SMOOTH_TRI6 = 5, &
                                       OSR_SMOOTH_TRI8 = 6, OSR_SMOOTH_LASERKDX = 10, &
                                       OSR_SMOOTH_DIGITAL = 11
  integer(c_int), parameter, public :: OSR_SMOOTH_ORDER_MAX = 15, OSR_SMOOTH_NCOEF = 31
  type, bind(c), public :: osr_smooth_params
    integer(c_int) :: type(3) = OSR_SMOOTH_NONE
    integer(c_int) :: order(3) = 0
    real(c_double) :: swfj(3, 15, 3) = 1.0_c_double
    real(c_double) :: laserkdx_omega0 = -1.0_c_double
    integer(c_int) :: laserkdx_dir = 1
    integer(c_int) :: reserved = 0
    real(c_double) :: digital_A(3) = -huge(1.0_c_double)
    real(c_double) :: digital_fc(3) = -huge(1.0_c_double)
  end type osr_smooth_params

  ! laser injection from an x wall (zpulse_wall, zpulse/os-zpulse-wall.f03): the namelist
  ! nl_zpulse_wall for direction 1 of a 2D box; defaults are the reference's
  integer(c_int), parameter, public :: OSR_TENV_POLYNOMIAL = 0, OSR_TENV_SIN2 = 1, &
                                       OSR_TENV_GAUSSIAN = 2
  integer(c_int), parameter, public :: OSR_PER_PLANE = 0, OSR_PER_HERMITE = 1
  integer(c_int), parameter, public :: OSR_PROP_FORWARD = 0, OSR_PROP_BACKWARD = 1
  type, bind(c), public :: osr_zpulse_wall
    real(c_double) :: a0 = 1.0_c_double, omega0 = 10.0_c_double
    real(c_double) :: phase = 0.0_c_double
    real(c_double) :: pol = 90.0_c_double
    integer(c_int) :: pol_type = 0
    integer(c_int) :: propagation = OSR_PROP_FORWARD
    integer(c_int) :: tenv_type = OSR_TENV_POLYNOMIAL
    integer(c_int) :: per_type = OSR_PER_PLANE
    real(c_double) :: tenv_fwhm = -1.0_c_double
    real(c_double) :: tenv_rise = 0.0_c_double, tenv_flat = 0.0_c_double, tenv_fall = 0.0_c_double
    real(c_double) :: tenv_duration = 0.0_c_double
    real(c_double) :: tenv_range = -huge(1.0_c_double)
    real(c_double) :: per_center(2) = -huge(1.0_c_double)
    real(c_double) :: per_w0 = 0.0_c_double, per_fwhm = 0.0_c_double
    real(c_double) :: per_focus = -huge(1.0_c_double)
    integer(c_int) :: per_tem_mode(2) = 0
    real(c_double) :: launch_time = 0.0_c_double
    real(c_double) :: xmin(3) = 0.0_c_double
  end type osr_zpulse_wall

  ! the in-box pulse (zpulse section) along x1, launched at set-up; defaults as the
  ! reference's and osr_zpulse_defaults
  type, bind(c), public :: osr_zpulse
    real(c_double) :: a0 = 1.0_c_double, omega0 = 10.0_c_double
    real(c_double) :: phase = 0.0_c_double
    real(c_double) :: pol = 90.0_c_double
    integer(c_int) :: pol_type = 0
    integer(c_int) :: propagation = OSR_PROP_FORWARD
    integer(c_int) :: lon_type = OSR_TENV_POLYNOMIAL
    integer(c_int) :: per_type = OSR_PER_PLANE
    real(c_double) :: lon_start = -huge(1.0_c_double)
    real(c_double) :: lon_fwhm = -1.0_c_double
    real(c_double) :: lon_rise = 0.0_c_double, lon_flat = 0.0_c_double, lon_fall = 0.0_c_double
    real(c_double) :: lon_duration = 0.0_c_double
    real(c_double) :: lon_x0 = -huge(1.0_c_double), lon_range = -huge(1.0_c_double)
    real(c_double) :: per_center(2) = -huge(1.0_c_double)
    real(c_double) :: per_w0 = 0.0_c_double, per_fwhm = 0.0_c_double
    real(c_double) :: per_focus = -huge(1.0_c_double)
    integer(c_int) :: per_tem_mode(2) = 0
    integer(c_int) :: no_div_corr = 0
    real(c_double) :: xmin(3) = 0.0_c_double
  end type osr_zpulse

  public :: osr_version, osr_device_count, osr_error_message
  public :: osr_advance_deposit, osr_yee_advance
  public :: hip_vadvance_deposit_2d_s1, hip_vadvance_deposit_2d_s2, hip_vadvance_deposit_2d_s3
  public :: hip_vadvance_deposit_3d_s1, hip_vadvance_deposit_3d_s2, hip_vadvance_deposit_3d_s3
  public :: osr_sim_create, osr_sim_destroy, osr_sim_add_species, osr_species_set, &
            osr_species_count, osr_species_get, osr_species_generate, osr_fields_set, &
            osr_fields_get, osr_fields_fill, osr_sim_run, osr_sim_steps, osr_sim_history, &
            osr_sim_gauss, osr_sim_relayouts, osr_sim_time_kernels, osr_sim_footprint, &
            osr_sim_launch_config
  public :: osr_smooth_setup, osr_smooth_vdf, osr_sim_set_smooth
  public :: osr_sim_write_hist, osr_sim_save_field
  public :: osr_sim_add_zpulse_wall, osr_sim_zpulse_column, osr_sim_init_fields, &
            osr_sim_add_zpulse

  interface
    function osr_last_error() bind(c, name="osr_last_error")
      import :: c_ptr
      type(c_ptr) :: osr_last_error
    end function
    function osr_version() bind(c, name="osr_version")
      import :: c_int
      integer(c_int) :: osr_version
    end function
    function osr_device_count() bind(c, name="osr_device_count")
      import :: c_int
      integer(c_int) :: osr_device_count
    end function

    ! ---- 1. stateless drop-ins
    function osr_advance_deposit(dims, order, pusher, ix, x, u, q, i0, i1, rqm, efield, bfield, &
                                 emf_size, emf_offset, j, j_size, j_offset, dx, dt, ene) &
        bind(c, name="osr_advance_deposit")
      import :: c_int, c_float, c_double
      integer(c_int), value :: dims, order, pusher, i0, i1
      integer(c_int), intent(inout) :: ix(*)
      real(c_float), intent(inout) :: x(*), u(*)
      real(c_float), intent(in) :: q(*), efield(*), bfield(*)
      real(c_float), value :: rqm
      integer(c_int), intent(in) :: emf_size(*), emf_offset(*), j_size(*), j_offset(*)
      real(c_float), intent(inout) :: j(*)
      real(c_double), intent(in) :: dx(*)
      real(c_double), value :: dt
      real(c_double), intent(inout) :: ene
      integer(c_int) :: osr_advance_deposit
    end function

    subroutine hip_vadvance_deposit_2d_s1(ix, x, u, q, i0, i1, rqm, efield, bfield, emf_size, &
                                          emf_offset, j, j_size, j_offset, dx, dt, ene) &
        bind(c, name="hip_vadvance_deposit_2d_s1_")
      import :: c_int, c_float, c_double
      integer(c_int) :: ix(*), i0, i1, emf_size(*), emf_offset(*), j_size(*), j_offset(*)
      real(c_float) :: x(*), u(*), q(*), rqm, efield(*), bfield(*), j(*)
      real(c_double) :: dx(*), dt, ene
    end subroutine
    subroutine hip_vadvance_deposit_2d_s2(ix, x, u, q, i0, i1, rqm, efield, bfield, emf_size, &
                                          emf_offset, j, j_size, j_offset, dx, dt, ene) &
        bind(c, name="hip_vadvance_deposit_2d_s2_")
      import :: c_int, c_float, c_double
      integer(c_int) :: ix(*), i0, i1, emf_size(*), emf_offset(*), j_size(*), j_offset(*)
      real(c_float) :: x(*), u(*), q(*), rqm, efield(*), bfield(*), j(*)
      real(c_double) :: dx(*), dt, ene
    end subroutine
    subroutine hip_vadvance_deposit_2d_s3(ix, x, u, q, i0, i1, rqm, efield, bfield, emf_size, &
                                          emf_offset, j, j_size, j_offset, dx, dt, ene) &
        bind(c, name="hip_vadvance_deposit_2d_s3_")
      import :: c_int, c_float, c_double
      integer(c_int) :: ix(*), i0, i1, emf_size(*), emf_offset(*), j_size(*), j_offset(*)
      real(c_float) :: x(*), u(*), q(*), rqm, efield(*), bfield(*), j(*)
      real(c_double) :: dx(*), dt, ene
    end subroutine
    subroutine hip_vadvance_deposit_3d_s1(ix, x, u, q, i0, i1, rqm, efield, bfield, emf_size, &
                                          emf_offset, j, j_size, j_offset, dx, dt, ene) &
        bind(c, name="hip_vadvance_deposit_3d_s1_")
      import :: c_int, c_float, c_double
      integer(c_int) :: ix(*), i0, i1, emf_size(*), emf_offset(*), j_size(*), j_offset(*)
      real(c_float) :: x(*), u(*), q(*), rqm, efield(*), bfield(*), j(*)
      real(c_double) :: dx(*), dt, ene
    end subroutine
    subroutine hip_vadvance_deposit_3d_s2(ix, x, u, q, i0, i1, rqm, efield, bfield, emf_size, &
                                          emf_offset, j, j_size, j_offset, dx, dt, ene) &
        bind(c, name="hip_vadvance_deposit_3d_s2_")
      import :: c_int, c_float, c_double
      integer(c_int) :: ix(*), i0, i1, emf_size(*), emf_offset(*), j_size(*), j_offset(*)
      real(c_float) :: x(*), u(*), q(*), rqm, efield(*), bfield(*), j(*)
      real(c_double) :: dx(*), dt, ene
    end subroutine
    subroutine hip_vadvance_deposit_3d_s3(ix, x, u, q, i0, i1, rqm, efield, bfield, emf_size, &
                                          emf_offset, j, j_size, j_offset, dx, dt, ene) &
        bind(c, name="hip_vadvance_deposit_3d_s3_")
      import :: c_int, c_float, c_double
      integer(c_int) :: ix(*), i0, i1, emf_size(*), emf_offset(*), j_size(*), j_offset(*)
      real(c_float) :: x(*), u(*), q(*), rqm, efield(*), bfield(*), j(*)
      real(c_double) :: dx(*), dt, ene
    end subroutine

    ! ---- 2. stateless Yee advance (periodic)
    function osr_yee_advance(dims, nx, ld, gcl, dx, dt, e, b, jay) bind(c, name="osr_yee_advance")
      import :: c_int, c_float, c_double
      integer(c_int), value :: dims
      integer(c_int), intent(in) :: nx(*), ld(*), gcl(*)
      real(c_double), intent(in) :: dx(*)
      real(c_double), value :: dt
      real(c_float), intent(inout) :: e(*), b(*)
      real(c_float), intent(in) :: jay(*)
      integer(c_int) :: osr_yee_advance
    end function

    ! ---- 3. resident simulation
    function osr_sim_create(desc, sim) bind(c, name="osr_sim_create")
      import :: c_int, c_ptr, osr_sim_desc
      type(osr_sim_desc), intent(in) :: desc
      type(c_ptr), intent(out) :: sim
      integer(c_int) :: osr_sim_create
    end function
    function osr_sim_destroy(sim) bind(c, name="osr_sim_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: sim
      integer(c_int) :: osr_sim_destroy
    end function
    function osr_sim_add_species(sim, rqm, pusher, sid) bind(c, name="osr_sim_add_species")
      import :: c_int, c_ptr, c_float
      type(c_ptr), value :: sim
      real(c_float), value :: rqm
      integer(c_int), value :: pusher
      integer(c_int), intent(out) :: sid
      integer(c_int) :: osr_sim_add_species
    end function
    function osr_species_set(sim, sid, np, ix, x, u, q) bind(c, name="osr_species_set")
      import :: c_int, c_ptr, c_float, c_int64_t
      type(c_ptr), value :: sim
      integer(c_int), value :: sid
      integer(c_int64_t), value :: np
      integer(c_int), intent(in) :: ix(*)
      real(c_float), intent(in) :: x(*), u(*), q(*)
      integer(c_int) :: osr_species_set
    end function
    function osr_species_count(sim, sid, np) bind(c, name="osr_species_count")
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: sim
      integer(c_int), value :: sid
      integer(c_int64_t), intent(out) :: np
      integer(c_int) :: osr_species_count
    end function
    function osr_species_get(sim, sid, ix, x, u, q) bind(c, name="osr_species_get")
      import :: c_int, c_ptr, c_float
      type(c_ptr), value :: sim
      integer(c_int), value :: sid
      integer(c_int), intent(out) :: ix(*)
      real(c_float), intent(out) :: x(*), u(*), q(*)
      integer(c_int) :: osr_species_get
    end function
    function osr_species_generate(sim, sid, ppc, density, ufl, uth, seed, random_pos) &
        bind(c, name="osr_species_generate")
      import :: c_int, c_ptr, c_double, c_int64_t
      type(c_ptr), value :: sim
      integer(c_int), value :: sid
      integer(c_int), intent(in) :: ppc(*)
      real(c_double), value :: density
      real(c_double), intent(in) :: ufl(*), uth(*)
      integer(c_int64_t), value :: seed
      integer(c_int), value :: random_pos
      integer(c_int) :: osr_species_generate
    end function
    function osr_fields_set(sim, e, b, ld, gcl) bind(c, name="osr_fields_set")
      import :: c_int, c_ptr, c_float
      type(c_ptr), value :: sim
      real(c_float), intent(in) :: e(*), b(*)
      integer(c_int), intent(in) :: ld(*), gcl(*)
      integer(c_int) :: osr_fields_set
    end function
    function osr_fields_get(sim, e, b, ld, gcl) bind(c, name="osr_fields_get")
      import :: c_int, c_ptr, c_float
      type(c_ptr), value :: sim
      real(c_float), intent(out) :: e(*), b(*)
      integer(c_int), intent(in) :: ld(*), gcl(*)
      integer(c_int) :: osr_fields_get
    end function
    function osr_fields_fill(sim, e0, b0) bind(c, name="osr_fields_fill")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: sim
      real(c_double), intent(in) :: e0(3), b0(3)
      integer(c_int) :: osr_fields_fill
    end function
    function osr_sim_run(sim, nsteps) bind(c, name="osr_sim_run")
      import :: c_int, c_ptr
      type(c_ptr), value :: sim
      integer(c_int), value :: nsteps
      integer(c_int) :: osr_sim_run
    end function
    function osr_sim_steps(sim, steps) bind(c, name="osr_sim_steps")
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: sim
      integer(c_int64_t), intent(out) :: steps
      integer(c_int) :: osr_sim_steps
    end function
    function osr_sim_history(sim, first, count, fld, par) bind(c, name="osr_sim_history")
      import :: c_int, c_ptr, c_double, c_int64_t
      type(c_ptr), value :: sim
      integer(c_int64_t), value :: first, count
      real(c_double), intent(out) :: fld(*), par(*)
      integer(c_int) :: osr_sim_history
    end function
    function osr_sim_gauss(sim, res, rho, maxabs) bind(c, name="osr_sim_gauss")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: sim
      type(c_ptr), value :: res, rho
      real(c_double), intent(out) :: maxabs
      integer(c_int) :: osr_sim_gauss
    end function
    function osr_sim_relayouts(sim, count) bind(c, name="osr_sim_relayouts")
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: sim
      integer(c_int64_t), intent(out) :: count
      integer(c_int) :: osr_sim_relayouts
    end function
    function osr_sim_time_kernels(sim, nsteps, ms) bind(c, name="osr_sim_time_kernels")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: sim
      integer(c_int), value :: nsteps
      real(c_double), intent(out) :: ms(6)
      integer(c_int) :: osr_sim_time_kernels
    end function
    function osr_sim_footprint(sim, bytes) bind(c, name="osr_sim_footprint")
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: sim
      integer(c_int64_t), intent(out) :: bytes
      integer(c_int) :: osr_sim_footprint
    end function
    function osr_sim_launch_config(sim, tile, waves, compiled_tile) &
        bind(c, name="osr_sim_launch_config")
      import :: c_int, c_ptr
      type(c_ptr), value :: sim
      integer(c_int), intent(out) :: tile(3), waves, compiled_tile
      integer(c_int) :: osr_sim_launch_config
    end function

    ! ---- 5. current smoothing
    function osr_smooth_setup(dims, p, dx, gamma, order, coef) bind(c, name="osr_smooth_setup")
      import :: c_int, c_float, c_double, osr_smooth_params
      integer(c_int), value :: dims
      type(osr_smooth_params), intent(in) :: p
      real(c_double), intent(in) :: dx(*)
      real(c_double), value :: gamma
      integer(c_int), intent(out) :: order(3)
      real(c_float), intent(out) :: coef(31, 3)
      integer(c_int) :: osr_smooth_setup
    end function
    function osr_smooth_vdf(dims, ld, fdim, order, coef, f) bind(c, name="osr_smooth_vdf")
      import :: c_int, c_float
      integer(c_int), value :: dims, fdim
      integer(c_int), intent(in) :: ld(*), order(*)
      real(c_float), intent(in) :: coef(31, 3)
      real(c_float), intent(inout) :: f(*)
      integer(c_int) :: osr_smooth_vdf
    end function
    function osr_sim_set_smooth(sim, order, coef) bind(c, name="osr_sim_set_smooth")
      import :: c_int, c_ptr, c_float
      type(c_ptr), value :: sim
      integer(c_int), intent(in) :: order(3)
      real(c_float), intent(in) :: coef(31, 3)
      integer(c_int) :: osr_sim_set_smooth
    end function

    ! ---- 6. output in the reference's formats (strings NUL-terminated: 'HIST'//c_null_char)
    function osr_sim_write_hist(sim, dir, names, first, count) bind(c, name="osr_sim_write_hist")
      import :: c_int, c_ptr, c_char, c_int64_t
      type(c_ptr), value :: sim
      character(kind=c_char), intent(in) :: dir(*)
      type(c_ptr), value :: names        ! char*[nspecies] or c_null_ptr
      integer(c_int64_t), value :: first, count
      integer(c_int) :: osr_sim_write_hist
    end function
    function osr_sim_save_field(sim, quant, ms_dir, xmin) bind(c, name="osr_sim_save_field")
      import :: c_int, c_ptr, c_char, c_double
      type(c_ptr), value :: sim
      integer(c_int), value :: quant
      character(kind=c_char), intent(in) :: ms_dir(*)
      real(c_double), intent(in) :: xmin(*)
      integer(c_int) :: osr_sim_save_field
    end function
    function osr_sim_init_fields(sim, sid, ufl1) bind(c, name="osr_sim_init_fields")
      import :: c_int, c_ptr, c_float
      type(c_ptr), value :: sim
      integer(c_int), value :: sid
      real(c_float), value :: ufl1
      integer(c_int) :: osr_sim_init_fields
    end function
    function osr_sim_add_zpulse_wall(sim, zp) bind(c, name="osr_sim_add_zpulse_wall")
      import :: c_int, c_ptr, osr_zpulse_wall
      type(c_ptr), value :: sim
      type(osr_zpulse_wall), intent(in) :: zp
      integer(c_int) :: osr_sim_add_zpulse_wall
    end function
    function osr_sim_add_zpulse(sim, zp) bind(c, name="osr_sim_add_zpulse")
      import :: c_int, c_ptr, osr_zpulse
      type(c_ptr), value :: sim
      type(osr_zpulse), intent(in) :: zp
      integer(c_int) :: osr_sim_add_zpulse
    end function
    function osr_sim_zpulse_column(sim, k, second, t, b12, b3, active) &
        bind(c, name="osr_sim_zpulse_column")
      import :: c_int, c_ptr, c_float, c_double
      type(c_ptr), value :: sim
      integer(c_int), value :: k, second
      real(c_double), value :: t
      real(c_float), intent(out) :: b12(*), b3(*)
      integer(c_int), intent(out) :: active
      integer(c_int) :: osr_sim_zpulse_column
    end function
  end interface

contains

  ! Text of the last error of the calling thread.
  function osr_error_message() result(msg)
    character(len=:), allocatable :: msg
    character(kind=c_char), pointer :: p(:)
    type(c_ptr) :: cp
    integer :: n
    interface
      function c_strlen(s) bind(c, name="strlen")
        import :: c_ptr, c_size_t
        type(c_ptr), value :: s
        integer(c_size_t) :: c_strlen
      end function
    end interface
    cp = osr_last_error()
    if (.not. c_associated(cp)) then
      msg = ""
      return
    end if
    n = int(c_strlen(cp))
    call c_f_pointer(cp, p, [n])
    allocate (character(len=n) :: msg)
    msg = transfer(p(1:n), msg)
  end function

end module m_osiris_amd
