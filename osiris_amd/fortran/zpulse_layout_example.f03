! zpulse_layout_example -- host-only check of the osr_zpulse_wall and osr_zpulse bindings
! of m_osiris_amd: prints the size of each bind(c) type and its default values (the
! reference's nl_zpulse_wall and zpulse defaults), and the same fields after
! osr_zpulse_wall_defaults / osr_zpulse_defaults fill a second copy through the C ABI.
program zpulse_layout_example
  use iso_c_binding
  use m_osiris_amd
  implicit none
  interface
    subroutine osr_zpulse_wall_defaults(zp) bind(c, name="osr_zpulse_wall_defaults")
      import :: osr_zpulse_wall
      type(osr_zpulse_wall), intent(out) :: zp
    end subroutine
    subroutine osr_zpulse_defaults(zp) bind(c, name="osr_zpulse_defaults")
      import :: osr_zpulse
      type(osr_zpulse), intent(out) :: zp
    end subroutine
  end interface
  type(osr_zpulse_wall) :: a, b
  type(osr_zpulse) :: p, q
  write (*, '(a,i0)') "size ", c_sizeof(a)
  call show("fortran", a)
  call osr_zpulse_wall_defaults(b)
  call show("c", b)
  write (*, '(a,i0)') "size2 ", c_sizeof(p)
  call show2("fortran2", p)
  call osr_zpulse_defaults(q)
  call show2("c2", q)
contains
  subroutine show2(tag, z)
    character(*), intent(in) :: tag
    type(osr_zpulse), intent(in) :: z
    write (*, '(a,1x,*(es24.16e3,1x))') tag, z%a0, z%omega0, z%phase, z%pol, z%lon_start, &
      z%lon_fwhm, z%lon_rise, z%lon_flat, z%lon_fall, z%lon_duration, z%lon_x0, z%lon_range, &
      z%per_center, z%per_w0, z%per_fwhm, z%per_focus, z%xmin
    write (*, '(a,1x,*(i0,1x))') tag, z%pol_type, z%propagation, z%lon_type, z%per_type, &
      z%per_tem_mode, z%no_div_corr
  end subroutine
  subroutine show(tag, z)
    character(*), intent(in) :: tag
    type(osr_zpulse_wall), intent(in) :: z
    write (*, '(a,1x,*(es24.16e3,1x))') tag, z%a0, z%omega0, z%phase, z%pol, z%tenv_fwhm, &
      z%tenv_rise, z%tenv_flat, z%tenv_fall, z%tenv_duration, z%tenv_range, z%per_center, &
      z%per_w0, z%per_fwhm, z%per_focus, z%launch_time, z%xmin
    write (*, '(a,1x,*(i0,1x))') tag, z%pol_type, z%propagation, z%tenv_type, z%per_type, &
      z%per_tem_mode
  end subroutine
end program zpulse_layout_example
