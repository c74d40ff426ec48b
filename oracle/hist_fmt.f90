! hist_fmt -- TEST INFRASTRUCTURE ONLY (the checker of the HIST writer).
!
! Writes HIST/fld_ene, par_ene and parNN_ene rows with the Fortran runtime, using the edit
! descriptors of the reference's writers (emf/diagnostics/os-emf-diagnostics.f03:60-81,
! particles/os-particles.f03:1736-1776, spec/diagnostics/os-spec-diagnostics.f03:1858-1882),
! so that the library's emulation of I6 / G15.8 / ES23.16 / A editing can be compared byte
! for byte.  Input (stdin): nspec nrows; nspec names (one per line); nspec particle counts;
! then per row: n t B1 B2 B3 E1 E2 E3 and nspec energies.  Output in the directory given
! as the first argument.
program hist_fmt
  implicit none
  integer, parameter :: dp = kind(1.0d0), i8 = selected_int_kind(15)
  integer :: nspec, nrows, r, s, n
  character(len=256) :: dir
  character(len=64), allocatable :: names(:)
  character(len=23), allocatable :: header(:)
  integer(i8), allocatable :: npar(:)
  real(dp) :: t, fld(6)
  real(dp), allocatable :: ene(:), ene_all(:)
  character(len=64) :: fmt
  character(len=300) :: fname

  call get_command_argument(1, dir)
  read (*, *) nspec, nrows
  allocate (names(nspec), header(nspec), npar(nspec), ene(nspec), ene_all(nspec + 1))
  do s = 1, nspec
    read (*, '(A)') names(s)
  end do
  read (*, *) npar
  do r = 1, nrows
    read (*, *) n, t, fld, ene
    ! fld_ene
    fname = trim(dir)//'/fld_ene'
    if (t == 0.0_dp) then
      open (unit=10, file=fname, status='REPLACE', form='formatted')
      write (10, '(A)') '! EM field energy per field component'
      write (10, '( A6, 1X,A15,6(1X,A23) )') 'Iter', 'Time    ', 'B1', 'B2', 'B3', 'E1', 'E2', 'E3'
    else
      open (unit=10, file=fname, position='append', form='formatted')
    end if
    write (10, '( I6, 1X,g15.8, 6(1X,es23.16) )') n, t, fld
    close (10)
    if (nspec == 0) cycle
    ! par_ene
    ene_all(1) = 0
    do s = 1, nspec
      ene_all(s + 1) = ene(s)
      ene_all(1) = ene_all(1) + ene_all(s + 1)
    end do
    fname = trim(dir)//'/par_ene'
    if (t == 0.0_dp) then
      open (unit=10, file=fname, status='REPLACE', form='formatted')
      write (10, '(A)') '! Total and per species particle kinetic energy'
      write (fmt, '(A,I0,A,I0,A)') '( A6, (1X,A15), ', nspec + 1, '(1X, A', 23, ') )'
      do s = 1, nspec
        header(s) = trim(names(s))
        header(s) = adjustr(header(s))
      end do
      write (10, fmt) 'Iter', 'Time    ', 'Total', header
    else
      open (unit=10, file=fname, position='append', form='formatted')
    end if
    write (fmt, '(A,I0,A)') '( I6, (1X,g15.8), ', nspec + 1, '(1X, es23.16) )'
    write (10, fmt) n, t, ene_all
    close (10)
    ! parNN_ene
    do s = 1, nspec
      write (fname, '(A,A,I0.2,A)') trim(dir), '/par', s, '_ene'
      if (t == 0.0_dp) then
        open (unit=10, file=fname, status='REPLACE', form='formatted')
        write (10, '(A,A,A)') '! Kinetic energy for species "', trim(names(s)), '"'
        write (10, '( A6, 2(1X,A15), (1X,A23) )') 'Iter', 'Time    ', 'Total Par.', 'Kin. Energy'
      else
        open (unit=10, file=fname, position='append', form='formatted')
      end if
      write (10, '( I6,1X,g15.8,1X,I14,1(1X,es23.16) )') n, t, npar(s), ene(s)
      close (10)
    end do
  end do
end program hist_fmt
