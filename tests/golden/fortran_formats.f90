! fortran_formats.f90 -- writes doubles read from a stream file with the edit
! descriptors of the reference's VTK writer (mcm_VTK.f90:209,212,215,232-246,
! 506): F15.8 for coordinates and velocities, E24.5 for scalars, E24.5 of the
! value rounded to default REAL (its von Mises stress), I6 and I5 for integers.
! Used only to make the formatting fixture of tests/test_output.py.
program fortran_formats
  implicit none
  integer :: n, i, u
  real(8), allocatable :: x(:)
  real :: s
  character(len=256) :: path
  call get_command_argument(1, path)
  open(newunit=u, file=trim(path), access='stream', form='unformatted', status='old')
  read(u) n
  allocate(x(n))
  read(u) x
  close(u)
  do i = 1, n
    s = real(x(i))
    write(*, '(F15.8)') x(i)
    write(*, '(E24.5)') x(i)
    write(*, '(E24.5)') s
  end do
  do i = 1, n
    if (x(i) /= x(i)) then
      write(*, '("1 ",I6)') 0
      write(*, '(I5)') 0
    else
      write(*, '("1 ",I6)') int(min(max(x(i) * 1.0d6, -2.0d9), 2.0d9))
      write(*, '(I5)') int(min(max(x(i) * 1.0d3, -2.0d9), 2.0d9))
    end if
  end do
end program fortran_formats
