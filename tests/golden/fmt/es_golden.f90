! Prints every value of a stream file of doubles with the edit descriptors of the reference's
! writers (es15.7 .out, es13.5 ASCII .vtk, es16.8 .vti/.vtk headers, es14.6 variogram table),
! one value per line, fields separated by '|'.  Compiled and run by make_fmt_golden.py.
program es_golden
  implicit none
  integer :: u, n, i
  real(8), allocatable :: v(:)
  character(len=256) :: fin
  call get_command_argument(1, fin)
  open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
  read(u) n
  allocate(v(n))
  read(u) v
  close(u)
  do i = 1, n
    write(*, '(es15.7,"|",es13.5,"|",es16.8,"|",es14.6)') v(i), v(i), v(i), v(i)
  end do
end program es_golden
