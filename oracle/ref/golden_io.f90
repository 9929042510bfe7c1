!> Writer of the fixture files of the reference drivers: a stream of
!> records, each a 32-character name, an int32 kind (1: float64, 2: int32),
!> an int64 count and the values (column-major order), little endian.
module m_golden_io
  implicit none
  private

  integer, parameter :: dp = kind(0.0d0)
  integer, parameter :: i8 = selected_int_kind(18)
  integer, parameter :: i4 = selected_int_kind(9)
  integer :: gu = -1

  public :: gio_open, gio_close, gio_real, gio_int

contains

  subroutine gio_open(fname)
    character(len=*), intent(in) :: fname
    open(newunit=gu, file=fname, access="stream", form="unformatted", &
         status="replace", action="write")
  end subroutine gio_open

  subroutine gio_close()
    close(gu)
    gu = -1
  end subroutine gio_close

  subroutine gio_real(name, x)
    character(len=*), intent(in) :: name
    real(dp), intent(in)         :: x(:)
    character(len=32)            :: nm
    nm = name
    write(gu) nm, 1_i4, int(size(x), i8), x
  end subroutine gio_real

  subroutine gio_int(name, x)
    character(len=*), intent(in) :: name
    integer, intent(in)          :: x(:)
    character(len=32)            :: nm
    nm = name
    write(gu) nm, 2_i4, int(size(x), i8), int(x, i4)
  end subroutine gio_int

end module m_golden_io
