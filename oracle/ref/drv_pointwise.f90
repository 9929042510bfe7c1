!> Golden vectors of the reference's point-wise routines, built from its own
!> sources: the slope limiters (afivo/src/m_af_limiters.f90), the transport
!> lookup table as transport_data_initialize forms it from old-style input
!> data (src/m_transport_data.f90:85-126 with table_set_column,
!> src/m_table_data.f90:82-118, and LT_create / get_x,
!> src/lookup_table_fortran/m_lookup_table.f90:186-327), the field table
!> of the standard chemistry (src/m_chemistry.f90:229-256, 345-372), and
!> lookups in them (LT_get_loc / LT_get_col, m_lookup_table.f90:330-406).
!> Usage: drv_pointwise <out-file>
program drv_pointwise
  use m_af_limiters
  use m_lookup_table
  use m_table_data
  use m_ref_common
  use m_golden_io
  implicit none

  integer, parameter  :: n_pairs = 4096, n_look = 3000
  real(dp)            :: a(n_pairs), b(n_pairs), lim(n_pairs)
  real(dp)            :: E(n_in), mu(n_in), dif(n_in), alpha(n_in), eta(n_in)
  real(dp)            :: x_look(n_look), y_look(n_look)
  real(dp)            :: frac_look(n_look), max_td, r1, r2, r3, r4
  real(dp), allocatable :: ion(:), att(:)
  integer             :: i, col, ix_look(n_look)
  integer(8)          :: rs
  type(LT_t)          :: td_tbl, chemtbl
  type(LT_loc_t)      :: loc
  character(len=256)  :: fname

  call get_command_argument(1, fname)
  call gio_open(trim(fname))
  rs = 20260101_8

  ! ---- limiters: random pairs spanning every branch, and edge cases
  do i = 1, n_pairs
     r1 = rnd(rs)
     r2 = rnd(rs)
     r3 = rnd(rs)
     r4 = rnd(rs)
     a(i) = (2 * r1 - 1) * 10.0_dp**(4 * r2 - 2)
     select case (mod(i, 8))
     case (0)
        b(i) = a(i) * (4 * r3 - 1)          ! ratios around the Koren breaks
     case (1)
        b(i) = 0                            ! one slope zero
     case (2)
        b(i) = a(i)                         ! equal slopes
     case (3)
        b(i) = a(i) * 4                     ! a/b = 1/4 exactly
     case (4)
        b(i) = a(i) / 2.5_dp                ! a/b = 2.5 (up to rounding)
     case (5)
        b(i) = -a(i) * r3                   ! opposite signs
     case default
        b(i) = (2 * r3 - 1) * 10.0_dp**(4 * r4 - 2)
     end select
  end do
  a(1) = 0; b(1) = 0
  a(2) = 0; b(2) = 1
  call gio_real("lim.a", a)
  call gio_real("lim.b", b)
  lim = af_limiter_koren(a, b)
  call gio_real("lim.koren", lim)
  lim = af_limiter_mc(a, b)
  call gio_real("lim.mc", lim)
  lim = af_limiter_gminmod43(a, b)
  call gio_real("lim.gminmod43", lim)

  ! ---- old-style transport data and the tables formed from it
  call transport_input(E, mu, dif, alpha, eta)
  call gio_real("in.E", E)
  call gio_real("in.mu", mu)
  call gio_real("in.dif", dif)
  call gio_real("in.alpha", alpha)
  call gio_real("in.eta", eta)
  call gio_real("in.N", [gas_n])
  call build_td_table(E, mu, dif, alpha, eta, td_tbl, max_td)
  call gio_real("td.x", td_tbl%x)
  call gio_real("td.x_min", [td_tbl%x_min])
  call gio_real("td.inv_fac", [td_tbl%inv_fac])
  call gio_real("td.cols", reshape(td_tbl%rows_cols, [size(td_tbl%rows_cols)]))

  ! the standard model's field table (src/m_chemistry.f90:229-256, 345-372):
  ! rates alpha * mu * E/N (SI) * N and eta * mu * E/N * N on the td grid,
  ! tabulated again on the same range
  ion = td_tbl%rows_cols(:, 3) * td_tbl%rows_cols(:, 1) * td_tbl%x * &
       Townsend_to_SI * gas_n
  att = td_tbl%rows_cols(:, 4) * td_tbl%rows_cols(:, 1) * td_tbl%x * &
       Townsend_to_SI * gas_n
  chemtbl = LT_create(td_tbl%x(1), td_tbl%x(td_tbl%n_points), table_size, 2, &
       table_xspacing)
  call table_set_column(chemtbl, 1, td_tbl%x, ion)
  call table_set_column(chemtbl, 2, td_tbl%x, att)
  call gio_real("chem.x_min", [chemtbl%x_min])
  call gio_real("chem.inv_fac", [chemtbl%inv_fac])
  call gio_real("chem.cols", reshape(chemtbl%rows_cols, [size(chemtbl%rows_cols)]))

  ! ---- lookups: random points below, inside and above the range, the
  ! grid points themselves and the range ends
  do i = 1, n_look
     select case (mod(i, 4))
     case (0)
        x_look(i) = td_tbl%x(1 + mod(i * 7, td_tbl%n_points))
     case default
        x_look(i) = (1.2_dp * rnd(rs) - 0.1_dp) * max_td
     end select
  end do
  x_look(1) = td_tbl%x_min
  x_look(2) = max_td
  x_look(3) = -1.0_dp
  x_look(4) = 2 * max_td
  call gio_real("look.x", x_look)
  do i = 1, n_look
     loc = LT_get_loc(td_tbl, x_look(i))
     ix_look(i) = loc%low_ix
     frac_look(i) = loc%low_frac
  end do
  call gio_int("look.low_ix", ix_look)
  call gio_real("look.low_frac", frac_look)
  do col = 1, 4
     y_look = LT_get_col(td_tbl, col, x_look)
     call gio_real("look.td" // achar(iachar('0') + col), y_look)
  end do
  do col = 1, 2
     y_look = LT_get_col(chemtbl, col, x_look)
     call gio_real("look.chem" // achar(iachar('0') + col), y_look)
  end do

  call gio_close()

end program drv_pointwise
