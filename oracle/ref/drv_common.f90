!> Shared parts of the reference drivers: the random numbers of the inputs
!> and the transport table as transport_data_initialize forms it from
!> old-style input data (src/m_transport_data.f90:85-126).
module m_ref_common
  use m_lookup_table
  use m_table_data
  use m_config
  use m_units_constants
  implicit none
  private

  integer, parameter, public :: dp = kind(0.0d0)
  real(dp), parameter, public :: SI_to_Townsend = 1e21_dp, Townsend_to_SI = 1e-21_dp
  !> gas number density at 1 bar and 300 K (src/m_gas.f90:174-176)
  real(dp), parameter, public :: gas_n = 1e5_dp / (UC_boltzmann_const * 300.0_dp)
  integer, parameter, public :: n_in = 1000

  public :: rnd, transport_input, build_td_table

contains

  !> Park-Miller minimal standard generator, values in (0, 1)
  real(dp) function rnd(s)
    integer(8), intent(inout) :: s
    s = mod(48271_8 * s, 2147483647_8)
    rnd = real(s, dp) / 2147483647.0_dp
  end function rnd

  !> Old-style transport data: mu 0.03, D 0.1, alpha(E) = 4.3e5
  !> exp(-1.9e7 / E) (0 at E = 0), eta 0, at 1000 fields from 0 to 3e7 V/m
  subroutine transport_input(E, mu, dif, alpha, eta)
    real(dp), intent(out) :: E(n_in), mu(n_in), dif(n_in), alpha(n_in), eta(n_in)
    integer :: i
    do i = 1, n_in
       E(i) = (i - 1) * (3e7_dp / (n_in - 1))
       mu(i) = 0.03_dp
       dif(i) = 0.1_dp
       if (E(i) > 0) then
          alpha(i) = 4.3e5_dp * exp(-1.9e7_dp / E(i))
       else
          alpha(i) = 0
       end if
       eta(i) = 0
    end do
  end subroutine transport_input

  !> td_tbl of transport_data_initialize (old style): mobility * N,
  !> diffusion * N, alpha / N and eta / N against E/N in Td, tabulated by
  !> table_set_column with the table_data defaults
  subroutine build_td_table(E, mu, dif, alpha, eta, td_tbl, max_td)
    real(dp), intent(in)    :: E(:), mu(:), dif(:), alpha(:), eta(:)
    type(LT_t), intent(out) :: td_tbl
    real(dp), intent(out)   :: max_td
    real(dp)                :: xx(size(E)), yy(size(E))
    type(CFG_t)             :: cfg

    ! the defaults of table_data (size 1000, linear input interpolation and
    ! spacing, from 0 Td to the data's last field)
    call table_data_initialize(cfg)
    xx = E * SI_to_Townsend / gas_n
    yy = mu * gas_n
    if (table_max_townsend < 0) then
       max_td = xx(size(xx))
    else
       max_td = table_max_townsend
    end if
    td_tbl = LT_create(table_min_townsend, max_td, table_size, 4, table_xspacing)
    call table_set_column(td_tbl, 1, xx, yy)
    yy = dif * gas_n
    call table_set_column(td_tbl, 2, xx, yy)
    yy = alpha / gas_n
    call table_set_column(td_tbl, 3, xx, yy)
    yy = eta / gas_n
    call table_set_column(td_tbl, 4, xx, yy)
  end subroutine build_td_table

end module m_ref_common
