!> A Fortran driver of libafgpu through the m_afgpu interface module, as
!> afivo-streamer's time loop would call it (src/streamer.f90:181-365):
!> create, sync_tree, set_params, set_tables, set_voltage, put_cc, the
!> initial field_compute, then n_steps steps; it writes every step's
!> results and the final electron density, potential and field.
!> Usage: drive_step <in-file> <out-file>
!> The input file (little-endian stream) holds int32 ndim, n_cell, coord,
!> precision, n_boxes, n_steps, n_rows; the tree arrays (int32 lvl, ix,
!> parent, children, neighbors, neighbor_mat; float64 r_min, dr); the bytes
!> of an afgpu_params_t as C lays it out; float64 td_x_min, td_inv_fac,
!> chem_x_min, chem_inv_fac, the 4 + 2 table columns, the voltage and the
!> initial electron (= ion) density of every box.
program drive_step
  use iso_c_binding
  use m_afgpu
  implicit none

  integer(c_int)              :: ndim, n_cell, coord, precision, n_boxes, n_steps, n_rows
  integer(c_int), allocatable :: lvl(:), ix(:), parent(:), children(:), nbs(:), nbmat(:), ids(:)
  real(c_double), allocatable :: r_min(:), dr(:), td(:), chem(:), ne(:), out(:)
  real(c_double)              :: td_x_min, td_inv_fac, chem_x_min, chem_inv_fac, voltage, res
  integer(c_int8_t), allocatable, target :: raw(:)
  type(afgpu_params_t), pointer :: prm
  type(afgpu_step_info_t), allocatable, target :: infos(:)
  type(c_ptr)                 :: h
  integer(c_int)              :: n_cyc, ncell, i, u
  character(len=512)          :: fin, fout
  type(afgpu_params_t)        :: size_probe

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(newunit=u, file=trim(fin), access="stream", form="unformatted", status="old", action="read")
  read(u) ndim, n_cell, coord, precision, n_boxes, n_steps, n_rows
  allocate(lvl(n_boxes), ix(ndim*n_boxes), parent(n_boxes), children(2**ndim*n_boxes), &
       nbs(2*ndim*n_boxes), nbmat(3**ndim*n_boxes), r_min(ndim*n_boxes), dr(ndim*n_boxes))
  read(u) lvl, ix, parent, children, nbs, nbmat, r_min, dr
  allocate(raw(c_sizeof(size_probe)))
  read(u) raw
  call c_f_pointer(c_loc(raw), prm)
  read(u) td_x_min, td_inv_fac, chem_x_min, chem_inv_fac
  allocate(td(4*n_rows), chem(2*n_rows))
  read(u) td, chem, voltage
  ncell = (n_cell + 2)**ndim
  allocate(ne(ncell*n_boxes), out(ncell*n_boxes), ids(n_boxes), infos(n_steps))
  read(u) ne
  close(u)
  ids = [(i, i = 1, n_boxes)]

  call check(afgpu_create(ndim, n_cell, coord, precision, h), "afgpu_create")
  call check(afgpu_sync_tree(h, n_boxes, lvl, ix, parent, children, nbs, nbmat, r_min, dr), &
       "afgpu_sync_tree")
  call check(afgpu_set_params(h, prm), "afgpu_set_params")
  call check(afgpu_set_tables(h, n_rows, td_x_min, td_inv_fac, td, chem_x_min, chem_inv_fac, chem), &
       "afgpu_set_tables")
  call check(afgpu_set_voltage(h, voltage), "afgpu_set_voltage")
  call check(afgpu_put_cc(h, afgpu_cc_electron, n_boxes, ids, ne), "afgpu_put_cc")
  call check(afgpu_put_cc(h, afgpu_cc_pos_ion, n_boxes, ids, ne), "afgpu_put_cc")
  call check(afgpu_field_compute(h, 0, 0, n_cyc, res), "afgpu_field_compute")
  do i = 1, n_steps
     call check(afgpu_step(h, infos(i)), "afgpu_step")
  end do

  open(newunit=u, file=trim(fout), access="stream", form="unformatted", status="replace", action="write")
  write(u) n_cyc, res, int(c_sizeof(infos(1)), c_int)
  do i = 1, n_steps
     block
       integer(c_int8_t), pointer :: b(:)
       call c_f_pointer(c_loc(infos(i)), b, [c_sizeof(infos(i))])
       write(u) b
     end block
  end do
  call check(afgpu_get_cc(h, afgpu_cc_electron, n_boxes, ids, out), "afgpu_get_cc")
  write(u) out
  call check(afgpu_get_cc(h, afgpu_cc_phi, n_boxes, ids, out), "afgpu_get_cc")
  write(u) out
  call check(afgpu_get_cc(h, afgpu_cc_efield, n_boxes, ids, out), "afgpu_get_cc")
  write(u) out
  close(u)
  call check(afgpu_destroy(h), "afgpu_destroy")

contains

  !> error stop on a non-zero status, as the reference does after its hypre
  !> calls (afivo/src/m_coarse_solver.f90:187-189)
  subroutine check(ierr, what)
    integer(c_int), intent(in)   :: ierr
    character(len=*), intent(in) :: what
    if (ierr /= afgpu_ok) then
       print *, what, " failed: ", ierr, " ", afgpu_error_message()
       error stop "libafgpu call failed"
    end if
  end subroutine check

end program drive_step
