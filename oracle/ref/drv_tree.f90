#include "cpp_macros.h"
!> Golden vectors of afivo's own tree routines, built from the reference's
!> sources, for the pins of the oracle's box and tree operations:
!> - ghost cells (af_gc_tree with af_gc_interp_lim, af_gc_interp, Neumann
!>   and Dirichlet boundaries, corners; afivo/src/m_af_ghostcell.f90:24-170,
!>   173-279, 394-612, 654-669, 860-924);
!> - restriction (af_restrict_tree / af_restrict_box, m_af_restrict.f90:50-136);
!> - two-layer halos (af_gc2_box / gc2_prolong_rb, m_af_ghostcell.f90:672-856);
!> - the constant Laplacian / Helmholtz stencils of the multigrid (the
!>   coefficients of mg_box_lpl_stencil, m_af_multigrid.f90:1246-1264) in
!>   stencil_gsrb_357 and stencil_apply_357, and linear prolongation by
!>   stencil_prolong_248 (m_af_stencil.f90:367-496, 695-815, 838-998);
!> - limited and linear prolongation (af_prolong_limit, af_prolong_linear,
!>   m_af_prolong.f90:311-420, 531-679);
!> - af_consistent_fluxes (m_af_core.f90:1257-1402);
!> - flux_upwind_tree (m_af_flux_schemes.f90:666-848) with the electron
!>   flux of the fluid model (src/m_fluid.f90:102-227, written here for the
!>   standard model: constant gas density, no energy equation, immobile
!>   ions) on the reference's transport table;
!> - flux_update_densities (m_af_flux_schemes.f90:320-436) without sources.
!> Usage: drv_tree <out-file> <xyz|cyl> <cells per box side> <levels> <salt>
!> (the salt varies which boxes the refinement rule picks)
program drv_tree
  use m_af_types
  use m_af_core
  use m_af_ghostcell
  use m_af_restrict
  use m_af_prolong
  use m_af_utils
  use m_af_flux_schemes
  use m_af_limiters
  use m_af_stencil
  use m_lookup_table
  use m_units_constants
  use m_ref_common, only: rnd, transport_input, build_td_table, n_in, gas_n, &
       SI_to_Townsend
  use m_golden_io
  implicit none

  integer, parameter :: nc_max = 16
  integer, parameter :: key_lpl0 = 1, key_lplL = 2, key_p248 = 3
  real(dp), parameter :: lambda2 = 5.0e7_dp
  real(dp), parameter :: dt_update = 2.5e-12_dp

  type(af_t)         :: tree
  type(ref_info_t)   :: ref_info
  type(LT_t)         :: td_tbl
  integer            :: i_a, i_b, i_c, i_phi, i_rhs, i_tmp, f_flux, f_field
  integer            :: coord, n, id, p_id, lvl, n_boxes, i, nc, max_lvl, salt
  integer, allocatable :: leaf_ids(:)
  real(dp), allocatable :: halos(:)
  real(dp), allocatable :: cc2(DTIMES(:), :)
  real(dp)           :: dt_lim(2), max_td
  real(dp)           :: E(n_in), mu(n_in), dif(n_in), alpha(n_in), eta(n_in)
  character(len=256) :: fname, arg

  call get_command_argument(1, fname)
  call get_command_argument(2, arg)
  coord = af_xyz
  if (arg == "cyl") coord = af_cyl
  call get_command_argument(3, arg)
  read(arg, *) nc
  call get_command_argument(4, arg)
  read(arg, *) max_lvl
  call get_command_argument(5, arg)
  read(arg, *) salt
  if (nc < 4 .or. nc > nc_max .or. max_lvl < 2) error stop "drv_tree: arguments"
  allocate(cc2(DTIMES(-1:nc+2), 1))

  call af_add_cc_variable(tree, "a", ix=i_a)
  call af_add_cc_variable(tree, "b", ix=i_b)
  call af_add_cc_variable(tree, "c", ix=i_c)
  call af_add_cc_variable(tree, "phi", ix=i_phi)
  call af_add_cc_variable(tree, "rhs", ix=i_rhs)
  call af_add_cc_variable(tree, "tmp", ix=i_tmp)
  call af_add_fc_variable(tree, "flux", ix=f_flux)
  call af_add_fc_variable(tree, "field", ix=f_field)
  ! a: a density (Neumann zero, limited interpolation at refinement
  ! boundaries, limited prolongation); b: like |E| (Neumann zero, linear
  ! interpolation); c: Dirichlet zero on the faces of the last dimension
  call af_set_cc_methods(tree, i_a, af_bc_neumann_zero, af_gc_interp_lim, &
       prolong=af_prolong_limit)
  call af_set_cc_methods(tree, i_b, af_bc_neumann_zero, af_gc_interp)
  call af_set_cc_methods(tree, i_c, bc_last_dirichlet, af_gc_interp)

  call af_init(tree, nc, [DTIMES(2e-3_dp)], [DTIMES(nc)], coord=coord)
  do n = 1, max_lvl
     call af_adjust_refinement(tree, ref_rule, ref_info, 0)
     if (ref_info%n_add == 0) exit
  end do
  n_boxes = tree%highest_id

  call transport_input(E, mu, dif, alpha, eta)
  call build_td_table(E, mu, dif, alpha, eta, td_tbl, max_td)

  call gio_open(trim(fname))
  call write_tree()
  call gio_real("lambda2", [lambda2])
  call gio_real("dt_update", [dt_update])
  call gio_real("td.x_min", [td_tbl%x_min])
  call gio_real("td.inv_fac", [td_tbl%inv_fac])
  call gio_real("td.cols", reshape(td_tbl%rows_cols, [size(td_tbl%rows_cols)]))
  call gio_real("gas_n", [gas_n])

  call init_values()
  call write_cc("init.a", i_a)
  call write_cc("init.b", i_b)
  call write_cc("init.c", i_c)
  call write_cc("init.phi", i_phi)
  call write_cc("init.rhs", i_rhs)
  call write_cc("init.tmp", i_tmp)
  call write_fc("init.flux", f_flux)
  call write_fc("init.field", f_field)

  ! ghost cells of every box, corners included
  call af_gc_tree(tree, [i_a, i_b, i_c])
  call write_cc("gc.a", i_a)
  call write_cc("gc.b", i_b)
  call write_cc("gc.c", i_c)

  call init_values()
  call af_restrict_tree(tree, [i_a])
  call write_cc("restrict.a", i_a)

  ! two-layer halos of the leaves, level by level in leaf order (each call
  ! also writes its first layer back into the box)
  call init_values()
  allocate(leaf_ids(0), halos(0))
  do lvl = 1, tree%highest_lvl
     do i = 1, size(tree%lvls(lvl)%leaves)
        id = tree%lvls(lvl)%leaves(i)
        call af_gc2_box(tree, id, [i_a], cc2)
        leaf_ids = [leaf_ids, id]
        halos = [halos, reshape(cc2, [size(cc2)])]
     end do
  end do
  call gio_int("gc2.ids", leaf_ids)
  call gio_real("gc2.halo", halos)
  call write_cc("gc2.a", i_a)

  ! stencils: the Laplacian (lambda 0) and a Helmholtz operator, linear
  ! prolongation
  call af_stencil_store(tree, key_lpl0, set_lpl0)
  call af_stencil_store(tree, key_lplL, set_lplL)
  call af_stencil_store(tree, key_p248, set_p248)

  call init_values()
  do id = 1, n_boxes
     call af_stencil_gsrb_box(tree%boxes(id), key_lpl0, 1, i_phi, i_rhs)
     call af_stencil_gsrb_box(tree%boxes(id), key_lpl0, 2, i_phi, i_rhs)
  end do
  call write_cc("gsrb0.phi", i_phi)
  call init_values()
  do id = 1, n_boxes
     call af_stencil_gsrb_box(tree%boxes(id), key_lplL, 1, i_phi, i_rhs)
     call af_stencil_gsrb_box(tree%boxes(id), key_lplL, 2, i_phi, i_rhs)
  end do
  call write_cc("gsrbL.phi", i_phi)

  call init_values()
  do id = 1, n_boxes
     call af_stencil_apply_box(tree%boxes(id), key_lpl0, i_phi, i_tmp)
  end do
  call write_cc("apply0.tmp", i_tmp)
  call init_values()
  do id = 1, n_boxes
     call af_stencil_apply_box(tree%boxes(id), key_lplL, i_phi, i_tmp)
  end do
  call write_cc("applyL.tmp", i_tmp)

  call init_values()
  do id = 1, n_boxes
     p_id = tree%boxes(id)%parent
     if (p_id > af_no_box) call af_stencil_prolong_box(tree%boxes(p_id), &
          tree%boxes(id), key_p248, i_phi, i_tmp, add=.true.)
  end do
  call write_cc("prolong.tmp", i_tmp)

  ! prolongation of every box from its parent, in id order
  call init_values()
  do id = 1, n_boxes
     p_id = tree%boxes(id)%parent
     if (p_id > af_no_box) call af_prolong_limit(tree%boxes(p_id), &
          tree%boxes(id), i_a, limiter=tree%cc_methods(i_a)%prolong_limiter)
  end do
  call write_cc("prolim.a", i_a)
  call init_values()
  do id = 1, n_boxes
     p_id = tree%boxes(id)%parent
     if (p_id > af_no_box) call af_prolong_linear(tree%boxes(p_id), &
          tree%boxes(id), i_b)
  end do
  call write_cc("prolin.b", i_b)

  call init_values()
  call af_consistent_fluxes(tree, [f_flux])
  call write_fc("consistent.flux", f_flux)

  ! the upwind flux of the electrons (a), with |E| in b and the face field
  call init_values()
  call flux_upwind_tree(tree, 1, [i_a], 0, [f_flux], 2, dt_lim, &
       electron_flux, electron_direction, no_line_modify, af_limiter_koren_t)
  call write_fc("flux.flux", f_flux)
  call write_cc("flux.a", i_a)
  call gio_real("flux.dt_lim", dt_lim)

  ! forward Euler's update from the (random) fluxes, without sources
  call init_values()
  call flux_update_densities(tree, dt_update, 1, [i_a], 1, [i_a], [f_flux], &
       0, 1, [0], [1.0_dp], 0, no_source, 1, dt_lim(1:1))
  call write_cc("update.a", i_a)

  call gio_close()

contains

  !> Deterministic hash of a box's level and index
  integer function box_hash(box)
    type(box_t), intent(in) :: box
    integer(8), parameter :: mult(3) = [73856093_8, 19349663_8, 83492791_8]
    integer(8) :: h
    integer    :: d
    h = 2654435_8 * box%lvl + 7919_8 * salt
    do d = 1, NDIM
       h = h + box%ix(d) * mult(d)
    end do
    box_hash = int(mod(h, 1000003_8))
  end function box_hash

  !> Level 1 refined everywhere, then about three in seven boxes per level
  subroutine ref_rule(box, cell_flags)
    type(box_t), intent(in) :: box
    integer, intent(out)    :: cell_flags(DTIMES(box%n_cell))
    if (box%lvl < 2) then
       cell_flags = af_do_ref
    else if (box%lvl < max_lvl .and. mod(box_hash(box), 7) < 3) then
       cell_flags = af_do_ref
    else
       cell_flags = af_keep_ref
    end if
  end subroutine ref_rule

  subroutine bc_last_dirichlet(box, nb, iv, coords, bc_val, bc_type)
    type(box_t), intent(in) :: box
    integer, intent(in)     :: nb
    integer, intent(in)     :: iv
    real(dp), intent(in)    :: coords(NDIM, box%n_cell**(NDIM-1))
    real(dp), intent(out)   :: bc_val(box%n_cell**(NDIM-1))
    integer, intent(out)    :: bc_type
    if (af_neighb_dim(nb) == NDIM) then
       bc_type = af_bc_dirichlet
    else
       bc_type = af_bc_neumann
    end if
    bc_val = 0.0_dp
  end subroutine bc_last_dirichlet

  !> Every cell of every box (ghost cells included) of every variable,
  !> drawn in id order
  subroutine init_values()
    integer(8) :: s
    integer    :: idb, q, nq, nf
    real(dp), allocatable :: v(:)
    s = 987654321_8
    nq = (nc+2)**NDIM
    nf = (nc+1)**NDIM * NDIM
    allocate(v(max(nq, nf)))
    do idb = 1, n_boxes
       associate (bx => tree%boxes(idb))
         do q = 1, nq                      ! densities over four decades
            v(q) = 1e15_dp * 10.0_dp**(4 * rnd(s) - 2)
         end do
         bx%cc(DTIMES(:), i_a) = reshape(v(1:nq), [DTIMES(nc+2)])
         do q = 1, nq                      ! fields across the whole table
            v(q) = 3.5e7_dp * rnd(s)
         end do
         bx%cc(DTIMES(:), i_b) = reshape(v(1:nq), [DTIMES(nc+2)])
         do q = 1, nq
            v(q) = 2 * rnd(s) - 1
         end do
         bx%cc(DTIMES(:), i_c) = reshape(v(1:nq), [DTIMES(nc+2)])
         do q = 1, nq
            v(q) = 100 * (2 * rnd(s) - 1)
         end do
         bx%cc(DTIMES(:), i_phi) = reshape(v(1:nq), [DTIMES(nc+2)])
         do q = 1, nq
            v(q) = 1e9_dp * (2 * rnd(s) - 1)
         end do
         bx%cc(DTIMES(:), i_rhs) = reshape(v(1:nq), [DTIMES(nc+2)])
         do q = 1, nq
            v(q) = 2 * rnd(s) - 1
         end do
         bx%cc(DTIMES(:), i_tmp) = reshape(v(1:nq), [DTIMES(nc+2)])
         do q = 1, nf
            v(q) = 1e21_dp * (2 * rnd(s) - 1)
         end do
         bx%fc(DTIMES(:), :, f_flux) = reshape(v(1:nf), [DTIMES(nc+1), NDIM])
         do q = 1, nf                      ! face fields of both signs
            v(q) = 3e7_dp * (2 * rnd(s) - 1)
         end do
         bx%fc(DTIMES(:), :, f_field) = reshape(v(1:nf), [DTIMES(nc+1), NDIM])
       end associate
    end do
  end subroutine init_values

  subroutine write_tree()
    integer, allocatable  :: lv(:), ix(:), par(:), ch(:), nb(:), nbm(:)
    real(dp), allocatable :: rmin(:), dr(:)
    integer :: idb, q, n_mat
    n_mat = 3**NDIM
    allocate(lv(n_boxes), ix(NDIM*n_boxes), par(n_boxes), ch(af_num_children*n_boxes), &
         nb(af_num_neighbors*n_boxes), nbm(n_mat*n_boxes), rmin(NDIM*n_boxes), dr(NDIM*n_boxes))
    ! element by element (the neighbour matrix as box%neighbor_mat(DTIMES(-1:1)), i fastest)
    do idb = 1, n_boxes
       lv(idb) = tree%boxes(idb)%lvl
       par(idb) = tree%boxes(idb)%parent
       do q = 1, NDIM
          ix((idb-1)*NDIM + q) = tree%boxes(idb)%ix(q)
          rmin((idb-1)*NDIM + q) = tree%boxes(idb)%r_min(q)
          dr((idb-1)*NDIM + q) = tree%boxes(idb)%dr(q)
       end do
       do q = 1, af_num_children
          ch((idb-1)*af_num_children + q) = tree%boxes(idb)%children(q)
       end do
       do q = 1, af_num_neighbors
          nb((idb-1)*af_num_neighbors + q) = tree%boxes(idb)%neighbors(q)
       end do
       do q = 1, n_mat
#if NDIM == 2
          nbm((idb-1)*n_mat + q) = tree%boxes(idb)%neighbor_mat(mod(q-1, 3) - 1, (q-1)/3 - 1)
#else
          nbm((idb-1)*n_mat + q) = tree%boxes(idb)%neighbor_mat(mod(q-1, 3) - 1, &
               mod((q-1)/3, 3) - 1, (q-1)/9 - 1)
#endif
       end do
    end do
    call gio_int("tree.coord", [coord])
    call gio_int("tree.lvl", lv)
    call gio_int("tree.ix", ix)
    call gio_int("tree.parent", par)
    call gio_int("tree.children", ch)
    call gio_int("tree.neighbors", nb)
    call gio_int("tree.nbmat", nbm)
    call gio_real("tree.r_min", rmin)
    call gio_real("tree.dr", dr)
  end subroutine write_tree

  subroutine write_cc(name, iv)
    character(len=*), intent(in) :: name
    integer, intent(in)          :: iv
    real(dp), allocatable        :: v(:)
    integer                      :: idb
    allocate(v(0))
    do idb = 1, n_boxes
       v = [v, reshape(tree%boxes(idb)%cc(DTIMES(:), iv), [(nc+2)**NDIM])]
    end do
    call gio_real(name, v)
  end subroutine write_cc

  subroutine write_fc(name, ifc)
    character(len=*), intent(in) :: name
    integer, intent(in)          :: ifc
    real(dp), allocatable        :: v(:)
    integer                      :: idb
    allocate(v(0))
    do idb = 1, n_boxes
       v = [v, reshape(tree%boxes(idb)%fc(DTIMES(:), :, ifc), [(nc+1)**NDIM * NDIM])]
    end do
    call gio_real(name, v)
  end subroutine write_fc

  !> The Laplacian's constant coefficients (the stencil mg_box_lpl_stencil
  !> stores), with Helmholtz term h2
  subroutine set_lpl(box, st, h2)
    type(box_t), intent(in)        :: box
    type(stencil_t), intent(inout) :: st
    real(dp), intent(in)           :: h2
    real(dp)                       :: inv_dr2(NDIM)
    integer                        :: idim
    st%shape = af_stencil_357
    st%stype = stencil_constant
    st%cylindrical_gradient = (box%coord_t == af_cyl)
    inv_dr2 = 1 / box%dr**2
    call af_stencil_allocate_coeff(st, box%n_cell)
    do idim = 1, NDIM
       st%c(2*idim:2*idim+1) = inv_dr2(idim)
    end do
    st%c(1) = -sum(st%c(2:)) - h2
  end subroutine set_lpl

  subroutine set_lpl0(box, st)
    type(box_t), intent(in)        :: box
    type(stencil_t), intent(inout) :: st
    call set_lpl(box, st, 0.0_dp)
  end subroutine set_lpl0

  subroutine set_lplL(box, st)
    type(box_t), intent(in)        :: box
    type(stencil_t), intent(inout) :: st
    call set_lpl(box, st, lambda2)
  end subroutine set_lplL

  !> Linear prolongation weights (the stencil
  !> mg_box_prolong_linear_stencil stores)
  subroutine set_p248(box, st)
    type(box_t), intent(in)        :: box
    type(stencil_t), intent(inout) :: st
    st%shape = af_stencil_p248
    st%stype = stencil_constant
    call af_stencil_allocate_coeff(st, box%n_cell)
#if NDIM == 2
    st%c(:) = [9, 3, 3, 1] / 16.0_dp
#elif NDIM == 3
    st%c(:) = [27, 9, 9, 3, 9, 3, 3, 1] / 64.0_dp
#endif
  end subroutine set_p248

  !> The electron flux of the fluid model (src/m_fluid.f90:102-209) for the
  !> standard model: constant gas density, no energy equation, no mobile
  !> ions; |E| in b, the face field in field
  subroutine electron_flux(nf, n_var, flux_dim, u, flux, cfl_sum, &
       n_other_dt, other_dt, box, line_ix, s_deriv)
    integer, intent(in)     :: nf, n_var, flux_dim
    real(dp), intent(in)    :: u(nf, n_var)
    real(dp), intent(out)   :: flux(nf, n_var)
    real(dp), intent(out)   :: cfl_sum(nf-1)
    integer, intent(in)     :: n_other_dt
    real(dp), intent(inout) :: other_dt(n_other_dt)
    type(box_t), intent(in) :: box
    integer, intent(in)     :: line_ix(NDIM-1)
    integer, intent(in)     :: s_deriv
    real(dp) :: E_cc(0:nf), E_x(nf), ne_cc(0:nf), v(nf), dc(nf)
    real(dp) :: tmp_fc(nf), N_inv(nf), mu(nf), sigma(nf), inv_dx, cfl_factor
    integer  :: ncl

    ncl = box%n_cell
    inv_dx = 1/box%dr(flux_dim)
    N_inv = 1/gas_n
    call flux_get_line_1fc(box, f_field, flux_dim, line_ix, E_x)
    call flux_get_line_1cc(box, i_a+s_deriv, flux_dim, line_ix, ne_cc)
    call flux_get_line_1cc(box, i_b, flux_dim, line_ix, E_cc)
    tmp_fc = 0.5_dp * (E_cc(0:ncl) + E_cc(1:ncl+1)) * SI_to_Townsend * N_inv
    mu = LT_get_col(td_tbl, 1, tmp_fc) * N_inv
    dc = LT_get_col(td_tbl, 2, tmp_fc) * N_inv
    v = -mu * E_x
    flux(:, 1) = v * u(:, 1) - dc * inv_dx * (ne_cc(1:ncl+1) - ne_cc(0:ncl))
    sigma = mu * u(:, 1)
    cfl_factor = 1.0_dp
    cfl_sum = cfl_factor * max(abs(v(2:)), abs(v(:nf-1))) * inv_dx + &
         2 * max(dc(2:), dc(:nf-1))  * inv_dx**2
    other_dt(1) = UC_eps0 / (UC_elem_charge * max(maxval(sigma), 1e-100_dp))
  end subroutine electron_flux

  !> Electrons drift against the field (flux_direction, src/m_fluid.f90:212-227)
  subroutine electron_direction(box, line_ix, s_deriv, n_var, flux_dim, &
       direction_positive)
    type(box_t), intent(in) :: box
    integer, intent(in)     :: line_ix(NDIM-1)
    integer, intent(in)     :: s_deriv
    integer, intent(in)     :: n_var
    integer, intent(in)     :: flux_dim
    logical, intent(out)    :: direction_positive(box%n_cell+1, n_var)
    real(dp)                :: E_x(box%n_cell+1)
    call flux_get_line_1fc(box, f_field, flux_dim, line_ix, E_x)
    direction_positive(:, 1) = (-1 * E_x > 0)
  end subroutine electron_direction

  subroutine no_line_modify(n_cc, n_var, cc_line, flux_dim, box, line_ix, s_deriv)
    integer, intent(in)     :: n_cc
    integer, intent(in)     :: n_var
    real(dp), intent(inout) :: cc_line(n_cc, n_var)
    integer, intent(in)     :: flux_dim
    type(box_t), intent(in) :: box
    integer, intent(in)     :: line_ix(NDIM-1)
    integer, intent(in)     :: s_deriv
  end subroutine no_line_modify

  subroutine no_source(box, dt, n_vars, i_cc, s_deriv, s_out, n_dt, dt_lim, mask)
    type(box_t), intent(inout) :: box
    real(dp), intent(in)       :: dt
    integer, intent(in)        :: n_vars
    integer, intent(in)        :: i_cc(n_vars)
    integer, intent(in)        :: s_deriv
    integer, intent(in)        :: s_out
    integer, intent(in)        :: n_dt
    real(dp), intent(inout)    :: dt_lim(n_dt)
    logical, intent(in)        :: mask(DTIMES(box%n_cell))
  end subroutine no_source

end program drv_tree
