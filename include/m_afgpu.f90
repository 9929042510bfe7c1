!> Fortran interface of libafgpu (include/afgpu.h), for afivo-streamer's
!> driver: the bind(c) mirrors of the C types and an interface block for
!> every entry point, in the form the reference already uses for foreign
!> code (afivo/src/m_coarse_solver.f90:46-56: type(c_ptr) handles, explicit
!> interfaces with import, an integer status that the caller turns into
!> error stop, :187-189).  Every entry point returns AFGPU_OK (0) on
!> success; afgpu_error_message returns the library's last message.
module m_afgpu
  use iso_c_binding
  implicit none
  private

  ! status codes
  integer(c_int), parameter, public :: afgpu_ok = 0, afgpu_err_arg = 1, &
       afgpu_err_hip = 2, afgpu_err_state = 3, afgpu_err_no_convergence = 4, &
       afgpu_err_all_rejected = 5, afgpu_err_dt_too_small = 6, &
       afgpu_err_unsupported = 7
  ! coordinates, precision forms, integrators, species boundary conditions
  integer(c_int), parameter, public :: afgpu_xyz = 1, afgpu_cyl = 2
  integer(c_int), parameter, public :: afgpu_prec_fp64 = 0, afgpu_prec_mixed = 1
  integer(c_int), parameter, public :: afgpu_forward_euler_t = 1, afgpu_heuns_method_t = 2
  integer(c_int), parameter, public :: afgpu_species_bc_neumann_zero = 0, &
       afgpu_species_bc_dirichlet_zero = 1
  ! cell-centred and face-centred variables of the device pools
  integer(c_int), parameter, public :: afgpu_cc_electron = 0, afgpu_cc_pos_ion = 3, &
       afgpu_cc_neg_ion = 6, afgpu_cc_phi = 9, afgpu_cc_phi_2 = 10, &
       afgpu_cc_efield = 11, afgpu_cc_rhs = 12, afgpu_cc_tmp = 13, &
       afgpu_cc_photo = 14, afgpu_cc_helmh = 15, afgpu_n_cc = 18
  integer(c_int), parameter, public :: afgpu_fc_flux = 0, afgpu_fc_field = 1, afgpu_n_fc = 2
  ! multigrid operators
  integer(c_int), parameter, public :: afgpu_op_poisson = 0, afgpu_op_helmh = 1, &
       afgpu_max_modes = 3
  ! refinement flags, diagnostics, shapes, chemistry
  integer(c_int), parameter, public :: afgpu_ref_derefine = -2, afgpu_ref_refine = 2, &
       afgpu_ref_max_items = 8
  integer(c_int), parameter, public :: afgpu_diag_max_vars = 8
  ! variable codes of the diagnostics and box I/O lists: species k of a
  ! reaction list is afgpu_var_species0 + k (AFGPU_VAR_SPECIES(k)), face
  ! variable ifc is afgpu_var_face0 + ifc (AFGPU_VAR_FACE(ifc))
  integer(c_int), parameter, public :: afgpu_var_species0 = 100, afgpu_var_face0 = 200
  integer(c_int), parameter, public :: afgpu_noalloc_scratch_words = 4194304
  integer(c_int), parameter, public :: afgpu_n_kernel_classes = 8
  integer(c_int), parameter, public :: afgpu_chem_max_species = 100, &
       afgpu_chem_max_reactions = 500, afgpu_chem_max_in = 8, &
       afgpu_chem_max_out = 8, afgpu_chem_max_coeff = 9
  integer(c_int), parameter, public :: afgpu_rate_tabulated_energy = 0, &
       afgpu_rate_tabulated_field = 1, afgpu_rate_constant = 2, &
       afgpu_rate_linear = 3, afgpu_rate_exp_v1 = 4, afgpu_rate_exp_v2 = 5, &
       afgpu_rate_k1 = 6, afgpu_rate_k3 = 8, afgpu_rate_k16 = 21
  ! execution options
  integer(c_int), parameter, public :: afgpu_opt_graphs = 1, afgpu_opt_direct_n = 2, &
       afgpu_opt_mixed_smooth = 3, afgpu_opt_block_smooth = 4, &
       afgpu_opt_batch_modes = 5, afgpu_opt_direct_mfma = 6, &
       afgpu_opt_photoi_vcycle = 7, afgpu_opt_flux32 = 8, &
       afgpu_opt_step_graph = 9, afgpu_opt_fuse_direct = 10, &
       afgpu_opt_direct_f32 = 11, afgpu_opt_box_lds = 12, afgpu_opt_direct_fold = 13

  !> afgpu_params_t
  type, bind(c), public :: afgpu_params_t
     real(c_double) :: gas_number_density
     integer(c_int) :: time_integrator
     real(c_double) :: dt_max, dt_min, dt_safety_factor, dt_cfl_number, &
          dt_max_growth_factor
     integer(c_int) :: dt_chemistry_limit_loss
     real(c_double) :: dt_chemistry_nmin
     integer(c_int) :: mg_num_vcycles
     real(c_double) :: mg_max_rel_residual
     integer(c_int) :: mg_n_cycle_down, mg_n_cycle_up
     real(c_double) :: domain_len_last
     integer(c_int) :: species_bc, photoi_enabled, photoi_per_steps, photoi_species
     real(c_double) :: photoi_eta_quench
     integer(c_int) :: photoi_n_modes
     real(c_double) :: photoi_lambda(3), photoi_coeff(3)
     real(c_double) :: photoi_max_rel_residual
     integer(c_int) :: photoi_max_fmg_cycles
  end type afgpu_params_t

  !> afgpu_step_info_t
  type, bind(c), public :: afgpu_step_info_t
     integer(c_int)       :: it
     real(c_double)       :: time, dt_used, dt_next, dt_lim
     integer(c_int)       :: n_rejected, n_vcycles
     real(c_double)       :: residual, rel_residual
     integer(c_int)       :: photoi_done
     integer(c_long_long) :: leaf_cells
     integer(c_int)       :: photoi_cycles
  end type afgpu_step_info_t

  !> afgpu_refine_t (C's [item][3] arrays are (3, item) here)
  type, bind(c), public :: afgpu_refine_t
     real(c_double) :: refine_adx, refine_adx_fac, derefine_dx
     real(c_double) :: refine_min_dx, refine_max_dx, refine_min_dens
     real(c_double) :: refine_init_time, refine_init_fac
     real(c_double) :: global_time
     integer(c_int) :: use_alpha_effective, buffer_width, max_lvl, n_seeds
     real(c_double) :: seed_r0(3, 8), seed_r1(3, 8)
     real(c_double) :: seed_width(8)
     integer(c_int) :: n_regions
     real(c_double) :: regions_dr(8), regions_tstop(8)
     real(c_double) :: regions_rmin(3, 8), regions_rmax(3, 8)
     integer(c_int) :: n_limits
     real(c_double) :: limits_dr(8)
     real(c_double) :: limits_rmin(3, 8), limits_rmax(3, 8)
  end type afgpu_refine_t

  !> afgpu_shape_t
  type, bind(c), public :: afgpu_shape_t
     integer(c_int) :: boxes_per_workgroup, workgroups_per_cu
     real(c_double) :: ms
  end type afgpu_shape_t

  !> afgpu_reaction_t
  type, bind(c), public :: afgpu_reaction_t
     integer(c_int) :: rate_type, lookup_col, n_in, n_out
     real(c_double) :: rate_factor
     real(c_double) :: rate_data(9)
     integer(c_int) :: ix_in(8), ix_out(8), multiplicity_out(8)
  end type afgpu_reaction_t

  public :: afgpu_error_message

  interface
     ! ---- lifetime ------------------------------------------------------
     integer(c_int) function afgpu_create(ndim, n_cell, coord, precision, h) &
          bind(c, name="afgpu_create")
       import
       integer(c_int), value :: ndim, n_cell, coord, precision
       type(c_ptr), intent(out) :: h
     end function afgpu_create

     integer(c_int) function afgpu_destroy(h) bind(c, name="afgpu_destroy")
       import
       type(c_ptr), value :: h
     end function afgpu_destroy

     type(c_ptr) function afgpu_last_error() bind(c, name="afgpu_last_error")
       import
     end function afgpu_last_error

     ! ---- tree and box data ----------------------------------------------
     integer(c_int) function afgpu_sync_tree(h, n_boxes, lvl, ix, parent, &
          children, neighbors, neighbor_mat, r_min, dr) bind(c, name="afgpu_sync_tree")
       import
       type(c_ptr), value :: h
       integer(c_int), value :: n_boxes
       integer(c_int), intent(in) :: lvl(*), ix(*), parent(*), children(*), &
            neighbors(*), neighbor_mat(*)
       real(c_double), intent(in) :: r_min(*), dr(*)
     end function afgpu_sync_tree

     integer(c_int) function afgpu_put_cc(h, iv, n_ids, ids, data) bind(c, name="afgpu_put_cc")
       import
       type(c_ptr), value :: h
       integer(c_int), value :: iv, n_ids
       integer(c_int), intent(in) :: ids(*)
       real(c_double), intent(in) :: data(*)
     end function afgpu_put_cc

     integer(c_int) function afgpu_get_cc(h, iv, n_ids, ids, data) bind(c, name="afgpu_get_cc")
       import
       type(c_ptr), value :: h
       integer(c_int), value :: iv, n_ids
       integer(c_int), intent(in) :: ids(*)
       real(c_double), intent(out) :: data(*)
     end function afgpu_get_cc

     integer(c_int) function afgpu_put_fc(h, ifc, n_ids, ids, data) bind(c, name="afgpu_put_fc")
       import
       type(c_ptr), value :: h
       integer(c_int), value :: ifc, n_ids
       integer(c_int), intent(in) :: ids(*)
       real(c_double), intent(in) :: data(*)
     end function afgpu_put_fc

     integer(c_int) function afgpu_get_fc(h, ifc, n_ids, ids, data) bind(c, name="afgpu_get_fc")
       import
       type(c_ptr), value :: h
       integer(c_int), value :: ifc, n_ids
       integer(c_int), intent(in) :: ids(*)
       real(c_double), intent(out) :: data(*)
     end function afgpu_get_fc

     ! ---- settings -------------------------------------------------------
     integer(c_int) function afgpu_set_params(h, p) bind(c, name="afgpu_set_params")
       import
       type(c_ptr), value :: h
       type(afgpu_params_t), intent(in) :: p
     end function afgpu_set_params

     integer(c_int) function afgpu_set_tables(h, n_rows, td_x_min, td_inv_fac, &
          td_cols, chem_x_min, chem_inv_fac, chem_cols) bind(c, name="afgpu_set_tables")
       import
       type(c_ptr), value :: h
       integer(c_int), value :: n_rows
       real(c_double), value :: td_x_min, td_inv_fac, chem_x_min, chem_inv_fac
       real(c_double), intent(in) :: td_cols(*), chem_cols(*)
     end function afgpu_set_tables

     integer(c_int) function afgpu_set_voltage(h, voltage) bind(c, name="afgpu_set_voltage")
       import
       type(c_ptr), value :: h
       real(c_double), value :: voltage
     end function afgpu_set_voltage

     integer(c_int) function afgpu_set_option(h, option, val) bind(c, name="afgpu_set_option")
       import
       type(c_ptr), value :: h
       integer(c_int), value :: option, val
     end function afgpu_set_option

     ! ---- multigrid and field --------------------------------------------
     integer(c_int) function afgpu_mg_fas_vcycle(h, op, set_residual, max_lvl) &
          bind(c, name="afgpu_mg_fas_vcycle")
       import
       type(c_ptr), value :: h
       integer(c_int), value :: op, set_residual, max_lvl
     end function afgpu_mg_fas_vcycle

     integer(c_int) function afgpu_mg_fas_fmg(h, op, set_residual, have_guess) &
          bind(c, name="afgpu_mg_fas_fmg")
       import
       type(c_ptr), value :: h
       integer(c_int), value :: op, set_residual, have_guess
     end function afgpu_mg_fas_fmg

     integer(c_int) function afgpu_tree_maxabs_cc(h, iv, val) bind(c, name="afgpu_tree_maxabs_cc")
       import
       type(c_ptr), value :: h
       integer(c_int), value :: iv
       real(c_double), intent(out) :: val
     end function afgpu_tree_maxabs_cc

     integer(c_int) function afgpu_gc_tree(h, iv) bind(c, name="afgpu_gc_tree")
       import
       type(c_ptr), value :: h
       integer(c_int), value :: iv
     end function afgpu_gc_tree

     integer(c_int) function afgpu_restrict_tree(h, iv) bind(c, name="afgpu_restrict_tree")
       import
       type(c_ptr), value :: h
       integer(c_int), value :: iv
     end function afgpu_restrict_tree

     integer(c_int) function afgpu_field_compute(h, s_in, have_guess, n_cycles, residual) &
          bind(c, name="afgpu_field_compute")
       import
       type(c_ptr), value :: h
       integer(c_int), value :: s_in, have_guess
       integer(c_int), intent(out) :: n_cycles
       real(c_double), intent(out) :: residual
     end function afgpu_field_compute

     integer(c_int) function afgpu_field_from_potential(h) bind(c, name="afgpu_field_from_potential")
       import
       type(c_ptr), value :: h
     end function afgpu_field_from_potential

     ! ---- time integration -------------------------------------------------
     integer(c_int) function afgpu_forward_euler(h, dt, s_deriv, n_prev, s_prev, w_prev, &
          s_out, i_step, n_steps, dt_lim) bind(c, name="afgpu_forward_euler")
       import
       type(c_ptr), value :: h
       real(c_double), value :: dt
       integer(c_int), value :: s_deriv, n_prev, s_out, i_step, n_steps
       integer(c_int), intent(in) :: s_prev(*)
       real(c_double), intent(in) :: w_prev(*)
       real(c_double), intent(out) :: dt_lim
     end function afgpu_forward_euler

     integer(c_int) function afgpu_advance(h, dt, dt_lim) bind(c, name="afgpu_advance")
       import
       type(c_ptr), value :: h
       real(c_double), value :: dt
       real(c_double), intent(out) :: dt_lim
     end function afgpu_advance

     integer(c_int) function afgpu_photoi_set_src(h) bind(c, name="afgpu_photoi_set_src")
       import
       type(c_ptr), value :: h
     end function afgpu_photoi_set_src

     integer(c_int) function afgpu_copy_current_state(h) bind(c, name="afgpu_copy_current_state")
       import
       type(c_ptr), value :: h
     end function afgpu_copy_current_state

     integer(c_int) function afgpu_restore_previous_state(h) bind(c, name="afgpu_restore_previous_state")
       import
       type(c_ptr), value :: h
     end function afgpu_restore_previous_state

     integer(c_int) function afgpu_step(h, info) bind(c, name="afgpu_step")
       import
       type(c_ptr), value :: h
       type(afgpu_step_info_t), intent(out) :: info
     end function afgpu_step

     integer(c_int) function afgpu_get_time_state(h, it, time, global_dt) &
          bind(c, name="afgpu_get_time_state")
       import
       type(c_ptr), value :: h
       integer(c_int), intent(out) :: it
       real(c_double), intent(out) :: time, global_dt
     end function afgpu_get_time_state

     integer(c_int) function afgpu_set_time_state(h, it, time, global_dt) &
          bind(c, name="afgpu_set_time_state")
       import
       type(c_ptr), value :: h
       integer(c_int), value :: it
       real(c_double), value :: time, global_dt
     end function afgpu_set_time_state

     integer(c_int) function afgpu_limit_step_dt(h, dt_cap) bind(c, name="afgpu_limit_step_dt")
       import
       type(c_ptr), value :: h
       real(c_double), value :: dt_cap
     end function afgpu_limit_step_dt

     integer(c_long_long) function afgpu_num_leaf_cells(h) bind(c, name="afgpu_num_leaf_cells")
       import
       type(c_ptr), value :: h
     end function afgpu_num_leaf_cells

     ! ---- level-1 factors (host only) ----------------------------------------
     integer(c_int) function afgpu_coarse_factors(ndim, n_cell, coord, r_min, dr, &
          bc_dirichlet, vinv, v, eig) bind(c, name="afgpu_coarse_factors")
       import
       integer(c_int), value :: ndim, n_cell, coord
       real(c_double), intent(in) :: r_min(*), dr(*)
       integer(c_int), intent(in) :: bc_dirichlet(*)
       real(c_double), intent(out) :: vinv(*), v(*), eig(*)
     end function afgpu_coarse_factors

     ! ---- refinement ---------------------------------------------------------------
     integer(c_int) function afgpu_refine_flags(h, rp, n_boxes, ref_flags) &
          bind(c, name="afgpu_refine_flags")
       import
       type(c_ptr), value :: h
       type(afgpu_refine_t), intent(in) :: rp
       integer(c_int), value :: n_boxes
       integer(c_int), intent(out) :: ref_flags(*)
     end function afgpu_refine_flags

     integer(c_int) function afgpu_restrict_boxes(h, n_ids, ids) bind(c, name="afgpu_restrict_boxes")
       import
       type(c_ptr), value :: h
       integer(c_int), value :: n_ids
       integer(c_int), intent(in) :: ids(*)
     end function afgpu_restrict_boxes

     integer(c_int) function afgpu_prolong_boxes(h, n_ids, ids) bind(c, name="afgpu_prolong_boxes")
       import
       type(c_ptr), value :: h
       integer(c_int), value :: n_ids
       integer(c_int), intent(in) :: ids(*)
     end function afgpu_prolong_boxes

     ! ---- diagnostics and box I/O ------------------------------------------------------
     integer(c_int) function afgpu_regression_values(h, n_vars, ivs, out) &
          bind(c, name="afgpu_regression_values")
       import
       type(c_ptr), value :: h
       integer(c_int), value :: n_vars
       integer(c_int), intent(in) :: ivs(*)
       real(c_double), intent(out) :: out(*)
     end function afgpu_regression_values

     integer(c_int) function afgpu_regression_values_noalloc(h, n_vars, ivs, out) &
          bind(c, name="afgpu_regression_values_noalloc")
       import
       type(c_ptr), value :: h
       integer(c_int), value :: n_vars
       integer(c_int), intent(in) :: ivs(*)
       real(c_double), intent(out) :: out(*)
     end function afgpu_regression_values_noalloc

     integer(c_int) function afgpu_format_regression_line(out_cnt, n_vals, vals, buf, buf_len) &
          bind(c, name="afgpu_format_regression_line")
       import
       integer(c_int), value :: out_cnt, n_vals, buf_len
       real(c_double), intent(in) :: vals(*)
       character(kind=c_char), intent(out) :: buf(*)
     end function afgpu_format_regression_line

     integer(c_int) function afgpu_write_regression_log(h, filename, out_cnt, n_vars, ivs, names) &
          bind(c, name="afgpu_write_regression_log")
       import
       type(c_ptr), value :: h
       character(kind=c_char), intent(in) :: filename(*)
       integer(c_int), value :: out_cnt, n_vars
       integer(c_int), intent(in) :: ivs(*)
       type(c_ptr), intent(in) :: names(*)   ! c_loc of null-terminated names
     end function afgpu_write_regression_log

     integer(c_int) function afgpu_get_boxes(h, n_vars, ivs, n_ids, ids, data) &
          bind(c, name="afgpu_get_boxes")
       import
       type(c_ptr), value :: h
       integer(c_int), value :: n_vars, n_ids
       integer(c_int), intent(in) :: ivs(*), ids(*)
       real(c_double), intent(out) :: data(*)
     end function afgpu_get_boxes

     integer(c_int) function afgpu_put_boxes(h, n_vars, ivs, n_ids, ids, data) &
          bind(c, name="afgpu_put_boxes")
       import
       type(c_ptr), value :: h
       integer(c_int), value :: n_vars, n_ids
       integer(c_int), intent(in) :: ivs(*), ids(*)
       real(c_double), intent(in) :: data(*)
     end function afgpu_put_boxes

     ! ---- measurement ------------------------------------------------------------------
     integer(c_int) function afgpu_kernel_timing(h, enable) bind(c, name="afgpu_kernel_timing")
       import
       type(c_ptr), value :: h
       integer(c_int), value :: enable
     end function afgpu_kernel_timing

     integer(c_int) function afgpu_kernel_stats(h, kind, reset, total_ms, launches, algo_bytes) &
          bind(c, name="afgpu_kernel_stats")
       import
       type(c_ptr), value :: h
       integer(c_int), value :: kind, reset
       real(c_double), intent(out) :: total_ms, algo_bytes
       integer(c_long_long), intent(out) :: launches
     end function afgpu_kernel_stats

     ! ---- ensembles of runs -------------------------------------------------------------
     integer(c_int) function afgpu_num_runs(h) bind(c, name="afgpu_num_runs")
       import
       type(c_ptr), value :: h
     end function afgpu_num_runs

     integer(c_int) function afgpu_set_run_voltages(h, n_runs, voltages) &
          bind(c, name="afgpu_set_run_voltages")
       import
       type(c_ptr), value :: h
       integer(c_int), value :: n_runs
       real(c_double), intent(in) :: voltages(*)
     end function afgpu_set_run_voltages

     integer(c_int) function afgpu_field_compute_runs(h, s_in, have_guess, n_runs, n_cycles, residual) &
          bind(c, name="afgpu_field_compute_runs")
       import
       type(c_ptr), value :: h
       integer(c_int), value :: s_in, have_guess, n_runs
       integer(c_int), intent(out) :: n_cycles(*)
       real(c_double), intent(out) :: residual(*)
     end function afgpu_field_compute_runs

     integer(c_int) function afgpu_step_ensemble(h, n_runs, info) bind(c, name="afgpu_step_ensemble")
       import
       type(c_ptr), value :: h
       integer(c_int), value :: n_runs
       type(afgpu_step_info_t), intent(out) :: info(*)
     end function afgpu_step_ensemble

     integer(c_int) function afgpu_get_run_time_state(h, run, it, time, global_dt) &
          bind(c, name="afgpu_get_run_time_state")
       import
       type(c_ptr), value :: h
       integer(c_int), value :: run
       integer(c_int), intent(out) :: it
       real(c_double), intent(out) :: time, global_dt
     end function afgpu_get_run_time_state

     integer(c_int) function afgpu_set_run_time_state(h, run, it, time, global_dt) &
          bind(c, name="afgpu_set_run_time_state")
       import
       type(c_ptr), value :: h
       integer(c_int), value :: run, it
       real(c_double), value :: time, global_dt
     end function afgpu_set_run_time_state

     ! ---- work decomposition --------------------------------------------------------------
     integer(c_int) function afgpu_autotune(h, n_classes, shapes) bind(c, name="afgpu_autotune")
       import
       type(c_ptr), value :: h
       integer(c_int), value :: n_classes
       type(afgpu_shape_t), intent(out) :: shapes(*)
     end function afgpu_autotune

     integer(c_int) function afgpu_get_shapes(h, n_classes, shapes) bind(c, name="afgpu_get_shapes")
       import
       type(c_ptr), value :: h
       integer(c_int), value :: n_classes
       type(afgpu_shape_t), intent(out) :: shapes(*)
     end function afgpu_get_shapes

     integer(c_int) function afgpu_set_shapes(h, n_classes, shapes) bind(c, name="afgpu_set_shapes")
       import
       type(c_ptr), value :: h
       integer(c_int), value :: n_classes
       type(afgpu_shape_t), intent(in) :: shapes(*)
     end function afgpu_set_shapes

     ! ---- chemistry ----------------------------------------------------------------------
     integer(c_int) function afgpu_chem_create(n_species, n_reactions, reactions, table_n, &
          table_x_min, table_x_max, n_fld_cols, fld_cols, energy_col, gas_temperature, c) &
          bind(c, name="afgpu_chem_create")
       import
       integer(c_int), value :: n_species, n_reactions, table_n, n_fld_cols
       type(afgpu_reaction_t), intent(in) :: reactions(*)
       real(c_double), value :: table_x_min, table_x_max, gas_temperature
       real(c_double), intent(in) :: fld_cols(*)
       type(c_ptr), value :: energy_col   ! c_null_ptr or c_loc of the column
       type(c_ptr), intent(out) :: c
     end function afgpu_chem_create

     integer(c_int) function afgpu_chem_destroy(c) bind(c, name="afgpu_chem_destroy")
       import
       type(c_ptr), value :: c
     end function afgpu_chem_destroy

     ! the array arguments below are device addresses (afgpu_device_alloc)
     integer(c_int) function afgpu_chem_rates(c, n_cells, fields, rates) bind(c, name="afgpu_chem_rates")
       import
       type(c_ptr), value :: c, fields, rates
       integer(c_int), value :: n_cells
     end function afgpu_chem_rates

     integer(c_int) function afgpu_chem_derivatives(c, n_cells, dens, rates, derivs) &
          bind(c, name="afgpu_chem_derivatives")
       import
       type(c_ptr), value :: c, dens, rates, derivs
       integer(c_int), value :: n_cells
     end function afgpu_chem_derivatives

     integer(c_int) function afgpu_chem_source(c, n_cells, fields, dens, derivs) &
          bind(c, name="afgpu_chem_source")
       import
       type(c_ptr), value :: c, fields, dens, derivs
       integer(c_int), value :: n_cells
     end function afgpu_chem_source

     integer(c_int) function afgpu_set_chemistry(h, chem, charges) bind(c, name="afgpu_set_chemistry")
       import
       type(c_ptr), value :: h, chem
       integer(c_int), intent(in) :: charges(*)
     end function afgpu_set_chemistry

     integer(c_int) function afgpu_get_species(h, k, state, n_ids, ids, data) &
          bind(c, name="afgpu_get_species")
       import
       type(c_ptr), value :: h
       integer(c_int), value :: k, state, n_ids
       integer(c_int), intent(in) :: ids(*)
       real(c_double), intent(out) :: data(*)
     end function afgpu_get_species

     integer(c_int) function afgpu_put_species(h, k, state, n_ids, ids, data) &
          bind(c, name="afgpu_put_species")
       import
       type(c_ptr), value :: h
       integer(c_int), value :: k, state, n_ids
       integer(c_int), intent(in) :: ids(*)
       real(c_double), intent(in) :: data(*)
     end function afgpu_put_species

     integer(c_int) function afgpu_device_alloc(bytes, ptr) bind(c, name="afgpu_device_alloc")
       import
       integer(c_size_t), value :: bytes
       type(c_ptr), intent(out) :: ptr
     end function afgpu_device_alloc

     integer(c_int) function afgpu_device_free(ptr) bind(c, name="afgpu_device_free")
       import
       type(c_ptr), value :: ptr
     end function afgpu_device_free

     integer(c_int) function afgpu_device_copy(dst, src, bytes, to_device) bind(c, name="afgpu_device_copy")
       import
       type(c_ptr), value :: dst, src
       integer(c_size_t), value :: bytes
       integer(c_int), value :: to_device
     end function afgpu_device_copy

     integer(c_size_t) function c_strlen(s) bind(c, name="strlen")
       import
       type(c_ptr), value :: s
     end function c_strlen
  end interface

  public :: afgpu_create, afgpu_destroy, afgpu_last_error, afgpu_sync_tree, &
       afgpu_put_cc, afgpu_get_cc, afgpu_put_fc, afgpu_get_fc, afgpu_set_params, &
       afgpu_set_tables, afgpu_set_voltage, afgpu_set_option, afgpu_mg_fas_vcycle, &
       afgpu_mg_fas_fmg, afgpu_tree_maxabs_cc, afgpu_gc_tree, afgpu_restrict_tree, &
       afgpu_field_compute, afgpu_field_from_potential, afgpu_forward_euler, &
       afgpu_advance, afgpu_photoi_set_src, afgpu_copy_current_state, &
       afgpu_restore_previous_state, afgpu_step, afgpu_get_time_state, &
       afgpu_set_time_state, afgpu_limit_step_dt, afgpu_num_leaf_cells, afgpu_coarse_factors, &
       afgpu_refine_flags, afgpu_restrict_boxes, afgpu_prolong_boxes, &
       afgpu_regression_values, afgpu_regression_values_noalloc, &
       afgpu_format_regression_line, afgpu_write_regression_log, afgpu_get_boxes, &
       afgpu_put_boxes, afgpu_kernel_timing, afgpu_kernel_stats, afgpu_num_runs, &
       afgpu_set_run_voltages, afgpu_field_compute_runs, afgpu_step_ensemble, &
       afgpu_get_run_time_state, afgpu_set_run_time_state, afgpu_autotune, &
       afgpu_get_shapes, afgpu_set_shapes, afgpu_chem_create, afgpu_chem_destroy, &
       afgpu_chem_rates, afgpu_chem_derivatives, afgpu_chem_source, &
       afgpu_set_chemistry, afgpu_get_species, afgpu_put_species, &
       afgpu_device_alloc, afgpu_device_free, afgpu_device_copy

contains

  !> The library's last error message as a Fortran string
  function afgpu_error_message() result(msg)
    character(len=:), allocatable :: msg
    type(c_ptr)                   :: p
    character(kind=c_char), pointer :: chars(:)
    integer(c_size_t)             :: n, i
    p = afgpu_last_error()
    if (.not. c_associated(p)) then
       msg = ""
       return
    end if
    n = c_strlen(p)
    call c_f_pointer(p, chars, [n])
    allocate(character(len=n) :: msg)
    do i = 1, n
       msg(i:i) = chars(i)
    end do
  end function afgpu_error_message

end module m_afgpu
