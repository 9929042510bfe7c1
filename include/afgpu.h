/*
 * afgpu.h -- C-ABI of the MI355X (gfx950) implementation of afivo-streamer's
 * per-time-step hot path: the afivo FAS multigrid V-cycle/FMG for the
 * electrostatic potential and the Helmholtz photoionization modes, the
 * Koren-limited drift-diffusion flux with transport-table lookup, and the
 * explicit density/source update.
 *
 * Plain C types only (no torch, no HIP types), so that a Fortran
 * `bind(C)` interface block, ctypes or any other FFI can bind it.
 *
 * Conventions, identical to afivo's box tree (afivo/src/m_af_types.f90):
 *   - box ids are 1-based, 0 (af_no_box) means "no box", a negative
 *     neighbour id (af_phys_boundary) means a physical boundary;
 *   - neighbour numbering lowx=1, highx=2, lowy=3, highy=4, lowz=5, highz=6
 *     (m_af_types.f90:164-172, 196-207); child numbering follows
 *     af_child_dix (m_af_types.f90:148, 181-183);
 *   - cell-centred box data is (nc+2)^D doubles with one ghost layer,
 *     column-major with i fastest, exactly box%cc(DTIMES(0:nc+1), iv)
 *     (m_af_types.f90:302, m_af_core.f90:549-551);
 *   - face data for one variable is D blocks of (nc+1)^D doubles,
 *     exactly box%fc(DTIMES(1:nc+1), 1:NDIM, ifc) (m_af_core.f90:551).
 *
 * Every entry returns an int status: AFGPU_OK on success, otherwise one of
 * the AFGPU_ERR_* codes; afgpu_last_error() returns the message.  The
 * Fortran wrapper turns a non-zero status into `error stop`, as the
 * reference does after its hypre calls (m_coarse_solver.f90:187-189).
 *
 * Threading: calls are made from serial driver code (as the reference
 * calls field_compute / af_advance outside OpenMP regions,
 * src/streamer.f90:263,327); a handle is not re-entrant.
 */
#ifndef AFGPU_H
#define AFGPU_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct afgpu_ctx afgpu_t;

/* ---- status codes ---------------------------------------------------- */
#define AFGPU_OK                 0
#define AFGPU_ERR_ARG            1 /* invalid argument / size            */
#define AFGPU_ERR_HIP            2 /* HIP runtime failure                */
#define AFGPU_ERR_STATE          3 /* call out of order (no tree, ...)   */
#define AFGPU_ERR_NO_CONVERGENCE 4 /* "No convergence in initial field computation" */
#define AFGPU_ERR_ALL_REJECTED   5 /* "All time steps were rejected"     */
#define AFGPU_ERR_DT_TOO_SMALL   6 /* "dt too small"                     */
#define AFGPU_ERR_UNSUPPORTED    7 /* configuration outside the built path */

/* ---- coordinate systems (af_xyz, af_cyl; m_af_types.f90:48-51) ------- */
#define AFGPU_XYZ 1
#define AFGPU_CYL 2

/* ---- precision forms ------------------------------------------------- */
#define AFGPU_PREC_FP64  0 /* plain FP64, same arithmetic as the CPU path  */
#define AFGPU_PREC_MIXED 1 /* FP32 multigrid inside FP64 defect correction */

/* ---- time integrators (m_af_advance.f90:10-11) ----------------------- */
#define AFGPU_FORWARD_EULER 1
#define AFGPU_HEUNS_METHOD  2

/* ---- species boundary conditions (src/m_streamer.f90:324-336) -------- */
#define AFGPU_SPECIES_BC_NEUMANN_ZERO   0
#define AFGPU_SPECIES_BC_DIRICHLET_ZERO 1

/* ---- cell-centred variables of the device box pool --------------------
 * Densities keep afivo's "iv + state" numbering of copies
 * (af_add_cc_variable n_copies, src/m_chemistry.f90:285-287): state s of
 * species k is AFGPU_CC_SPECIES(k) + s.  With forward Euler, state 1 is the
 * backup used by copy_current_state; with Heun, state 1 is the predictor
 * and state 2 the backup (src/streamer.f90:659-688). */
#define AFGPU_CC_ELECTRON   0  /* e, e_2, e_3    */
#define AFGPU_CC_POS_ION    3  /* M+, ...        */
#define AFGPU_CC_NEG_ION    6  /* M-, ...        */
#define AFGPU_CC_PHI        9  /* phi            */
#define AFGPU_CC_PHI_2      10 /* phi backup     */
#define AFGPU_CC_EFIELD     11 /* electric_fld   */
#define AFGPU_CC_RHS        12 /* rhs            */
#define AFGPU_CC_TMP        13 /* tmp            */
#define AFGPU_CC_PHOTO      14 /* photo          */
#define AFGPU_CC_HELMH      15 /* helmh_1..helmh_3 (15, 16, 17) */
#define AFGPU_N_CC          18
#define AFGPU_CC_SPECIES(k) (3 * (k))

/* ---- face-centred variables ------------------------------------------ */
#define AFGPU_FC_FLUX  0 /* flux_elec                      */
#define AFGPU_FC_FIELD 1 /* electric_fld (face components) */
#define AFGPU_N_FC     2

/* ---- multigrid operators --------------------------------------------- */
#define AFGPU_OP_POISSON 0 /* mg for phi, bc field_bc_homogeneous         */
#define AFGPU_OP_HELMH   1 /* mg_helm(n) is AFGPU_OP_HELMH + n - 1         */
#define AFGPU_MAX_MODES  3

/* Scalar settings of the path: the values the reference reads from its
 * config (m_streamer.f90, m_dt.f90, m_photoi*.f90, m_gas.f90). */
typedef struct {
  double gas_number_density;   /* N = 1e5 p / (k_B T)  (m_gas.f90:174-176)   */
  int    time_integrator;      /* AFGPU_FORWARD_EULER or AFGPU_HEUNS_METHOD   */
  double dt_max;               /* m_dt.f90:39                                 */
  double dt_min;               /* m_dt.f90:42                                 */
  double dt_safety_factor;     /* m_dt.f90:27                                 */
  double dt_cfl_number;        /* m_dt.f90:30 (0.5 default)                   */
  double dt_max_growth_factor; /* m_dt.f90:45                                 */
  int    dt_chemistry_limit_loss; /* m_dt.f90:36                              */
  double dt_chemistry_nmin;    /* m_dt.f90:33 (<= 0: unused)                  */
  int    mg_num_vcycles;       /* ST_multigrid_num_vcycles (m_streamer.f90:151)*/
  double mg_max_rel_residual;  /* ST_multigrid_max_rel_residual               */
  int    mg_n_cycle_down;      /* mg_t%n_cycle_down (2)                       */
  int    mg_n_cycle_up;        /* mg_t%n_cycle_up (2)                         */
  double domain_len_last;      /* ST_domain_len(NDIM), for the residual floor */
  int    species_bc;           /* AFGPU_SPECIES_BC_*                          */
  int    photoi_enabled;
  int    photoi_per_steps;     /* photoi%per_steps                            */
  int    photoi_species;       /* species index that gains the source too
                                  (photoi%species): 0 = e, 1 = M+, 2 = M- of the
                                  standard model, or a reaction list's index;
                                  checked against the model at step time */
  double photoi_eta_quench;    /* photoi_eta * quench_fac (m_photoi.f90:175-184) */
  int    photoi_n_modes;       /* <= AFGPU_MAX_MODES                          */
  double photoi_lambda[AFGPU_MAX_MODES]; /* 1/m, after pressure scaling       */
  double photoi_coeff[AFGPU_MAX_MODES];  /* 1/m^2, after pressure scaling     */
  double photoi_max_rel_residual; /* photoi_helmh%max_rel_residual (1e-2)     */
  int    photoi_max_fmg_cycles;   /* max_fmg_cycles (10)                      */
} afgpu_params_t;

/* Result of one time step (afgpu_step). */
typedef struct {
  int    it;            /* iteration number after the step                */
  double time;          /* simulation time after the step                 */
  double dt_used;       /* accepted dt                                    */
  double dt_next;       /* global_dt for the next step                    */
  double dt_lim;        /* min over stages of the stage dt limits         */
  int    n_rejected;    /* rejected attempts in this step                 */
  int    n_vcycles;     /* V-cycles of the final field_compute            */
  double residual;      /* max |residual| of the final field_compute      */
  double rel_residual;  /* residual / max |rhs|                           */
  int    photoi_done;   /* photoionization recomputed in this step        */
  long long leaf_cells; /* leaf cells advanced (af_num_cells_used)        */
  int    photoi_cycles; /* FMG cycles of the Helmholtz modes in this step,
                           summed over the modes (0 without photoionization) */
} afgpu_step_info_t;

/* ---- lifetime ---------------------------------------------------------
 * Replaces tree/mg set-up: af_init (m_af_core.f90:138) and mg_init
 * (m_af_multigrid.f90:43).  ndim 2 or 3; n_cell even, 8 or 16;
 * coord AFGPU_XYZ or AFGPU_CYL (2D only). */
int afgpu_create(int ndim, int n_cell, int coord, int precision, afgpu_t **out);
int afgpu_destroy(afgpu_t *h);
const char *afgpu_last_error(void);

/* Mirror of the AMR tree after af_init / af_adjust_refinement / restart
 * (box_t fields, m_af_types.f90:286-324).  Arrays are indexed by id-1 and
 * hold n_boxes entries of: lvl (0 = unused slot), ix[D], parent,
 * children[2^D], neighbors[2D], neighbor_mat[3^D] (i fastest, as
 * box%neighbor_mat(DTIMES(-1:1))), r_min[D], dr[D].  Level lists
 * (lvl_t ids/leaves/parents) are derived.  Box data is kept for box ids
 * whose slot survives; new slots start at zero. */
int afgpu_sync_tree(afgpu_t *h, int n_boxes, const int *lvl, const int *ix,
                    const int *parent, const int *children,
                    const int *neighbors, const int *neighbor_mat,
                    const double *r_min, const double *dr);

/* Box data transfer: n_ids boxes, data[n][(nc+2)^D] (cc) or
 * data[n][D][(nc+1)^D] (fc).  Replaces reading/writing box%cc / box%fc. */
int afgpu_put_cc(afgpu_t *h, int iv, int n_ids, const int *ids, const double *data);
int afgpu_get_cc(afgpu_t *h, int iv, int n_ids, const int *ids, double *data);
int afgpu_put_fc(afgpu_t *h, int ifc, int n_ids, const int *ids, const double *data);
int afgpu_get_fc(afgpu_t *h, int ifc, int n_ids, const int *ids, double *data);

/* Settings, transport tables and applied voltage.
 * td_cols: [4][n_rows] columns mobility*N, diffusion*N, alpha/N, eta/N of
 * td_tbl (src/m_transport_data.f90:85-126); chem_cols: [2][n_rows]
 * ionization and attachment rate columns of chemtbl_fld
 * (src/m_chemistry.f90:233-256, 345-372).  Lookup as LT_get_col
 * (src/lookup_table_fortran/m_lookup_table.f90:330-406). */
int afgpu_set_params(afgpu_t *h, const afgpu_params_t *p);
int afgpu_set_tables(afgpu_t *h, int n_rows,
                     double td_x_min, double td_inv_fac, const double *td_cols,
                     double chem_x_min, double chem_inv_fac, const double *chem_cols);
/* field_set_voltage result (src/m_field.f90:551-586): Dirichlet value at
 * the high boundary of the last dimension. */
int afgpu_set_voltage(afgpu_t *h, double voltage);

/* ---- multigrid (afivo/src/m_af_multigrid.f90) --------------------------
 * op: AFGPU_OP_POISSON or AFGPU_OP_HELMH + n.  Replaces
 * mg_fas_vcycle(tree, mg, set_residual, highest_lvl) (:185) and
 * mg_fas_fmg(tree, mg, set_residual, have_guess) (:137).  max_lvl <= 0
 * means tree%highest_lvl. */
int afgpu_mg_fas_vcycle(afgpu_t *h, int op, int set_residual, int max_lvl);
int afgpu_mg_fas_fmg(afgpu_t *h, int op, int set_residual, int have_guess);
/* af_tree_maxabs_cc (afivo/src/m_af_utils.f90:773): max |cc| over leaf
 * interior cells. */
int afgpu_tree_maxabs_cc(afgpu_t *h, int iv, double *out);
/* af_gc_tree (afivo/src/m_af_ghostcell.f90:24) and af_restrict_tree
 * (afivo/src/m_af_restrict.f90:50) for one variable, with the variable's
 * own methods; used before refinement on all_densities
 * (src/streamer.f90:399-400).  iv may also be AFGPU_VAR_SPECIES(k), a
 * species of the reaction list set with afgpu_set_chemistry (state 0),
 * with the densities' methods. */
int afgpu_gc_tree(afgpu_t *h, int iv);
int afgpu_restrict_tree(afgpu_t *h, int iv);

/* ---- field (src/m_field.f90) -------------------------------------------
 * field_compute(tree, mg, s_in, time, have_guess) (:448): rhs from the
 * charged species of state s_in, V-cycles (or FMG without guess) to the
 * residual threshold, then field_from_potential.  Outputs the number of
 * V-cycles (or FMG cycles) and the final max |residual|. */
int afgpu_field_compute(afgpu_t *h, int s_in, int have_guess,
                        int *n_cycles, double *residual);
/* field_from_potential (:531): E_fc = -grad phi, |E| and its ghost cells. */
int afgpu_field_from_potential(afgpu_t *h);

/* ---- fluid (src/m_fluid.f90, afivo/src/m_af_flux_schemes.f90) ----------
 * forward_euler (src/m_fluid.f90:21): flux_upwind_tree (Koren, table
 * lookup, af_consistent_fluxes) + flux_update_densities with
 * add_source_terms; returns dt_lim = min(dt_max, minval(dt_limits)). */
int afgpu_forward_euler(afgpu_t *h, double dt, int s_deriv, int n_prev,
                        const int *s_prev, const double *w_prev, int s_out,
                        int i_step, int n_steps, double *dt_lim);
/* af_advance(tree, dt, dt_lim, time, all_densities, integrator,
 * forward_euler) (afivo/src/m_af_advance.f90:121). */
int afgpu_advance(afgpu_t *h, double dt, double *dt_lim);

/* ---- refinement flags --------------------------------------------------
 * The box refinement flags that af_adjust_refinement
 * (afivo/src/m_af_core.f90:697-735) derives with the streamer's
 * default_refinement (src/m_refine.f90:198-298), from the device's |E| and
 * n_e, so that only the flags (one int per box) leave the device:
 * consistent_ref_flags (m_af_core.f90:929-1013) evaluates every leaf and
 * every box with a leaf child, cell_to_ref_flags (:1091-1145) turns the
 * cell flags into box flags and, within buffer_width cells of a box side,
 * into refinement of the neighbour there; then the level cap,
 * ensure_two_one_balance (:1016-1057) and handle_derefinement_flags
 * (:1059-1088).  On return ref_flags[id - 1] holds the reference's final
 * flag of box id: AFGPU_REF_REFINE and AFGPU_REF_DEREFINE are acted on,
 * the other values (-1, 0, 1) ignored, as there.  The densities of refined
 * boxes are read as they are: restrict n_e first (afgpu_restrict_tree), as
 * src/streamer.f90:399-400 does.  Boxes tagged for an electrode
 * (m_refine.f90:267-270) do not occur on this path. */
#define AFGPU_REF_DEREFINE (-2)
#define AFGPU_REF_REFINE   2
#define AFGPU_REF_MAX_ITEMS 8
typedef struct {
  double refine_adx;       /* refine if alpha * dx > refine_adx */
  double refine_adx_fac;   /* alpha(f E) / f */
  double derefine_dx;      /* derefine only below this spacing */
  double refine_min_dx, refine_max_dx, refine_min_dens;
  double refine_init_time, refine_init_fac;
  double global_time;
  int use_alpha_effective; /* alpha - eta */
  int buffer_width;        /* refine_buffer_width, 0 .. n_cell */
  int max_lvl;             /* af_max_lvl (30) */
  int n_seeds;             /* seeds of the initial conditions (init_conds) */
  double seed_r0[AFGPU_REF_MAX_ITEMS][3], seed_r1[AFGPU_REF_MAX_ITEMS][3];
  double seed_width[AFGPU_REF_MAX_ITEMS];
  int n_regions;           /* refine_regions_* */
  double regions_dr[AFGPU_REF_MAX_ITEMS], regions_tstop[AFGPU_REF_MAX_ITEMS];
  double regions_rmin[AFGPU_REF_MAX_ITEMS][3], regions_rmax[AFGPU_REF_MAX_ITEMS][3];
  int n_limits;            /* refine_limits_* */
  double limits_dr[AFGPU_REF_MAX_ITEMS];
  double limits_rmin[AFGPU_REF_MAX_ITEMS][3], limits_rmax[AFGPU_REF_MAX_ITEMS][3];
} afgpu_refine_t;
int afgpu_refine_flags(afgpu_t *h, const afgpu_refine_t *rp, int n_boxes, int *ref_flags);

/* The data of boxes removed or added by af_adjust_refinement, on the
 * device, for the variables with methods (photo with photoionization, |E|,
 * phi, the electron and ion densities, rhs; src/streamer.f90:84-108):
 * - afgpu_restrict_boxes: auto_restrict (afivo/src/m_af_core.f90:825-840),
 *   the children of each listed box restricted into it (af_restrict_box,
 *   m_af_restrict.f90:62-136), before the driver removes the children;
 * - afgpu_prolong_boxes: auto_prolong (m_af_core.f90:843-881) for the
 *   listed new boxes after afgpu_sync_tree with the new tree: level by
 *   level, each from its parent (af_prolong_limit, m_af_prolong.f90:311-420,
 *   for the densities and rhs; af_prolong_linear, :531-679, for the
 *   others), then their ghost cells (af_gc_box with corners).
 * Existing boxes keep their data across afgpu_sync_tree, so the new boxes'
 * values never pass through the host. */
int afgpu_restrict_boxes(afgpu_t *h, int n_ids, const int *ids);
int afgpu_prolong_boxes(afgpu_t *h, int n_ids, const int *ids);

/* Diagnostics from the device pool.
 * - afgpu_regression_values: the statistics of output_regression_log
 *   (src/m_output.f90:788-842) for n_vars cell-centred variables (at most
 *   AFGPU_DIAG_MAX_VARS), per run of the tree (one for a plain tree):
 *   out[(r * n_vars + v) * 3 + 0] = af_tree_sum_cc / af_total_volume,
 *   [.. + 1] the same with power 2, [.. + 2] af_tree_max_cc
 *   (afivo/src/m_af_utils.f90:758-769, 966-1026; m_af_types.f90:805-825).
 *   The sums are deterministic but in another order than the reference's
 *   (whose own order depends on its OpenMP thread count); the maxima are
 *   exact.
 * - afgpu_format_regression_line: host only (no GPU use); one line of the
 *   log in the reference's format "(I0, n E20.8e3)" without the newline;
 *   returns the length, or a negative status if buf is too short.
 * - afgpu_write_regression_log: output_regression_log itself on a plain
 *   tree: at out_cnt == 0 the file is (re)created with the header
 *   "it time dt sum(name) ... sum(name^2) ... max(name) ...", then one line
 *   (out_cnt, time, dt and the values) is appended, time and dt being the
 *   handle's time state. */
#define AFGPU_DIAG_MAX_VARS 8
/* In the variable lists of afgpu_regression_values, afgpu_write_regression_log
 * and afgpu_get_boxes / afgpu_put_boxes: species k of a reaction list
 * (state 0), including the species beyond the three standard pools. */
#define AFGPU_VAR_SPECIES(k) (100 + (k))
int afgpu_regression_values(afgpu_t *h, int n_vars, const int *ivs, double *out);
/* afgpu_regression_values_noalloc: the same statistics as
 * afgpu_regression_values, bit for bit, from a call that allocates no
 * device memory.  It has no counterpart in the reference: it is for
 * callers that cannot allocate (code under HIP graph capture, tight output
 * loops, harnesses that forbid hipMalloc).  Its scratch (the per-leaf-box
 * terms and the per-run results, (n_leaf_boxes + n_runs) * n_vars * 3
 * doubles) lives in a static device array of AFGPU_NOALLOC_SCRATCH_WORDS
 * doubles (32 MiB, one copy per GPU, held until the process ends); larger
 * requests return AFGPU_ERR_ARG.  All calls on one GPU share that array:
 * calls that overlap on one GPU (several threads, several handles at once)
 * race. */
#define AFGPU_NOALLOC_SCRATCH_WORDS (1 << 22)
int afgpu_regression_values_noalloc(afgpu_t *h, int n_vars, const int *ivs, double *out);
int afgpu_format_regression_line(int out_cnt, int n_vals, const double *vals, char *buf, int buf_len);
int afgpu_write_regression_log(afgpu_t *h, const char *filename, int out_cnt, int n_vars, const int *ivs,
                               const char *const *names);

/* Box data for checkpoints and restarts, in the order af_write_tree writes
 * it (afivo/src/m_af_output.f90:90-114: box by box, and within a box the
 * variables in turn, each cell variable as box%cc(DTIMES(:), iv), each
 * face variable as box%fc(DTIMES(:), :, ifc)); af_read_tree reads the same
 * order (m_af_output.f90:269-295).  A variable is a cell variable index,
 * AFGPU_VAR_SPECIES(k), or AFGPU_VAR_FACE(ifc) for the face variable ifc
 * (AFGPU_FC_FLUX, AFGPU_FC_FIELD); the reference writes a box's cell
 * variables before its face variables, so a checkpoint lists them in that
 * order.  data holds, per box, (n_cell + 2)^D doubles per cell variable
 * and D (n_cell + 1)^D per face variable, the boxes in the order of ids;
 * at most AFGPU_DIAG_MAX_VARS variables per call.  One gather (or scatter)
 * on the device and one copy over the host link. */
#define AFGPU_VAR_FACE(ifc) (200 + (ifc))
int afgpu_get_boxes(afgpu_t *h, int n_vars, const int *ivs, int n_ids, const int *ids, double *data);
int afgpu_put_boxes(afgpu_t *h, int n_vars, const int *ivs, int n_ids, const int *ids, const double *data);

/* photoi_set_src (src/m_photoi.f90:153) with the Helmholtz method
 * (src/m_photoi_helmh.f90:162). */
int afgpu_photoi_set_src(afgpu_t *h);

/* copy_current_state / restore_previous_state (src/streamer.f90:659-688). */
int afgpu_copy_current_state(afgpu_t *h);
int afgpu_restore_previous_state(afgpu_t *h);

/* One iteration of the time loop of src/streamer.f90:181-365 without
 * output and refinement: photoionization every photoi_per_steps, step
 * attempts with rejection/restore, field_compute, next dt.  The first call
 * takes dt = dt_min (src/streamer.f90:459). */
int afgpu_step(afgpu_t *h, afgpu_step_info_t *info);

/* Host-only helper (no GPU use): factors of the exact level-1 solve used in
 * place of hypre PFMG (afivo/src/m_coarse_solver.f90:421).  Level 1 must be
 * one box with physical boundaries on all faces; r_min/dr are that box's,
 * bc_dirichlet[2D] is 1 for a Dirichlet face and 0 for a Neumann face.
 * Outputs per dimension d: vinv/v (nc x nc, row major) and eig (nc), with
 * T_d = V_d diag(eig_d) V_d^-1 the 1-D operator including boundary
 * folding. */
int afgpu_coarse_factors(int ndim, int n_cell, int coord, const double *r_min,
                         const double *dr, const int *bc_dirichlet,
                         double *vinv, double *v, double *eig);

/* Time-loop scalars (it, time, global_dt), as written and read for a
 * restart by write_sim_data / read_sim_data (src/streamer.f90:541-577). */
int afgpu_get_time_state(afgpu_t *h, int *it, double *time, double *global_dt);
int afgpu_set_time_state(afgpu_t *h, int it, double time, double global_dt);

/* Leaf-cell count of the current tree (af_num_cells_used). */
/* The output-time clipping of the time loop (src/streamer.f90:222-229):
 * the next step's dt becomes min(dt, dt_cap), while global_dt, from which
 * the following step's dt grows, stays as it is.  dt_cap >= 0; a step that
 * has not started yet starts at dt_min, as the first step does. */
int afgpu_limit_step_dt(afgpu_t *h, double dt_cap);
long long afgpu_num_leaf_cells(afgpu_t *h);

/* Kernel-level timing: when enabled, every launch of kernel class `kind`
 * (0 = GSRB half-sweep, 1 = flux, 2 = density update) is bracketed by HIP
 * events on the library's stream.  Reports the summed duration (ms), the
 * launch count and the algorithmic bytes (interior words moved) of those
 * launches since the last reset. */
int afgpu_kernel_timing(afgpu_t *h, int enable);
int afgpu_kernel_stats(afgpu_t *h, int kind, int reset, double *total_ms,
                       long long *launches, double *algo_bytes);

/* Execution options. */
#define AFGPU_OPT_GRAPHS 1 /* capture V-cycle / FMG launch sequences as HIP graphs (default 1) */
#define AFGPU_OPT_DIRECT_N 2 /* mixed precision: largest uniformly refined level (cells per axis,
                                16..128) that the inner cycle solves exactly; below 16 the
                                inner cycle descends to level 1 (default 128) */
#define AFGPU_OPT_MIXED_SMOOTH 3 /* mixed precision: GSRB sweeps before and after the coarse
                                    correction in the inner cycle; 0 = the params' counts
                                    (default 0) */
#define AFGPU_OPT_BLOCK_SMOOTH 4 /* mixed precision: inner-cycle smoothing phases as one
                                    launch per level, each box swept in LDS with its
                                    same-level neighbours' values held (default 0) */
#define AFGPU_OPT_BATCH_MODES 5 /* mixed precision: the photoionization modes' cycles run as
                                   parallel branches of one captured graph (default 1) */
#define AFGPU_OPT_DIRECT_MFMA 6 /* mixed precision: uniform-level transforms (64 or 128 cells
                                   per axis) on the FP64 matrix cores (1) or as vector FMA
                                   tiles (0) (default 1) */
#define AFGPU_OPT_PHOTOI_VCYCLE 7 /* mixed precision: the photoionization modes' inner cycle is
                                     one V-cycle (1) or one FMG cycle (0) per outer iteration
                                     (default 0) */
#define AFGPU_OPT_FLUX32 8 /* mixed precision: the drift-diffusion flux stencil in FP32, the
                              density update accumulating in FP64 (default 1) */
#define AFGPU_OPT_STEP_GRAPH 9 /* ensembles: each step's first attempt and its field solve as one
                                  captured graph with the stopping tests on the device (1), or
                                  one captured graph per cycle with the tests on the host (0);
                                  the results are identical (default 0, measured faster) */
#define AFGPU_OPT_FUSE_DIRECT 10 /* mixed precision: the uniform-level solve's last forward
                                    transform, division and first backward transform as one pass
                                    (1) or two (0); identical results (default 0: the one pass,
                                    at one 100 KB-LDS workgroup per CU, measured slower) */
#define AFGPU_OPT_DIRECT_F32 11 /* mixed precision: the uniform-level solve of the FP32 inner
                                   cycle on the matrix cores (64 or 128 cells per axis) in FP32
                                   (1) or FP64 (0); the FP64 outer correction sets the accuracy
                                   either way (default 0: the coarsest-grid solve stays FP64;
                                   0 whenever AFGPU_OPT_FUSE_DIRECT is 1) */
#define AFGPU_OPT_BOX_LDS 12 /* mixed precision, 3D boxes of 16^3 cells and 2D boxes of 8^2:
                               the inner cycle's half-sweeps stage the whole box in LDS (1),
                               or only what they read (the other colour's block as it is
                               stored, the rhs, the face inputs), by DMA (0); identical
                               results (default 0) */
#define AFGPU_OPT_DIRECT_FOLD 13 /* mixed precision: the FP64 transforms (128 cells per axis) of
                                   axes whose 1-D operator is symmetric under reversal (the same
                                   boundary condition at both ends, Cartesian) split the modes by
                                   parity: half the products; results differ in rounding only
                                   (default 1; AFGPU_OPT_FUSE_DIRECT applies with it off) */
int afgpu_set_option(afgpu_t *h, int option, int value);

/* ---- ensembles of independent runs --------------------------------------
 * The reference runs a parameter study as separate processes, one per run
 * (programs/parameter_study_2d/run_study.sh:1-5), each executing the time
 * loop of src/streamer.f90:181-432.  Here a forest synchronised with
 * afgpu_sync_tree holds many runs in one handle: every level-1 box roots a
 * run of its own (run r = the r-th level-1 box in id order; the level-1
 * boxes must share their geometry, their faces all physical boundaries).
 * Every run keeps the reference's semantics: its own applied voltage,
 * time step, rejections and stopping tests, so each run computes what it
 * would compute alone.  With photoionization each run solves the modes to
 * its own stopping test (photoi_max_rel_residual of its own max |rhs|) at
 * its own step count (photoi_per_steps); afgpu_photoi_set_src covers every
 * run.
 */
int afgpu_num_runs(afgpu_t *h);
/* Applied voltage of every run (afgpu_set_voltage sets all runs). */
int afgpu_set_run_voltages(afgpu_t *h, int n_runs, const double *voltages);
/* field_compute for every run; per-run V-cycle (or FMG) counts and final
 * max |residual| (arrays of n_runs). */
int afgpu_field_compute_runs(afgpu_t *h, int s_in, int have_guess, int n_runs,
                             int *n_cycles, double *residual);
/* One time step of every run (afgpu_step per run); info has n_runs
 * entries.  afgpu_step itself takes single runs only. */
int afgpu_step_ensemble(afgpu_t *h, int n_runs, afgpu_step_info_t *info);
/* Time-loop scalars of run `run` as afgpu_step_ensemble advances them (a
 * restart of an ensemble, write_sim_data / read_sim_data per run,
 * src/streamer.f90:541-577). */
int afgpu_get_run_time_state(afgpu_t *h, int run, int *it, double *time, double *global_dt);
int afgpu_set_run_time_state(afgpu_t *h, int run, int it, double time, double global_dt);

/* ---- work decomposition and its autotuner --------------------------------
 * The list kernels fall into AFGPU_N_KERNEL_CLASSES classes (0 half-sweeps,
 * 1 residuals, 2 restriction, 3 prolongation and coarse rhs, 4 ghost layers,
 * 5 box operations and reductions, 6 drift-diffusion flux, 7 update, field
 * and sources).  Each class has a shape: how many boxes a workgroup takes at
 * a time (1, 2, 4 or 8; above 1 only for boxes of 8 x 8 cells, one
 * wavefront each) and how many workgroups per CU the persistent grid has
 * (0: one workgroup per batch of boxes).  Shapes change no result. */
#define AFGPU_N_KERNEL_CLASSES 8
typedef struct {
  int    boxes_per_workgroup;
  int    workgroups_per_cu;
  double ms;  /* autotune: the class's kernel time in one step with this shape */
} afgpu_shape_t;
/* Measure every candidate shape on the handle's own workload (one eager
 * time step per candidate from a snapshot of the current state, restored
 * after each, so the simulation is unchanged), keep the fastest shape of
 * every class, and report them (n_classes entries; may be 0). */
int afgpu_autotune(afgpu_t *h, int n_classes, afgpu_shape_t *shapes);
int afgpu_get_shapes(afgpu_t *h, int n_classes, afgpu_shape_t *shapes);
int afgpu_set_shapes(afgpu_t *h, int n_classes, const afgpu_shape_t *shapes);

/* ---- General chemistry networks (src/m_chemistry.f90) ----------------------
 * get_rates (m_chemistry.f90:672-768) for every rate type the reference
 * parses except the electron-energy table, and get_derivatives (:770-804)
 * for an arbitrary reaction list, on arrays resident in device memory
 * (the pointers of afgpu_chem_rates / _derivatives / _source are device
 * addresses).  Arrays are column-major like the reference's: cell index
 * fastest, fields[n_cells] in Td, dens[n_species][n_cells],
 * rates[n_reactions][n_cells], derivs[n_species][n_cells].
 * A reaction is the compiled form read_reactions / parse_reaction
 * (m_chemistry.f90:857-1142, 1156-1278) build: gas species and "M" folded
 * into rate_factor, the input species repeated by multiplicity, the output
 * species with their multiplicities; indices are 0-based plasma species. */
#define AFGPU_CHEM_MAX_SPECIES 100 /* max_num_species, m_chemistry.f90:129 */
#define AFGPU_CHEM_MAX_REACTIONS 500 /* max_num_reactions, m_chemistry.f90:132 */
#define AFGPU_CHEM_MAX_IN      8
#define AFGPU_CHEM_MAX_OUT     8
#define AFGPU_CHEM_MAX_COEFF   9   /* rate_max_num_coeff */
/* rate types, with the reference's codes (m_chemistry.f90:66-126) */
#define AFGPU_RATE_TABULATED_ENERGY 0  /* not supported: needs the energy equation */
#define AFGPU_RATE_TABULATED_FIELD  1
#define AFGPU_RATE_CONSTANT         2
#define AFGPU_RATE_LINEAR           3
#define AFGPU_RATE_EXP_V1           4
#define AFGPU_RATE_EXP_V2           5
#define AFGPU_RATE_K1               6
#define AFGPU_RATE_K3               8
#define AFGPU_RATE_K4               9
#define AFGPU_RATE_K5               10
#define AFGPU_RATE_K6               11
#define AFGPU_RATE_K7               12
#define AFGPU_RATE_K8               13
#define AFGPU_RATE_K9               14
#define AFGPU_RATE_K10              15
#define AFGPU_RATE_K11              16
#define AFGPU_RATE_K12              17
#define AFGPU_RATE_K13              18
#define AFGPU_RATE_K14              19
#define AFGPU_RATE_K15              20
#define AFGPU_RATE_K16              21
typedef struct {
  int    rate_type;
  int    lookup_col;                        /* column of the field table (0-based) */
  int    n_in, n_out;
  double rate_factor;                       /* c0 */
  double rate_data[AFGPU_CHEM_MAX_COEFF];   /* c(1:n_coeff) */
  int    ix_in[AFGPU_CHEM_MAX_IN];          /* repeated by multiplicity */
  int    ix_out[AFGPU_CHEM_MAX_OUT];
  int    multiplicity_out[AFGPU_CHEM_MAX_OUT];
} afgpu_reaction_t;
typedef struct afgpu_chem afgpu_chem_t;
/* table_n rows on the uniform grid of LT_create(x_min, x_max, table_n)
 * (chemtbl_fld and td_tbl share it, m_chemistry.f90:344-345): fld_cols holds n_fld_cols columns of
 * table_n values; energy_col (td_tbl's mean energy in eV, for the types
 * that use Te: K1, K3, K16) may be NULL when no reaction needs it;
 * gas_temperature is m_gas's. */
int afgpu_chem_create(int n_species, int n_reactions, const afgpu_reaction_t *reactions, int table_n,
                      double table_x_min, double table_x_max, int n_fld_cols, const double *fld_cols,
                      const double *energy_col, double gas_temperature, afgpu_chem_t **out);
int afgpu_chem_destroy(afgpu_chem_t *c);
/* get_rates: the rate coefficients of every reaction in every cell */
int afgpu_chem_rates(afgpu_chem_t *c, int n_cells, const double *fields, double *rates);
/* get_derivatives: rates become the reaction rates (times the input
 * densities), derivs the species' time derivatives */
int afgpu_chem_derivatives(afgpu_chem_t *c, int n_cells, const double *dens, double *rates, double *derivs);
/* both in one pass, without the rates array: the same derivs */
int afgpu_chem_source(afgpu_chem_t *c, int n_cells, const double *fields, const double *dens, double *derivs);
/* The time step with a reaction list in place of the standard model
 * (add_source_terms, src/m_fluid.f90:298-466, for the list given to
 * chemistry_initialize, src/m_chemistry.f90:196-404).  Species 0 are the
 * electrons (charge -1, the transported species); species 1 and 2 live in
 * the pools of the positive and negative ions (AFGPU_CC_SPECIES(1), (2)),
 * further species (at most 16 in all) in pools of their own; charges[k] is
 * species k's charge (NULL with three species: -1, +1, -1).  The rhs sums
 * the charged species in species order (field_set_rhs); photoi_species
 * indexes the list; afgpu_restrict_boxes / afgpu_prolong_boxes cover every
 * species (the densities' methods); in a forest of runs every run takes
 * the list.  NULL returns to the standard model.  The chemistry handle must
 * stay alive while the step uses it. */
int afgpu_set_chemistry(afgpu_t *h, const afgpu_chem_t *chem, const int *charges);
/* A species' cells in a state (box by box, as afgpu_get_cc / _put_cc). */
int afgpu_get_species(afgpu_t *h, int k, int state, int n_ids, const int *ids, double *data);
int afgpu_put_species(afgpu_t *h, int k, int state, int n_ids, const int *ids, const double *data);
/* Device memory for such arrays, from the library's own HIP runtime, for
 * callers without one: allocation, release, and a synchronous copy
 * (to_device 1: host to device; 0: device to host). */
int afgpu_device_alloc(size_t bytes, void **ptr);
int afgpu_device_free(void *ptr);
int afgpu_device_copy(void *dst, const void *src, size_t bytes, int to_device);

#ifdef __cplusplus
}
#endif

#endif /* AFGPU_H */
