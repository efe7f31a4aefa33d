!> Fortran binding of the a5g C ABI (include/a5g.h) for the streamer drivers.
!>
!> The raw interfaces (c_*) bind every entry point of the C ABI with
!> iso_c_binding. The public wrappers keep Afivo's and the drivers' names and
!> argument meanings (afivo/src/m_aX_core.f90, m_aX_mg.f90, m_aX_gc.f90,
!> m_aX_utils.f90, src/streamer_2d.f90) so a driver keeps its control flow:
!>   - a2_*/a3_* and mg2_*/mg3_* are the same procedures (the tree handle
!>     carries the dimension);
!>   - callbacks become small integers: ghost-cell rules (a5_gc_*), the
!>     multigrid face rules (a5_bc_*) and refinement flags by id;
!>   - the drivers' box procedures (fld_from_pot, fluxes_koren,
!>     update_solution, average_dens and compute_fld's rhs) are st_* calls.
!> A non-zero status from the library stops the program with its message,
!> which is the reference's fail-fast behaviour (stop "msg").
module m_a5g
  use iso_c_binding
  implicit none
  private

  integer, parameter :: dp = kind(0.0d0)

  ! ghost-cell rule at refinement boundaries (a5g_gc_tree: rb)
  integer, parameter, public :: a5_gc_sides_rb   = 0 ! mg$D_sides_rb
  integer, parameter, public :: a5_gc_interp     = 1 ! a$D_gc_interp
  integer, parameter, public :: a5_gc_interp_lim = 2 ! a$D_gc_interp_lim
  ! ghost-cell rule at physical boundaries (a5g_gc_tree: bc)
  integer, parameter, public :: a5_gc_mg_bc      = 0 ! the multigrid face rules
  integer, parameter, public :: a5_gc_neumann    = 1 ! a$D_gc_neumann
  integer, parameter, public :: a5_gc_dirichlet  = 2 ! a$D_gc_dirichlet
  ! multigrid face rules (the drivers' sides_bc_pot)
  integer, parameter, public :: a5_bc_copy  = 0 ! ghost = interior (Neumann)
  integer, parameter, public :: a5_bc_neg   = 1 ! ghost = -interior (grounded)
  integer, parameter, public :: a5_bc_value = 2 ! ghost = 2 V - interior
  ! ghost = 2 V - interior over the whole ghost row of a 2D y face, corners
  ! included (the cylindrical driver's top face, src/streamer_cyl.f90:408-410)
  integer, parameter, public :: a5_bc_row_value = 3
  ! refinement flags (afivo/src/m_afivo_types.f90)
  integer, parameter, public :: a5_rm_ref = -1, a5_kp_ref = 0, a5_do_ref = 1
  ! coordinate systems
  integer, parameter, public :: a5_xyz = 0, a5_cyl = 1

  !> A tree whose cell data lives on the GPU
  type, public :: a5g_t
     type(c_ptr) :: h      = c_null_ptr
     integer     :: dim    = 0
     integer     :: n_cell = 0
  end type a5g_t

  !> Multigrid options (mg$D_t, afivo/src/m_aX_mg.f90:20-52); the sides_bc
  !> procedure becomes one rule and value per physical face
  type, public :: a5g_mg_t
     integer  :: i_phi        = -1
     integer  :: i_tmp        = -1
     integer  :: i_rhs        = -1
     integer  :: i_eps        = -1
     integer  :: n_cycle_down = 2
     integer  :: n_cycle_up   = 2
     integer  :: n_cycle_base = 2
     integer  :: bc_kind(6)   = a5_bc_copy
     real(dp) :: bc_val(6)    = 0.0_dp
     logical  :: auto_ops     = .true.
  end type a5g_mg_t

  !> ref_info_t (afivo/src/m_aX_core.f90:540-601)
  type, public :: a5g_lvl_info_t
     integer, allocatable :: add(:)
     integer, allocatable :: rm(:)
  end type a5g_lvl_info_t

  type, public :: a5g_ref_info_t
     integer                           :: n_add = 0
     integer                           :: n_rm  = 0
     type(a5g_lvl_info_t), allocatable :: lvls(:)
  end type a5g_ref_info_t

  !> a$D_loc_t (afivo/src/m_aX_t.f90): a leaf box and a cell of it; ix(1:dim)
  !> is used
  type, public :: a5g_loc_t
     integer :: id    = -1
     integer :: ix(3) = -1
  end type a5g_loc_t

  ! per-box rules of st_reduction_loc (get_streamer_properties,
  ! src/streamer_cyl.f90:831-903)
  integer, parameter, public :: a5_rl_face_mean = 1 ! box_maxfld_r / box_maxfld_z
  integer, parameter, public :: a5_rl_coord_max = 2 ! box_fld_maxr / box_fld_maxz

  ! ------------------------------------------------------------------
  ! raw interfaces, one per entry point of include/a5g.h
  ! ------------------------------------------------------------------
  interface
     function c_last_error() bind(C, name="a5g_last_error") result(p)
       import :: c_ptr
       type(c_ptr) :: p
     end function c_last_error

     function c_init(t, dim, n_cell, n_cc, n_fc, dr, r_min, coarsen_to, &
          lvls_max, coord, n_boxes) bind(C, name="a5g_init") result(s)
       import :: c_ptr, c_int, c_double
       type(c_ptr)           :: t
       integer(c_int), value :: dim, n_cell, n_cc, n_fc, coarsen_to, lvls_max
       integer(c_int), value :: coord, n_boxes
       real(c_double), value :: dr
       real(c_double)        :: r_min(*)
       integer(c_int)        :: s
     end function c_init

     function c_destroy(t) bind(C, name="a5g_destroy") result(s)
       import :: c_ptr, c_int
       type(c_ptr), value :: t
       integer(c_int)     :: s
     end function c_destroy

     function c_set_base(t, n, ix_list, nb_list) bind(C, name="a5g_set_base") result(s)
       import :: c_ptr, c_int
       type(c_ptr), value    :: t
       integer(c_int), value :: n
       integer(c_int)        :: ix_list(*), nb_list(*)
       integer(c_int)        :: s
     end function c_set_base

     function c_adjust_refinement(t, n, flags, n_add, n_rm) &
          bind(C, name="a5g_adjust_refinement") result(s)
       import :: c_ptr, c_int
       type(c_ptr), value    :: t
       integer(c_int), value :: n
       integer(c_int)        :: flags(*), n_add, n_rm
       integer(c_int)        :: s
     end function c_adjust_refinement

     function c_ref_info_lvl(t, lvl, add, rm, n_add, n_rm) &
          bind(C, name="a5g_ref_info_lvl") result(s)
       import :: c_ptr, c_int
       type(c_ptr), value    :: t
       integer(c_int), value :: lvl
       type(c_ptr), value    :: add, rm
       integer(c_int)        :: n_add, n_rm
       integer(c_int)        :: s
     end function c_ref_info_lvl

     function c_tidy_up(t, max_frac, goal_frac, n_clean_min, reorder) &
          bind(C, name="a5g_tidy_up") result(s)
       import :: c_ptr, c_int, c_double
       type(c_ptr), value    :: t
       real(c_double), value :: max_frac, goal_frac
       integer(c_int), value :: n_clean_min, reorder
       integer(c_int)        :: s
     end function c_tidy_up

     function c_info(t, out8) bind(C, name="a5g_info") result(s)
       import :: c_ptr, c_int
       type(c_ptr), value :: t
       integer(c_int)     :: out8(8)
       integer(c_int)     :: s
     end function c_info

     function c_get_boxes(t, n, lvl, ix, parent, children, neighbors, in_use, &
          n_cell, dr, r_min) bind(C, name="a5g_get_boxes") result(s)
       import :: c_ptr, c_int, c_double
       type(c_ptr), value    :: t
       integer(c_int), value :: n
       integer(c_int)        :: lvl(*), ix(*), parent(*), children(*)
       integer(c_int)        :: neighbors(*), in_use(*), n_cell(*)
       real(c_double)        :: dr(*), r_min(*)
       integer(c_int)        :: s
     end function c_get_boxes

     function c_level_counts(t, lvl, counts3) bind(C, name="a5g_level_counts") result(s)
       import :: c_ptr, c_int
       type(c_ptr), value    :: t
       integer(c_int), value :: lvl
       integer(c_int)        :: counts3(3)
       integer(c_int)        :: s
     end function c_level_counts

     function c_level_lists(t, lvl, ids, leaves, parents) &
          bind(C, name="a5g_level_lists") result(s)
       import :: c_ptr, c_int
       type(c_ptr), value    :: t
       integer(c_int), value :: lvl
       integer(c_int)        :: ids(*), leaves(*), parents(*)
       integer(c_int)        :: s
     end function c_level_lists

     function c_set_cc(t, id, iv, data) bind(C, name="a5g_set_cc") result(s)
       import :: c_ptr, c_int, c_double
       type(c_ptr), value    :: t
       integer(c_int), value :: id, iv
       real(c_double)        :: data(*)
       integer(c_int)        :: s
     end function c_set_cc

     function c_get_cc(t, id, iv, data) bind(C, name="a5g_get_cc") result(s)
       import :: c_ptr, c_int, c_double
       type(c_ptr), value    :: t
       integer(c_int), value :: id, iv
       real(c_double)        :: data(*)
       integer(c_int)        :: s
     end function c_get_cc

     function c_set_cc_all(t, iv, data) bind(C, name="a5g_set_cc_all") result(s)
       import :: c_ptr, c_int, c_double
       type(c_ptr), value    :: t
       integer(c_int), value :: iv
       real(c_double)        :: data(*)
       integer(c_int)        :: s
     end function c_set_cc_all

     function c_get_cc_all(t, iv, data) bind(C, name="a5g_get_cc_all") result(s)
       import :: c_ptr, c_int, c_double
       type(c_ptr), value    :: t
       integer(c_int), value :: iv
       real(c_double)        :: data(*)
       integer(c_int)        :: s
     end function c_get_cc_all

     function c_set_fc(t, id, dim, ifc, data) bind(C, name="a5g_set_fc") result(s)
       import :: c_ptr, c_int, c_double
       type(c_ptr), value    :: t
       integer(c_int), value :: id, dim, ifc
       real(c_double)        :: data(*)
       integer(c_int)        :: s
     end function c_set_fc

     function c_get_fc(t, id, dim, ifc, data) bind(C, name="a5g_get_fc") result(s)
       import :: c_ptr, c_int, c_double
       type(c_ptr), value    :: t
       integer(c_int), value :: id, dim, ifc
       real(c_double)        :: data(*)
       integer(c_int)        :: s
     end function c_get_fc

     function c_sync(t) bind(C, name="a5g_sync") result(s)
       import :: c_ptr, c_int
       type(c_ptr), value :: t
       integer(c_int)     :: s
     end function c_sync

     function c_mg_setup(t, i_phi, i_tmp, i_rhs, i_eps, n_down, n_up, n_base, &
          bc_kind, bc_val, auto_ops) bind(C, name="a5g_mg_setup") result(s)
       import :: c_ptr, c_int, c_double
       type(c_ptr), value    :: t
       integer(c_int), value :: i_phi, i_tmp, i_rhs, i_eps, n_down, n_up, n_base
       integer(c_int)        :: bc_kind(*)
       real(c_double)        :: bc_val(*)
       integer(c_int), value :: auto_ops
       integer(c_int)        :: s
     end function c_mg_setup

     function c_mg_set_bc_value(t, nb, value) bind(C, name="a5g_mg_set_bc_value") result(s)
       import :: c_ptr, c_int, c_double
       type(c_ptr), value    :: t
       integer(c_int), value :: nb
       real(c_double), value :: value
       integer(c_int)        :: s
     end function c_mg_set_bc_value

     function c_mg_set_helmholtz(t, kappa2) bind(C, name="a5g_mg_set_helmholtz") result(s)
       import :: c_ptr, c_int, c_double
       type(c_ptr), value    :: t
       real(c_double), value :: kappa2
       integer(c_int)        :: s
     end function c_mg_set_helmholtz

     function c_fas_fmg(t, set_residual, first_call) bind(C, name="a5g_fas_fmg") result(s)
       import :: c_ptr, c_int
       type(c_ptr), value    :: t
       integer(c_int), value :: set_residual, first_call
       integer(c_int)        :: s
     end function c_fas_fmg

     function c_fas_vcycle(t, max_lvl, set_residual) bind(C, name="a5g_fas_vcycle") result(s)
       import :: c_ptr, c_int
       type(c_ptr), value    :: t
       integer(c_int), value :: max_lvl, set_residual
       integer(c_int)        :: s
     end function c_fas_vcycle

     function c_fas_fmg_adjoint(t, adj, vbar) bind(C, name="a5g_fas_fmg_adjoint") result(s)
       import :: c_ptr, c_int, c_double
       type(c_ptr), value    :: t
       integer(c_int)        :: adj(*)
       real(c_double)        :: vbar(*)
       integer(c_int)        :: s
     end function c_fas_fmg_adjoint

     function c_fas_vcycle_adjoint(t, max_lvl, adj, vbar) &
          bind(C, name="a5g_fas_vcycle_adjoint") result(s)
       import :: c_ptr, c_int, c_double
       type(c_ptr), value    :: t
       integer(c_int), value :: max_lvl
       integer(c_int)        :: adj(*)
       real(c_double)        :: vbar(*)
       integer(c_int)        :: s
     end function c_fas_vcycle_adjoint

     function c_set_shape(t, kernel, nc, cls, shape) bind(C, name="a5g_set_shape") result(s)
       import :: c_ptr, c_int
       type(c_ptr), value    :: t
       integer(c_int), value :: kernel, nc, cls, shape
       integer(c_int)        :: s
     end function c_set_shape

     function c_get_shape(t, kernel, nc, cls, shape) bind(C, name="a5g_get_shape") result(s)
       import :: c_ptr, c_int
       type(c_ptr), value    :: t
       integer(c_int), value :: kernel, nc, cls
       integer(c_int)        :: shape
       integer(c_int)        :: s
     end function c_get_shape

     function c_autotune(t, reps, max_rec, rec, n_rec) bind(C, name="a5g_autotune") result(s)
       import :: c_ptr, c_int, c_double
       type(c_ptr), value    :: t
       integer(c_int), value :: reps, max_rec
       real(c_double)        :: rec(*)
       integer(c_int)        :: n_rec
       integer(c_int)        :: s
     end function c_autotune

     function c_mg_solve_mixed(t, tol, max_outer, n_inner, zero_start, n_outer, &
          rel_res) bind(C, name="a5g_mg_solve_mixed") result(s)
       import :: c_ptr, c_int, c_double
       type(c_ptr), value    :: t
       real(c_double), value :: tol
       integer(c_int), value :: max_outer, n_inner, zero_start
       integer(c_int)        :: n_outer
       real(c_double)        :: rel_res
       integer(c_int)        :: s
     end function c_mg_solve_mixed

     function c_mg_residual_norms(t, max_res, max_rhs) &
          bind(C, name="a5g_mg_residual_norms") result(s)
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: t
       real(c_double)     :: max_res, max_rhs
       integer(c_int)     :: s
     end function c_mg_residual_norms

     function c_gc_tree(t, iv, rb, bc) bind(C, name="a5g_gc_tree") result(s)
       import :: c_ptr, c_int
       type(c_ptr), value    :: t
       integer(c_int), value :: iv, rb, bc
       integer(c_int)        :: s
     end function c_gc_tree

     function c_restrict_tree(t, iv) bind(C, name="a5g_restrict_tree") result(s)
       import :: c_ptr, c_int
       type(c_ptr), value    :: t
       integer(c_int), value :: iv
       integer(c_int)        :: s
     end function c_restrict_tree

     function c_tree_copy_cc(t, iv_from, iv_to) bind(C, name="a5g_tree_copy_cc") result(s)
       import :: c_ptr, c_int
       type(c_ptr), value    :: t
       integer(c_int), value :: iv_from, iv_to
       integer(c_int)        :: s
     end function c_tree_copy_cc

     function c_tree_minmax_cc(t, iv, vmin, vmax) bind(C, name="a5g_tree_minmax_cc") result(s)
       import :: c_ptr, c_int, c_double
       type(c_ptr), value    :: t
       integer(c_int), value :: iv
       real(c_double)        :: vmin, vmax
       integer(c_int)        :: s
     end function c_tree_minmax_cc

     function c_table_create(t, x_min, x_max, n_rows, n_cols) &
          bind(C, name="a5g_table_create") result(s)
       import :: c_ptr, c_int, c_double
       type(c_ptr), value    :: t
       real(c_double), value :: x_min, x_max
       integer(c_int), value :: n_rows, n_cols
       integer(c_int)        :: s
     end function c_table_create

     function c_table_set_col(t, col, n, xx, yy) bind(C, name="a5g_table_set_col") result(s)
       import :: c_ptr, c_int, c_double
       type(c_ptr), value    :: t
       integer(c_int), value :: col, n
       real(c_double)        :: xx(*), yy(*)
       integer(c_int)        :: s
     end function c_table_set_col

     function c_table_get(t, rows_cols, x_min, inv_dx) bind(C, name="a5g_table_get") result(s)
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: t
       real(c_double)     :: rows_cols(*), x_min, inv_dx
       integer(c_int)     :: s
     end function c_table_get

     function c_set_table(t, n_rows, n_cols, x_min, inv_dx, rows_cols) &
          bind(C, name="a5g_set_table") result(s)
       import :: c_ptr, c_int, c_double
       type(c_ptr), value    :: t
       integer(c_int), value :: n_rows, n_cols
       real(c_double), value :: x_min, inv_dx
       real(c_double)        :: rows_cols(*)
       integer(c_int)        :: s
     end function c_set_table

     function c_streamer_setup(t, vars) bind(C, name="a5g_streamer_setup") result(s)
       import :: c_ptr, c_int
       type(c_ptr), value :: t
       integer(c_int)     :: vars(16)
       integer(c_int)     :: s
     end function c_streamer_setup

     function c_space_charge(t, fac) bind(C, name="a5g_space_charge") result(s)
       import :: c_ptr, c_int, c_double
       type(c_ptr), value    :: t
       real(c_double), value :: fac
       integer(c_int)        :: s
     end function c_space_charge

     function c_fld_from_pot(t) bind(C, name="a5g_fld_from_pot") result(s)
       import :: c_ptr, c_int
       type(c_ptr), value :: t
       integer(c_int)     :: s
     end function c_fld_from_pot

     function c_fluxes_koren(t, dt) bind(C, name="a5g_fluxes_koren") result(s)
       import :: c_ptr, c_int, c_double
       type(c_ptr), value    :: t
       real(c_double), value :: dt
       integer(c_int)        :: s
     end function c_fluxes_koren

     function c_consistent_fluxes(t, n, f_ixs) bind(C, name="a5g_consistent_fluxes") result(s)
       import :: c_ptr, c_int
       type(c_ptr), value    :: t
       integer(c_int), value :: n
       integer(c_int)        :: f_ixs(*)
       integer(c_int)        :: s
     end function c_consistent_fluxes

     function c_update_solution(t, dt) bind(C, name="a5g_update_solution") result(s)
       import :: c_ptr, c_int, c_double
       type(c_ptr), value    :: t
       real(c_double), value :: dt
       integer(c_int)        :: s
     end function c_update_solution

     function c_average_dens(t) bind(C, name="a5g_average_dens") result(s)
       import :: c_ptr, c_int
       type(c_ptr), value :: t
       integer(c_int)     :: s
     end function c_average_dens

     function c_photoi_helmholtz(t, coeff, n_terms, amp, kappa, i_phi, i_rhs, i_tmp, tol, &
          max_cycles, cycles) bind(C, name="a5g_photoi_helmholtz") result(s)
       import :: c_ptr, c_int, c_double
       type(c_ptr), value    :: t
       real(c_double), value :: coeff, tol
       integer(c_int), value :: n_terms, i_phi, i_rhs, i_tmp, max_cycles
       real(c_double)        :: amp(*), kappa(*)
       integer(c_int)        :: cycles
       integer(c_int)        :: s
     end function c_photoi_helmholtz

     function c_photoi_helmholtz_terms(t, coeff, n_terms, amp, kappa, i_terms, i_rhs, i_tmp, &
          warm, tol, max_cycles, cycles) bind(C, name="a5g_photoi_helmholtz_terms") result(s)
       import :: c_ptr, c_int, c_double
       type(c_ptr), value    :: t
       real(c_double), value :: coeff, tol
       integer(c_int), value :: n_terms, warm, max_cycles
       real(c_double)        :: amp(*), kappa(*)
       integer(c_int)        :: i_terms(*), i_rhs(*), i_tmp(*)
       integer(c_int)        :: cycles
       integer(c_int)        :: s
     end function c_photoi_helmholtz_terms

     function c_box_summaries(t, out) bind(C, name="a5g_box_summaries") result(s)
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: t
       real(c_double)     :: out(*)
       integer(c_int)     :: s
     end function c_box_summaries

     function c_set_ref_rules(t, n_seed, seeds, n_reg, regions, init_window, crv_refine, &
          n_base_last) bind(C, name="a5g_set_ref_rules") result(s)
       import :: c_ptr, c_int, c_double
       type(c_ptr), value    :: t
       integer(c_int), value :: n_seed, n_reg, n_base_last
       real(c_double)        :: seeds(*), regions(*)
       real(c_double), value :: init_window, crv_refine
       integer(c_int)        :: s
     end function c_set_ref_rules

     function c_ref_flags(t, times, flags) bind(C, name="a5g_ref_flags") result(s)
       import :: c_ptr, c_int, c_double
       type(c_ptr), value    :: t
       real(c_double)        :: times(*)
       integer(c_int)        :: flags(*)
       integer(c_int)        :: s
     end function c_ref_flags

     function c_instance_maxima(t, iv, vmax, lvl_max) bind(C, name="a5g_instance_maxima") &
          result(s)
       import :: c_ptr, c_int, c_double
       type(c_ptr), value    :: t
       integer(c_int), value :: iv
       real(c_double)        :: vmax(*)
       integer(c_int)        :: lvl_max(*)
       integer(c_int)        :: s
     end function c_instance_maxima

     function c_update_solution_adjoint(t, dt, adj) &
          bind(C, name="a5g_update_solution_adjoint") result(s)
       import :: c_ptr, c_int, c_double
       type(c_ptr), value    :: t
       real(c_double), value :: dt
       integer(c_int)        :: adj(6)
       integer(c_int)        :: s
     end function c_update_solution_adjoint

     function c_consistent_fluxes_adjoint(t, lf) &
          bind(C, name="a5g_consistent_fluxes_adjoint") result(s)
       import :: c_ptr, c_int
       type(c_ptr), value    :: t
       integer(c_int), value :: lf
       integer(c_int)        :: s
     end function c_consistent_fluxes_adjoint

     function c_fluxes_koren_adjoint(t, dt, adj) &
          bind(C, name="a5g_fluxes_koren_adjoint") result(s)
       import :: c_ptr, c_int, c_double
       type(c_ptr), value    :: t
       real(c_double), value :: dt
       integer(c_int)        :: adj(6)
       integer(c_int)        :: s
     end function c_fluxes_koren_adjoint

     function c_table_adjoint(t, clear, rows_cols) bind(C, name="a5g_table_adjoint") result(s)
       import :: c_ptr, c_int
       type(c_ptr), value    :: t
       integer(c_int), value :: clear
       type(c_ptr), value    :: rows_cols
       integer(c_int)        :: s
     end function c_table_adjoint

     function c_tree_copy_cc_adjoint(t, lfrom, lto) &
          bind(C, name="a5g_tree_copy_cc_adjoint") result(s)
       import :: c_ptr, c_int
       type(c_ptr), value    :: t
       integer(c_int), value :: lfrom, lto
       integer(c_int)        :: s
     end function c_tree_copy_cc_adjoint

     function c_average_dens_adjoint(t, adj) bind(C, name="a5g_average_dens_adjoint") result(s)
       import :: c_ptr, c_int
       type(c_ptr), value :: t
       integer(c_int)     :: adj(4)
       integer(c_int)     :: s
     end function c_average_dens_adjoint

     function c_space_charge_adjoint(t, fac, adj) &
          bind(C, name="a5g_space_charge_adjoint") result(s)
       import :: c_ptr, c_int, c_double
       type(c_ptr), value    :: t
       real(c_double), value :: fac
       integer(c_int)        :: adj(3)
       integer(c_int)        :: s
     end function c_space_charge_adjoint

     function c_restrict_tree_adjoint(t, il) bind(C, name="a5g_restrict_tree_adjoint") result(s)
       import :: c_ptr, c_int
       type(c_ptr), value    :: t
       integer(c_int), value :: il
       integer(c_int)        :: s
     end function c_restrict_tree_adjoint

     function c_gc_tree_adjoint(t, iv, il, rb, bc) bind(C, name="a5g_gc_tree_adjoint") result(s)
       import :: c_ptr, c_int
       type(c_ptr), value    :: t
       integer(c_int), value :: iv, il, rb, bc
       integer(c_int)        :: s
     end function c_gc_tree_adjoint

     function c_fld_from_pot_adjoint(t, adj) bind(C, name="a5g_fld_from_pot_adjoint") result(s)
       import :: c_ptr, c_int
       type(c_ptr), value :: t
       integer(c_int)     :: adj(3)
       integer(c_int)     :: s
     end function c_fld_from_pot_adjoint

     function c_state_store(t, slot, n_cc, ivs, n_fc, fvs, op) &
          bind(C, name="a5g_state_store") result(s)
       import :: c_ptr, c_int
       type(c_ptr), value    :: t
       integer(c_int), value :: slot, n_cc, n_fc, op
       integer(c_int)        :: ivs(*), fvs(*)
       integer(c_int)        :: s
     end function c_state_store

     function c_box_summaries_noalloc(t, out) bind(C, name="a5g_box_summaries_noalloc") &
          result(s)
       import :: c_ptr, c_int, c_double
       type(c_ptr), value :: t
       real(c_double)     :: out(*)
       integer(c_int)     :: s
     end function c_box_summaries_noalloc

     function c_prolong1_to(t, n, ids, iv) bind(C, name="a5g_prolong1_to") result(s)
       import :: c_ptr, c_int
       type(c_ptr), value    :: t
       integer(c_int), value :: n, iv
       integer(c_int)        :: ids(*)
       integer(c_int)        :: s
     end function c_prolong1_to

     function c_fill_cc(t, n, ids, iv, value) bind(C, name="a5g_fill_cc") result(s)
       import :: c_ptr, c_int, c_double
       type(c_ptr), value    :: t
       integer(c_int), value :: n, iv
       integer(c_int)        :: ids(*)
       real(c_double), value :: value
       integer(c_int)        :: s
     end function c_fill_cc

     function c_gc_boxes(t, n, ids, iv, rb, bc) bind(C, name="a5g_gc_boxes") result(s)
       import :: c_ptr, c_int
       type(c_ptr), value    :: t
       integer(c_int), value :: n, iv, rb, bc
       integer(c_int)        :: ids(*)
       integer(c_int)        :: s
     end function c_gc_boxes

     function c_tree_maxdiff_cc(t, iv_a, iv_b, maxdiff, maxabs_b) &
          bind(C, name="a5g_tree_maxdiff_cc") result(s)
       import :: c_ptr, c_int, c_double
       type(c_ptr), value    :: t
       integer(c_int), value :: iv_a, iv_b
       real(c_double)        :: maxdiff, maxabs_b
       integer(c_int)        :: s
     end function c_tree_maxdiff_cc

     function c_set_gaussians(t, iv, n_blob, blobs, base) &
          bind(C, name="a5g_set_gaussians") result(s)
       import :: c_ptr, c_int, c_double
       type(c_ptr), value    :: t
       integer(c_int), value :: iv, n_blob
       real(c_double)        :: blobs(*)
       real(c_double), value :: base
       integer(c_int)        :: s
     end function c_set_gaussians

     function c_set_gaussians_instances(t, iv, n_blob, blobs, base) &
          bind(C, name="a5g_set_gaussians_instances") result(s)
       import :: c_ptr, c_int, c_double
       type(c_ptr), value    :: t
       integer(c_int), value :: iv, n_blob
       real(c_double)        :: blobs(*), base(*)
       integer(c_int)        :: s
     end function c_set_gaussians_instances

     function c_get_loc(t, r, max_lvl, id, ix) bind(C, name="a5g_get_loc") result(s)
       import :: c_ptr, c_int, c_double
       type(c_ptr), value    :: t
       real(c_double)        :: r(*)
       integer(c_int), value :: max_lvl
       integer(c_int)        :: id, ix(*)
       integer(c_int)        :: s
     end function c_get_loc

     function c_reduction_loc(t, kind, var, d, thr, init_val, val, id, ix) &
          bind(C, name="a5g_reduction_loc") result(s)
       import :: c_ptr, c_int, c_double
       type(c_ptr), value    :: t
       integer(c_int), value :: kind, var, d
       real(c_double), value :: thr, init_val
       real(c_double)        :: val
       integer(c_int)        :: id, ix(*)
       integer(c_int)        :: s
     end function c_reduction_loc

     function c_cc_axis(t, n_cc, ixs_cc, n_fc, ixs_fc, cap, out, n_points) &
          bind(C, name="a5g_cc_axis") result(s)
       import :: c_ptr, c_int, c_double
       type(c_ptr), value    :: t
       integer(c_int), value :: n_cc, n_fc, cap
       integer(c_int)        :: ixs_cc(*), ixs_fc(*)
       real(c_double)        :: out(*)
       integer(c_int)        :: n_points
       integer(c_int)        :: s
     end function c_cc_axis

     function c_leaf_cells(t, n_cc, ixs_cc, n_fc, ixs_fc, cap, out, n_cells) &
          bind(C, name="a5g_leaf_cells") result(s)
       import :: c_ptr, c_int, c_double, c_long_long
       type(c_ptr), value         :: t
       integer(c_int), value      :: n_cc, n_fc
       integer(c_int)             :: ixs_cc(*), ixs_fc(*)
       integer(c_long_long), value :: cap
       real(c_double)             :: out(*)
       integer(c_long_long)       :: n_cells
       integer(c_int)             :: s
     end function c_leaf_cells

     function c_write_vtk(t, filename, n_cc, cc_names, ixs_cc, n_fc, fc_names, ixs_fc, &
          n_cycle, time) bind(C, name="a5g_write_vtk") result(s)
       import :: c_ptr, c_int, c_double, c_char
       type(c_ptr), value     :: t
       character(kind=c_char) :: filename(*)
       integer(c_int), value  :: n_cc, n_fc, n_cycle
       type(c_ptr)            :: cc_names(*), fc_names(*)
       type(c_ptr), value     :: ixs_cc
       integer(c_int)         :: ixs_fc(*)
       real(c_double), value  :: time
       integer(c_int)         :: s
     end function c_write_vtk

     function c_dump_leaves(t, filename, n_var, ivs, cycle, time) &
          bind(C, name="a5g_dump_leaves") result(s)
       import :: c_ptr, c_int, c_double, c_char
       type(c_ptr), value     :: t
       character(kind=c_char) :: filename(*)
       integer(c_int), value  :: n_var, cycle
       integer(c_int)         :: ivs(*)
       real(c_double), value  :: time
       integer(c_int)         :: s
     end function c_dump_leaves

     function c_set_instances(t, n_inst, stride) bind(C, name="a5g_set_instances") result(s)
       import :: c_ptr, c_int, c_double
       type(c_ptr), value    :: t
       integer(c_int), value :: n_inst
       real(c_double), value :: stride
       integer(c_int)        :: s
     end function c_set_instances

     function c_set_instance_values(t, what, n, values) &
          bind(C, name="a5g_set_instance_values") result(s)
       import :: c_ptr, c_int, c_double
       type(c_ptr), value    :: t
       integer(c_int), value :: what, n
       real(c_double)        :: values(*)
       integer(c_int)        :: s
     end function c_set_instance_values

     function c_set_active_instances(t, n, active) &
          bind(C, name="a5g_set_active_instances") result(s)
       import :: c_ptr, c_int
       type(c_ptr), value    :: t
       integer(c_int), value :: n
       integer(c_int)        :: active(*)
       integer(c_int)        :: s
     end function c_set_active_instances

     function c_set_instance_groups(t, n, group) &
          bind(C, name="a5g_set_instance_groups") result(s)
       import :: c_ptr, c_int
       type(c_ptr), value    :: t
       integer(c_int), value :: n
       integer(c_int)        :: group(*)
       integer(c_int)        :: s
     end function c_set_instance_groups

     function c_fas_fmg_groups(t, set_residual, first_call) &
          bind(C, name="a5g_fas_fmg_groups") result(s)
       import :: c_ptr, c_int
       type(c_ptr), value    :: t
       integer(c_int), value :: set_residual, first_call
       integer(c_int)        :: s
     end function c_fas_fmg_groups

     function c_prof_enable(t, on) bind(C, name="a5g_prof_enable") result(s)
       import :: c_ptr, c_int
       type(c_ptr), value    :: t
       integer(c_int), value :: on
       integer(c_int)        :: s
     end function c_prof_enable

     function c_prof_read(t, kind, lvl, n_launch, ms, bytes) &
          bind(C, name="a5g_prof_read") result(s)
       import :: c_ptr, c_int, c_double
       type(c_ptr), value    :: t
       integer(c_int), value :: kind, lvl
       integer(c_int)        :: n_launch
       real(c_double)        :: ms, bytes
       integer(c_int)        :: s
     end function c_prof_read
  end interface

  ! Afivo's names for the 2D and 3D trees
  interface a2_destroy;           module procedure t_destroy;           end interface
  interface a3_destroy;           module procedure t_destroy;           end interface
  interface a2_set_base;          module procedure t_set_base;          end interface
  interface a3_set_base;          module procedure t_set_base;          end interface
  interface a2_adjust_refinement; module procedure t_adjust_refinement; end interface
  interface a3_adjust_refinement; module procedure t_adjust_refinement; end interface
  interface a2_tidy_up;           module procedure t_tidy_up;           end interface
  interface a3_tidy_up;           module procedure t_tidy_up;           end interface
  interface a2_gc_tree;           module procedure t_gc_tree;           end interface
  interface a3_gc_tree;           module procedure t_gc_tree;           end interface
  interface a2_gc_boxes;          module procedure t_gc_boxes;          end interface
  interface a3_gc_boxes;          module procedure t_gc_boxes;          end interface
  interface a2_restrict_tree;     module procedure t_restrict_tree;     end interface
  interface a3_restrict_tree;     module procedure t_restrict_tree;     end interface
  interface a2_tree_copy_cc;      module procedure t_tree_copy_cc;      end interface
  interface a3_tree_copy_cc;      module procedure t_tree_copy_cc;      end interface
  interface a2_tree_max_cc;       module procedure t_tree_max_cc;       end interface
  interface a3_tree_max_cc;       module procedure t_tree_max_cc;       end interface
  interface a2_tree_min_cc;       module procedure t_tree_min_cc;       end interface
  interface a3_tree_min_cc;       module procedure t_tree_min_cc;       end interface
  interface a2_consistent_fluxes; module procedure t_consistent_fluxes; end interface
  interface a3_consistent_fluxes; module procedure t_consistent_fluxes; end interface
  interface a2_prolong1_to;       module procedure t_prolong1_to;       end interface
  interface a3_prolong1_to;       module procedure t_prolong1_to;       end interface
  interface mg2_init_mg;          module procedure t_mg_init;           end interface
  interface mg3_init_mg;          module procedure t_mg_init;           end interface
  interface mg2_fas_fmg;          module procedure t_fas_fmg;           end interface
  interface mg3_fas_fmg;          module procedure t_fas_fmg;           end interface
  interface mg2_fas_vcycle;       module procedure t_fas_vcycle;        end interface
  interface mg3_fas_vcycle;       module procedure t_fas_vcycle;        end interface
  interface a2_get_loc;           module procedure t_get_loc;           end interface
  interface a3_get_loc;           module procedure t_get_loc;           end interface
  interface a2_write_vtk;         module procedure t_write_vtk;         end interface
  interface a3_write_vtk;         module procedure t_write_vtk;         end interface

  public :: a2_init, a3_init
  public :: a2_destroy, a3_destroy, a2_set_base, a3_set_base
  public :: a2_adjust_refinement, a3_adjust_refinement, a2_tidy_up, a3_tidy_up
  public :: a2_gc_tree, a3_gc_tree, a2_gc_boxes, a3_gc_boxes
  public :: a2_restrict_tree, a3_restrict_tree, a2_tree_copy_cc, a3_tree_copy_cc
  public :: a2_tree_max_cc, a3_tree_max_cc, a2_tree_min_cc, a3_tree_min_cc
  public :: a2_consistent_fluxes, a3_consistent_fluxes, a2_prolong1_to, a3_prolong1_to
  public :: mg2_init_mg, mg3_init_mg, mg2_fas_fmg, mg3_fas_fmg
  public :: mg2_fas_vcycle, mg3_fas_vcycle, mg_fas_fmg_adjoint, mg_fas_vcycle_adjoint
  public :: a5g_max_id, a5g_max_lvl, a5g_min_dr, a5g_level_leaves, a5g_box_info
  public :: a5g_set_cc, a5g_get_cc, a5g_fill_cc, a5g_sync
  public :: a5g_mixed_solve, a5g_residual_norms, a5g_set_bc_value
  public :: st_setup, st_table_create, st_table_set_col, st_table_get
  public :: st_space_charge, st_fld_from_pot, st_fluxes_koren, st_update_solution
  public :: st_average_dens, st_box_summaries, st_photoi_helmholtz, a5g_set_helmholtz
  public :: st_box_summaries_noalloc
  public :: st_transport_adjoint, st_table_adjoint
  public :: st_copy_cc_adjoint, st_average_dens_adjoint, st_space_charge_adjoint
  public :: st_restrict_tree_adjoint, st_gc_tree_adjoint, st_fld_from_pot_adjoint
  public :: a5g_state_store
  public :: st_photoi_helmholtz_terms
  public :: a2_get_loc, a3_get_loc, st_reduction_loc, st_get_cc_axis
  public :: a2_write_vtk, a3_write_vtk, a5g_dump_leaves
  public :: a5g_set_instances, a5g_set_instance_values, a5g_set_active_instances
  public :: a5g_set_instance_groups, mg_fas_fmg_groups

contains

  !> Stop with the library's message when a call failed
  subroutine check(s, what)
    integer(c_int), intent(in)   :: s
    character(len=*), intent(in) :: what
    character(kind=c_char), pointer :: msg(:)
    character(len=:), allocatable   :: text
    integer :: n

    if (s == 0) return
    call c_f_pointer(c_last_error(), msg, [4096])
    n = 0
    do while (n < 4096)
       if (msg(n+1) == c_null_char) exit
       n = n + 1
    end do
    allocate(character(len=n) :: text)
    text = transfer(msg(1:n), text)
    write(*, "(A)") "a5g error in " // what // ": " // text
    error stop 1
  end subroutine check

  integer(c_int) function lg(x)
    logical, intent(in) :: x
    lg = merge(1, 0, x)
  end function lg

  ! ------------------------------------------------------------------
  ! tree (afivo/src/m_aX_core.f90)
  ! ------------------------------------------------------------------

  !> a2_init (m_aX_core.f90:47-116) for a 2D tree on the GPU
  subroutine a2_init(tree, n_cell, n_var_cell, n_var_face, dr, coarsen_to, &
       n_boxes, coord, lvls_max, r_min)
    type(a5g_t), intent(inout)     :: tree
    integer, intent(in)            :: n_cell, n_var_cell, n_var_face
    real(dp), intent(in)           :: dr
    integer, intent(in), optional  :: coarsen_to, n_boxes, coord, lvls_max
    real(dp), intent(in), optional :: r_min(2)
    call t_init(tree, 2, n_cell, n_var_cell, n_var_face, dr, coarsen_to, &
         n_boxes, coord, lvls_max, r_min)
  end subroutine a2_init

  !> a3_init for a 3D tree on the GPU
  subroutine a3_init(tree, n_cell, n_var_cell, n_var_face, dr, coarsen_to, &
       n_boxes, lvls_max, r_min)
    type(a5g_t), intent(inout)     :: tree
    integer, intent(in)            :: n_cell, n_var_cell, n_var_face
    real(dp), intent(in)           :: dr
    integer, intent(in), optional  :: coarsen_to, n_boxes, lvls_max
    real(dp), intent(in), optional :: r_min(3)
    call t_init(tree, 3, n_cell, n_var_cell, n_var_face, dr, coarsen_to, &
         n_boxes, a5_xyz, lvls_max, r_min)
  end subroutine a3_init

  subroutine t_init(tree, dim, n_cell, n_var_cell, n_var_face, dr, coarsen_to, &
       n_boxes, coord, lvls_max, r_min)
    type(a5g_t), intent(inout)     :: tree
    integer, intent(in)            :: dim, n_cell, n_var_cell, n_var_face
    real(dp), intent(in)           :: dr
    integer, intent(in), optional  :: coarsen_to, n_boxes, coord, lvls_max
    real(dp), intent(in), optional :: r_min(:)
    integer  :: ct, nb, cs, lm
    real(dp) :: rm(3)

    ! defaults of a$D_init (m_aX_core.f90:47-66)
    ct = -1; if (present(coarsen_to)) ct = coarsen_to
    nb = 1000; if (present(n_boxes)) nb = n_boxes
    cs = a5_xyz; if (present(coord)) cs = coord
    lm = 30; if (present(lvls_max)) lm = lvls_max
    rm = 0.0_dp; if (present(r_min)) rm(1:dim) = r_min(1:dim)
    call check(c_init(tree%h, dim, n_cell, n_var_cell, n_var_face, dr, rm, ct, &
         lm, cs, nb), "a$D_init")
    tree%dim = dim
    tree%n_cell = n_cell
  end subroutine t_init

  subroutine t_destroy(tree)
    type(a5g_t), intent(inout) :: tree
    call check(c_destroy(tree%h), "a$D_destroy")
    tree%h = c_null_ptr
  end subroutine t_destroy

  !> a$D_set_base (m_aX_core.f90:137-220): ix_list(dim, n), nb_list(2 dim, n)
  subroutine t_set_base(tree, ix_list, nb_list)
    type(a5g_t), intent(inout) :: tree
    integer, intent(in)        :: ix_list(:, :), nb_list(:, :)
    call check(c_set_base(tree%h, size(ix_list, 2), ix_list, nb_list), "a$D_set_base")
  end subroutine t_set_base

  !> a$D_adjust_refinement (m_aX_core.f90:473-537). ref_flags(id) holds what
  !> the user's set_ref_flags would have set (a5_rm_ref / a5_kp_ref /
  !> a5_do_ref) for boxes 1..size(ref_flags).
  subroutine t_adjust_refinement(tree, ref_flags, ref_info)
    type(a5g_t), intent(inout)          :: tree
    integer, intent(in)                 :: ref_flags(:)
    type(a5g_ref_info_t), intent(inout) :: ref_info
    integer(c_int) :: n_add, n_rm, na, nr
    integer        :: lvl, max_lvl
    integer(c_int), allocatable, target :: add(:), rm(:)

    call check(c_adjust_refinement(tree%h, size(ref_flags), ref_flags, n_add, n_rm), &
         "a$D_adjust_refinement")
    ref_info%n_add = n_add
    ref_info%n_rm  = n_rm
    max_lvl = a5g_max_lvl(tree)
    if (allocated(ref_info%lvls)) deallocate(ref_info%lvls)
    allocate(ref_info%lvls(max_lvl + 1))
    do lvl = 1, max_lvl + 1
       call check(c_ref_info_lvl(tree%h, lvl, c_null_ptr, c_null_ptr, na, nr), "ref_info")
       allocate(add(max(na, 1)), rm(max(nr, 1)))
       call check(c_ref_info_lvl(tree%h, lvl, c_loc(add), c_loc(rm), na, nr), "ref_info")
       ref_info%lvls(lvl)%add = add(1:na)
       ref_info%lvls(lvl)%rm  = rm(1:nr)
       deallocate(add, rm)
    end do
  end subroutine t_adjust_refinement

  !> a$D_tidy_up (m_aX_core.f90:224-325)
  subroutine t_tidy_up(tree, max_frac, goal_frac, n_clean_min, reorder)
    type(a5g_t), intent(inout) :: tree
    real(dp), intent(in)       :: max_frac, goal_frac
    integer, intent(in)        :: n_clean_min
    logical, intent(in)        :: reorder
    call check(c_tidy_up(tree%h, max_frac, goal_frac, n_clean_min, lg(reorder)), &
         "a$D_tidy_up")
  end subroutine t_tidy_up

  integer function a5g_max_id(tree)
    type(a5g_t), intent(in) :: tree
    integer(c_int) :: out8(8)
    call check(c_info(tree%h, out8), "info")
    a5g_max_id = out8(1)
  end function a5g_max_id

  integer function a5g_max_lvl(tree)
    type(a5g_t), intent(in) :: tree
    integer(c_int) :: out8(8)
    call check(c_info(tree%h, out8), "info")
    a5g_max_lvl = out8(2)
  end function a5g_max_lvl

  !> a$D_min_dr (m_aX_utils.f90:196-200) from the finest box in use
  real(dp) function a5g_min_dr(tree)
    type(a5g_t), intent(in) :: tree
    integer, allocatable  :: lvl(:), in_use(:), n_cell(:)
    real(dp), allocatable :: dr(:), r_min(:, :)
    integer :: n
    call a5g_box_info(tree, lvl, in_use, n_cell, dr, r_min)
    n = size(lvl)
    a5g_min_dr = minval(dr, mask=(in_use /= 0))
  end function a5g_min_dr

  !> Level, use flag, cell count, dr and r_min of boxes 1..max_id
  subroutine a5g_box_info(tree, lvl, in_use, n_cell, dr, r_min)
    type(a5g_t), intent(in)               :: tree
    integer, allocatable, intent(out)     :: lvl(:), in_use(:), n_cell(:)
    real(dp), allocatable, intent(out)    :: dr(:), r_min(:, :)
    integer, allocatable :: ix(:, :), parent(:), children(:, :), nbs(:, :)
    integer :: n, d

    n = a5g_max_id(tree)
    d = tree%dim
    allocate(lvl(n), in_use(n), n_cell(n), dr(n), r_min(d, n))
    allocate(ix(d, n), parent(n), children(2**d, n), nbs(2*d, n))
    call check(c_get_boxes(tree%h, n, lvl, ix, parent, children, nbs, in_use, &
         n_cell, dr, r_min), "get_boxes")
  end subroutine a5g_box_info

  !> tree%lvls(lvl)%leaves
  subroutine a5g_level_leaves(tree, lvl, leaves)
    type(a5g_t), intent(in)           :: tree
    integer, intent(in)               :: lvl
    integer, allocatable, intent(out) :: leaves(:)
    integer(c_int) :: cnt(3)
    integer, allocatable :: ids(:), parents(:)
    call check(c_level_counts(tree%h, lvl, cnt), "level_counts")
    allocate(ids(max(cnt(1), 1)), leaves(max(cnt(2), 1)), parents(max(cnt(3), 1)))
    call check(c_level_lists(tree%h, lvl, ids, leaves, parents), "level_lists")
    leaves = leaves(1:cnt(2))
  end subroutine a5g_level_leaves

  ! ------------------------------------------------------------------
  ! cell data
  ! ------------------------------------------------------------------

  !> cc(:, :[, :], iv) of box id, ghosts included, in Afivo's order
  subroutine a5g_set_cc(tree, id, iv, cc)
    type(a5g_t), intent(inout) :: tree
    integer, intent(in)        :: id, iv
    real(dp), intent(in)       :: cc(*)
    call check(c_set_cc(tree%h, id, iv, cc), "set_cc")
  end subroutine a5g_set_cc

  subroutine a5g_get_cc(tree, id, iv, cc)
    type(a5g_t), intent(inout) :: tree
    integer, intent(in)        :: id, iv
    real(dp), intent(out)      :: cc(*)
    call check(c_get_cc(tree%h, id, iv, cc), "get_cc")
  end subroutine a5g_get_cc

  subroutine a5g_fill_cc(tree, ids, iv, value)
    type(a5g_t), intent(inout) :: tree
    integer, intent(in)        :: ids(:), iv
    real(dp), intent(in)       :: value
    call check(c_fill_cc(tree%h, size(ids), ids, iv, value), "fill_cc")
  end subroutine a5g_fill_cc

  subroutine a5g_sync(tree)
    type(a5g_t), intent(inout) :: tree
    call check(c_sync(tree%h), "sync")
  end subroutine a5g_sync

  ! ------------------------------------------------------------------
  ! multigrid (afivo/src/m_aX_mg.f90)
  ! ------------------------------------------------------------------

  !> mg$D_init_mg (m_aX_mg.f90:130-154) with the options in mg
  subroutine t_mg_init(tree, mg)
    type(a5g_t), intent(inout)  :: tree
    type(a5g_mg_t), intent(in)  :: mg
    call check(c_mg_setup(tree%h, mg%i_phi, mg%i_tmp, mg%i_rhs, mg%i_eps, &
         mg%n_cycle_down, mg%n_cycle_up, mg%n_cycle_base, mg%bc_kind, mg%bc_val, &
         lg(mg%auto_ops)), "mg$D_init_mg")
  end subroutine t_mg_init

  !> New value of a face rule a5_bc_value (the drivers' applied voltage)
  subroutine a5g_set_bc_value(tree, nb, value)
    type(a5g_t), intent(inout) :: tree
    integer, intent(in)        :: nb
    real(dp), intent(in)       :: value
    call check(c_mg_set_bc_value(tree%h, nb, value), "mg_set_bc_value")
  end subroutine a5g_set_bc_value

  !> Helmholtz operator lpl(phi) - kappa2 phi for the following multigrid
  !> calls (0 restores the Laplacian); the custom box_op / box_gsrb pattern of
  !> afivo/examples/tutorial/implicit.f90:147-193
  subroutine a5g_set_helmholtz(tree, kappa2)
    type(a5g_t), intent(inout) :: tree
    real(dp), intent(in)       :: kappa2
    call check(c_mg_set_helmholtz(tree%h, kappa2), "mg_set_helmholtz")
  end subroutine a5g_set_helmholtz

  !> mg$D_fas_fmg (m_aX_mg.f90:164-203)
  subroutine t_fas_fmg(tree, mg, set_residual, first_call)
    type(a5g_t), intent(inout) :: tree
    type(a5g_mg_t), intent(in) :: mg
    logical, intent(in)        :: set_residual, first_call
    call check(c_fas_fmg(tree%h, lg(set_residual), lg(first_call)), "mg$D_fas_fmg")
  end subroutine t_fas_fmg

  !> Reverse mode of mg$D_fas_fmg(tree, mg, .false., .false.): adj = cell
  !> variables lp, lt, lr, lw holding the adjoints of the potential, tmp and
  !> rhs the solve leaves (lw: work space); on return those of the values
  !> it started from, and vbar(2 dim) the adjoints of the face values V.
  !> No counterpart in the reference.
  subroutine mg_fas_fmg_adjoint(tree, adj, vbar)
    type(a5g_t), intent(inout) :: tree
    integer, intent(in)        :: adj(4)
    real(dp), intent(out)      :: vbar(:)
    integer(c_int)             :: a(4)
    real(c_double)             :: v(6)
    a = adj
    v = 0
    call check(c_fas_fmg_adjoint(tree%h, a, v), "mg$D_fas_fmg adjoint")
    vbar = v(1:size(vbar))
  end subroutine mg_fas_fmg_adjoint

  !> Reverse mode of mg$D_fas_vcycle(tree, mg, .false., max_lvl)
  !> (m_aX_mg.f90:208-255): adj and vbar as for mg_fas_fmg_adjoint.
  !> No counterpart in the reference.
  subroutine mg_fas_vcycle_adjoint(tree, max_lvl, adj, vbar)
    type(a5g_t), intent(inout) :: tree
    integer, intent(in)        :: max_lvl
    integer, intent(in)        :: adj(4)
    real(dp), intent(out)      :: vbar(:)
    integer(c_int)             :: a(4)
    real(c_double)             :: v(6)
    a = adj
    v = 0
    call check(c_fas_vcycle_adjoint(tree%h, int(max_lvl, c_int), a, v), "mg$D_fas_vcycle adjoint")
    vbar = v(1:size(vbar))
  end subroutine mg_fas_vcycle_adjoint

  !> mg$D_fas_vcycle (m_aX_mg.f90:208-255)
  subroutine t_fas_vcycle(tree, mg, max_lvl, set_residual)
    type(a5g_t), intent(inout) :: tree
    type(a5g_mg_t), intent(in) :: mg
    integer, intent(in)        :: max_lvl
    logical, intent(in)        :: set_residual
    call check(c_fas_vcycle(tree%h, max_lvl, lg(set_residual)), "mg$D_fas_vcycle")
  end subroutine t_fas_vcycle

  !> FP32 multigrid inside FP64 defect correction, to max|r| <= tol max|rhs|
  subroutine a5g_mixed_solve(tree, tol, max_outer, n_inner, zero_start, n_outer, rel_res)
    type(a5g_t), intent(inout) :: tree
    real(dp), intent(in)       :: tol
    integer, intent(in)        :: max_outer, n_inner
    logical, intent(in)        :: zero_start
    integer, intent(out)       :: n_outer
    real(dp), intent(out)      :: rel_res
    integer(c_int) :: no
    call check(c_mg_solve_mixed(tree%h, tol, max_outer, n_inner, lg(zero_start), no, &
         rel_res), "mg_solve_mixed")
    n_outer = no
  end subroutine a5g_mixed_solve

  subroutine a5g_residual_norms(tree, max_res, max_rhs)
    type(a5g_t), intent(inout) :: tree
    real(dp), intent(out)      :: max_res, max_rhs
    call check(c_mg_residual_norms(tree%h, max_res, max_rhs), "mg_residual_norms")
  end subroutine a5g_residual_norms

  ! ------------------------------------------------------------------
  ! ghost cells, restriction, prolongation, copies, reductions
  ! ------------------------------------------------------------------

  !> a$D_gc_tree (m_aX_gc.f90:28-46); rb = a5_gc_sides_rb / _interp /
  !> _interp_lim, bc = a5_gc_mg_bc / _neumann / _dirichlet
  subroutine t_gc_tree(tree, iv, rb, bc)
    type(a5g_t), intent(inout) :: tree
    integer, intent(in)        :: iv, rb, bc
    call check(c_gc_tree(tree%h, iv, rb, bc), "a$D_gc_tree")
  end subroutine t_gc_tree

  !> a$D_gc_box (m_aX_gc.f90:50-68) for a list of boxes
  subroutine t_gc_boxes(tree, ids, iv, rb, bc)
    type(a5g_t), intent(inout) :: tree
    integer, intent(in)        :: ids(:), iv, rb, bc
    call check(c_gc_boxes(tree%h, size(ids), ids, iv, rb, bc), "a$D_gc_box")
  end subroutine t_gc_boxes

  !> a$D_restrict_tree (m_aX_restrict.f90:54-64)
  subroutine t_restrict_tree(tree, iv)
    type(a5g_t), intent(inout) :: tree
    integer, intent(in)        :: iv
    call check(c_restrict_tree(tree%h, iv), "a$D_restrict_tree")
  end subroutine t_restrict_tree

  !> a$D_tree_copy_cc (m_aX_utils.f90:445-453)
  subroutine t_tree_copy_cc(tree, iv_from, iv_to)
    type(a5g_t), intent(inout) :: tree
    integer, intent(in)        :: iv_from, iv_to
    call check(c_tree_copy_cc(tree%h, iv_from, iv_to), "a$D_tree_copy_cc")
  end subroutine t_tree_copy_cc

  !> a$D_tree_max_cc (m_aX_utils.f90:591-619)
  subroutine t_tree_max_cc(tree, iv, cc_max)
    type(a5g_t), intent(inout) :: tree
    integer, intent(in)        :: iv
    real(dp), intent(out)      :: cc_max
    real(dp) :: lo
    call check(c_tree_minmax_cc(tree%h, iv, lo, cc_max), "a$D_tree_max_cc")
  end subroutine t_tree_max_cc

  !> a$D_tree_min_cc (m_aX_utils.f90:559-587)
  subroutine t_tree_min_cc(tree, iv, cc_min)
    type(a5g_t), intent(inout) :: tree
    integer, intent(in)        :: iv
    real(dp), intent(out)      :: cc_min
    real(dp) :: hi
    call check(c_tree_minmax_cc(tree%h, iv, cc_min, hi), "a$D_tree_min_cc")
  end subroutine t_tree_min_cc

  !> a$D_consistent_fluxes (m_aX_core.f90:818-930)
  subroutine t_consistent_fluxes(tree, f_ixs)
    type(a5g_t), intent(inout) :: tree
    integer, intent(in)        :: f_ixs(:)
    call check(c_consistent_fluxes(tree%h, size(f_ixs), f_ixs), "a$D_consistent_fluxes")
  end subroutine t_consistent_fluxes

  !> a$D_prolong1_to (m_aX_prolong.f90:99-191) for a list of new boxes
  subroutine t_prolong1_to(tree, ids, iv)
    type(a5g_t), intent(inout) :: tree
    integer, intent(in)        :: ids(:), iv
    call check(c_prolong1_to(tree%h, size(ids), ids, iv), "a$D_prolong1_to")
  end subroutine t_prolong1_to

  ! ------------------------------------------------------------------
  ! streamer (src/m_streamer.f90, src/streamer_{2d,3d,cyl}.f90)
  ! ------------------------------------------------------------------

  !> Variable indices (src/m_streamer.f90:16-43) and the dielectric rule of
  !> update_solution (2D/3D drivers: .true., cylindrical driver: .false.)
  subroutine st_setup(tree, i_elec, i_pion, i_elec_old, i_pion_old, i_phi, i_fld, &
       i_rhs, i_pho, i_eps, f_elec, f_fld, i_mobility, i_diffusion, i_alpha, &
       i_eta, dielectric_rule)
    type(a5g_t), intent(inout) :: tree
    integer, intent(in) :: i_elec, i_pion, i_elec_old, i_pion_old, i_phi, i_fld
    integer, intent(in) :: i_rhs, i_pho, i_eps, f_elec, f_fld
    integer, intent(in) :: i_mobility, i_diffusion, i_alpha, i_eta
    logical, intent(in) :: dielectric_rule
    integer(c_int) :: v(16)
    v = [i_elec, i_pion, i_elec_old, i_pion_old, i_phi, i_fld, i_rhs, i_pho, i_eps, &
         f_elec, f_fld, i_mobility, i_diffusion, i_alpha, i_eta, lg(dielectric_rule)]
    call check(c_streamer_setup(tree%h, v), "streamer_setup")
  end subroutine st_setup

  !> LT_create (fosito/m_lookup_table.f90:46-58) of the transport table
  subroutine st_table_create(tree, x_min, x_max, n_rows, n_cols)
    type(a5g_t), intent(inout) :: tree
    real(dp), intent(in)       :: x_min, x_max
    integer, intent(in)        :: n_rows, n_cols
    call check(c_table_create(tree%h, x_min, x_max, n_rows, n_cols), "LT_create")
  end subroutine st_table_create

  !> LT_set_col (fosito/m_lookup_table.f90:61-72)
  subroutine st_table_set_col(tree, col, xx, yy)
    type(a5g_t), intent(inout) :: tree
    integer, intent(in)        :: col
    real(dp), intent(in)       :: xx(:), yy(:)
    call check(c_table_set_col(tree%h, col, size(xx), xx, yy), "LT_set_col")
  end subroutine st_table_set_col

  !> The table as built: rows_cols(n_rows, n_cols), x_min, inv_dx
  subroutine st_table_get(tree, rows_cols, x_min, inv_dx)
    type(a5g_t), intent(inout) :: tree
    real(dp), intent(out)      :: rows_cols(:, :), x_min, inv_dx
    call check(c_table_get(tree%h, rows_cols, x_min, inv_dx), "table_get")
  end subroutine st_table_get

  !> compute_fld's source term (src/streamer_2d.f90:318-329)
  subroutine st_space_charge(tree, fac)
    type(a5g_t), intent(inout) :: tree
    real(dp), intent(in)       :: fac
    call check(c_space_charge(tree%h, fac), "space_charge")
  end subroutine st_space_charge

  !> fld_from_pot on every box (src/streamer_2d.f90:346-381)
  subroutine st_fld_from_pot(tree)
    type(a5g_t), intent(inout) :: tree
    call check(c_fld_from_pot(tree%h), "fld_from_pot")
  end subroutine st_fld_from_pot

  !> fluxes_koren on the leaves (src/streamer_2d.f90:384-489)
  subroutine st_fluxes_koren(tree, dt)
    type(a5g_t), intent(inout) :: tree
    real(dp), intent(in)       :: dt
    call check(c_fluxes_koren(tree%h, dt), "fluxes_koren")
  end subroutine st_fluxes_koren

  !> update_solution on the leaves (src/streamer_2d.f90:499-542)
  subroutine st_update_solution(tree, dt)
    type(a5g_t), intent(inout) :: tree
    real(dp), intent(in)       :: dt
    call check(c_update_solution(tree%h, dt), "update_solution")
  end subroutine st_update_solution

  !> Photoionization source in i_pho by the three-term Helmholtz model, in
  !> place of set_photoionization's Monte-Carlo photons
  !> (src/streamer_2d.f90:544-589): coeff = eta p_q / (p + p_q), amp(j) =
  !> A_j p_O2^2 and kappa(j) = lambda_j p_O2 in SI units; i_phi, i_rhs and
  !> i_tmp are the work variables of the solves.
  subroutine st_photoi_helmholtz(tree, coeff, amp, kappa, i_phi, i_rhs, i_tmp, tol, &
       max_cycles, cycles)
    type(a5g_t), intent(inout) :: tree
    real(dp), intent(in)       :: coeff, amp(:), kappa(:), tol
    integer, intent(in)        :: i_phi, i_rhs, i_tmp, max_cycles
    integer, intent(out)       :: cycles
    integer(c_int)             :: cyc
    if (size(amp) /= size(kappa)) error stop "st_photoi_helmholtz: amp and kappa differ"
    call check(c_photoi_helmholtz(tree%h, coeff, size(amp), amp, kappa, i_phi, i_rhs, &
         i_tmp, tol, max_cycles, cyc), "photoi_helmholtz")
    cycles = cyc
  end subroutine st_photoi_helmholtz

  !> The same model with term j in variables of its own: solution
  !> i_terms(j), right-hand side i_rhs(j), multigrid scratch i_tmp(j); the
  !> terms' solves run side by side. warm = .true. starts each term's
  !> V-cycles from the solution an earlier call left in i_terms(j) (and
  !> prolonged to new boxes) instead of an FMG from zero.
  subroutine st_photoi_helmholtz_terms(tree, coeff, amp, kappa, i_terms, i_rhs, i_tmp, warm, &
       tol, max_cycles, cycles)
    type(a5g_t), intent(inout) :: tree
    real(dp), intent(in)       :: coeff, amp(:), kappa(:), tol
    integer, intent(in)        :: i_terms(:), i_rhs(:), i_tmp(:), max_cycles
    logical, intent(in)        :: warm
    integer, intent(out)       :: cycles
    integer(c_int)             :: cyc, iw
    integer(c_int)             :: it(size(i_terms)), ir(size(i_terms)), im(size(i_terms))
    if (size(amp) /= size(kappa) .or. size(amp) /= size(i_terms) .or. &
         size(amp) /= size(i_rhs) .or. size(amp) /= size(i_tmp)) &
         error stop "st_photoi_helmholtz_terms: amp, kappa and the variables differ"
    it = i_terms
    ir = i_rhs
    im = i_tmp
    iw = merge(1, 0, warm)
    call check(c_photoi_helmholtz_terms(tree%h, coeff, size(amp), amp, kappa, it, ir, im, &
         iw, tol, max_cycles, cyc), "photoi_helmholtz_terms")
    cycles = cyc
  end subroutine st_photoi_helmholtz_terms

  !> average_dens on every box (src/streamer_2d.f90:492-496)
  subroutine st_average_dens(tree)
    type(a5g_t), intent(inout) :: tree
    call check(c_average_dens(tree%h), "average_dens")
  end subroutine st_average_dens

  !> Per-box maxima read by set_ref_flags (src/streamer_2d.f90:186-192):
  !> summ(:, id) = (dr^2 max|rhs|, max fld, max elec), boxes of level >= 1
  subroutine st_box_summaries(tree, summ)
    type(a5g_t), intent(inout)         :: tree
    real(dp), allocatable, intent(out) :: summ(:, :)
    allocate(summ(3, a5g_max_id(tree)))
    call check(c_box_summaries(tree%h, summ), "box_summaries")
  end subroutine st_box_summaries

  !> Reverse mode of one substep of the time loop (fluxes_koren,
  !> a$D_consistent_fluxes, update_solution; src/streamer_2d.f90:384-542) on
  !> a frozen mesh and field: adj = cell variables le, li, lE, lpho and face
  !> variables lF, lfld. le, li hold the adjoints of the new densities and
  !> get those of the old ones; lF must start at zero. No counterpart in the
  !> reference.
  subroutine st_transport_adjoint(tree, dt, adj)
    type(a5g_t), intent(inout) :: tree
    real(dp), intent(in)       :: dt
    integer, intent(in)        :: adj(6)
    integer(c_int)             :: a(6)
    a = adj
    call check(c_update_solution_adjoint(tree%h, dt, a), "update_solution adjoint")
    call check(c_consistent_fluxes_adjoint(tree%h, a(5)), "consistent_fluxes adjoint")
    call check(c_fluxes_koren_adjoint(tree%h, dt, a), "fluxes_koren adjoint")
  end subroutine st_transport_adjoint

  !> The adjoints of the transport-table rows accumulated so far,
  !> rows_cols(n_rows, n_cols); clear them after reading when clear
  subroutine st_table_adjoint(tree, rows_cols, clear)
    type(a5g_t), intent(inout)        :: tree
    real(dp), intent(inout), target   :: rows_cols(:, :)
    logical, intent(in)               :: clear
    call check(c_table_adjoint(tree%h, merge(1, 0, clear), c_loc(rows_cols)), "table adjoint")
  end subroutine st_table_adjoint

  !> Reverse mode of a$D_tree_copy_cc(from, to): lfrom += lto, lto = 0
  subroutine st_copy_cc_adjoint(tree, lfrom, lto)
    type(a5g_t), intent(inout) :: tree
    integer, intent(in)        :: lfrom, lto
    call check(c_tree_copy_cc_adjoint(tree%h, lfrom, lto), "tree_copy_cc adjoint")
  end subroutine st_copy_cc_adjoint

  !> Reverse mode of average_dens: adj = le, le_old, li, li_old
  subroutine st_average_dens_adjoint(tree, adj)
    type(a5g_t), intent(inout) :: tree
    integer, intent(in)        :: adj(4)
    integer(c_int)             :: a(4)
    a = adj
    call check(c_average_dens_adjoint(tree%h, a), "average_dens adjoint")
  end subroutine st_average_dens_adjoint

  !> Reverse mode of compute_fld's rhs = fac (n_e - n_i): adj = lrhs, le, li
  subroutine st_space_charge_adjoint(tree, fac, adj)
    type(a5g_t), intent(inout) :: tree
    real(dp), intent(in)       :: fac
    integer, intent(in)        :: adj(3)
    integer(c_int)             :: a(3)
    a = adj
    call check(c_space_charge_adjoint(tree%h, fac, a), "space charge adjoint")
  end subroutine st_space_charge_adjoint

  !> Reverse mode of a$D_restrict_tree for the adjoint variable il
  subroutine st_restrict_tree_adjoint(tree, il)
    type(a5g_t), intent(inout) :: tree
    integer, intent(in)        :: il
    call check(c_restrict_tree_adjoint(tree%h, il), "restrict_tree adjoint")
  end subroutine st_restrict_tree_adjoint

  !> Reverse mode of a$D_gc_tree(iv, rb, bc), rb 1 (interp) or 2
  !> (interp_lim), bc 1 (Neumann) or 2 (Dirichlet); il the adjoint of iv
  subroutine st_gc_tree_adjoint(tree, iv, il, rb, bc)
    type(a5g_t), intent(inout) :: tree
    integer, intent(in)        :: iv, il, rb, bc
    call check(c_gc_tree_adjoint(tree%h, iv, il, rb, bc), "gc_tree adjoint")
  end subroutine st_gc_tree_adjoint

  !> Reverse mode of fld_from_pot: adj = lE (cell), lfld (face), lphi (cell)
  subroutine st_fld_from_pot_adjoint(tree, adj)
    type(a5g_t), intent(inout) :: tree
    integer, intent(in)        :: adj(3)
    integer(c_int)             :: a(3)
    a = adj
    call check(c_fld_from_pot_adjoint(tree%h, a), "fld_from_pot adjoint")
  end subroutine st_fld_from_pot_adjoint

  !> Whole cell variables ivs and face variables fvs: op 0 saves them in the
  !> device store slot `slot`, op 1 restores them from it, op 2 sets them to
  !> zero
  subroutine a5g_state_store(tree, slot, ivs, fvs, op)
    type(a5g_t), intent(inout) :: tree
    integer, intent(in)        :: slot, ivs(:), fvs(:), op
    integer(c_int)             :: a(max(1, size(ivs))), b(max(1, size(fvs)))
    a = 0
    b = 0
    a(1:size(ivs)) = ivs
    b(1:size(fvs)) = fvs
    call check(c_state_store(tree%h, slot, size(ivs), a, size(fvs), b, op), "state store")
  end subroutine a5g_state_store

  !> st_box_summaries without any device allocation (the library's fixed
  !> scratch, at most 131072 boxes; the tree must have been uploaded since
  !> its last change). summ(3, max_id) is the caller's.
  subroutine st_box_summaries_noalloc(tree, summ)
    type(a5g_t), intent(inout) :: tree
    real(dp), intent(inout)    :: summ(:, :)
    if (size(summ, 1) /= 3 .or. size(summ, 2) < a5g_max_id(tree)) &
         error stop "st_box_summaries_noalloc: summ must be (3, max_id)"
    call check(c_box_summaries_noalloc(tree%h, summ), "box_summaries_noalloc")
  end subroutine st_box_summaries_noalloc

  ! ------------------------------------------------------------------
  ! streamer diagnostics (src/streamer_cyl.f90:682-956)
  ! ------------------------------------------------------------------

  !> a$D_get_loc (afivo/src/m_aX_utils.f90:244-281)
  function t_get_loc(tree, rr, max_lvl) result(loc)
    type(a5g_t), intent(in)       :: tree
    real(dp), intent(in)          :: rr(:)
    integer, intent(in), optional :: max_lvl
    type(a5g_loc_t)               :: loc
    integer(c_int) :: m
    m = 0
    if (present(max_lvl)) m = max(max_lvl, 1)
    call check(c_get_loc(tree%h, rr, m, loc%id, loc%ix), "a$D_get_loc")
  end function t_get_loc

  !> a$D_reduction_loc (afivo/src/m_aX_utils.f90:499-555) with reduce_max and
  !> the per-box rule kind (a5_rl_face_mean on face variable var,
  !> a5_rl_coord_max on cells whose cc(var) > thr) along dimension d.
  !> out_loc keeps its value when no box exceeds init_val, as an intent(out)
  !> location without default initialisation does in the reference.
  subroutine st_reduction_loc(tree, kind, var, d, thr, init_val, out_val, out_loc)
    type(a5g_t), intent(inout)     :: tree
    integer, intent(in)            :: kind, var, d
    real(dp), intent(in)           :: thr, init_val
    real(dp), intent(out)          :: out_val
    type(a5g_loc_t), intent(inout) :: out_loc
    integer(c_int) :: id, ix(3)
    call check(c_reduction_loc(tree%h, kind, var, d, thr, init_val, out_val, id, ix), &
         "a$D_reduction_loc")
    if (id /= -1) then
       out_loc%id = id
       out_loc%ix = ix
    end if
  end subroutine st_reduction_loc

  !> get_cc_axis (src/streamer_cyl.f90:905-956)
  subroutine st_get_cc_axis(tree, ixs_cc, ixs_fc, axis_data)
    type(a5g_t), intent(inout)           :: tree
    integer, intent(in)                  :: ixs_cc(:), ixs_fc(:)
    real(dp), allocatable, intent(inout) :: axis_data(:, :)
    integer(c_int) :: n
    real(dp)       :: none(1)
    call check(c_cc_axis(tree%h, size(ixs_cc), ixs_cc, size(ixs_fc), ixs_fc, 0, none, n), &
         "get_cc_axis")
    if (allocated(axis_data)) deallocate(axis_data)
    allocate(axis_data(size(ixs_cc) + size(ixs_fc) + 1, n))
    if (n > 0) call check(c_cc_axis(tree%h, size(ixs_cc), ixs_cc, size(ixs_fc), ixs_fc, n, &
         axis_data, n), "get_cc_axis")
  end subroutine st_get_cc_axis

  ! ------------------------------------------------------------------
  ! output
  ! ------------------------------------------------------------------

  !> a$D_write_vtk (afivo/src/m_aX_io.f90:20-209): the leaves as a VTK
  !> unstructured grid in <dir>/<filename>.vtu
  subroutine t_write_vtk(tree, filename, cc_names, n_cycle, time, ixs_cc, &
       fc_names, ixs_fc, dir)
    type(a5g_t), intent(inout)             :: tree
    character(len=*), intent(in)           :: filename, cc_names(:)
    integer, intent(in)                    :: n_cycle
    real(dp), intent(in)                   :: time
    integer, intent(in), optional          :: ixs_cc(:), ixs_fc(:)
    character(len=*), intent(in), optional :: fc_names(:), dir
    character(len=:), allocatable          :: fname
    character(kind=c_char), allocatable, target :: ccb(:, :), fcb(:, :)
    type(c_ptr), allocatable               :: ccp(:), fcp(:)
    integer(c_int), allocatable, target    :: icc(:)
    integer(c_int), allocatable            :: ifc(:)
    type(c_ptr)                            :: picc
    integer                                :: i, n_fc

    if (present(ixs_cc)) then
       if (size(ixs_cc) /= size(cc_names)) &
            stop "a$D_write_vtk: size(cc_names) /= size(ixs_cc)"
       icc = ixs_cc
       picc = c_loc(icc)
    else
       picc = c_null_ptr
    end if
    n_fc = 0
    allocate(ifc(1), fcp(1))
    ifc = 0
    fcp = c_null_ptr
    if (present(fc_names)) then
       if (.not. present(ixs_fc)) stop "a$D_write_vtk: ixs_fc not present (but fc_names is)"
       if (size(ixs_fc) * tree%dim /= size(fc_names)) &
            stop "a$D_write_vtk: size(fc_names) /= size(ixs_fc) * $D"
       n_fc = size(ixs_fc)
       ifc = ixs_fc
       call c_strings(fc_names, fcb, fcp)
    end if
    call c_strings(cc_names, ccb, ccp)
    fname = trim(filename)
    if (present(dir)) then
       i = len_trim(dir)
       if (i > 0) then
          if (dir(i:i) == "/") then
             fname = trim(dir) // fname
          else
             fname = trim(dir) // "/" // fname
          end if
       end if
    end if
    call check(c_write_vtk(tree%h, fname // c_null_char, size(cc_names), ccp, picc, n_fc, &
         fcp, ifc, n_cycle, time), "a$D_write_vtk")
  end subroutine t_write_vtk

  !> NUL-terminated copies of trimmed names and pointers to them
  subroutine c_strings(names, buf, ptrs)
    character(len=*), intent(in)                                  :: names(:)
    character(kind=c_char), allocatable, target, intent(inout)    :: buf(:, :)
    type(c_ptr), allocatable, intent(inout)                       :: ptrs(:)
    integer :: i, j, n
    allocate(buf(len(names) + 1, size(names)), ptrs(max(size(names), 1)))
    ptrs = c_null_ptr
    do i = 1, size(names)
       n = len_trim(names(i))
       do j = 1, n
          buf(j, i) = names(i)(j:j)
       end do
       buf(n + 1, i) = c_null_char
       ptrs(i) = c_loc(buf(1, i))
    end do
  end subroutine c_strings

  !> The raw leaf-box dump of the variables ivs (see a5g_dump_leaves)
  subroutine a5g_dump_leaves(tree, filename, ivs, cycle, time)
    type(a5g_t), intent(inout)   :: tree
    character(len=*), intent(in) :: filename
    integer, intent(in)          :: ivs(:), cycle
    real(dp), intent(in)         :: time
    call check(c_dump_leaves(tree%h, trim(filename) // c_null_char, size(ivs), ivs, cycle, &
         time), "dump_leaves")
  end subroutine a5g_dump_leaves

  ! ------------------------------------------------------------------
  ! independent runs in one tree
  ! ------------------------------------------------------------------

  !> n_inst runs stacked along the last dimension, one every stride
  subroutine a5g_set_instances(tree, n_inst, stride)
    type(a5g_t), intent(inout) :: tree
    integer, intent(in)        :: n_inst
    real(dp), intent(in)       :: stride
    call check(c_set_instances(tree%h, n_inst, stride), "set_instances")
  end subroutine a5g_set_instances

  !> Per-run time steps (what = 0) or values V of the rule 2V - c on face what
  subroutine a5g_set_instance_values(tree, what, values)
    type(a5g_t), intent(inout) :: tree
    integer, intent(in)        :: what
    real(dp), intent(in)       :: values(:)
    call check(c_set_instance_values(tree%h, what, size(values), values), &
         "set_instance_values")
  end subroutine a5g_set_instance_values

  !> Restrict the tree-wide calls to the runs with active(g)
  subroutine a5g_set_active_instances(tree, active)
    type(a5g_t), intent(inout) :: tree
    logical, intent(in)        :: active(:)
    integer(c_int) :: a(size(active))
    a = merge(1, 0, active)
    call check(c_set_active_instances(tree%h, size(a), a), "set_active_instances")
  end subroutine a5g_set_active_instances

  !> Groups of runs for mg_fas_fmg_groups: group(g) >= 0 is run g's group,
  !> -1 leaves it out; an empty array clears the groups
  subroutine a5g_set_instance_groups(tree, group)
    type(a5g_t), intent(inout) :: tree
    integer, intent(in)        :: group(:)
    integer(c_int) :: a(max(size(group), 1))
    a = 0
    a(1:size(group)) = group
    call check(c_set_instance_groups(tree%h, size(group), a), "set_instance_groups")
  end subroutine a5g_set_instance_groups

  !> mg$D_fas_fmg (afivo/src/m_aX_mg.f90:164-203) of every group of runs, each
  !> up to its own finest level, the groups side by side
  subroutine mg_fas_fmg_groups(tree, set_residual, first_call)
    type(a5g_t), intent(inout) :: tree
    logical, intent(in)        :: set_residual, first_call
    call check(c_fas_fmg_groups(tree%h, merge(1, 0, set_residual), merge(1, 0, first_call)), &
         "fas_fmg_groups")
  end subroutine mg_fas_fmg_groups

end module m_a5g
