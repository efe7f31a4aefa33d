/* a5g — MI355X (gfx950) implementation of the a5_streamer per-timestep hot
 * path: Afivo's FAS multigrid Poisson solve, ghost cells, restriction and
 * prolongation over quadtree/octree box lists, and the streamer's
 * drift-diffusion flux/source update.
 *
 * Plain C ABI: opaque handle, plain pointers and sizes, int status (0 = ok;
 * otherwise a5g_last_error() describes the failure). No exception crosses
 * the boundary. Each entry point names the reference routine it stands in
 * for; argument meaning and 1-based ids / variable indices follow Afivo
 * (afivo/src/m_aX_t.f90:136-205): box id 0 = no box, negative neighbour ids
 * are physical boundaries, a box's cell block is cc(0:nc+1,0:nc+1[,0:nc+1])
 * with i fastest. One handle is used by one host thread at a time; all GPU
 * work is ordered on the handle's own HIP stream (independent photoionization
 * terms run on side streams that it waits for before it goes on).
 */
#ifndef A5G_H
#define A5G_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct a5g_tree a5g_tree;

/* Message of the last failure on this host thread ("" if none). */
const char *a5g_last_error(void);

/* ---- tree (afivo/src/m_aX_core.f90) --------------------------------- */

/* a$D_init, m_aX_core.f90:47-116. coarsen_to <= 0: no extra coarse levels.
 * coord 0 = Cartesian, 1 = cylindrical (2D only). */
int a5g_init(a5g_tree **t, int dim, int n_cell, int n_var_cell,
             int n_var_face, double dr, const double *r_min, int coarsen_to,
             int lvls_max, int coord, int n_boxes);
/* a$D_destroy, m_aX_core.f90:120-133 (also frees GPU memory) */
int a5g_destroy(a5g_tree *t);
/* a$D_set_base, m_aX_core.f90:137-220; ix_list[n][dim], nb_list[n][2*dim] */
int a5g_set_base(a5g_tree *t, int n, const int *ix_list, const int *nb_list);
/* a$D_adjust_refinement, m_aX_core.f90:473-537. The user's set_ref_flags
 * callback is replaced by flags[id-1] (id = 1..n_flags): -1 derefine,
 * 0 keep, 1 refine; boxes beyond n_flags keep their refinement. */
int a5g_adjust_refinement(a5g_tree *t, int n_flags, const int *flags,
                          int *n_add, int *n_rm);
/* ref_info_t%lvls(lvl)%add / %rm of the last adjustment
 * (m_aX_core.f90:540-601); pass NULL arrays to query the counts only. */
int a5g_ref_info_lvl(a5g_tree *t, int lvl, int *add, int *rm, int *n_add,
                     int *n_rm);
/* a$D_tidy_up, m_aX_core.f90:224-325 (Morton reorder; moves GPU data) */
int a5g_tidy_up(a5g_tree *t, double max_frac, double goal_frac,
                int n_clean_min, int reorder);
/* out8 = max_id, max_lvl, min_lvl, lvls_max, storage size, n_cell, coord,
 * ready (the fields of a$D_t, m_aX_t.f90:173-186) */
int a5g_info(a5g_tree *t, int *out8);
/* Box records 1..n (box$D_t, m_aX_t.f90:136-161); arrays are [n][...] */
int a5g_get_boxes(a5g_tree *t, int n, int *lvl, int *ix, int *parent,
                  int *children, int *neighbors, int *in_use, int *n_cell,
                  double *dr, double *r_min);
/* tree%lvls(lvl): counts3 = sizes of ids, leaves, parents */
int a5g_level_counts(a5g_tree *t, int lvl, int *counts3);
int a5g_level_lists(a5g_tree *t, int lvl, int *ids, int *leaves,
                    int *parents);

/* ---- data movement (host <-> HBM) ----------------------------------- */

/* cc(..., iv) of one box, ghosts included ((nc+2)^dim values) */
int a5g_set_cc(a5g_tree *t, int id, int iv, const double *data);
int a5g_get_cc(a5g_tree *t, int id, int iv, double *data);
/* cc(..., iv) of boxes 1..max_id, each at stride (n_cell+2)^dim */
int a5g_set_cc_all(a5g_tree *t, int iv, const double *data);
int a5g_get_cc_all(a5g_tree *t, int iv, double *data);
/* fx/fy/fz(..., ifc) of one box: dim = 1, 2, 3 */
int a5g_set_fc(a5g_tree *t, int id, int dim, int ifc, const double *data);
int a5g_get_fc(a5g_tree *t, int id, int dim, int ifc, double *data);
/* Wait for all GPU work of the handle */
int a5g_sync(a5g_tree *t);

/* ---- multigrid (afivo/src/m_aX_mg.f90) ------------------------------ */

/* mg$D_t options + mg$D_init_mg, m_aX_mg.f90:20-52, 130-154.
 * bc_kind per physical face (2*dim entries, order lx hx ly hy [lz hz]):
 * 0 = Neumann zero, 1 = Dirichlet zero, 2 = value bc_val (ghost 2V - c):
 * the three forms of the drivers' sides_bc callbacks
 * (src/streamer_2d.f90:620-638). auto_ops = 1 installs mg$D_auto_* as the
 * drivers do; 0 keeps the plain Laplacian defaults. */
int a5g_mg_setup(a5g_tree *t, int i_phi, int i_tmp, int i_rhs, int i_eps,
                 int n_cycle_down, int n_cycle_up, int n_cycle_base,
                 const int *bc_kind, const double *bc_val, int auto_ops);
/* Change the value of a physical-face rule (the applied voltage) */
int a5g_mg_set_bc_value(a5g_tree *t, int nb, double value);
/* Helmholtz operator lpl(phi) - kappa2 phi for the multigrid calls below
 * (smoother (sum of the neighbours - h^2 rhs) / (2D + kappa2 h^2)); 0 restores
 * the Laplacian. No counterpart in the reference's drivers: it is the custom
 * mg%box_op / mg%box_gsrb pair that Afivo's tutorial installs for an implicit
 * operator (afivo/examples/tutorial/implicit.f90:147-193). Not combined with
 * a permittivity (auto_ops with i_eps > 0). */
int a5g_mg_set_helmholtz(a5g_tree *t, double kappa2);
/* mg$D_fas_fmg, m_aX_mg.f90:164-203 */
int a5g_fas_fmg(a5g_tree *t, int set_residual, int first_call);
/* mg$D_fas_vcycle, m_aX_mg.f90:208-255 */
int a5g_fas_vcycle(a5g_tree *t, int max_lvl, int set_residual);
/* Reverse mode of a5g_fas_fmg(t, 0, 0) (mg$D_fas_fmg, m_aX_mg.f90:164-203,
 * with first_call false and no residual): the solve is an affine map of the
 * potential, tmp and rhs it starts from and of the values V of the rule
 * 2V - c; this applies its transpose. adj[4] = cell variables lp, lt, lr,
 * lw: on entry lp, lt, lr hold the adjoints of the potential, tmp and rhs
 * the solve leaves (every cell, ghosts included); on return the adjoints of
 * the values it started from; lw is work space. vbar[2 dim] receives the
 * adjoints of V per face. The transpose of each operation of the
 * reference's schedule, in reverse order; no state of the forward solve is
 * needed. The operators of the forward solve per box (the Laplacian, with
 * the Helmholtz term of a5g_mg_set_helmholtz, or the permittivity forms
 * mg$D_box_lpld / gsrb_lpld / corr_lpld on boxes tagged with eps); one run
 * per tree.
 * No counterpart in the reference, which computes no derivatives. */
int a5g_fas_fmg_adjoint(a5g_tree *t, const int *adj, double *vbar);
/* Reverse mode of a5g_fas_vcycle(t, max_lvl, 0) (mg$D_fas_vcycle,
 * m_aX_mg.f90:208-255, without the residual), an affine map of the same
 * values: the same adj[4] and vbar as a5g_fas_fmg_adjoint, the same
 * operators. With it the reverse of a whole solve (an FMG and V-cycles) is
 * the V-cycles' reverses, last first, then the FMG's. No counterpart in the
 * reference, which computes no derivatives. */
int a5g_fas_vcycle_adjoint(a5g_tree *t, int max_lvl, const int *adj, double *vbar);
/* Launch shapes of the box kernels (no reference counterpart: the
 * reference's loops over boxes have no launch geometry). kernel 0 is the
 * multigrid half-sweep, 1 the ghost fill, 2 the residual + restriction, 3
 * the coarse rhs, 4 the correction; for boxes of nc cells a side (2..32) on
 * levels of size class cls (0: < 1024 boxes, 1: < 16384, 2: more), shape 0
 * is the default (lanes per box from the cells the kernel treats, 256-lane
 * workgroups, one workgroup per group of boxes), 2 half the lanes per box
 * and twice the boxes per workgroup (kernels 1-4: boxes of up to 8 cells a
 * side); for the half-sweep also 1, the default on a persistent grid (8
 * workgroups per compute unit walking the list), and 3, shape 2 on a
 * persistent grid. Every shape computes the same values. */
int a5g_set_shape(a5g_tree *t, int kernel, int nc, int cls, int shape);
int a5g_get_shape(a5g_tree *t, int kernel, int nc, int cls, int *shape);
/* Autotuner: times every shape of every kernel above on the tree's current
 * levels (reps launches per shape and level on the multigrid's potential,
 * tmp and rhs, which are restored afterwards) and keeps, per kernel, box
 * size and size class present, the fastest when it beats the default by
 * more than 3%.
 * Writes up to max_rec rows of 8 doubles (kernel, dim, nc, size class,
 * shape, boxes timed, microseconds per launch, 1 if chosen) and their
 * count to *n_rec. */
int a5g_autotune(a5g_tree *t, int reps, int max_rec, double *rec, int *n_rec);
/* Mixed-precision solve of the same problem to max|r| <= tol max|rhs| over
 * the leaves: an FP64 defect-correction loop whose correction A e = r (zero
 * boundary data) is solved by FP32 FAS multigrid with the schedule of
 * mg$D_fas_fmg / mg$D_fas_vcycle — the first correction by one FMG plus
 * n_inner - 1 V-cycles, later ones by n_inner V-cycles from e = 0 — while
 * the residual, the update phi += e, the restriction of phi to the parents
 * and the ghost cells of phi stay FP64. zero_start = 1 starts from phi = 0.
 * No reference counterpart: the reference solves in FP64 only
 * (afivo/src/m_aX_t.f90:11). */
int a5g_mg_solve_mixed(a5g_tree *t, double tol, int max_outer, int n_inner,
                       int zero_start, int *n_outer, double *rel_res);
/* max over leaf cells of |rhs - A phi| (scratch: i_tmp), and of |rhs| */
int a5g_mg_residual_norms(a5g_tree *t, double *max_res, double *max_rhs);

/* ---- ghost cells, restriction, copies (m_aX_gc/restrict/utils) -------- */

/* a$D_gc_tree, m_aX_gc.f90:28-46. rb: 0 = mg$D_sides_rb, 1 = a$D_gc_interp,
 * 2 = a$D_gc_interp_lim; bc: 0 = the multigrid face rules, 1 =
 * a$D_gc_neumann, 2 = a$D_gc_dirichlet */
int a5g_gc_tree(a5g_tree *t, int iv, int rb, int bc);
/* a$D_restrict_tree, m_aX_restrict.f90:54-64 */
int a5g_restrict_tree(a5g_tree *t, int iv);
/* a$D_tree_copy_cc, m_aX_utils.f90:445-453 */
int a5g_tree_copy_cc(a5g_tree *t, int iv_from, int iv_to);
/* a$D_tree_min_cc / a$D_tree_max_cc, m_aX_utils.f90:559-619 */
int a5g_tree_minmax_cc(a5g_tree *t, int iv, double *vmin, double *vmax);

/* ---- transport lookup table (fosito/m_lookup_table.f90) --------------- */

/* LT_create: a table of n_rows on the uniform grid [x_min, x_max] */
int a5g_table_create(a5g_tree *t, double x_min, double x_max, int n_rows,
                     int n_cols);
/* LT_set_col: column col (1-based) from scattered data (xx, yy), resampled
 * by linear interpolation with end clamping (LT_lin_interp_list) */
int a5g_table_set_col(a5g_tree *t, int col, int n, const double *xx,
                      const double *yy);
/* The table as built: rows_cols(n_rows, n_cols) column-major, x_min,
 * inv_dx = (n_rows - 1) / (x_max - x_min) */
int a5g_table_get(a5g_tree *t, double *rows_cols, double *x_min,
                  double *inv_dx);
/* A table built by the caller (e.g. the drivers' ST_load_transport_data) */
int a5g_set_table(a5g_tree *t, int n_rows, int n_cols, double x_min,
                  double inv_dx, const double *rows_cols);

/* ---- streamer physics (src/streamer_{2d,3d,cyl}.f90) ---------------- */

/* Variable indices used by the calls below (src/m_streamer.f90:16-43),
 * vars[16] = i_elec, i_pion, i_elec_old, i_pion_old, i_phi, i_fld, i_rhs,
 * i_pho (<= 0: no photoionisation source), i_eps, f_elec, f_fld, and the
 * table columns mobility, diffusion, alpha, eta, then 1 when the update
 * moves electrons in dielectric cells to the ions (2D/3D drivers) or 0
 * (cylindrical driver, src/streamer_cyl.f90:531-570). */
int a5g_streamer_setup(a5g_tree *t, const int *vars);
/* compute_fld's source term on leaves: rhs = fac (n_e - n_i), whole boxes
 * (src/streamer_2d.f90:318-329) */
int a5g_space_charge(a5g_tree *t, double fac);
/* fld_from_pot on every box (src/streamer_2d.f90:346-381,
 * src/streamer_3d.f90:348-395) */
int a5g_fld_from_pot(a5g_tree *t);
/* fluxes_koren on leaves (src/streamer_2d.f90:384-489,
 * src/streamer_3d.f90:396-554, src/streamer_cyl.f90:416-521) */
int a5g_fluxes_koren(a5g_tree *t, double dt);
/* a$D_consistent_fluxes, afivo/src/m_aX_core.f90:818-930 */
int a5g_consistent_fluxes(a5g_tree *t, int n, const int *f_ixs);
/* update_solution on leaves (src/streamer_2d.f90:499-542,
 * src/streamer_3d.f90:566-615, src/streamer_cyl.f90:531-570) */
int a5g_update_solution(a5g_tree *t, double dt);
/* average_dens on every box, whole boxes (src/streamer_2d.f90:492-496) */
int a5g_average_dens(a5g_tree *t);
/* Photoionization source into i_pho by the Helmholtz model, replacing
 * set_photoionization's Monte-Carlo photons (src/streamer_2d.f90:544-565,
 * src/m_photons.f90:207-627), which have no GPU counterpart here:
 * I = |E| mu alpha n_e coeff, clamped at 0 (set_photoi_rate, :567-589,
 * coeff = eta p_q / (p + p_q)), then for j < n_terms the multigrid solve of
 * (lpl - kappa_j^2) S_j = -amp_j I with S_j = 0 on the domain boundary, and
 * i_pho = sum_j S_j on leaves. amp_j = A_j p_O2^2 (m^-2), kappa_j =
 * lambda_j p_O2 (m^-1) of the three-term fit of Bourdon et al. 2007. Each
 * solve is one FMG plus V-cycles until max|r| <= tol max|rhs|; i_phi, i_rhs
 * and i_tmp are its work variables. The multigrid options of the field solve
 * are kept. *cycles = V-cycles run in total. */
int a5g_photoi_helmholtz(a5g_tree *t, double coeff, int n_terms, const double *amp,
                         const double *kappa, int i_phi, int i_rhs, int i_tmp,
                         double tol, int max_cycles, int *cycles);
/* The same model with every term in variables of its own: S_j in
 * i_terms[j], its right-hand side -amp_j I in i_rhs[j] and its multigrid
 * scratch in i_tmp[j] (1 <= n_terms <= 4; all 3 n_terms variables and i_pho
 * distinct). The terms then share no data and their solves run side by side
 * on the GPU. S_j stays in i_terms[j] for the next call. warm = 0: each term
 * is solved as above (FMG from zero, then V-cycles). warm != 0: i_terms[j]
 * holds the solution of an earlier call, prolonged to boxes created since
 * (a$D_prolong1_to); it is restricted to every parent with the multigrid's
 * restriction (mg$D_box_rstr_lpl, afivo/src/m_aX_mg.f90:851-893), its ghost
 * cells are filled level by level upwards, and V-cycles alone bring max|r|
 * under tol max|rhs_j|. Same replacement of set_photoionization
 * (src/streamer_2d.f90:544-565) as a5g_photoi_helmholtz; the source moves
 * little in one time step, so an earlier solution is close and the FMG is
 * not needed. i_pho = sum_j S_j, added in term order. */
int a5g_photoi_helmholtz_terms(a5g_tree *t, double coeff, int n_terms, const double *amp,
                               const double *kappa, const int *i_terms, const int *i_rhs,
                               const int *i_tmp, int warm, double tol, int max_cycles,
                               int *cycles);
/* The maxima set_ref_flags reads per box (src/streamer_2d.f90:186-192):
 * out[id-1] = (dr^2 max|rhs|, max fld, max elec) for boxes of levels >= 1 */
int a5g_box_summaries(a5g_tree *t, double *out);
/* set_ref_flags (src/streamer_2d.f90:177-216; the 3D and cylindrical
 * drivers differ in init_window and crv_refine) evaluated on the device for
 * every box of levels >= 1 of the active runs, from the maxima of
 * a5g_box_summaries, in the arithmetic of the drivers' rule (flang's NORM2
 * for the seed distance, GM_dist_line of src/m_geom.f90:29-46), so each
 * flag equals the host rule's. The reference evaluates the rule in its
 * adjust_refinement callback; here the flags come back for
 * a5g_adjust_refinement. Rules, set once: per run n_seed seeds of 2 dim + 1
 * values (r0, r1, width), n_reg refinement regions shared by all runs of
 * 1 + 2 dim values (level below which a box whose centre lies in [lo, hi]
 * is refined; a driver variant for large runs, not in the reference),
 * the time below which the seed rule applies, crv_refine, and the base
 * boxes per run along the last dimension (a run's coordinates are those of
 * a one-run tree). */
int a5g_set_ref_rules(a5g_tree *t, int n_seed, const double *seeds, int n_reg,
                      const double *regions, double init_window, double crv_refine,
                      int n_base_last);
/* times[g]: run g's time (one value without instances); flags[id-1] = -1,
 * 0 or 1 for ids 1..max_id, 0 for boxes not flagged (levels < 1, runs not
 * active). */
int a5g_ref_flags(a5g_tree *t, const double *times, int *flags);
/* Per run (one value without instances): the largest cc(iv) over the
 * interiors of its leaf boxes of levels >= 1 (-inf if none), and its
 * highest level holding boxes: the maxima get_max_dt reads
 * (src/streamer_2d.f90:271-303). */
int a5g_instance_maxima(a5g_tree *t, int iv, double *vmax, int *lvl_max);
/* ---- reverse mode of the transport substep (frozen mesh) ------------- */

/* The adjoints of one substep of the drivers' time loop, fluxes_koren, then
 * a$D_consistent_fluxes, then update_solution (src/streamer_2d.f90:384-542,
 * src/streamer_3d.f90:396-615, src/streamer_cyl.f90:416-570,
 * afivo/src/m_aX_core.f90:818-930), with the mesh and the field held fixed.
 * No counterpart in the reference, which computes no derivatives. Run them
 * in the reverse order on the state the forward calls read (the fluxes of
 * the forward pass in the flux variable). adj[6] = cell variables le, li,
 * lE, lpho and face variables lF, lfld, which hold the adjoints of the
 * densities, |E|, the photoionization source, the electron flux and the
 * face field. Each call takes the adjoints of its outputs and adds those of
 * its inputs, taking the branches (upwind side, limiter case, flux clamp,
 * table row, dielectric rule) the forward pass takes on the same data; the
 * adjoints of the table rows accumulate in the handle
 * (a5g_table_adjoint).
 *   update_solution: le, li hold the adjoints of n_e', n_i' on the leaves
 *     and get those of n_e, n_i; adds to lF (start from zero), lE, lpho
 *     (with photoionization) and the alpha and eta rows; */
int a5g_update_solution_adjoint(a5g_tree *t, double dt, const int *adj);
/*   consistent fluxes: the adjoints of the coarse faces the forward call
 *     overwrote go to the fine faces it averaged, and are cleared; */
int a5g_consistent_fluxes_adjoint(a5g_tree *t, int lF);
/*   fluxes_koren: consumes lF on the leaves (clearing it) into le (cells,
 *     first-layer ghosts and the cells the second ghost layer is taken
 *     from), lE, lfld and the mobility and diffusion rows. */
int a5g_fluxes_koren_adjoint(a5g_tree *t, double dt, const int *adj);
/* The table-row adjoints accumulated so far, rows_cols(n_rows, n_cols)
 * column-major (NULL: not read); clear = 1 then sets them to zero. */
int a5g_table_adjoint(a5g_tree *t, int clear, double *rows_cols);

/* ---- reverse mode of the rest of a time step (frozen mesh) ------------ */

/* The adjoints of the remaining calls of the drivers' time step
 * (src/streamer_2d.f90:98-129 with compute_fld, :307-381). No counterpart
 * in the reference, which computes no derivatives. Each call takes the
 * adjoints of the values its forward call writes (and clears them: the
 * forward call overwrites those values) and adds those of the values it
 * reads. Calls whose forward call branches on its data (the limited
 * interpolation, the field norm) are linearised on the state in the primal
 * variables, which must be the state the forward call read.
 *   a$D_tree_copy_cc(from, to) (m_aX_utils.f90:432-453): lfrom += lto on
 *     whole boxes, lto = 0; */
int a5g_tree_copy_cc_adjoint(a5g_tree *t, int lfrom, int lto);
/*   average_dens (src/streamer_2d.f90:492-496): adj[4] = le, le_old, li,
 *     li_old; */
int a5g_average_dens_adjoint(a5g_tree *t, const int *adj);
/*   compute_fld's rhs = fac (n_e - n_i) on the leaves
 *     (src/streamer_2d.f90:318-329): adj[3] = lrhs, le, li; */
int a5g_space_charge_adjoint(a5g_tree *t, double fac, const int *adj);
/*   a$D_restrict_tree (m_aX_restrict.f90:54-142): parents' adjoints go to
 *     their children, coarsest first; */
int a5g_restrict_tree_adjoint(a5g_tree *t, int il);
/*   a$D_gc_tree(iv, rb, bc) (m_aX_gc.f90:28-46) with rb 1 a$D_gc_interp or
 *     2 a$D_gc_interp_lim and bc 1 a$D_gc_neumann or 2 a$D_gc_dirichlet:
 *     every face ghost's adjoint in il goes to the interior cells it was made
 *     from; the limiter's case is taken from iv; */
int a5g_gc_tree_adjoint(a5g_tree *t, int iv, int il, int rb, int bc);
/*   fld_from_pot (src/streamer_2d.f90:346-381, src/streamer_3d.f90:348-395):
 *     adj[3] = lE (cell), lfld (face), lphi (cell): the adjoints of the
 *     interior |E| and of the face field go to the potential (cells and
 *     ghosts), linearised at the field the forward call left. */
int a5g_fld_from_pot_adjoint(a5g_tree *t, const int *adj);
/* A device store of whole variables for checkpoint-recompute: op = 0
 * copies the cell variables ivs[n_cc] and face variables fvs[n_fc] of every
 * box slot into store slot `slot` (0..4096, grown as needed, freed with the
 * handle); op = 1 writes them back; op = 2 sets the variables to zero
 * (slot unused). A load needs the same variables and an unchanged mesh.
 * No counterpart in the reference, which keeps no states. */
int a5g_state_store(a5g_tree *t, int slot, int n_cc, const int *ivs, int n_fc, const int *fvs,
                    int op);

/* a5g_box_summaries without any device allocation, for callers that cannot
 * allocate (code under stream capture, serving loops, harnesses that forbid
 * hipMalloc); no counterpart in the reference. The per-box maxima go through
 * a fixed device scratch of the library (3 doubles for each of up to 131072
 * boxes, held by each GPU until the process ends), shared by every call on
 * that GPU: calls that overlap on one GPU race. Fails when max_id exceeds
 * 131072, and when the tree changed since its last upload (after
 * a5g_adjust_refinement or a5g_tidy_up, a call that may allocate, such as
 * a5g_box_summaries or any kernel call, has to come first). */
int a5g_box_summaries_noalloc(a5g_tree *t, double *out);
/* a$D_prolong1_to (afivo/src/m_aX_prolong.f90:99-191, add = .false.) for
 * the boxes ids[0..n-1] (prolong_to_new_boxes, src/streamer_2d.f90:592-617) */
int a5g_prolong1_to(a5g_tree *t, int n_ids, const int *ids, int iv);
/* cc(..., iv) = value on whole boxes (e.g. set_box_eps with eps = 1) */
int a5g_fill_cc(a5g_tree *t, int n_ids, const int *ids, int iv, double value);
/* a$D_gc_box for the boxes ids[0..n-1], level by level (rules as in
 * a5g_gc_tree) */
int a5g_gc_boxes(a5g_tree *t, int n_ids, const int *ids, int iv, int rb,
                 int bc);

/* max over leaf cells of |cc(iv_a) - cc(iv_b)| and of |cc(iv_b)| (the
 * error of one solution against another) */
int a5g_tree_maxdiff_cc(a5g_tree *t, int iv_a, int iv_b, double *maxdiff,
                        double *maxabs_b);

/* ---- initial conditions --------------------------------------------- */

/* cc(..., iv) = base + sum_k amp_k exp(-|x - c_k|^2 / w_k^2) at every cell
 * centre of every box, ghosts included; blobs[k] = (c_1..c_dim, w, amp).
 * The Gaussian seed of m_geom.f90:82-86 (GM_gaussian) written for blobs. */
int a5g_set_gaussians(a5g_tree *t, int iv, int n_blob, const double *blobs,
                      double base);
/* a5g_set_gaussians per run of a tree with instances: run g's boxes take
 * base[g] and blobs[g * n_blob .. (g + 1) * n_blob - 1] (centres in the
 * tree's coordinates). The seeds of set_init_cond (src/streamer_cyl.f90,
 * m_geom.f90:82-86) for every run of a batch in one call. */
int a5g_set_gaussians_instances(a5g_tree *t, int iv, int n_blob, const double *blobs,
                                const double *base);

/* ---- independent runs in one tree ------------------------------------ */

/* Instances: n_inst independent runs in one tree, the base boxes of run g
 * placed along the last dimension in [g * stride, (g + 1) * stride) with no
 * neighbour links between runs (physical boundaries). Every kernel then
 * treats all runs in one launch; the time step, the Koren clamp factor and
 * the boundary values of the rule 2V - c are taken per run
 * (a5g_set_instance_values). A box belongs to run
 * floor((r_min_last + dr / 4) / stride). n_inst = 0 returns to one run.
 * No reference counterpart: the reference runs one streamer per process;
 * this is the batched form of the cylindrical runs. */
int a5g_set_instances(a5g_tree *t, int n_inst, double stride);
/* Per-run values, one per instance: what = 0 the time step (the argument dt
 * of a5g_fluxes_koren and a5g_update_solution, which is then ignored),
 * what = nb in 1..2*dim the value V of the rule 2V - c on face nb (instead
 * of the value a5g_mg_setup / a5g_mg_set_bc_value gave). */
int a5g_set_instance_values(a5g_tree *t, int what, int n, const double *values);
/* Restrict every following call that works on the tree's level lists
 * (multigrid, ghost cells, restriction, fluxes, update, reductions) to the
 * runs with active[g] != 0; active NULL makes all runs active. Calls given
 * explicit box lists are not affected. */
int a5g_set_active_instances(a5g_tree *t, int n, const int *active);
/* Groups of runs for a5g_fas_fmg_groups: group[g] >= 0 is run g's group, -1
 * leaves it out; n = 0 clears the groups. No reference counterpart (one run
 * per process there). */
int a5g_set_instance_groups(a5g_tree *t, int n, const int *group);
/* mg$D_fas_fmg (afivo/src/m_aX_mg.f90:164-203) of every group of runs, each
 * up to its own finest level as if it were the only active group, the
 * groups side by side on the GPU (they share no boxes). Gives each run what
 * a5g_fas_fmg gives it with only its group active. */
int a5g_fas_fmg_groups(a5g_tree *t, int set_residual, int first_call);

/* ---- streamer diagnostics -------------------------------------------- */

/* a$D_get_loc (afivo/src/m_aX_utils.f90:244-281): the first level-1 box in
 * list order that holds point r (faces included), then down the children
 * holding it while the box's level is below max_lvl (max_lvl < 1: lvls_max,
 * the absent argument); ix = ceiling((r - r_min) / dr) per dimension
 * (a$D_cc_ix, :301-306). id = -1 and ix = -1 outside the level-1 boxes.
 * Host tree only: no GPU call. */
int a5g_get_loc(a5g_tree *t, const double *r, int max_lvl, int *id, int *ix);
/* a$D_reduction_loc (afivo/src/m_aX_utils.f90:499-555) with reduce_max for
 * the per-box procedures of get_streamer_properties
 * (src/streamer_cyl.f90:831-903), over the leaves on the GPU.
 *   kind 1 (box_maxfld_r, box_maxfld_z): per box the first cell in array
 *     order (x fastest) of the largest |f(ix) + f(ix + e_d)| of face variable
 *     var along dimension d (1-based); the box value is half of it.
 *   kind 2 (box_fld_maxr, box_fld_maxz): per box the largest cell-centre
 *     coordinate d over cells whose cc(var) exceeds thr, first such cell; 0
 *     and no cell where no coordinate above 0 qualifies.
 * Leaves are taken level by level in list order, and a box replaces the
 * running value (init_val at first) only where it is larger: the outcome of
 * the reference on one OpenMP thread. Writes the value, the box id and cell
 * (id = -1, ix = -1 when no box exceeded init_val). */
int a5g_reduction_loc(a5g_tree *t, int kind, int var, int d, double thr,
                      double init_val, double *val, int *id, int *ix);
/* get_cc_axis (src/streamer_cyl.f90:905-956), 2D: walking up the line x = 0
 * from y = 0 through the leaf boxes a$D_get_loc finds there, each box's
 * first column of cells gives nc points: y of the cell centre, cc(1, j) of
 * ixs_cc, 0.5 (fy(1, j) + fy(1, j+1)) of ixs_fc. Sets *n_points. With out
 * NULL or cap = 0 nothing else happens; otherwise writes
 * out(n_cc + n_fc + 1, *n_points) column-major, failing when *n_points
 * exceeds cap. */
int a5g_cc_axis(a5g_tree *t, int n_cc, const int *ixs_cc, int n_fc,
                const int *ixs_fc, int cap, double *out, int *n_points);

/* ---- output ----------------------------------------------------------- */

/* The interior cells of the leaves in the order of a$D_write_vtk
 * (afivo/src/m_aX_io.f90:20-209): leaves level by level in list order,
 * cells x fastest. Columns of *n_cells values: cc(ixs_cc(v)) for each cell
 * variable, then, when n_fc > 0, one column per dimension d holding
 * 0.5 * sum of f_d over the cell's two faces along d and over all face
 * variables ixs_fc, in array element order. Sets *n_cells. With out NULL or
 * cap = 0 nothing else happens; otherwise writes out(*n_cells, columns)
 * column-major, failing when *n_cells * columns exceeds cap. */
int a5g_leaf_cells(a5g_tree *t, int n_cc, const int *ixs_cc, int n_fc,
                   const int *ixs_fc, long long cap, double *out,
                   long long *n_cells);
/* a$D_write_vtk (afivo/src/m_aX_io.f90:20-209) with m_vtk
 * (afivo/src/m_vtk.f90): the leaves as a VTK XML unstructured grid of
 * pixels (2D) or voxels (3D) in <filename>.vtu, raw appended data. Cell
 * variables ixs_cc (NULL: all, n_cc then equals the number of cell
 * variables) named cc_names; with n_fc > 0 face variables ixs_fc whose
 * n_fc * dim names fc_names label the face-average columns as the reference
 * does (each column of dimension d holds the sum over all face variables).
 * The TIME value is written with the fewest digits that read back to the
 * same double (the reference writes it list-directed). */
int a5g_write_vtk(a5g_tree *t, const char *filename, int n_cc,
                  const char *const *cc_names, const int *ixs_cc, int n_fc,
                  const char *const *fc_names, const int *ixs_fc, int n_cycle,
                  double time);
/* Raw leaf-box dump, the fixture and viewer format of this library (no
 * counterpart in the reference, whose Silo output cannot be built here).
 * Little endian: "A5GLEAF1", int32 dim, n_cell, n_leaves, n_var, coord,
 * cycle, float64 time, int32 ivs[n_var]; per leaf (order of a5g_leaf_cells)
 * int32 id, lvl, ix[3], float64 dr, r_min[3]; then float64
 * data[n_var][n_leaves][n_cell^dim], cells x fastest, interior only. */
int a5g_dump_leaves(a5g_tree *t, const char *filename, int n_var, const int *ivs,
                    int cycle, double time);

/* ---- measurement ----------------------------------------------------- */

/* Start (on = 1) or stop timing every kernel launch with HIP events on the
 * handle's stream; enabling clears earlier records. */
int a5g_prof_enable(a5g_tree *t, int on);
/* Launch count, summed duration (ms) and algorithmic HBM bytes of the
 * recorded launches of one kernel class on level lvl (-999: all levels).
 * Classes: 0 Gauss-Seidel half-sweep, 1 ghost cells, 2 residual +
 * restriction, 3 coarse rhs, 4 correction, 5 copy, 6 other multigrid
 * kernels, 7 space charge, 8 field from potential, 9 Koren fluxes,
 * 10 consistent fluxes, 11 density update, 12 averaging, 13 restriction of
 * the densities, 14 photoionization rate. The consecutive half-sweeps of
 * one smoothing call are timed by one event pair, every other launch by its
 * own. */
int a5g_prof_read(a5g_tree *t, int kind, int lvl, int *n_launch, double *ms,
                  double *bytes);

#ifdef __cplusplus
}
#endif

#endif
