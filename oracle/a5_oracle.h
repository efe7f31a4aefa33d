/* CPU restatement of the a5_streamer hot path (test infrastructure only).
 *
 * This is the oracle that the HIP library is checked against. It restates,
 * in plain C with one scalar loop per reference loop and the reference's
 * operand order, the Afivo tree core (afivo/src/m_aX_core.f90), ghost cells
 * (m_aX_gc.f90), prolongation (m_aX_prolong.f90), restriction
 * (m_aX_restrict.f90), multigrid (m_aX_mg.f90) and the streamer per-box
 * physics (src/streamer_{2d,3d,cyl}.f90, src/m_streamer.f90,
 * fosito/m_lookup_table.f90). Only tests, the smoke check and the CPU
 * baseline leg of bench.py load it; the product never does.
 *
 * How each part is pinned (the tests named run on the CPU):
 *  - tree core, ghost cells, restriction, prolongation, multigrid, Koren
 *    limiter, lookup table, NORM2, Morton keys, point location and the
 *    located reduction: bit for bit against the reference compiled from its
 *    own sources (oracle/_ref; tests/test_oracle_ref.py,
 *    tests/test_oracle_diag.py);
 *  - the streamer box procedures (space charge, field, fluxes, update,
 *    averaging, photoionization rate, refinement maxima): bit for bit
 *    against the reference drivers' own internal procedures and main loop,
 *    built from src/streamer_{2d,3d,cyl}.f90 into oracle/_ref as modules
 *    (tests/test_oracle_ref_drivers.py), on random data (every branch of
 *    the limiter and of the flux clamp) and over whole runs; also
 *    conservation (tests/test_driver_host.py);
 *  - the Helmholtz operator and photoionization (no reference
 *    counterpart): smoother closed form, direct sparse solve, manufactured
 *    solution and far field (tests/test_oracle_helmholtz.py);
 *  - the axis walk and the output arrays: the reference's own a2_get_loc
 *    and m_vtk, geometry and brute force (tests/test_oracle_diag.py,
 *    tests/test_oracle_output.py);
 *  - the reverse mode of the transport substep (no reference
 *    counterpart): central differences of the restated forward substep
 *    (itself pinned above) along 20 seeded directions, at two step sizes
 *    whose error bounds follow the limiter's and the table's kinks
 *    (tests/test_oracle_adjoint.py);
 *  - the reverse mode of the rest of a time step and of the FAS-FMG solve
 *    (copy, average, space charge, restriction, ghost cells, field from
 *    potential, o_fas_fmg_adj; no reference counterpart): the dot-product
 *    identity <w, F(x)> = <J^T w, x> against the forward routines above,
 *    which are positively homogeneous in what they read (affine in the
 *    boundary values for the FMG), a central difference off the ray for the
 *    field, and the gradient of whole windows of steps against central
 *    differences of the forward pass (tests/test_oracle_step_adjoint.py).
 * No function is left without such a check.
 *
 * Conventions follow the reference: box ids are 1-based, 0 = no box,
 * negative neighbour ids are physical boundaries, variable indices are
 * 1-based. A box's cell-centred block is cc(0:nc+1, 0:nc+1[, 0:nc+1]) with
 * i fastest, stored for every variable of a box contiguously.
 */
#ifndef A5_ORACLE_H
#define A5_ORACLE_H

#include <stddef.h>

typedef struct otree otree;

otree *o_tree_new(int dim, int n_cell, int n_cc, int n_fc, double dr,
                  const double *r_min, int coarsen_to, int lvls_max, int coord,
                  int n_boxes);
void o_tree_free(otree *t);
void o_set_base(otree *t, int n, const int *ix_list, const int *nb_list);
void o_adjust(otree *t, int n_flags, const int *flags, int *n_add, int *n_rm);
void o_tidy_up(otree *t, double max_frac, double goal_frac, int n_clean_min,
               int reorder);
void o_info(const otree *t, int *out8);
void o_get_boxes(const otree *t, int n, int *lvl, int *ix, int *parent,
                 int *children, int *neighbors, int *in_use, int *n_cell,
                 double *dr, double *r_min);
void o_level_counts(const otree *t, int lvl, int *counts3);
void o_level_lists(const otree *t, int lvl, int *ids, int *leaves,
                   int *parents);
int o_ref_info_lvl(const otree *t, int lvl, int *add, int *rm, int *n_add,
                   int *n_rm);

void o_set_cc(otree *t, int id, int iv, const double *data);
void o_get_cc(const otree *t, int id, int iv, double *data);
void o_set_fc(otree *t, int id, int dim, int ifc, const double *data);
void o_get_fc(const otree *t, int id, int dim, int ifc, double *data);

/* Multigrid options (mg$D_t, m_aX_mg.f90:20-52). bc_kind per physical face:
 * 0 = Neumann (copy), 1 = Dirichlet zero (-c), 2 = value (2V - c). */
void o_mg_setup(otree *t, int i_phi, int i_tmp, int i_rhs, int i_eps,
                int n_down, int n_up, int n_base, const int *bc_kind,
                const double *bc_val, int auto_ops);
void o_fas_fmg(otree *t, int set_residual, int first_call);
/* Helmholtz operator lpl(phi) - kappa2 phi (0: the Laplacian) */
void o_mg_set_bc_value(otree *t, int nb, double value);
void o_mg_set_helmholtz(otree *t, double kappa2);
void o_fas_vcycle(otree *t, int max_lvl, int set_residual);
void o_box_gsrb(otree *t, int id, int cntr);
void o_box_op(otree *t, int id, int i_out);
void o_box_rstr(otree *t, int id, int iv);
void o_box_corr(otree *t, int id);
/* rb: 0 = mg sides_rb, 1 = gc_interp, 2 = gc_interp_lim;
 * bc: 0 = mg boundary rule, 1 = Neumann, 2 = Dirichlet zero */
void o_gc_box(otree *t, int id, int iv, int rb, int bc);
void o_gc_tree(otree *t, int iv, int rb, int bc);
void o_gc2_box(otree *t, int id, int iv, double *out);
void o_restrict_tree(otree *t, int iv);
void o_consistent_fluxes(otree *t, int n, const int *f_ixs);
void o_prolong1_to(otree *t, int id, int iv, int add);
void o_prolong0_to(otree *t, int id, int iv);
void o_tree_copy_cc(otree *t, int iv_from, int iv_to);
void o_tree_minmax_cc(const otree *t, int iv, double *vmin, double *vmax);
double o_min_dr(const otree *t);
/* a$D_get_loc: leaf box and cell holding point r (max_lvl < 1: no limit) */
void o_get_loc(const otree *t, const double *r, int max_lvl, int *id, int *ix);
/* a$D_reduction_loc with max for the box procedures of
   get_streamer_properties: kind 1 = 0.5|f(ix) + f(ix+e_d)| of face
   variable var, kind 2 = cell-centre coordinate d where cc(var) > thr */
void o_reduction_loc(const otree *t, int kind, int var, int d, double thr, double init_val,
                     double *val, int *id, int *ix);
/* a$D_write_vtk's cell data (returns n_cells; out may be NULL) and its
   node coordinates, connectivity, offsets and cell types */
long long o_leaf_cells(const otree *t, int n_cc, const int *ixs_cc, int n_fc,
                       const int *ixs_fc, double *out);
void o_vtk_geometry(const otree *t, double *coords, int *connects, int *offsets, int *types);
/* get_cc_axis (2D): returns the number of points */
int o_cc_axis(const otree *t, int n_cc, const int *ixs_cc, int n_fc, const int *ixs_fc,
              double *out, int cap);
/* max|rhs - A phi| over leaves (uses i_tmp as scratch, like residual_box) */
double o_max_residual(otree *t);

/* Koren-limited flux limiter, src/m_streamer.f90:339-362 */
double o_koren_mlim(double a, double b);

/* ---- streamer physics (src/streamer_{2d,3d,cyl}.f90, fosito) ---------- */

/* Uniform lookup table, fosito/m_lookup_table.f90 (LT_table_t). */
typedef struct otable otable;
otable *o_lt_create(double x_min, double x_max, int n_rows, int n_cols);
void o_lt_free(otable *t);
/* LT_set_col: linear resampling of (xx, yy) onto the table grid */
void o_lt_set_col(otable *t, int col, int n, const double *xx, const double *yy);
double o_lt_get_col(const otable *t, int col, double x);
/* rows_cols(n_rows, n_cols), column-major, and x_min / inv_dx */
void o_lt_data(const otable *t, double *rows_cols, double *x_min, double *inv_dx);

/* overwrite the table's rows_cols(n_rows, n_cols) (column-major) */
void o_lt_set_rows(otable *t, const double *rc);

/* norm2 of the flang runtime (scaled sum of squares) */
double o_norm2(int n, const double *x);

/* Variable indices of the streamer models (src/m_streamer.f90:16-43). */
typedef struct {
  int elec, pion, elec_old, pion_old, phi, fld, rhs, pho, eps;
  int f_elec, f_fld;
  int t_mob, t_dif, t_alpha, t_eta; /* table columns */
  int dielectric_rule;              /* 2D/3D update: n_e -> n_i in eps > 1 */
  int photoi;                       /* add i_pho to the source term */
} o_vars;

/* photon production rate (set_photoi_rate) into v->pho on leaves */
void o_photoi_rate(otree *t, const o_vars *v, const otable *tbl, double coeff);
void o_axpy_leaves(otree *t, int iin, int iout, double s, int add);
/* Helmholtz photoionization source into v->pho (three-term model) */
int o_photoi_helmholtz(otree *t, const o_vars *v, const otable *tbl, double coeff,
                       int n_terms, const double *amp, const double *kappa, int i_phi,
                       int i_rhs, int i_tmp, double tol, int max_cycles, int *cycles);
/* the same with term j in variables of its own (solution i_terms[j], rhs
 * i_rhs[j], scratch i_tmp[j]); warm: V-cycles from the solution an earlier
 * call left in i_terms[j], instead of an FMG from zero */
int o_photoi_helmholtz_terms(otree *t, const o_vars *v, const otable *tbl, double coeff,
                             int n_terms, const double *amp, const double *kappa,
                             const int *i_terms, const int *i_rhs, const int *i_tmp, int warm,
                             double tol, int max_cycles, int *cycles);
/* compute_fld's right-hand side on leaves, whole boxes: rhs = fac (ne - ni) */
void o_space_charge(otree *t, const o_vars *v, double fac);
/* fld_from_pot on every box */
void o_fld_from_pot(otree *t, const o_vars *v);
/* fluxes_koren on leaves */
void o_fluxes_koren(otree *t, const o_vars *v, const otable *tbl, double dt);
/* update_solution on leaves */
void o_update_solution(otree *t, const o_vars *v, const otable *tbl, double dt);
/* average_dens on every box, whole boxes */
void o_average_dens(otree *t, const o_vars *v);
/* Reverse mode of one transport substep on a frozen mesh (no reference
 * counterpart). Adjoint variables: cell variables le, li (n_e, n_i), lE
 * (|E|), lpho (photoionization source) and face variables lF (electron
 * flux), lfld (field). Each routine takes the adjoints of its outputs and
 * adds those of its inputs; tgrad (n_cols x n_rows, column-major, or NULL)
 * accumulates the adjoints of the table rows.
 *  update_solution_adj: in le, li the adjoints of n_e', n_i' on leaf
 *    interiors; out: those of n_e, n_i, plus lF, lE, lpho, tgrad(alpha, eta);
 *  consistent_fluxes_adj: moves lF from the coarse faces the forward pass
 *    overwrote to the fine faces it averaged, clearing the coarse ones;
 *  fluxes_koren_adj: consumes lF on leaves (clearing it) into le (cells and
 *    first- and second-layer ghost sources), lE, lfld, tgrad(mobility,
 *    diffusion). */
typedef struct {
  int le, li, lE, lpho; /* cell variables */
  int lF, lfld;         /* face variables */
} o_adj;
void o_update_solution_adj(otree *t, const o_vars *v, const otable *tbl, double dt,
                           const o_adj *a, double *tgrad);
void o_consistent_fluxes_adj(otree *t, int lF);
void o_fluxes_koren_adj(otree *t, const o_vars *v, const otable *tbl, double dt,
                        const o_adj *a, double *tgrad);
/* Reverse mode of the rest of a time step (no reference counterpart), in
 * the forms of the library's a5g_*_adjoint calls: each takes the adjoints
 * of what its forward routine writes (clearing them) and adds those of
 * what it reads. gc_tree_adj covers rb 1 (interp) and 2 (interp_lim, the
 * limiter's case taken from the primal iv) with bc 1 (Neumann) and 2
 * (Dirichlet); fld_from_pot_adj is linearised at the field the forward
 * routine left. */
void o_tree_copy_cc_adj(otree *t, int lfrom, int lto);
void o_average_dens_adj(otree *t, int le, int leo, int li, int lio);
void o_space_charge_adj(otree *t, double fac, int lr, int le, int li);
void o_restrict_tree_adj(otree *t, int il);
void o_gc_tree_adj(otree *t, int iv, int il, int rb, int bc);
void o_fld_from_pot_adj(otree *t, const o_vars *v, int lE, int lfld, int lphi);
/* o_fas_fmg(t, 0, 0) transposed: adj = lp, lt, lr, lw (the adjoints of the
 * potential, tmp and rhs it leaves, in; of those it starts from, out; lw
 * work); vbar[nb-1] the adjoint of face nb's value V. The operators are
 * the forward solve's per box (Laplacian with its Helmholtz term, or the
 * permittivity forms on tagged boxes). Returns 0. */
int o_fas_fmg_adj(otree *t, const int *adj, double *vbar);
/* o_fas_vcycle(t, max_lvl, 0) transposed; adj and vbar as above */
int o_fas_vcycle_adj(otree *t, int max_lvl, const int *adj, double *vbar);
/* whole variables of box slots 0..max_id to (load 0) or from (load 1) buf,
 * which holds o_store_size doubles */
void o_store_vars(otree *t, int n_cc, const int *ivs, int n_fc, const int *fvs, double *buf,
                  int load);
size_t o_store_size(const otree *t, int n_cc, int n_fc);
/* per-box maxima for the refinement criterion: out[id-1] = (dr^2 max|rhs|,
 * max fld, max elec) over the interior, boxes of levels >= 1 */
void o_box_summaries(const otree *t, const o_vars *v, double *out);

#endif
