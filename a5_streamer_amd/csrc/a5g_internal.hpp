// Internal entry points shared by the a5g translation units.
#pragma once

#include <vector>

#include "a5g_pool.hpp"

namespace a5g {

void gc_level(Pool& P, int lvl, int iv, int rb, int bc);
void gc_tree(Pool& P, int iv, int rb, int bc);
void restrict_tree(Pool& P, int iv);
void tree_copy_cc(Pool& P, int from, int to);
void minmax_boxes(Pool& P, int iv, bool absval, bool whole, bool leaves_only,
                  double* lo, double* hi);
void mg_fmg(Pool& P, bool set_residual, bool first_call);
void mg_fmg_groups(Pool& P, bool set_residual, bool first_call);
void mg_vcycle(Pool& P, int max_lvl, bool set_residual);
void autotune(Pool& P, int reps, std::vector<double>& rec);
void fmg_adjoint(Pool& P, const int* adj, double* vbar_host);
void vcycle_adjoint(Pool& P, int max_lvl, const int* adj, double* vbar_host);
// stores the ghost cells a multigrid call left to the next reader
void flush_ghosts(Pool& P);
void mg_residual_norms(Pool& P, double* max_res, double* max_rhs);
void maxdiff_leaves(Pool& P, int ia, int ib, double* maxdiff, double* maxabs_b);
void mg_solve_mixed(Pool& P, double tol, int max_outer, int n_inner, int* outer_done,
                    double* rel_res);

void set_gaussians(Pool& P, int iv, int nblob, const double* blobs, double base,
                   const double* bases = nullptr);
void gc_list(Pool& P, int lvl, const int* d_ids, int n, int iv, int rb, int bc);

void set_table(Pool& P, int n_rows, int n_cols, double x_min, double inv_dx, const double* rc);
void space_charge(Pool& P, const StVars& v, double fac);
void fld_from_pot(Pool& P, const StVars& v);
void fluxes_koren(Pool& P, const StVars& v, double dt);
void consistent_fluxes(Pool& P, int f_elec);
void update_solution(Pool& P, const StVars& v, double dt);
void average_dens(Pool& P, const StVars& v);
void box_summaries(Pool& P, const StVars& v, double* out_host);
void set_ref_rules(Pool& P, int n_seed, const double* seeds, int n_reg, const double* regions,
                   double window, double crv_refine, int n_last);
void ref_flags(Pool& P, const StVars& v, const double* times, int* flags_host);
void instance_maxima(Pool& P, int iv, double* vmax, int* lvl_max);
// reverse mode of the transport substep (a5g_*_adjoint)
void update_solution_adj(Pool& P, const StVars& v, double dt, const AdjVars& av);
void consistent_fluxes_adj(Pool& P, int lF);
void fluxes_koren_adj(Pool& P, const StVars& v, double dt, const AdjVars& av);
void table_adjoint(Pool& P, bool clear, double* out_host);
// reverse mode of the rest of the step (a5g_adjoint.hip)
void copy_cc_adj(Pool& P, int lfrom, int lto);
void average_dens_adj(Pool& P, int le, int leo, int li, int lio);
void space_charge_adj(Pool& P, double fac, int lr, int le, int li);
void restrict_tree_adj(Pool& P, int il);
void gc_tree_adj(Pool& P, int iv, int il, int rb, int bc);
void fld_from_pot_adj(Pool& P, const StVars& v, int lE, int lfld, int lphi);
void state_store(Pool& P, int slot, int n_cc, const int* ivs, int n_fc, const int* fvs, int op);
void store_release(Pool& P);
// the same with the library's fixed device scratch instead of allocations
// (a5g_box_summaries_noalloc); noalloc_init looks the scratch up for the
// current device
constexpr int kNoAllocMaxBoxes = 131072;
void noalloc_init();
void box_summaries_noalloc(Pool& P, const StVars& v, double* out_host);
void photoi_rate(Pool& P, const StVars& v, double coeff);
void axpy_leaves(Pool& P, int iin, int iout, double s, bool add);
void photoi_helmholtz(Pool& P, const StVars& v, double coeff, int n_terms, const double* amp,
                      const double* kappa, int i_phi, int i_rhs, int i_tmp, double tol,
                      int max_cycles, int* cycles_done);
void photoi_helmholtz_terms(Pool& P, const StVars& v, double coeff, int n_terms,
                            const double* amp, const double* kappa, const int* i_terms,
                            const int* i_rhs, const int* i_tmp, bool warm, double tol,
                            int max_cycles, int* cycles_done);
void prolong1_boxes(Pool& P, int n_ids, const int* ids, int iv);
void fill_boxes(Pool& P, int n_ids, const int* ids, int iv, double value);
void gc_boxes(Pool& P, int n_ids, const int* ids, int iv, int rb, int bc);
void reduction_loc(Pool& P, int kind, int var, int d, double thr, double init_val, double* val,
                   int* id, int* ix);
int cc_axis(Pool& P, int n_cc, const int* ixs_cc, int n_fc, const int* ixs_fc, int cap,
            double* out);
long long leaf_cells(Pool& P, int n_cc, const int* ixs_cc, int n_fc, const int* ixs_fc,
                     std::vector<double>& host);
void write_vtk(Pool& P, const char* filename, int n_cc, const char* const* cc_names,
               const int* ixs_cc, int n_fc, const char* const* fc_names, const int* ixs_fc,
               int n_cycle, double time);
void dump_leaves(Pool& P, const char* filename, int n_var, const int* ivs, int cycle,
                 double time);

}  // namespace a5g
