// Internal entry points shared by the a5g translation units.
#pragma once

#include "a5g_pool.hpp"

namespace a5g {

void gc_level(Pool& P, int lvl, int iv, int rb, int bc);
void gc_tree(Pool& P, int iv, int rb, int bc);
void restrict_tree(Pool& P, int iv);
void tree_copy_cc(Pool& P, int from, int to);
void minmax_boxes(Pool& P, int iv, bool absval, bool whole, bool leaves_only,
                  double* lo, double* hi);
void mg_fmg(Pool& P, bool set_residual, bool first_call);
void mg_vcycle(Pool& P, int max_lvl, bool set_residual);
void mg_residual_norms(Pool& P, double* max_res, double* max_rhs);

void set_gaussians(Pool& P, int iv, int nblob, const double* blobs, double base);

}  // namespace a5g
