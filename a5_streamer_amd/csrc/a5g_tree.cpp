// Host-side tree core; see a5g_tree.hpp. Each routine names the reference
// routine whose integer behaviour it reproduces.
#include "a5g_tree.hpp"

#include <algorithm>
#include <cmath>
#include <stdexcept>

namespace a5g {

int Tree::box_len(int id) const {
  int n = ncell[id] + 2;
  return dim == 2 ? n * n : n * n * n;
}

void Tree::resize(int new_cap) {
  // a$D_resize_box_storage, m_aX_core.f90:448-464
  lvl.resize(new_cap + 1, 0);
  parent.resize(new_cap + 1, 0);
  in_use.resize(new_cap + 1, 0);
  ncell.resize(new_cap + 1, 0);
  ix.resize((size_t)(new_cap + 1) * 3, 0);
  children.resize((size_t)(new_cap + 1) * 8, 0);
  nbs.resize((size_t)(new_cap + 1) * 6, 0);
  dr.resize(new_cap + 1, 0.0);
  rmin.resize((size_t)(new_cap + 1) * 3, 0.0);
  tag.resize(new_cap + 1, kTagInit);
  cap = new_cap;
}

void Tree::init(int dim_, int n_cell_, int n_cc_, int n_fc_, double dr_,
                const double* r_min, int coarsen_to, int lvls_max_, int coord_,
                int n_boxes) {
  // a$D_init, m_aX_core.f90:47-116
  if (dim_ != 2 && dim_ != 3) throw std::invalid_argument("dim must be 2 or 3");
  if (n_cell_ < 2) throw std::invalid_argument("n_cell should be >= 2");
  if (n_cell_ & 1) throw std::invalid_argument("n_cell should be even");
  if (n_cc_ <= 0) throw std::invalid_argument("n_var_cell should be > 0");
  if (n_boxes <= 0) throw std::invalid_argument("n_boxes should be > 0");
  if (lvls_max_ <= 0) throw std::invalid_argument("lvls_max should be > 0");
  if (dim_ == 3 && coord_ == 1)
    throw std::invalid_argument("cannot have 3d cyl coords");
  *this = Tree();
  dim = dim_;
  n_cell = n_cell_;
  n_cc = n_cc_;
  n_fc = n_fc_;
  coord = coord_;
  lvls_max = lvls_max_;
  dr_base = dr_;
  for (int d = 0; d < dim; d++) r_base[d] = r_min ? r_min[d] : 0.0;
  nch = 1 << dim;
  nnb = 2 * dim;
  if (coarsen_to > 0) {
    min_lvl = 1 - (int)std::lround(std::log((double)n_cell / coarsen_to) / std::log(2.0));
    if ((1 << (1 - min_lvl)) * coarsen_to != n_cell)
      throw std::invalid_argument("cannot coarsen to given value");
  } else {
    min_lvl = 1;
  }
  int nl = lvls_max + 1 - min_lvl + 1;
  ids.assign(nl, {});
  leaves.assign(nl, {});
  parents.assign(nl, {});
  resize(n_boxes);
  max_id = 0;
  max_lvl = 0;
}

void Tree::get_free_ids(int n, int* out) {
  // get_free_ids, m_aX_core.f90:605-617
  for (int i = 0; i < n; i++) out[i] = max_id + 1 + i;
  max_id += n;
}

void Tree::set_leaves_parents(int l) {
  // set_leaves_parents, m_aX_core.f90:328-359
  auto& li = lv_ids(l);
  auto& lv = lv_leaves(l);
  auto& pa = lv_parents(l);
  lv.clear();
  pa.clear();
  for (int id : li) (has_children(id) ? pa : lv).push_back(id);
}

void Tree::set_base(int n, const int* ix_list, const int* nb_in) {
  // a$D_set_base, m_aX_core.f90:137-220
  for (int i = 0; i < n * dim; i++)
    if (ix_list[i] < 1) throw std::invalid_argument("need all ix_list > 0");
  if (max_id > 0) throw std::invalid_argument("this tree already has boxes");
  std::vector<int> nbl(nb_in, nb_in + (size_t)n * nnb);
  for (int i = 1; i <= n; i++)
    for (int nb = 1; nb <= nnb; nb++) {
      int nb_id = nbl[(i - 1) * nnb + nb - 1];
      if (nb_id > kNoBox && nb_id != i) nbl[(nb_id - 1) * nnb + nb_rev(nb) - 1] = i;
    }
  for (int v : nbl)
    if (v == kNoBox) throw std::invalid_argument("unresolved neighbors");
  if (n > cap) resize(n);
  for (int l = min_lvl; l <= 1; l++) {
    auto& li = lv_ids(l);
    li.resize(n);
    get_free_ids(n, li.data());
    if (max_id > cap) resize(max_id);
    int offset = li[0] - 1;
    for (int i = 1; i <= n; i++) {
      int id = li[i - 1];
      lvl[id] = l;
      for (int d = 0; d < dim; d++) {
        ix[id * 3 + d] = ix_list[(i - 1) * dim + d];
        rmin[id * 3 + d] = r_base[d] + (ix_list[(i - 1) * dim + d] - 1) * dr_base * n_cell;
      }
      dr[id] = dr_base * std::pow(0.5, l - 1);
      ncell[id] = n_cell / (1 << (1 - l));
      parent[id] = kNoBox;
      for (int c = 0; c < 8; c++) children[id * 8 + c] = kNoBox;
      for (int nb = 1; nb <= nnb; nb++) {
        int v = nbl[(i - 1) * nnb + nb - 1];
        nbs[id * 6 + nb - 1] = v > kNoBox ? v + offset : v;
      }
      in_use[id] = 1;
    }
    if (l == 1) lv_leaves(l) = li;
    else lv_parents(l) = li;
    if (l > min_lvl) {
      auto& pi = lv_ids(l - 1);
      for (int i = 0; i < n; i++) {
        children[pi[i] * 8] = li[i];
        parent[li[i]] = pi[i];
      }
    }
  }
  max_lvl = 1;
  ready = true;
}

int Tree::find_nb(int id, int nb) const {
  // find_nb_$Dd, m_aX_core.f90:427-445
  int p_id = parent[id];
  const int* x = &ix[id * 3];
  int c_ix = dim == 2 ? 4 - 2 * (x[1] & 1) - (x[0] & 1)
                      : 8 - 4 * (x[2] & 1) - 2 * (x[1] & 1) - (x[0] & 1);
  int d = nb_dim(nb);
  if ((ch_dix(c_ix, d) == 0) == nb_low(nb)) p_id = nbs[p_id * 6 + nb - 1];
  return children[p_id * 8 + ch_rev(c_ix, d) - 1];
}

void Tree::set_nbs(int id) {
  // set_nbs_$Dd, m_aX_core.f90:399-413
  for (int nb = 1; nb <= nnb; nb++) {
    if (nbs[id * 6 + nb - 1] == kNoBox) {
      int nb_id = find_nb(id, nb);
      if (nb_id > kNoBox) {
        nbs[id * 6 + nb - 1] = nb_id;
        nbs[nb_id * 6 + nb_rev(nb) - 1] = id;
      }
    }
  }
}

void Tree::add_children(int id, const int* c_ids) {
  // add_children, m_aX_core.f90:753-788
  for (int i = 1; i <= nch; i++) children[id * 8 + i - 1] = c_ids[i - 1];
  for (int i = 1; i <= nch; i++) {
    int c = c_ids[i - 1];
    for (int d = 0; d < dim; d++) {
      ix[c * 3 + d] = 2 * ix[id * 3 + d] - 1 + ch_dix(i, d);
      rmin[c * 3 + d] = rmin[id * 3 + d] + 0.5 * dr[id] * ch_dix(i, d) * ncell[id];
    }
    lvl[c] = lvl[id] + 1;
    parent[c] = id;
    for (int k = 0; k < 8; k++) children[c * 8 + k] = kNoBox;
    for (int k = 0; k < 6; k++) nbs[c * 6 + k] = kNoBox;
    ncell[c] = ncell[id];
    dr[c] = 0.5 * dr[id];
    in_use[c] = 1;
    tag[c] = kTagInit;  // m_aX_core.f90:769
  }
  for (int nb = 1; nb <= nnb; nb++) {
    int v = nbs[id * 6 + nb - 1];
    if (v < kNoBox) {
      // children adjacent to face nb (m_aX_t.f90 a$D_ch_adj_nb)
      for (int c = 1; c <= nch; c++)
        if (ch_dix(c, nb_dim(nb)) == nb_hi01(nb)) nbs[c_ids[c - 1] * 6 + nb - 1] = v;
    }
  }
}

void Tree::remove_children(int id) {
  // remove_children, m_aX_core.f90:729-750
  for (int ic = 0; ic < nch; ic++) {
    int c = children[id * 8 + ic];
    for (int nb = 1; nb <= nnb; nb++) {
      int nb_id = nbs[c * 6 + nb - 1];
      if (nb_id > kNoBox) nbs[nb_id * 6 + nb_rev(nb) - 1] = kNoBox;
    }
    in_use[c] = 0;
  }
  for (int ic = 0; ic < 8; ic++) children[id * 8 + ic] = kNoBox;
}

void Tree::set_child_ids(int l) {
  // set_child_ids, m_aX_core.f90:792-809
  auto& pa = lv_parents(l);
  auto& c = lv_ids(l + 1);
  c.resize((size_t)nch * pa.size());
  for (size_t i = 0; i < pa.size(); i++)
    for (int k = 0; k < nch; k++) c[i * nch + k] = children[pa[i] * 8 + k];
}

void Tree::consistent_ref_flags(std::vector<int>& f, int n_in, const int* in) {
  // consistent_ref_flags, m_aX_core.f90:624-726
  int n = max_id;
  f.assign(n + 1, -2147483647);
  for (int l = 1; l <= max_lvl; l++)
    for (int id : lv_ids(l)) f[id] = id <= n_in ? in[id - 1] : kKpRef;
  for (int i = 1; i <= n; i++) {
    if (f[i] > kDoRef) f[i] = kKpRef;
    if (f[i] < kRmRef) f[i] = kKpRef;
  }
  for (int id : lv_ids(lvls_max))
    if (f[id] == kDoRef) f[id] = kKpRef;
  for (int l = max_lvl; l >= 1; l--) {
    for (int id : lv_leaves(l)) {
      if (f[id] > kKpRef) {
        f[id] = kRefine;
        for (int nb = 1; nb <= nnb; nb++)
          if (nbs[id * 6 + nb - 1] == kNoBox) f[nbs[parent[id] * 6 + nb - 1]] = kRefine;
      } else if (f[id] == kRmRef) {
        for (int nb = 1; nb <= nnb; nb++) {
          int nb_id = nbs[id * 6 + nb - 1];
          if (nb_id > kNoBox && (has_children(nb_id) || f[nb_id] > kKpRef)) {
            f[id] = kKpRef;
            break;
          }
        }
      }
    }
  }
  for (int l = max_lvl - 1; l >= 1; l--) {
    for (int id : lv_parents(l)) {
      bool all_rm = true;
      for (int c = 0; c < nch; c++)
        if (f[children[id * 8 + c]] != kRmRef) all_rm = false;
      f[id] = (all_rm && f[id] <= kKpRef) ? kDerefine : kKpRef;
    }
  }
}

void Tree::set_ref_info(const std::vector<int>& f, int n) {
  // set_ref_info, m_aX_core.f90:540-601
  int nl = max_lvl + 1;
  ri_add.assign(nl, {});
  ri_rm.assign(nl, {});
  for (int id = 1; id <= n; id++) {
    int l = lvl[id];
    if (l < 1 || l + 1 > nl) continue;
    if (f[id] == kRefine)
      for (int c = 0; c < nch; c++) ri_add[l].push_back(children[id * 8 + c]);
    else if (f[id] == kDerefine)
      for (int c = 0; c < nch; c++) ri_rm[l].push_back(children[id * 8 + c]);
  }
}

void Tree::adjust_refinement(int n_flags, const int* flags, int* n_add, int* n_rm) {
  // a$D_adjust_refinement, m_aX_core.f90:473-537
  if (!ready) throw std::logic_error("Tree not ready");
  int max_id_prev = max_id;
  std::vector<int> f;
  consistent_ref_flags(f, n_flags, flags);
  int nref = 0, nder = 0;
  for (int i = 1; i <= max_id_prev; i++) {
    nref += f[i] == kRefine;
    nder += f[i] == kDerefine;
  }
  int req = max_id_prev + nch * nref;
  if (req > cap) resize(req);
  int l;
  for (l = 1; l <= lvls_max - 1; l++) {
    auto li = lv_ids(l);  // copy: children lists change below
    for (int id : li) {
      if (id > max_id_prev) continue;
      if (f[id] == kRefine) {
        int c_ids[8];
        get_free_ids(nch, c_ids);
        add_children(id, c_ids);
      } else if (f[id] == kDerefine) {
        remove_children(id);
      }
    }
    set_leaves_parents(l);
    set_child_ids(l);
    for (int id : lv_ids(l + 1))
      if (id > max_id_prev) set_nbs(id);
    if (lv_ids(l + 1).empty()) break;
  }
  max_lvl = l;
  set_leaves_parents(max_lvl + 1);
  set_ref_info(f, max_id_prev);
  *n_add = nch * nref;
  *n_rm = nch * nder;
}

static uint64_t morton_key(const int* x, int dim) {
  // morton_from_ix2 / morton_from_ix3, m_morton.f90:33-59
  uint64_t m = 0;
  for (int b = 0; b < 21; b++)
    for (int d = 0; d < dim; d++)
      m |= (((uint64_t)(x[d] - 1) >> b) & 1ull) << (b * dim + d);
  return m;
}

bool Tree::tidy_up(double max_frac, double goal_frac, int n_clean_min, bool reorder) {
  // a$D_tidy_up, m_aX_core.f90:224-325
  if (!ready) throw std::logic_error("Tree not ready");
  if (goal_frac > max_frac) throw std::invalid_argument("need goal_frac_used < max_frac_used");
  if (max_frac > 1.0) throw std::invalid_argument("need max_frac_used < 1");
  if (n_clean_min < 1) throw std::invalid_argument("need n_clean_min > 0");
  int n_used = 0;
  for (int i = 1; i <= max_id; i++) n_used += in_use[i];
  int old_size = cap;
  double frac = n_used / (double)old_size;
  int n_clean = (int)std::lround((goal_frac - frac) * old_size);
  int new_size = old_size;
  if (!reorder && (max_id > old_size * max_frac || (frac < goal_frac && n_clean > n_clean_min)))
    new_size = std::max(1, (int)std::lround(n_used / goal_frac));
  last_moves.clear();
  if (new_size == old_size && !reorder) return false;

  std::vector<int> map(max_id + 1, 0), order(n_used + 1, 0);
  int n_stored = 0;
  for (int l = min_lvl; l <= max_lvl; l++) {
    auto& li = lv_ids(l);
    std::vector<std::pair<uint64_t, int>> k(li.size());
    for (size_t i = 0; i < li.size(); i++) k[i] = {morton_key(&ix[li[i] * 3], dim), (int)i};
    std::stable_sort(k.begin(), k.end(),
                     [](const auto& a, const auto& b) { return a.first < b.first; });
    std::vector<int> sorted(li.size());
    for (size_t i = 0; i < li.size(); i++) sorted[i] = li[k[i].second];
    for (size_t i = 0; i < li.size(); i++) {
      order[n_stored + i + 1] = sorted[i];
      map[sorted[i]] = n_stored + (int)i + 1;
      li[i] = n_stored + (int)i + 1;
    }
    n_stored += (int)li.size();
  }
  Tree old = *this;
  int ncap = reorder ? std::max(old_size, n_used) : std::max(new_size, n_used);
  resize(0);
  resize(ncap);
  for (int n = 1; n <= n_used; n++) {
    int o = order[n];
    lvl[n] = old.lvl[o];
    for (int d = 0; d < 3; d++) {
      ix[n * 3 + d] = old.ix[o * 3 + d];
      rmin[n * 3 + d] = old.rmin[o * 3 + d];
    }
    parent[n] = map[old.parent[o]];
    for (int c = 0; c < 8; c++) children[n * 8 + c] = map[old.children[o * 8 + c]];
    for (int b = 0; b < 6; b++) {
      int v = old.nbs[o * 6 + b];
      nbs[n * 6 + b] = v > kNoBox ? map[v] : v;
    }
    in_use[n] = 1;
    ncell[n] = old.ncell[o];
    dr[n] = old.dr[o];
    tag[n] = old.tag[o];
  }
  for (int l = min_lvl; l <= max_lvl; l++) set_leaves_parents(l);
  max_id = n_used;
  last_moves = order;
  return true;
}

// a$D_get_loc (afivo/src/m_aX_utils.f90:244-281): the first level-1 box in
// list order holding r (a$D_r_inside, :203-216, faces included), then down
// through child_that_contains (:284-298) while the box has children and its
// level is below max_lvl; the cell is a$D_cc_ix (:301-306).
int Tree::get_loc(const double* r, int max_lvl, int* ix) const {
  const int lmax = max_lvl < 1 ? lvls_max : max_lvl;
  int id = -1;
  for (int c : lv_ids(1)) {
    bool in = true;
    for (int d = 0; d < dim; d++) {
      const double lo = rmin[c * 3 + d], hi = lo + dr[c] * ncell[c];
      if (!(r[d] >= lo && r[d] <= hi)) in = false;
    }
    if (in) {
      id = c;
      break;
    }
  }
  if (id < 0) {
    for (int d = 0; d < dim; d++) ix[d] = -1;
    return -1;
  }
  while (lvl[id] < lmax && has_children(id)) {
    int i_ch = 1;
    for (int d = 0; d < dim; d++) {
      const double cntr = rmin[id * 3 + d] + dr[id] * (ncell[id] >> 1);
      if (r[d] > cntr) i_ch += 1 << d;
    }
    id = children[id * 8 + i_ch - 1];
  }
  for (int d = 0; d < dim; d++) ix[d] = (int)std::ceil((r[d] - rmin[id * 3 + d]) / dr[id]);
  return id;
}

}  // namespace a5g
