// Host-side quadtree/octree core of the a5g library.
//
// Bit-exact restatement of Afivo's integer tree logic (afivo/src/m_aX_core.f90):
// base creation with coarse sub-levels, refinement with 2:1 balance and
// free-id allocation order, neighbour search through parents, leaf/parent
// lists and Morton compaction (tidy_up). Box ids, levels, neighbour and
// child ids are therefore the ones the reference produces on the same flags.
//
// The tree owns only topology. Cell data lives on the GPU in the pool
// (a5g_pool.hip); every topology change is mirrored to device index maps.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace a5g {

constexpr int kNoBox = 0;
// multigrid operator class of a box (mg$D_set_box_tag, m_aX_mg.f90:673-712):
// set at the first multigrid call that meets the box, kept until the box is
// created again by a refinement
constexpr int kTagInit = -1, kTagNormal = 0, kTagCeps = 1, kTagVeps = 2;
constexpr int kRmRef = -1, kKpRef = 0, kDoRef = 1, kDerefine = -2, kRefine = 2;

struct Tree {
  int dim = 2, n_cell = 8, n_cc = 1, n_fc = 0, coord = 0;
  int lvls_max = 30, min_lvl = 1, max_lvl = 0, max_id = 0;
  int cap = 0;  // size of the reference's boxes(:) array
  bool ready = false;
  double dr_base = 1, r_base[3] = {0, 0, 0};
  int nch = 4, nnb = 4;

  // Per box, index = id (1-based); entry 0 is unused.
  std::vector<int> lvl, parent, in_use, ncell;
  std::vector<int> ix;        // [id*3 + d]
  std::vector<int> children;  // [id*8 + c]
  std::vector<int> nbs;       // [id*6 + nb]
  std::vector<double> dr, rmin;  // rmin[id*3 + d]
  std::vector<int> tag;          // kTag*

  // Per level, index lvl - min_lvl for lvl in min_lvl .. lvls_max+1
  std::vector<std::vector<int>> ids, leaves, parents;
  // ref_info: added/removed ids for levels 1 .. n (index lvl-1)
  std::vector<std::vector<int>> ri_add, ri_rm;

  // Box moves made by the last tidy_up: new id -> old id (1-based; 0 none).
  std::vector<int> last_moves;

  void init(int dim, int n_cell, int n_cc, int n_fc, double dr,
            const double* r_min, int coarsen_to, int lvls_max, int coord,
            int n_boxes);
  void set_base(int n, const int* ix_list, const int* nb_list);
  // flags[id-1] for id = 1..n_flags: the user refinement choice per box
  void adjust_refinement(int n_flags, const int* flags, int* n_add, int* n_rm);
  // returns true when boxes were renumbered (last_moves is then valid)
  bool tidy_up(double max_frac, double goal_frac, int n_clean_min, bool reorder);

  std::vector<int>& lv_ids(int l) { return ids[l - min_lvl]; }
  std::vector<int>& lv_leaves(int l) { return leaves[l - min_lvl]; }
  std::vector<int>& lv_parents(int l) { return parents[l - min_lvl]; }
  const std::vector<int>& lv_ids(int l) const { return ids[l - min_lvl]; }
  const std::vector<int>& lv_leaves(int l) const { return leaves[l - min_lvl]; }
  const std::vector<int>& lv_parents(int l) const { return parents[l - min_lvl]; }
  bool has_children(int id) const { return children[id * 8] != kNoBox; }
  // leaf box (levels up to max_lvl; max_lvl < 1: lvls_max) and cell ix
  // holding point r; -1 (and ix = -1) outside the level-1 boxes
  int get_loc(const double* r, int max_lvl, int* ix) const;
  int box_len(int id) const;

 private:
  void resize(int new_cap);
  void get_free_ids(int n, int* out);
  void set_leaves_parents(int l);
  int find_nb(int id, int nb) const;
  void set_nbs(int id);
  void add_children(int id, const int* c_ids);
  void remove_children(int id);
  void set_child_ids(int l);
  void consistent_ref_flags(std::vector<int>& f, int n_in, const int* in);
  void set_ref_info(const std::vector<int>& f, int n);
};

// Topology tables of m_aX_t.f90:41-132, as closed forms.
inline int ch_dix(int c, int d) { return ((c - 1) >> d) & 1; }
inline int ch_rev(int c, int d) { return ((c - 1) ^ (1 << d)) + 1; }
inline int nb_dim(int nb) { return (nb - 1) / 2; }
inline bool nb_low(int nb) { return ((nb - 1) & 1) == 0; }
inline int nb_hi01(int nb) { return (nb - 1) & 1; }
inline int nb_rev(int nb) { return ((nb - 1) ^ 1) + 1; }

}  // namespace a5g
