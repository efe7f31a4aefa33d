// GPU-resident box pool of one a5g tree: cell/face data in HBM, device copies
// of the topology and level lists, and the multigrid options.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>

#include <string>
#include <vector>

#include "a5g_tree.hpp"

namespace a5g {

// Everything a kernel needs to address boxes; passed by value.
struct Dev {
  double* cc = nullptr;  // [n_cc][cap+1][bs]; box id at slot id
  double* fc = nullptr;  // [n_fc][cap+1][dim][fs]
  float* cc32 = nullptr;  // FP32 scratch variables [n32][cap+1][bs]
  size_t vs32 = 0;
  size_t vs = 0;         // per-variable stride of cc ((cap+1) * bs)
  size_t fvs = 0;        // per-variable stride of fc ((cap+1) * dim * fs)
  int bs = 0;            // box stride: (n_cell+2)^dim
  int fs = 0;            // face-array stride: (n_cell+1) * n_cell^(dim-1)
  int coord = 0;
  const int* lvl = nullptr;
  const int* ix = nullptr;        // [id*3 + d]
  const int* parent = nullptr;
  const int* children = nullptr;  // [id*8 + c]
  const int* nbs = nullptr;       // [id*6 + nb-1]
  const int* ncell = nullptr;
  const double* dr = nullptr;
  const double* rmin = nullptr;   // [id*3 + d]
  int bc_kind[6] = {0, 0, 0, 0, 0, 0};
  double bc_val[6] = {0, 0, 0, 0, 0, 0};
  // multigrid operator class per box (kTag*), set for the multigrid passes
  // when the automatic operators and a permittivity variable are in use
  const int* tag = nullptr;
  int i_eps = -1;
  // Helmholtz term of the multigrid operator: lpl(phi) - helm * phi (0: the
  // Laplacian)
  double helm = 0;
  // Independent runs in one tree (instances), stacked along the last
  // dimension one every inst_stride (n_inst = 0: a single run, the scalar
  // values apply). inst_val holds kInstVals values per instance.
  int n_inst = 0;
  double inst_stride = 0;
  const double* inst_val = nullptr;
};

// Per-instance values: the time step, the Koren flux clamp factor
// -0.8 eps0 / (e dt) and the value V of the rule 2V - c on faces 1..6.
constexpr int kInstVals = 8;
constexpr int kInstDt = 0, kInstKorenFac = 1, kInstBc = 2;

// Physical-face rules of the potential (bc_kind): copy, -c, 2V - c on the
// face, and 2V - c over the whole ghost row of a 2D y face, corners included
constexpr int kBcCopy = 0, kBcNeg = 1, kBc2V = 2, kBcRow2V = 3;

constexpr int kNoDefer = -1000000;

// Per-launch timing with HIP events on the pool's stream (kernel classes
// below), used by the benchmark to measure kernel durations in place.
enum ProfKind { PK_GSRB = 0, PK_GC, PK_RESID_RSTR, PK_COARSE_RHS, PK_CORR, PK_COPY, PK_OTHER,
                // the streamer's box procedures
                PK_SPACE_CHARGE, PK_FIELD, PK_KOREN, PK_FLUX_FIX, PK_UPDATE, PK_AVERAGE,
                PK_RESTRICT, PK_PHOTOI_RATE, PK_COUNT };

struct ProfRec {
  int kind, lvl;
  double bytes;  // algorithmic bytes of the launches
  hipEvent_t a, b;
  int count;     // launches between the two events
};

// Variable indices of the streamer models (src/m_streamer.f90:16-43): cell
// variables, face variables, lookup-table columns and two driver options.
struct StVars {
  int elec = 1, pion = 2, elec_old = 3, pion_old = 4, phi = 5, fld = 6, rhs = 7, pho = -1,
      eps = 9;
  int f_elec = 1, f_fld = 2;
  int t_mob = 1, t_dif = 2, t_alpha = 3, t_eta = 4;
  int dielectric_rule = 1;
};

// adjoint variables of the reverse-mode transport substep: cell variables
// le, li (densities), lE (|E|), lpho (photoionization source); face
// variables lF (electron flux), lfld (field)
struct AdjVars {
  int le = -1, li = -1, lE = -1, lpho = -1, lF = -1, lfld = -1;
};

struct MgOpts {
  int i_phi = -1, i_tmp = -1, i_rhs = -1, i_eps = -1;
  int n_down = 2, n_up = 2, n_base = 2;
  int auto_ops = 0;
  double helm = 0;  // Helmholtz coefficient kappa^2 (a5g_mg_set_helmholtz)
  bool ready = false;
};

constexpr int kRedSlots = 256;
// independent solves that may run side by side (photoionization terms)
constexpr int kMaxTerms = 4;

// Launch shapes of the box kernels (a5g_set_shape, a5g_autotune): per
// kernel, dimension, box size (log2 of the cells a side) and level-size
// class, an index into the kernel's menu (0: the default shape).
// kernels: the multigrid half-sweep (all four shapes), the ghost fill, the
// residual + restriction, the coarse rhs and the correction (shapes 0 and 2)
constexpr int kShapeGsrb = 0, kShapeGc = 1, kShapeResidRstr = 2, kShapeCoarseRhs = 3,
              kShapeCorr = 4;
constexpr int kShapeKernels = 5;
constexpr int kSizeClasses = 3;   // levels of < 1024, < 16384 and more boxes
constexpr int kShapes = 4;        // bit 0: persistent grid, bit 1: half the lanes per box
inline int size_class(int n) { return n < 1024 ? 0 : (n < 16384 ? 1 : 2); }
inline int log2i(int v) {
  int l = 0;
  while ((1 << l) < v) l++;
  return l;
}

struct Pool {
  Tree tree;
  // No GPU visible: the tree core still works (topology calls), every
  // call that touches cell data fails.
  bool host_only = false;
  hipStream_t stream = nullptr;
  // the handle's own stream; `stream` is a side stream while independent
  // work is issued there
  hipStream_t home = nullptr;
  int dev_cap = 0;  // boxes allocated on the device (slots 0..dev_cap)
  double* cc = nullptr;
  double* fc = nullptr;
  // FP32 variables of the mixed-precision solve (scratch: their contents do
  // not survive a capacity change or a tidy_up)
  float* cc32 = nullptr;
  int n32 = 0, cap32 = 0;
  // device topology
  int *d_lvl = nullptr, *d_ix = nullptr, *d_parent = nullptr, *d_children = nullptr;
  int *d_nbs = nullptr, *d_ncell = nullptr, *d_tag = nullptr;
  double *d_dr = nullptr, *d_rmin = nullptr;
  int* d_lists = nullptr;  // concatenated level lists
  size_t lists_cap = 0;
  // host offsets into d_lists, per level: ids, leaves, parents
  std::vector<size_t> off_ids, off_leaves, off_parents;
  bool topo_dirty = true;
  bool lists_dirty = false;  // only the level lists need an upload (a mask change)
  uint64_t topo_version = 0;  // counts topology uploads
  // (parent, direction) pairs of a$D_consistent_fluxes per level, on the
  // device, for topology version flux_pairs_version
  int2* d_flux_pairs = nullptr;
  size_t flux_pairs_cap = 0;
  std::vector<size_t> flux_pairs_off, flux_pairs_n;
  uint64_t flux_pairs_version = ~0ull;
  // reduction scratch: kRedSlots pairs of ordered-integer keys, one pair per
  // group of workgroups, so that atomics do not all meet on one address
  double* d_red = nullptr;
  // scratch of the diagnostics (a5g_diag.hip): diag_cap doubles, then as
  // many ints, then 4 doubles
  double* d_diag = nullptr;
  int diag_cap = 0;
  void ensure_diag(int n);
  // Device buffer for the short-lived data of one call (per-box outputs,
  // id lists uploaded for a kernel): grown when a call needs more, kept
  // between calls and freed in release(), so that the stepping loop does
  // not allocate and free (hipFree waits for the whole device)
  void* d_work = nullptr;
  size_t work_bytes = 0;
  void* work(size_t bytes);
  MgOpts mg;
  StVars sv;
  // transport lookup table on the device: rows_cols(n_rows, n_cols)
  double* d_table = nullptr;
  int tbl_rows = 0, tbl_cols = 0;
  // adjoints of the table rows accumulated by the reverse-mode kernels
  double* d_tgrad = nullptr;
  size_t tgrad_n = 0;
  double tbl_xmin = 0, tbl_inv_dx = 0;
  std::vector<double> h_table;  // host copy while it is being built
  bool prof_on = false;
  std::vector<ProfRec> prof;
  std::vector<hipEvent_t> ev_pool;
  // events around `count` consecutive launches of one kernel class
  int prof_begin(int kind, int lvl, double bytes, int count = 1);
  void prof_end(int idx);
  // The refinement rules of set_ref_flags for a5g_ref_flags
  // (a5g_set_ref_rules): per run the seeds (r0, r1, width), the refinement
  // regions shared by all runs (lvl, lo, hi), the time window of the seed
  // rule, crv_refine and the base boxes per run along the last dimension.
  // h_rules holds the seeds, then the regions; d_rules the same followed by
  // one time per run, uploaded by each a5g_ref_flags.
  std::vector<double> h_rules;
  bool rules_set = false;
  int rules_n_seed = 0, rules_n_reg = 0, rules_n_last = 1, rules_runs = 1;
  double rules_window = 0, rules_crv = 0;
  double* d_rules = nullptr;
  size_t rules_cap = 0;
  unsigned char shape_tab[kShapeKernels][2][6][kSizeClasses] = {};
  int shape(int kernel, int dim, int nc, int n) const {
    return shape_tab[kernel][dim - 2][log2i(nc)][size_class(n)];
  }
  // workgroups of a persistent grid: 8 per compute unit
  int n_cu = 256;
  int persist_grid(int nblk) const {
    const int g = std::min(nblk, 8 * n_cu);
    return g >= 8 ? g - g % 8 : g;
  }
  int bc_kind[6] = {0, 0, 0, 0, 0, 0};
  double bc_val[6] = {0, 0, 0, 0, 0, 0};
  // Multigrid call sequences captured as HIP graphs and replayed while the
  // state their kernel arguments were taken from (device pointers, level
  // lists, options) is unchanged; graphs_state fingerprints that state.
  bool graphs_on = true;
  // Bytes of potential and rhs of a level above which its half-sweeps
  // stream (non-temporal rhs loads and stores of the new colour): three
  // times the 256 MB Infinity Cache. A5G_STREAM_BYTES at a5g_init sets
  // another threshold (0: every level of 16^3 or 32^2 boxes streams).
  double stream_bytes = 768.0 * (1 << 20);
  // The finest level whose ghost cells of variable defer_iv the last
  // multigrid call left unstored (kNoDefer: none); flush_ghosts stores
  // them. A5G_DEFER_GHOSTS=0 at a5g_init stores them at once instead.
  int defer_lvl = kNoDefer;
  int defer_iv = -1;
  bool defer_on = true;
  // smoothing calls of small levels in one workgroup's LDS (k_gsrb_small);
  // A5G_SMALL_LEVELS=0 at a5g_init launches every half-sweep instead
  bool small_on = true;
  // LDS bytes one workgroup may use (hipDeviceAttributeMaxSharedMemoryPerBlock)
  int max_lds = 64 * 1024;
  // the most LDS a level of the small-level smoother may take
  // (A5G_SMALL_LDS at a5g_init; default 64 KB)
  long small_lds = 64 * 1024;
  // residual + restriction of 16^3 FP64 boxes in two z slabs per box
  // (A5G_RSTR_SPLIT=1 at a5g_init; by default one workgroup per box, which
  // measures faster: the slabs double the workgroups per compute unit but
  // re-read a plane each and filter the staging loop)
  bool rstr_split = false;
  uint64_t graphs_state = 0;
  std::vector<std::pair<uint64_t, hipGraphExec_t>> graphs;
  // waits for every stream of the handle, then destroys the graphs (one may
  // still run on a side stream otherwise)
  void clear_graphs();
  // side streams for independent solves run side by side, created on first
  // use; the handle's stream waits for them after every such section
  hipStream_t side[kMaxTerms - 1] = {};
  hipEvent_t ev_fork = nullptr;
  hipEvent_t ev_join[kMaxTerms - 1] = {};
  void ensure_side(int n);
  void sync_streams();
  // 2 kRedSlots reduction keys per independent solve (allocated once and
  // kept: captured graphs hold the pointer)
  unsigned long long* d_term_red = nullptr;
  // adjoints of the boundary values of the multigrid's reverse mode (8
  // doubles, allocated once and kept: a captured graph holds the pointer)
  double* d_vbar = nullptr;
  void ensure_term_red();

  // instances (a5g_set_instances): the device level lists hold only the
  // boxes of active instances
  int n_inst = 0;
  double inst_stride = 0;
  double* d_inst = nullptr;
  std::vector<double> h_inst;
  std::vector<char> inst_active;
  std::vector<std::vector<int>> act_ids, act_leaves, act_parents;
  int inst_of(int id) const;
  bool all_active() const;
  // the level lists the device holds (all boxes, or those of active
  // instances); valid after sync_topology
  const std::vector<int>& lids(int l) const {
    if (cur_group >= 0) return g_ids[cur_group][l - tree.min_lvl];
    return n_inst && !topo_dirty && !lists_dirty ? act_ids[l - tree.min_lvl] : tree.lv_ids(l);
  }
  const std::vector<int>& lleaves(int l) const {
    if (cur_group >= 0) return g_leaves[cur_group][l - tree.min_lvl];
    return n_inst && !topo_dirty && !lists_dirty ? act_leaves[l - tree.min_lvl] : tree.lv_leaves(l);
  }
  const std::vector<int>& lparents(int l) const {
    if (cur_group >= 0) return g_parents[cur_group][l - tree.min_lvl];
    return n_inst && !topo_dirty && !lists_dirty ? act_parents[l - tree.min_lvl] : tree.lv_parents(l);
  }
  // Groups of instances (a5g_set_instance_groups): the group of every
  // instance (-1: none). After sync_topology every group's level lists sit
  // at offsets of their own in d_lists, and cur_group >= 0 (set only while
  // one group's work is issued) makes the accessors return them, so groups
  // of runs can be solved side by side.
  std::vector<int> inst_group;
  int n_groups = 0;
  int cur_group = -1;
  std::vector<std::vector<std::vector<int>>> g_ids, g_leaves, g_parents;
  std::vector<std::vector<size_t>> g_off_ids, g_off_leaves, g_off_parents;

  // device store of whole variables (a5g_state_store): per slot one buffer
  // with the copies, the box slots it covers, the variable lists and the
  // topology version it was taken at
  struct Slot {
    double* d = nullptr;
    size_t bytes = 0;
    int max_id = 0;
    std::vector<int> key;
    uint64_t topo = 0;
  };
  std::vector<Slot> store;

  int bs() const;
  int fs() const;
  Dev dev();
  void ensure_capacity(int cap);  // grows GPU storage, keeping data
  void sync_topology();           // uploads topology and level lists if dirty
  void apply_moves(const std::vector<int>& moves, int n_used);  // after tidy_up
  void release();
  // level list on the device
  const int* ids(int l) const {
    const int k = l - tree.min_lvl;
    return d_lists + (cur_group >= 0 ? g_off_ids[cur_group][k] : off_ids[k]);
  }
  const int* leaves(int l) const {
    const int k = l - tree.min_lvl;
    return d_lists + (cur_group >= 0 ? g_off_leaves[cur_group][k] : off_leaves[k]);
  }
  const int* parents(int l) const {
    const int k = l - tree.min_lvl;
    return d_lists + (cur_group >= 0 ? g_off_parents[cur_group][k] : off_parents[k]);
  }
  int n_ids(int l) const { return (int)lids(l).size(); }
  int n_leaves(int l) const { return (int)lleaves(l).size(); }
  int n_parents(int l) const { return (int)lparents(l).size(); }
  // the finest level of the multigrid: with instances the finest level
  // holding boxes of an active instance (after sync_topology)
  int mg_max_lvl() const {
    if (!n_inst) return tree.max_lvl;
    int m = tree.min_lvl;
    for (int l = tree.min_lvl; l <= tree.max_lvl; l++)
      if (n_ids(l) > 0) m = l;
    return m;
  }
  int nc_lvl(int l) const {
    return l >= 1 ? tree.n_cell : tree.n_cell / (1 << (1 - l));
  }
};

}  // namespace a5g
