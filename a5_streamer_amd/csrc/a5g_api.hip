// The device pool of the tree's boxes and the C ABI of the a5g library
// (include/a5g.h).
#include <hip/hip_runtime.h>

#include <cstdlib>

#include <algorithm>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/a5g.h"
#include "a5g_internal.hpp"
#include "a5g_kernels.hpp"

namespace a5g {

static void hcheck(hipError_t e, const char* what) {
  if (e != hipSuccess)
    throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

int Pool::bs() const {
  int n = tree.n_cell + 2;
  return tree.dim == 2 ? n * n : n * n * n;
}

int Pool::fs() const {
  int nc = tree.n_cell;
  return tree.dim == 2 ? (nc + 1) * nc : (nc + 1) * nc * nc;
}

Dev Pool::dev() {
  Dev a;
  a.cc = cc;
  a.fc = fc;
  a.bs = bs();
  a.fs = fs();
  a.vs = (size_t)(dev_cap + 1) * a.bs;
  a.fvs = (size_t)(dev_cap + 1) * tree.dim * a.fs;
  a.cc32 = cc32;
  a.vs32 = (size_t)(cap32 + 1) * a.bs;
  a.coord = tree.coord;
  a.lvl = d_lvl;
  a.ix = d_ix;
  a.parent = d_parent;
  a.children = d_children;
  a.nbs = d_nbs;
  a.ncell = d_ncell;
  a.dr = d_dr;
  a.rmin = d_rmin;
  for (int i = 0; i < 6; i++) {
    a.bc_kind[i] = bc_kind[i];
    a.bc_val[i] = bc_val[i];
  }
  a.n_inst = n_inst;
  a.inst_stride = inst_stride;
  a.inst_val = d_inst;
  return a;
}

// The instance of a box: its position along the last dimension, a quarter
// cell inside the box so that rounding of r_min cannot move it across
int Pool::inst_of(int id) const {
  const int d = tree.dim - 1;
  const int g = (int)std::floor((tree.rmin[id * 3 + d] + 0.25 * tree.dr[id]) / inst_stride);
  return std::min(std::max(g, 0), n_inst - 1);
}

bool Pool::all_active() const {
  for (char c : inst_active)
    if (!c) return false;
  return true;
}

template <typename T>
static void realloc_dev(T** p, size_t n) {
  if (*p) hcheck(hipFree(*p), "hipFree");
  *p = nullptr;
  hcheck(hipMalloc((void**)p, std::max<size_t>(n, 1) * sizeof(T)), "hipMalloc");
}

void Pool::ensure_capacity(int need) {
  if (host_only) return;
  if (need <= dev_cap && cc) return;
  int ncap = std::max(need, dev_cap + dev_cap / 2);
  const size_t b = bs(), f = fs();
  const int ncc = tree.n_cc, nfc = std::max(tree.n_fc, 1), D = tree.dim;
  double* ncc_p = nullptr;
  double* nfc_p = nullptr;
  hcheck(hipMalloc((void**)&ncc_p, (size_t)ncc * (ncap + 1) * b * sizeof(double)), "hipMalloc cc");
  hcheck(hipMalloc((void**)&nfc_p, (size_t)nfc * (ncap + 1) * D * f * sizeof(double)), "hipMalloc fc");
  hcheck(hipMemsetAsync(ncc_p, 0, (size_t)ncc * (ncap + 1) * b * sizeof(double), stream), "memset");
  hcheck(hipMemsetAsync(nfc_p, 0, (size_t)nfc * (ncap + 1) * D * f * sizeof(double), stream), "memset");
  if (cc) {
    for (int v = 0; v < ncc; v++)
      hcheck(hipMemcpyAsync(ncc_p + (size_t)v * (ncap + 1) * b, cc + (size_t)v * (dev_cap + 1) * b,
                            (size_t)(dev_cap + 1) * b * sizeof(double), hipMemcpyDeviceToDevice,
                            stream), "copy cc");
    for (int v = 0; v < nfc; v++)
      hcheck(hipMemcpyAsync(nfc_p + (size_t)v * (ncap + 1) * D * f,
                            fc + (size_t)v * (dev_cap + 1) * D * f,
                            (size_t)(dev_cap + 1) * D * f * sizeof(double),
                            hipMemcpyDeviceToDevice, stream), "copy fc");
    hcheck(hipStreamSynchronize(stream), "sync");
    hcheck(hipFree(cc), "free");
    hcheck(hipFree(fc), "free");
  }
  cc = ncc_p;
  fc = nfc_p;
  dev_cap = ncap;
  if (cc32) hcheck(hipFree(cc32), "hipFree");
  cc32 = nullptr;
  n32 = cap32 = 0;
  realloc_dev(&d_lvl, ncap + 1);
  realloc_dev(&d_ix, (size_t)(ncap + 1) * 3);
  realloc_dev(&d_parent, ncap + 1);
  realloc_dev(&d_children, (size_t)(ncap + 1) * 8);
  realloc_dev(&d_nbs, (size_t)(ncap + 1) * 6);
  realloc_dev(&d_ncell, ncap + 1);
  realloc_dev(&d_tag, ncap + 1);
  realloc_dev(&d_dr, ncap + 1);
  realloc_dev(&d_rmin, (size_t)(ncap + 1) * 3);
  topo_dirty = true;
}

void Pool::sync_topology() {
  if (host_only) throw std::runtime_error("no GPU visible: cell data and kernels are unavailable");
  if (!topo_dirty && !lists_dirty) return;
  // an upload belongs on the handle's stream, before any side-by-side section
  if (stream != home)
    throw std::logic_error("topology upload on a side stream: upload it before the fork");
  auto up = [&](void* dst, const void* src, size_t bytes) {
    hcheck(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, stream), "topology upload");
  };
  if (topo_dirty) {
    ensure_capacity(std::max(tree.max_id, tree.cap));
    const int n = tree.max_id + 1;
    up(d_lvl, tree.lvl.data(), n * sizeof(int));
    up(d_ix, tree.ix.data(), (size_t)n * 3 * sizeof(int));
    up(d_parent, tree.parent.data(), n * sizeof(int));
    up(d_children, tree.children.data(), (size_t)n * 8 * sizeof(int));
    up(d_nbs, tree.nbs.data(), (size_t)n * 6 * sizeof(int));
    up(d_ncell, tree.ncell.data(), n * sizeof(int));
    up(d_tag, tree.tag.data(), n * sizeof(int));
    up(d_dr, tree.dr.data(), n * sizeof(double));
    up(d_rmin, tree.rmin.data(), (size_t)n * 3 * sizeof(double));
  }
  // level lists (with instances, only the boxes of active instances)
  const int nl = (int)tree.ids.size();
  off_ids.assign(nl, 0);
  off_leaves.assign(nl, 0);
  off_parents.assign(nl, 0);
  const bool filter = n_inst > 0 && !all_active();
  auto pick = [&](const std::vector<std::vector<int>>& src, std::vector<std::vector<int>>& dst) {
    dst.assign(src.size(), {});
    for (size_t l = 0; l < src.size(); l++) {
      if (!filter) {
        dst[l] = src[l];
        continue;
      }
      for (int id : src[l])
        if (inst_active[inst_of(id)]) dst[l].push_back(id);
    }
  };
  if (n_inst) {
    pick(tree.ids, act_ids);
    pick(tree.leaves, act_leaves);
    pick(tree.parents, act_parents);
  }
  const auto& L_ids = n_inst ? act_ids : tree.ids;
  const auto& L_leaves = n_inst ? act_leaves : tree.leaves;
  const auto& L_parents = n_inst ? act_parents : tree.parents;
  static thread_local std::vector<int> host;
  host.clear();
  for (int l = 0; l < nl; l++) {
    off_ids[l] = host.size();
    host.insert(host.end(), L_ids[l].begin(), L_ids[l].end());
    off_leaves[l] = host.size();
    host.insert(host.end(), L_leaves[l].begin(), L_leaves[l].end());
    off_parents[l] = host.size();
    host.insert(host.end(), L_parents[l].begin(), L_parents[l].end());
  }
  // every group's lists after the active ones
  const int ng = n_inst ? n_groups : 0;
  auto split = [&](const std::vector<std::vector<int>>& src,
                   std::vector<std::vector<std::vector<int>>>& dst) {
    dst.assign(ng, std::vector<std::vector<int>>(nl));
    if (ng == 0) return;  // no groups: inst_group may be empty
    for (int l = 0; l < nl; l++)
      for (int id : src[l]) {
        const int g = inst_group[inst_of(id)];
        if (g >= 0) dst[g][l].push_back(id);
      }
  };
  split(tree.ids, g_ids);
  split(tree.leaves, g_leaves);
  split(tree.parents, g_parents);
  g_off_ids.assign(ng, std::vector<size_t>(nl));
  g_off_leaves.assign(ng, std::vector<size_t>(nl));
  g_off_parents.assign(ng, std::vector<size_t>(nl));
  for (int g = 0; g < ng; g++)
    for (int l = 0; l < nl; l++) {
      g_off_ids[g][l] = host.size();
      host.insert(host.end(), g_ids[g][l].begin(), g_ids[g][l].end());
      g_off_leaves[g][l] = host.size();
      host.insert(host.end(), g_leaves[g][l].begin(), g_leaves[g][l].end());
      g_off_parents[g][l] = host.size();
      host.insert(host.end(), g_parents[g][l].begin(), g_parents[g][l].end());
    }
  if (host.size() > lists_cap || !d_lists) {
    lists_cap = std::max(host.size(), lists_cap * 3 / 2) + 64;
    realloc_dev(&d_lists, lists_cap);
  }
  if (!host.empty()) up(d_lists, host.data(), host.size() * sizeof(int));
  hcheck(hipStreamSynchronize(stream), "topology sync");
  topo_dirty = false;
  lists_dirty = false;
  topo_version++;
}

// new slot n <- old slot map[n] for every variable
__global__ void k_gather(const double* __restrict__ src, double* __restrict__ dst,
                         const int* __restrict__ map, int n_used, size_t per_box,
                         size_t old_vs, size_t new_vs, int n_var) {
  const size_t total = (size_t)n_used * per_box;
  for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < total;
       q += (size_t)gridDim.x * blockDim.x) {
    const int nn = (int)(q / per_box) + 1;
    const size_t w = q % per_box;
    const int o = map[nn];
    for (int v = 0; v < n_var; v++)
      dst[(size_t)v * new_vs + (size_t)nn * per_box + w] =
          src[(size_t)v * old_vs + (size_t)o * per_box + w];
  }
}

void Pool::apply_moves(const std::vector<int>& moves, int n_used) {
  if (host_only) return;
  const int ncap = std::max(dev_cap, std::max(tree.cap, n_used));
  const size_t b = bs(), f = (size_t)fs() * tree.dim;
  const int ncc = tree.n_cc, nfc = std::max(tree.n_fc, 1);
  int* d_map = nullptr;
  realloc_dev(&d_map, moves.size());
  hcheck(hipMemcpyAsync(d_map, moves.data(), moves.size() * sizeof(int), hipMemcpyHostToDevice,
                        stream), "moves");
  double *ncc_p = nullptr, *nfc_p = nullptr;
  hcheck(hipMalloc((void**)&ncc_p, (size_t)ncc * (ncap + 1) * b * sizeof(double)), "hipMalloc");
  hcheck(hipMalloc((void**)&nfc_p, (size_t)nfc * (ncap + 1) * f * sizeof(double)), "hipMalloc");
  hcheck(hipMemsetAsync(ncc_p, 0, (size_t)ncc * (ncap + 1) * b * sizeof(double), stream), "memset");
  hcheck(hipMemsetAsync(nfc_p, 0, (size_t)nfc * (ncap + 1) * f * sizeof(double), stream), "memset");
  if (n_used > 0) {
    hipLaunchKernelGGL(k_gather, dim3(1024), dim3(256), 0, stream, cc, ncc_p, d_map, n_used, b,
                       (size_t)(dev_cap + 1) * b, (size_t)(ncap + 1) * b, ncc);
    hipLaunchKernelGGL(k_gather, dim3(1024), dim3(256), 0, stream, fc, nfc_p, d_map, n_used, f,
                       (size_t)(dev_cap + 1) * f, (size_t)(ncap + 1) * f, nfc);
    hcheck(hipGetLastError(), "gather");
  }
  hcheck(hipStreamSynchronize(stream), "gather sync");
  hcheck(hipFree(cc), "free");
  hcheck(hipFree(fc), "free");
  hcheck(hipFree(d_map), "free");
  cc = ncc_p;
  fc = nfc_p;
  if (cc32) hcheck(hipFree(cc32), "hipFree");
  cc32 = nullptr;
  n32 = cap32 = 0;
  if (ncap != dev_cap) {
    dev_cap = ncap;
    realloc_dev(&d_lvl, ncap + 1);
    realloc_dev(&d_ix, (size_t)(ncap + 1) * 3);
    realloc_dev(&d_parent, ncap + 1);
    realloc_dev(&d_children, (size_t)(ncap + 1) * 8);
    realloc_dev(&d_nbs, (size_t)(ncap + 1) * 6);
    realloc_dev(&d_ncell, ncap + 1);
    realloc_dev(&d_tag, ncap + 1);
    realloc_dev(&d_dr, ncap + 1);
    realloc_dev(&d_rmin, (size_t)(ncap + 1) * 3);
  }
  topo_dirty = true;
}

int Pool::prof_begin(int kind, int lvl, double bytes, int count) {
  if (!prof_on) return -1;
  auto take = [&]() {
    hipEvent_t e;
    if (!ev_pool.empty()) {
      e = ev_pool.back();
      ev_pool.pop_back();
    } else {
      hcheck(hipEventCreate(&e), "hipEventCreate");
    }
    return e;
  };
  ProfRec r{kind, lvl, bytes, take(), take(), count};
  hcheck(hipEventRecord(r.a, stream), "hipEventRecord");
  prof.push_back(r);
  return (int)prof.size() - 1;
}

void Pool::prof_end(int idx) {
  if (idx < 0) return;
  hcheck(hipEventRecord(prof[idx].b, stream), "hipEventRecord");
}

void Pool::clear_graphs() {
  if (!graphs.empty()) {
    if (home) (void)hipStreamSynchronize(home);
    if (stream && stream != home) (void)hipStreamSynchronize(stream);
    for (auto s : side)
      if (s) (void)hipStreamSynchronize(s);
  }
  for (auto& g : graphs) (void)hipGraphExecDestroy(g.second);
  graphs.clear();
}

void Pool::ensure_side(int n) {
  if (n > kMaxTerms - 1) throw std::invalid_argument("too many side streams");
  if (!ev_fork) hcheck(hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming), "event");
  for (int k = 0; k < n; k++) {
    if (!side[k]) hcheck(hipStreamCreateWithFlags(&side[k], hipStreamNonBlocking), "side stream");
    if (!ev_join[k])
      hcheck(hipEventCreateWithFlags(&ev_join[k], hipEventDisableTiming), "event");
  }
}

void Pool::ensure_term_red() {
  if (d_term_red) return;
  hcheck(hipMalloc((void**)&d_term_red, (size_t)kMaxTerms * 2 * kRedSlots * sizeof(unsigned long long)),
         "hipMalloc term reductions");
}

void Pool::sync_streams() {
  if (home) hcheck(hipStreamSynchronize(home), "stream sync");
  if (stream && stream != home) hcheck(hipStreamSynchronize(stream), "stream sync");
  for (auto s : side)
    if (s) hcheck(hipStreamSynchronize(s), "stream sync");
}

void Pool::ensure_diag(int n) {
  if (n <= diag_cap) return;
  if (d_diag) hcheck(hipFree(d_diag), "hipFree diag");
  d_diag = nullptr;
  diag_cap = 0;
  const int cap = std::max(n, 1024);
  hcheck(hipMalloc((void**)&d_diag, ((size_t)cap + (cap + 1) / 2 + 4) * sizeof(double)),
         "hipMalloc diag");
  diag_cap = cap;
}

void* Pool::work(size_t bytes) {
  if (bytes > work_bytes || !d_work) {
    // growing waits for the whole device, which a section running side by
    // side on the side streams must not do (it would serialise them, and
    // break a capture)
    if (stream != home)
      throw std::logic_error("work buffer growth on a side stream: grow it before the fork");
    // no kernel of an earlier call may still read the old buffer: hipFree
    // waits for the device
    if (d_work) hcheck(hipFree(d_work), "hipFree work");
    d_work = nullptr;
    work_bytes = 0;
    const size_t want = std::max(bytes + bytes / 2, (size_t)1 << 16);
    hcheck(hipMalloc(&d_work, want), "hipMalloc work");
    work_bytes = want;
  }
  return d_work;
}

void Pool::release() {
  clear_graphs();
  for (auto& r : prof) {
    (void)hipEventDestroy(r.a);
    (void)hipEventDestroy(r.b);
  }
  for (auto& e : ev_pool) (void)hipEventDestroy(e);
  prof.clear();
  ev_pool.clear();
  auto f = [](void* p) {
    if (p) (void)hipFree(p);
  };
  f(d_table);
  d_table = nullptr;
  f(cc32);
  cc32 = nullptr;
  f(cc); f(fc); f(d_lvl); f(d_ix); f(d_parent); f(d_children); f(d_nbs); f(d_ncell);
  f(d_tag); f(d_dr); f(d_rmin); f(d_lists); f(d_red); f(d_flux_pairs); f(d_diag); f(d_inst);
  f(d_term_red);
  d_term_red = nullptr;
  f(d_vbar);
  d_vbar = nullptr;
  store_release(*this);
  f(d_work);
  d_work = nullptr;
  work_bytes = 0;
  f(d_tgrad);
  d_tgrad = nullptr;
  f(d_rules);
  d_rules = nullptr;
  rules_cap = 0;
  rules_set = false;
  tgrad_n = 0;
  d_inst = nullptr;
  d_flux_pairs = nullptr;
  d_diag = nullptr;
  diag_cap = 0;
  for (int k = 0; k < kMaxTerms - 1; k++) {
    if (side[k]) (void)hipStreamDestroy(side[k]);
    if (ev_join[k]) (void)hipEventDestroy(ev_join[k]);
    side[k] = nullptr;
    ev_join[k] = nullptr;
  }
  if (ev_fork) (void)hipEventDestroy(ev_fork);
  ev_fork = nullptr;
  if (home) (void)hipStreamDestroy(home);
  cc = fc = nullptr;
  stream = nullptr;
  home = nullptr;
}

}  // namespace a5g

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------

struct a5g_tree {
  a5g::Pool P;
};

static double* a5g_facep_host(const a5g::Dev& a, int ifc, int id, int d, int dim) {
  return a.fc + (size_t)(ifc - 1) * a.fvs + ((size_t)id * dim + d) * a.fs;
}

static thread_local std::string g_err;

#define API_BEGIN \
  try {           \
    g_err.clear();
#define API_END                    \
  return 0;                        \
  }                                \
  catch (const std::exception& e) { \
    g_err = e.what();              \
    return 1;                      \
  }                                \
  catch (...) {                    \
    g_err = "unknown error";       \
    return 1;                      \
  }

#define NEED(t) \
  if (!(t)) throw std::invalid_argument("null a5g_tree handle");
// ghost cells the last multigrid call left unstored are stored before any
// other call sees the data (flush_ghosts)
#define FLUSH(t) \
  if (!(t)->P.host_only) a5g::flush_ghosts((t)->P);

static void check_id(const a5g::Tree& T, int id) {
  if (id < 1 || id > T.max_id || !T.in_use[id])
    throw std::out_of_range("box id " + std::to_string(id) + " is not in use");
}
// Host conversion between the reference's lexicographic box block
// cc(0:nc+1, ...) and the colour-split device layout (cidx in a5g_kernels.hpp).
static void split_perm(int dim, int nc, std::vector<int>& perm) {
  const int n = nc + 2;
  const int bl = dim == 2 ? n * n : n * n * n;
  perm.resize(bl);
  for (int q = 0; q < bl; q++) {
    const int i = q % n, j = (q / n) % n, k = dim == 3 ? q / (n * n) : 0;
    perm[q] = dim == 2 ? a5g::cidx<2>(nc, i, j, k) : a5g::cidx<3>(nc, i, j, k);
  }
}
static void lex_to_split(int dim, int nc, const double* in, double* out) {
  static thread_local std::vector<int> perm;
  static thread_local int key = -1;
  if (key != dim * 1000 + nc) {
    split_perm(dim, nc, perm);
    key = dim * 1000 + nc;
  }
  for (size_t q = 0; q < perm.size(); q++) out[perm[q]] = in[q];
}
static void split_to_lex(int dim, int nc, const double* in, double* out) {
  static thread_local std::vector<int> perm;
  static thread_local int key = -1;
  if (key != dim * 1000 + nc) {
    split_perm(dim, nc, perm);
    key = dim * 1000 + nc;
  }
  for (size_t q = 0; q < perm.size(); q++) out[q] = in[perm[q]];
}

static void need_gpu(const a5g::Pool& P) {
  if (P.host_only) throw std::runtime_error("no GPU visible: cell data and kernels are unavailable");
}
static void check_iv(const a5g::Tree& T, int iv) {
  if (iv < 1 || iv > T.n_cc)
    throw std::out_of_range("cell variable " + std::to_string(iv) + " out of range");
}

extern "C" {

const char* a5g_last_error(void) { return g_err.c_str(); }

int a5g_init(a5g_tree** t, int dim, int n_cell, int n_var_cell, int n_var_face, double dr,
             const double* r_min, int coarsen_to, int lvls_max, int coord, int n_boxes) {
  API_BEGIN
  if (!t) throw std::invalid_argument("null output handle");
  *t = nullptr;
  auto* h = new a5g_tree();
  try {
    h->P.tree.init(dim, n_cell, n_var_cell, n_var_face, dr, r_min, coarsen_to,
                   lvls_max > 0 ? lvls_max : 30, coord, n_boxes > 0 ? n_boxes : 100);
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
      (void)hipGetLastError();
      h->P.host_only = true;
    } else {
      a5g::hcheck(hipStreamCreateWithFlags(&h->P.stream, hipStreamNonBlocking), "stream");
      h->P.home = h->P.stream;
      // A5G_GRAPHS=0 runs every multigrid call as plain launches
      const char* g = std::getenv("A5G_GRAPHS");
      h->P.graphs_on = !(g && g[0] == '0');
      // A5G_DEFER_GHOSTS=0 stores the finest level's ghost cells at the end
      // of every multigrid call, as the reference does, instead of leaving
      // them to the next reader
      const char* dg = std::getenv("A5G_DEFER_GHOSTS");
      h->P.defer_on = !(dg && dg[0] == '0');
      // A5G_SMALL_LEVELS=0 launches every half-sweep of the small coarse
      // levels instead of one workgroup per smoothing call
      const char* sl = std::getenv("A5G_SMALL_LEVELS");
      h->P.small_on = !(sl && sl[0] == '0');
      const char* rs = std::getenv("A5G_RSTR_SPLIT");
      h->P.rstr_split = rs && rs[0] == '1';
      const char* sx = std::getenv("A5G_SMALL_LDS");
      if (sx && sx[0]) {
        char* end = nullptr;
        const long v = std::strtol(sx, &end, 10);
        if (end == sx || *end != 0 || v < 0)
          throw std::invalid_argument("A5G_SMALL_LDS must be a byte count >= 0");
        h->P.small_lds = v;
      }
      const char* sb = std::getenv("A5G_STREAM_BYTES");
      if (sb && sb[0]) {
        char* end = nullptr;
        const double v = std::strtod(sb, &end);
        if (end == sb || *end != 0 || !(v >= 0))
          throw std::invalid_argument("A5G_STREAM_BYTES must be a byte count >= 0");
        h->P.stream_bytes = v;
      }
      a5g::hcheck(hipMalloc((void**)&h->P.d_red, 2 * a5g::kRedSlots * sizeof(unsigned long long)),
                  "hipMalloc");
      a5g::noalloc_init();
      {
        int d0 = 0, lds = 0;
        if (hipGetDevice(&d0) == hipSuccess &&
            hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, d0) ==
                hipSuccess &&
            lds > 0)
          h->P.max_lds = lds;
        (void)hipGetLastError();
      }
      int dev = 0, n_cu = 0;
      if (hipGetDevice(&dev) == hipSuccess &&
          hipDeviceGetAttribute(&n_cu, hipDeviceAttributeMultiprocessorCount, dev) ==
              hipSuccess &&
          n_cu > 0)
        h->P.n_cu = n_cu;
      h->P.ensure_capacity(h->P.tree.cap);
    }
  } catch (...) {
    h->P.release();
    delete h;
    throw;
  }
  *t = h;
  API_END
}

int a5g_destroy(a5g_tree* t) {
  API_BEGIN
  NEED(t);
  t->P.release();
  delete t;
  API_END
}

int a5g_set_base(a5g_tree* t, int n, const int* ix_list, const int* nb_list) {
  API_BEGIN
  NEED(t);
  FLUSH(t);
  t->P.tree.set_base(n, ix_list, nb_list);
  t->P.ensure_capacity(t->P.tree.cap);
  t->P.topo_dirty = true;
  API_END
}

int a5g_adjust_refinement(a5g_tree* t, int n_flags, const int* flags, int* n_add, int* n_rm) {
  API_BEGIN
  NEED(t);
  FLUSH(t);
  int a = 0, r = 0;
  t->P.tree.adjust_refinement(n_flags, flags, &a, &r);
  t->P.ensure_capacity(std::max(t->P.tree.cap, t->P.tree.max_id));
  t->P.topo_dirty = true;
  if (n_add) *n_add = a;
  if (n_rm) *n_rm = r;
  API_END
}

int a5g_ref_info_lvl(a5g_tree* t, int lvl, int* add, int* rm, int* n_add, int* n_rm) {
  API_BEGIN
  NEED(t);
  const auto& T = t->P.tree;
  int na = 0, nr = 0;
  if (lvl >= 1 && lvl <= (int)T.ri_add.size()) {
    na = (int)T.ri_add[lvl - 1].size();
    nr = (int)T.ri_rm[lvl - 1].size();
    if (add) std::copy(T.ri_add[lvl - 1].begin(), T.ri_add[lvl - 1].end(), add);
    if (rm) std::copy(T.ri_rm[lvl - 1].begin(), T.ri_rm[lvl - 1].end(), rm);
  }
  if (n_add) *n_add = na;
  if (n_rm) *n_rm = nr;
  API_END
}

int a5g_tidy_up(a5g_tree* t, double max_frac, double goal_frac, int n_clean_min, int reorder) {
  API_BEGIN
  NEED(t);
  FLUSH(t);
  auto& P = t->P;
  if (P.tree.tidy_up(max_frac, goal_frac, n_clean_min, reorder != 0)) {
    P.apply_moves(P.tree.last_moves, P.tree.max_id);
    P.topo_dirty = true;
  }
  API_END
}

int a5g_info(a5g_tree* t, int* out) {
  API_BEGIN
  NEED(t);
  const auto& T = t->P.tree;
  out[0] = T.max_id;
  out[1] = T.max_lvl;
  out[2] = T.min_lvl;
  out[3] = T.lvls_max;
  out[4] = T.cap;
  out[5] = T.n_cell;
  out[6] = T.coord;
  out[7] = T.ready ? 1 : 0;
  API_END
}

int a5g_get_boxes(a5g_tree* t, int n, int* lvl, int* ix, int* parent, int* children,
                  int* neighbors, int* in_use, int* n_cell, double* dr, double* r_min) {
  API_BEGIN
  NEED(t);
  const auto& T = t->P.tree;
  const int D = T.dim, nch = T.nch, nnb = T.nnb;
  for (int id = 1; id <= n; id++) {
    const bool u = id <= T.max_id && T.in_use[id];
    in_use[id - 1] = u;
    lvl[id - 1] = u ? T.lvl[id] : 0;
    parent[id - 1] = u ? T.parent[id] : 0;
    n_cell[id - 1] = u ? T.ncell[id] : 0;
    dr[id - 1] = u ? T.dr[id] : 0;
    for (int d = 0; d < D; d++) {
      ix[(id - 1) * D + d] = u ? T.ix[id * 3 + d] : 0;
      r_min[(id - 1) * D + d] = u ? T.rmin[id * 3 + d] : 0;
    }
    for (int c = 0; c < nch; c++) children[(id - 1) * nch + c] = u ? T.children[id * 8 + c] : 0;
    for (int b = 0; b < nnb; b++) neighbors[(id - 1) * nnb + b] = u ? T.nbs[id * 6 + b] : 0;
  }
  API_END
}

int a5g_level_counts(a5g_tree* t, int lvl, int* c) {
  API_BEGIN
  NEED(t);
  const auto& T = t->P.tree;
  c[0] = c[1] = c[2] = 0;
  if (lvl >= T.min_lvl && lvl <= T.lvls_max + 1) {
    c[0] = (int)T.lv_ids(lvl).size();
    c[1] = (int)T.lv_leaves(lvl).size();
    c[2] = (int)T.lv_parents(lvl).size();
  }
  API_END
}

int a5g_level_lists(a5g_tree* t, int lvl, int* ids, int* leaves, int* parents) {
  API_BEGIN
  NEED(t);
  const auto& T = t->P.tree;
  if (lvl < T.min_lvl || lvl > T.lvls_max + 1) throw std::out_of_range("level out of range");
  if (ids) std::copy(T.lv_ids(lvl).begin(), T.lv_ids(lvl).end(), ids);
  if (leaves) std::copy(T.lv_leaves(lvl).begin(), T.lv_leaves(lvl).end(), leaves);
  if (parents) std::copy(T.lv_parents(lvl).begin(), T.lv_parents(lvl).end(), parents);
  API_END
}

int a5g_set_cc(a5g_tree* t, int id, int iv, const double* data) {
  API_BEGIN
  NEED(t);
  FLUSH(t);
  auto& P = t->P;
  need_gpu(P);
  check_id(P.tree, id);
  check_iv(P.tree, iv);
  P.ensure_capacity(P.tree.max_id);
  a5g::Dev a = P.dev();
  double* dst = a.cc + (size_t)(iv - 1) * a.vs + (size_t)id * a.bs;
  std::vector<double> tmp(P.tree.box_len(id));
  lex_to_split(P.tree.dim, P.tree.ncell[id], data, tmp.data());
  a5g::hcheck(hipMemcpyAsync(dst, tmp.data(), tmp.size() * sizeof(double),
                             hipMemcpyHostToDevice, P.stream), "set_cc");
  a5g::hcheck(hipStreamSynchronize(P.stream), "set_cc sync");
  API_END
}

int a5g_get_cc(a5g_tree* t, int id, int iv, double* data) {
  API_BEGIN
  NEED(t);
  FLUSH(t);
  auto& P = t->P;
  need_gpu(P);
  check_id(P.tree, id);
  check_iv(P.tree, iv);
  a5g::Dev a = P.dev();
  const double* src = a.cc + (size_t)(iv - 1) * a.vs + (size_t)id * a.bs;
  std::vector<double> tmp(P.tree.box_len(id));
  a5g::hcheck(hipMemcpyAsync(tmp.data(), src, tmp.size() * sizeof(double),
                             hipMemcpyDeviceToHost, P.stream), "get_cc");
  a5g::hcheck(hipStreamSynchronize(P.stream), "get_cc sync");
  split_to_lex(P.tree.dim, P.tree.ncell[id], tmp.data(), data);
  API_END
}

int a5g_set_cc_all(a5g_tree* t, int iv, const double* data) {
  API_BEGIN
  NEED(t);
  FLUSH(t);
  auto& P = t->P;
  need_gpu(P);
  check_iv(P.tree, iv);
  P.ensure_capacity(P.tree.max_id);
  a5g::Dev a = P.dev();
  const auto& T = P.tree;
  std::vector<double> tmp((size_t)T.max_id * a.bs, 0.0);
  for (int id = 1; id <= T.max_id; id++)
    if (T.in_use[id])
      lex_to_split(T.dim, T.ncell[id], data + (size_t)(id - 1) * a.bs,
                   tmp.data() + (size_t)(id - 1) * a.bs);
  a5g::hcheck(hipMemcpyAsync(a.cc + (size_t)(iv - 1) * a.vs + a.bs, tmp.data(),
                             tmp.size() * sizeof(double), hipMemcpyHostToDevice, P.stream),
              "set_cc_all");
  a5g::hcheck(hipStreamSynchronize(P.stream), "set_cc_all sync");
  API_END
}

int a5g_get_cc_all(a5g_tree* t, int iv, double* data) {
  API_BEGIN
  NEED(t);
  FLUSH(t);
  auto& P = t->P;
  need_gpu(P);
  check_iv(P.tree, iv);
  a5g::Dev a = P.dev();
  const auto& T = P.tree;
  std::vector<double> tmp((size_t)T.max_id * a.bs);
  a5g::hcheck(hipMemcpyAsync(tmp.data(), a.cc + (size_t)(iv - 1) * a.vs + a.bs,
                             tmp.size() * sizeof(double), hipMemcpyDeviceToHost, P.stream),
              "get_cc_all");
  a5g::hcheck(hipStreamSynchronize(P.stream), "get_cc_all sync");
  for (int id = 1; id <= T.max_id; id++) {
    double* o = data + (size_t)(id - 1) * a.bs;
    if (T.in_use[id]) split_to_lex(T.dim, T.ncell[id], tmp.data() + (size_t)(id - 1) * a.bs, o);
    else std::fill(o, o + a.bs, 0.0);
  }
  API_END
}

int a5g_set_fc(a5g_tree* t, int id, int dim, int ifc, const double* data) {
  API_BEGIN
  NEED(t);
  FLUSH(t);
  auto& P = t->P;
  need_gpu(P);
  check_id(P.tree, id);
  if (ifc < 1 || ifc > P.tree.n_fc || dim < 1 || dim > P.tree.dim)
    throw std::out_of_range("face variable out of range");
  a5g::Dev a = P.dev();
  a5g::hcheck(hipMemcpyAsync(a5g_facep_host(a, ifc, id, dim - 1, P.tree.dim), data,
                             a.fs * sizeof(double), hipMemcpyHostToDevice, P.stream), "set_fc");
  a5g::hcheck(hipStreamSynchronize(P.stream), "set_fc sync");
  API_END
}

int a5g_get_fc(a5g_tree* t, int id, int dim, int ifc, double* data) {
  API_BEGIN
  NEED(t);
  FLUSH(t);
  auto& P = t->P;
  need_gpu(P);
  check_id(P.tree, id);
  if (ifc < 1 || ifc > P.tree.n_fc || dim < 1 || dim > P.tree.dim)
    throw std::out_of_range("face variable out of range");
  a5g::Dev a = P.dev();
  a5g::hcheck(hipMemcpyAsync(data, a5g_facep_host(a, ifc, id, dim - 1, P.tree.dim),
                             a.fs * sizeof(double), hipMemcpyDeviceToHost, P.stream), "get_fc");
  a5g::hcheck(hipStreamSynchronize(P.stream), "get_fc sync");
  API_END
}

int a5g_sync(a5g_tree* t) {
  API_BEGIN
  NEED(t);
  FLUSH(t);
  need_gpu(t->P);
  a5g::hcheck(hipStreamSynchronize(t->P.stream), "sync");
  API_END
}

int a5g_mg_setup(a5g_tree* t, int i_phi, int i_tmp, int i_rhs, int i_eps, int n_down, int n_up,
                 int n_base, const int* bc_kind, const double* bc_val, int auto_ops) {
  API_BEGIN
  NEED(t);
  FLUSH(t);
  auto& P = t->P;
  // mg$D_init_mg, m_aX_mg.f90:130-154
  if (i_phi < 0) throw std::invalid_argument("mg_init_mg: i_phi not set");
  if (i_tmp < 0) throw std::invalid_argument("mg_init_mg: i_tmp not set");
  if (i_rhs < 0) throw std::invalid_argument("mg_init_mg: i_rhs not set");
  check_iv(P.tree, i_phi);
  check_iv(P.tree, i_tmp);
  check_iv(P.tree, i_rhs);
  if (i_eps > 0) check_iv(P.tree, i_eps);
  if (!bc_kind) throw std::invalid_argument("mg_init_mg: sides_bc not set");
  P.mg.i_phi = i_phi;
  P.mg.i_tmp = i_tmp;
  P.mg.i_rhs = i_rhs;
  P.mg.i_eps = i_eps;
  P.mg.n_down = n_down < 0 ? 2 : n_down;
  P.mg.n_up = n_up < 0 ? 2 : n_up;
  P.mg.n_base = n_base < 0 ? 2 : n_base;
  P.mg.auto_ops = auto_ops;
  P.mg.helm = 0;  // a new setup starts from the Laplacian
  for (int nb = 0; nb < 2 * P.tree.dim; nb++) {
    if (bc_kind[nb] < 0 || bc_kind[nb] > 3) throw std::invalid_argument("bad boundary kind");
    if (bc_kind[nb] == a5g::kBcRow2V && (P.tree.dim != 2 || nb < 2))
      throw std::invalid_argument("boundary kind 3 (whole ghost row) is for the y faces of a 2D tree");
    P.bc_kind[nb] = bc_kind[nb];
    P.bc_val[nb] = bc_val ? bc_val[nb] : 0.0;
  }
  P.mg.ready = true;
  API_END
}

int a5g_mg_set_helmholtz(a5g_tree* t, double kappa2) {
  API_BEGIN
  NEED(t);
  FLUSH(t);
  auto& P = t->P;
  if (!(kappa2 >= 0)) throw std::invalid_argument("mg_set_helmholtz: need kappa2 >= 0");
  if (kappa2 > 0 && P.mg.auto_ops && P.mg.i_eps > 0)
    throw std::invalid_argument("mg_set_helmholtz: not combined with a permittivity");
  P.mg.helm = kappa2;
  API_END
}

int a5g_mg_set_bc_value(a5g_tree* t, int nb, double value) {
  API_BEGIN
  NEED(t);
  FLUSH(t);
  if (nb < 1 || nb > 2 * t->P.tree.dim) throw std::out_of_range("face index out of range");
  t->P.bc_val[nb - 1] = value;
  API_END
}

int a5g_fas_fmg(a5g_tree* t, int set_residual, int first_call) {
  API_BEGIN
  NEED(t);
  a5g::mg_fmg(t->P, set_residual != 0, first_call != 0);
  API_END
}

int a5g_fas_vcycle(a5g_tree* t, int max_lvl, int set_residual) {
  API_BEGIN
  NEED(t);
  a5g::mg_vcycle(t->P, max_lvl, set_residual != 0);
  API_END
}

int a5g_fas_fmg_adjoint(a5g_tree* t, const int* adj, double* vbar) {
  API_BEGIN
  NEED(t);
  FLUSH(t);
  need_gpu(t->P);
  if (!adj || !vbar) throw std::invalid_argument("null argument");
  a5g::fmg_adjoint(t->P, adj, vbar);
  API_END
}

int a5g_fas_vcycle_adjoint(a5g_tree* t, int max_lvl, const int* adj, double* vbar) {
  API_BEGIN
  NEED(t);
  FLUSH(t);
  need_gpu(t->P);
  if (!adj || !vbar) throw std::invalid_argument("null argument");
  a5g::vcycle_adjoint(t->P, max_lvl, adj, vbar);
  API_END
}

int a5g_set_shape(a5g_tree* t, int kernel, int nc, int cls, int shape) {
  API_BEGIN
  NEED(t);
  FLUSH(t);
  auto& P = t->P;
  if (kernel < 0 || kernel >= a5g::kShapeKernels) throw std::out_of_range("unknown kernel");
  if (nc < 2 || nc > 32 || (nc & (nc - 1))) throw std::invalid_argument("nc must be 2..32, a power of 2");
  if (cls < 0 || cls >= a5g::kSizeClasses) throw std::out_of_range("size class out of range");
  if (shape < 0 || shape >= a5g::kShapes) throw std::out_of_range("shape out of range");
  P.clear_graphs();
  P.shape_tab[kernel][P.tree.dim - 2][a5g::log2i(nc)][cls] = (unsigned char)shape;
  API_END
}

int a5g_get_shape(a5g_tree* t, int kernel, int nc, int cls, int* shape) {
  API_BEGIN
  NEED(t);
  if (!shape) throw std::invalid_argument("null output");
  if (kernel < 0 || kernel >= a5g::kShapeKernels) throw std::out_of_range("unknown kernel");
  if (nc < 2 || nc > 32 || (nc & (nc - 1))) throw std::invalid_argument("nc must be 2..32, a power of 2");
  if (cls < 0 || cls >= a5g::kSizeClasses) throw std::out_of_range("size class out of range");
  *shape = t->P.shape_tab[kernel][t->P.tree.dim - 2][a5g::log2i(nc)][cls];
  API_END
}

int a5g_autotune(a5g_tree* t, int reps, int max_rec, double* rec, int* n_rec) {
  API_BEGIN
  NEED(t);
  FLUSH(t);
  need_gpu(t->P);
  if (!n_rec || (max_rec > 0 && !rec)) throw std::invalid_argument("null output");
  std::vector<double> r;
  a5g::autotune(t->P, reps, r);
  const int n = (int)(r.size() / 8);
  *n_rec = n;
  for (int i = 0; i < std::min(n, max_rec); i++)
    for (int k = 0; k < 8; k++) rec[i * 8 + k] = r[i * 8 + k];
  API_END
}

int a5g_mg_solve_mixed(a5g_tree* t, double tol, int max_outer, int n_inner, int zero_start,
                       int* n_outer, double* rel_res) {
  API_BEGIN
  NEED(t);
  FLUSH(t);
  need_gpu(t->P);
  auto& P = t->P;
  if (zero_start) {
    P.sync_topology();
    std::vector<int> all;
    for (int l = P.tree.min_lvl; l <= P.tree.max_lvl; l++)
      all.insert(all.end(), P.lids(l).begin(), P.lids(l).end());
    a5g::fill_boxes(P, (int)all.size(), all.data(), P.mg.i_phi, 0.0);
    a5g::gc_tree(P, P.mg.i_phi, 0, 0);
  }
  int it = 0;
  double rel = 0;
  a5g::mg_solve_mixed(P, tol, max_outer, n_inner, &it, &rel);
  if (n_outer) *n_outer = it;
  if (rel_res) *rel_res = rel;
  API_END
}

int a5g_mg_residual_norms(a5g_tree* t, double* max_res, double* max_rhs) {
  API_BEGIN
  NEED(t);
  FLUSH(t);
  a5g::mg_residual_norms(t->P, max_res, max_rhs);
  API_END
}

int a5g_gc_tree(a5g_tree* t, int iv, int rb, int bc) {
  API_BEGIN
  NEED(t);
  FLUSH(t);
  check_iv(t->P.tree, iv);
  if (rb < 0 || rb > 2 || bc < 0 || bc > 2) throw std::invalid_argument("bad ghost rule");
  a5g::gc_tree(t->P, iv, rb, bc);
  API_END
}

int a5g_restrict_tree(a5g_tree* t, int iv) {
  API_BEGIN
  NEED(t);
  FLUSH(t);
  check_iv(t->P.tree, iv);
  a5g::restrict_tree(t->P, iv);
  API_END
}

int a5g_tree_copy_cc(a5g_tree* t, int iv_from, int iv_to) {
  API_BEGIN
  NEED(t);
  FLUSH(t);
  check_iv(t->P.tree, iv_from);
  check_iv(t->P.tree, iv_to);
  a5g::tree_copy_cc(t->P, iv_from, iv_to);
  API_END
}

int a5g_tree_maxdiff_cc(a5g_tree* t, int iv_a, int iv_b, double* maxdiff, double* maxabs_b) {
  API_BEGIN
  NEED(t);
  FLUSH(t);
  need_gpu(t->P);
  check_iv(t->P.tree, iv_a);
  check_iv(t->P.tree, iv_b);
  a5g::maxdiff_leaves(t->P, iv_a, iv_b, maxdiff, maxabs_b);
  API_END
}

int a5g_tree_minmax_cc(a5g_tree* t, int iv, double* vmin, double* vmax) {
  API_BEGIN
  NEED(t);
  FLUSH(t);
  check_iv(t->P.tree, iv);
  a5g::minmax_boxes(t->P, iv, false, false, true, vmin, vmax);
  API_END
}

// ---- transport table (fosito/m_lookup_table.f90, m_find_index.f90) -------

// FI_adaptive_r (fosito/m_find_index.f90:56-67): smallest 1-based ix with
// val <= list(ix); linear search below 40 entries, bisection above
static int fi_adaptive(int n, const double* list, double val) {
  if (n < 40) {
    int ix;
    for (ix = 1; ix <= n; ix++)
      if (val <= list[ix - 1]) break;
    return ix;
  }
  int lo = 1, hi = n;
  while (lo < hi) {
    const int mid = lo + (hi - lo) / 2;
    if (val <= list[mid - 1]) hi = mid;
    else lo = mid + 1;
  }
  if (val > list[lo - 1]) lo = n + 1;
  return lo;
}

// LT_lin_interp_list (fosito/m_lookup_table.f90:201-221)
static double lin_interp_list(int n, const double* x, const double* y, double xv) {
  if (xv <= x[0]) return y[0];
  if (xv >= x[n - 1]) return y[n - 1];
  const int ix = fi_adaptive(n, x, xv);
  const double temp = (xv - x[ix - 2]) / (x[ix - 1] - x[ix - 2]);
  return (1 - temp) * y[ix - 2] + temp * y[ix - 1];
}

int a5g_table_create(a5g_tree* t, double x_min, double x_max, int n_rows, int n_cols) {
  // LT_create, fosito/m_lookup_table.f90:46-60
  API_BEGIN
  NEED(t);
  if (x_max <= x_min) throw std::invalid_argument("table: x_max should be > x_min");
  if (n_rows <= 1) throw std::invalid_argument("table: n_rows should be bigger than 1");
  if (n_cols < 1) throw std::invalid_argument("table: n_cols should be >= 1");
  auto& P = t->P;
  P.tbl_rows = n_rows;
  P.tbl_cols = n_cols;
  P.tbl_xmin = x_min;
  P.tbl_inv_dx = (n_rows - 1) / (x_max - x_min);
  P.h_table.assign((size_t)n_rows * n_cols, 0.0);
  if (!P.host_only) a5g::set_table(P, n_rows, n_cols, P.tbl_xmin, P.tbl_inv_dx, P.h_table.data());
  API_END
}

int a5g_table_set_col(a5g_tree* t, int col, int n, const double* xx, const double* yy) {
  // LT_set_col / get_spaced_data / LT_get_xdata,
  // fosito/m_lookup_table.f90:62-72, 180-198
  API_BEGIN
  NEED(t);
  auto& P = t->P;
  if (P.h_table.empty()) throw std::logic_error("table: call a5g_table_create first");
  if (col < 1 || col > P.tbl_cols) throw std::out_of_range("table column out of range");
  if (n < 1) throw std::invalid_argument("table: empty data");
  const double dx = 1 / P.tbl_inv_dx;
  for (int ix = 1; ix <= P.tbl_rows; ix++) {
    const double xv = P.tbl_xmin + (ix - 1) * dx;
    P.h_table[(size_t)(col - 1) * P.tbl_rows + ix - 1] = lin_interp_list(n, xx, yy, xv);
  }
  if (!P.host_only)
    a5g::set_table(P, P.tbl_rows, P.tbl_cols, P.tbl_xmin, P.tbl_inv_dx, P.h_table.data());
  API_END
}

int a5g_table_get(a5g_tree* t, double* rows_cols, double* x_min, double* inv_dx) {
  API_BEGIN
  NEED(t);
  auto& P = t->P;
  if (P.h_table.empty()) throw std::logic_error("table: none set");
  std::copy(P.h_table.begin(), P.h_table.end(), rows_cols);
  *x_min = P.tbl_xmin;
  *inv_dx = P.tbl_inv_dx;
  API_END
}

int a5g_set_table(a5g_tree* t, int n_rows, int n_cols, double x_min, double inv_dx,
                  const double* rows_cols) {
  API_BEGIN
  NEED(t);
  auto& P = t->P;
  P.h_table.assign(rows_cols, rows_cols + (size_t)n_rows * n_cols);
  P.tbl_rows = n_rows;
  P.tbl_cols = n_cols;
  P.tbl_xmin = x_min;
  P.tbl_inv_dx = inv_dx;
  if (!P.host_only) a5g::set_table(P, n_rows, n_cols, x_min, inv_dx, rows_cols);
  API_END
}

// ---- streamer physics -------------------------------------------------------

int a5g_streamer_setup(a5g_tree* t, const int* v) {
  API_BEGIN
  NEED(t);
  auto& P = t->P;
  a5g::StVars s;
  s.elec = v[0]; s.pion = v[1]; s.elec_old = v[2]; s.pion_old = v[3]; s.phi = v[4];
  s.fld = v[5]; s.rhs = v[6]; s.pho = v[7]; s.eps = v[8];
  s.f_elec = v[9]; s.f_fld = v[10];
  s.t_mob = v[11]; s.t_dif = v[12]; s.t_alpha = v[13]; s.t_eta = v[14];
  s.dielectric_rule = v[15];
  for (int iv : {s.elec, s.pion, s.elec_old, s.pion_old, s.phi, s.fld, s.rhs, s.eps})
    check_iv(P.tree, iv);
  if (s.pho > 0) check_iv(P.tree, s.pho);
  if (s.f_elec < 1 || s.f_elec > P.tree.n_fc || s.f_fld < 1 || s.f_fld > P.tree.n_fc)
    throw std::out_of_range("face variable out of range");
  P.sv = s;
  API_END
}

static void need_table(const a5g::Pool& P, const a5g::StVars& v) {
  if (!P.d_table) throw std::logic_error("no transport table on the GPU");
  const int mx = std::max(std::max(v.t_mob, v.t_dif), std::max(v.t_alpha, v.t_eta));
  if (mx > P.tbl_cols) throw std::out_of_range("table column out of range");
}

int a5g_space_charge(a5g_tree* t, double fac) {
  API_BEGIN
  NEED(t);
  FLUSH(t);
  need_gpu(t->P);
  a5g::space_charge(t->P, t->P.sv, fac);
  API_END
}

int a5g_fld_from_pot(a5g_tree* t) {
  API_BEGIN
  NEED(t);
  FLUSH(t);
  need_gpu(t->P);
  a5g::fld_from_pot(t->P, t->P.sv);
  API_END
}

int a5g_fluxes_koren(a5g_tree* t, double dt) {
  API_BEGIN
  NEED(t);
  FLUSH(t);
  need_gpu(t->P);
  need_table(t->P, t->P.sv);
  if (!(dt > 0)) throw std::invalid_argument("fluxes_koren: dt must be > 0");
  a5g::fluxes_koren(t->P, t->P.sv, dt);
  API_END
}

int a5g_consistent_fluxes(a5g_tree* t, int n, const int* f_ixs) {
  API_BEGIN
  NEED(t);
  FLUSH(t);
  need_gpu(t->P);
  for (int i = 0; i < n; i++) {
    if (f_ixs[i] < 1 || f_ixs[i] > t->P.tree.n_fc)
      throw std::out_of_range("face variable out of range");
    a5g::consistent_fluxes(t->P, f_ixs[i]);
  }
  API_END
}

int a5g_update_solution(a5g_tree* t, double dt) {
  API_BEGIN
  NEED(t);
  FLUSH(t);
  need_gpu(t->P);
  need_table(t->P, t->P.sv);
  a5g::update_solution(t->P, t->P.sv, dt);
  API_END
}

int a5g_photoi_helmholtz(a5g_tree* t, double coeff, int n_terms, const double* amp,
                         const double* kappa, int i_phi, int i_rhs, int i_tmp, double tol,
                         int max_cycles, int* cycles) {
  API_BEGIN
  NEED(t);
  FLUSH(t);
  auto& P = t->P;
  need_gpu(P);
  need_table(P, P.sv);
  if (P.sv.pho <= 0) throw std::invalid_argument("photoionization: i_pho not set");
  if (!amp || !kappa) throw std::invalid_argument("photoionization: null coefficients");
  check_iv(P.tree, i_phi);
  check_iv(P.tree, i_rhs);
  check_iv(P.tree, i_tmp);
  a5g::photoi_helmholtz(P, P.sv, coeff, n_terms, amp, kappa, i_phi, i_rhs, i_tmp, tol,
                        max_cycles, cycles);
  API_END
}

int a5g_photoi_helmholtz_terms(a5g_tree* t, double coeff, int n_terms, const double* amp,
                               const double* kappa, const int* i_terms, const int* i_rhs,
                               const int* i_tmp, int warm, double tol, int max_cycles,
                               int* cycles) {
  API_BEGIN
  NEED(t);
  FLUSH(t);
  auto& P = t->P;
  need_gpu(P);
  need_table(P, P.sv);
  if (P.sv.pho <= 0) throw std::invalid_argument("photoionization: i_pho not set");
  if (!amp || !kappa || !i_terms || !i_rhs || !i_tmp)
    throw std::invalid_argument("photoionization: null coefficients or variables");
  if (n_terms < 1 || n_terms > a5g::kMaxTerms)
    throw std::invalid_argument("photoionization: need 1 to 4 terms");
  for (int j = 0; j < n_terms; j++) {
    check_iv(P.tree, i_terms[j]);
    check_iv(P.tree, i_rhs[j]);
    check_iv(P.tree, i_tmp[j]);
  }
  a5g::photoi_helmholtz_terms(P, P.sv, coeff, n_terms, amp, kappa, i_terms, i_rhs, i_tmp,
                              warm != 0, tol, max_cycles, cycles);
  API_END
}

int a5g_average_dens(a5g_tree* t) {
  API_BEGIN
  NEED(t);
  FLUSH(t);
  need_gpu(t->P);
  a5g::average_dens(t->P, t->P.sv);
  API_END
}

int a5g_box_summaries(a5g_tree* t, double* out) {
  API_BEGIN
  NEED(t);
  FLUSH(t);
  need_gpu(t->P);
  a5g::box_summaries(t->P, t->P.sv, out);
  API_END
}

int a5g_set_ref_rules(a5g_tree* t, int n_seed, const double* seeds, int n_reg,
                      const double* regions, double init_window, double crv_refine,
                      int n_base_last) {
  API_BEGIN
  NEED(t);
  if (n_seed < 0 || n_reg < 0 || (n_seed > 0 && !seeds) || (n_reg > 0 && !regions))
    throw std::invalid_argument("bad seed or region list");
  if (n_base_last < 1) throw std::invalid_argument("n_base_last must be >= 1");
  a5g::set_ref_rules(t->P, n_seed, seeds, n_reg, regions, init_window, crv_refine, n_base_last);
  API_END
}

int a5g_ref_flags(a5g_tree* t, const double* times, int* flags) {
  API_BEGIN
  NEED(t);
  FLUSH(t);
  need_gpu(t->P);
  if (!times || !flags) throw std::invalid_argument("null argument");
  a5g::ref_flags(t->P, t->P.sv, times, flags);
  API_END
}

int a5g_instance_maxima(a5g_tree* t, int iv, double* vmax, int* lvl_max) {
  API_BEGIN
  NEED(t);
  FLUSH(t);
  need_gpu(t->P);
  check_iv(t->P.tree, iv);
  if (!vmax || !lvl_max) throw std::invalid_argument("null argument");
  a5g::instance_maxima(t->P, iv, vmax, lvl_max);
  API_END
}

static a5g::AdjVars adj_vars(const a5g::Tree& T, const int* adj, bool photoi) {
  if (!adj) throw std::invalid_argument("adjoint variables not given");
  a5g::AdjVars av;
  av.le = adj[0];
  av.li = adj[1];
  av.lE = adj[2];
  av.lpho = adj[3];
  av.lF = adj[4];
  av.lfld = adj[5];
  check_iv(T, av.le);
  check_iv(T, av.li);
  check_iv(T, av.lE);
  if (photoi) check_iv(T, av.lpho);
  for (int f : {av.lF, av.lfld})
    if (f < 1 || f > T.n_fc) throw std::out_of_range("face variable out of range");
  return av;
}

int a5g_update_solution_adjoint(a5g_tree* t, double dt, const int* adj) {
  API_BEGIN
  NEED(t);
  FLUSH(t);
  need_gpu(t->P);
  a5g::update_solution_adj(t->P, t->P.sv, dt, adj_vars(t->P.tree, adj, t->P.sv.pho > 0));
  API_END
}

int a5g_consistent_fluxes_adjoint(a5g_tree* t, int lF) {
  API_BEGIN
  NEED(t);
  FLUSH(t);
  need_gpu(t->P);
  if (lF < 1 || lF > t->P.tree.n_fc) throw std::out_of_range("face variable out of range");
  a5g::consistent_fluxes_adj(t->P, lF);
  API_END
}

int a5g_fluxes_koren_adjoint(a5g_tree* t, double dt, const int* adj) {
  API_BEGIN
  NEED(t);
  FLUSH(t);
  need_gpu(t->P);
  a5g::fluxes_koren_adj(t->P, t->P.sv, dt, adj_vars(t->P.tree, adj, false));
  API_END
}

int a5g_table_adjoint(a5g_tree* t, int clear, double* rows_cols) {
  API_BEGIN
  NEED(t);
  need_gpu(t->P);
  a5g::table_adjoint(t->P, clear != 0, rows_cols);
  API_END
}

int a5g_tree_copy_cc_adjoint(a5g_tree* t, int lfrom, int lto) {
  API_BEGIN
  NEED(t);
  FLUSH(t);
  need_gpu(t->P);
  a5g::copy_cc_adj(t->P, lfrom, lto);
  API_END
}

int a5g_average_dens_adjoint(a5g_tree* t, const int* adj) {
  API_BEGIN
  NEED(t);
  FLUSH(t);
  need_gpu(t->P);
  if (!adj) throw std::invalid_argument("null argument");
  a5g::average_dens_adj(t->P, adj[0], adj[1], adj[2], adj[3]);
  API_END
}

int a5g_space_charge_adjoint(a5g_tree* t, double fac, const int* adj) {
  API_BEGIN
  NEED(t);
  FLUSH(t);
  need_gpu(t->P);
  if (!adj) throw std::invalid_argument("null argument");
  a5g::space_charge_adj(t->P, fac, adj[0], adj[1], adj[2]);
  API_END
}

int a5g_restrict_tree_adjoint(a5g_tree* t, int il) {
  API_BEGIN
  NEED(t);
  FLUSH(t);
  need_gpu(t->P);
  a5g::restrict_tree_adj(t->P, il);
  API_END
}

int a5g_gc_tree_adjoint(a5g_tree* t, int iv, int il, int rb, int bc) {
  API_BEGIN
  NEED(t);
  FLUSH(t);
  need_gpu(t->P);
  a5g::gc_tree_adj(t->P, iv, il, rb, bc);
  API_END
}

int a5g_fld_from_pot_adjoint(a5g_tree* t, const int* adj) {
  API_BEGIN
  NEED(t);
  FLUSH(t);
  need_gpu(t->P);
  if (!adj) throw std::invalid_argument("null argument");
  a5g::fld_from_pot_adj(t->P, t->P.sv, adj[0], adj[1], adj[2]);
  API_END
}

int a5g_state_store(a5g_tree* t, int slot, int n_cc, const int* ivs, int n_fc, const int* fvs,
                    int op) {
  API_BEGIN
  NEED(t);
  FLUSH(t);
  need_gpu(t->P);
  a5g::state_store(t->P, slot, n_cc, ivs, n_fc, fvs, op);
  API_END
}

int a5g_box_summaries_noalloc(a5g_tree* t, double* out) {
  API_BEGIN
  NEED(t);
  FLUSH(t);
  need_gpu(t->P);
  if (!out) throw std::invalid_argument("box_summaries_noalloc: null output");
  a5g::box_summaries_noalloc(t->P, t->P.sv, out);
  API_END
}

int a5g_get_loc(a5g_tree* t, const double* r, int max_lvl, int* id, int* ix) {
  API_BEGIN
  NEED(t);
  if (!r || !id || !ix) throw std::invalid_argument("get_loc: null argument");
  if (!t->P.tree.ready) throw std::runtime_error("get_loc: the base is not set");
  *id = t->P.tree.get_loc(r, max_lvl, ix);
  API_END
}

int a5g_reduction_loc(a5g_tree* t, int kind, int var, int d, double thr, double init_val,
                      double* val, int* id, int* ix) {
  API_BEGIN
  NEED(t);
  FLUSH(t);
  need_gpu(t->P);
  const auto& T = t->P.tree;
  if (!val || !id || !ix) throw std::invalid_argument("reduction_loc: null argument");
  if (d < 1 || d > T.dim) throw std::out_of_range("reduction_loc: dimension out of range");
  if (kind == 1) {
    if (var < 1 || var > T.n_fc)
      throw std::out_of_range("face variable " + std::to_string(var) + " out of range");
  } else if (kind == 2) {
    check_iv(T, var);
  } else {
    throw std::invalid_argument("reduction_loc: kind must be 1 or 2");
  }
  a5g::reduction_loc(t->P, kind, var, d - 1, thr, init_val, val, id, ix);
  API_END
}

int a5g_cc_axis(a5g_tree* t, int n_cc, const int* ixs_cc, int n_fc, const int* ixs_fc, int cap,
                double* out, int* n_points) {
  API_BEGIN
  NEED(t);
  FLUSH(t);
  const auto& T = t->P.tree;
  if (!n_points) throw std::invalid_argument("cc_axis: null n_points");
  if (n_cc < 0 || n_fc < 0 || (n_cc && !ixs_cc) || (n_fc && !ixs_fc))
    throw std::invalid_argument("cc_axis: bad variable lists");
  for (int v = 0; v < n_cc; v++) check_iv(T, ixs_cc[v]);
  for (int v = 0; v < n_fc; v++)
    if (ixs_fc[v] < 1 || ixs_fc[v] > T.n_fc)
      throw std::out_of_range("face variable " + std::to_string(ixs_fc[v]) + " out of range");
  if (out) need_gpu(t->P);
  *n_points = a5g::cc_axis(t->P, n_cc, ixs_cc, n_fc, ixs_fc, cap, out);
  API_END
}

static void check_out_vars(const a5g::Tree& T, int n_cc, const int* ixs_cc, int n_fc,
                           const int* ixs_fc) {
  if (n_cc < 0 || n_fc < 0 || (n_cc && !ixs_cc) || (n_fc && !ixs_fc))
    throw std::invalid_argument("bad variable lists");
  for (int v = 0; v < n_cc; v++) check_iv(T, ixs_cc[v]);
  for (int v = 0; v < n_fc; v++)
    if (ixs_fc[v] < 1 || ixs_fc[v] > T.n_fc)
      throw std::out_of_range("face variable " + std::to_string(ixs_fc[v]) + " out of range");
}

int a5g_leaf_cells(a5g_tree* t, int n_cc, const int* ixs_cc, int n_fc, const int* ixs_fc,
                   long long cap, double* out, long long* n_cells) {
  API_BEGIN
  NEED(t);
  FLUSH(t);
  const auto& T = t->P.tree;
  if (!n_cells) throw std::invalid_argument("leaf_cells: null n_cells");
  check_out_vars(T, n_cc, ixs_cc, n_fc, ixs_fc);
  long long n_leaves = 0;
  if (t->P.n_inst && !t->P.host_only) t->P.sync_topology();
  for (int l = 1; l <= T.max_lvl; l++) n_leaves += (long long)t->P.lleaves(l).size();
  const long long cpb = T.dim == 2 ? (long long)T.n_cell * T.n_cell
                                   : (long long)T.n_cell * T.n_cell * T.n_cell;
  *n_cells = n_leaves * cpb;
  if (!out || cap == 0) return 0;
  need_gpu(t->P);
  const long long cols = n_cc + (n_fc ? T.dim : 0);
  if (*n_cells * cols > cap)
    throw std::length_error("leaf_cells: " + std::to_string(*n_cells * cols) +
                            " values, room for " + std::to_string(cap));
  std::vector<double> h;
  a5g::leaf_cells(t->P, n_cc, ixs_cc, n_fc, ixs_fc, h);
  std::copy(h.begin(), h.end(), out);
  API_END
}

int a5g_write_vtk(a5g_tree* t, const char* filename, int n_cc, const char* const* cc_names,
                  const int* ixs_cc, int n_fc, const char* const* fc_names, const int* ixs_fc,
                  int n_cycle, double time) {
  API_BEGIN
  NEED(t);
  FLUSH(t);
  need_gpu(t->P);
  const auto& T = t->P.tree;
  if (!T.ready) throw std::runtime_error("Tree not ready");
  if (!filename) throw std::invalid_argument("write_vtk: null filename");
  std::vector<int> all;
  if (!ixs_cc) {
    if (n_cc != T.n_cc)
      throw std::invalid_argument("a$D_write_vtk: size(cc_names) /= n_var_cell");
    for (int v = 1; v <= T.n_cc; v++) all.push_back(v);
    ixs_cc = all.data();
  }
  check_out_vars(T, n_cc, ixs_cc, n_fc, ixs_fc);
  if ((n_cc && !cc_names) || (n_fc && !fc_names))
    throw std::invalid_argument("write_vtk: null names");
  a5g::write_vtk(t->P, filename, n_cc, cc_names, ixs_cc, n_fc, fc_names, ixs_fc, n_cycle, time);
  API_END
}

int a5g_dump_leaves(a5g_tree* t, const char* filename, int n_var, const int* ivs, int cycle,
                    double time) {
  API_BEGIN
  NEED(t);
  FLUSH(t);
  need_gpu(t->P);
  if (!filename) throw std::invalid_argument("dump_leaves: null filename");
  check_out_vars(t->P.tree, n_var, ivs, 0, nullptr);
  a5g::dump_leaves(t->P, filename, n_var, ivs, cycle, time);
  API_END
}

int a5g_set_instances(a5g_tree* t, int n_inst, double stride) {
  API_BEGIN
  NEED(t);
  FLUSH(t);
  auto& P = t->P;
  if (n_inst < 0) throw std::invalid_argument("set_instances: negative count");
  if (n_inst > 0 && !(stride > 0)) throw std::invalid_argument("set_instances: stride must be > 0");
  if (P.d_inst) a5g::hcheck(hipFree(P.d_inst), "hipFree");
  P.d_inst = nullptr;
  P.n_inst = n_inst;
  P.inst_group.clear();
  P.n_groups = 0;
  P.inst_stride = n_inst ? stride : 0.0;
  P.h_inst.assign((size_t)n_inst * a5g::kInstVals, 0.0);
  P.inst_active.assign(n_inst, 1);
  if (n_inst && !P.host_only) {
    a5g::hcheck(hipMalloc((void**)&P.d_inst, P.h_inst.size() * sizeof(double)), "hipMalloc");
    a5g::hcheck(hipMemcpyAsync(P.d_inst, P.h_inst.data(), P.h_inst.size() * sizeof(double),
                          hipMemcpyHostToDevice, P.stream), "instances");
    a5g::hcheck(hipStreamSynchronize(P.stream), "instances sync");
  }
  P.topo_dirty = true;
  API_END
}

int a5g_set_instance_values(a5g_tree* t, int what, int n, const double* values) {
  API_BEGIN
  NEED(t);
  FLUSH(t);
  auto& P = t->P;
  need_gpu(P);
  if (!P.n_inst) throw std::logic_error("set_instance_values: no instances set");
  if (n != P.n_inst || !values) throw std::invalid_argument("set_instance_values: one value per instance");
  if (what < 0 || what > 2 * P.tree.dim) throw std::out_of_range("set_instance_values: what");
  // the constants and the expression of fluxes_koren's clamp factor
  const double eps0 = 8.8541878176e-12, qe = 1.6022e-19;
  for (int g = 0; g < n; g++) {
    double* v = &P.h_inst[(size_t)g * a5g::kInstVals];
    if (what == 0) {
      v[a5g::kInstDt] = values[g];
      v[a5g::kInstKorenFac] = -0.8 * eps0 / (qe * values[g]);
    } else {
      v[a5g::kInstBc + what - 1] = values[g];
    }
  }
  a5g::hcheck(hipMemcpyAsync(P.d_inst, P.h_inst.data(), P.h_inst.size() * sizeof(double),
                        hipMemcpyHostToDevice, P.stream), "instance values");
  a5g::hcheck(hipStreamSynchronize(P.stream), "instance values sync");
  API_END
}

int a5g_set_active_instances(a5g_tree* t, int n, const int* active) {
  API_BEGIN
  NEED(t);
  FLUSH(t);
  auto& P = t->P;
  if (!P.n_inst) throw std::logic_error("set_active_instances: no instances set");
  if (active && n != P.n_inst) throw std::invalid_argument("set_active_instances: one flag per instance");
  for (int g = 0; g < P.n_inst; g++) P.inst_active[g] = active ? (active[g] != 0) : 1;
  P.lists_dirty = true;
  API_END
}

int a5g_set_instance_groups(a5g_tree* t, int n, const int* group) {
  API_BEGIN
  NEED(t);
  FLUSH(t);
  auto& P = t->P;
  if (n != 0 && n != P.n_inst)
    throw std::invalid_argument("instance groups: need one entry per instance (or n = 0)");
  if (n > 0 && !group) throw std::invalid_argument("instance groups: null array");
  std::vector<int> g(group, group + n);
  int ng = 0;
  for (int v : g) {
    if (v < -1) throw std::invalid_argument("instance groups: a group is -1 or >= 0");
    ng = std::max(ng, v + 1);
  }
  if (g != P.inst_group) {
    P.inst_group = g;
    P.n_groups = ng;
    P.lists_dirty = true;
  }
  API_END
}

int a5g_fas_fmg_groups(a5g_tree* t, int set_residual, int first_call) {
  API_BEGIN
  NEED(t);
  FLUSH(t);
  need_gpu(t->P);
  a5g::mg_fmg_groups(t->P, set_residual != 0, first_call != 0);
  API_END
}

int a5g_prolong1_to(a5g_tree* t, int n_ids, const int* ids, int iv) {
  API_BEGIN
  NEED(t);
  FLUSH(t);
  need_gpu(t->P);
  check_iv(t->P.tree, iv);
  a5g::prolong1_boxes(t->P, n_ids, ids, iv);
  API_END
}

int a5g_fill_cc(a5g_tree* t, int n_ids, const int* ids, int iv, double value) {
  API_BEGIN
  NEED(t);
  FLUSH(t);
  need_gpu(t->P);
  check_iv(t->P.tree, iv);
  a5g::fill_boxes(t->P, n_ids, ids, iv, value);
  API_END
}

int a5g_gc_boxes(a5g_tree* t, int n_ids, const int* ids, int iv, int rb, int bc) {
  API_BEGIN
  NEED(t);
  FLUSH(t);
  need_gpu(t->P);
  check_iv(t->P.tree, iv);
  if (rb < 0 || rb > 2 || bc < 0 || bc > 2) throw std::invalid_argument("bad ghost rule");
  a5g::gc_boxes(t->P, n_ids, ids, iv, rb, bc);
  API_END
}

int a5g_set_gaussians(a5g_tree* t, int iv, int n_blob, const double* blobs, double base) {
  API_BEGIN
  NEED(t);
  FLUSH(t);
  need_gpu(t->P);
  check_iv(t->P.tree, iv);
  if (n_blob < 0 || (n_blob > 0 && !blobs)) throw std::invalid_argument("bad blob list");
  a5g::set_gaussians(t->P, iv, n_blob, blobs, base);
  API_END
}

int a5g_set_gaussians_instances(a5g_tree* t, int iv, int n_blob, const double* blobs,
                                const double* base) {
  API_BEGIN
  NEED(t);
  FLUSH(t);
  need_gpu(t->P);
  check_iv(t->P.tree, iv);
  if (t->P.n_inst == 0) throw std::invalid_argument("no instances set");
  if (n_blob < 0 || (n_blob > 0 && !blobs) || !base) throw std::invalid_argument("bad blob list");
  a5g::set_gaussians(t->P, iv, n_blob, blobs, 0.0, base);
  API_END
}

int a5g_prof_enable(a5g_tree* t, int on) {
  API_BEGIN
  NEED(t);
  need_gpu(t->P);
  auto& P = t->P;
  a5g::hcheck(hipStreamSynchronize(P.stream), "sync");
  for (auto& r : P.prof) {
    P.ev_pool.push_back(r.a);
    P.ev_pool.push_back(r.b);
  }
  P.prof.clear();
  P.prof_on = on != 0;
  API_END
}

int a5g_prof_read(a5g_tree* t, int kind, int lvl, int* n_launch, double* ms, double* bytes) {
  API_BEGIN
  NEED(t);
  need_gpu(t->P);
  auto& P = t->P;
  a5g::hcheck(hipStreamSynchronize(P.stream), "sync");
  int cnt = 0;
  double tot = 0, by = 0;
  for (auto& r : P.prof) {
    if (r.kind != kind || (lvl != -999 && r.lvl != lvl)) continue;
    float m = 0;
    a5g::hcheck(hipEventElapsedTime(&m, r.a, r.b), "hipEventElapsedTime");
    cnt += r.count;
    tot += m;
    by += r.bytes;
  }
  *n_launch = cnt;
  *ms = tot;
  *bytes = by;
  API_END
}

}  // extern "C"
