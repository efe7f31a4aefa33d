// Streamer diagnostics and output of the a5g library: the located maximum
// over the leaves (a$D_reduction_loc, afivo/src/m_aX_utils.f90:499-555) with
// the per-box procedures of get_streamer_properties (src/streamer_cyl.f90:
// 831-903), the axis profile of get_cc_axis (:905-956), and the leaf cells
// of a$D_write_vtk (afivo/src/m_aX_io.f90:20-209) written as a VTK
// unstructured grid or as a raw leaf-box dump.
//
// The reduction runs in two launches: one workgroup per leaf box finds the
// box's value and cell, then one workgroup picks the box. Both keep the
// outcome of the reference's loop on one thread: inside a box the first cell
// in array order (x fastest) of the largest value, across boxes the first
// leaf (levels upwards, leaves in list order) whose value is the largest,
// provided it exceeds the initial value.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "a5g_internal.hpp"
#include "a5g_kernels.hpp"

namespace a5g {

static void chk_d(hipError_t e, const char* what) {
  if (e != hipSuccess) throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

// (value, index) pairs: the larger value wins, the smaller index on a tie;
// index -1 marks "no candidate" and loses to any candidate
__device__ __forceinline__ void rl_take(double& bv, int& bi, double v, int i) {
  if (i < 0) return;
  if (bi < 0 || v > bv || (v == bv && i < bi)) {
    bv = v;
    bi = i;
  }
}

template <int NT>
__device__ __forceinline__ void rl_block(double& bv, int& bi) {
  for (int o = 32; o > 0; o >>= 1) {
    const double v = __shfl_xor(bv, o, 64);
    const int i = __shfl_xor(bi, o, 64);
    rl_take(bv, bi, v, i);
  }
  __shared__ double sv[NT / 64];
  __shared__ int si[NT / 64];
  if ((threadIdx.x & 63) == 0) {
    sv[threadIdx.x >> 6] = bv;
    si[threadIdx.x >> 6] = bi;
  }
  __syncthreads();
  if (threadIdx.x == 0)
    for (int w = 1; w < NT / 64; w++) rl_take(bv, bi, sv[w], si[w]);
}

// One workgroup per leaf: kind 1 = 0.5 |f(p) + f(p + e_d)| of face variable
// var (box_maxfld_r/_z), kind 2 = the cell-centre coordinate d (a$D_r_cc,
// afivo/src/m_aX_utils.f90:309-314) over cells with cc(var) > thr, where
// only coordinates above 0 count (box_fld_maxr/_maxz start from val = 0).
// Writes the box value and linear cell index (-1: none) at slot base + li.
template <int D>
__global__ __launch_bounds__(256) void k_rloc_box(Dev a, const int* __restrict__ list, int n,
                                                  int nc, int kind, int var, int d, double thr,
                                                  int base, double* __restrict__ vals,
                                                  int* __restrict__ cells) {
  const int li = blockIdx.x;
  if (li >= n) return;
  const int id = list[li];
  const int total = D == 2 ? nc * nc : nc * nc * nc;
  double bv = 0;
  int bi = -1;
  if (kind == 1) {
    const double* f = facep(a, var, id, d, D);
    for (int q = threadIdx.x; q < total; q += blockDim.x) {
      int p[3] = {q % nc + 1, (q / nc) % nc + 1, D == 3 ? q / (nc * nc) + 1 : 1};
      const double lo = f[fidx<D>(nc, d, p[0], p[1], p[2])];
      p[d]++;
      const double hi = f[fidx<D>(nc, d, p[0], p[1], p[2])];
      rl_take(bv, bi, fabs(lo + hi), q);
    }
  } else {
    const double* c = boxp(a, var, id);
    const double r0 = a.rmin[id * 3 + d], h = a.dr[id];
    for (int q = threadIdx.x; q < total; q += blockDim.x) {
      const int p[3] = {q % nc + 1, (q / nc) % nc + 1, D == 3 ? q / (nc * nc) + 1 : 0};
      if (c[cidx_in<D>(nc, p[0], p[1], p[2])] > thr) {
        const double rc = r0 + (p[d] - 0.5) * h;
        if (rc > 0) rl_take(bv, bi, rc, q);
      }
    }
  }
  rl_block<256>(bv, bi);
  if (threadIdx.x == 0) {
    vals[base + li] = bi < 0 ? 0.0 : (kind == 1 ? 0.5 * bv : bv);
    cells[base + li] = bi;
  }
}

// One workgroup over the n box results in leaf order: the first box of the
// largest value, if that value exceeds init_val. out = (value, slot, cell);
// a box without a cell (kind 2, nothing above thr) still has its value 0.
__global__ __launch_bounds__(1024) void k_rloc_pick(const double* __restrict__ vals,
                                                    const int* __restrict__ cells, int n,
                                                    double init_val, double* __restrict__ out) {
  double bv = 0;
  int bi = -1;
  for (int s = threadIdx.x; s < n; s += blockDim.x) rl_take(bv, bi, vals[s], s);
  rl_block<1024>(bv, bi);
  if (threadIdx.x == 0) {
    const bool won = bi >= 0 && bv > init_val;
    out[0] = won ? bv : init_val;
    out[1] = won ? (double)bi : -1.0;
    out[2] = won ? (double)cells[bi] : -1.0;
  }
}

void reduction_loc(Pool& P, int kind, int var, int d, double thr, double init_val, double* val,
                   int* id_out, int* ix) {
  P.sync_topology();
  const auto& T = P.tree;
  const int D = T.dim;
  std::vector<int> order;  // leaf ids in the reference's order
  for (int l = T.min_lvl; l <= T.max_lvl; l++)
    for (int id : P.lleaves(l)) order.push_back(id);
  const int n = (int)order.size();
  for (int e = 0; e < D; e++) ix[e] = -1;
  *val = init_val;
  *id_out = -1;
  if (!n) return;
  P.ensure_diag(n);
  double* vals = P.d_diag;
  int* cells = reinterpret_cast<int*>(P.d_diag + P.diag_cap);
  double* out = P.d_diag + P.diag_cap + (P.diag_cap + 1) / 2;
  int base = 0;
  for (int l = T.min_lvl; l <= T.max_lvl; l++) {
    const int m = P.n_leaves(l);
    if (!m) continue;
    if (D == 2)
      hipLaunchKernelGGL((k_rloc_box<2>), dim3(m), dim3(256), 0, P.stream, P.dev(), P.leaves(l), m,
                         P.nc_lvl(l), kind, var, d, thr, base, vals, cells);
    else
      hipLaunchKernelGGL((k_rloc_box<3>), dim3(m), dim3(256), 0, P.stream, P.dev(), P.leaves(l), m,
                         P.nc_lvl(l), kind, var, d, thr, base, vals, cells);
    chk_d(hipGetLastError(), "reduction_loc");
    base += m;
  }
  hipLaunchKernelGGL(k_rloc_pick, dim3(1), dim3(1024), 0, P.stream, vals, cells, n, init_val, out);
  chk_d(hipGetLastError(), "reduction_loc pick");
  double h[3];
  chk_d(hipMemcpyAsync(h, out, sizeof h, hipMemcpyDeviceToHost, P.stream), "reduction_loc copy");
  chk_d(hipStreamSynchronize(P.stream), "reduction_loc sync");
  *val = h[0];
  const int slot = (int)h[1];
  if (slot < 0) return;
  const int id = order[slot];
  const int nc = T.ncell[id];
  const int q = (int)h[2];
  *id_out = id;
  if (q < 0) return;
  ix[0] = q % nc + 1;
  ix[1] = (q / nc) % nc + 1;
  if (D == 3) ix[2] = q / (nc * nc) + 1;
}

// get_cc_axis gather: per listed box and row j, cc(1, j, ixs_cc) and
// 0.5 (fy(1, j) + fy(1, j+1)) of ixs_fc into out[(b*nc + j-1)*nv + v]
__global__ __launch_bounds__(256) void k_axis(Dev a, const int* __restrict__ ids, int n, int nc,
                                              const int* __restrict__ vars, int n_cc, int n_fc,
                                              double* __restrict__ out) {
  const int nv = n_cc + n_fc;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * nc * nv) return;
  const int v = t % nv, j = (t / nv) % nc + 1, b = t / (nv * nc);
  const int id = ids[b];
  double r;
  if (v < n_cc) {
    r = boxp(a, vars[v], id)[cidx_in<2>(nc, 1, j, 0)];
  } else {
    const double* f = facep(a, vars[v], id, 1, 2);
    r = 0.5 * (f[fidx<2>(nc, 1, 1, j, 0)] + f[fidx<2>(nc, 1, 1, j + 1, 0)]);
  }
  out[t] = r;
}

int cc_axis(Pool& P, int n_cc, const int* ixs_cc, int n_fc, const int* ixs_fc, int cap,
            double* out) {
  const auto& T = P.tree;
  if (T.dim != 2) throw std::invalid_argument("cc_axis: the axis profile is 2D");
  const int nc = T.n_cell;
  std::vector<int> boxes;
  double r[2] = {0.0, 0.0};
  for (;;) {
    int ixl[2];
    const int id = T.get_loc(r, 0, ixl);
    if (id == -1) break;
    boxes.push_back(id);
    r[1] = T.rmin[id * 3 + 1] + (nc + 1) * T.dr[id];
  }
  const int n_pts = (int)boxes.size() * nc;
  if (!out || cap == 0 || boxes.empty()) return n_pts;
  if (n_pts > cap)
    throw std::length_error("cc_axis: " + std::to_string(n_pts) + " points, room for " +
                            std::to_string(cap));
  P.sync_topology();
  const int nv = n_cc + n_fc, rows = nv + 1;
  std::vector<int> hv(ixs_cc, ixs_cc + n_cc);
  hv.insert(hv.end(), ixs_fc, ixs_fc + n_fc);
  const size_t n_out = (size_t)n_pts * nv;
  // scratch: the gathered values, then the box ids and variable numbers
  P.ensure_diag((int)(n_out + boxes.size() + hv.size()));
  double* d_out = P.d_diag;
  int* d_ids = reinterpret_cast<int*>(P.d_diag + n_out);
  int* d_vars = d_ids + boxes.size();
  chk_d(hipMemcpyAsync(d_ids, boxes.data(), boxes.size() * sizeof(int), hipMemcpyHostToDevice,
                       P.stream), "cc_axis ids");
  if (nv)
    chk_d(hipMemcpyAsync(d_vars, hv.data(), hv.size() * sizeof(int), hipMemcpyHostToDevice,
                         P.stream), "cc_axis vars");
  std::vector<double> h(n_out);
  if (nv) {
    const int total = (int)n_out;
    hipLaunchKernelGGL(k_axis, dim3((total + 255) / 256), dim3(256), 0, P.stream, P.dev(), d_ids,
                       (int)boxes.size(), nc, d_vars, n_cc, n_fc, d_out);
    chk_d(hipGetLastError(), "cc_axis");
    chk_d(hipMemcpyAsync(h.data(), d_out, n_out * sizeof(double), hipMemcpyDeviceToHost,
                         P.stream), "cc_axis copy");
  }
  chk_d(hipStreamSynchronize(P.stream), "cc_axis sync");
  for (size_t b = 0; b < boxes.size(); b++) {
    const int id = boxes[b];
    const double box_z = T.rmin[id * 3 + 1], box_dz = T.dr[id];
    for (int j = 1; j <= nc; j++) {
      const size_t p = b * nc + (j - 1);
      double* o = out + p * rows;
      o[0] = box_z + (j - 0.5) * box_dz;
      for (int v = 0; v < nv; v++) o[1 + v] = h[p * nv + v];
    }
  }
  return n_pts;
}

}  // namespace a5g

// ---------------------------------------------------------------------------
// Output: the interior cells of the leaves in the order of a$D_write_vtk
// (afivo/src/m_aX_io.f90:20-209): leaves level by level in list order, cells
// x fastest. One column of n_cells values per cell variable, then, with face
// variables, one column per dimension d holding 0.5 times the sum, in array
// element order, of the section f_d(two faces of the cell along d, ifc_used).
// ---------------------------------------------------------------------------

namespace a5g {

template <int D>
__global__ __launch_bounds__(256) void k_leaf_cells(Dev a, const int* __restrict__ list, int n,
                                                    int nc, long long base,
                                                    const int* __restrict__ vars, int n_cc,
                                                    int n_fc, long long n_cells,
                                                    double* __restrict__ out) {
  const int cpb = D == 2 ? nc * nc : nc * nc * nc;
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)n * cpb) return;
  const int li = (int)(t / cpb), q = (int)(t - (long long)li * cpb);
  const int id = list[li];
  const int p[3] = {q % nc + 1, (q / nc) % nc + 1, D == 3 ? q / (nc * nc) + 1 : 1};
  const long long c = base + t;
  for (int v = 0; v < n_cc; v++)
    out[v * n_cells + c] = boxp(a, vars[v], id)[cidx_in<D>(nc, p[0], p[1], D == 3 ? p[2] : 0)];
  if (!n_fc) return;
  for (int d = 0; d < D; d++) {
    int h[3] = {p[0], p[1], p[2]};
    h[d]++;
    double s = 0.0;
    for (int v = 0; v < n_fc; v++) {
      const double* f = facep(a, vars[n_cc + v], id, d, D);
      s = s + f[fidx<D>(nc, d, p[0], p[1], p[2])];
      s = s + f[fidx<D>(nc, d, h[0], h[1], h[2])];
    }
    out[(n_cc + d) * n_cells + c] = 0.5 * s;
  }
}

long long leaf_cells(Pool& P, int n_cc, const int* ixs_cc, int n_fc, const int* ixs_fc,
                     std::vector<double>& host) {
  P.sync_topology();
  const auto& T = P.tree;
  const int D = T.dim, nc = T.n_cell;
  const int cpb = D == 2 ? nc * nc : nc * nc * nc;
  long long n_leaves = 0;
  for (int l = 1; l <= T.max_lvl; l++) n_leaves += P.n_leaves(l);
  const long long n_cells = n_leaves * cpb;
  const int cols = n_cc + (n_fc ? D : 0);
  host.assign((size_t)(n_cells * cols), 0.0);
  if (!n_cells || !cols) return n_cells;
  const long long n_out = n_cells * cols;
  const int nv = n_cc + n_fc;
  if (n_out + nv > (long long)INT32_MAX) throw std::length_error("leaf_cells: output too large");
  P.ensure_diag((int)(n_out + nv));
  double* d_out = P.d_diag;
  int* d_vars = reinterpret_cast<int*>(P.d_diag + n_out);
  std::vector<int> hv(ixs_cc, ixs_cc + n_cc);
  hv.insert(hv.end(), ixs_fc, ixs_fc + n_fc);
  chk_d(hipMemcpyAsync(d_vars, hv.data(), hv.size() * sizeof(int), hipMemcpyHostToDevice,
                       P.stream), "leaf_cells vars");
  long long base = 0;
  for (int l = 1; l <= T.max_lvl; l++) {
    const int m = P.n_leaves(l);
    if (!m) continue;
    const long long work = (long long)m * cpb;
    const dim3 grid((unsigned)((work + 255) / 256));
    if (D == 2)
      hipLaunchKernelGGL((k_leaf_cells<2>), grid, dim3(256), 0, P.stream, P.dev(), P.leaves(l), m,
                         nc, base, d_vars, n_cc, n_fc, n_cells, d_out);
    else
      hipLaunchKernelGGL((k_leaf_cells<3>), grid, dim3(256), 0, P.stream, P.dev(), P.leaves(l), m,
                         nc, base, d_vars, n_cc, n_fc, n_cells, d_out);
    chk_d(hipGetLastError(), "leaf_cells");
    base += work;
  }
  chk_d(hipMemcpyAsync(host.data(), d_out, (size_t)n_out * sizeof(double),
                       hipMemcpyDeviceToHost, P.stream), "leaf_cells copy");
  chk_d(hipStreamSynchronize(P.stream), "leaf_cells sync");
  return n_cells;
}

}  // namespace a5g

namespace a5g {

// A VTK XML file laid out as the reference's m_vtk writes it
// (afivo/src/m_vtk.f90): the XML part with two-space indentation, then one
// raw appended block holding, per array, an int32 byte count and the data.
struct VtkOut {
  std::string xml;
  std::vector<char> app;
  int indent = 0;
  long long offset = 0;
  void line(const std::string& s) { xml += std::string(indent, ' ') + s + "\n"; }
  void open(const std::string& tag) {
    line("<" + tag + ">");
    indent += 2;
  }
  void close(const std::string& tag) {
    indent -= 2;
    line("</" + tag + ">");
  }
  template <typename V>
  void append(const V* p, size_t n) {
    const int32_t nb = (int32_t)(n * sizeof(V));
    const char* b = reinterpret_cast<const char*>(&nb);
    app.insert(app.end(), b, b + 4);
    const char* q = reinterpret_cast<const char*>(p);
    app.insert(app.end(), q, q + n * sizeof(V));
    offset += 4 + (long long)(n * sizeof(V));
  }
};

// The value line of TIME: the reference writes it list-directed (a leading
// blank, the indentation, a separator blank, the value); here the value has
// the fewest significant digits that read back to the same double.
static std::string num_line(int indent, double v) {
  char b[64];
  for (int p = 1; p <= 17; p++) {
    std::snprintf(b, sizeof b, "%.*G", p, v);
    if (std::strtod(b, nullptr) == v) break;
  }
  return std::string(indent + 2, ' ') + b;
}

void write_vtk(Pool& P, const char* filename, int n_cc, const char* const* cc_names,
               const int* ixs_cc, int n_fc, const char* const* fc_names, const int* ixs_fc,
               int n_cycle, double time) {
  P.sync_topology();
  const auto& T = P.tree;
  const int D = T.dim, bc = T.n_cell, bn = T.n_cell + 1;
  const int n_ch = 1 << D;
  const int nodes_per_box = D == 2 ? bn * bn : bn * bn * bn;
  const int cells_per_box = D == 2 ? bc * bc : bc * bc * bc;
  std::vector<int> leaves;
  for (int l = 1; l <= T.max_lvl; l++)
    for (int id : P.lleaves(l)) leaves.push_back(id);
  const long long n_grids = (long long)leaves.size();
  const long long n_nodes = nodes_per_box * n_grids, n_cells = cells_per_box * n_grids;
  if (n_ch * n_cells > (long long)INT32_MAX / 4)
    throw std::length_error("write_vtk: more cells than an int32 VTK block holds");
  std::vector<double> cells;
  leaf_cells(P, n_cc, ixs_cc, n_fc, ixs_fc, cells);
  // node coordinates (Float32, three components), connectivity, offsets, types
  std::vector<float> wr((size_t)(3 * n_nodes), 0.0f);
  std::vector<int32_t> connects((size_t)(n_ch * n_cells)), offsets((size_t)n_cells),
      types((size_t)n_cells, D == 2 ? 8 : 11);
  for (long long g = 0; g < n_grids; g++) {
    const int id = leaves[g];
    const double* r0 = &T.rmin[id * 3];
    const double dr = T.dr[id];
    const long long node_ix = g * nodes_per_box, cell_ix = g * cells_per_box;
    const int kn = D == 3 ? bn : 1, kc = D == 3 ? bc : 1;
    for (int k = 1; k <= kn; k++)
      for (int j = 1; j <= bn; j++)
        for (int i = 1; i <= bn; i++) {
          const long long nn = node_ix + (long long)(k - 1) * bn * bn + (j - 1) * bn + (i - 1);
          const int w[3] = {i - 1, j - 1, k - 1};
          for (int d = 0; d < D; d++) wr[3 * nn + d] = (float)(r0[d] + w[d] * dr);
        }
    for (int k = 1; k <= kc; k++)
      for (int j = 1; j <= bc; j++)
        for (int i = 1; i <= bc; i++) {
          const long long n_ix = node_ix + (long long)(k - 1) * bn * bn + (j - 1) * bn + i - 1;
          const long long c_ix = cell_ix + (long long)(k - 1) * bc * bc + (j - 1) * bc + i;
          offsets[c_ix - 1] = (int32_t)(n_ch * c_ix);
          int32_t* cn = &connects[n_ch * (c_ix - 1)];
          const long long bn2 = (long long)bn * bn;
          const long long q[8] = {n_ix, n_ix + 1, n_ix + bn, n_ix + bn + 1,
                                  n_ix + bn2, n_ix + bn2 + 1, n_ix + bn2 + bn, n_ix + bn2 + bn + 1};
          for (int c = 0; c < n_ch; c++) cn[c] = (int32_t)q[c];
        }
  }
  VtkOut v;
  v.xml = "<?xml version=\"1.0\"?>\n";
  v.line("<VTKFile type=\"UnstructuredGrid\" version=\"0.1\" byte_order=\"LittleEndian\">");
  v.indent = 2;
  v.open("UnstructuredGrid");
  v.open("FieldData");
  v.line("<DataArray type=\"Float64\" Name=\"TIME\" NumberOfTuples=\"1\" format=\"ascii\">");
  v.indent += 2;
  v.xml += num_line(v.indent, time) + "\n";
  v.close("DataArray");
  v.line("<DataArray type=\"Int32\" Name=\"CYCLE\" NumberOfTuples=\"1\" format=\"ascii\">");
  v.indent += 2;
  v.xml += std::string(v.indent + 2, ' ') + std::to_string(n_cycle) + "\n";
  v.close("DataArray");
  v.close("FieldData");
  v.line("<Piece NumberOfPoints=\"" + std::to_string(n_nodes) + "\" NumberOfCells=\"" +
         std::to_string(n_cells) + "\">");
  v.indent += 2;
  v.open("Points");
  v.line("<DataArray type=\"Float32\" NumberOfComponents=\"3\" Name=\"Points\" "
         "format=\"appended\" offset=\"" + std::to_string(v.offset) + "\"/>");
  v.append(wr.data(), wr.size());
  v.close("Points");
  v.open("Cells");
  v.line("<DataArray type=\"Int32\" Name=\"connectivity\" format=\"appended\" offset=\"" +
         std::to_string(v.offset) + "\"/>");
  v.append(connects.data(), connects.size());
  v.line("<DataArray type=\"Int32\" Name=\"offsets\" format=\"appended\" offset=\"" +
         std::to_string(v.offset) + "\"/>");
  v.append(offsets.data(), offsets.size());
  v.line("<DataArray type=\"Int32\" Name=\"types\" format=\"appended\" offset=\"" +
         std::to_string(v.offset) + "\"/>");
  v.append(types.data(), types.size());
  v.close("Cells");
  v.open("CellData");
  // cell variables, then per face variable one column per dimension; every
  // face column of dimension d holds the same sum over all face variables
  for (int c = 0; c < n_cc + n_fc * D; c++) {
    const char* name = c < n_cc ? cc_names[c] : fc_names[c - n_cc];
    const int col = c < n_cc ? c : n_cc + (c - n_cc) % D;
    v.line("<DataArray type=\"Float64\" Name=\"" + std::string(name) +
           "\" NumberOfComponents=\"1\" format=\"appended\" offset=\"" +
           std::to_string(v.offset) + "\"/>");
    v.append(cells.data() + (size_t)col * n_cells, (size_t)n_cells);
  }
  v.close("CellData");
  v.close("Piece");
  v.close("UnstructuredGrid");
  v.line("<AppendedData encoding=\"raw\">");
  const std::string fname = std::string(filename) + ".vtu";
  FILE* f = std::fopen(fname.c_str(), "wb");
  if (!f) throw std::runtime_error("write_vtk: cannot open " + fname);
  bool ok = std::fwrite(v.xml.data(), 1, v.xml.size(), f) == v.xml.size();
  ok = ok && std::fputc('_', f) != EOF;
  ok = ok && std::fwrite(v.app.data(), 1, v.app.size(), f) == v.app.size();
  std::string tail = "\n" + std::string(v.indent, ' ') + "</AppendedData>\n</VTKFile>\n";
  ok = ok && std::fwrite(tail.data(), 1, tail.size(), f) == tail.size();
  ok = (std::fclose(f) == 0) && ok;
  if (!ok) throw std::runtime_error("write_vtk: cannot write " + fname);
}

// Raw leaf-box dump (little endian): "A5GLEAF1", int32 dim, n_cell,
// n_leaves, n_var, coord, cycle, float64 time, int32 var[n_var]; per leaf
// int32 id, lvl, ix[3] and float64 dr, r_min[3]; then float64
// data[n_var][n_leaves][n_cell^dim], cells x fastest, interior only.
void dump_leaves(Pool& P, const char* filename, int n_var, const int* ivs, int cycle,
                 double time) {
  P.sync_topology();
  const auto& T = P.tree;
  std::vector<int> leaves;
  for (int l = 1; l <= T.max_lvl; l++)
    for (int id : P.lleaves(l)) leaves.push_back(id);
  std::vector<double> cells;
  leaf_cells(P, n_var, ivs, 0, nullptr, cells);
  FILE* f = std::fopen(filename, "wb");
  if (!f) throw std::runtime_error(std::string("dump_leaves: cannot open ") + filename);
  bool ok = std::fwrite("A5GLEAF1", 1, 8, f) == 8;
  const int32_t head[6] = {T.dim, T.n_cell, (int32_t)leaves.size(), n_var, T.coord, cycle};
  ok = ok && std::fwrite(head, 4, 6, f) == 6;
  ok = ok && std::fwrite(&time, 8, 1, f) == 1;
  std::vector<int32_t> iv(ivs, ivs + n_var);
  ok = ok && std::fwrite(iv.data(), 4, iv.size(), f) == iv.size();
  for (int id : leaves) {
    const int32_t rec[5] = {id, T.lvl[id], T.ix[id * 3], T.ix[id * 3 + 1], T.ix[id * 3 + 2]};
    const double geo[4] = {T.dr[id], T.rmin[id * 3], T.rmin[id * 3 + 1], T.rmin[id * 3 + 2]};
    ok = ok && std::fwrite(rec, 4, 5, f) == 5 && std::fwrite(geo, 8, 4, f) == 4;
  }
  ok = ok && std::fwrite(cells.data(), 8, cells.size(), f) == cells.size();
  ok = (std::fclose(f) == 0) && ok;
  if (!ok) throw std::runtime_error(std::string("dump_leaves: cannot write ") + filename);
}

}  // namespace a5g
