// Streamer diagnostics of the a5g library: the located maximum over the
// leaves (a$D_reduction_loc, afivo/src/m_aX_utils.f90:499-555) with the
// per-box procedures of get_streamer_properties (src/streamer_cyl.f90:
// 831-903), and the axis profile of get_cc_axis (:905-956).
//
// The reduction runs in two launches: one workgroup per leaf box finds the
// box's value and cell, then one workgroup picks the box. Both keep the
// outcome of the reference's loop on one thread: inside a box the first cell
// in array order (x fastest) of the largest value, across boxes the first
// leaf (levels upwards, leaves in list order) whose value is the largest,
// provided it exceeds the initial value.
#include <hip/hip_runtime.h>

#include <cmath>
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
    for (int id : T.lv_leaves(l)) order.push_back(id);
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
