// Device helpers shared by the a5g kernels: box addressing, child offsets and
// the face ghost-cell rules of Afivo (afivo/src/m_aX_gc.f90, m_aX_mg.f90).
//
// Floating-point expressions keep the reference's operand order; the library
// is compiled with -ffp-contract=off so none is fused into an FMA, which makes
// the FP64 path bit-identical to the reference schedule.
#pragma once

#include <hip/hip_runtime.h>

#include "a5g_layout.hpp"
#include "a5g_pool.hpp"

namespace a5g {

__device__ __forceinline__ int d_nb_dim(int nb) { return (nb - 1) >> 1; }
__device__ __forceinline__ bool d_nb_low(int nb) { return ((nb - 1) & 1) == 0; }
__device__ __forceinline__ int d_nb_dix(int nb, int d) {
  return d_nb_dim(nb) != d ? 0 : (d_nb_low(nb) ? -1 : 1);
}

__device__ __forceinline__ double* boxp(const Dev& a, int iv, int id) {
  return a.cc + (size_t)(iv - 1) * a.vs + (size_t)id * a.bs;
}

// The same box slot in the FP64 (cc) or FP32 (cc32) variable arrays
template <typename T>
__device__ __forceinline__ T* boxp_t(const Dev& a, int iv, int id);
template <>
__device__ __forceinline__ double* boxp_t<double>(const Dev& a, int iv, int id) {
  return boxp(a, iv, id);
}
template <>
__device__ __forceinline__ float* boxp_t<float>(const Dev& a, int iv, int id) {
  return a.cc32 + (size_t)(iv - 1) * a.vs32 + (size_t)id * a.bs;
}

__device__ __forceinline__ double* facep(const Dev& a, int ifc, int id, int d, int dim) {
  return a.fc + (size_t)(ifc - 1) * a.fvs + ((size_t)id * dim + d) * a.fs;
}

// The instance of box id (Pool::inst_of): its position along the last
// dimension, a quarter cell inside the box
template <int D>
__device__ __forceinline__ int inst_of(const Dev& a, int id) {
  const int g = (int)floor((a.rmin[id * 3 + D - 1] + 0.25 * a.dr[id]) / a.inst_stride);
  return min(max(g, 0), a.n_inst - 1);
}

// index into a face array fx(nc+1,nc[,nc]) / fy / fz of a box (1-based)
template <int D>
__device__ __forceinline__ int fidx(int nc, int d, int i, int j, int k) {
  const int n0 = nc + (d == 0), n1 = nc + (d == 1);
  return D == 2 ? (i - 1) + n0 * (j - 1) : (i - 1) + n0 * ((j - 1) + n1 * (k - 1));
}

// a$D_get_child_offset, m_aX_utils.f90:227-239
template <int D>
__device__ __forceinline__ void child_off(const Dev& a, int id, int nc, int nb, int off[3]) {
  const int l = a.lvl[id];
#pragma unroll
  for (int d = 0; d < 3; d++) {
    if (d >= D) { off[d] = 0; continue; }
    if (l > 1) {
      off[d] = ((a.ix[id * 3 + d] - 1) & 1) * (nc >> 1);
      if (nb > 0) off[d] -= d_nb_dix(nb, d) * nc;
    } else {
      off[d] = 0;
      if (nb > 0) off[d] -= d_nb_dix(nb, d) * (nc >> 1);
    }
  }
}

// The value V of the rule 2V - c on physical face nb of box id: the run's
// own with instances, else the tree's
template <int D>
__device__ __forceinline__ double face_value(const Dev& a, int id, int nb) {
  return a.n_inst ? a.inst_val[inst_of<D>(a, id) * kInstVals + kInstBc + nb - 1]
                  : a.bc_val[nb - 1];
}

// Value of the ghost cell g = (gi, gj, gk) on face nb of box id for the
// multigrid potential: neighbour copy (sides_from_nb, m_aX_gc.f90:368-402),
// refinement-boundary extrapolation (mg$D_sides_rb, m_aX_mg.f90:262-356,
// with a$D_gc_prolong0 / a$D_prolong0_to, m_aX_gc.f90:71-86,
// m_aX_prolong.f90:37-78) or the physical rule (copy, -c, 2V - c).
// own(i, j, k) gives the box's own interior cells (from HBM or from an LDS
// copy); neighbours and the parent are read from HBM.
template <int D, typename T, typename U>
__device__ __forceinline__ T ghost_mg_u(const Dev& a, U own, int iv, int id, int nc, int nb,
                                        int gi, int gj, int gk) {
  const int nb_id = a.nbs[id * 6 + nb - 1];
  const int d = d_nb_dim(nb);
  int g[3] = {gi, gj, gk};
  int c[3] = {gi, gj, gk};
  const bool low = d_nb_low(nb);
  c[d] = low ? 1 : nc;
  if (nb_id > 0) {
    int s[3] = {gi, gj, gk};
    s[d] = low ? nc : 1;
    return boxp_t<T>(a, iv, nb_id)[cidx_in<D>(nc, s[0], s[1], s[2])];
  } else if (nb_id == 0) {
    const int pid = a.parent[id];
    const int pnc = a.ncell[pid];
    int off[3];
    child_off<D>(a, id, nc, 0, off);
    int q[3];
#pragma unroll
    for (int e = 0; e < 3; e++) q[e] = (e < D) ? off[e] + ((g[e] + 1) >> 1) : 0;
    const T P = boxp_t<T>(a, iv, pid)[cidx<D>(pnc, q[0], q[1], q[2])];
    const T cv = own(c[0], c[1], c[2]);
    // shifted interior cells, in the order dim 0, 1, 2 of the reference
    T t[3];
#pragma unroll
    for (int e = 0; e < D; e++) {
      int s[3] = {c[0], c[1], c[2]};
      if (e == d) s[e] += low ? 1 : -1;
      else s[e] += -1 + 2 * (s[e] & 1);
      t[e] = own(s[0], s[1], s[2]);
    }
    if (D == 2) return T(0.5) * P + cv - T(0.25) * (t[0] + t[1]);
    return T(0.5) * P + T(1.25) * cv - T(0.25) * (t[0] + t[1] + t[2]);
  } else {
    const T cv = own(c[0], c[1], c[2]);
    const int kind = a.bc_kind[nb - 1];
    if (kind == 0) return cv;
    if (kind == 1) return -cv;
    return T(2) * (T)face_value<D>(a, id, nb) - cv;
  }
}

// ghost_mg_u with the box's own cells in HBM (or an LDS copy in cidx order)
template <int D, typename T = double>
__device__ __forceinline__ T ghost_mg(const Dev& a, const T* p, int iv, int id, int nc, int nb,
                                      int gi, int gj, int gk) {
  return ghost_mg_u<D, T>(a, [&](int x, int y, int z) { return p[cidx_in<D>(nc, x, y, z)]; },
                          iv, id, nc, nb, gi, gj, gk);
}

}  // namespace a5g
