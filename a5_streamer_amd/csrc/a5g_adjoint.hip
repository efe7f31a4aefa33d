// Reverse mode of the rest of a streamer time step on a frozen mesh (no
// counterpart in the reference, which computes no derivatives): the tree
// copy, the trapezoidal average, the restriction of the densities, the
// interpolated ghost cells, the space-charge right-hand side and the field
// from the potential, in the forms the drivers call them
// (src/streamer_2d.f90:98-129,307-381). With the reverse transport substep
// (a5g_streamer.hip) and the multigrid's reverse (a5g_mg.hip) they make the
// whole step's transpose. Each kernel takes the adjoints of the values its
// forward operation writes (which it clears: the forward pass overwrites
// them) and adds those of the values it reads. A forward operation that
// branches on its data (the limited interpolation, the field norm) is
// linearised on the state the caller leaves in the primal variables, which
// must be the state the forward operation read.
//
// Also here: a device store of whole variables (slots), which the
// checkpoint-recompute of a differentiated window keeps its states in.
#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>
#include <vector>

#include "a5g_internal.hpp"
#include "a5g_kernels.hpp"

namespace a5g {

#define ADJ_BOXES                                   \
  const int bpb_ = blockDim.x / tpb;                \
  const int slot_ = threadIdx.x / tpb;              \
  const int lt = threadIdx.x - slot_ * tpb;         \
  const int li_ = blockIdx.x * bpb_ + slot_;        \
  const bool act = li_ < n;                         \
  const int id = act ? list[li_] : 0;

namespace {

void achk(hipError_t e, const char* w) {
  if (e != hipSuccess) throw std::runtime_error(std::string(w) + ": " + hipGetErrorString(e));
}

int up2(int v) {
  int p = 1;
  while (p < v) p <<= 1;
  return p;
}

// lanes per box (a power of two from 64 to 256, at least the work of one
// box when that fits) and workgroups of 256 lanes for n boxes
struct Launch {
  int tpb, grid;
};
Launch launch_for(int n, int work) {
  const int tpb = std::min(256, std::max(64, up2(work)));
  return {tpb, (n + 256 / tpb - 1) / (256 / tpb)};
}

int ncells(int D, int nc, bool ghosts) {
  const int m = ghosts ? nc + 2 : nc;
  return D == 2 ? m * m : m * m * m;
}

}  // namespace

// ---------------------------------------------------------------------------
// whole-slot transposes over the boxes of a list
// ---------------------------------------------------------------------------

// a$D_tree_copy_cc (afivo/src/m_aX_utils.f90:432-453), to = from on whole
// boxes: lfrom += lto, lto = 0
__global__ __launch_bounds__(256) void k_adj_copy(Dev a, const int* __restrict__ list, int n,
                                                  int bl, int tpb, int lfrom, int lto) {
  ADJ_BOXES
  if (!act) return;
  double* Lf = boxp(a, lfrom, id);
  double* Lt = boxp(a, lto, id);
  for (int q = lt; q < bl; q += tpb) {
    Lf[q] += Lt[q];
    Lt[q] = 0.0;
  }
}

// average_dens (src/streamer_2d.f90:492-496), n = 0.5 (n + n_old) on whole
// boxes: l_old += 0.5 l, l = 0.5 l, for the electrons and the ions
__global__ __launch_bounds__(256) void k_adj_average(Dev a, const int* __restrict__ list, int n,
                                                     int bl, int tpb, int le, int leo, int li,
                                                     int lio) {
  ADJ_BOXES
  if (!act) return;
  double* Le = boxp(a, le, id);
  double* Leo = boxp(a, leo, id);
  double* Li = boxp(a, li, id);
  double* Lio = boxp(a, lio, id);
  for (int q = lt; q < bl; q += tpb) {
    const double xe = Le[q], xi = Li[q];
    Leo[q] += 0.5 * xe;
    Le[q] = 0.5 * xe;
    Lio[q] += 0.5 * xi;
    Li[q] = 0.5 * xi;
  }
}

// compute_fld's rhs = fac (n_e - n_i) on whole leaf boxes
// (src/streamer_2d.f90:318-329): le += fac lr, li -= fac lr, lr = 0
__global__ __launch_bounds__(256) void k_adj_space_charge(Dev a, const int* __restrict__ list,
                                                          int n, int bl, int tpb, int lr, int le,
                                                          int li, double fac) {
  ADJ_BOXES
  if (!act) return;
  double* Lr = boxp(a, lr, id);
  double* Le = boxp(a, le, id);
  double* Li = boxp(a, li, id);
  for (int q = lt; q < bl; q += tpb) {
    const double r = Lr[q];
    Le[q] += fac * r;
    Li[q] -= fac * r;
    Lr[q] = 0.0;
  }
}

// ---------------------------------------------------------------------------
// a$D_restrict_box (afivo/src/m_aX_restrict.f90:67-142) of the children in
// list, transposed: each parent cell's adjoint goes to the 2^D child cells it
// averaged, with the forward pass's weights (the radius-weighted ones in
// cylindrical coordinates, with the radius of m_aX_restrict.f90:106), and is
// cleared. A parent cell is the average of one child's cells, so every
// value has one writer.
// ---------------------------------------------------------------------------
template <int D>
__global__ __launch_bounds__(256) void k_adj_restrict(Dev a, const int* __restrict__ list, int n,
                                                      int nc, int tpb, int il) {
  ADJ_BOXES
  if (!act) return;
  const int pid = a.parent[id];
  const int pnc = a.ncell[pid];
  const int hnc = nc >> 1;
  int off[3];
  child_off<D>(a, id, nc, 0, off);
  double* C = boxp(a, il, id);
  double* Pp = boxp(a, il, pid);
  const int total = D == 2 ? hnc * hnc : hnc * hnc * hnc;
  for (int q = lt; q < total; q += tpb) {
    const int ci = q % hnc + 1;
    const int cj = (q / hnc) % hnc + 1;
    const int ck = D == 3 ? q / (hnc * hnc) + 1 : 0;
    const int fi = 2 * ci - 1, fj = 2 * cj - 1, fk = D == 3 ? 2 * ck - 1 : 0;
    double* pp = Pp + cidx<D>(pnc, off[0] + ci, off[1] + cj, D == 3 ? off[2] + ck : 0);
    const double lam = *pp;
    *pp = 0.0;
    if (D == 2 && a.coord == 1) {
      const double dr16 = 0.0625 * a.dr[pid];
      const double rr = a.rmin[pid * 3] + (ci - 0.5) * a.dr[pid];
      const double rfac = dr16 / rr;
      const double w1 = (0.25 - rfac) * lam, w2 = (0.25 + rfac) * lam;
      C[cidx<D>(nc, fi, fj, 0)] += w1;
      C[cidx<D>(nc, fi, fj + 1, 0)] += w1;
      C[cidx<D>(nc, fi + 1, fj, 0)] += w2;
      C[cidx<D>(nc, fi + 1, fj + 1, 0)] += w2;
    } else {
      const double w = (D == 2 ? 0.25 : 0.125) * lam;
      for (int dk = 0; dk < (D == 3 ? 2 : 1); dk++)
        for (int dj = 0; dj < 2; dj++)
          for (int di = 0; di < 2; di++) C[cidx<D>(nc, fi + di, fj + dj, fk + dk)] += w;
    }
  }
}

// ---------------------------------------------------------------------------
// a$D_gc_box with a$D_gc_interp (rb = 1, afivo/src/m_aX_gc.f90:90-185) or
// a$D_gc_interp_lim (rb = 2, :190-295) and a$D_gc_neumann (bc = 1,
// :298-330) or a$D_gc_dirichlet (bc = 2, :333-365), transposed over the
// boxes in list: every face ghost's adjoint goes to the interior cells it
// was made from (a same-level neighbour's adjacent cell; the coarse
// neighbour's two (2D) or three (3D) cells and the box's own cell at a
// refinement boundary; the box's own cell on a physical face) and is
// cleared. The limiter's case is recomputed from iv, with the forward
// pass's arithmetic. Ghosts are made from interior cells only, so the
// scatter never meets a ghost; two boxes can send to one cell, hence the
// atomics.
// ---------------------------------------------------------------------------
template <int D>
__global__ __launch_bounds__(256) void k_adj_gc_interp(Dev a, const int* __restrict__ list, int n,
                                                       int nc, int tpb, int iv, int il, int rb,
                                                       int bc) {
  ADJ_BOXES
  if (!act) return;
  const double* p = boxp(a, iv, id);
  double* L = boxp(a, il, id);
  const int per_face = D == 2 ? nc : nc * nc;
  const double sixth = 1.0 / 6.0, third = 1.0 / 3.0;
  for (int q = lt; q < 2 * D * per_face; q += tpb) {
    const int nb = q / per_face + 1;
    const int t = q - (nb - 1) * per_face;
    const int d = d_nb_dim(nb);
    const bool low = d_nb_low(nb);
    const int ta = d == 0 ? 1 : 0, tb = d == 2 ? 1 : 2;
    int g[3] = {0, 0, 0};
    g[d] = low ? 0 : nc + 1;
    g[ta] = t % nc + 1;
    if (D == 3) g[tb] = t / nc + 1;
    double* gp = L + cidx<D>(nc, g[0], g[1], g[2]);
    const double lam = *gp;
    *gp = 0.0;
    if (lam == 0.0) continue;
    const int nb_id = a.nbs[id * 6 + nb - 1];
    if (nb_id > 0) {
      int s[3] = {g[0], g[1], g[2]};
      s[d] = low ? nc : 1;
      atomicAdd(boxp(a, il, nb_id) + cidx_in<D>(nc, s[0], s[1], s[2]), lam);
      continue;
    }
    if (nb_id < 0) {
      int c[3] = {g[0], g[1], g[2]};
      c[d] = low ? 1 : nc;
      atomicAdd(L + cidx_in<D>(nc, c[0], c[1], c[2]), bc == 1 ? lam : -lam);
      continue;
    }
    // coarse-neighbour interpolation: the cells of the parent's neighbour
    // pn next to the face, and the box's own cell f beside the ghost
    const int pid = a.parent[id];
    const int pn = a.nbs[pid * 6 + nb - 1];
    const int pnc = a.ncell[pn];
    int off[3];
    child_off<D>(a, id, nc, nb, off);
    const int ixc = low ? nc : 1, ixf = low ? 1 : nc;
    const double* qv = boxp(a, iv, pn);
    double* Lq = boxp(a, il, pn);
    int f[3] = {g[0], g[1], g[2]};
    f[d] = ixf;
    const int fi = cidx_in<D>(nc, f[0], f[1], f[2]);
    const double fv = p[fi];
    if (D == 2) {
      const int x = g[ta];
      const int c1 = off[ta] + ((x + 1) >> 1), c2 = c1 + 1 - 2 * (x & 1);
      int s1[3] = {0, 0, 0}, s2[3] = {0, 0, 0};
      s1[d] = s2[d] = ixc;
      s1[ta] = c1;
      s2[ta] = c2;
      const int q1 = cidx<D>(pnc, s1[0], s1[1], 0), q2 = cidx<D>(pnc, s2[0], s2[1], 0);
      const double v1 = qv[q1], v2 = qv[q2];
      const double v = 0.5 * v1 + sixth * v2 + third * fv;
      if (rb == 2 && v > 2.0 * v1) {
        atomicAdd(Lq + q1, 2.0 * lam);
      } else {
        atomicAdd(Lq + q1, 0.5 * lam);
        atomicAdd(Lq + q2, sixth * lam);
        atomicAdd(L + fi, third * lam);
      }
    } else {
      const int x = g[ta], y = g[tb];
      const int a1 = off[ta] + ((x + 1) >> 1), a2 = a1 + 1 - 2 * (x & 1);
      const int b1 = off[tb] + ((y + 1) >> 1), b2 = b1 + 1 - 2 * (y & 1);
      int s1[3], s2[3], s3[3];
      s1[d] = s2[d] = s3[d] = ixc;
      s1[ta] = a1;
      s1[tb] = b1;
      if (d == 2) {
        s2[ta] = a1; s2[tb] = b2;
        s3[ta] = a2; s3[tb] = b1;
      } else {
        s2[ta] = a2; s2[tb] = b1;
        s3[ta] = a1; s3[tb] = b2;
      }
      const int q1 = cidx<D>(pnc, s1[0], s1[1], s1[2]);
      const int q2 = cidx<D>(pnc, s2[0], s2[1], s2[2]);
      const int q3 = cidx<D>(pnc, s3[0], s3[1], s3[2]);
      const double v1 = qv[q1], v2 = qv[q2], v3 = qv[q3];
      const double v = third * v1 + sixth * v2 + sixth * v3 + third * fv;
      if (rb == 2 && v > 2.0 * v1) {
        atomicAdd(Lq + q1, 2.0 * lam);
      } else {
        atomicAdd(Lq + q1, third * lam);
        atomicAdd(Lq + q2, sixth * lam);
        atomicAdd(Lq + q3, sixth * lam);
        atomicAdd(L + fi, third * lam);
      }
    }
  }
}

// ---------------------------------------------------------------------------
// fld_from_pot (src/streamer_2d.f90:346-381, src/streamer_3d.f90:348-395)
// of the boxes in list, transposed. The forward pass sets every face field
// F from the potential (inv_dr (phi_a - phi_b) inside the box,
// 2 inv_dr (phi_a - phi_b) eps_w / (eps_a + eps_b) on its boundary faces,
// eps_w the permittivity on the face's own side of the box) and then
// |E| = sqrt(sum_d (F_lo + F_hi)^2 / 4) on the interior cells, whose
// derivative in either face of direction d is (F_lo + F_hi) / (4 |E|)
// (taken as zero where |E| = 0). Phase 1 adds each face's share of the
// interior |E| adjoints to the face adjoint lfld; phase 2 sends the face
// adjoints to the potential of the two cells of every face (every cell of
// the box with its faces gathered, ghosts included, so no atomics);
// phase 3 clears lfld and the interior of lE.
// ---------------------------------------------------------------------------
template <int D>
__device__ __forceinline__ double fld_coef(const double* e, double inv_dr, int nc, int pos,
                                          int ia, int ib) {
  if (pos == 1) return 2 * inv_dr * e[ia] / (e[ib] + e[ia]);
  if (pos == nc + 1) return 2 * inv_dr * e[ib] / (e[ib] + e[ia]);
  return inv_dr;
}

template <int D>
__global__ __launch_bounds__(256) void k_adj_fld(Dev a, const int* __restrict__ list, int n,
                                                 int nc, int tpb, int ieps, int ifld, int ffld,
                                                 int lE, int lfld, int lphi) {
  ADJ_BOXES
  const int per = (nc + 1) * (D == 2 ? nc : nc * nc);
  double inv_dr = 0;
  const double* E = nullptr;
  const double* eps = nullptr;
  double* LE = nullptr;
  const double* F[3] = {nullptr, nullptr, nullptr};
  double* LF[3] = {nullptr, nullptr, nullptr};
  if (act) {
    inv_dr = 1 / a.dr[id];
    E = boxp(a, ifld, id);
    eps = boxp(a, ieps, id);
    LE = boxp(a, lE, id);
    for (int d = 0; d < D; d++) {
      F[d] = facep(a, ffld, id, d, D);
      LF[d] = facep(a, lfld, id, d, D);
    }
    for (int q = lt; q < D * per; q += tpb) {
      const int d = q / per, w = q - d * per;
      const int n0 = nc + (d == 0), n1 = nc + (d == 1);
      int c[3] = {w % n0 + 1, (w / n0) % n1 + 1, D == 3 ? w / (n0 * n1) + 1 : 0};
      const int pos = c[d];
      double lam = LF[d][w];
      for (int side = 0; side < 2; side++) {
        // the interior cell below (side 0) or above (side 1) the face
        int x[3] = {c[0], c[1], c[2]};
        x[d] = pos - 1 + side;
        if (x[d] < 1 || x[d] > nc) continue;
        const int ci = cidx_in<D>(nc, x[0], x[1], x[2]);
        const double le = LE[ci], ev = E[ci];
        if (le == 0.0 || !(ev > 0.0)) continue;
        int h[3] = {x[0], x[1], x[2]};
        h[d] += 1;
        const double s = F[d][fidx<D>(nc, d, x[0], x[1], x[2])] + F[d][fidx<D>(nc, d, h[0], h[1], h[2])];
        lam += le * (0.25 * s / ev);
      }
      LF[d][w] = lam;
    }
  }
  __syncthreads();
  if (act) {
    double* Lp = boxp(a, lphi, id);
    const int m = nc + 2;
    const int bl = D == 2 ? m * m : m * m * m;
    for (int q = lt; q < bl; q += tpb) {
      int c[3] = {q % m, (q / m) % m, D == 3 ? q / (m * m) : 0};
      double acc = 0.0;
      for (int d = 0; d < D; d++) {
        bool tang = true;
        for (int e2 = 0; e2 < D; e2++)
          if (e2 != d && (c[e2] < 1 || c[e2] > nc)) tang = false;
        if (!tang) continue;
        const int xd = c[d];
        if (xd <= nc) {
          // the cell is phi_a of the face at pos = xd + 1
          int f[3] = {c[0], c[1], c[2]};
          f[d] = xd + 1;
          int b[3] = {c[0], c[1], c[2]};
          b[d] = xd + 1;
          const double k = fld_coef<D>(eps, inv_dr, nc, xd + 1, cidx<D>(nc, c[0], c[1], c[2]),
                                       cidx<D>(nc, b[0], b[1], b[2]));
          acc += k * LF[d][fidx<D>(nc, d, f[0], f[1], f[2])];
        }
        if (xd >= 1) {
          // the cell is phi_b of the face at pos = xd
          int f[3] = {c[0], c[1], c[2]};
          int aa[3] = {c[0], c[1], c[2]};
          aa[d] = xd - 1;
          const double k = fld_coef<D>(eps, inv_dr, nc, xd, cidx<D>(nc, aa[0], aa[1], aa[2]),
                                       cidx<D>(nc, c[0], c[1], c[2]));
          acc -= k * LF[d][fidx<D>(nc, d, f[0], f[1], f[2])];
        }
      }
      if (acc != 0.0) Lp[cidx<D>(nc, c[0], c[1], c[2])] += acc;
    }
  }
  __syncthreads();
  if (act) {
    for (int q = lt; q < D * per; q += tpb) {
      const int d = q / per, w = q - d * per;
      LF[d][w] = 0.0;
    }
    const int total = D == 2 ? nc * nc : nc * nc * nc;
    for (int q = lt; q < total; q += tpb) {
      const int i = q % nc + 1, j = (q / nc) % nc + 1, k = D == 3 ? q / (nc * nc) + 1 : 0;
      LE[cidx_in<D>(nc, i, j, k)] = 0.0;
    }
  }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------

static void check_vars(const Pool& P, std::initializer_list<int> ivs) {
  for (int iv : ivs)
    if (iv < 1 || iv > P.tree.n_cc) throw std::out_of_range("cell variable out of range");
}

void copy_cc_adj(Pool& P, int lfrom, int lto) {
  check_vars(P, {lfrom, lto});
  if (lfrom == lto) throw std::invalid_argument("copy adjoint: the two variables must differ");
  P.sync_topology();
  const int D = P.tree.dim;
  for (int lvl = P.tree.min_lvl; lvl <= P.tree.max_lvl; lvl++) {
    const int n = P.n_ids(lvl);
    if (!n) continue;
    const int bl = ncells(D, P.nc_lvl(lvl), true);
    const Launch g = launch_for(n, bl);
    hipLaunchKernelGGL(k_adj_copy, dim3(g.grid), dim3(256), 0, P.stream, P.dev(), P.ids(lvl), n,
                       bl, g.tpb, lfrom, lto);
    achk(hipGetLastError(), "copy adjoint");
  }
}

void average_dens_adj(Pool& P, int le, int leo, int li, int lio) {
  check_vars(P, {le, leo, li, lio});
  P.sync_topology();
  const int D = P.tree.dim;
  for (int lvl = P.tree.min_lvl; lvl <= P.tree.max_lvl; lvl++) {
    const int n = P.n_ids(lvl);
    if (!n) continue;
    const int bl = ncells(D, P.nc_lvl(lvl), true);
    const Launch g = launch_for(n, bl);
    hipLaunchKernelGGL(k_adj_average, dim3(g.grid), dim3(256), 0, P.stream, P.dev(), P.ids(lvl),
                       n, bl, g.tpb, le, leo, li, lio);
    achk(hipGetLastError(), "average adjoint");
  }
}

void space_charge_adj(Pool& P, double fac, int lr, int le, int li) {
  check_vars(P, {lr, le, li});
  P.sync_topology();
  const int D = P.tree.dim;
  for (int lvl = 1; lvl <= P.tree.max_lvl; lvl++) {
    const int n = P.n_leaves(lvl);
    if (!n) continue;
    const int bl = ncells(D, P.nc_lvl(lvl), true);
    const Launch g = launch_for(n, bl);
    hipLaunchKernelGGL(k_adj_space_charge, dim3(g.grid), dim3(256), 0, P.stream, P.dev(),
                       P.leaves(lvl), n, bl, g.tpb, lr, le, li, fac);
    achk(hipGetLastError(), "space charge adjoint");
  }
}

void restrict_tree_adj(Pool& P, int il) {
  check_vars(P, {il});
  P.sync_topology();
  const int D = P.tree.dim;
  // the forward pass restricts finest first; its transpose goes coarsest
  // first, so that a parent's adjoint holds its own parent's share
  for (int lvl = P.tree.min_lvl; lvl <= P.tree.max_lvl - 1; lvl++) {
    const int n = P.n_ids(lvl + 1);
    if (!n) continue;
    const int nc = P.nc_lvl(lvl + 1);
    const Launch g = launch_for(n, ncells(D, nc / 2, false));
    if (D == 2)
      hipLaunchKernelGGL((k_adj_restrict<2>), dim3(g.grid), dim3(256), 0, P.stream, P.dev(),
                         P.ids(lvl + 1), n, nc, g.tpb, il);
    else
      hipLaunchKernelGGL((k_adj_restrict<3>), dim3(g.grid), dim3(256), 0, P.stream, P.dev(),
                         P.ids(lvl + 1), n, nc, g.tpb, il);
    achk(hipGetLastError(), "restrict adjoint");
  }
}

void gc_tree_adj(Pool& P, int iv, int il, int rb, int bc) {
  check_vars(P, {iv, il});
  if (iv == il) throw std::invalid_argument("ghost-cell adjoint: il must differ from iv");
  if (rb != 1 && rb != 2)
    throw std::invalid_argument("ghost-cell adjoint: rb must be 1 (interp) or 2 (interp_lim)");
  if (bc != 1 && bc != 2)
    throw std::invalid_argument("ghost-cell adjoint: bc must be 1 (Neumann) or 2 (Dirichlet)");
  P.sync_topology();
  const int D = P.tree.dim;
  for (int lvl = P.tree.max_lvl; lvl >= P.tree.min_lvl; lvl--) {
    const int n = P.n_ids(lvl);
    if (!n) continue;
    const int nc = P.nc_lvl(lvl);
    const Launch g = launch_for(n, 2 * D * (D == 2 ? nc : nc * nc));
    if (D == 2)
      hipLaunchKernelGGL((k_adj_gc_interp<2>), dim3(g.grid), dim3(256), 0, P.stream, P.dev(),
                         P.ids(lvl), n, nc, g.tpb, iv, il, rb, bc);
    else
      hipLaunchKernelGGL((k_adj_gc_interp<3>), dim3(g.grid), dim3(256), 0, P.stream, P.dev(),
                         P.ids(lvl), n, nc, g.tpb, iv, il, rb, bc);
    achk(hipGetLastError(), "ghost-cell adjoint");
  }
}

void fld_from_pot_adj(Pool& P, const StVars& v, int lE, int lfld, int lphi) {
  check_vars(P, {lE, lphi});
  if (lfld < 1 || lfld > P.tree.n_fc) throw std::out_of_range("face variable out of range");
  if (lE == v.fld || lphi == v.phi || lfld == v.f_fld)
    throw std::invalid_argument("field adjoint: the adjoints must be variables of their own");
  P.sync_topology();
  const int D = P.tree.dim;
  for (int lvl = P.tree.min_lvl; lvl <= P.tree.max_lvl; lvl++) {
    const int n = P.n_ids(lvl);
    if (!n) continue;
    const int nc = P.nc_lvl(lvl);
    const Launch g = launch_for(n, ncells(D, nc, true));
    if (D == 2)
      hipLaunchKernelGGL((k_adj_fld<2>), dim3(g.grid), dim3(256), 0, P.stream, P.dev(),
                         P.ids(lvl), n, nc, g.tpb, v.eps, v.fld, v.f_fld, lE, lfld, lphi);
    else
      hipLaunchKernelGGL((k_adj_fld<3>), dim3(g.grid), dim3(256), 0, P.stream, P.dev(),
                         P.ids(lvl), n, nc, g.tpb, v.eps, v.fld, v.f_fld, lE, lfld, lphi);
    achk(hipGetLastError(), "field adjoint");
  }
}

// ---------------------------------------------------------------------------
// Device store of whole variables: slot k holds copies of the cell
// variables ivs and face variables fvs over the box slots 0..max_id, taken
// at save (op 0) and written back at load (op 1). The mesh must not change
// between the two (the slot remembers max_id, the variable lists and the
// topology version, and refuses otherwise). op 2 sets the variables to
// zero (no slot involved).
// ---------------------------------------------------------------------------
void state_store(Pool& P, int slot, int n_cc, const int* ivs, int n_fc, const int* fvs,
                 int op) {
  if (op < 0 || op > 2) throw std::invalid_argument("store: op must be 0 (save), 1 (load), 2 (clear)");
  if (n_cc < 0 || n_fc < 0 || (n_cc && !ivs) || (n_fc && !fvs))
    throw std::invalid_argument("store: bad variable lists");
  for (int k = 0; k < n_cc; k++) check_vars(P, {ivs[k]});
  for (int k = 0; k < n_fc; k++)
    if (fvs[k] < 1 || fvs[k] > P.tree.n_fc) throw std::out_of_range("face variable out of range");
  P.sync_topology();
  const size_t boxes = (size_t)P.tree.max_id + 1;
  const size_t cbytes = boxes * P.bs() * sizeof(double);
  const size_t fbytes = boxes * P.tree.dim * P.fs() * sizeof(double);
  if (op == 2) {
    // the variables set to zero over every box slot (never inside a
    // captured section: these are plain stream operations)
    const Dev a = P.dev();
    for (int k = 0; k < n_cc; k++)
      achk(hipMemsetAsync(a.cc + (size_t)(ivs[k] - 1) * a.vs, 0, cbytes, P.stream), "store clear");
    for (int k = 0; k < n_fc; k++)
      achk(hipMemsetAsync(a.fc + (size_t)(fvs[k] - 1) * a.fvs, 0, fbytes, P.stream),
           "store clear");
    return;
  }
  const bool load = op == 1;
  if (slot < 0 || slot > 4096) throw std::out_of_range("store slot out of range");
  const size_t need = n_cc * cbytes + n_fc * fbytes;
  if ((int)P.store.size() <= slot) P.store.resize(slot + 1);
  Pool::Slot& s = P.store[slot];
  std::vector<int> key(ivs, ivs + n_cc);
  key.push_back(-1);
  key.insert(key.end(), fvs, fvs + n_fc);
  const Dev a = P.dev();
  if (load) {
    if (!s.d || s.max_id != P.tree.max_id || s.key != key || s.topo != P.topo_version)
      throw std::logic_error("store: slot " + std::to_string(slot) +
                             " holds no state of these variables on this mesh");
  } else {
    if (s.bytes < need) {
      if (s.d) achk(hipFree(s.d), "hipFree store");
      s.d = nullptr;
      s.bytes = 0;
      achk(hipMalloc((void**)&s.d, need), "hipMalloc store");
      s.bytes = need;
    }
    s.max_id = P.tree.max_id;
    s.key = key;
    s.topo = P.topo_version;
  }
  char* base = (char*)s.d;
  for (int k = 0; k < n_cc; k++) {
    double* v = a.cc + (size_t)(ivs[k] - 1) * a.vs;
    void* st = base + k * cbytes;
    achk(hipMemcpyAsync(load ? (void*)v : st, load ? st : (void*)v, cbytes,
                        hipMemcpyDeviceToDevice, P.stream),
         "store copy");
  }
  base += n_cc * cbytes;
  for (int k = 0; k < n_fc; k++) {
    double* v = a.fc + (size_t)(fvs[k] - 1) * a.fvs;
    void* st = base + k * fbytes;
    achk(hipMemcpyAsync(load ? (void*)v : st, load ? st : (void*)v, fbytes,
                        hipMemcpyDeviceToDevice, P.stream),
         "store copy");
  }
}

void store_release(Pool& P) {
  for (auto& s : P.store)
    if (s.d) (void)hipFree(s.d);
  P.store.clear();
}

}  // namespace a5g
