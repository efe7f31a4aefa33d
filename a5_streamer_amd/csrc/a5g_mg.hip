// Multigrid, ghost-cell, restriction and copy kernels of the a5g library, and
// the host orchestration of Afivo's FAS V-cycle and FMG over box lists.
//
// Layout: every variable is one HBM array of boxes, box `id` at slot id, each
// box the reference's cc(0:nc+1, ...) block with i fastest. A pass over a
// level is one launch over the level's device id list; a workgroup of 256
// lanes takes one box (or 4 boxes of 64 lanes for small 2D boxes).
//
// Schedule: the passes and their order are those of m_aX_mg.f90, so the FP64
// result equals the reference's bit for bit. Two fusions keep that property:
//  * a Gauss-Seidel half-sweep recomputes the face ghosts it reads from the
//    current state instead of reading the ghost layer written by a separate
//    pass (the ghost layer is written once, after the last half-sweep); the
//    first half-sweep of a call reads the stored ghosts, as the reference does;
//  * residual, restriction of residual and potential, and the restore of the
//    fine tmp (update_coarse) are one pass; the residual stays in registers.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <stdexcept>

#include "a5g_internal.hpp"
#include "a5g_kernels.hpp"

namespace a5g {

// ---------------------------------------------------------------------------
// launch geometry
// ---------------------------------------------------------------------------

static int pow2ceil(int v) {
  int p = 1;
  while (p < v) p <<= 1;
  return p;
}

struct Geo {
  int tpb, grid;
};

static Geo geo(int n_boxes, int work) {
  int tpb = std::min(256, std::max(64, pow2ceil(work)));
  int bpb = 256 / tpb;
  return {tpb, (n_boxes + bpb - 1) / bpb};
}

#define BOX_PROLOGUE                                        \
  const int bpb_ = blockDim.x / tpb;                        \
  const int slot_ = threadIdx.x / tpb;                      \
  const int lt = threadIdx.x - slot_ * tpb;                 \
  const int li_ = blockIdx.x * bpb_ + slot_;                \
  if (li_ >= n) return;                                     \
  const int id = list[li_];

// ---------------------------------------------------------------------------
// Gauss-Seidel red-black half-sweep
// mg$D_box_gsrb_lpl (m_aX_mg.f90:769-811), mg2_box_gsrb_clpl (:1413-1439)
// ---------------------------------------------------------------------------

template <int D, bool CYL>
__global__ __launch_bounds__(256) void k_gsrb(Dev a, const int* __restrict__ list, int n,
                                              int nc, int cntr, int fused, int tpb,
                                              int i_phi, int i_rhs) {
  BOX_PROLOGUE
  double* p = boxp(a, i_phi, id);
  const double* r = boxp(a, i_rhs, id);
  const double dx2 = a.dr[id] * a.dr[id];
  const int half = nc >> 1;
  const int ncol = (D == 2 ? nc * nc : nc * nc * nc) >> 1;
  const int ioff = CYL ? (a.ix[id * 3] - 1) * nc : 0;
  for (int q = lt; q < ncol; q += tpb) {
    const int m = q % half, row = q / half;
    const int j = row % nc + 1;
    const int k = D == 3 ? row / nc + 1 : 0;
    const int i0 = 2 - ((cntr ^ (D == 3 ? k + j : j)) & 1);
    const int i = i0 + 2 * m;
    // neighbour value at (i+di, j+dj, k+dk); ghosts recomputed when fused
    auto nbv = [&](int di, int dj, int dk, int nb) -> double {
      const int ii = i + di, jj = j + dj, kk = k + dk;
      const bool ghost = (ii < 1 || ii > nc || jj < 1 || jj > nc ||
                          (D == 3 && (kk < 1 || kk > nc)));
      if (ghost && fused) return ghost_mg<D>(a, p, i_phi, id, nc, nb, ii, jj, kk);
      return p[cidx<D>(nc, ii, jj, kk)];
    };
    const int c0 = cidx<D>(nc, i, j, k);
    if (D == 2) {
      if (CYL) {
        const double den = i + ioff - 0.5;
        const double rf1 = (double)(i + ioff - 1) / den;
        const double rf2 = (double)(i + ioff) / den;
        const double v = 0.25 * (rf1 * nbv(-1, 0, 0, 1) + rf2 * nbv(1, 0, 0, 2) +
                                 nbv(0, 1, 0, 4) + nbv(0, -1, 0, 3) - dx2 * r[c0]);
        p[c0] = v;
      } else {
        const double v = 0.25 * (nbv(1, 0, 0, 2) + nbv(-1, 0, 0, 1) + nbv(0, 1, 0, 4) +
                                 nbv(0, -1, 0, 3) - dx2 * r[c0]);
        p[c0] = v;
      }
    } else {
      const double sixth = 1 / 6.0;
      const double v = sixth * (nbv(1, 0, 0, 2) + nbv(-1, 0, 0, 1) + nbv(0, 1, 0, 4) +
                                nbv(0, -1, 0, 3) + nbv(0, 0, 1, 6) + nbv(0, 0, -1, 5) -
                                dx2 * r[c0]);
      p[c0] = v;
    }
  }
}

// Consecutive entries of a box list go to the same XCD (the hardware deals
// workgroups round-robin over the 8 XCDs), so Morton-adjacent boxes share an
// L2 and neighbour faces are L2 hits.
__device__ __forceinline__ int xcd_block() {
  const int b = blockIdx.x, g = gridDim.x, per = g >> 3;
  return b < (per << 3) ? (b & 7) * per + (b >> 3) : b;
}

// Box-size-specialised half-sweep. The other colour's half of the box (ghost
// slots included) is staged in LDS; when fused, that colour's face ghost
// slots are first recomputed from the current state (the values the
// reference's ghost pass after the previous half-sweep would store). Then
// every lane updates cells of colour cntr&1 from LDS and rhs.
template <int D, bool CYL, int NC>
__global__ __launch_bounds__(256) void k_gsrb_lds(Dev a, const int* __restrict__ list, int n,
                                                  int cntr, int fused, int i_phi, int i_rhs) {
  constexpr int N2 = NC + 2;
  constexpr int BL = D == 2 ? N2 * N2 : N2 * N2 * N2;
  constexpr int HALF = BL / 2;
  constexpr int NCOL = (D == 2 ? NC * NC : NC * NC * NC) / 2;
  constexpr int TPB = NCOL > 128 ? 256 : (NCOL > 64 ? 128 : 64);
  constexpr int BPB = 256 / TPB;
  constexpr int PER_FACE = D == 2 ? NC : NC * NC;
  constexpr int NGH = 2 * D * PER_FACE;
  constexpr int HNC = NC / 2;
  __shared__ double sh[BPB * HALF];
  const int slot = threadIdx.x / TPB, lt = threadIdx.x % TPB;
  const int li = xcd_block() * BPB + slot;
  const bool act = li < n;
  const int id = act ? list[li] : 0;
  const int c = cntr & 1, oc = c ^ 1;
  double* s = sh + slot * HALF;
  double* p = act ? boxp(a, i_phi, id) : nullptr;
  if (act) {
    for (int q = lt; q < HALF; q += TPB) s[q] = p[oc * HALF + q];
    if (fused) {
      for (int q = lt; q < NGH; q += TPB) {
        const int nb = q / PER_FACE + 1;
        const int t = q - (nb - 1) * PER_FACE;
        const int d = d_nb_dim(nb);
        const int ta = d == 0 ? 1 : 0, tb = d == 2 ? 1 : 2;
        int g[3] = {0, 0, 0};
        g[d] = d_nb_low(nb) ? 0 : NC + 1;
        g[ta] = t % NC + 1;
        if (D == 3) g[tb] = t / NC + 1;
        if (((g[0] + g[1] + g[2]) & 1) != oc) continue;
        const int lex = D == 2 ? g[0] + N2 * g[1] : g[0] + N2 * (g[1] + N2 * g[2]);
        s[lex >> 1] = ghost_mg<D>(a, p, i_phi, id, NC, nb, g[0], g[1], g[2]);
      }
    }
  }
  __syncthreads();
  if (!act) return;
  const double* r = boxp(a, i_rhs, id) + c * HALF;
  double* pc = p + c * HALF;
  const double dx2 = a.dr[id] * a.dr[id];
  const int ioff = CYL ? (a.ix[id * 3] - 1) * NC : 0;
  for (int q = lt; q < NCOL; q += TPB) {
    const int mm = q % HNC, row = q / HNC;
    const int j = row % NC + 1;
    const int k = D == 3 ? row / NC + 1 : 0;
    const int i = 2 * mm + 2 - ((c + j + k) & 1);
    const int lex = D == 2 ? i + N2 * j : i + N2 * (j + N2 * k);
    const int x = lex >> 1;
    const double xp = s[(lex + 1) >> 1], xm = s[(lex - 1) >> 1];
    const double yp = s[(lex + N2) >> 1], ym = s[(lex - N2) >> 1];
    double v;
    if (D == 2) {
      if (CYL) {
        const double den = i + ioff - 0.5;
        const double rf1 = (double)(i + ioff - 1) / den;
        const double rf2 = (double)(i + ioff) / den;
        v = 0.25 * (rf1 * xm + rf2 * xp + yp + ym - dx2 * r[x]);
      } else {
        v = 0.25 * (xp + xm + yp + ym - dx2 * r[x]);
      }
    } else {
      const double zp = s[(lex + N2 * N2) >> 1], zm = s[(lex - N2 * N2) >> 1];
      const double sixth = 1 / 6.0;
      v = sixth * (xp + xm + yp + ym + zp + zm - dx2 * r[x]);
    }
    pc[x] = v;
  }
}

template <int D, bool CYL>
static bool launch_gsrb_lds(Pool& P, const int* list, int n, int nc, int cntr, int fused) {
  auto go = [&](auto kern, int ncol) {
    const int tpb = ncol > 128 ? 256 : (ncol > 64 ? 128 : 64);
    const int bpb = 256 / tpb;
    hipLaunchKernelGGL(kern, dim3((n + bpb - 1) / bpb), dim3(256), 0, P.stream, P.dev(), list, n,
                       cntr, fused, P.mg.i_phi, P.mg.i_rhs);
  };
  const int ncol = (D == 2 ? nc * nc : nc * nc * nc) / 2;
  switch (nc) {
    case 2: go(k_gsrb_lds<D, CYL, 2>, ncol); return true;
    case 4: go(k_gsrb_lds<D, CYL, 4>, ncol); return true;
    case 8: go(k_gsrb_lds<D, CYL, 8>, ncol); return true;
    case 16: go(k_gsrb_lds<D, CYL, 16>, ncol); return true;
    case 32:
      if (D == 2) {
        go(k_gsrb_lds<2, CYL, 32>, ncol);
        return true;
      }
      return false;
    default: return false;
  }
}

// ---------------------------------------------------------------------------
// Face ghost cells of one variable: a$D_gc_box (m_aX_gc.f90:50-68) with
// rb = 0 mg$D_sides_rb, 1 a$D_gc_interp (:90-185), 2 a$D_gc_interp_lim
// (:190-295); bc = 0 multigrid face rules, 1 Neumann (:298-330),
// 2 Dirichlet zero (:333-365).
// ---------------------------------------------------------------------------

template <int D>
__global__ __launch_bounds__(256) void k_gc(Dev a, const int* __restrict__ list, int n,
                                            int nc, int tpb, int iv, int rb, int bc) {
  BOX_PROLOGUE
  double* p = boxp(a, iv, id);
  const int per_face = D == 2 ? nc : nc * nc;
  const int total = 2 * D * per_face;
  const double sixth = 1 / 6.0, third = 1 / 3.0;
  for (int q = lt; q < total; q += tpb) {
    const int nb = q / per_face + 1;
    const int t = q - (nb - 1) * per_face;
    const int d = d_nb_dim(nb);
    const bool low = d_nb_low(nb);
    const int ta = d == 0 ? 1 : 0, tb = d == 2 ? 1 : 2;
    int g[3] = {0, 0, 0};
    g[d] = low ? 0 : nc + 1;
    g[ta] = t % nc + 1;
    if (D == 3) g[tb] = t / nc + 1;
    const int nb_id = a.nbs[id * 6 + nb - 1];
    double v;
    if (nb_id > 0 || (nb_id == 0 && rb == 0) || (nb_id < 0 && bc == 0)) {
      v = ghost_mg<D>(a, p, iv, id, nc, nb, g[0], g[1], g[2]);
    } else if (nb_id < 0) {
      int c[3] = {g[0], g[1], g[2]};
      c[d] = low ? 1 : nc;
      const double cv = p[cidx<D>(nc, c[0], c[1], c[2])];
      v = bc == 1 ? cv : -cv;
    } else {
      // coarse-neighbour interpolation
      const int pid = a.parent[id];
      const int pn = a.nbs[pid * 6 + nb - 1];
      const int pnc = a.ncell[pn];
      int off[3];
      child_off<D>(a, id, nc, nb, off);
      const int ixc = low ? nc : 1, ixf = low ? 1 : nc;
      const double* qv = boxp(a, iv, pn);
      int f[3] = {g[0], g[1], g[2]};
      f[d] = ixf;
      const double fv = p[cidx<D>(nc, f[0], f[1], f[2])];
      if (D == 2) {
        const int x = g[ta];
        const int c1 = off[ta] + ((x + 1) >> 1), c2 = c1 + 1 - 2 * (x & 1);
        int s1[3], s2[3];
        s1[d] = s2[d] = ixc;
        s1[ta] = c1;
        s2[ta] = c2;
        const double v1 = qv[cidx<D>(pnc, s1[0], s1[1], 0)];
        const double v2 = qv[cidx<D>(pnc, s2[0], s2[1], 0)];
        v = 0.5 * v1 + sixth * v2 + third * fv;
        if (rb == 2 && v > 2 * v1) v = 2 * v1;
      } else {
        // the two tangential shifts enter in the order of the reference:
        // x-faces vary j then k, y-faces i then k, z-faces j then i
        const int x = g[ta], y = g[tb];
        const int a1 = off[ta] + ((x + 1) >> 1), a2 = a1 + 1 - 2 * (x & 1);
        const int b1 = off[tb] + ((y + 1) >> 1), b2 = b1 + 1 - 2 * (y & 1);
        int s1[3], s2[3], s3[3];
        s1[d] = s2[d] = s3[d] = ixc;
        s1[ta] = a1; s1[tb] = b1;
        if (d == 2) {
          s2[ta] = a1; s2[tb] = b2;
          s3[ta] = a2; s3[tb] = b1;
        } else {
          s2[ta] = a2; s2[tb] = b1;
          s3[ta] = a1; s3[tb] = b2;
        }
        const double v1 = qv[cidx<D>(pnc, s1[0], s1[1], s1[2])];
        const double v2 = qv[cidx<D>(pnc, s2[0], s2[1], s2[2])];
        const double v3 = qv[cidx<D>(pnc, s3[0], s3[1], s3[2])];
        v = third * v1 + sixth * v2 + sixth * v3 + third * fv;
        if (rb == 2 && v > 2 * v1) v = 2 * v1;
      }
    }
    p[cidx<D>(nc, g[0], g[1], g[2])] = v;
  }
}

// ---------------------------------------------------------------------------
// Laplacian of phi at interior cell c0: mg$D_box_lpl (m_aX_mg.f90:814-848),
// mg2_box_clpl (:1442-1464)
// ---------------------------------------------------------------------------

template <int D, bool CYL>
__device__ __forceinline__ double lpl_at(const double* p, int nc, int i, int j, int k,
                                         double inv_dr_sq, int ioff) {
  if (D == 2) {
    if (CYL) {
      const double den = i + ioff - 0.5;
      const double rf1 = (double)(i + ioff - 1) / den;
      const double rf2 = (double)(i + ioff) / den;
      return (rf1 * p[cidx<D>(nc, i - 1, j, 0)] + rf2 * p[cidx<D>(nc, i + 1, j, 0)] +
              p[cidx<D>(nc, i, j - 1, 0)] + p[cidx<D>(nc, i, j + 1, 0)] -
              4 * p[cidx<D>(nc, i, j, 0)]) * inv_dr_sq;
    }
    return inv_dr_sq * (p[cidx<D>(nc, i - 1, j, 0)] + p[cidx<D>(nc, i + 1, j, 0)] +
                        p[cidx<D>(nc, i, j - 1, 0)] + p[cidx<D>(nc, i, j + 1, 0)] -
                        4 * p[cidx<D>(nc, i, j, 0)]);
  }
  return inv_dr_sq * (p[cidx<D>(nc, i - 1, j, k)] + p[cidx<D>(nc, i + 1, j, k)] +
                      p[cidx<D>(nc, i, j - 1, k)] + p[cidx<D>(nc, i, j + 1, k)] +
                      p[cidx<D>(nc, i, j, k - 1)] + p[cidx<D>(nc, i, j, k + 1)] -
                      6 * p[cidx<D>(nc, i, j, k)]);
}

// ---------------------------------------------------------------------------
// Residual + restriction to the parent (update_coarse / set_coarse_phi_rhs,
// m_aX_mg.f90:430-523, with residual_box :551-565 and mg$D_box_rstr_lpl
// :851-893). One lane per parent cell covered by the child: the 2^D fine
// residuals stay in registers. keep_res = 1 stores them in the child's tmp
// (set_coarse_phi_rhs); 0 leaves tmp as it was (update_coarse restores it).
// ---------------------------------------------------------------------------

template <int D, bool CYL>
__global__ __launch_bounds__(256) void k_resid_rstr(Dev a, const int* __restrict__ list,
                                                    int n, int nc, int tpb, int i_phi,
                                                    int i_rhs, int i_tmp, int keep_res) {
  BOX_PROLOGUE
  const int hnc = nc >> 1;
  const int pid = a.parent[id];
  const int pnc = a.ncell[pid];
  int off[3];
  child_off<D>(a, id, nc, 0, off);
  const double* p = boxp(a, i_phi, id);
  const double* r = boxp(a, i_rhs, id);
  double* t = boxp(a, i_tmp, id);
  double* pp = boxp(a, i_phi, pid);
  double* pt = boxp(a, i_tmp, pid);
  const double inv_dr_sq = 1 / (a.dr[id] * a.dr[id]);
  const int ioff = CYL ? (a.ix[id * 3] - 1) * nc : 0;
  const int total = D == 2 ? hnc * hnc : hnc * hnc * hnc;
  for (int q = lt; q < total; q += tpb) {
    const int ci = q % hnc + 1;
    const int cj = (q / hnc) % hnc + 1;
    const int ck = D == 3 ? q / (hnc * hnc) + 1 : 0;
    const int fi = 2 * ci - 1, fj = 2 * cj - 1, fk = D == 3 ? 2 * ck - 1 : 0;
    double sr = 0, sp = 0;
#pragma unroll
    for (int dk = 0; dk < (D == 3 ? 2 : 1); dk++)
#pragma unroll
      for (int dj = 0; dj < 2; dj++)
#pragma unroll
        for (int di = 0; di < 2; di++) {
          const int ii = fi + di, jj = fj + dj, kk = fk + dk;
          const int c = cidx<D>(nc, ii, jj, kk);
          const double res = r[c] - lpl_at<D, CYL>(p, nc, ii, jj, kk, inv_dr_sq, ioff);
          if (keep_res) t[c] = res;
          sr = sr + res;
          sp = sp + p[c];
        }
    const int pc = cidx<D>(pnc, off[0] + ci, off[1] + cj, D == 3 ? off[2] + ck : 0);
    const double w = D == 2 ? 0.25 : 0.125;
    pt[pc] = w * sr;
    pp[pc] = w * sp;
  }
}

// Plain restriction of one variable to the parent (mg$D_box_rstr_lpl), and
// for init_phi_rhs (m_aX_mg.f90:526-549) the clearing of the child's phi.
template <int D>
__global__ __launch_bounds__(256) void k_rstr_clear(Dev a, const int* __restrict__ list,
                                                    int n, int nc, int tpb, int iv,
                                                    int i_clear) {
  BOX_PROLOGUE
  const int hnc = nc >> 1;
  const int pid = a.parent[id];
  const int pnc = a.ncell[pid];
  int off[3];
  child_off<D>(a, id, nc, 0, off);
  const double* c = boxp(a, iv, id);
  double* pv = boxp(a, iv, pid);
  const int total = D == 2 ? hnc * hnc : hnc * hnc * hnc;
  for (int q = lt; q < total; q += tpb) {
    const int ci = q % hnc + 1;
    const int cj = (q / hnc) % hnc + 1;
    const int ck = D == 3 ? q / (hnc * hnc) + 1 : 0;
    const int fi = 2 * ci - 1, fj = 2 * cj - 1, fk = D == 3 ? 2 * ck - 1 : 0;
    double s = 0;
#pragma unroll
    for (int dk = 0; dk < (D == 3 ? 2 : 1); dk++)
#pragma unroll
      for (int dj = 0; dj < 2; dj++)
#pragma unroll
        for (int di = 0; di < 2; di++) s = s + c[cidx<D>(nc, fi + di, fj + dj, fk + dk)];
    pv[cidx<D>(pnc, off[0] + ci, off[1] + cj, D == 3 ? off[2] + ck : 0)] =
        (D == 2 ? 0.25 : 0.125) * s;
  }
  if (i_clear > 0) {
    // runs after all reads of this box (other lanes read `iv`, not i_clear)
    double* z = boxp(a, i_clear, id);
    const int bl = D == 2 ? (nc + 2) * (nc + 2) : (nc + 2) * (nc + 2) * (nc + 2);
    for (int q = lt; q < bl; q += tpb) z[q] = 0;
  }
}

// a$D_restrict_box (m_aX_restrict.f90:67-142) for all children of the given
// parents: Cartesian average, or in 2D cylindrical coordinates the
// radius-weighted average with the radius of parent cell (i, j) as written
// at m_aX_restrict.f90:106.
template <int D>
__global__ __launch_bounds__(256) void k_restrict_box(Dev a, const int* __restrict__ list,
                                                      int n, int nc, int tpb, int iv) {
  BOX_PROLOGUE
  const int pid = a.parent[id];
  const int pnc = a.ncell[pid];
  const int hnc = nc >> 1;
  int off[3];
  child_off<D>(a, id, nc, 0, off);
  const double* c = boxp(a, iv, id);
  double* pv = boxp(a, iv, pid);
  const int total = D == 2 ? hnc * hnc : hnc * hnc * hnc;
  for (int q = lt; q < total; q += tpb) {
    const int ci = q % hnc + 1;
    const int cj = (q / hnc) % hnc + 1;
    const int ck = D == 3 ? q / (hnc * hnc) + 1 : 0;
    const int fi = 2 * ci - 1, fj = 2 * cj - 1, fk = D == 3 ? 2 * ck - 1 : 0;
    const int pc = cidx<D>(pnc, off[0] + ci, off[1] + cj, D == 3 ? off[2] + ck : 0);
    if (D == 2 && a.coord == 1) {
      const double dr16 = 0.0625 * a.dr[pid];
      const double rr = a.rmin[pid * 3] + (ci - 0.5) * a.dr[pid];
      const double rfac = dr16 / rr;
      double s1 = 0, s2 = 0;
      s1 = s1 + c[cidx<D>(nc, fi, fj, 0)];
      s1 = s1 + c[cidx<D>(nc, fi, fj + 1, 0)];
      s2 = s2 + c[cidx<D>(nc, fi + 1, fj, 0)];
      s2 = s2 + c[cidx<D>(nc, fi + 1, fj + 1, 0)];
      pv[pc] = (0.25 - rfac) * s1 + (0.25 + rfac) * s2;
    } else {
      double s = 0;
#pragma unroll
      for (int dk = 0; dk < (D == 3 ? 2 : 1); dk++)
#pragma unroll
        for (int dj = 0; dj < 2; dj++)
#pragma unroll
          for (int di = 0; di < 2; di++) s = s + c[cidx<D>(nc, fi + di, fj + dj, fk + dk)];
      pv[pc] = (D == 2 ? 0.25 : 0.125) * s;
    }
  }
}

// ---------------------------------------------------------------------------
// FAS coarse right-hand side on parents: rhs = A phi (interior), then
// rhs += tmp over the whole box (a$D_box_add_cc), then optionally tmp = phi
// over the whole box (a$D_box_copy_cc); m_aX_mg.f90:478-485, 516-522.
// ---------------------------------------------------------------------------

template <int D, bool CYL>
__global__ __launch_bounds__(256) void k_coarse_rhs(Dev a, const int* __restrict__ list,
                                                    int n, int nc, int tpb, int i_phi,
                                                    int i_rhs, int i_tmp, int copy_tmp) {
  BOX_PROLOGUE
  const double* p = boxp(a, i_phi, id);
  double* r = boxp(a, i_rhs, id);
  double* t = boxp(a, i_tmp, id);
  const double inv_dr_sq = 1 / (a.dr[id] * a.dr[id]);
  const int ioff = CYL ? (a.ix[id * 3] - 1) * nc : 0;
  const int n2 = nc + 2;
  const int bl = D == 2 ? n2 * n2 : n2 * n2 * n2;
  for (int q = lt; q < bl; q += tpb) {
    const int i = q % n2, j = (q / n2) % n2, k = D == 3 ? q / (n2 * n2) : 0;
    const int x = cidx<D>(nc, i, j, k);
    const bool inner = i >= 1 && i <= nc && j >= 1 && j <= nc &&
                       (D == 2 || (k >= 1 && k <= nc));
    double rv = inner ? lpl_at<D, CYL>(p, nc, i, j, k, inv_dr_sq, ioff) : r[x];
    rv = rv + t[x];
    r[x] = rv;
    if (copy_tmp) t[x] = p[x];
  }
}

// correct_children, first half (m_aX_mg.f90:385-392): tmp = phi - tmp on the
// whole parent box
__global__ __launch_bounds__(256) void k_sub_whole(Dev a, const int* __restrict__ list,
                                                   int n, int bl, int tpb, int i_phi,
                                                   int i_tmp) {
  BOX_PROLOGUE
  const double* p = boxp(a, i_phi, id);
  double* t = boxp(a, i_tmp, id);
  for (int q = lt; q < bl; q += tpb) t[q] = p[q] - t[q];
}

// correct_children, second half: mg$D_box_corr_lpl (m_aX_mg.f90:714-766) on
// every child box in the list
template <int D>
__global__ __launch_bounds__(256) void k_corr(Dev a, const int* __restrict__ list, int n,
                                              int nc, int tpb, int i_phi, int i_tmp) {
  BOX_PROLOGUE
  const int pid = a.parent[id];
  const int pnc = a.ncell[pid];
  int off[3];
  child_off<D>(a, id, nc, 0, off);
  double* f = boxp(a, i_phi, id);
  const double* c = boxp(a, i_tmp, pid);
  const int total = D == 2 ? nc * nc : nc * nc * nc;
  for (int q = lt; q < total; q += tpb) {
    const int i = q % nc + 1, j = (q / nc) % nc + 1, k = D == 3 ? q / (nc * nc) + 1 : 0;
    const int i1 = off[0] + ((i + 1) >> 1), i2 = i1 + 1 - 2 * (i & 1);
    const int j1 = off[1] + ((j + 1) >> 1), j2 = j1 + 1 - 2 * (j & 1);
    const int x = cidx<D>(nc, i, j, k);
    if (D == 2) {
      f[x] = f[x] + 0.5 * c[cidx<D>(pnc, i1, j1, 0)] +
             0.25 * (c[cidx<D>(pnc, i2, j1, 0)] + c[cidx<D>(pnc, i1, j2, 0)]);
    } else {
      const int k1 = off[2] + ((k + 1) >> 1), k2 = k1 + 1 - 2 * (k & 1);
      f[x] = f[x] + 0.25 * (c[cidx<D>(pnc, i1, j1, k1)] + c[cidx<D>(pnc, i2, j1, k1)] +
                            c[cidx<D>(pnc, i1, j2, k1)] + c[cidx<D>(pnc, i1, j1, k2)]);
    }
  }
}

// whole-box copy (a$D_box_copy_cc, m_aX_utils.f90:421-429)
__global__ __launch_bounds__(256) void k_copy_whole(Dev a, const int* __restrict__ list,
                                                    int n, int bl, int tpb, int from,
                                                    int to) {
  BOX_PROLOGUE
  const double* s = boxp(a, from, id);
  double* d = boxp(a, to, id);
  for (int q = lt; q < bl; q += tpb) d[q] = s[q];
}

// residual_box (m_aX_mg.f90:551-565): tmp = rhs - A phi on the interior
template <int D, bool CYL>
__global__ __launch_bounds__(256) void k_residual(Dev a, const int* __restrict__ list,
                                                  int n, int nc, int tpb, int i_phi,
                                                  int i_rhs, int i_tmp) {
  BOX_PROLOGUE
  const double* p = boxp(a, i_phi, id);
  const double* r = boxp(a, i_rhs, id);
  double* t = boxp(a, i_tmp, id);
  const double inv_dr_sq = 1 / (a.dr[id] * a.dr[id]);
  const int ioff = CYL ? (a.ix[id * 3] - 1) * nc : 0;
  const int total = D == 2 ? nc * nc : nc * nc * nc;
  for (int q = lt; q < total; q += tpb) {
    const int i = q % nc + 1, j = (q / nc) % nc + 1, k = D == 3 ? q / (nc * nc) + 1 : 0;
    const int x = cidx<D>(nc, i, j, k);
    t[x] = r[x] - lpl_at<D, CYL>(p, nc, i, j, k, inv_dr_sq, ioff);
  }
}

// ---------------------------------------------------------------------------
// reductions over leaf interiors: max |v| (mode 0), min and max (mode 1).
// Wavefront shuffles, then one atomic per wavefront on the ordered-integer
// image of the double.
// ---------------------------------------------------------------------------

__device__ __forceinline__ unsigned long long ord_key(double v) {
  unsigned long long u = __double_as_longlong(v);
  return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}

__device__ __forceinline__ double wave_max(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
  return v;
}
__device__ __forceinline__ double wave_min(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
  return v;
}

template <int D>
__global__ __launch_bounds__(256) void k_minmax(Dev a, const int* __restrict__ list, int n,
                                                int nc, int tpb, int iv, int absval,
                                                int whole, unsigned long long* out) {
  const int bpb_ = blockDim.x / tpb;
  const int slot_ = threadIdx.x / tpb;
  const int lt = threadIdx.x - slot_ * tpb;
  const int li_ = blockIdx.x * bpb_ + slot_;
  double lo = 1.7976931348623157e308, hi = -1.7976931348623157e308;
  if (li_ < n) {
    const int id = list[li_];
    const double* p = boxp(a, iv, id);
    const int m = whole ? nc + 2 : nc, o = whole ? 0 : 1;
    const int total = D == 2 ? m * m : m * m * m;
    for (int q = lt; q < total; q += tpb) {
      const int i = q % m + o, j = (q / m) % m + o, k = D == 3 ? q / (m * m) + o : 0;
      double v = p[cidx<D>(nc, i, j, k)];
      if (absval) v = fabs(v);
      lo = fmin(lo, v);
      hi = fmax(hi, v);
    }
  }
  lo = wave_min(lo);
  hi = wave_max(hi);
  __shared__ double wl[4], wh[4];
  if ((threadIdx.x & 63) == 0) {
    wl[threadIdx.x >> 6] = lo;
    wh[threadIdx.x >> 6] = hi;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < (int)(blockDim.x >> 6); w++) {
      lo = fmin(lo, wl[w]);
      hi = fmax(hi, wh[w]);
    }
    atomicMin(&out[0], ord_key(lo));
    atomicMax(&out[1], ord_key(hi));
  }
}

// ---------------------------------------------------------------------------
// host orchestration
// ---------------------------------------------------------------------------

static void check(hipError_t e, const char* what) {
  if (e != hipSuccess)
    throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(e));
}

static void post_launch(const char* what) { check(hipGetLastError(), what); }

template <int D>
static void gsrb_boxes(Pool& P, int lvl, int n_cycle) {
  // gsrb_boxes, m_aX_mg.f90:403-426
  const int n = P.n_ids(lvl);
  if (n == 0) return;
  const int nc = P.nc_lvl(lvl);
  const int ncol = (D == 2 ? nc * nc : nc * nc * nc) / 2;
  Geo g = geo(n, ncol);
  Dev a = P.dev();
  const bool cyl = D == 2 && P.tree.coord == 1 && P.mg.auto_ops;
  const double cells = (double)n * (D == 2 ? nc * nc : nc * nc * nc);
  for (int it = 1; it <= 2 * n_cycle; it++) {
    const int fused = it > 1;
    // algorithmic bytes of a half-sweep: 1.5 words per cell
    const int pr = P.prof_begin(PK_GSRB, lvl, 1.5 * 8 * cells);
    const bool done = cyl ? launch_gsrb_lds<2, true>(P, P.ids(lvl), n, nc, it, fused)
                          : launch_gsrb_lds<D, false>(P, P.ids(lvl), n, nc, it, fused);
    if (done) {
    } else if (cyl)
      hipLaunchKernelGGL((k_gsrb<2, true>), dim3(g.grid), dim3(256), 0, P.stream, a,
                         P.ids(lvl), n, nc, it, fused, g.tpb, P.mg.i_phi, P.mg.i_rhs);
    else
      hipLaunchKernelGGL((k_gsrb<D, false>), dim3(g.grid), dim3(256), 0, P.stream, a,
                         P.ids(lvl), n, nc, it, fused, g.tpb, P.mg.i_phi, P.mg.i_rhs);
    post_launch("gsrb");
    P.prof_end(pr);
  }
  gc_level(P, lvl, P.mg.i_phi, 0, 0);
}

void gc_level(Pool& P, int lvl, int iv, int rb, int bc) {
  gc_list(P, lvl, P.ids(lvl), P.n_ids(lvl), iv, rb, bc);
}

void gc_list(Pool& P, int lvl, const int* list, int n, int iv, int rb, int bc) {
  if (n == 0) return;
  const int nc = P.nc_lvl(lvl);
  const int D = P.tree.dim;
  const int work = 2 * D * (D == 2 ? nc : nc * nc);
  Geo g = geo(n, work);
  Dev a = P.dev();
  const int pr = P.prof_begin(PK_GC, lvl, 0.0);
  if (D == 2)
    hipLaunchKernelGGL((k_gc<2>), dim3(g.grid), dim3(256), 0, P.stream, a, list, n, nc, g.tpb,
                       iv, rb, bc);
  else
    hipLaunchKernelGGL((k_gc<3>), dim3(g.grid), dim3(256), 0, P.stream, a, list, n, nc, g.tpb,
                       iv, rb, bc);
  post_launch("gc");
  P.prof_end(pr);
}

template <int D>
static void resid_rstr(Pool& P, int lvl, int keep_res) {
  const int n = P.n_ids(lvl);
  if (n == 0) return;
  const int nc = P.nc_lvl(lvl);
  const int hnc = nc / 2;
  Geo g = geo(n, D == 2 ? hnc * hnc : hnc * hnc * hnc);
  Dev a = P.dev();
  const bool cyl = D == 2 && P.tree.coord == 1 && P.mg.auto_ops;
  const double cells = (double)n * (D == 2 ? nc * nc : nc * nc * nc);
  const int pr = P.prof_begin(PK_RESID_RSTR, lvl, (2.0 + 2.0 / (1 << D)) * 8 * cells);
  if (cyl)
    hipLaunchKernelGGL((k_resid_rstr<2, true>), dim3(g.grid), dim3(256), 0, P.stream, a,
                       P.ids(lvl), n, nc, g.tpb, P.mg.i_phi, P.mg.i_rhs, P.mg.i_tmp, keep_res);
  else
    hipLaunchKernelGGL((k_resid_rstr<D, false>), dim3(g.grid), dim3(256), 0, P.stream, a,
                       P.ids(lvl), n, nc, g.tpb, P.mg.i_phi, P.mg.i_rhs, P.mg.i_tmp, keep_res);
  post_launch("resid_rstr");
  P.prof_end(pr);
}

template <int D>
static void coarse_rhs(Pool& P, int lvl, int copy_tmp) {
  const int n = P.n_parents(lvl);
  if (n == 0) return;
  const int nc = P.nc_lvl(lvl);
  const int bl = D == 2 ? (nc + 2) * (nc + 2) : (nc + 2) * (nc + 2) * (nc + 2);
  Geo g = geo(n, bl);
  Dev a = P.dev();
  const bool cyl = D == 2 && P.tree.coord == 1 && P.mg.auto_ops;
  const double cells = (double)n * (D == 2 ? nc * nc : nc * nc * nc);
  const int pr = P.prof_begin(PK_COARSE_RHS, lvl, 4.0 * 8 * cells);
  if (cyl)
    hipLaunchKernelGGL((k_coarse_rhs<2, true>), dim3(g.grid), dim3(256), 0, P.stream, a,
                       P.parents(lvl), n, nc, g.tpb, P.mg.i_phi, P.mg.i_rhs, P.mg.i_tmp,
                       copy_tmp);
  else
    hipLaunchKernelGGL((k_coarse_rhs<D, false>), dim3(g.grid), dim3(256), 0, P.stream, a,
                       P.parents(lvl), n, nc, g.tpb, P.mg.i_phi, P.mg.i_rhs, P.mg.i_tmp,
                       copy_tmp);
  post_launch("coarse_rhs");
  P.prof_end(pr);
}

template <int D>
static void update_coarse(Pool& P, int lvl) {
  // update_coarse, m_aX_mg.f90:430-486
  resid_rstr<D>(P, lvl, 0);
  gc_level(P, lvl - 1, P.mg.i_phi, 0, 0);
  coarse_rhs<D>(P, lvl - 1, 1);
}

template <int D>
static void set_coarse_phi_rhs(Pool& P, int lvl) {
  // set_coarse_phi_rhs, m_aX_mg.f90:490-523
  if (lvl == P.tree.max_lvl) gc_level(P, lvl, P.mg.i_phi, 0, 0);
  resid_rstr<D>(P, lvl, 1);
  gc_level(P, lvl - 1, P.mg.i_phi, 0, 0);
  coarse_rhs<D>(P, lvl - 1, 0);
}

template <int D>
static void correct_children(Pool& P, int lvl) {
  // correct_children over the parents of lvl, m_aX_mg.f90:374-401
  const int n = P.n_parents(lvl);
  if (n == 0) return;
  const int nc = P.nc_lvl(lvl);
  const int bl = D == 2 ? (nc + 2) * (nc + 2) : (nc + 2) * (nc + 2) * (nc + 2);
  Dev a = P.dev();
  Geo g = geo(n, bl);
  const double pcells = (double)n * (D == 2 ? nc * nc : nc * nc * nc);
  int pr = P.prof_begin(PK_CORR, lvl, 3.0 * 8 * pcells);
  hipLaunchKernelGGL(k_sub_whole, dim3(g.grid), dim3(256), 0, P.stream, a, P.parents(lvl), n,
                     bl, g.tpb, P.mg.i_phi, P.mg.i_tmp);
  post_launch("corr_prep");
  P.prof_end(pr);
  // every child of these parents is a box of level lvl+1
  const int nch = P.n_ids(lvl + 1);
  const int cnc = P.nc_lvl(lvl + 1);
  Geo g2 = geo(nch, D == 2 ? cnc * cnc : cnc * cnc * cnc);
  const double ccells = (double)nch * (D == 2 ? cnc * cnc : cnc * cnc * cnc);
  pr = P.prof_begin(PK_CORR, lvl + 1, (2.0 + 1.0 / (1 << D)) * 8 * ccells);
  hipLaunchKernelGGL((k_corr<D>), dim3(g2.grid), dim3(256), 0, P.stream, a, P.ids(lvl + 1), nch,
                     cnc, g2.tpb, P.mg.i_phi, P.mg.i_tmp);
  post_launch("corr");
  P.prof_end(pr);
}

static void copy_level(Pool& P, int lvl, int from, int to) {
  const int n = P.n_ids(lvl);
  if (n == 0) return;
  const int nc = P.nc_lvl(lvl);
  const int bl = P.tree.dim == 2 ? (nc + 2) * (nc + 2) : (nc + 2) * (nc + 2) * (nc + 2);
  Geo g = geo(n, bl);
  const int pr = P.prof_begin(PK_COPY, lvl,
                              2.0 * 8 * n * (double)(P.tree.dim == 2 ? nc * nc : nc * nc * nc));
  hipLaunchKernelGGL(k_copy_whole, dim3(g.grid), dim3(256), 0, P.stream, P.dev(), P.ids(lvl),
                     n, bl, g.tpb, from, to);
  post_launch("copy");
  P.prof_end(pr);
}

template <int D>
static void residual_level(Pool& P, int lvl, bool leaves_only) {
  const int n = leaves_only ? P.n_leaves(lvl) : P.n_ids(lvl);
  if (n == 0) return;
  const int nc = P.nc_lvl(lvl);
  Geo g = geo(n, D == 2 ? nc * nc : nc * nc * nc);
  const int* list = leaves_only ? P.leaves(lvl) : P.ids(lvl);
  const bool cyl = D == 2 && P.tree.coord == 1 && P.mg.auto_ops;
  if (cyl)
    hipLaunchKernelGGL((k_residual<2, true>), dim3(g.grid), dim3(256), 0, P.stream, P.dev(),
                       list, n, nc, g.tpb, P.mg.i_phi, P.mg.i_rhs, P.mg.i_tmp);
  else
    hipLaunchKernelGGL((k_residual<D, false>), dim3(g.grid), dim3(256), 0, P.stream, P.dev(),
                       list, n, nc, g.tpb, P.mg.i_phi, P.mg.i_rhs, P.mg.i_tmp);
  post_launch("residual");
}

template <int D>
static void init_phi_rhs(Pool& P) {
  // init_phi_rhs, m_aX_mg.f90:526-549
  for (int lvl = P.tree.max_lvl; lvl >= P.tree.min_lvl + 1; lvl--) {
    const int n = P.n_ids(lvl);
    if (n == 0) continue;
    const int nc = P.nc_lvl(lvl);
    Geo g = geo(n, D == 2 ? (nc + 2) * (nc + 2) : (nc + 2) * (nc + 2) * (nc + 2));
    hipLaunchKernelGGL((k_rstr_clear<D>), dim3(g.grid), dim3(256), 0, P.stream, P.dev(),
                       P.ids(lvl), n, nc, g.tpb, P.mg.i_rhs, P.mg.i_phi);
    post_launch("init_phi_rhs");
  }
}

template <int D>
static void fas_vcycle(Pool& P, int max_lvl, bool set_residual) {
  // mg$D_fas_vcycle, m_aX_mg.f90:208-255
  const int min_lvl = P.tree.min_lvl;
  for (int lvl = max_lvl; lvl >= min_lvl + 1; lvl--) {
    gsrb_boxes<D>(P, lvl, P.mg.n_down);
    update_coarse<D>(P, lvl);
  }
  gsrb_boxes<D>(P, min_lvl, P.mg.n_base);
  for (int lvl = min_lvl + 1; lvl <= max_lvl; lvl++) {
    correct_children<D>(P, lvl - 1);
    gc_level(P, lvl, P.mg.i_phi, 0, 0);
    gsrb_boxes<D>(P, lvl, P.mg.n_up);
  }
  if (set_residual)
    for (int lvl = min_lvl; lvl <= max_lvl; lvl++) residual_level<D>(P, lvl, false);
}

template <int D>
static void fas_fmg(Pool& P, bool set_residual, bool first_call) {
  // mg$D_fas_fmg, m_aX_mg.f90:164-203
  const int min_lvl = P.tree.min_lvl;
  if (first_call) init_phi_rhs<D>(P);
  else
    for (int lvl = P.tree.max_lvl; lvl >= min_lvl + 1; lvl--) set_coarse_phi_rhs<D>(P, lvl);
  for (int lvl = min_lvl; lvl <= P.tree.max_lvl; lvl++) {
    copy_level(P, lvl, P.mg.i_phi, P.mg.i_tmp);
    if (lvl > min_lvl) {
      correct_children<D>(P, lvl - 1);
      gc_level(P, lvl, P.mg.i_phi, 0, 0);
    }
    fas_vcycle<D>(P, lvl, set_residual && lvl == P.tree.max_lvl);
  }
}

static void mg_check(Pool& P) {
  // check_mg, m_aX_mg.f90:156-159
  if (!P.mg.ready) throw std::logic_error("check_mg: you haven't called mg_init_mg");
  P.sync_topology();
  if (P.mg.auto_ops && P.mg.i_eps > 0) {
    // mg$D_set_box_tag (m_aX_mg.f90:673-712) gives every box the plain
    // Laplacian when eps is 1 everywhere; boxes with eps /= 1 (the lpld
    // operator family) are not implemented on the GPU.
    double lo, hi;
    minmax_boxes(P, P.mg.i_eps, false, true, false, &lo, &hi);
    if (lo != 1.0 || hi != 1.0)
      throw std::runtime_error("multigrid: boxes with permittivity /= 1 are not supported");
  }
}

void mg_fmg(Pool& P, bool set_residual, bool first_call) {
  mg_check(P);
  if (P.tree.dim == 2) fas_fmg<2>(P, set_residual, first_call);
  else fas_fmg<3>(P, set_residual, first_call);
}

void mg_vcycle(Pool& P, int max_lvl, bool set_residual) {
  mg_check(P);
  if (max_lvl < P.tree.min_lvl || max_lvl > P.tree.max_lvl)
    throw std::invalid_argument("fas_vcycle: max_lvl outside the tree");
  if (P.tree.dim == 2) fas_vcycle<2>(P, max_lvl, set_residual);
  else fas_vcycle<3>(P, max_lvl, set_residual);
}

void minmax_boxes(Pool& P, int iv, bool absval, bool whole, bool leaves_only,
                  double* lo, double* hi) {
  P.sync_topology();
  unsigned long long init[2] = {~0ull, 0ull};
  check(hipMemcpyAsync(P.d_red, init, sizeof(init), hipMemcpyHostToDevice, P.stream),
        "minmax init");
  const int D = P.tree.dim;
  for (int lvl = P.tree.min_lvl; lvl <= P.tree.max_lvl; lvl++) {
    const int n = leaves_only ? P.n_leaves(lvl) : P.n_ids(lvl);
    if (n == 0) continue;
    const int* list = leaves_only ? P.leaves(lvl) : P.ids(lvl);
    const int nc = P.nc_lvl(lvl);
    Geo g = geo(n, D == 2 ? nc * nc : nc * nc * nc);
    if (D == 2)
      hipLaunchKernelGGL((k_minmax<2>), dim3(g.grid), dim3(256), 0, P.stream, P.dev(), list,
                         n, nc, g.tpb, iv, (int)absval, (int)whole,
                         (unsigned long long*)P.d_red);
    else
      hipLaunchKernelGGL((k_minmax<3>), dim3(g.grid), dim3(256), 0, P.stream, P.dev(), list,
                         n, nc, g.tpb, iv, (int)absval, (int)whole,
                         (unsigned long long*)P.d_red);
    post_launch("minmax");
  }
  unsigned long long out[2];
  check(hipMemcpyAsync(out, P.d_red, sizeof(out), hipMemcpyDeviceToHost, P.stream),
        "minmax read");
  check(hipStreamSynchronize(P.stream), "minmax sync");
  auto unkey = [](unsigned long long k) {
    unsigned long long u = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
    double v;
    std::memcpy(&v, &u, 8);
    return v;
  };
  *lo = unkey(out[0]);
  *hi = unkey(out[1]);
}

void mg_residual_norms(Pool& P, double* max_res, double* max_rhs) {
  mg_check(P);
  for (int lvl = P.tree.min_lvl; lvl <= P.tree.max_lvl; lvl++) {
    if (P.tree.dim == 2) residual_level<2>(P, lvl, true);
    else residual_level<3>(P, lvl, true);
  }
  double lo;
  minmax_boxes(P, P.mg.i_tmp, true, false, true, &lo, max_res);
  minmax_boxes(P, P.mg.i_rhs, true, false, true, &lo, max_rhs);
}

void gc_tree(Pool& P, int iv, int rb, int bc) {
  // a$D_gc_tree, m_aX_gc.f90:28-46
  P.sync_topology();
  for (int lvl = P.tree.min_lvl; lvl <= P.tree.max_lvl; lvl++) gc_level(P, lvl, iv, rb, bc);
}

void restrict_tree(Pool& P, int iv) {
  // a$D_restrict_tree, m_aX_restrict.f90:54-64: children of the parents of
  // lvl are exactly the boxes of lvl+1
  P.sync_topology();
  for (int lvl = P.tree.max_lvl - 1; lvl >= P.tree.min_lvl; lvl--) {
    const int n = P.n_ids(lvl + 1);
    if (n == 0) continue;
    const int nc = P.nc_lvl(lvl + 1);
    const int hnc = nc / 2;
    if (P.tree.dim == 2) {
      Geo g = geo(n, hnc * hnc);
      hipLaunchKernelGGL((k_restrict_box<2>), dim3(g.grid), dim3(256), 0, P.stream, P.dev(),
                         P.ids(lvl + 1), n, nc, g.tpb, iv);
    } else {
      Geo g = geo(n, hnc * hnc * hnc);
      hipLaunchKernelGGL((k_restrict_box<3>), dim3(g.grid), dim3(256), 0, P.stream, P.dev(),
                         P.ids(lvl + 1), n, nc, g.tpb, iv);
    }
    post_launch("restrict");
  }
}

void tree_copy_cc(Pool& P, int from, int to) {
  P.sync_topology();
  for (int lvl = P.tree.min_lvl; lvl <= P.tree.max_lvl; lvl++) copy_level(P, lvl, from, to);
}

}  // namespace a5g
