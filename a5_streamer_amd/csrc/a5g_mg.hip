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
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

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
// Laplacian of phi at interior cell c0: mg$D_box_lpl (m_aX_mg.f90:814-848),
// mg2_box_clpl (:1442-1464)
// ---------------------------------------------------------------------------

// The operators read the potential through u(i, j, k), so that a kernel can
// take it from HBM (cidx order) or from an LDS copy in another order.
template <int D, bool CYL, typename T, typename U>
__device__ __forceinline__ T lpl_u(U u, int i, int j, int k, T inv_dr_sq, int ioff) {
  if (D == 2) {
    if (CYL) {
      const double den = i + ioff - 0.5;
      const T rf1 = (T)((double)(i + ioff - 1) / den);
      const T rf2 = (T)((double)(i + ioff) / den);
      return (rf1 * u(i - 1, j, 0) + rf2 * u(i + 1, j, 0) + u(i, j - 1, 0) + u(i, j + 1, 0) -
              T(4) * u(i, j, 0)) * inv_dr_sq;
    }
    return inv_dr_sq * (u(i - 1, j, 0) + u(i + 1, j, 0) + u(i, j - 1, 0) + u(i, j + 1, 0) -
                        T(4) * u(i, j, 0));
  }
  return inv_dr_sq * (u(i - 1, j, k) + u(i + 1, j, k) + u(i, j - 1, k) + u(i, j + 1, k) +
                      u(i, j, k - 1) + u(i, j, k + 1) - T(6) * u(i, j, k));
}

template <int D, bool CYL, typename T>
__device__ __forceinline__ T lpl_at(const T* p, int nc, int i, int j, int k, T inv_dr_sq,
                                    int ioff) {
  return lpl_u<D, CYL, T>([&](int x, int y, int z) { return p[cidx<D>(nc, x, y, z)]; }, i, j,
                          k, inv_dr_sq, ioff);
}

// Operator with a permittivity eps: mg$D_box_lpld (m_aX_mg.f90:951-1001)
// and mg2_box_clpld (:1467-1495) at interior cell (i, j, k). Neighbours in the
// order i-1, i+1, j-1, j+1, k-1, k+1 of the reference's u(:) and a(:).
template <int D, bool CYL, typename T, typename U>
__device__ __forceinline__ T lpld_u(const double* e, int nc, U u, int i, int j, int k,
                                    T inv_dr_sq, int ioff) {
  const int di[6] = {-1, 1, 0, 0, 0, 0}, dj[6] = {0, 0, -1, 1, 0, 0}, dk[6] = {0, 0, 0, 0, -1, 1};
  const T u0 = u(i, j, k), a0 = (T)e[cidx_in<D>(nc, i, j, k)];
  T rf[2] = {T(1), T(1)};
  if (CYL) {
    const double den = i + ioff - 0.5;
    rf[0] = (T)((double)(i + ioff - 1) / den);
    rf[1] = (T)((double)(i + ioff) / den);
  }
  T s = 0;
#pragma unroll
  for (int m = 0; m < 2 * D; m++) {
    const T am = (T)e[cidx<D>(nc, i + di[m], j + dj[m], k + dk[m])];
    const T um = u(i + di[m], j + dj[m], k + dk[m]);
    if (CYL && m < 2) s = s + rf[m] * a0 * am / (a0 + am) * (um - u0);
    else s = s + a0 * am / (a0 + am) * (um - u0);
  }
  return inv_dr_sq * T(2) * s;
}

// The multigrid operator of box id: the permittivity form for boxes tagged
// constant- or variable-eps (mg$D_auto_op, m_aX_mg.f90:603-635), else the
// Laplacian.
// The Laplacian, or with a Helmholtz coefficient lpl(u) - helm * u (the
// operator of the photoionization solves)
template <int D, bool CYL, typename T, typename U>
__device__ __forceinline__ T lplh_u(const Dev& a, U u, int i, int j, int k, T inv_dr_sq,
                                    int ioff) {
  const T v = lpl_u<D, CYL, T>(u, i, j, k, inv_dr_sq, ioff);
  if (a.helm != 0) return v - (T)a.helm * u(i, j, k);
  return v;
}

template <int D, bool CYL, typename T, typename U>
__device__ __forceinline__ T op_u(const Dev& a, int id, int nc, U u, int i, int j, int k,
                                  T inv_dr_sq, int ioff) {
  if (a.tag != nullptr && a.tag[id] > kTagNormal)
    return lpld_u<D, CYL, T>(boxp(a, a.i_eps, id), nc, u, i, j, k, inv_dr_sq, ioff);
  return lplh_u<D, CYL, T>(a, u, i, j, k, inv_dr_sq, ioff);
}

template <int D, bool CYL, typename T>
__device__ __forceinline__ T op_at(const Dev& a, int id, const T* p, int nc, int i, int j, int k,
                                   T inv_dr_sq, int ioff) {
  return op_u<D, CYL, T>(a, id, nc, [&](int x, int y, int z) { return p[cidx<D>(nc, x, y, z)]; },
                         i, j, k, inv_dr_sq, ioff);
}

// One Gauss-Seidel update with a permittivity: mg$D_box_gsrb_lpld
// (m_aX_mg.f90:895-948), mg2_box_gsrb_clpld (:1499-1531). u[m], am[m]:
// potential and eps of the neighbours in the order i-1, i+1, j-1, j+1, k-1,
// k+1.
template <int D, bool CYL, typename T>
__device__ __forceinline__ T gsrb_lpld(const T* u, const T* am, T a0, T dx2, T rhs, T rf1,
                                       T rf2) {
  T c[2 * D];
#pragma unroll
  for (int m = 0; m < 2 * D; m++) c[m] = T(2) * a0 * am[m] / (a0 + am[m]);
  T s1 = 0, s2 = 0;
  if (CYL) {
    const T rf[4] = {rf1, rf2, T(1), T(1)};
#pragma unroll
    for (int m = 0; m < 2 * D; m++) s1 = s1 + c[m] * rf[m] * u[m];
#pragma unroll
    for (int m = 0; m < 2 * D; m++) s2 = s2 + c[m] * rf[m];
  } else {
#pragma unroll
    for (int m = 0; m < 2 * D; m++) s1 = s1 + c[m] * u[m];
#pragma unroll
    for (int m = 0; m < 2 * D; m++) s2 = s2 + c[m];
  }
  return (s1 - dx2 * rhs) / s2;
}

// ---------------------------------------------------------------------------
// Gauss-Seidel red-black half-sweep
// mg$D_box_gsrb_lpl (m_aX_mg.f90:769-811), mg2_box_gsrb_clpl (:1413-1439)
// ---------------------------------------------------------------------------

template <int D, bool CYL, typename T>
__global__ __launch_bounds__(256) void k_gsrb(Dev a, const int* __restrict__ list, int n,
                                              int nc, int cntr, int fused, int tpb,
                                              int i_phi, int i_rhs) {
  BOX_PROLOGUE
  T* p = boxp_t<T>(a, i_phi, id);
  const T* r = boxp_t<T>(a, i_rhs, id);
  const T dx2 = (T)(a.dr[id] * a.dr[id]);
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
    auto nbv = [&](int di, int dj, int dk, int nb) -> T {
      const int ii = i + di, jj = j + dj, kk = k + dk;
      const bool ghost = (ii < 1 || ii > nc || jj < 1 || jj > nc ||
                          (D == 3 && (kk < 1 || kk > nc)));
      if (ghost && fused) return ghost_mg<D, T>(a, p, i_phi, id, nc, nb, ii, jj, kk);
      return p[cidx<D>(nc, ii, jj, kk)];
    };
    const int c0 = cidx<D>(nc, i, j, k);
    if (a.tag != nullptr && a.tag[id] > kTagNormal) {
      const double* ev = boxp(a, a.i_eps, id);
      T uu[6], am[6];
      uu[0] = nbv(-1, 0, 0, 1); uu[1] = nbv(1, 0, 0, 2);
      uu[2] = nbv(0, -1, 0, 3); uu[3] = nbv(0, 1, 0, 4);
      am[0] = (T)ev[cidx<D>(nc, i - 1, j, k)]; am[1] = (T)ev[cidx<D>(nc, i + 1, j, k)];
      am[2] = (T)ev[cidx<D>(nc, i, j - 1, k)]; am[3] = (T)ev[cidx<D>(nc, i, j + 1, k)];
      if (D == 3) {
        uu[4] = nbv(0, 0, -1, 5); uu[5] = nbv(0, 0, 1, 6);
        am[4] = (T)ev[cidx<D>(nc, i, j, k - 1)]; am[5] = (T)ev[cidx<D>(nc, i, j, k + 1)];
      }
      T rf1 = T(1), rf2 = T(1);
      if (CYL) {
        const double den = i + ioff - 0.5;
        rf1 = (T)((double)(i + ioff - 1) / den);
        rf2 = (T)((double)(i + ioff) / den);
      }
      p[c0] = gsrb_lpld<D, CYL, T>(uu, am, (T)ev[c0], dx2, r[c0], rf1, rf2);
    } else {
      T num;
      if (D == 2) {
        if (CYL) {
          const double den = i + ioff - 0.5;
          const T rf1 = (T)((double)(i + ioff - 1) / den);
          const T rf2 = (T)((double)(i + ioff) / den);
          num = rf1 * nbv(-1, 0, 0, 1) + rf2 * nbv(1, 0, 0, 2) + nbv(0, 1, 0, 4) +
                nbv(0, -1, 0, 3) - dx2 * r[c0];
        } else {
          num = nbv(1, 0, 0, 2) + nbv(-1, 0, 0, 1) + nbv(0, 1, 0, 4) + nbv(0, -1, 0, 3) -
                dx2 * r[c0];
        }
      } else {
        num = nbv(1, 0, 0, 2) + nbv(-1, 0, 0, 1) + nbv(0, 1, 0, 4) + nbv(0, -1, 0, 3) +
              nbv(0, 0, 1, 6) + nbv(0, 0, -1, 5) - dx2 * r[c0];
      }
      if (a.helm != 0) p[c0] = num / (T(2 * D) + (T)a.helm * dx2);
      else p[c0] = (D == 2 ? T(0.25) : T(1) / T(6)) * num;
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

// xcd_block of workgroup slot v of a one-shot grid of nblk workgroups, for
// a persistent grid whose size is a multiple of 8 (slot v then runs on the
// XCD that workgroup v of the one-shot grid would run on)
__device__ __forceinline__ int xcd_slot(int v, int nblk) {
  const int per = nblk >> 3;
  return v < (per << 3) ? (v & 7) * per + (v >> 3) : v;
}

// Box-size-specialised half-sweep. The other colour's interior cells and
// face ghosts are staged in LDS, in an order used only there: slot
// lex(i, j, k) / 2 of the whole (NC+2)^D block, one colour, in which the
// neighbours of a cell are fixed offsets. The ghosts are the stored ones on
// the first half-sweep of a call; later half-sweeps recompute them from the
// current state (the values the reference's ghost pass after the previous
// half-sweep would store). Then every lane updates cells of colour cntr&1
// from LDS and rhs (loaded into registers before the barrier).
//
// Launch shapes (the autotuner's menu, kGsrbShapes): HALVE gives each box
// half the lanes of the default (each lane then updates twice the cells) and
// twice the boxes per 256-lane workgroup; PERSIST launches at most
// persist_grid workgroups, each walking the list's workgroup slots with the
// grid's stride (slot v of the one-shot grid, so the XCD mapping and every
// value are those of the one-shot form).
template <int D, int NC, bool HALVE>
struct GsrbShape {
  static constexpr int NCOL = (D == 2 ? NC * NC : NC * NC * NC) / 2;
  static constexpr int TPB0 = NCOL > 128 ? 256 : (NCOL > 64 ? 128 : 64);
  static constexpr int TPB = HALVE ? TPB0 / 2 : TPB0;
  static constexpr int BPB = 256 / TPB;
};

template <int D, bool CYL, int NC, typename T, bool EPS, bool NT, bool HALVE = false,
          bool PERSIST = false>
__global__ __launch_bounds__(256) void k_gsrb_lds(Dev a, const int* __restrict__ list, int n,
                                                  int cntr, int fused, int i_phi, int i_rhs) {
  constexpr int N2 = NC + 2;
  constexpr int BL = D == 2 ? N2 * N2 : N2 * N2 * N2;
  constexpr int HALF = BL / 2;
  constexpr int NCOL = GsrbShape<D, NC, HALVE>::NCOL;
  constexpr int TPB = GsrbShape<D, NC, HALVE>::TPB;
  constexpr int BPB = GsrbShape<D, NC, HALVE>::BPB;
  constexpr int PER_FACE = D == 2 ? NC : NC * NC;
  constexpr int HNC = NC / 2;
  constexpr int NIT = (NCOL + TPB - 1) / TPB;
  constexpr int NGC = D * PER_FACE;  // face ghosts of one colour
  constexpr int NITG = (NGC + TPB - 1) / TPB;
  constexpr int GPF = PER_FACE / 2;
  __shared__ T sh[BPB * HALF];
  const int slot = threadIdx.x / TPB, lt = threadIdx.x % TPB;
  const int nblk = (n + BPB - 1) / BPB;
  const int c = cntr & 1, oc = c ^ 1;
  T* s = sh + slot * HALF;
  // the interior cell of colour col at slot q of its half; returns its
  // lexicographic index in the block
  auto cell_of = [&](int q, int col, int& i, int& j, int& k) {
    icell<D>(NC, q, col, i, j, k);
    return D == 2 ? i + N2 * j : i + N2 * (j + N2 * k);
  };
  for (int vb = PERSIST ? (int)blockIdx.x : 0;;) {
  const int li = (PERSIST ? xcd_slot(vb, nblk) : xcd_block()) * BPB + slot;
  const bool act = li < n;
  const int id = act ? list[li] : 0;
  T* p = act ? boxp_t<T>(a, i_phi, id) : nullptr;
  T rv[NIT];
  if (act) {
    // Every global load (interior cells of colour oc, rhs of colour c, the
    // ghost values or the neighbour cells they copy) is issued before any
    // LDS store; the interior and ghost slots are disjoint.
    T sv[NIT];
    int sa[NIT];
    const T* po = p + oc * HALF;
#pragma unroll
    for (int u = 0; u < NIT; u++) {
      const int q = lt + u * TPB;
      sa[u] = -1;
      if (q < NCOL) {
        int i, j, k;
        sa[u] = cell_of(q, oc, i, j, k) >> 1;
        sv[u] = po[q];
      }
    }
    const T* r = boxp_t<T>(a, i_rhs, id) + c * HALF;
#pragma unroll
    for (int u = 0; u < NIT; u++) {
      const int q = lt + u * TPB;
      if (q < NCOL) rv[u] = NT ? __builtin_nontemporal_load(r + q) : r[q];
    }
    T gv[NITG];
    int gs[NITG], gnb[NITG], gg[NITG][3];
#pragma unroll
    for (int u = 0; u < NITG; u++) {
      const int q = lt + u * TPB;
      gs[u] = -1;
      if (q < NGC) {
        const int nb = q / GPF + 1;
        const int t = q - (nb - 1) * GPF;
        const int d = d_nb_dim(nb);
        const int ta = d == 0 ? 1 : 0, tb = d == 2 ? 1 : 2;
        int g[3] = {0, 0, 0};
        g[d] = d_nb_low(nb) ? 0 : NC + 1;
        // ghosts of colour oc: along ta every other cell of each tb row
        const int y = D == 3 ? t / HNC + 1 : 0;
        const int w = D == 3 ? t % HNC : t;
        if (D == 3) g[tb] = y;
        g[ta] = 2 * w + 2 - ((oc + g[d] + y) & 1);
        gs[u] = (D == 2 ? g[0] + N2 * g[1] : g[0] + N2 * (g[1] + N2 * g[2])) >> 1;
        gnb[u] = 0;
        gg[u][0] = g[0]; gg[u][1] = g[1]; gg[u][2] = g[2];
        if (!fused) {
          // the first half-sweep of a call reads the stored ghost layer
          gv[u] = p[cidx<D>(NC, g[0], g[1], g[2])];
        } else {
          const int nb_id = a.nbs[id * 6 + nb - 1];
          if (nb_id > 0) {
            int sc[3] = {g[0], g[1], g[2]};
            sc[d] = d_nb_low(nb) ? NC : 1;
            gv[u] = boxp_t<T>(a, i_phi, nb_id)[cidx_in<D>(NC, sc[0], sc[1], sc[2])];
          } else {
            gnb[u] = nb;
          }
        }
      }
    }
#pragma unroll
    for (int u = 0; u < NIT; u++)
      if (sa[u] >= 0) s[sa[u]] = sv[u];
#pragma unroll
    for (int u = 0; u < NITG; u++) {
      if (gs[u] < 0) continue;
      // refinement-boundary and physical rules from the box's own cells
      if (gnb[u] > 0) gv[u] = ghost_mg<D, T>(a, p, i_phi, id, NC, gnb[u], gg[u][0], gg[u][1], gg[u][2]);
      s[gs[u]] = gv[u];
    }
  }
  __syncthreads();
  if (act) {
  T* pc = p + c * HALF;
  const T dx2 = (T)(a.dr[id] * a.dr[id]);
  const int ioff = CYL ? (a.ix[id * 3] - 1) * NC : 0;
  // boxes tagged constant- or variable-eps (mg$D_auto_gsrb,
  // m_aX_mg.f90:568-600); eps is read from HBM
  const double* ev =
      (EPS && a.tag != nullptr && a.tag[id] > kTagNormal) ? boxp(a, a.i_eps, id) : nullptr;
#pragma unroll
  for (int u = 0; u < NIT; u++) {
    const int q = lt + u * TPB;
    if (q >= NCOL) break;
    int i, j, k;
    const int lex = cell_of(q, c, i, j, k);
    const T xp = s[(lex + 1) >> 1], xm = s[(lex - 1) >> 1];
    const T yp = s[(lex + N2) >> 1], ym = s[(lex - N2) >> 1];
    T v;
    if (EPS && ev != nullptr) {
      T uu[6], am[6];
      uu[0] = xm; uu[1] = xp; uu[2] = ym; uu[3] = yp;
      am[0] = (T)ev[cidx<D>(NC, i - 1, j, k)]; am[1] = (T)ev[cidx<D>(NC, i + 1, j, k)];
      am[2] = (T)ev[cidx<D>(NC, i, j - 1, k)]; am[3] = (T)ev[cidx<D>(NC, i, j + 1, k)];
      if (D == 3) {
        uu[4] = s[(lex - N2 * N2) >> 1]; uu[5] = s[(lex + N2 * N2) >> 1];
        am[4] = (T)ev[cidx<D>(NC, i, j, k - 1)]; am[5] = (T)ev[cidx<D>(NC, i, j, k + 1)];
      }
      T rf1 = T(1), rf2 = T(1);
      if (CYL) {
        const double den = i + ioff - 0.5;
        rf1 = (T)((double)(i + ioff - 1) / den);
        rf2 = (T)((double)(i + ioff) / den);
      }
      v = gsrb_lpld<D, CYL, T>(uu, am, (T)ev[c * HALF + q], dx2, rv[u], rf1, rf2);
    } else {
      // (sum of the neighbours - h^2 rhs) / 2D, or with a Helmholtz term
      // divided by 2D + helm h^2
      T num;
      if (D == 2) {
        if (CYL) {
          const double den = i + ioff - 0.5;
          const T rf1 = (T)((double)(i + ioff - 1) / den);
          const T rf2 = (T)((double)(i + ioff) / den);
          num = rf1 * xm + rf2 * xp + yp + ym - dx2 * rv[u];
        } else {
          num = xp + xm + yp + ym - dx2 * rv[u];
        }
      } else {
        const T zp = s[(lex + N2 * N2) >> 1], zm = s[(lex - N2 * N2) >> 1];
        num = xp + xm + yp + ym + zp + zm - dx2 * rv[u];
      }
      if (a.helm != 0) v = num / (T(2 * D) + (T)a.helm * dx2);
      else v = (D == 2 ? T(0.25) : T(1) / T(6)) * num;
    }
    if (NT) __builtin_nontemporal_store(v, pc + q);
    else pc[q] = v;
  }
  }
  if (!PERSIST) break;
  vb += gridDim.x;
  if (vb >= nblk) break;
  __syncthreads();  // the next slot's loads overwrite the LDS
  }
}

// One half-sweep in launch shape `shape` (kShapes: bit 0 persistent grid,
// bit 1 half the lanes per box)
// ---------------------------------------------------------------------------
// A whole smoothing call (gsrb_boxes: 2 n_cycle half-sweeps, each followed by
// the ghost pass, and the final stored ghost pass when store) on a level
// small enough that its potential and rhs fit in the LDS of one workgroup:
// the coarse levels of a V-cycle, where a launch per half-sweep costs more
// than the work. The level is loaded once (the potential with its stored
// ghost layer, the interior of the rhs), every half-sweep and ghost pass
// runs in LDS with a barrier between them, and the interior (and the ghost
// layer when store) goes back at the end. Each value is computed from the
// same operands in the same order as by k_gsrb and k_gc, so the result is
// bit for bit theirs. For levels whose boxes have no refinement boundary
// and no permittivity tags (the host checks that: small_level_lds). The
// whole-row rule of a y face (2V - c on its corner ghosts too, from the x
// ghosts of the same pass, as k_gc writes them) is applied in the final
// stored pass.
// LDS: the potential, lexicographic per box ((nc+2)^D per box), then the
// rhs interiors (nc^D per box), then the neighbour slots (2 D per box, -1 a
// physical face).
// ---------------------------------------------------------------------------

template <int D, bool CYL, typename T>
__global__ __launch_bounds__(1024) void k_gsrb_small(Dev a, const int* __restrict__ list, int n,
                                                     int nc, int n_half, int first_fused,
                                                     int store, int i_phi, int i_rhs) {
  extern __shared__ __align__(16) unsigned char small_lds[];
  const int m = nc + 2;
  const int bl = D == 2 ? m * m : m * m * m;
  const int ni = D == 2 ? nc * nc : nc * nc * nc;
  const int nf = D == 2 ? nc : nc * nc;  // cells of a face
  T* sp = (T*)small_lds;
  T* sr = sp + (size_t)n * bl;
  int* snb = (int*)(sr + (size_t)n * ni);
  const int tid = threadIdx.x, nt = blockDim.x;
  auto L = [&](int i, int j, int k) { return D == 2 ? i + m * j : i + m * (j + m * k); };
  // load: the whole potential block, the rhs interior, the neighbour slots
  for (int q = tid; q < n * bl; q += nt) {
    const int b = q / bl, c = q - b * bl;
    const int i = c % m, j = (c / m) % m, k = D == 3 ? c / (m * m) : 0;
    sp[q] = boxp_t<T>(a, i_phi, list[b])[cidx<D>(nc, i, j, k)];
  }
  for (int q = tid; q < n * ni; q += nt) {
    const int b = q / ni, c = q - b * ni;
    const int i = c % nc + 1, j = (c / nc) % nc + 1, k = D == 3 ? c / (nc * nc) + 1 : 0;
    sr[q] = boxp_t<T>(a, i_rhs, list[b])[cidx_in<D>(nc, i, j, k)];
  }
  for (int q = tid; q < n * 2 * D; q += nt) {
    const int b = q / (2 * D), nb = q - b * 2 * D + 1;
    const int nb_id = a.nbs[list[b] * 6 + nb - 1];
    int slot = -1;
    if (nb_id > 0)
      for (int o = 0; o < n; o++)
        if (list[o] == nb_id) slot = o;
    snb[q] = slot;
  }
  __syncthreads();
  // the face ghosts of every box from the current state (k_gc's rules for
  // a box with same-level neighbours and physical faces)
  auto ghosts = [&]() {
    for (int q = tid; q < n * 2 * D * nf; q += nt) {
      const int b = q / (2 * D * nf), r = q - b * 2 * D * nf;
      const int nb = r / nf + 1, t = r - (nb - 1) * nf;
      const int d = d_nb_dim(nb);
      const bool low = d_nb_low(nb);
      const int ta = d == 0 ? 1 : 0, tb = d == 2 ? 1 : 2;
      int g[3] = {0, 0, 0};
      g[d] = low ? 0 : nc + 1;
      g[ta] = t % nc + 1;
      if (D == 3) g[tb] = t / nc + 1;
      const int ns = snb[b * 2 * D + nb - 1];
      T v;
      if (ns >= 0) {
        int sc[3] = {g[0], g[1], g[2]};
        sc[d] = low ? nc : 1;
        v = sp[(size_t)ns * bl + L(sc[0], sc[1], sc[2])];
      } else {
        int c[3] = {g[0], g[1], g[2]};
        c[d] = low ? 1 : nc;
        const T cv = sp[(size_t)b * bl + L(c[0], c[1], c[2])];
        const int kind = a.bc_kind[nb - 1];
        v = kind == 0 ? cv : (kind == 1 ? -cv : T(2) * (T)a.bc_val[nb - 1] - cv);
      }
      sp[(size_t)b * bl + L(g[0], g[1], g[2])] = v;
    }
  };
  const int half = nc >> 1;
  const int ncol = ni >> 1;
  for (int cntr = 1; cntr <= n_half; cntr++) {
    if (cntr > 1 || first_fused) {
      ghosts();
      __syncthreads();
    }
    for (int q = tid; q < n * ncol; q += nt) {
      const int b = q / ncol, w = q - b * ncol;
      const int id = list[b];
      const int mm = w % half, row = w / half;
      const int j = row % nc + 1;
      const int k = D == 3 ? row / nc + 1 : 0;
      const int i0 = 2 - ((cntr ^ (D == 3 ? k + j : j)) & 1);
      const int i = i0 + 2 * mm;
      const T* u = sp + (size_t)b * bl;
      auto nbv = [&](int di, int dj, int dk) { return u[L(i + di, j + dj, k + dk)]; };
      const T dx2 = (T)(a.dr[id] * a.dr[id]);
      const T rhs = sr[(size_t)b * ni + (i - 1) + nc * ((j - 1) + (D == 3 ? nc * (k - 1) : 0))];
      T num;
      if (D == 2) {
        if (CYL) {
          const int ioff = (a.ix[id * 3] - 1) * nc;
          const double den = i + ioff - 0.5;
          const T rf1 = (T)((double)(i + ioff - 1) / den);
          const T rf2 = (T)((double)(i + ioff) / den);
          num = rf1 * nbv(-1, 0, 0) + rf2 * nbv(1, 0, 0) + nbv(0, 1, 0) + nbv(0, -1, 0) -
                dx2 * rhs;
        } else {
          num = nbv(1, 0, 0) + nbv(-1, 0, 0) + nbv(0, 1, 0) + nbv(0, -1, 0) - dx2 * rhs;
        }
      } else {
        num = nbv(1, 0, 0) + nbv(-1, 0, 0) + nbv(0, 1, 0) + nbv(0, -1, 0) + nbv(0, 0, 1) +
              nbv(0, 0, -1) - dx2 * rhs;
      }
      T* pc = sp + (size_t)b * bl + L(i, j, k);
      if (a.helm != 0) *pc = num / (T(2 * D) + (T)a.helm * dx2);
      else *pc = (D == 2 ? T(0.25) : T(1) / T(6)) * num;
    }
    __syncthreads();
  }
  // the corner ghosts of the whole-row rule on a y face of box b
  auto row_corner = [&](int b, int nb) {
    return D == 2 && snb[b * 2 * D + nb - 1] < 0 && a.bc_kind[nb - 1] == kBcRow2V;
  };
  if (store) {
    ghosts();
    __syncthreads();
    if (D == 2) {
      for (int q = tid; q < n * 4; q += nt) {
        const int b = q >> 2, nb = 3 + ((q >> 1) & 1), side = q & 1;
        if (!row_corner(b, nb)) continue;
        const int gx = side ? nc + 1 : 0, gy = nb == 3 ? 0 : nc + 1, cy = nb == 3 ? 1 : nc;
        T* u = sp + (size_t)b * bl;
        u[L(gx, gy, 0)] = T(2) * (T)a.bc_val[nb - 1] - u[L(gx, cy, 0)];
      }
      __syncthreads();
    }
  }
  // back to HBM: the interior, and the ghosts the stored pass wrote
  for (int q = tid; q < n * bl; q += nt) {
    const int b = q / bl, c = q - b * bl;
    const int i = c % m, j = (c / m) % m, k = D == 3 ? c / (m * m) : 0;
    const int out = (i < 1 || i > nc) + (j < 1 || j > nc) + (D == 3 && (k < 1 || k > nc));
    const bool corner = D == 2 && store && out == 2 && row_corner(b, j == 0 ? 3 : 4);
    if (out == 0 || (store && out == 1) || corner)
      boxp_t<T>(a, i_phi, list[b])[cidx<D>(nc, i, j, k)] = sp[q];
  }
}

template <int D, bool CYL, int NC, typename T, bool EPS, bool NT>
static void go_gsrb(Pool& P, const Dev& a, const int* list, int n, int cntr, int fused,
                    int i_phi, int i_rhs, int shape) {
  const bool persist = shape & 1, halve = shape & 2;
  const int bpb = halve ? GsrbShape<D, NC, true>::BPB : GsrbShape<D, NC, false>::BPB;
  const int nblk = (n + bpb - 1) / bpb;
  const int grid = persist ? P.persist_grid(nblk) : nblk;
  auto kern = halve ? (persist ? k_gsrb_lds<D, CYL, NC, T, EPS, NT, true, true>
                               : k_gsrb_lds<D, CYL, NC, T, EPS, NT, true, false>)
                    : (persist ? k_gsrb_lds<D, CYL, NC, T, EPS, NT, false, true>
                               : k_gsrb_lds<D, CYL, NC, T, EPS, NT, false, false>);
  hipLaunchKernelGGL(kern, dim3(grid), dim3(256), 0, P.stream, a, list, n, cntr, fused, i_phi,
                     i_rhs);
}

template <int D, bool CYL, typename T, bool EPS>
static bool launch_gsrb_lds_t(Pool& P, const Dev& a, const int* list, int n, int nc, int cntr,
                              int fused, int i_phi, int i_rhs) {
  const int sh = P.shape(0, D, nc, n);
  // A level whose potential and rhs far exceed the 256 MB Infinity Cache
  // streams: the rhs is read and the new colour written with non-temporal
  // accesses, which keeps L2 for the neighbour faces the next boxes read.
  // Smaller levels stay cached from one half-sweep to the next.
  const double box_cells = D == 2 ? (nc + 2.0) * (nc + 2) : (nc + 2.0) * (nc + 2) * (nc + 2);
  const double level_bytes = 2.0 * n * box_cells * sizeof(T);
  const bool nt = level_bytes > P.stream_bytes;
  switch (nc) {
    case 2: go_gsrb<D, CYL, 2, T, EPS, false>(P, a, list, n, cntr, fused, i_phi, i_rhs, sh); return true;
    case 4: go_gsrb<D, CYL, 4, T, EPS, false>(P, a, list, n, cntr, fused, i_phi, i_rhs, sh); return true;
    case 8: go_gsrb<D, CYL, 8, T, EPS, false>(P, a, list, n, cntr, fused, i_phi, i_rhs, sh); return true;
    case 16:
      if (nt) go_gsrb<D, CYL, 16, T, EPS, true>(P, a, list, n, cntr, fused, i_phi, i_rhs, sh);
      else go_gsrb<D, CYL, 16, T, EPS, false>(P, a, list, n, cntr, fused, i_phi, i_rhs, sh);
      return true;
    case 32:
      if (D == 2) {
        if (nt) go_gsrb<2, CYL, 32, T, EPS, true>(P, a, list, n, cntr, fused, i_phi, i_rhs, sh);
        else go_gsrb<2, CYL, 32, T, EPS, false>(P, a, list, n, cntr, fused, i_phi, i_rhs, sh);
        return true;
      }
      return false;
    default: return false;
  }
}

// The permittivity branch is compiled in only when the operator tags are in
// use, so the common kernel keeps its registers for loads in flight.
template <int D, bool CYL, typename T>
static bool launch_gsrb_lds(Pool& P, const Dev& a, const int* list, int n, int nc, int cntr,
                            int fused, int i_phi, int i_rhs) {
  if (a.tag != nullptr)
    return launch_gsrb_lds_t<D, CYL, T, true>(P, a, list, n, nc, cntr, fused, i_phi, i_rhs);
  return launch_gsrb_lds_t<D, CYL, T, false>(P, a, list, n, nc, cntr, fused, i_phi, i_rhs);
}

// ---------------------------------------------------------------------------
// Face ghost cells of one variable: a$D_gc_box (m_aX_gc.f90:50-68) with
// rb = 0 mg$D_sides_rb, 1 a$D_gc_interp (:90-185), 2 a$D_gc_interp_lim
// (:190-295); bc = 0 multigrid face rules, 1 Neumann (:298-330),
// 2 Dirichlet zero (:333-365).
// ---------------------------------------------------------------------------

// A ghost value of k_gc for the faces without a same-level neighbour
// (nb_id <= 0): refinement boundary or physical boundary.
template <int D, typename T>
__device__ __forceinline__ T gc_boundary(const Dev& a, const T* p, int iv, int id, int nc, int nb,
                                         const int g[3], int rb, int bc) {
  const int nb_id = a.nbs[id * 6 + nb - 1];
  const int d = d_nb_dim(nb);
  const bool low = d_nb_low(nb);
  const int ta = d == 0 ? 1 : 0, tb = d == 2 ? 1 : 2;
  const T sixth = T(1) / T(6), third = T(1) / T(3);
  if ((nb_id == 0 && rb == 0) || (nb_id < 0 && bc == 0))
    return ghost_mg<D, T>(a, p, iv, id, nc, nb, g[0], g[1], g[2]);
  if (nb_id < 0) {
    int c[3] = {g[0], g[1], g[2]};
    c[d] = low ? 1 : nc;
    const T cv = p[cidx_in<D>(nc, c[0], c[1], c[2])];
    return bc == 1 ? cv : -cv;
  }
  // coarse-neighbour interpolation
  const int pid = a.parent[id];
  const int pn = a.nbs[pid * 6 + nb - 1];
  const int pnc = a.ncell[pn];
  int off[3];
  child_off<D>(a, id, nc, nb, off);
  const int ixc = low ? nc : 1, ixf = low ? 1 : nc;
  const T* qv = boxp_t<T>(a, iv, pn);
  int f[3] = {g[0], g[1], g[2]};
  f[d] = ixf;
  const T fv = p[cidx_in<D>(nc, f[0], f[1], f[2])];
  T v;
  if (D == 2) {
    const int x = g[ta];
    const int c1 = off[ta] + ((x + 1) >> 1), c2 = c1 + 1 - 2 * (x & 1);
    int s1[3], s2[3];
    s1[d] = s2[d] = ixc;
    s1[ta] = c1;
    s2[ta] = c2;
    const T v1 = qv[cidx<D>(pnc, s1[0], s1[1], 0)];
    const T v2 = qv[cidx<D>(pnc, s2[0], s2[1], 0)];
    v = T(0.5) * v1 + sixth * v2 + third * fv;
    if (rb == 2 && v > T(2) * v1) v = T(2) * v1;
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
    const T v1 = qv[cidx<D>(pnc, s1[0], s1[1], s1[2])];
    const T v2 = qv[cidx<D>(pnc, s2[0], s2[1], s2[2])];
    const T v3 = qv[cidx<D>(pnc, s3[0], s3[1], s3[2])];
    v = third * v1 + sixth * v2 + sixth * v3 + third * fv;
    if (rb == 2 && v > T(2) * v1) v = T(2) * v1;
  }
  return v;
}

// Every ghost read here is an interior cell (of the box, a neighbour, the
// parent or its neighbour) and every write a ghost slot of the box, so the
// neighbour copies of four ghosts per lane are loaded before any store.
template <int D, typename T>
__global__ __launch_bounds__(256) void k_gc(Dev a, const int* __restrict__ list, int n,
                                            int nc, int tpb, int iv, int rb, int bc) {
  BOX_PROLOGUE
  T* p = boxp_t<T>(a, iv, id);
  const int per_face = D == 2 ? nc : nc * nc;
  const int total = 2 * D * per_face;
  constexpr int B = 4;
  for (int q0 = lt; q0 < total; q0 += B * tpb) {
    T v[B];
    int x[B], gg[B][3], bnb[B];
#pragma unroll
    for (int u = 0; u < B; u++) {
      const int q = q0 + u * tpb;
      x[u] = -1;
      bnb[u] = 0;
      if (q < total) {
        const int nb = q / per_face + 1;
        const int t = q - (nb - 1) * per_face;
        const int d = d_nb_dim(nb);
        const bool low = d_nb_low(nb);
        const int ta = d == 0 ? 1 : 0, tb = d == 2 ? 1 : 2;
        int g[3] = {0, 0, 0};
        g[d] = low ? 0 : nc + 1;
        g[ta] = t % nc + 1;
        if (D == 3) g[tb] = t / nc + 1;
        x[u] = cidx<D>(nc, g[0], g[1], g[2]);
        gg[u][0] = g[0]; gg[u][1] = g[1]; gg[u][2] = g[2];
        const int nb_id = a.nbs[id * 6 + nb - 1];
        if (nb_id > 0) {
          // sides_from_nb (m_aX_gc.f90:368-402)
          int s[3] = {g[0], g[1], g[2]};
          s[d] = low ? nc : 1;
          v[u] = boxp_t<T>(a, iv, nb_id)[cidx_in<D>(nc, s[0], s[1], s[2])];
        } else {
          bnb[u] = nb;
        }
      }
    }
#pragma unroll
    for (int u = 0; u < B; u++) {
      if (x[u] < 0) continue;
      if (bnb[u] > 0) v[u] = gc_boundary<D, T>(a, p, iv, id, nc, bnb[u], gg[u], rb, bc);
      p[x[u]] = v[u];
    }
  }
  // The rule 2V - c over a whole y ghost row (the cylindrical driver's
  // sides_bc_pot, src/streamer_cyl.f90:408-410, cc(:, nc+1) = 2V - cc(:, nc))
  // also writes the two corner ghosts, from the x ghosts that the same
  // a2_gc_box call has just filled (faces go in the order lx, hx, ly, hy).
  // Lanes 0, 1 take the corners of face ly, lanes 2, 3 those of hy; each
  // recomputes its x ghost from the interior cells it is made of.
  if (D == 2 && bc == 0 && lt < 4) {
    const int nb = 3 + (lt >> 1);
    if (a.nbs[id * 6 + nb - 1] < 0 && a.bc_kind[nb - 1] == kBcRow2V) {
      const int side = lt & 1;
      const int xnb = 1 + side;
      const int g[3] = {side ? nc + 1 : 0, nb == 3 ? 1 : nc, 0};
      const int xid = a.nbs[id * 6 + xnb - 1];
      T xg;
      if (xid > 0) {
        xg = boxp_t<T>(a, iv, xid)[cidx_in<D>(nc, side ? 1 : nc, g[1], 0)];
      } else {
        xg = gc_boundary<D, T>(a, p, iv, id, nc, xnb, g, rb, bc);
      }
      p[cidx<D>(nc, g[0], nb == 3 ? 0 : nc + 1, 0)] = T(2) * (T)face_value<D>(a, id, nb) - xg;
    }
  }
}

// ---------------------------------------------------------------------------
// Residual + restriction to the parent (update_coarse / set_coarse_phi_rhs,
// m_aX_mg.f90:430-523, with residual_box :551-565 and mg$D_box_rstr_lpl
// :851-893). One lane per parent cell covered by the child: the 2^D fine
// residuals stay in registers. keep_res = 1 stores them in the child's tmp
// (set_coarse_phi_rhs); 0 leaves tmp as it was (update_coarse restores it).
// ---------------------------------------------------------------------------

template <int D, bool CYL, typename T>
__global__ __launch_bounds__(256) void k_resid_rstr(Dev a, const int* __restrict__ list,
                                                    int n, int nc, int tpb, int i_phi,
                                                    int i_rhs, int i_tmp, int keep_res) {
  BOX_PROLOGUE
  const int hnc = nc >> 1;
  const int pid = a.parent[id];
  const int pnc = a.ncell[pid];
  int off[3];
  child_off<D>(a, id, nc, 0, off);
  const T* p = boxp_t<T>(a, i_phi, id);
  const T* r = boxp_t<T>(a, i_rhs, id);
  T* t = boxp_t<T>(a, i_tmp, id);
  T* pp = boxp_t<T>(a, i_phi, pid);
  T* pt = boxp_t<T>(a, i_tmp, pid);
  const T inv_dr_sq = (T)(1 / (a.dr[id] * a.dr[id]));
  const int ioff = CYL ? (a.ix[id * 3] - 1) * nc : 0;
  const int total = D == 2 ? hnc * hnc : hnc * hnc * hnc;
  for (int q = lt; q < total; q += tpb) {
    const int ci = q % hnc + 1;
    const int cj = (q / hnc) % hnc + 1;
    const int ck = D == 3 ? q / (hnc * hnc) + 1 : 0;
    const int fi = 2 * ci - 1, fj = 2 * cj - 1, fk = D == 3 ? 2 * ck - 1 : 0;
    T sr = 0, sp = 0;
#pragma unroll
    for (int dk = 0; dk < (D == 3 ? 2 : 1); dk++)
#pragma unroll
      for (int dj = 0; dj < 2; dj++)
#pragma unroll
        for (int di = 0; di < 2; di++) {
          const int ii = fi + di, jj = fj + dj, kk = fk + dk;
          const int c = cidx<D>(nc, ii, jj, kk);
          const T res = r[c] - op_at<D, CYL, T>(a, id, p, nc, ii, jj, kk, inv_dr_sq, ioff);
          if (keep_res) t[c] = res;
          sr = sr + res;
          sp = sp + p[c];
        }
    const int pc = cidx<D>(pnc, off[0] + ci, off[1] + cj, D == 3 ? off[2] + ck : 0);
    const T w = D == 2 ? T(0.25) : T(0.125);
    pt[pc] = w * sr;
    pp[pc] = w * sp;
  }
}

// The same pass with the child's potential staged in LDS, in lexicographic
// order of the whole (NC+2)^D block, and its face ghosts recomputed there
// from the current state (the values the ghost pass that ends the
// reference's smoothing, fill_gc_phi after gsrb_boxes, would store: nothing
// changes the child, its neighbours or the uncovered parent cells in
// between), so that pass is not needed before this one.
// HALVE (a launch shape of the autotuner's menu): half the lanes per box,
// twice the boxes per workgroup.
// SPLIT > 1 (3D): each box is shared by SPLIT workgroups, slab h taking the
// fine planes k = h NC/SPLIT + 1 .. (h+1) NC/SPLIT and staging those planes
// with one more on either side (the z neighbours of its first and last
// plane) and the face ghosts they read. A slab holds whole pairs of planes
// (odd, even), so the refinement-boundary extrapolation of its x and y face
// ghosts reads planes inside it. The staging then takes (NC/SPLIT + 2)
// planes instead of NC + 2: for 16^3 FP64 boxes 26 KB of LDS instead of
// 47 KB, twice the workgroups per compute unit. Every value is the same.
template <int D, bool CYL, int NC, typename T, bool EPS, bool HALVE = false, int SPLIT = 1>
__global__ __launch_bounds__(256) void k_resid_rstr_lds(Dev a, const int* __restrict__ list,
                                                        int n, int i_phi, int i_rhs, int i_tmp,
                                                        int keep_res) {
  static_assert(SPLIT == 1 || (D == 3 && (NC / 2) % SPLIT == 0), "z slabs of whole plane pairs");
  constexpr int N2 = NC + 2;
  constexpr int HALF = (D == 2 ? N2 * N2 : N2 * N2 * N2) / 2;
  constexpr int HNC = NC / 2;
  constexpr int KS = NC / SPLIT;  // fine planes per slab
  constexpr int BLS = D == 2 ? N2 * N2 : N2 * N2 * (SPLIT == 1 ? N2 : KS + 2);
  constexpr int NCOL = (D == 2 ? NC * NC : NC * NC * NC) / 2;
  constexpr int TOTAL = (D == 2 ? HNC * HNC : HNC * HNC * HNC) / SPLIT;
  constexpr int TPB0 = TOTAL > 128 ? 256 : (TOTAL > 64 ? 128 : 64);
  constexpr int TPB = HALVE ? TPB0 / 2 : TPB0;
  constexpr int BPB = 256 / TPB;
  constexpr int PER_FACE = D == 2 ? NC : NC * NC;
  constexpr int NGH = 2 * D * PER_FACE;
  __shared__ T sh[BPB * BLS];
  const int slot = threadIdx.x / TPB, lt = threadIdx.x % TPB;
  const int lw = xcd_block() * BPB + slot;  // (box, slab) work item
  const bool act = lw < n * SPLIT;
  const int li = lw / SPLIT, h = lw - li * SPLIT;
  const int id = act ? list[li] : 0;
  // fine planes of the slab, and the first staged plane
  const int kf0 = SPLIT == 1 ? 1 : h * KS + 1, kf1 = SPLIT == 1 ? NC : kf0 + KS - 1;
  const int kbase = SPLIT == 1 ? 0 : kf0 - 1;
  constexpr int NCELL = 2 * NCOL;
  constexpr int NITC = (NCELL + TPB - 1) / TPB;
  constexpr int NITG = (NGH + TPB - 1) / TPB;
  T* s = sh + slot * BLS;
  // The tile in LDS: a whole box (SPLIT 1) split by parity into 2^D blocks
  // of (N2/2)^D, cell (i, j, k) in block (i&1, j&1, k&1) at (i>>1, j>>1,
  // k>>1). A lane restricts to one parent cell, so neighbouring lanes take
  // fine cells two apart: in lexicographic order their stencil reads were
  // two doubles apart and conflicted in the LDS banks (3.9 stall cycles per
  // active LDS cycle); here they are adjacent. Slabs (SPLIT > 1) keep the
  // lexicographic order.
  constexpr int H2 = N2 / 2;
  constexpr int PBLK = D == 2 ? H2 * H2 : H2 * H2 * H2;
  auto L = [&](int i, int j, int k) {
    if (SPLIT == 1) {
      const int sub = (i & 1) + 2 * (j & 1) + (D == 3 ? 4 * (k & 1) : 0);
      return sub * PBLK + (i >> 1) + H2 * ((j >> 1) + (D == 3 ? H2 * (k >> 1) : 0));
    }
    return D == 2 ? i + N2 * j : i + N2 * (j + N2 * (k - kbase));
  };
  auto own = [&](int i, int j, int k) { return s[L(i, j, k)]; };
  // whether a face ghost is read by this slab
  auto ghost_needed = [&](int nb, int gk) {
    if (SPLIT == 1) return true;
    const int d = d_nb_dim(nb);
    if (d < 2) return gk >= kf0 && gk <= kf1;
    return d_nb_low(nb) ? kf0 == 1 : kf1 == NC;
  };
  // Face ghosts that copy a neighbour's cell are loaded first, then the
  // interior cells (each colour a contiguous run in HBM) in batches; all
  // these LDS slots are disjoint. Ghosts of the other rules read the staged
  // interior, after a barrier.
  int gq[NITG];
  // rhs of the 2^D fine cells of each parent cell a lane restricts, loaded
  // with the staging loads so that the compute phase reads only LDS
  constexpr int NITP = (TOTAL + TPB - 1) / TPB;
  constexpr int NF = D == 2 ? 4 : 8;
  T rq[NITP][NF];
  const int ck0 = SPLIT == 1 ? 0 : h * (HNC / SPLIT);  // parent planes before the slab
  auto fine = [&](int q, int f, int& ii, int& jj, int& kk) {
    const int ci = q % HNC + 1, cj = (q / HNC) % HNC + 1;
    const int ck = D == 3 ? q / (HNC * HNC) + 1 + ck0 : 0;
    ii = 2 * ci - 1 + (f & 1);
    jj = 2 * cj - 1 + ((f >> 1) & 1);
    kk = D == 3 ? 2 * ck - 1 + (f >> 2) : 0;
  };
  if (act) {
    const T* p = boxp_t<T>(a, i_phi, id);
    const T* r = boxp_t<T>(a, i_rhs, id);
#pragma unroll
    for (int u = 0; u < NITP; u++) {
      const int q = lt + u * TPB;
      if (q < TOTAL) {
#pragma unroll
        for (int f = 0; f < NF; f++) {
          int ii, jj, kk;
          fine(q, f, ii, jj, kk);
          rq[u][f] = r[cidx_in<D>(NC, ii, jj, kk)];
        }
      }
    }
    T gv[NITG];
    int gs[NITG];
#pragma unroll
    for (int u = 0; u < NITG; u++) {
      const int q = lt + u * TPB;
      gs[u] = -1;
      gq[u] = -1;
      if (q < NGH) {
        const int nb = q / PER_FACE + 1;
        const int t = q - (nb - 1) * PER_FACE;
        const int d = d_nb_dim(nb);
        const int ta = d == 0 ? 1 : 0, tb = d == 2 ? 1 : 2;
        int g[3] = {0, 0, 0};
        g[d] = d_nb_low(nb) ? 0 : NC + 1;
        g[ta] = t % NC + 1;
        if (D == 3) g[tb] = t / NC + 1;
        if (ghost_needed(nb, g[2])) {
          const int nb_id = a.nbs[id * 6 + nb - 1];
          if (nb_id > 0) {
            int sc[3] = {g[0], g[1], g[2]};
            sc[d] = d_nb_low(nb) ? NC : 1;
            gv[u] = boxp_t<T>(a, i_phi, nb_id)[cidx_in<D>(NC, sc[0], sc[1], sc[2])];
            gs[u] = L(g[0], g[1], g[2]);
          } else {
            gq[u] = q;
          }
        }
      }
    }
    constexpr int B = NITC < 8 ? NITC : 8;
#pragma unroll
    for (int i0 = 0; i0 < NITC; i0 += B) {
      T v[B];
      int at[B];
#pragma unroll
      for (int u = 0; u < B; u++) {
        const int e = lt + (i0 + u) * TPB;
        at[u] = -1;
        if (i0 + u < NITC && e < NCELL) {
          // the interior cell of colour col at slot q of its half
          const int col = e >= NCOL, q = e - col * NCOL;
          int i, j, k;
          icell<D>(NC, q, col, i, j, k);
          if (SPLIT == 1 || (k >= kf0 - 1 && k <= kf1 + 1)) {
            at[u] = L(i, j, k);
            v[u] = p[col * HALF + q];
          }
        }
      }
#pragma unroll
      for (int u = 0; u < B; u++)
        if (at[u] >= 0) s[at[u]] = v[u];
    }
#pragma unroll
    for (int u = 0; u < NITG; u++)
      if (gs[u] >= 0) s[gs[u]] = gv[u];
  }
  __syncthreads();
  if (act) {
#pragma unroll
    for (int u = 0; u < NITG; u++) {
      const int q = gq[u];
      if (q < 0) continue;
      const int nb = q / PER_FACE + 1;
      const int t = q - (nb - 1) * PER_FACE;
      const int d = d_nb_dim(nb);
      const int ta = d == 0 ? 1 : 0, tb = d == 2 ? 1 : 2;
      int g[3] = {0, 0, 0};
      g[d] = d_nb_low(nb) ? 0 : NC + 1;
      g[ta] = t % NC + 1;
      if (D == 3) g[tb] = t / NC + 1;
      // reads interior cells of s only, writes a ghost slot
      s[L(g[0], g[1], g[2])] = ghost_mg_u<D, T>(a, own, i_phi, id, NC, nb, g[0], g[1], g[2]);
    }
  }
  __syncthreads();
  if (!act) return;
  const int pid = a.parent[id];
  const int pnc = a.ncell[pid];
  int off[3];
  child_off<D>(a, id, NC, 0, off);
  T* tv = boxp_t<T>(a, i_tmp, id);
  T* pp = boxp_t<T>(a, i_phi, pid);
  T* pt = boxp_t<T>(a, i_tmp, pid);
  const T inv_dr_sq = (T)(1 / (a.dr[id] * a.dr[id]));
  const int ioff = CYL ? (a.ix[id * 3] - 1) * NC : 0;
#pragma unroll
  for (int u = 0; u < NITP; u++) {
    const int q = lt + u * TPB;
    if (q >= TOTAL) break;
    const int ci = q % HNC + 1;
    const int cj = (q / HNC) % HNC + 1;
    const int ck = D == 3 ? q / (HNC * HNC) + 1 + ck0 : 0;
    // Fortran's sum over cc(i_f:i_f+1, j_f:j_f+1[, k_f:k_f+1]) runs in array
    // order, i fastest (m_aX_mg.f90:872-890)
    T sr = 0, sp = 0;
#pragma unroll
    for (int f = 0; f < NF; f++) {
      int ii, jj, kk;
      fine(q, f, ii, jj, kk);
      const T res = rq[u][f] - (EPS ? op_u<D, CYL, T>(a, id, NC, own, ii, jj, kk, inv_dr_sq, ioff)
                                    : lplh_u<D, CYL, T>(a, own, ii, jj, kk, inv_dr_sq, ioff));
      if (keep_res) tv[cidx_in<D>(NC, ii, jj, kk)] = res;
      sr = sr + res;
      sp = sp + own(ii, jj, kk);
    }
    const int pc = cidx_in<D>(pnc, off[0] + ci, off[1] + cj, D == 3 ? off[2] + ck : 0);
    const T w = D == 2 ? T(0.25) : T(0.125);
    pt[pc] = w * sr;
    pp[pc] = w * sp;
  }
}

// Plain restriction of one variable to the parent (mg$D_box_rstr_lpl), and
// for init_phi_rhs (m_aX_mg.f90:526-549) the clearing of the child's phi.
template <int D, typename T>
__global__ __launch_bounds__(256) void k_rstr_clear(Dev a, const int* __restrict__ list,
                                                    int n, int nc, int tpb, int iv,
                                                    int i_clear) {
  BOX_PROLOGUE
  const int hnc = nc >> 1;
  const int pid = a.parent[id];
  const int pnc = a.ncell[pid];
  int off[3];
  child_off<D>(a, id, nc, 0, off);
  const T* c = boxp_t<T>(a, iv, id);
  T* pv = boxp_t<T>(a, iv, pid);
  const int total = D == 2 ? hnc * hnc : hnc * hnc * hnc;
  for (int q = lt; q < total; q += tpb) {
    const int ci = q % hnc + 1;
    const int cj = (q / hnc) % hnc + 1;
    const int ck = D == 3 ? q / (hnc * hnc) + 1 : 0;
    const int fi = 2 * ci - 1, fj = 2 * cj - 1, fk = D == 3 ? 2 * ck - 1 : 0;
    T s = 0;
#pragma unroll
    for (int dk = 0; dk < (D == 3 ? 2 : 1); dk++)
#pragma unroll
      for (int dj = 0; dj < 2; dj++)
#pragma unroll
        for (int di = 0; di < 2; di++) s = s + c[cidx<D>(nc, fi + di, fj + dj, fk + dk)];
    pv[cidx<D>(pnc, off[0] + ci, off[1] + cj, D == 3 ? off[2] + ck : 0)] =
        (D == 2 ? T(0.25) : T(0.125)) * s;
  }
  if (i_clear > 0) {
    // runs after all reads of this box (other lanes read `iv`, not i_clear)
    T* z = boxp_t<T>(a, i_clear, id);
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
                                                      int n, int nc, int tpb, int iv,
                                                      bool plain) {
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
    if (D == 2 && a.coord == 1 && !plain) {
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

template <int D, bool CYL, typename T>
__global__ __launch_bounds__(256) void k_coarse_rhs(Dev a, const int* __restrict__ list,
                                                    int n, int nc, int tpb, int i_phi,
                                                    int i_rhs, int i_tmp, int copy_tmp) {
  BOX_PROLOGUE
  const T* p = boxp_t<T>(a, i_phi, id);
  T* r = boxp_t<T>(a, i_rhs, id);
  T* t = boxp_t<T>(a, i_tmp, id);
  const T inv_dr_sq = (T)(1 / (a.dr[id] * a.dr[id]));
  const int ioff = CYL ? (a.ix[id * 3] - 1) * nc : 0;
  const int n2 = nc + 2;
  const int bl = D == 2 ? n2 * n2 : n2 * n2 * n2;
  for (int q = lt; q < bl; q += tpb) {
    const int i = q % n2, j = (q / n2) % n2, k = D == 3 ? q / (n2 * n2) : 0;
    const int x = cidx<D>(nc, i, j, k);
    const bool inner = i >= 1 && i <= nc && j >= 1 && j <= nc &&
                       (D == 2 || (k >= 1 && k <= nc));
    T rv = inner ? op_at<D, CYL, T>(a, id, p, nc, i, j, k, inv_dr_sq, ioff) : r[x];
    rv = rv + t[x];
    r[x] = rv;
    if (copy_tmp) t[x] = p[x];
  }
}

// The same pass with phi staged in LDS (lexicographic order, interior and
// face ghosts) for the operator; the interior and the ghost slots of each
// colour are visited as contiguous runs. fused = 1: the face ghosts of phi
// are recomputed from the current state (neighbour copy, refinement-boundary
// extrapolation, physical rule) instead of read, so that no ghost pass has
// to store them after the restriction; with copy_tmp they also go to tmp's
// face ghost slots, as the stored values would have. Every global load of a
// box (its potential, tmp of its interior, rhs and tmp of its ghost slots,
// the face ghosts or the neighbour cells they copy) is issued before any
// store: on a level of a few boxes the kernel's time is one workgroup's
// chain of memory latencies, which loads taken one loop iteration at a time
// (each behind a store to rhs or tmp) would lengthen many times over.
template <int D, bool CYL, int NC, typename T, bool EPS, bool HALVE = false>
__global__ __launch_bounds__(256) void k_coarse_rhs_lds(Dev a, const int* __restrict__ list,
                                                        int n, int i_phi, int i_rhs, int i_tmp,
                                                        int copy_tmp, int fused) {
  constexpr int N2 = NC + 2;
  constexpr int BL = D == 2 ? N2 * N2 : N2 * N2 * N2;
  constexpr int HALF = BL / 2;
  constexpr int NCOL = (D == 2 ? NC * NC : NC * NC * NC) / 2;
  constexpr int NGC = HALF - NCOL;  // ghost slots of one colour
  constexpr int PER_FACE = D == 2 ? NC : NC * NC;
  constexpr int NGH = 2 * D * PER_FACE;
  constexpr int TPB0 = BL > 128 ? 256 : (BL > 64 ? 128 : 64);
  constexpr int TPB = HALVE ? TPB0 / 2 : TPB0;
  constexpr int BPB = 256 / TPB;
  constexpr int NITC = (2 * NCOL + TPB - 1) / TPB;  // interior cells per lane
  constexpr int NITH = (NGH + TPB - 1) / TPB;       // face ghosts per lane
  constexpr int NITX = (2 * NGC + TPB - 1) / TPB;   // ghost slots per lane
  __shared__ T sh[BPB * BL];
  const int slot = threadIdx.x / TPB, lt = threadIdx.x % TPB;
  const int li = xcd_block() * BPB + slot;
  const bool act = li < n;
  const int id = act ? list[li] : 0;
  T* s = sh + slot * BL;
  auto L = [](int i, int j, int k) { return D == 2 ? i + N2 * j : i + N2 * (j + N2 * k); };
  auto own = [&](int i, int j, int k) { return s[L(i, j, k)]; };
  auto cell = [](int e, int& i, int& j, int& k) {
    const int col = e >= NCOL, q = e - col * NCOL;
    icell<D>(NC, q, col, i, j, k);
    return col * HALF + q;
  };
  auto face = [](int q, int g[3]) {
    const int nb = q / PER_FACE + 1;
    const int t = q - (nb - 1) * PER_FACE;
    const int d = d_nb_dim(nb);
    const int ta = d == 0 ? 1 : 0, tb = d == 2 ? 1 : 2;
    g[0] = g[1] = g[2] = 0;
    g[d] = d_nb_low(nb) ? 0 : NC + 1;
    g[ta] = t % NC + 1;
    if (D == 3) g[tb] = t / NC + 1;
    return nb;
  };
  auto gslot = [](int e) {
    const int col = e >= NGC;
    return col * HALF + NCOL + (e - col * NGC);
  };
  const T* p = act ? boxp_t<T>(a, i_phi, id) : nullptr;
  T* r = act ? boxp_t<T>(a, i_rhs, id) : nullptr;
  T* t = act ? boxp_t<T>(a, i_tmp, id) : nullptr;
  T pv[NITC], tv[NITC], gv[NITH], rg[NITX], tg[NITX], pg[NITX];
  int gnb[NITH];
  if (act) {
#pragma unroll
    for (int u = 0; u < NITC; u++) {
      const int e = lt + u * TPB;
      if (e < 2 * NCOL) {
        int i, j, k;
        const int x = cell(e, i, j, k);
        pv[u] = p[x];
        tv[u] = t[x];
      }
    }
#pragma unroll
    for (int u = 0; u < NITH; u++) {
      const int q = lt + u * TPB;
      gnb[u] = 0;
      if (q < NGH) {
        int g[3];
        const int nb = face(q, g);
        if (!fused) {
          gv[u] = p[cidx<D>(NC, g[0], g[1], g[2])];
        } else {
          const int nb_id = a.nbs[id * 6 + nb - 1];
          if (nb_id > 0) {
            const int d = d_nb_dim(nb);
            int sc[3] = {g[0], g[1], g[2]};
            sc[d] = d_nb_low(nb) ? NC : 1;
            gv[u] = boxp_t<T>(a, i_phi, nb_id)[cidx_in<D>(NC, sc[0], sc[1], sc[2])];
          } else {
            gnb[u] = nb;
          }
        }
      }
    }
#pragma unroll
    for (int u = 0; u < NITX; u++) {
      const int e = lt + u * TPB;
      if (e < 2 * NGC) {
        const int x = gslot(e);
        rg[u] = r[x];
        tg[u] = t[x];
        if (copy_tmp) pg[u] = p[x];
      }
    }
#pragma unroll
    for (int u = 0; u < NITC; u++) {
      const int e = lt + u * TPB;
      if (e < 2 * NCOL) {
        int i, j, k;
        cell(e, i, j, k);
        s[L(i, j, k)] = pv[u];
      }
    }
#pragma unroll
    for (int u = 0; u < NITH; u++) {
      const int q = lt + u * TPB;
      if (q < NGH) {
        int g[3];
        face(q, g);
        // refinement-boundary and physical rules from the box's own cells
        // (read from HBM: the LDS copy is not complete before the barrier)
        if (gnb[u] > 0) gv[u] = ghost_mg<D, T>(a, p, i_phi, id, NC, gnb[u], g[0], g[1], g[2]);
        s[L(g[0], g[1], g[2])] = gv[u];
      }
    }
  }
  __syncthreads();
  if (act) {
    const T inv_dr_sq = (T)(1 / (a.dr[id] * a.dr[id]));
    const int ioff = CYL ? (a.ix[id * 3] - 1) * NC : 0;
#pragma unroll
    for (int u = 0; u < NITC; u++) {
      const int e = lt + u * TPB;
      if (e >= 2 * NCOL) break;
      int i, j, k;
      const int x = cell(e, i, j, k);
      T rv = EPS ? op_u<D, CYL, T>(a, id, NC, own, i, j, k, inv_dr_sq, ioff)
                 : lplh_u<D, CYL, T>(a, own, i, j, k, inv_dr_sq, ioff);
      rv = rv + tv[u];
      r[x] = rv;
      if (copy_tmp) t[x] = pv[u];
    }
#pragma unroll
    for (int u = 0; u < NITX; u++) {
      const int e = lt + u * TPB;
      if (e >= 2 * NGC) break;
      const int x = gslot(e);
      r[x] = rg[u] + tg[u];
      if (copy_tmp) t[x] = pg[u];
    }
  }
  if (!(copy_tmp && fused)) return;
  // tmp's face ghosts take the recomputed values, after the loop above has
  // copied the stored (stale) ones (possibly from another lane)
  __syncthreads();
  if (act) {
#pragma unroll
    for (int u = 0; u < NITH; u++) {
      const int q = lt + u * TPB;
      if (q < NGH) {
        int g[3];
        face(q, g);
        t[cidx<D>(NC, g[0], g[1], g[2])] = gv[u];
      }
    }
    // the corner ghosts the rule 2V - c over a whole y ghost row writes
    // (k_gc), from the x ghosts recomputed above
    if (D == 2 && lt < 4) {
      const int nb = 3 + (lt >> 1), side = lt & 1;
      if (a.nbs[id * 6 + nb - 1] < 0 && a.bc_kind[nb - 1] == kBcRow2V) {
        const int gx = side ? NC + 1 : 0, cy = nb == 3 ? 1 : NC;
        t[cidx<D>(NC, gx, nb == 3 ? 0 : NC + 1, 0)] =
            T(2) * (T)face_value<D>(a, id, nb) - s[L(gx, cy, 0)];
      }
    }
  }
}

template <int D, bool CYL, typename T, bool EPS>
static bool launch_coarse_rhs_lds_t(Pool& P, const Dev& a, const int* list, int n, int nc,
                                    int i_phi, int i_rhs, int i_tmp, int copy_tmp, int fused) {
  const bool halve = nc <= 8 && (P.shape(kShapeCoarseRhs, D, nc, n) & 2);
  auto go = [&](auto kern, int bl) {
    const int tpb = (bl > 128 ? 256 : (bl > 64 ? 128 : 64)) >> (halve ? 1 : 0);
    const int bpb = 256 / tpb;
    hipLaunchKernelGGL(kern, dim3((n + bpb - 1) / bpb), dim3(256), 0, P.stream, a, list, n,
                       i_phi, i_rhs, i_tmp, copy_tmp, fused);
  };
  const int bl = D == 2 ? (nc + 2) * (nc + 2) : (nc + 2) * (nc + 2) * (nc + 2);
  switch (nc) {
    case 4:
      if (halve) go(k_coarse_rhs_lds<D, CYL, 4, T, EPS, true>, bl);
      else go(k_coarse_rhs_lds<D, CYL, 4, T, EPS>, bl);
      return true;
    case 8:
      if (halve) go(k_coarse_rhs_lds<D, CYL, 8, T, EPS, true>, bl);
      else go(k_coarse_rhs_lds<D, CYL, 8, T, EPS>, bl);
      return true;
    case 16: go(k_coarse_rhs_lds<D, CYL, 16, T, EPS>, bl); return true;
    case 32:
      if (D == 2) {
        go(k_coarse_rhs_lds<2, CYL, 32, T, EPS>, bl);
        return true;
      }
      return false;
    default: return false;
  }
}

// correct_children, first half (m_aX_mg.f90:385-392): tmp = phi - tmp on the
// whole parent box
template <typename T>
__global__ __launch_bounds__(256) void k_sub_whole(Dev a, const int* __restrict__ list,
                                                   int n, int bl, int tpb, int i_phi,
                                                   int i_tmp) {
  BOX_PROLOGUE
  const T* p = boxp_t<T>(a, i_phi, id);
  T* t = boxp_t<T>(a, i_tmp, id);
  for (int q = lt; q < bl; q += tpb) t[q] = p[q] - t[q];
}

// correct_children, second half: mg$D_box_corr_lpl (m_aX_mg.f90:714-766) on
// every child box in the list
template <int D, typename T>
__global__ __launch_bounds__(256) void k_corr(Dev a, const int* __restrict__ list, int n,
                                              int nc, int tpb, int i_phi, int i_tmp) {
  BOX_PROLOGUE
  const int pid = a.parent[id];
  const int pnc = a.ncell[pid];
  int off[3];
  child_off<D>(a, id, nc, 0, off);
  T* f = boxp_t<T>(a, i_phi, id);
  const T* c = boxp_t<T>(a, i_tmp, pid);
  const double* ep = (a.tag != nullptr && a.tag[id] > kTagNormal) ? boxp(a, a.i_eps, pid) : nullptr;
  const int total = D == 2 ? nc * nc : nc * nc * nc;
  for (int q = lt; q < total; q += tpb) {
    const int i = q % nc + 1, j = (q / nc) % nc + 1, k = D == 3 ? q / (nc * nc) + 1 : 0;
    const int i1 = off[0] + ((i + 1) >> 1), i2 = i1 + 1 - 2 * (i & 1);
    const int j1 = off[1] + ((j + 1) >> 1), j2 = j1 + 1 - 2 * (j & 1);
    const int x = cidx<D>(nc, i, j, k);
    if (ep != nullptr) {
      // mg$D_box_corr_lpld, m_aX_mg.f90:1003-1071 (parent's eps)
      if (D == 2) {
        const int q0 = cidx<D>(pnc, i1, j1, 0);
        const int qa[2] = {cidx<D>(pnc, i2, j1, 0), cidx<D>(pnc, i1, j2, 0)};
        const T u0 = c[q0], a0 = (T)ep[q0];
        T sm = 0;
#pragma unroll
        for (int m = 0; m < 2; m++) {
          const T am = (T)ep[qa[m]];
          sm = sm + (a0 * u0 + am * c[qa[m]]) / (a0 + am);
        }
        f[x] = f[x] + T(0.5) * sm;
      } else {
        const int k1 = off[2] + ((k + 1) >> 1), k2 = k1 + 1 - 2 * (k & 1);
        const int q0 = cidx<D>(pnc, i1, j1, k1);
        const int qa[3] = {cidx<D>(pnc, i2, j1, k1), cidx<D>(pnc, i1, j2, k1),
                           cidx<D>(pnc, i1, j1, k2)};
        const T u0 = c[q0], a0 = (T)ep[q0];
        T sm = 0;
#pragma unroll
        for (int m = 0; m < 3; m++) {
          const T am = (T)ep[qa[m]];
          sm = sm + (a0 * u0 + am * (T(1.5) * c[qa[m]] - T(0.5) * u0)) / (a0 + am);
        }
        f[x] = f[x] + (T(1) / T(3)) * sm;
      }
    } else if (D == 2) {
      f[x] = f[x] + T(0.5) * c[cidx<D>(pnc, i1, j1, 0)] +
             T(0.25) * (c[cidx<D>(pnc, i2, j1, 0)] + c[cidx<D>(pnc, i1, j2, 0)]);
    } else {
      const int k1 = off[2] + ((k + 1) >> 1), k2 = k1 + 1 - 2 * (k & 1);
      f[x] = f[x] + T(0.25) * (c[cidx<D>(pnc, i1, j1, k1)] + c[cidx<D>(pnc, i2, j1, k1)] +
                               c[cidx<D>(pnc, i1, j2, k1)] + c[cidx<D>(pnc, i1, j1, k2)]);
    }
  }
}

// The same correction with the parent's region under the child staged in
// LDS: (NC/2 + 2)^D values of the parent's tmp, ghosts included, read once
// per child box instead of once per child cell.
template <int D, int NC, typename T, bool HALVE = false>
__global__ __launch_bounds__(256) void k_corr_lds(Dev a, const int* __restrict__ list, int n,
                                                  int i_phi, int i_tmp) {
  constexpr int R = NC / 2 + 2;
  constexpr int RB = D == 2 ? R * R : R * R * R;
  constexpr int TOTAL = D == 2 ? NC * NC : NC * NC * NC;
  constexpr int TPB0 = TOTAL > 128 ? 256 : (TOTAL > 64 ? 128 : 64);
  constexpr int TPB = HALVE ? TPB0 / 2 : TPB0;
  constexpr int BPB = 256 / TPB;
  constexpr int NIT = (TOTAL + TPB - 1) / TPB;
  constexpr int NCOL = TOTAL / 2;
  constexpr int HALF = (D == 2 ? (NC + 2) * (NC + 2) : (NC + 2) * (NC + 2) * (NC + 2)) / 2;
  __shared__ T sh[BPB * RB];
  const int slot = threadIdx.x / TPB, lt = threadIdx.x % TPB;
  const int li = xcd_block() * BPB + slot;
  const bool act = li < n;
  const int id = act ? list[li] : 0;
  T* q = sh + slot * RB;
  int off[3] = {0, 0, 0};
  int pid = 0, pnc = 0;
  // child cells in slot order (each colour's interior a contiguous run),
  // except that the x-boundary cells go in row order: neighbouring lanes
  // then read parent values from different LDS banks. Their potential is
  // loaded with the staging loads.
  constexpr int NROW = D == 2 ? NC : NC * NC;
  auto cell = [&](int t, int& i, int& j, int& k) {
    const int col = t >= NCOL;
    int qs = t - col * NCOL;
    if (qs < NROW) {
      j = qs % NC + 1;
      k = D == 3 ? qs / NC + 1 : 0;
      i = ((1 + j + k) & 1) == col ? 1 : NC;
      qs = isl<D>(NC, i, j, k);
    } else {
      icell<D>(NC, qs, col, i, j, k);
    }
    return col * HALF + qs;
  };
  T fv[NIT];
  T* f = act ? boxp_t<T>(a, i_phi, id) : nullptr;
  if (act) {
    pid = a.parent[id];
    pnc = a.ncell[pid];
    child_off<D>(a, id, NC, 0, off);
#pragma unroll
    for (int u = 0; u < NIT; u++) {
      const int t = lt + u * TPB;
      int i, j, k;
      if (t < TOTAL) fv[u] = f[cell(t, i, j, k)];
    }
    const T* c = boxp_t<T>(a, i_tmp, pid);
    for (int t = lt; t < RB; t += TPB) {
      const int x = t % R, y = (t / R) % R, z = D == 3 ? t / (R * R) : 0;
      q[t] = c[cidx<D>(pnc, off[0] + x, off[1] + y, D == 3 ? off[2] + z : 0)];
    }
  }
  __syncthreads();
  if (!act) return;
  const double* ep = (a.tag != nullptr && a.tag[id] > kTagNormal) ? boxp(a, a.i_eps, pid) : nullptr;
  auto L = [&](int x, int y, int z) { return D == 2 ? x + R * y : x + R * (y + R * z); };
#pragma unroll
  for (int u = 0; u < NIT; u++) {
    const int t = lt + u * TPB;
    if (t >= TOTAL) break;
    int i, j, k;
    const int x = cell(t, i, j, k);
    // local parent indices: (i+1)/2 is the closest coarse cell, the other
    // one lies on the side of the fine cell
    const int i1 = (i + 1) >> 1, i2 = i1 + 1 - 2 * (i & 1);
    const int j1 = (j + 1) >> 1, j2 = j1 + 1 - 2 * (j & 1);
    if (ep != nullptr) {
      if (D == 2) {
        const int p0 = cidx<D>(pnc, off[0] + i1, off[1] + j1, 0);
        const int pa[2] = {cidx<D>(pnc, off[0] + i2, off[1] + j1, 0),
                           cidx<D>(pnc, off[0] + i1, off[1] + j2, 0)};
        const T ua[2] = {q[L(i2, j1, 0)], q[L(i1, j2, 0)]};
        const T u0 = q[L(i1, j1, 0)], a0 = (T)ep[p0];
        T sm = 0;
#pragma unroll
        for (int m = 0; m < 2; m++) {
          const T am = (T)ep[pa[m]];
          sm = sm + (a0 * u0 + am * ua[m]) / (a0 + am);
        }
        f[x] = fv[u] + T(0.5) * sm;
      } else {
        const int k1 = (k + 1) >> 1, k2 = k1 + 1 - 2 * (k & 1);
        const int p0 = cidx<D>(pnc, off[0] + i1, off[1] + j1, off[2] + k1);
        const int pa[3] = {cidx<D>(pnc, off[0] + i2, off[1] + j1, off[2] + k1),
                           cidx<D>(pnc, off[0] + i1, off[1] + j2, off[2] + k1),
                           cidx<D>(pnc, off[0] + i1, off[1] + j1, off[2] + k2)};
        const T ua[3] = {q[L(i2, j1, k1)], q[L(i1, j2, k1)], q[L(i1, j1, k2)]};
        const T u0 = q[L(i1, j1, k1)], a0 = (T)ep[p0];
        T sm = 0;
#pragma unroll
        for (int m = 0; m < 3; m++) {
          const T am = (T)ep[pa[m]];
          sm = sm + (a0 * u0 + am * (T(1.5) * ua[m] - T(0.5) * u0)) / (a0 + am);
        }
        f[x] = fv[u] + (T(1) / T(3)) * sm;
      }
    } else if (D == 2) {
      f[x] = fv[u] + T(0.5) * q[L(i1, j1, 0)] + T(0.25) * (q[L(i2, j1, 0)] + q[L(i1, j2, 0)]);
    } else {
      const int k1 = (k + 1) >> 1, k2 = k1 + 1 - 2 * (k & 1);
      f[x] = fv[u] + T(0.25) * (q[L(i1, j1, k1)] + q[L(i2, j1, k1)] + q[L(i1, j2, k1)] +
                               q[L(i1, j1, k2)]);
    }
  }
}

template <int D, typename T>
static bool launch_corr_lds(Pool& P, const Dev& a, const int* list, int n, int nc, int i_phi,
                            int i_tmp) {
  const bool halve = nc <= 8 && (P.shape(kShapeCorr, D, nc, n) & 2);
  auto go = [&](auto kern, int total) {
    const int tpb = (total > 128 ? 256 : (total > 64 ? 128 : 64)) >> (halve ? 1 : 0);
    const int bpb = 256 / tpb;
    hipLaunchKernelGGL(kern, dim3((n + bpb - 1) / bpb), dim3(256), 0, P.stream, a, list, n,
                       i_phi, i_tmp);
  };
  const int total = D == 2 ? nc * nc : nc * nc * nc;
  switch (nc) {
    case 8:
      if (halve) go(k_corr_lds<D, 8, T, true>, total);
      else go(k_corr_lds<D, 8, T>, total);
      return true;
    case 16: go(k_corr_lds<D, 16, T>, total); return true;
    case 32:
      if (D == 2) {
        go(k_corr_lds<2, 32, T>, total);
        return true;
      }
      return false;
    default: return false;
  }
}

// whole-box copy (a$D_box_copy_cc, m_aX_utils.f90:421-429)
template <typename T>
__global__ __launch_bounds__(256) void k_copy_whole(Dev a, const int* __restrict__ list,
                                                    int n, int bl, int tpb, int from,
                                                    int to) {
  BOX_PROLOGUE
  const T* s = boxp_t<T>(a, from, id);
  T* d = boxp_t<T>(a, to, id);
  for (int q = lt; q < bl; q += tpb) d[q] = s[q];
}

// residual_box (m_aX_mg.f90:551-565): tmp = rhs - A phi on the interior
template <int D, bool CYL, typename T>
__global__ __launch_bounds__(256) void k_residual(Dev a, const int* __restrict__ list,
                                                  int n, int nc, int tpb, int i_phi,
                                                  int i_rhs, int i_tmp) {
  BOX_PROLOGUE
  const T* p = boxp_t<T>(a, i_phi, id);
  const T* r = boxp_t<T>(a, i_rhs, id);
  T* t = boxp_t<T>(a, i_tmp, id);
  const T inv_dr_sq = (T)(1 / (a.dr[id] * a.dr[id]));
  const int ioff = CYL ? (a.ix[id * 3] - 1) * nc : 0;
  const int total = D == 2 ? nc * nc : nc * nc * nc;
  for (int q = lt; q < total; q += tpb) {
    const int i = q % nc + 1, j = (q / nc) % nc + 1, k = D == 3 ? q / (nc * nc) + 1 : 0;
    const int x = cidx<D>(nc, i, j, k);
    t[x] = r[x] - op_at<D, CYL, T>(a, id, p, nc, i, j, k, inv_dr_sq, ioff);
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
    const int sl = 2 * (blockIdx.x % kRedSlots);
    atomicMin(&out[sl], ord_key(lo));
    atomicMax(&out[sl + 1], ord_key(hi));
  }
}

// max over leaf interiors of |rhs - A phi| (resid) or of |rhs|, into
// out[blockIdx.x % kRedSlots] as ordered-integer keys. The residual is the
// expression of k_residual, so the maximum equals that of the stored
// residual. Used by the convergence checks of solves that run side by side,
// each with slots of its own.
template <int D, bool CYL>
__global__ __launch_bounds__(256) void k_leaf_absmax(Dev a, const int* __restrict__ list, int n,
                                                     int nc, int tpb, int i_phi, int i_rhs,
                                                     int resid, unsigned long long* out) {
  const int bpb_ = blockDim.x / tpb;
  const int slot_ = threadIdx.x / tpb;
  const int lt = threadIdx.x - slot_ * tpb;
  const int li_ = blockIdx.x * bpb_ + slot_;
  double hi = 0.0;
  if (li_ < n) {
    const int id = list[li_];
    const double* p = boxp(a, i_phi, id);
    const double* r = boxp(a, i_rhs, id);
    const double inv_dr_sq = 1 / (a.dr[id] * a.dr[id]);
    const int ioff = CYL ? (a.ix[id * 3] - 1) * nc : 0;
    const int total = D == 2 ? nc * nc : nc * nc * nc;
    for (int q = lt; q < total; q += tpb) {
      const int i = q % nc + 1, j = (q / nc) % nc + 1, k = D == 3 ? q / (nc * nc) + 1 : 0;
      const int x = cidx<D>(nc, i, j, k);
      const double v =
          resid ? r[x] - op_at<D, CYL, double>(a, id, p, nc, i, j, k, inv_dr_sq, ioff) : r[x];
      hi = fmax(hi, fabs(v));
    }
  }
  hi = wave_max(hi);
  __shared__ double wh[4];
  if ((threadIdx.x & 63) == 0) wh[threadIdx.x >> 6] = hi;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < (int)(blockDim.x >> 6); w++) hi = fmax(hi, wh[w]);
    atomicMax(&out[blockIdx.x % kRedSlots], ord_key(hi));
  }
}

__global__ void k_fill_u64(unsigned long long* p, int n, unsigned long long v) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) p[i] = v;
}

// Range of one variable over whole boxes (ghosts included): out[2 li] = min,
// out[2 li + 1] = max for box list[li]. One workgroup per box.
template <int D>
__global__ __launch_bounds__(256) void k_box_range(Dev a, const int* __restrict__ list, int n,
                                                   int iv, double* out) {
  const int li = blockIdx.x;
  if (li >= n) return;
  const int id = list[li];
  const double* p = boxp(a, iv, id);
  const int m = a.ncell[id] + 2;
  const int total = D == 2 ? m * m : m * m * m;
  double lo = p[0], hi = p[0];
  for (int q = threadIdx.x; q < total; q += blockDim.x) {
    lo = fmin(lo, p[q]);
    hi = fmax(hi, p[q]);
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
    out[2 * li] = lo;
    out[2 * li + 1] = hi;
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

// One multigrid problem: its variables, its precision and whether the
// physical boundary values are replaced by zero (the correction equation of
// the mixed-precision solve, whose boundary data is homogeneous).
struct Ctx {
  int i_phi, i_tmp, i_rhs;
  bool homogeneous;
  bool cyl;  // cylindrical operators (2D, automatic operators installed)
  bool eps;  // per-box operator tags (automatic operators and an eps variable)
  int n_down, n_up, n_base;  // smoothing cycles (mg$D_t%n_cycle_*)
  double helm;               // Helmholtz coefficient (0: Laplacian)
  // The finest level of a V-cycle over the whole tree leaves its ghost
  // cells to the next reader (defer_top), and the call after it starts from
  // such a level (top_stale); see flush_ghosts.
  bool defer_top = false;
  bool top_stale = false;
};

static Ctx ctx64(const Pool& P) {
  return {P.mg.i_phi, P.mg.i_tmp, P.mg.i_rhs, false,
          P.tree.dim == 2 && P.tree.coord == 1 && P.mg.auto_ops != 0,
          P.mg.auto_ops != 0 && P.mg.i_eps > 0, P.mg.n_down, P.mg.n_up, P.mg.n_base,
          P.mg.helm};
}

static Dev dev_for(Pool& P, const Ctx& c) {
  Dev a = P.dev();
  if (c.homogeneous) {
    for (int i = 0; i < 6; i++) a.bc_val[i] = 0.0;
    a.n_inst = 0;  // the correction problem has zero boundary values everywhere
  }
  if (c.eps) {
    a.tag = P.d_tag;
    a.i_eps = P.mg.i_eps;
  }
  a.helm = c.helm;
  return a;
}

static int cells_of(int D, int nc, bool ghosts) {
  const int m = ghosts ? nc + 2 : nc;
  return D == 2 ? m * m : m * m * m;
}

template <int D, typename T>
static void gc_list_t(Pool& P, const Dev& a, int lvl, const int* list, int n, int iv, int rb,
                      int bc) {
  if (n == 0) return;
  const int nc = P.nc_lvl(lvl);
  Geo g = geo(n, 2 * D * (D == 2 ? nc : nc * nc));
  if ((P.shape(kShapeGc, D, nc, n) & 2) && g.tpb > 16) {
    // half the lanes per box, twice the boxes per workgroup
    g.tpb /= 2;
    g.grid = (n + 256 / g.tpb - 1) / (256 / g.tpb);
  }
  const int pr = P.prof_begin(PK_GC, lvl, 0.0);
  hipLaunchKernelGGL((k_gc<D, T>), dim3(g.grid), dim3(256), 0, P.stream, a, list, n, nc, g.tpb,
                     iv, rb, bc);
  post_launch("gc");
  P.prof_end(pr);
}

void gc_level(Pool& P, int lvl, int iv, int rb, int bc) {
  gc_list(P, lvl, P.ids(lvl), P.n_ids(lvl), iv, rb, bc);
}

void gc_list(Pool& P, int lvl, const int* list, int n, int iv, int rb, int bc) {
  Dev a = P.dev();
  if (P.tree.dim == 2) gc_list_t<2, double>(P, a, lvl, list, n, iv, rb, bc);
  else gc_list_t<3, double>(P, a, lvl, list, n, iv, rb, bc);
}

template <int D, typename T>
static void fill_gc_phi(Pool& P, const Ctx& c, int lvl) {
  // fill_gc_phi, m_aX_mg.f90:358-371
  gc_list_t<D, T>(P, dev_for(P, c), lvl, P.ids(lvl), P.n_ids(lvl), c.i_phi, 0, 0);
}

// gsrb_boxes, m_aX_mg.f90:403-426: 2 n_cycle half-sweeps, each followed by
// the ghost pass. Ghosts are recomputed inside every half-sweep after the
// first; the first reads the stored ghosts unless first_fused (the caller
// skipped a ghost pass that would have stored exactly what is recomputed).
// store = false leaves the stored ghost layer as it was, for callers whose
// next reader of these ghosts recomputes them (the residual pass).
// Whether a smoothing call on lvl can run in one workgroup's LDS
// (k_gsrb_small): its boxes' potential and rhs fit, none has a refinement
// boundary or a permittivity tag, no run of instances, and the face rules
// are copy, -c and 2V - c. Returns the LDS bytes (0: no).
template <typename T>
static size_t small_level_lds(const Pool& P, const Ctx& c, const Dev& a, int lvl, int n, int nc) {
  if (!P.small_on || a.n_inst) return 0;
  const int D = P.tree.dim;
  const size_t bytes = (size_t)n * (cells_of(D, nc, true) + cells_of(D, nc, false)) * sizeof(T) +
                       (size_t)n * 2 * D * sizeof(int);
  if (bytes > (size_t)std::min<long>(P.max_lds, P.small_lds)) return 0;
  for (int id : P.lids(lvl)) {
    if (c.eps && P.tree.tag[id] > kTagNormal) return 0;
    for (int nb = 0; nb < 2 * D; nb++)
      if (P.tree.nbs[id * 6 + nb] == 0) return 0;
  }
  return bytes;
}

template <int D, typename T>
static void gsrb_boxes(Pool& P, const Ctx& c, int lvl, int n_cycle, bool first_fused = false,
                       bool store = true) {
  const int n = P.n_ids(lvl);
  if (n == 0) return;
  const int nc = P.nc_lvl(lvl);
  const int ncol = cells_of(D, nc, false) / 2;
  Geo g = geo(n, ncol);
  Dev a = dev_for(P, c);
  const double cells = (double)n * cells_of(D, nc, false);
  // one event pair around the consecutive half-sweeps (an event between two
  // launches holds the second back by a few microseconds); algorithmic bytes
  // of a half-sweep: 1.5 words per cell
  const int pr = n_cycle > 0
                     ? P.prof_begin(PK_GSRB, lvl, 2 * n_cycle * 1.5 * sizeof(T) * cells, 2 * n_cycle)
                     : -1;
  if (n_cycle > 0) {
    const size_t lds = small_level_lds<T>(P, c, a, lvl, n, nc);
    if (lds > 0) {
      const int work = std::max(n * ncol, n * 2 * D * (D == 2 ? nc : nc * nc));
      const int nt = std::min(1024, std::max(64, pow2ceil(work)));
      // more than 64 KB of dynamic LDS is granted per kernel on request
      // (once per instantiation, before any capture can hold the launch)
      auto allow = [&](const void* fn) {
        static std::vector<const void*> done;
        if (lds > 65536 && std::find(done.begin(), done.end(), fn) == done.end()) {
          check(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, P.max_lds),
                "LDS of the small-level smoother");
          done.push_back(fn);
        }
      };
      allow(c.cyl ? (const void*)k_gsrb_small<2, true, T> : (const void*)k_gsrb_small<D, false, T>);
      if (c.cyl)
        hipLaunchKernelGGL((k_gsrb_small<2, true, T>), dim3(1), dim3(nt), lds, P.stream, a,
                           P.ids(lvl), n, nc, 2 * n_cycle, first_fused ? 1 : 0, store ? 1 : 0,
                           c.i_phi, c.i_rhs);
      else
        hipLaunchKernelGGL((k_gsrb_small<D, false, T>), dim3(1), dim3(nt), lds, P.stream, a,
                           P.ids(lvl), n, nc, 2 * n_cycle, first_fused ? 1 : 0, store ? 1 : 0,
                           c.i_phi, c.i_rhs);
      post_launch("gsrb_small");
      P.prof_end(pr);
      return;
    }
  }
  for (int it = 1; it <= 2 * n_cycle; it++) {
    const int fused = it > 1 || first_fused;
    const bool done =
        c.cyl ? launch_gsrb_lds<2, true, T>(P, a, P.ids(lvl), n, nc, it, fused, c.i_phi, c.i_rhs)
              : launch_gsrb_lds<D, false, T>(P, a, P.ids(lvl), n, nc, it, fused, c.i_phi,
                                             c.i_rhs);
    if (done) {
    } else if (c.cyl)
      hipLaunchKernelGGL((k_gsrb<2, true, T>), dim3(g.grid), dim3(256), 0, P.stream, a,
                         P.ids(lvl), n, nc, it, fused, g.tpb, c.i_phi, c.i_rhs);
    else
      hipLaunchKernelGGL((k_gsrb<D, false, T>), dim3(g.grid), dim3(256), 0, P.stream, a,
                         P.ids(lvl), n, nc, it, fused, g.tpb, c.i_phi, c.i_rhs);
    post_launch("gsrb");
  }
  P.prof_end(pr);
  if (store || n_cycle == 0) fill_gc_phi<D, T>(P, c, lvl);
}

template <int D, bool CYL, typename T, bool EPS>
static bool launch_resid_rstr_lds_t(Pool& P, const Dev& a, const int* list, int n, int nc,
                                    int i_phi, int i_rhs, int i_tmp, int keep_res) {
  const bool halve = nc <= 8 && (P.shape(kShapeResidRstr, D, nc, n) & 2);
  auto go = [&](auto kern, int total) {
    const int tpb = (total > 128 ? 256 : (total > 64 ? 128 : 64)) >> (halve ? 1 : 0);
    const int bpb = 256 / tpb;
    hipLaunchKernelGGL(kern, dim3((n + bpb - 1) / bpb), dim3(256), 0, P.stream, a, list, n,
                       i_phi, i_rhs, i_tmp, keep_res);
  };
  const int hnc = nc / 2;
  const int total = D == 2 ? hnc * hnc : hnc * hnc * hnc;
  switch (nc) {
    case 2:
      if (halve) go(k_resid_rstr_lds<D, CYL, 2, T, EPS, true>, total);
      else go(k_resid_rstr_lds<D, CYL, 2, T, EPS>, total);
      return true;
    case 4:
      if (halve) go(k_resid_rstr_lds<D, CYL, 4, T, EPS, true>, total);
      else go(k_resid_rstr_lds<D, CYL, 4, T, EPS>, total);
      return true;
    case 8:
      if (halve) go(k_resid_rstr_lds<D, CYL, 8, T, EPS, true>, total);
      else go(k_resid_rstr_lds<D, CYL, 8, T, EPS>, total);
      return true;
    case 16:
      if constexpr (D == 3 && sizeof(T) == 8) {
        if (P.rstr_split) {
          // two z slabs per box (26 KB of LDS each instead of 47 KB), one
          // 256-lane workgroup per slab
          hipLaunchKernelGGL((k_resid_rstr_lds<D, CYL, 16, T, EPS, false, 2>), dim3(2 * n),
                             dim3(256), 0, P.stream, a, list, n, i_phi, i_rhs, i_tmp, keep_res);
          return true;
        }
      }
      go(k_resid_rstr_lds<D, CYL, 16, T, EPS>, total);
      return true;
    case 32:
      if (D == 2) {
        go(k_resid_rstr_lds<2, CYL, 32, T, EPS>, total);
        return true;
      }
      return false;
    default: return false;
  }
}

template <int D, bool CYL, typename T>
static bool launch_resid_rstr_lds(Pool& P, const Dev& a, const int* list, int n, int nc,
                                  int i_phi, int i_rhs, int i_tmp, int keep_res) {
  if (a.tag != nullptr)
    return launch_resid_rstr_lds_t<D, CYL, T, true>(P, a, list, n, nc, i_phi, i_rhs, i_tmp,
                                                    keep_res);
  return launch_resid_rstr_lds_t<D, CYL, T, false>(P, a, list, n, nc, i_phi, i_rhs, i_tmp,
                                                   keep_res);
}

// Residual + restriction of level lvl into its parents. The face ghosts of
// phi on lvl are recomputed inside the pass, so callers need not have
// stored them since the last change of phi on lvl.
template <int D, typename T>
static void resid_rstr(Pool& P, const Ctx& c, int lvl, int keep_res) {
  const int n = P.n_ids(lvl);
  if (n == 0) return;
  const int nc = P.nc_lvl(lvl);
  const int hnc = nc / 2;
  Geo g = geo(n, D == 2 ? hnc * hnc : hnc * hnc * hnc);
  Dev a = dev_for(P, c);
  const double cells = (double)n * cells_of(D, nc, false);
  int pr = P.prof_begin(PK_RESID_RSTR, lvl, (2.0 + 2.0 / (1 << D)) * sizeof(T) * cells);
  const bool done =
      c.cyl ? launch_resid_rstr_lds<2, true, T>(P, a, P.ids(lvl), n, nc, c.i_phi, c.i_rhs,
                                                c.i_tmp, keep_res)
            : launch_resid_rstr_lds<D, false, T>(P, a, P.ids(lvl), n, nc, c.i_phi, c.i_rhs,
                                                 c.i_tmp, keep_res);
  if (!done) {
    P.prof_end(pr);
    fill_gc_phi<D, T>(P, c, lvl);
    pr = P.prof_begin(PK_RESID_RSTR, lvl, (2.0 + 2.0 / (1 << D)) * sizeof(T) * cells);
    if (c.cyl)
      hipLaunchKernelGGL((k_resid_rstr<2, true, T>), dim3(g.grid), dim3(256), 0, P.stream, a,
                         P.ids(lvl), n, nc, g.tpb, c.i_phi, c.i_rhs, c.i_tmp, keep_res);
    else
      hipLaunchKernelGGL((k_resid_rstr<D, false, T>), dim3(g.grid), dim3(256), 0, P.stream, a,
                         P.ids(lvl), n, nc, g.tpb, c.i_phi, c.i_rhs, c.i_tmp, keep_res);
  }
  post_launch("resid_rstr");
  P.prof_end(pr);
}

// The FAS right-hand side on the parents of level lvl. fused: the parents'
// face ghosts of phi are recomputed in the pass (the LDS kernels); where no
// such kernel exists they are stored first. Returns whether it recomputed.
template <int D, typename T>
static void coarse_rhs(Pool& P, const Ctx& c, int lvl, int copy_tmp, bool fused = false) {
  const int n = P.n_parents(lvl);
  if (n == 0) return;
  const int nc = P.nc_lvl(lvl);
  Geo g = geo(n, cells_of(D, nc, true));
  Dev a = dev_for(P, c);
  const double cells = (double)n * cells_of(D, nc, false);
  const int f = fused ? 1 : 0;
  const bool lds = nc == 4 || nc == 8 || nc == 16 || (D == 2 && nc == 32);
  if (fused && !lds) fill_gc_phi<D, T>(P, c, lvl);
  const int pr = P.prof_begin(PK_COARSE_RHS, lvl, 4.0 * sizeof(T) * cells);
  bool done;
  if (c.cyl)
    done = c.eps ? launch_coarse_rhs_lds_t<2, true, T, true>(P, a, P.parents(lvl), n, nc,
                                                              c.i_phi, c.i_rhs, c.i_tmp, copy_tmp, f)
                 : launch_coarse_rhs_lds_t<2, true, T, false>(P, a, P.parents(lvl), n, nc,
                                                               c.i_phi, c.i_rhs, c.i_tmp,
                                                               copy_tmp, f);
  else
    done = c.eps ? launch_coarse_rhs_lds_t<D, false, T, true>(P, a, P.parents(lvl), n, nc,
                                                               c.i_phi, c.i_rhs, c.i_tmp,
                                                               copy_tmp, f)
                 : launch_coarse_rhs_lds_t<D, false, T, false>(P, a, P.parents(lvl), n, nc,
                                                                c.i_phi, c.i_rhs, c.i_tmp,
                                                                copy_tmp, f);
  if (!done) {
    if (c.cyl)
      hipLaunchKernelGGL((k_coarse_rhs<2, true, T>), dim3(g.grid), dim3(256), 0, P.stream, a,
                         P.parents(lvl), n, nc, g.tpb, c.i_phi, c.i_rhs, c.i_tmp, copy_tmp);
    else
      hipLaunchKernelGGL((k_coarse_rhs<D, false, T>), dim3(g.grid), dim3(256), 0, P.stream, a,
                         P.parents(lvl), n, nc, g.tpb, c.i_phi, c.i_rhs, c.i_tmp, copy_tmp);
  }
  post_launch("coarse_rhs");
  P.prof_end(pr);
}

template <int D, typename T>
static void update_coarse(Pool& P, const Ctx& c, int lvl) {
  // update_coarse, m_aX_mg.f90:430-486. The reference stores the ghost cells
  // of phi on lvl - 1 after the restriction; here their two readers, the
  // coarse rhs and the first down half-sweep of lvl - 1, recompute them
  // from the same state, so they are not stored (the up-smoothing of
  // lvl - 1 stores them before anything else reads them).
  resid_rstr<D, T>(P, c, lvl, 0);
  coarse_rhs<D, T>(P, c, lvl - 1, 1, true);
}

template <int D, typename T>
static void set_coarse_phi_rhs(Pool& P, const Ctx& c, int lvl, int max_lvl) {
  // set_coarse_phi_rhs, m_aX_mg.f90:490-523
  if (lvl == max_lvl) fill_gc_phi<D, T>(P, c, lvl);
  resid_rstr<D, T>(P, c, lvl, 1);
  fill_gc_phi<D, T>(P, c, lvl - 1);
  coarse_rhs<D, T>(P, c, lvl - 1, 0);
}

template <int D, typename T>
static void correct_children(Pool& P, const Ctx& c, int lvl) {
  // correct_children over the parents of lvl, m_aX_mg.f90:374-401
  const int n = P.n_parents(lvl);
  if (n == 0) return;
  const int nc = P.nc_lvl(lvl);
  const int bl = cells_of(D, nc, true);
  Dev a = dev_for(P, c);
  Geo g = geo(n, bl);
  int pr = P.prof_begin(PK_CORR, lvl, 3.0 * sizeof(T) * n * cells_of(D, nc, false));
  hipLaunchKernelGGL((k_sub_whole<T>), dim3(g.grid), dim3(256), 0, P.stream, a, P.parents(lvl),
                     n, bl, g.tpb, c.i_phi, c.i_tmp);
  post_launch("corr_prep");
  P.prof_end(pr);
  // every child of these parents is a box of level lvl+1
  const int nch = P.n_ids(lvl + 1);
  const int cnc = P.nc_lvl(lvl + 1);
  Geo g2 = geo(nch, cells_of(D, cnc, false));
  pr = P.prof_begin(PK_CORR, lvl + 1,
                    (2.0 + 1.0 / (1 << D)) * sizeof(T) * nch * cells_of(D, cnc, false));
  if (!launch_corr_lds<D, T>(P, a, P.ids(lvl + 1), nch, cnc, c.i_phi, c.i_tmp))
    hipLaunchKernelGGL((k_corr<D, T>), dim3(g2.grid), dim3(256), 0, P.stream, a,
                       P.ids(lvl + 1), nch, cnc, g2.tpb, c.i_phi, c.i_tmp);
  post_launch("corr");
  P.prof_end(pr);
}

template <typename T>
static void copy_level_t(Pool& P, int lvl, int from, int to) {
  const int n = P.n_ids(lvl);
  if (n == 0) return;
  const int nc = P.nc_lvl(lvl);
  const int D = P.tree.dim;
  const int bl = cells_of(D, nc, true);
  Geo g = geo(n, bl);
  const int pr = P.prof_begin(PK_COPY, lvl, 2.0 * sizeof(T) * n * cells_of(D, nc, false));
  hipLaunchKernelGGL((k_copy_whole<T>), dim3(g.grid), dim3(256), 0, P.stream, P.dev(),
                     P.ids(lvl), n, bl, g.tpb, from, to);
  post_launch("copy");
  P.prof_end(pr);
}

static void copy_level(Pool& P, int lvl, int from, int to) { copy_level_t<double>(P, lvl, from, to); }

template <int D, typename T>
static void residual_level(Pool& P, const Ctx& c, int lvl, bool leaves_only) {
  const int n = leaves_only ? P.n_leaves(lvl) : P.n_ids(lvl);
  if (n == 0) return;
  const int nc = P.nc_lvl(lvl);
  Geo g = geo(n, cells_of(D, nc, false));
  const int* list = leaves_only ? P.leaves(lvl) : P.ids(lvl);
  Dev a = dev_for(P, c);
  if (c.cyl)
    hipLaunchKernelGGL((k_residual<2, true, T>), dim3(g.grid), dim3(256), 0, P.stream, a, list, n,
                       nc, g.tpb, c.i_phi, c.i_rhs, c.i_tmp);
  else
    hipLaunchKernelGGL((k_residual<D, false, T>), dim3(g.grid), dim3(256), 0, P.stream, a, list,
                       n, nc, g.tpb, c.i_phi, c.i_rhs, c.i_tmp);
  post_launch("residual");
}

template <int D, typename T>
static void init_phi_rhs(Pool& P, const Ctx& c) {
  // init_phi_rhs, m_aX_mg.f90:526-549
  for (int lvl = P.tree.max_lvl; lvl >= P.tree.min_lvl + 1; lvl--) {
    const int n = P.n_ids(lvl);
    if (n == 0) continue;
    const int nc = P.nc_lvl(lvl);
    Geo g = geo(n, cells_of(D, nc, true));
    hipLaunchKernelGGL((k_rstr_clear<D, T>), dim3(g.grid), dim3(256), 0, P.stream,
                       dev_for(P, c), P.ids(lvl), n, nc, g.tpb, c.i_rhs, c.i_phi);
    post_launch("init_phi_rhs");
  }
}

template <int D, typename T>
static void fas_vcycle(Pool& P, const Ctx& c, int max_lvl, bool set_residual) {
  // mg$D_fas_vcycle, m_aX_mg.f90:208-255
  // Two ghost passes of the reference are folded into their readers: the one
  // after the down-smoothing (the residual pass recomputes those ghosts) and
  // the one after the correction (the first up-smoothing half-sweep does).
  // Until the up-smoothing stores them again, the stored ghosts of a level
  // are read by nobody.
  // The FP32 correction of the mixed solve goes further: its ghost cells are
  // never read from outside, so a level without parents (whose ghosts no
  // correction reads) does not store them after the up-smoothing, and the
  // next down-smoothing of that level recomputes them in its first
  // half-sweep.
  // In FP64 the finest level of the tree does the same across calls
  // (defer_top): its ghosts are stored by the next call that reads them
  // (flush_ghosts), or recomputed by the first half-sweep of the next
  // V-cycle (top_stale).
  const int min_lvl = P.tree.min_lvl;
  auto lazy = [&](int lvl) { return c.homogeneous && P.n_parents(lvl) == 0; };
  const bool defer = c.defer_top && !set_residual && max_lvl == P.tree.max_lvl &&
                     P.n_parents(max_lvl) == 0;
  for (int lvl = max_lvl; lvl >= min_lvl + 1; lvl--) {
    // below max_lvl the stored ghosts are stale (update_coarse)
    gsrb_boxes<D, T>(P, c, lvl, c.n_down, lvl < max_lvl || lazy(lvl) || c.top_stale, false);
    update_coarse<D, T>(P, c, lvl);
  }
  gsrb_boxes<D, T>(P, c, min_lvl, c.n_base, min_lvl < max_lvl || c.top_stale,
                   !(min_lvl == max_lvl && defer));
  for (int lvl = min_lvl + 1; lvl <= max_lvl; lvl++) {
    correct_children<D, T>(P, c, lvl - 1);
    gsrb_boxes<D, T>(P, c, lvl, c.n_up, true, !(lvl == max_lvl && (lazy(lvl) || defer)));
  }
  if (set_residual)
    for (int lvl = min_lvl; lvl <= max_lvl; lvl++) residual_level<D, T>(P, c, lvl, false);
}

template <int D, typename T>
static void fas_fmg(Pool& P, const Ctx& c, bool set_residual, bool first_call) {
  // mg$D_fas_fmg, m_aX_mg.f90:164-203
  // tree%max_lvl of the reference; with instances the finest level of the
  // active ones, so that each run takes the V-cycles its own tree would
  const int min_lvl = P.tree.min_lvl, max_lvl = P.mg_max_lvl();
  if (first_call) init_phi_rhs<D, T>(P, c);
  else
    for (int lvl = max_lvl; lvl >= min_lvl + 1; lvl--) set_coarse_phi_rhs<D, T>(P, c, lvl, max_lvl);
  for (int lvl = min_lvl; lvl <= max_lvl; lvl++) {
    copy_level_t<T>(P, lvl, c.i_phi, c.i_tmp);
    if (lvl > min_lvl) {
      correct_children<D, T>(P, c, lvl - 1);
      fill_gc_phi<D, T>(P, c, lvl);
    }
    fas_vcycle<D, T>(P, c, lvl, set_residual && lvl == max_lvl);
  }
}

// mg$D_set_box_tag (m_aX_mg.f90:673-712) for the boxes that meet the
// multigrid for the first time since they were created: variable-eps when eps
// varies over the box (ghosts included), constant-eps when it is one value
// other than 1, else normal. The multigrid never writes eps, so tagging every
// such box at the start of a call gives the tags the reference sets at the
// box's first operator call.
static void ensure_tags(Pool& P) {
  auto& T = P.tree;
  std::vector<int> todo;
  for (int id = 1; id <= T.max_id; id++)
    if (T.in_use[id] && T.tag[id] == kTagInit) todo.push_back(id);
  if (todo.empty()) return;
  const int n = (int)todo.size();
  // the pool's work buffer: the ranges first, then the id list
  double* d_out = (double*)P.work(2 * (size_t)n * sizeof(double) + n * sizeof(int));
  int* d_list = (int*)(d_out + 2 * (size_t)n);
  std::vector<double> out(2 * (size_t)n);
  check(hipMemcpyAsync(d_list, todo.data(), n * sizeof(int), hipMemcpyHostToDevice, P.stream),
        "tag list");
  if (T.dim == 2)
    hipLaunchKernelGGL(k_box_range<2>, dim3(n), dim3(256), 0, P.stream, P.dev(), d_list, n,
                       P.mg.i_eps, d_out);
  else
    hipLaunchKernelGGL(k_box_range<3>, dim3(n), dim3(256), 0, P.stream, P.dev(), d_list, n,
                       P.mg.i_eps, d_out);
  post_launch("box_range");
  check(hipMemcpyAsync(out.data(), d_out, out.size() * sizeof(double), hipMemcpyDeviceToHost,
                       P.stream), "tag range");
  check(hipStreamSynchronize(P.stream), "tag sync");
  for (int q = 0; q < n; q++) {
    const double lo = out[2 * q], hi = out[2 * q + 1];
    int tag = kTagNormal;
    if (hi > lo) tag = kTagVeps;
    else if (lo < 1 || lo > 1) tag = kTagCeps;
    T.tag[todo[q]] = tag;
  }
  check(hipMemcpyAsync(P.d_tag, T.tag.data(), (T.max_id + 1) * sizeof(int),
                       hipMemcpyHostToDevice, P.stream), "tag upload");
}

// ---------------------------------------------------------------------------
// HIP graphs of whole multigrid calls. A call sequence (FMG, V-cycle, the
// FP32 correction of the mixed solve) is captured on its first use and
// replayed afterwards. Its kernel arguments are fixed at capture, so the
// cache is dropped whenever the fingerprint of what they derive from changes:
// the pool's device pointers and strides, the face rules, the level lists and
// their offsets, and the multigrid options. Timed runs (events on) and
// A5G_GRAPHS=0 use plain launches.
// ---------------------------------------------------------------------------

static uint64_t fnv(uint64_t h, uint64_t v) {
  for (int b = 0; b < 8; b++) {
    h ^= (v >> (8 * b)) & 0xff;
    h *= 1099511628211ull;
  }
  return h;
}

static uint64_t state_fingerprint(Pool& P) {
  uint64_t h = 1469598103934665603ull;
  auto ptr = [&](const void* q) { h = fnv(h, (uint64_t)(uintptr_t)q); };
  auto num = [&](uint64_t v) { h = fnv(h, v); };
  ptr(P.cc); ptr(P.fc); ptr(P.cc32); num(P.n32); num(P.cap32); num(P.dev_cap);
  ptr(P.d_lvl); ptr(P.d_ix); ptr(P.d_parent); ptr(P.d_children); ptr(P.d_nbs);
  ptr(P.d_ncell); ptr(P.d_tag); ptr(P.d_dr); ptr(P.d_rmin); ptr(P.d_lists);
  num(P.tree.coord);
  num(P.n_inst); ptr(P.d_inst);
  {
    uint64_t sb;
    std::memcpy(&sb, &P.inst_stride, 8);
    num(sb);
  }
  return h;
}

// What else a captured call's arguments derive from and that switches back
// and forth (the field solve and the photoionization terms, groups of
// instances): the face rules, the multigrid options, and the counts and
// offsets of the level lists. Graphs of several such states are kept.
static uint64_t variant_fingerprint(Pool& P) {
  uint64_t h = 1469598103934665603ull;
  auto num = [&](uint64_t v) { h = fnv(h, v); };
  for (int i = 0; i < 6; i++) {
    num(P.bc_kind[i]);
    uint64_t b;
    std::memcpy(&b, &P.bc_val[i], 8);
    num(b);
  }
  num(P.mg.i_phi); num(P.mg.i_tmp); num(P.mg.i_rhs); num(P.mg.i_eps);
  num(P.mg.n_down); num(P.mg.n_up); num(P.mg.n_base); num(P.mg.auto_ops);
  {
    uint64_t hb;
    std::memcpy(&hb, &P.mg.helm, 8);
    num(hb);
  }
  num((uint64_t)(int64_t)P.tree.min_lvl); num(P.tree.max_lvl);
  for (int l = P.tree.min_lvl; l <= P.tree.max_lvl; l++) {
    // offsets of the lists in use (the active ones, or one group's)
    num((uint64_t)(P.ids(l) - P.d_lists));
    num((uint64_t)(P.leaves(l) - P.d_lists));
    num((uint64_t)(P.parents(l) - P.d_lists));
    num(P.n_ids(l)); num(P.n_leaves(l)); num(P.n_parents(l));
  }
  return h;
}

// graphs kept for earlier level-list states (instances switched on and off)
constexpr size_t kMaxGraphs = 48;

template <class F>
static void run_captured(Pool& P, uint64_t key, F&& f) {
  if (!P.graphs_on || P.prof_on) {
    f();
    return;
  }
  const uint64_t st = state_fingerprint(P);
  if (st != P.graphs_state) {
    P.clear_graphs();
    P.graphs_state = st;
  }
  key = fnv(variant_fingerprint(P), key);
  for (auto& g : P.graphs)
    if (g.first == key) {
      check(hipGraphLaunch(g.second, P.stream), "graph launch");
      return;
    }
  if (P.graphs.size() >= kMaxGraphs) {
    if (P.stream != P.home) {
      // inside a side-by-side section evicting would wait for every stream
      // (branches makes room before the fork): run this call plainly
      f();
      return;
    }
    // the oldest graph may still run on a side stream
    P.sync_streams();
    (void)hipGraphExecDestroy(P.graphs.front().second);
    P.graphs.erase(P.graphs.begin());
  }
  check(hipStreamBeginCapture(P.stream, hipStreamCaptureModeThreadLocal), "begin capture");
  try {
    f();
  } catch (...) {
    hipGraph_t g = nullptr;
    (void)hipStreamEndCapture(P.stream, &g);
    if (g) (void)hipGraphDestroy(g);
    throw;
  }
  hipGraph_t g = nullptr;
  check(hipStreamEndCapture(P.stream, &g), "end capture");
  hipGraphExec_t ex = nullptr;
  const hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  check(e, "graph instantiate");
  P.graphs.push_back({key, ex});
  check(hipGraphLaunch(ex, P.stream), "graph launch");
}

static uint64_t call_key(int kind, int a0, int a1, int a2) {
  return ((uint64_t)kind << 48) | ((uint64_t)(a0 & 0xffff) << 32) |
         ((uint64_t)(a1 & 0xffff) << 16) | (uint64_t)(a2 & 0xffff);
}

static void mg_check(Pool& P) {
  // check_mg, m_aX_mg.f90:156-159
  if (!P.mg.ready) throw std::logic_error("check_mg: you haven't called mg_init_mg");
  P.sync_topology();
  if (P.mg.auto_ops && P.mg.i_eps > 0) ensure_tags(P);
}

// Ghost cells of the potential left unstored by the last multigrid call
// (a5g_fas_fmg / a5g_fas_vcycle over the whole tree): the reference stores
// them at the end of every V-cycle (fill_gc_phi after the up-smoothing,
// m_aX_mg.f90:418-423), but inside a solve nothing reads those of the
// finest level before the next V-cycle's first half-sweep, which recomputes
// them from the same state. So the store is left to whoever comes next: a
// following V-cycle over the tree recomputes, and every other entry point
// first calls flush_ghosts, which stores exactly what the skipped pass would
// have stored (nothing has changed the level, its neighbours or its parents
// in between). The state seen through the C ABI is the reference's.
void flush_ghosts(Pool& P) {
  if (P.defer_lvl == kNoDefer) return;
  const int lvl = P.defer_lvl, iv = P.defer_iv;
  P.defer_lvl = kNoDefer;
  gc_list(P, lvl, P.ids(lvl), P.n_ids(lvl), iv, 0, 0);
}

// whether a call over the whole tree may leave the finest level's ghosts
static bool may_defer(const Pool& P, bool set_residual) {
  return !set_residual && P.n_inst == 0 && P.cur_group < 0 && P.defer_on;
}

void mg_fmg(Pool& P, bool set_residual, bool first_call) {
  mg_check(P);
  flush_ghosts(P);
  Ctx c = ctx64(P);
  c.defer_top = may_defer(P, set_residual);
  run_captured(P, call_key(1, set_residual, first_call, c.defer_top), [&] {
    if (P.tree.dim == 2) fas_fmg<2, double>(P, c, set_residual, first_call);
    else fas_fmg<3, double>(P, c, set_residual, first_call);
  });
  if (c.defer_top && P.n_parents(P.tree.max_lvl) == 0) {
    P.defer_lvl = P.tree.max_lvl;
    P.defer_iv = c.i_phi;
  }
}

void mg_vcycle(Pool& P, int max_lvl, bool set_residual) {
  mg_check(P);
  if (max_lvl < P.tree.min_lvl || max_lvl > P.tree.max_lvl)
    throw std::invalid_argument("fas_vcycle: max_lvl outside the tree");
  Ctx c = ctx64(P);
  if (P.defer_lvl != kNoDefer && !(P.defer_lvl == max_lvl && P.defer_iv == c.i_phi))
    flush_ghosts(P);
  c.top_stale = P.defer_lvl == max_lvl;
  P.defer_lvl = kNoDefer;
  c.defer_top = may_defer(P, set_residual) && max_lvl == P.tree.max_lvl &&
                P.n_parents(max_lvl) == 0;
  run_captured(P, call_key(2, max_lvl - P.tree.min_lvl, set_residual,
                           (c.top_stale ? 1 : 0) | (c.defer_top ? 2 : 0)),
               [&] {
                 if (P.tree.dim == 2) fas_vcycle<2, double>(P, c, max_lvl, set_residual);
                 else fas_vcycle<3, double>(P, c, max_lvl, set_residual);
               });
  if (c.defer_top) {
    P.defer_lvl = max_lvl;
    P.defer_iv = c.i_phi;
  }
}

// The slots of the reduction scratch: set every pair to (lo_init, hi_init)
// before the launches, combine them after.
static void red_init(Pool& P, unsigned long long lo_init, unsigned long long hi_init) {
  static thread_local std::vector<unsigned long long> init;
  init.assign(2 * kRedSlots, 0);
  for (int s = 0; s < kRedSlots; s++) {
    init[2 * s] = lo_init;
    init[2 * s + 1] = hi_init;
  }
  check(hipMemcpyAsync(P.d_red, init.data(), init.size() * sizeof(unsigned long long),
                       hipMemcpyHostToDevice, P.stream), "reduction init");
}

// min over the even slots and max over the odd ones (or max of both when
// both_max), after the stream has drained
static void red_read(Pool& P, bool both_max, unsigned long long* k0, unsigned long long* k1) {
  std::vector<unsigned long long> out(2 * kRedSlots);
  check(hipMemcpyAsync(out.data(), P.d_red, out.size() * sizeof(unsigned long long),
                       hipMemcpyDeviceToHost, P.stream), "reduction read");
  check(hipStreamSynchronize(P.stream), "reduction sync");
  unsigned long long a = out[0], b = out[1];
  for (int s = 1; s < kRedSlots; s++) {
    a = both_max ? std::max(a, out[2 * s]) : std::min(a, out[2 * s]);
    b = std::max(b, out[2 * s + 1]);
  }
  *k0 = a;
  *k1 = b;
}

void minmax_boxes(Pool& P, int iv, bool absval, bool whole, bool leaves_only,
                  double* lo, double* hi) {
  P.sync_topology();
  red_init(P, ~0ull, 0ull);
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
  red_read(P, false, &out[0], &out[1]);
  auto unkey = [](unsigned long long k) {
    unsigned long long u = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
    double v;
    std::memcpy(&v, &u, 8);
    return v;
  };
  *lo = unkey(out[0]);
  *hi = unkey(out[1]);
}


// max over leaf interiors of |a - b| and of |b|
template <int D>
__global__ __launch_bounds__(256) void k_maxdiff(Dev a, const int* __restrict__ list, int n,
                                                 int nc, int tpb, int ia, int ib,
                                                 unsigned long long* out) {
  const int bpb_ = blockDim.x / tpb;
  const int slot_ = threadIdx.x / tpb;
  const int lt = threadIdx.x - slot_ * tpb;
  const int li_ = blockIdx.x * bpb_ + slot_;
  double md = 0, mb = 0;
  if (li_ < n) {
    const int id = list[li_];
    const double* x = boxp(a, ia, id);
    const double* y = boxp(a, ib, id);
    const int total = D == 2 ? nc * nc : nc * nc * nc;
    for (int q = lt; q < total; q += tpb) {
      const int i = q % nc + 1, j = (q / nc) % nc + 1, k = D == 3 ? q / (nc * nc) + 1 : 0;
      const int c = cidx<D>(nc, i, j, k);
      md = fmax(md, fabs(x[c] - y[c]));
      mb = fmax(mb, fabs(y[c]));
    }
  }
  md = wave_max(md);
  mb = wave_max(mb);
  __shared__ double w1[4], w2[4];
  if ((threadIdx.x & 63) == 0) {
    w1[threadIdx.x >> 6] = md;
    w2[threadIdx.x >> 6] = mb;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < (int)(blockDim.x >> 6); w++) {
      md = fmax(md, w1[w]);
      mb = fmax(mb, w2[w]);
    }
    const int sl = 2 * (blockIdx.x % kRedSlots);
    atomicMax(&out[sl], ord_key(md));
    atomicMax(&out[sl + 1], ord_key(mb));
  }
}

void maxdiff_leaves(Pool& P, int ia, int ib, double* maxdiff, double* maxabs_b) {
  P.sync_topology();
  red_init(P, 0ull, 0ull);
  const int D = P.tree.dim;
  for (int lvl = P.tree.min_lvl; lvl <= P.tree.max_lvl; lvl++) {
    const int n = P.n_leaves(lvl);
    if (n == 0) continue;
    const int nc = P.nc_lvl(lvl);
    Geo g = geo(n, cells_of(D, nc, false));
    if (D == 2)
      hipLaunchKernelGGL((k_maxdiff<2>), dim3(g.grid), dim3(256), 0, P.stream, P.dev(),
                         P.leaves(lvl), n, nc, g.tpb, ia, ib, (unsigned long long*)P.d_red);
    else
      hipLaunchKernelGGL((k_maxdiff<3>), dim3(g.grid), dim3(256), 0, P.stream, P.dev(),
                         P.leaves(lvl), n, nc, g.tpb, ia, ib, (unsigned long long*)P.d_red);
    post_launch("maxdiff");
  }
  unsigned long long out[2];
  red_read(P, true, &out[0], &out[1]);
  double v[2];
  for (int m = 0; m < 2; m++) {
    unsigned long long u = (out[m] >> 63) ? (out[m] & 0x7fffffffffffffffull) : ~out[m];
    std::memcpy(&v[m], &u, 8);
  }
  *maxdiff = v[0];
  *maxabs_b = v[1];
}

void mg_residual_norms(Pool& P, double* max_res, double* max_rhs) {
  mg_check(P);
  const Ctx c = ctx64(P);
  for (int lvl = P.tree.min_lvl; lvl <= P.tree.max_lvl; lvl++) {
    if (P.tree.dim == 2) residual_level<2, double>(P, c, lvl, true);
    else residual_level<3, double>(P, c, lvl, true);
  }
  double lo;
  minmax_boxes(P, P.mg.i_tmp, true, false, true, &lo, max_res);
  minmax_boxes(P, P.mg.i_rhs, true, false, true, &lo, max_rhs);
}

// ---------------------------------------------------------------------------
// Mixed precision: FP32 FAS multigrid on the correction equation inside an
// FP64 defect-correction loop. The FP64 residual, the coarsest-level data of
// the outer problem and the update of phi stay FP64.
// ---------------------------------------------------------------------------

// FP64 residual r = rhs - A phi on leaf interiors, stored as FP32 right-hand
// side of the correction problem; its max |r| is reduced on the way
template <int D, bool CYL>
__global__ __launch_bounds__(256) void k_resid_to32(Dev a, const int* __restrict__ list, int n,
                                                    int nc, int tpb, int i_phi, int i_rhs,
                                                    int j_rhs, unsigned long long* mx) {
  const int bpb_ = blockDim.x / tpb;
  const int slot_ = threadIdx.x / tpb;
  const int lt = threadIdx.x - slot_ * tpb;
  const int li_ = blockIdx.x * bpb_ + slot_;
  double m = 0;
  if (li_ < n) {
    const int id = list[li_];
    const double* p = boxp(a, i_phi, id);
    const double* r = boxp(a, i_rhs, id);
    float* o = boxp_t<float>(a, j_rhs, id);
    const double inv_dr_sq = 1 / (a.dr[id] * a.dr[id]);
    const int ioff = CYL ? (a.ix[id * 3] - 1) * nc : 0;
    const int total = D == 2 ? nc * nc : nc * nc * nc;
    for (int q = lt; q < total; q += tpb) {
      const int i = q % nc + 1, j = (q / nc) % nc + 1, k = D == 3 ? q / (nc * nc) + 1 : 0;
      const int x = cidx<D>(nc, i, j, k);
      const double res = r[x] - op_at<D, CYL, double>(a, id, p, nc, i, j, k, inv_dr_sq, ioff);
      o[x] = (float)res;
      m = fmax(m, fabs(res));
    }
  }
  m = wave_max(m);
  __shared__ double wm[4];
  if ((threadIdx.x & 63) == 0) wm[threadIdx.x >> 6] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < (int)(blockDim.x >> 6); w++) m = fmax(m, wm[w]);
    atomicMax(&mx[blockIdx.x % kRedSlots], ord_key(m));
  }
}

// phi += e on box interiors
template <int D>
__global__ __launch_bounds__(256) void k_add32(Dev a, const int* __restrict__ list, int n,
                                               int nc, int tpb, int i_phi, int j_e) {
  BOX_PROLOGUE
  double* p = boxp(a, i_phi, id);
  const float* e = boxp_t<float>(a, j_e, id);
  const int total = D == 2 ? nc * nc : nc * nc * nc;
  for (int q = lt; q < total; q += tpb) {
    const int i = q % nc + 1, j = (q / nc) % nc + 1, k = D == 3 ? q / (nc * nc) + 1 : 0;
    const int x = cidx<D>(nc, i, j, k);
    p[x] = p[x] + (double)e[x];
  }
}

// whole FP32 boxes set to zero
__global__ __launch_bounds__(256) void k_zero32(Dev a, const int* __restrict__ list, int n,
                                                int bl, int tpb, int j) {
  BOX_PROLOGUE
  float* p = boxp_t<float>(a, j, id);
  for (int q = lt; q < bl; q += tpb) p[q] = 0.0f;
}

static void ensure32(Pool& P, int n32) {
  const int need = std::max(P.dev_cap, 1);
  if (P.cc32 && P.n32 >= n32 && P.cap32 == need) return;
  if (P.cc32) check(hipFree(P.cc32), "hipFree");
  P.cc32 = nullptr;
  check(hipMalloc((void**)&P.cc32, (size_t)n32 * (need + 1) * P.bs() * sizeof(float)),
        "hipMalloc cc32");
  P.n32 = n32;
  P.cap32 = need;
}

static void restrict_levels(Pool& P, int iv, bool plain);

template <int D>
static void mixed_solve(Pool& P, double tol, int max_outer, int n_inner, int* outer_done,
                        double* rel_res) {
  const Ctx c64 = ctx64(P);
  // FP32 variables: 1 = correction e, 2 = tmp, 3 = rhs of the correction
  const Ctx c32{1, 2, 3, true, c64.cyl, c64.eps, c64.n_down, c64.n_up, c64.n_base, c64.helm};
  ensure32(P, 3);
  const int min_lvl = P.tree.min_lvl, max_lvl = P.tree.max_lvl;
  double rhs_lo, rhs_max;
  minmax_boxes(P, c64.i_rhs, true, false, true, &rhs_lo, &rhs_max);
  // the FP64 residual takes each box's operator (permittivity tags)
  Dev a = dev_for(P, c64);
  int it = 0;
  double rel = 0;
  for (;; it++) {
    // r = rhs - A phi (FP64) -> FP32 rhs of the correction, and max |r|
    red_init(P, 0ull, 0ull);
    for (int lvl = min_lvl; lvl <= max_lvl; lvl++) {
      const int n = P.n_leaves(lvl);
      if (n == 0) continue;
      const int nc = P.nc_lvl(lvl);
      Geo g = geo(n, cells_of(D, nc, false));
      const int pr = P.prof_begin(PK_OTHER, lvl, (2.0 * 8 + 4) * n * cells_of(D, nc, false));
      if (c64.cyl)
        hipLaunchKernelGGL((k_resid_to32<2, true>), dim3(g.grid), dim3(256), 0, P.stream, a,
                           P.leaves(lvl), n, nc, g.tpb, c64.i_phi, c64.i_rhs, c32.i_rhs,
                           (unsigned long long*)P.d_red);
      else
        hipLaunchKernelGGL((k_resid_to32<D, false>), dim3(g.grid), dim3(256), 0, P.stream, a,
                           P.leaves(lvl), n, nc, g.tpb, c64.i_phi, c64.i_rhs, c32.i_rhs,
                           (unsigned long long*)P.d_red);
      post_launch("resid_to32");
      P.prof_end(pr);
    }
    // the residual maxima sit in the first kRedSlots words
    unsigned long long k0, k1, key;
    red_read(P, true, &k0, &k1);
    key = std::max(k0, k1);
    unsigned long long u = (key >> 63) ? (key & 0x7fffffffffffffffull) : ~key;
    double rmax;
    std::memcpy(&rmax, &u, 8);
    rel = rhs_max > 0 ? rmax / rhs_max : rmax;
    if (rel <= tol || it >= max_outer) break;
    // e = 0 everywhere, then FMG (+ V-cycles) on A e = r with zero boundary
    // data; phi += e (FP64) on the leaves; parents then take the restriction
    // of the new leaf values, which is where the FAS iteration leaves them at
    // convergence (update_coarse, m_aX_mg.f90:430-486), so that the
    // refinement-boundary ghosts of the next residual see the same composite
    // problem as the FP64 solve. Then the FP64 ghost layer, coarse to fine.
    const bool first = it == 0;
    run_captured(P, call_key(3, first, n_inner, max_lvl - min_lvl), [&] {
      for (int lvl = min_lvl; lvl <= max_lvl; lvl++) {
        const int n = P.n_ids(lvl);
        if (n == 0) continue;
        const int bl = cells_of(D, P.nc_lvl(lvl), true);
        Geo g = geo(n, bl);
        hipLaunchKernelGGL(k_zero32, dim3(g.grid), dim3(256), 0, P.stream, a, P.ids(lvl), n,
                           bl, g.tpb, c32.i_phi);
        post_launch("zero32");
      }
      // the first correction is the whole solution: FMG from the coarse
      // levels; later ones start from e = 0 with V-cycles
      if (first) {
        fas_fmg<D, float>(P, c32, false, true);
        for (int v = 1; v < n_inner; v++) fas_vcycle<D, float>(P, c32, max_lvl, false);
      } else {
        for (int v = 0; v < n_inner; v++) fas_vcycle<D, float>(P, c32, max_lvl, false);
      }
      for (int lvl = min_lvl; lvl <= max_lvl; lvl++) {
        const int n = P.n_leaves(lvl);
        if (n == 0) continue;
        const int nc = P.nc_lvl(lvl);
        Geo g = geo(n, cells_of(D, nc, false));
        const int pr = P.prof_begin(PK_OTHER, lvl, (2.0 * 8 + 4) * n * cells_of(D, nc, false));
        hipLaunchKernelGGL((k_add32<D>), dim3(g.grid), dim3(256), 0, P.stream, a, P.leaves(lvl),
                           n, nc, g.tpb, c64.i_phi, c32.i_phi);
        post_launch("add32");
        P.prof_end(pr);
      }
      restrict_levels(P, c64.i_phi, true);
      for (int lvl = min_lvl; lvl <= max_lvl; lvl++) fill_gc_phi<D, double>(P, c64, lvl);
    });
  }
  *outer_done = it;
  *rel_res = rel;
}

void mg_solve_mixed(Pool& P, double tol, int max_outer, int n_inner, int* outer_done,
                    double* rel_res) {
  mg_check(P);
  if (n_inner < 1) throw std::invalid_argument("mixed solve: n_inner must be >= 1");
  if (P.tree.dim == 2) mixed_solve<2>(P, tol, max_outer, n_inner, outer_done, rel_res);
  else mixed_solve<3>(P, tol, max_outer, n_inner, outer_done, rel_res);
}

// Photoionization by the three-term Helmholtz model (Bourdon et al. 2007,
// Plasma Sources Sci. Technol. 16, 656) in place of the reference's Monte-Carlo
// photons (set_photoionization, src/streamer_2d.f90:544-565): the production
// rate I (photoi_rate) is turned into the source sum_j S_j, where
// (lpl - kappa_j^2) S_j = -amp_j I with S_j = 0 on the domain boundary. Each
// term is one FMG plus V-cycles until max|r| <= tol max|rhs| on the leaves.
// The work variables carry phi, rhs and tmp of these solves; the caller's
// multigrid options are restored afterwards. The leaf rhs of term j + 1 is
// the leaf rhs of term j times amp_{j+1} / amp_j (a FAS solve leaves the rhs
// of leaves unchanged), so that i_pho can sum the terms from the start.
void photoi_helmholtz(Pool& P, const StVars& v, double coeff, int n_terms, const double* amp,
                      const double* kappa, int i_phi, int i_rhs, int i_tmp, double tol,
                      int max_cycles, int* cycles_done) {
  if (n_terms < 1) throw std::invalid_argument("photoionization: need at least one term");
  for (int j = 0; j < n_terms; j++)
    if (!(amp[j] > 0) || !(kappa[j] >= 0))
      throw std::invalid_argument("photoionization: need amp > 0 and kappa >= 0");
  if (i_phi == v.pho || i_rhs == v.pho || i_tmp == v.pho)
    throw std::invalid_argument("photoionization: work variables must differ from i_pho");
  P.sync_topology();
  photoi_rate(P, v, coeff);
  const MgOpts saved = P.mg;
  int bk[6];
  double bv[6];
  for (int i = 0; i < 6; i++) {
    bk[i] = P.bc_kind[i];
    bv[i] = P.bc_val[i];
    P.bc_kind[i] = 1;  // S = 0 on the boundary: ghost = -c
    P.bc_val[i] = 0;
  }
  // the operator family of the field solve (cylindrical when it has the
  // automatic operators), without permittivity
  P.mg.i_phi = i_phi;
  P.mg.i_rhs = i_rhs;
  P.mg.i_tmp = i_tmp;
  P.mg.i_eps = -1;
  P.mg.ready = true;
  int total = 0;
  try {
    axpy_leaves(P, v.pho, i_rhs, -amp[0], false);
    axpy_leaves(P, v.pho, v.pho, 0.0, false);  // i_pho = 0 on leaves
    for (int j = 0; j < n_terms; j++) {
      if (j > 0) axpy_leaves(P, i_rhs, i_rhs, amp[j] / amp[j - 1], false);
      P.mg.helm = kappa[j] * kappa[j];
      const Ctx c = ctx64(P);
      // captured like the field solve's calls; the Helmholtz coefficient
      // keeps each term's graphs apart
      run_captured(P, call_key(1, 0, 1, 0), [&] {
        if (P.tree.dim == 2) fas_fmg<2, double>(P, c, false, true);
        else fas_fmg<3, double>(P, c, false, true);
      });
      const int mx = P.mg_max_lvl();
      for (int cyc = 0;; cyc++) {
        double res, rhs;
        mg_residual_norms(P, &res, &rhs);
        if (res <= tol * rhs) break;
        if (cyc >= max_cycles)
          throw std::runtime_error("photoionization: Helmholtz solve did not converge");
        run_captured(P, call_key(2, mx - P.tree.min_lvl, 0, 0), [&] {
          if (P.tree.dim == 2) fas_vcycle<2, double>(P, c, mx, false);
          else fas_vcycle<3, double>(P, c, mx, false);
        });
        total++;
      }
      axpy_leaves(P, i_phi, v.pho, 1.0, true);
    }
  } catch (...) {
    P.mg = saved;
    for (int i = 0; i < 6; i++) {
      P.bc_kind[i] = bk[i];
      P.bc_val[i] = bv[i];
    }
    throw;
  }
  P.mg = saved;
  for (int i = 0; i < 6; i++) {
    P.bc_kind[i] = bk[i];
    P.bc_val[i] = bv[i];
  }
  if (cycles_done) *cycles_done = total;
}

// Maximum over leaf interiors of |rhs - A phi| (resid) or |rhs| for the
// problem c, into the kRedSlots keys at out (cleared by a kernel first, so
// that a graph can hold the whole check).
template <int D>
static void leaf_absmax(Pool& P, const Ctx& c, bool resid, unsigned long long* out) {
  hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(256), 0, P.stream, out, kRedSlots, 0ull);
  post_launch("fill");
  const Dev a = dev_for(P, c);
  for (int lvl = P.tree.min_lvl; lvl <= P.tree.max_lvl; lvl++) {
    const int n = P.n_leaves(lvl);
    if (n == 0) continue;
    const int nc = P.nc_lvl(lvl);
    Geo g = geo(n, cells_of(D, nc, false));
    if (c.cyl)
      hipLaunchKernelGGL((k_leaf_absmax<2, true>), dim3(g.grid), dim3(256), 0, P.stream, a,
                         P.leaves(lvl), n, nc, g.tpb, c.i_phi, c.i_rhs, (int)resid, out);
    else
      hipLaunchKernelGGL((k_leaf_absmax<D, false>), dim3(g.grid), dim3(256), 0, P.stream, a,
                         P.leaves(lvl), n, nc, g.tpb, c.i_phi, c.i_rhs, (int)resid, out);
    post_launch("leaf_absmax");
  }
}

// Runs f(j) for every j of js, each on its own stream: the first on the
// handle's stream, the others on side streams that wait for everything queued
// before and that the handle's stream waits for afterwards. Work of the
// different j must touch different variables. Graphs captured inside f run on
// the stream they were captured on, so independent terms overlap on the GPU.
template <class F>
static void branches(Pool& P, const std::vector<int>& js, F&& f) {
  if (js.empty()) return;
  if (js.size() == 1) {
    f(js[0]);
    return;
  }
  P.ensure_side((int)js.size() - 1);
  // Work that must wait for the device, or upload on the handle's stream,
  // happens here, before the fork: the topology upload, and room in the
  // graph cache for the captures of the section (a side stream never
  // evicts; sync_topology and Pool::work refuse to run on one).
  P.sync_topology();
  const size_t room = 8 * js.size();
  if (P.graphs_on && P.graphs.size() + room > kMaxGraphs) {
    P.sync_streams();
    while (!P.graphs.empty() && P.graphs.size() + room > kMaxGraphs) {
      (void)hipGraphExecDestroy(P.graphs.front().second);
      P.graphs.erase(P.graphs.begin());
    }
  }
  hipStream_t main = P.stream;
  check(hipEventRecord(P.ev_fork, main), "fork event");
  for (size_t k = 1; k < js.size(); k++)
    check(hipStreamWaitEvent(P.side[k - 1], P.ev_fork, 0), "fork wait");
  try {
    f(js[0]);
    for (size_t k = 1; k < js.size(); k++) {
      P.stream = P.side[k - 1];
      f(js[k]);
      P.stream = main;
      check(hipEventRecord(P.ev_join[k - 1], P.side[k - 1]), "join event");
      check(hipStreamWaitEvent(main, P.ev_join[k - 1], 0), "join wait");
    }
  } catch (...) {
    P.stream = main;
    (void)hipDeviceSynchronize();
    throw;
  }
}

// The Helmholtz photoionization with each term in variables of its own: S_j
// in i_terms[j], its right-hand side -amp_j I in i_rhs[j] and its multigrid
// scratch in i_tmp[j]. The terms then share no data, so their solves run
// side by side (branches): on a mesh whose levels are mostly small, a
// V-cycle is a chain of short launches and three chains overlap. Each term
// goes through the same steps as a solve on its own: the start (cold: FMG
// from zero, the reference's mg$D_fas_fmg with first_call; warm: the
// solution an earlier call left in i_terms[j], restricted to every parent
// with the multigrid's restriction and its ghost cells filled level by level
// upwards), then V-cycles until max|r| <= tol max|rhs_j| on the leaves.
// i_pho = sum_j S_j, added in term order.
void photoi_helmholtz_terms(Pool& P, const StVars& v, double coeff, int n_terms,
                            const double* amp, const double* kappa, const int* i_terms,
                            const int* i_rhs, const int* i_tmp, bool warm, double tol,
                            int max_cycles, int* cycles_done) {
  if (n_terms < 1 || n_terms > kMaxTerms)
    throw std::invalid_argument("photoionization: need 1 to " + std::to_string(kMaxTerms) +
                                " terms");
  for (int j = 0; j < n_terms; j++)
    if (!(amp[j] > 0) || !(kappa[j] >= 0))
      throw std::invalid_argument("photoionization: need amp > 0 and kappa >= 0");
  {
    std::vector<int> all;
    for (int j = 0; j < n_terms; j++) {
      all.push_back(i_terms[j]);
      all.push_back(i_rhs[j]);
      all.push_back(i_tmp[j]);
    }
    all.push_back(v.pho);
    std::vector<int> sorted = all;
    std::sort(sorted.begin(), sorted.end());
    if (std::adjacent_find(sorted.begin(), sorted.end()) != sorted.end())
      throw std::invalid_argument("photoionization: the term, rhs and tmp variables of all "
                                  "terms and i_pho must all differ");
  }
  P.sync_topology();
  photoi_rate(P, v, coeff);
  const MgOpts saved = P.mg;
  int bk[6];
  double bv[6];
  for (int i = 0; i < 6; i++) {
    bk[i] = P.bc_kind[i];
    bv[i] = P.bc_val[i];
    P.bc_kind[i] = 1;  // S = 0 on the boundary: ghost = -c
    P.bc_val[i] = 0;
  }
  P.mg.i_eps = -1;
  P.mg.ready = true;
  auto term = [&](int j) {
    P.mg.i_phi = i_terms[j];
    P.mg.i_rhs = i_rhs[j];
    P.mg.i_tmp = i_tmp[j];
    P.mg.helm = kappa[j] * kappa[j];
    return ctx64(P);
  };
  const auto restore = [&] {
    P.mg = saved;
    for (int i = 0; i < 6; i++) {
      P.bc_kind[i] = bk[i];
      P.bc_val[i] = bv[i];
    }
  };
  int total = 0;
  try {
    for (int j = 0; j < n_terms; j++) axpy_leaves(P, v.pho, i_rhs[j], -amp[j], false);
    axpy_leaves(P, v.pho, v.pho, 0.0, false);  // i_pho = 0 on leaves
    const int mx = P.mg_max_lvl();
    std::vector<int> active;
    for (int j = 0; j < n_terms; j++) active.push_back(j);
    // Each term's slots: kRedSlots keys of max|r|, then kRedSlots of
    // max|rhs|. The checks run inside the branches, so one copy and one wait
    // per section read them all.
    P.ensure_term_red();
    auto res_slots = [&](int j) { return P.d_term_red + (size_t)j * 2 * kRedSlots; };
    std::vector<unsigned long long> keys((size_t)kMaxTerms * 2 * kRedSlots);
    auto read_keys = [&] {
      check(hipMemcpyAsync(keys.data(), P.d_term_red, keys.size() * sizeof(unsigned long long),
                           hipMemcpyDeviceToHost, P.stream), "check read");
      check(hipStreamSynchronize(P.stream), "check sync");
    };
    auto key_max = [&](int j, int half) {
      unsigned long long k = 0;
      for (int sl = 0; sl < kRedSlots; sl++)
        k = std::max(k, keys[(size_t)j * 2 * kRedSlots + half * kRedSlots + sl]);
      if (k == 0) return 0.0;  // no leaf cell
      const unsigned long long u = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
      double d;
      std::memcpy(&d, &u, 8);
      return d;
    };
    // captured like the field solve's calls; each term's variables and
    // Helmholtz coefficient keep its graphs apart
    branches(P, active, [&](int j) {
      const Ctx c = term(j);
      run_captured(P, call_key(warm ? 4 : 5, mx - P.tree.min_lvl, j, 0), [&] {
        if (warm) {
          restrict_levels(P, c.i_phi, true);
          for (int lvl = P.tree.min_lvl; lvl <= mx; lvl++) {
            if (P.tree.dim == 2) fill_gc_phi<2, double>(P, c, lvl);
            else fill_gc_phi<3, double>(P, c, lvl);
          }
        } else {
          if (P.tree.dim == 2) fas_fmg<2, double>(P, c, false, true);
          else fas_fmg<3, double>(P, c, false, true);
        }
        // the leaf rhs of a FAS solve does not change: its maximum once
        if (P.tree.dim == 2) {
          leaf_absmax<2>(P, c, false, res_slots(j) + kRedSlots);
          leaf_absmax<2>(P, c, true, res_slots(j));
        } else {
          leaf_absmax<3>(P, c, false, res_slots(j) + kRedSlots);
          leaf_absmax<3>(P, c, true, res_slots(j));
        }
      });
    });
    read_keys();
    double rhs_max[kMaxTerms];
    for (int j = 0; j < n_terms; j++) rhs_max[j] = key_max(j, 1);
    for (int cyc = 0;; cyc++) {
      std::vector<int> next;
      for (int j : active)
        if (!(key_max(j, 0) <= tol * rhs_max[j])) next.push_back(j);
      active.swap(next);
      if (active.empty()) break;
      if (cyc >= max_cycles)
        throw std::runtime_error("photoionization: Helmholtz solve did not converge");
      branches(P, active, [&](int j) {
        const Ctx c = term(j);
        run_captured(P, call_key(6, mx - P.tree.min_lvl, j, 0), [&] {
          if (P.tree.dim == 2) {
            fas_vcycle<2, double>(P, c, mx, false);
            leaf_absmax<2>(P, c, true, res_slots(j));
          } else {
            fas_vcycle<3, double>(P, c, mx, false);
            leaf_absmax<3>(P, c, true, res_slots(j));
          }
        });
      });
      read_keys();
      total += (int)active.size();
    }
    for (int j = 0; j < n_terms; j++) axpy_leaves(P, i_terms[j], v.pho, 1.0, true);
  } catch (...) {
    restore();
    throw;
  }
  restore();
  if (cycles_done) *cycles_done = total;
}

// mg$D_fas_fmg of every group of instances (a5g_set_instance_groups), each
// group up to its own finest level, as if it were the only active one. The
// groups share no boxes, so up to kMaxTerms of them run side by side.
void mg_fmg_groups(Pool& P, bool set_residual, bool first_call) {
  mg_check(P);
  if (P.n_inst == 0 || P.n_groups < 1)
    throw std::logic_error("fas_fmg_groups: no instance groups set");
  std::vector<int> gs;
  for (int g = 0; g < P.n_groups; g++) {
    bool any = false;
    for (const auto& l : P.g_ids[g]) any = any || !l.empty();
    if (any) gs.push_back(g);
  }
  const Ctx c = ctx64(P);
  for (size_t w = 0; w < gs.size(); w += kMaxTerms) {
    const std::vector<int> wave(gs.begin() + w, gs.begin() + std::min(gs.size(), w + kMaxTerms));
    branches(P, wave, [&](int g) {
      P.cur_group = g;
      try {
        run_captured(P, call_key(1, set_residual, first_call, 0), [&] {
          if (P.tree.dim == 2) fas_fmg<2, double>(P, c, set_residual, first_call);
          else fas_fmg<3, double>(P, c, set_residual, first_call);
        });
      } catch (...) {
        P.cur_group = -1;
        throw;
      }
      P.cur_group = -1;
    });
  }
}

void gc_tree(Pool& P, int iv, int rb, int bc) {
  // a$D_gc_tree, m_aX_gc.f90:28-46
  P.sync_topology();
  for (int lvl = P.tree.min_lvl; lvl <= P.tree.max_lvl; lvl++) gc_level(P, lvl, iv, rb, bc);
}

// Restriction of iv from every level to its parents, finest first. plain
// selects the multigrid's own restriction (mg$D_box_rstr_lpl,
// m_aX_mg.f90:851-893: the 2^D average in every coordinate system) instead of
// a$D_restrict_box (radius-weighted in cylindrical coordinates).
static void restrict_levels(Pool& P, int iv, bool plain) {
  for (int lvl = P.tree.max_lvl - 1; lvl >= P.tree.min_lvl; lvl--) {
    const int n = P.n_ids(lvl + 1);
    if (n == 0) continue;
    const int nc = P.nc_lvl(lvl + 1);
    const int hnc = nc / 2;
    // 1 + 2^-D words per child cell (read the child, write the parent)
    const int pr = P.prof_begin(PK_RESTRICT, lvl + 1,
                                (1.0 + 1.0 / (1 << P.tree.dim)) * 8 * n *
                                    cells_of(P.tree.dim, nc, false));
    if (P.tree.dim == 2) {
      Geo g = geo(n, hnc * hnc);
      hipLaunchKernelGGL((k_restrict_box<2>), dim3(g.grid), dim3(256), 0, P.stream, P.dev(),
                         P.ids(lvl + 1), n, nc, g.tpb, iv, plain);
    } else {
      Geo g = geo(n, hnc * hnc * hnc);
      hipLaunchKernelGGL((k_restrict_box<3>), dim3(g.grid), dim3(256), 0, P.stream, P.dev(),
                         P.ids(lvl + 1), n, nc, g.tpb, iv, plain);
    }
    post_launch("restrict");
    P.prof_end(pr);
  }
}

void restrict_tree(Pool& P, int iv) {
  // a$D_restrict_tree, m_aX_restrict.f90:54-64: children of the parents of
  // lvl are exactly the boxes of lvl+1
  P.sync_topology();
  restrict_levels(P, iv, false);
}

void tree_copy_cc(Pool& P, int from, int to) {
  P.sync_topology();
  for (int lvl = P.tree.min_lvl; lvl <= P.tree.max_lvl; lvl++) copy_level(P, lvl, from, to);
}

// ---------------------------------------------------------------------------
// Launch-shape autotuner: for every kernel of the menu and every (box size,
// level-size class) that the tree's levels hold, each shape is timed on
// those levels (reps launches each after two untimed ones, on the
// multigrid's potential, tmp and rhs, with HIP events on the handle's
// stream), and the fastest is kept when it beats the default by more than
// 3%. The half-sweep has four shapes; the ghost fill, the residual +
// restriction, the coarse rhs and the correction have the default and half
// the lanes per box (for boxes of 8 cells a side or fewer, where the default
// leaves lanes idle). The three variables are saved to a device copy first
// and restored afterwards; every shape computes the same values, so only
// the time differs. Rows of rec: kernel, dim, nc, size class, shape, boxes
// timed, microseconds per launch, 1 if chosen.
// ---------------------------------------------------------------------------

template <int D>
static void autotune_t(Pool& P, int reps, std::vector<double>& rec) {
  const Ctx c = ctx64(P);
  const Dev a = dev_for(P, c);
  hipEvent_t e0, e1;
  check(hipEventCreate(&e0), "event");
  check(hipEventCreate(&e1), "event");
  const int min_lvl = P.tree.min_lvl;
  for (int kern = 0; kern < kShapeKernels; kern++) {
    // the boxes one launch of this kernel treats on level lvl (0: none)
    auto n_of = [&](int lvl) {
      const int nc = P.nc_lvl(lvl);
      switch (kern) {
        case kShapeGsrb: return P.n_ids(lvl);
        case kShapeGc: return nc <= 8 ? P.n_ids(lvl) : 0;
        case kShapeResidRstr: return lvl > min_lvl && nc <= 8 ? P.n_ids(lvl) : 0;
        case kShapeCoarseRhs: return nc >= 4 && nc <= 8 ? P.n_parents(lvl) : 0;
        default: return lvl > min_lvl && nc == 8 ? P.n_ids(lvl) : 0;
      }
    };
    auto launch = [&](int lvl, int it) {
      const int n = n_of(lvl), nc = P.nc_lvl(lvl);
      bool ok = true;
      switch (kern) {
        case kShapeGsrb:
          ok = c.cyl ? launch_gsrb_lds<2, true, double>(P, a, P.ids(lvl), n, nc, it, 1, c.i_phi,
                                                        c.i_rhs)
                     : launch_gsrb_lds<D, false, double>(P, a, P.ids(lvl), n, nc, it, 1,
                                                         c.i_phi, c.i_rhs);
          break;
        case kShapeGc:
          gc_list_t<D, double>(P, a, lvl, P.ids(lvl), n, c.i_phi, 0, 0);
          break;
        case kShapeResidRstr:
          ok = c.cyl ? launch_resid_rstr_lds<2, true, double>(P, a, P.ids(lvl), n, nc, c.i_phi,
                                                              c.i_rhs, c.i_tmp, 0)
                     : launch_resid_rstr_lds<D, false, double>(P, a, P.ids(lvl), n, nc,
                                                               c.i_phi, c.i_rhs, c.i_tmp, 0);
          break;
        case kShapeCoarseRhs:
          if (c.cyl)
            ok = c.eps ? launch_coarse_rhs_lds_t<2, true, double, true>(
                             P, a, P.parents(lvl), n, nc, c.i_phi, c.i_rhs, c.i_tmp, 1, 1)
                       : launch_coarse_rhs_lds_t<2, true, double, false>(
                             P, a, P.parents(lvl), n, nc, c.i_phi, c.i_rhs, c.i_tmp, 1, 1);
          else
            ok = c.eps ? launch_coarse_rhs_lds_t<D, false, double, true>(
                             P, a, P.parents(lvl), n, nc, c.i_phi, c.i_rhs, c.i_tmp, 1, 1)
                       : launch_coarse_rhs_lds_t<D, false, double, false>(
                             P, a, P.parents(lvl), n, nc, c.i_phi, c.i_rhs, c.i_tmp, 1, 1);
          break;
        default:
          ok = launch_corr_lds<D, double>(P, a, P.ids(lvl), n, nc, c.i_phi, c.i_tmp);
      }
      if (!ok) throw std::runtime_error("autotune: no kernel for this box size");
      post_launch("autotune");
    };
    std::vector<std::vector<int>> groups(6 * kSizeClasses);
    for (int lvl = min_lvl; lvl <= P.tree.max_lvl; lvl++) {
      const int n = n_of(lvl);
      if (n > 0) groups[log2i(P.nc_lvl(lvl)) * kSizeClasses + size_class(n)].push_back(lvl);
    }
    const int shapes[4] = {0, 2, 1, 3};
    const int n_shapes = kern == kShapeGsrb ? 4 : 2;
    for (int key = 0; key < (int)groups.size(); key++) {
      if (groups[key].empty()) continue;
      const int l2 = key / kSizeClasses, cls = key % kSizeClasses;
      double t[kShapes] = {0, 0, 0, 0};
      int boxes = 0;
      for (int si = 0; si < n_shapes; si++) {
        const int sh = shapes[si];
        P.shape_tab[kern][D - 2][l2][cls] = (unsigned char)sh;
        boxes = 0;
        for (int lvl : groups[key]) {
          boxes += n_of(lvl);
          launch(lvl, 1);
          launch(lvl, 2);
          check(hipEventRecord(e0, P.stream), "event");
          for (int r = 0; r < reps; r++) launch(lvl, r + 1);
          check(hipEventRecord(e1, P.stream), "event");
          check(hipEventSynchronize(e1), "event sync");
          float ms = 0;
          check(hipEventElapsedTime(&ms, e0, e1), "event time");
          t[sh] += 1e3 * ms / reps;
        }
      }
      int best = 0;
      for (int si = 1; si < n_shapes; si++)
        if (t[shapes[si]] < t[best]) best = shapes[si];
      if (!(t[best] < 0.97 * t[0])) best = 0;
      P.shape_tab[kern][D - 2][l2][cls] = (unsigned char)best;
      for (int si = 0; si < n_shapes; si++) {
        const int sh = shapes[si];
        const double row[8] = {(double)kern, (double)D, (double)(1 << l2), (double)cls,
                               (double)sh, (double)boxes, t[sh], sh == best ? 1.0 : 0.0};
        rec.insert(rec.end(), row, row + 8);
      }
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
}

void autotune(Pool& P, int reps, std::vector<double>& rec) {
  if (P.mg.i_phi <= 0 || P.mg.i_rhs <= 0 || P.mg.i_tmp <= 0)
    throw std::runtime_error("autotune: multigrid not set up");
  if (reps < 1) throw std::invalid_argument("autotune: reps must be >= 1");
  mg_check(P);  // topology uploaded, operator tags set
  // captured graphs hold the old shapes
  P.clear_graphs();
  const Dev a = P.dev();
  const int ivs[3] = {P.mg.i_phi, P.mg.i_tmp, P.mg.i_rhs};
  double* save = nullptr;
  check(hipMalloc((void**)&save, 3 * a.vs * sizeof(double)), "hipMalloc");
  auto copy_all = [&](bool back) {
    for (int k = 0; k < 3; k++) {
      double* v = a.cc + (size_t)(ivs[k] - 1) * a.vs;
      double* sv = save + (size_t)k * a.vs;
      const hipError_t e = hipMemcpyAsync(back ? v : sv, back ? sv : v, a.vs * sizeof(double),
                                          hipMemcpyDeviceToDevice, P.stream);
      if (e != hipSuccess) return e;
    }
    return hipStreamSynchronize(P.stream);
  };
  check(copy_all(false), "save");
  try {
    if (P.tree.dim == 2) autotune_t<2>(P, reps, rec);
    else autotune_t<3>(P, reps, rec);
  } catch (...) {
    (void)copy_all(true);
    (void)hipFree(save);
    throw;
  }
  check(copy_all(true), "restore");
  check(hipFree(save), "hipFree");
}


// ---------------------------------------------------------------------------
// Reverse mode of the FAS multigrid (a5g_fas_fmg_adjoint). mg$D_fas_fmg
// (m_aX_mg.f90:164-203) with first_call false and without the residual is an
// affine map of the potential, tmp and rhs it starts from and of the
// boundary values V of the rule 2V - c (every operation is linear in them:
// the Gauss-Seidel half-sweeps, the ghost rules, residual, restriction,
// coarse rhs, correction, copies). Its adjoint is the transpose of each
// operation, applied in the reverse order of the reference's schedule
// (not of this library's fused one, which computes the same values), and
// needs no state of the forward solve. Adjoint variables: lp (potential),
// lt (tmp), lr (rhs), lw (work: the residual restricted by update_coarse);
// the adjoints of V go to vbar (one per face). The operators are those the
// forward solve used per box: the Laplacian, or for boxes tagged with a
// permittivity mg$D_box_lpld, gsrb_lpld and corr_lpld (their weights depend
// on eps only, which the window holds fixed), and the Helmholtz term
// lpl(u) - helm u of the photoionization solves. The transposes scatter with
// atomics, so the sums arrive in an order of their own.
// ---------------------------------------------------------------------------

// One face ghost's rule (a$D_gc_box with mg$D_sides_rb and the physical
// rules) transposed: the adjoint gb of ghost g of face nb of box id (nb_id its
// neighbour entry) is added to the values the ghost was made from.
template <int D>
__device__ __forceinline__ void adj_ghost(const Dev& a, int il, int id, int nc, int nb,
                                          int nb_id, const int g[3], double gb,
                                          double* __restrict__ vbar) {
  double* L = boxp(a, il, id);
  const int d = d_nb_dim(nb);
  const bool low = d_nb_low(nb);
  const int cix = low ? 1 : nc, dix = low ? 1 : -1;
  const int ta = d == 0 ? 1 : 0, tb = d == 2 ? 1 : 2;
  int c[3] = {g[0], g[1], g[2]};
  c[d] = cix;
  if (nb_id > 0) {
    // sides_from_nb (m_aX_gc.f90:368-402): the neighbour's adjacent layer
    int s[3] = {c[0], c[1], c[2]};
    s[d] = low ? nc : 1;
    atomicAdd(boxp(a, il, nb_id) + cidx<D>(nc, s[0], s[1], s[2]), gb);
  } else if (nb_id == 0) {
    // mg$D_sides_rb (m_aX_mg.f90:262-356): 0.5 times the parent's cell
    // (a$D_gc_prolong0) plus an extrapolation from the box's own cells
    const int pid = a.parent[id];
    const int pnc = a.ncell[pid];
    int off[3];
    child_off<D>(a, id, nc, 0, off);
    int pc[3] = {0, 0, 0};
    for (int e = 0; e < D; e++) pc[e] = off[e] + ((g[e] + 1) >> 1);
    atomicAdd(boxp(a, il, pid) + cidx<D>(pnc, pc[0], pc[1], pc[2]), 0.5 * gb);
    atomicAdd(L + cidx<D>(nc, c[0], c[1], c[2]), (D == 2 ? 1.0 : 1.25) * gb);
    int e1[3] = {c[0], c[1], c[2]};
    e1[d] += dix;
    atomicAdd(L + cidx<D>(nc, e1[0], e1[1], e1[2]), -0.25 * gb);
    int e2[3] = {c[0], c[1], c[2]};
    e2[ta] += -1 + 2 * (c[ta] & 1);
    atomicAdd(L + cidx<D>(nc, e2[0], e2[1], e2[2]), -0.25 * gb);
    if (D == 3) {
      int e3[3] = {c[0], c[1], c[2]};
      e3[tb] += -1 + 2 * (c[tb] & 1);
      atomicAdd(L + cidx<D>(nc, e3[0], e3[1], e3[2]), -0.25 * gb);
    }
  } else {
    // the drivers' rules: copy, -c, 2V - c (also the face ghosts of the
    // whole-row rule)
    const int kind = a.bc_kind[nb - 1];
    double* cp = L + cidx<D>(nc, c[0], c[1], c[2]);
    if (kind == kBcCopy) {
      atomicAdd(cp, gb);
    } else if (kind == kBcNeg) {
      atomicAdd(cp, -gb);
    } else {
      atomicAdd(cp, -gb);
      atomicAdd(vbar + nb - 1, 2 * gb);
    }
  }
}

// a$D_gc_box's rules on one box (m_aX_gc.f90:50-68, with mg$D_sides_rb and
// the physical rules), transposed: every face ghost's adjoint is sent to the
// values it was made from and cleared. The forward pass makes face ghosts
// from interior cells only, except the corner ghosts of the whole-row rule
// 2V - c, which it writes last from the x ghosts of the same pass; so those
// corners go first (into their x ghosts), then every face ghost at once.
// Several boxes share a workgroup (tpb lanes each).
template <int D>
__global__ __launch_bounds__(256) void k_adj_gc(Dev a, const int* __restrict__ list, int n,
                                                int nc, int tpb, int il, double* __restrict__ vbar) {
  const int slot = threadIdx.x / tpb;
  const int lt = threadIdx.x - slot * tpb;
  const int li_ = blockIdx.x * (blockDim.x / tpb) + slot;
  const bool act = li_ < n;
  const int id = act ? list[li_] : 0;
  double* L = act ? boxp(a, il, id) : nullptr;
  if (D == 2 && act && lt < 4) {
    // a(gx, gy) = 2V - a(gx, cy) on the y faces with the whole-row rule
    const int nb = 3 + (lt >> 1);
    if (a.nbs[id * 6 + nb - 1] < 0 && a.bc_kind[nb - 1] == kBcRow2V) {
      const bool low = d_nb_low(nb);
      const int gy = low ? 0 : nc + 1, cy = low ? 1 : nc;
      const int gx = (lt & 1) ? nc + 1 : 0;
      double* g = L + cidx<D>(nc, gx, gy, 0);
      const double gb = *g;
      *g = 0;
      L[cidx<D>(nc, gx, cy, 0)] -= gb;
      atomicAdd(vbar + nb - 1, 2 * gb);
    }
  }
  __syncthreads();
  if (!act) return;
  const int per_face = D == 2 ? nc : nc * nc;
  for (int q = lt; q < 2 * D * per_face; q += tpb) {
    const int nb = q / per_face + 1;
    const int t = q - (nb - 1) * per_face;
    const int d = d_nb_dim(nb);
    const int nb_id = a.nbs[id * 6 + nb - 1];
    const int ta = d == 0 ? 1 : 0, tb = d == 2 ? 1 : 2;
    int g[3] = {0, 0, 0};
    g[d] = d_nb_low(nb) ? 0 : nc + 1;
    g[ta] = t % nc + 1;
    if (D == 3) g[tb] = t / nc + 1;
    double* gp = L + cidx<D>(nc, g[0], g[1], g[2]);
    const double gb = *gp;
    *gp = 0;
    if (gb == 0) continue;
    adj_ghost<D>(a, il, id, nc, nb, nb_id, g, gb, vbar);
  }
}

// One half-sweep (mg$D_box_gsrb_lpl, m_aX_mg.f90:769-811; mg2_box_gsrb_clpl,
// :1413-1439) transposed: each updated cell's adjoint goes to its
// neighbours and to its rhs, and is cleared
// the cidx of the neighbours of interior cell (i, j, k) in the order i-1,
// i+1, j-1, j+1, k-1, k+1 of the reference's u(:) and a(:)
template <int D>
__device__ __forceinline__ void nb_cells(int nc, int i, int j, int k, int nbx[6]) {
  nbx[0] = cidx<D>(nc, i - 1, j, k);
  nbx[1] = cidx<D>(nc, i + 1, j, k);
  nbx[2] = cidx<D>(nc, i, j - 1, k);
  nbx[3] = cidx<D>(nc, i, j + 1, k);
  nbx[4] = D == 3 ? cidx<D>(nc, i, j, k - 1) : 0;
  nbx[5] = D == 3 ? cidx<D>(nc, i, j, k + 1) : 0;
}

// GC: the ghost pass that precedes this half-sweep in the forward schedule
// is transposed here too: a contribution to a face ghost goes straight
// through that ghost's rule (adj_ghost) instead of into the ghost slot for a
// k_adj_gc launch. The rule sends it to the cell itself (written only by
// this lane, after its read), to cells of the other colour and to the
// parent, none of which this launch reads; the transpose is linear, so the
// sum is that of the two launches.
template <int D, bool CYL, bool GC = false>
__global__ __launch_bounds__(256) void k_adj_gsrb(Dev a, const int* __restrict__ list, int n,
                                                  int nc, int cntr, int tpb, int il, int ir,
                                                  double* __restrict__ vbar) {
  BOX_PROLOGUE
  double* L = boxp(a, il, id);
  double* R = boxp(a, ir, id);
  const double* ev =
      (a.tag != nullptr && a.tag[id] > kTagNormal) ? boxp(a, a.i_eps, id) : nullptr;
  const double dx2 = a.dr[id] * a.dr[id];
  const int half = nc >> 1;
  const int ncol = (D == 2 ? nc * nc : nc * nc * nc) >> 1;
  const int ioff = CYL ? (a.ix[id * 3] - 1) * nc : 0;
  const double w = D == 2 ? 0.25 : 1.0 / 6.0;
  for (int q = lt; q < ncol; q += tpb) {
    const int m = q % half, row = q / half;
    const int j = row % nc + 1;
    const int k = D == 3 ? row / nc + 1 : 0;
    const int i = 2 - ((cntr ^ (D == 3 ? k + j : j)) & 1) + 2 * m;
    double* yp = L + cidx<D>(nc, i, j, k);
    const double yb = *yp;
    *yp = 0;
    if (yb == 0) continue;
    // add v to the adjoint of (x, y, z), a neighbour of (i, j, k)
    auto put = [&](int x, int y, int z, double v) {
      if (GC) {
        const int nb = x < 1 ? 1 : x > nc ? 2 : y < 1 ? 3 : y > nc ? 4 : (D == 3 && z < 1) ? 5
                     : (D == 3 && z > nc) ? 6 : 0;
        if (nb > 0) {
          const int g[3] = {x, y, z};
          adj_ghost<D>(a, il, id, nc, nb, a.nbs[id * 6 + nb - 1], g, v, vbar);
          return;
        }
      }
      atomicAdd(L + cidx<D>(nc, x, y, z), v);
    };
    if (ev != nullptr) {
      // mg$D_box_gsrb_lpld: (sum c_m rf_m u_m - h^2 rhs) / sum c_m rf_m with
      // c_m = 2 eps_0 eps_m / (eps_0 + eps_m)
      const int c0 = cidx<D>(nc, i, j, k);
      int nbx[6];
      nb_cells<D>(nc, i, j, k, nbx);
      double c[6], s2 = 0;
      const double a0 = ev[c0];
#pragma unroll
      for (int m = 0; m < 2 * D; m++) {
        c[m] = 2 * a0 * ev[nbx[m]] / (a0 + ev[nbx[m]]);
        if (CYL && m < 2) {
          const double den = i + ioff - 0.5;
          c[m] = c[m] * ((double)(i + ioff - 1 + m) / den);
        }
        s2 = s2 + c[m];
      }
      const int di[6] = {-1, 1, 0, 0, 0, 0}, dj[6] = {0, 0, -1, 1, 0, 0};
      const int dk[6] = {0, 0, 0, 0, -1, 1};
#pragma unroll
      for (int m = 0; m < 2 * D; m++) put(i + di[m], j + dj[m], k + dk[m], c[m] / s2 * yb);
      R[c0] += -dx2 / s2 * yb;
      continue;
    }
    // (sum of the neighbours - h^2 rhs) / 2D, or with a Helmholtz term
    // divided by 2D + helm h^2
    const double wc = a.helm != 0 ? 1 / (2 * D + a.helm * dx2) : w;
    double wxm = wc, wxp = wc;
    if (CYL) {
      const double den = i + ioff - 0.5;
      wxm = wc * ((double)(i + ioff - 1) / den);
      wxp = wc * ((double)(i + ioff) / den);
    }
    put(i - 1, j, k, wxm * yb);
    put(i + 1, j, k, wxp * yb);
    put(i, j - 1, k, wc * yb);
    put(i, j + 1, k, wc * yb);
    if (D == 3) {
      put(i, j, k - 1, wc * yb);
      put(i, j, k + 1, wc * yb);
    }
    R[cidx<D>(nc, i, j, k)] += -wc * dx2 * yb;
  }
}

// dst += s A^T src for the box operator A (mg$D_box_lpl, m_aX_mg.f90:814-848;
// mg2_box_clpl, :1442-1464): src on the interior; then src = 0 on the
// interior when clear_src
template <int D, bool CYL>
__global__ __launch_bounds__(256) void k_adj_op(Dev a, const int* __restrict__ list, int n,
                                                int nc, int tpb, int isrc, double sgn, int idst,
                                                int clear_src) {
  BOX_PROLOGUE
  double* S = boxp(a, isrc, id);
  double* Dp = boxp(a, idst, id);
  const double* ev =
      (a.tag != nullptr && a.tag[id] > kTagNormal) ? boxp(a, a.i_eps, id) : nullptr;
  const double inv = 1 / (a.dr[id] * a.dr[id]);
  const int ioff = CYL ? (a.ix[id * 3] - 1) * nc : 0;
  const int total = D == 2 ? nc * nc : nc * nc * nc;
  for (int q = lt; q < total; q += tpb) {
    const int i = q % nc + 1, j = (q / nc) % nc + 1, k = D == 3 ? q / (nc * nc) + 1 : 0;
    const int c0 = cidx<D>(nc, i, j, k);
    const double sb = sgn * S[c0];
    if (clear_src) S[c0] = 0;
    if (sb == 0) continue;
    if (ev != nullptr) {
      // mg$D_box_lpld: 2 inv sum_m rf_m eps_0 eps_m / (eps_0 + eps_m) (u_m - u_0)
      int nbx[6];
      nb_cells<D>(nc, i, j, k, nbx);
      const double a0 = ev[c0];
      double sum = 0;
#pragma unroll
      for (int m = 0; m < 2 * D; m++) {
        double cm = a0 * ev[nbx[m]] / (a0 + ev[nbx[m]]);
        if (CYL && m < 2) {
          const double den = i + ioff - 0.5;
          cm = ((double)(i + ioff - 1 + m) / den) * cm;
        }
        atomicAdd(Dp + nbx[m], 2 * inv * cm * sb);
        sum = sum + cm;
      }
      atomicAdd(Dp + c0, -2 * inv * sum * sb);
      continue;
    }
    double wxm = inv, wxp = inv;
    if (CYL) {
      const double den = i + ioff - 0.5;
      wxm = ((double)(i + ioff - 1) / den) * inv;
      wxp = ((double)(i + ioff) / den) * inv;
    }
    atomicAdd(Dp + cidx<D>(nc, i - 1, j, k), wxm * sb);
    atomicAdd(Dp + cidx<D>(nc, i + 1, j, k), wxp * sb);
    atomicAdd(Dp + cidx<D>(nc, i, j - 1, k), inv * sb);
    atomicAdd(Dp + cidx<D>(nc, i, j + 1, k), inv * sb);
    if (D == 3) {
      atomicAdd(Dp + cidx<D>(nc, i, j, k - 1), inv * sb);
      atomicAdd(Dp + cidx<D>(nc, i, j, k + 1), inv * sb);
    }
    atomicAdd(Dp + c0, -(D == 2 ? 4.0 : 6.0) * inv * sb);
    if (a.helm != 0) atomicAdd(Dp + c0, -a.helm * sb);  // lpl(u) - helm u
  }
}

// mg$D_box_rstr_lpl (m_aX_mg.f90:851-893) of the children in list,
// transposed: the parent cells' adjoint in variable ip goes to the child's
// fine cells in variable idst (added, or assigned when assign) and is
// cleared
template <int D>
__global__ __launch_bounds__(256) void k_adj_rstr(Dev a, const int* __restrict__ list, int n,
                                                  int nc, int tpb, int ip, int idst,
                                                  int assign) {
  BOX_PROLOGUE
  const int pid = a.parent[id];
  const int pnc = a.ncell[pid];
  const int hnc = nc >> 1;
  int off[3];
  child_off<D>(a, id, nc, 0, off);
  double* Pp = boxp(a, ip, pid);
  double* C = boxp(a, idst, id);
  const int total = D == 2 ? hnc * hnc : hnc * hnc * hnc;
  const double w = D == 2 ? 0.25 : 0.125;
  for (int q = lt; q < total; q += tpb) {
    const int ci = q % hnc + 1, cj = (q / hnc) % hnc + 1, ck = D == 3 ? q / (hnc * hnc) + 1 : 0;
    double* pp = Pp + cidx<D>(pnc, off[0] + ci, off[1] + cj, D == 3 ? off[2] + ck : 0);
    const double v = w * *pp;
    *pp = 0;
    for (int f = 0; f < (D == 2 ? 4 : 8); f++) {
      const int fi = 2 * ci - 1 + (f & 1), fj = 2 * cj - 1 + ((f >> 1) & 1);
      const int fk = D == 3 ? 2 * ck - 1 + (f >> 2) : 0;
      double* cp = C + cidx<D>(nc, fi, fj, fk);
      *cp = assign ? v : *cp + v;
    }
  }
}

// mg$D_box_corr_lpl (m_aX_mg.f90:714-766) of the children in list,
// transposed: the child's potential adjoint (il, unchanged) is added to the
// adjoint of the parent's tmp (ilt) with the prolongation's weights
template <int D>
__global__ __launch_bounds__(256) void k_adj_corr(Dev a, const int* __restrict__ list, int n,
                                                  int nc, int tpb, int il, int ilt) {
  BOX_PROLOGUE
  const int pid = a.parent[id];
  const int pnc = a.ncell[pid];
  int off[3];
  child_off<D>(a, id, nc, 0, off);
  const double* F = boxp(a, il, id);
  double* Q = boxp(a, ilt, pid);
  // mg$D_box_corr_lpld for a child tagged with a permittivity (the parent's
  // eps weights the prolongation)
  const double* ep =
      (a.tag != nullptr && a.tag[id] > kTagNormal) ? boxp(a, a.i_eps, pid) : nullptr;
  const int total = D == 2 ? nc * nc : nc * nc * nc;
  for (int q = lt; q < total; q += tpb) {
    const int i = q % nc + 1, j = (q / nc) % nc + 1, k = D == 3 ? q / (nc * nc) + 1 : 0;
    const double fb = F[cidx<D>(nc, i, j, k)];
    if (fb == 0) continue;
    const int ic1 = off[0] + ((i + 1) >> 1), ic2 = ic1 + 1 - 2 * (i & 1);
    const int jc1 = off[1] + ((j + 1) >> 1), jc2 = jc1 + 1 - 2 * (j & 1);
    if (ep != nullptr) {
      // 2D: f += 1/2 sum_m (a0 u0 + a_m u_m) / (a0 + a_m); 3D: f += 1/3 sum_m
      // (a0 u0 + a_m (3/2 u_m - 1/2 u0)) / (a0 + a_m)
      const int kc1 = D == 3 ? off[2] + ((k + 1) >> 1) : 0;
      const int kc2 = D == 3 ? kc1 + 1 - 2 * (k & 1) : 0;
      const int p0 = cidx<D>(pnc, ic1, jc1, kc1);
      const int pm[3] = {cidx<D>(pnc, ic2, jc1, kc1), cidx<D>(pnc, ic1, jc2, kc1),
                         D == 3 ? cidx<D>(pnc, ic1, jc1, kc2) : 0};
      const double a0 = ep[p0];
      double w0 = 0;
#pragma unroll
      for (int m = 0; m < D; m++) {
        const double am = ep[pm[m]];
        atomicAdd(Q + pm[m], 0.5 * am / (a0 + am) * fb);
        w0 = w0 + (D == 2 ? 0.5 * a0 : (a0 - 0.5 * am) / 3.0) / (a0 + am);
      }
      atomicAdd(Q + p0, w0 * fb);
      continue;
    }
    if (D == 2) {
      atomicAdd(Q + cidx<D>(pnc, ic1, jc1, 0), 0.5 * fb);
      atomicAdd(Q + cidx<D>(pnc, ic2, jc1, 0), 0.25 * fb);
      atomicAdd(Q + cidx<D>(pnc, ic1, jc2, 0), 0.25 * fb);
    } else {
      const int kc1 = off[2] + ((k + 1) >> 1), kc2 = kc1 + 1 - 2 * (k & 1);
      atomicAdd(Q + cidx<D>(pnc, ic1, jc1, kc1), 0.25 * fb);
      atomicAdd(Q + cidx<D>(pnc, ic2, jc1, kc1), 0.25 * fb);
      atomicAdd(Q + cidx<D>(pnc, ic1, jc2, kc1), 0.25 * fb);
      atomicAdd(Q + cidx<D>(pnc, ic1, jc1, kc2), 0.25 * fb);
    }
  }
}

// whole-box transposes: mode 0, tmp = phi: lp += lt, lt = 0; mode 1,
// tmp = phi - tmp: lp += lt, lt = -lt; mode 2, rhs += tmp: lt += lr;
// mode 3, modes 0 and 2 in turn: lp += lt, lt = lr
__global__ __launch_bounds__(256) void k_adj_whole(Dev a, const int* __restrict__ list, int n,
                                                   int bl, int tpb, int mode, int il, int ilt,
                                                   int ilr) {
  BOX_PROLOGUE
  double* Lp = boxp(a, il, id);
  double* Lt = boxp(a, ilt, id);
  const double* Lr = boxp(a, ilr, id);
  for (int q = lt; q < bl; q += tpb) {
    if (mode == 2) {
      Lt[q] += Lr[q];
    } else if (mode == 3) {
      // mode 0 then mode 2 of the same cell: tmp = phi, then rhs += tmp
      Lp[q] += Lt[q];
      Lt[q] = Lr[q];
    } else {
      const double tb = Lt[q];
      Lp[q] += tb;
      Lt[q] = mode == 0 ? 0.0 : -tb;
    }
  }
}

// dst += src on the interior of the boxes in list; then src = 0 there when
// clear_src
template <int D>
__global__ __launch_bounds__(256) void k_adj_int_add(Dev a, const int* __restrict__ list, int n,
                                                     int nc, int tpb, int isrc, int idst,
                                                     int clear_src) {
  BOX_PROLOGUE
  double* S = boxp(a, isrc, id);
  double* Dp = boxp(a, idst, id);
  const int total = D == 2 ? nc * nc : nc * nc * nc;
  for (int q = lt; q < total; q += tpb) {
    const int i = q % nc + 1, j = (q / nc) % nc + 1, k = D == 3 ? q / (nc * nc) + 1 : 0;
    const int c0 = cidx<D>(nc, i, j, k);
    Dp[c0] += S[c0];
    if (clear_src) S[c0] = 0;
  }
}

namespace {
struct AdjMg {
  int lp, lt, lr, lw;
  double* vbar;
};
}  // namespace

template <int D>
static void adj_gc_level(Pool& P, const Dev& a, const AdjMg& m, int lvl) {
  const int n = P.n_ids(lvl);
  if (!n) return;
  const int nc = P.nc_lvl(lvl);
  const Geo g = geo(n, 2 * D * (D == 2 ? nc : nc * nc));
  hipLaunchKernelGGL(k_adj_gc<D>, dim3(g.grid), dim3(256), 0, P.stream, a, P.ids(lvl), n, nc,
                     g.tpb, m.lp, m.vbar);
  post_launch("adj_gc");
}

template <int D>
static void adj_gsrb_boxes(Pool& P, const Ctx& c, const Dev& a, const AdjMg& m, int lvl,
                           int n_cycle, bool gc_before = false) {
  const int n = P.n_ids(lvl);
  if (!n) return;
  const int nc = P.nc_lvl(lvl);
  const Geo g = geo(n, cells_of(D, nc, false) / 2);
  // forward: (half-sweep it, ghost pass) for it = 1 .. 2 n_cycle; the last
  // ghost pass is transposed on its own, each earlier one inside the
  // transpose of the half-sweep that follows it. gc_before: the caller's
  // ghost pass just before the call (the one after a correction) goes
  // inside the first half-sweep's transpose as well.
  adj_gc_level<D>(P, a, m, lvl);
  for (int it = 2 * n_cycle; it >= 1; it--) {
    const bool gc = it > 1 || gc_before;
    if (c.cyl) {
      if (gc)
        hipLaunchKernelGGL((k_adj_gsrb<2, true, true>), dim3(g.grid), dim3(256), 0, P.stream, a,
                           P.ids(lvl), n, nc, it, g.tpb, m.lp, m.lr, m.vbar);
      else
        hipLaunchKernelGGL((k_adj_gsrb<2, true>), dim3(g.grid), dim3(256), 0, P.stream, a,
                           P.ids(lvl), n, nc, it, g.tpb, m.lp, m.lr, m.vbar);
    } else {
      if (gc)
        hipLaunchKernelGGL((k_adj_gsrb<D, false, true>), dim3(g.grid), dim3(256), 0, P.stream,
                           a, P.ids(lvl), n, nc, it, g.tpb, m.lp, m.lr, m.vbar);
      else
        hipLaunchKernelGGL((k_adj_gsrb<D, false>), dim3(g.grid), dim3(256), 0, P.stream, a,
                           P.ids(lvl), n, nc, it, g.tpb, m.lp, m.lr, m.vbar);
    }
    post_launch("adj_gsrb");
  }
}

template <int D>
static void adj_op(Pool& P, const Ctx& c, const Dev& a, const int* list, int n, int nc,
                   int isrc, double sgn, int idst, bool clear) {
  if (!n) return;
  const Geo g = geo(n, cells_of(D, nc, false));
  if (c.cyl)
    hipLaunchKernelGGL((k_adj_op<2, true>), dim3(g.grid), dim3(256), 0, P.stream, a, list, n, nc,
                       g.tpb, isrc, sgn, idst, clear ? 1 : 0);
  else
    hipLaunchKernelGGL((k_adj_op<D, false>), dim3(g.grid), dim3(256), 0, P.stream, a, list, n,
                       nc, g.tpb, isrc, sgn, idst, clear ? 1 : 0);
  post_launch("adj_op");
}

static void adj_whole(Pool& P, const Dev& a, const int* list, int n, int nc, int mode,
                      const AdjMg& m) {
  if (!n) return;
  const int bl = cells_of(P.tree.dim, nc, true);
  const Geo g = geo(n, bl);
  hipLaunchKernelGGL(k_adj_whole, dim3(g.grid), dim3(256), 0, P.stream, a, list, n, bl, g.tpb,
                     mode, m.lp, m.lt, m.lr);
  post_launch("adj_whole");
}

template <int D>
static void adj_rstr(Pool& P, const Dev& a, int lvl, int ip, int idst, bool assign) {
  const int n = P.n_ids(lvl);
  if (!n) return;
  const int nc = P.nc_lvl(lvl);
  const Geo g = geo(n, cells_of(D, nc / 2, false));
  hipLaunchKernelGGL(k_adj_rstr<D>, dim3(g.grid), dim3(256), 0, P.stream, a, P.ids(lvl), n, nc,
                     g.tpb, ip, idst, assign ? 1 : 0);
  post_launch("adj_rstr");
}

template <int D>
static void adj_interior_add(Pool& P, const Dev& a, int lvl, int isrc, int idst, bool clear) {
  const int n = P.n_ids(lvl);
  if (!n) return;
  const int nc = P.nc_lvl(lvl);
  const Geo g = geo(n, cells_of(D, nc, false));
  hipLaunchKernelGGL(k_adj_int_add<D>, dim3(g.grid), dim3(256), 0, P.stream, a, P.ids(lvl), n,
                     nc, g.tpb, isrc, idst, clear ? 1 : 0);
  post_launch("adj_int_add");
}

// correct_children over the parents of lvl (m_aX_mg.f90:374-401), transposed
template <int D>
static void adj_correct_children(Pool& P, const Dev& a, const AdjMg& m, int lvl) {
  const int np = P.n_parents(lvl);
  if (!np) return;
  const int nch = P.n_ids(lvl + 1), cnc = P.nc_lvl(lvl + 1);
  const Geo g = geo(nch, cells_of(D, cnc, false));
  hipLaunchKernelGGL(k_adj_corr<D>, dim3(g.grid), dim3(256), 0, P.stream, a, P.ids(lvl + 1),
                     nch, cnc, g.tpb, m.lp, m.lt);
  post_launch("adj_corr");
  adj_whole(P, a, P.parents(lvl), np, P.nc_lvl(lvl), 1, m);
}

// update_coarse (m_aX_mg.f90:430-486; the fine tmp is restored there) and
// set_coarse_phi_rhs (:490-523; the fine tmp keeps the residual), transposed
template <int D>
static void adj_update_coarse(Pool& P, const Ctx& c, const Dev& a, const AdjMg& m, int lvl,
                              bool keep_res) {
  const int np = P.n_parents(lvl - 1);
  const int pnc = P.nc_lvl(lvl - 1);
  // tmp = phi (update_coarse only), then rhs += tmp
  adj_whole(P, a, P.parents(lvl - 1), np, pnc, keep_res ? 2 : 3, m);
  adj_op<D>(P, c, a, P.parents(lvl - 1), np, pnc, m.lr, 1.0, m.lp, true);  // rhs = A phi
  adj_gc_level<D>(P, a, m, lvl - 1);
  const int n = P.n_ids(lvl), nc = P.nc_lvl(lvl);
  adj_rstr<D>(P, a, lvl, m.lp, m.lp, false);  // phi_p = R phi
  if (keep_res) {
    adj_rstr<D>(P, a, lvl, m.lt, m.lt, false);  // tmp_p = R tmp
    // tmp = rhs - A phi on the interior
    adj_op<D>(P, c, a, P.ids(lvl), n, nc, m.lt, -1.0, m.lp, false);
    adj_interior_add<D>(P, a, lvl, m.lt, m.lr, true);
  } else {
    adj_rstr<D>(P, a, lvl, m.lt, m.lw, true);  // tmp_p = R (rhs - A phi)
    adj_op<D>(P, c, a, P.ids(lvl), n, nc, m.lw, -1.0, m.lp, false);
    adj_interior_add<D>(P, a, lvl, m.lw, m.lr, false);
  }
  if (keep_res && lvl == P.tree.max_lvl) adj_gc_level<D>(P, a, m, lvl);
}

template <int D>
static void adj_vcycle(Pool& P, const Ctx& c, const Dev& a, const AdjMg& m, int max_lvl) {
  const int min_lvl = P.tree.min_lvl;
  for (int lvl = max_lvl; lvl >= min_lvl + 1; lvl--) {
    // the up-smoothing and the ghost pass after the correction before it
    adj_gsrb_boxes<D>(P, c, a, m, lvl, c.n_up, true);
    adj_correct_children<D>(P, a, m, lvl - 1);
  }
  adj_gsrb_boxes<D>(P, c, a, m, min_lvl, c.n_base);
  for (int lvl = min_lvl + 1; lvl <= max_lvl; lvl++) {
    adj_update_coarse<D>(P, c, a, m, lvl, false);
    adj_gsrb_boxes<D>(P, c, a, m, lvl, c.n_down);
  }
}

template <int D>
static void fmg_adjoint_t(Pool& P, const AdjMg& m) {
  const Ctx c = ctx64(P);
  // with the permittivity tags and eps when the operators use them
  const Dev a = dev_for(P, c);
  const int min_lvl = P.tree.min_lvl, max_lvl = P.tree.max_lvl;
  for (int lvl = max_lvl; lvl >= min_lvl; lvl--) {
    adj_vcycle<D>(P, c, a, m, lvl);
    if (lvl > min_lvl) {
      adj_gc_level<D>(P, a, m, lvl);
      adj_correct_children<D>(P, a, m, lvl - 1);
    }
    adj_whole(P, a, P.ids(lvl), P.n_ids(lvl), P.nc_lvl(lvl), 0, m);  // tmp = phi
  }
  for (int lvl = min_lvl + 1; lvl <= max_lvl; lvl++) adj_update_coarse<D>(P, c, a, m, lvl, true);
}

// Reverse mode of one FAS V-cycle up to max_lvl without the residual
// (mg$D_fas_vcycle, m_aX_mg.f90:208-255): adj_vcycle, with the checks,
// variables and vbar buffer of fmg_adjoint.
template <int D>
static void vcycle_adjoint_t(Pool& P, const AdjMg& m, int max_lvl) {
  const Ctx c = ctx64(P);
  const Dev a = dev_for(P, c);
  adj_vcycle<D>(P, c, a, m, max_lvl);
}

void vcycle_adjoint(Pool& P, int max_lvl, const int* adj, double* vbar_host) {
  if (!P.mg.ready) throw std::logic_error("vcycle_adjoint: you haven't called mg_init_mg");
  if (P.n_inst) throw std::invalid_argument("vcycle_adjoint: one run per tree only");
  mg_check(P);
  if (max_lvl < P.tree.min_lvl || max_lvl > P.tree.max_lvl)
    throw std::invalid_argument("vcycle_adjoint: max_lvl outside the tree");
  const int nv = P.tree.n_cc;
  for (int k = 0; k < 4; k++)
    if (adj[k] < 1 || adj[k] > nv) throw std::out_of_range("vcycle_adjoint: bad adjoint variable");
  if (!P.d_vbar) check(hipMalloc((void**)&P.d_vbar, 8 * sizeof(double)), "hipMalloc vbar");
  const AdjMg m{adj[0], adj[1], adj[2], adj[3], P.d_vbar};
  run_captured(P, fnv(call_key(10, adj[0], adj[1], max_lvl - P.tree.min_lvl),
                      fnv((uint64_t)adj[2], (uint64_t)adj[3])),
               [&] {
                 hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, P.stream,
                                    (unsigned long long*)P.d_vbar, 8, 0ull);
                 post_launch("fill");
                 if (P.tree.dim == 2) vcycle_adjoint_t<2>(P, m, max_lvl);
                 else vcycle_adjoint_t<3>(P, m, max_lvl);
               });
  check(hipMemcpyAsync(vbar_host, P.d_vbar, 2 * P.tree.dim * sizeof(double),
                       hipMemcpyDeviceToHost, P.stream),
        "vbar");
  check(hipStreamSynchronize(P.stream), "vcycle_adjoint sync");
}

void fmg_adjoint(Pool& P, const int* adj, double* vbar_host) {
  if (!P.mg.ready) throw std::logic_error("fmg_adjoint: you haven't called mg_init_mg");
  if (P.n_inst) throw std::invalid_argument("fmg_adjoint: one run per tree only");
  mg_check(P);
  const int nv = P.tree.n_cc;
  for (int k = 0; k < 4; k++)
    if (adj[k] < 1 || adj[k] > nv) throw std::out_of_range("fmg_adjoint: bad adjoint variable");
  // the adjoints of V in a buffer of their own, kept until release (a
  // captured graph holds its address)
  if (!P.d_vbar) check(hipMalloc((void**)&P.d_vbar, 8 * sizeof(double)), "hipMalloc vbar");
  double* d_vbar = P.d_vbar;
  const AdjMg m{adj[0], adj[1], adj[2], adj[3], d_vbar};
  // the launches depend on the tree, its lists and the variables only: one
  // graph per variable set, replayed on later calls
  run_captured(P, fnv(call_key(9, adj[0], adj[1], adj[2]), (uint64_t)adj[3]), [&] {
    hipLaunchKernelGGL(k_fill_u64, dim3(1), dim3(64), 0, P.stream, (unsigned long long*)d_vbar,
                       8, 0ull);
    post_launch("fill");
    if (P.tree.dim == 2) fmg_adjoint_t<2>(P, m);
    else fmg_adjoint_t<3>(P, m);
  });
  check(hipMemcpyAsync(vbar_host, d_vbar, 2 * P.tree.dim * sizeof(double), hipMemcpyDeviceToHost,
                       P.stream),
        "vbar");
  check(hipStreamSynchronize(P.stream), "fmg_adjoint sync");
}

}  // namespace a5g
