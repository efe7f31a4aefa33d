// Streamer physics kernels of the a5g library: space-charge right-hand side,
// field from potential, Koren-limited drift-diffusion fluxes, consistent
// fluxes at refinement boundaries, the density update with the
// impact-ionisation source, trapezoidal averaging, per-box refinement
// summaries and prolongation to new boxes. They restate the per-box
// procedures of src/streamer_{2d,3d,cyl}.f90 with the reference's operand
// order (the library is built with -ffp-contract=off), so the FP64 results
// equal the reference's bit for bit.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "a5g_internal.hpp"
#include "a5g_kernels.hpp"

namespace a5g {

#define BOXES_PER_BLOCK                                     \
  const int bpb_ = blockDim.x / tpb;                        \
  const int slot_ = threadIdx.x / tpb;                      \
  const int lt = threadIdx.x - slot_ * tpb;                 \
  const int li_ = blockIdx.x * bpb_ + slot_;                \
  const bool act = li_ < n;                                 \
  const int id = act ? list[li_] : 0;

static int pow2c(int v) {
  int p = 1;
  while (p < v) p <<= 1;
  return p;
}
struct G2 {
  int tpb, grid;
};
static G2 g2(int n, int work) {
  int tpb = std::min(256, std::max(64, pow2c(work)));
  return {tpb, (n + 256 / tpb - 1) / (256 / tpb)};
}

// ---------------------------------------------------------------------------
// lookup table (fosito/m_lookup_table.f90:100-161): the location is
// low_ix = ceiling((x - x_min) inv_dx), low_frac = low_ix - frac, clamped to
// the table; a column value is low_frac t(low_ix) + (1 - low_frac) t(low_ix+1)
// ---------------------------------------------------------------------------

struct Tbl {
  const double* rc = nullptr;  // [col][row], 0-based
  int n_rows = 0;
  double x_min = 0, inv_dx = 0;
};

__device__ __forceinline__ void lt_loc(const Tbl& t, double x, int& li, double& lf) {
  const double frac = (x - t.x_min) * t.inv_dx;
  li = (int)ceil(frac);
  lf = li - frac;
  if (li < 1) {
    li = 1;
    lf = 1;
  } else if (li >= t.n_rows) {
    li = t.n_rows - 1;
    lf = 0;
  }
}

__device__ __forceinline__ double lt_col(const Tbl& t, int col, int li, double lf) {
  const double* c = t.rc + (size_t)(col - 1) * t.n_rows;
  return lf * c[li - 1] + (1 - lf) * c[li];
}

// ---------------------------------------------------------------------------
// rhs = fac (n_e - n_i) on whole leaf boxes (compute_fld,
// src/streamer_2d.f90:318-329)
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(256) void k_space_charge(Dev a, const int* __restrict__ list, int n,
                                                      int bl, int tpb, int ie, int ip, int ir,
                                                      double fac) {
  BOXES_PER_BLOCK
  if (!act) return;
  const double* e = boxp(a, ie, id);
  const double* p = boxp(a, ip, id);
  double* r = boxp(a, ir, id);
  for (int q = lt; q < bl; q += tpb) r[q] = fac * (e[q] - p[q]);
}

// ---------------------------------------------------------------------------
// fld_from_pot (src/streamer_2d.f90:346-381, src/streamer_3d.f90:348-395):
// face fields, with the permittivity-weighted form on the box faces, then the
// cell field norm
// ---------------------------------------------------------------------------

template <int D>
__global__ __launch_bounds__(256) void k_fld(Dev a, const int* __restrict__ list, int n, int nc,
                                             int tpb, int iphi, int ieps, int ifld, int ffld) {
  BOXES_PER_BLOCK
  const int per = (nc + 1) * (D == 2 ? nc : nc * nc);
  if (act) {
    const double inv_dr = 1 / a.dr[id];
    const double* p = boxp(a, iphi, id);
    const double* e = boxp(a, ieps, id);
    for (int q = lt; q < D * per; q += tpb) {
      const int d = q / per, w = q - d * per;
      const int n0 = nc + (d == 0), n1 = nc + (d == 1);
      const int i = w % n0 + 1, j = (w / n0) % n1 + 1, k = D == 3 ? w / (n0 * n1) + 1 : 0;
      int lo[3] = {i, j, k};
      lo[d] -= 1;
      const int ia = cidx<D>(nc, lo[0], lo[1], lo[2]);
      const int ib = cidx<D>(nc, i, j, k);
      const int fpos = d == 0 ? i : (d == 1 ? j : k);
      double v;
      if (fpos == 1) v = 2 * inv_dr * (p[ia] - p[ib]) * e[ia] / (e[ib] + e[ia]);
      else if (fpos == nc + 1) v = 2 * inv_dr * (p[ia] - p[ib]) * e[ib] / (e[ib] + e[ia]);
      else v = inv_dr * (p[ia] - p[ib]);
      facep(a, ffld, id, d, D)[w] = v;
    }
  }
  __syncthreads();
  if (!act) return;
  double* fl = boxp(a, ifld, id);
  const double* F0 = facep(a, ffld, id, 0, D);
  const double* F1 = facep(a, ffld, id, 1, D);
  const double* F2 = D == 3 ? facep(a, ffld, id, 2, D) : nullptr;
  const int total = D == 2 ? nc * nc : nc * nc * nc;
  for (int q = lt; q < total; q += tpb) {
    const int i = q % nc + 1, j = (q / nc) % nc + 1, k = D == 3 ? q / (nc * nc) + 1 : 0;
    const double sx = F0[fidx<D>(nc, 0, i, j, k)] + F0[fidx<D>(nc, 0, i + 1, j, k)];
    const double sy = F1[fidx<D>(nc, 1, i, j, k)] + F1[fidx<D>(nc, 1, i, j + 1, k)];
    if (D == 2) {
      fl[cidx<D>(nc, i, j, k)] = sqrt(0.25 * (sx * sx) + 0.25 * (sy * sy));
    } else {
      const double sz = F2[fidx<D>(nc, 2, i, j, k)] + F2[fidx<D>(nc, 2, i, j, k + 1)];
      fl[cidx<D>(nc, i, j, k)] = 0.5 * sqrt(sx * sx + sy * sy + sz * sz);
    }
  }
}

// ---------------------------------------------------------------------------
// Koren limiter, src/m_streamer.f90:339-362
// ---------------------------------------------------------------------------

__device__ __forceinline__ double koren_mlim(double x, double y) {
  const double sixth = 1 / 6.0;
  const double aa = x * x, ab = x * y;
  if (ab <= 0) return 0;
  if (aa <= 0.25 * ab) return x;
  if (aa <= 2.5 * ab) return sixth * (y + 2 * x);
  return y;
}

// Second ghost layer value beyond face nb at tangential cell (t1, t2) of box
// id: a$D_gc2_box with a$D_gc2_prolong1 / a$D_gc2_neumann
// (afivo/src/m_aX_gc.f90:407-622)
template <int D>
__device__ double gc2_value(const Dev& a, int iv, int id, int nc, int nb, int t1, int t2) {
  const int d = d_nb_dim(nb);
  const int ta = d == 0 ? 1 : 0, tb = d == 2 ? 1 : 2;
  const bool low = d_nb_low(nb);
  const int nb_id = a.nbs[id * 6 + nb - 1];
  int c[3] = {0, 0, 0};
  c[ta] = t1;
  if (D == 3) c[tb] = t2;
  if (nb_id > 0) {
    c[d] = low ? nc - 1 : 2;
    return boxp(a, iv, nb_id)[cidx<D>(nc, c[0], c[1], c[2])];
  }
  if (nb_id < 0) {
    c[d] = low ? 2 : nc - 1;
    return boxp(a, iv, id)[cidx<D>(nc, c[0], c[1], c[2])];
  }
  const int pid = a.parent[id];
  const int pn = a.nbs[pid * 6 + nb - 1];
  const int pnc = a.ncell[pn];
  int off[3];
  child_off<D>(a, id, nc, nb, off);
  const double* q = boxp(a, iv, pn);
  const int ix = low ? -1 : nc + 2;
  const int n1 = off[d] + ((ix + 1) >> 1), n2 = n1 + 1 - 2 * (ix & 1);
  const int c1 = off[ta] + ((t1 + 1) >> 1), c2 = c1 + 1 - 2 * (t1 & 1);
  if (D == 2) {
    int s0[3], s1[3], s2[3];
    s0[d] = n1; s0[ta] = c1;
    s1[d] = n2; s1[ta] = c1;
    s2[d] = n1; s2[ta] = c2;
    // the shifted cells enter in the order dim 0 then dim 1
    const double v0 = q[cidx<D>(pnc, s0[0], s0[1], 0)];
    double va = q[cidx<D>(pnc, s1[0], s1[1], 0)];
    double vb = q[cidx<D>(pnc, s2[0], s2[1], 0)];
    if (d == 1) {
      const double tmp = va;
      va = vb;
      vb = tmp;
    }
    return 0.5 * v0 + 0.25 * va + 0.25 * vb;
  }
  const int e1 = off[tb] + ((t2 + 1) >> 1), e2 = e1 + 1 - 2 * (t2 & 1);
  int P[4][3];
  for (int m = 0; m < 4; m++) {
    P[m][d] = n1;
    P[m][ta] = c1;
    P[m][tb] = e1;
  }
  P[1][d] = n2;
  P[2][ta] = c2;
  P[3][tb] = e2;
  // sum order: base, then the cell shifted in dim 0, dim 1, dim 2
  double v[4];
  v[0] = q[cidx<D>(pnc, P[0][0], P[0][1], P[0][2])];
  int m = 1;
  for (int dd = 0; dd < 3; dd++) {
    const int w = dd == d ? 1 : (dd == ta ? 2 : 3);
    v[m++] = q[cidx<D>(pnc, P[w][0], P[w][1], P[w][2])];
  }
  return 0.25 * v[0] + 0.25 * v[1] + 0.25 * v[2] + 0.25 * v[3];
}

// fluxes_koren (src/streamer_2d.f90:384-489, src/streamer_3d.f90:396-554,
// src/streamer_cyl.f90:416-521) on leaves: one lane per face
template <int D>
__global__ __launch_bounds__(256) void k_koren(Dev a, Tbl tb, const int* __restrict__ list,
                                               int n, int nc, int tpb, int ie, int ifld,
                                               int fel, int ffld, int cmob, int cdif,
                                               double fac_all) {
  BOXES_PER_BLOCK
  if (!act) return;
  // the clamp factor of the box's own run when the tree holds several
  const double fac = a.n_inst ? a.inst_val[inst_of<D>(a, id) * kInstVals + kInstKorenFac]
                              : fac_all;
  const double inv_dr = 1 / a.dr[id];
  const double* ne = boxp(a, ie, id);
  const double* E = boxp(a, ifld, id);
  const int per = (nc + 1) * (D == 2 ? nc : nc * nc);
  for (int q = lt; q < D * per; q += tpb) {
    const int d = q / per, w = q - d * per;
    const int n0 = nc + (d == 0), n1 = nc + (d == 1);
    int c[3];
    c[0] = w % n0 + 1;
    c[1] = (w / n0) % n1 + 1;
    c[2] = D == 3 ? w / (n0 * n1) + 1 : 0;
    int m1[3] = {c[0], c[1], c[2]};
    m1[d] -= 1;
    const int pos = c[d];
    const int ic = cidx<D>(nc, c[0], c[1], c[2]);
    const int im = cidx<D>(nc, m1[0], m1[1], m1[2]);
    const double fld_avg = 0.5 * (E[ic] + E[im]);
    int li;
    double lf;
    lt_loc(tb, fld_avg, li, lf);
    const double mob = lt_col(tb, cmob, li, lf);
    const double dif = lt_col(tb, cdif, li, lf);
    const double fld = facep(a, ffld, id, d, D)[w];
    const double vd = -mob * fld;
    const double n0v = ne[ic], nm1 = ne[im];
    const double gradc = n0v - nm1;
    const int ta = d == 0 ? 1 : 0, tbb = d == 2 ? 1 : 2;
    double f;
    if (vd < 0.0) {
      double tmp;
      if (pos == nc + 1) {
        tmp = gc2_value<D>(a, ie, id, nc, 2 * d + 2, c[ta], D == 3 ? c[tbb] : 0);
      } else {
        int p1[3] = {c[0], c[1], c[2]};
        p1[d] += 1;
        tmp = ne[cidx<D>(nc, p1[0], p1[1], p1[2])];
      }
      const double gradn = tmp - n0v;
      f = vd * (n0v - koren_mlim(gradc, gradn));
      if (f < fac * fld) f = fac * fld;
    } else {
      double tmp;
      if (pos == 1) {
        tmp = gc2_value<D>(a, ie, id, nc, 2 * d + 1, c[ta], D == 3 ? c[tbb] : 0);
      } else {
        int m2[3] = {c[0], c[1], c[2]};
        m2[d] -= 2;
        tmp = ne[cidx<D>(nc, m2[0], m2[1], m2[2])];
      }
      const double gradp = nm1 - tmp;
      f = vd * (nm1 + koren_mlim(gradc, gradp));
      if (f > fac * fld) f = fac * fld;
    }
    facep(a, fel, id, d, D)[w] = f - dif * gradc * inv_dr;
  }
}

// a$D_consistent_fluxes / flux_from_children (afivo/src/m_aX_core.f90:818-930):
// one lane per coarse face cell of a (parent, direction) pair
template <int D>
__global__ __launch_bounds__(256) void k_flux_fix(Dev a, const int2* __restrict__ pairs, int n,
                                                  int nc, int fel) {
  const int hnc = nc >> 1;
  const int per = D == 2 ? nc : nc * nc;  // coarse face cells per pair
  const long total = (long)n * per;
  for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < total;
       q += (long)gridDim.x * blockDim.x) {
    const int pi = (int)(q / per), w = (int)(q - (long)pi * per);
    const int id = pairs[pi].x, nb = pairs[pi].y;
    const int d = d_nb_dim(nb);
    const int ta = d == 0 ? 1 : 0, tb = d == 2 ? 1 : 2;
    const int nb_id = a.nbs[id * 6 + nb - 1];
    const bool low = d_nb_low(nb);
    const int ii = low ? 1 : nc + 1, inb = low ? nc + 1 : 1;
    // coarse face cell (a_, b_) in 1..nc, and the child that covers it
    const int a_ = w % nc + 1, b_ = D == 3 ? w / nc + 1 : 1;
    const int ca = (a_ - 1) / hnc, cb = D == 3 ? (b_ - 1) / hnc : 0;
    int cdix[3] = {0, 0, 0};
    cdix[d] = low ? 0 : 1;
    cdix[ta] = ca;
    if (D == 3) cdix[tb] = cb;
    const int ich = 1 + cdix[0] + 2 * cdix[1] + 4 * cdix[2];
    const int cid = a.children[id * 8 + ich - 1];
    const int sa = 2 * (a_ - ca * hnc) - 1, sb = D == 3 ? 2 * (b_ - cb * hnc) - 1 : 1;
    const double* src = facep(a, fel, cid, d, D);
    double* dst = facep(a, fel, nb_id, d, D);
    int s[3], t[3];
    s[d] = ii;
    s[ta] = sa;
    t[d] = inb;
    t[ta] = a_;
    if (D == 2) {
      const double v0 = src[fidx<D>(nc, d, s[0], s[1], 0)];
      s[ta] += 1;
      const double v1 = src[fidx<D>(nc, d, s[0], s[1], 0)];
      dst[fidx<D>(nc, d, t[0], t[1], 0)] = 0.5 * (v0 + v1);
    } else {
      s[tb] = sb;
      t[tb] = b_;
      const double v0 = src[fidx<D>(nc, d, s[0], s[1], s[2])];
      s[ta] += 1;
      const double v1 = src[fidx<D>(nc, d, s[0], s[1], s[2])];
      s[ta] -= 1;
      s[tb] += 1;
      const double v2 = src[fidx<D>(nc, d, s[0], s[1], s[2])];
      s[ta] += 1;
      const double v3 = src[fidx<D>(nc, d, s[0], s[1], s[2])];
      dst[fidx<D>(nc, d, t[0], t[1], t[2])] = 0.25 * (v0 + v1 + v2 + v3);
    }
  }
}

// NORM2 as the flang runtime evaluates it (running maximum m, scaled sum)
template <int N>
__device__ __forceinline__ double norm2_flang(const double* x) {
  double mx = 0, s = 0;
#pragma unroll
  for (int i = 0; i < N; i++) {
    const double ax = fabs(x[i]);
    if (mx == 0) {
      mx = ax;
    } else if (ax > mx) {
      const double q = mx / ax, qq = q * q;
      s = s * qq;
      s = s + qq;
      mx = ax;
    } else {
      const double q = ax / mx;
      s = s + q * q;
    }
  }
  return mx * sqrt(1 + s);
}

// update_solution (src/streamer_2d.f90:499-542, src/streamer_3d.f90:566-615,
// src/streamer_cyl.f90:531-570) on leaves: one lane per cell
template <int D, bool CYL>
__global__ __launch_bounds__(256) void k_update(Dev a, Tbl tb, const int* __restrict__ list,
                                                int n, int nc, int tpb, int ie, int ip,
                                                int ifld, int ieps, int ipho, int fel,
                                                int calpha, int ceta, int diel,
                                                double dt_all) {
  BOXES_PER_BLOCK
  if (!act) return;
  // the time step of the box's own run when the tree holds several
  const double dt = a.n_inst ? a.inst_val[inst_of<D>(a, id) * kInstVals + kInstDt] : dt_all;
  const double inv_dr = 1 / a.dr[id];
  double* ne = boxp(a, ie, id);
  double* ni = boxp(a, ip, id);
  const double* E = boxp(a, ifld, id);
  const double* eps = boxp(a, ieps, id);
  const double* ph = ipho > 0 ? boxp(a, ipho, id) : nullptr;
  const double* F0 = facep(a, fel, id, 0, D);
  const double* F1 = facep(a, fel, id, 1, D);
  const double* F2 = D == 3 ? facep(a, fel, id, 2, D) : nullptr;
  const int ioff = CYL ? (a.ix[id * 3] - 1) * nc : 0;
  const int total = D == 2 ? nc * nc : nc * nc * nc;
  for (int q = lt; q < total; q += tpb) {
    const int i = q % nc + 1, j = (q / nc) % nc + 1, k = D == 3 ? q / (nc * nc) + 1 : 0;
    const int c = cidx<D>(nc, i, j, k);
    int li;
    double lf;
    lt_loc(tb, E[c], li, lf);
    const double alpha = lt_col(tb, calpha, li, lf);
    const double eta = lt_col(tb, ceta, li, lf);
    const double fxl = F0[fidx<D>(nc, 0, i, j, k)], fxh = F0[fidx<D>(nc, 0, i + 1, j, k)];
    const double fyl = F1[fidx<D>(nc, 1, i, j, k)], fyh = F1[fidx<D>(nc, 1, i, j + 1, k)];
    double src, sflux;
    if (D == 2) {
      const double df[2] = {fxl + fxh, fyl + fyh};
      src = 0.5 * norm2_flang<2>(df) * (alpha - eta);
      if (ph) src = src + ph[c];
      if (CYL) {
        const double den = i + ioff - 0.5;
        const double rf1 = (double)(i + ioff - 1) / den, rf2 = (double)(i + ioff) / den;
        sflux = (fyl - fyh + rf1 * fxl - rf2 * fxh) * inv_dr;
      } else {
        sflux = (fxl - fxh + fyl - fyh) * inv_dr;
      }
    } else {
      const double fzl = F2[fidx<D>(nc, 2, i, j, k)], fzh = F2[fidx<D>(nc, 2, i, j, k + 1)];
      const double df[3] = {fxl + fxh, fyl + fyh, fzl + fzh};
      src = 0.5 * norm2_flang<3>(df) * (alpha - eta);
      if (ph) src = src + ph[c];
      sflux = (fxl - fxh + fyl - fyh + fzl - fzh) * inv_dr;
    }
    double e = ne[c] + (src + sflux) * dt;
    double p = ni[c] + src * dt;
    if (diel && eps[c] > 1.0) {
      p = p - e;
      e = 0;
    }
    ne[c] = e;
    ni[c] = p;
  }
}

// average_dens (src/streamer_2d.f90:492-496): whole boxes
__global__ __launch_bounds__(256) void k_average(Dev a, const int* __restrict__ list, int n,
                                                 int bl, int tpb, int ie, int ieo, int ip,
                                                 int ipo) {
  BOXES_PER_BLOCK
  if (!act) return;
  double* e = boxp(a, ie, id);
  const double* eo = boxp(a, ieo, id);
  double* p = boxp(a, ip, id);
  const double* po = boxp(a, ipo, id);
  for (int q = lt; q < bl; q += tpb) {
    e[q] = 0.5 * (e[q] + eo[q]);
    p[q] = 0.5 * (p[q] + po[q]);
  }
}

// out[0..n) = 0
__global__ void k_zero(double* __restrict__ out, size_t n) {
  for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n;
       q += (size_t)gridDim.x * blockDim.x)
    out[q] = 0.0;
}

// ---------------------------------------------------------------------------
// Reverse mode of one transport substep on a frozen mesh: update_solution,
// a$D_consistent_fluxes and fluxes_koren (no counterpart in the reference,
// which has no derivatives). Each kernel takes the adjoints of its outputs
// and adds those of its inputs, taking every branch (upwind side, limiter
// case, flux clamp, table row, dielectric rule) the forward pass takes on
// the same data. The adjoint variables: cell variables le, li (densities),
// lE (|E|), lpho (photoionization source); face variables lF (electron flux),
// lfld (field); table rows in tg[col][row].
// ---------------------------------------------------------------------------

// d/dx of a table column at x (lt_col at lt_loc(x)): the row interval's
// slope, 0 where the location is clamped to the table ends
__device__ __forceinline__ double lt_dcol(const Tbl& t, int col, double x, int li) {
  const double frac = (x - t.x_min) * t.inv_dx;
  const int raw = (int)ceil(frac);
  if (raw < 1 || raw >= t.n_rows) return 0.0;
  const double* c = t.rc + (size_t)(col - 1) * t.n_rows;
  return (c[li] - c[li - 1]) * t.inv_dx;
}

__device__ __forceinline__ void lt_col_adj(const Tbl& t, double* tg, int col, int li, double lf,
                                           double lam) {
  if (lam == 0.0) return;
  double* g = tg + (size_t)(col - 1) * t.n_rows;
  atomicAdd(g + li - 1, lf * lam);
  atomicAdd(g + li, (1 - lf) * lam);
}

// Table-row adjoints of two columns (c0, c1) gathered per workgroup: every
// cell of a level feeds the few rows its field falls in, so global atomics
// from all lanes would meet on a handful of addresses. With the two
// columns' rows in LDS (lds, 2 n_rows doubles: kTgLdsRows rows at most),
// the lanes add there, each lane keeps the range of rows it touched, and
// the workgroup adds the touched range to the global rows once at the end
// (tg_flush). Without LDS room the rows are added to global memory
// directly.
constexpr int kTgLdsRows = 2048;
struct TgAcc {
  double* lds;  // nullptr: global atomics
  double* tg;
  int n_rows, c0, c1;
  int lo = 1 << 30, hi = -1;  // rows this lane touched (0-based)
  __device__ void add(int which, int li, double lf, double lam) {
    if (lam == 0.0) return;
    if (!lds) {
      double* g = tg + (size_t)((which ? c1 : c0) - 1) * n_rows;
      atomicAdd(g + li - 1, lf * lam);
      atomicAdd(g + li, (1 - lf) * lam);
      return;
    }
    double* g = lds + which * n_rows;
    atomicAdd(g + li - 1, lf * lam);
    atomicAdd(g + li, (1 - lf) * lam);
    lo = min(lo, li - 1);
    hi = max(hi, li);
  }
};

// zero the LDS rows and the workgroup's row range (all lanes, then a barrier)
__device__ __forceinline__ void tg_begin(double* lds, int n_rows, int* range) {
  if (lds)
    for (int q = threadIdx.x; q < 2 * n_rows; q += blockDim.x) lds[q] = 0.0;
  if (threadIdx.x == 0) {
    range[0] = 1 << 30;
    range[1] = -1;
  }
  __syncthreads();
}

// the workgroup's touched rows to global memory (all lanes)
__device__ __forceinline__ void tg_flush(TgAcc& acc, int* range) {
  if (!acc.lds) return;
  if (acc.hi >= 0) {
    atomicMin(range, acc.lo);
    atomicMax(range + 1, acc.hi);
  }
  __syncthreads();
  const int lo = range[0], hi = range[1];
  const int w = hi - lo + 1;
  for (int q = threadIdx.x; q < 2 * max(w, 0); q += blockDim.x) {
    const int which = q / w, r = lo + q - which * w;
    const double v = acc.lds[which * acc.n_rows + r];
    if (v != 0.0) atomicAdd(acc.tg + (size_t)((which ? acc.c1 : acc.c0) - 1) * acc.n_rows + r, v);
  }
}

// update_solution reverse: one lane per leaf cell. le, li hold the adjoints
// of n_e', n_i' on entry and those of n_e, n_i on exit; each face of lF gets
// the contributions of the (one or two) cells beside it, so with lF zero on
// entry the result does not depend on their order.
template <int D, bool CYL>
__device__ void update_adj_box(const Dev& a, const Tbl& tb, int id, int nc, int tpb, int lt,
                               int ifld, int ieps, int fel, int calpha, int ceta, int diel,
                               int photoi, int le, int li, int lE, int lph, int lF, TgAcc& acc,
                               double dt_all);

template <int D, bool CYL>
__global__ __launch_bounds__(256) void k_update_adj(Dev a, Tbl tb, const int* __restrict__ list,
                                                    int n, int nc, int tpb, int ifld, int ieps,
                                                    int fel, int calpha, int ceta, int diel,
                                                    int photoi, int le, int li, int lE, int lph,
                                                    int lF, double* __restrict__ tg,
                                                    double dt_all, int use_lds) {
  extern __shared__ double s_tg[];
  __shared__ int s_range[2];
  TgAcc acc{use_lds ? s_tg : nullptr, tg, tb.n_rows, calpha, ceta};
  tg_begin(acc.lds, tb.n_rows, s_range);
  const int bpb = blockDim.x / tpb;
  const int slot = threadIdx.x / tpb;
  const int lt = threadIdx.x - slot * tpb;
  const int n_groups = (n + bpb - 1) / bpb;
  for (int grp = blockIdx.x; grp < n_groups; grp += gridDim.x) {
    const int li_ = grp * bpb + slot;
    if (li_ < n) update_adj_box<D, CYL>(a, tb, list[li_], nc, tpb, lt, ifld, ieps, fel, calpha,
                                        ceta, diel, photoi, le, li, lE, lph, lF, acc, dt_all);
  }
  tg_flush(acc, s_range);
}

template <int D, bool CYL>
__device__ void update_adj_box(const Dev& a, const Tbl& tb, int id, int nc, int tpb, int lt,
                               int ifld, int ieps, int fel, int calpha, int ceta, int diel,
                               int photoi, int le, int li, int lE, int lph, int lF, TgAcc& acc,
                               double dt_all) {
  const double dt = a.n_inst ? a.inst_val[inst_of<D>(a, id) * kInstVals + kInstDt] : dt_all;
  const double inv_dr = 1 / a.dr[id];
  const double* E = boxp(a, ifld, id);
  const double* eps = boxp(a, ieps, id);
  double* Le = boxp(a, le, id);
  double* Li = boxp(a, li, id);
  double* LEf = boxp(a, lE, id);
  double* Lph = photoi ? boxp(a, lph, id) : nullptr;
  const double* F[3];
  double* LF[3];
  for (int d = 0; d < D; d++) {
    F[d] = facep(a, fel, id, d, D);
    LF[d] = facep(a, lF, id, d, D);
  }
  const int ioff = CYL ? (a.ix[id * 3] - 1) * nc : 0;
  const int total = D == 2 ? nc * nc : nc * nc * nc;
  for (int q = lt; q < total; q += tpb) {
    const int i = q % nc + 1, j = (q / nc) % nc + 1, k = D == 3 ? q / (nc * nc) + 1 : 0;
    const int c = cidx<D>(nc, i, j, k);
    double le_pre = Le[c], lp_pre = Li[c];
    if (diel && eps[c] > 1.0) le_pre = -lp_pre;  // n_i' = n_i* - n_e*, n_e' = 0
    Le[c] = le_pre;
    Li[c] = lp_pre;
    const double lsrc = le_pre * dt + lp_pre * dt;
    const double s = le_pre * dt * inv_dr;
    int ix;
    double lf;
    lt_loc(tb, E[c], ix, lf);
    const double alpha = lt_col(tb, calpha, ix, lf);
    const double eta = lt_col(tb, ceta, ix, lf);
    int fl[3], fh[3];
    double df[3] = {0, 0, 0};
    for (int d = 0; d < D; d++) {
      int h[3] = {i, j, k};
      h[d] += 1;
      fl[d] = fidx<D>(nc, d, i, j, k);
      fh[d] = fidx<D>(nc, d, h[0], h[1], h[2]);
      df[d] = F[d][fl[d]] + F[d][fh[d]];
    }
    const double nrm = D == 2 ? norm2_flang<2>(df) : norm2_flang<3>(df);
    const double lnrm = lsrc * 0.5 * (alpha - eta);
    const double lal = lsrc * 0.5 * nrm, let = -lsrc * 0.5 * nrm;
    if (Lph) Lph[c] += lsrc;
    LEf[c] += lal * lt_dcol(tb, calpha, E[c], ix) + let * lt_dcol(tb, ceta, E[c], ix);
    acc.add(0, ix, lf, lal);
    acc.add(1, ix, lf, let);
    for (int d = 0; d < D; d++) {
      const double ldf = nrm > 0 ? lnrm * df[d] / nrm : 0.0;
      double wl = 1.0, wh = 1.0;
      if (CYL && d == 0) {
        const double den = i + ioff - 0.5;
        wl = (double)(i + ioff - 1) / den;
        wh = (double)(i + ioff) / den;
      }
      atomicAdd(LF[d] + fl[d], ldf + wl * s);
      atomicAdd(LF[d] + fh[d], ldf - wh * s);
    }
  }
}

// consistent fluxes reverse: one lane per coarse face cell of a (parent,
// direction) pair; its adjoint goes to the fine faces it averaged and is
// cleared (each fine face feeds one coarse face)
template <int D>
__global__ __launch_bounds__(256) void k_flux_fix_adj(Dev a, const int2* __restrict__ pairs,
                                                      int n, int nc, int lF) {
  const int hnc = nc >> 1;
  const int per = D == 2 ? nc : nc * nc;
  const long total = (long)n * per;
  for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < total;
       q += (long)gridDim.x * blockDim.x) {
    const int pi = (int)(q / per), w = (int)(q - (long)pi * per);
    const int id = pairs[pi].x, nb = pairs[pi].y;
    const int d = d_nb_dim(nb);
    const int ta = d == 0 ? 1 : 0, tb = d == 2 ? 1 : 2;
    const int nb_id = a.nbs[id * 6 + nb - 1];
    const bool low = d_nb_low(nb);
    const int ii = low ? 1 : nc + 1, inb = low ? nc + 1 : 1;
    const int a_ = w % nc + 1, b_ = D == 3 ? w / nc + 1 : 1;
    const int ca = (a_ - 1) / hnc, cb = D == 3 ? (b_ - 1) / hnc : 0;
    int cdix[3] = {0, 0, 0};
    cdix[d] = low ? 0 : 1;
    cdix[ta] = ca;
    if (D == 3) cdix[tb] = cb;
    const int ich = 1 + cdix[0] + 2 * cdix[1] + 4 * cdix[2];
    const int cid = a.children[id * 8 + ich - 1];
    const int sa = 2 * (a_ - ca * hnc) - 1, sb = D == 3 ? 2 * (b_ - cb * hnc) - 1 : 1;
    double* src = facep(a, lF, cid, d, D);
    double* dst = facep(a, lF, nb_id, d, D);
    int s[3] = {0, 0, 0}, t[3] = {0, 0, 0};
    s[d] = ii;
    s[ta] = sa;
    t[d] = inb;
    t[ta] = a_;
    if (D == 3) {
      s[tb] = sb;
      t[tb] = b_;
    }
    const int g = fidx<D>(nc, d, t[0], t[1], t[2]);
    const double lam = dst[g];
    dst[g] = 0;
    if (D == 2) {
      src[fidx<D>(nc, d, s[0], s[1], 0)] += 0.5 * lam;
      s[ta] += 1;
      src[fidx<D>(nc, d, s[0], s[1], 0)] += 0.5 * lam;
    } else {
      for (int m = 0; m < 4; m++) {
        int s0[3] = {s[0], s[1], s[2]};
        s0[ta] += m & 1;
        s0[tb] += m >> 1;
        src[fidx<D>(nc, d, s0[0], s0[1], s0[2])] += 0.25 * lam;
      }
    }
  }
}

// the adjoint of one second-layer ghost value of gc2_value: lam goes to the
// cells it was taken from (a neighbour's, the box's own, or the parent's
// neighbour's)
template <int D>
__device__ void gc2_adj(const Dev& a, int iv, int id, int nc, int nb, int t1, int t2, double lam) {
  const int d = d_nb_dim(nb);
  const int ta = d == 0 ? 1 : 0, tb = d == 2 ? 1 : 2;
  const bool low = d_nb_low(nb);
  const int nb_id = a.nbs[id * 6 + nb - 1];
  int c[3] = {0, 0, 0};
  c[ta] = t1;
  if (D == 3) c[tb] = t2;
  if (nb_id > 0) {
    c[d] = low ? nc - 1 : 2;
    atomicAdd(boxp(a, iv, nb_id) + cidx<D>(nc, c[0], c[1], c[2]), lam);
    return;
  }
  if (nb_id < 0) {
    c[d] = low ? 2 : nc - 1;
    atomicAdd(boxp(a, iv, id) + cidx<D>(nc, c[0], c[1], c[2]), lam);
    return;
  }
  const int pid = a.parent[id];
  const int pn = a.nbs[pid * 6 + nb - 1];
  const int pnc = a.ncell[pn];
  int off[3];
  child_off<D>(a, id, nc, nb, off);
  double* q = boxp(a, iv, pn);
  const int ix = low ? -1 : nc + 2;
  const int n1 = off[d] + ((ix + 1) >> 1), n2 = n1 + 1 - 2 * (ix & 1);
  const int c1 = off[ta] + ((t1 + 1) >> 1), c2 = c1 + 1 - 2 * (t1 & 1);
  if (D == 2) {
    int s0[3] = {0, 0, 0}, s1[3] = {0, 0, 0}, s2[3] = {0, 0, 0};
    s0[d] = n1; s0[ta] = c1;
    s1[d] = n2; s1[ta] = c1;
    s2[d] = n1; s2[ta] = c2;
    atomicAdd(q + cidx<D>(pnc, s0[0], s0[1], 0), 0.5 * lam);
    atomicAdd(q + cidx<D>(pnc, s1[0], s1[1], 0), 0.25 * lam);
    atomicAdd(q + cidx<D>(pnc, s2[0], s2[1], 0), 0.25 * lam);
    return;
  }
  const int e1 = off[tb] + ((t2 + 1) >> 1), e2 = e1 + 1 - 2 * (t2 & 1);
  int P4[4][3];
  for (int m = 0; m < 4; m++) {
    P4[m][d] = n1;
    P4[m][ta] = c1;
    P4[m][tb] = e1;
  }
  P4[1][d] = n2;
  P4[2][ta] = c2;
  P4[3][tb] = e2;
  for (int m = 0; m < 4; m++) atomicAdd(q + cidx<D>(pnc, P4[m][0], P4[m][1], P4[m][2]), 0.25 * lam);
}

__device__ __forceinline__ void koren_mlim_d(double x, double y, double& dx, double& dy) {
  const double aa = x * x, ab = x * y;
  if (ab <= 0) { dx = 0; dy = 0; }
  else if (aa <= 0.25 * ab) { dx = 1; dy = 0; }
  else if (aa <= 2.5 * ab) { dx = 1 / 3.0; dy = 1 / 6.0; }
  else { dx = 0; dy = 1; }
}

// fluxes_koren reverse: one lane per leaf face; consumes lF (clearing it)
template <int D>
__device__ void koren_adj_box(const Dev& a, const Tbl& tb, int id, int nc, int tpb, int lt,
                              int ie, int ifld, int ffld, int cmob, int cdif, int le, int lE,
                              int lF, int lfld, TgAcc& acc, double fac_all);

template <int D>
__global__ __launch_bounds__(256) void k_koren_adj(Dev a, Tbl tb, const int* __restrict__ list,
                                                   int n, int nc, int tpb, int ie, int ifld,
                                                   int ffld, int cmob, int cdif, int le, int lE,
                                                   int lF, int lfld, double* __restrict__ tg,
                                                   double fac_all, int use_lds) {
  extern __shared__ double s_tg[];
  __shared__ int s_range[2];
  TgAcc acc{use_lds ? s_tg : nullptr, tg, tb.n_rows, cmob, cdif};
  tg_begin(acc.lds, tb.n_rows, s_range);
  const int bpb = blockDim.x / tpb;
  const int slot = threadIdx.x / tpb;
  const int lt = threadIdx.x - slot * tpb;
  const int n_groups = (n + bpb - 1) / bpb;
  for (int grp = blockIdx.x; grp < n_groups; grp += gridDim.x) {
    const int li_ = grp * bpb + slot;
    if (li_ < n) koren_adj_box<D>(a, tb, list[li_], nc, tpb, lt, ie, ifld, ffld, cmob, cdif, le,
                                  lE, lF, lfld, acc, fac_all);
  }
  tg_flush(acc, s_range);
}

template <int D>
__device__ void koren_adj_box(const Dev& a, const Tbl& tb, int id, int nc, int tpb, int lt,
                              int ie, int ifld, int ffld, int cmob, int cdif, int le, int lE,
                              int lF, int lfld, TgAcc& acc, double fac_all) {
  const double fac = a.n_inst ? a.inst_val[inst_of<D>(a, id) * kInstVals + kInstKorenFac]
                              : fac_all;
  const double inv_dr = 1 / a.dr[id];
  const double* ne = boxp(a, ie, id);
  const double* E = boxp(a, ifld, id);
  double* Ln = boxp(a, le, id);
  double* LEf = boxp(a, lE, id);
  const int per = (nc + 1) * (D == 2 ? nc : nc * nc);
  for (int q = lt; q < D * per; q += tpb) {
    const int d = q / per, w = q - d * per;
    double* LFd = facep(a, lF, id, d, D);
    const double lam = LFd[w];
    LFd[w] = 0;
    if (lam == 0.0) continue;
    const int n0 = nc + (d == 0), n1 = nc + (d == 1);
    int c[3];
    c[0] = w % n0 + 1;
    c[1] = (w / n0) % n1 + 1;
    c[2] = D == 3 ? w / (n0 * n1) + 1 : 0;
    int m1[3] = {c[0], c[1], c[2]};
    m1[d] -= 1;
    const int pos = c[d];
    const int ic = cidx<D>(nc, c[0], c[1], c[2]);
    const int im = cidx<D>(nc, m1[0], m1[1], m1[2]);
    const double fld_avg = 0.5 * (E[ic] + E[im]);
    int li;
    double lf;
    lt_loc(tb, fld_avg, li, lf);
    const double mob = lt_col(tb, cmob, li, lf);
    const double dif = lt_col(tb, cdif, li, lf);
    const double fld = facep(a, ffld, id, d, D)[w];
    const double vd = -mob * fld;
    const double n0v = ne[ic], nm1 = ne[im];
    const double gradc = n0v - nm1;
    const int ta = d == 0 ? 1 : 0, tbb = d == 2 ? 1 : 2;
    const int t2 = D == 3 ? c[tbb] : 0;
    const double ldif = -lam * gradc * inv_dr;
    double lgradc = -lam * dif * inv_dr;
    double lvd = 0, lfl = 0;
    if (vd < 0.0) {
      const bool ghost = pos == nc + 1;
      int p1[3] = {c[0], c[1], c[2]};
      p1[d] += 1;
      const double tmp =
          ghost ? gc2_value<D>(a, ie, id, nc, 2 * d + 2, c[ta], t2) : ne[cidx<D>(nc, p1[0], p1[1], p1[2])];
      const double gradn = tmp - n0v;
      const double lim = koren_mlim(gradc, gradn);
      const double f = vd * (n0v - lim);
      if (f < fac * fld) {
        lfl += lam * fac;
      } else {
        double da, db;
        koren_mlim_d(gradc, gradn, da, db);
        lvd = lam * (n0v - lim);
        const double llim = -lam * vd;
        lgradc += llim * da;
        const double lgn = llim * db;
        atomicAdd(Ln + ic, lam * vd - lgn);
        if (ghost) gc2_adj<D>(a, le, id, nc, 2 * d + 2, c[ta], t2, lgn);
        else atomicAdd(Ln + cidx<D>(nc, p1[0], p1[1], p1[2]), lgn);
      }
    } else {
      const bool ghost = pos == 1;
      int m2[3] = {c[0], c[1], c[2]};
      m2[d] -= 2;
      const double tmp =
          ghost ? gc2_value<D>(a, ie, id, nc, 2 * d + 1, c[ta], t2) : ne[cidx<D>(nc, m2[0], m2[1], m2[2])];
      const double gradp = nm1 - tmp;
      const double lim = koren_mlim(gradc, gradp);
      const double f = vd * (nm1 + lim);
      if (f > fac * fld) {
        lfl += lam * fac;
      } else {
        double da, db;
        koren_mlim_d(gradc, gradp, da, db);
        lvd = lam * (nm1 + lim);
        const double llim = lam * vd;
        lgradc += llim * da;
        const double lgp = llim * db;
        atomicAdd(Ln + im, lam * vd + lgp);
        if (ghost) gc2_adj<D>(a, le, id, nc, 2 * d + 1, c[ta], t2, -lgp);
        else atomicAdd(Ln + cidx<D>(nc, m2[0], m2[1], m2[2]), -lgp);
      }
    }
    atomicAdd(Ln + ic, lgradc);
    atomicAdd(Ln + im, -lgradc);
    const double lmob = -lvd * fld;
    lfl += -lvd * mob;
    facep(a, lfld, id, d, D)[w] += lfl;
    const double lx = lmob * lt_dcol(tb, cmob, fld_avg, li) + ldif * lt_dcol(tb, cdif, fld_avg, li);
    atomicAdd(LEf + ic, 0.5 * lx);
    atomicAdd(LEf + im, 0.5 * lx);
    acc.add(0, li, lf, lmob);
    acc.add(1, li, lf, ldif);
  }
}

// per-box maxima of set_ref_flags (src/streamer_2d.f90:186-192):
// (dr^2 max|rhs|, max fld, max elec) over the interior
template <int D>
__global__ __launch_bounds__(256) void k_summary(Dev a, const int* __restrict__ list, int n,
                                                 int nc, int ir, int ifld, int ie,
                                                 double* __restrict__ out) {
  const int li = blockIdx.x;
  if (li >= n) return;
  const int id = list[li];
  const double* r = boxp(a, ir, id);
  const double* E = boxp(a, ifld, id);
  const double* ne = boxp(a, ie, id);
  double mr = -1.7976931348623157e308, me = mr, mn = mr;
  const int total = D == 2 ? nc * nc : nc * nc * nc;
  for (int q = threadIdx.x; q < total; q += blockDim.x) {
    const int i = q % nc + 1, j = (q / nc) % nc + 1, k = D == 3 ? q / (nc * nc) + 1 : 0;
    const int c = cidx<D>(nc, i, j, k);
    mr = fmax(mr, fabs(r[c]));
    me = fmax(me, E[c]);
    mn = fmax(mn, ne[c]);
  }
  for (int o = 32; o > 0; o >>= 1) {
    mr = fmax(mr, __shfl_xor(mr, o, 64));
    me = fmax(me, __shfl_xor(me, o, 64));
    mn = fmax(mn, __shfl_xor(mn, o, 64));
  }
  __shared__ double sm[3][4];
  if ((threadIdx.x & 63) == 0) {
    sm[0][threadIdx.x >> 6] = mr;
    sm[1][threadIdx.x >> 6] = me;
    sm[2][threadIdx.x >> 6] = mn;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < (int)(blockDim.x >> 6); w++) {
      mr = fmax(mr, sm[0][w]);
      me = fmax(me, sm[1][w]);
      mn = fmax(mn, sm[2][w]);
    }
    out[(size_t)(id - 1) * 3 + 0] = a.dr[id] * a.dr[id] * mr;
    out[(size_t)(id - 1) * 3 + 1] = me;
    out[(size_t)(id - 1) * 3 + 2] = mn;
  }
}

// a$D_prolong1_to (afivo/src/m_aX_prolong.f90:99-191), add = false, for every
// box of the list
template <int D>
__global__ __launch_bounds__(256) void k_prolong1(Dev a, const int* __restrict__ list, int n,
                                                  int nc, int tpb, int iv) {
  BOXES_PER_BLOCK
  if (!act) return;
  const int hnc = nc >> 1;
  const int pid = a.parent[id];
  const int pnc = a.ncell[pid];
  int off[3];
  child_off<D>(a, id, nc, 0, off);
  double* c = boxp(a, iv, id);
  const double* p = boxp(a, iv, pid);
  const int total = D == 2 ? hnc * hnc : hnc * hnc * hnc;
  for (int q = lt; q < total; q += tpb) {
    const int i = q % hnc + 1, j = (q / hnc) % hnc + 1, k = D == 3 ? q / (hnc * hnc) + 1 : 0;
    const int ic = i + off[0], jc = j + off[1], kc = D == 3 ? k + off[2] : 0;
    const int fi = 2 * i - 1, fj = 2 * j - 1, fk = D == 3 ? 2 * k - 1 : 0;
    if (D == 2) {
      const double f0 = 0.5 * p[cidx<D>(pnc, ic, jc, 0)];
      const double flx = 0.25 * p[cidx<D>(pnc, ic - 1, jc, 0)];
      const double fhx = 0.25 * p[cidx<D>(pnc, ic + 1, jc, 0)];
      const double fly = 0.25 * p[cidx<D>(pnc, ic, jc - 1, 0)];
      const double fhy = 0.25 * p[cidx<D>(pnc, ic, jc + 1, 0)];
      // the children start from zero (prolong1 without add)
      c[cidx<D>(nc, fi, fj, 0)] = f0 + flx + fly + 0.0;
      c[cidx<D>(nc, fi + 1, fj, 0)] = f0 + fhx + fly + 0.0;
      c[cidx<D>(nc, fi, fj + 1, 0)] = f0 + flx + fhy + 0.0;
      c[cidx<D>(nc, fi + 1, fj + 1, 0)] = f0 + fhx + fhy + 0.0;
    } else {
      const double f0 = 0.25 * p[cidx<D>(pnc, ic, jc, kc)];
      const double flx = 0.25 * p[cidx<D>(pnc, ic - 1, jc, kc)];
      const double fhx = 0.25 * p[cidx<D>(pnc, ic + 1, jc, kc)];
      const double fly = 0.25 * p[cidx<D>(pnc, ic, jc - 1, kc)];
      const double fhy = 0.25 * p[cidx<D>(pnc, ic, jc + 1, kc)];
      const double flz = 0.25 * p[cidx<D>(pnc, ic, jc, kc - 1)];
      const double fhz = 0.25 * p[cidx<D>(pnc, ic, jc, kc + 1)];
      c[cidx<D>(nc, fi, fj, fk)] = f0 + flx + fly + flz + 0.0;
      c[cidx<D>(nc, fi + 1, fj, fk)] = f0 + fhx + fly + flz + 0.0;
      c[cidx<D>(nc, fi, fj + 1, fk)] = f0 + flx + fhy + flz + 0.0;
      c[cidx<D>(nc, fi + 1, fj + 1, fk)] = f0 + fhx + fhy + flz + 0.0;
      c[cidx<D>(nc, fi, fj, fk + 1)] = f0 + flx + fly + fhz + 0.0;
      c[cidx<D>(nc, fi + 1, fj, fk + 1)] = f0 + fhx + fly + fhz + 0.0;
      c[cidx<D>(nc, fi, fj + 1, fk + 1)] = f0 + flx + fhy + fhz + 0.0;
      c[cidx<D>(nc, fi + 1, fj + 1, fk + 1)] = f0 + fhx + fhy + fhz + 0.0;
    }
  }
}

// cc(..., iv) = value over whole boxes
__global__ __launch_bounds__(256) void k_fill(Dev a, const int* __restrict__ list, int n,
                                              int bl, int tpb, int iv, double value) {
  BOXES_PER_BLOCK
  if (!act) return;
  double* p = boxp(a, iv, id);
  for (int q = lt; q < bl; q += tpb) p[q] = value;
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------

// Photon production rate of the photoionization model (set_photoi_rate,
// src/streamer_2d.f90:567-589, with src/streamer_3d.f90 and _cyl alike):
// |E| mu alpha n_e coeff on leaf interiors, clamped at zero; coeff is eta
// times the quenching factor p_q / (p + p_q).
template <int D>
__global__ __launch_bounds__(256) void k_photoi_rate(Dev a, Tbl tb, const int* __restrict__ list,
                                                     int n, int nc, int tpb, int ifld, int ie,
                                                     int ipho, int cmob, int calpha,
                                                     double coeff) {
  BOXES_PER_BLOCK
  if (!act) return;
  const double* E = boxp(a, ifld, id);
  const double* ne = boxp(a, ie, id);
  double* out = boxp(a, ipho, id);
  const int total = D == 2 ? nc * nc : nc * nc * nc;
  for (int q = lt; q < total; q += tpb) {
    const int i = q % nc + 1, j = (q / nc) % nc + 1, k = D == 3 ? q / (nc * nc) + 1 : 0;
    const int c = cidx_in<D>(nc, i, j, k);
    int li;
    double lf;
    lt_loc(tb, E[c], li, lf);
    const double alpha = lt_col(tb, calpha, li, lf);
    const double mob = lt_col(tb, cmob, li, lf);
    double v = E[c] * mob * alpha * ne[c] * coeff;
    if (v < 0) v = 0;
    out[c] = v;
  }
}

// out = s * in (add = 0) or out = out + s * in (add = 1) on leaf interiors
template <int D>
__global__ __launch_bounds__(256) void k_axpy_in(Dev a, const int* __restrict__ list, int n,
                                                 int nc, int tpb, int iin, int iout, double s,
                                                 int add) {
  BOXES_PER_BLOCK
  if (!act) return;
  const double* x = boxp(a, iin, id);
  double* y = boxp(a, iout, id);
  const int total = D == 2 ? nc * nc : nc * nc * nc;
  for (int q = lt; q < total; q += tpb) {
    const int i = q % nc + 1, j = (q / nc) % nc + 1, k = D == 3 ? q / (nc * nc) + 1 : 0;
    const int c = cidx_in<D>(nc, i, j, k);
    y[c] = add ? y[c] + s * x[c] : s * x[c];
  }
}

// ---------------------------------------------------------------------------

static void chk(hipError_t e, const char* w) {
  if (e != hipSuccess) throw std::runtime_error(std::string(w) + ": " + hipGetErrorString(e));
}

static Tbl device_table(Pool& P) {
  if (!P.d_table) throw std::logic_error("no transport table: call a5g_set_table first");
  Tbl t;
  t.rc = P.d_table;
  t.n_rows = P.tbl_rows;
  t.x_min = P.tbl_xmin;
  t.inv_dx = P.tbl_inv_dx;
  return t;
}

void set_table(Pool& P, int n_rows, int n_cols, double x_min, double inv_dx, const double* rc) {
  if (n_rows < 2 || n_cols < 1) throw std::invalid_argument("table needs >= 2 rows and >= 1 column");
  if (P.d_table) chk(hipFree(P.d_table), "hipFree");
  P.d_table = nullptr;
  chk(hipMalloc((void**)&P.d_table, (size_t)n_rows * n_cols * sizeof(double)), "hipMalloc table");
  chk(hipMemcpyAsync(P.d_table, rc, (size_t)n_rows * n_cols * sizeof(double),
                     hipMemcpyHostToDevice, P.stream), "table upload");
  chk(hipStreamSynchronize(P.stream), "table sync");
  P.tbl_rows = n_rows;
  P.tbl_cols = n_cols;
  P.tbl_xmin = x_min;
  P.tbl_inv_dx = inv_dx;
}

static int box_cells(int D, int nc, bool ghosts) {
  const int m = ghosts ? nc + 2 : nc;
  return D == 2 ? m * m : m * m * m;
}

void space_charge(Pool& P, const StVars& v, double fac) {
  P.sync_topology();
  const int D = P.tree.dim;
  for (int lvl = 1; lvl <= P.tree.max_lvl; lvl++) {
    const int n = P.n_leaves(lvl);
    if (!n) continue;
    const int nc = P.nc_lvl(lvl), bl = box_cells(D, nc, true);
    G2 g = g2(n, bl);
    // algorithmic bytes: 3 words per leaf cell (read n_e, n_i, write rhs)
    const int pr = P.prof_begin(PK_SPACE_CHARGE, lvl, 3.0 * 8 * n * box_cells(D, nc, false));
    hipLaunchKernelGGL(k_space_charge, dim3(g.grid), dim3(256), 0, P.stream, P.dev(),
                       P.leaves(lvl), n, bl, g.tpb, v.elec, v.pion, v.rhs, fac);
    chk(hipGetLastError(), "space_charge");
    P.prof_end(pr);
  }
}

void fld_from_pot(Pool& P, const StVars& v) {
  P.sync_topology();
  const int D = P.tree.dim;
  for (int lvl = P.tree.min_lvl; lvl <= P.tree.max_lvl; lvl++) {
    const int n = P.n_ids(lvl);
    if (!n) continue;
    const int nc = P.nc_lvl(lvl);
    G2 g = g2(n, D * (nc + 1) * (D == 2 ? nc : nc * nc));
    // 2 + D words per cell (read phi and eps, write D face fields and |E|)
    const int pr = P.prof_begin(PK_FIELD, lvl, (2.0 + D) * 8 * n * box_cells(D, nc, false));
    if (D == 2)
      hipLaunchKernelGGL((k_fld<2>), dim3(g.grid), dim3(256), 0, P.stream, P.dev(), P.ids(lvl),
                         n, nc, g.tpb, v.phi, v.eps, v.fld, v.f_fld);
    else
      hipLaunchKernelGGL((k_fld<3>), dim3(g.grid), dim3(256), 0, P.stream, P.dev(), P.ids(lvl),
                         n, nc, g.tpb, v.phi, v.eps, v.fld, v.f_fld);
    chk(hipGetLastError(), "fld_from_pot");
    P.prof_end(pr);
  }
}

void fluxes_koren(Pool& P, const StVars& v, double dt) {
  P.sync_topology();
  const int D = P.tree.dim;
  Tbl tb = device_table(P);
  const double eps0 = 8.8541878176e-12, qe = 1.6022e-19;
  const double fac = -0.8 * eps0 / (qe * dt);
  for (int lvl = P.tree.min_lvl; lvl <= P.tree.max_lvl; lvl++) {
    const int n = P.n_leaves(lvl);
    if (!n) continue;
    const int nc = P.nc_lvl(lvl);
    G2 g = g2(n, D * (nc + 1) * (D == 2 ? nc : nc * nc));
    // 2 + 2D words per leaf cell (read n_e, |E| and D face fields, write D
    // face fluxes)
    const int pr = P.prof_begin(PK_KOREN, lvl, (2.0 + 2 * D) * 8 * n * box_cells(D, nc, false));
    if (D == 2)
      hipLaunchKernelGGL((k_koren<2>), dim3(g.grid), dim3(256), 0, P.stream, P.dev(), tb,
                         P.leaves(lvl), n, nc, g.tpb, v.elec, v.fld, v.f_elec, v.f_fld, v.t_mob,
                         v.t_dif, fac);
    else
      hipLaunchKernelGGL((k_koren<3>), dim3(g.grid), dim3(256), 0, P.stream, P.dev(), tb,
                         P.leaves(lvl), n, nc, g.tpb, v.elec, v.fld, v.f_elec, v.f_fld, v.t_mob,
                         v.t_dif, fac);
    chk(hipGetLastError(), "fluxes_koren");
    P.prof_end(pr);
  }
}

// a$D_consistent_fluxes: (parent, direction) pairs whose neighbour is a
// leaf, per level; the list is rebuilt when the topology changes
static void ensure_flux_pairs(Pool& P) {
  const auto& T = P.tree;
  const int nl = T.max_lvl - T.min_lvl;
  if (P.flux_pairs_version != P.topo_version) {
    std::vector<int2> pairs;
    P.flux_pairs_off.assign(std::max(nl, 0), 0);
    P.flux_pairs_n.assign(std::max(nl, 0), 0);
    for (int lvl = T.min_lvl; lvl <= T.max_lvl - 1; lvl++) {
      P.flux_pairs_off[lvl - T.min_lvl] = pairs.size();
      for (int id : P.lparents(lvl))
        for (int nb = 1; nb <= T.nnb; nb++) {
          const int nb_id = T.nbs[id * 6 + nb - 1];
          if (nb_id > 0 && !T.has_children(nb_id)) pairs.push_back(make_int2(id, nb));
        }
      P.flux_pairs_n[lvl - T.min_lvl] = pairs.size() - P.flux_pairs_off[lvl - T.min_lvl];
    }
    if (pairs.size() > P.flux_pairs_cap || !P.d_flux_pairs) {
      if (P.d_flux_pairs) chk(hipFree(P.d_flux_pairs), "hipFree");
      P.d_flux_pairs = nullptr;
      P.flux_pairs_cap = std::max<size_t>(pairs.size(), 64);
      chk(hipMalloc((void**)&P.d_flux_pairs, P.flux_pairs_cap * sizeof(int2)), "hipMalloc");
    }
    if (!pairs.empty()) {
      chk(hipMemcpyAsync(P.d_flux_pairs, pairs.data(), pairs.size() * sizeof(int2),
                         hipMemcpyHostToDevice, P.stream), "pairs");
      chk(hipStreamSynchronize(P.stream), "pairs sync");
    }
    P.flux_pairs_version = P.topo_version;
  }
}

void consistent_fluxes(Pool& P, int f_elec) {
  P.sync_topology();
  ensure_flux_pairs(P);
  const auto& T = P.tree;
  const int D = T.dim;
  for (int lvl = T.min_lvl; lvl <= T.max_lvl - 1; lvl++) {
    const size_t np = P.flux_pairs_n[lvl - T.min_lvl];
    if (np == 0) continue;
    const int2* d_pairs = P.d_flux_pairs + P.flux_pairs_off[lvl - T.min_lvl];
    const int nc = P.nc_lvl(lvl);
    const long work = (long)np * (D == 2 ? nc : nc * nc);
    const int grid = (int)std::min<long>((work + 255) / 256, 65535);
    // one coarse face written per 2^(D-1) fine faces read
    const int pr = P.prof_begin(PK_FLUX_FIX, lvl, (1.0 + (1 << (D - 1))) * 8 * work / (1 << (D - 1)));
    if (D == 2)
      hipLaunchKernelGGL((k_flux_fix<2>), dim3(grid), dim3(256), 0, P.stream, P.dev(), d_pairs,
                         (int)np, nc, f_elec);
    else
      hipLaunchKernelGGL((k_flux_fix<3>), dim3(grid), dim3(256), 0, P.stream, P.dev(), d_pairs,
                         (int)np, nc, f_elec);
    chk(hipGetLastError(), "consistent_fluxes");
    P.prof_end(pr);
  }
}

void update_solution(Pool& P, const StVars& v, double dt) {
  P.sync_topology();
  const int D = P.tree.dim;
  Tbl tb = device_table(P);
  const bool cyl = D == 2 && P.tree.coord == 1;
  for (int lvl = P.tree.min_lvl; lvl <= P.tree.max_lvl; lvl++) {
    const int n = P.n_leaves(lvl);
    if (!n) continue;
    const int nc = P.nc_lvl(lvl);
    G2 g = g2(n, box_cells(D, nc, false));
    // 6 + D words per leaf cell (5 + D cylindrical): read n_e, n_i, |E|, eps
    // and the 2D face fluxes counted once per cell, write n_e, n_i
    const int pr = P.prof_begin(PK_UPDATE, lvl,
                                ((cyl ? 5.0 : 6.0) + D) * 8 * n * box_cells(D, nc, false));
    if (D == 3)
      hipLaunchKernelGGL((k_update<3, false>), dim3(g.grid), dim3(256), 0, P.stream, P.dev(), tb,
                         P.leaves(lvl), n, nc, g.tpb, v.elec, v.pion, v.fld, v.eps, v.pho,
                         v.f_elec, v.t_alpha, v.t_eta, v.dielectric_rule, dt);
    else if (cyl)
      hipLaunchKernelGGL((k_update<2, true>), dim3(g.grid), dim3(256), 0, P.stream, P.dev(), tb,
                         P.leaves(lvl), n, nc, g.tpb, v.elec, v.pion, v.fld, v.eps, v.pho,
                         v.f_elec, v.t_alpha, v.t_eta, v.dielectric_rule, dt);
    else
      hipLaunchKernelGGL((k_update<2, false>), dim3(g.grid), dim3(256), 0, P.stream, P.dev(), tb,
                         P.leaves(lvl), n, nc, g.tpb, v.elec, v.pion, v.fld, v.eps, v.pho,
                         v.f_elec, v.t_alpha, v.t_eta, v.dielectric_rule, dt);
    chk(hipGetLastError(), "update_solution");
    P.prof_end(pr);
  }
}

void average_dens(Pool& P, const StVars& v) {
  P.sync_topology();
  const int D = P.tree.dim;
  for (int lvl = P.tree.min_lvl; lvl <= P.tree.max_lvl; lvl++) {
    const int n = P.n_ids(lvl);
    if (!n) continue;
    const int bl = box_cells(D, P.nc_lvl(lvl), true);
    G2 g = g2(n, bl);
    // 6 words per cell (read 4, write 2)
    const int pr = P.prof_begin(PK_AVERAGE, lvl, 6.0 * 8 * n * box_cells(D, P.nc_lvl(lvl), false));
    hipLaunchKernelGGL(k_average, dim3(g.grid), dim3(256), 0, P.stream, P.dev(), P.ids(lvl), n,
                       bl, g.tpb, v.elec, v.elec_old, v.pion, v.pion_old);
    chk(hipGetLastError(), "average_dens");
    P.prof_end(pr);
  }
}

void photoi_rate(Pool& P, const StVars& v, double coeff) {
  P.sync_topology();
  const int D = P.tree.dim;
  Tbl tb = device_table(P);
  for (int lvl = P.tree.min_lvl; lvl <= P.tree.max_lvl; lvl++) {
    const int n = P.n_leaves(lvl);
    if (!n) continue;
    const int nc = P.nc_lvl(lvl);
    G2 g = g2(n, box_cells(D, nc, false));
    const int pr = P.prof_begin(PK_PHOTOI_RATE, lvl, 3.0 * 8 * n * box_cells(D, nc, false));
    if (D == 3)
      hipLaunchKernelGGL((k_photoi_rate<3>), dim3(g.grid), dim3(256), 0, P.stream, P.dev(), tb,
                         P.leaves(lvl), n, nc, g.tpb, v.fld, v.elec, v.pho, v.t_mob, v.t_alpha,
                         coeff);
    else
      hipLaunchKernelGGL((k_photoi_rate<2>), dim3(g.grid), dim3(256), 0, P.stream, P.dev(), tb,
                         P.leaves(lvl), n, nc, g.tpb, v.fld, v.elec, v.pho, v.t_mob, v.t_alpha,
                         coeff);
    chk(hipGetLastError(), "photoi_rate");
    P.prof_end(pr);
  }
}

void axpy_leaves(Pool& P, int iin, int iout, double s, bool add) {
  P.sync_topology();
  const int D = P.tree.dim;
  for (int lvl = P.tree.min_lvl; lvl <= P.tree.max_lvl; lvl++) {
    const int n = P.n_leaves(lvl);
    if (!n) continue;
    const int nc = P.nc_lvl(lvl);
    G2 g = g2(n, box_cells(D, nc, false));
    if (D == 3)
      hipLaunchKernelGGL((k_axpy_in<3>), dim3(g.grid), dim3(256), 0, P.stream, P.dev(),
                         P.leaves(lvl), n, nc, g.tpb, iin, iout, s, (int)add);
    else
      hipLaunchKernelGGL((k_axpy_in<2>), dim3(g.grid), dim3(256), 0, P.stream, P.dev(),
                         P.leaves(lvl), n, nc, g.tpb, iin, iout, s, (int)add);
    chk(hipGetLastError(), "axpy_leaves");
  }
}

// The per-box maxima of every box of levels >= 1 into d_out (3 max_id
// values, zero for other boxes), then copied to the host.
static void summaries_into(Pool& P, const StVars& v, double* d_out, double* out_host) {
  const auto& T = P.tree;
  const int D = T.dim;
  const size_t n_out = (size_t)std::max(T.max_id, 1) * 3;
  hipLaunchKernelGGL(k_zero, dim3((unsigned)std::min<size_t>((n_out + 255) / 256, 1024)),
                     dim3(256), 0, P.stream, d_out, n_out);
  chk(hipGetLastError(), "box_summaries zero");
  for (int lvl = 1; lvl <= T.max_lvl; lvl++) {
    const int n = P.n_ids(lvl);
    if (!n) continue;
    if (D == 2)
      hipLaunchKernelGGL((k_summary<2>), dim3(n), dim3(256), 0, P.stream, P.dev(), P.ids(lvl),
                         n, P.nc_lvl(lvl), v.rhs, v.fld, v.elec, d_out);
    else
      hipLaunchKernelGGL((k_summary<3>), dim3(n), dim3(256), 0, P.stream, P.dev(), P.ids(lvl),
                         n, P.nc_lvl(lvl), v.rhs, v.fld, v.elec, d_out);
    chk(hipGetLastError(), "box_summaries");
  }
  chk(hipMemcpyAsync(out_host, d_out, (size_t)T.max_id * 3 * sizeof(double),
                     hipMemcpyDeviceToHost, P.stream), "summaries");
  chk(hipStreamSynchronize(P.stream), "summaries sync");
}

void box_summaries(Pool& P, const StVars& v, double* out_host) {
  P.sync_topology();
  const size_t n_out = (size_t)std::max(P.tree.max_id, 1) * 3;
  summaries_into(P, v, (double*)P.work(n_out * sizeof(double)), out_host);
}

// ---------------------------------------------------------------------------
// Scratch of the no-allocation summaries (a5g_box_summaries_noalloc): the
// 3 maxima of up to kNoAllocMaxBoxes boxes in a file-scope device array,
// which each GPU the library runs on holds until the process ends. Host
// code reaches it only through the address hipGetSymbolAddress returns for
// the device (noalloc_ptr), looked up once per device by a5g_init.
// ---------------------------------------------------------------------------

static __device__ __align__(256) double scratch[3 * kNoAllocMaxBoxes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_noalloc_mu;
double* g_noalloc_ptr[kMaxDevices] = {};
}  // namespace

void noalloc_init() {
  int dev = 0;
  chk(hipGetDevice(&dev), "hipGetDevice");
  if (dev < 0 || dev >= kMaxDevices) throw std::runtime_error("device index out of range");
  std::lock_guard<std::mutex> lk(g_noalloc_mu);
  if (g_noalloc_ptr[dev]) return;
  void* p = nullptr;
  chk(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress scratch");
  g_noalloc_ptr[dev] = (double*)p;
}

void box_summaries_noalloc(Pool& P, const StVars& v, double* out_host) {
  if (P.topo_dirty || P.lists_dirty)
    throw std::logic_error(
        "box_summaries_noalloc: the tree changed since its last upload; a call that may "
        "allocate (a5g_box_summaries or any kernel call) has to upload it first");
  if (P.tree.max_id > kNoAllocMaxBoxes)
    throw std::length_error("box_summaries_noalloc: " + std::to_string(P.tree.max_id) +
                            " boxes, the scratch holds " + std::to_string(kNoAllocMaxBoxes));
  int dev = 0;
  chk(hipGetDevice(&dev), "hipGetDevice");
  double* ptr = (dev >= 0 && dev < kMaxDevices) ? g_noalloc_ptr[dev] : nullptr;
  if (!ptr) throw std::logic_error("box_summaries_noalloc: no scratch on this device (a5g_init "
                                   "was called with another device current)");
  summaries_into(P, v, ptr, out_host);
}

// device copy of a host id list, grouped by level; calls fn(level, ptr, n)
template <typename F>
static void by_level(Pool& P, int n_ids, const int* ids, F fn) {
  const auto& T = P.tree;
  for (int lvl = T.min_lvl; lvl <= T.max_lvl; lvl++) {
    std::vector<int> sel;
    for (int i = 0; i < n_ids; i++) {
      const int id = ids[i];
      if (id < 1 || id > T.max_id || !T.in_use[id])
        throw std::out_of_range("box id " + std::to_string(id) + " is not in use");
      if (T.lvl[id] == lvl) sel.push_back(id);
    }
    if (sel.empty()) continue;
    int* d = (int*)P.work(sel.size() * sizeof(int));
    chk(hipMemcpyAsync(d, sel.data(), sel.size() * sizeof(int), hipMemcpyHostToDevice, P.stream),
        "ids");
    fn(lvl, d, (int)sel.size());
    chk(hipStreamSynchronize(P.stream), "sync");
  }
}

void prolong1_boxes(Pool& P, int n_ids, const int* ids, int iv) {
  P.sync_topology();
  const int D = P.tree.dim;
  by_level(P, n_ids, ids, [&](int lvl, const int* d, int n) {
    if (lvl <= P.tree.min_lvl) throw std::invalid_argument("prolong1: box without parent");
    const int nc = P.nc_lvl(lvl), hnc = nc / 2;
    G2 g = g2(n, D == 2 ? hnc * hnc : hnc * hnc * hnc);
    if (D == 2)
      hipLaunchKernelGGL((k_prolong1<2>), dim3(g.grid), dim3(256), 0, P.stream, P.dev(), d, n,
                         nc, g.tpb, iv);
    else
      hipLaunchKernelGGL((k_prolong1<3>), dim3(g.grid), dim3(256), 0, P.stream, P.dev(), d, n,
                         nc, g.tpb, iv);
    chk(hipGetLastError(), "prolong1");
  });
}

void fill_boxes(Pool& P, int n_ids, const int* ids, int iv, double value) {
  P.sync_topology();
  const int D = P.tree.dim;
  by_level(P, n_ids, ids, [&](int lvl, const int* d, int n) {
    const int bl = box_cells(D, P.nc_lvl(lvl), true);
    G2 g = g2(n, bl);
    hipLaunchKernelGGL(k_fill, dim3(g.grid), dim3(256), 0, P.stream, P.dev(), d, n, bl, g.tpb, iv,
                       value);
    chk(hipGetLastError(), "fill");
  });
}

void gc_boxes(Pool& P, int n_ids, const int* ids, int iv, int rb, int bc) {
  P.sync_topology();
  by_level(P, n_ids, ids, [&](int lvl, const int* d, int n) { gc_list(P, lvl, d, n, iv, rb, bc); });
}

// host side of the reverse-mode kernels

static double* table_grad(Pool& P) {
  if (!P.d_table) throw std::logic_error("no transport table: call a5g_set_table first");
  const size_t n = (size_t)P.tbl_rows * P.tbl_cols;
  if (!P.d_tgrad || P.tgrad_n != n) {
    if (P.d_tgrad) chk(hipFree(P.d_tgrad), "hipFree");
    P.d_tgrad = nullptr;
    chk(hipMalloc((void**)&P.d_tgrad, n * sizeof(double)), "hipMalloc table adjoint");
    P.tgrad_n = n;
    hipLaunchKernelGGL(k_zero, dim3((unsigned)std::min<size_t>((n + 255) / 256, 1024)), dim3(256),
                       0, P.stream, P.d_tgrad, n);
    chk(hipGetLastError(), "table adjoint zero");
  }
  return P.d_tgrad;
}

void table_adjoint(Pool& P, bool clear, double* out_host) {
  double* g = table_grad(P);
  if (out_host) {
    chk(hipMemcpyAsync(out_host, g, P.tgrad_n * sizeof(double), hipMemcpyDeviceToHost, P.stream),
        "table adjoint");
    chk(hipStreamSynchronize(P.stream), "table adjoint sync");
  }
  if (clear) {
    hipLaunchKernelGGL(k_zero, dim3((unsigned)std::min<size_t>((P.tgrad_n + 255) / 256, 1024)),
                       dim3(256), 0, P.stream, g, P.tgrad_n);
    chk(hipGetLastError(), "table adjoint zero");
  }
}

// Launch of the reverse kernels that gather table-row adjoints: with the
// two columns' rows fitting in LDS, 4 workgroups per compute unit at most
// walk the box groups (so each workgroup adds its touched rows to global
// memory once for many boxes); otherwise the one-shot grid with global
// atomics.
struct TgLaunch {
  int grid, use_lds;
  size_t lds_bytes;
};
static TgLaunch tg_launch(const Pool& P, int grid) {
  if (P.tbl_rows > kTgLdsRows) return {grid, 0, 0};
  return {std::min(grid, 4 * P.n_cu), 1, (size_t)2 * P.tbl_rows * sizeof(double)};
}

void update_solution_adj(Pool& P, const StVars& v, double dt, const AdjVars& av) {
  P.sync_topology();
  const int D = P.tree.dim;
  Tbl tb = device_table(P);
  double* tg = table_grad(P);
  const bool cyl = D == 2 && P.tree.coord == 1;
  const int photoi = v.pho > 0 ? 1 : 0;
  for (int lvl = P.tree.min_lvl; lvl <= P.tree.max_lvl; lvl++) {
    const int n = P.n_leaves(lvl);
    if (!n) continue;
    const int nc = P.nc_lvl(lvl);
    G2 g = g2(n, box_cells(D, nc, false));
    const TgLaunch tl = tg_launch(P, g.grid);
    auto go = [&](auto kern) {
      hipLaunchKernelGGL(kern, dim3(tl.grid), dim3(256), tl.lds_bytes, P.stream, P.dev(), tb,
                         P.leaves(lvl), n, nc, g.tpb, v.fld, v.eps, v.f_elec, v.t_alpha, v.t_eta,
                         v.dielectric_rule, photoi, av.le, av.li, av.lE, av.lpho, av.lF, tg, dt,
                         tl.use_lds);
    };
    if (D == 3) go(k_update_adj<3, false>);
    else if (cyl) go(k_update_adj<2, true>);
    else go(k_update_adj<2, false>);
    chk(hipGetLastError(), "update_solution_adj");
  }
}

void consistent_fluxes_adj(Pool& P, int lF) {
  P.sync_topology();
  ensure_flux_pairs(P);
  const auto& T = P.tree;
  const int D = T.dim;
  for (int lvl = T.min_lvl; lvl <= T.max_lvl - 1; lvl++) {
    const size_t np = P.flux_pairs_n[lvl - T.min_lvl];
    if (np == 0) continue;
    const int2* d_pairs = P.d_flux_pairs + P.flux_pairs_off[lvl - T.min_lvl];
    const int nc = P.nc_lvl(lvl);
    const long work = (long)np * (D == 2 ? nc : nc * nc);
    const int grid = (int)std::min<long>((work + 255) / 256, 65535);
    if (D == 2)
      hipLaunchKernelGGL((k_flux_fix_adj<2>), dim3(grid), dim3(256), 0, P.stream, P.dev(), d_pairs,
                         (int)np, nc, lF);
    else
      hipLaunchKernelGGL((k_flux_fix_adj<3>), dim3(grid), dim3(256), 0, P.stream, P.dev(), d_pairs,
                         (int)np, nc, lF);
    chk(hipGetLastError(), "consistent_fluxes_adj");
  }
}

void fluxes_koren_adj(Pool& P, const StVars& v, double dt, const AdjVars& av) {
  P.sync_topology();
  const int D = P.tree.dim;
  Tbl tb = device_table(P);
  double* tg = table_grad(P);
  const double eps0 = 8.8541878176e-12, qe = 1.6022e-19;
  const double fac = -0.8 * eps0 / (qe * dt);
  for (int lvl = P.tree.min_lvl; lvl <= P.tree.max_lvl; lvl++) {
    const int n = P.n_leaves(lvl);
    if (!n) continue;
    const int nc = P.nc_lvl(lvl);
    G2 g = g2(n, D * (nc + 1) * (D == 2 ? nc : nc * nc));
    const TgLaunch tl = tg_launch(P, g.grid);
    auto go = [&](auto kern) {
      hipLaunchKernelGGL(kern, dim3(tl.grid), dim3(256), tl.lds_bytes, P.stream, P.dev(), tb,
                         P.leaves(lvl), n, nc, g.tpb, v.elec, v.fld, v.f_fld, v.t_mob, v.t_dif,
                         av.le, av.lE, av.lF, av.lfld, tg, fac, tl.use_lds);
    };
    if (D == 2) go(k_koren_adj<2>);
    else go(k_koren_adj<3>);
    chk(hipGetLastError(), "fluxes_koren_adj");
  }
}

}  // namespace a5g

namespace a5g {

// ---------------------------------------------------------------------------
// set_ref_flags on the device (src/streamer_2d.f90:177-216, with the 3D and
// cylindrical drivers' crv_refine and seed windows): the flag of every box of
// the active runs from its maxima (k_summary), in the arithmetic of the
// host's rule, so that each flag is the host's. A box of run g is placed in
// g's own coordinates the way a one-run tree places it: its level-1
// ancestor at (ix - 1) dr box_size along the last dimension (ix counted
// within the run), then each child's offset 0.5 dr dix n_cell added level by
// level.
// ---------------------------------------------------------------------------

// GM_dist_line (src/m_geom.f90:29-46), the norm by norm2_flang
template <int D>
__device__ double dist_line_d(const double* r, const double* r0, const double* r1) {
  double line_len2 = 0.0, temp = 0.0, d[3];
  for (int k = 0; k < D; k++) {
    const double t = r1[k] - r0[k];
    line_len2 = line_len2 + t * t;
  }
  for (int k = 0; k < D; k++) temp = temp + (r[k] - r0[k]) * (r1[k] - r0[k]);
  if (temp <= 0.0) {
    for (int k = 0; k < D; k++) d[k] = r[k] - r0[k];
  } else if (temp >= line_len2) {
    for (int k = 0; k < D; k++) d[k] = r[k] - r1[k];
  } else {
    for (int k = 0; k < D; k++) d[k] = r[k] - (r0[k] + temp / line_len2 * (r1[k] - r0[k]));
  }
  return norm2_flang<D>(d);
}

__global__ __launch_bounds__(256) void k_fill_i32(int* __restrict__ p, size_t n, int v) {
  for (size_t q = blockIdx.x * (size_t)blockDim.x + threadIdx.x; q < n;
       q += (size_t)gridDim.x * blockDim.x)
    p[q] = v;
}

template <int D>
__global__ __launch_bounds__(256) void k_ref_flags(Dev a, Tbl tb, const int* __restrict__ list,
                                                   int n, const double* __restrict__ sum,
                                                   const double* __restrict__ rules, int n_seed,
                                                   int n_reg, int n_runs, int n_last,
                                                   double window, double crv_refine,
                                                   int c_alpha, int* __restrict__ flags) {
  const int li = blockIdx.x * blockDim.x + threadIdx.x;
  if (li >= n) return;
  const int id = list[li];
  const int g = a.n_inst ? inst_of<D>(a, id) : 0;
  const double dr = a.dr[id];
  const int nc = a.ncell[id];
  const double crv_phi = sum[(size_t)(id - 1) * 3], max_fld = sum[(size_t)(id - 1) * 3 + 1];
  int lti;
  double ltf;
  lt_loc(tb, max_fld, lti, ltf);
  const double adx = dr * lt_col(tb, c_alpha, lti, ltf);
  int f = 0;
  if (adx < 0.1 && dr < 2.0e-5) f = -1;
  if (adx < 0.1 && crv_phi < 4.0 && dr < 5.0e-5) f = -1;
  // the box's r_min in its run's coordinates
  double rl[3] = {a.rmin[id * 3], a.rmin[id * 3 + 1], a.rmin[id * 3 + 2]};
  if (a.n_inst) {
    int chain[64];
    int m = 0, b = id;
    while (a.lvl[b] > 1 && m < 64) {
      chain[m++] = b;
      b = a.parent[b];
    }
    const int ixl = a.ix[b * 3 + D - 1] - g * n_last;
    double loc = 0.0 + (double)(ixl - 1) * a.dr[b] * (double)a.ncell[b];
    for (int q = m - 1; q >= 0; q--) {
      const int c = chain[q], p = a.parent[c];
      const int dix = (a.ix[c * 3 + D - 1] - 1) & 1;
      loc = loc + 0.5 * a.dr[p] * (double)dix * (double)a.ncell[p];
    }
    rl[D - 1] = loc;
  }
  double centre[3];
  for (int k = 0; k < D; k++) centre[k] = rl[k] + 0.5 * nc * dr;
  const double* seeds = rules;
  const double* regions = rules + (size_t)n_runs * n_seed * (2 * D + 1);
  const double* times = regions + (size_t)n_reg * (1 + 2 * D);
  if (times[g] < window) {
    const double boxlen = nc * dr;
    for (int k = 0; k < n_seed; k++) {
      const double* sd = seeds + ((size_t)g * n_seed + k) * (2 * D + 1);
      const double w = sd[2 * D];
      if (dist_line_d<D>(centre, sd, sd + D) - w < boxlen && dr > 0.2 * w) f = 1;
    }
  }
  if (adx > 1.0 && crv_phi > crv_refine) f = 1;
  for (int r = 0; r < n_reg; r++) {
    const double* rg = regions + (size_t)r * (1 + 2 * D);
    if (a.lvl[id] >= rg[0]) continue;
    bool in = true;
    for (int k = 0; k < D; k++) in = in && centre[k] >= rg[1 + k] && centre[k] <= rg[1 + D + k];
    if (in) f = 1;
  }
  flags[id - 1] = f;
}

// Per run: the largest value of cc(iv) over the interiors of its leaf boxes
// of levels >= 1 (ordered-integer keys) and its highest level holding boxes
template <int D>
__global__ __launch_bounds__(256) void k_inst_max(Dev a, const int* __restrict__ list, int n,
                                                  int nc, int iv, int lvl, int leaves,
                                                  unsigned long long* __restrict__ vkey,
                                                  int* __restrict__ lmax) {
  const int li = blockIdx.x;
  if (li >= n) return;
  const int id = list[li];
  const int g = a.n_inst ? inst_of<D>(a, id) : 0;
  if (threadIdx.x == 0) atomicMax(&lmax[g], lvl);
  if (!leaves) return;
  const double* v = boxp(a, iv, id);
  double m = -__builtin_huge_val();
  const int total = D == 2 ? nc * nc : nc * nc * nc;
  for (int q = threadIdx.x; q < total; q += blockDim.x) {
    const int i = q % nc + 1, j = (q / nc) % nc + 1, k = D == 3 ? q / (nc * nc) + 1 : 0;
    m = fmax(m, v[cidx<D>(nc, i, j, k)]);
  }
  for (int o = 32; o > 0; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
  if ((threadIdx.x & 63) == 0) {
    unsigned long long u = __double_as_longlong(m);
    atomicMax(&vkey[g], (u >> 63) ? ~u : (u | 0x8000000000000000ull));
  }
}

void set_ref_rules(Pool& P, int n_seed, const double* seeds, int n_reg, const double* regions,
                   double window, double crv_refine, int n_last) {
  const int D = P.tree.dim;
  const int runs = P.n_inst ? P.n_inst : 1;
  const size_t ns = (size_t)runs * n_seed * (2 * D + 1), nr = (size_t)n_reg * (1 + 2 * D);
  P.h_rules.assign(seeds, seeds + ns);
  P.h_rules.insert(P.h_rules.end(), regions, regions + nr);
  P.rules_n_seed = n_seed;
  P.rules_n_reg = n_reg;
  P.rules_runs = runs;
  P.rules_window = window;
  P.rules_crv = crv_refine;
  P.rules_n_last = n_last;
  P.rules_set = true;
}

void ref_flags(Pool& P, const StVars& v, const double* times, int* flags_host) {
  if (!P.rules_set) throw std::logic_error("ref_flags: call a5g_set_ref_rules first");
  const int runs = P.n_inst ? P.n_inst : 1;
  if (runs != P.rules_runs) throw std::logic_error("ref_flags: the rules were set for another number of runs");
  P.sync_topology();
  const auto& T = P.tree;
  const int D = T.dim;
  const Tbl tb = device_table(P);
  const size_t n_id = (size_t)std::max(T.max_id, 1);
  // the work buffer: maxima (3 per id), then the flags
  double* d_sum = (double*)P.work(n_id * 3 * sizeof(double) + n_id * sizeof(int));
  int* d_flags = (int*)(d_sum + n_id * 3);
  // the rules and this call's times, in one upload (kept alive until the
  // synchronisation at the end)
  std::vector<double> buf(P.h_rules);
  buf.insert(buf.end(), times, times + runs);
  if (buf.size() > P.rules_cap) {
    if (P.d_rules) chk(hipFree(P.d_rules), "hipFree rules");
    P.d_rules = nullptr;
    P.rules_cap = 0;
    chk(hipMalloc((void**)&P.d_rules, buf.size() * sizeof(double)), "hipMalloc rules");
    P.rules_cap = buf.size();
  }
  chk(hipMemcpyAsync(P.d_rules, buf.data(), buf.size() * sizeof(double), hipMemcpyHostToDevice,
                     P.stream),
      "rules upload");
  hipLaunchKernelGGL(k_fill_i32, dim3((unsigned)std::min<size_t>((n_id + 255) / 256, 4096)),
                     dim3(256), 0, P.stream, d_flags, n_id, 0);
  chk(hipGetLastError(), "ref_flags fill");
  for (int lvl = 1; lvl <= T.max_lvl; lvl++) {
    const int n = P.n_ids(lvl);
    if (!n) continue;
    if (D == 2)
      hipLaunchKernelGGL((k_summary<2>), dim3(n), dim3(256), 0, P.stream, P.dev(), P.ids(lvl),
                         n, P.nc_lvl(lvl), v.rhs, v.fld, v.elec, d_sum);
    else
      hipLaunchKernelGGL((k_summary<3>), dim3(n), dim3(256), 0, P.stream, P.dev(), P.ids(lvl),
                         n, P.nc_lvl(lvl), v.rhs, v.fld, v.elec, d_sum);
    chk(hipGetLastError(), "ref_flags summaries");
    const dim3 grid((n + 255) / 256);
    if (D == 2)
      hipLaunchKernelGGL((k_ref_flags<2>), grid, dim3(256), 0, P.stream, P.dev(), tb, P.ids(lvl),
                         n, d_sum, P.d_rules, P.rules_n_seed, P.rules_n_reg, runs,
                         P.rules_n_last, P.rules_window, P.rules_crv, v.t_alpha, d_flags);
    else
      hipLaunchKernelGGL((k_ref_flags<3>), grid, dim3(256), 0, P.stream, P.dev(), tb, P.ids(lvl),
                         n, d_sum, P.d_rules, P.rules_n_seed, P.rules_n_reg, runs,
                         P.rules_n_last, P.rules_window, P.rules_crv, v.t_alpha, d_flags);
    chk(hipGetLastError(), "ref_flags");
  }
  chk(hipMemcpyAsync(flags_host, d_flags, (size_t)T.max_id * sizeof(int), hipMemcpyDeviceToHost,
                     P.stream),
      "flags download");
  chk(hipStreamSynchronize(P.stream), "ref_flags sync");
}

void instance_maxima(Pool& P, int iv, double* vmax, int* lvl_max) {
  P.sync_topology();
  const auto& T = P.tree;
  const int D = T.dim;
  const int runs = P.n_inst ? P.n_inst : 1;
  unsigned long long* d_key =
      (unsigned long long*)P.work(runs * (sizeof(unsigned long long) + sizeof(int)));
  int* d_lvl = (int*)(d_key + runs);
  hipLaunchKernelGGL(k_zero, dim3(1), dim3(256), 0, P.stream, (double*)d_key, (size_t)runs);
  hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(256), 0, P.stream, d_lvl, (size_t)runs, 0);
  chk(hipGetLastError(), "instance_maxima fill");
  for (int lvl = 1; lvl <= T.max_lvl; lvl++) {
    for (int leaves = 0; leaves < 2; leaves++) {
      const int n = leaves ? P.n_leaves(lvl) : P.n_ids(lvl);
      if (!n) continue;
      const int* list = leaves ? P.leaves(lvl) : P.ids(lvl);
      if (D == 2)
        hipLaunchKernelGGL((k_inst_max<2>), dim3(n), dim3(256), 0, P.stream, P.dev(), list, n,
                           P.nc_lvl(lvl), iv, lvl, leaves, d_key, d_lvl);
      else
        hipLaunchKernelGGL((k_inst_max<3>), dim3(n), dim3(256), 0, P.stream, P.dev(), list, n,
                           P.nc_lvl(lvl), iv, lvl, leaves, d_key, d_lvl);
      chk(hipGetLastError(), "instance_maxima");
    }
  }
  std::vector<unsigned long long> key(runs);
  chk(hipMemcpyAsync(key.data(), d_key, runs * sizeof(unsigned long long), hipMemcpyDeviceToHost,
                     P.stream),
      "maxima download");
  chk(hipMemcpyAsync(lvl_max, d_lvl, runs * sizeof(int), hipMemcpyDeviceToHost, P.stream),
      "levels download");
  chk(hipStreamSynchronize(P.stream), "instance_maxima sync");
  for (int g = 0; g < runs; g++) {
    if (key[g] == 0) {
      vmax[g] = -HUGE_VAL;
      continue;
    }
    const unsigned long long u = (key[g] >> 63) ? (key[g] & 0x7fffffffffffffffull) : ~key[g];
    std::memcpy(&vmax[g], &u, 8);
  }
}

}  // namespace a5g
