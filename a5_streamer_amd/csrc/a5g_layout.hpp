// Cell order inside a box slot of the a5g pool, shared by the kernels and
// the host conversion to and from the reference's cc(0:nc+1, ...) order.
// Compiled by hipcc for both sides, and by a plain C++ compiler for the
// layout's own host test.
#pragma once

#if defined(__HIPCC__)
#define A5G_HD __host__ __device__ __forceinline__
#else
#define A5G_HD inline
#endif

namespace a5g {

// Flat index of cell (i, j, k), ghosts included (0 <= i,j,k <= nc+1), in a
// box with nc (even) cells per dimension. Boxes are stored colour-split: the
// cells with i+j+k even fill the first half of the slot, the odd ones the
// second half. Within a colour half:
//  * the nc^D/2 interior cells come first. Each interior row (j, k) holds
//    nc/2 cells of a colour, one of them on the x boundary (i = 1 or nc).
//    The boundary cells come first: those at i = 1 (slot = row / 2), then
//    those at i = nc (slot = nrow / 2 + row / 2); then the other nc/2 - 1
//    cells of each row in row order (isl below);
//  * the ghost cells follow, in lexicographic order of the ghost pairs (a row
//    on the ghost layer holds (nc+2)/2 pairs, any other row one pair: its
//    cells i = 0 and nc+1).
// A red-black half-sweep then reads the other colour's interior and writes
// its own as contiguous runs that share no cache line with a ghost, and the
// face a neighbour reads is one dense run in every direction.
template <int D>
A5G_HD int isl(int nc, int i, int j, int k) {
  const int row = D == 2 ? j - 1 : (j - 1) + nc * (k - 1);
  const int nrow = D == 2 ? nc : nc * nc;
  if (i == 1) return row >> 1;
  if (i == nc) return (nrow >> 1) + (row >> 1);
  return nrow + row * ((nc >> 1) - 1) + ((i - 2) >> 1);
}

template <int D>
A5G_HD int cidx(int nc, int i, int j, int k) {
  const int n = nc + 2;
  const int half = D == 2 ? (n * n) >> 1 : (n * n * n) >> 1;
  const int base = ((i + j + k) & 1) ? half : 0;
  const bool in_i = (unsigned)(i - 1) < (unsigned)nc;
  const bool in_j = (unsigned)(j - 1) < (unsigned)nc;
  const bool in_k = D == 2 || (unsigned)(k - 1) < (unsigned)nc;
  if (in_i && in_j && in_k) return base + isl<D>(nc, i, j, k);
  const int n_in = D == 2 ? (nc * nc) >> 1 : (nc * nc * nc) >> 1;
  const int hr = n >> 1;  // pairs in a row on the ghost layer
  // ghost pairs of a plane (2D: the box) before row j, plus the pair of (i, j)
  const int in_plane = j == 0 ? (i >> 1) : (j == n - 1 ? hr + nc + (i >> 1) : hr + (j - 1));
  if (D == 2) return base + n_in + in_plane;
  const int full = n * hr, mid = 2 * hr + nc;  // ghost pairs of a ghost / inner plane
  const int gp = k == 0       ? j * hr + (i >> 1)
                 : k == n - 1 ? full + nc * mid + j * hr + (i >> 1)
                              : full + (k - 1) * mid + in_plane;
  return base + n_in + gp;
}

// cidx for a cell known to be interior (1 <= i,j,k <= nc)
template <int D>
A5G_HD int cidx_in(int nc, int i, int j, int k) {
  const int n = nc + 2;
  const int half = D == 2 ? (n * n) >> 1 : (n * n * n) >> 1;
  return (((i + j + k) & 1) ? half : 0) + isl<D>(nc, i, j, k);
}

// The interior cell of colour col at slot q of that colour's half
// (0 <= q < nc^D/2): the inverse of isl.
template <int D>
A5G_HD void icell(int nc, int q, int col, int& i, int& j, int& k) {
  const int nrow = D == 2 ? nc : nc * nc;
  if (q < nrow) {
    // a boundary cell: rows 2p and 2p+1 share slot p of their side; the
    // colour picks the row
    const bool low = q < (nrow >> 1);
    const int r2 = (low ? q : q - (nrow >> 1)) << 1;
    k = D == 3 ? r2 / nc + 1 : 0;
    i = low ? 1 : nc;
    const int j0 = r2 % nc + 1;
    j = ((i + j0 + k) & 1) == col ? j0 : j0 + 1;
    return;
  }
  const int w = (nc >> 1) - 1;  // cells of a colour in a row, boundary cell excluded
  const int row = (q - nrow) / w, m = q - nrow - row * w;
  j = row % nc + 1;
  k = D == 3 ? row / nc + 1 : 0;
  const bool odd = ((1 + j + k) & 1) == col;  // the colour's cells have odd i
  i = odd ? 3 + 2 * m : 2 + 2 * m;
}

}  // namespace a5g
