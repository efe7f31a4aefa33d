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
// second half. Within a colour half, the nc^D/2 interior cells come first,
// in lexicographic order of the interior row pairs (i = 2m+1, 2m+2), so the
// q-th interior cell of either colour sits at q; the ghost cells follow, in
// lexicographic order of the ghost pairs (a row on the ghost layer holds
// (nc+2)/2 pairs, any other row one pair: its cells i = 0 and nc+1).
//  * A red-black half-sweep reads the other colour's interior and writes
//    its own colour's interior as contiguous runs, and no ghost slot shares
//    a cache line with them.
//  * The ghost layer of a colour is one contiguous run after its interior.
template <int D>
A5G_HD int cidx(int nc, int i, int j, int k) {
  const int n = nc + 2;
  const int half = D == 2 ? (n * n) >> 1 : (n * n * n) >> 1;
  const int base = ((i + j + k) & 1) ? half : 0;
  const bool in_i = (unsigned)(i - 1) < (unsigned)nc;
  const bool in_j = (unsigned)(j - 1) < (unsigned)nc;
  const bool in_k = D == 2 || (unsigned)(k - 1) < (unsigned)nc;
  if (in_i && in_j && in_k) {
    const int q = D == 2 ? (i - 1) + nc * (j - 1) : (i - 1) + nc * ((j - 1) + nc * (k - 1));
    return base + (q >> 1);
  }
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
  const int q = D == 2 ? (i - 1) + nc * (j - 1) : (i - 1) + nc * ((j - 1) + nc * (k - 1));
  return (((i + j + k) & 1) ? half : 0) + (q >> 1);
}

}  // namespace a5g
