// Host check of the box cell order (a5g_layout.hpp): for every box size the
// library supports, cidx is a bijection of the (nc+2)^D block onto the slot,
// colour c fills slots [c*half, (c+1)*half), its interior cells come first
// with the x-boundary cells of the rows at slot = row / 2 (i = 1) and
// nrow / 2 + row / 2 (i = nc), so each x face of a colour is one dense run,
// icell inverts the interior order, and cidx_in agrees with cidx on the
// interior.
#include <cstdio>
#include <vector>

#include "../a5_streamer_amd/csrc/a5g_layout.hpp"

template <int D>
static int check(int nc) {
  const int n = nc + 2, bl = D == 2 ? n * n : n * n * n, half = bl / 2;
  const int n_in = (D == 2 ? nc * nc : nc * nc * nc) / 2;
  std::vector<int> seen(bl, 0);
  int bad = 0;
  for (int k = 0; k < (D == 3 ? n : 1); k++)
    for (int j = 0; j < n; j++)
      for (int i = 0; i < n; i++) {
        const int x = a5g::cidx<D>(nc, i, j, k);
        const int col = (i + j + k) & 1;
        const bool in = i >= 1 && i <= nc && j >= 1 && j <= nc && (D == 2 || (k >= 1 && k <= nc));
        if (x < col * half || x >= (col + 1) * half) bad++;
        else seen[x]++;
        if (in) {
          const int q = x - col * half, row = D == 2 ? j - 1 : (j - 1) + nc * (k - 1);
          int ii, jj, kk;
          a5g::icell<D>(nc, q, col, ii, jj, kk);
          if (ii != i || jj != j || (D == 3 && kk != k)) bad++;
          const int nrow = D == 2 ? nc : nc * nc;
          if (i == 1 && q != row / 2) bad++;
          if (i == nc && q != nrow / 2 + row / 2) bad++;
          if (a5g::cidx_in<D>(nc, i, j, k) != x) bad++;
        } else if (x - col * half < n_in) {
          bad++;
        }
      }
  for (int x = 0; x < bl; x++) bad += seen[x] != 1;
  return bad;
}

int main() {
  int bad = 0;
  for (int nc = 2; nc <= 32; nc *= 2) bad += check<2>(nc);
  for (int nc = 2; nc <= 32; nc *= 2) bad += check<3>(nc);
  printf("layout check: %d errors\n", bad);
  return bad != 0;
}
