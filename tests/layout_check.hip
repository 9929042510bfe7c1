// Host-side check of the box-pool layouts (csrc/afgpu_common.h, Lay) and
// the index helpers of the multigrid kernels (csrc/afgpu_mg.h): built and
// run by tests/test_layout.py, no GPU involved.  Prints one line per box
// shape and exits non-zero on any mismatch.
#include <cstdio>
#include <vector>

#include "afgpu_mg.h"

using namespace afg;

template <int D, int NC>
int check() {
  using G = Geo<D, NC>;
  using L = Lay<D, NC, float>;
  static_assert(L::STRIDE == G::NCELL, "the colour-split layout keeps the box size");
  int bad = 0;
  // every cell of the box (ghosts included) has its own position in [0, STRIDE);
  // interior cells come first, ghosts after them
  std::vector<int> seen(L::STRIDE, 0);
  const int k0 = (D == 3) ? 0 : 0, k1 = (D == 3) ? NC + 1 : 0;
  for (int k = k0; k <= k1; k++)
    for (int j = 0; j <= NC + 1; j++)
      for (int i = 0; i <= NC + 1; i++) {
        const int x = L::at(i, j, k);
        if (x < 0 || x >= L::STRIDE) {
          bad++;
          continue;
        }
        if (seen[x]++) bad++;
        const bool in = i >= 1 && i <= NC && j >= 1 && j <= NC && (D == 2 || (k >= 1 && k <= NC));
        if (in && (L::in_at(i, j, k) != x || x >= G::NINT)) bad++;
        if (!in && x < G::NINT) bad++;
        if (LayStd<D, NC>::at(i, j, k) != G::idx(i, j, k)) bad++;
      }
  // colour_cell enumerates each colour's block in order
  for (int c = 0; c < 2; c++)
    for (int q = 0; q < G::NINT / 2; q++) {
      int i, j, k;
      colour_cell<D, NC>(c, q, i, j, k);
      if (((i + j + k) & 1) != c || L::in_at(i, j, k) != c * L::HALF + q) bad++;
    }
  // the half-sweep's thread q of half-sweep n updates colour_cell(n & 1, q)
  for (int n = 1; n <= 2; n++)
    for (int q = 0; q < G::NINT / 2; q++) {
      const int mm = q % (NC / 2);
      const int j = (q / (NC / 2)) % NC + 1;
      const int k = (D == 3) ? q / ((NC / 2) * NC) + 1 : 0;
      const int i = 2 - ((n ^ (D == 3 ? k + j : j)) & 1) + 2 * mm;
      int a, b, e;
      colour_cell<D, NC>(n & 1, q, a, b, e);
      if (a != i || b != j || e != k) bad++;
    }
  // neighbours of interior cells, in both layouts
  for (int k = (D == 3) ? 1 : 0; k <= ((D == 3) ? NC : 0); k++)
    for (int j = 1; j <= NC; j++)
      for (int i = 1; i <= NC; i++)
        for (int d = 0; d < D; d++)
          for (int o = -1; o <= 1; o += 2) {
            int p[3] = {i, j, k};
            p[d] += o;
            if (L::nb_at(i, j, k, d, o) != L::at(p[0], p[1], p[2])) bad++;
            if (LayStd<D, NC>::nb_at(i, j, k, d, o) != G::idx(p[0], p[1], p[2])) bad++;
          }
  // the 2^D children of each parent cell
  for (int k = (D == 3) ? 1 : 0; k <= ((D == 3) ? NC : 0); k += 2)
    for (int j = 1; j <= NC; j += 2)
      for (int i = 1; i <= NC; i += 2) {
        int p[8], ps[8];
        child_pos<D, NC, L>(i, j, k, p);
        child_pos<D, NC, LayStd<D, NC>>(i, j, k, ps);
        for (int c = 0; c < (D == 3 ? 8 : 4); c++) {
          const int a = i + (c & 1), b = j + ((c >> 1) & 1), e = (D == 3) ? k + ((c >> 2) & 1) : 0;
          if (p[c] != L::in_at(a, b, e) || ps[c] != G::idx(a, b, e)) bad++;
        }
      }
  printf("D=%d NC=%d stride %d mismatches %d\n", D, NC, L::STRIDE, bad);
  return bad;
}

int main() {
  const int bad = check<2, 4>() + check<2, 8>() + check<2, 16>() + check<2, 32>() + check<3, 4>() + check<3, 8>() +
                  check<3, 16>();
  return bad != 0;
}
