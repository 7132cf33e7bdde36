// Checks spark::boundary_cell / boundary_count: for boxes of 1 to 3
// dimensions and every extent from 1 to 6 (and 16), the enumeration visits
// each cell with an index 0 or n-1 in some active dimension exactly once,
// and no other cell.  Prints "ok <boxes>" or the first failure.
#include <cstdio>
#include <vector>

#include "spark_regions.hpp"

int main() {
  int boxes = 0;
  const int ext[] = {1, 2, 3, 4, 5, 6, 16};
  for (int nd = 1; nd <= 3; nd++)
    for (int a : ext)
      for (int b : ext)
        for (int c : ext) {
          int n[3] = {a, nd > 1 ? b : 1, nd > 2 ? c : 1};
          if ((nd < 2 && b != 1) || (nd < 3 && c != 1)) continue;
          std::vector<int> seen(n[0] * n[1] * n[2], 0);
          const int cnt = spark::boundary_count(nd, n);
          for (int t = 0; t < cnt; t++) {
            int q[3];
            spark::boundary_cell(nd, n, t, q);
            for (int d = 0; d < 3; d++)
              if (q[d] < 0 || q[d] >= n[d]) {
                printf("fail: nd %d box %d %d %d t %d out of range\n", nd, n[0], n[1], n[2], t);
                return 1;
              }
            seen[q[0] + n[0] * (q[1] + n[1] * q[2])]++;
          }
          for (int k = 0; k < n[2]; k++)
            for (int j = 0; j < n[1]; j++)
              for (int i = 0; i < n[0]; i++) {
                const int q[3] = {i, j, k};
                bool bnd = false;
                for (int d = 0; d < nd; d++) bnd = bnd || q[d] == 0 || q[d] == n[d] - 1;
                const int s = seen[i + n[0] * (j + n[1] * k)];
                if (s != (bnd ? 1 : 0)) {
                  printf("fail: nd %d box %d %d %d cell %d %d %d seen %d\n", nd, n[0], n[1], n[2], i, j, k, s);
                  return 1;
                }
              }
          boxes++;
        }
  printf("ok %d\n", boxes);
  return 0;
}
