// spark_regions.hpp -- stage regions of the batched forms as boxes of a
// block's guard-celled index space (host code), and the boundary cells of a
// box (host and device).
#pragma once
#include <stdint.h>

#include <utility>
#include <vector>

// the boundary-cell enumeration below is also used by device code
#if defined(__HIPCC__)
#define SPARK_HD __host__ __device__
#else
#define SPARK_HD
#endif

namespace spark {

// Boundary cells of an n[0] x n[1] x n[2] box of nd dimensions (n[d] = 1 for
// d >= nd), enumerated without the interior: first the outer layer(s) of
// dimension nd-1 over the whole cross section, then those of dimension nd-2
// over the cross section less the layers already counted, and so on; a
// dimension of one cell has one layer, otherwise two.  Within a layer the
// first dimension runs fastest.
SPARK_HD inline int boundary_layer_count(int nd, const int n[3], int d, int* layers, int* rest) {
  *layers = n[d] > 1 ? 2 : 1;
  int r = 1;
  for (int e = 0; e < nd; e++)
    if (e != d) r *= e < d ? n[e] : (n[e] > 2 ? n[e] - 2 : 0);
  *rest = r;
  return *layers * r;
}

SPARK_HD inline int boundary_count(int nd, const int n[3]) {
  int tot = 0, l, r;
  for (int d = nd - 1; d >= 0; d--) tot += boundary_layer_count(nd, n, d, &l, &r);
  return tot;
}

// cell c[0..2] (0-based in the box) of boundary cell t, 0 <= t < boundary_count
SPARK_HD inline void boundary_cell(int nd, const int n[3], int t, int c[3]) {
  c[0] = c[1] = c[2] = 0;
  for (int d = nd - 1; d >= 0; d--) {
    int l, r;
    const int cnt = boundary_layer_count(nd, n, d, &l, &r);
    if (t >= cnt) { t -= cnt; continue; }
    c[d] = (t / r) == 0 ? 0 : n[d] - 1;
    int q = t % r;
    for (int e = 0; e < nd; e++) {
      if (e == d) continue;
      const int m = e < d ? n[e] : n[e] - 2;
      c[e] = (e < d ? 0 : 1) + q % m;
      q /= m;
    }
    return;
  }
}

// a box of a block's guard-celled index space, bounds inclusive
struct Box {
  int x0, x1, y0, y1, z0, z1;
  double cost() const { return (double)(x1 - x0 + 1) * (y1 - y0 + 1) * (z1 - z0 + 3); }
};

// Stage region of a block that reads some guard sectors from same-level
// leaves: its interior and the sectors it computes (bit 13 + ox + 3 oy +
// 9 oz of `local`), each to depth e, as disjoint boxes of the guard-celled
// index space (g guard cells, nx x ny x nz interior).  Per column sector
// (ox, oy) the z extent is the union of its computed sectors (two runs when
// only the middle one is read from a leaf); column sectors of one row with
// equal z runs merge along x, and rows with equal boxes merge along y.
inline void sector_boxes(int nx, int ny, int nz, int g, int e, uint32_t local, std::vector<Box>& out) {
  auto span = [&](int o, int nn, int& lo, int& hi) {
    if (o < 0) { lo = g - e; hi = g - 1; }
    else if (o == 0) { lo = g; hi = g + nn - 1; }
    else { lo = g + nn; hi = g + nn + e - 1; }
  };
  std::vector<Box> prev;
  auto flush = [&](std::vector<Box>& row) { for (auto& r : row) out.push_back(r); row.clear(); };
  for (int oy = -1; oy <= 1; oy++) {
    int y0, y1;
    span(oy, ny, y0, y1);
    if (y1 < y0) continue;
    std::vector<Box> row;
    std::vector<std::pair<int, int>> runs_prev;
    int run_x0 = 0, run_x1 = -1;
    for (int ox = -1; ox <= 1; ox++) {
      int x0, x1;
      span(ox, nx, x0, x1);
      if (x1 < x0) continue;
      std::vector<std::pair<int, int>> runs;
      for (int oz = -1; oz <= 1; oz++) {
        const bool comp = (ox == 0 && oy == 0 && oz == 0) || ((local >> (13 + ox + 3 * oy + 9 * oz)) & 1u);
        int z0, z1;
        span(oz, nz, z0, z1);
        if (!comp || z1 < z0) continue;
        if (!runs.empty() && runs.back().second + 1 == z0) runs.back().second = z1;
        else runs.push_back({z0, z1});
      }
      if (run_x1 >= run_x0 && runs == runs_prev) { run_x1 = x1; continue; }
      for (auto& zr : runs_prev) row.push_back(Box{run_x0, run_x1, y0, y1, zr.first, zr.second});
      runs_prev = runs;
      run_x0 = x0; run_x1 = x1;
    }
    for (auto& zr : runs_prev) row.push_back(Box{run_x0, run_x1, y0, y1, zr.first, zr.second});
    // merge with the previous row when the boxes match in x and z
    bool same = !prev.empty() && prev.size() == row.size();
    for (size_t t = 0; same && t < row.size(); t++)
      same = row[t].x0 == prev[t].x0 && row[t].x1 == prev[t].x1 && row[t].z0 == prev[t].z0 && row[t].z1 == prev[t].z1;
    if (same) {
      for (auto& r : prev) r.y1 = y1;
    } else {
      flush(prev);
      prev = row;
    }
  }
  flush(prev);
}

}  // namespace spark
