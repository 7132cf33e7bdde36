// Checks spark::sector_boxes (the stage regions of blocks that read some
// guard sectors from same-level leaves): for every computed-sector mask
// given, and the three RK3 depths, the boxes are disjoint and cover exactly
// the interior plus the computed sectors to that depth.  Masks: the 64
// unions of whole guard faces (a level change across faces), then
// pseudo-random masks.  Prints "ok <cases>" or the first mismatch.
#include <cstdio>
#include <vector>

#include "spark_regions.hpp"

int main() {
  const int g = 10, n = 16, NB = n + 2 * g;
  const int depths[3] = {6, 3, 0};
  unsigned long long seed = 20261202ULL;
  long cases = 0;
  for (int t = 0; t < 2064; t++) {
    uint32_t m = 0;
    if (t < 64) {
      for (int f = 0; f < 6; f++) {
        if (!((t >> f) & 1)) continue;
        const int d = f / 2, s = (f & 1) ? 1 : -1;
        for (int oz = -1; oz <= 1; oz++)
          for (int oy = -1; oy <= 1; oy++)
            for (int ox = -1; ox <= 1; ox++) {
              const int o[3] = {ox, oy, oz};
              if (o[d] == s) m |= 1u << (13 + ox + 3 * oy + 9 * oz);
            }
      }
    } else {
      seed = seed * 6364136223846793005ULL + 1442695040888963407ULL;
      m = (uint32_t)(seed >> 20) & ((1u << 27) - 1);
    }
    m &= ~(1u << 13);
    for (int e : depths) {
      std::vector<spark::Box> out;
      spark::sector_boxes(n, n, n, g, e, m, out);
      // cover count of every cell of the guard-celled box
      std::vector<int> cnt((size_t)NB * NB * NB, 0);
      for (const auto& r : out) {
        if (r.x0 < 0 || r.y0 < 0 || r.z0 < 0 || r.x1 >= NB || r.y1 >= NB || r.z1 >= NB) {
          std::printf("box outside the block: mask %x depth %d\n", m, e);
          return 1;
        }
        for (int z = r.z0; z <= r.z1; z++)
          for (int y = r.y0; y <= r.y1; y++)
            for (int x = r.x0; x <= r.x1; x++) cnt[x + NB * (y + NB * z)]++;
      }
      for (int z = 0; z < NB; z++)
        for (int y = 0; y < NB; y++)
          for (int x = 0; x < NB; x++) {
            const auto side = [&](int c) { return c < g - e || c >= g + n + e ? 2 : c < g ? -1 : (c >= g + n ? 1 : 0); };
            const int ox = side(x), oy = side(y), oz = side(z);
            int want = 0;
            if (ox != 2 && oy != 2 && oz != 2) {
              const int s = 13 + ox + 3 * oy + 9 * oz;
              want = (s == 13 || ((m >> s) & 1u)) ? 1 : 0;
            }
            if (cnt[x + NB * (y + NB * z)] != want) {
              std::printf("mask %x depth %d cell %d %d %d covered %d times, want %d\n", m, e, x, y, z,
                          cnt[x + NB * (y + NB * z)], want);
              return 1;
            }
          }
      if (out.size() > 27) {
        std::printf("mask %x depth %d: %zu boxes\n", m, e, out.size());
        return 1;
      }
      cases++;
    }
  }
  std::printf("ok %ld\n", cases);
  return 0;
}
