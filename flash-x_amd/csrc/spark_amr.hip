// spark_amr.hip -- block-hierarchy services on the device: guard fill over
// leaves of different refinement levels and the fine/coarse flux-register
// exchange that feeds hy_rk_correctFluxes.
//
// Every block has n^3 interior cells.  A block at level L (1-based) has its
// first interior cell at org[3*b..] in the global cell index space of level
// L; the level-1 grid is nb0[0] x nb0[1] x nb0[2] blocks.  The leaf map has
// one entry per finest-level block slot and names the leaf covering it, so
// the leaf under any cell of any level is one table read away.
#include <hip/hip_runtime.h>

#include "spark_launch.hpp"

namespace spark {
namespace {

__device__ __forceinline__ int64_t acix(int NX, int i, int j, int k) {
  return (int64_t)i + (int64_t)NX * ((int64_t)j + (int64_t)NX * (int64_t)k);
}

__device__ __forceinline__ int leaf_at(const AmrGrid& G, int L, const int gi[3]) {
  const int sh = G.lmax - L;
  const int s0 = (gi[0] << sh) / G.n, s1 = (gi[1] << sh) / G.n, s2 = (gi[2] << sh) / G.n;
  return G.map[s0 + (int64_t)G.md[0] * (s1 + (int64_t)G.md[1] * s2)];
}

// Guard fill (Grid_fillGuardCells over a Paramesh hierarchy): each guard
// cell of depth <= depth takes, with its global index clamped into the
// domain (OUTFLOW, Grid_bcApplyToRegion.F90:526-532), the value of the leaf
// covering it: a copy from a same-level leaf, the containing coarse cell of
// a coarser leaf (injection), or the mean of the 8 cells of a finer leaf.
__global__ void __launch_bounds__(256) k_amr_fill(DevParams P, const BlockMeta* __restrict__ meta, AmrGrid G,
                                                  double* __restrict__ U, VarList vars, int depth) {
  const int bid = blockIdx.y;
  const BlockMeta m = meta[bid];
  const int n = G.n, g = P.ng, L = m.level;
  const int ext = n + 2 * depth;
  const int total = ext * ext * ext;
  const int o[3] = {G.org[3 * bid], G.org[3 * bid + 1], G.org[3 * bid + 2]};
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < total; t += gridDim.x * blockDim.x) {
    const int loc[3] = {g - depth + t % ext, g - depth + (t / ext) % ext, g - depth + t / (ext * ext)};
    bool inside = true;
    int gi[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
      if (loc[d] < g || loc[d] >= g + n) inside = false;
      const int x = o[d] + loc[d] - g;
      const int xmax = ((G.nb0[d] * n) << (L - 1)) - 1;
      gi[d] = x < 0 ? 0 : (x > xmax ? xmax : x);
    }
    if (inside) continue;
    const int s = leaf_at(G, L, gi);
    const BlockMeta ms = meta[s];
    const int os[3] = {G.org[3 * s], G.org[3 * s + 1], G.org[3 * s + 2]};
    const int64_t dst = m.base + acix(m.NX, loc[0], loc[1], loc[2]);
    if (ms.level < L - 1 || ms.level > L + 1) continue;  // excluded when the hierarchy is declared
    if (ms.level <= L) {
      const int sh = L - ms.level;
      const int64_t src =
          ms.base + acix(ms.NX, (gi[0] >> sh) - os[0] + g, (gi[1] >> sh) - os[1] + g, (gi[2] >> sh) - os[2] + g);
      for (int q = 0; q < vars.n; q++) U[dst + vars.v[q] * m.cstride] = U[src + vars.v[q] * ms.cstride];
    } else {
      const int64_t src = ms.base + acix(ms.NX, 2 * gi[0] - os[0] + g, 2 * gi[1] - os[1] + g, 2 * gi[2] - os[2] + g);
      const int64_t sx = 1, sy = ms.NX, sz = (int64_t)ms.NX * ms.NX;
      for (int q = 0; q < vars.n; q++) {
        const double* p = U + src + vars.v[q] * ms.cstride;
        double sum = 0.0;
        sum = sum + p[0];
        sum = sum + p[sx];
        sum = sum + p[sy];
        sum = sum + p[sx + sy];
        sum = sum + p[sz];
        sum = sum + p[sx + sz];
        sum = sum + p[sy + sz];
        sum = sum + p[sx + sy + sz];
        U[dst + vars.v[q] * m.cstride] = 0.125 * sum;
      }
    }
  }
}

// Flux-register exchange for direction dir: on each boundary face of a
// block whose neighbour across is a finer leaf, the correction is the mean
// of the 4 fine faces (red_f = 0.25 for flux densities,
// PM4_package/source/amr_restrict_red.F90, summed in that order) minus the
// block's own face flux (Grid_getFluxCorrData_block.F90:182-216); every
// other boundary face gets 0.  fb / fc: per block (NFLUX, e0, e1, e2) face
// arrays, one flux at a time (stride e0*e1*e2), at BlockMeta::bbase.
__global__ void __launch_bounds__(256) k_flux_exchange(DevParams P, const BlockMeta* __restrict__ meta, AmrGrid G,
                                                       const double* __restrict__ fb, double* __restrict__ fc,
                                                       int dir) {
  const int bid = blockIdx.y;
  const BlockMeta m = meta[bid];
  const int n = G.n, L = m.level;
  const int o[3] = {G.org[3 * bid], G.org[3 * bid + 1], G.org[3 * bid + 2]};
  int e[3] = {n, n, n};
  e[dir] += 1;
  const int64_t bstride = (int64_t)e[0] * e[1] * e[2];
  const int d2 = (dir == 0) ? 1 : 0, d3 = (dir == 2) ? 1 : 2;
  const int nf = F_NUM + P.ns;
  const int gmax = ((G.nb0[dir] * n) << (L - 1)) - 1;
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < 2 * n * n; t += gridDim.x * blockDim.x) {
    const int side = t / (n * n);
    const int a2 = t % n, a3 = (t / n) % n;
    int f[3];
    f[dir] = side ? n : 0;
    f[d2] = a2;
    f[d3] = a3;
    const int64_t co = m.bbase + f[0] + (int64_t)e[0] * (f[1] + (int64_t)e[1] * f[2]);
    const int across = side ? o[dir] + n : o[dir] - 1;
    int s = -1;
    if (across >= 0 && across <= gmax) {
      int gi[3];
      gi[dir] = across;
      gi[d2] = o[d2] + a2;
      gi[d3] = o[d3] + a3;
      s = leaf_at(G, L, gi);
      if (meta[s].level != L + 1) s = -1;
    }
    if (s < 0) {
      for (int q = 0; q < nf; q++) fc[co + q * bstride] = 0.0;
      continue;
    }
    const BlockMeta ms = meta[s];
    int ff[3];
    ff[dir] = side ? 0 : n;
    ff[d2] = 2 * (o[d2] + a2) - G.org[3 * s + d2];
    ff[d3] = 2 * (o[d3] + a3) - G.org[3 * s + d3];
    int64_t fo[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
      int tt[3] = {ff[0], ff[1], ff[2]};
      tt[d2] += q & 1;
      tt[d3] += q >> 1;
      fo[q] = ms.bbase + tt[0] + (int64_t)e[0] * (tt[1] + (int64_t)e[1] * tt[2]);
    }
    for (int q = 0; q < nf; q++) {
      const int64_t qo = q * bstride;
      const double rest = (fb[fo[0] + qo] + fb[fo[1] + qo] + fb[fo[2] + qo] + fb[fo[3] + qo]) * 0.25;
      fc[co + qo] = rest - fb[co + qo];
    }
  }
}

inline int grid_x(int total, int cap) {
  int g = (total + 255) / 256;
  return g < 1 ? 1 : (g > cap ? cap : g);
}

}  // namespace

hipError_t amr_fill(const PlainLaunch& L, const AmrGrid& G, double* U, const VarList& vars, int depth) {
  const int ext = G.n + 2 * depth;
  const int box = ext * ext * ext;
  const int nin = G.n * G.n * G.n;
  if (L.prof) L.prof->begin(L.stream, "k_amr_fill", 16.0 * L.nblk * (double)(box - nin) * vars.n, 0.0);
  hipLaunchKernelGGL(k_amr_fill, dim3(grid_x(box, 512), L.nblk), dim3(256), 0, L.stream, L.P, L.meta, G, U, vars,
                     depth);
  if (L.prof) L.prof->end(L.stream);
  return hipGetLastError();
}

hipError_t amr_flux_exchange(const PlainLaunch& L, const AmrGrid& G, double* const fb[3], double* const fc[3]) {
  for (int dir = 0; dir < 3; dir++) {
    hipLaunchKernelGGL(k_flux_exchange, dim3(grid_x(2 * G.n * G.n, 64), L.nblk), dim3(256), 0, L.stream, L.P, L.meta,
                       G, fb[dir], fc[dir], dir);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

}  // namespace spark
