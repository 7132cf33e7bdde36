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
  int s0, s1, s2;
  if (G.nshift >= 0) {  // n a power of two (the usual 8, 16, 32): shifts, no integer divide
    s0 = (gi[0] << sh) >> G.nshift;
    s1 = (gi[1] << sh) >> G.nshift;
    s2 = (gi[2] << sh) >> G.nshift;
  } else {
    s0 = (gi[0] << sh) / G.n;
    s1 = (gi[1] << sh) / G.n;
    s2 = (gi[2] << sh) / G.n;
  }
  return G.map[s0 + (int64_t)G.md[0] * (s1 + (int64_t)G.md[1] * s2)];
}

// global cell index of level L wrapped (PERIODIC), clamped (OUTFLOW,
// Grid_bcApplyToRegion.F90:526-532) or mirrored (REFLECT) into the domain;
// returns the mask of the dimensions whose normal velocity changes sign
__device__ __forceinline__ int amr_domain(const AmrGrid& G, int L, int gi[3]) {
  int flip = 0;
#pragma unroll
  for (int d = 0; d < 3; d++) {
    const int N = (G.nb0[d] * G.n) << (L - 1);
    // guard cells and prolongation neighbours lie less than one period
    // outside the domain, so one add or subtract wraps them
    if (G.bc[d] == AMR_BC_PERIODIC) gi[d] = gi[d] < 0 ? gi[d] + N : (gi[d] >= N ? gi[d] - N : gi[d]);
    else if (G.bc[d] == AMR_BC_REFLECT && (gi[d] < 0 || gi[d] > N - 1)) {
      gi[d] = gi[d] < 0 ? -1 - gi[d] : 2 * N - 1 - gi[d];
      flip |= 1 << d;
    } else gi[d] = gi[d] < 0 ? 0 : (gi[d] > N - 1 ? N - 1 : gi[d]);
  }
  return flip;
}

__device__ __forceinline__ double flipped(double x, int flip, int v) {
  return (v >= V_VELX && v <= V_VELZ && ((flip >> (v - V_VELX)) & 1)) ? -x : x;
}

// where the value of a level-L cell comes from: the containing cell of a
// same-level or coarser leaf, or the 8 cells of a finer leaf (averaged)
struct Src {
  int64_t off;
  bool avg;
};

// every block of a hierarchy has the same shape, so the strides of a
// source block are those of the block being filled
struct Shape {
  int64_t sy, sz, cs;
};

__device__ __forceinline__ Src amr_src(const BlockMeta* __restrict__ meta, const AmrGrid& G, int g, int L,
                                       const int gi[3]) {
  const int s = leaf_at(G, L, gi);
  const BlockMeta& ms = meta[s];
  const int lv = ms.level, NX = ms.NX;
  int c[3];
  const bool coarse = lv <= L;
#pragma unroll
  for (int d = 0; d < 3; d++) {
    const int o = G.org[3 * s + d];
    c[d] = coarse ? (gi[d] >> (L - lv)) - o + g : 2 * gi[d] - o + g;
    // a leaf more than one level finer is excluded when the hierarchy is
    // declared; keep the read inside the block regardless
    c[d] = c[d] < 0 ? 0 : (c[d] > NX - 2 ? NX - 2 : c[d]);
  }
  Src r;
  r.avg = !coarse;
  r.off = ms.base + acix(NX, c[0], c[1], c[2]);
  return r;
}

__device__ __forceinline__ double src_val(const double* __restrict__ U, const Src& r, const Shape& S, int v) {
  const double* p = U + r.off + v * S.cs;
  if (!r.avg) return p[0];
  double sum = 0.0;
  sum = sum + p[0];
  sum = sum + p[1];
  sum = sum + p[S.sy];
  sum = sum + p[1 + S.sy];
  sum = sum + p[S.sz];
  sum = sum + p[1 + S.sz];
  sum = sum + p[S.sy + S.sz];
  sum = sum + p[1 + S.sy + S.sz];
  return 0.125 * sum;
}

// monotonized-central slope of a coarse cell from its two differences
__device__ __forceinline__ double mc_slope(double a, double b) {
  if (a * b <= 0.0) return 0.0;
  const double t = __builtin_fabs(a) < __builtin_fabs(b) ? 2.0 * __builtin_fabs(a) : 2.0 * __builtin_fabs(b);
  const double c = 0.5 * __builtin_fabs(a + b);
  const double m = t < c ? t : c;
  return a > 0.0 ? m : -m;
}

// Guard fill (Grid_fillGuardCells over a Paramesh hierarchy): each guard
// cell of depth <= depth takes, with its global index wrapped, clamped or
// mirrored into the domain (the normal velocity flips sign across a
// reflecting face), the value of the leaf covering it: a copy from a
// same-level leaf, the mean of the 8 cells of a finer leaf, and from a
// coarser leaf a limited linear prolongation (the parent coarse cell plus,
// per dimension, the monotonized-central slope of its face neighbours times
// +-1/4; piecewise-constant injection feeds an odd-even mode into the fine
// blocks that grows at fine/coarse faces under gravity).
// smask (optional): per block the guard sectors (bit 13 + ox + 3 oy + 9 oz)
// to fill to `depth`; the other sectors are filled only to `keep` layers
__global__ void __launch_bounds__(256) k_amr_fill(DevParams P, const BlockMeta* __restrict__ meta,
                                                  const int* __restrict__ blist, AmrGrid G,
                                                  double* __restrict__ U, VarList vars, int depth,
                                                  const uint32_t* __restrict__ smask, int keep) {
  const int bid = blist ? blist[blockIdx.y] : (int)blockIdx.y;
  const BlockMeta m = meta[bid];
  const int n = G.n, g = P.ng, L = m.level;
  const int ext = n + 2 * depth;
  const int total = ext * ext * ext;
  const int o[3] = {G.org[3 * bid], G.org[3 * bid + 1], G.org[3 * bid + 2]};
  const Shape S{m.NX, (int64_t)m.NX * m.NX, m.cstride};
  const uint32_t mk = smask ? smask[bid] : ~0u;
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < total; t += gridDim.x * blockDim.x) {
    const int loc[3] = {g - depth + t % ext, g - depth + (t / ext) % ext, g - depth + t / (ext * ext)};
    bool inside = true;
    int gi[3], sec = 13, dep = 0;
#pragma unroll
    for (int d = 0; d < 3; d++) {
      const int sd = d == 0 ? 1 : d == 1 ? 3 : 9;
      if (loc[d] < g) { inside = false; sec -= sd; dep = max(dep, g - loc[d]); }
      else if (loc[d] >= g + n) { inside = false; sec += sd; dep = max(dep, loc[d] - g - n + 1); }
      gi[d] = o[d] + loc[d] - g;
    }
    if (inside) continue;
    if (dep > keep && !((mk >> sec) & 1u)) continue;
    const int flip = amr_domain(G, L, gi);
    const int64_t dst = m.base + acix(m.NX, loc[0], loc[1], loc[2]);
    const int s = leaf_at(G, L, gi);
    if (meta[s].level != L - 1) {
      // copy / restriction: the loads of a chunk of variables are issued
      // before its stores (a store to U could alias a later load, so the
      // compiler keeps them in program order)
      const Src r = amr_src(meta, G, g, L, gi);
      constexpr int CHV = 8;
      for (int q0 = 0; q0 < vars.n; q0 += CHV) {
        double val[CHV];
        // a slot past vars.n re-reads the last variable, so no load is
        // conditional
#pragma unroll
        for (int u = 0; u < CHV; u++) {
          const int v = vars.v[q0 + u < vars.n ? q0 + u : vars.n - 1];
          val[u] = flipped(src_val(U, r, S, v), flip, v);
        }
#pragma unroll
        for (int u = 0; u < CHV; u++)
          if (q0 + u < vars.n) U[dst + vars.v[q0 + u] * m.cstride] = val[u];
      }
      continue;
    }
    int gc[3] = {gi[0] >> 1, gi[1] >> 1, gi[2] >> 1};
    const Src c0 = amr_src(meta, G, g, L - 1, gc);
    Src sm[3], sp[3];
    int fm[3], fp[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
      int gm[3] = {gc[0], gc[1], gc[2]}, gp[3] = {gc[0], gc[1], gc[2]};
      gm[d] -= 1;
      gp[d] += 1;
      fm[d] = amr_domain(G, L - 1, gm);
      fp[d] = amr_domain(G, L - 1, gp);
      sm[d] = amr_src(meta, G, g, L - 1, gm);
      sp[d] = amr_src(meta, G, g, L - 1, gp);
    }
    // two variables at a time: their 14 coarse loads issue before the
    // first slope waits (a slot past vars.n re-reads the last variable)
    for (int q0 = 0; q0 < vars.n; q0 += 2) {
      double cc[2], cmv[2][3], cpv[2][3];
#pragma unroll
      for (int h = 0; h < 2; h++) {
        const int v = vars.v[q0 + h < vars.n ? q0 + h : vars.n - 1];
        cc[h] = src_val(U, c0, S, v);
#pragma unroll
        for (int d = 0; d < 3; d++) {
          cmv[h][d] = flipped(src_val(U, sm[d], S, v), fm[d], v);
          cpv[h][d] = flipped(src_val(U, sp[d], S, v), fp[d], v);
        }
      }
#pragma unroll
      for (int h = 0; h < 2; h++) {
        if (q0 + h < vars.n) {
          const int v = vars.v[q0 + h];
          const double c = cc[h];
          double val = c;
#pragma unroll
          for (int d = 0; d < 3; d++) {
            const double sl = mc_slope(c - cmv[h][d], cpv[h][d] - c);
            val = val + sl * ((gi[d] & 1) ? 0.25 : -0.25);
          }
          U[dst + v * m.cstride] = flipped(val, flip, v);
        }
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
__global__ void __launch_bounds__(256) k_flux_exchange(DevParams P, const BlockMeta* __restrict__ meta,
                                                       const int* __restrict__ blist, AmrGrid G,
                                                       const double* __restrict__ fb, double* __restrict__ fc,
                                                       int dir) {
  const int bid = blist ? blist[blockIdx.y] : (int)blockIdx.y;
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
    int across = side ? o[dir] + n : o[dir] - 1;
    if (G.bc[dir] == AMR_BC_PERIODIC) across = (across + gmax + 1) % (gmax + 1);
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

hipError_t amr_fill(const PlainLaunch& L, const AmrGrid& G, double* U, const VarList& vars, int depth,
                    const uint32_t* smask, int keep) {
  const int ext = G.n + 2 * depth;
  const int box = ext * ext * ext;
  const int nin = G.n * G.n * G.n;
  if (L.prof) L.prof->begin(L.stream, "k_amr_fill", 16.0 * L.nblk * (double)(box - nin) * vars.n, 0.0);
  hipLaunchKernelGGL(k_amr_fill, dim3(grid_x(box, 512), L.nblk), dim3(256), 0, L.stream, L.P, L.meta, L.blist, G, U,
                     vars, depth, smask, keep);
  if (L.prof) L.prof->end(L.stream);
  return hipGetLastError();
}

hipError_t amr_flux_exchange(const PlainLaunch& L, const AmrGrid& G, double* const fb[3], double* const fc[3]) {
  // algorithmic bytes: read own + up to 4 fine faces, write the correction
  const double faces = 2.0 * G.n * G.n * L.nblk * 3, nf = 5.0 + L.P.ns;
  if (L.prof) L.prof->begin(L.stream, "k_flux_exchange", 8.0 * faces * nf * 3, faces * nf * 4);
  for (int dir = 0; dir < 3; dir++) {
    hipLaunchKernelGGL(k_flux_exchange, dim3(grid_x(2 * G.n * G.n, 64), L.nblk), dim3(256), 0, L.stream, L.P, L.meta,
                       L.blist, G, fb[dir], fc[dir], dir);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return e;
  }
  if (L.prof) L.prof->end(L.stream);
  return hipSuccess;
}

}  // namespace spark
