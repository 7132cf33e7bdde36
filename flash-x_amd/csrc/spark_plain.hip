// spark_plain.hip -- the plain FP64 form of the Spark RK stage on gfx950.
//
// One launch per kernel per stage over all blocks of one shape (blockIdx.y
// walks the block list).  The dataflow follows the reference per stage
// (Spark/Hydro.F90-mc:144-194): flattening -> face fluxes per direction ->
// conserved update + floors + Gamma EOS.  Each sweep turns its face fluxes
// into its term of the update's flux divergence per cell (the x sweep
// writes -dt/dx (F+ - F-), the y sweep subtracts dt/dy (G+ - G-), in the
// update's own order) and the last sweep finishes the sum and applies the
// update, so no face array reaches HBM.
// Built with -ffp-contract=off: every FP64 expression rounds exactly as the
// Fortran line it restates, so results match the CPU restatement bit for
// bit (IEEE division and square root on both sides).
#include <map>
#include <mutex>
#include <tuple>

#include "spark_batched.hpp"
#include "spark_launch.hpp"
#include "spark_physics.hpp"
#include "spark_regions.hpp"

namespace spark {

struct Lims { int lo[3], hi[3]; };

// Workgroups of a launch are dispatched round-robin over the 8 XCDs
// (linear id mod 8), each with its own L2.  For the cell-chunk kernels
// (grid.x = chunks of a block's region, grid.y = blocks) this remaps the
// linear id so that each XCD runs one contiguous range of (chunk, block)
// pairs: consecutive chunks of a block, which read each other's face and
// state rows (k+1 planes, j+1 rows), then share an L2.
__device__ __forceinline__ void xcd_chunk(int& cx, int& cy) {
  const int gx = (int)gridDim.x, G = gx * (int)gridDim.y;
  const int L = (int)blockIdx.x + gx * (int)blockIdx.y;
  const int q = G >> 3, r = G & 7, x = L & 7;
  const int l = x * q + (x < r ? x : r) + (L >> 3);
  cx = l % gx;
  cy = l / gx;
}

__device__ __forceinline__ Lims stage_lims(const DevParams& P, const BlockMeta& m, int ngoff) {
  Lims L;
  const int n[3] = {m.nx, m.ny, m.nz};
#pragma unroll
  for (int d = 0; d < 3; d++) {
    if (d < P.ndim) { L.lo[d] = P.ng - ngoff; L.hi[d] = P.ng + n[d] - 1 + ngoff; }
    else { L.lo[d] = 0; L.hi[d] = 0; }
  }
  return L;
}

__device__ __forceinline__ int64_t cix(const BlockMeta& m, int i, int j, int k) {
  return (int64_t)i + (int64_t)m.NX * ((int64_t)j + (int64_t)m.NY * (int64_t)k);
}
__device__ __forceinline__ int64_t fix(const BlockMeta& m, int i, int j, int k) {
  return (int64_t)i + (int64_t)(m.NX + 1) * ((int64_t)j + (int64_t)(m.NY + 1) * (int64_t)k);
}

// ---------------------------------------------------------------------------
// flattening over limits +/- 1 (hy_rk_getFaceFlux.F90-mc:200-276); the
// dimension count is a template parameter, so the loads of all directions
// issue before the first of them is used
template <int NDIM>
__global__ void __launch_bounds__(256) k_flatten(DevParams P, const BlockMeta* __restrict__ meta,
                                                 const int* __restrict__ blist, const double* __restrict__ S,
                                                 double* __restrict__ flat, int ngoff) {
  int cx, cy;
  xcd_chunk(cx, cy);
  const BlockMeta m = meta[blist[cy]];
  Lims L = stage_lims(P, m, ngoff);
  const int kx = 1, ky = (NDIM > 1), kz = (NDIM == 3);
  const int lo0 = L.lo[0] - kx, lo1 = L.lo[1] - ky, lo2 = L.lo[2] - kz;
  const int ex = L.hi[0] + kx - lo0 + 1, ey = L.hi[1] + ky - lo1 + 1, ez = L.hi[2] + kz - lo2 + 1;
  const int total = ex * ey * ez;
  const double* sp = S + m.base;
  const int64_t cs = m.cstride;
  const int64_t str[3] = {1, (int64_t)m.NX, (int64_t)m.NX * m.NY};
  const int velv[3] = {V_VELX, V_VELY, V_VELZ};
  const double betaMin = 0.75, betaMax = 0.85, Zmin = 0.25, Zmax = 0.75;
  for (int t = cx * blockDim.x + threadIdx.x; t < total; t += gridDim.x * blockDim.x) {
    int i = lo0 + t % ex, j = lo1 + (t / ex) % ey, k = lo2 + t / (ex * ey);
    int64_t c = cix(m, i, j, k);
    const double* p = sp + V_PRES * cs;
    const double p0 = p[c], g0 = sp[V_GAMC * cs + c];
    double pm2[3], pm1[3], pp1[3], pp2[3], u0[3], up1[3];
#pragma unroll
    for (int d = 0; d < NDIM; d++) {
      const int64_t s = str[d];
      const double* u = sp + velv[d] * cs;
      pm2[d] = p[c - 2 * s];
      pm1[d] = p[c - s];
      pp1[d] = p[c + s];
      pp2[d] = p[c + 2 * s];
      u0[d] = u[c];
      up1[d] = u[c + s];
    }
    double mn = 0.;
#pragma unroll
    for (int d = 0; d < NDIM; d++) {
      double dp1 = __builtin_fabs(pp1[d] - pm1[d]);
      double beta = dp1 / fmax2(Num<double>::flat_floor, __builtin_fabs(pp2[d] - pm2[d]));
      double Z = dp1 / (g0 * p0);
      double ft = fmax2(0., fmin2(1., (betaMax - beta) / (betaMax - betaMin)));
      if (up1[d] < u0[d]) ft = fmax2(ft, fmin2(1., (Zmax - Z) / (Zmax - Zmin)));
      else ft = 1.0;
      mn = (d == 0) ? ft : fmin2(mn, ft);
    }
    flat[m.obase + c] = mn;
  }
}

// ---------------------------------------------------------------------------
// conserved update + floors of one cell (hy_rk_updateSoln.F90-mc:211-287)
// and the Gamma EOS in MODE_DENS_EI (Eos_wrapped over the same limits).
// inner_of(n) gives flux n's divergence term of the cell before sources,
// -dt/dx (facP F+ - facM F-) - dt/dy (G+ - G-) - dt/dz (H+ - H-), in that
// order.  Every load of the cell (the U0 and stage-input state, the
// divergence terms) is issued before the first use waits: scalar slots
// past ns re-read scalar ns-1, so no load sits behind a branch.  u0 and us
// may alias (the last stage writes U): only the cell's own values are read
// and written.
template <int NSM, class InnerF>
__device__ __forceinline__ void update_cell(const DevParams& P, const BlockMeta& m, int i, int j, int64_t c,
                                            const double* u0, const double* ui, double* us, InnerF inner_of,
                                            double dt, double c1, double c2, double c3, int final_stage) {
  // the first stage's input is U0 itself: its values are loaded once
  const bool same = u0 == ui;
  const int ns = P.ns;
  const double dx = m.dx;
  const int64_t cs = m.cstride;
  // hydro part: its 10 state loads and 5 divergence terms issue together
  double A[F_NUM], B[F_NUM], div[F_NUM];
  {
    const int sv[5] = {V_DENS, V_VELX, V_VELY, V_VELZ, V_ENER};
#pragma unroll
    for (int q = 0; q < F_NUM; q++) {
      A[q] = u0[sv[q] * cs + c];
      div[q] = inner_of(q);
    }
    if (same) {
#pragma unroll
      for (int q = 0; q < F_NUM; q++) B[q] = A[q];
    } else {
#pragma unroll
      for (int q = 0; q < F_NUM; q++) B[q] = ui[sv[q] * cs + c];
    }
  }
  const double d0 = A[F_MASS];
#pragma unroll
  for (int q = 1; q < F_NUM; q++) A[q] = d0 * A[q];
  const double ds = B[F_MASS];
#pragma unroll
  for (int q = 1; q < F_NUM; q++) B[q] = ds * B[q];
  double Sg[F_NUM];
#pragma unroll
  for (int n = 0; n < F_NUM; n++) Sg[n] = 0.;
  if (P.use_grav) {
    Sg[F_XMOM] = B[F_MASS] * P.grav[0];
    Sg[F_YMOM] = B[F_MASS] * P.grav[1];
    Sg[F_ZMOM] = B[F_MASS] * P.grav[2];
    Sg[F_ENER] = B[F_XMOM] * P.grav[0] + B[F_YMOM] * P.grav[1] + B[F_ZMOM] * P.grav[2];
  }
  // geometric factors and sources (geoFacs, hy_rk_updateSoln.F90-mc:98-175)
  double facM, facP, rc, Sgeo[F_NUM] = {0., 0., 0., 0., 0.};
  if (geo_facs(P, m.id, i, j, dx, facM, facP, rc) && P.geom != 0)
    geo_sources(P, m.id, i, dx, rc, ui[V_DENS * cs + c], ui[V_VELX * cs + c], ui[V_VELZ * cs + c],
                ui[V_PRES * cs + c], B[F_YMOM], Sgeo);
#pragma unroll
  for (int n = 0; n < F_NUM; n++) {
    const double inner = div[n] + dt * (Sgeo[n] + Sg[n]);
    B[n] = c1 * A[n] + c2 * B[n] + c3 * inner;
  }
  const double dens = fmax2(B[F_MASS], P.smlrho);
  const double vx = B[F_XMOM] / dens, vy = B[F_YMOM] / dens, vz = B[F_ZMOM] / dens;
  const double ener = fmax2(P.smalle, B[F_ENER] / dens);
  const double ekin = .5 * (vx * vx + vy * vy + vz * vz);
  const double eint = fmax2(P.smalle, ener - ekin);
  // mass scalars, four at a time: the 8 state loads and 4 divergence terms
  // of a group issue together (slots past ns re-read scalar ns-1); the
  // scheduling barriers keep one group's values in registers at a time
  double Xs[NSM > 0 ? NSM : 1];  // the mass fractions (species first) the EOS mixes
#pragma unroll
  for (int v0 = 0; v0 < NSM; v0 += 4) {
    __builtin_amdgcn_sched_barrier(0);
    double a4[4], b4[4], d4[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const int v = v0 + q < ns ? v0 + q : ns - 1;
      const int64_t o = (int64_t)(V_NPROP + v) * cs + c;
      a4[q] = u0[o];
      d4[q] = inner_of(F_NUM + v);
    }
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const int v = v0 + q < ns ? v0 + q : ns - 1;
      b4[q] = same ? a4[q] : ui[(int64_t)(V_NPROP + v) * cs + c];
    }
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const int v = v0 + q;
      if (v < ns) {
        const double As = a4[q] * d0;
        double Bs = b4[q] * ds;
        const double inner = d4[q] + dt * (0.0 + 0.0);
        Bs = c1 * As + c2 * Bs + c3 * inner;
        Xs[v] = Bs / dens;
        us[(V_NPROP + v) * cs + c] = Xs[v];
      }
    }
  }
  __builtin_amdgcn_sched_barrier(0);
  EosOut e = eos_dens_ei(P, dens, vx, vy, vz, ener, eint, NSM > 0 ? Xs : nullptr);
  us[V_DENS * cs + c] = dens;
  us[V_VELX * cs + c] = vx;
  us[V_VELY * cs + c] = vy;
  us[V_VELZ * cs + c] = vz;
  us[V_PRES * cs + c] = e.pres;
  us[V_GAMC * cs + c] = e.gamc;
  us[V_EINT * cs + c] = e.eint;
  us[V_ENER * cs + c] = e.ener;
  if (final_stage) {
    us[V_TEMP * cs + c] = e.temp;
    us[V_GAME * cs + c] = e.game;
  }
}

// what a sweep needs for its part of the stage update
struct UpdArgs {
  const double* U0;  // step input (U); may alias Sout
  double* Sout;      // stage output state
  double* div;       // the divergence terms of the sweeps so far (5 + ns per cell)
  const double* dtp;
  double c1, c2, c3;
  int final_stage;
};

// ---------------------------------------------------------------------------
// face fluxes for one sweep direction.  A workgroup owns `lanes` pencils
// (DIR 0: rows of x-pencils; DIR 1/2: x-columns of y/z pencils, so adjacent
// lanes always touch adjacent x cells and every global access coalesces).
// Each thread reconstructs one cell; the i-1/2 face of that cell takes the
// i+1/2 state of the cell below through LDS.  Workgroups are ordered so
// that each XCD runs a contiguous range of (pencil chunk, plane, block)
// triples (xcd_chunk), so neighbouring chunks, which share cache lines at
// their edges, share an L2.  The 2R+1 stencil values come straight from
// global memory (L1/L2 serve the repeats), the next variable's while the
// current one is reconstructed: staging them in LDS, one variable per
// barrier, was measured slower (the kernel is bound by FP64 VALU issue,
// not by its loads).
template <int DIR, int RECON, int RIEMANN, int NSM, int STEP>
__global__ void __launch_bounds__(1024) __attribute__((amdgpu_waves_per_eu(STEP == 3 ? 4 : 5))) k_flux(DevParams P, const BlockMeta* __restrict__ meta,
                                               const int* __restrict__ blist, const double* __restrict__ S,
                                               const double* __restrict__ Ushok, const double* __restrict__ flat,
                                               int ngoff, int lanes, int* __restrict__ err,
                                               double* __restrict__ fb, double weight, int add, UpdArgs UA) {
  constexpr int R = ReconRadius<RECON>::R;
  constexpr int NFM = F_NUM + NSM;
  extern __shared__ double sm[];
  const int ns = P.ns;
  int bx, by;
  xcd_chunk(bx, by);
  const BlockMeta m = meta[blist[by]];
  Lims L = stage_lims(P, m, ngoff);
  constexpr int T1 = (DIR == 0) ? 1 : 0;
  constexpr int T2 = (DIR == 2) ? 1 : 2;
  const int Lr = L.hi[DIR] - L.lo[DIR] + 3;
  const int nt1 = L.hi[T1] - L.lo[T1] + 1;
  const int npen = nt1 * (L.hi[T2] - L.lo[T2] + 1);
  const int NT = blockDim.x;
  const int tid = threadIdx.x;
  int s, a;
  if (DIR == 0) { s = tid % Lr; a = tid / Lr; }
  else { a = tid % lanes; s = tid / lanes; }
  // pencils are numbered over the cross-section, T1 fastest, and dealt to
  // workgroups `lanes` at a time, so a workgroup's pencils may continue into
  // the next plane and no chunk of a plane is left mostly empty
  const int pen = bx * lanes + a;
  const bool valid = (a < lanes) && (pen < npen) && (s < Lr);
  int ijk[3];
  ijk[DIR] = L.lo[DIR] - 1 + s;
  ijk[T1] = L.lo[T1] + pen % nt1;
  ijk[T2] = L.lo[T2] + pen / nt1;
  const int64_t c = cix(m, ijk[0], ijk[1], ijk[2]);
  const int64_t st = (DIR == 0) ? 1 : (DIR == 1) ? (int64_t)m.NX : (int64_t)m.NX * m.NY;
  const double* sp = S + m.base;
  const int64_t cs = m.cstride;

  // hydro variables first: their face states and centred values go through
  // LDS, the Riemann solve runs; the mass scalars follow through the same
  // LDS slots (the solve reads only the hydro states, the scalar fluxes
  // only the solve's mass flux), so neither the LDS buffer nor the
  // registers hold all variables at once
  const double fl_c = (valid && RECON == RECON_WENO5 && P.use_flat) ? flat[m.obase + c] : 1.0;
  // variables in order dens, velx, vely, velz, pres, gamc, eint (rhoe =
  // dens * eint per cell), then the scalars; the stencil of variable v+1 is
  // loaded before variable v is reconstructed (software pipelining: the
  // scheduling barrier keeps the loads ahead of the reconstruction)
  auto var_off = [&](int v) -> int64_t {
    const int sv[7] = {V_DENS, V_VELX, V_VELY, V_VELZ, V_PRES, V_GAMC, V_EINT};
    return (int64_t)(v < R_NUM ? sv[v < R_NUM ? v : 0] : V_NPROP + (v - R_NUM < ns ? v - R_NUM : ns - 1)) * cs;
  };
  auto stencil = [&](int v, double (&u)[5]) {
#pragma unroll
    for (int q = 0; q < 5; q++) u[q] = (q >= 2 - R && q <= 2 + R) ? sp[var_off(v) + c + (q - 2) * st] : 0.0;
  };
  double umr[R_NUM];
  if (valid) {
    double ud[5], cur[5], nxt[5], up, um;
    stencil(0, nxt);
#pragma unroll
    for (int v = 0; v < R_NUM; v++) {
#pragma unroll
      for (int q = 0; q < 5; q++) cur[q] = nxt[q];
      if (v + 1 < R_NUM) stencil(v + 1, nxt);
      __builtin_amdgcn_sched_barrier(0);
      if (v == R_DENS) {
#pragma unroll
        for (int q = 0; q < 5; q++) ud[q] = cur[q];
      }
      if (v == R_RHOE) {
#pragma unroll
        for (int q = -R; q <= R; q++) cur[q + 2] = ud[q + 2] * cur[q + 2];
      }
      recon1<RECON>(cur, fl_c, up, um);
      sm[v * NT + tid] = up;
      umr[v] = um;
      if (v == R_DENS) sm[(R_NUM + 0) * NT + tid] = cur[2];
      else if (v <= R_VELZ) sm[(R_NUM + v) * NT + tid] = cur[2];
      else if (v == R_RHOE) sm[(R_NUM + 4) * NT + tid] = cur[2];
    }
    sm[(R_NUM + 5) * NT + tid] = Ushok[m.base + V_SHOK * cs + c];
  }
  __syncthreads();
  const bool face = valid && s >= 1;
  const int tn = (DIR == 0) ? tid - 1 : tid - lanes;
  double F[NFM];
  if (face) {
    double left[R_NUM], right[R_NUM];
#pragma unroll
    for (int v = 0; v < R_NUM; v++) {
      left[v] = sm[v * NT + tn];
      right[v] = umr[v];
    }
    left[R_DENS] = fmax2(P.smlrho, left[R_DENS]);
    left[R_PRES] = fmax2(P.smallp, left[R_PRES]);
    right[R_DENS] = fmax2(P.smlrho, right[R_DENS]);
    right[R_PRES] = fmax2(P.smallp, right[R_PRES]);
    const bool inShock = (sm[(R_NUM + 5) * NT + tn] != 0.0) || (sm[(R_NUM + 5) * NT + tid] != 0.0);
    int ierr;
    if (RIEMANN == RIEMANN_HLL) ierr = hll<DIR, double>(left, right, F);
    else ierr = hllc<DIR, double>(left, right, P.hybrid != 0, inShock, P.tiny, F);
    if (ierr) {
#pragma unroll
      for (int n = 0; n < F_NUM; n++) F[n] = 0.;
      atomicAdd(err, 1);
    }
    // artificial viscosity (Hydro.ini:285-301)
    {
      const double dL = sm[(R_NUM + 0) * NT + tn], dH = sm[(R_NUM + 0) * NT + tid];
      const double xL = sm[(R_NUM + 1) * NT + tn], xH = sm[(R_NUM + 1) * NT + tid];
      const double yL = sm[(R_NUM + 2) * NT + tn], yH = sm[(R_NUM + 2) * NT + tid];
      const double zL = sm[(R_NUM + 3) * NT + tn], zH = sm[(R_NUM + 3) * NT + tid];
      const double eL = sm[(R_NUM + 4) * NT + tn], eH = sm[(R_NUM + 4) * NT + tid];
      const double nL = (DIR == 0) ? xL : (DIR == 1) ? yL : zL;
      const double nH = (DIR == 0) ? xH : (DIR == 1) ? yH : zH;
      double cvisc = P.cvisc * fmax2(-(nH - nL), 0.);
      double VenerLo = dL * 0.5 * (xL * xL + yL * yL + zL * zL) + eL;
      double VenerHi = dH * 0.5 * (xH * xH + yH * yH + zH * zH) + eH;
      F[F_MASS] = F[F_MASS] + cvisc * (dL - dH);
      F[F_XMOM] = F[F_XMOM] + cvisc * (dL * xL - dH * xH);
      F[F_YMOM] = F[F_YMOM] + cvisc * (dL * yL - dH * yH);
      F[F_ZMOM] = F[F_ZMOM] + cvisc * (dL * zL - dH * zH);
      F[F_ENER] = F[F_ENER] + cvisc * (VenerLo - VenerHi);
    }
  }
  if (NSM > 0) {
    __syncthreads();  // every read of the hydro exchange is done
    double ums[NSM > 0 ? NSM : 1];
    if (valid) {
      double cur[5], nxt[5], up, um;
      stencil(R_NUM, nxt);
#pragma unroll
      for (int v = 0; v < NSM; v++) {
        if (v < ns) {
#pragma unroll
          for (int q = 0; q < 5; q++) cur[q] = nxt[q];
          if (v + 1 < NSM) stencil(R_NUM + v + 1, nxt);
          __builtin_amdgcn_sched_barrier(0);
          recon1<RECON>(cur, fl_c, up, um);
          sm[v * NT + tid] = up;
          ums[v] = um;
        }
      }
    }
    __syncthreads();
    if (face) {
      double xl[NSM > 0 ? NSM : 1], xr[NSM > 0 ? NSM : 1];
#pragma unroll
      for (int v = 0; v < NSM; v++) {
        if (v < ns) {
          xl[v] = sm[v * NT + tn];
          xr[v] = ums[v];
        }
      }
      if (P.nspec > 0) {
        // species: clamp to [smallx, 1] and renormalise (Spark/Hydro.ini:227-235)
        double sl = 0.0, sr = 0.0;
#pragma unroll
        for (int v = 0; v < NSM; v++) {
          if (v < P.nspec) {
            xl[v] = fmax2(P.smallx, fmin2(1.0, xl[v]));
            xr[v] = fmax2(P.smallx, fmin2(1.0, xr[v]));
            sl = sl + xl[v];
            sr = sr + xr[v];
          }
        }
        const double il = 1. / sl, ir = 1. / sr;
#pragma unroll
        for (int v = 0; v < NSM; v++) {
          if (v < P.nspec) {
            xl[v] = xl[v] * il;
            xr[v] = xr[v] * ir;
          }
        }
      }
#pragma unroll
      for (int v = 0; v < NSM; v++) {
        if (v < ns) {
          if (F[F_MASS] > 0.) F[F_NUM + v] = xl[v] * F[F_MASS];
          else F[F_NUM + v] = xr[v] * F[F_MASS];
        }
      }
    }
  }
  if (face) {
    // flux registers (addFluxes, Hydro_funcs.F90-mc:481-507) of the block's
    // two boundary faces of this direction: the register hand-off reads no
    // other face (Grid_putFluxData_block.F90:200-202, 278-279), so the
    // interior entries are never accumulated and stay zero
    if (fb != nullptr) {
      const int g = P.ng;
      const int nn[3] = {m.nx, m.ny, m.nz};
      bool bnd = ijk[DIR] == g || ijk[DIR] == g + nn[DIR];
#pragma unroll
      for (int d = 0; d < 3; d++)
        if (d != DIR && d < P.ndim) bnd = bnd && ijk[d] >= g && ijk[d] < g + nn[d];
      if (bnd) {
        int e[3], fi[3];
#pragma unroll
        for (int d = 0; d < 3; d++) {
          e[d] = d < P.ndim ? nn[d] : 1;
          fi[d] = d < P.ndim ? ijk[d] - g : 0;
        }
        e[DIR] += 1;
        const int64_t bstride = (int64_t)e[0] * e[1] * e[2];
        const int64_t o = m.bbase + fi[0] + (int64_t)e[0] * (fi[1] + (int64_t)e[1] * fi[2]);
        // every old entry is loaded before the first store (the compiler
        // cannot tell the entries apart, so a load after a store would wait
        // for it); a slot past F_NUM + ns re-reads the last entry
        double prev[NFM];
#pragma unroll
        for (int n = 0; n < NFM; n++) prev[n] = 0.0;
        if (add) {
#pragma unroll
          for (int n = 0; n < NFM; n++) prev[n] = fb[o + (n < F_NUM + ns ? n : F_NUM + ns - 1) * bstride];
        }
#pragma unroll
        for (int n = 0; n < NFM; n++) {
          if (n < F_NUM + ns) {
            const double w = weight * F[n];
            fb[o + n * bstride] = add ? prev[n] + w : w;
          }
        }
      }
    }
  }
  // the face fluxes of this direction through LDS (cell s owns its low
  // face, thread s + 1 computed its high face), then per cell of the region
  // (s = 1 .. Lr - 2) this direction's divergence term: STEP 1 (x, not the
  // last direction) writes it, STEP 2 (y before z) adds it to the x term,
  // STEP 3 (the last direction) adds it to the earlier terms and applies
  // the update
  __syncthreads();  // every read of the exchange buffer is done
  if (face) {
#pragma unroll
    for (int n = 0; n < NFM; n++)
      if (n < F_NUM + ns) sm[n * NT + tid] = F[n];
  }
  __syncthreads();
  if (!valid || s < 1 || s > Lr - 2) return;
  const int th = (DIR == 0) ? tid + 1 : tid + lanes;
  const double dt = *UA.dtp;
  const double dd = DIR == 0 ? m.dx : DIR == 1 ? m.dy : m.dz;
  double facM = 1.0, facP = 1.0, rc = 0.0;
  if (DIR == 0) (void)geo_facs(P, m.id, ijk[0], ijk[1], m.dx, facM, facP, rc);
  double* dv = UA.div + m.dbase + c;
  // flux n's term, added to the earlier directions' (prev) as the update
  // writes it: -(dt/dx (facP F+ - facM F-)), then - dt/dy (G+ - G-), ...
  auto term = [&](int n, double prev) -> double {
    const double lo = sm[n * NT + tid], hi = sm[n * NT + th];
    if (DIR == 0) return -(dt / dd * (facP * hi - facM * lo));
    return prev - dt / dd * (hi - lo);
  };
  if (STEP == 3) {
    auto inner_of = [&](int n) -> double { return term(n, DIR == 0 ? 0.0 : dv[n * m.cstride]); };
    update_cell<NSM>(P, m, ijk[0], ijk[1], c, UA.U0 + m.base, sp, UA.Sout + m.base, inner_of, dt, UA.c1, UA.c2,
                     UA.c3, UA.final_stage);
  } else {
    // every earlier term is loaded before the first store (a slot past
    // F_NUM + ns re-reads the last)
    double prev[NFM];
#pragma unroll
    for (int n = 0; n < NFM; n++) prev[n] = STEP == 2 ? dv[(n < F_NUM + ns ? n : F_NUM + ns - 1) * m.cstride] : 0.0;
#pragma unroll
    for (int n = 0; n < NFM; n++)
      if (n < F_NUM + ns) dv[n * m.cstride] = term(n, prev[n]);
  }
}

// ---------------------------------------------------------------------------
// shockDetect (Hydro_funcs.F90-mc:600-706): pass 1 zeroes SHOK and stores
// the sound speed over the guard-celled box, pass 2 flags shocks over
// interior +/- (NGUARD-1).
__global__ void __launch_bounds__(256) k_shock_vc(DevParams P, const BlockMeta* __restrict__ meta,
                                                  const int* __restrict__ blist, double* __restrict__ U,
                                                  double* __restrict__ vc, int band) {
  int cx, cy;
  xcd_chunk(cx, cy);
  const BlockMeta m = meta[blist[cy]];
  Lims L = stage_lims(P, m, band + 1);
  const int ex = L.hi[0] - L.lo[0] + 1, ey = L.hi[1] - L.lo[1] + 1, ez = L.hi[2] - L.lo[2] + 1;
  const int total = ex * ey * ez;
  double* up = U + m.base;
  const int64_t cs = m.cstride;
  for (int t = cx * blockDim.x + threadIdx.x; t < total; t += gridDim.x * blockDim.x) {
    const int i = L.lo[0] + t % ex, j = L.lo[1] + (t / ex) % ey, k = L.lo[2] + t / (ex * ey);
    const int64_t c = cix(m, i, j, k);
    up[V_SHOK * cs + c] = 0.;
    vc[m.obase + c] = Num<double>::sqr(up[V_GAMC * cs + c] * up[V_PRES * cs + c] / fmax2(up[V_DENS * cs + c], P.tiny));
  }
}

__global__ void __launch_bounds__(256) k_shock(DevParams P, const BlockMeta* __restrict__ meta,
                                               const int* __restrict__ blist, double* __restrict__ U,
                                               const double* __restrict__ vc, int band) {
  int cx, cy;
  xcd_chunk(cx, cy);
  const BlockMeta m = meta[blist[cy]];
  Lims L = stage_lims(P, m, band);
  const int ex = L.hi[0] - L.lo[0] + 1, ey = L.hi[1] - L.lo[1] + 1, ez = L.hi[2] - L.lo[2] + 1;
  const int total = ex * ey * ez;
  double* up = U + m.base;
  const double* cp = vc + m.obase;
  const int64_t cs = m.cstride;
  const int ky = (P.ndim > 1), kz = (P.ndim == 3);
  const int64_t sx = 1, sy = m.NX, sz = (int64_t)m.NX * m.NY;
  const double beta = 0.5, delta = 0.1;
  for (int t = cx * blockDim.x + threadIdx.x; t < total; t += gridDim.x * blockDim.x) {
    const int i = L.lo[0] + t % ex, j = L.lo[1] + (t / ex) % ey, k = L.lo[2] + t / (ex * ey);
    const int64_t c = cix(m, i, j, k);
    const double* p = up + V_PRES * cs;
    double minP = 0., minC = 0.;
    bool first = true;
    for (int kk = -kz; kk <= kz; kk++)
      for (int jj = -ky; jj <= ky; jj++)
        for (int ii = -1; ii <= 1; ii++) {
          const int64_t cc = cix(m, i + ii, j + jj, k + kk);
          if (first) { minP = p[cc]; minC = cp[cc]; first = false; }
          else { minP = fmin2(minP, p[cc]); minC = fmin2(minC, cp[cc]); }
        }
    double divv = up[V_VELX * cs + c + sx] - up[V_VELX * cs + c - sx];
    if (P.ndim > 1) divv = divv + up[V_VELY * cs + c + sy] - up[V_VELY * cs + c - sy];
    if (P.ndim == 3) divv = divv + up[V_VELZ * cs + c + sz] - up[V_VELZ * cs + c - sz];
    divv = 0.5 * divv;
    double gradPx = 0.5 * (p[c + sx] - p[c - sx]);
    double gradPy = 0., gradPz = 0.;
    if (P.ndim > 1) gradPy = 0.5 * (p[c + sy] - p[c - sy]);
    if (P.ndim == 3) gradPz = 0.5 * (p[c + sz] - p[c - sz]);
    const bool SW1 = (__builtin_fabs(gradPx) + __builtin_fabs(gradPy) + __builtin_fabs(gradPz) >= beta * minP);
    const bool SW2 = (-delta * minC >= divv);
    if (SW1 && SW2) up[V_SHOK * cs + c] = 1.;
  }
}

// shockDetect of 3D blocks in one pass: TW x TH column tiles of the
// (band + 1) box marched over z, one thread per column, with three z planes
// (tile + 1-cell ring) of pressure, sound speed and the three velocities in
// LDS, so every value is loaded once and the sound speed computed once per
// cell.  Cells of the band box get the flag, the outer layer 0, as the
// two-pass kernels above give.  Tiles are as wide as the box up to SHK_T
// per side.  NB: cells outside the interior are read through a neighbour
// table (nbr: 27 pool offsets per block, every guard sector a same-level
// leaf) from the leaves around the block, so its own guard cells need no
// fill; the flags are still written into the block's own box.
constexpr int SHK_T = 20;
constexpr int SHK_F = 5;  // staged fields: pres, sound speed, velx, vely, velz
template <bool NB>
__global__ void __launch_bounds__(1024) k_shock3(DevParams P, const BlockMeta* __restrict__ meta,
                                                 const int* __restrict__ blist, double* __restrict__ U, int band,
                                                 int TW, int TH, const int64_t* __restrict__ nbr) {
  extern __shared__ double shk_lds[];
  __shared__ int64_t s_nb[27];
  const int PW = TW + 2, PS = PW * (TH + 2);
  double* sp = shk_lds;           // [3][PS] pressure
  double* sc = shk_lds + 3 * PS;  // [3][PS] sound speed
  double* sv = shk_lds + 6 * PS;  // [3 velocities][3][PS]
  const int bid = blist[blockIdx.y];
  const BlockMeta m = meta[bid];
  const Lims L1 = stage_lims(P, m, band + 1), L0 = stage_lims(P, m, band);
  if (NB) {
    if (threadIdx.x < 27) s_nb[threadIdx.x] = nbr[(int64_t)bid * 27 + threadIdx.x];
    __syncthreads();
  }
  const int ex = L1.hi[0] - L1.lo[0] + 1;
  const int ntx = (ex + TW - 1) / TW;
  const int x0 = L1.lo[0] + (int)(blockIdx.x % ntx) * TW, y0 = L1.lo[1] + (int)(blockIdx.x / ntx) * TH;
  const int tid = threadIdx.x;
  const int i = x0 + tid % TW, j = y0 + tid / TW;
  const bool own = tid < TW * TH && i <= L1.hi[0] && j <= L1.hi[1];
  const int64_t cs = m.cstride;
  const double beta = 0.5, delta = 0.1;
  // plane k into its slot: cells of the (band + 1) box only (the others
  // are never read by a flagged cell)
  auto load = [&](int k) {
    const int s = (k - L1.lo[2]) % 3;
    for (int u = tid; u < PS; u += blockDim.x) {
      const int ii = x0 - 1 + u % PW, jj = y0 - 1 + u / PW;
      double pv = 0., cv = 0., vx = 0., vy = 0., vz = 0.;
      if (ii >= L1.lo[0] && ii <= L1.hi[0] && jj >= L1.lo[1] && jj <= L1.hi[1] && k <= L1.hi[2]) {
        const int64_t c = NB ? cell_at(m, s_nb, P.ng, ii, jj, k) : m.base + cix(m, ii, jj, k);
        pv = U[V_PRES * cs + c];
        cv = Num<double>::sqr(U[V_GAMC * cs + c] * pv / fmax2(U[V_DENS * cs + c], P.tiny));
        vx = U[V_VELX * cs + c];
        vy = U[V_VELY * cs + c];
        vz = U[V_VELZ * cs + c];
      }
      sp[s * PS + u] = pv;
      sc[s * PS + u] = cv;
      sv[(0 * 3 + s) * PS + u] = vx;
      sv[(1 * 3 + s) * PS + u] = vy;
      sv[(2 * 3 + s) * PS + u] = vz;
    }
  };
  load(L1.lo[2]);
  for (int k = L1.lo[2]; k <= L1.hi[2]; k++) {
    load(k + 1);
    __syncthreads();
    if (own) {
      double shok = 0.;
      if (i >= L0.lo[0] && i <= L0.hi[0] && j >= L0.lo[1] && j <= L0.hi[1] && k >= L0.lo[2] && k <= L0.hi[2]) {
        const int q = (j - y0 + 1) * PW + (i - x0 + 1);
        double minP = 0., minC = 0.;
        bool first = true;
        for (int kk = -1; kk <= 1; kk++) {
          const int s = (k + kk - L1.lo[2]) % 3;
          for (int jj = -1; jj <= 1; jj++)
            for (int ii = -1; ii <= 1; ii++) {
              const double pv = sp[s * PS + q + jj * PW + ii], cv = sc[s * PS + q + jj * PW + ii];
              if (first) { minP = pv; minC = cv; first = false; }
              else { minP = fmin2(minP, pv); minC = fmin2(minC, cv); }
            }
        }
        // slots of levels k-1, k, k+1 (k-1 >= L1.lo: a band cell is inside L1)
        const int sm = (k - 1 - L1.lo[2]) % 3, s0 = (k - L1.lo[2]) % 3, sq = (k + 1 - L1.lo[2]) % 3;
        const double* vxp = sv + (0 * 3 + s0) * PS + q;
        const double* vyp = sv + (1 * 3 + s0) * PS + q;
        const double* p0 = sp + s0 * PS + q;
        double divv = vxp[1] - vxp[-1];
        divv = divv + vyp[PW] - vyp[-PW];
        divv = divv + sv[(2 * 3 + sq) * PS + q] - sv[(2 * 3 + sm) * PS + q];
        divv = 0.5 * divv;
        const double gradPx = 0.5 * (p0[1] - p0[-1]);
        const double gradPy = 0.5 * (p0[PW] - p0[-PW]);
        const double gradPz = 0.5 * (sp[sq * PS + q] - sp[sm * PS + q]);
        const bool SW1 = (__builtin_fabs(gradPx) + __builtin_fabs(gradPy) + __builtin_fabs(gradPz) >= beta * minP);
        const bool SW2 = (-delta * minC >= divv);
        if (SW1 && SW2) shok = 1.;
      }
      U[m.base + V_SHOK * cs + cix(m, i, j, k)] = shok;
    }
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// CFL dt (Hydro_computeDt.F90:77-148).  Pass 1: per (block, chunk) the first
// cell in (k,j,i) order with the largest dt_ltemp*cfl.  Pass 2 (one
// workgroup): per block dt = cfl/dt_temp, then the minimum over blocks with
// the first block winning ties (Driver_computeDt's strict '<').
struct DtPart { double key, lt; long long idx; };

__global__ void __launch_bounds__(256) k_dt_part(DevParams P, const BlockMeta* __restrict__ meta,
                                                 const int* __restrict__ blist, const double* __restrict__ U,
                                                 DtPart* __restrict__ out, int chunks) {
  __shared__ double skey[4], slt[4];
  __shared__ long long sidx[4];
  const int b = blockIdx.y;
  const BlockMeta m = meta[blist[b]];
  Lims L = stage_lims(P, m, 0);
  const int ex = L.hi[0] - L.lo[0] + 1, ey = L.hi[1] - L.lo[1] + 1, ez = L.hi[2] - L.lo[2] + 1;
  const int total = ex * ey * ez;
  const double* up = U + m.base;
  const int64_t cs = m.cstride;
  const double dxi = 1.0 / m.dx, dyi = 1.0 / m.dy, dzi = 1.0 / m.dz;
  double bk = 0., bl = 0.;
  long long bi = 0x7fffffffffffffffLL;
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < total; t += chunks * blockDim.x) {
    const int i = L.lo[0] + t % ex, j = L.lo[1] + (t / ex) % ey, k = L.lo[2] + t / (ex * ey);
    const int64_t c = cix(m, i, j, k);
    const double snd2 = up[V_GAMC * cs + c] * up[V_PRES * cs + c] / up[V_DENS * cs + c];
    const double cf = Num<double>::sqr(snd2);
    double lt = (__builtin_fabs(up[V_VELX * cs + c]) + cf) * dxi;
    // spherical: y is theta, delyinv = 1/(r dtheta) (Hydro_computeDt.F90:104-111)
    const double dyc = (P.geom == 2 && P.ndim > 1) ? 1.0 / (P.geo[(int64_t)m.id * geo_stride(P.geo_nx, P.geo_ny) + i] * m.dy)
                                                    : dyi;
    if (P.ndim > 1) lt = fmax2(lt, (__builtin_fabs(up[V_VELY * cs + c]) + cf) * dyc);
    if (P.ndim > 2) lt = fmax2(lt, (__builtin_fabs(up[V_VELZ * cs + c]) + cf) * dzi);
    const double key = lt * P.cfl;
    if (key > bk || (key == bk && (long long)c < bi)) { bk = key; bl = lt; bi = c; }
  }
  // the combine is a total order (larger key, then smaller cell index), so
  // any reduction tree gives the result of the reference's first-cell scan:
  // a butterfly over the 64 lanes of each wavefront, then the 4 wavefront
  // results through LDS
  for (int off = 32; off > 0; off >>= 1) {
    const double k2 = __shfl_xor(bk, off, 64), l2 = __shfl_xor(bl, off, 64);
    const long long i2 = __shfl_xor(bi, off, 64);
    if (k2 > bk || (k2 == bk && i2 < bi)) { bk = k2; bl = l2; bi = i2; }
  }
  const int wv = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) { skey[wv] = bk; slt[wv] = bl; sidx[wv] = bi; }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int q = 1; q < (int)(blockDim.x >> 6); q++)
      if (skey[q] > bk || (skey[q] == bk && sidx[q] < bi)) { bk = skey[q]; bl = slt[q]; bi = sidx[q]; }
    DtPart r; r.key = bk; r.lt = bl; r.idx = bi;
    out[(size_t)blist[b] * chunks + blockIdx.x] = r;
  }
}

// A block's chunks are all loaded before the first is compared (the
// comparison picks the largest key, then the smallest index: a total order,
// so the result does not depend on the order or the thread count).
constexpr int DT_FINAL_THREADS = 1024;
constexpr int DT_MAX_CHUNKS = 16;
__global__ void __launch_bounds__(DT_FINAL_THREADS) k_dt_final(DevParams P, const DtPart* __restrict__ parts,
                                                               int nblocks, int chunks, double* __restrict__ dt_out,
                                                               long long* __restrict__ loc_out,
                                                               double* __restrict__ dt_hist,
                                                               int* __restrict__ hist_slot, DtCtl* __restrict__ ctl) {
  __shared__ double sdt[DT_FINAL_THREADS / 64];
  __shared__ int sb[DT_FINAL_THREADS / 64];
  __shared__ long long sc[DT_FINAL_THREADS / 64];
  double best = 1.7976931348623157e308;
  int bb = 0x7fffffff;
  long long bc = 0;
  for (int b = threadIdx.x; b < nblocks; b += blockDim.x) {
    double bk = 0., bl = 0.;
    long long bi = 0x7fffffffffffffffLL;
    DtPart r[DT_MAX_CHUNKS];
#pragma unroll
    for (int q = 0; q < DT_MAX_CHUNKS; q++)
      if (q < chunks) r[q] = parts[(size_t)b * chunks + q];
#pragma unroll
    for (int q = 0; q < DT_MAX_CHUNKS; q++)
      if (q < chunks && (r[q].key > bk || (r[q].key == bk && r[q].idx < bi))) {
        bk = r[q].key;
        bl = r[q].lt;
        bi = r[q].idx;
      }
    const double dtb = (bl != 0.0) ? P.cfl / bl : 1.7976931348623157e308;
    if (dtb < best || (dtb == best && b < bb)) { best = dtb; bb = b; bc = bi; }
  }
  // smaller dt, then the lower block index (Driver_computeDt's strict '<'
  // over blocks in order): a total order, reduced per wavefront by a
  // butterfly and then over the wavefronts
  for (int off = 32; off > 0; off >>= 1) {
    const double d2 = __shfl_xor(best, off, 64);
    const int b2 = __shfl_xor(bb, off, 64);
    const long long c2 = __shfl_xor(bc, off, 64);
    if (d2 < best || (d2 == best && b2 < bb)) { best = d2; bb = b2; bc = c2; }
  }
  const int wv = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) { sdt[wv] = best; sb[wv] = bb; sc[wv] = bc; }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int q = 1; q < (int)(blockDim.x >> 6); q++)
      if (sdt[q] < best || (sdt[q] == best && sb[q] < bb)) { best = sdt[q]; bb = sb[q]; bc = sc[q]; }
    double dt = best;
    if (ctl) {
      // Driver_evolveAll's first dt, then Driver_computeDt.F90:450-460:
      // dtNew = min(dtNew, dtOld * factor); dtNew = min(max(dtNew, dtMin), dtMax)
      if (ctl->first > 0.) {
        dt = ctl->first;
        ctl->first = 0.;
      } else {
        if (ctl->factor > 0.) dt = fmin2(dt, ctl->prev * ctl->factor);
        if (ctl->dtmin > 0.) dt = fmax2(dt, ctl->dtmin);
        if (ctl->dtmax > 0.) dt = fmin2(dt, ctl->dtmax);
      }
      ctl->prev = dt;
    }
    dt_out[0] = dt;
    if (dt_hist) {
      // the slot counter lives on the device so that a captured step graph
      // records each replay's dt in the next slot
      const int h = *hist_slot;
      dt_hist[h] = dt;
      *hist_slot = h + 1;
    }
    if (bb < nblocks) {
      loc_out[0] = bb;
      loc_out[1] = bc;
    }
  }
}

// ---------------------------------------------------------------------------
// Uniform block grid guard fill: same-level neighbour copy and OUTFLOW at
// the domain faces (Grid_bcApplyToRegion.F90:526-532), i.e. every guard
// cell takes the value at its global index clamped into the domain.
// Blocks are numbered x-fastest over an nb[0] x nb[1] x nb[2] grid.
__global__ void __launch_bounds__(256) k_ug_fill(DevParams P, const BlockMeta* __restrict__ meta, int nbx,
                                                 int nby, int nbz, double* __restrict__ U, VarList vars, int depth) {
  const int bid = blockIdx.y;
  const BlockMeta m = meta[bid];
  const int n[3] = {m.nx, (P.ndim > 1) ? m.ny : 1, (P.ndim > 2) ? m.nz : 1};
  const int gg[3] = {P.ng, (P.ndim > 1) ? P.ng : 0, (P.ndim > 2) ? P.ng : 0};
  const int dd[3] = {depth, (P.ndim > 1) ? depth : 0, (P.ndim > 2) ? depth : 0};
  const int ext[3] = {n[0] + 2 * dd[0], n[1] + 2 * dd[1], n[2] + 2 * dd[2]};
  const int total = ext[0] * ext[1] * ext[2];
  const int nb[3] = {nbx, nby, nbz};
  const int bi[3] = {bid % nbx, (bid / nbx) % nby, bid / (nbx * nby)};
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < total; t += gridDim.x * blockDim.x) {
    const int loc[3] = {gg[0] - dd[0] + t % ext[0], gg[1] - dd[1] + (t / ext[0]) % ext[1],
                        gg[2] - dd[2] + t / (ext[0] * ext[1])};
    bool inside = true;
    int sb[3], sl[3];
#pragma unroll
    for (int d = 0; d < 3; d++) {
      if (loc[d] < gg[d] || loc[d] >= gg[d] + n[d]) inside = false;
      int g = bi[d] * n[d] + (loc[d] - gg[d]);
      const int gmax = nb[d] * n[d] - 1;
      g = g < 0 ? 0 : (g > gmax ? gmax : g);
      sb[d] = g / n[d];
      sl[d] = g % n[d] + gg[d];
    }
    if (inside) continue;
    const BlockMeta ms = meta[sb[0] + nbx * (sb[1] + nby * sb[2])];
    const int64_t src = ms.base + cix(ms, sl[0], sl[1], sl[2]);
    const int64_t dst = m.base + cix(m, loc[0], loc[1], loc[2]);
    for (int q = 0; q < vars.n; q++) U[dst + vars.v[q] * m.cstride] = U[src + vars.v[q] * ms.cstride];
  }
}

// ---------------------------------------------------------------------------
// flux correction of every block's boundary cells (correctZone + per-cell
// EOS, hy_rk_correctFluxes.F90:126-338).  The reference sweeps the low and
// high layer of x, then of y, then of z, each sweep rebuilding the
// conserved state from the primitives the previous one wrote.  Here a
// thread owns one boundary cell and runs its x, y, z corrections in that
// order with the primitives carried in registers (a store and reload of a
// double is exact), so every cell gets the values of the three sweeps and
// the state is read and written once.  Lanes map to consecutive x cells.
// With a face mask (bit 2d+s: low / high face of dimension d), only cells
// on a masked face are corrected (each through all of its faces).
template <int NSM>
__global__ void __launch_bounds__(256) k_correct(DevParams P, const BlockMeta* __restrict__ meta,
                                                 const int* __restrict__ blist, double* __restrict__ U,
                                                 const double* __restrict__ fbx, const double* __restrict__ fby,
                                                 const double* __restrict__ fbz, const double* __restrict__ dtp,
                                                 const unsigned char* __restrict__ fmask) {
  constexpr int NFM = F_NUM + NSM;
  const int bid = blist[blockIdx.y];
  const BlockMeta m = meta[bid];
  Lims L = stage_lims(P, m, 0);
  const int n[3] = {L.hi[0] - L.lo[0] + 1, L.hi[1] - L.lo[1] + 1, L.hi[2] - L.lo[2] + 1};
  const int ns = NSM > 0 ? P.ns : 0, nd = P.ndim;
  const int nb = boundary_count(nd, n);  // the interior cells get no thread
  const int mask = fmask ? fmask[bid] : 0x3f;
  const double dt = *dtp;
  const int64_t cs = m.cstride;
  double* up = U + m.base;
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < nb; t += gridDim.x * blockDim.x) {
    int c[3];
    boundary_cell(nd, n, t, c);
    bool edge = false;
    for (int d = 0; d < nd; d++)
      edge = edge || (c[d] == 0 && (mask >> (2 * d) & 1)) || (c[d] == n[d] - 1 && (mask >> (2 * d + 1) & 1));
    if (!edge) continue;
    const int64_t cc = cix(m, L.lo[0] + c[0], L.lo[1] + c[1], L.lo[2] + c[2]);
    double dens = up[V_DENS * cs + cc];
    double vx = up[V_VELX * cs + cc], vy = up[V_VELY * cs + cc], vz = up[V_VELZ * cs + cc];
    double ener = up[V_ENER * cs + cc];
    double X[NSM > 0 ? NSM : 1];  // mass scalars (a slot past ns re-reads scalar ns-1)
#pragma unroll
    for (int v = 0; v < NSM; v++) X[v] = up[(V_NPROP + (v < ns ? v : ns - 1)) * cs + cc];
    EosOut o;
    for (int d = 0; d < nd; d++) {
      const double* fb = (d == 0) ? fbx : (d == 1) ? fby : fbz;
      int e[3] = {n[0], n[1], n[2]};
      e[d] += 1;
      const int64_t bstride = (int64_t)e[0] * e[1] * e[2];
      const double del = (d == 0) ? m.dx : (d == 1) ? m.dy : m.dz;
      for (int side = 0; side < 2; side++) {
        if (c[d] != (side ? n[d] - 1 : 0)) continue;
        int f[3] = {c[0], c[1], c[2]};
        if (side) f[d] += 1;
        const int64_t bo = m.bbase + (int64_t)f[0] + (int64_t)e[0] * ((int64_t)f[1] + (int64_t)e[1] * f[2]);
        // the face's register entries, all loaded before the first is used
        double b[NFM];
#pragma unroll
        for (int q = 0; q < NFM; q++) b[q] = fb[bo + (q < F_NUM + ns ? q : F_NUM + ns - 1) * bstride];
        double Us[NFM];
        const double d0 = dens;
        Us[F_MASS] = d0;
        Us[F_XMOM] = d0 * vx;
        Us[F_YMOM] = d0 * vy;
        Us[F_ZMOM] = d0 * vz;
        Us[F_ENER] = d0 * ener;
#pragma unroll
        for (int v = 0; v < NSM; v++) Us[F_NUM + v] = X[v] * d0;
        // geoFacs of the x faces (hy_rk_correctFluxes.F90:137-168, 341-358)
        double fac = 1.0;
        if (d == 0 && P.geom != 0) {
          double facM, facP, rc;
          geo_facs(P, m.id, L.lo[0] + c[0], L.lo[1] + c[1], m.dx, facM, facP, rc);
          fac = side ? facP : facM;
        }
#pragma unroll
        for (int q = 0; q < NFM; q++) {
          const double dF = side ? b[q] : -b[q];
          Us[q] = Us[q] - dt / del * (fac * dF);
        }
        dens = fmax2(Us[F_MASS], P.smlrho);
        vx = Us[F_XMOM] / dens;
        vy = Us[F_YMOM] / dens;
        vz = Us[F_ZMOM] / dens;
        const double en = fmax2(P.smalle, Us[F_ENER] / dens);
        const double ekin = .5 * (vx * vx + vy * vy + vz * vz);
        const double eint = fmax2(P.smalle, en - ekin);
#pragma unroll
        for (int v = 0; v < NSM; v++) X[v] = Us[F_NUM + v] / dens;
        o = eos_dens_ei(P, dens, vx, vy, vz, en, eint, X);
        ener = o.ener;
      }
    }
#pragma unroll
    for (int v = 0; v < NSM; v++)
      if (v < ns) up[(V_NPROP + v) * cs + cc] = X[v];
    up[V_DENS * cs + cc] = dens;
    up[V_VELX * cs + cc] = vx;
    up[V_VELY * cs + cc] = vy;
    up[V_VELZ * cs + cc] = vz;
    up[V_PRES * cs + cc] = o.pres;
    up[V_GAMC * cs + cc] = o.gamc;
    up[V_EINT * cs + cc] = o.eint;
    up[V_ENER * cs + cc] = o.ener;
    up[V_TEMP * cs + cc] = o.temp;
    up[V_GAME * cs + cc] = o.game;
  }
}

// ---------------------------------------------------------------------------
// AoS (unk: variable fastest, host-packed into a contiguous staging buffer)
// <-> SoA transposes.  var_index[v] gives the unk slot of canonical var v.
__global__ void __launch_bounds__(256) k_aos_to_soa(const BlockMeta* __restrict__ meta, const double* __restrict__ aos,
                                                    const int64_t* __restrict__ aos_off, int nunk,
                                                    const int* __restrict__ var_index, int nvar,
                                                    double* __restrict__ U) {
  const BlockMeta m = meta[blockIdx.y];
  const int total = m.NX * m.NY * m.NZ;
  const double* src = aos + aos_off[blockIdx.y];
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < total; t += gridDim.x * blockDim.x)
    for (int v = 0; v < nvar; v++) U[m.base + v * m.cstride + t] = src[(int64_t)t * nunk + var_index[v]];
}

__global__ void __launch_bounds__(256) k_soa_to_aos(DevParams P, const BlockMeta* __restrict__ meta,
                                                    const double* __restrict__ U, double* __restrict__ aos,
                                                    const int64_t* __restrict__ aos_off, int nunk,
                                                    const int* __restrict__ var_index, int nvar, int interior) {
  const BlockMeta m = meta[blockIdx.y];
  const int total = m.NX * m.NY * m.NZ;
  double* dst = aos + aos_off[blockIdx.y];
  Lims L = stage_lims(P, m, 0);
  for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < total; t += gridDim.x * blockDim.x) {
    if (interior) {
      const int i = t % m.NX, j = (t / m.NX) % m.NY, k = t / (m.NX * m.NY);
      if (i < L.lo[0] || i > L.hi[0] || j < L.lo[1] || j > L.hi[1] || k < L.lo[2] || k > L.hi[2]) continue;
    }
    for (int v = 0; v < nvar; v++) dst[(int64_t)t * nunk + var_index[v]] = U[m.base + v * m.cstride + t];
  }
}

// ---------------------------------------------------------------------------
// host-side launchers
static inline int grid_x_for(int total, int cap) {
  int g = (total + 255) / 256;
  return g < 1 ? 1 : (g > cap ? cap : g);
}

// waves of one kernel resident on a CU at a block size and dynamic LDS
// size (the runtime's occupancy calculator, cached per kernel and shape)
static int resident_waves(const void* kern, int threads, size_t shm) {
  static std::mutex mu;
  static std::map<std::tuple<const void*, int, size_t>, int> cache;
  std::lock_guard<std::mutex> g(mu);
  const auto key = std::make_tuple(kern, threads, shm);
  const auto it = cache.find(key);
  if (it != cache.end()) return it->second;
  int nb = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kern, threads, shm) != hipSuccess || nb < 1) nb = 1;
  const int res = nb * ((threads + 63) / 64);
  cache.emplace(key, res);
  return res;
}

template <int DIR, int RECON, int RIEMANN, int NSM, int STEP>
static hipError_t launch_flux_t(const PlainLaunch& L, int ngoff, const UpdArgs& UA) {
  // geometry of this stage (all blocks in the list share one shape)
  const int n[3] = {L.nx, L.ny, L.nz};
  int lo[3], hi[3];
  for (int d = 0; d < 3; d++) {
    if (d < L.P.ndim) { lo[d] = L.P.ng - ngoff; hi[d] = L.P.ng + n[d] - 1 + ngoff; }
    else { lo[d] = 0; hi[d] = 0; }
  }
  const int T1 = (DIR == 0) ? 1 : 0, T2 = (DIR == 2) ? 1 : 2;
  const int Lr = hi[DIR] - lo[DIR] + 3;
  const int nt1 = hi[T1] - lo[T1] + 1, nt2 = hi[T2] - lo[T2] + 1;
  // LDS slots per thread: the hydro face states, 5 centred values and the
  // shock flag; then the scalars' face states; then the face fluxes
  constexpr int LDSV = (R_NUM + 6 > F_NUM + NSM) ? R_NUM + 6 : F_NUM + NSM;
  // pencils per workgroup.  The LDS exchange buffer stays within 80 KiB and
  // the workgroup within 1024 threads.  A workgroup of l pencils runs
  // w = ceil(l Lr / 64) waves; workgroups of 2, 4 or 8 waves run best (3, 5,
  // 6, 7 or 9 measured 5-45 % slower per launch in a sweep over pencil
  // counts), and y/z sweeps keep at least 8 x columns per row of loads.
  // Among those counts: the fewest waves over the whole cross-section per
  // wave resident on a CU (so a count whose workgroups leave wave slots of
  // the CU empty loses), the largest count within 3 % of the best.  Without
  // such a count, the fewest waves.
  const int npen = nt1 * nt2;
  const int cap = 81920 / (int)(sizeof(double) * LDSV * Lr);
  const int lmax = std::max(1, std::min(std::min(1024 / Lr, cap), npen));
  const int lmin = DIR == 0 ? 1 : std::min(8, lmax);
  auto wg_waves = [&](int l) { return (l * Lr + 63) / 64; };
  auto waves_of = [&](int l) { return ((npen + l - 1) / l) * wg_waves(l); };
  auto shape_ok = [&](int l) {
    const int w = wg_waves(l);
    return w == 2 || w == 4 || w == 8;
  };
  // cost in units of 1/1000 wave per resident wave
  auto cost = [&](int l) {
    const int res = resident_waves(reinterpret_cast<const void*>(&k_flux<DIR, RECON, RIEMANN, NSM, STEP>), l * Lr,
                                   sizeof(double) * (size_t)LDSV * l * Lr);
    return (int64_t)waves_of(l) * 1000 / std::max(res, 1);
  };
  int64_t best = INT64_MAX;
  for (int l = lmin; l <= lmax; l++)
    if (shape_ok(l)) best = std::min(best, cost(l));
  int lanes = lmax;
  if (best < INT64_MAX) {
    for (int l = lmax; l >= lmin; l--)
      if (shape_ok(l) && cost(l) * 100 <= best * 103) { lanes = l; break; }
  } else {
    int fewest = 1 << 30;
    for (int l = 1; l <= lmax; l++)
      if (waves_of(l) <= fewest) { fewest = waves_of(l); lanes = l; }
  }
  const int chunks = (npen + lanes - 1) / lanes;
  dim3 grid(chunks, L.nblk);
  dim3 block(lanes * Lr);
  const size_t shm = sizeof(double) * (size_t)LDSV * block.x;
  if (L.prof) {
    // algorithmic traffic: every state value of the stencil box read once
    // (dens velx vely velz pres gamc eint shok, scalars, flat); per region
    // cell the divergence terms (5 + ns) written (STEP 1), read and written
    // (STEP 2) or read (STEP 3 after another direction); STEP 3 also reads
    // U0 and the stage input (5 + ns each) and writes the stage output
    // (8 + ns, + temp and game at the last stage).  Flops: 117 per variable
    // per reconstructed cell (WENO5-Z) + 175 per face (HLLC, viscosity,
    // floors) + 3 per flux per cell for the divergence term, + the update
    const double cells_in = (double)(Lr + 2 * ReconRadius<RECON>::R) * nt1 * nt2;
    const double recon_cells = (double)Lr * nt1 * nt2;
    const double faces = (double)(Lr - 1) * nt1 * nt2;
    const double box = (double)(Lr - 2) * nt1 * nt2;
    const int ns = L.P.ns;
    const double ndiv = (STEP == 1 ? 1 : STEP == 2 ? 2 : DIR > 0 ? 1 : 0);
    double bytes = 8.0 * L.nblk * (cells_in * (8 + ns + (RECON == RECON_WENO5 && L.P.use_flat)) +
                                   box * ndiv * (5 + ns));
    double flops = (double)L.nblk * (recon_cells * (7 + ns) * 117.0 + faces * (175.0 + ns) + box * 3.0 * (5 + ns));
    if (STEP == 3) {
      bytes += 8.0 * L.nblk * box * (2 * (5 + ns) + 8 + ns + 2 * UA.final_stage);
      flops += (double)L.nblk * box * (60.0 + 4.0 * (5 + ns));
    }
    static const char* names[3] = {"k_flux_x", "k_flux_y", "k_flux_z"};
    L.prof->begin(L.stream, names[DIR], bytes, flops);
  }
  hipLaunchKernelGGL((k_flux<DIR, RECON, RIEMANN, NSM, STEP>), grid, block, shm, L.stream, L.P, L.meta, L.blist,
                     L.Sin, L.U0, L.flat, ngoff, lanes, L.err, L.fb[DIR], L.weight, L.add_flux, UA);
  if (L.prof) L.prof->end(L.stream);
  return hipGetLastError();
}

template <int DIR, int NSM, int STEP>
static hipError_t launch_flux_r(const PlainLaunch& L, int ngoff, const UpdArgs& UA) {
  const int r = L.P.recon, q = L.P.riemann;
  if (q == RIEMANN_HLL) {
    switch (r) {
      case RECON_MP5: return launch_flux_t<DIR, RECON_MP5, RIEMANN_HLL, NSM, STEP>(L, ngoff, UA);
      case RECON_PLM: return launch_flux_t<DIR, RECON_PLM, RIEMANN_HLL, NSM, STEP>(L, ngoff, UA);
      case RECON_FOG: return launch_flux_t<DIR, RECON_FOG, RIEMANN_HLL, NSM, STEP>(L, ngoff, UA);
      default: return launch_flux_t<DIR, RECON_WENO5, RIEMANN_HLL, NSM, STEP>(L, ngoff, UA);
    }
  }
  switch (r) {
    case RECON_MP5: return launch_flux_t<DIR, RECON_MP5, RIEMANN_HLLC, NSM, STEP>(L, ngoff, UA);
    case RECON_PLM: return launch_flux_t<DIR, RECON_PLM, RIEMANN_HLLC, NSM, STEP>(L, ngoff, UA);
    case RECON_FOG: return launch_flux_t<DIR, RECON_FOG, RIEMANN_HLLC, NSM, STEP>(L, ngoff, UA);
    default: return launch_flux_t<DIR, RECON_WENO5, RIEMANN_HLLC, NSM, STEP>(L, ngoff, UA);
  }
}

hipError_t plain_stage(const PlainLaunch& L, int ngoff, double c1, double c2, double c3, int final_stage) {
  hipError_t e;
  const int n[3] = {L.nx, L.ny, L.nz};
  int boxf = 1;
  for (int d = 0; d < L.P.ndim; d++) boxf *= n[d] + 2 * ngoff + 2;
  if (L.P.use_flat && L.P.recon == RECON_WENO5) {
    if (L.prof) {
      int bx = 1, bxr = 1;
      for (int d = 0; d < L.P.ndim; d++) { bx *= n[d] + 2 * ngoff + 2; bxr *= n[d] + 2 * ngoff + 6; }
      L.prof->begin(L.stream, "k_flatten", 8.0 * L.nblk * ((double)bxr * (2 + L.P.ndim) + bx),
                    (double)L.nblk * bx * 25.0 * L.P.ndim);
    }
    const dim3 fg(grid_x_for(boxf, 512), L.nblk), fb(256);
    if (L.P.ndim == 3)
      hipLaunchKernelGGL(k_flatten<3>, fg, fb, 0, L.stream, L.P, L.meta, L.blist, L.Sin, L.flat, ngoff);
    else if (L.P.ndim == 2)
      hipLaunchKernelGGL(k_flatten<2>, fg, fb, 0, L.stream, L.P, L.meta, L.blist, L.Sin, L.flat, ngoff);
    else
      hipLaunchKernelGGL(k_flatten<1>, fg, fb, 0, L.stream, L.P, L.meta, L.blist, L.Sin, L.flat, ngoff);
    if (L.prof) L.prof->end(L.stream);
    if ((e = hipGetLastError()) != hipSuccess) return e;
  }
  // the last sweep applies the update; it reads the stage input around each
  // cell while writing the stage output, so the two must be different
  // buffers
  if (L.Sin == L.Sout) return hipErrorInvalidValue;
  const UpdArgs UA{L.U0, L.Sout, L.div, L.dtp, c1, c2, c3, final_stage};
  const int nd = L.P.ndim;
  if (L.P.ns == 0) {
    if (nd == 1) return launch_flux_r<0, 0, 3>(L, ngoff, UA);
    if ((e = launch_flux_r<0, 0, 1>(L, ngoff, UA)) != hipSuccess) return e;
    if (nd == 2) return launch_flux_r<1, 0, 3>(L, ngoff, UA);
    if ((e = launch_flux_r<1, 0, 2>(L, ngoff, UA)) != hipSuccess) return e;
    return launch_flux_r<2, 0, 3>(L, ngoff, UA);
  }
  if (nd == 1) return launch_flux_r<0, 8, 3>(L, ngoff, UA);
  if ((e = launch_flux_r<0, 8, 1>(L, ngoff, UA)) != hipSuccess) return e;
  if (nd == 2) return launch_flux_r<1, 8, 3>(L, ngoff, UA);
  if ((e = launch_flux_r<1, 8, 2>(L, ngoff, UA)) != hipSuccess) return e;
  return launch_flux_r<2, 8, 3>(L, ngoff, UA);
}

// band: SHOK is computed over interior +/- band (the reference uses
// NGUARD-1; only the cells the stage-1 face fluxes read are needed).
hipError_t plain_shock_detect(const PlainLaunch& L, double* U, double* scratch, int band, const int64_t* nbr) {
  const int n[3] = {L.nx, L.P.ndim > 1 ? L.ny : 1, L.P.ndim > 2 ? L.nz : 1};
  int b1 = 1, b0 = 1;
  for (int d = 0; d < L.P.ndim; d++) { b1 *= n[d] + 2 * band + 2; b0 *= n[d] + 2 * band; }
  const double nc1 = (double)b1 * L.nblk, nc0 = (double)b0 * L.nblk;
  if (L.P.ndim == 3) {
    const int e1 = n[0] + 2 * band + 2, e2 = n[1] + 2 * band + 2;
    const int tw = e1 < SHK_T ? e1 : SHK_T, th = e2 < SHK_T ? e2 : SHK_T;
    const int tiles = ((e1 + tw - 1) / tw) * ((e2 + th - 1) / th);
    const int threads = (tw * th + 63) / 64 * 64;
    const size_t lds = sizeof(double) * 3 * SHK_F * (tw + 2) * (th + 2);
    if (L.prof) L.prof->begin(L.stream, "k_shock", 8.0 * nc1 * 7, nc1 * 6 + nc0 * 70);
    if (nbr)
      hipLaunchKernelGGL(k_shock3<true>, dim3(tiles, L.nblk), dim3(threads), lds, L.stream, L.P, L.meta, L.blist, U,
                         band, tw, th, nbr);
    else
      hipLaunchKernelGGL(k_shock3<false>, dim3(tiles, L.nblk), dim3(threads), lds, L.stream, L.P, L.meta, L.blist, U,
                         band, tw, th, nbr);
    if (L.prof) L.prof->end(L.stream);
    return hipGetLastError();
  }
  if (nbr) return hipErrorInvalidValue;  // the neighbour-table path is 3D only
  if (L.prof) L.prof->begin(L.stream, "k_shock_vc", 8.0 * nc1 * 5, nc1 * 6);
  hipLaunchKernelGGL(k_shock_vc, dim3(grid_x_for(b1, 512), L.nblk), dim3(256), 0, L.stream, L.P, L.meta,
                     L.blist, U, scratch, band);
  if (L.prof) L.prof->end(L.stream);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  if (L.prof) L.prof->begin(L.stream, "k_shock", 8.0 * (nc1 * 5 + nc0), nc0 * 70);
  hipLaunchKernelGGL(k_shock, dim3(grid_x_for(b0, 512), L.nblk), dim3(256), 0, L.stream, L.P, L.meta,
                     L.blist, U, scratch, band);
  if (L.prof) L.prof->end(L.stream);
  return hipGetLastError();
}

hipError_t plain_dt_part(const PlainLaunch& L, const double* U, void* parts, int chunks) {
  const double nc = (double)L.nx * (L.P.ndim > 1 ? L.ny : 1) * (L.P.ndim > 2 ? L.nz : 1) * L.nblk;
  if (L.prof) L.prof->begin(L.stream, "k_dt_part", 8.0 * nc * (3 + L.P.ndim), nc * (6 + 4 * L.P.ndim));
  hipLaunchKernelGGL(k_dt_part, dim3(chunks, L.nblk), dim3(256), 0, L.stream, L.P, L.meta, L.blist, U,
                     (DtPart*)parts, chunks);
  if (L.prof) L.prof->end(L.stream);
  return hipGetLastError();
}

hipError_t plain_dt_final(hipStream_t st, const DevParams& P, const void* parts, int nblk_total, int chunks,
                          double* dt_out, long long* loc_out, double* dt_hist, int* hist_slot, DtCtl* ctl) {
  if (chunks < 1 || chunks > DT_MAX_CHUNKS) return hipErrorInvalidValue;
  hipLaunchKernelGGL(k_dt_final, dim3(1), dim3(DT_FINAL_THREADS), 0, st, P, (const DtPart*)parts, nblk_total, chunks,
                     dt_out, loc_out, dt_hist, hist_slot, ctl);
  return hipGetLastError();
}

__global__ void k_set_scalar(double* p, double v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) p[0] = v;
}

__global__ void k_set_int(int* p, int v) {
  if (threadIdx.x == 0 && blockIdx.x == 0) p[0] = v;
}

hipError_t plain_set_int(hipStream_t st, int* p, int v) {
  hipLaunchKernelGGL(k_set_int, dim3(1), dim3(64), 0, st, p, v);
  return hipGetLastError();
}

hipError_t plain_set_scalar(hipStream_t st, double* p, double v) {
  hipLaunchKernelGGL(k_set_scalar, dim3(1), dim3(64), 0, st, p, v);
  return hipGetLastError();
}

size_t plain_dt_part_size() { return sizeof(DtPart); }

hipError_t plain_ug_fill(const PlainLaunch& L, int nbx, int nby, int nbz, double* U, const VarList& vars, int depth) {
  const int n[3] = {L.nx, L.P.ndim > 1 ? L.ny : 1, L.P.ndim > 2 ? L.nz : 1};
  int box = 1, nin = 1;
  for (int d = 0; d < 3; d++) {
    box *= n[d] + (d < L.P.ndim ? 2 * depth : 0);
    nin *= n[d];
  }
  if (L.prof) L.prof->begin(L.stream, "k_ug_fill", 16.0 * L.nblk * (double)(box - nin) * vars.n, 0.0);
  hipLaunchKernelGGL(k_ug_fill, dim3(grid_x_for(box, 512), L.nblk), dim3(256), 0, L.stream, L.P, L.meta,
                     nbx, nby, nbz, U, vars, depth);
  if (L.prof) L.prof->end(L.stream);
  return hipGetLastError();
}

hipError_t plain_correct(const PlainLaunch& L, double* U, const unsigned char* fmask) {
  const int total = L.nx * (L.P.ndim > 1 ? L.ny : 1) * (L.P.ndim > 2 ? L.nz : 1);
  const int n3[3] = {L.nx, L.P.ndim > 1 ? L.ny : 1, L.P.ndim > 2 ? L.nz : 1};
  int inner = 1;
  const int n[3] = {L.nx, L.ny, L.nz};
  for (int d = 0; d < 3; d++) inner *= (d < L.P.ndim && n[d] > 2) ? n[d] - 2 : (d < L.P.ndim ? 0 : 1);
  // boundary cells: read 5+ns state fields and the registers, write 10+ns
  const double cells = (double)(total - inner) * L.nblk, nf = 5.0 + L.P.ns;
  if (L.prof) L.prof->begin(L.stream, "k_correct", 8.0 * cells * (nf + nf + 10.0 + L.P.ns), cells * (nf * 4 + 40.0));
  const int nbc = boundary_count(L.P.ndim, n3);
  if (L.P.ns == 0)
    hipLaunchKernelGGL(k_correct<0>, dim3(grid_x_for(nbc, 64), L.nblk), dim3(256), 0, L.stream, L.P, L.meta, L.blist,
                       U, L.fb[0], L.fb[1], L.fb[2], L.dtp, fmask);
  else
    hipLaunchKernelGGL(k_correct<8>, dim3(grid_x_for(nbc, 64), L.nblk), dim3(256), 0, L.stream, L.P, L.meta, L.blist,
                       U, L.fb[0], L.fb[1], L.fb[2], L.dtp, fmask);
  if (L.prof) L.prof->end(L.stream);
  return hipGetLastError();
}

hipError_t plain_aos_to_soa(hipStream_t st, const BlockMeta* meta, int nblk, int maxcells, const double* aos,
                            const int64_t* aos_off, int nunk, const int* var_index, int nvar, double* U) {
  hipLaunchKernelGGL(k_aos_to_soa, dim3(grid_x_for(maxcells, 512), nblk), dim3(256), 0, st, meta, aos, aos_off, nunk,
                     var_index, nvar, U);
  return hipGetLastError();
}

hipError_t plain_soa_to_aos(hipStream_t st, const DevParams& P, const BlockMeta* meta, int nblk, int maxcells,
                            const double* U, double* aos, const int64_t* aos_off, int nunk, const int* var_index,
                            int nvar, int interior) {
  hipLaunchKernelGGL(k_soa_to_aos, dim3(grid_x_for(maxcells, 512), nblk), dim3(256), 0, st, P, meta, U, aos, aos_off,
                     nunk, var_index, nvar, interior);
  return hipGetLastError();
}

}  // namespace spark
