// spark_batched.hip -- the batched form of the Spark RK stage: one fused,
// persistent launch per RK stage over every block of the batch.
//
// This file is compiled twice:
//   SPARK_FUSED_REAL=float   FP32 reconstruction / Riemann arithmetic
//                            (the FP32-flux form, -ffp-contract=fast)
//   SPARK_FUSED_REAL=double  FP64 throughout, every expression in the
//                            reference's order (-ffp-contract=off), so it
//                            reproduces the plain FP64 form bit for bit
// Conserved-variable accumulation, floors, EOS and flux-register sums are
// FP64 in both.
//
// Work items are (block, x-y tile); a workgroup takes items from per-XCD
// queues (an XCD's queue holds whole blocks, so the halos that tiles of one
// block share stay in that XCD's L2) and steals from the other queues when
// its own runs dry.  Inside an item the workgroup marches up z:
//   * plane k (tile + 3 halo cells in x and y) is staged in LDS and every
//     thread keeps a 5-deep z window of its own column in registers;
//   * flattening, reconstruction, Riemann solve, artificial viscosity and
//     scalar fluxes run in Real; x/y face states and fluxes go through LDS,
//     the z face state/flux of the previous level stays in registers;
//   * the flux differences are accumulated in FP64 in the reference order
//     (x, y, z, source; hy_rk_updateSoln.F90-mc:257-265), then the update,
//     floors and Gamma EOS (hy_rk_updateSoln.F90-mc:278-287, Eos_wrapped).
// Stage input and output states are distinct buffers, so no tile reads a
// cell another tile writes.
#include <type_traits>

#include "spark_physics.hpp"
#include "spark_batched.hpp"

#ifndef SPARK_FUSED_REAL
#define SPARK_FUSED_REAL float
#endif

namespace spark {
namespace {

using Real = SPARK_FUSED_REAL;
constexpr bool kF64 = std::is_same<Real, double>::value;

__device__ __forceinline__ int64_t bcix(const BlockMeta& m, int i, int j, int k) {
  return (int64_t)i + (int64_t)m.NX * ((int64_t)j + (int64_t)m.NY * (int64_t)k);
}

// pencil slot v of a cell (R_ order: dens velx vely velz pres gamc rhoe,
// then scalars).  The FP32 form takes gamc = gamma (the Gamma EOS sets
// every cell's gamc to gamma); slots past the scalar count read as 0.
__device__ __forceinline__ Real load_prim(const double* __restrict__ s, int64_t cs, int64_t c, int v, int ns,
                                          double gamma) {
  switch (v) {
    case R_DENS: return (Real)s[V_DENS * cs + c];
    case R_VELX: return (Real)s[V_VELX * cs + c];
    case R_VELY: return (Real)s[V_VELY * cs + c];
    case R_VELZ: return (Real)s[V_VELZ * cs + c];
    case R_PRES: return (Real)s[V_PRES * cs + c];
    case R_GAMC: return kF64 ? (Real)s[V_GAMC * cs + c] : (Real)gamma;
    case R_RHOE: return (Real)(s[V_DENS * cs + c] * s[V_EINT * cs + c]);
    default: return (v - R_NUM < ns) ? (Real)s[(V_NPROP + v - R_NUM) * cs + c] : Real(0.);
  }
}

// Miller-Colella flattening term along one direction (hy_rk_getFaceFlux.F90-mc:226-236)
__device__ __forceinline__ Real flat_dir(Real pm2, Real pm1, Real p0, Real pp1, Real pp2, Real g0, Real u0, Real up1) {
  const Real betaMin = Real(0.75), betaMax = Real(0.85), Zmin = Real(0.25), Zmax = Real(0.75);
  const Real dp1 = __builtin_fabs(pp1 - pm1);
  const Real beta = Num<Real>::div(dp1, fmax2(Num<Real>::flat_floor, (Real)__builtin_fabs(pp2 - pm2)));
  const Real Z = Num<Real>::div(dp1, g0 * p0);
  Real ft = fmax2(Real(0.), fmin2(Real(1.), Num<Real>::div(betaMax - beta, betaMax - betaMin)));
  if (up1 < u0) ft = fmax2(ft, fmin2(Real(1.), Num<Real>::div(Zmax - Z, Zmax - Zmin)));
  else ft = Real(1.0);
  return ft;
}

struct RParams {
  Real gamma, cvisc, smlrho, smallp, tiny;
};

// face flux from the left/right face states (ensurePhysicalState,
// Riemann solve, artificial viscosity, scalar upwinding).  c* hold the two
// cells' centred dens, velx, vely, velz, rhoe.  (Hydro.ini:222-247,285-301)
template <int DIR, int RIEMANN, int NSM>
__device__ __forceinline__ void face_flux(const RParams& rp, int ns, Real* L, Real* R, const Real* cl, const Real* cr,
                                          Real* F, int* err) {
  L[R_DENS] = fmax2(rp.smlrho, L[R_DENS]);
  L[R_PRES] = fmax2(rp.smallp, L[R_PRES]);
  R[R_DENS] = fmax2(rp.smlrho, R[R_DENS]);
  R[R_PRES] = fmax2(rp.smallp, R[R_PRES]);
  if (!kF64) { L[R_GAMC] = rp.gamma; R[R_GAMC] = rp.gamma; }
  int ierr;
  if (RIEMANN == RIEMANN_HLL) ierr = hll<DIR, Real>(L, R, F);
  else ierr = hllc<DIR, Real>(L, R, false, false, rp.tiny, F);
  if (ierr) {
#pragma unroll
    for (int n = 0; n < F_NUM; n++) F[n] = Real(0.);
    atomicAdd(err, 1);
  }
  const Real nL = cl[1 + DIR], nH = cr[1 + DIR];
  const Real cvisc = rp.cvisc * fmax2(-(nH - nL), Real(0.));
  const Real VenerLo = cl[0] * Real(0.5) * (cl[1] * cl[1] + cl[2] * cl[2] + cl[3] * cl[3]) + cl[4];
  const Real VenerHi = cr[0] * Real(0.5) * (cr[1] * cr[1] + cr[2] * cr[2] + cr[3] * cr[3]) + cr[4];
  F[F_MASS] = F[F_MASS] + cvisc * (cl[0] - cr[0]);
  F[F_XMOM] = F[F_XMOM] + cvisc * (cl[0] * cl[1] - cr[0] * cr[1]);
  F[F_YMOM] = F[F_YMOM] + cvisc * (cl[0] * cl[2] - cr[0] * cr[2]);
  F[F_ZMOM] = F[F_ZMOM] + cvisc * (cl[0] * cl[3] - cr[0] * cr[3]);
  F[F_ENER] = F[F_ENER] + cvisc * (VenerLo - VenerHi);
#pragma unroll
  for (int s = 0; s < NSM; s++) {
    if (s < ns) {
      if (F[F_MASS] > Real(0.)) F[F_NUM + s] = L[R_NUM + s] * F[F_MASS];
      else F[F_NUM + s] = R[R_NUM + s] * F[F_MASS];
    }
  }
}

template <int NSM, int RECON, int RIEMANN>
__global__ void __launch_bounds__(FUSED_THREADS) k_fused(FusedArgs A) {
  constexpr int NV = R_NUM + NSM;
  constexpr int NF = F_NUM + NSM;
  constexpr int NT = FUSED_THREADS;
  extern __shared__ __align__(16) unsigned char lds_raw[];
  Real* lds = reinterpret_cast<Real*>(lds_raw);
  __shared__ int s_item;
  const int tid = threadIdx.x;
  const DevParams& P = A.P;
  const RParams rp{(Real)P.gamma, (Real)P.cvisc, (Real)P.smlrho, (Real)P.smallp, (Real)P.tiny};
  const double dt = *A.dtp;
  const int ns = P.ns;
  int q = blockIdx.x % FUSED_XCDS;  // the queue of this workgroup's XCD first
  int tried = 0;
  Real* plane = lds;  // [NV][PW*PH]
  for (;;) {
    if (tid == 0) {
      int it = -1;
      while (tried < FUSED_XCDS) {
        const int pos = atomicAdd(&A.counters[q], 1);
        if (pos < A.qlen[q]) { it = A.qoff[q] + pos; break; }
        q = (q + 1) % FUSED_XCDS;
        tried++;
      }
      s_item = it;
    }
    __syncthreads();
    const int it = s_item;
    __syncthreads();
    if (it < 0) break;
    const FusedItem I = A.items[it];
    const BlockMeta m = A.meta[I.blk];
    const int W = I.tx + 2, H = I.ty + 2, PW = I.tx + 6, PH = I.ty + 6;
    const int psz = PW * PH;
    Real* xex = lds + NV * psz;  // [NV][NT] x face states u+, then x fluxes
    Real* yex = xex + NV * NT;   // [NV][NT] y face states u+, then y fluxes
    const bool active = tid < W * H;
    const int px = tid % W, py = tid / W;
    const int i = I.x0 - 1 + px, j = I.y0 - 1 + py;
    const bool rowx = active && py >= 1 && py <= I.ty;  // x faces / x reconstruction rows
    const bool coly = active && px >= 1 && px <= I.tx;  // y faces / y reconstruction columns
    const bool inner = rowx && coly;                    // cells this item updates
    const int64_t cs = m.cstride;
    const int64_t sz = (int64_t)m.NX * m.NY;
    const double* sin = A.Sin + m.base;
    const double* u0p = A.U0 + m.base;
    double* sout = A.Sout + m.base;
    const double dtdx = dt / m.dx, dtdy = dt / m.dy, dtdz = dt / m.dz;
    const int64_t colc = bcix(m, active ? i : I.x0, active ? j : I.y0, 0);

    Real win[NV][5];  // z window k-2 .. k+2 of this thread's column
    if (active) {
#pragma unroll
      for (int w = 1; w < 5; w++) {
        const int64_t c = colc + (int64_t)(I.z0 - 4 + w) * sz;
#pragma unroll
        for (int v = 0; v < NV; v++) win[v][w] = load_prim(sin, cs, c, v, ns, P.gamma);
      }
    }
    Real upz[NV];     // z face state u+ of level k-1
    Real Hprev[NF];   // z flux at the low face of level k-1
    double dUxy[NF];  // -dt/dx dF - dt/dy dG of level k-1
#pragma unroll
    for (int v = 0; v < NV; v++) upz[v] = Real(0.);
#pragma unroll
    for (int n = 0; n < NF; n++) { Hprev[n] = Real(0.); dUxy[n] = 0.; }

    for (int k = I.z0 - 1; k <= I.z1 + 1; k++) {
      const bool xyk = (k >= I.z0) && (k <= I.z1);
      // advance the z window to k-2..k+2 and stage plane k
      if (active) {
#pragma unroll
        for (int v = 0; v < NV; v++) {
#pragma unroll
          for (int w = 0; w < 4; w++) win[v][w] = win[v][w + 1];
        }
        const int64_t c = colc + (int64_t)(k + 2) * sz;
#pragma unroll
        for (int v = 0; v < NV; v++) win[v][4] = load_prim(sin, cs, c, v, ns, P.gamma);
        const int pc = (py + 2) * PW + (px + 2);
#pragma unroll
        for (int v = 0; v < NV; v++) plane[v * psz + pc] = win[v][2];
      }
      for (int qq = tid; qq < psz; qq += NT) {
        const int qx = qq % PW, qy = qq / PW;
        if (qx >= 2 && qx < W + 2 && qy >= 2 && qy < H + 2) continue;
        const int64_t c = bcix(m, I.x0 - 3 + qx, I.y0 - 3 + qy, k);
#pragma unroll
        for (int v = 0; v < NV; v++) plane[v * psz + qq] = load_prim(sin, cs, c, v, ns, P.gamma);
      }
      __syncthreads();

      const int pc = (py + 2) * PW + (px + 2);
      Real flat = Real(1.);
      if (RECON == RECON_WENO5 && P.use_flat && active) {
        const Real* pp = plane + R_PRES * psz;
        const Real g0 = plane[R_GAMC * psz + pc];
        const Real fx = flat_dir(pp[pc - 2], pp[pc - 1], pp[pc], pp[pc + 1], pp[pc + 2], g0,
                                 plane[R_VELX * psz + pc], plane[R_VELX * psz + pc + 1]);
        const Real fy = flat_dir(pp[pc - 2 * PW], pp[pc - PW], pp[pc], pp[pc + PW], pp[pc + 2 * PW], g0,
                                 plane[R_VELY * psz + pc], plane[R_VELY * psz + pc + PW]);
        const Real fz = flat_dir(win[R_PRES][0], win[R_PRES][1], win[R_PRES][2], win[R_PRES][3], win[R_PRES][4], g0,
                                 win[R_VELZ][2], win[R_VELZ][3]);
        flat = fmin2(fmin2(fx, fy), fz);
      }
      // x and y reconstruction of level k from the plane
      Real umx[NV], umy[NV];
      if (xyk && rowx) {
#pragma unroll
        for (int v = 0; v < NV; v++) {
          if (v < R_NUM + ns && (kF64 || v != R_GAMC)) {
            const Real* r = plane + v * psz + pc;
            Real u[5] = {r[-2], r[-1], r[0], r[1], r[2]};
            Real up;
            recon1<RECON, Real>(u, flat, up, umx[v]);
            xex[v * NT + tid] = up;
          }
        }
      }
      if (xyk && coly) {
#pragma unroll
        for (int v = 0; v < NV; v++) {
          if (v < R_NUM + ns && (kF64 || v != R_GAMC)) {
            const Real* r = plane + v * psz + pc;
            Real u[5] = {r[-2 * PW], r[-PW], r[0], r[PW], r[2 * PW]};
            Real up;
            recon1<RECON, Real>(u, flat, up, umy[v]);
            yex[v * NT + tid] = up;
          }
        }
      }
      __syncthreads();
      Real Fx[NF], Fy[NF];
      const bool fxok = xyk && rowx && px >= 1;
      const bool fyok = xyk && coly && py >= 1;
      if (fxok) {
        Real Ls[NV], cl[5], cr[5];
#pragma unroll
        for (int v = 0; v < NV; v++) Ls[v] = (v < R_NUM + ns && (kF64 || v != R_GAMC)) ? xex[v * NT + tid - 1] : Real(0.);
        const int c0 = pc - 1;
        cl[0] = plane[R_DENS * psz + c0]; cl[1] = plane[R_VELX * psz + c0]; cl[2] = plane[R_VELY * psz + c0];
        cl[3] = plane[R_VELZ * psz + c0]; cl[4] = plane[R_RHOE * psz + c0];
        cr[0] = plane[R_DENS * psz + pc]; cr[1] = plane[R_VELX * psz + pc]; cr[2] = plane[R_VELY * psz + pc];
        cr[3] = plane[R_VELZ * psz + pc]; cr[4] = plane[R_RHOE * psz + pc];
        face_flux<0, RIEMANN, NSM>(rp, ns, Ls, umx, cl, cr, Fx, A.err);
      }
      if (fyok) {
        Real Ls[NV], cl[5], cr[5];
#pragma unroll
        for (int v = 0; v < NV; v++) Ls[v] = (v < R_NUM + ns && (kF64 || v != R_GAMC)) ? yex[v * NT + tid - W] : Real(0.);
        const int c0 = pc - PW;
        cl[0] = plane[R_DENS * psz + c0]; cl[1] = plane[R_VELX * psz + c0]; cl[2] = plane[R_VELY * psz + c0];
        cl[3] = plane[R_VELZ * psz + c0]; cl[4] = plane[R_RHOE * psz + c0];
        cr[0] = plane[R_DENS * psz + pc]; cr[1] = plane[R_VELX * psz + pc]; cr[2] = plane[R_VELY * psz + pc];
        cr[3] = plane[R_VELZ * psz + pc]; cr[4] = plane[R_RHOE * psz + pc];
        face_flux<1, RIEMANN, NSM>(rp, ns, Ls, umy, cl, cr, Fy, A.err);
      }
      __syncthreads();
      if (fxok) {
#pragma unroll
        for (int n = 0; n < NF; n++) xex[n * NT + tid] = Fx[n];
      }
      if (fyok) {
#pragma unroll
        for (int n = 0; n < NF; n++) yex[n * NT + tid] = Fy[n];
      }
      __syncthreads();
      // flux registers of interior faces (addFluxes; each face written once)
      if (A.fb[0] != nullptr && xyk) {
        const int g = P.ng;
        if (fxok && (px <= I.tx || i == I.xhi + 1) && i >= g && i <= g + m.nx && j >= g && j < g + m.ny && k >= g &&
            k < g + m.nz) {
          const int ex = m.nx + 1, ey = m.ny;
          const int64_t tot = (int64_t)ex * ey * m.nz;
          const int64_t o = m.bbase + (i - g) + (int64_t)ex * ((j - g) + (int64_t)ey * (k - g));
          for (int n = 0; n < F_NUM + ns; n++) {
            const double w = A.weight * (double)Fx[n];
            A.fb[0][o + n * tot] = A.add_flux ? A.fb[0][o + n * tot] + w : w;
          }
        }
        if (fyok && (py <= I.ty || j == I.yhi + 1) && i >= g && i < g + m.nx && j >= g && j <= g + m.ny && k >= g &&
            k < g + m.nz) {
          const int ex = m.nx, ey = m.ny + 1;
          const int64_t tot = (int64_t)ex * ey * m.nz;
          const int64_t o = m.bbase + (i - g) + (int64_t)ex * ((j - g) + (int64_t)ey * (k - g));
          for (int n = 0; n < F_NUM + ns; n++) {
            const double w = A.weight * (double)Fy[n];
            A.fb[1][o + n * tot] = A.add_flux ? A.fb[1][o + n * tot] + w : w;
          }
        }
      }
      // z reconstruction of level k (register window) and the face k-1/2
      Real upzk[NV], umz[NV];
      if (inner) {
#pragma unroll
        for (int v = 0; v < NV; v++) {
          if (v < R_NUM + ns && (kF64 || v != R_GAMC)) {
            Real u[5] = {win[v][0], win[v][1], win[v][2], win[v][3], win[v][4]};
            recon1<RECON, Real>(u, flat, upzk[v], umz[v]);
          } else {
            upzk[v] = Real(0.);
            umz[v] = Real(0.);
          }
        }
      }
      if (inner && k >= I.z0) {
        Real Hz[NF], cl[5], cr[5];
        cl[0] = win[R_DENS][1]; cl[1] = win[R_VELX][1]; cl[2] = win[R_VELY][1]; cl[3] = win[R_VELZ][1];
        cl[4] = win[R_RHOE][1];
        cr[0] = win[R_DENS][2]; cr[1] = win[R_VELX][2]; cr[2] = win[R_VELY][2]; cr[3] = win[R_VELZ][2];
        cr[4] = win[R_RHOE][2];
        face_flux<2, RIEMANN, NSM>(rp, ns, upz, umz, cl, cr, Hz, A.err);
        if (A.fb[2] != nullptr) {
          const int g = P.ng;
          if (i >= g && i < g + m.nx && j >= g && j < g + m.ny && k >= g && k <= g + m.nz) {
            const int64_t tot = (int64_t)m.nx * m.ny * (m.nz + 1);
            const int64_t o = m.bbase + (i - g) + (int64_t)m.nx * ((j - g) + (int64_t)m.ny * (k - g));
            for (int n = 0; n < F_NUM + ns; n++) {
              const double w = A.weight * (double)Hz[n];
              A.fb[2][o + n * tot] = A.add_flux ? A.fb[2][o + n * tot] + w : w;
            }
          }
        }
        if (k - 1 >= I.z0) {
          // the update of cell (i, j, k-1), FP64 (hy_rk_updateSoln.F90-mc:227-287)
          const int64_t c = colc + (int64_t)(k - 1) * sz;
          double Au[NF], Bu[NF];
          const double d0 = u0p[V_DENS * cs + c];
          Au[F_MASS] = d0;
          Au[F_XMOM] = d0 * u0p[V_VELX * cs + c];
          Au[F_YMOM] = d0 * u0p[V_VELY * cs + c];
          Au[F_ZMOM] = d0 * u0p[V_VELZ * cs + c];
          Au[F_ENER] = d0 * u0p[V_ENER * cs + c];
          const double ds = sin[V_DENS * cs + c];
          Bu[F_MASS] = ds;
          Bu[F_XMOM] = ds * sin[V_VELX * cs + c];
          Bu[F_YMOM] = ds * sin[V_VELY * cs + c];
          Bu[F_ZMOM] = ds * sin[V_VELZ * cs + c];
          Bu[F_ENER] = ds * sin[V_ENER * cs + c];
#pragma unroll
          for (int s = 0; s < NSM; s++) {
            Au[F_NUM + s] = (s < ns) ? u0p[(V_NPROP + s) * cs + c] * d0 : 0.;
            Bu[F_NUM + s] = (s < ns) ? sin[(V_NPROP + s) * cs + c] * ds : 0.;
          }
          double Sg[NF];
#pragma unroll
          for (int n = 0; n < NF; n++) Sg[n] = 0.;
          if (P.use_grav) {
            Sg[F_XMOM] = Bu[F_MASS] * P.grav[0];
            Sg[F_YMOM] = Bu[F_MASS] * P.grav[1];
            Sg[F_ZMOM] = Bu[F_MASS] * P.grav[2];
            Sg[F_ENER] = Bu[F_XMOM] * P.grav[0] + Bu[F_YMOM] * P.grav[1] + Bu[F_ZMOM] * P.grav[2];
          }
#pragma unroll
          for (int n = 0; n < NF; n++) {
            double inner_n = dUxy[n] - dtdz * ((double)Hz[n] - (double)Hprev[n]);
            inner_n = inner_n + dt * (0.0 + Sg[n]);
            Bu[n] = A.c1 * Au[n] + A.c2 * Bu[n] + A.c3 * inner_n;
          }
          const double dens = fmax2(Bu[F_MASS], P.smlrho);
          const double vx = Bu[F_XMOM] / dens, vy = Bu[F_YMOM] / dens, vz = Bu[F_ZMOM] / dens;
          const double ener = fmax2(P.smalle, Bu[F_ENER] / dens);
          const double ekin = .5 * (vx * vx + vy * vy + vz * vz);
          const double eint = fmax2(P.smalle, ener - ekin);
#pragma unroll
          for (int s = 0; s < NSM; s++)
            if (s < ns) sout[(V_NPROP + s) * cs + c] = Bu[F_NUM + s] / dens;
          EosOut e = eos_dens_ei(P, dens, vx, vy, vz, ener, eint);
          sout[V_DENS * cs + c] = dens;
          sout[V_VELX * cs + c] = vx;
          sout[V_VELY * cs + c] = vy;
          sout[V_VELZ * cs + c] = vz;
          sout[V_PRES * cs + c] = e.pres;
          sout[V_GAMC * cs + c] = P.gamma;
          sout[V_EINT * cs + c] = e.eint;
          sout[V_ENER * cs + c] = e.ener;
          if (A.final_stage) {
            sout[V_TEMP * cs + c] = e.temp;
            sout[V_GAME * cs + c] = e.game;
          }
        }
#pragma unroll
        for (int n = 0; n < NF; n++) Hprev[n] = Hz[n];
      }
      if (inner) {
#pragma unroll
        for (int v = 0; v < NV; v++) upz[v] = upzk[v];
      }
      // x/y part of level k's update, consumed at the next level
      if (inner && xyk) {
#pragma unroll
        for (int n = 0; n < NF; n++) {
          double d = -(dtdx * ((double)xex[n * NT + tid + 1] - (double)xex[n * NT + tid]));
          dUxy[n] = d - dtdy * ((double)yex[n * NT + tid + W] - (double)yex[n * NT + tid]);
        }
      }
      __syncthreads();
    }
  }
}

// reset of the per-XCD queue heads (a kernel, so it can sit in a graph)
__global__ void k_reset_counters(int* c, int n) {
  if ((int)threadIdx.x < n) c[threadIdx.x] = 0;
}

template <int NSM, int RECON, int RIEMANN>
hipError_t launch_t(const FusedArgs& A, int grid, size_t lds, hipStream_t st) {
  hipLaunchKernelGGL((k_fused<NSM, RECON, RIEMANN>), dim3(grid), dim3(FUSED_THREADS), lds, st, A);
  return hipGetLastError();
}

template <int NSM>
hipError_t launch_r(const FusedArgs& A, int grid, size_t lds, hipStream_t st) {
  const bool hll = A.P.riemann == RIEMANN_HLL;
  switch (A.P.recon) {
    case RECON_MP5: return hll ? launch_t<NSM, RECON_MP5, RIEMANN_HLL>(A, grid, lds, st)
                               : launch_t<NSM, RECON_MP5, RIEMANN_HLLC>(A, grid, lds, st);
    case RECON_PLM: return hll ? launch_t<NSM, RECON_PLM, RIEMANN_HLL>(A, grid, lds, st)
                               : launch_t<NSM, RECON_PLM, RIEMANN_HLLC>(A, grid, lds, st);
    case RECON_FOG: return hll ? launch_t<NSM, RECON_FOG, RIEMANN_HLL>(A, grid, lds, st)
                               : launch_t<NSM, RECON_FOG, RIEMANN_HLLC>(A, grid, lds, st);
    default: return hll ? launch_t<NSM, RECON_WENO5, RIEMANN_HLL>(A, grid, lds, st)
                        : launch_t<NSM, RECON_WENO5, RIEMANN_HLLC>(A, grid, lds, st);
  }
}

}  // namespace

#define FUSED_NAME2(a, b) a##b
#define FUSED_NAME(a, b) FUSED_NAME2(a, b)
#define FUSED_SFX SPARK_FUSED_SUFFIX

size_t FUSED_NAME(fused_lds_bytes, FUSED_SFX)(int ns, int max_plane) {
  const int nv = R_NUM + (ns > 0 ? 8 : 0);
  return sizeof(Real) * (size_t)nv * ((size_t)max_plane + 2 * FUSED_THREADS);
}

hipError_t FUSED_NAME(fused_reset, FUSED_SFX)(hipStream_t st, int* counters) {
  hipLaunchKernelGGL(k_reset_counters, dim3(1), dim3(64), 0, st, counters, FUSED_XCDS);
  return hipGetLastError();
}

hipError_t FUSED_NAME(fused_stage, FUSED_SFX)(const FusedArgs& A, int grid, size_t lds, hipStream_t st) {
  return (A.P.ns == 0) ? launch_r<0>(A, grid, lds, st) : launch_r<8>(A, grid, lds, st);
}

hipError_t FUSED_NAME(fused_occupancy, FUSED_SFX)(int ns, size_t lds, int* blocks_per_cu) {
  if (ns == 0)
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, k_fused<0, RECON_WENO5, RIEMANN_HLLC>,
                                                        FUSED_THREADS, lds);
  return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, k_fused<8, RECON_WENO5, RIEMANN_HLLC>,
                                                      FUSED_THREADS, lds);
}

}  // namespace spark
