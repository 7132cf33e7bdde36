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
// its own runs dry.  Inside an item the workgroup marches up z.  Each phase
// of a level maps threads onto its own task space, so the lanes stay busy
// although the task counts differ (tile tx x ty, C = tx*ty columns):
//   P1  stage plane k (tile + 3 halo cells) in LDS; thread t < C owns
//       column t and keeps a 5-deep z window of it in registers
//   P2  flattening of the tile and its 1-cell ring (hy_rk_getFaceFlux.F90-mc:200-276)
//   P3  reconstruction: z of the C columns (registers), x of the
//       (tx+2) x ty and y of the tx x (ty+2) cells (LDS stencils)
//   P4  Riemann solve + artificial viscosity: z faces of the columns, the
//       (tx+1) x ty x faces and tx x (ty+1) y faces
//   P5  FP64 update, floors and Gamma EOS of level k-1 (the z flux above it
//       is known now), then the x/y flux differences of level k
//   mass scalars follow in chunks of four through the same buffers
//       (recon, upwinded fluxes, update), so the LDS and register footprint
//       of the hydro part does not grow with the scalar count.
// x and y task ranges are padded to whole wavefronts, so no wavefront mixes
// the two directions.  The flux differences are accumulated in FP64 in the
// reference order (x, y, z, source; hy_rk_updateSoln.F90-mc:257-265).
// Stage input and output states are distinct buffers, so no tile reads a
// cell another tile writes.
#include <type_traits>

#include "spark_physics.hpp"
#include "spark_batched.hpp"

#ifndef SPARK_FUSED_REAL
#define SPARK_FUSED_REAL float
#endif

#define FUSED_NAME2(a, b) a##b
#define FUSED_NAME(a, b) FUSED_NAME2(a, b)
#define FUSED_SFX SPARK_FUSED_SUFFIX

namespace spark {
namespace {

using Real = SPARK_FUSED_REAL;
constexpr bool kF64 = std::is_same<Real, double>::value;
constexpr int NT = FUSED_THREADS;
// hydro slots staged per cell: dens velx vely velz pres [gamc] rhoe.  The
// FP32 form takes gamc = gamma (the Gamma EOS sets every cell's gamc to
// gamma), so it neither stages nor reconstructs it.
constexpr int NH = kF64 ? 7 : 6;
constexpr int CH = FUSED_SCALAR_CHUNK;  // mass scalars per chunk (the species all in the first)
constexpr int SL_RHOE = NH - 1;

__host__ __device__ __forceinline__ constexpr int hslot(int s) { return (!kF64 && s >= R_GAMC) ? s + 1 : s; }
__host__ __device__ __forceinline__ constexpr int rup64(int n) { return (n + 63) & ~63; }

// LDS carve-up of one work item, in Reals
struct Lay {
  int plane, flat, xm, xp, ym, yp, fx, fy, mx, my, total;
};
__host__ __device__ __forceinline__ Lay lds_layout(int tx, int ty, int nvt) {
  Lay L;
  const int psz = (tx + 6) * (ty + 6);
  const int nrx = (tx + 2) * ty, nry = tx * (ty + 2), nfx = (tx + 1) * ty, nfy = tx * (ty + 1);
  int o = 0;
  L.plane = o; o += nvt * psz;
  L.flat = o; o += (tx + 2) * (ty + 2);
  L.xm = o; o += NH * nrx;
  L.xp = o; o += NH * nrx;
  L.ym = o; o += NH * nry;
  L.yp = o; o += NH * nry;
  L.fx = o; o += NH * nfx;  // flux rows: 5 hydro fluxes, or one chunk of scalars
  L.fy = o; o += NH * nfy;
  L.mx = o; o += nfx;       // mass flux of the x / y faces (scalar upwinding)
  L.my = o; o += nfy;
  L.total = o;
  return L;
}

// division of small non-negative task indices by a per-item divisor
// through one float multiply and a correction (no integer divide sequence)
struct FDiv {
  int d;
  float inv;
};
__device__ __forceinline__ FDiv fdiv(int d) { return FDiv{d, 1.0f / (float)d}; }
__device__ __forceinline__ int fquot(int w, const FDiv& f, int& r) {
  int q = (int)((float)w * f.inv);
  r = w - q * f.d;
  if (r < 0) { q--; r += f.d; }
  else if (r >= f.d) { q++; r -= f.d; }
  return q;
}

// staged slot v of cell c: hydro slots in the pencil order, then scalars
__device__ __forceinline__ Real load_slot(const double* __restrict__ s, int64_t cs, int64_t c, int v, double gamma) {
  if (v >= NH) return (Real)s[(V_NPROP + v - NH) * cs + c];
  switch (hslot(v)) {
    case R_DENS: return (Real)s[V_DENS * cs + c];
    case R_VELX: return (Real)s[V_VELX * cs + c];
    case R_VELY: return (Real)s[V_VELY * cs + c];
    case R_VELZ: return (Real)s[V_VELZ * cs + c];
    case R_PRES: return (Real)s[V_PRES * cs + c];
    case R_GAMC: return kF64 ? (Real)s[V_GAMC * cs + c] : (Real)gamma;
    default: return (Real)(s[V_DENS * cs + c] * s[V_EINT * cs + c]);
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

// Riemann template value of the fused kernel beside RIEMANN_HLLC / _HLL:
// HLLC with the hybrid switch to HLL at faces next to a shock-flagged cell
// (use_hybridRiemann, hy_rk_getFaceFlux.F90-mc:258-274)
constexpr int RIEMANN_HLLC_HYBRID = 2;

// SHOK of the two cells on either side of a face, from the step's input
// state; the cells are at most one layer outside the stage region, inside
// the block's own box, where the step's shock detection has written SHOK
__device__ __forceinline__ bool face_in_shock(const double* __restrict__ U0, const BlockMeta& m, int64_t lo,
                                              int64_t hi) {
  const double* sh = U0 + m.base + (int64_t)V_SHOK * m.cstride;
  return sh[lo] != 0.0 || sh[hi] != 0.0;
}

// hydro face flux from the left/right face states (ensurePhysicalState,
// Riemann solve, artificial viscosity; Hydro.ini:222-235,285-301).  c*
// hold the two cells' centred dens, velx, vely, velz, rhoe.
template <int DIR, int RIEMANN>
__device__ __forceinline__ void hydro_face(const RParams& rp, Real* L, Real* R, const Real* cl, const Real* cr, Real* F,
                                           int* err, bool inShock) {
  L[R_DENS] = fmax2(rp.smlrho, L[R_DENS]);
  L[R_PRES] = fmax2(rp.smallp, L[R_PRES]);
  R[R_DENS] = fmax2(rp.smlrho, R[R_DENS]);
  R[R_PRES] = fmax2(rp.smallp, R[R_PRES]);
  if (!kF64) { L[R_GAMC] = rp.gamma; R[R_GAMC] = rp.gamma; }
  int ierr;
  if (RIEMANN == RIEMANN_HLL) ierr = hll<DIR, Real>(L, R, F);
  else ierr = hllc<DIR, Real>(L, R, RIEMANN == RIEMANN_HLLC_HYBRID, inShock, rp.tiny, F);
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
}

// face states from the staged slots: L from a u+ buffer, R from a u- buffer
__device__ __forceinline__ void face_states(const Real* __restrict__ bp, const Real* __restrict__ bm, int nr, int il,
                                            int ir, Real* L, Real* R) {
#pragma unroll
  for (int s = 0; s < NH; s++) {
    L[hslot(s)] = bp[s * nr + il];
    R[hslot(s)] = bm[s * nr + ir];
  }
  if (!kF64) { L[R_GAMC] = Real(0.); R[R_GAMC] = Real(0.); }
}

// species face states: clamp to [smallx, 1] and renormalise, left and right
// (Hydro.ini:227-235); x[j], j < nspec (all in the first scalar chunk)
__device__ __forceinline__ void species_faces(const DevParams& P, Real* xl, Real* xr) {
  if (P.nspec <= 0) return;
  const Real lo = (Real)P.smallx;
  Real sl = Real(0.), sr = Real(0.);
#pragma unroll
  for (int j = 0; j < CH; j++) {
    if (j < P.nspec) {
      xl[j] = fmax2(lo, fmin2(Real(1.), xl[j]));
      xr[j] = fmax2(lo, fmin2(Real(1.), xr[j]));
      sl = sl + xl[j];
      sr = sr + xr[j];
    }
  }
  const Real il = Num<Real>::rcp(sl), ir = Num<Real>::rcp(sr);
#pragma unroll
  for (int j = 0; j < CH; j++) {
    if (j < P.nspec) {
      xl[j] = xl[j] * il;
      xr[j] = xr[j] * ir;
    }
  }
}

__device__ __forceinline__ void centred(const Real* __restrict__ plane, int psz, int pc, Real* c) {
  c[0] = plane[0 * psz + pc];
  c[1] = plane[1 * psz + pc];
  c[2] = plane[2 * psz + pc];
  c[3] = plane[3 * psz + pc];
  c[4] = plane[SL_RHOE * psz + pc];
}

// flux-register accumulation of one face (addFluxes, Hydro_funcs.F90-mc:175-201):
// entries n0 .. n0+cnt-1 (cnt <= N) of the face take weight * f[i]; with
// add, every old entry is loaded before the first store (the compiler
// cannot tell the entries apart, so a load after a store would wait for it)
template <int N>
__device__ __forceinline__ void reg_add_n(double* fb, int64_t o, int64_t tot, int n0, int cnt, double weight, int add,
                                          const Real* f, int fstride = 1) {
  double w[N], prev[N];
#pragma unroll
  for (int i = 0; i < N; i++) {
    w[i] = weight * (double)f[i * fstride];
    prev[i] = 0.0;
  }
  if (add) {
#pragma unroll
    for (int i = 0; i < N; i++)
      if (i < cnt) prev[i] = fb[o + (n0 + i) * tot];
  }
#pragma unroll
  for (int i = 0; i < N; i++)
    if (i < cnt) fb[o + (n0 + i) * tot] = add ? prev[i] + w[i] : w[i];
}

// x (DIR 0) or y (DIR 1) face w of the tile at level k: states from the
// LDS reconstruction buffers, Riemann solve + viscosity, flux rows and the
// mass flux to LDS, flux registers for interior faces the item owns.
template <int DIR, int RIEMANN>
__device__ __forceinline__ void xy_face(const FusedArgs& A, const BlockMeta& m, const FusedItem& I, const RParams& rp,
                                        int g, int k, const Real* bp, const Real* bm, int nr, const Real* plane, int psz,
                                        int PW, const FDiv& dv, Real* fo, Real* mo, int nf, int w) {
  const int tx = I.tx, ty = I.ty;
  Real Ls[R_NUM], Rs[R_NUM], cl[5], cr[5], F[F_NUM];
  int fi;
  const int fj = fquot(w, dv, fi);  // the face is the low face of cell (fi, fj) of the tile
  const int il = DIR == 0 ? fj * (tx + 2) + fi : fj * tx + fi;
  face_states(bp, bm, nr, il, il + (DIR == 0 ? 1 : tx), Ls, Rs);
  const int pc = (fj + 3) * PW + (fi + 3);
  centred(plane, psz, pc - (DIR == 0 ? 1 : PW), cl);
  centred(plane, psz, pc, cr);
  bool shk = false;
  if (RIEMANN == RIEMANN_HLLC_HYBRID) {
    const int64_t hi = bcix(m, I.x0 + fi, I.y0 + fj, k);
    shk = face_in_shock(A.U0, m, hi - (DIR == 0 ? 1 : m.NX), hi);
  }
  hydro_face<DIR, RIEMANN>(rp, Ls, Rs, cl, cr, F, A.err, shk);
#pragma unroll
  for (int n = 0; n < F_NUM; n++) fo[n * nf + w] = F[n];
  mo[w] = F[F_MASS];
  if (A.fb[0] != nullptr) {
    const int i = I.x0 + fi, j = I.y0 + fj;
    // a boundary face belongs to the item holding the interior cell beside it
    const bool own = DIR == 0 ? (i == g ? fi < tx : fi > 0) : (j == g ? fj < ty : fj > 0);
    // the block's boundary faces only (the register hand-off reads no other)
    const bool in = DIR == 0 ? ((i == g || i == g + m.nx) && j >= g && j < g + m.ny)
                             : (i >= g && i < g + m.nx && (j == g || j == g + m.ny));
    if (own && in && k >= g && k < g + m.nz) {
      const int ex = m.nx + (DIR == 0), ey = m.ny + (DIR == 1);
      const int64_t tot = (int64_t)ex * ey * m.nz;
      const int64_t o = m.bbase + (i - g) + (int64_t)ex * ((j - g) + (int64_t)ey * (k - g));
      reg_add_n<F_NUM>(A.fb[DIR], o, tot, 0, F_NUM, A.weight, A.add_flux, F);
    }
  }
}

template <int NSM, int RECON, int RIEMANN>
// waves per SIMD the register allocation targets: the FP32 form fits 2
// (<= 256 VGPRs, no spills, with or without scalars); the FP64 form keeps
// all its registers (accumulation registers included) at 1 rather than
// spilling
#ifndef FUSED_WAVES_OF
#define FUSED_WAVES_OF(nsm) (kF64 ? 1 : 2)
#endif
#ifndef FUSED_XY_UNROLL
#define FUSED_XY_UNROLL 6
#endif
#ifndef FUSED_EARLY_Z
#define FUSED_EARLY_Z kF64
#endif
#ifndef EARLY_Z_BARRIER
#define EARLY_Z_BARRIER
#endif
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(FUSED_WAVES_OF(NSM)))) FUSED_NAME(k_fused, FUSED_SFX)(FusedArgs A) {
  constexpr int NFM = F_NUM + NSM;
  constexpr int NSA = NSM > 0 ? NSM : 1;
  extern __shared__ __align__(16) unsigned char lds_raw[];
  Real* lds = reinterpret_cast<Real*>(lds_raw);
  __shared__ int s_item;
  __shared__ int64_t s_nb[27];
  const int tid = threadIdx.x;
  const DevParams& P = A.P;
  const RParams rp{(Real)P.gamma, (Real)P.cvisc, (Real)P.smlrho, (Real)P.smallp, (Real)P.tiny};
  const double dt = *A.dtp;
  const int ns = NSM > 0 ? P.ns : 0;
  const int nvt = NH + ns;
  const bool use_flat = RECON == RECON_WENO5 && P.use_flat;
  const int g = P.ng;
  int q = blockIdx.x % FUSED_XCDS;  // the queue of this workgroup's XCD first
  int tried = 0;
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
    const int tx = I.tx, ty = I.ty, PW = tx + 6, psz = PW * (ty + 6);
    const int C = tx * ty, Cp = rup64(C);
    const int nrx = (tx + 2) * ty, nry = tx * (ty + 2), nfx = (tx + 1) * ty, nfy = tx * (ty + 1);
    const Lay L = lds_layout(tx, ty, nvt);
    const FDiv d_tx = fdiv(tx), d_tx1 = fdiv(tx + 1), d_tx2 = fdiv(tx + 2), d_pw = fdiv(PW);
    Real* plane = lds + L.plane;
    Real* flatb = lds + L.flat;
    Real* xm = lds + L.xm;
    Real* xp = lds + L.xp;
    Real* ym = lds + L.ym;
    Real* yp = lds + L.yp;
    Real* fxb = lds + L.fx;
    Real* fyb = lds + L.fy;
    Real* mxb = lds + L.mx;
    Real* myb = lds + L.my;
    const int64_t cs = m.cstride;
    const int64_t sz = (int64_t)m.NX * m.NY;
    const double* sin = A.Sin + m.base;
    const double* S = A.Sin;  // pool base: cell_at() offsets include the block's base
    // the item's neighbour offsets, cached in LDS (read on every halo load)
    const int64_t* nbt = nullptr;
    if (A.nbr) {
      if (tid < 27) s_nb[tid] = A.nbr[(int64_t)I.blk * 27 + tid];
      __syncthreads();
      if (s_nb[13] >= 0) nbt = s_nb;  // -1: the block reads its own guard cells
    }
    const double* u0p = A.U0 + m.base;
    double* sout = A.Sout + m.base;
    const double dtdx = dt / m.dx, dtdy = dt / m.dy, dtdz = dt / m.dz;
    const bool col = tid < C;
    int cx = 0, cy = 0;
    if (col) cy = fquot(tid, d_tx, cx);
    const int ci = I.x0 + cx, cj = I.y0 + cy;
    const int64_t colc = bcix(m, ci, cj, 0);
    const int pcc = (cy + 3) * PW + (cx + 3);  // the column's cell in the plane
    const int fcc = (cy + 1) * (tx + 2) + (cx + 1);
    const bool regs = A.fb[0] != nullptr;
    const bool colin = ci >= g && ci < g + m.nx && cj >= g && cj < g + m.ny;

    Real win[NH][5];    // z window k-2 .. k+2 of the column, hydro slots
    Real upz[NH], upzs[NSA];  // u+ of level k-1 (z faces)
    Real Hprev[NFM];          // z flux at the low face of level k-1
    double dUxy[NFM];         // -dt/dx dF - dt/dy dG of level k-1
#pragma unroll
    for (int v = 0; v < NH; v++) {
      upz[v] = Real(0.);
#pragma unroll
      for (int w = 0; w < 5; w++) win[v][w] = Real(0.);
    }
#pragma unroll
    for (int s = 0; s < NSA; s++) upzs[s] = Real(0.);
#pragma unroll
    for (int n = 0; n < NFM; n++) { Hprev[n] = Real(0.); dUxy[n] = 0.; }
    if (col) {
      // levels z0-3 .. z0; the first pass (k = z0-1) shifts them down and
      // loads z0+1
#pragma unroll
      for (int w = 1; w < 5; w++) {
        const int64_t c = cell_at(m, nbt, g, ci, cj, I.z0 - 4 + w);
#pragma unroll
        for (int v = 0; v < NH; v++) win[v][w] = load_slot(S, cs, c, v, P.gamma);
      }
    }

    for (int k = I.z0 - 1; k <= I.z1 + 1; k++) {
      const bool xyk = (k >= I.z0) && (k <= I.z1);
      // ---- P1: advance the z window to k-2..k+2, stage plane k
      if (col) {
        const int64_t c = cell_at(m, nbt, g, ci, cj, k + 2);
#pragma unroll
        for (int v = 0; v < NH; v++) {
#pragma unroll
          for (int w = 0; w < 4; w++) win[v][w] = win[v][w + 1];
          win[v][4] = load_slot(S, cs, c, v, P.gamma);
          plane[v * psz + pcc] = win[v][2];
        }
        // scalars: the plane only (their z stencils are read again when the
        // scalar chunks are reconstructed); all loads issue before the first
        // store waits on them
        const int64_t ck = cell_at(m, nbt, g, ci, cj, k);
#pragma unroll
        for (int s0 = 0; s0 < NSM; s0 += 4) {
          Real sv[4];
#pragma unroll
          for (int q = 0; q < 4; q++) sv[q] = (Real)S[(V_NPROP + (s0 + q < ns ? s0 + q : ns - 1)) * cs + ck];
#pragma unroll
          for (int q = 0; q < 4; q++)
            if (s0 + q < ns) plane[(NH + s0 + q) * psz + pcc] = sv[q];
        }
      }
      {
        // the 3-cell halo ring: 3 full rows below and above, 3 + 3 cells in
        // each of the ty middle rows
        const int nhalo = 6 * PW + 6 * ty;
        for (int u = tid; u < nhalo; u += NT) {
          int qx, qy, h = u;
          if (h < 3 * PW) { qy = fquot(h, d_pw, qx); }
          else if ((h -= 3 * PW) < 3 * PW) { qy = ty + 3 + fquot(h, d_pw, qx); }
          else {
            h -= 3 * PW;
            qy = 3 + h / 6;
            const int cc = h % 6;
            qx = cc < 3 ? cc : tx + cc;
          }
          const int64_t c = cell_at(m, nbt, g, I.x0 - 3 + qx, I.y0 - 3 + qy, k);
          const int pc = qy * PW + qx;
          // the hydro slots, then the scalars, each group loaded before
          // the first of its stores waits
          Real hv[NH > NSA ? NH : NSA];
#pragma unroll
          for (int v = 0; v < NH; v++) hv[v] = load_slot(S, cs, c, v, P.gamma);
#pragma unroll
          for (int v = 0; v < NH; v++) plane[v * psz + pc] = hv[v];
          // (slots past ns load scalar ns-1 again, so no load is conditional)
          if (NSM > 0) {
#pragma unroll
            for (int v = 0; v < NSA; v++) hv[v] = (Real)S[(V_NPROP + (v < ns ? v : ns - 1)) * cs + c];
#pragma unroll
            for (int v = 0; v < NSA; v++)
              if (v < ns) plane[(NH + v) * psz + pc] = hv[v];
          }
        }
      }
      __syncthreads();

      // ---- P2: flattening of the tile cells (all levels) and of the ring
      // around the tile (levels with x/y reconstruction)
      if (use_flat) {
        const int nflat = xyk ? (tx + 2) * (ty + 2) : C;
        const Real* pp = plane + 4 * psz;
        for (int u = tid; u < nflat; u += NT) {
          int px, py;
          const bool own = u < C;  // u == tid: the column's own cell
          if (own) { py = fquot(u, d_tx, px) + 1; px += 1; }
          else {
            const int h = u - C;
            if (h < tx + 2) { px = h; py = 0; }
            else if (h < 2 * (tx + 2)) { px = h - (tx + 2); py = ty + 1; }
            else { const int r = h - 2 * (tx + 2); py = 1 + r / 2; px = (r & 1) ? tx + 1 : 0; }
          }
          const int pc = (py + 2) * PW + (px + 2);
          const Real g0 = kF64 ? plane[R_GAMC * psz + pc] : rp.gamma;
          const Real fx = flat_dir(pp[pc - 2], pp[pc - 1], pp[pc], pp[pc + 1], pp[pc + 2], g0, plane[1 * psz + pc],
                                   plane[1 * psz + pc + 1]);
          const Real fy = flat_dir(pp[pc - 2 * PW], pp[pc - PW], pp[pc], pp[pc + PW], pp[pc + 2 * PW], g0,
                                   plane[2 * psz + pc], plane[2 * psz + pc + PW]);
          Real fz;
          if (own) {
            fz = flat_dir(win[4][0], win[4][1], win[4][2], win[4][3], win[4][4], g0, win[3][2], win[3][3]);
          } else {
            const int x = I.x0 - 1 + px, y = I.y0 - 1 + py;
            const double* pg = S + V_PRES * cs;
            const double* vg = S + V_VELZ * cs;
            const int64_t c0 = cell_at(m, nbt, g, x, y, k), c1 = cell_at(m, nbt, g, x, y, k + 1);
            fz = flat_dir((Real)pg[cell_at(m, nbt, g, x, y, k - 2)], (Real)pg[cell_at(m, nbt, g, x, y, k - 1)],
                          (Real)pg[c0], (Real)pg[c1], (Real)pg[cell_at(m, nbt, g, x, y, k + 2)], g0, (Real)vg[c0],
                          (Real)vg[c1]);
          }
          flatb[py * (tx + 2) + px] = fmin2(fmin2(fx, fy), fz);
        }
      }
      __syncthreads();

      // ---- P3: reconstruction of the hydro slots
      Real upzk[NH], umz[NH];
#pragma unroll
      for (int v = 0; v < NH; v++) { upzk[v] = Real(0.); umz[v] = Real(0.); }
      {
        const int nrxp = rup64(nrx);
        const int utot = Cp + (xyk ? nrxp + nry : 0);
        for (int u = tid; u < utot; u += NT) {
          if (u < Cp) {
            if (u < C) {
              const Real fl = use_flat ? flatb[fcc] : Real(1.);
#pragma unroll
              for (int v = 0; v < NH; v++) recon1<RECON, Real>(win[v], fl, upzk[v], umz[v]);
            }
            continue;
          }
          int w = u - Cp;
          const bool isx = w < nrxp;
          if (isx && w >= nrx) continue;
          if (!isx) w -= nrxp;
          int rx, ry, st, nr;
          Real *om, *op;
          if (isx) { ry = fquot(w, d_tx2, rx); rx -= 1; st = 1; nr = nrx; om = xm; op = xp; }
          else { ry = fquot(w, d_tx, rx) - 1; st = PW; nr = nry; om = ym; op = yp; }
          const int pc = (ry + 3) * PW + (rx + 3);
          const Real fl = use_flat ? flatb[(ry + 1) * (tx + 2) + (rx + 1)] : Real(1.);
#pragma unroll FUSED_XY_UNROLL
          for (int v = 0; v < NH; v++) {
            const Real* r = plane + v * psz + pc;
            const Real u5[5] = {r[-2 * st], r[-st], r[0], r[st], r[2 * st]};
            Real up, um;
            recon1<RECON, Real>(u5, fl, up, um);
            om[v * nr + w] = um;
            op[v * nr + w] = up;
          }
        }
      }
      __syncthreads();

      // ---- P4: hydro face fluxes
      Real Hz[F_NUM];
#pragma unroll
      for (int n = 0; n < F_NUM; n++) Hz[n] = Real(0.);
      {
        const int nfxp = rup64(nfx);
        const int utot = Cp + (xyk ? nfxp : 0);
        for (int u = tid; u < utot; u += NT) {
          if (u < Cp) {
            if (u < C && k >= I.z0) {
              Real Ls[R_NUM], Rs[R_NUM], cl[5], cr[5];
#pragma unroll
              for (int s = 0; s < NH; s++) {
                Ls[hslot(s)] = upz[s];
                Rs[hslot(s)] = umz[s];
              }
              if (!kF64) { Ls[R_GAMC] = Real(0.); Rs[R_GAMC] = Real(0.); }
              cl[0] = win[0][1]; cl[1] = win[1][1]; cl[2] = win[2][1]; cl[3] = win[3][1]; cl[4] = win[SL_RHOE][1];
              cr[0] = win[0][2]; cr[1] = win[1][2]; cr[2] = win[2][2]; cr[3] = win[3][2]; cr[4] = win[SL_RHOE][2];
              bool shk = false;
              if (RIEMANN == RIEMANN_HLLC_HYBRID) {
                const int64_t hi = colc + (int64_t)k * sz;
                shk = face_in_shock(A.U0, m, hi - sz, hi);
              }
              hydro_face<2, RIEMANN>(rp, Ls, Rs, cl, cr, Hz, A.err, shk);
              if (regs && colin && (k == g || k == g + m.nz)) {
                const int64_t tot = (int64_t)m.nx * m.ny * (m.nz + 1);
                const int64_t o = m.bbase + (ci - g) + (int64_t)m.nx * ((cj - g) + (int64_t)m.ny * (k - g));
                reg_add_n<F_NUM>(A.fb[2], o, tot, 0, F_NUM, A.weight, A.add_flux, Hz);
              }
            }
            continue;
          }
          const int w = u - Cp;
          if (w < nfx) xy_face<0, RIEMANN>(A, m, I, rp, g, k, xp, xm, nrx, plane, psz, PW, d_tx1, fxb, mxb, nfx, w);
        }
        // y faces continue the same thread sequence (one face direction per loop)
        if (xyk) {
          const int u0 = Cp + nfxp;
          for (int w = ((tid - u0) % NT + NT) % NT; w < nfy; w += NT)
            xy_face<1, RIEMANN>(A, m, I, rp, g, k, yp, ym, nry, plane, psz, PW, d_tx, fyb, myb, nfy, w);
        }
      }
      __syncthreads();

      // ---- P5: update of level k-1 (hydro part), x/y differences of level k
      const bool upd = col && k - 1 >= I.z0 && k - 1 <= I.z1;
      double d0 = 0., ds = 0., dens1 = 1.;
      // Multigamma (FP64 form): the EOS of level k-1 waits for its updated
      // species (first scalar chunk); dens vx vy vz ener eint meanwhile
      const bool defer_eos = kF64 && P.eos == 1;
      double mg[6] = {0., 0., 0., 0., 0., 0.};
      if (upd) {
        // FP64 (hy_rk_updateSoln.F90-mc:227-287)
        const int64_t c = colc + (int64_t)(k - 1) * sz;
        double Au[F_NUM], Bu[F_NUM];
        d0 = u0p[V_DENS * cs + c];
        Au[F_MASS] = d0;
        Au[F_XMOM] = d0 * u0p[V_VELX * cs + c];
        Au[F_YMOM] = d0 * u0p[V_VELY * cs + c];
        Au[F_ZMOM] = d0 * u0p[V_VELZ * cs + c];
        Au[F_ENER] = d0 * u0p[V_ENER * cs + c];
        ds = sin[V_DENS * cs + c];
        Bu[F_MASS] = ds;
        Bu[F_XMOM] = ds * sin[V_VELX * cs + c];
        Bu[F_YMOM] = ds * sin[V_VELY * cs + c];
        Bu[F_ZMOM] = ds * sin[V_VELZ * cs + c];
        Bu[F_ENER] = ds * sin[V_ENER * cs + c];
        double Sg[F_NUM];
#pragma unroll
        for (int n = 0; n < F_NUM; n++) Sg[n] = 0.;
        if (P.use_grav) {
          Sg[F_XMOM] = Bu[F_MASS] * P.grav[0];
          Sg[F_YMOM] = Bu[F_MASS] * P.grav[1];
          Sg[F_ZMOM] = Bu[F_MASS] * P.grav[2];
          Sg[F_ENER] = Bu[F_XMOM] * P.grav[0] + Bu[F_YMOM] * P.grav[1] + Bu[F_ZMOM] * P.grav[2];
        }
#pragma unroll
        for (int n = 0; n < F_NUM; n++) {
          double inner_n = dUxy[n] - dtdz * ((double)Hz[n] - (double)Hprev[n]);
          inner_n = inner_n + dt * (0.0 + Sg[n]);
          Bu[n] = A.c1 * Au[n] + A.c2 * Bu[n] + A.c3 * inner_n;
        }
        const double dens = fmax2(Bu[F_MASS], P.smlrho);
        const double vx = Bu[F_XMOM] / dens, vy = Bu[F_YMOM] / dens, vz = Bu[F_ZMOM] / dens;
        const double ener = fmax2(P.smalle, Bu[F_ENER] / dens);
        const double ekin = .5 * (vx * vx + vy * vy + vz * vz);
        const double eint = fmax2(P.smalle, ener - ekin);
        sout[V_DENS * cs + c] = dens;
        sout[V_VELX * cs + c] = vx;
        sout[V_VELY * cs + c] = vy;
        sout[V_VELZ * cs + c] = vz;
        if (defer_eos) {
          mg[0] = dens; mg[1] = vx; mg[2] = vy; mg[3] = vz; mg[4] = ener; mg[5] = eint;
        } else {
          EosOut e = eos_dens_ei(P, dens, vx, vy, vz, ener, eint);
          sout[V_PRES * cs + c] = e.pres;
          sout[V_GAMC * cs + c] = P.gamma;
          sout[V_EINT * cs + c] = e.eint;
          sout[V_ENER * cs + c] = e.ener;
          if (A.final_stage) {
            sout[V_TEMP * cs + c] = e.temp;
            sout[V_GAME * cs + c] = e.game;
          }
        }
        dens1 = dens;
      }
      if (col) {
        if (k >= I.z0) {
#pragma unroll
          for (int n = 0; n < F_NUM; n++) Hprev[n] = Hz[n];
        }
        if (xyk) {
          const int ix = cy * (tx + 1) + cx, iy = cy * tx + cx;
#pragma unroll
          for (int n = 0; n < F_NUM; n++) {
            const double d = -(dtdx * ((double)fxb[n * nfx + ix + 1] - (double)fxb[n * nfx + ix]));
            dUxy[n] = d - dtdy * ((double)fyb[n * nfy + iy + tx] - (double)fyb[n * nfy + iy]);
          }
        }
#pragma unroll
        for (int v = 0; v < NH; v++) upz[v] = upzk[v];
      }

      // ---- mass scalars, CH at a time (Hydro.ini:238-247)
#pragma unroll
      for (int c0 = 0; c0 < NSM; c0 += CH) {
        if (c0 < ns) {
          const int nc = ns - c0 < CH ? ns - c0 : CH;
          // S1: reconstruction
          Real upsk[CH], ums[CH];
#pragma unroll
          for (int j = 0; j < CH; j++) { upsk[j] = Real(0.); ums[j] = Real(0.); }
          {
            const int nrxp = rup64(nrx);
            const int utot = Cp + (xyk ? nrxp + nry : 0);
            // the column's z stencils (task u = tid < C); a slot past nc
            // loads scalar nc-1 again.  The FP64 form loads them first and
            // runs the thread's x/y tasks (u >= Cp, stencils in LDS) while
            // they are in flight; the FP32 form (two waves per SIMD, no
            // registers to spare) loads and reconstructs them in its column
            // task
            Real u5s[CH][5];
            auto load_z = [&]() {
              int64_t zc[5];
#pragma unroll
              for (int w = 0; w < 5; w++) zc[w] = cell_at(m, nbt, g, ci, cj, k - 2 + w);
#pragma unroll
              for (int j = 0; j < CH; j++) {
                const double* sp = S + (int64_t)(V_NPROP + c0 + (j < nc ? j : nc - 1)) * cs;
#pragma unroll
                for (int w = 0; w < 5; w++) u5s[j][w] = (Real)sp[zc[w]];
              }
              EARLY_Z_BARRIER;
            };
            auto recon_z = [&]() {
              const Real fl = use_flat ? flatb[fcc] : Real(1.);
#pragma unroll
              for (int j = 0; j < CH; j++)
                if (j < nc) recon1<RECON, Real>(u5s[j], fl, upsk[j], ums[j]);
            };
            constexpr bool early = FUSED_EARLY_Z;
            if (early && col) load_z();
            for (int u = early && tid < Cp ? tid + NT : tid; u < utot; u += NT) {
              if (u < Cp) {
                if (u < C) {
                  load_z();
                  recon_z();
                }
                continue;
              }
              int w = u - Cp;
              const bool isx = w < nrxp;
              if (isx && w >= nrx) continue;
              if (!isx) w -= nrxp;
              int rx, ry, st, nr;
              Real *om, *op;
              if (isx) { ry = fquot(w, d_tx2, rx); rx -= 1; st = 1; nr = nrx; om = xm; op = xp; }
              else { ry = fquot(w, d_tx, rx) - 1; st = PW; nr = nry; om = ym; op = yp; }
              const int pc = (ry + 3) * PW + (rx + 3);
              const Real fl = use_flat ? flatb[(ry + 1) * (tx + 2) + (rx + 1)] : Real(1.);
              for (int j = 0; j < nc; j++) {
                const Real* r = plane + (NH + c0 + j) * psz + pc;
                const Real u5[5] = {r[-2 * st], r[-st], r[0], r[st], r[2 * st]};
                Real up, um;
                recon1<RECON, Real>(u5, fl, up, um);
                om[j * nr + w] = um;
                op[j * nr + w] = up;
              }
            }
            if (early && col) recon_z();
          }
          __syncthreads();
          // S2: upwinded scalar fluxes
          Real Hs[CH];
#pragma unroll
          for (int j = 0; j < CH; j++) Hs[j] = Real(0.);
          if (col && k >= I.z0) {
            Real xl[CH], xr[CH];
#pragma unroll
            for (int j = 0; j < CH; j++) { xl[j] = upzs[c0 + j]; xr[j] = ums[j]; }
            if (c0 == 0) species_faces(P, xl, xr);
#pragma unroll
            for (int j = 0; j < CH; j++) {
              if (j < nc) {
                if (Hz[F_MASS] > Real(0.)) Hs[j] = xl[j] * Hz[F_MASS];
                else Hs[j] = xr[j] * Hz[F_MASS];
              }
            }
            if (regs && colin && (k == g || k == g + m.nz)) {
              const int64_t tot = (int64_t)m.nx * m.ny * (m.nz + 1);
              const int64_t o = m.bbase + (ci - g) + (int64_t)m.nx * ((cj - g) + (int64_t)m.ny * (k - g));
              reg_add_n<CH>(A.fb[2], o, tot, F_NUM + c0, nc, A.weight, A.add_flux, Hs);
            }
          }
          if (xyk) {
            for (int u = tid; u < nfx + nfy; u += NT) {
              const bool isx = u < nfx;
              const int w = isx ? u : u - nfx;
              int fi, fj, il, ir, nr;
              const Real *bp, *bm;
              Real fm;
              if (isx) {
                fj = fquot(w, d_tx1, fi);
                il = fj * (tx + 2) + fi; ir = il + 1; nr = nrx; bp = xp; bm = xm; fm = mxb[w];
              } else {
                fj = fquot(w, d_tx, fi);
                il = fj * tx + fi; ir = il + tx; nr = nry; bp = yp; bm = ym; fm = myb[w];
              }
              Real* fo = isx ? fxb : fyb;
              const int nf = isx ? nfx : nfy;
              if (c0 == 0 && P.nspec > 0) {
                Real xl[CH], xr[CH];
#pragma unroll
                for (int j = 0; j < CH; j++) {
                  xl[j] = j < nc ? bp[j * nr + il] : Real(0.);
                  xr[j] = j < nc ? bm[j * nr + ir] : Real(0.);
                }
                species_faces(P, xl, xr);
#pragma unroll
                for (int j = 0; j < CH; j++)
                  if (j < nc) fo[j * nf + w] = fm > Real(0.) ? xl[j] * fm : xr[j] * fm;
              } else {
                for (int j = 0; j < nc; j++) {
                  Real f;
                  if (fm > Real(0.)) f = bp[j * nr + il] * fm;
                  else f = bm[j * nr + ir] * fm;
                  fo[j * nf + w] = f;
                }
              }
              if (regs) {
                const int i = I.x0 + fi, jj = I.y0 + fj;
                if (isx) {
                  if ((i == g ? fi < tx : fi > 0) && (i == g || i == g + m.nx) && jj >= g && jj < g + m.ny &&
                      k >= g && k < g + m.nz) {
                    const int ex = m.nx + 1, ey = m.ny;
                    const int64_t tot = (int64_t)ex * ey * m.nz;
                    const int64_t o = m.bbase + (i - g) + (int64_t)ex * ((jj - g) + (int64_t)ey * (k - g));
                    reg_add_n<CH>(A.fb[0], o, tot, F_NUM + c0, nc, A.weight, A.add_flux, fo + w, nf);
                  }
                } else {
                  if ((jj == g ? fj < ty : fj > 0) && i >= g && i < g + m.nx && (jj == g || jj == g + m.ny) &&
                      k >= g && k < g + m.nz) {
                    const int ex = m.nx, ey = m.ny + 1;
                    const int64_t tot = (int64_t)ex * ey * m.nz;
                    const int64_t o = m.bbase + (i - g) + (int64_t)ex * ((jj - g) + (int64_t)ey * (k - g));
                    reg_add_n<CH>(A.fb[1], o, tot, F_NUM + c0, nc, A.weight, A.add_flux, fo + w, nf);
                  }
                }
              }
            }
          }
          __syncthreads();
          // S3: scalar update of level k-1, x/y differences of level k
          if (upd) {
            const int64_t c = colc + (int64_t)(k - 1) * sz;
            double Xn[8] = {0., 0., 0., 0., 0., 0., 0., 0.};  // updated species of the first chunk
            // the chunk's U0 and stage-input loads issue before the first
            // store (which the compiler cannot tell apart from them); a slot
            // past nc re-reads scalar nc-1
            double a4[CH], b4[CH];
#pragma unroll
            for (int j = 0; j < CH; j++) {
              const int64_t o = (int64_t)(V_NPROP + c0 + (j < nc ? j : nc - 1)) * cs + c;
              a4[j] = u0p[o];
              b4[j] = sin[o];
            }
#pragma unroll
            for (int j = 0; j < CH; j++) {
              if (j < nc) {
                const int n = F_NUM + c0 + j;
                const double Au = a4[j] * d0;
                double Bu = b4[j] * ds;
                double inner_n = dUxy[n] - dtdz * ((double)Hs[j] - (double)Hprev[n]);
                inner_n = inner_n + dt * (0.0 + 0.0);
                Bu = A.c1 * Au + A.c2 * Bu + A.c3 * inner_n;
                const double x = Bu / dens1;
                sout[(V_NPROP + c0 + j) * cs + c] = x;
                if (c0 == 0) Xn[j] = x;
              }
            }
            if (defer_eos && c0 == 0) {
              // the EOS of level k-1 with its updated species (Eos_wrapped
              // after updateSoln)
              const EosOut e = eos_dens_ei(P, mg[0], mg[1], mg[2], mg[3], mg[4], mg[5], Xn);
              sout[V_PRES * cs + c] = e.pres;
              sout[V_GAMC * cs + c] = e.gamc;
              sout[V_EINT * cs + c] = e.eint;
              sout[V_ENER * cs + c] = e.ener;
              if (A.final_stage) {
                sout[V_TEMP * cs + c] = e.temp;
                sout[V_GAME * cs + c] = e.game;
              }
            }
          }
          if (col) {
#pragma unroll
            for (int j = 0; j < CH; j++) {
              if (j < nc) {
                const int n = F_NUM + c0 + j;
                if (k >= I.z0) Hprev[n] = Hs[j];
                if (xyk) {
                  const int ix = cy * (tx + 1) + cx, iy = cy * tx + cx;
                  const double d = -(dtdx * ((double)fxb[j * nfx + ix + 1] - (double)fxb[j * nfx + ix]));
                  dUxy[n] = d - dtdy * ((double)fyb[j * nfy + iy + tx] - (double)fyb[j * nfy + iy]);
                }
                upzs[c0 + j] = upsk[j];
              }
            }
          }
        }
      }
    }
  }
}

// reset of the per-XCD queue heads (a kernel, so it can sit in a graph)
__global__ void k_reset_counters(int* c, int n) {
  if ((int)threadIdx.x < n) c[threadIdx.x] = 0;
}

template <int NSM, int RECON, int RIEMANN>
hipError_t launch_t(const FusedArgs& A, int grid, size_t lds, hipStream_t st) {
  hipLaunchKernelGGL((FUSED_NAME(k_fused, FUSED_SFX)<NSM, RECON, RIEMANN>), dim3(grid), dim3(FUSED_THREADS), lds, st,
                     A);
  return hipGetLastError();
}

template <int NSM>
hipError_t launch_r(const FusedArgs& A, int grid, size_t lds, hipStream_t st) {
  // the hybrid switch only changes HLLC (HLL is what it switches to)
  const int rs = A.P.riemann == RIEMANN_HLL ? RIEMANN_HLL : A.P.hybrid ? RIEMANN_HLLC_HYBRID : RIEMANN_HLLC;
#define FUSED_R(RC)                                                                   \
  (rs == RIEMANN_HLL    ? launch_t<NSM, RC, RIEMANN_HLL>(A, grid, lds, st)           \
   : rs == RIEMANN_HLLC ? launch_t<NSM, RC, RIEMANN_HLLC>(A, grid, lds, st)          \
                        : launch_t<NSM, RC, RIEMANN_HLLC_HYBRID>(A, grid, lds, st))
  switch (A.P.recon) {
    case RECON_MP5: return FUSED_R(RECON_MP5);
    case RECON_PLM: return FUSED_R(RECON_PLM);
    case RECON_FOG: return FUSED_R(RECON_FOG);
    default: return FUSED_R(RECON_WENO5);
  }
#undef FUSED_R
}

}  // namespace

size_t FUSED_NAME(fused_lds_bytes, FUSED_SFX)(int ns, int tx, int ty) {
  return sizeof(Real) * (size_t)lds_layout(tx, ty, NH + ns).total;
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
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(
        blocks_per_cu, FUSED_NAME(k_fused, FUSED_SFX)<0, RECON_WENO5, RIEMANN_HLLC>, FUSED_THREADS, lds);
  return hipOccupancyMaxActiveBlocksPerMultiprocessor(
      blocks_per_cu, FUSED_NAME(k_fused, FUSED_SFX)<8, RECON_WENO5, RIEMANN_HLLC>, FUSED_THREADS, lds);
}

// static LDS of the kernel (work-item index, neighbour table): a launch's
// dynamic LDS must leave room for it under the per-workgroup limit
size_t FUSED_NAME(fused_static_lds, FUSED_SFX)(int ns) {
  hipFuncAttributes fa;
  hipError_t e = ns == 0
                     ? hipFuncGetAttributes(&fa, reinterpret_cast<const void*>(
                                                     FUSED_NAME(k_fused, FUSED_SFX)<0, RECON_WENO5, RIEMANN_HLLC>))
                     : hipFuncGetAttributes(&fa, reinterpret_cast<const void*>(
                                                     FUSED_NAME(k_fused, FUSED_SFX)<8, RECON_WENO5, RIEMANN_HLLC>));
  const size_t own = sizeof(int) + 27 * sizeof(int64_t);
  if (e != hipSuccess) return 1024;  // a safe upper bound of the kernel's own variables
  return fa.sharedSizeBytes > own ? fa.sharedSizeBytes : own;
}

}  // namespace spark
