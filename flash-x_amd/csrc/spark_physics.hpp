// spark_physics.hpp -- device-side physics of the Spark RK stage, templated
// on the arithmetic type T (double for the plain FP64 form, float for the
// FP32-flux form).  For T = double every expression keeps the evaluation
// order of the Fortran it restates, and the translation unit that uses it
// is compiled with -ffp-contract=off, so results round exactly like the
// reference's operation sequence.
//
// Reference (Flash-X/source/physics/Hydro/HydroMain/Spark/):
//   reconstruct.F90-mc:310-394 + Hydro.ini:202-220 (WENO5, WENO-Z weights)
//   mp5/reconstruct.F90:28-100, tvd2/reconstruct.F90:28-54, fog/reconstruct.F90
//   riemann.F90-mc:42-282 (HLLC), tvd2/riemann.F90:63-141 (HLL)
//   prim2con.F90:317-345, prim2flx.F90:381-463
//   Hydro.ini:222-235 (ensurePhysicalState), :238-247 (scalar fluxes),
//   :285-301 (artificial viscosity)
//   hy_rk_updateSoln.F90-mc:194-304, Eos/EosMain/Gamma/Eos.F90:238-247,
//   Eos/EosMain/Eos_getData.F90:114-169, Eos/EosMain/Eos_putData.F90:207-224
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace spark {

// unk variable order (alphabetical, as setup.py sorts Spark + Gamma)
enum Var : int {
  V_DENS = 0, V_EINT = 1, V_ENER = 2, V_GAMC = 3, V_GAME = 4, V_PRES = 5,
  V_SHOK = 6, V_TEMP = 7, V_VELX = 8, V_VELY = 9, V_VELZ = 10, V_NPROP = 11
};
// pencil primitive order, Spark.h:12-18
enum Rec : int { R_DENS = 0, R_VELX = 1, R_VELY = 2, R_VELZ = 3, R_PRES = 4, R_GAMC = 5, R_RHOE = 6, R_NUM = 7 };
// conserved/flux order, Spark.h:30-34
enum Flx : int { F_MASS = 0, F_XMOM = 1, F_YMOM = 2, F_ZMOM = 3, F_ENER = 4, F_NUM = 5 };

enum Recon : int { RECON_WENO5 = 0, RECON_MP5 = 1, RECON_PLM = 2, RECON_FOG = 3 };
enum Riemann : int { RIEMANN_HLLC = 0, RIEMANN_HLL = 1 };

template <typename T> struct Num;
template <> struct Num<double> {
  static __device__ __forceinline__ double div(double a, double b) { return a / b; }
  static __device__ __forceinline__ double rcp(double a) { return 1. / a; }
  static __device__ __forceinline__ double sqr(double a) { return __builtin_sqrt(a); }
  static constexpr double weno_eps = 1.e-36;
  static constexpr double flat_floor = 1e-99;
  static constexpr bool clamp_ratio = false;
  static constexpr double ratio_max = 0.0;
};
// FP32: hardware reciprocal and square root (1 ulp); the WENO-Z ratio is
// clamped so that tau/(beta+eps) cannot overflow the float range, and the
// flattening floor is raised above the float underflow threshold.
template <> struct Num<float> {
  static __device__ __forceinline__ float div(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
  static __device__ __forceinline__ float rcp(float a) { return __builtin_amdgcn_rcpf(a); }
  static __device__ __forceinline__ float sqr(float a) { return __builtin_amdgcn_sqrtf(a); }
  static constexpr float weno_eps = 1.e-36f;
  static constexpr float flat_floor = 1e-30f;
  static constexpr bool clamp_ratio = true;
  static constexpr float ratio_max = 1e30f;
};

// Fortran MAX/MIN of two non-NaN reals
template <typename T> __device__ __forceinline__ T fmax2(T a, T b) { return (b > a) ? b : a; }
template <typename T> __device__ __forceinline__ T fmin2(T a, T b) { return (b < a) ? b : a; }

// ---------------------------------------------------------------------------
// WENO5 reconstruction with WENO-Z weights (reconstruct.F90-mc:330-391)
template <typename T>
__device__ __forceinline__ void weno5(T um2, T um1, T u0, T up1, T up2, T flat, T& oup, T& oum) {
  const T c1p1_0 = T(2. / 6.), c1p1_1 = T(-7. / 6.), c1p1_2 = T(11. / 6.);
  const T c1p2_0 = T(-1. / 6.), c1p2_1 = T(5. / 6.), c1p2_2 = T(2. / 6.);
  const T c1p3_0 = T(2. / 6.), c1p3_1 = T(5. / 6.), c1p3_2 = T(-1. / 6.);
  const T c1m1_0 = T(-1. / 6.), c1m1_1 = T(5. / 6.), c1m1_2 = T(2. / 6.);
  const T c1m2_0 = T(2. / 6.), c1m2_1 = T(5. / 6.), c1m2_2 = T(-1. / 6.);
  const T c1m3_0 = T(11. / 6.), c1m3_1 = T(-7. / 6.), c1m3_2 = T(2. / 6.);
  const T n13o12 = T(13. / 12.);
  const T eps = Num<T>::weno_eps;
  T W5p1 = c1p1_0 * um2 + c1p1_1 * um1 + c1p1_2 * u0;
  T W5p2 = c1p2_0 * um1 + c1p2_1 * u0 + c1p2_2 * up1;
  T W5p3 = c1p3_0 * u0 + c1p3_1 * up1 + c1p3_2 * up2;
  T W5m1 = c1m1_0 * um2 + c1m1_1 * um1 + c1m1_2 * u0;
  T W5m2 = c1m2_0 * um1 + c1m2_1 * u0 + c1m2_2 * up1;
  T W5m3 = c1m3_0 * u0 + c1m3_1 * up1 + c1m3_2 * up2;
  T t, s;
  t = um2 - T(2.) * um1 + u0;
  s = um2 - T(4.) * um1 + T(3.) * u0;
  T b1 = n13o12 * (t * t) + T(0.25) * (s * s);
  t = um1 - T(2.) * u0 + up1;
  s = um1 - up1;
  T b2 = n13o12 * (t * t) + T(0.25) * (s * s);
  t = u0 - T(2.) * up1 + up2;
  s = T(3.) * u0 - T(4.) * up1 + up2;
  T b3 = n13o12 * (t * t) + T(0.25) * (s * s);
  T tau = __builtin_fabs(b1 - b3);
  T r1 = Num<T>::div(tau, b1 + eps);
  T r2 = Num<T>::div(tau, b2 + eps);
  T r3 = Num<T>::div(tau, b3 + eps);
  if constexpr (Num<T>::clamp_ratio) {
    r1 = fmin2(r1, Num<T>::ratio_max);
    r2 = fmin2(r2, Num<T>::ratio_max);
    r3 = fmin2(r3, Num<T>::ratio_max);
  }
  T A1 = T(0.1) * (T(1.) + r1), A2 = T(0.6) * (T(1.) + r2), A3 = T(0.3) * (T(1.) + r3);
  T inv = Num<T>::rcp(A1 + A2 + A3);
  T up = (A1 * inv) * W5p1 + (A2 * inv) * W5p2 + (A3 * inv) * W5p3;
  up = flat * up + (T(1.) - flat) * u0;
  A1 = T(0.3) * (T(1.) + r1);
  A2 = T(0.6) * (T(1.) + r2);
  A3 = T(0.1) * (T(1.) + r3);
  inv = Num<T>::rcp(A1 + A2 + A3);
  T um = (A1 * inv) * W5m1 + (A2 * inv) * W5m2 + (A3 * inv) * W5m3;
  um = flat * um + (T(1.) - flat) * u0;
  if ((up - u0) * (u0 - um) <= T(0.)) { up = u0; um = u0; }
  oup = up;
  oum = um;
}

template <typename T> __device__ __forceinline__ T minmod(T a, T b) {
  return T(.5) * (__builtin_copysign(T(1.), a) + __builtin_copysign(T(1.), b)) *
         fmin2(__builtin_fabs(a), __builtin_fabs(b));
}
template <typename T> __device__ __forceinline__ T median3(T a, T b, T c) { return a + minmod(b - a, c - a); }

// MP5 (mp5/reconstruct.F90:47-86); d2mp/d2mm are evaluated before both
// limiter branches (the reference reads them in the second branch even
// when the first branch did not set them).
template <typename T>
__device__ __forceinline__ void mp5(T um2, T um1, T u0, T up1, T up2, T& oup, T& oum) {
  const T B1 = T(1. / 60.), B2 = T(4. / 3.), alpha = T(4.), eps = T(0.);
  T Uplus = B1 * (T(2.) * um2 - T(13.) * um1 + T(47.) * u0 + T(27.) * up1 - T(3.) * up2);
  T Uminus = B1 * (T(-3.) * um2 + T(27.) * um1 + T(47.) * u0 - T(13.) * up1 + T(2.) * up2);
  T del_p = up1 - u0;
  T del_m = u0 - um1;
  T d2m = um2 - T(2.) * um1 + u0;
  T d2 = um1 - T(2.) * u0 + up1;
  T d2p = u0 - T(2.) * up1 + up2;
  T d2mp = minmod(minmod(T(4) * d2 - d2p, T(4) * d2p - d2), minmod(d2, d2p));
  T d2mm = minmod(minmod(T(4) * d2m - d2, T(4) * d2 - d2m), minmod(d2m, d2));
  T ump = u0 + minmod(del_p, alpha * del_m);
  if ((Uplus - u0) * (Uplus - ump) >= eps) {
    T uul = u0 + alpha * del_m;
    T umd = T(0.5) * (u0 + up1) - T(0.5) * d2mp;
    T ulc = u0 + T(0.5) * del_m + B2 * d2mm;
    T umin = fmax2(fmin2(fmin2(u0, up1), umd), fmin2(fmin2(u0, uul), ulc));
    T umax = fmin2(fmax2(fmax2(u0, up1), umd), fmax2(fmax2(u0, uul), ulc));
    Uplus = median3(umin, Uplus, umax);
  }
  ump = u0 - minmod(del_m, alpha * del_p);
  if ((u0 - Uminus) * (ump - Uminus) >= eps) {
    T uul = u0 - alpha * del_p;
    T umd = T(0.5) * (u0 + um1) - T(0.5) * d2mm;
    T ulc = u0 - T(0.5) * del_p + B2 * d2mp;
    T umin = fmax2(fmin2(fmin2(u0, um1), umd), fmin2(fmin2(u0, uul), ulc));
    T umax = fmin2(fmax2(fmax2(u0, um1), umd), fmax2(fmax2(u0, uul), ulc));
    Uminus = median3(umin, Uminus, umax);
  }
  oup = Uplus;
  oum = Uminus;
}

// PLM with minmod slopes (tvd2/reconstruct.F90:43-52)
template <typename T> __device__ __forceinline__ void plm(T um1, T u0, T up1, T& oup, T& oum) {
  T delbar = minmod(up1 - u0, u0 - um1);
  oup = u0 + T(0.5) * delbar;
  oum = u0 - T(0.5) * delbar;
}

// stencil radius per reconstruction
template <int RECON> struct ReconRadius { static constexpr int R = 2; };
template <> struct ReconRadius<RECON_PLM> { static constexpr int R = 1; };
template <> struct ReconRadius<RECON_FOG> { static constexpr int R = 0; };

template <int RECON, typename T>
__device__ __forceinline__ void recon1(const T* u /* u[-2..2] at u[0..4] */, T flat, T& up, T& um) {
  if constexpr (RECON == RECON_WENO5) weno5<T>(u[0], u[1], u[2], u[3], u[4], flat, up, um);
  else if constexpr (RECON == RECON_MP5) mp5<T>(u[0], u[1], u[2], u[3], u[4], up, um);
  else if constexpr (RECON == RECON_PLM) plm<T>(u[1], u[2], u[3], up, um);
  else { up = u[2]; um = u[2]; }
}

// ---------------------------------------------------------------------------
// prim2con / prim2flx (hydro branch)
template <typename T> __device__ __forceinline__ void prim2con(const T* V, T* U) {
  const T B2 = T(0.);
  T u2 = V[R_VELX] * V[R_VELX] + V[R_VELY] * V[R_VELY] + V[R_VELZ] * V[R_VELZ];
  U[F_MASS] = V[R_DENS];
  U[F_XMOM] = V[R_DENS] * V[R_VELX];
  U[F_YMOM] = V[R_DENS] * V[R_VELY];
  U[F_ZMOM] = V[R_DENS] * V[R_VELZ];
  U[F_ENER] = T(0.5) * V[R_DENS] * u2 + V[R_RHOE] + T(0.5) * B2;
}

template <int DIR, typename T> __device__ __forceinline__ void prim2flx(const T* V, T* F) {
  T u2 = V[R_VELX] * V[R_VELX] + V[R_VELY] * V[R_VELY] + V[R_VELZ] * V[R_VELZ];
  T E = T(0.5) * V[R_DENS] * u2 + V[R_RHOE];
  T Ptot = V[R_PRES];
  T vn = V[R_VELX + DIR];
  F[F_MASS] = V[R_DENS] * vn;
  F[F_XMOM] = F[F_MASS] * V[R_VELX];
  F[F_YMOM] = F[F_MASS] * V[R_VELY];
  F[F_ZMOM] = F[F_MASS] * V[R_VELZ];
  F[F_XMOM + DIR] = F[F_MASS] * V[R_VELX + DIR] + Ptot;
  F[F_ENER] = (E + Ptot) * vn;
}

// HLLC (riemann.F90-mc:60-274).  Returns 1 for a negative squared sound
// speed (the reference sets ierr and leaves Fstar unset; callers zero it).
template <int DIR, typename T>
__device__ __forceinline__ int hllc(const T* VL, const T* VR, bool hybrid, bool inShock, T tiny, T* F) {
  T velNL = VL[R_VELX + DIR];
  T velNR = VR[R_VELX + DIR];
  T aL2 = Num<T>::div(VL[R_GAMC] * VL[R_PRES], VL[R_DENS]);
  T aR2 = Num<T>::div(VR[R_GAMC] * VR[R_PRES], VR[R_DENS]);
  const T magNL = T(0.), magNR = T(0.);
  if ((VL[R_DENS] < tiny && VL[R_DENS] > T(0.)) || (VR[R_DENS] < tiny && VR[R_DENS] > T(0.)) ||
      (VL[R_PRES] < tiny && VL[R_PRES] > T(0.)) || (VR[R_PRES] < tiny && VR[R_PRES] > T(0.))) {
#pragma unroll
    for (int n = 0; n < 5; n++) F[n] = T(0.);
    return 0;
  } else if (aL2 < T(0.) || aR2 < T(0.)) {
    return 1;
  }
  T cfL = Num<T>::sqr(aL2);
  T cfR = Num<T>::sqr(aR2);
  T SL = fmin2(velNL - cfL, velNR - cfR);
  T SR = fmax2(velNL + cfL, velNR + cfR);
  T totalPresL = VL[R_PRES];
  T totalPresR = VR[R_PRES];
  T UL[5], UR[5], FL[5], FR[5], UCsL[5], UCsR[5];
  prim2con(VL, UL);
  prim2con(VR, UR);
  prim2flx<DIR>(VL, FL);
  prim2flx<DIR>(VR, FR);
  T qStar = VR[R_DENS] * velNR * (SR - velNR) - VL[R_DENS] * velNL * (SL - velNL) + totalPresL - totalPresR -
            magNL * magNL + magNR * magNR;
  qStar = Num<T>::div(qStar, VR[R_DENS] * (SR - velNR) - VL[R_DENS] * (SL - velNL));
  T numerL = SL - velNL;
  T denomL = SL - qStar;
  T numerR = SR - velNR;
  T denomR = SR - qStar;
  T pStar = VL[R_DENS] * numerL * (qStar - velNL) + totalPresL;
  T dStarL = Num<T>::div(UL[F_MASS] * numerL, denomL);
  T dStarR = Num<T>::div(UR[F_MASS] * numerR, denomR);
  UCsL[F_MASS] = dStarL;
  UCsL[F_ENER] = Num<T>::div(UL[F_ENER] * numerL, denomL) + Num<T>::div((pStar * qStar - totalPresL * velNL), denomL);
  UCsR[F_MASS] = dStarR;
  UCsR[F_ENER] = Num<T>::div(UR[F_ENER] * numerR, denomR) + Num<T>::div((pStar * qStar - totalPresR * velNR), denomR);
#pragma unroll
  for (int d = 0; d < 3; d++) {
    if (d == DIR) {
      UCsL[F_XMOM + d] = dStarL * qStar;
      UCsR[F_XMOM + d] = dStarR * qStar;
    } else {
      UCsL[F_XMOM + d] = Num<T>::div(UL[F_XMOM + d] * numerL, denomL);
      UCsR[F_XMOM + d] = Num<T>::div(UR[F_XMOM + d] * numerR, denomR);
    }
  }
  if (SL >= T(0.)) {
#pragma unroll
    for (int n = 0; n < 5; n++) F[n] = FL[n];
  } else if ((SL < T(0.)) && (qStar >= T(0.))) {
#pragma unroll
    for (int n = 0; n < 5; n++) F[n] = FL[n] + SL * (UCsL[n] - UL[n]);
  } else if ((qStar < T(0.)) && (SR >= T(0.))) {
#pragma unroll
    for (int n = 0; n < 5; n++) F[n] = FR[n] + SR * (UCsR[n] - UR[n]);
  } else {
#pragma unroll
    for (int n = 0; n < 5; n++) F[n] = FR[n];
  }
  if (hybrid && inShock) {
    if (SL > T(0.)) {
#pragma unroll
      for (int n = 0; n < 5; n++) F[n] = FL[n];
    } else if (SR < T(0.)) {
#pragma unroll
      for (int n = 0; n < 5; n++) F[n] = FR[n];
    } else {
#pragma unroll
      for (int n = 0; n < 5; n++)
        F[n] = Num<T>::div(SR * FL[n] - SL * FR[n] + SR * SL * (UR[n] - UL[n]), SR - SL);
    }
  }
  return 0;
}

// HLL(E) (tvd2/riemann.F90:86-139)
template <int DIR, typename T> __device__ __forceinline__ int hll(const T* VL, const T* VR, T* F) {
  const T tiny = T(1.e-32);
#pragma unroll
  for (int n = 0; n < 5; n++) F[n] = T(0.);
  T velNL = VL[R_VELX + DIR];
  T velNR = VR[R_VELX + DIR];
  T aL2 = Num<T>::div(VL[R_GAMC] * VL[R_PRES], VL[R_DENS]);
  T aR2 = Num<T>::div(VR[R_GAMC] * VR[R_PRES], VR[R_DENS]);
  if ((VL[R_DENS] < tiny && VL[R_DENS] > T(0.)) || (VR[R_DENS] < tiny && VR[R_DENS] > T(0.)) ||
      (VL[R_PRES] < tiny && VL[R_PRES] > T(0.)) || (VR[R_PRES] < tiny && VR[R_PRES] > T(0.))) {
    return 0;
  } else if (aL2 < T(0.) || aR2 < T(0.)) {
    return 1;
  }
  T cfL = Num<T>::sqr(aL2);
  T cfR = Num<T>::sqr(aR2);
  T SL = fmin2(velNL - cfL, velNR - cfR);
  T SR = fmax2(velNL + cfL, velNR + cfR);
  T UL[5], UR[5], FL[5], FR[5];
  prim2con(VL, UL);
  prim2con(VR, UR);
  prim2flx<DIR>(VL, FL);
  prim2flx<DIR>(VR, FR);
  if (SL > T(0.)) {
#pragma unroll
    for (int n = 0; n < 5; n++) F[n] = FL[n];
  } else if (SR < T(0.)) {
#pragma unroll
    for (int n = 0; n < 5; n++) F[n] = FR[n];
  } else {
#pragma unroll
    for (int n = 0; n < 5; n++) F[n] = Num<T>::div(SR * FL[n] - SL * FR[n] + SR * SL * (UR[n] - UL[n]), SR - SL);
  }
  return 0;
}

// ---------------------------------------------------------------------------
// Parameters passed by value to every kernel
struct DevParams {
  double gamma, cvisc, smlrho, smallp, smalle, smallx, eintSwitch, tiny, cfl;
  double gam1inv, ggprodinv, abar;  // EOS constants (host-computed like Gamma/Eos.F90:177,240-241)
  double grav[3];
  int ndim, ng, ns, use_grav, use_flat, hybrid, recon, riemann;
  int nspec, eos;       // species among the scalars; 0 gamma law, 1 Multigamma
  double gasConst;
  double mg_gm1[8];     // 1/(gamma_s - 1) of each species
  double mg_ainv[8];    // 1/A_s
  // geometry: 0 Cartesian, 1 cylindrical (r, z), 2 spherical (r[, theta]).
  // geo holds per block (BlockMeta::id), over the guard-celled box of
  // geo_nx x geo_ny cells: the cell-centre, left-edge and right-edge radius
  // (geo_nx each), then the area of each cell's low x face and each cell's
  // volume (geo_nx * geo_ny each, x fastest) -- the Grid's hy_xCenter,
  // hy_xLeft, hy_xRight, hy_farea and hy_cvol (Hydro.ini:40-56)
  int geom, geo_nx, geo_ny;
  const double* geo;
};
__host__ __device__ __forceinline__ int64_t geo_stride(int nx, int ny) { return 3 * (int64_t)nx + 2 * (int64_t)nx * ny; }

// geometric sources of the update (geoFacs, hy_rk_updateSoln.F90-mc:131-
// 170) from the star state V (vd, vx, vz, p) and the star conserved theta
// momentum uy: cylindrical alpha = 1, spherical alpha = 2 plus (rho v)^2 /
// rho / r and the whole source scaled by dx / dx_sph,
// dx_sph = (r_R^3 - r_L^3) / (3 r_c^2); the azimuthal momentum is the z one
__device__ __forceinline__ void geo_sources(const DevParams& P, int id, int i, double dx, double rc, double vd,
                                            double vx, double vz, double pres, double uy, double* Sgeo) {
  const double* g = P.geo + (int64_t)id * geo_stride(P.geo_nx, P.geo_ny);
  const double alpha = P.geom == 2 ? 2. : 1.;
  Sgeo[F_XMOM] = (vd * vz * vz + alpha * pres) / rc;
  Sgeo[F_ZMOM] = -(vd * vz * vx / rc);
  if (P.geom == 2) {
    const double rl = g[P.geo_nx + i], rr = g[2 * P.geo_nx + i];
    const double dx_sph = (rr * rr * rr - rl * rl * rl) / (3. * (rc * rc));
    Sgeo[F_XMOM] = Sgeo[F_XMOM] + uy * uy / vd / rc;
#pragma unroll
    for (int n = 0; n < F_NUM; n++) Sgeo[n] = Sgeo[n] * dx / dx_sph;
  }
}

// geoFacs for the x faces of cell (i, j) (hy_rk_updateSoln.F90-mc:98-175,
// hy_rk_correctFluxes.F90:341-358): facM, facP = farea * dx / cvol of the
// low and high face, and the cell-centre radius; all factors 1 (and rc 0)
// in Cartesian geometry.  Returns false where the reference zeroes them
// (a cell centre at negative radius).
__device__ __forceinline__ bool geo_facs(const DevParams& P, int id, int i, int j, double dx, double& facM,
                                         double& facP, double& rc) {
  facM = 1.0;
  facP = 1.0;
  rc = 0.0;
  if (P.geom == 0) return true;
  const double* g = P.geo + (int64_t)id * geo_stride(P.geo_nx, P.geo_ny);
  const double* area = g + 3 * P.geo_nx;
  const double* vol = area + (int64_t)P.geo_nx * P.geo_ny;
  const int64_t q = (int64_t)j * P.geo_nx + i;
  rc = g[i];
  facM = area[q] * dx / vol[q];
  facP = area[q + 1] * dx / vol[q];
  if (rc < 0.0) {
    facM = 0.;
    facP = 0.;
    return false;
  }
  return true;
}

// Per-block table entry (device resident)
struct BlockMeta {
  int64_t base;     // offset (doubles) of variable 0 in the state pools
  int64_t cstride;  // variable stride (doubles) in the state pools
  int64_t obase;    // offset in one-variable cell pools (flattening, scratch)
  int64_t dbase;    // offset in the plain form's divergence-term pool (5 + ns
                    // fields of cstride doubles per block)
  int64_t bbase;    // offset in the flux-register pools (interior faces)
  int nx, ny, nz;   // interior cells
  int NX, NY, NZ;   // guard-celled extents
  double dx, dy, dz;
  int level, id;
};

// EOS in MODE_DENS_EI on one cell held in registers (Eos_getData.F90:114-169
// -> Eos -> Eos_putData.F90:207-224).  The gamma law (Gamma/Eos.F90:238-247)
// or, with species mass fractions X, the Multigamma mixture
// (Multigamma/Eos.F90: abar = 1/sum(X/A), gamc = 1 + 1/(abar sum(X/(gamma-1)/A)),
// p = rho e (gamc - 1), T = e (gamc - 1) abar / R).
struct EosOut { double pres, eint, ener, temp, game, gamc; };
__device__ __forceinline__ EosOut eos_dens_ei(const DevParams& P, double dens, double vx, double vy, double vz,
                                               double ener, double eint_in, const double* X = nullptr) {
  double ekin = 0.5 * (vx * vx + vy * vy + vz * vz);
  double eint = eint_in;
  if (ener - ekin > fmax2(P.smalle, P.eintSwitch * ekin)) eint = ener - ekin;
  eint = fmax2(eint, P.smalle);
  EosOut o;
  if (P.eos == 1 && X != nullptr) {
    double abarInv = 0.0, rt = 0.0;
#pragma unroll
    for (int s = 0; s < 8; s++)
      if (s < P.nspec) abarInv = abarInv + X[s] * P.mg_ainv[s];
    const double abar = 1.e0 / abarInv;
#pragma unroll
    for (int s = 0; s < 8; s++)
      if (s < P.nspec) rt = rt + (X[s] * P.mg_gm1[s]) * P.mg_ainv[s];
    o.gamc = 1.0e0 + 1.0e0 / (rt * abar);
    o.pres = dens * eint * (o.gamc - 1.e0);
    o.temp = eint * (o.gamc - 1.e0) * abar / P.gasConst;
  } else {
    o.gamc = P.gamma;
    o.pres = dens * eint * P.gam1inv;
    o.temp = eint * P.ggprodinv * P.abar;
  }
  o.eint = eint;
  o.ener = eint + ekin;
  o.game = o.pres / (eint * dens) + 1;
  return o;
}

}  // namespace spark
