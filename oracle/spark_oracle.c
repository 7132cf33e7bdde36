/*
 * spark_oracle.c -- CPU restatement (FP64, plain C) of the Flash-X Spark
 * unsplit-hydro RK stage and the Gamma EOS closure.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP path.  Nothing in the
 * product links or loads this file.  Built with -O2 -ffp-contract=off so
 * that every expression rounds exactly as written; each expression keeps
 * the evaluation order of the Fortran line it restates (Fortran evaluates
 * a*b*c as (a*b)*c and sums left to right, as C does).
 *
 * Reference files restated (paths relative to Flash-X/source/):
 *   physics/Hydro/HydroMain/Spark/{Hydro.F90-mc, Hydro.ini, Hydro_funcs.F90-mc,
 *     hy_rk_getFaceFlux.F90-mc, reconstruct.F90-mc, riemann.F90-mc,
 *     prim2con.F90, prim2flx.F90, hy_rk_updateSoln.F90-mc,
 *     hy_rk_correctFluxes.F90, Hydro_computeDt.F90, mp5/reconstruct.F90,
 *     tvd2/reconstruct.F90, tvd2/riemann.F90, fog/reconstruct.F90,
 *     Inhost/hy_local.ini}
 *   physics/Eos/EosMain/{Eos_wrapped.F90, Eos_getData.F90, Eos_putData.F90,
 *     Gamma/Eos.F90, Gamma/Eos_init.F90}
 *
 * Pinning status of each part (see tests/test_oracle_*.py):
 *   prim2con/prim2flx, HLLC, HLL, WENO5-Z, MP5, PLM, FOG, Gamma EOS, update,
 *   dt, flux correction: pinned by closed forms, invariants and hand cases.
 *   flattening, shockDetect: hand-computed cases on special inputs.
 *   artificial viscosity: the face flux of a supersonic first-order pencil
 *   equals F(V_{i-1}) + cvisc max(-du, 0) (U_{i-1} - U_i) (nothing added
 *   on expansion).  mass-scalar fluxes: rho X follows the upwind RK Taylor
 *   polynomial for u > 0 and u < 0.  gravity source: on a uniform state one
 *   RK3 step gives d(rho u) = rho g dt and dE = m0.g dt + rho |g|^2 dt^2 / 2.
 *   Planted mistakes in these three parts (tests/test_oracle_mutants.py)
 *   each fail one of those pins.
 *   Whole-solver trajectories: pinned by the Sod exact solution (Toro ch.4)
 *   and the Sedov analytic profile held by the reference
 *   (Simulation/SimulationMain/Sedov/sedovSolution.dat).
 */
#include "spark_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* pencil (snake) variable indices, Spark/Spark.h:12-18 */
enum { HY_DENS = 0, HY_VELX = 1, HY_VELY = 2, HY_VELZ = 3, HY_PRES = 4, HY_GAMC = 5, HY_RHOE = 6,
       HY_NUM_VARS = 7 };
/* conserved/flux indices, Spark/Spark.h:30-34 */
enum { HY_MASS = 0, HY_XMOM = 1, HY_YMOM = 2, HY_ZMOM = 3, HY_ENER = 4, HY_NUM_FLUX = 5 };

#define MAXREC 32
#define MAXFLX 32

/* Fortran MAX/MIN of two non-NaN reals */
static inline double fmax2(double a, double b) { return (b > a) ? b : a; }
static inline double fmin2(double a, double b) { return (b < a) ? b : a; }

int orc_nunk(const orc_params* P) { return ORC_NPROP + P->nscalars; }
int orc_nflux(const orc_params* P) { return HY_NUM_FLUX + P->nscalars; }

void orc_gc_dims(const orc_params* P, const orc_block* b, int* NX, int* NY, int* NZ) {
  int g = P->nguard;
  *NX = b->nx + 2 * g;
  *NY = (P->ndim >= 2) ? b->ny + 2 * g : 1;
  *NZ = (P->ndim == 3) ? b->nz + 2 * g : 1;
}

static inline size_t cidx(int NX, int NY, int i, int j, int k) {
  return (size_t)i + (size_t)NX * ((size_t)j + (size_t)NY * (size_t)k);
}
/* face arrays are dimensioned (NX+1, NY+1, NZ+1) */
static inline size_t fidx(int NX, int NY, int i, int j, int k) {
  return (size_t)i + (size_t)(NX + 1) * ((size_t)j + (size_t)(NY + 1) * (size_t)k);
}

static void interior_limits(const orc_params* P, const orc_block* b, int lim[3][2]) {
  int n[3] = {b->nx, b->ny, b->nz};
  for (int d = 0; d < 3; d++) {
    if (d < P->ndim) { lim[d][0] = P->nguard; lim[d][1] = P->nguard + n[d] - 1; }
    else { lim[d][0] = 0; lim[d][1] = 0; }
  }
}

/* setLims, Spark/Hydro_funcs.F90-mc:584-598 */
static void set_lims(const orc_params* P, const orc_block* b, int ngoff, int lim[3][2]) {
  interior_limits(P, b, lim);
  for (int d = 0; d < P->ndim; d++) { lim[d][0] -= ngoff; lim[d][1] += ngoff; }
}

/* ------------------------------------------------------------------ */
/* prim2con, Spark/prim2con.F90:317-345 (hydro branch)                 */
void orc_prim2con(const double* V, double* U) {
  double B2 = 0.0;
  double u2 = V[HY_VELX] * V[HY_VELX] + V[HY_VELY] * V[HY_VELY] + V[HY_VELZ] * V[HY_VELZ];
  U[HY_MASS] = V[HY_DENS];
  U[HY_XMOM] = V[HY_DENS] * V[HY_VELX];
  U[HY_YMOM] = V[HY_DENS] * V[HY_VELY];
  U[HY_ZMOM] = V[HY_DENS] * V[HY_VELZ];
  U[HY_ENER] = 0.5 * V[HY_DENS] * u2 + V[HY_RHOE] + 0.5 * B2;
}

/* prim2flx, Spark/prim2flx.F90:381-463 (hydro branch, :440-459) */
void orc_prim2flx(int dir, const double* V, double* F) {
  double u2 = V[HY_VELX] * V[HY_VELX] + V[HY_VELY] * V[HY_VELY] + V[HY_VELZ] * V[HY_VELZ];
  double E = 0.5 * V[HY_DENS] * u2 + V[HY_RHOE];
  double Ptot = V[HY_PRES];
  double vn = V[HY_VELX + dir];
  F[HY_MASS] = V[HY_DENS] * vn;
  F[HY_XMOM] = F[HY_MASS] * V[HY_VELX];
  F[HY_YMOM] = F[HY_MASS] * V[HY_VELY];
  F[HY_ZMOM] = F[HY_MASS] * V[HY_VELZ];
  F[HY_XMOM + dir] = F[HY_MASS] * V[HY_VELX + dir] + Ptot;
  F[HY_ENER] = (E + Ptot) * vn;
}

/* HLLC, Spark/riemann.F90-mc:42-282 (no GLM).  Returns ierr. */
int orc_riemann_hllc(const orc_params* P, int dir, const double* VL, const double* VR,
                     int inShock, double* F) {
  const double tiny = P->tiny;
  double velNL = VL[HY_VELX + dir];
  double velNR = VR[HY_VELX + dir];
  double aL2 = VL[HY_GAMC] * VL[HY_PRES] / VL[HY_DENS];
  double aR2 = VR[HY_GAMC] * VR[HY_PRES] / VR[HY_DENS];
  double magNL = 0.0, magNR = 0.0;
  if ((VL[HY_DENS] < tiny && VL[HY_DENS] > 0.) || (VR[HY_DENS] < tiny && VR[HY_DENS] > 0.) ||
      (VL[HY_PRES] < tiny && VL[HY_PRES] > 0.) || (VR[HY_PRES] < tiny && VR[HY_PRES] > 0.)) {
    for (int n = 0; n < HY_NUM_FLUX; n++) F[n] = 0.;
    return 0;
  } else if (aL2 < 0. || aR2 < 0.) {
    return 1;
  }
  double cfL = sqrt(aL2);
  double cfR = sqrt(aR2);
  double SL = fmin2(velNL - cfL, velNR - cfR);
  double SR = fmax2(velNL + cfL, velNR + cfR);
  double totalPresL = VL[HY_PRES];
  double totalPresR = VR[HY_PRES];
  double UL[5], UR[5], FL[5], FR[5], UCsL[5], UCsR[5];
  orc_prim2con(VL, UL);
  orc_prim2con(VR, UR);
  orc_prim2flx(dir, VL, FL);
  orc_prim2flx(dir, VR, FR);

  double qStar = VR[HY_DENS] * velNR * (SR - velNR) - VL[HY_DENS] * velNL * (SL - velNL) +
                 totalPresL - totalPresR - magNL * magNL + magNR * magNR;
  qStar = qStar / (VR[HY_DENS] * (SR - velNR) - VL[HY_DENS] * (SL - velNL));
  double numerL = SL - velNL;
  double denomL = SL - qStar;
  double numerR = SR - velNR;
  double denomR = SR - qStar;
  double pStar = VL[HY_DENS] * numerL * (qStar - velNL) + totalPresL;
  double dStarL = UL[HY_MASS] * numerL / denomL;
  double dStarR = UR[HY_MASS] * numerR / denomR;
  UCsL[HY_MASS] = dStarL;
  UCsL[HY_ENER] = UL[HY_ENER] * numerL / denomL + ((pStar * qStar - totalPresL * velNL)) / denomL;
  UCsR[HY_MASS] = dStarR;
  UCsR[HY_ENER] = UR[HY_ENER] * numerR / denomR + ((pStar * qStar - totalPresR * velNR)) / denomR;
  for (int d = 0; d < 3; d++) {
    if (d == dir) {
      UCsL[HY_XMOM + d] = dStarL * qStar;
      UCsR[HY_XMOM + d] = dStarR * qStar;
    } else {
      UCsL[HY_XMOM + d] = UL[HY_XMOM + d] * numerL / denomL;
      UCsR[HY_XMOM + d] = UR[HY_XMOM + d] * numerR / denomR;
    }
  }
  if (SL >= 0.) {
    for (int n = 0; n < 5; n++) F[n] = FL[n];
  } else if ((SL < 0.) && (qStar >= 0.)) {
    for (int n = 0; n < 5; n++) F[n] = FL[n] + SL * (UCsL[n] - UL[n]);
  } else if ((qStar < 0.) && (SR >= 0.)) {
    for (int n = 0; n < 5; n++) F[n] = FR[n] + SR * (UCsR[n] - UR[n]);
  } else {
    for (int n = 0; n < 5; n++) F[n] = FR[n];
  }
  if (P->use_hybrid_riemann && inShock) { /* HLLE inside shocks, :266-274 */
    if (SL > 0.) {
      for (int n = 0; n < 5; n++) F[n] = FL[n];
    } else if (SR < 0.) {
      for (int n = 0; n < 5; n++) F[n] = FR[n];
    } else {
      for (int n = 0; n < 5; n++)
        F[n] = (SR * FL[n] - SL * FR[n] + SR * SL * (UR[n] - UL[n])) / (SR - SL);
    }
  }
  return 0;
}

/* HLL(E), Spark/tvd2/riemann.F90:63-141 */
int orc_riemann_hll(const orc_params* P, int dir, const double* VL, const double* VR, double* F) {
  const double tiny = 1.e-32;
  (void)P;
  for (int n = 0; n < 5; n++) F[n] = 0.0;
  double velNL = VL[HY_VELX + dir];
  double velNR = VR[HY_VELX + dir];
  double aL2 = VL[HY_GAMC] * VL[HY_PRES] / VL[HY_DENS];
  double aR2 = VR[HY_GAMC] * VR[HY_PRES] / VR[HY_DENS];
  double cfL, cfR;
  if ((VL[HY_DENS] < tiny && VL[HY_DENS] > 0.) || (VR[HY_DENS] < tiny && VR[HY_DENS] > 0.) ||
      (VL[HY_PRES] < tiny && VL[HY_PRES] > 0.) || (VR[HY_PRES] < tiny && VR[HY_PRES] > 0.)) {
    for (int n = 0; n < 5; n++) F[n] = 0.;
    return 0;
  } else if (aL2 < 0. || aR2 < 0.) {
    return 1;
  } else {
    cfL = sqrt(aL2);
    cfR = sqrt(aR2);
  }
  double SL = fmin2(velNL - cfL, velNR - cfR);
  double SR = fmax2(velNL + cfL, velNR + cfR);
  double UL[5], UR[5], FL[5], FR[5];
  orc_prim2con(VL, UL);
  orc_prim2con(VR, UR);
  orc_prim2flx(dir, VL, FL);
  orc_prim2flx(dir, VR, FR);
  if (SL > 0.) {
    for (int n = 0; n < 5; n++) F[n] = FL[n];
  } else if (SR < 0.) {
    for (int n = 0; n < 5; n++) F[n] = FR[n];
  } else {
    for (int n = 0; n < 5; n++)
      F[n] = (SR * FL[n] - SL * FR[n] + SR * SL * (UR[n] - UL[n])) / (SR - SL);
  }
  return 0;
}

/* ------------------------------------------------------------------ */
/* Reconstruction.  pen is laid out pen[v*plen + i].                   */

/* WENO5 with WENO-Z weights, Spark/reconstruct.F90-mc:310-394 and the
 * coefficients of [weno5_coeffs], Spark/Hydro.ini:202-220 */
static void recon_weno5(int nrecon, const double* pen, int plen, int i1, double flat,
                        double* uP, double* uM) {
  const double c1p1[3] = {2. / 6., -7. / 6., 11. / 6.};
  const double c1p2[3] = {-1. / 6., 5. / 6., 2. / 6.};
  const double c1p3[3] = {2. / 6., 5. / 6., -1. / 6.};
  const double c2p[3] = {0.1, 0.6, 0.3};
  const double c1m1[3] = {-1. / 6., 5. / 6., 2. / 6.};
  const double c1m2[3] = {2. / 6., 5. / 6., -1. / 6.};
  const double c1m3[3] = {11. / 6., -7. / 6., 2. / 6.};
  const double c2m[3] = {0.3, 0.6, 0.1};
  const double epsilon = 1.e-36;
  const double n13o12 = 13. / 12.;
  for (int v = 0; v < nrecon; v++) {
    const double* u = pen + (size_t)v * plen;
    double um2 = u[i1 - 2], um1 = u[i1 - 1], u0 = u[i1], up1 = u[i1 + 1], up2 = u[i1 + 2];
    double W5p1 = c1p1[0] * um2 + c1p1[1] * um1 + c1p1[2] * u0;
    double W5p2 = c1p2[0] * um1 + c1p2[1] * u0 + c1p2[2] * up1;
    double W5p3 = c1p3[0] * u0 + c1p3[1] * up1 + c1p3[2] * up2;
    double W5m1 = c1m1[0] * um2 + c1m1[1] * um1 + c1m1[2] * u0;
    double W5m2 = c1m2[0] * um1 + c1m2[1] * u0 + c1m2[2] * up1;
    double W5m3 = c1m3[0] * u0 + c1m3[1] * up1 + c1m3[2] * up2;
    double t, s;
    t = um2 - 2. * um1 + u0;
    s = um2 - 4. * um1 + 3. * u0;
    double b1 = n13o12 * (t * t) + 0.25 * (s * s);
    t = um1 - 2. * u0 + up1;
    s = um1 - up1;
    double b2 = n13o12 * (t * t) + 0.25 * (s * s);
    t = u0 - 2. * up1 + up2;
    s = 3. * u0 - 4. * up1 + up2;
    double b3 = n13o12 * (t * t) + 0.25 * (s * s);
    double tau = fabs(b1 - b3);
    double r1 = tau / (b1 + epsilon), r2 = tau / (b2 + epsilon), r3 = tau / (b3 + epsilon);
    double A1 = c2p[0] * (1. + r1), A2 = c2p[1] * (1. + r2), A3 = c2p[2] * (1. + r3);
    double inv = 1. / (A1 + A2 + A3);
    double w1 = A1 * inv, w2 = A2 * inv, w3 = A3 * inv;
    double up = w1 * W5p1 + w2 * W5p2 + w3 * W5p3;
    up = flat * up + (1. - flat) * u0;
    A1 = c2m[0] * (1. + r1); A2 = c2m[1] * (1. + r2); A3 = c2m[2] * (1. + r3);
    inv = 1. / (A1 + A2 + A3);
    w1 = A1 * inv; w2 = A2 * inv; w3 = A3 * inv;
    double um = w1 * W5m1 + w2 * W5m2 + w3 * W5m3;
    um = flat * um + (1. - flat) * u0;
    if ((up - u0) * (u0 - um) <= 0.) { up = u0; um = u0; }
    uP[v] = up;
    uM[v] = um;
  }
}

/* minmod, Spark/hy_rk_getFaceFlux.F90-mc:285-289 and mp5/reconstruct.F90:90-94 */
static inline double minmod(double a, double b) {
  return .5 * (copysign(1., a) + copysign(1., b)) * fmin2(fabs(a), fabs(b));
}
/* median, mp5/reconstruct.F90:96-100 */
static inline double median3(double a, double b, double c) { return a + minmod(b - a, c - a); }

/* MP5 (Suresh & Huynh), Spark/mp5/reconstruct.F90:28-88.  The reference
 * computes d2mp/d2mm only inside the i+1/2 limiter branch and then reads
 * them in the i-1/2 branch; here they are computed before both branches
 * (the published algorithm), so both branches see this cell's values. */
static void recon_mp5(int nrecon, const double* pen, int plen, int i, double* uP, double* uM) {
  const double B1 = 1. / 60., B2 = 4. / 3., alpha = 4., eps = 0.;
  for (int g = 0; g < nrecon; g++) {
    const double* U = pen + (size_t)g * plen;
    double Uplus = B1 * (2. * U[i - 2] - 13. * U[i - 1] + 47. * U[i] + 27. * U[i + 1] - 3. * U[i + 2]);
    double Uminus = B1 * (-3. * U[i - 2] + 27. * U[i - 1] + 47. * U[i] - 13. * U[i + 1] + 2. * U[i + 2]);
    double del_p = U[i + 1] - U[i];
    double del_m = U[i] - U[i - 1];
    double d2m = U[i - 2] - 2. * U[i - 1] + U[i];
    double d2 = U[i - 1] - 2. * U[i] + U[i + 1];
    double d2p = U[i] - 2. * U[i + 1] + U[i + 2];
    double d2mp = minmod(minmod(4 * d2 - d2p, 4 * d2p - d2), minmod(d2, d2p));
    double d2mm = minmod(minmod(4 * d2m - d2, 4 * d2 - d2m), minmod(d2m, d2));
    double ump = U[i] + minmod(del_p, alpha * del_m);
    if ((Uplus - U[i]) * (Uplus - ump) >= eps) {
      double uul = U[i] + alpha * del_m;
      double umd = 0.5 * (U[i] + U[i + 1]) - 0.5 * d2mp;
      double ulc = U[i] + 0.5 * del_m + B2 * d2mm;
      double umin = fmax2(fmin2(fmin2(U[i], U[i + 1]), umd), fmin2(fmin2(U[i], uul), ulc));
      double umax = fmin2(fmax2(fmax2(U[i], U[i + 1]), umd), fmax2(fmax2(U[i], uul), ulc));
      Uplus = median3(umin, Uplus, umax);
    }
    ump = U[i] - minmod(del_m, alpha * del_p);
    if ((U[i] - Uminus) * (ump - Uminus) >= eps) {
      double uul = U[i] - alpha * del_p;
      double umd = 0.5 * (U[i] + U[i - 1]) - 0.5 * d2mm;
      double ulc = U[i] - 0.5 * del_p + B2 * d2mp;
      double umin = fmax2(fmin2(fmin2(U[i], U[i - 1]), umd), fmin2(fmin2(U[i], uul), ulc));
      double umax = fmin2(fmax2(fmax2(U[i], U[i - 1]), umd), fmax2(fmax2(U[i], uul), ulc));
      Uminus = median3(umin, Uminus, umax);
    }
    uP[g] = Uplus;
    uM[g] = Uminus;
  }
}

/* PLM with minmod slopes, Spark/tvd2/reconstruct.F90:28-54 */
static void recon_plm(int nrecon, const double* pen, int plen, int ind, double* uP, double* uM) {
  for (int g = 0; g < nrecon; g++) {
    const double* d = pen + (size_t)g * plen;
    double del_p = d[ind + 1] - d[ind];
    double del_m = d[ind] - d[ind - 1];
    double delbar = minmod(del_p, del_m);
    uP[g] = d[ind] + 0.5 * delbar;
    uM[g] = d[ind] - 0.5 * delbar;
  }
}

/* first-order Godunov, Spark/fog/reconstruct.F90:28-45 */
static void recon_fog(int nrecon, const double* pen, int plen, int ind, double* uP, double* uM) {
  for (int g = 0; g < nrecon; g++) {
    uP[g] = pen[(size_t)g * plen + ind];
    uM[g] = pen[(size_t)g * plen + ind];
  }
}

void orc_recon(const orc_params* P, int nrecon, const double* pen, int i1, int plen, double flat,
               double* uP, double* uM) {
  switch (P->recon) {
    case ORC_RECON_MP5: recon_mp5(nrecon, pen, plen, i1, uP, uM); break;
    case ORC_RECON_PLM: recon_plm(nrecon, pen, plen, i1, uP, uM); break;
    case ORC_RECON_FOG: recon_fog(nrecon, pen, plen, i1, uP, uM); break;
    default: recon_weno5(nrecon, pen, plen, i1, flat, uP, uM); break;
  }
}

/* ------------------------------------------------------------------ */
/* flattening (Miller & Colella), Spark/hy_rk_getFaceFlux.F90-mc:200-276 */
void orc_flattening(const orc_params* P, const orc_block* b, const double* S, const int lim[3][2],
                    double* flat3d) {
  int NX, NY, NZ;
  orc_gc_dims(P, b, &NX, &NY, &NZ);
  const int nunk = orc_nunk(P);
  const double betaMin = 0.75, betaMax = 0.85, Zmin = 0.25, Zmax = 0.75;
  int kx = 1, ky = (P->ndim > 1) ? 1 : 0, kz = (P->ndim == 3) ? 1 : 0;
  int str[3] = {1, NX, NX * NY};
  const int velv[3] = {ORC_VELX, ORC_VELY, ORC_VELZ};
#define SV(var, c) S[(size_t)(var) + (size_t)nunk * (c)]
  for (int k = lim[2][0] - kz; k <= lim[2][1] + kz; k++)
    for (int j = lim[1][0] - ky; j <= lim[1][1] + ky; j++)
      for (int i = lim[0][0] - kx; i <= lim[0][1] + kx; i++) {
        size_t c = cidx(NX, NY, i, j, k);
        double ft[3] = {1.0, 1.0, 1.0};
        for (int d = 0; d < P->ndim; d++) {
          size_t s = (size_t)str[d];
          double beta = fabs(SV(ORC_PRES, c + s) - SV(ORC_PRES, c - s)) /
                        fmax2(1e-99, fabs(SV(ORC_PRES, c + 2 * s) - SV(ORC_PRES, c - 2 * s)));
          double Z = fabs(SV(ORC_PRES, c + s) - SV(ORC_PRES, c - s)) /
                     (SV(ORC_GAMC, c) * SV(ORC_PRES, c));
          ft[d] = fmax2(0., fmin2(1., (betaMax - beta) / (betaMax - betaMin)));
          if (SV(velv[d], c + s) < SV(velv[d], c)) {
            ft[d] = fmax2(ft[d], fmin2(1., (Zmax - Z) / (Zmax - Zmin)));
          } else {
            ft[d] = 1.0;
          }
        }
        double m = ft[0];
        for (int d = 1; d < P->ndim; d++) m = fmin2(m, ft[d]);
        flat3d[c] = m;
      }
#undef SV
}

/* ------------------------------------------------------------------ */
/* hy_rk_getFaceFlux body (Spark/hy_rk_getFaceFlux.F90-mc:97-183) with the
 * Inhost pencil macros (Spark/Inhost/hy_local.ini:83-127,163-199) */
void orc_get_face_flux(const orc_params* P, const orc_block* b, const double* S,
                       const double* flat3d, const int lim[3][2], double* flx, double* fly,
                       double* flz, int* nerr) {
  int NX, NY, NZ;
  orc_gc_dims(P, b, &NX, &NY, &NZ);
  const int nunk = orc_nunk(P), nrec = HY_NUM_VARS + P->nscalars, nf = orc_nflux(P);
  const int ns = P->nscalars;
  int N[3] = {NX, NY, NZ};
  int maxlen = NX > NY ? NX : NY;
  if (NZ > maxlen) maxlen = NZ;
  double* pen = (double*)malloc(sizeof(double) * (size_t)nrec * maxlen);
  double* pflat = (double*)malloc(sizeof(double) * maxlen);
  double* pshck = (double*)malloc(sizeof(double) * maxlen);
  double* fl[3] = {flx, fly, flz};
  double uP[MAXREC], uM[MAXREC], left[MAXREC], right[MAXREC], F[MAXFLX];
  for (int dir = 0; dir < P->ndim; dir++) {
    int d1 = dir, d2 = (dir == 0) ? 1 : 0, d3 = (dir == 2) ? 1 : 2;
    int plen = N[d1];
    for (int i3 = lim[d3][0]; i3 <= lim[d3][1]; i3++)
      for (int i2 = lim[d2][0]; i2 <= lim[d2][1]; i2++) {
        /* setSnake over the whole pencil, Spark/Hydro.ini:166-199 */
        for (int i = 0; i < plen; i++) {
          int ijk[3];
          ijk[d1] = i; ijk[d2] = i2; ijk[d3] = i3;
          size_t c = cidx(NX, NY, ijk[0], ijk[1], ijk[2]);
          const double* v = S + (size_t)nunk * c;
          pen[(size_t)HY_DENS * plen + i] = v[ORC_DENS];
          pen[(size_t)HY_VELX * plen + i] = v[ORC_VELX];
          pen[(size_t)HY_VELY * plen + i] = v[ORC_VELY];
          pen[(size_t)HY_VELZ * plen + i] = v[ORC_VELZ];
          pen[(size_t)HY_PRES * plen + i] = v[ORC_PRES];
          pen[(size_t)HY_GAMC * plen + i] = v[ORC_GAMC];
          pen[(size_t)HY_RHOE * plen + i] = v[ORC_DENS] * v[ORC_EINT];
          for (int s = 0; s < ns; s++) pen[(size_t)(HY_NUM_VARS + s) * plen + i] = v[ORC_NPROP + s];
          pshck[i] = v[ORC_SHOK];
          pflat[i] = flat3d ? flat3d[c] : 1.0;
        }
        int lo = lim[d1][0], hi = lim[d1][1];
        orc_recon(P, nrec, pen, lo - 1, plen, pflat[lo - 1], uP, uM);
        for (int i1 = lo; i1 <= hi + 1; i1++) {
          for (int v = 0; v < nrec; v++) left[v] = uP[v];
          orc_recon(P, nrec, pen, i1, plen, pflat[i1], uP, uM);
          for (int v = 0; v < nrec; v++) right[v] = uM[v];
          /* hy_ensurePhysicalState, Spark/Hydro.ini:222-235 */
          left[HY_DENS] = fmax2(P->smlrho, left[HY_DENS]);
          left[HY_PRES] = fmax2(P->smallp, left[HY_PRES]);
          right[HY_DENS] = fmax2(P->smlrho, right[HY_DENS]);
          right[HY_PRES] = fmax2(P->smallp, right[HY_PRES]);
          if (P->nspecies > 0) {
            /* limit and renormalise the species (:227-235) */
            double* st[2] = {left, right};
            for (int q = 0; q < 2; q++) {
              double* spc = st[q] + HY_NUM_VARS;
              double sum = 0.0;
              for (int s = 0; s < P->nspecies; s++) spc[s] = fmax2(P->smallx, fmin2(1.0, spc[s]));
              for (int s = 0; s < P->nspecies; s++) sum = sum + spc[s];
              const double spcSumInv = 1. / sum;
              for (int s = 0; s < P->nspecies; s++) spc[s] = spc[s] * spcSumInv;
            }
          }
          int inShock = (pshck[i1 - 1] != 0.0) || (pshck[i1] != 0.0);
          int ierr;
          if (P->riemann == ORC_RIEMANN_HLL) ierr = orc_riemann_hll(P, dir, left, right, F);
          else ierr = orc_riemann_hllc(P, dir, left, right, inShock, F);
          if (ierr) {
            for (int n = 0; n < HY_NUM_FLUX; n++) F[n] = 0.0;
            if (nerr) (*nerr)++;
          }
          /* hy_avisc, Spark/Hydro.ini:285-301 */
          {
            const double* pd = pen + (size_t)HY_DENS * plen;
            const double* px = pen + (size_t)HY_VELX * plen;
            const double* py = pen + (size_t)HY_VELY * plen;
            const double* pz = pen + (size_t)HY_VELZ * plen;
            const double* pe = pen + (size_t)HY_RHOE * plen;
            const double* pn = pen + (size_t)(HY_VELX + dir) * plen;
            double cvisc = P->cvisc * fmax2(-(pn[i1] - pn[i1 - 1]), 0.);
            double VenerLo = pd[i1 - 1] * 0.5 *
                                 (px[i1 - 1] * px[i1 - 1] + py[i1 - 1] * py[i1 - 1] + pz[i1 - 1] * pz[i1 - 1]) +
                             pe[i1 - 1];
            double VenerHi = pd[i1] * 0.5 * (px[i1] * px[i1] + py[i1] * py[i1] + pz[i1] * pz[i1]) + pe[i1];
            F[HY_MASS] = F[HY_MASS] + cvisc * (pd[i1 - 1] - pd[i1]);
            F[HY_XMOM] = F[HY_XMOM] + cvisc * (pd[i1 - 1] * px[i1 - 1] - pd[i1] * px[i1]);
            F[HY_YMOM] = F[HY_YMOM] + cvisc * (pd[i1 - 1] * py[i1 - 1] - pd[i1] * py[i1]);
            F[HY_ZMOM] = F[HY_ZMOM] + cvisc * (pd[i1 - 1] * pz[i1 - 1] - pd[i1] * pz[i1]);
            F[HY_ENER] = F[HY_ENER] + cvisc * (VenerLo - VenerHi);
          }
          /* hy_mscalarFluxes, Spark/Hydro.ini:238-247 */
          for (int s = 0; s < ns; s++) {
            if (F[HY_MASS] > 0.) F[HY_NUM_FLUX + s] = left[HY_NUM_VARS + s] * F[HY_MASS];
            else F[HY_NUM_FLUX + s] = right[HY_NUM_VARS + s] * F[HY_MASS];
          }
          /* hy_saveFluxes: face i1 is the low face of cell i1 */
          int ijk[3];
          ijk[d1] = i1; ijk[d2] = i2; ijk[d3] = i3;
          double* out = fl[dir] + (size_t)nf * fidx(NX, NY, ijk[0], ijk[1], ijk[2]);
          for (int n = 0; n < nf; n++) out[n] = F[n];
        }
      }
  }
  free(pen);
  free(pflat);
  free(pshck);
}

/* ------------------------------------------------------------------ */
/* EOS (Gamma, or Multigamma over the species), MODE_DENS_EI, on one cell of an unk-layout state:
 * Eos_getData.F90:114-169 -> Gamma/Eos.F90:238-247 -> Eos_putData.F90:207-224.
 * full=0 skips the TEMP/GAME outputs (used when only the stage chain is
 * needed; both are pure outputs that nothing on the path reads). */
void orc_eos_dens_ei_cell(const orc_params* P, double* v, int full) {
  double ekin = 0.5 * (v[ORC_VELX] * v[ORC_VELX] + v[ORC_VELY] * v[ORC_VELY] + v[ORC_VELZ] * v[ORC_VELZ]);
  double eint = v[ORC_EINT];
  if (v[ORC_ENER] - ekin > fmax2(P->smalle, P->eintSwitch * ekin)) eint = v[ORC_ENER] - ekin;
  eint = fmax2(eint, P->smalle);
  double dens = v[ORC_DENS];
  double pres, temp, gamc;
  if (P->eos == 1) {
    /* Multigamma/Eos.F90 (ORIGINAL_GAMC_AVERAGE): abar from
     * Multispecies_getSumInv(A, X) (Multispecies_getSumInv.F90:143-152),
     * rt from the weights X/(gamma-1) (eos_initMgamma.F90:15), then
     * MODE_DENS_EI */
    const double* X = v + ORC_NPROP;
    double abarInv = 0.0, rt = 0.0;
    for (int s = 0; s < P->nspecies; s++) abarInv = abarInv + X[s] * (1.e0 / P->species_a[s]);
    const double abar = 1.e0 / abarInv;
    for (int s = 0; s < P->nspecies; s++) {
      const double gammam1j = 1. / (P->species_gamma[s] - 1.);
      const double weight = X[s] * gammam1j;
      rt = rt + weight * (1.e0 / P->species_a[s]);
    }
    gamc = 1.0e0 + 1.0e0 / (rt * abar);
    pres = dens * eint * (gamc - 1.e0);
    temp = eint * (gamc - 1.e0) * abar / P->gasConst;
  } else {
    double eos_gammam1 = 1.0 / (P->gamma - 1.0);  /* Gamma/Eos_init.F90:86 */
    double ggprod = eos_gammam1 * P->gasConst;
    double ggprodinv = 1. / ggprod;
    double gam1inv = 1. / eos_gammam1;
    gamc = P->gamma;
    pres = dens * eint * gam1inv;
    temp = eint * ggprodinv * P->abar;
  }
  v[ORC_PRES] = pres;
  v[ORC_GAMC] = gamc;
  v[ORC_EINT] = eint;
  v[ORC_ENER] = eint + ekin;
  if (full) {
    v[ORC_TEMP] = temp;
    v[ORC_GAME] = pres / (eint * dens) + 1;
  }
}

void orc_eos_wrapped(const orc_params* P, const orc_block* b, double* S, const int lim[3][2]) {
  int NX, NY, NZ;
  orc_gc_dims(P, b, &NX, &NY, &NZ);
  const int nunk = orc_nunk(P);
  for (int k = lim[2][0]; k <= lim[2][1]; k++)
    for (int j = lim[1][0]; j <= lim[1][1]; j++)
      for (int i = lim[0][0]; i <= lim[0][1]; i++)
        orc_eos_dens_ei_cell(P, S + (size_t)nunk * cidx(NX, NY, i, j, k), 1);
}

/* ------------------------------------------------------------------ */
/* update_solution, Spark/hy_rk_updateSoln.F90-mc:194-304 (Cartesian),
 * gravSources :179-190 with a constant acceleration. */
/* geoFacs (hy_rk_updateSoln.F90-mc:98-175, hy_rk_correctFluxes.F90:341-358)
 * for cylindrical geometry: facM, facP = farea * dx / cvol of the x faces
 * of cell i; 0 and no sources where the cell centre is at negative radius.
 * Returns 1 where the sources apply. */
static int geo_facs(const orc_block* b, int NX, int NY, int i, int j, double dx, double* facM, double* facP,
                    double* rc) {
  *facM = 1.0;
  *facP = 1.0;
  *rc = 0.0;
  if (!b->geo) return 0;
  const double* area = b->geo + 3 * NX;
  const double* vol = area + (size_t)NX * NY;
  const size_t q = (size_t)j * NX + i;
  *rc = b->geo[i];
  *facM = area[q] * dx / vol[q];
  *facP = area[q + 1] * dx / vol[q];
  if (*rc < 0.0) {
    *facM = 0.;
    *facP = 0.;
    return 0;
  }
  return 1;
}

void orc_update_soln(const orc_params* P, const orc_block* b, const double* U0v, double* S,
                     const int lim[3][2], double dt, const double coeffs[3], const double* flx,
                     const double* fly, const double* flz) {
  int NX, NY, NZ;
  orc_gc_dims(P, b, &NX, &NY, &NZ);
  const int nunk = orc_nunk(P), nf = orc_nflux(P), ns = P->nscalars;
  const double dx = b->dx, dy = b->dy, dz = b->dz;
  for (int k = lim[2][0]; k <= lim[2][1]; k++)
    for (int j = lim[1][0]; j <= lim[1][1]; j++)
      for (int i = lim[0][0]; i <= lim[0][1]; i++) {
        const double* V0 = U0v + (size_t)nunk * cidx(NX, NY, i, j, k);
        double* Vs = S + (size_t)nunk * cidx(NX, NY, i, j, k);
        const double* Fm = flx + (size_t)nf * fidx(NX, NY, i, j, k);
        const double* Fp = flx + (size_t)nf * fidx(NX, NY, i + 1, j, k);
        const double *Gm = NULL, *Gp = NULL, *Hm = NULL, *Hp = NULL;
        if (P->ndim > 1) {
          Gm = fly + (size_t)nf * fidx(NX, NY, i, j, k);
          Gp = fly + (size_t)nf * fidx(NX, NY, i, j + 1, k);
        }
        if (P->ndim == 3) {
          Hm = flz + (size_t)nf * fidx(NX, NY, i, j, k);
          Hp = flz + (size_t)nf * fidx(NX, NY, i, j, k + 1);
        }
        double U0[MAXFLX], Us[MAXFLX], Sg[MAXFLX], St[MAXFLX];
        U0[HY_MASS] = V0[ORC_DENS];
        U0[HY_XMOM] = V0[ORC_DENS] * V0[ORC_VELX];
        U0[HY_YMOM] = V0[ORC_DENS] * V0[ORC_VELY];
        U0[HY_ZMOM] = V0[ORC_DENS] * V0[ORC_VELZ];
        U0[HY_ENER] = V0[ORC_DENS] * V0[ORC_ENER];
        for (int s = 0; s < ns; s++) U0[HY_NUM_FLUX + s] = V0[ORC_NPROP + s] * V0[ORC_DENS];
        Us[HY_MASS] = Vs[ORC_DENS];
        Us[HY_XMOM] = Vs[ORC_DENS] * Vs[ORC_VELX];
        Us[HY_YMOM] = Vs[ORC_DENS] * Vs[ORC_VELY];
        Us[HY_ZMOM] = Vs[ORC_DENS] * Vs[ORC_VELZ];
        Us[HY_ENER] = Vs[ORC_DENS] * Vs[ORC_ENER];
        for (int s = 0; s < ns; s++) Us[HY_NUM_FLUX + s] = Vs[ORC_NPROP + s] * Vs[ORC_DENS];
        for (int n = 0; n < nf; n++) Sg[n] = 0.;
        if (P->use_gravity) {
          Sg[HY_XMOM] = Us[HY_MASS] * P->grav[0];
          Sg[HY_YMOM] = Us[HY_MASS] * P->grav[1];
          Sg[HY_ZMOM] = Us[HY_MASS] * P->grav[2];
          Sg[HY_ENER] = Us[HY_XMOM] * P->grav[0] + Us[HY_YMOM] * P->grav[1] + Us[HY_ZMOM] * P->grav[2];
        }
        /* Stot = Sgeo + Sgrav; cylindrical: alpha = 1, the azimuthal
         * momentum is the z one, V the star state (geoFacs :131-157) */
        double Sgeo[MAXFLX], facM, facP, rc;
        for (int n = 0; n < nf; n++) Sgeo[n] = 0.;
        if (geo_facs(b, NX, NY, i, j, dx, &facM, &facP, &rc)) {
          const int sph = b->geometry == 2;
          const double alpha = sph ? 2. : 1.;
          Sgeo[HY_XMOM] = (Vs[ORC_DENS] * Vs[ORC_VELZ] * Vs[ORC_VELZ] + alpha * Vs[ORC_PRES]) / rc;
          Sgeo[HY_ZMOM] = Vs[ORC_DENS] * Vs[ORC_VELZ] * Vs[ORC_VELX] / rc;
          Sgeo[HY_ZMOM] = -Sgeo[HY_ZMOM];
          if (sph) {
            /* the polar momentum is the y one; dx_sph (:126) */
            const double rl = b->geo[NX + i], rr = b->geo[2 * NX + i];
            const double dx_sph = (rr * rr * rr - rl * rl * rl) / (3. * (rc * rc));
            Sgeo[HY_XMOM] = Sgeo[HY_XMOM] + Us[HY_YMOM] * Us[HY_YMOM] / Vs[ORC_DENS] / rc;
            for (int n = 0; n < nf; n++) Sgeo[n] = Sgeo[n] * dx / dx_sph;
          }
        }
        for (int n = 0; n < nf; n++) St[n] = Sgeo[n] + Sg[n];
        for (int n = 0; n < nf; n++) {
          double inner = -(dt / dx * (facP * Fp[n] - facM * Fm[n]));
          if (P->ndim > 1) inner = inner - dt / dy * (Gp[n] - Gm[n]);
          if (P->ndim == 3) inner = inner - dt / dz * (Hp[n] - Hm[n]);
          inner = inner + dt * St[n];
          Us[n] = coeffs[0] * U0[n] + coeffs[1] * Us[n] + coeffs[2] * inner;
        }
        Vs[ORC_DENS] = fmax2(Us[HY_MASS], P->smlrho);
        Vs[ORC_VELX] = Us[HY_XMOM] / Vs[ORC_DENS];
        Vs[ORC_VELY] = Us[HY_YMOM] / Vs[ORC_DENS];
        Vs[ORC_VELZ] = Us[HY_ZMOM] / Vs[ORC_DENS];
        Vs[ORC_ENER] = fmax2(P->smalle, Us[HY_ENER] / Vs[ORC_DENS]);
        double ekin = .5 * (Vs[ORC_VELX] * Vs[ORC_VELX] + Vs[ORC_VELY] * Vs[ORC_VELY] + Vs[ORC_VELZ] * Vs[ORC_VELZ]);
        Vs[ORC_EINT] = fmax2(P->smalle, Vs[ORC_ENER] - ekin);
        for (int s = 0; s < ns; s++) Vs[ORC_NPROP + s] = Us[HY_NUM_FLUX + s] / Vs[ORC_DENS];
      }
}

/* ------------------------------------------------------------------ */
/* shockDetect, Spark/Hydro_funcs.F90-mc:600-706, over limits = interior
 * +/- (NGUARD-1); writes SHOK into U (whole block zeroed first). */
void orc_shock_detect(const orc_params* P, const orc_block* b, double* U) {
  int NX, NY, NZ;
  orc_gc_dims(P, b, &NX, &NY, &NZ);
  const int nunk = orc_nunk(P);
  size_t ncell = (size_t)NX * NY * NZ;
  const double beta = 0.5, delta = 0.1;
  double* Vc = (double*)malloc(sizeof(double) * ncell);
  for (size_t c = 0; c < ncell; c++) U[(size_t)ORC_SHOK + nunk * c] = 0.;
  for (size_t c = 0; c < ncell; c++) {
    const double* v = U + (size_t)nunk * c;
    Vc[c] = sqrt(v[ORC_GAMC] * v[ORC_PRES] / fmax2(v[ORC_DENS], P->tiny));
  }
  int lim[3][2];
  set_lims(P, b, P->nguard - 1, lim);
  int ky = (P->ndim > 1), kz = (P->ndim == 3);
  size_t sx = 1, sy = (size_t)NX, sz = (size_t)NX * NY;
#define UV(var, c) U[(size_t)(var) + (size_t)nunk * (c)]
  for (int k = lim[2][0]; k <= lim[2][1]; k++)
    for (int j = lim[1][0]; j <= lim[1][1]; j++)
      for (int i = lim[0][0]; i <= lim[0][1]; i++) {
        size_t c = cidx(NX, NY, i, j, k);
        double minP = DBL_MAX, minC = DBL_MAX;
        int first = 1;
        for (int kk = -kz; kk <= kz; kk++)
          for (int jj = -ky; jj <= ky; jj++)
            for (int ii = -1; ii <= 1; ii++) {
              size_t cc = cidx(NX, NY, i + ii, j + jj, k + kk);
              if (first) { minP = UV(ORC_PRES, cc); minC = Vc[cc]; first = 0; }
              else { minP = fmin2(minP, UV(ORC_PRES, cc)); minC = fmin2(minC, Vc[cc]); }
            }
        double divv = UV(ORC_VELX, c + sx) - UV(ORC_VELX, c - sx);
        if (P->ndim > 1) divv = divv + UV(ORC_VELY, c + sy) - UV(ORC_VELY, c - sy);
        if (P->ndim == 3) divv = divv + UV(ORC_VELZ, c + sz) - UV(ORC_VELZ, c - sz);
        divv = 0.5 * divv;
        double gradPx = 0.5 * (UV(ORC_PRES, c + sx) - UV(ORC_PRES, c - sx));
        double gradPy = 0., gradPz = 0.;
        if (P->ndim > 1) gradPy = 0.5 * (UV(ORC_PRES, c + sy) - UV(ORC_PRES, c - sy));
        if (P->ndim == 3) gradPz = 0.5 * (UV(ORC_PRES, c + sz) - UV(ORC_PRES, c - sz));
        int SW1 = (fabs(gradPx) + fabs(gradPy) + fabs(gradPz) >= beta * minP);
        int SW2 = (-delta * minC >= divv);
        if (SW1 && SW2) UV(ORC_SHOK, c) = 1.;
      }
#undef UV
  free(Vc);
}

/* ------------------------------------------------------------------ */
/* Hydro_computeDt, Spark/Hydro_computeDt.F90:27-150 (Cartesian, ugrid=0).
 * minloc gets the 0-based GC-box index of the limiting cell. */
double orc_compute_dt(const orc_params* P, const orc_block* b, const double* U, int minloc[3]) {
  int NX, NY, NZ;
  orc_gc_dims(P, b, &NX, &NY, &NZ);
  const int nunk = orc_nunk(P);
  int lim[3][2];
  interior_limits(P, b, lim);
  double delxinv = 1.0 / b->dx, delyinv = 1.0 / b->dy, delzinv = 1.0 / b->dz;
  double dt_temp = 0.;
  int loc[3] = {0, 0, 0};
  for (int k = lim[2][0]; k <= lim[2][1]; k++)
    for (int j = lim[1][0]; j <= lim[1][1]; j++)
      for (int i = lim[0][0]; i <= lim[0][1]; i++) {
        const double* v = U + (size_t)nunk * cidx(NX, NY, i, j, k);
        double snd2 = v[ORC_GAMC] * v[ORC_PRES] / v[ORC_DENS];
        double lt = (fabs(v[ORC_VELX]) + sqrt(snd2)) * delxinv;
        /* spherical: y is theta, delyinv = 1/(x(i) dy) (Hydro_computeDt.F90:108-110) */
        if (b->geo && b->geometry == 2 && P->ndim > 1) delyinv = 1.0 / (b->geo[i] * b->dy);
        if (P->ndim > 1) lt = fmax2(lt, (fabs(v[ORC_VELY]) + sqrt(snd2)) * delyinv);
        if (P->ndim > 2) lt = fmax2(lt, (fabs(v[ORC_VELZ]) + sqrt(snd2)) * delzinv);
        if (lt * P->cfl > dt_temp * P->cfl) {
          dt_temp = lt;
          loc[0] = i; loc[1] = j; loc[2] = k;
        }
      }
  if (minloc) { minloc[0] = loc[0]; minloc[1] = loc[1]; minloc[2] = loc[2]; }
  if (dt_temp != 0.0) return P->cfl / dt_temp;
  return DBL_MAX;
}

/* ------------------------------------------------------------------ */
/* RK tables, Spark/Hydro.ini:2-48.  The shipped [RK3] block uses
 * onethird/twothirds/onesixth without defining them; they are 1/3, 2/3, 1/6. */
void orc_rk_tables(const orc_params* P, int* last_stage, double coeff[3][3], int limits_array[3],
                   double weights[3]) {
  const int ns = P->nstencil;
  if (P->rk_order == 3) {
    const double onethird = 1. / 3., twothirds = 2. / 3., onesixth = 1. / 6.;
    double c[3][3] = {{1., 0., 1.}, {0.75, 0.25, 0.25}, {onethird, twothirds, twothirds}};
    memcpy(coeff, c, sizeof(c));
    *last_stage = 3;
    limits_array[0] = 2 * ns; limits_array[1] = ns; limits_array[2] = 0;
    weights[0] = onesixth; weights[1] = onesixth; weights[2] = twothirds;
  } else {
    double c[3][3] = {{1., 0., 1.}, {0.5, 0.5, 0.5}, {0., 0., 0.}};
    memcpy(coeff, c, sizeof(c));
    *last_stage = 2;
    limits_array[0] = ns; limits_array[1] = 0; limits_array[2] = 0;
    weights[0] = 0.5; weights[1] = 0.5; weights[2] = 0.;
  }
}

/* addFluxes, Spark/Hydro_funcs.F90-mc:481-507.  Buffers cover the interior
 * faces: X (nx+1, ny, nz), Y (nx, ny+1, nz), Z (nx, ny, nz+1). */
static void add_fluxes(const orc_params* P, const orc_block* b, const double* fl[3], double* fb[3],
                       double weight, int addFlux) {
  int NX, NY, NZ;
  orc_gc_dims(P, b, &NX, &NY, &NZ);
  const int nf = orc_nflux(P);
  int lim[3][2];
  interior_limits(P, b, lim);
  int n[3] = {lim[0][1] - lim[0][0] + 1, lim[1][1] - lim[1][0] + 1, lim[2][1] - lim[2][0] + 1};
  for (int dir = 0; dir < P->ndim; dir++) {
    int e[3] = {n[0], n[1], n[2]};
    e[dir] += 1;
    for (int k = 0; k < e[2]; k++)
      for (int j = 0; j < e[1]; j++)
        for (int i = 0; i < e[0]; i++) {
          const double* f = fl[dir] + (size_t)nf * fidx(NX, NY, lim[0][0] + i, lim[1][0] + j, lim[2][0] + k);
          double* o = fb[dir] + (size_t)nf * ((size_t)i + (size_t)e[0] * ((size_t)j + (size_t)e[1] * k));
          for (int v = 0; v < nf; v++) {
            if (addFlux) o[v] = o[v] + weight * f[v];
            else o[v] = weight * f[v];
          }
        }
  }
}

void orc_stage(const orc_params* P, const orc_block* b, const double* U0, double* S, int stage,
               int ngoff, double dt, double* fbx, double* fby, double* fbz, int* nerr) {
  int NX, NY, NZ;
  orc_gc_dims(P, b, &NX, &NY, &NZ);
  const int nf = orc_nflux(P);
  int last;
  double coeff[3][3], weights[3];
  int la[3];
  orc_rk_tables(P, &last, coeff, la, weights);
  size_t ncell = (size_t)NX * NY * NZ;
  size_t nface = (size_t)(NX + 1) * (NY + 1) * (NZ + 1);
  double* flat = (double*)malloc(sizeof(double) * ncell);
  double* fl[3];
  for (int d = 0; d < 3; d++) fl[d] = (double*)calloc((size_t)nf * nface, sizeof(double));
  int lim[3][2];
  set_lims(P, b, ngoff, lim);
  if (P->use_flattening && P->recon == ORC_RECON_WENO5) {
    orc_flattening(P, b, S, lim, flat);
  } else {
    for (size_t c = 0; c < ncell; c++) flat[c] = 1.0;
  }
  orc_get_face_flux(P, b, S, flat, lim, fl[0], fl[1], fl[2], nerr);
  if (fbx) {
    double* fb[3] = {fbx, fby, fbz};
    const double* flc[3] = {fl[0], fl[1], fl[2]};
    add_fluxes(P, b, flc, fb, weights[stage - 1], stage > 1);
  }
  orc_update_soln(P, b, U0, S, lim, dt, coeff[stage - 1], fl[0], fl[1], fl[2]);
  orc_eos_wrapped(P, b, S, lim);
  free(flat);
  for (int d = 0; d < 3; d++) free(fl[d]);
}

int orc_hydro_block(const orc_params* P, const orc_block* b, double* U, double dt, double* fbx,
                    double* fby, double* fbz) {
  int NX, NY, NZ;
  orc_gc_dims(P, b, &NX, &NY, &NZ);
  const int nunk = orc_nunk(P);
  size_t n = (size_t)nunk * NX * NY * NZ;
  int last, la[3];
  double coeff[3][3], weights[3];
  orc_rk_tables(P, &last, coeff, la, weights);
  orc_shock_detect(P, b, U);
  double* S = (double*)malloc(sizeof(double) * n);
  double* U0 = (double*)malloc(sizeof(double) * n);
  memcpy(S, U, sizeof(double) * n);  /* saveState, Hydro_funcs.F90-mc:512-541 */
  memcpy(U0, U, sizeof(double) * n);
  int nerr = 0;
  for (int stage = 1; stage <= last; stage++)
    orc_stage(P, b, U0, S, stage, la[stage - 1], dt, fbx, fby, fbz, &nerr);
  /* updateState, Hydro_funcs.F90-mc:544-581 */
  int lim[3][2];
  interior_limits(P, b, lim);
  for (int k = lim[2][0]; k <= lim[2][1]; k++)
    for (int j = lim[1][0]; j <= lim[1][1]; j++)
      for (int i = lim[0][0]; i <= lim[0][1]; i++) {
        size_t c = (size_t)nunk * cidx(NX, NY, i, j, k);
        memcpy(U + c, S + c, sizeof(double) * nunk);
      }
  free(S);
  free(U0);
  return nerr;
}

/* ------------------------------------------------------------------ */
/* correctZone + per-cell EOS, Spark/hy_rk_correctFluxes.F90:126-338 */
static void correct_zone(const orc_params* P, double* V, const double* dFlux, double dt, double dx, double fac) {
  const int nf = orc_nflux(P), ns = P->nscalars;
  double Us[MAXFLX];
  Us[HY_MASS] = V[ORC_DENS];
  Us[HY_XMOM] = V[ORC_DENS] * V[ORC_VELX];
  Us[HY_YMOM] = V[ORC_DENS] * V[ORC_VELY];
  Us[HY_ZMOM] = V[ORC_DENS] * V[ORC_VELZ];
  Us[HY_ENER] = V[ORC_DENS] * V[ORC_ENER];
  for (int s = 0; s < ns; s++) Us[HY_NUM_FLUX + s] = V[ORC_NPROP + s] * V[ORC_DENS];
  for (int n = 0; n < nf; n++) Us[n] = Us[n] - dt / dx * (fac * dFlux[n]);
  V[ORC_DENS] = fmax2(Us[HY_MASS], P->smlrho);
  V[ORC_VELX] = Us[HY_XMOM] / V[ORC_DENS];
  V[ORC_VELY] = Us[HY_YMOM] / V[ORC_DENS];
  V[ORC_VELZ] = Us[HY_ZMOM] / V[ORC_DENS];
  V[ORC_ENER] = fmax2(P->smalle, Us[HY_ENER] / V[ORC_DENS]);
  double ekin = .5 * (V[ORC_VELX] * V[ORC_VELX] + V[ORC_VELY] * V[ORC_VELY] + V[ORC_VELZ] * V[ORC_VELZ]);
  V[ORC_EINT] = fmax2(P->smalle, V[ORC_ENER] - ekin);
  for (int s = 0; s < ns; s++) V[ORC_NPROP + s] = Us[HY_NUM_FLUX + s] / V[ORC_DENS];
  orc_eos_dens_ei_cell(P, V, 1);
}

void orc_correct_fluxes(const orc_params* P, const orc_block* b, double* U, double dt,
                        const double* fbx, const double* fby, const double* fbz) {
  int NX, NY, NZ;
  orc_gc_dims(P, b, &NX, &NY, &NZ);
  const int nunk = orc_nunk(P), nf = orc_nflux(P);
  int lim[3][2];
  interior_limits(P, b, lim);
  int n[3] = {lim[0][1] - lim[0][0] + 1, lim[1][1] - lim[1][0] + 1, lim[2][1] - lim[2][0] + 1};
  double dF[MAXFLX];
  const double del[3] = {b->dx, b->dy, b->dz};
  const double* fb[3] = {fbx, fby, fbz};
  for (int dir = 0; dir < P->ndim; dir++) {
    int e[3] = {n[0], n[1], n[2]};
    e[dir] += 1;
    int d2 = (dir == 0) ? 1 : 0, d3 = (dir == 2) ? 1 : 2;
    /* loop order: x: (k, j), y: (k, i), z: (j, i) -- the faces of one
     * direction touch distinct cells, so the order within a pass is free */
    for (int a3 = 0; a3 < n[d3]; a3++)
      for (int a2 = 0; a2 < n[d2]; a2++) {
        for (int side = 0; side < 2; side++) {
          int c[3];
          c[d2] = a2; c[d3] = a3;
          c[dir] = side ? n[dir] - 1 : 0;
          int f[3] = {c[0], c[1], c[2]};
          if (side) f[dir] += 1;
          const double* buf = fb[dir] + (size_t)nf * ((size_t)f[0] + (size_t)e[0] * ((size_t)f[1] + (size_t)e[1] * f[2]));
          for (int v = 0; v < nf; v++) dF[v] = side ? buf[v] : -buf[v];
          double* V = U + (size_t)nunk * cidx(NX, NY, lim[0][0] + c[0], lim[1][0] + c[1], lim[2][0] + c[2]);
          /* fac = facM on the low x face, facP on the high one (:137-168) */
          double fac = 1.0;
          if (dir == 0 && b->geo) {
            double facM, facP, rc;
            geo_facs(b, NX, NY, lim[0][0] + c[0], lim[1][0] + c[1], b->dx, &facM, &facP, &rc);
            fac = side ? facP : facM;
          }
          correct_zone(P, V, dF, dt, del[dir], fac);
        }
      }
  }
}

/* ------------------------------------------------------------------ */
/* Uniform block grid: same-level neighbour copy, OUTFLOW at domain faces
 * (Grid_bcApplyToRegion.F90:526-532 copies the first interior layer into
 * every guard layer).  Applying the x, y and z faces over full transverse
 * extents makes every guard cell the value at its global index clamped
 * into the domain. */
void orc_ug_fill_guards(const orc_params* P, const orc_block* b, int nbx, int nby, int nbz,
                        double* const* blocks, int nvar) {
  int NX, NY, NZ;
  orc_gc_dims(P, b, &NX, &NY, &NZ);
  const int g = P->nguard;
  int n[3] = {b->nx, (P->ndim > 1) ? b->ny : 1, (P->ndim > 2) ? b->nz : 1};
  int nb[3] = {nbx, nby, nbz};
  int gg[3] = {g, (P->ndim > 1) ? g : 0, (P->ndim > 2) ? g : 0};
  int nblk = nbx * nby * nbz;
#pragma omp parallel for schedule(static)
  for (int bid = 0; bid < nblk; bid++) {
    int bi[3] = {bid % nbx, (bid / nbx) % nby, bid / (nbx * nby)};
    double* dst = blocks[bid];
    for (int k = 0; k < NZ; k++)
      for (int j = 0; j < NY; j++)
        for (int i = 0; i < NX; i++) {
          int loc[3] = {i, j, k};
          int inside = 1;
          for (int d = 0; d < 3; d++)
            if (loc[d] < gg[d] || loc[d] >= gg[d] + n[d]) inside = 0;
          if (inside) continue;
          int sb[3], sl[3];
          for (int d = 0; d < 3; d++) {
            int gidx = bi[d] * n[d] + (loc[d] - gg[d]);
            int gmax = nb[d] * n[d] - 1;
            if (gidx < 0) gidx = 0;
            if (gidx > gmax) gidx = gmax;
            sb[d] = gidx / n[d];
            sl[d] = gidx % n[d] + gg[d];
          }
          const double* src = blocks[sb[0] + nbx * (sb[1] + nby * sb[2])];
          memcpy(dst + (size_t)nvar * cidx(NX, NY, i, j, k),
                 src + (size_t)nvar * cidx(NX, NY, sl[0], sl[1], sl[2]), sizeof(double) * nvar);
        }
  }
}

static orc_block with_geo(const orc_block* b, int i);

double orc_ug_compute_dt(const orc_params* P, const orc_block* b, int nblocks, double* const* blocks) {
  double dt = DBL_MAX;
  for (int i = 0; i < nblocks; i++) {
    const orc_block bi = with_geo(b, i);
    double d = orc_compute_dt(P, &bi, blocks[i], NULL);
    if (d < dt) dt = d;
  }
  return dt;
}

/* The driver's limits on each step's dt in the advance loops below
 * (Driver_computeDt.F90:450-460, and dr_dtInit for the first step of
 * Driver_evolveAll): dt_init > 0 is the next step's dt; later steps take
 * min(dt_cfl, dt_prev * change_factor) (factor > 0), then max with dt_min
 * and min with dt_max (each when > 0).  All zero: the CFL dt. */
static double g_dt_first = 0.0, g_dt_prev = DBL_MAX, g_dt_factor = 0.0, g_dt_min = 0.0, g_dt_max = 0.0;

static const double* const* g_geo = NULL;
static int g_geo_n = 0;

void orc_set_geometry(const double* const* geo, int nblocks) {
  g_geo = geo;
  g_geo_n = geo ? nblocks : 0;
}

/* block i's descriptor with its geometry rows */
static orc_block with_geo(const orc_block* b, int i) {
  orc_block c = *b;
  if (g_geo && i < g_geo_n) c.geo = g_geo[i];
  return c;
}

void orc_set_dt_control(double dt_init, double change_factor, double dt_min, double dt_max) {
  g_dt_first = dt_init;
  g_dt_prev = dt_init > 0.0 ? dt_init : DBL_MAX;
  g_dt_factor = change_factor;
  g_dt_min = dt_min;
  g_dt_max = dt_max;
}

static double driver_dt(double dt) {
  if (g_dt_first > 0.0) {
    dt = g_dt_first;
    g_dt_first = 0.0;
  } else {
    if (g_dt_factor > 0.0) dt = fmin2(dt, g_dt_prev * g_dt_factor);
    if (g_dt_min > 0.0) dt = fmax2(dt, g_dt_min);
    if (g_dt_max > 0.0) dt = fmin2(dt, g_dt_max);
  }
  g_dt_prev = dt;
  return dt;
}

int orc_ug_advance(const orc_params* P, const orc_block* b, int nbx, int nby, int nbz,
                   double* const* blocks, int nsteps, int refill, int nthreads, double* dts_out) {
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif
  int NX, NY, NZ;
  orc_gc_dims(P, b, &NX, &NY, &NZ);
  const int nunk = orc_nunk(P);
  const int nblk = nbx * nby * nbz;
  size_t n = (size_t)nunk * NX * NY * NZ;
  int last, la[3];
  double coeff[3][3], weights[3];
  orc_rk_tables(P, &last, coeff, la, weights);
  int nerr = 0;
  double** S = NULL;
  double** U0 = NULL;
  if (refill) {
    S = (double**)malloc(sizeof(double*) * nblk);
    U0 = (double**)malloc(sizeof(double*) * nblk);
    for (int i = 0; i < nblk; i++) {
      S[i] = (double*)malloc(sizeof(double) * n);
      U0[i] = (double*)malloc(sizeof(double) * n);
    }
  }
  for (int step = 0; step < nsteps; step++) {
    double dt = driver_dt(orc_ug_compute_dt(P, b, nblk, blocks));
    if (dts_out) dts_out[step] = dt;
    orc_ug_fill_guards(P, b, nbx, nby, nbz, blocks, nunk);
    if (!refill) {
#pragma omp parallel for schedule(dynamic) reduction(+ : nerr)
      for (int i = 0; i < nblk; i++) {
        const orc_block bi = with_geo(b, i);
        nerr += orc_hydro_block(P, &bi, blocks[i], dt, NULL, NULL, NULL);
      }
    } else {
#pragma omp parallel for schedule(static)
      for (int i = 0; i < nblk; i++) {
        orc_shock_detect(P, b, blocks[i]);
        memcpy(S[i], blocks[i], sizeof(double) * n);
        memcpy(U0[i], blocks[i], sizeof(double) * n);
      }
      for (int stage = 1; stage <= last; stage++) {
#pragma omp parallel for schedule(dynamic) reduction(+ : nerr)
        for (int i = 0; i < nblk; i++) {
          int e = 0;
          const orc_block bi = with_geo(b, i);
          orc_stage(P, &bi, U0[i], S[i], stage, 0, dt, NULL, NULL, NULL, &e);
          nerr += e;
        }
        if (stage < last) orc_ug_fill_guards(P, b, nbx, nby, nbz, S, nunk);
      }
      int lim[3][2];
      interior_limits(P, b, lim);
#pragma omp parallel for schedule(static)
      for (int i = 0; i < nblk; i++)
        for (int k = lim[2][0]; k <= lim[2][1]; k++)
          for (int j = lim[1][0]; j <= lim[1][1]; j++)
            for (int ii = lim[0][0]; ii <= lim[0][1]; ii++) {
              size_t c = (size_t)nunk * cidx(NX, NY, ii, j, k);
              memcpy(blocks[i] + c, S[i] + c, sizeof(double) * nunk);
            }
    }
  }
  if (refill) {
    for (int i = 0; i < nblk; i++) { free(S[i]); free(U0[i]); }
    free(S);
    free(U0);
  }
  return nerr;
}

int orc_batch_advance(const orc_params* P, const orc_block* bl, int nblocks, double* const* blocks,
                      int nsteps, int nthreads, double* dts_out) {
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif
  int nerr = 0;
  for (int step = 0; step < nsteps; step++) {
    double dt = DBL_MAX;
    for (int i = 0; i < nblocks; i++) {
      double d = orc_compute_dt(P, &bl[i], blocks[i], NULL);
      if (d < dt) dt = d;
    }
    dt = driver_dt(dt);
    if (dts_out) dts_out[step] = dt;
#pragma omp parallel for schedule(dynamic) reduction(+ : nerr)
    for (int i = 0; i < nblocks; i++) nerr += orc_hydro_block(P, &bl[i], blocks[i], dt, NULL, NULL, NULL);
  }
  return nerr;
}

/* ------------------------------------------------------------------ */
/* Block hierarchy: guard fill, flux-register exchange, advance.        */

typedef struct amr_map {
  int lmax, md[3];
  int* leaf; /* md[0]*md[1]*md[2]: block covering each finest-level block slot */
} amr_map;

static int amr_build_map(const orc_amr* A, int nblocks, amr_map* M) {
  M->lmax = 1;
  for (int i = 0; i < nblocks; i++)
    if (A->level[i] > M->lmax) M->lmax = A->level[i];
  for (int d = 0; d < 3; d++) M->md[d] = A->nb0[d] << (M->lmax - 1);
  size_t tot = (size_t)M->md[0] * M->md[1] * M->md[2];
  M->leaf = (int*)malloc(sizeof(int) * tot);
  for (size_t t = 0; t < tot; t++) M->leaf[t] = -1;
  for (int i = 0; i < nblocks; i++) {
    const int sh = M->lmax - A->level[i];
    int lo[3], hi[3];
    for (int d = 0; d < 3; d++) {
      lo[d] = (A->origin[3 * i + d] / A->n) << sh;
      hi[d] = lo[d] + (1 << sh);
    }
    for (int z = lo[2]; z < hi[2]; z++)
      for (int y = lo[1]; y < hi[1]; y++)
        for (int x = lo[0]; x < hi[0]; x++) M->leaf[x + (size_t)M->md[0] * (y + (size_t)M->md[1] * z)] = i;
  }
  return 0;
}

/* leaf covering global cell gi (level L indices) */
static int amr_leaf_at(const orc_amr* A, const amr_map* M, int L, const int gi[3]) {
  int s[3];
  for (int d = 0; d < 3; d++) s[d] = (gi[d] << (M->lmax - L)) / A->n;
  return M->leaf[s[0] + (size_t)M->md[0] * (s[1] + (size_t)M->md[1] * s[2])];
}

/* global cell index of level L wrapped (PERIODIC), clamped (OUTFLOW,
 * Grid_bcApplyToRegion.F90:526-532) or mirrored (REFLECT) into the domain;
 * returns the mask of the dimensions whose normal velocity changes sign */
static int amr_domain(const orc_amr* A, int L, int gi[3]) {
  int flip = 0;
  for (int d = 0; d < 3; d++) {
    const int N = A->nb0[d] * A->n << (L - 1);
    if (A->bc[d] == ORC_BC_PERIODIC) gi[d] = ((gi[d] % N) + N) % N;
    else if (A->bc[d] == ORC_BC_REFLECT && (gi[d] < 0 || gi[d] > N - 1)) {
      gi[d] = gi[d] < 0 ? -1 - gi[d] : 2 * N - 1 - gi[d];
      flip |= 1 << d;
    } else gi[d] = gi[d] < 0 ? 0 : (gi[d] > N - 1 ? N - 1 : gi[d]);
  }
  return flip;
}

static inline double flipped(double x, int flip, int v) {
  return (v >= ORC_VELX && v <= ORC_VELZ && ((flip >> (v - ORC_VELX)) & 1)) ? -x : x;
}

/* variable v at level-L cell gi (inside the domain) from the leaf covering
 * it: the containing cell of a same-level or coarser leaf, the mean of the
 * 8 cells of a leaf one level finer.  A leaf two or more levels finer has
 * no such cell set (the hierarchy is not nested deeply enough for the fill
 * depth): *bad is set and nothing is read. */
static double amr_value(const orc_amr* A, const amr_map* M, double* const* blocks, int nvar, int g, int NX, int L,
                        const int gi[3], int v, int* bad) {
  const int s = amr_leaf_at(A, M, L, gi);
  const int Ls = A->level[s];
  const int* os = A->origin + 3 * s;
  const double* src = blocks[s];
  if (Ls > L + 1) {
    *bad = 1;
    return 0.0;
  }
  if (Ls <= L) {
    int sl[3];
    for (int d = 0; d < 3; d++) sl[d] = (gi[d] >> (L - Ls)) - os[d] + g;
    return src[(size_t)nvar * cidx(NX, NX, sl[0], sl[1], sl[2]) + v];
  }
  int f0[3];
  for (int d = 0; d < 3; d++) f0[d] = 2 * gi[d] - os[d] + g;
  double sum = 0.0;
  for (int c = 0; c < 8; c++)
    sum = sum + src[(size_t)nvar * cidx(NX, NX, f0[0] + (c & 1), f0[1] + ((c >> 1) & 1), f0[2] + ((c >> 2) & 1)) + v];
  return 0.125 * sum;
}

/* monotonized-central slope of a coarse cell from its two differences */
static double mc_slope(double a, double b) {
  if (a * b <= 0.0) return 0.0;
  const double t = fabs(a) < fabs(b) ? 2.0 * fabs(a) : 2.0 * fabs(b);
  const double c = 0.5 * fabs(a + b);
  const double m = t < c ? t : c;
  return a > 0.0 ? m : -m;
}

int orc_amr_fill_guards(const orc_params* P, const orc_amr* A, int nblocks, double* const* blocks, int nvar,
                        int depth) {
  amr_map M;
  amr_build_map(A, nblocks, &M);
  const int n = A->n, g = P->nguard, NX = n + 2 * g;
  int nbad = 0;
#pragma omp parallel for schedule(dynamic) reduction(+ : nbad)
  for (int b = 0; b < nblocks; b++) {
    int bad = 0;
    const int L = A->level[b];
    const int* o = A->origin + 3 * b;
    double* dst = blocks[b];
    for (int k = g - depth; k < g + n + depth; k++)
      for (int j = g - depth; j < g + n + depth; j++)
        for (int i = g - depth; i < g + n + depth; i++) {
          const int loc[3] = {i, j, k};
          int inside = 1, gi[3];
          for (int d = 0; d < 3; d++) {
            if (loc[d] < g || loc[d] >= g + n) inside = 0;
            gi[d] = o[d] + loc[d] - g;
          }
          if (inside) continue;
          const int flip = amr_domain(A, L, gi);
          const int s = amr_leaf_at(A, &M, L, gi);
          double* out = dst + (size_t)nvar * cidx(NX, NX, i, j, k);
          if (A->level[s] != L - 1) {
            for (int v = 0; v < nvar; v++)
              out[v] = flipped(amr_value(A, &M, blocks, nvar, g, NX, L, gi, v, &bad), flip, v);
            continue;
          }
          /* a coarser leaf: limited linear prolongation from the parent
           * coarse cell and its 6 face neighbours (MC slopes) */
          int gc[3];
          for (int d = 0; d < 3; d++) gc[d] = gi[d] >> 1;
          for (int v = 0; v < nvar; v++) {
            const double c0 = amr_value(A, &M, blocks, nvar, g, NX, L - 1, gc, v, &bad);
            double val = c0;
            for (int d = 0; d < 3; d++) {
              int gm[3] = {gc[0], gc[1], gc[2]}, gp[3] = {gc[0], gc[1], gc[2]};
              gm[d] -= 1;
              gp[d] += 1;
              const int fm = amr_domain(A, L - 1, gm), fp = amr_domain(A, L - 1, gp);
              const double cm = flipped(amr_value(A, &M, blocks, nvar, g, NX, L - 1, gm, v, &bad), fm, v);
              const double cp = flipped(amr_value(A, &M, blocks, nvar, g, NX, L - 1, gp, v, &bad), fp, v);
              const double sl = mc_slope(c0 - cm, cp - c0);
              val = val + sl * ((gi[d] & 1) ? 0.25 : -0.25);
            }
            out[v] = flipped(val, flip, v);
          }
        }
    nbad += bad;
  }
  free(M.leaf);
  return nbad;
}

void orc_flux_exchange(const orc_params* P, const orc_amr* A, int nblocks, double* const* fb, double* const* fc) {
  amr_map M;
  amr_build_map(A, nblocks, &M);
  const int n = A->n, nf = orc_nflux(P);
#pragma omp parallel for schedule(dynamic)
  for (int b = 0; b < nblocks; b++) {
    const int L = A->level[b];
    const int* o = A->origin + 3 * b;
    for (int dir = 0; dir < 3; dir++) {
      int e[3] = {n, n, n};
      e[dir] += 1;
      const int d2 = (dir == 0) ? 1 : 0, d3 = (dir == 2) ? 1 : 2;
      double* out = fc[3 * b + dir];
      memset(out, 0, sizeof(double) * (size_t)nf * e[0] * e[1] * e[2]);
      const double* own = fb[3 * b + dir];
      for (int side = 0; side < 2; side++) {
        const int gmax = (A->nb0[dir] * n << (L - 1)) - 1;
        int across = side ? o[dir] + n : o[dir] - 1;
        if (A->bc[dir] == ORC_BC_PERIODIC) across = (across + gmax + 1) % (gmax + 1);
        else if (across < 0 || across > gmax) continue;
        for (int a3 = 0; a3 < n; a3++)
          for (int a2 = 0; a2 < n; a2++) {
            int gi[3];
            gi[dir] = across;
            gi[d2] = o[d2] + a2;
            gi[d3] = o[d3] + a3;
            const int s = amr_leaf_at(A, &M, L, gi);
            if (A->level[s] != L + 1) continue;
            const int* os = A->origin + 3 * s;
            int ff[3], f[3];
            ff[dir] = side ? 0 : n;
            ff[d2] = 2 * gi[d2] - os[d2];
            ff[d3] = 2 * gi[d3] - os[d3];
            f[dir] = side ? n : 0;
            f[d2] = a2;
            f[d3] = a3;
            const double* fine = fb[3 * s + dir];
            const size_t co = (size_t)nf * ((size_t)f[0] + (size_t)e[0] * ((size_t)f[1] + (size_t)e[1] * f[2]));
            size_t fo[4];
            for (int q = 0; q < 4; q++) {
              int t[3] = {ff[0], ff[1], ff[2]};
              t[d2] += q & 1;
              t[d3] += q >> 1;
              fo[q] = (size_t)nf * ((size_t)t[0] + (size_t)e[0] * ((size_t)t[1] + (size_t)e[1] * t[2]));
            }
            for (int v = 0; v < nf; v++) {
              const double rest = (fine[fo[0] + v] + fine[fo[1] + v] + fine[fo[2] + v] + fine[fo[3] + v]) * 0.25;
              out[co + v] = rest - own[co + v];
            }
          }
      }
    }
  }
  free(M.leaf);
}

int orc_amr_advance(const orc_params* P, const orc_amr* A, const orc_block* bl, int nblocks, double* const* blocks,
                    int nsteps, int flux_correct, int nthreads, double* dts_out) {
#ifdef _OPENMP
  if (nthreads > 0) omp_set_num_threads(nthreads);
#else
  (void)nthreads;
#endif
  const int nunk = orc_nunk(P), nf = orc_nflux(P), n = A->n;
  int last, la[3];
  double coeff[3][3], weights[3];
  orc_rk_tables(P, &last, coeff, la, weights);
  const int depth = la[0] + P->nstencil;
  double** fb = NULL;
  double** fc = NULL;
  if (flux_correct) {
    fb = (double**)malloc(sizeof(double*) * 3 * nblocks);
    fc = (double**)malloc(sizeof(double*) * 3 * nblocks);
    for (int i = 0; i < 3 * nblocks; i++) {
      size_t sz = (size_t)nf * (n + 1) * n * n;
      fb[i] = (double*)calloc(sz, sizeof(double));
      fc[i] = (double*)calloc(sz, sizeof(double));
    }
  }
  int nerr = 0;
  for (int step = 0; step < nsteps && nerr >= 0; step++) {
    if (orc_amr_fill_guards(P, A, nblocks, blocks, nunk, depth)) {
      nerr = -1;  /* a guard layer lies over a leaf two levels finer */
      break;
    }
    double dt = DBL_MAX;
    for (int i = 0; i < nblocks; i++) {
      double d = orc_compute_dt(P, &bl[i], blocks[i], NULL);
      if (d < dt) dt = d;
    }
    dt = driver_dt(dt);
    if (dts_out) dts_out[step] = dt;
#pragma omp parallel for schedule(dynamic) reduction(+ : nerr)
    for (int i = 0; i < nblocks; i++)
      nerr += orc_hydro_block(P, &bl[i], blocks[i], dt, fb ? fb[3 * i] : NULL, fb ? fb[3 * i + 1] : NULL,
                              fb ? fb[3 * i + 2] : NULL);
    if (flux_correct) {
      orc_flux_exchange(P, A, nblocks, fb, fc);
#pragma omp parallel for schedule(dynamic)
      for (int i = 0; i < nblocks; i++) orc_correct_fluxes(P, &bl[i], blocks[i], dt, fc[3 * i], fc[3 * i + 1], fc[3 * i + 2]);
    }
  }
  if (flux_correct) {
    for (int i = 0; i < 3 * nblocks; i++) { free(fb[i]); free(fc[i]); }
    free(fb);
    free(fc);
  }
  return nerr;
}
