/*
 * spark_oracle.h -- CPU restatement of the Flash-X Spark unsplit-hydro RK
 * stage (Flash-X/source/physics/Hydro/HydroMain/Spark).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker the HIP path is
 * compared against.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it; no product code links it.
 *
 * Block data use the reference unk layout: unk(var, i, j, k), variable
 * index fastest (Paramesh_for_FlashX/.../physicaldata.F90:80), 0-based here.
 * Variable order is the alphabetical order setup.py gives Spark + Gamma
 * (dens eint ener gamc game pres shok temp velx vely velz), then the mass
 * scalars.  Face fluxes use (flux, i, j, k) with the face i the low face of
 * cell i, as hy_saveFluxes stores them (Spark/Inhost/hy_local.ini:118-127).
 */
#ifndef SPARK_ORACLE_H
#define SPARK_ORACLE_H
#ifdef __cplusplus
extern "C" {
#endif

enum {
  ORC_DENS = 0, ORC_EINT = 1, ORC_ENER = 2, ORC_GAMC = 3, ORC_GAME = 4,
  ORC_PRES = 5, ORC_SHOK = 6, ORC_TEMP = 7, ORC_VELX = 8, ORC_VELY = 9,
  ORC_VELZ = 10, ORC_NPROP = 11
};

enum { ORC_RECON_WENO5 = 0, ORC_RECON_MP5 = 1, ORC_RECON_PLM = 2, ORC_RECON_FOG = 3 };
enum { ORC_RIEMANN_HLLC = 0, ORC_RIEMANN_HLL = 1 };

typedef struct orc_params {
  double gamma;        /* Eos/EosMain/Config:43            */
  double cfl;          /* Hydro/HydroMain/Config:26        */
  double cvisc;        /* Spark/Config:17                  */
  double smlrho, smallp, smalle, smallx; /* Grid Config floors */
  double eintSwitch;   /* Eos/EosMain/Config:39            */
  double tiny;         /* hy_tiny, Spark/Hydro_data.F90:85 */
  double gasConst;     /* PhysicalConstants_init.F90:186   */
  double abar;         /* eos_singleSpeciesA               */
  double grav[3];      /* constant gravity (source term only) */
  int ndim, nguard, nstencil, rk_order;
  int recon, riemann;
  int use_flattening, use_hybrid_riemann;
  int nscalars;
  int use_gravity;
  /* species: the first nspecies scalars (Spark/Hydro.ini:227-235); eos 1 =
   * Multigamma (Eos/EosMain/Multigamma/Eos.F90) with per-species A, gamma */
  int nspecies;
  int eos;
  double species_a[8];
  double species_gamma[8];
} orc_params;

/* One block: interior sizes (1 for unused dims) and its deltas.  geo:
 * NULL in Cartesian geometry; otherwise (x = radius), over the guard-celled
 * box of NX x NY cells: cell-centre, left-edge and right-edge radius (NX
 * each), then the area of each cell's low x face and each cell's volume
 * (NX * NY each, x fastest) -- the Grid arrays hy_xCenter, hy_xLeft,
 * hy_xRight, hy_farea, hy_cvol of Spark/Hydro.ini:40-56 -- with geometry 1
 * cylindrical, 2 spherical. */
typedef struct orc_block {
  int nx, ny, nz;
  double dx, dy, dz;
  const double* geo;
  int geometry;
} orc_block;

/* helpers */
int orc_nunk(const orc_params* P);
int orc_nflux(const orc_params* P);
void orc_gc_dims(const orc_params* P, const orc_block* b, int* NX, int* NY, int* NZ);

/* Individual reference functions (exposed for unit tests). */
void orc_prim2con(const double* V, double* U);                /* prim2con.F90:34-62 */
void orc_prim2flx(int dir, const double* V, double* F);       /* prim2flx.F90:36-118 */
int  orc_riemann_hllc(const orc_params* P, int dir, const double* VL, const double* VR,
                      int inShock, double* F);                /* riemann.F90-mc:42-282 */
int  orc_riemann_hll(const orc_params* P, int dir, const double* VL, const double* VR,
                     double* F);                               /* tvd2/riemann.F90:63-141 */
void orc_recon(const orc_params* P, int nrecon, const double* pen, int i1, int plen,
               double flat, double* uP, double* uM);           /* reconstruct.F90-mc:20-104 */
void orc_eos_dens_ei_cell(const orc_params* P, double* v, int full);  /* Eos_wrapped path */

/* Block-level kernels. lim[d][0..1] are 0-based GC-box limits. */
void orc_flattening(const orc_params* P, const orc_block* b, const double* S,
                    const int lim[3][2], double* flat3d);
void orc_get_face_flux(const orc_params* P, const orc_block* b, const double* S,
                       const double* flat3d, const int lim[3][2],
                       double* flx, double* fly, double* flz, int* nerr);
void orc_update_soln(const orc_params* P, const orc_block* b, const double* U0, double* S,
                     const int lim[3][2], double dt, const double coeffs[3],
                     const double* flx, const double* fly, const double* flz);
void orc_eos_wrapped(const orc_params* P, const orc_block* b, double* S, const int lim[3][2]);
void orc_shock_detect(const orc_params* P, const orc_block* b, double* U);
double orc_compute_dt(const orc_params* P, const orc_block* b, const double* U, int minloc[3]);

/* RK tables (Spark/Hydro.ini:2-48, with the RK3 constants defined). */
void orc_rk_tables(const orc_params* P, int* last_stage, double coeff[3][3],
                   int limits_array[3], double weights[3]);

/* One stage on one block: ngoff extra layers, optional flux-buffer sum. */
void orc_stage(const orc_params* P, const orc_block* b, const double* U0, double* S,
               int stage, int ngoff, double dt,
               double* fbx, double* fby, double* fbz, int* nerr);

/* Whole Hydro() body for one block, telescoping the stages
 * (Spark/Hydro.F90-mc:113-201).  U is updated in place (interior). */
int orc_hydro_block(const orc_params* P, const orc_block* b, double* U, double dt,
                    double* fbx, double* fby, double* fbz);

/* Flux correction on one block (hy_rk_correctFluxes.F90:46-338).
 * fb[xyz] are (restricted fine - coarse) on coarse-side f/c faces, 0 elsewhere. */
void orc_correct_fluxes(const orc_params* P, const orc_block* b, double* U, double dt,
                        const double* fbx, const double* fby, const double* fbz);

/* Uniform block grid (one refinement level) helpers: guard fill by
 * same-level neighbour copy with OUTFLOW at the domain faces, and a full
 * multi-step driver in either parity (telescoping) or refill mode. */
void orc_ug_fill_guards(const orc_params* P, const orc_block* b, int nbx, int nby, int nbz,
                        double* const* blocks, int nvar);
/* Per-block geometry rows (orc_block.geo) for the advance loops, which
 * share one block descriptor: geo[i] for block i, or NULL (Cartesian). */
void orc_set_geometry(const double* const* geo, int nblocks);
/* The driver's dt limits for the advance loops (Driver_computeDt.F90:450-460
 * and dr_dtInit); all zero: the CFL dt (the default). */
void orc_set_dt_control(double dt_init, double change_factor, double dt_min, double dt_max);
int orc_ug_advance(const orc_params* P, const orc_block* b, int nbx, int nby, int nbz,
                   double* const* blocks, int nsteps, int refill, int nthreads,
                   double* dts_out);
double orc_ug_compute_dt(const orc_params* P, const orc_block* b, int nblocks,
                         double* const* blocks);

/* Independent blocks (frozen guards): nsteps of orc_hydro_block with a
 * shared CFL dt recomputed each step.  blocks[i] each has its own shape. */
int orc_batch_advance(const orc_params* P, const orc_block* bl, int nblocks,
                      double* const* blocks, int nsteps, int nthreads, double* dts_out);

/* Block hierarchy of leaves (Paramesh-like): every block has n^3 interior
 * cells; level[i] is 1-based; origin[3*i..] is the block's first interior
 * cell in the global index space of its level; the level-1 grid has
 * nb0[0] x nb0[1] x nb0[2] blocks; neighbouring leaves differ by at most
 * one level. */
typedef struct orc_amr {
  int n;
  int nb0[3];
  const int* level;
  const int* origin;
  int bc[3]; /* per dimension: ORC_BC_OUTFLOW, ORC_BC_PERIODIC or ORC_BC_REFLECT */
} orc_amr;
enum { ORC_BC_OUTFLOW = 0, ORC_BC_PERIODIC = 1, ORC_BC_REFLECT = 2 };

/* Guard fill over the hierarchy: a guard cell takes, with its global index
 * first wrapped (PERIODIC), clamped (OUTFLOW) or mirrored (REFLECT, the
 * normal velocity changing sign) into the domain, the value of the leaf
 * covering its position -- a copy from a same-level
 * leaf, the mean of the 8 fine cells from a finer leaf, and from a coarser
 * leaf a limited linear prolongation: the parent coarse cell plus, per
 * dimension, a monotonized-central slope of its face neighbours times
 * +-1/4.  depth guard layers of all nvar variables.  Returns the number of
 * blocks with a guard cell (or a prolongation neighbour) over a leaf two or
 * more levels finer, a hierarchy not nested deeply enough for the depth;
 * those values are not filled (0: all filled). */
int orc_amr_fill_guards(const orc_params* P, const orc_amr* A, int nblocks, double* const* blocks, int nvar,
                        int depth);

/* Flux-register exchange (Grid_putFluxData_block, Grid_communicateFluxes,
 * Grid_getFluxCorrData_block with consv_flux_densities, red_f = 0.25,
 * PM4_package/source/amr_restrict_red.F90, Grid_getFluxCorrData_block.F90:182-216):
 * fc = (restricted fine - own) on coarse-side faces of fine/coarse
 * boundaries, 0 on every other boundary face.  fb / fc hold 3 arrays per
 * block (x, y, z interior faces, flux index fastest). */
void orc_flux_exchange(const orc_params* P, const orc_amr* A, int nblocks, double* const* fb, double* const* fc);

/* nsteps of the hierarchy: guard fill, shared CFL dt, Hydro() on every
 * leaf, then (flux_correct) the exchange and hy_rk_correctFluxes.  Returns
 * the count of Riemann errors, or -1 when the guard fill refuses the
 * hierarchy (see orc_amr_fill_guards). */
int orc_amr_advance(const orc_params* P, const orc_amr* A, const orc_block* bl, int nblocks, double* const* blocks,
                    int nsteps, int flux_correct, int nthreads, double* dts_out);

#ifdef __cplusplus
}
#endif
#endif
