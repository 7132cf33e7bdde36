// spark_capi.cpp -- the extern "C" boundary (include/spark_hip.h): batch
// bookkeeping, device memory layout and the launch sequence of one RK step.
//
// Device layout of a batch (all pools allocated once per batch):
//   state pools U (solution) and S (stage state): per block, per variable,
//     a contiguous guard-celled box (x fastest), variable stride padded to
//     64 doubles; variable order dens eint ener gamc game pres shok temp
//     velx vely velz scalars.
//   one-variable cell pool: flattening coefficients / sound-speed scratch.
//   face pools (one per direction): per block, per flux variable, faces on
//     an (NX+1)(NY+1)(NZ+1) box; face i is the low face of cell i.
//   flux-register pools (flux_correct only): interior faces per block.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cfloat>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/spark_hip.h"
#include "spark_batched.hpp"
#include "spark_regions.hpp"
#include "spark_launch.hpp"

using spark::BlockMeta;
using spark::DevParams;
using spark::PlainLaunch;

namespace {

thread_local std::string g_err;
int g_device = -1;
hipStream_t g_stream = nullptr;
// device address of each device's fixed staging area (spark_pack_noalloc),
// looked up once per device at spark_init
constexpr int kMaxDevices = 64;
void* g_noalloc[kMaxDevices] = {};

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define HIPCHK(expr)                                                                           \
  do {                                                                                         \
    hipError_t e_ = (expr);                                                                    \
    if (e_ != hipSuccess) return fail(SPARK_ERR_HIP, "%s: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                                      __FILE__, __LINE__);                                     \
  } while (0)

inline int64_t round_up(int64_t a, int64_t m) { return (a + m - 1) / m * m; }

struct ShapeClass {
  int nx, ny, nz;
  std::vector<int> ids;
  int* ids_d = nullptr;
};

}  // namespace

struct spark_batch {
  spark_params p;
  DevParams dp;
  int nblk = 0;
  int nvar = 0;
  std::vector<spark_block_desc> desc;
  std::vector<BlockMeta> meta_h;
  BlockMeta* meta_d = nullptr;
  std::vector<ShapeClass> classes;
  int64_t pool = 0, opool = 0, dpool = 0, bpool = 0;  // sizes in doubles
  double* U = nullptr;
  double* S = nullptr;
  double* flat = nullptr;
  double* div = nullptr;  // plain form: the sweeps' divergence terms per cell
  double* fb[3] = {nullptr, nullptr, nullptr};
  double* dtp = nullptr;
  long long* loc = nullptr;
  double* dt_hist = nullptr;
  int hist_cap = 0;
  void* dtparts = nullptr;
  int dtchunks = 1;
  int* err = nullptr;
  int ug[3] = {0, 0, 0};
  bool has_ug = false;
  bool step_open = false;
  int64_t max_cells = 0;
  spark::Prof prof;
  // stage buffers: S (stage 1) and S2 (stage 2 of RK3); star: the output
  // of the last stage run before the final one (spark_unpack which = 1)
  double* S2 = nullptr;
  double* star = nullptr;
  bool fused = false;
  bool fused64 = false;  // batched form in FP64 arithmetic
  int cus = 0;            // compute units of the device
  size_t lds_limit = 0;   // LDS bytes a workgroup may allocate
  int64_t* nbr_d = nullptr;  // uniform grid: 27 neighbour pool offsets per block
  bool amr = false;          // block hierarchy declared (spark_set_amr_grid)
  // batched forms on a block hierarchy: blocks whose 26 neighbours are leaves
  // of their own level read their halos from them (nbr_d row), every stage
  // over the interior only (amr_remap_blk 1); blocks with some such
  // neighbours read those and telescope over the other guard sectors
  // (amr_remap_blk 2, amr_local: bit 13 + ox + 3 oy + 9 oz set for a sector
  // they compute); the others telescope over all their guard cells (0)
  std::vector<char> amr_remap_blk;
  std::vector<uint32_t> amr_local;
  uint32_t* amr_smask_d = nullptr;  // per block the guard sectors it computes (all for telescoping blocks)
  spark::DtCtl* dtctl_d = nullptr;  // the driver's dt limits for spark_advance (spark_set_dt_control)
  double* geo_d = nullptr;          // per block radius / low x-face area / volume rows (spark_set_geometry)
  bool dt_control = false;
  int amr_remap_count = 0;
  int amr_partial_count = 0;
  bool same_level = true;      // SPARK_OPT_SAME_LEVEL_HALOS
  int* remap_ids_d = nullptr;  // the remapping blocks
  int* tele_ids_d = nullptr;   // the telescoping blocks
  int n_tele = 0;
  // per block the faces (bit 2d+s) with a finer leaf across, and the blocks
  // that have one (the FP32-flux form corrects only the cells on them)
  unsigned char* amr_fmask_d = nullptr;
  int* corr_ids_d = nullptr;
  int n_corr = 0;
  // fc is zero on every face without a finer leaf across (set at
  // allocation and kept by the exchange): the exchange then only visits
  // the blocks that have such a face.  Host corrections or a new hierarchy
  // clear it until the next full exchange.
  bool fc_clean = false;
  spark::AmrGrid amr_g{};
  int* amr_map_d = nullptr;
  int* amr_org_d = nullptr;
  double* fc[3] = {nullptr, nullptr, nullptr};  // flux corrections (fine/coarse faces)
  size_t stage_lds[3] = {0, 0, 0};
  int stage_grid[3] = {0, 0, 0};
  struct StagePlan {
    std::vector<spark::FusedItem> items;
    spark::FusedItem* items_d = nullptr;
    int* qoff_d = nullptr;
    int* qlen_d = nullptr;
  };
  StagePlan plan[3];
  int* counters = nullptr;  // FUSED_XCDS per stage
  std::map<std::pair<int, int>, std::pair<int, int>> tiles;  // (class, ngoff) -> (tx, ty)
  bool planned = false;
  bool autotune = true;
  int* tune_err = nullptr;    // Riemann-error counter of tuning launches
  double* tune_dt = nullptr;  // dt = 0 of tuning launches (the batch's dt stays untouched)
  // graph of one device-resident step (dt, guard fill, stages)
  bool use_graph = true;
  int* hist_slot = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t gexec = nullptr;
  int last_stage = 3;
  double coeff[3][3];
  int limits_array[3];
  double weights[3];
};

namespace {

void rk_tables(const spark_params& p, int* last, double c[3][3], int la[3], double w[3]) {
  // Spark/Hydro.ini:2-48 (RK3 constants 1/3, 2/3, 1/6 supplied)
  const int ns = p.nstencil;
  if (p.rk_order == 3) {
    const double onethird = 1. / 3., twothirds = 2. / 3., onesixth = 1. / 6.;
    const double t[3][3] = {{1., 0., 1.}, {0.75, 0.25, 0.25}, {onethird, twothirds, twothirds}};
    memcpy(c, t, sizeof(t));
    *last = 3;
    la[0] = 2 * ns; la[1] = ns; la[2] = 0;
    w[0] = onesixth; w[1] = onesixth; w[2] = twothirds;
  } else {
    const double t[3][3] = {{1., 0., 1.}, {0.5, 0.5, 0.5}, {0., 0., 0.}};
    memcpy(c, t, sizeof(t));
    *last = 2;
    la[0] = ns; la[1] = 0; la[2] = 0;
    w[0] = 0.5; w[1] = 0.5; w[2] = 0.;
  }
}

PlainLaunch make_launch(spark_batch* b, const ShapeClass& c) {
  PlainLaunch L;
  L.P = b->dp;
  L.stream = g_stream;
  L.meta = b->meta_d;
  L.blist = c.ids_d;
  L.nblk = (int)c.ids.size();
  L.nx = c.nx; L.ny = c.ny; L.nz = c.nz;
  L.U0 = b->U;
  L.Sin = b->S;
  L.Sout = b->S;
  L.flat = b->flat;
  L.div = b->div;
  for (int d = 0; d < 3; d++) L.fb[d] = b->p.flux_correct ? b->fb[d] : nullptr;
  L.weight = 0.;
  L.add_flux = 0;
  L.dtp = b->dtp;
  L.err = b->err;
  L.prof = b->prof.on ? &b->prof : nullptr;
  return L;
}

int required_guard(const spark_params& p) {
  int la[3], last;
  double c[3][3], w[3];
  rk_tables(p, &last, c, la, w);
  const int maxoff = (p.guard_mode == SPARK_GUARD_REFILL) ? 0 : la[0];
  // the fused kernel of the batched forms stages a 3-cell halo whatever
  // the reconstruction's own stencil
  const bool fused = p.form == SPARK_FORM_FP32_FLUX || p.form == SPARK_FORM_BATCHED_FP64;
  return maxoff + (fused ? std::max(p.nstencil, 3) : p.nstencil);
}

bool fused_remap(const spark_batch* b);
void drop_graph(spark_batch* b);

// batched forms on a block hierarchy: same-level halos in use
bool amr_remap_on(const spark_batch* b) {
  return b->same_level && b->amr_remap_count + b->amr_partial_count > 0;
}

int run_dt(spark_batch* b, int hist_slot) {
  for (auto& c : b->classes) {
    PlainLaunch L = make_launch(b, c);
    HIPCHK(spark::plain_dt_part(L, b->U, b->dtparts, b->dtchunks));
  }
  // the driver's limits apply to the steps of spark_advance only (hist_slot
  // >= 0); spark_compute_dt returns the CFL dt as Hydro_computeDt does
  HIPCHK(spark::plain_dt_final(g_stream, b->dp, b->dtparts, b->nblk, b->dtchunks, b->dtp, b->loc,
                               hist_slot >= 0 ? b->dt_hist : nullptr, b->hist_slot,
                               hist_slot >= 0 && b->dt_control ? b->dtctl_d : nullptr));
  return SPARK_OK;
}

// Guard fill of a uniform block grid.  full: every variable over every
// guard layer (Grid_fillGuardCells semantics).  Otherwise only what the
// stages read: dens eint ener gamc pres vel(3) and the scalars, over the
// guard layers the next stage's stencils reach.
int run_fill(spark_batch* b, double* pool, bool full, int depth) {
  PlainLaunch L = make_launch(b, b->classes[0]);
  L.blist = nullptr;
  L.nblk = b->nblk;
  spark::VarList vl;
  vl.n = 0;
  if (full) {
    for (int v = 0; v < b->nvar; v++) vl.v[vl.n++] = v;
    depth = b->p.nguard;
  } else {
    const int need[8] = {spark::V_DENS, spark::V_EINT, spark::V_ENER, spark::V_GAMC,
                         spark::V_PRES, spark::V_VELX, spark::V_VELY, spark::V_VELZ};
    for (int v : need) vl.v[vl.n++] = v;
    for (int s = 0; s < b->p.nscalars; s++) vl.v[vl.n++] = spark::V_NPROP + s;
  }
  if (b->amr && !full && amr_remap_on(b)) {
    // remapping blocks read no guard cell of their own in the stages.
    // Those whose guard sectors all face same-level leaves get none: their
    // shockDetect reads the leaves around them through the neighbour table.
    // The others get what shockDetect of their interior reads (and of the
    // layer around it, which the hybrid Riemann solver reads SHOK of)
    if (b->n_tele > 0) {
      // blocks that read some sectors from same-level leaves: those sectors
      // only to the layer shockDetect of the interior reads (the hybrid
      // solver reads SHOK of computed cells next to them: all layers then)
      L.blist = b->tele_ids_d;
      L.nblk = b->n_tele;
      HIPCHK(spark::amr_fill(L, b->amr_g, pool, vl, depth, b->p.use_hybrid_riemann ? nullptr : b->amr_smask_d, 1));
    }
  } else if (b->amr) {
    HIPCHK(spark::amr_fill(L, b->amr_g, pool, vl, depth));
  } else {
    HIPCHK(spark::plain_ug_fill(L, b->ug[0], b->ug[1], b->ug[2], pool, vl, depth));
  }
  return SPARK_OK;
}

// the fine/coarse flux-register exchange: over every block until fc is
// known to be zero off the faces toward finer leaves, then only over the
// blocks that have such a face
int run_exchange(spark_batch* b) {
  PlainLaunch L = make_launch(b, b->classes[0]);
  if (b->fc_clean) {
    if (b->n_corr == 0) return SPARK_OK;
    L.blist = b->corr_ids_d;
    L.nblk = b->n_corr;
  } else {
    L.blist = nullptr;
    L.nblk = b->nblk;
  }
  HIPCHK(spark::amr_flux_exchange(L, b->amr_g, b->fb, b->fc));
  b->fc_clean = true;
  return SPARK_OK;
}

// the exchange and hy_rk_correctFluxes on every block
// (Spark/Hydro.F90-mc:205-236), dt from the device
int run_amr_correct(spark_batch* b) {
  int rc;
  if ((rc = run_exchange(b))) return rc;
  if (b->fused && !b->fused64) {
    // FP32-flux form: only the cells on a face toward a finer leaf (their
    // whole x, y, z sequence).  Everywhere else the correction is zero and
    // the reference's pass is the identity up to one FP64 rounding of the
    // primitive -> conserved -> primitive round trip, which the FP64 forms
    // reproduce and this form leaves out.
    if (b->n_corr > 0) {
      PlainLaunch Lc = make_launch(b, b->classes[0]);
      Lc.blist = b->corr_ids_d;
      Lc.nblk = b->n_corr;
      for (int d = 0; d < 3; d++) Lc.fb[d] = b->fc[d];
      HIPCHK(spark::plain_correct(Lc, b->U, b->amr_fmask_d));
    }
    return SPARK_OK;
  }
  for (auto& c : b->classes) {
    PlainLaunch Lc = make_launch(b, c);
    for (int d = 0; d < 3; d++) Lc.fb[d] = b->fc[d];
    HIPCHK(spark::plain_correct(Lc, b->U));
  }
  return SPARK_OK;
}

// first-stage guard reach: stencil extent of stage 1 beyond the interior
int stage1_reach(const spark_batch* b) {
  // the remapping batched forms only need the layers shockDetect reads: one
  // for the interior, two when the hybrid solver reads SHOK one layer out
  if (fused_remap(b)) return b->p.use_hybrid_riemann ? 2 : 1;
  return (b->p.guard_mode == SPARK_GUARD_REFILL ? 0 : b->limits_array[0]) + b->p.nstencil;
}

int run_step_begin(spark_batch* b) {
  // SHOK over the cells the stages' hybrid Riemann solves read: the stage-1
  // region plus one layer (the reference computes NGUARD-1 layers; the
  // outer layers are never read before the next refill).  Without the
  // hybrid solver no form reads SHOK, so it is computed for the interior,
  // where it is part of the solution; guard-layer SHOK, which only the
  // hybrid solver reads, is left as it was.  The batched forms read SHOK
  // from the block's own box, so remapping blocks (no guard cells of their
  // own in the stages) get it one layer out.
  const bool hyb = b->p.use_hybrid_riemann != 0;
  const int band = !hyb ? 0 : (b->p.guard_mode == SPARK_GUARD_REFILL ? 0 : b->limits_array[0]) + 1;
  if (amr_remap_on(b)) {
    // a hierarchy has one block shape
    PlainLaunch L = make_launch(b, b->classes[0]);
    if (b->n_tele > 0) {
      L.blist = b->tele_ids_d;
      L.nblk = b->n_tele;
      HIPCHK(spark::plain_shock_detect(L, b->U, b->flat, band));
    }
    if (b->amr_remap_count > 0) {
      L.blist = b->remap_ids_d;
      L.nblk = b->amr_remap_count;
      HIPCHK(spark::plain_shock_detect(L, b->U, b->flat, hyb ? 1 : 0, b->nbr_d));
    }
  } else {
    for (auto& c : b->classes) {
      PlainLaunch L = make_launch(b, c);
      HIPCHK(spark::plain_shock_detect(L, b->U, b->flat, band));
    }
  }
  b->step_open = true;
  return SPARK_OK;
}

// stage input/output buffers: U -> S -> S2 -> U (RK3) or U -> S -> U
// (RK2).  A stage reads its input around each cell while it writes its
// output (the plain form's last sweep applies the update, the fused form's
// tiles read neighbours' cells), so input and output never alias.
const double* stage_in(const spark_batch* b, int stage) {
  if (stage == 1) return b->U;
  return stage == 2 ? b->S : b->S2;
}
double* stage_out(spark_batch* b, int stage) {
  if (stage == b->last_stage) return b->U;
  return stage == 1 ? b->S : b->S2;
}

int stage_ngoff(const spark_batch* b, int stage) {
  return b->p.guard_mode == SPARK_GUARD_REFILL ? 0 : b->limits_array[stage - 1];
}

// stage region of one block in the batched forms: blocks that read all
// their halos from same-level neighbours update only their interior (the
// region of a block that reads some is built by partial_boxes())
int blk_ngoff(const spark_batch* b, int blk, int stage) {
  if (amr_remap_on(b) && b->amr_remap_blk[blk] == 1) return 0;
  return stage_ngoff(b, stage);
}

int fused_plan(spark_batch* b);
int fused_tune(spark_batch* b);

size_t fused_lds(const spark_batch* b, int tx, int ty) {
  return b->fused64 ? spark::fused_lds_bytes_f64(b->p.nscalars, tx, ty)
                    : spark::fused_lds_bytes_f32(b->p.nscalars, tx, ty);
}

// persistent grid for a given LDS size: resident workgroups of the whole
// device, a multiple of the XCD count
int fused_grid_for(const spark_batch* b, size_t lds) {
  int occ = 0;
  hipError_t e = b->fused64 ? spark::fused_occupancy_f64(b->p.nscalars, lds, &occ)
                            : spark::fused_occupancy_f32(b->p.nscalars, lds, &occ);
  if (e != hipSuccess || occ < 1) occ = 1;
  return std::max(spark::FUSED_XCDS, (b->cus * occ) / spark::FUSED_XCDS * spark::FUSED_XCDS);
}

hipError_t fused_launch(const spark_batch* b, const spark::FusedArgs& A, int grid, size_t lds) {
  hipError_t e = b->fused64 ? spark::fused_reset_f64(g_stream, A.counters) : spark::fused_reset_f32(g_stream, A.counters);
  if (e != hipSuccess) return e;
  return b->fused64 ? spark::fused_stage_f64(A, grid, lds, g_stream) : spark::fused_stage_f32(A, grid, lds, g_stream);
}

// the batched forms on a uniform grid in refill mode read halo cells from
// the neighbouring blocks, so no guard fill runs between stages
bool fused_remap(const spark_batch* b) {
  return b->fused && b->has_ug && b->p.guard_mode == SPARK_GUARD_REFILL;
}

spark::FusedArgs fused_args(spark_batch* b, int stage) {
  spark::FusedArgs A;
  memset(&A, 0, sizeof(A));
  A.P = b->dp;
  A.meta = b->meta_d;
  A.items = b->plan[stage - 1].items_d;
  A.qoff = b->plan[stage - 1].qoff_d;
  A.qlen = b->plan[stage - 1].qlen_d;
  A.counters = b->counters + (stage - 1) * spark::FUSED_XCDS;
  A.U0 = b->U;
  A.Sin = stage_in(b, stage);
  A.Sout = stage_out(b, stage);
  for (int d = 0; d < 3; d++) A.fb[d] = b->p.flux_correct ? b->fb[d] : nullptr;
  A.nbr = (fused_remap(b) || amr_remap_on(b)) ? b->nbr_d : nullptr;
  A.weight = b->weights[stage - 1];
  A.add_flux = stage > 1;
  A.dtp = b->dtp;
  A.c1 = b->coeff[stage - 1][0];
  A.c2 = b->coeff[stage - 1][1];
  A.c3 = b->coeff[stage - 1][2];
  A.final_stage = stage == b->last_stage;
  A.err = b->err;
  return A;
}

// saveState's copy is folded away: stage 1 reads the solution U directly
// (hy_starState == Uin at stage 1) and the last stage writes U directly
// (updateState copies exactly the last stage's region, the interior).
int run_stage(spark_batch* b, int stage, int ngoff) {
  const double* co = b->coeff[stage - 1];
  if (stage != b->last_stage) b->star = stage_out(b, stage);
  if (b->fused) {
    if (ngoff != stage_ngoff(b, stage))
      return fail(SPARK_ERR_ARG, "batched form: stage %d runs with ngoff %d", stage, stage_ngoff(b, stage));
    int rc;
    if (!b->planned && (rc = fused_plan(b))) return rc;
    spark::FusedArgs A = fused_args(b, stage);
    const auto& pl = b->plan[stage - 1];
    double cells = 0.;
    for (const auto& it : pl.items) cells += (double)it.tx * it.ty * (it.z1 - it.z0 + 1);
    // algorithmic traffic: read U0 and the stage input (5+ns conserved-side
    // fields each), write the stage output (8+ns); flops as the plain form
    b->prof.begin(g_stream, "k_fused_stage", cells * 8.0 * (2 * (5 + b->p.nscalars) + 8 + b->p.nscalars),
                  cells * (3.0 * ((7 + b->p.nscalars) * 117.0 + 175.0) + 150.0));
    HIPCHK(fused_launch(b, A, b->stage_grid[stage - 1], b->stage_lds[stage - 1]));
    b->prof.end(g_stream);
    return SPARK_OK;
  }
  for (auto& c : b->classes) {
    PlainLaunch L = make_launch(b, c);
    L.weight = b->weights[stage - 1];
    L.add_flux = stage > 1;
    L.Sin = stage_in(b, stage);
    L.Sout = stage_out(b, stage);
    HIPCHK(spark::plain_stage(L, ngoff, co[0], co[1], co[2], stage == b->last_stage));
  }
  return SPARK_OK;
}

int run_step_end(spark_batch* b) {
  b->step_open = false;
  return SPARK_OK;
}

int run_full_step(spark_batch* b) {
  int rc;
  if ((rc = run_step_begin(b))) return rc;
  const bool refill = b->p.guard_mode == SPARK_GUARD_REFILL;
  for (int s = 1; s <= b->last_stage; s++) {
    if ((rc = run_stage(b, s, refill ? 0 : b->limits_array[s - 1]))) return rc;
    if (refill && s < b->last_stage && !fused_remap(b))
      if ((rc = run_fill(b, stage_out(b, s), false, b->p.nstencil))) return rc;
  }
  return run_step_end(b);
}

void free_batch(spark_batch* b) {
  if (!b) return;
  if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  if (b->nbr_d) (void)hipFree(b->nbr_d);
  if (b->remap_ids_d) (void)hipFree(b->remap_ids_d);
  if (b->tele_ids_d) (void)hipFree(b->tele_ids_d);
  if (b->amr_smask_d) (void)hipFree(b->amr_smask_d);
  if (b->dtctl_d) (void)hipFree(b->dtctl_d);
  if (b->geo_d) (void)hipFree(b->geo_d);
  if (b->amr_fmask_d) (void)hipFree(b->amr_fmask_d);
  if (b->corr_ids_d) (void)hipFree(b->corr_ids_d);
  if (b->amr_map_d) (void)hipFree(b->amr_map_d);
  if (b->amr_org_d) (void)hipFree(b->amr_org_d);
  for (int d = 0; d < 3; d++)
    if (b->fc[d]) (void)hipFree(b->fc[d]);
  for (auto& pl : b->plan) {
    if (pl.items_d) (void)hipFree(pl.items_d);
    if (pl.qoff_d) (void)hipFree(pl.qoff_d);
    if (pl.qlen_d) (void)hipFree(pl.qlen_d);
  }
  if (b->S2) (void)hipFree(b->S2);
  if (b->counters) (void)hipFree(b->counters);
  if (b->tune_err) (void)hipFree(b->tune_err);
  if (b->tune_dt) (void)hipFree(b->tune_dt);
  if (b->hist_slot) (void)hipFree(b->hist_slot);
  for (auto& r : b->prof.recs) {
    if (r.e0) (void)hipEventDestroy(r.e0);
    if (r.e1) (void)hipEventDestroy(r.e1);
  }
  auto F = [](void* p) { if (p) (void)hipFree(p); };
  F(b->meta_d);
  for (auto& c : b->classes) F(c.ids_d);
  F(b->U); F(b->S); F(b->flat);
  F(b->div);
  for (int d = 0; d < 3; d++) F(b->fb[d]);
  F(b->dtp); F(b->loc); F(b->dt_hist); F(b->dtparts); F(b->err);
  delete b;
}

// ---------------------------------------------------------------------------
// batched form: work items and tile autotuning

// candidate (tx, ty) tiles of an ex x ey region: near-equal splits whose
// columns fit the workgroup (tx * ty <= threads) and whose LDS fits
std::vector<std::pair<int, int>> tile_candidates(const spark_batch* b, int ex, int ey) {
  std::vector<std::pair<int, int>> out;
  std::vector<int> xs, ys;
  for (int parts = 1; parts <= 8; parts++) {
    xs.push_back((ex + parts - 1) / parts);
    ys.push_back((ey + parts - 1) / parts);
  }
  std::sort(xs.begin(), xs.end());
  xs.erase(std::unique(xs.begin(), xs.end()), xs.end());
  std::sort(ys.begin(), ys.end());
  ys.erase(std::unique(ys.begin(), ys.end()), ys.end());
  for (int tx : xs)
    for (int ty : ys) {
      if (tx < 2 || ty < 2) continue;
      if (tx * ty > spark::FUSED_THREADS) continue;
      if (fused_lds(b, tx, ty) > b->lds_limit) continue;
      out.push_back({tx, ty});
    }
  // rank by work items per region, then by the halo cells a tile stages
  auto tiles = [&](const std::pair<int, int>& t) {
    return ((ex + t.first - 1) / t.first) * ((ey + t.second - 1) / t.second);
  };
  std::sort(out.begin(), out.end(), [&](const std::pair<int, int>& a, const std::pair<int, int>& c) {
    const int ta = tiles(a), tc = tiles(c);
    if (ta != tc) return ta < tc;
    return (a.first + 6) * (a.second + 6) < (c.first + 6) * (c.second + 6);
  });
  if (out.size() > 6) out.resize(6);
  return out;
}

using spark::Box;

// the work items of a box: (tx x ty) column tiles marched over its z range
void box_items(int blk, const Box& r, int tx, int ty, std::vector<spark::FusedItem>& out) {
  for (int y0 = r.y0; y0 <= r.y1; y0 += ty)
    for (int x0 = r.x0; x0 <= r.x1; x0 += tx) {
      spark::FusedItem it;
      memset(&it, 0, sizeof(it));
      it.blk = blk;
      it.x0 = x0; it.y0 = y0;
      it.tx = std::min(tx, r.x1 - x0 + 1);
      it.ty = std::min(ty, r.y1 - y0 + 1);
      it.z0 = r.z0; it.z1 = r.z1;
      out.push_back(it);
    }
}

Box region_box(const BlockMeta& m, int g, int ngoff) {
  return Box{g - ngoff, g + m.nx - 1 + ngoff, g - ngoff, g + m.ny - 1 + ngoff, g - ngoff, g + m.nz - 1 + ngoff};
}

// stage region of a block that reads some guard sectors from same-level
// leaves (sector_boxes over its computed-sector mask)
void partial_boxes(const spark_batch* b, int blk, int e, std::vector<Box>& out) {
  const BlockMeta& m = b->meta_h[blk];
  spark::sector_boxes(m.nx, m.ny, m.nz, b->p.nguard, e, b->amr_local[blk], out);
}

// build the per-XCD queues of one stage from the chosen tiles.  A block's
// items stay in one queue; blocks go to the least-loaded queue, largest first.
// force_ngoff >= 0 (tuning): every block over that region.
int build_stage_plan(spark_batch* b, int stage, const std::vector<int>* only_blocks, int force_ngoff) {
  std::vector<int> blocks;
  if (only_blocks) blocks = *only_blocks;
  else for (int i = 0; i < b->nblk; i++) blocks.push_back(i);
  std::vector<std::pair<double, int>> cost;
  std::map<int, int> cls_of;
  for (size_t ci = 0; ci < b->classes.size(); ci++)
    for (int id : b->classes[ci].ids) cls_of[id] = (int)ci;
  std::vector<int> ngoff_of(b->nblk, 0);
  std::map<int, std::vector<Box>> boxes;  // blocks that read some sectors from same-level leaves
  for (int id : blocks) {
    const BlockMeta& m = b->meta_h[id];
    if (force_ngoff < 0 && amr_remap_on(b) && b->amr_remap_blk[id] == 2) {
      auto& bx = boxes[id];
      partial_boxes(b, id, stage_ngoff(b, stage), bx);
      double c = 0.;
      for (const auto& r : bx) c += r.cost();
      cost.push_back({c, id});
      continue;
    }
    const int ngoff = force_ngoff >= 0 ? force_ngoff : blk_ngoff(b, id, stage);
    ngoff_of[id] = ngoff;
    cost.push_back({region_box(m, b->p.nguard, ngoff).cost(), id});
  }
  std::sort(cost.begin(), cost.end(), [](const std::pair<double, int>& a, const std::pair<double, int>& c) {
    return a.first > c.first || (a.first == c.first && a.second < c.second);
  });
  std::vector<std::vector<spark::FusedItem>> q(spark::FUSED_XCDS);
  std::vector<double> load(spark::FUSED_XCDS, 0.);
  for (auto& pr : cost) {
    const int id = pr.second;
    const BlockMeta& m = b->meta_h[id];
    const int qi = (int)(std::min_element(load.begin(), load.end()) - load.begin());
    load[qi] += pr.first;
    auto bx = boxes.find(id);
    if (bx != boxes.end()) {
      for (const auto& r : bx->second) {
        const int w = r.x1 - r.x0 + 1, h = r.y1 - r.y0 + 1;
        // the interior columns take the tile tuned for interior regions
        auto tt = b->tiles.find(std::make_pair(cls_of[id], 0));
        std::pair<int, int> t;
        if (w == m.nx && h == m.ny && tt != b->tiles.end()) {
          t = tt->second;
        } else {
          auto cands = tile_candidates(b, w, h);
          if (cands.empty()) return fail(SPARK_ERR_ARG, "no tile fits a %d x %d region", w, h);
          t = cands[0];
        }
        box_items(id, r, t.first, t.second, q[qi]);
      }
      continue;
    }
    const int ngoff = ngoff_of[id];
    const auto key = std::make_pair(cls_of[id], ngoff);
    auto tt = b->tiles.find(key);
    if (tt == b->tiles.end()) return fail(SPARK_ERR_STATE, "no tile chosen for class %d ngoff %d", key.first, ngoff);
    box_items(id, region_box(m, b->p.nguard, ngoff), tt->second.first, tt->second.second, q[qi]);
  }
  auto& pl = b->plan[stage - 1];
  pl.items.clear();
  int qoff[spark::FUSED_XCDS], qlen[spark::FUSED_XCDS];
  for (int i = 0; i < spark::FUSED_XCDS; i++) {
    qoff[i] = (int)pl.items.size();
    qlen[i] = (int)q[i].size();
    pl.items.insert(pl.items.end(), q[i].begin(), q[i].end());
  }
  if (pl.items_d) (void)hipFree(pl.items_d);
  if (!pl.qoff_d) HIPCHK(hipMalloc((void**)&pl.qoff_d, sizeof(int) * spark::FUSED_XCDS));
  if (!pl.qlen_d) HIPCHK(hipMalloc((void**)&pl.qlen_d, sizeof(int) * spark::FUSED_XCDS));
  HIPCHK(hipMalloc((void**)&pl.items_d, sizeof(spark::FusedItem) * std::max<size_t>(1, pl.items.size())));
  HIPCHK(hipMemcpy(pl.items_d, pl.items.data(), sizeof(spark::FusedItem) * pl.items.size(), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(pl.qoff_d, qoff, sizeof(qoff), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(pl.qlen_d, qlen, sizeof(qlen), hipMemcpyHostToDevice));
  size_t lds = 0;
  for (const auto& it : pl.items) lds = std::max(lds, fused_lds(b, it.tx, it.ty));
  b->stage_lds[stage - 1] = lds;
  b->stage_grid[stage - 1] = fused_grid_for(b, b->stage_lds[stage - 1]);
  return SPARK_OK;
}

// Autotuner: for every (shape class, stage region) time the candidate
// tiles on the batch's own blocks of that class and keep the fastest.
// Runs once, on the first step, with the real data packed.
int fused_tune(spark_batch* b) {
  std::vector<int> ngoffs;
  for (int s = 1; s <= b->last_stage; s++)
    if (std::find(ngoffs.begin(), ngoffs.end(), stage_ngoff(b, s)) == ngoffs.end()) ngoffs.push_back(stage_ngoff(b, s));
  if (amr_remap_on(b) && std::find(ngoffs.begin(), ngoffs.end(), 0) == ngoffs.end()) ngoffs.push_back(0);
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  for (size_t ci = 0; ci < b->classes.size(); ci++) {
    const ShapeClass& c = b->classes[ci];
    // time on the whole class: a sample smaller than the device would favour
    // tiles that make more, smaller work items
    std::vector<int> sample(c.ids.begin(), c.ids.end());
    for (int ngoff : ngoffs) {
      auto cands = tile_candidates(b, c.nx + 2 * ngoff, c.ny + 2 * ngoff);
      if (cands.empty()) return fail(SPARK_ERR_ARG, "no tile fits a %d x %d region", c.nx + 2 * ngoff, c.ny + 2 * ngoff);
      const auto key = std::make_pair((int)ci, ngoff);
      if (!b->autotune || cands.size() == 1) {
        b->tiles[key] = cands[0];
        continue;
      }
      // the stage that uses this region, for its coefficients
      int stage = 1;
      for (int s = 1; s <= b->last_stage; s++)
        if (stage_ngoff(b, s) == ngoff) { stage = s; break; }
      float best = 1e30f;
      std::pair<int, int> pick = cands[0];
      for (auto& cand : cands) {
        b->tiles[key] = cand;
        int rc = build_stage_plan(b, stage, &sample, ngoff);
        if (rc) return rc;
        spark::FusedArgs A = fused_args(b, stage);
        A.Sin = b->U;                     // read the packed solution,
        A.Sout = b->S2 ? b->S2 : b->S;    // write a scratch stage buffer
        for (int d = 0; d < 3; d++) A.fb[d] = nullptr;
        A.err = b->tune_err;
        A.dtp = b->tune_dt;
        float t_best = 1e30f;
        for (int rep = 0; rep < 3; rep++) {
          HIPCHK(hipEventRecord(e0, g_stream));
          HIPCHK(fused_launch(b, A, b->stage_grid[stage - 1], b->stage_lds[stage - 1]));
          HIPCHK(hipEventRecord(e1, g_stream));
          HIPCHK(hipEventSynchronize(e1));
          float ms = 0.f;
          HIPCHK(hipEventElapsedTime(&ms, e0, e1));
          if (rep > 0) t_best = std::min(t_best, ms);
        }
        if (t_best < best) { best = t_best; pick = cand; }
      }
      b->tiles[key] = pick;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return SPARK_OK;
}

int fused_plan(spark_batch* b) {
  int rc;
  if ((rc = fused_tune(b))) return rc;
  for (int s = 1; s <= b->last_stage; s++)
    if ((rc = build_stage_plan(b, s, nullptr, -1))) return rc;
  b->planned = true;
  return SPARK_OK;
}

void drop_graph(spark_batch* b) {
  if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->gexec = nullptr;
  b->graph = nullptr;
}

}  // namespace

extern "C" {

int spark_params_default(spark_params* p) {
  if (!p) return fail(SPARK_ERR_ARG, "spark_params_default: null pointer");
  memset(p, 0, sizeof(*p));
  p->gamma = 1.6667;
  p->cfl = 0.8;
  p->cvisc = 0.1;
  p->smlrho = 1e-10;
  p->smallp = 1e-10;
  p->smalle = 1e-10;
  p->smallx = 1e-10;
  p->eintSwitch = 0.0;
  p->tiny = 1.e-32;
  p->gasConst = 8.3144598E7;
  p->abar = 1.0;
  p->ndim = 3;
  p->nguard = 6;
  p->nstencil = 3;
  p->rk_order = 2;
  p->recon = SPARK_RECON_WENO5;
  p->riemann = SPARK_RIEMANN_HLLC;
  p->use_flattening = 1;
  p->use_hybrid_riemann = 0;
  p->nscalars = 0;
  p->use_gravity = 0;
  p->flux_correct = 0;
  p->form = SPARK_FORM_PLAIN_FP64;
  p->guard_mode = SPARK_GUARD_TELESCOPE;
  p->nunk_vars = SPARK_NPROP;
  for (int v = 0; v < SPARK_NPROP + SPARK_MAX_SCALARS; v++) p->var_index[v] = v;
  p->nspecies = 0;
  p->eos = SPARK_EOS_GAMMA;
  for (int s = 0; s < SPARK_MAX_SCALARS; s++) {
    p->species_a[s] = 1.0;
    p->species_gamma[s] = p->gamma;
  }
  return SPARK_OK;
}

int spark_init(int device) {
  int n = 0;
  HIPCHK(hipGetDeviceCount(&n));
  if (device < 0 || device >= n) return fail(SPARK_ERR_ARG, "spark_init: device %d of %d", device, n);
  HIPCHK(hipSetDevice(device));
  if (g_stream == nullptr || g_device != device) {
    if (g_stream) (void)hipStreamDestroy(g_stream);
    HIPCHK(hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking));
  }
  g_device = device;
  if (device < kMaxDevices && g_noalloc[device] == nullptr) {
    void* ptr = nullptr;
    HIPCHK(spark::noalloc_scratch_address(&ptr));
    if (ptr == nullptr) return fail(SPARK_ERR_HIP, "spark_init: no device address for the staging area");
    g_noalloc[device] = ptr;
  }
  return SPARK_OK;
}

int spark_finalize(void) {
  if (g_stream) {
    HIPCHK(hipStreamSynchronize(g_stream));
    HIPCHK(hipStreamDestroy(g_stream));
  }
  g_stream = nullptr;
  g_device = -1;
  return SPARK_OK;
}

int spark_batch_create(const spark_params* p, int nblocks, const spark_block_desc* blocks, spark_batch** out) {
  if (!p || !blocks || !out) return fail(SPARK_ERR_ARG, "spark_batch_create: null argument");
  *out = nullptr;
  if (g_stream == nullptr) return fail(SPARK_ERR_STATE, "spark_batch_create: call spark_init first");
  if (nblocks <= 0) return fail(SPARK_ERR_ARG, "spark_batch_create: nblocks = %d", nblocks);
  if (p->ndim < 1 || p->ndim > 3) return fail(SPARK_ERR_ARG, "ndim = %d", p->ndim);
  if (p->nscalars < 0 || p->nscalars > SPARK_MAX_SCALARS)
    return fail(SPARK_ERR_ARG, "nscalars = %d (0..%d)", p->nscalars, SPARK_MAX_SCALARS);
  if (p->rk_order != 2 && p->rk_order != 3) return fail(SPARK_ERR_ARG, "rk_order = %d", p->rk_order);
  if (p->recon < 0 || p->recon > 3) return fail(SPARK_ERR_ARG, "recon = %d", p->recon);
  if (p->riemann < 0 || p->riemann > 1) return fail(SPARK_ERR_ARG, "riemann = %d", p->riemann);
  if (p->form < SPARK_FORM_PLAIN_FP64 || p->form > SPARK_FORM_BATCHED_FP64)
    return fail(SPARK_ERR_ARG, "form = %d", p->form);
  const int radius = (p->recon == SPARK_RECON_PLM) ? 1 : (p->recon == SPARK_RECON_FOG) ? 0 : 2;
  if (p->nstencil < radius + 1) return fail(SPARK_ERR_ARG, "nstencil = %d too small for the reconstruction", p->nstencil);
  if (p->nguard < required_guard(*p))
    return fail(SPARK_ERR_ARG, "nguard = %d, this RK order / guard mode needs >= %d", p->nguard, required_guard(*p));
  const int nvar = SPARK_NPROP + p->nscalars;
  if (p->nunk_vars < nvar) return fail(SPARK_ERR_ARG, "nunk_vars = %d < %d", p->nunk_vars, nvar);
  for (int v = 0; v < nvar; v++)
    if (p->var_index[v] < 0 || p->var_index[v] >= p->nunk_vars)
      return fail(SPARK_ERR_ARG, "var_index[%d] = %d out of range", v, p->var_index[v]);
  if (p->nspecies < 0 || p->nspecies > p->nscalars)
    return fail(SPARK_ERR_ARG, "nspecies = %d (0..nscalars = %d)", p->nspecies, p->nscalars);
  if (p->eos != SPARK_EOS_GAMMA && p->eos != SPARK_EOS_MULTIGAMMA) return fail(SPARK_ERR_ARG, "eos = %d", p->eos);
  if (p->eos == SPARK_EOS_MULTIGAMMA) {
    if (p->nspecies < 1) return fail(SPARK_ERR_ARG, "the Multigamma EOS needs nspecies >= 1");
    for (int s = 0; s < p->nspecies; s++)
      if (!(p->species_a[s] > 0.0) || !(p->species_gamma[s] > 1.0))
        return fail(SPARK_ERR_ARG, "species %d: A = %g, gamma = %g", s, p->species_a[s], p->species_gamma[s]);
  }
  if (p->form == SPARK_FORM_FP32_FLUX || p->form == SPARK_FORM_BATCHED_FP64) {
    if (p->ndim != 3) return fail(SPARK_ERR_ARG, "the batched form is 3D only (ndim = %d)", p->ndim);
    // species are clamped and renormalised together, within the first
    // chunk of scalars the fused kernel stages
    if (p->nspecies > spark::FUSED_SCALAR_CHUNK)
      return fail(SPARK_ERR_ARG, "the batched forms take at most %d species (nspecies = %d)",
                  spark::FUSED_SCALAR_CHUNK, p->nspecies);
    // the FP32 reconstruction takes gamc = gamma and so needs the gamma law
    if (p->form == SPARK_FORM_FP32_FLUX && p->eos != SPARK_EOS_GAMMA)
      return fail(SPARK_ERR_ARG, "the FP32-flux form implements the gamma law only (eos = %d)", p->eos);
  }

  spark_batch* b = new spark_batch();
  b->p = *p;
  b->nblk = nblocks;
  b->nvar = nvar;
  b->desc.assign(blocks, blocks + nblocks);
  rk_tables(b->p, &b->last_stage, b->coeff, b->limits_array, b->weights);

  DevParams& d = b->dp;
  d.gamma = p->gamma; d.cvisc = p->cvisc; d.smlrho = p->smlrho; d.smallp = p->smallp; d.smalle = p->smalle;
  d.smallx = p->smallx; d.eintSwitch = p->eintSwitch; d.tiny = p->tiny; d.cfl = p->cfl;
  {
    // Gamma/Eos_init.F90:86 and Gamma/Eos.F90:177,240-241
    const double eos_gammam1 = 1.0 / (p->gamma - 1.0);
    const double ggprod = eos_gammam1 * p->gasConst;
    d.ggprodinv = 1. / ggprod;
    d.gam1inv = 1. / eos_gammam1;
  }
  d.abar = p->abar;
  d.grav[0] = p->grav[0]; d.grav[1] = p->grav[1]; d.grav[2] = p->grav[2];
  d.nspec = p->nspecies;
  d.eos = p->eos;
  d.gasConst = p->gasConst;
  for (int s = 0; s < SPARK_MAX_SCALARS; s++) {
    // Multigamma/eos_initMgamma.F90:15 and the 1/A of Multispecies_getSumInv.F90:150
    d.mg_gm1[s] = s < p->nspecies ? 1. / (p->species_gamma[s] - 1.) : 0.0;
    d.mg_ainv[s] = s < p->nspecies ? 1.e0 / p->species_a[s] : 0.0;
  }
  d.ndim = p->ndim; d.ng = p->nguard; d.ns = p->nscalars;
  d.use_grav = p->use_gravity; d.use_flat = p->use_flattening; d.hybrid = p->use_hybrid_riemann;
  d.recon = p->recon; d.riemann = p->riemann;
  d.geom = 0; d.geo_nx = 0; d.geo_ny = 0; d.geo = nullptr;  // Cartesian until spark_set_geometry

  const int g = p->nguard;
  const int nf = 5 + p->nscalars;
  std::map<std::tuple<int, int, int>, int> cls;
  int64_t max_int = 0;
  for (int i = 0; i < nblocks; i++) {
    const spark_block_desc& bd = blocks[i];
    int n[3] = {bd.nx, p->ndim > 1 ? bd.ny : 1, p->ndim > 2 ? bd.nz : 1};
    for (int q = 0; q < p->ndim; q++)
      if (n[q] < 1) { free_batch(b); return fail(SPARK_ERR_ARG, "block %d has an empty dimension", i); }
    BlockMeta m;
    memset(&m, 0, sizeof(m));
    m.nx = n[0]; m.ny = n[1]; m.nz = n[2];
    m.NX = n[0] + 2 * g;
    m.NY = p->ndim > 1 ? n[1] + 2 * g : 1;
    m.NZ = p->ndim > 2 ? n[2] + 2 * g : 1;
    m.dx = bd.dx; m.dy = bd.dy; m.dz = bd.dz;
    m.level = bd.level; m.id = i;
    const int64_t ncell = (int64_t)m.NX * m.NY * m.NZ;
    m.cstride = round_up(ncell, 64);
    m.base = b->pool;
    b->pool += (int64_t)nvar * m.cstride;
    m.obase = b->opool;
    b->opool += m.cstride;
    m.dbase = b->dpool;
    b->dpool += (int64_t)nf * m.cstride;
    int64_t reg = 1;
    for (int q = 0; q < 3; q++) reg *= n[q] + (q < p->ndim ? 1 : 0);
    m.bbase = b->bpool;
    b->bpool += (int64_t)nf * round_up(reg, 64);
    b->max_cells = std::max<int64_t>(b->max_cells, ncell);
    max_int = std::max<int64_t>(max_int, (int64_t)n[0] * n[1] * n[2]);
    b->meta_h.push_back(m);
    auto key = std::make_tuple(n[0], n[1], n[2]);
    auto it = cls.find(key);
    if (it == cls.end()) {
      cls[key] = (int)b->classes.size();
      ShapeClass c;
      c.nx = n[0]; c.ny = n[1]; c.nz = n[2];
      b->classes.push_back(c);
      it = cls.find(key);
    }
    b->classes[it->second].ids.push_back(i);
  }
  // CFL partials: a block's cells split into chunks (one workgroup each)
  // only until the batch has about 4096 workgroups; fewer partials keep the
  // one-workgroup final reduction short
  b->dtchunks = (int)std::min<int64_t>({16, std::max<int64_t>(1, (max_int + 255) / 256),
                                        std::max<int64_t>(1, (4096 + b->nblk - 1) / b->nblk)});

#define ALLOC(ptr, bytes)                                                              \
  do {                                                                                 \
    hipError_t e_ = hipMalloc((void**)&(ptr), (bytes));                                \
    if (e_ != hipSuccess) {                                                            \
      free_batch(b);                                                                   \
      return fail(SPARK_ERR_HIP, "hipMalloc(%zu bytes): %s", (size_t)(bytes), hipGetErrorString(e_)); \
    }                                                                                  \
  } while (0)
  ALLOC(b->meta_d, sizeof(BlockMeta) * nblocks);
  for (auto& c : b->classes) ALLOC(c.ids_d, sizeof(int) * c.ids.size());
  ALLOC(b->U, sizeof(double) * b->pool);
  ALLOC(b->S, sizeof(double) * b->pool);
  ALLOC(b->flat, sizeof(double) * b->opool);
  ALLOC(b->S2, sizeof(double) * b->pool);
  // the plain form's divergence terms (the sweeps before the last write
  // them; 1D has no earlier sweep)
  const bool plain_form = p->form != SPARK_FORM_FP32_FLUX && p->form != SPARK_FORM_BATCHED_FP64;
  if (plain_form && p->ndim > 1) ALLOC(b->div, sizeof(double) * b->dpool);
  if (p->flux_correct) {
    for (int q = 0; q < 3; q++) ALLOC(b->fb[q], sizeof(double) * b->bpool);
    for (int q = 0; q < 3; q++) ALLOC(b->fc[q], sizeof(double) * b->bpool);
    // only boundary faces are ever accumulated: the interior ones read 0;
    // corrections start at 0 everywhere
    for (int q = 0; q < 3; q++) {
      hipError_t e_ = hipMemset(b->fb[q], 0, sizeof(double) * b->bpool);
      if (e_ == hipSuccess) e_ = hipMemset(b->fc[q], 0, sizeof(double) * b->bpool);
      if (e_ != hipSuccess) {
        free_batch(b);
        return fail(SPARK_ERR_HIP, "hipMemset: %s", hipGetErrorString(e_));
      }
    }
    b->fc_clean = true;
  }
  ALLOC(b->dtp, sizeof(double));
  ALLOC(b->loc, sizeof(long long) * 2);
  b->hist_cap = 4096;
  ALLOC(b->dt_hist, sizeof(double) * b->hist_cap);
  ALLOC(b->dtparts, spark::plain_dt_part_size() * (size_t)nblocks * b->dtchunks);
  ALLOC(b->err, sizeof(int));
  ALLOC(b->hist_slot, sizeof(int));
  if (p->form == SPARK_FORM_FP32_FLUX || p->form == SPARK_FORM_BATCHED_FP64) {
    b->fused = true;
    b->fused64 = p->form == SPARK_FORM_BATCHED_FP64;
    ALLOC(b->counters, sizeof(int) * 3 * spark::FUSED_XCDS);
    ALLOC(b->tune_err, sizeof(int));
    ALLOC(b->tune_dt, sizeof(double));
  }
#undef ALLOC
  HIPCHK(hipMemcpyAsync(b->meta_d, b->meta_h.data(), sizeof(BlockMeta) * nblocks, hipMemcpyHostToDevice, g_stream));
  for (auto& c : b->classes)
    HIPCHK(hipMemcpyAsync(c.ids_d, c.ids.data(), sizeof(int) * c.ids.size(), hipMemcpyHostToDevice, g_stream));
  HIPCHK(hipMemsetAsync(b->U, 0, sizeof(double) * b->pool, g_stream));
  HIPCHK(hipMemsetAsync(b->S, 0, sizeof(double) * b->pool, g_stream));
  HIPCHK(hipMemsetAsync(b->flat, 0, sizeof(double) * b->opool, g_stream));
  HIPCHK(hipMemsetAsync(b->S2, 0, sizeof(double) * b->pool, g_stream));
  if (b->div) HIPCHK(hipMemsetAsync(b->div, 0, sizeof(double) * b->dpool, g_stream));
  if (p->flux_correct)
    for (int q = 0; q < 3; q++) HIPCHK(hipMemsetAsync(b->fb[q], 0, sizeof(double) * b->bpool, g_stream));
  HIPCHK(hipMemsetAsync(b->err, 0, sizeof(int), g_stream));
  HIPCHK(hipMemsetAsync(b->hist_slot, 0, sizeof(int), g_stream));
  if (b->fused) {
    HIPCHK(hipMemsetAsync(b->counters, 0, sizeof(int) * 3 * spark::FUSED_XCDS, g_stream));
    HIPCHK(hipMemsetAsync(b->tune_err, 0, sizeof(int), g_stream));
    HIPCHK(hipMemsetAsync(b->tune_dt, 0, sizeof(double), g_stream));
    int lim = 0;
    HIPCHK(hipDeviceGetAttribute(&b->cus, hipDeviceAttributeMultiprocessorCount, g_device));
    HIPCHK(hipDeviceGetAttribute(&lim, hipDeviceAttributeMaxSharedMemoryPerBlock, g_device));
    // the dynamic LDS of a launch plus the kernel's own __shared__ variables
    const size_t own = b->fused64 ? spark::fused_static_lds_f64(b->p.nscalars) : spark::fused_static_lds_f32(b->p.nscalars);
    b->lds_limit = (size_t)lim > own + 16 ? (size_t)lim - own - 16 : 0;
    if (fused_lds(b, 4, 4) > b->lds_limit) {
      const size_t need = fused_lds(b, 4, 4), have = b->lds_limit;
      free_batch(b);
      return fail(SPARK_ERR_ARG, "batched form: %zu bytes of LDS needed, the device allows %zu", need, have);
    }
  }
  HIPCHK(hipMemsetAsync(b->loc, 0, sizeof(long long) * 2, g_stream));
  HIPCHK(hipStreamSynchronize(g_stream));
  *out = b;
  return SPARK_OK;
}

int spark_batch_destroy(spark_batch* b) {
  if (!b) return SPARK_OK;
  if (g_stream) HIPCHK(hipStreamSynchronize(g_stream));
  free_batch(b);
  return SPARK_OK;
}

// AoS <-> SoA transfer in rounds of blocks through a device staging area of
// `cap` doubles (`stage`), with device room for `maxb` block offsets
// (`off_d`) and the variable map (`vidx_d`); off_h is host room for maxb
// offsets.  Every round ends with the stream synchronised, so the staging
// area is free again when the next round writes it.
static int transfer_rounds(spark_batch* b, double* const* host, int to_device, int interior_only, double* pool,
                           double* stage, int64_t cap, int64_t* off_d, int* vidx_d, int maxb, int64_t* off_h) {
  const int nunk = b->p.nunk_vars;
  HIPCHK(hipMemcpy(vidx_d, b->p.var_index, sizeof(int) * b->nvar, hipMemcpyHostToDevice));
  int rc = SPARK_OK;
  int i0 = 0;
  while (i0 < b->nblk && rc == SPARK_OK) {
    int64_t used = 0;
    int i1 = i0;
    while (i1 < b->nblk && i1 - i0 < maxb) {
      const BlockMeta& m = b->meta_h[i1];
      const int64_t sz = (int64_t)nunk * m.NX * m.NY * m.NZ;
      if (used + sz > cap && i1 > i0) break;
      off_h[i1 - i0] = used;
      used += sz;
      i1++;
    }
    hipError_t e = hipMemcpy(off_d, off_h, sizeof(int64_t) * (i1 - i0), hipMemcpyHostToDevice);
    for (int i = i0; i < i1 && e == hipSuccess; i++) {
      const BlockMeta& m = b->meta_h[i];
      const int64_t sz = (int64_t)nunk * m.NX * m.NY * m.NZ;
      // unpack also stages the host block first, so unk slots outside the
      // variable map (and guards when interior_only) keep their host values
      e = hipMemcpy(stage + off_h[i - i0], host[i], sizeof(double) * sz, hipMemcpyHostToDevice);
    }
    if (e == hipSuccess) {
      if (to_device)
        e = spark::plain_aos_to_soa(g_stream, b->meta_d + i0, i1 - i0, (int)b->max_cells, stage, off_d, nunk, vidx_d,
                                    b->nvar, pool);
      else
        e = spark::plain_soa_to_aos(g_stream, b->dp, b->meta_d + i0, i1 - i0, (int)b->max_cells, pool, stage, off_d,
                                    nunk, vidx_d, b->nvar, interior_only);
    }
    if (e == hipSuccess) e = hipStreamSynchronize(g_stream);
    for (int i = i0; i < i1 && e == hipSuccess && !to_device; i++) {
      const BlockMeta& m = b->meta_h[i];
      const int64_t sz = (int64_t)nunk * m.NX * m.NY * m.NZ;
      e = hipMemcpy(host[i], stage + off_h[i - i0], sizeof(double) * sz, hipMemcpyDeviceToHost);
    }
    if (e != hipSuccess) rc = fail(SPARK_ERR_HIP, "pack/unpack: %s", hipGetErrorString(e));
    i0 = i1;
  }
  return rc;
}

static int transfer(spark_batch* b, double* const* host, int to_device, int interior_only, double* pool) {
  const int nunk = b->p.nunk_vars;
  // stage at most ~256 MiB of AoS data per round trip
  const int64_t cap = std::max<int64_t>((int64_t)nunk * b->max_cells, (int64_t)32 << 20);
  double* stage = nullptr;
  int64_t* off_d = nullptr;
  int* vidx_d = nullptr;
  HIPCHK(hipMalloc((void**)&stage, sizeof(double) * cap));
  HIPCHK(hipMalloc((void**)&off_d, sizeof(int64_t) * b->nblk));
  HIPCHK(hipMalloc((void**)&vidx_d, sizeof(int) * b->nvar));
  std::vector<int64_t> off_h(b->nblk);
  const int rc = transfer_rounds(b, host, to_device, interior_only, pool, stage, cap, off_d, vidx_d, b->nblk,
                                 off_h.data());
  (void)hipFree(stage);
  (void)hipFree(off_d);
  (void)hipFree(vidx_d);
  return rc;
}

int spark_pack(spark_batch* b, const double* const* unk_blocks) {
  if (!b || !unk_blocks) return fail(SPARK_ERR_ARG, "spark_pack: null argument");
  for (int i = 0; i < b->nblk; i++)
    if (!unk_blocks[i]) return fail(SPARK_ERR_ARG, "spark_pack: block %d pointer is null", i);
  return transfer(b, const_cast<double* const*>(unk_blocks), 1, 0, b->U);
}

int spark_pack_noalloc(spark_batch* b, const double* const* unk_blocks) {
  if (!b || !unk_blocks) return fail(SPARK_ERR_ARG, "spark_pack_noalloc: null argument");
  for (int i = 0; i < b->nblk; i++)
    if (!unk_blocks[i]) return fail(SPARK_ERR_ARG, "spark_pack_noalloc: block %d pointer is null", i);
  if (g_device < 0 || g_device >= kMaxDevices || g_noalloc[g_device] == nullptr)
    return fail(SPARK_ERR_STATE, "spark_pack_noalloc: no staging area on device %d (call spark_init)", g_device);
  const int64_t cap = (int64_t)((spark::NOALLOC_BYTES - spark::NOALLOC_STAGE) / sizeof(double));
  for (int i = 0; i < b->nblk; i++) {
    const BlockMeta& m = b->meta_h[i];
    const int64_t sz = (int64_t)b->p.nunk_vars * m.NX * m.NY * m.NZ;
    if (sz > cap)
      return fail(SPARK_ERR_ARG,
                  "spark_pack_noalloc: block %d holds %lld unk values, more than the %lld of the fixed %zu MiB "
                  "staging area (use spark_pack)", i, (long long)sz, (long long)cap, spark::NOALLOC_BYTES >> 20);
  }
  unsigned char* area = static_cast<unsigned char*>(g_noalloc[g_device]);
  int64_t off_h[spark::NOALLOC_MAXB];
  return transfer_rounds(b, const_cast<double* const*>(unk_blocks), 1, 0, b->U,
                         reinterpret_cast<double*>(area + spark::NOALLOC_STAGE), cap,
                         reinterpret_cast<int64_t*>(area + 256), reinterpret_cast<int*>(area), spark::NOALLOC_MAXB,
                         off_h);
}

int spark_unpack(spark_batch* b, double* const* unk_blocks, int interior_only, int which) {
  if (!b || !unk_blocks) return fail(SPARK_ERR_ARG, "spark_unpack: null argument");
  for (int i = 0; i < b->nblk; i++)
    if (!unk_blocks[i]) return fail(SPARK_ERR_ARG, "spark_unpack: block %d pointer is null", i);
  HIPCHK(hipStreamSynchronize(g_stream));
  return transfer(b, unk_blocks, 0, interior_only, which ? (b->star ? b->star : b->S) : b->U);
}

int spark_set_uniform_grid(spark_batch* b, int nbx, int nby, int nbz) {
  if (!b) return fail(SPARK_ERR_ARG, "spark_set_uniform_grid: null batch");
  if (nbx < 1 || nby < 1 || nbz < 1 || (int64_t)nbx * nby * nbz != b->nblk)
    return fail(SPARK_ERR_ARG, "uniform grid %d x %d x %d does not match %d blocks", nbx, nby, nbz, b->nblk);
  if (b->classes.size() != 1) return fail(SPARK_ERR_ARG, "uniform grid needs blocks of one shape");
  if (b->amr) return fail(SPARK_ERR_STATE, "spark_set_uniform_grid: the batch has a block hierarchy");
  if ((b->p.ndim < 2 && nby != 1) || (b->p.ndim < 3 && nbz != 1))
    return fail(SPARK_ERR_ARG, "uniform grid extent in an unused dimension");
  if (b->fused && b->p.guard_mode == SPARK_GUARD_REFILL) {
    const BlockMeta& m0 = b->meta_h[0];
    if (m0.nx < 3 || m0.ny < 3 || m0.nz < 3)
      return fail(SPARK_ERR_ARG, "batched form on a uniform grid needs blocks of >= 3 cells per side");
  }
  b->ug[0] = nbx; b->ug[1] = nby; b->ug[2] = nbz;
  // neighbour table (blocks numbered x fastest)
  std::vector<int64_t> nb((size_t)b->nblk * 27, spark::NBR_EDGE);
  for (int id = 0; id < b->nblk; id++) {
    const int bx = id % nbx, by = (id / nbx) % nby, bz = id / (nbx * nby);
    for (int oz = -1; oz <= 1; oz++)
      for (int oy = -1; oy <= 1; oy++)
        for (int ox = -1; ox <= 1; ox++) {
          const int x = bx + ox, y = by + oy, z = bz + oz;
          if (x < 0 || x >= nbx || y < 0 || y >= nby || z < 0 || z >= nbz) continue;
          nb[(size_t)id * 27 + 13 + ox + 3 * oy + 9 * oz] = b->meta_h[x + nbx * (y + nby * z)].base;
        }
  }
  if (!b->nbr_d) HIPCHK(hipMalloc((void**)&b->nbr_d, sizeof(int64_t) * nb.size()));
  HIPCHK(hipMemcpy(b->nbr_d, nb.data(), sizeof(int64_t) * nb.size(), hipMemcpyHostToDevice));
  b->has_ug = true;
  drop_graph(b);
  return SPARK_OK;
}

int spark_fill_guardcells(spark_batch* b) {
  if (!b) return fail(SPARK_ERR_ARG, "spark_fill_guardcells: null batch");
  if (!b->has_ug && !b->amr) return fail(SPARK_ERR_STATE, "spark_fill_guardcells: no block grid declared");
  return run_fill(b, b->U, true, b->p.nguard);
}

int spark_compute_dt(spark_batch* b, double* dt, int minloc[5]) {
  if (!b || !dt) return fail(SPARK_ERR_ARG, "spark_compute_dt: null argument");
  int rc = run_dt(b, -1);
  if (rc) return rc;
  long long loc[2];
  HIPCHK(hipMemcpyAsync(dt, b->dtp, sizeof(double), hipMemcpyDeviceToHost, g_stream));
  HIPCHK(hipMemcpyAsync(loc, b->loc, sizeof(loc), hipMemcpyDeviceToHost, g_stream));
  HIPCHK(hipStreamSynchronize(g_stream));
  if (minloc) {
    const int bid = (int)loc[0];
    const BlockMeta& m = b->meta_h[bid < 0 || bid >= b->nblk ? 0 : bid];
    const long long c = loc[1];
    minloc[0] = (int)(c % m.NX) + 1;
    minloc[1] = (int)((c / m.NX) % m.NY) + 1;
    minloc[2] = (int)(c / ((long long)m.NX * m.NY)) + 1;
    minloc[3] = m.level;
    minloc[4] = bid;
  }
  return SPARK_OK;
}

int spark_hydro_step(spark_batch* b, double dt, double dtOld) {
  (void)dtOld;  // only the Poisson-gravity extrapolation reads dtOld (hy_rk_updateSoln.F90-mc:291-292)
  if (!b) return fail(SPARK_ERR_ARG, "spark_hydro_step: null batch");
  if (b->p.guard_mode == SPARK_GUARD_REFILL && !b->has_ug)
    return fail(SPARK_ERR_STATE, "refill guard mode needs spark_set_uniform_grid");
  HIPCHK(spark::plain_set_scalar(g_stream, b->dtp, dt));
  int rc;
  if ((b->has_ug || b->amr) && (rc = run_fill(b, b->U, false, stage1_reach(b)))) return rc;
  if ((rc = run_full_step(b))) return rc;
  if (b->amr && b->p.flux_correct) return run_amr_correct(b);
  return SPARK_OK;
}

int spark_step_begin(spark_batch* b) {
  if (!b) return fail(SPARK_ERR_ARG, "spark_step_begin: null batch");
  return run_step_begin(b);
}

int spark_stage(spark_batch* b, int stage, double dt, int ngoff) {
  if (!b) return fail(SPARK_ERR_ARG, "spark_stage: null batch");
  if (!b->step_open) return fail(SPARK_ERR_STATE, "spark_stage: call spark_step_begin first");
  if (stage < 1 || stage > b->last_stage) return fail(SPARK_ERR_ARG, "stage %d of %d", stage, b->last_stage);
  if (ngoff < 0 || ngoff + b->p.nstencil > b->p.nguard)
    return fail(SPARK_ERR_ARG, "ngoff = %d needs nguard >= %d", ngoff, ngoff + b->p.nstencil);
  HIPCHK(spark::plain_set_scalar(g_stream, b->dtp, dt));
  return run_stage(b, stage, ngoff);
}

int spark_step_end(spark_batch* b) {
  if (!b) return fail(SPARK_ERR_ARG, "spark_step_end: null batch");
  if (!b->step_open) return fail(SPARK_ERR_STATE, "spark_step_end: no open step");
  return run_step_end(b);
}

static int one_step(spark_batch* b) {
  int rc = run_dt(b, 0);
  if (rc) return rc;
  if ((b->has_ug || b->amr) && (rc = run_fill(b, b->U, false, stage1_reach(b)))) return rc;
  if ((rc = run_full_step(b))) return rc;
  if (b->amr && b->p.flux_correct) return run_amr_correct(b);
  return SPARK_OK;
}


int spark_advance(spark_batch* b, int nsteps, double* dts) {
  if (!b) return fail(SPARK_ERR_ARG, "spark_advance: null batch");
  if (nsteps < 0) return fail(SPARK_ERR_ARG, "nsteps = %d", nsteps);
  if (b->p.guard_mode == SPARK_GUARD_REFILL && !b->has_ug)
    return fail(SPARK_ERR_STATE, "refill guard mode needs spark_set_uniform_grid");
  if (nsteps > b->hist_cap) {
    drop_graph(b);
    (void)hipFree(b->dt_hist);
    b->dt_hist = nullptr;
    HIPCHK(hipMalloc((void**)&b->dt_hist, sizeof(double) * nsteps));
    b->hist_cap = nsteps;
  }
  int rc;
  if (b->fused && !b->planned && (rc = fused_plan(b))) return rc;
  HIPCHK(spark::plain_set_int(g_stream, b->hist_slot, 0));
  const bool graph = b->use_graph && !b->prof.on && nsteps > 0;
  if (graph && b->gexec == nullptr && b->amr && b->p.flux_correct && !b->fc_clean) {
    // zero the corrections outside the graph so that the captured exchange
    // visits only the blocks with a finer face
    for (int d = 0; d < 3; d++) HIPCHK(hipMemsetAsync(b->fc[d], 0, sizeof(double) * b->bpool, g_stream));
    b->fc_clean = true;
  }
  if (graph && b->gexec == nullptr) {
    // capture one device-resident step; every parameter of it lives on the
    // device (dt, the dt-history slot, the queue heads), so replays chain
    HIPCHK(hipStreamBeginCapture(g_stream, hipStreamCaptureModeThreadLocal));
    rc = one_step(b);
    hipGraph_t gr = nullptr;
    hipError_t e = hipStreamEndCapture(g_stream, &gr);
    if (rc) {
      if (gr) (void)hipGraphDestroy(gr);
      return rc;
    }
    if (e != hipSuccess) return fail(SPARK_ERR_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
    b->graph = gr;
    HIPCHK(hipGraphInstantiate(&b->gexec, b->graph, nullptr, nullptr, 0));
  }
  for (int s = 0; s < nsteps; s++) {
    if (graph) {
      HIPCHK(hipGraphLaunch(b->gexec, g_stream));
    } else if ((rc = one_step(b))) {
      return rc;
    }
  }
  if (dts && nsteps > 0) {
    HIPCHK(hipMemcpyAsync(dts, b->dt_hist, sizeof(double) * nsteps, hipMemcpyDeviceToHost, g_stream));
    HIPCHK(hipStreamSynchronize(g_stream));
  }
  return SPARK_OK;
}

int spark_set_geometry(spark_batch* b, int geometry, const double* const* r_center, const double* const* x_face_area,
                       const double* const* cell_volume, const double* const* r_left, const double* const* r_right) {
  if (!b) return fail(SPARK_ERR_ARG, "spark_set_geometry: null batch");
  if (geometry == SPARK_GEOM_CARTESIAN) {
    b->dp.geom = 0;
    b->dp.geo = nullptr;
    drop_graph(b);
    return SPARK_OK;
  }
  if (geometry != SPARK_GEOM_CYLINDRICAL && geometry != SPARK_GEOM_SPHERICAL)
    return fail(SPARK_ERR_ARG, "spark_set_geometry: geometry %d", geometry);
  if (b->fused) return fail(SPARK_ERR_ARG, "spark_set_geometry: the batched forms are Cartesian only");
  // the hierarchy's prolongation and flux restriction are the Cartesian ones
  if (b->amr) return fail(SPARK_ERR_STATE, "spark_set_geometry: the block hierarchy is Cartesian only");
  // the Grid's cell volumes exist for cylindrical geometry in 2D and
  // spherical in 1D and 2D (Grid_getCellVolumes.F90:49-55)
  if (geometry == SPARK_GEOM_CYLINDRICAL && b->p.ndim != 2)
    return fail(SPARK_ERR_ARG, "spark_set_geometry: cylindrical geometry is 2D (r, z) only");
  if (geometry == SPARK_GEOM_SPHERICAL && b->p.ndim > 2)
    return fail(SPARK_ERR_ARG, "spark_set_geometry: spherical geometry is 1D (r) or 2D (r, theta) only");
  if (b->classes.size() != 1) return fail(SPARK_ERR_ARG, "spark_set_geometry: blocks of one shape only");
  if (!r_center || !x_face_area || !cell_volume || !r_left || !r_right)
    return fail(SPARK_ERR_ARG, "spark_set_geometry: null argument");
  const int NX = b->meta_h[0].NX, NY = b->meta_h[0].NY;
  const int64_t S = spark::geo_stride(NX, NY), nxy = (int64_t)NX * NY;
  std::vector<double> h((size_t)(b->nblk * S));
  for (int i = 0; i < b->nblk; i++) {
    if (!r_center[i] || !x_face_area[i] || !cell_volume[i] || !r_left[i] || !r_right[i])
      return fail(SPARK_ERR_ARG, "spark_set_geometry: block %d: null array", i);
    double* o = h.data() + (size_t)(i * S);
    for (int q = 0; q < NX; q++) {
      o[q] = r_center[i][q];
      o[NX + q] = r_left[i][q];
      o[2 * NX + q] = r_right[i][q];
    }
    for (int64_t q = 0; q < nxy; q++) {
      if (!(cell_volume[i][q] != 0.0)) return fail(SPARK_ERR_ARG, "block %d: cell volume %lld is zero", i, (long long)q);
      o[3 * NX + q] = x_face_area[i][q];
      o[3 * NX + nxy + q] = cell_volume[i][q];
    }
  }
  if (b->geo_d) (void)hipFree(b->geo_d);
  b->geo_d = nullptr;
  HIPCHK(hipMalloc((void**)&b->geo_d, sizeof(double) * h.size()));
  HIPCHK(hipMemcpy(b->geo_d, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice));
  b->dp.geom = geometry == SPARK_GEOM_CYLINDRICAL ? 1 : 2;
  b->dp.geo_nx = NX;
  b->dp.geo_ny = NY;
  b->dp.geo = b->geo_d;
  drop_graph(b);  // the kernels take the parameters by value
  return SPARK_OK;
}

int spark_set_dt_control(spark_batch* b, double dt_init, double change_factor, double dt_min, double dt_max) {
  if (!b) return fail(SPARK_ERR_ARG, "spark_set_dt_control: null batch");
  const double v[4] = {dt_init, change_factor, dt_min, dt_max};
  for (double x : v)
    if (!(x >= 0.0) || !std::isfinite(x)) return fail(SPARK_ERR_ARG, "spark_set_dt_control: %g is not a finite value >= 0", x);
  if (dt_max > 0.0 && dt_min > dt_max) return fail(SPARK_ERR_ARG, "spark_set_dt_control: dt_min %g > dt_max %g", dt_min, dt_max);
  const bool on = dt_init > 0.0 || change_factor > 0.0 || dt_min > 0.0 || dt_max > 0.0;
  if (on && !b->dtctl_d) {
    HIPCHK(hipMalloc((void**)&b->dtctl_d, sizeof(spark::DtCtl)));
    drop_graph(b);  // the captured dt kernel takes the control block's address
  }
  if (on != b->dt_control) drop_graph(b);
  b->dt_control = on;
  if (on) {
    // no previous dt before the first step: no growth limit on it
    const spark::DtCtl c{dt_init > 0.0 ? dt_init : DBL_MAX, dt_init, change_factor, dt_min, dt_max};
    HIPCHK(hipMemcpyAsync(b->dtctl_d, &c, sizeof(c), hipMemcpyHostToDevice, g_stream));
    HIPCHK(hipStreamSynchronize(g_stream));
  }
  return SPARK_OK;
}

int spark_set_option(spark_batch* b, int option, int value) {
  if (!b) return fail(SPARK_ERR_ARG, "spark_set_option: null batch");
  switch (option) {
    case SPARK_OPT_GRAPH:
      b->use_graph = value != 0;
      if (!b->use_graph) drop_graph(b);
      return SPARK_OK;
    case SPARK_OPT_AUTOTUNE:
      b->autotune = value != 0;
      return SPARK_OK;
    case SPARK_OPT_SAME_LEVEL_HALOS:
      b->same_level = value != 0;
      b->planned = false;
      drop_graph(b);
      return SPARK_OK;
    default:
      return fail(SPARK_ERR_ARG, "spark_set_option: unknown option %d", option);
  }
}

int spark_batched_tiles(spark_batch* b, int* out, int max_rows, int* nrows) {
  if (!b || !out || !nrows) return fail(SPARK_ERR_ARG, "spark_batched_tiles: bad argument");
  if (!b->fused) return fail(SPARK_ERR_STATE, "spark_batched_tiles: not a batched FP32-flux batch");
  int rc;
  if (!b->planned && (rc = fused_plan(b))) return rc;
  int n = 0;
  for (const auto& kv : b->tiles) {
    if (n < max_rows) {
      const ShapeClass& c = b->classes[kv.first.first];
      int* r = out + 6 * n;
      r[0] = c.nx; r[1] = c.ny; r[2] = c.nz; r[3] = kv.first.second; r[4] = kv.second.first; r[5] = kv.second.second;
    }
    n++;
  }
  *nrows = n;
  return SPARK_OK;
}

static int reg_extent(const spark_batch* b, int i, int dir, int e[3]) {
  const BlockMeta& m = b->meta_h[i];
  e[0] = m.nx; e[1] = m.ny; e[2] = m.nz;
  e[dir] += 1;
  return e[0] * e[1] * e[2];
}

int spark_get_flux_registers(spark_batch* b, double* const* fbx, double* const* fby, double* const* fbz) {
  if (!b) return fail(SPARK_ERR_ARG, "spark_get_flux_registers: null batch");
  if (!b->p.flux_correct) return fail(SPARK_ERR_STATE, "flux registers need flux_correct = 1");
  double* const* outs[3] = {fbx, fby, fbz};
  const int nf = 5 + b->p.nscalars;
  HIPCHK(hipStreamSynchronize(g_stream));
  std::vector<double> tmp;
  for (int dir = 0; dir < b->p.ndim; dir++) {
    if (!outs[dir]) return fail(SPARK_ERR_ARG, "flux register array %d is null", dir);
    for (int i = 0; i < b->nblk; i++) {
      int e[3];
      const int tot = reg_extent(b, i, dir, e);
      tmp.resize((size_t)nf * tot);
      HIPCHK(hipMemcpy(tmp.data(), b->fb[dir] + b->meta_h[i].bbase, sizeof(double) * nf * tot, hipMemcpyDeviceToHost));
      double* o = outs[dir][i];
      for (int t = 0; t < tot; t++)
        for (int n = 0; n < nf; n++) o[(size_t)t * nf + n] = tmp[(size_t)n * tot + t];
    }
  }
  return SPARK_OK;
}

int spark_set_flux_corrections(spark_batch* b, const double* const* fbx, const double* const* fby,
                               const double* const* fbz) {
  if (!b) return fail(SPARK_ERR_ARG, "spark_set_flux_corrections: null batch");
  if (!b->p.flux_correct) return fail(SPARK_ERR_STATE, "flux corrections need flux_correct = 1");
  b->fc_clean = false;
  drop_graph(b);
  const double* const* ins[3] = {fbx, fby, fbz};
  const int nf = 5 + b->p.nscalars;
  HIPCHK(hipStreamSynchronize(g_stream));
  std::vector<double> tmp;
  for (int dir = 0; dir < b->p.ndim; dir++) {
    if (!ins[dir]) return fail(SPARK_ERR_ARG, "flux correction array %d is null", dir);
    for (int i = 0; i < b->nblk; i++) {
      int e[3];
      const int tot = reg_extent(b, i, dir, e);
      tmp.resize((size_t)nf * tot);
      const double* in = ins[dir][i];
      for (int t = 0; t < tot; t++)
        for (int n = 0; n < nf; n++) tmp[(size_t)n * tot + t] = in[(size_t)t * nf + n];
      HIPCHK(hipMemcpy(b->fc[dir] + b->meta_h[i].bbase, tmp.data(), sizeof(double) * nf * tot, hipMemcpyHostToDevice));
    }
  }
  return SPARK_OK;
}

int spark_flux_correct(spark_batch* b, double dt) {
  if (!b) return fail(SPARK_ERR_ARG, "spark_flux_correct: null batch");
  if (!b->p.flux_correct) return fail(SPARK_ERR_STATE, "spark_flux_correct needs flux_correct = 1");
  HIPCHK(spark::plain_set_scalar(g_stream, b->dtp, dt));
  for (auto& c : b->classes) {
    PlainLaunch L = make_launch(b, c);
    for (int d = 0; d < 3; d++) L.fb[d] = b->fc[d];
    HIPCHK(spark::plain_correct(L, b->U));
  }
  return SPARK_OK;
}

int spark_flux_exchange(spark_batch* b) {
  if (!b) return fail(SPARK_ERR_ARG, "spark_flux_exchange: null batch");
  if (!b->p.flux_correct) return fail(SPARK_ERR_STATE, "spark_flux_exchange needs flux_correct = 1");
  if (!b->amr) return fail(SPARK_ERR_STATE, "spark_flux_exchange: no block hierarchy declared");
  return run_exchange(b);
}

int spark_set_amr_grid(spark_batch* b, int nbx0, int nby0, int nbz0, const int* origins, const int* bc) {
  if (!b || !origins) return fail(SPARK_ERR_ARG, "spark_set_amr_grid: null argument");
  if (b->p.ndim != 3) return fail(SPARK_ERR_ARG, "spark_set_amr_grid: 3D only");
  if (b->p.guard_mode != SPARK_GUARD_TELESCOPE)
    return fail(SPARK_ERR_ARG, "spark_set_amr_grid: the block hierarchy is filled once per step (telescope mode)");
  if (nbx0 < 1 || nby0 < 1 || nbz0 < 1) return fail(SPARK_ERR_ARG, "spark_set_amr_grid: bad root grid");
  if (b->has_ug) return fail(SPARK_ERR_STATE, "spark_set_amr_grid: the batch has a uniform grid");
  if (b->dp.geom != 0) return fail(SPARK_ERR_STATE, "spark_set_amr_grid: the block hierarchy is Cartesian only");
  if (b->classes.size() != 1) return fail(SPARK_ERR_ARG, "spark_set_amr_grid: blocks of one shape only");
  const int n = b->meta_h[0].nx;
  if (b->meta_h[0].ny != n || b->meta_h[0].nz != n || (n & 1))
    return fail(SPARK_ERR_ARG, "spark_set_amr_grid: blocks must be cubic with an even cell count");
  int lmax = 1;
  for (int i = 0; i < b->nblk; i++) {
    const int L = b->meta_h[i].level;
    if (L < 1 || L > 12) return fail(SPARK_ERR_ARG, "block %d: level %d out of range", i, L);
    lmax = std::max(lmax, L);
  }
  const int nb0[3] = {nbx0, nby0, nbz0};
  int md[3];
  for (int d = 0; d < 3; d++) md[d] = nb0[d] << (lmax - 1);
  std::vector<int> map((size_t)md[0] * md[1] * md[2], -1);
  for (int i = 0; i < b->nblk; i++) {
    const int L = b->meta_h[i].level, sh = lmax - L;
    int lo[3];
    for (int d = 0; d < 3; d++) {
      const int o = origins[3 * i + d];
      if (o < 0 || o % n || o / n >= (nb0[d] << (L - 1)))
        return fail(SPARK_ERR_ARG, "block %d: origin %d not a block position of level %d", i, o, L);
      lo[d] = (o / n) << sh;
    }
    for (int z = lo[2]; z < lo[2] + (1 << sh); z++)
      for (int y = lo[1]; y < lo[1] + (1 << sh); y++)
        for (int x = lo[0]; x < lo[0] + (1 << sh); x++) {
          int& slot = map[x + (size_t)md[0] * (y + (size_t)md[1] * z)];
          if (slot >= 0) return fail(SPARK_ERR_ARG, "blocks %d and %d overlap", slot, i);
          slot = i;
        }
  }
  for (size_t t = 0; t < map.size(); t++)
    if (map[t] < 0) return fail(SPARK_ERR_ARG, "the blocks do not cover the domain");
  // every leaf under a block's guard region is within one level of it (the
  // guard fill and the flux exchange handle coarser-by-one, same and
  // finer-by-one sources only)
  int per[3], bcs[3];
  for (int d = 0; d < 3; d++) {
    bcs[d] = bc ? bc[d] : SPARK_BC_OUTFLOW;
    if (bcs[d] < SPARK_BC_OUTFLOW || bcs[d] > SPARK_BC_REFLECT)
      return fail(SPARK_ERR_ARG, "spark_set_amr_grid: boundary code %d in dimension %d", bcs[d], d);
    per[d] = bcs[d] == SPARK_BC_PERIODIC;
    // a periodic guard cell (and its prolongation stencil) must lie within
    // one period of the domain
    if (per[d] && nb0[d] * n < b->p.nguard + 1)
      return fail(SPARK_ERR_ARG, "spark_set_amr_grid: periodic dimension %d spans %d cells, fewer than nguard + 1 = %d",
                  d, nb0[d] * n, b->p.nguard + 1);
  }
  for (int i = 0; i < b->nblk; i++) {
    const int L = b->meta_h[i].level, sh = lmax - L, g = b->p.nguard;
    int lo[3], hi[3];
    for (int d = 0; d < 3; d++) {
      const int cmax = (nb0[d] * n << (L - 1)) - 1;
      int c0 = origins[3 * i + d] - g, c1 = origins[3 * i + d] + n - 1 + g;
      if (!per[d]) { c0 = std::max(0, c0); c1 = std::min(cmax, c1); }
      // slot range (floor division), wrapped below for periodic dimensions
      lo[d] = (int)std::floor((double)((int64_t)c0 << sh) / n);
      hi[d] = (int)std::floor((double)((((int64_t)c1 + 1) << sh) - 1) / n) + 1;
      if (!per[d]) hi[d] = std::min(hi[d], md[d]);
    }
    for (int z = lo[2]; z < hi[2]; z++)
      for (int y = lo[1]; y < hi[1]; y++)
        for (int x = lo[0]; x < hi[0]; x++) {
          const int X = ((x % md[0]) + md[0]) % md[0], Y = ((y % md[1]) + md[1]) % md[1], Z = ((z % md[2]) + md[2]) % md[2];
          const int L2 = b->meta_h[map[X + (size_t)md[0] * (Y + (size_t)md[1] * Z)]].level;
          if (L2 - L > 1 || L - L2 > 1)
            return fail(SPARK_ERR_ARG, "block %d (level %d): a leaf of level %d lies under its guard cells", i, L, L2);
        }
  }
  // batched forms: a guard sector (face, edge or corner region of the guard
  // band) across which lies a leaf of the block's own level (periodic wrap
  // allowed) holds nothing a stage needs that the leaf's interior does not,
  // so the block reads that sector from the leaf, every stage.  A block
  // whose 26 sectors all qualify updates only its interior; one with some
  // telescopes over the others (it computes them at each stage's depth);
  // one with none telescopes over its own guard cells.  The stage values
  // this gives are those of telescoping: the same cells from the same
  // inputs, since the guard fill is a function of the cell's position and
  // the leaves' data.
  b->amr_remap_blk.assign(b->nblk, 0);
  b->amr_local.assign(b->nblk, 0);
  b->amr_remap_count = 0;
  b->amr_partial_count = 0;
  if (b->fused && n >= b->p.nguard) {
    std::vector<int64_t> nb((size_t)b->nblk * 27, -1);
    const uint32_t all_sectors = ((1u << 27) - 1) & ~(1u << 13);
    for (int i = 0; i < b->nblk; i++) {
      const BlockMeta& mi = b->meta_h[i];
      const int L = mi.level, sh = lmax - L;
      int bp[3], nbl[3];
      for (int d = 0; d < 3; d++) { bp[d] = origins[3 * i + d] / n; nbl[d] = nb0[d] << (L - 1); }
      int64_t row[27];
      uint32_t local = 0;
      for (int oz = -1; oz <= 1; oz++)
        for (int oy = -1; oy <= 1; oy++)
          for (int ox = -1; ox <= 1; ox++) {
            const int s = 13 + ox + 3 * oy + 9 * oz;
            if (s == 13) { row[s] = mi.base; continue; }
            const int o[3] = {ox, oy, oz};
            int q[3];
            bool in = true;
            for (int d = 0; d < 3; d++) {
              q[d] = bp[d] + o[d];
              if (q[d] < 0 || q[d] >= nbl[d]) {
                if (!per[d]) in = false;
                q[d] = (q[d] + nbl[d]) % nbl[d];
              }
            }
            int j = -1;
            if (in) j = map[((size_t)q[0] << sh) + (size_t)md[0] * (((size_t)q[1] << sh) + (size_t)md[1] * ((size_t)q[2] << sh))];
            if (j >= 0 && b->meta_h[j].level == L) {
              row[s] = b->meta_h[j].base;
            } else {
              // computed here: the entry cell_at() shifts back into this block
              row[s] = mi.base + (int64_t)ox * mi.nx + (int64_t)mi.NX * ((int64_t)oy * mi.ny + (int64_t)mi.NY * oz * mi.nz);
              local |= 1u << s;
            }
          }
      if (local == all_sectors) continue;  // telescopes over its own guard cells
      std::copy(row, row + 27, nb.begin() + (size_t)i * 27);
      b->amr_local[i] = local;
      b->amr_remap_blk[i] = local ? 2 : 1;
      (local ? b->amr_partial_count : b->amr_remap_count)++;
    }
    if (b->amr_remap_count + b->amr_partial_count > 0) {
      std::vector<int> rl, tl;
      for (int i = 0; i < b->nblk; i++) (b->amr_remap_blk[i] == 1 ? rl : tl).push_back(i);
      auto F = [](void* p) { if (p) (void)hipFree(p); };
      F(b->nbr_d); F(b->remap_ids_d); F(b->tele_ids_d);
      b->nbr_d = nullptr; b->remap_ids_d = nullptr; b->tele_ids_d = nullptr;
      HIPCHK(hipMalloc((void**)&b->nbr_d, sizeof(int64_t) * nb.size()));
      HIPCHK(hipMemcpy(b->nbr_d, nb.data(), sizeof(int64_t) * nb.size(), hipMemcpyHostToDevice));
      if (!rl.empty()) {
        HIPCHK(hipMalloc((void**)&b->remap_ids_d, sizeof(int) * rl.size()));
        HIPCHK(hipMemcpy(b->remap_ids_d, rl.data(), sizeof(int) * rl.size(), hipMemcpyHostToDevice));
      }
      b->n_tele = (int)tl.size();
      if (!tl.empty()) {
        HIPCHK(hipMalloc((void**)&b->tele_ids_d, sizeof(int) * tl.size()));
        HIPCHK(hipMemcpy(b->tele_ids_d, tl.data(), sizeof(int) * tl.size(), hipMemcpyHostToDevice));
      }
      std::vector<uint32_t> sm(b->nblk);
      for (int i = 0; i < b->nblk; i++) sm[i] = b->amr_remap_blk[i] == 2 ? b->amr_local[i] : all_sectors;
      F(b->amr_smask_d);
      b->amr_smask_d = nullptr;
      HIPCHK(hipMalloc((void**)&b->amr_smask_d, sizeof(uint32_t) * sm.size()));
      HIPCHK(hipMemcpy(b->amr_smask_d, sm.data(), sizeof(uint32_t) * sm.size(), hipMemcpyHostToDevice));
    }
    b->planned = false;
  }
  // faces with a finer leaf across (the region across a face is one block
  // position of the block's level: one leaf, or finer leaves only)
  {
    std::vector<unsigned char> fm(b->nblk, 0);
    std::vector<int> cl;
    for (int i = 0; i < b->nblk; i++) {
      const int L = b->meta_h[i].level, sh = lmax - L;
      for (int d = 0; d < 3; d++)
        for (int sd = 0; sd < 2; sd++) {
          int q[3];
          bool in = true;
          for (int e = 0; e < 3; e++) {
            const int nbl = nb0[e] << (L - 1);
            q[e] = origins[3 * i + e] / n + (e == d ? (sd ? 1 : -1) : 0);
            if (q[e] < 0 || q[e] >= nbl) {
              if (per[e]) q[e] = (q[e] + nbl) % nbl;
              else in = false;
            }
          }
          if (!in) continue;
          const int j = map[((size_t)q[0] << sh) + (size_t)md[0] * (((size_t)q[1] << sh) + (size_t)md[1] * ((size_t)q[2] << sh))];
          if (b->meta_h[j].level > L) fm[i] |= (unsigned char)(1 << (2 * d + sd));
        }
      if (fm[i]) cl.push_back(i);
    }
    auto F = [](void* p) { if (p) (void)hipFree(p); };
    F(b->amr_fmask_d); F(b->corr_ids_d);
    b->amr_fmask_d = nullptr; b->corr_ids_d = nullptr;
    HIPCHK(hipMalloc((void**)&b->amr_fmask_d, b->nblk));
    HIPCHK(hipMemcpy(b->amr_fmask_d, fm.data(), b->nblk, hipMemcpyHostToDevice));
    b->n_corr = (int)cl.size();
    b->fc_clean = false;
    if (!cl.empty()) {
      HIPCHK(hipMalloc((void**)&b->corr_ids_d, sizeof(int) * cl.size()));
      HIPCHK(hipMemcpy(b->corr_ids_d, cl.data(), sizeof(int) * cl.size(), hipMemcpyHostToDevice));
    }
  }
  if (!b->amr_map_d) HIPCHK(hipMalloc((void**)&b->amr_map_d, sizeof(int) * map.size()));
  else {
    (void)hipFree(b->amr_map_d);
    HIPCHK(hipMalloc((void**)&b->amr_map_d, sizeof(int) * map.size()));
  }
  if (!b->amr_org_d) HIPCHK(hipMalloc((void**)&b->amr_org_d, sizeof(int) * 3 * b->nblk));
  HIPCHK(hipMemcpy(b->amr_map_d, map.data(), sizeof(int) * map.size(), hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(b->amr_org_d, origins, sizeof(int) * 3 * b->nblk, hipMemcpyHostToDevice));
  b->amr_g.lmax = lmax;
  b->amr_g.n = n;
  b->amr_g.nshift = -1;
  for (int q = 0; q < 31; q++)
    if ((1 << q) == n) b->amr_g.nshift = q;
  for (int d = 0; d < 3; d++) { b->amr_g.nb0[d] = nb0[d]; b->amr_g.md[d] = md[d]; b->amr_g.bc[d] = bcs[d]; }
  b->amr_g.map = b->amr_map_d;
  b->amr_g.org = b->amr_org_d;
  b->amr = true;
  drop_graph(b);
  return SPARK_OK;
}

static void prof_clear(spark_batch* b) {
  for (auto& r : b->prof.recs) {
    if (r.e0) (void)hipEventDestroy(r.e0);
    if (r.e1) (void)hipEventDestroy(r.e1);
  }
  b->prof.recs.clear();
}

int spark_profile(spark_batch* b, int enable) {
  if (!b) return fail(SPARK_ERR_ARG, "spark_profile: null batch");
  HIPCHK(hipStreamSynchronize(g_stream));
  prof_clear(b);
  b->prof.err = hipSuccess;
  b->prof.on = enable != 0;
  return SPARK_OK;
}

int spark_profile_read(spark_batch* b, spark_kernel_stat* out, int max, int* n) {
  if (!b || !out || !n || max < 0) return fail(SPARK_ERR_ARG, "spark_profile_read: bad argument");
  HIPCHK(hipStreamSynchronize(g_stream));
  HIPCHK(b->prof.err);
  std::vector<spark_kernel_stat> acc;
  for (auto& r : b->prof.recs) {
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, r.e0, r.e1));
    spark_kernel_stat* k = nullptr;
    for (auto& a : acc)
      if (strncmp(a.name, r.name, sizeof(a.name)) == 0) k = &a;
    if (!k) {
      spark_kernel_stat z;
      memset(&z, 0, sizeof(z));
      snprintf(z.name, sizeof(z.name), "%s", r.name);
      acc.push_back(z);
      k = &acc.back();
    }
    k->launches += 1;
    k->total_ms += ms;
    k->bytes += r.bytes;
    k->flops += r.flops;
  }
  *n = (int)acc.size();
  for (int i = 0; i < (int)acc.size() && i < max; i++) out[i] = acc[i];
  return SPARK_OK;
}

int spark_synchronize(spark_batch* b) {
  (void)b;
  if (g_stream) HIPCHK(hipStreamSynchronize(g_stream));
  return SPARK_OK;
}

int spark_get_counters(spark_batch* b, long long* counters, int n) {
  if (!b || !counters || n < 1) return fail(SPARK_ERR_ARG, "spark_get_counters: bad argument");
  int e = 0;
  HIPCHK(hipMemcpyAsync(&e, b->err, sizeof(int), hipMemcpyDeviceToHost, g_stream));
  HIPCHK(hipStreamSynchronize(g_stream));
  counters[0] = e;
  for (int i = 1; i < n; i++) counters[i] = 0;
  return SPARK_OK;
}

int spark_last_error(char* buf, int len) {
  if (!buf || len <= 0) return SPARK_ERR_ARG;
  snprintf(buf, (size_t)len, "%s", g_err.c_str());
  return SPARK_OK;
}

}  // extern "C"
