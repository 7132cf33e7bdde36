// spark_launch.hpp -- host-side launch interface between the C-ABI layer
// (spark_capi.cpp) and the kernel translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "spark_physics.hpp"

namespace spark {

// Optional per-launch timing: HIP events recorded on the launch stream
// around each kernel, with the launch's algorithmic bytes and flops.
struct ProfRec {
  const char* name;
  hipEvent_t e0, e1;
  double bytes, flops;
};
struct Prof {
  bool on = false;
  hipError_t err = hipSuccess;  // first event error, reported when read
  std::vector<ProfRec> recs;
  void begin(hipStream_t st, const char* name, double bytes, double flops) {
    if (!on) return;
    ProfRec r{name, nullptr, nullptr, bytes, flops};
    hipError_t e = hipEventCreate(&r.e0);
    if (e == hipSuccess) e = hipEventCreate(&r.e1);
    if (e == hipSuccess) e = hipEventRecord(r.e0, st);
    if (e != hipSuccess && err == hipSuccess) err = e;
    recs.push_back(r);
  }
  void end(hipStream_t st) {
    if (!on || recs.empty()) return;
    hipError_t e = hipEventRecord(recs.back().e1, st);
    if (e != hipSuccess && err == hipSuccess) err = e;
  }
};

// Everything one plain-form launch sequence needs for a list of blocks that
// share one interior shape.
struct PlainLaunch {
  DevParams P;
  hipStream_t stream;
  const BlockMeta* meta;  // device block table (all blocks of the batch)
  const int* blist;       // device list of block ids in this shape class
  int nblk;               // length of blist
  int nx, ny, nz;         // interior shape of the class
  const double* U0;       // step input (hy_tmpState role); also holds SHOK
  const double* Sin;      // stage input state (U at stage 1, S after)
  double* Sout;           // stage output state (U at the last stage, S before)
  double* flat;           // flattening coefficients / scratch (cell pool)
  double* div;            // divergence terms of the sweeps so far (cell pool, 5 + ns fields)
  double* fb[3];          // flux-register pools (nullptr when not correcting)
  double weight;          // flux-register weight of this stage
  int add_flux;           // 0: overwrite (stage 1), 1: accumulate
  const double* dtp;      // device scalar holding dt
  int* err;               // device counter of negative a^2 events
  Prof* prof;             // per-launch timing (nullptr: off)
};

hipError_t plain_stage(const PlainLaunch& L, int ngoff, double c1, double c2, double c3, int final_stage);
// nbr (3D, optional): per block 27 pool offsets of the same-level leaves
// around it; the block's guard cells are then read from those leaves
hipError_t plain_shock_detect(const PlainLaunch& L, double* U, double* scratch, int band,
                              const int64_t* nbr = nullptr);
hipError_t plain_dt_part(const PlainLaunch& L, const double* U, void* parts, int chunks);
// the driver's limits on the step's dt (Driver_computeDt.F90:450-460 and
// the first step's dtinit, Driver_evolveAll), kept on the device: `first` > 0
// is taken as the next dt (then cleared), later dts are the CFL dt limited
// to prev * factor (factor > 0) and clamped to [dtmin, dtmax] (each when > 0)
struct DtCtl {
  double prev, first, factor, dtmin, dtmax;
};
hipError_t plain_dt_final(hipStream_t st, const DevParams& P, const void* parts, int nblk_total, int chunks,
                          double* dt_out, long long* loc_out, double* dt_hist, int* hist_slot, DtCtl* ctl);
hipError_t plain_set_scalar(hipStream_t st, double* p, double v);
hipError_t plain_set_int(hipStream_t st, int* p, int v);
size_t plain_dt_part_size();
// variables a guard fill copies (canonical indices)
struct VarList {
  int n;
  int v[24];
};
hipError_t plain_ug_fill(const PlainLaunch& L, int nbx, int nby, int nbz, double* U, const VarList& vars, int depth);
hipError_t plain_correct(const PlainLaunch& L, double* U, const unsigned char* fmask = nullptr);

// Block hierarchy (leaves of n^3 cells over a level-1 grid of nb0 blocks)
struct AmrGrid {
  int lmax;       // finest level (1-based)
  int n;          // interior cells per side of every block
  int nshift;     // log2(n) when n is a power of two, else -1
  int nb0[3];     // level-1 blocks per dimension
  int md[3];      // leaf map extent: nb0 << (lmax - 1)
  const int* map; // finest-level block slot -> leaf block
  const int* org; // 3 per block: first interior cell, global index at its level
  int bc[3];      // per dimension: AMR_BC_OUTFLOW, AMR_BC_PERIODIC, AMR_BC_REFLECT
};
enum { AMR_BC_OUTFLOW = 0, AMR_BC_PERIODIC = 1, AMR_BC_REFLECT = 2 };
// smask (optional, per block): guard sectors filled to `depth`; the others
// only to `keep` layers
hipError_t amr_fill(const PlainLaunch& L, const AmrGrid& G, double* U, const VarList& vars, int depth,
                    const uint32_t* smask = nullptr, int keep = 0);
hipError_t amr_flux_exchange(const PlainLaunch& L, const AmrGrid& G, double* const fb[3], double* const fc[3]);
hipError_t plain_aos_to_soa(hipStream_t st, const BlockMeta* meta, int nblk, int maxcells, const double* aos,
                            const int64_t* aos_off, int nunk, const int* var_index, int nvar, double* U);
hipError_t plain_soa_to_aos(hipStream_t st, const DevParams& P, const BlockMeta* meta, int nblk, int maxcells,
                            const double* U, double* aos, const int64_t* aos_off, int nunk, const int* var_index,
                            int nvar, int interior);

// fixed staging area of spark_pack_noalloc (spark_noalloc.hip): 64 MiB per
// device, of which the AoS staging of one round of blocks starts at
// NOALLOC_STAGE; a round holds at most NOALLOC_MAXB blocks
constexpr size_t NOALLOC_BYTES = size_t(64) << 20;
constexpr int NOALLOC_MAXB = 1024;
constexpr size_t NOALLOC_STAGE = 256 + 8 * (size_t)NOALLOC_MAXB;
// device address of this device's area (call with the device current)
hipError_t noalloc_scratch_address(void** ptr);

}  // namespace spark
