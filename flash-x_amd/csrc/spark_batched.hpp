// spark_batched.hpp -- host/device interface of the batched FP32-flux form.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "spark_physics.hpp"

namespace spark {

constexpr int FUSED_THREADS = 256;    // workgroup size of the fused stage kernel
constexpr int FUSED_XCDS = 8;         // MI355X: 8 XCDs, workgroups dispatched round-robin
// mass scalars staged together; species (clamped and renormalised
// together at faces) must all fall in the first chunk
constexpr int FUSED_SCALAR_CHUNK = 4;

// one work item: a (tx x ty) column tile of one block's stage region,
// marched over z0..z1 (guard-celled box coordinates)
struct FusedItem {
  int blk;
  int x0, y0, tx, ty;
  int z0, z1;
  int pad[5];
};

// neighbour-table entry of a uniform grid's guard sector past the domain edge
constexpr int64_t NBR_EDGE = INT64_MIN;

__device__ __forceinline__ int64_t bcix(const BlockMeta& m, int i, int j, int k) {
  return (int64_t)i + (int64_t)m.NX * ((int64_t)j + (int64_t)m.NY * (int64_t)k);
}

// pool offset of cell (i, j, k) of a block's guard-celled box.  With a
// neighbour table a cell outside the interior is taken from the adjacent
// block of its guard sector, or, past a uniform grid's domain edge
// (NBR_EDGE), from the block's own boundary layer (outflow: the guard copies
// the first interior layer, Grid_bcApplyToRegion.F90:526-532), one
// dimension at a time as the guard fill clamps the global index.  A sector
// the block computes itself holds its own base shifted by the sector's
// offset, so the same arithmetic lands in its own guard cells.
__device__ __forceinline__ int64_t cell_at(const BlockMeta& m, const int64_t* __restrict__ nbt, int g, int i, int j,
                                           int k) {
  if (nbt == nullptr) return m.base + bcix(m, i, j, k);
  int ox = 0, oy = 0, oz = 0;
  if (i < g) ox = -1; else if (i >= g + m.nx) ox = 1;
  if (j < g) oy = -1; else if (j >= g + m.ny) oy = 1;
  if (k < g) oz = -1; else if (k >= g + m.nz) oz = 1;
  if ((ox | oy | oz) == 0) return m.base + bcix(m, i, j, k);
  if (ox != 0) {
    if (nbt[13 + ox] == NBR_EDGE) { i = ox < 0 ? g : g + m.nx - 1; ox = 0; }
    else i -= ox * m.nx;
  }
  if (oy != 0) {
    if (nbt[13 + 3 * oy] == NBR_EDGE) { j = oy < 0 ? g : g + m.ny - 1; oy = 0; }
    else j -= oy * m.ny;
  }
  if (oz != 0) {
    if (nbt[13 + 9 * oz] == NBR_EDGE) { k = oz < 0 ? g : g + m.nz - 1; oz = 0; }
    else k -= oz * m.nz;
  }
  return nbt[13 + ox + 3 * oy + 9 * oz] + bcix(m, i, j, k);
}

struct FusedArgs {
  DevParams P;
  const BlockMeta* meta;
  const FusedItem* items;  // grouped by queue
  const int* qoff;         // [FUSED_XCDS] first item of each queue
  const int* qlen;         // [FUSED_XCDS] items in each queue
  int* counters;           // [FUSED_XCDS] queue heads (reset before every launch)
  const double* U0;        // step input (conserved U0 of the update)
  const double* Sin;       // stage input state
  double* Sout;            // stage output state (never equal to Sin)
  double* fb[3];           // flux registers (nullptr when not correcting)
  // per block the 27 pool offsets (BlockMeta::base) of the blocks around
  // it: cells outside the block's interior are read from those neighbours
  // (uniform grid, guards not refilled between stages: NBR_EDGE where the
  // domain ends, outflow).  A sector the block computes itself (hierarchy:
  // no leaf of its level across) holds base + ox nx + NX (oy ny + NY oz nz),
  // which cell_at() turns into the block's own guard cell.  A row whose
  // centre entry is -1, or nullptr: the block reads its own guard cells.
  const int64_t* nbr;
  double weight;
  int add_flux;
  const double* dtp;
  double c1, c2, c3;
  int final_stage;
  int* err;
};

// the fused stage kernel in FP32 (_f32) and FP64 (_f64) arithmetic; the LDS
// of a work item depends on its tile (tx x ty, tx * ty <= FUSED_THREADS)
size_t fused_lds_bytes_f32(int ns, int tx, int ty);
hipError_t fused_reset_f32(hipStream_t st, int* counters);
hipError_t fused_stage_f32(const FusedArgs& A, int grid, size_t lds, hipStream_t st);
hipError_t fused_occupancy_f32(int ns, size_t lds, int* blocks_per_cu);
size_t fused_static_lds_f32(int ns);  // the kernel's own __shared__ bytes beside the dynamic LDS
size_t fused_lds_bytes_f64(int ns, int tx, int ty);
hipError_t fused_reset_f64(hipStream_t st, int* counters);
hipError_t fused_stage_f64(const FusedArgs& A, int grid, size_t lds, hipStream_t st);
hipError_t fused_occupancy_f64(int ns, size_t lds, int* blocks_per_cu);
size_t fused_static_lds_f64(int ns);

}  // namespace spark
