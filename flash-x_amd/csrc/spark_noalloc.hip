// spark_noalloc.hip -- the fixed device staging area of spark_pack_noalloc().
//
// spark_pack() stages the host's AoS blocks in device memory it allocates
// and frees on every call.  spark_pack_noalloc() stages them here instead:
// a file-scope device array of NOALLOC_BYTES that each GPU the library runs
// on holds until the process ends.  Host code never passes the array's name
// to a kernel or a copy (on the host it is only a stand-in symbol); it asks
// the runtime for the array's device address once per device, at
// spark_init(), and uses that address.
//
// Layout of the area (byte offsets):
//   [0, 256)                    variable map (int, at most 19 entries)
//   [256, 256 + 8 NOALLOC_MAXB) per-block offsets (int64) of one round
//   [NOALLOC_STAGE, end)        AoS staging of one round of blocks
#include <hip/hip_runtime.h>

#include "spark_launch.hpp"

namespace spark {

static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

hipError_t noalloc_scratch_address(void** ptr) {
  *ptr = nullptr;
  return hipGetSymbolAddress(ptr, HIP_SYMBOL(scratch));
}

}  // namespace spark
