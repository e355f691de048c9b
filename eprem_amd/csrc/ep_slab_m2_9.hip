// ep_slab_m2_9.hip -- the persistent streaming launch (k_stream_slab) for
// 2 to 9 mu cells, in a translation unit of its own so that it compiles
// in parallel with the per-pass kernels (ep_stream_m2_9.hip); the kernels
// are in ep_stream.h.
#include "ep_stream.h"

namespace ep {

hipError_t launch_stream_slab_m2_9(const Dev &D, int do_seed, int sub, int fuse, int tune, int one, hipStream_t st) {
  switch (D.M) {
    EP_SLAB_CASE(2) EP_SLAB_CASE(3) EP_SLAB_CASE(4) EP_SLAB_CASE(5) EP_SLAB_CASE(6) EP_SLAB_CASE(7) EP_SLAB_CASE(8) EP_SLAB_CASE(9)
    default:
      return hipErrorInvalidValue;
  }
}

}  // namespace ep
