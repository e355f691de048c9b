// ep_slab_m10_17.hip -- the persistent streaming launch (k_stream_slab) for
// 10 to 17 mu cells, in a translation unit of its own so that it compiles
// in parallel with the per-pass kernels (ep_stream_m10_17.hip); the kernels
// are in ep_stream.h.
#include "ep_stream.h"

namespace ep {

hipError_t launch_stream_slab_m10_17(const Dev &D, int do_seed, int sub, int fuse, int tune, int one, hipStream_t st) {
  switch (D.M) {
    EP_SLAB_CASE(10) EP_SLAB_CASE(11) EP_SLAB_CASE(12) EP_SLAB_CASE(13) EP_SLAB_CASE(14) EP_SLAB_CASE(15) EP_SLAB_CASE(16) EP_SLAB_CASE(17)
    default:
      return hipErrorInvalidValue;
  }
}

}  // namespace ep
