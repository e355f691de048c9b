// ep_slab_m18_25.hip -- the persistent streaming launch (k_stream_slab) for
// 18 to 25 mu cells, in a translation unit of its own so that it compiles
// in parallel with the per-pass kernels (ep_stream_m18_25.hip); the kernels
// are in ep_stream.h.
#include "ep_stream.h"

namespace ep {

hipError_t launch_stream_slab_m18_25(const Dev &D, int do_seed, int sub, int fuse, int tune, int one, hipStream_t st) {
  switch (D.M) {
    EP_SLAB_CASE(18) EP_SLAB_CASE(19) EP_SLAB_CASE(20) EP_SLAB_CASE(21) EP_SLAB_CASE(22) EP_SLAB_CASE(23) EP_SLAB_CASE(24) EP_SLAB_CASE(25)
    default:
      return hipErrorInvalidValue;
  }
}

}  // namespace ep
