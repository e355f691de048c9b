// ep_slab_m26_32.hip -- the persistent streaming launch (k_stream_slab) for
// 26 to 32 mu cells, in a translation unit of its own so that it compiles
// in parallel with the per-pass kernels (ep_stream_m26_32.hip); the kernels
// are in ep_stream.h.
#include "ep_stream.h"

namespace ep {

hipError_t launch_stream_slab_m26_32(const Dev &D, int do_seed, int sub, int fuse, int tune, int one, hipStream_t st) {
  switch (D.M) {
    EP_SLAB_CASE(26) EP_SLAB_CASE(27) EP_SLAB_CASE(28) EP_SLAB_CASE(29) EP_SLAB_CASE(30) EP_SLAB_CASE(31) EP_SLAB_CASE(32)
    default:
      return hipErrorInvalidValue;
  }
}

}  // namespace ep
