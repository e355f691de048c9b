// ep_stream_m2_9.hip -- the streaming passes for 2 to 9 mu cells (one of
// four translation units, split by mu count so that they compile in
// parallel; the kernels are in ep_stream.h).
#include "ep_stream.h"

namespace ep {

hipError_t launch_stream_pass_m2_9(const Dev &D, int pass, int flip, int do_seed, int sub, int chunks,
                                   int fuse, hipStream_t st) {
  switch (D.M) {
    EP_PASS_CASE(2) EP_PASS_CASE(3) EP_PASS_CASE(4) EP_PASS_CASE(5) EP_PASS_CASE(6) EP_PASS_CASE(7) EP_PASS_CASE(8) EP_PASS_CASE(9)
    default:
      return hipErrorInvalidValue;
  }
}

}  // namespace ep
