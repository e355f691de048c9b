// ep_stream_m10_17.hip -- the streaming passes for 10 to 17 mu cells (one of
// four translation units, split by mu count so that they compile in
// parallel; the kernels are in ep_stream.h).
#include "ep_stream.h"

namespace ep {

hipError_t launch_stream_pass_m10_17(const Dev &D, int pass, int flip, int do_seed, int sub, int chunks,
                                   int fuse, hipStream_t st) {
  switch (D.M) {
    EP_PASS_CASE(10) EP_PASS_CASE(11) EP_PASS_CASE(12) EP_PASS_CASE(13) EP_PASS_CASE(14) EP_PASS_CASE(15) EP_PASS_CASE(16) EP_PASS_CASE(17)
    default:
      return hipErrorInvalidValue;
  }
}

}  // namespace ep
