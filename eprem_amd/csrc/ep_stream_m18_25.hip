// ep_stream_m18_25.hip -- the streaming passes for 18 to 25 mu cells (one of
// four translation units, split by mu count so that they compile in
// parallel; the kernels are in ep_stream.h).
#include "ep_stream.h"

namespace ep {

hipError_t launch_stream_pass_m18_25(const Dev &D, int pass, int flip, int do_seed, int sub, int chunks,
                                   int fuse, hipStream_t st) {
  switch (D.M) {
    EP_PASS_CASE(18) EP_PASS_CASE(19) EP_PASS_CASE(20) EP_PASS_CASE(21) EP_PASS_CASE(22) EP_PASS_CASE(23) EP_PASS_CASE(24) EP_PASS_CASE(25)
    default:
      return hipErrorInvalidValue;
  }
}

}  // namespace ep
