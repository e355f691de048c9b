// ep_stream_m26_32.hip -- the streaming passes for 26 to 32 mu cells (one of
// four translation units, split by mu count so that they compile in
// parallel; the kernels are in ep_stream.h).
#include "ep_stream.h"

namespace ep {

hipError_t launch_stream_pass_m26_32(const Dev &D, int pass, int flip, int do_seed, int sub, int chunks,
                                   int fuse, hipStream_t st) {
  switch (D.M) {
    EP_PASS_CASE(26) EP_PASS_CASE(27) EP_PASS_CASE(28) EP_PASS_CASE(29) EP_PASS_CASE(30) EP_PASS_CASE(31) EP_PASS_CASE(32)
    default:
      return hipErrorInvalidValue;
  }
}

}  // namespace ep
