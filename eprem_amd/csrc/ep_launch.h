// ep_launch.h -- host-side launchers of the kernels in ep_kernels.hip.
#pragma once
#include <hip/hip_runtime.h>

#include "ep_internal.h"

namespace ep {
hipError_t launch_stream_prep(const Dev &D, hipStream_t st);
hipError_t launch_node_prep(const Dev &D, hipStream_t st);
hipError_t launch_seed_check(const Dev &D, hipStream_t st);
hipError_t launch_stream_pass(const Dev &D, int pass, int flip, int do_seed, int substep, int chunks,
                              int fuse_focus, hipStream_t st);
// the same for M in 2..9, 10..17, 18..25, 26..32 (ep_stream_m*.hip)
hipError_t launch_stream_pass_m2_9(const Dev &D, int pass, int flip, int do_seed, int substep, int chunks,
                                   int fuse_focus, hipStream_t st);
hipError_t launch_stream_pass_m10_17(const Dev &D, int pass, int flip, int do_seed, int substep, int chunks,
                                     int fuse_focus, hipStream_t st);
hipError_t launch_stream_pass_m18_25(const Dev &D, int pass, int flip, int do_seed, int substep, int chunks,
                                     int fuse_focus, hipStream_t st);
hipError_t launch_stream_pass_m26_32(const Dev &D, int pass, int flip, int do_seed, int substep, int chunks,
                                     int fuse_focus, hipStream_t st);
// one persistent launch per EP substep over whole slabs (ep_stream.h,
// k_stream_slab); lane_nodes 2 or 3.  tune: state-preserving timing run;
// one >= 0: only stream `one` (the unbatched form)
hipError_t launch_stream_slab(const Dev &D, int do_seed, int sub, int fuse, int tune, int one, hipStream_t st);
hipError_t launch_stream_slab_m2_9(const Dev &D, int do_seed, int sub, int fuse, int tune, int one, hipStream_t st);
hipError_t launch_stream_slab_m10_17(const Dev &D, int do_seed, int sub, int fuse, int tune, int one, hipStream_t st);
hipError_t launch_stream_slab_m18_25(const Dev &D, int do_seed, int sub, int fuse, int tune, int one, hipStream_t st);
hipError_t launch_stream_slab_m26_32(const Dev &D, int do_seed, int sub, int fuse, int tune, int one, hipStream_t st);
hipError_t launch_reset_queue(const Dev &D, hipStream_t st);
hipError_t launch_task_prep(const Dev &D, hipStream_t st);
hipError_t launch_stream_finish(const Dev &D, int substep, hipStream_t st);
// fused: focusing of full streams was done by the streaming passes;
// spare: write the spare buffer and skip the heavy nodes (tile-size search)
// heavy = 0: leave the heavy nodes to a separate launch_focus_heavy
hipError_t launch_focus(const Dev &D, int substep, int fused, hipStream_t st, int spare = 0, int heavy = 1);
hipError_t launch_change(const Dev &D, int substep, hipStream_t st);
// focusing / change with the RK3 integrator, alg 2 (upwind) or 3 (WENO3),
// on every node (ep_rk.hip; FP64 storage)
hipError_t launch_focus_rk(const Dev &D, int alg, int substep, hipStream_t st);
hipError_t launch_change_rk(const Dev &D, int alg, int substep, hipStream_t st);
// focusing of the heavy nodes only (nsub_f >= kHeavyFocus), in place
hipError_t launch_focus_heavy(const Dev &D, int substep, hipStream_t st);
// focusing then change of every node in one pass (ep_node.hip); heavy
// nodes are only changed (launch_focus_heavy first)
hipError_t launch_focus_change(const Dev &D, int substep, int fused, int spare, hipStream_t st);
// streams s0 .. s0+count-1 (count < 0: to the last stream); the stage holds
// them from its start
hipError_t launch_from_reference(const Dev &D, const double *stage, double *dstbuf, hipStream_t st, int s0 = 0,
                                 int count = -1);
hipError_t launch_to_reference(const Dev &D, const double *srcbuf, double *stage, hipStream_t st, int s0 = 0,
                               int count = -1);
hipError_t launch_normalize(const Dev &D, hipStream_t st);
// to_fp32: FP64 -> float storage, else back; logf: the float side holds ln f
hipError_t launch_convert(const Dev &D, int to_fp32, int logf, hipStream_t st);
hipError_t launch_flip_all(const Dev &D, hipStream_t st);
hipError_t launch_shell_links(const Dev &D, hipStream_t st);
hipError_t launch_shell_diffuse(const Dev &D, int sub, hipStream_t st);
hipError_t launch_shell_drift(const Dev &D, int sub, hipStream_t st);
hipError_t launch_fill(double *p, size_t n, double v, hipStream_t st);
hipError_t launch_init_seed(const Dev &D, hipStream_t st);
hipError_t launch_dpp_probe(const double *in, double *out, double *out2, hipStream_t st);
hipError_t launch_copy_probe(const void *src, void *dst, size_t bytes, hipStream_t st);
hipError_t launch_reset_diag(DevDiag *d, hipStream_t st);
hipError_t launch_fastmath_probe(const double *tab, const double *x, int64_t n, double *lg, double *ex,
                                 hipStream_t st);
// learned-surrogate form (ep_sur.hip): the network applications of one
// substep (levels = the largest application count), their residuals and
// the accept / fall-back decision; the Dev copies the plain streaming and
// node stage then run with
hipError_t launch_sur_substep(const Dev &D, int substep, int levels, hipStream_t st);
// the network's scalar inputs of every kept node, once per global step
hipError_t launch_sur_prep(const Dev &D, hipStream_t st);
Dev sur_stream_dev(const Dev &D);
Dev sur_node_dev(const Dev &D);
int sur_check_step();
}  // namespace ep
