/*
 * bpl_amd.h -- C ABI of the MI355X (gfx950) BP-Layers hot path.
 *
 * One shared library (libbpl_amd.so) exports the SAD matching-cost ops and the
 * min-sum loopy-belief-propagation ops of VLOGroup/bp-layers.  Every entry point
 * takes device pointers allocated by the caller, 64-bit sizes and a HIP stream
 * (passed as void*; NULL = the default stream), launches asynchronously on that
 * stream and returns 0 on success or a negative BPL_E* code.  The library never
 * allocates device memory; scratch space is passed in as a caller-owned workspace
 * whose size is queried first.  bpl_last_error() returns a thread-local message
 * for the last failing call on the calling thread.
 *
 * Tensor layouts are the reference's (all float32 and contiguous unless noted):
 *   features      [N,C,H,W]
 *   stereo cost   [N,H,W,D]             flow volumes [N,H,W,K]
 *   LBP cost      [N,H,W,C]             messages     [N,4,H,W,C]
 *   edge weights  [N,4,H,W]             message slot order LEFT=0 RIGHT=1 UP=2 DOWN=3
 *   jump          stereo [Nj,4,K,H,W] (Nj in {1,N}); semantic global [1,4,C,C];
 *                 semantic pixel [4,H,W,C,C]
 *   argmin        uint8 [N,4,H,W,C]: winning sender label per receiver label, or
 *                 255 when the value came from the truncation term, whose label is
 *                 then gargmin (uint8 [N,4,H,W]) at the same receiver slot.
 */
#ifndef BPL_AMD_H
#define BPL_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define BPL_OK 0
#define BPL_E_ARG -1      /* invalid size, pointer or parameter */
#define BPL_E_LIMIT -2    /* a size beyond what the kernels support */
#define BPL_E_HIP -3      /* a HIP runtime error (launch failure, ...) */

/* pairwise-term variants of the LBP op */
#define BPL_LBP_STEREO 0          /* ops/lbp_stereo: per-pixel K-vector over label differences */
#define BPL_LBP_SEMANTIC_GLOBAL 1 /* ops/lbp_semantic_pw: one C x C matrix per direction      */
#define BPL_LBP_SEMANTIC_PIXEL 2  /* ops/lbp_semantic_pw_pixel: C x C matrix per pixel+dir.  */

/* backward modes */
#define BPL_GRAD_REFERENCE 0  /* the reference's backward_minsum semantics              */
#define BPL_GRAD_EXACT 1      /* exact adjoint of one iteration (input-message gradient) */
/*
 * Flags OR-ed into BPL_GRAD_EXACT for a chain of iterations whose backward calls feed each
 * other (iteration t's grad_msg is iteration t-1's in_grad).  In the exact adjoint the
 * LEFT and RIGHT slots of grad_msg are zero away from the row ends (the forward reads those
 * input slots only at x = 0 and x = W-1), so the chain need not write or read them:
 *   BPL_GRAD_IN_H_ROW_ENDS   in_grad's LEFT slot is read only at x = 0 and its RIGHT slot
 *                            only at x = W-1; everywhere else they are taken as zero (their
 *                            memory is not read);
 *   BPL_GRAD_OUT_H_ROW_ENDS  grad_msg's LEFT slot is written only at x = 0 and its RIGHT
 *                            slot only at x = W-1; everywhere else their memory is left as
 *                            it was (the exact value there is zero).
 * A grad_msg written with OUT_H_ROW_ENDS is a valid in_grad for IN_H_ROW_ENDS.  No
 * counterpart in the reference (whose grad_messages is a copy of in_grad).
 */
#define BPL_GRAD_IN_H_ROW_ENDS 0x10
#define BPL_GRAD_OUT_H_ROW_ENDS 0x20
/*
 * Likewise for the UP and DOWN slots, which are equal in the exact adjoint away from rows 0
 * and H-1 (both are the horizontal passes' h-gradient; only the pass-through of a vertical
 * pass's first pixel adds to one of them there):
 *   BPL_GRAD_OUT_UD_SHARED   grad_msg's DOWN slot is written only at rows 0 and H-1; its
 *                            other rows are left as they were (they equal the UP slot's);
 *   BPL_GRAD_IN_UD_SHARED    in_grad's DOWN slot is read only at rows 0 and H-1; its other
 *                            rows are read from the UP slot.
 * A grad_msg written with OUT_UD_SHARED is a valid in_grad for IN_UD_SHARED.
 */
#define BPL_GRAD_OUT_UD_SHARED 0x80
#define BPL_GRAD_IN_UD_SHARED 0x100
/*
 * BPL_GRAD_ACCUMULATE, OR-ed into either mode: grad_cost, grad_jump and grad_edge are added
 * to (their contents plus this iteration's gradient) instead of overwritten, so the calls of
 * a chain of iterations sum the parameter gradients in place -- what autograd does across the
 * reference's iterations (ops/lbp_stereo/message_passing_op_cuda.py:165-167).  grad_msg is
 * written as without the flag.  Kinds BPL_LBP_STEREO and BPL_LBP_SEMANTIC_GLOBAL.
 */
#define BPL_GRAD_ACCUMULATE 0x40

/* Library identification and diagnostics. */
int bpl_version(void);
const char *bpl_last_error(void);

/*
 * Kernel timing (instrumentation; no counterpart in the reference, for benchmarks and
 * profiling tools).  While enabled, every kernel launch of the library is bracketed by a
 * pair of HIP events recorded on the launch stream.  bpl_timing_read() waits for the
 * recorded events and returns the accumulated duration and launch count of one kernel
 * class; bpl_timing_reset() clears the totals.  Disabled by default; when enabled it adds
 * two event records per launch and no synchronisation.
 */
#define BPL_KERNEL_LBP_FWD_H 0      /* forward, horizontal phase (RIGHT and LEFT passes) */
#define BPL_KERNEL_LBP_FWD_V 1      /* forward, vertical phase (DOWN and UP passes)      */
#define BPL_KERNEL_LBP_BWD_V 2      /* backward, vertical phase                           */
#define BPL_KERNEL_LBP_BWD_H 3      /* backward, horizontal phase                         */
#define BPL_KERNEL_LBP_BWD_COMBINE 4
#define BPL_KERNEL_LBP_BWD_REDUCE 5 /* semantic-global matrix gradient reduction          */
#define BPL_KERNEL_SAD_STEREO_FWD 6
#define BPL_KERNEL_SAD_STEREO_BWD 7
#define BPL_KERNEL_SAD_FLOW_FWD 8
#define BPL_KERNEL_SAD_FLOW_BWD 9
#define BPL_KERNEL_SOFTMAX 10
#define BPL_KERNEL_BELIEFS 11
#define BPL_KERNEL_SAD_STEREO_SOFTMAX 12
#define BPL_KERNEL_COUNT 13
int bpl_timing_enable(int on);
int bpl_timing_reset(void);
int bpl_timing_read(int kernel, double *total_ms, int64_t *launches);
const char *bpl_timing_name(int kernel);

/*
 * Stereo SAD cost volume.  Replaces pytorch_cuda_stereo_sad_op.forward
 * (reference ops/sad/src/stereo_sad.cpp:10-15; kernel ops/sad/src/stereo_sad_kernel.cu:16-62).
 * cost[n,y,x,i] = sum_c |f0[n,c,y,x] - f1[n,c,y,x-d]|, d = int(i*step) + min_disp,
 * or 40 where x-d is outside the image; D = int((max_disp-min_disp+1)/step).
 */
int bpl_sad_stereo_fwd(const float *f0, const float *f1, float *cost,
                       int64_t N, int64_t C, int64_t H, int64_t W,
                       int min_disp, int max_disp, float step, void *stream);

/*
 * Stereo SAD (step 1) with the matching softmax fused in: y = sign * softmax(scale * SAD) over
 * the D = max_disp - min_disp + 1 disparities of each pixel, y [N,H,W,D], D <= 256.  Replaces
 * the composition StereoMatchingSadFunction -> -cost -> TemperatureSoftmax of the reference's
 * StereoMatchingSad.forward (matching.py:22-25, :45-47; corenet.py:97-104) with scale = -1/T
 * and sign = 1 (sign = -1 gives the message-passing cost -prob,
 * ops/lbp_stereo/message_passing_op_cuda.py:162).  The SAD volume stays on chip; y equals
 * bpl_sad_stereo_fwd followed by bpl_softmax_labels bit for bit.
 */
int bpl_sad_stereo_softmax_fwd(const float *f0, const float *f1, float *y,
                               int64_t N, int64_t C, int64_t H, int64_t W,
                               int min_disp, int max_disp, float scale, float sign, void *stream);

/*
 * Gradient of the stereo SAD (step 1).  Replaces pytorch_cuda_stereo_sad_op.backward
 * (reference ops/sad/src/stereo_sad.cpp:17-24; kernel stereo_sad_kernel.cu:64-108).
 * Deterministic gather over disparities; df0/df1 are fully overwritten.
 */
int bpl_sad_stereo_bwd(const float *f0, const float *f1, const float *grad_cost,
                       float *df0, float *df1,
                       int64_t N, int64_t C, int64_t H, int64_t W,
                       int min_disp, int max_disp, void *stream);

/*
 * Fused flow SAD over a (sws+1)^2 window with min-projections.  Replaces
 * pytorch_cuda_flow_mp_sad_op.forward (reference ops/flow_mp_sad/src/flow_mp_sad.cpp:10-16;
 * kernel flow_mp_sad_kernel.cu:10-124).  Outputs [N,H,W,sws+1]; argmins as float label ids.
 */
int bpl_sad_flow_fwd(const float *f0, const float *f1,
                     float *cv_u, float *cv_v, float *u_star, float *v_star,
                     int64_t N, int64_t C, int64_t H, int64_t W,
                     int sws, int offset_u, int offset_v, int block_u, int block_v,
                     void *stream);

/*
 * Gradient of the flow SAD through the stored argmins.  Replaces
 * pytorch_cuda_flow_mp_sad_op.backward (reference flow_mp_sad.cpp:18-30; kernel
 * flow_mp_sad_kernel.cu:126-200).  Deterministic gather (no float atomics).
 */
int bpl_sad_flow_bwd(const float *f0, const float *f1, const float *grad_u,
                     const float *grad_v, const float *u_star, const float *v_star,
                     float *df0, float *df1,
                     int64_t N, int64_t C, int64_t H, int64_t W, int sws, void *stream);

/*
 * One min-sum LBP iteration: passes RIGHT, LEFT, DOWN, UP.  Replaces
 * pytorch_cuda_lbp_op.forward_minsum of the three reference builds
 * (ops/lbp_{stereo,semantic_pw,semantic_pw_pixel}/src/lbp.cpp:16-27; kernels
 * lbp_min_sum_kernel.cu:9-213 and host loop :424-500).
 * msg_in and msg_out may be the same buffer (in-place, as in the reference).
 * jump_k is K for the stereo variant and ignored otherwise; jump_n is the leading
 * dimension of a stereo jump (1 = shared by the batch, or N).
 * scale (float [N,4,H,W], sum_c exp(message)) may be NULL; gargmin may be NULL when
 * delta >= C (no truncation term); argmin and gargmin may both be NULL for a forward-only
 * solve (no backward possible).
 */
int bpl_lbp_fwd(int kind, const float *cost, const float *jump, int64_t jump_k, int64_t jump_n,
                const float *edge, const float *msg_in, float *msg_out,
                uint8_t *argmin, uint8_t *gargmin, float *scale,
                int64_t N, int64_t H, int64_t W, int64_t C, int delta, void *stream);

/*
 * bpl_lbp_fwd with a workspace; no counterpart in the reference (a faster form of the same
 * iteration, for its callers that can provide device memory).  For the stereo kind with
 * jump_k == 6 and delta == 3 the horizontal phase also writes the plane
 * (cost + M_R) + M_L into the workspace and the vertical phase reads it in place of the cost
 * and both horizontal messages (16 bytes per pixel-label fewer reads): outputs are bit for bit
 * those of bpl_lbp_fwd.  Other calls, or a NULL workspace, run bpl_lbp_fwd.  The workspace
 * must hold bpl_lbp_fwd_workspace_size bytes and overlap no tensor of the call.
 * flags: BPL_FWD_MSG_IN_ZERO declares msg_in all zeros (the first iteration from zero
 * messages): the P1/P2 sweeps then skip reading it (16 bytes per pixel-label fewer in the
 * horizontal phase), with the same results; a msg_in that is not zero gives undefined results.
 */
#define BPL_FWD_MSG_IN_ZERO 0x1
int64_t bpl_lbp_fwd_workspace_size(int kind, int64_t N, int64_t H, int64_t W, int64_t C);
int bpl_lbp_fwd_ws(int kind, const float *cost, const float *jump, int64_t jump_k, int64_t jump_n,
                   const float *edge, const float *msg_in, float *msg_out,
                   uint8_t *argmin, uint8_t *gargmin, float *scale,
                   int64_t N, int64_t H, int64_t W, int64_t C, int delta, void *workspace,
                   int64_t workspace_bytes, int flags, void *stream);

/* Workspace bytes needed by bpl_lbp_bwd for these sizes. */
int64_t bpl_lbp_bwd_workspace_size(int kind, int64_t N, int64_t H, int64_t W, int64_t C);

/*
 * Backward of one LBP iteration.  Replaces pytorch_cuda_lbp_op.backward_minsum
 * (reference ops/lbp_stereo/src/lbp.cpp:29-46 and its semantic twins; kernels
 * lbp_min_sum_kernel.cu:215-417, host loop :513-621).
 * messages: the message tensor the probabilities exp(m) are read from (the
 * reference passes the in-place-updated tensor); argmin/gargmin as written by
 * bpl_lbp_fwd.  Outputs are fully overwritten (or, BPL_GRAD_ACCUMULATE, the first three
 * added to): grad_cost [N,H,W,C]; grad_jump
 * (shape of jump); grad_edge [N,4,H,W] (SEMANTIC_PIXEL in the reference mode: zeros, as
 * the reference, and jump may be NULL); grad_msg [N,4,H,W,C] (BPL_GRAD_REFERENCE: a copy
 * of in_grad; BPL_GRAD_EXACT: the gradient w.r.t. the iteration's input messages).
 * In the reference mode grad_msg may be in_grad itself: nothing is then written to it (an
 * autograd caller can pass its incoming gradient through instead of copying it).
 * grad_msg may be NULL when the input-message gradient is not wanted (the first iteration
 * from constant initial messages): the other outputs are unchanged and no message-gradient
 * bytes are written.
 * delta: the delta the forward ran with, or -1 if unknown (an extra argument of this ABI,
 * not of the reference).  It only selects kernels: a value in [0,3] lets the stereo kernel
 * assume every non-truncation argmin lies within 3 labels of its receiver, which holds
 * for argmins written by bpl_lbp_fwd with that delta where the window offsets are zero.
 */
int bpl_lbp_bwd(int kind, int mode, const float *jump, int64_t jump_k, int64_t jump_n,
                const float *edge, const float *in_grad, const float *messages,
                const uint8_t *argmin, const uint8_t *gargmin,
                float *grad_cost, float *grad_jump, float *grad_edge, float *grad_msg,
                void *workspace, int64_t workspace_bytes,
                int64_t N, int64_t H, int64_t W, int64_t C, int delta, void *stream);

/*
 * bpl_lbp_bwd without a workspace argument: no counterpart in the reference.  For callers that
 * cannot allocate device memory (code under HIP-graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  The workspace is a scratch array of BPL_LBP_BWD_NOALLOC_MAX_BYTES
 * built into the library, one per GPU (held from the library's first use on that GPU until the
 * process ends); sizes whose bpl_lbp_bwd_workspace_size exceeds it are rejected with
 * BPL_E_LIMIT.  The entry point calls no allocating routine.  It uses the array of the current
 * device, so the stream must belong to that device.  Every call on one GPU shares that array:
 * calls that overlap on one GPU (on different streams, or from different threads) race and
 * must be ordered by the caller.  Results are those of bpl_lbp_bwd, bit for bit.
 */
#define BPL_LBP_BWD_NOALLOC_MAX_BYTES (128LL << 20)
int bpl_lbp_bwd_noalloc(int kind, int mode, const float *jump, int64_t jump_k, int64_t jump_n,
                        const float *edge, const float *in_grad, const float *messages,
                        const uint8_t *argmin, const uint8_t *gargmin,
                        float *grad_cost, float *grad_jump, float *grad_edge, float *grad_msg,
                        int64_t N, int64_t H, int64_t W, int64_t C, int delta, void *stream);

/*
 * y = sign * softmax(scale * x) over the contiguous last dimension of a [rows, C] tensor
 * (C <= 256).  With scale = -1/T and sign = -1 this is the reference's matching softmax
 * softmax(-SAD/T) (matching.py:22-25, corenet.py:97-104) followed by the LBP input
 * cost = -prob (ops/lbp_stereo/message_passing_op_cuda.py:162).  One wave per row.
 */
int bpl_softmax_labels(const float *x, float *y, int64_t rows, int64_t C, float scale, float sign,
                       void *stream);

/*
 * Beliefs softmin((cost + sum_s messages_s) / T) over labels (reference
 * ops/lbp_stereo/message_passing_op_cuda.py:170-173) into beliefs [N,H,W,C] (may be NULL)
 * and the winner-take-all label argmax_c (first index, inference_op.py:20-24) into
 * label [N,H,W] as float (may be NULL).  C <= 256.
 */
int bpl_lbp_beliefs(const float *cost, const float *messages, float *beliefs, float *label,
                    int64_t N, int64_t H, int64_t W, int64_t C, float T, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* BPL_AMD_H */
