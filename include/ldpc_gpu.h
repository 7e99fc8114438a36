/*
 * MI355X (gfx950) LDPC decoder for DVB-S2 / S2X / T2 codes: C-ABI.
 *
 * Layered offset min-sum on int8 LLRs, bit-exact with libdvbrx's CPU
 * decoder (libdvbrx/src/ldpc.c), run by hand-written HIP kernels.  The
 * reference decodes one batch of <= 64 frames of one code per call
 * (drx_ldpc_decode, libdvbrx/include/dvbrx/ldpc.h:32-36, one frame per SIMD
 * byte lane); this interface takes any number of frames of mixed codes in
 * one call, described by a frame table, with per-frame results.  The
 * reference's four drx_ldpc_* entry points are provided on top of it (see
 * include/dvbrx/ldpc.h).
 *
 * All pointers named d_* are device memory.  Plain C, no HIP types: the
 * stream is handed out as an opaque pointer (a hipStream_t).
 */
#ifndef LDPC_GPU_H
#define LDPC_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Error codes (negative). */
#define LDPC_GPU_OK 0
#define LDPC_GPU_EINVAL (-1)   /* bad argument (unknown code, misaligned offset, ...) */
#define LDPC_GPU_ENOMEM (-2)   /* device or host allocation failed */
#define LDPC_GPU_EDEVICE (-3)  /* HIP runtime error; see ldpc_gpu_last_error() */

typedef struct ldpc_gpu_ctx ldpc_gpu_ctx;

/* One frame of a batch.  Frames are independent unless they share a group. */
typedef struct {
    int32_t code;     /* ldpc_gpu_code_lookup() */
    int32_t max_iter; /* iteration cap, >= 0 (the reference's max_iter) */
    int32_t nbits;    /* decoded bits written (kldpc), 1..K; 0 means K */
    int32_t group;    /* -1: the frame stops when its own syndrome holds;
                         >= 0: all frames with this group number iterate until
                         every one of them holds its syndrome at the same
                         iteration, as the frames of one drx_ldpc_decode call
                         (libdvbrx/src/ldpc.c:627-637).  Frames of a group
                         must share max_iter. */
    uint64_t llr_off; /* byte offset of the frame's N LLRs in d_llr (multiple of 8) */
    uint64_t out_off; /* byte offset of its ceil(nbits/8) output bytes in d_out */
} ldpc_gpu_frame;

/* Context: device, stream, device-resident code tables and scratch.  One
 * thread at a time per context; distinct contexts are independent. */
int ldpc_gpu_create(int device, ldpc_gpu_ctx** out);
void ldpc_gpu_destroy(ldpc_gpu_ctx* ctx);
void* ldpc_gpu_stream(ldpc_gpu_ctx* ctx); /* hipStream_t */
const char* ldpc_gpu_last_error(ldpc_gpu_ctx* ctx);

/* Codes.  ldpc_gpu_code_lookup replaces find_table (libdvbrx/src/ldpc.c:188-197):
 * (frame size, rate name as drx_fec_lookup: "1_2", "2_3_t2", ...) -> code
 * id, or -1.  ldpc_gpu_fec_lookup is drx_fec_lookup (libdvbrx/src/bch.c:25-41). */
int ldpc_gpu_code_count(void);
int ldpc_gpu_code_lookup(int short_frame, const char* rate);
int ldpc_gpu_code_params(int code, int* n, int* k, int* q, int* links_total, int* level_steps);
/* Largest information degree of a layer of the code (check degree - 2). */
int ldpc_gpu_code_max_degree(int code);
/* Kernel instance that decodes the code with per-frame early stop, as the
 * largest layer information degree it covers (2 x slots per lane - 2),
 * plus 100 for the instance that also takes layers several slots shorter;
 * the key of the launches ldpc_gpu_last_launches reports.  LDPC_GPU_EINVAL
 * for an unknown code.  Diagnostics only, no counterpart in the
 * reference. */
int ldpc_gpu_code_kernel_degree(int code);
int ldpc_gpu_fec_lookup(int short_frame, const char* rate, int* nldpc, int* kldpc, int* kbch, int* t);

/* Device memory helpers (synchronous copies on the context's stream). */
void* ldpc_gpu_malloc(ldpc_gpu_ctx* ctx, size_t bytes);
int ldpc_gpu_free(ldpc_gpu_ctx* ctx, void* d_ptr);
int ldpc_gpu_upload(ldpc_gpu_ctx* ctx, void* d_dst, const void* src, size_t bytes);
int ldpc_gpu_download(ldpc_gpu_ctx* ctx, void* dst, const void* d_src, size_t bytes);
int ldpc_gpu_synchronize(ldpc_gpu_ctx* ctx);
/* Page-locked host memory and copies enqueued on the context's stream
 * without waiting (no counterpart in the reference, whose receiver hands
 * its FEC stage host buffers): a receiver keeps its input and result
 * buffers pinned, so their transfers run at full link rate and overlap
 * other contexts' decoding.  A buffer given to an async copy must stay as
 * it is (upload) or unread (download) until the stream has passed the
 * copy (ldpc_gpu_query returns 1, or ldpc_gpu_synchronize).  The async
 * copies accept pageable memory too, then with the runtime's staging. */
void* ldpc_gpu_host_alloc(ldpc_gpu_ctx* ctx, size_t bytes);
int ldpc_gpu_host_free(ldpc_gpu_ctx* ctx, void* h_ptr);
int ldpc_gpu_upload_async(ldpc_gpu_ctx* ctx, void* d_dst, const void* src, size_t bytes);
int ldpc_gpu_download_async(ldpc_gpu_ctx* ctx, void* dst, const void* d_src, size_t bytes);
/* Non-blocking completion test (no counterpart in the reference, whose
 * decode call returns when the batch is done): 1 if every call enqueued on
 * the context has finished, 0 if work is pending, < 0 on error.  Lets one
 * host thread keep batches in flight on several contexts at once. */
int ldpc_gpu_query(ldpc_gpu_ctx* ctx);

/*
 * Decode a batch (replaces drx_ldpc_decode, libdvbrx/src/ldpc.c:614-638).
 *
 * d_llr:   int8 channel LLRs, codeword order, positive for a 0 bit.
 * d_out:   hard decisions of the first nbits bits, packed MSB first.
 * d_iters: per frame, iterations run (for a group: the group's count).
 * d_synd:  per frame, 1 if every parity check holds at the end.
 *
 * Without groups the call only enqueues work on the context's stream
 * (results valid after ldpc_gpu_synchronize); with groups it returns
 * after the batch is decoded.  Returns LDPC_GPU_OK or an error code.
 */
int ldpc_gpu_decode_batch(ldpc_gpu_ctx* ctx, const ldpc_gpu_frame* frames, int nframes,
                          const int8_t* d_llr, uint8_t* d_out, int32_t* d_iters, uint8_t* d_synd);

/*
 * ldpc_gpu_decode_batch without any allocation: no counterpart in the
 * reference; for callers that cannot allocate device or host memory
 * during decoding (tight serving loops, harnesses that forbid it).  Same
 * arguments and results as ldpc_gpu_decode_batch for frames with their own
 * early stop (every group -1; groups are rejected with LDPC_GPU_EINVAL).
 *
 * The call's device scratch (frame table, lists, counters and the
 * decoder's message blocks) is a fixed 96 MiB array of the library on each
 * device, set up when a context is created there and held until the
 * process ends; at most 65536 frames per call (LDPC_GPU_EINVAL beyond).
 * The launches run one after the other on the context's stream with as
 * many workgroups as that scratch holds, so large batches decode more
 * slowly than with ldpc_gpu_decode_batch.  Calls on one device share the
 * array: two calls whose work overlaps on the device (different contexts
 * or streams of one GPU) race; calls on one context are ordered.  The
 * frame table is copied before the call returns; the decode is enqueued on
 * the context's stream (results valid after ldpc_gpu_synchronize).
 */
int ldpc_gpu_decode_batch_static(ldpc_gpu_ctx* ctx, const ldpc_gpu_frame* frames, int nframes, const int8_t* d_llr,
                                 uint8_t* d_out, int32_t* d_iters, uint8_t* d_synd);

/* Same on host buffers: copies in, decodes, copies out (synchronous). */
int ldpc_gpu_decode_host(ldpc_gpu_ctx* ctx, const ldpc_gpu_frame* frames, int nframes,
                         const int8_t* llr, size_t llr_bytes, uint8_t* out, size_t out_bytes,
                         int32_t* iters, uint8_t* synd);

/*
 * Learned offset min-sum decoding (no counterpart in the reference, whose
 * offset is fixed at 1, libdvbrx/src/ldpc.c:353; the tables come from
 * include/ldpc_nms.h's trainer).  ldpc_gpu_set_nms stores, for one code,
 * T rows (T <= LDPC_NMS_MAX_ITERS) of LDPC_NMS_DEGS scales S (0..63) and
 * offsets B (-1024..1023), both in 1/16 units, indexed by check degree;
 * the check node's magnitude becomes
 * clamp((min(m, 127) * S[t][dc] - B[t][dc]) >> 4, 0, 127) at iteration t
 * (row min(t, T-1)).  T = 0 removes the table.
 *
 * ldpc_gpu_decode_batch_nms decodes frames (group -1 only) with their
 * code's table for up to frames[f].max_iter learned iterations; a frame
 * whose syndrome still fails is decoded again from its LLRs by the plain
 * decoder for up to fallback_max_iter iterations (0: no fallback), in the
 * same kernel.  d_iters[f] = learned + plain iterations; d_fallback (may
 * be NULL) = 1 where the fallback ran.  Enqueues only, as
 * ldpc_gpu_decode_batch.
 */
int ldpc_gpu_set_nms(ldpc_gpu_ctx* ctx, int code, int T, const uint8_t* S, const int16_t* B);
int ldpc_gpu_decode_batch_nms(ldpc_gpu_ctx* ctx, const ldpc_gpu_frame* frames, int nframes, const int8_t* d_llr,
                              uint8_t* d_out, int32_t* d_iters, uint8_t* d_synd, int fallback_max_iter,
                              uint8_t* d_fallback);

/*
 * Prepared batches (no counterpart in the reference, whose batches are
 * built per call): a fixed frame table (group -1 only) bound to fixed
 * device buffers, prepared once - frame table, launch layout and scratch
 * of its own - with its launches captured as one hipGraph that
 * ldpc_gpu_plan_launch replays on the context's stream (enqueue only).
 * The caller refills d_llr between launches; results as
 * ldpc_gpu_decode_batch.
 *
 * ldpc_gpu_plan_tune is the autotuner: it times the shapes common to all
 * launches (the persistent grid per launch, i.e. frames in flight, at
 * 100 / 75 / 50 % of what fits; the launches over 2, 3 or 4 forked streams
 * or one after another) over reps replays, keeps the fastest, then tunes each launch (one per kernel
 * instance and frame size) in turn, longest first: its workgroups per CU
 * (1, or 2 where built: plain decoding of short frames on instances 2, 4,
 * 6 and 106) and its grid scale, keeping a change where the whole replay
 * gets faster by more than 0.5 %.  ldpc_gpu_plan_info reports the start
 * variant chosen (-1 once a launch changed), per variant its grid scale,
 * its forked stream count (0: one after another; the `concurrent` array)
 * and its timing (ms per replay, < 0 if not timed); ldpc_gpu_plan_shape the shape of each
 * launch: kernel instance (the bound ldpc_gpu_last_launches reports),
 * short frames or not, workgroups per CU, grid scale, workgroups launched
 * and frames; it returns the count.
 */
typedef struct ldpc_gpu_plan ldpc_gpu_plan;
int ldpc_gpu_plan_create(ldpc_gpu_ctx* ctx, const ldpc_gpu_frame* frames, int nframes, const int8_t* d_llr,
                         uint8_t* d_out, int32_t* d_iters, uint8_t* d_synd, ldpc_gpu_plan** out);
int ldpc_gpu_plan_launch(ldpc_gpu_plan* plan);
int ldpc_gpu_plan_tune(ldpc_gpu_plan* plan, int reps);
float ldpc_gpu_plan_last_ms(ldpc_gpu_plan* plan);
int ldpc_gpu_plan_info(ldpc_gpu_plan* plan, int* variant, float* scale, int* concurrent, float* ms, int max);
int ldpc_gpu_plan_shape(ldpc_gpu_plan* plan, int* di, int* short_frames, int* occ, float* scale, int* grid,
                        int* frames, int max);
/* Use launch shape `variant` (0 .. count returned by ldpc_gpu_plan_info - 1)
 * for all launches. */
int ldpc_gpu_plan_select(ldpc_gpu_plan* plan, int variant);
void ldpc_gpu_plan_destroy(ldpc_gpu_plan* plan);

/*
 * Max-log demapping of received symbols into the decoder's input (replaces
 * drx_demap_frame, libdvbrx/src/demap.c:491-532, which demaps one frame
 * per call; drx_demap_new's constellation lookup, :164-203, is by PLS code
 * here).  nframes frames of MODCOD pls (PLS code, pilot bit ignored):
 *   d_sym  nframes x ldpc_gpu_demap_symbols(pls) complex symbols as (re, im)
 *          float pairs, at unit average energy (the AGC's output);
 *   d_n0   nframes noise variances;
 *   d_llr  nframes x nldpc int8 LLRs (64800 or 16200 each, back to back),
 *          codeword order (bit interleaver undone), positive for a 0 bit.
 * Identical LLRs to the reference's AVX2/FMA build.  Enqueues only.
 */
int ldpc_gpu_demap_symbols(int pls);
int ldpc_gpu_demap(ldpc_gpu_ctx* ctx, int pls, int nframes, const float* d_sym, const float* d_n0, int8_t* d_llr);

/*
 * BCH decoding of LDPC decoder output (replaces drx_bch_decode,
 * libdvbrx/src/bch.c:189-290, one codeword per call there; drx_bch_new's
 * per-FEC set-up, :106-172, is by (frame size, rate) here, rate names as
 * drx_fec_lookup).  nframes codewords, each nbch (= kldpc) bits packed MSB
 * first at d_cw + f * stride (stride >= nbch / 8 bytes: the decoder's
 * output with nbits = kldpc), corrected in place.  d_result[f]: the number
 * of corrected bits, 0 for a codeword, -1 if uncorrectable (the codeword is
 * left as it was), as the reference returns.  Enqueues only.
 */
int ldpc_gpu_bch_decode(ldpc_gpu_ctx* ctx, int short_frame, const char* rate, int nframes, uint8_t* d_cw,
                        size_t stride, int32_t* d_result);

/*
 * DVB-T2 FEC blocks of one PLP (the reference's FEC thread, decode_job,
 * libdvbrx/src/dvbt2.c:2188-2400, and the per-PLP maps of setup_plp_fec,
 * :1953-2082; there one frame's blocks per call from a pthread, here a
 * whole batch per launch).  The PLP: cod 0..5 (1/2 3/5 2/3 3/4 4/5 5/6),
 * mod 0..3 (QPSK, 16-, 64-, 256-QAM), fec_type 1 normal / 0 short,
 * rotation 0/1, and the time interleaver's length and type (with type 0
 * the cell deinterleaver's cyclic shift restarts with each TI-block), as
 * in L1-post (struct plp, dvbt2.c:216-221).
 */
typedef struct {
    int32_t cod, mod, fec_type, rotation, til_length, til_type;
} ldpc_gpu_t2_plp;

/* nldpc, kldpc, kbch and cells per block; LDPC_GPU_EINVAL if unsupported
 * (the checks of setup_plp_fec, dvbt2.c:1963). */
int ldpc_gpu_t2_geometry(const ldpc_gpu_t2_plp* plp, int* nldpc, int* kldpc, int* kbch, int* ncells);

/* The maps the front end applies, computed on the host (no device call):
 * perm (ncells entries) the cell interleaver's L0 (setup_plp_fec,
 * dvbt2.c:2052-2081), bitsrc (nldpc entries) for codeword bit n the soft
 * value cell * m + plane it takes (plane 2j: bit j of I, 2j + 1: bit j of
 * Q; dvbt2.c:1969-2050, :2364-2385), shift (nblocks entries) each block's
 * cyclic shift (dvbt2.c:2230-2254).  Null pointers are skipped.  For
 * transmit-side models and checks. */
int ldpc_gpu_t2_maps(const ldpc_gpu_t2_plp* plp, int nblocks, int32_t* perm, uint16_t* bitsrc, int32_t* shift);

/*
 * Front end (dvbt2.c:2213-2386): nblocks blocks of ncells time-
 * deinterleaved cells, d_cells as (re, im) float pairs (8-byte aligned)
 * and d_csi their channel state -> d_llr nblocks x nldpc int8 soft values
 * in codeword order (4-byte aligned; positive for a 0 bit), d_err/d_sig
 * the MER error and signal sums of each block.  Cell deinterleaving,
 * derotation with per-axis noise weights, the noise estimate, max-log
 * soft values and bit deinterleaving: identical to the reference's x86-64
 * FMA build.  Enqueues only.
 */
int ldpc_gpu_t2_demap(ldpc_gpu_ctx* ctx, const ldpc_gpu_t2_plp* plp, int nblocks, const float* d_cells,
                      const float* d_csi, int8_t* d_llr, float* d_err, float* d_sig);

/*
 * The whole job: the front end into d_llr (nblocks x nldpc, 8-byte
 * aligned), LDPC decoding with max_iter (the reference's FEC thread uses
 * 30, 15 when frames queue up, dvbt2.c:127-129, :2457) into d_cw (kldpc / 8
 * bytes a block, packed MSB first; d_iters and d_synd per block as
 * ldpc_gpu_decode_batch), then BCH in place: d_bch[b] >= 0 is the
 * reference's job->ok, and the block's BBFRAME is its first kbch / 8
 * bytes.  lanes > 0 decodes blocks b / lanes together as one call of the
 * reference's decoder does (32 in its AVX2 build, 64 with AVX-512BW), for
 * bit-identical output; lanes <= 0 gives every block its own early stop.
 * Returns after decoding when lanes > 0, else enqueues.
 */
int ldpc_gpu_t2_decode(ldpc_gpu_ctx* ctx, const ldpc_gpu_t2_plp* plp, int nblocks, const float* d_cells,
                       const float* d_csi, int max_iter, int lanes, int8_t* d_llr, uint8_t* d_cw, int32_t* d_iters,
                       uint8_t* d_synd, int32_t* d_bch, float* d_err, float* d_sig);

/*
 * decode_job's own interface on host memory (dvbt2.c:2188-2400): the job's
 * cells and channel state in (job->buf, job->csi: nblocks x ncells (re, im)
 * floats and floats), its results out: bb nblocks x kbch / 8 BBFRAME bytes
 * (job->bb), ok[b] = 1 when BCH decoded block b (job->ok), err / sig the
 * MER sums (job->err, job->sig).  Copies in, runs ldpc_gpu_t2_decode,
 * copies out; synchronous.  Device buffers are allocated per call.
 */
int ldpc_gpu_t2_decode_host(ldpc_gpu_ctx* ctx, const ldpc_gpu_t2_plp* plp, int nblocks, const float* cells,
                            const float* csi, int max_iter, int lanes, uint8_t* bb, uint8_t* ok, float* err,
                            float* sig);

/* Kernel choice for frames with their own early stop (group -1):
 * LDPC_GPU_KERNEL_AUTO runs two frames of one code per workgroup in packed
 * 16-bit arithmetic; LDPC_GPU_KERNEL_SINGLE one frame per workgroup (the
 * kernel that also serves groups).  Both give identical results. */
#define LDPC_GPU_KERNEL_AUTO 0
#define LDPC_GPU_KERNEL_SINGLE 1
int ldpc_gpu_set_kernel(ldpc_gpu_ctx* ctx, int kernel);

/* Frames in groups (group >= 0, the reference's lock-step batches) are
 * decoded in two passes of the two-frame kernel: each frame with its own
 * early stop, then every frame again to its group's count (the largest of
 * its frames' first passing iterations, which no group can stop before),
 * checked there.  Where a group's frames do not all hold their checks at
 * that count (a converged frame lost a check, rare), the call is decoded
 * again by the exact one-frame path.  The count of such calls on this
 * context (no counterpart in the reference).  LDPC_GPU_KERNEL_SINGLE takes
 * the exact path directly. */
int ldpc_gpu_group_fallbacks(ldpc_gpu_ctx* ctx);

/* Launch mode of ldpc_gpu_decode_batch (no counterpart in the reference).
 * 1 (the default): the launches of one call - one per kernel instance and
 * frame size - run concurrently, spread by expected work over three forked
 * streams (one hardware queue each at GPU_MAX_HW_QUEUES 4).  2 .. 4: over
 * that many (4 suits batches of about four launches, as a receiver's
 * mixed batches are).  0: one after the other on the context's stream, so
 * that the context occupies a single hardware queue.  Plans created
 * afterwards use the same count for their concurrent shapes. */
int ldpc_gpu_set_concurrent(ldpc_gpu_ctx* ctx, int concurrent);

/* Launch shape of the context's decodes (no counterpart in the reference):
 * for the launch of kernel instance di (the bound ldpc_gpu_last_launches
 * reports) over short (short_frames 1) or normal frames, the workgroups per
 * CU it is built for (1, or 2 for plain decoding on instances 2, 4, 6 and
 * 106: two frames more in flight per CU; learned decoding keeps 1) and the
 * share of the resident workgroups it launches (0 < scale <= 1).  Default:
 * 2 for the short-frame launches of those instances, 1 elsewhere, scale 1.  Used by
 * ldpc_gpu_decode_batch, ldpc_gpu_decode_batch_static and plans created
 * afterwards; ldpc_gpu_plan_shape's output can be set here.  Results do not
 * depend on the shape.  LDPC_GPU_EINVAL for an instance or build that does
 * not exist. */
int ldpc_gpu_set_launch_shape(ldpc_gpu_ctx* ctx, int di, int short_frames, int occ, float scale);

/* Timing of the last ldpc_gpu_decode_batch's decode kernel on the stream it
 * ran on (HIP events around the launch), in milliseconds; < 0 if none. */
float ldpc_gpu_last_kernel_ms(ldpc_gpu_ctx* ctx);

/* Per-launch detail of the last ldpc_gpu_decode_batch: frames are decoded
 * by one persistent kernel launch per kernel instance (the layer degree
 * bound of ldpc_gpu_code_kernel_degree; for frames in groups, the one-frame
 * kernel's classes 12 and 28) and frame size.  Fills up to max entries
 * with that bound, the frames of the launch and its time in ms; returns
 * the count. */
int ldpc_gpu_last_launches(ldpc_gpu_ctx* ctx, int* di, int* nframes, float* ms, int max);
/* The same launches, in the same order: the workgroups per CU each was
 * built for (ldpc_gpu_set_launch_shape) and the workgroups it launched. */
int ldpc_gpu_last_launch_shapes(ldpc_gpu_ctx* ctx, int* occ, int* grid, int max);

#ifdef __cplusplus
}
#endif

#endif
