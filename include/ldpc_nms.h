/*
 * Learned (neural) offset min-sum for the MI355X LDPC decoder: training on
 * the GPU, and the GPU channel that feeds it.  C-ABI.
 *
 * The reference decoder has no learned form: its check node uses a fixed
 * offset of 1 (libdvbrx/src/ldpc.c:353).  The learned form replaces that
 * magnitude, per iteration t and check degree dc, by
 *     float (training):  clamp(alpha[t][dc] * min(m, 127) - beta[t][dc], 0, 127)
 *     integer (decoding): clamp((min(m, 127) * S[t][dc] - B[t][dc]) >> 4, 0, 127)
 * (arithmetic shift) with S = clamp(rint(16 alpha), 0, 63) and
 * B = clamp(rint(16 beta), -1024, 1023): scale and offset in 1/16 units;
 * S = 16, B = 16 is the reference.  Everything else (int8 saturation,
 * messages clamped to [-32, 31], layer order) is the reference's.
 *
 * Training uses the integer forward with the quantised parameters and the
 * float model's gradient (straight-through), so what is trained is what
 * ldpc_gpu_decode_batch_nms (include/ldpc_gpu.h) decodes with.
 *
 * Parameter arrays are T rows of LDPC_NMS_DEGS entries, indexed by check
 * degree (edges of the node).  d_* pointers are device memory.
 */
#ifndef LDPC_NMS_H
#define LDPC_NMS_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LDPC_NMS_DEGS 33
#define LDPC_NMS_MAX_ITERS 32

typedef struct ldpc_nms_trainer ldpc_nms_trainer;

/* Trainer for one code (ldpc_gpu_code_lookup id) on one device: T
 * unrolled iterations (1..LDPC_NMS_MAX_ITERS), scratch for steps of up
 * to max_frames frames.  Returns 0 or an LDPC_GPU_* error code. */
int ldpc_nms_trainer_create(int device, int code, int T, int max_frames, ldpc_nms_trainer** out);
void ldpc_nms_trainer_destroy(ldpc_nms_trainer* t);
void* ldpc_nms_trainer_stream(ldpc_nms_trainer* t); /* hipStream_t */
const char* ldpc_nms_trainer_error(ldpc_nms_trainer* t);
int ldpc_nms_trainer_synchronize(ldpc_nms_trainer* t);

/*
 * One forward + backward pass (enqueued on the trainer's stream):
 *   d_llr     nframes x N int8 channel LLRs, codeword order
 *   d_cw      codeword bits, one byte each, rows of N; frame f uses row d_row[f]
 *   d_params  2 x T x LDPC_NMS_DEGS floats: alpha rows, then beta rows
 *   d_w       T loss weights (iteration t's posteriors weigh d_w[t])
 *   temp      logit scale: bit 1 has probability sigmoid(-temp * L)
 *   d_loss    1 float: sum over t of d_w[t] x mean binary cross-entropy
 *   d_grad    2 x T x LDPC_NMS_DEGS floats: gradient of d_loss
 *   d_post    NULL, or nframes x N int8: posteriors after T iterations
 */
int ldpc_nms_train_step(ldpc_nms_trainer* t, const int8_t* d_llr, const uint8_t* d_cw, const int32_t* d_row,
                        int nframes, const float* d_params, const float* d_w, float temp, float* d_loss,
                        float* d_grad, int8_t* d_post);

/* Adam (Kingma & Ba) on n parameters: m, v moments, step >= 1.  The
 * first n/2 parameters (alpha) are then kept in [0.25, 1.5], the others
 * (beta) in [-8, 16]: min-sum overestimates the magnitudes, a learned
 * correction scales them down or offsets them. */
int ldpc_nms_adam(ldpc_nms_trainer* t, float* d_param, const float* d_grad, float* d_m, float* d_v, int n,
                  float lr, float beta1, float beta2, float eps, int step);

/* Quantised integer table of float parameters (host arrays, T rows). */
void ldpc_nms_quantize(const float* alpha, const float* beta, int T, uint8_t* S, int16_t* B);

/*
 * GPU channel for training and evaluation data: codeword rows (one byte
 * per bit) -> the MODCOD's constellation (PLS code, as drx_pls_lookup)
 * with its bit interleaver -> complex AWGN at d_esn0_db[f] (dB, per
 * frame) -> the library's max-log demapper (ldpc_gpu_demap's kernel: the
 * reference's int8 LLRs for those symbols), positive for bit 0.  pls < 0:
 * +-1 per bit, llr = sat127(rint(4 y / sigma^2)).  Noise from a
 * counter-based generator keyed by (seed, frame, symbol).  n: codeword
 * length of the rows.  Runs on stream (a hipStream_t, or NULL for the
 * default stream); MODCOD frames go through a symbol buffer of the call
 * and the call returns when the LLRs are written.
 */
int ldpc_gpu_channel(int pls, int n, const uint8_t* d_cw, const int32_t* d_row, int nframes,
                     const float* d_esn0_db, uint64_t seed, int8_t* d_llr, void* stream);

/* The received symbols of ldpc_gpu_channel's MODCOD frames (pls >= 0), the
 * input of its demapping: d_sym nframes x ldpc_gpu_demap_symbols(pls)
 * (re, im) float pairs at unit average energy, d_n0 the frames' noise
 * variances, as ldpc_gpu_demap takes them.  Enqueued on stream. */
int ldpc_gpu_channel_symbols(int pls, int n, const uint8_t* d_cw, const int32_t* d_row, int nframes,
                             const float* d_esn0_db, uint64_t seed, float* d_sym, float* d_n0, void* stream);

#ifdef __cplusplus
}
#endif

#endif
