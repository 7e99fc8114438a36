/*
 * CPU oracle for the LDPC decode path.  TEST INFRASTRUCTURE ONLY: it is the
 * checker that the HIP decoder is compared against, never part of the
 * product.  Only tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline leg may load it.
 *
 * A scalar restatement of libdvbrx/src/ldpc.c (layered offset min-sum on
 * int8 LLRs).  Several frames decoded together share one early stop, as in
 * the reference's drx_ldpc_decode (ldpc.c:614-638), but without its limit
 * of 64 frames per call.
 */
#ifndef LDPC_ORACLE_H
#define LDPC_ORACLE_H

#include <stdint.h>

typedef struct oracle_ldpc oracle_ldpc;

/* Code of (frame size, rate) as drx_ldpc_new (ldpc.c:199-260); kldpc is the
 * number of decoded bits written per frame.  NULL on unknown rate. */
oracle_ldpc* oracle_ldpc_new(int short_frame, const char* rate, int kldpc);
void oracle_ldpc_free(oracle_ldpc* d);
int oracle_ldpc_n(const oracle_ldpc* d);

/* Group decode: nframes frames of N int8 LLRs (contiguous), decoded bits
 * packed MSB first, ceil(kldpc/8) bytes per frame.  Returns the shared
 * iteration count, or -1 if some frame still fails a check after max_iter
 * (ldpc.c:627-637). */
int oracle_ldpc_decode(oracle_ldpc* d, const int8_t* llr, uint8_t* out, int nframes, int max_iter);

/* Per-frame decode: each frame alone (the reference called with
 * nframes = 1).  iters[f] = updates run, ok[f] = syndrome satisfied. */
void oracle_ldpc_decode_frames(oracle_ldpc* d, const int8_t* llr, uint8_t* out, int nframes,
                               int max_iter, int32_t* iters, uint8_t* ok);

/* Syndrome of hard decisions: 1 if every parity check holds for the LLR
 * signs of one frame (same test as bad(), ldpc.c:286-305). */
int oracle_ldpc_check(oracle_ldpc* d, const int8_t* llr);

/* One check node of the reference (check_node, ldpc.c:321-362), exposed for
 * brute-force tests: deg posteriors L[] and old messages m[] in, new
 * posteriors and messages out.  first: messages are all zero. */
void oracle_check_node(int deg, const int8_t* L, const int8_t* m, int first, int8_t* L_out,
                       int8_t* m_out);

/* nframes x T plain updates of the reference, no early stop; posteriors in
 * codeword order (N per frame). */
void oracle_ldpc_run(oracle_ldpc* d, const int8_t* llr, int nframes, int T, int8_t* post);

/* ---- learned (neural) offset min-sum ------------------------------------
 *
 * Parameters are per iteration t and per check degree dc (the node's edge
 * count, row index 0..ORACLE_NMS_DEGS-1); iterations past the table's T
 * rows use row T-1.  Integer form: the check node's magnitudes become
 *     mn = clamp((min(m, 127) * S[t][dc] - B[t][dc]) >> 4, 0, 127)
 * (scale and offset in 1/16 units, arithmetic shift) in place of
 * max(min(m, 127) - 1, 0); S = 16, B = 16 is the reference.
 * Float form (training): mag = clamp(alpha * min(mu, 127) - beta, 0, 127),
 * everything else as the integer decoder (saturation to [-128, 127],
 * stored messages clamped to [-32, 31]).
 */
#define ORACLE_NMS_DEGS 33

/* Quantisation of a float (alpha, beta) pair: S = clamp(rint(16 alpha), 0,
 * 63), B = clamp(rint(16 beta), -1024, 1023). */
int oracle_nms_scale_q(float alpha);
int oracle_nms_offset_q(float beta);

/* One learned check node (integer form), as oracle_check_node. */
void oracle_nms_check_node(int deg, const int8_t* L, const int8_t* m, int first, int S, int B,
                           int8_t* L_out, int8_t* m_out);

/* Per-frame learned decode: up to learned_max learned iterations with the
 * syndrome early stop; a frame still failing is decoded again from its
 * channel LLRs by the plain decoder for up to fallback_max iterations
 * (0: no fallback).  iters[f] = learned + plain iterations, ok[f] the
 * final syndrome, fell_back[f] (may be NULL) 1 if the fallback ran. */
void oracle_nms_decode_frames(oracle_ldpc* d, const uint8_t* S, const int16_t* B, int T, const int8_t* llr,
                              uint8_t* out, int nframes, int learned_max, int fallback_max, int32_t* iters,
                              uint8_t* ok, uint8_t* fell_back);

/* Training step of the float model over T unrolled iterations (no early
 * stop) on nframes frames: llr (nframes x N, float), bits (nframes x N
 * codeword bits 0/1), alpha/beta (T x ORACLE_NMS_DEGS).  Loss: sum over t
 * of w[t] times the mean over frames and bits of the binary cross-entropy
 * of the posteriors after iteration t, read as logits temp * L (L > 0:
 * bit 0).  mode 0: float magnitudes; mode 1: the integer forward with
 * S, B quantised from alpha, beta (posteriors equal oracle_nms_decode's)
 * and the float model's gradient (straight-through).  Returns the loss;
 * galpha/gbeta (T x ORACLE_NMS_DEGS, may be NULL) receive its gradient,
 * post (nframes x N, may be NULL) the final posteriors in codeword order. */
double oracle_nms_train(oracle_ldpc* d, int mode, const float* llr, const uint8_t* bits, int nframes, int T,
                        const float* alpha, const float* beta, const float* w, float temp, float* galpha,
                        float* gbeta, float* post);

/* ---- max-log demapper (libdvbrx/src/demap.c:164-203, :491-532) --------
 *
 * One frame of MODCOD pls (PLS code, pilot bit ignored): sym holds the
 * frame's symbols as (re, im) float pairs at unit average energy,
 * oracle_demap_symbols(pls) of them; n0 the noise variance; llr receives
 * nldpc int8 LLRs in codeword order (positive for a 0 bit), as the
 * reference's AVX2/FMA build computes them.  Returns 0, or -1 for an
 * unknown pls.  oracle_demap_ring: 1 if the constellation takes the
 * ring metric (8 rings of 32 phases), 0 if not, -1 unknown. */
int oracle_demap_symbols(int pls);
int oracle_demap_ring(int pls);
/* the constellation at unit average energy, by label: returns the point
 * count (re, im hold up to 256), -1 unknown */
int oracle_demap_points(int pls, float* re, float* im);
int oracle_demap_frame(int pls, const float* sym, float n0, int8_t* llr);

/* ---- BCH decoder (libdvbrx/src/bch.c:106-172, :189-290) ---------------
 *
 * One codeword of the FEC (short_frame, rate as drx_fec_lookup): nbch bits
 * packed MSB first (the LDPC decoder's output), corrected in place.
 * Returns the number of corrected bits, 0 for a codeword, -1 if more
 * errors than the code corrects were found (cw unchanged), -2 for an
 * unknown rate.  oracle_bch_nbch: nbch, -1 unknown. */
int oracle_bch_nbch(int short_frame, const char* rate);
int oracle_bch_decode(int short_frame, const char* rate, uint8_t* cw);

/* ---- DVB-T2 FEC block front end (libdvbrx/src/dvbt2.c:1953-2082,
 * :2188-2390) ---------------------------------------------------------
 *
 * A PLP's FEC blocks as the reference's FEC thread takes them: cod 0..5
 * (1/2 3/5 2/3 3/4 4/5 5/6), mod 0..3 (QPSK, 16-, 64-, 256-QAM), fec_type
 * 1 normal / 0 short, rotation 0/1, and the time interleaver's length and
 * type (the cell deinterleaver's shift restarts with each TI-block when
 * the type is 0).
 *
 * oracle_t2_geometry: nldpc, ncells (= nldpc / bits per cell) and the LDPC
 * code name ("2_3_t2" and "3_5_t2" for the T2-only codes); -1 unsupported.
 * oracle_t2_cellperm: the cell interleaver's basic permutation L0
 * (ncells entries).  oracle_t2_shift: each block's cyclic shift P.
 * oracle_t2_bitsrc: for each codeword bit n, the soft value it takes:
 * cell * m + plane (plane 2j: bit j of I, 2j + 1: bit j of Q) -- the
 * demultiplexer, column twist and parity interleaving undone.
 * oracle_t2_demap: nblocks blocks of ncells complex cells (re, im) and
 * their channel state -> nldpc int8 soft values per block in codeword
 * order (positive for a 0 bit) and the block's MER error and signal sums,
 * as the reference's x86-64 build with FMA contraction computes them.
 * Returns 0, -1 unsupported. */
int oracle_t2_geometry(int cod, int mod, int fec_type, int* nldpc, int* ncells, char* code, int code_len);
int oracle_t2_cellperm(int ncells, int32_t* perm);
int oracle_t2_shift(int nblocks, int til_length, int til_type, int ncells, int32_t* P);
int oracle_t2_bitsrc(int cod, int mod, int fec_type, uint16_t* src);
int oracle_t2_demap(int cod, int mod, int fec_type, int rotation, int til_length, int til_type, int nblocks,
                    const float* cells, const float* csi, int8_t* llr, float* err, float* sig);

#endif
