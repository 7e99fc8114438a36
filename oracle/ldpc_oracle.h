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

#endif
