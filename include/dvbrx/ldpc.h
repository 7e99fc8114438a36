/*
 * Drop-in for libdvbrx's LDPC decoder API (libdvbrx/include/dvbrx/ldpc.h:14-36),
 * decoding on an MI355X through include/ldpc_gpu.h.  Same four entry
 * points, same types, same batch semantics: the frames of one call iterate
 * together until all satisfy their checks (libdvbrx/src/ldpc.c:614-638).
 * Callers that relink against libldpcgpu.so instead of libdvbrx's ldpc.c
 * need no source change.
 */
#ifndef DVBRX_LDPC_GPU_DROPIN_H
#define DVBRX_LDPC_GPU_DROPIN_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#ifndef DVBRX_FEC_H
/* Same layout as libdvbrx/include/dvbrx/fec.h:13-19 (filled by the caller's
 * drx_fec_lookup, or by ldpc_gpu_fec_lookup). */
typedef struct {
    int short_frame; /* 0: normal (64800 bits), 1: short (16200 bits) */
    int nldpc;       /* LDPC codeword length */
    int kldpc;       /* = nbch, decoded bits written per frame */
    int kbch;        /* BBFRAME length */
    int t;           /* BCH error-correcting capability */
} drx_fec_params;
#endif

typedef struct drx_ldpc drx_ldpc;

/* Frames per call: 64, the AVX-512 build's lane count.  Callers size fixed
 * arrays by it (libdvbrx/src/rx.c:14, rxmt.c:23, dvbt2.c:604). */
int drx_ldpc_lanes(void);

/* rate as drx_fec_lookup ("1_2", "2_3_t2", ...).  NULL on an unknown rate
 * or when no GPU is usable. */
drx_ldpc* drx_ldpc_new(const drx_fec_params* p, const char* rate);
void drx_ldpc_free(drx_ldpc* d);

/* Decode nframes (<= drx_ldpc_lanes()) codewords: llr[f] holds the nldpc
 * LLRs of frame f (positive for a 0 bit), out[f] receives the kldpc
 * decoded bits packed MSB first.  Returns the iterations run, or -1 if some
 * frame still violates a check after max_iter iterations (or nframes is
 * above drx_ldpc_lanes(), or the device failed). */
int drx_ldpc_decode(drx_ldpc* d, const int8_t* const* llr, uint8_t* const* out, int nframes, int max_iter);

#ifdef __cplusplus
}
#endif

#endif
