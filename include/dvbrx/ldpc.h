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

/* libdvbrx's ldpc.h includes its fec.h (libdvbrx/include/dvbrx/ldpc.h:9),
 * and its callers reach drx_fec_lookup and drx_bch_* through it.  Where
 * that header is on the include path it is used as it is; elsewhere the
 * same type and functions are declared here, and libldpcgpu.so defines
 * them (weak symbols: linked beside libdvbrx's bch.c, the reference's
 * win). */
#if defined(__has_include)
#if __has_include("dvbrx/fec.h")
#include "dvbrx/fec.h"
#endif
#endif

#ifndef DVBRX_FEC_H
/* Same layout as libdvbrx/include/dvbrx/fec.h:13-19. */
typedef struct {
    int short_frame; /* 0: normal (64800 bits), 1: short (16200 bits) */
    int nldpc;       /* LDPC codeword length */
    int kldpc;       /* = nbch, decoded bits written per frame */
    int kbch;        /* BBFRAME length */
    int t;           /* BCH error-correcting capability */
} drx_fec_params;

/* libdvbrx/include/dvbrx/fec.h:25: parameters of a rate ("3_4", "9_10",
 * "2_9_VLSNR", ...); 0 on success. */
int drx_fec_lookup(int short_frame, const char* rate, drx_fec_params* p);

/* libdvbrx/include/dvbrx/fec.h:27-37: BCH decoding of one codeword in
 * place (kldpc bits, MSB first); the corrected bit count, or -1 if
 * uncorrectable (the codeword is left untouched).  Decoded on the GPU
 * (ldpc_gpu_bch_decode, one codeword a call). */
typedef struct drx_bch drx_bch;
drx_bch* drx_bch_new(const drx_fec_params* p);
void drx_bch_free(drx_bch* b);
int drx_bch_decode(drx_bch* b, uint8_t* cw);
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
