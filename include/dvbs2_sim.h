/*
 * Frame source for tests, benchmarks and decoder training: DVB-S2/S2X/T2
 * BCH + LDPC encoder, Gray/APSK mapping with the bit interleaver, complex
 * AWGN and an int8 max-log demapper with libdvbrx's LLR scale.
 *
 * The LLR contract is the reference demapper's (libdvbrx/src/demap.c:32-34,
 * :225-268, :491-500): llr = sat127(rint(2/n0 * (d1 - d0))), d0/d1 the
 * squared distances to the nearest point with the bit at 0/1, positive for
 * a 0 bit, codeword order.  The "unit" channel (pls < 0) is the one of
 * tests/test_ldpc.c:64-73: +-1 per bit, noise sigma^2 = 1/EsN0,
 * llr = sat127(rint(4 y / sigma^2)).
 */
#ifndef DVBS2_SIM_H
#define DVBS2_SIM_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int short_frame;   /* ignored when pls >= 0 */
    char rate[24];     /* drx_fec_lookup naming; ignored when pls >= 0 */
    int pls;           /* MODCOD PLS code (pilot bit ignored), or -1: unit channel */
    double esn0_db;    /* Es/N0 in dB; >= 300 means noiseless */
    uint64_t llr_off;  /* byte offset of the frame's nldpc LLRs */
    uint64_t ref_off;  /* byte offset of the frame's kldpc/8 reference bytes */
    uint64_t stream;   /* RNG stream (frame identity) */
} sim_frame_spec;

/* Generate frames: random kbch info bits -> BCH -> LDPC -> channel ->
 * int8 LLRs at llr + llr_off; the BCH codeword (kldpc bits, MSB first) at
 * ref + ref_off when ref is not NULL.  Frames are independent given
 * (seed, stream), so the result does not depend on nthreads.  Returns 0,
 * or -1 for an unknown rate / MODCOD. */
int sim_generate(const sim_frame_spec* spec, int nframes, uint64_t seed, int8_t* llr,
                 uint8_t* ref, int nthreads);

/* FEC parameters of a rate (drx_fec_lookup, libdvbrx/src/bch.c:25-41). */
int sim_fec_lookup(int short_frame, const char* rate, int* nldpc, int* kldpc, int* kbch, int* t);

/* MODCOD of a PLS code (drx_pls_lookup, libdvbrx/src/demap.c:61-76): rate
 * name, frame size and bits per symbol.  Returns 0 if known. */
int sim_pls_lookup(int pls, char* rate, int rate_len, int* short_frame, int* bits,
                   char* constellation, int const_len);

/* Quasi-error-free Es/N0 of a MODCOD (dB): EN 302 307-1 Table 13 for
 * DVB-S2 PLS codes, Shannon bound + 1 dB for DVB-S2X (libdvbrx/src/pl.c:68-87). */
double sim_qef_db(int pls);

/* Encoders on unpacked bits (one byte per bit, 0/1). */
int sim_bch_encode(int short_frame, const char* rate, const uint8_t* info, uint8_t* cw);
int sim_ldpc_encode(int short_frame, const char* rate, const uint8_t* info, uint8_t* cw);

#ifdef __cplusplus
}
#endif

#endif
