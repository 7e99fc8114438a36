/*
 * The drop-in, driven from C exactly as libdvbrx's callers drive
 * drx_ldpc_decode (libdvbrx/tests/test_ldpc.c:64-100, rx.c:132): batches of
 * drx_ldpc_lanes() frames, each call returning the batch's iteration count
 * or -1.  Links libldpcgpu.so (include/dvbrx/ldpc.h) in place of ldpc.c.
 *
 *   dropin_ldpc <short 0/1> <rate> <pls or -1> <EsN0 dB> <frames> <max_iter> <out file> [<bch file>]
 *
 * Frames come from the frame source (libdvbs2sim.so): encoder, channel,
 * int8 LLRs.  Prints the per-batch results and the frame error count; the
 * decoded bytes of all frames go to <out file>.  With <bch file>: the
 * parameters come from drx_fec_lookup, and each decoded frame, with f mod
 * (t + 3) bits flipped (bit 131 k + 7 f for the k-th), goes through
 * drx_bch_decode as rx.c does after the LDPC decoder; the results are
 * printed and the codewords written to <bch file>.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "dvbrx/ldpc.h"
#include "dvbs2_sim.h"

int main(int argc, char** argv)
{
    if (argc != 8 && argc != 9) {
        fprintf(stderr, "usage: %s short rate pls esn0_db frames max_iter out [bch_out]\n", argv[0]);
        return 2;
    }
    drx_fec_params p;
    memset(&p, 0, sizeof(p));
    p.short_frame = atoi(argv[1]);
    const char* rate = argv[2];
    const int pls = (int)strtol(argv[3], NULL, 0), nf = atoi(argv[5]), max_iter = atoi(argv[6]);
    if (sim_fec_lookup(p.short_frame, rate, &p.nldpc, &p.kldpc, &p.kbch, &p.t))
        return 2;
    if (argc == 9) {
        drx_fec_params q;
        if (drx_fec_lookup(p.short_frame, rate, &q) || memcmp(&p, &q, sizeof(p))) {
            fprintf(stderr, "drx_fec_lookup differs from the frame source's parameters\n");
            return 4;
        }
    }
    const int kb = p.kldpc / 8;
    sim_frame_spec* spec = calloc((size_t)nf, sizeof(*spec));
    int8_t* llr = malloc((size_t)nf * p.nldpc);
    uint8_t* ref = malloc((size_t)nf * kb);
    uint8_t* out = calloc((size_t)nf, (size_t)kb);
    if (!spec || !llr || !ref || !out)
        return 1;
    for (int f = 0; f < nf; f++) {
        spec[f].short_frame = p.short_frame;
        snprintf(spec[f].rate, sizeof(spec[f].rate), "%s", rate);
        spec[f].pls = pls;
        spec[f].esn0_db = atof(argv[4]);
        spec[f].llr_off = (uint64_t)f * p.nldpc;
        spec[f].ref_off = (uint64_t)f * kb;
        spec[f].stream = (uint64_t)f;
    }
    if (sim_generate(spec, nf, 2024, llr, ref, 0))
        return 1;

    drx_ldpc* d = drx_ldpc_new(&p, rate);
    if (!d) {
        fprintf(stderr, "drx_ldpc_new failed (no GPU?)\n");
        return 3;
    }
    const int lanes = drx_ldpc_lanes();
    printf("lanes %d\nbatches", lanes);
    for (int f0 = 0; f0 < nf; f0 += lanes) {
        const int n = nf - f0 < lanes ? nf - f0 : lanes;
        const int8_t* in[64];
        uint8_t* o[64];
        for (int k = 0; k < n; k++) {
            in[k] = llr + (size_t)(f0 + k) * p.nldpc;
            o[k] = out + (size_t)(f0 + k) * kb;
        }
        printf(" %d", drx_ldpc_decode(d, in, o, n, max_iter));
    }
    int errors = 0;
    for (int f = 0; f < nf; f++)
        errors += memcmp(out + (size_t)f * kb, ref + (size_t)f * kb, (size_t)kb) != 0;
    printf("\nframes %d in_error %d\n", nf, errors);
    drx_ldpc_free(d);
    if (argc == 9) {
        drx_bch* b = drx_bch_new(&p);
        uint8_t* cw = malloc((size_t)nf * kb);
        if (!b || !cw)
            return 3;
        memcpy(cw, out, (size_t)nf * kb);
        printf("bch");
        for (int f = 0; f < nf; f++) {
            uint8_t* c = cw + (size_t)f * kb;
            for (int k = 0; k < f % (p.t + 3); k++) {
                const int bit = (131 * k + 7 * f) % p.kldpc;
                c[bit / 8] ^= (uint8_t)(0x80 >> (bit % 8));
            }
            printf(" %d", drx_bch_decode(b, c));
        }
        printf("\n");
        drx_bch_free(b);
        FILE* fb = fopen(argv[8], "wb");
        if (!fb || fwrite(cw, 1, (size_t)nf * kb, fb) != (size_t)nf * kb)
            return 1;
        fclose(fb);
        free(cw);
    }
    FILE* fo = fopen(argv[7], "wb");
    if (!fo || fwrite(out, 1, (size_t)nf * kb, fo) != (size_t)nf * kb)
        return 1;
    fclose(fo);
    return 0;
}
