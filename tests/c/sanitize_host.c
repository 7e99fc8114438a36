/*
 * The host C code under AddressSanitizer and UndefinedBehaviorSanitizer
 * (built by tests/test_sanitize.py with gcc -fsanitize=address,undefined,
 * errors fatal): the code registry and its split-kernel tables for every
 * code, the frame source (encoders, channel, demapper), the oracle (LDPC
 * group and per-frame decoding, the learned decoder and its training
 * step, BCH, the max-log demapper, the DVB-T2 front end) and the drop-in
 * shim's host paths (FEC lookup; on a host without a GPU, the failing
 * creations).  Exits 0 when every check holds.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "dvbrx/ldpc.h"
#include "dvbs2_sim.h"
#include "ldpc_codes.h"
#include "ldpc_oracle.h"

#define CHECK(c)                                                          \
    do {                                                                  \
        if (!(c)) {                                                       \
            fprintf(stderr, "%s:%d: check failed: %s\n", __FILE__, __LINE__, #c); \
            return 1;                                                     \
        }                                                                 \
    } while (0)

static int decode_case(int short_frame, const char* rate, int pls, double esn0, int nf)
{
    int nldpc, kldpc, kbch, t;
    CHECK(sim_fec_lookup(short_frame, rate, &nldpc, &kldpc, &kbch, &t) == 0);
    const int kb = (kldpc + 7) / 8;
    sim_frame_spec* spec = calloc((size_t)nf, sizeof(*spec));
    int8_t* llr = malloc((size_t)nf * nldpc);
    uint8_t* ref = malloc((size_t)nf * kb);
    uint8_t* out = malloc((size_t)nf * kb);
    int32_t* it = malloc(sizeof(int32_t) * nf);
    uint8_t* ok = malloc((size_t)nf);
    CHECK(spec && llr && ref && out && it && ok);
    for (int f = 0; f < nf; f++) {
        spec[f].short_frame = short_frame;
        snprintf(spec[f].rate, sizeof(spec[f].rate), "%s", rate);
        spec[f].pls = pls;
        spec[f].esn0_db = esn0;
        spec[f].llr_off = (uint64_t)f * nldpc;
        spec[f].ref_off = (uint64_t)f * kb;
        spec[f].stream = (uint64_t)f;
    }
    CHECK(sim_generate(spec, nf, 5, llr, ref, 2) == 0);
    oracle_ldpc* d = oracle_ldpc_new(short_frame, rate, kldpc);
    CHECK(d);
    oracle_ldpc_decode_frames(d, llr, out, nf, 20, it, ok);
    const int r = oracle_ldpc_decode(d, llr, out, nf, 20);
    CHECK(r >= -1 && r <= 20);
    /* learned decoder with the plain table (S = B = 16) and its training
     * step on one frame */
    uint8_t S[2 * ORACLE_NMS_DEGS];
    int16_t B[2 * ORACLE_NMS_DEGS];
    float al[2 * ORACLE_NMS_DEGS], be[2 * ORACLE_NMS_DEGS], ga[2 * ORACLE_NMS_DEGS], gb[2 * ORACLE_NMS_DEGS];
    for (int k = 0; k < 2 * ORACLE_NMS_DEGS; k++) {
        S[k] = 16;
        B[k] = 16;
        al[k] = 1.0f;
        be[k] = 1.0f;
    }
    oracle_nms_decode_frames(d, S, B, 2, llr, out, 1, 5, 5, it, ok, NULL);
    float* fl = malloc(sizeof(float) * nldpc);
    uint8_t* bits = calloc((size_t)nldpc, 1);
    float w[2] = { 1.0f, 1.0f };
    CHECK(fl && bits);
    for (int v = 0; v < nldpc; v++)
        fl[v] = llr[v];
    const double loss = oracle_nms_train(d, 1, fl, bits, 1, 2, al, be, w, 0.5f, ga, gb, NULL);
    CHECK(isfinite(loss));
    /* BCH on the first frame's decoded bits, with errors */
    out[0] ^= 0x81;
    CHECK(oracle_bch_decode(short_frame, rate, out) >= -1);
    oracle_ldpc_free(d);
    free(fl);
    free(bits);
    free(spec);
    free(llr);
    free(ref);
    free(out);
    free(it);
    free(ok);
    return 0;
}

int main(void)
{
    /* every code of the registry, its expansion and split tables */
    const int nc = lg_code_count();
    CHECK(nc > 50);
    for (int id = 0; id < nc; id++) {
        const lg_code* c = lg_code_get(id);
        CHECK(c && c->sk_tab && c->sk_words > 0);
    }
    CHECK(lg_code_get(nc) == NULL && lg_code_lookup(0, "7_7") < 0);
    int kb, nb, t;
    CHECK(lg_fec_lookup(1, "8_9", &kb, &nb, &t) == 0 && lg_fec_rate(1, kb, nb, t) != NULL);
    /* the drop-in's host paths */
    drx_fec_params p;
    CHECK(drx_fec_lookup(0, "1_2", &p) == 0 && p.kldpc == 32400 && drx_fec_lookup(0, "7_7", &p) == -1);
    drx_ldpc* dl = drx_ldpc_new(&p, "1_2");
    drx_ldpc_free(dl);
    drx_bch* db = drx_bch_new(&p);
    drx_bch_free(db);
    CHECK(drx_ldpc_new(&p, "no_rate") == NULL && drx_ldpc_new(NULL, "1_2") == NULL);
    /* frame source and oracle: unit channel and max-log demapped MODCODs */
    if (decode_case(1, "1_2", -1, 1.3, 3) || decode_case(1, "8_9", 0x42, 7.0, 2) || decode_case(1, "1_4", 0x06, -1.5, 2) ||
        decode_case(0, "2_3", 0x48, 9.5, 1) || decode_case(1, "3_5_t2", -1, 2.5, 2))
        return 1;
    /* max-log demapper of a 32APSK MODCOD */
    const int ns = oracle_demap_symbols(0x62);
    CHECK(ns > 0);
    float* sym = malloc(sizeof(float) * 2 * ns);
    int8_t* dl8 = malloc((size_t)ns * 5);
    CHECK(sym && dl8);
    for (int k = 0; k < 2 * ns; k++)
        sym[k] = (float)((k * 37 % 101) - 50) / 60.0f;
    CHECK(oracle_demap_frame(0x62, sym, 0.2f, dl8) == 0);
    free(sym);
    free(dl8);
    /* DVB-T2 front end: 64-QAM 2/3 short, rotated, 3 blocks of 2 TI-blocks */
    int nldpc, ncells;
    char code[32];
    CHECK(oracle_t2_geometry(2, 2, 0, &nldpc, &ncells, code, sizeof(code)) == 0);
    const int nbl = 3;
    float* cells = malloc(sizeof(float) * 2 * ncells * nbl);
    float* csi = malloc(sizeof(float) * ncells * nbl);
    int8_t* tl = malloc((size_t)nldpc * nbl);
    float err[3], sig[3];
    CHECK(cells && csi && tl);
    for (int k = 0; k < 2 * ncells * nbl; k++)
        cells[k] = (float)((k * 13 % 29) - 14) / 10.0f;
    for (int k = 0; k < ncells * nbl; k++)
        csi[k] = 1.0f;
    CHECK(oracle_t2_demap(2, 2, 0, 1, 2, 0, nbl, cells, csi, tl, err, sig) == 0);
    free(cells);
    free(csi);
    free(tl);
    printf("sanitized host run ok: %d codes\n", nc);
    return 0;
}
