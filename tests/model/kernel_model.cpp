/*
 * CPU model of the HIP decoder's schedule (test infrastructure, never part
 * of the product): the same code registry (circulant layers, dependency
 * levels and masks), the same per-node arithmetic (dvb-rx_amd/csrc/ldpc_node.h)
 * and the same posterior / message layout as the decode kernel, run on one
 * core.  Within each phase and level the nodes are processed in decreasing
 * j: the kernel runs them concurrently, so the result must not depend on
 * their order; a wrong level or mask would change the result against the
 * oracle, which processes them in increasing j.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "ldpc_codes.h"
#include "ldpc_node.h"
#include "ldpc_pk.h"

extern "C" {
int kmodel_decode(int code, const int8_t* llr, int max_iter, int nbits, uint8_t* out, int* ok);
int kmodel_check_levels(int code);
}

template <int DEG>
static void model_layer(const lg_code* c, int i, int8_t* post, uint32_t* msg, int first)
{
    const int K = c->k, q = c->q;
    const uint32_t* ent = c->ent + c->layer_start[i];
    const uint64_t* meta = c->meta + (size_t)i * 360;
    static LgNode<DEG> nodes[360];
    uint32_t mw[360][LG_MSG_WORDS(DEG)];
    const int W = LG_MSG_WORDS(DEG);
    uint32_t* m = msg + (size_t)c->msg_off[i] * 360;
    for (int j = 0; j < 360; j++)
        for (int w = 0; w < W; w++)
            mw[j][w] = first ? 0 : m[w * 360 + j];
    for (int j = 359; j >= 0; j--) {
        lg_phase_a<DEG>(nodes[j], post, ent, i, j, K, q, LG_META_DEP(meta[j]), mw[j], first);
        if (LG_META_LEVEL(meta[j]) == 0)
            lg_write_succ<DEG>(nodes[j], post, LG_META_SUCC(meta[j]));
    }
    for (int l = 1; l < c->nlev[i]; l++)
        for (int j = 359; j >= 0; j--)
            if (LG_META_LEVEL(meta[j]) == l) {
                lg_step<DEG>(nodes[j], post, LG_META_DEP(meta[j]), mw[j], first);
                lg_write_succ<DEG>(nodes[j], post, LG_META_SUCC(meta[j]));
            }
    for (int j = 359; j >= 0; j--) {
        uint32_t o[LG_MSG_WORDS(DEG)];
        lg_phase_c<DEG>(nodes[j], post, LG_META_SUCC(meta[j]), o);
        for (int w = 0; w < W; w++)
            m[w * 360 + j] = o[w];
    }
}

template <int DEG>
static int model_check(const lg_code* c, int i, const int8_t* post)
{
    int good = 1;
    for (int j = 0; j < 360; j++)
        good &= lg_node_check_t<DEG>(post, c->ent + c->layer_start[i], DEG, i, j, c->k, c->q);
    return good;
}

#define DEG_CASES(X)                                                                                      \
    X(1) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12) X(13) X(14) X(15) X(16) X(17) X(18) X(19) \
        X(20) X(21) X(22) X(23) X(24) X(25) X(26) X(27) X(28)

static int layer_any(const lg_code* c, int i, int8_t* post, uint32_t* msg, int first)
{
    switch (c->layer_start[i + 1] - c->layer_start[i]) {
#define X(D)                                 \
    case D:                                  \
        model_layer<D>(c, i, post, msg, first); \
        return 0;
        DEG_CASES(X)
#undef X
    }
    return -1;
}

static int check_any(const lg_code* c, int i, const int8_t* post)
{
    switch (c->layer_start[i + 1] - c->layer_start[i]) {
#define X(D) \
    case D:  \
        return model_check<D>(c, i, post);
        DEG_CASES(X)
#undef X
    }
    return 0;
}

/* Decode one frame with per-frame early stop.  Returns iterations run;
 * *ok = syndrome holds.  out: ceil(nbits/8) bytes, MSB first. */
int kmodel_decode(int code, const int8_t* llr, int max_iter, int nbits, uint8_t* out, int* ok)
{
    const lg_code* c = lg_code_get(code);
    if (!c)
        return -1;
    const int K = c->k, R = c->r, q = c->q;
    int8_t* post = (int8_t*)malloc((size_t)c->n);
    uint32_t* msg = (uint32_t*)calloc((size_t)c->msg_words, sizeof(uint32_t));
    if (!post || !msg) {
        free(post);
        free(msg);
        return -1;
    }
    memcpy(post, llr, (size_t)K);
    for (int t = 0; t < R; t++)
        post[K + 360 * (t % q) + t / q] = llr[K + t];
    int it = 0, good;
    for (;;) {
        good = 1;
        for (int i = 0; i < q && good; i++)
            good &= check_any(c, i, post);
        if (good || it >= max_iter)
            break;
        for (int i = 0; i < q; i++)
            if (layer_any(c, i, post, msg, it == 0)) {
                free(post);
                free(msg);
                return -1;
            }
        it++;
    }
    if (nbits <= 0 || nbits > K)
        nbits = K;
    memset(out, 0, (size_t)(nbits + 7) / 8);
    for (int v = 0; v < nbits; v++)
        if (post[v] < 0)
            out[v / 8] |= (uint8_t)(0x80 >> (v % 8));
    *ok = good;
    free(post);
    free(msg);
    return it;
}

/* ---- the two-frame packed path (dvb-rx_amd/csrc/ldpc_pk.h) --------------- */

/* learned parameters of the two slots for this layer: rows (S, B) by
 * check degree; NULL rows: the plain decoder's arithmetic (pk_mags) */
struct PkRows {
    const uint8_t* S[2];
    const int16_t* B[2];
};

/* hand-over rounds run and level steps they replace (statistics) */
static long g_pk_rounds, g_pk_levels;

extern "C" void kmodel_pk_rounds(long* rounds, long* levels)
{
    *rounds = g_pk_rounds;
    *levels = g_pk_levels;
}

/* DEG: the kernel instance's layer degree bound (DM), the layer's own
 * degree D <= DEG is a run-time value as in the kernel */
template <int DEG, bool NMS>
static void pk_model_layer(const lg_code* c, int i, uint16_t* post, uint32_t* msg, uint32_t sel_even,
                           uint32_t sel_odd, const PkRows& rows)
{
    const int K = c->k, q = c->q;
    const uint64_t* meta = c->meta + (size_t)i * 360;
    static PkNode<DEG> nodes[360];
    const int W = LG_PK_WORDS(c->layer_start[i + 1] - c->layer_start[i]);
    static uint32_t mw[360][LG_PK_WORDS(28)];
    uint32_t* m = msg + (size_t)c->pk_msg_off[i] * 360;
    lg_pk ps[360], pb[360];
    for (int j = 0; j < 360; j++) {
        for (int w = 0; w < W; w++)
            mw[j][w] = m[w * 360 + j];
        ps[j] = lg_pk{ 16, 16 };
        pb[j] = lg_pk{ 16, 16 };
        if (NMS) {
            const int dc = c->layer_start[i + 1] - c->layer_start[i] + 2 - (i == 0 && j == 0);
            ps[j] = lg_pk{ (short)rows.S[0][dc], (short)rows.S[1][dc] };
            pb[j] = lg_pk{ (short)rows.B[0][dc], (short)rows.B[1][dc] };
        }
    }
    const uint32_t* xent = c->xent + c->layer_start[i];
    const uint64_t* pos = c->pos + (size_t)i * 360;
    const uint32_t dupm = c->dupm[i];
    const int D = c->layer_start[i + 1] - c->layer_start[i];
    for (int j = 359; j >= 0; j--) {
        pk_addrs<DEG>(nodes[j].ad, xent, D, j);
        if (c->nlev[i] > 1)
            pk_phase_a<DEG, true>(nodes[j], post, D, i, j, K, q, LG_META_DEP(meta[j]), mw[j], sel_even, sel_odd);
        else
            pk_phase_a<DEG, false>(nodes[j], post, D, i, j, K, q, 0, mw[j], sel_even, sel_odd);
    }
    /* hand-over rounds as the kernel runs them: all nodes of a round
     * together (here in decreasing j), two buffers, stop at the first
     * round that reads what the round before read, or after levels - 1 */
    if (c->nlev[i] > 1) {
        static PkPart part[360];
        static uint16_t hb[2][LG_MAX_XFER * 360];
        for (int j = 359; j >= 0; j--) {
            part[j] = pk_part(nodes[j]);
            pk_fold_dep0<DEG>(nodes[j], D, LG_META_DEP(meta[j]));
            pk_put_hand<DEG, NMS>(nodes[j], hb[0], xent, D, dupm, LG_META_SUCC(meta[j]), pos[j], ps[j], pb[j]);
        }
        for (int r = 1; r < c->nlev[i]; r++) {
            uint32_t chg = 0;
            for (int j = 359; j >= 0; j--) {
                pk_restore(nodes[j], part[j]);
                chg |= pk_take_hand<DEG>(nodes[j], hb[(r - 1) & 1], xent, D, dupm, LG_META_DEP(meta[j]), pos[j],
                                         mw[j], sel_even, sel_odd);
            }
            g_pk_rounds++;
            if (r == c->nlev[i] - 1)
                break;
            for (int j = 359; j >= 0; j--)
                pk_put_hand<DEG, NMS>(nodes[j], hb[r & 1], xent, D, dupm, LG_META_SUCC(meta[j]), pos[j], ps[j],
                                      pb[j]);
            if (!chg)
                break;
        }
        g_pk_levels += c->nlev[i] - 1;
    }
    for (int j = 359; j >= 0; j--) {
        if (c->nlev[i] > 1)
            pk_phase_c<DEG, true, NMS>(nodes[j], post, D, LG_META_SUCC(meta[j]), m + j, 360, ps[j], pb[j]);
        else
            pk_phase_c<DEG, false, NMS>(nodes[j], post, D, 0, m + j, 360, ps[j], pb[j]);
    }
}

/* the kernel instance of the code: the smallest layer degree bound that
 * covers its layers (the library's choice, 4 8 12 16 20 28) */
static int pk_dm(const lg_code* c)
{
    static const int dms[] = { 4, 8, 12, 16, 20, 28 };
    for (int d : dms)
        if (c->max_layer_deg <= d)
            return d;
    return -1;
}

#define PK_DM_CASES(X) X(4) X(8) X(12) X(16) X(20) X(28)

static int pk_layer_any(const lg_code* c, int i, uint16_t* post, uint32_t* msg, uint32_t se, uint32_t so,
                        const PkRows* rows)
{
    static const PkRows none = { { nullptr, nullptr }, { nullptr, nullptr } };
    switch (pk_dm(c)) {
#define X(D)                                                         \
    case D:                                                          \
        if (rows)                                                    \
            pk_model_layer<D, true>(c, i, post, msg, se, so, *rows); \
        else                                                         \
            pk_model_layer<D, false>(c, i, post, msg, se, so, none); \
        return 0;
        PK_DM_CASES(X)
#undef X
    }
    return -1;
}

static uint32_t pk_check_any(const lg_code* c, int i, int j, const uint16_t* post)
{
    const uint32_t* e = c->xent + c->layer_start[i];
    const int D = c->layer_start[i + 1] - c->layer_start[i];
    switch (pk_dm(c)) {
#define X(DMX) \
    case DMX:  \
        return pk_check<DMX>(post, e, D, i, j, c->k, c->q);
        PK_DM_CASES(X)
#undef X
    }
    return 0;
}

/* Two frames of one code in the two slots; frame B enters slot 1 `delay`
 * iterations after frame A started in slot 0 (as when a slot is refilled
 * while the other one runs).  Each slot stops on its own syndrome.
 * With a learned table (S, B: T rows of 33, NULL: plain decoder) a slot
 * runs up to max_iter learned iterations (row min(it, T-1)); if its
 * syndrome still fails and fallback > 0 the frame is loaded again and
 * decoded by the plain arithmetic for up to fallback iterations, as the
 * kernel's learned mode does. */
static int pk_pair(int code, const int8_t* llr_a, const int8_t* llr_b, int delay, int max_iter, int nbits,
                   const uint8_t* S, const int16_t* B, int T, int fallback, uint8_t* out_a, uint8_t* out_b,
                   int* it_out, int* ok_out, int* fb_out)
{
    const lg_code* c = lg_code_get(code);
    if (!c)
        return -1;
    const int K = c->k, q = c->q, N = c->n;
    uint16_t* post = (uint16_t*)calloc((size_t)N, sizeof(uint16_t));
    uint32_t* msg = (uint32_t*)calloc((size_t)c->pk_msg_words, sizeof(uint32_t));
    if (!post || !msg) {
        free(post);
        free(msg);
        return -1;
    }
    static uint8_t plainS[33];
    static int16_t plainB[33];
    for (int d = 0; d < 33; d++) {
        plainS[d] = 16;
        plainB[d] = 16;
    }
    const int8_t* src[2] = { llr_a, llr_b };
    uint8_t* dst[2] = { out_a, out_b };
    int it[2] = { 0, 0 }, active[2] = { 0, 0 }, started[2] = { 0, 0 }, phase[2] = { 0, 0 }, lit[2] = { 0, 0 };
    if (nbits <= 0 || nbits > K)
        nbits = K;
    auto load = [&](int s) {
        for (int v = 0; v < N; v++) {
            const int pos = v < K ? v : K + 360 * ((v - K) % q) + (v - K) / q;
            post[pos] = (uint16_t)((post[pos] & (s ? 0x00ff : 0xff00)) | (uint16_t)(uint8_t)src[s][v] << (8 * s));
        }
        it[s] = 0;
    };
    for (int t = 0;; t++) {
        for (int s = 0; s < 2; s++) {
            if (started[s] || t < (s ? delay : 0))
                continue;
            load(s);
            started[s] = active[s] = 1;
        }
        uint32_t good = 3;
        for (int i = 0; i < q; i++)
            for (int j = 0; j < 360; j++)
                good &= pk_check_any(c, i, j, post);
        for (int s = 0; s < 2; s++) {
            if (!active[s])
                continue;
            const int ok = (good >> s) & 1;
            const int cap = S && phase[s] == 1 ? fallback : max_iter;
            if (ok || it[s] >= cap) {
                if (S && !ok && phase[s] == 0 && fallback > 0) {
                    lit[s] = it[s];
                    phase[s] = 1;
                    load(s);
                    continue;
                }
                memset(dst[s], 0, (size_t)(nbits + 7) / 8);
                for (int v = 0; v < nbits; v++)
                    if ((int8_t)(post[v] >> (8 * s)) < 0)
                        dst[s][v / 8] |= (uint8_t)(0x80 >> (v % 8));
                it_out[s] = lit[s] + it[s];
                ok_out[s] = ok;
                if (fb_out)
                    fb_out[s] = phase[s];
                active[s] = 0;
            }
        }
        if (!active[0] && !active[1] && started[0] && started[1])
            break;
        const uint32_t se = pk_msg_sel(0, !active[0] || it[0] == 0, !active[1] || it[1] == 0);
        const uint32_t so = pk_msg_sel(1, !active[0] || it[0] == 0, !active[1] || it[1] == 0);
        PkRows rows;
        for (int s = 0; s < 2; s++) {
            const int learned = S && active[s] && phase[s] == 0;
            const int r = it[s] < T ? it[s] : T - 1;
            rows.S[s] = learned ? S + (size_t)r * 33 : plainS;
            rows.B[s] = learned ? B + (size_t)r * 33 : plainB;
        }
        if (active[0] || active[1])
            for (int i = 0; i < q; i++)
                if (pk_layer_any(c, i, post, msg, se, so, S ? &rows : nullptr)) {
                    free(post);
                    free(msg);
                    return -1;
                }
        for (int s = 0; s < 2; s++)
            it[s] += active[s];
    }
    free(post);
    free(msg);
    return 0;
}

extern "C" int kmodel_pk_pair(int code, const int8_t* llr_a, const int8_t* llr_b, int delay, int max_iter,
                              int nbits, uint8_t* out_a, uint8_t* out_b, int* it_out, int* ok_out)
{
    return pk_pair(code, llr_a, llr_b, delay, max_iter, nbits, nullptr, nullptr, 0, 0, out_a, out_b, it_out, ok_out,
                   nullptr);
}

extern "C" int kmodel_pk_pair_nms(int code, const int8_t* llr_a, const int8_t* llr_b, int delay, const uint8_t* S,
                                  const int16_t* B, int T, int max_iter, int fallback, int nbits, uint8_t* out_a,
                                  uint8_t* out_b, int* it_out, int* ok_out, int* fb_out)
{
    return pk_pair(code, llr_a, llr_b, delay, max_iter, nbits, S, B, T, fallback, out_a, out_b, it_out, ok_out,
                   fb_out);
}

/* The schedule's own invariants: no two nodes of one level of a layer
 * share a variable, and the masks name exactly the shared edges.  Returns
 * the number of violations. */
int kmodel_check_levels(int code)
{
    const lg_code* c = lg_code_get(code);
    if (!c)
        return -1;
    int bad = 0;
    int* seen = (int*)malloc(sizeof(int) * (size_t)c->n);
    int* count = (int*)malloc(sizeof(int) * (size_t)c->n);
    if (!seen || !count) {
        free(seen);
        free(count);
        return -1;
    }
    for (int i = 0; i < c->q; i++) {
        const int st = c->layer_start[i], deg = c->layer_start[i + 1] - st;
        for (int l = 0; l < c->nlev[i]; l++) {
            for (int v = 0; v < c->n; v++)
                seen[v] = -1;
            for (int j = 0; j < 360; j++) {
                if (c->level[360 * i + j] != l)
                    continue;
                for (int e = 0; e < lg_node_degree(deg, i, j); e++) {
                    const int ad = lg_edge_addr(c->ent + st, deg, e, i, j, c->k, c->q);
                    if (seen[ad] >= 0)
                        bad++;
                    seen[ad] = j;
                }
            }
        }
        /* masks: an info edge is dependent iff some earlier node of the
         * layer touches its bit, successor-shared iff a later one does */
        for (int v = 0; v < c->n; v++)
            count[v] = 0;
        for (int j = 0; j < 360; j++)
            for (int e = 0; e < deg; e++)
                count[lg_edge_addr(c->ent + st, deg, e, i, j, c->k, c->q)]++;
        for (int v = 0; v < c->n; v++)
            seen[v] = 0;
        for (int j = 0; j < 360; j++) {
            const uint64_t m = c->meta[(size_t)360 * i + j];
            if (LG_META_LEVEL(m) != c->level[360 * i + j])
                bad++;
            for (int e = 0; e < deg; e++) {
                const int ad = lg_edge_addr(c->ent + st, deg, e, i, j, c->k, c->q);
                const int before = seen[ad], after = count[ad] - seen[ad] - 1;
                if (((LG_META_DEP(m) >> e) & 1) != (before > 0) || ((LG_META_SUCC(m) >> e) & 1) != (after > 0))
                    bad++;
                seen[ad]++;
            }
        }
    }
    free(seen);
    free(count);
    return bad;
}
