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
#include "ldpc_sk.h"

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

/* ---- the two-frame split path (dvb-rx_amd/csrc/ldpc_pk.h, ldpc_sk.h) ----- */

/* hand-over rounds run and level steps they replace (statistics) */
static long g_pk_rounds, g_pk_levels;

extern "C" void kmodel_pk_rounds(long* rounds, long* levels)
{
    *rounds = g_pk_rounds;
    *levels = g_pk_levels;
}

/* learned parameters of the two slots for this layer: rows (S, B) by
 * check degree; NULL rows: the plain decoder's arithmetic */
struct PkRows {
    const uint8_t* S[2];
    const int16_t* B[2];
};

/* the lane pairs exchange minimums and sign (the kernel's DPP swap) */
template <int KS>
static void sk_model_pair(SkHalf<KS>* n)
{
    for (int j = 0; j < 360; j++) {
        SkHalf<KS>& a = n[2 * j];
        SkHalf<KS>& b = n[2 * j + 1];
        const lg_pk a0 = a.min0, a1 = a.min1, as = a.s;
        sk_combine(a, b.min0, b.min1, b.s);
        sk_combine(b, a0, a1, as);
    }
}

/* One layer of the split kernel (ldpc_sk_kern.hip) on the CPU: 720 lanes,
 * each phase over all lanes (in decreasing lane order: the kernel runs
 * them at once, the result must not depend on the order), the code's own
 * split tables. */
template <int KS, int KA, bool NMS>
static void sk_model_layer(const lg_code* c, int i, uint16_t* post, uint32_t* msg, uint32_t se, uint32_t so,
                           const PkRows& rows)
{
    const uint32_t* L = c->sk_lay + 4 * i;
    const int T = (int)(L[0] & 0xff), nlev = (int)((L[0] >> 8) & 0xff), kde = (int)(L[0] >> 24);
    const int WA = SK_WORDS(KS);
    const bool tail = nlev > 1;
    static SkHalf<KS> n[720];
    static uint32_t mw[720][8], tm[720], hw[720][8];
    static lg_pk ps[720], nb[720];
    static PkPart part[720];
    uint32_t* m = msg + L[1];
    for (int t = 0; t < 720; t++) {
        const uint32_t* ar = c->sk_tab + L[2] + (size_t)t * SK_RA(KS);
        const uint32_t* tr = c->sk_tab + L[3] + (size_t)t * SK_RB(KS);
        for (int w = 0; w < 8; w++)
            hw[t][w] = 0;
        for (int w = 0; w < WA; w++)
            mw[t][w] = m[(size_t)t * WA + w];
        for (int w = 0; tail && w <= SK_KD(KS) / 2; w++)
            hw[t][w] = tr[1 + w];
        tm[t] = tail ? tr[0] : 0;
        sk_addrs(n[t], ar);
        const int j = t >> 1;
        ps[t] = lg_pk{ 16, 16 };
        nb[t] = lg_pk{ -16, -16 };
        if (NMS) {
            const int dc = T - (i == 0 && j == 0);
            ps[t] = lg_pk{ (short)rows.S[0][dc], (short)rows.S[1][dc] };
            nb[t] = lg_pk{ (short)-rows.B[0][dc], (short)-rows.B[1][dc] };
        }
    }
    for (int t = 719; t >= 0; t--) {
        const uint32_t skip0 = (i == 0 && t == 1) ? ~0u : 0u;
        uint32_t raw[KS];
        sk_read<KS>(raw, post - SK_POST_OFF / 2, n[t]);
        if (tail)
            sk_phase_a<KS, KA, true>(n[t], raw, T, t & 1, skip0, tm[t] & 0xffff, mw[t], se, so);
        else
            sk_phase_a<KS, KA, false>(n[t], raw, T, t & 1, skip0, 0, mw[t], se, so);
    }
    if (tail) {
        static uint16_t hb[2][LG_MAX_XFER * 360];
        for (int t = 719; t >= 0; t--) {
            part[t] = sk_part(n[t]);
            sk_fold_dep0(n[t], tm[t] & 0xffff, kde);
        }
        sk_model_pair(n);
        for (int t = 719; t >= 0; t--)
            sk_put_hand<KS, NMS>(n[t], hb[0], hw[t], tm[t] >> 16, kde, ps[t], nb[t]);
        static uint32_t old[720][KS];
        for (int r = 1; r < nlev; r++) {
            uint32_t chg = 0;
            for (int t = 719; t >= 0; t--) {
                sk_restore(n[t], part[t]);
                sk_take_hand(n[t], hb[(r - 1) & 1], hw[t], tm[t] & 0xffff, kde, mw[t], se, so);
                sk_prev_hand<KS>(old[t], hb[(r - 1) & 1], hw[t], tm[t] >> 16, kde);
            }
            sk_model_pair(n);
            g_pk_rounds++;
            if (r == nlev - 1)
                break;
            for (int t = 719; t >= 0; t--)
                chg |= sk_put_hand_chg<KS, NMS>(n[t], hb[r & 1], hw[t], tm[t] >> 16, kde, old[t], ps[t], nb[t]);
            if (!chg)
                break;
        }
        g_pk_levels += nlev - 1;
    } else {
        sk_model_pair(n);
    }
    for (int t = 719; t >= 0; t--) {
        if (tail)
            sk_phase_c<KS, true, NMS>(n[t], post - SK_POST_OFF / 2, tm[t] >> 16, m + (size_t)t * WA, ps[t], nb[t]);
        else
            sk_phase_c<KS, false, NMS>(n[t], post - SK_POST_OFF / 2, 0, m + (size_t)t * WA, ps[t], nb[t]);
    }
}

#define SK_MODEL_INSTS(X) X(2, 0) X(3, 0) X(4, 0) X(4, 1) X(5, 0) X(6, 0) X(7, 0) X(7, 1) X(8, 0) X(9, 0) \
    X(10, 0) X(10, 1) X(11, 0) X(14, 0) X(15, 0)

static int pk_layer_any(const lg_code* c, int i, uint16_t* post, uint32_t* msg, uint32_t se, uint32_t so,
                        const PkRows* rows)
{
    static const PkRows none = { { nullptr, nullptr }, { nullptr, nullptr } };
#define X(KSX, W)                                                                                     \
    if (c->sk_ks == KSX && c->sk_wide == W) {                                                       \
        if (rows)                                                                                   \
            sk_model_layer<KSX, W ? 1 : KSX - 1, true>(c, i, post, msg, se, so, *rows);             \
        else                                                                                        \
            sk_model_layer<KSX, W ? 1 : KSX - 1, false>(c, i, post, msg, se, so, none);            \
        return 0;                                                                                   \
    }
    SK_MODEL_INSTS(X)
#undef X
    return -1;
}

/* syndrome of layer i for both slots (bit f: frame f holds every check) */
template <int KS, int KA>
static uint32_t sk_model_check(const lg_code* c, int i, const uint16_t* post)
{
    const uint32_t* L = c->sk_lay + 4 * i;
    const int T = (int)(L[0] & 0xff);
    uint32_t good = 3;
    for (int j = 0; j < 360; j++) {
        uint32_t par[2], nz[2];
        for (int h = 0; h < 2; h++) {
            const int t = 2 * j + h;
            SkHalf<KS> n;
            sk_addrs(n, c->sk_tab + L[2] + (size_t)t * SK_RA(KS));
            sk_check<KS, KA>(post - SK_POST_OFF / 2, n.ad, T, h, (i == 0 && t == 1) ? ~0u : 0u, par[h], nz[h]);
        }
        good &= sk_check_bits(par[0] ^ par[1], nz[0] & nz[1]);
    }
    return good;
}

static uint32_t pk_check_layer(const lg_code* c, int i, const uint16_t* post)
{
#define X(KSX, W)                                                        \
    if (c->sk_ks == KSX && c->sk_wide == W)                            \
        return sk_model_check<KSX, W ? 1 : KSX - 1>(c, i, post);
    SK_MODEL_INSTS(X)
#undef X
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
    /* the kernel's LDS: control words, then the posteriors (and the spare
     * cell of empty slots) at SK_POST_OFF, which the address records count
     * from */
    uint16_t* lds = (uint16_t*)calloc(SK_POST_OFF / 2 + (size_t)N + 1, sizeof(uint16_t));
    uint16_t* post = lds ? lds + SK_POST_OFF / 2 : nullptr;
    uint32_t* msg = (uint32_t*)calloc((size_t)c->sk_msg_words, sizeof(uint32_t));
    if (!lds || !msg) {
        free(lds);
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
            good &= pk_check_layer(c, i, post);
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
                    free(lds);
                    free(msg);
                    return -1;
                }
        for (int s = 0; s < 2; s++)
            it[s] += active[s];
    }
    free(lds);
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
