/*
 * LDPC code registry and layer-schedule compiler, see ldpc_codes.h.
 */
#include "ldpc_codes.h"

#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#include "dvbs2_table_types.h"
#include "dvbs2_tables.inc"

#define NCODES ((int)(sizeof(LG_CODES) / sizeof(*LG_CODES)))

static lg_code g_codes[sizeof(LG_CODES) / sizeof(*LG_CODES)];
static pthread_once_t g_once = PTHREAD_ONCE_INIT;
static int g_build_failed;

/* Split-kernel tables (lg_code.sk_*), from the expanded code: per layer
 * and lane an address record (the LDS byte offsets of the lane's slot
 * posteriors), and for layers with nodes sharing bits a tail record (slot
 * masks and the hand-over cells of the two-frame kernel's rounds).  The
 * records of one layer are consecutive, lane after lane, so a lane reads
 * its record with a few wide loads at a fixed offset from the layer's. */
static int build_split(lg_code* c)
{
    const int q = c->q, K = c->k;
    int dmax = 0, dmin = 1 << 30;
    for (int i = 0; i < q; i++) {
        const int d = c->layer_start[i + 1] - c->layer_start[i];
        if (d > dmax)
            dmax = d;
        if (d < dmin)
            dmin = d;
    }
    const int ks = (dmax + 3) / 2;
    if (ks > 16)
        return -1;
    const int wa = SK_WORDS(ks), ra = SK_RA(ks), rb = SK_RB(ks);
    c->sk_ks = ks;
    c->sk_wide = (dmin + 2) / 2 < ks - 1;
    int words = 0;
    for (int i = 0; i < q; i++)
        words += LG_LANES * (ra + (c->nlev[i] > 1 ? rb : 0));
    c->sk_words = words;
    c->sk_msg_words = q * LG_LANES * wa;
    c->sk_lay = calloc((size_t)q * 4, sizeof(uint32_t));
    c->sk_tab = calloc((size_t)words, sizeof(uint32_t));
    if (!c->sk_lay || !c->sk_tab)
        return -1;
    /* barriers: layer i must wait for the workgroup (bit 16) unless it
     * shares no information group with the layers since the last barrier
     * (a parity bit passes between the two lanes of one node, in one
     * wave); the first layer of an iteration and the layers with nodes
     * sharing bits, and the one after them, always wait */
    uint8_t* used = calloc((size_t)(K / LG_M) + 1, 1);
    if (!used)
        return -1;
    int off = 0;
    for (int i = 0; i < q; i++) {
        const int s = c->layer_start[i], d = c->layer_start[i + 1] - s, T = d + 2;
        const int tail = c->nlev[i] > 1;
        /* information edges by position: in layers with nodes sharing bits
         * the edges of the shared groups first (positions 2 .. 1 + S), so
         * the rounds touch slots 1 .. kde - 1 only (the node's arithmetic
         * does not depend on the order of its edges) */
        int order[LG_MAX_CN_DEG], S = 0;
        for (int e = 0; e < d; e++)
            if (tail && ((c->dupm[i] >> e) & 1))
                order[S++] = e;
        for (int e = 0, o = S; e < d; e++)
            if (!(tail && ((c->dupm[i] >> e) & 1)))
                order[o++] = e;
        const int kde = tail ? 1 + (S + 1) / 2 : 0;
        if (kde - 1 > SK_KD(ks)) {
            free(used);
            return -1;
        }
        int bar = i == 0 || tail || c->nlev[i - 1] > 1;
        for (int e = 0; e < d && !bar; e++)
            bar = used[(c->ent[s + e] >> 16) / LG_M];
        if (bar)
            memset(used, 0, (size_t)(K / LG_M) + 1);
        for (int e = 0; e < d; e++)
            used[(c->ent[s + e] >> 16) / LG_M] = 1;
        uint32_t* L = c->sk_lay + 4 * i;
        const int arec = off, trec = off + LG_LANES * ra;
        off += LG_LANES * (ra + (tail ? rb : 0));
        L[0] = (uint32_t)T | (uint32_t)c->nlev[i] << 8 | (uint32_t)bar << 16 | (uint32_t)kde << 24;
        L[1] = (uint32_t)(i * LG_LANES * wa);
        L[2] = (uint32_t)arec;
        L[3] = tail ? (uint32_t)trec : 0u;
        for (int t = 0; t < LG_LANES; t++) {
            const int j = t >> 1, h = t & 1;
            const uint64_t mt = c->meta[(size_t)i * LG_M + j], ps = c->pos[(size_t)i * LG_M + j];
            uint32_t dep = 0, succ = 0;
            uint32_t* ar = c->sk_tab + arec + (size_t)t * ra;
            uint32_t* tr = c->sk_tab + trec + (size_t)t * rb;
            for (int k = 0; k < ks; k++) {
                const int p = 2 * k + h;
                uint32_t ad = (uint32_t)c->n, cell = 0;
                if (p == 0) {
                    ad = (uint32_t)(K + 360 * i + j);
                } else if (p == 1) {
                    if (i)
                        ad = (uint32_t)(K + 360 * (i - 1) + j);
                    else if (j)
                        ad = (uint32_t)(K + 360 * (q - 1) + j - 1);
                } else if (p < T) {
                    const int e = order[p - 2];
                    const uint32_t xe = c->xent[s + e];
                    const int m = (j - LG_XENT_SHIFT(xe) + LG_M) % LG_M;
                    ad = (xe >> 16) + (uint32_t)m;
                    if (tail && ((c->dupm[i] >> e) & 1)) {
                        const int r = (int)((ps >> (2 * e)) & 3);
                        cell = (uint32_t)((LG_XENT_T0(xe) + r) * LG_M + m);
                        dep |= ((LG_META_DEP(mt) >> e) & 1u) << k;
                        succ |= ((LG_META_SUCC(mt) >> e) & 1u) << k;
                    }
                }
                if (ad > (uint32_t)c->n || cell > 0xffffu || (cell && (k == 0 || k >= kde))) {
                    free(used);
                    return -1;
                }
                ar[k] = SK_POST_OFF + 2 * ad;
                if (tail && k <= SK_KD(ks))
                    tr[1 + k / 2] |= cell << (16 * (k & 1));
            }
            if (tail)
                tr[0] = dep | succ << 16;
        }
    }
    free(used);
    return 0;
}

/* Layer i holds the table addresses x with x mod q == i; node j of that
 * layer is check q*j + i, and the bit 360*row + m reaches check
 * (x + m*q) mod R (ldpc.c:240-249), i.e. node (x div q + m) mod 360. */
static int build_code(lg_code* c, int id)
{
    const struct lg_code_def* t = &LG_CODES[id];
    memset(c, 0, sizeof(*c));
    c->id = id;
    c->name = t->name;
    c->n = t->n;
    c->k = t->k;
    c->r = t->n - t->k;
    c->q = c->r / LG_M;
    c->short_frame = t->n == 16200;
    c->links_total = t->links_total;
    c->links_max_cn = t->links_max_cn;
    const int q = c->q;

    int naddr = 0;
    for (int g = 0; g < t->ngroups; g++)
        naddr += LG_GROUPS[t->group_off + g].deg * LG_GROUPS[t->group_off + g].rows;
    c->nent = naddr;
    c->layer_start = calloc((size_t)q + 1, sizeof(uint16_t));
    c->ent = malloc(sizeof(uint32_t) * (size_t)naddr);
    c->level = calloc((size_t)c->r, 1);
    c->nlev = calloc((size_t)q, 1);
    int* cnt = calloc((size_t)q, sizeof(int));
    if (!c->layer_start || !c->ent || !c->level || !c->nlev || !cnt) {
        free(cnt);
        return -1;
    }
    /* count per layer, then place (stable: table order within a layer) */
    const uint16_t* a = LG_ADDR + t->addr_off;
    for (int i = 0; i < naddr; i++)
        cnt[a[i] % q]++;
    for (int i = 0; i < q; i++)
        c->layer_start[i + 1] = (uint16_t)(c->layer_start[i] + cnt[i]);
    memset(cnt, 0, sizeof(int) * (size_t)q);
    int row = 0, p = 0;
    for (int g = 0; g < t->ngroups; g++) {
        const int deg = LG_GROUPS[t->group_off + g].deg;
        for (int rr = 0; rr < LG_GROUPS[t->group_off + g].rows; rr++, row++) {
            for (int e = 0; e < deg; e++, p++) {
                const int x = a[p], li = x % q;
                c->ent[c->layer_start[li] + cnt[li]++] = (uint32_t)(row * LG_M) << 16 | (uint32_t)(x / q);
            }
        }
    }
    free(cnt);
    if (row * LG_M != c->k)
        return -1;

    /* levels: owner[b] = last node of the current layer touching bit b */
    int16_t* owner = malloc(sizeof(int16_t) * (size_t)c->k);
    if (!owner)
        return -1;
    for (int i = 0; i < q; i++) {
        const int s = c->layer_start[i], d = c->layer_start[i + 1] - s;
        if (d > c->max_layer_deg)
            c->max_layer_deg = d;
        if (d + 2 > LG_MAX_CN_DEG) {
            free(owner);
            return -1;
        }
        for (int e = 0; e < d; e++) {
            const int base = (int)(c->ent[s + e] >> 16);
            for (int m = 0; m < LG_M; m++)
                owner[base + m] = -1;
        }
        int maxl = 0;
        uint8_t* lv = c->level + (size_t)i * LG_M;
        for (int j = 0; j < LG_M; j++) {
            int l = 0;
            for (int e = 0; e < d; e++) {
                const int base = (int)(c->ent[s + e] >> 16), sh = (int)(c->ent[s + e] & 0xffff);
                const int b = base + (j - sh + LG_M) % LG_M;
                if (owner[b] >= 0 && lv[owner[b]] + 1 > l)
                    l = lv[owner[b]] + 1;
                owner[b] = (int16_t)j;
            }
            if (l > 255) {
                free(owner);
                return -1;
            }
            lv[j] = (uint8_t)l;
            if (l)
                c->dep_nodes++;
            if (l > maxl)
                maxl = l;
        }
        c->nlev[i] = (uint8_t)(maxl + 1);
        c->level_steps += maxl + 1;
        if (maxl + 1 > c->max_levels)
            c->max_levels = maxl + 1;
    }
    /* dependency masks: edge e of node j is dependent if an earlier node of
     * the layer touches the same bit, successor-shared if a later one does */
    c->meta = calloc((size_t)c->r, sizeof(uint64_t));
    c->msg_off = malloc(sizeof(int) * (size_t)q);
    c->pk_msg_off = malloc(sizeof(int) * (size_t)q);
    if (!c->meta || !c->msg_off || !c->pk_msg_off) {
        free(owner);
        return -1;
    }
    int rows = 0, pk_rows = 0;
    for (int i = 0; i < q; i++) {
        const int s = c->layer_start[i], d = c->layer_start[i + 1] - s;
        c->msg_off[i] = rows;
        rows += (d + 2 + 3) / 4;
        c->pk_msg_off[i] = pk_rows;
        pk_rows += (d + 2 + 1) / 2;
        uint64_t* mt = c->meta + (size_t)i * LG_M;
        for (int j = 0; j < LG_M; j++)
            mt[j] = c->level[(size_t)i * LG_M + j];
        if (c->nlev[i] == 1)
            continue;
        for (int e = 0; e < d; e++) {
            const int base = (int)(c->ent[s + e] >> 16);
            for (int m = 0; m < LG_M; m++)
                owner[base + m] = -1;
        }
        /* owner[b] = last node so far touching b, with its edge index in
         * the high bits */
        for (int j = 0; j < LG_M; j++) {
            for (int e = 0; e < d; e++) {
                const int base = (int)(c->ent[s + e] >> 16), sh = (int)(c->ent[s + e] & 0xffff);
                const int b = base + (j - sh + LG_M) % LG_M;
                if (owner[b] >= 0) {
                    const int pj = owner[b] & 0x1ff, pe = owner[b] >> 9;
                    mt[j] |= (uint64_t)1 << (8 + e);
                    mt[pj] |= (uint64_t)1 << (36 + pe);
                }
                owner[b] = (int16_t)(j | e << 9);
            }
        }
    }
    c->msg_words = rows * LG_M;
    c->pk_msg_words = pk_rows * LG_M;

    /* hand-over slots: per layer, the groups that occur k > 1 times get
     * k - 1 consecutive hand-over indices; each node's rank on such an
     * edge is the number of earlier nodes of the layer touching its bit */
    c->xent = malloc(sizeof(uint32_t) * (size_t)naddr);
    c->pos = calloc((size_t)c->r, sizeof(uint64_t));
    c->dupm = calloc((size_t)q, sizeof(uint32_t));
    if (!c->xent || !c->pos || !c->dupm) {
        free(owner);
        return -1;
    }
    for (int i = 0; i < q; i++) {
        const int s = c->layer_start[i], d = c->layer_start[i + 1] - s;
        int nx = 0;
        for (int e = 0; e < d; e++) {
            const uint32_t base = c->ent[s + e] >> 16;
            int k = 0, first = e;
            for (int f = 0; f < d; f++) {
                if ((c->ent[s + f] >> 16) == base) {
                    k++;
                    if (f < first)
                        first = f;
                }
            }
            if (k > 4) {
                free(owner);
                return -1;
            }
            uint32_t t0 = 0;
            if (k > 1) {
                c->dupm[i] |= 1u << e;
                if (first == e) {
                    t0 = (uint32_t)nx;
                    nx += k - 1;
                } else {
                    t0 = LG_XENT_T0(c->xent[s + first]);
                }
            }
            c->xent[s + e] = c->ent[s + e] | (uint32_t)(k - 1) << 9 | t0 << 11;
        }
        if (nx > LG_MAX_XFER || (nx > 0) != (c->nlev[i] > 1)) {
            free(owner);
            return -1;
        }
        if (nx > c->max_xfer)
            c->max_xfer = nx;
        if (!nx)
            continue;
        /* rank: count of earlier nodes touching the bit (owner[] counts) */
        for (int e = 0; e < d; e++) {
            const int base = (int)(c->ent[s + e] >> 16);
            for (int m = 0; m < LG_M; m++)
                owner[base + m] = 0;
        }
        for (int j = 0; j < LG_M; j++) {
            for (int e = 0; e < d; e++) {
                const int base = (int)(c->ent[s + e] >> 16), sh = (int)(c->ent[s + e] & 0xffff);
                const int b = base + (j - sh + LG_M) % LG_M;
                c->pos[(size_t)i * LG_M + j] |= (uint64_t)(owner[b] & 3) << (2 * e);
                owner[b]++;
            }
        }
    }
    free(owner);
    if (build_split(c))
        return -1;
    /* every edge counted: info edges plus the 2R-1 of the staircase */
    if ((long)naddr * LG_M + 2L * c->r - 1 != c->links_total)
        return -1;
    return 0;
}

static void build_all(void)
{
    for (int i = 0; i < NCODES; i++)
        if (build_code(&g_codes[i], i))
            g_build_failed = 1;
}

int lg_code_count(void) { return NCODES; }

const lg_code* lg_code_get(int id)
{
    if (id < 0 || id >= NCODES)
        return NULL;
    pthread_once(&g_once, build_all);
    if (g_build_failed)
        return NULL;
    return &g_codes[id];
}

int lg_code_lookup(int short_frame, const char* rate)
{
    if (!rate)
        return -1;
    for (size_t i = 0; i < sizeof(LG_RATES) / sizeof(*LG_RATES); i++)
        if (LG_RATES[i].short_frame == (short_frame != 0) && strcmp(LG_RATES[i].rate, rate) == 0)
            return LG_RATES[i].code;
    return -1;
}

/* "2_3_t2" / "3_5_t2" (the DVB-T2 codes) take the BCH parameters of their
 * DVB-S2 rate, as libdvbrx/src/dvbt2.c:1963-1967 does. */
static int same_rate(const char* row, const char* rate)
{
    const size_t n = strlen(rate);
    if (n > 3 && strcmp(rate + n - 3, "_t2") == 0)
        return strlen(row) == n - 3 && strncmp(row, rate, n - 3) == 0;
    return strcmp(row, rate) == 0;
}

int lg_fec_lookup(int short_frame, const char* rate, int* kbch, int* nbch, int* t)
{
    if (!rate)
        return -1;
    for (size_t i = 0; i < sizeof(LG_FEC) / sizeof(*LG_FEC); i++) {
        if (LG_FEC[i].short_frame == (short_frame != 0) && same_rate(LG_FEC[i].rate, rate)) {
            if (kbch)
                *kbch = LG_FEC[i].kbch;
            if (nbch)
                *nbch = LG_FEC[i].nbch;
            if (t)
                *t = LG_FEC[i].t;
            return 0;
        }
    }
    return -1;
}

const char* lg_fec_rate(int short_frame, int kbch, int nbch, int t)
{
    for (size_t i = 0; i < sizeof(LG_FEC) / sizeof(*LG_FEC); i++)
        if (LG_FEC[i].short_frame == (short_frame != 0) && LG_FEC[i].kbch == kbch && LG_FEC[i].nbch == nbch &&
            LG_FEC[i].t == t)
            return LG_FEC[i].rate;
    return NULL;
}

int lg_fec_index(int short_frame, const char* rate)
{
    int n = 0;
    if (!rate)
        return -1;
    for (size_t i = 0; i < sizeof(LG_FEC) / sizeof(*LG_FEC); i++) {
        if (LG_FEC[i].short_frame != (short_frame != 0))
            continue;
        if (same_rate(LG_FEC[i].rate, rate))
            return n;
        n++;
    }
    return -1;
}

const struct lg_modcod_row* lg_modcod_get(int pls)
{
    pls &= 0xFE;
    for (size_t i = 0; i < sizeof(LG_MODCODS) / sizeof(*LG_MODCODS); i++)
        if (LG_MODCODS[i].pls == pls)
            return &LG_MODCODS[i];
    return NULL;
}

const float* lg_points(void) { return LG_POINTS; }
