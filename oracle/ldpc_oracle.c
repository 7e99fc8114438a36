/*
 * CPU oracle for the LDPC decode path (TEST INFRASTRUCTURE ONLY, see
 * ldpc_oracle.h).  Scalar restatement of libdvbrx/src/ldpc.c: every
 * function cites the lines it follows.  Frames are "lanes" as in the
 * reference: data[v * lanes + f] is the posterior of variable v in frame f.
 */
#include "ldpc_oracle.h"

#include <stdlib.h>
#include <string.h>

#include "../dvb-rx_amd/csrc/dvbs2_table_types.h"
#include "../dvb-rx_amd/csrc/dvbs2_tables.inc"

/* ---- byte arithmetic: the scalar spec of ldpc.c:115-163 ----------------- */

static inline int8_t sat8(int x) { return (int8_t)(x > 127 ? 127 : x < -128 ? -128 : x); }
static inline int8_t qadd(int8_t a, int8_t b) { return sat8(a + b); }
static inline int8_t qsub(int8_t a, int8_t b) { return sat8(a - b); }
/* |a| as an unsigned byte: -128 gives 128 (ldpc.c:156) */
static inline uint8_t absu(int8_t a) { return (uint8_t)(a < 0 ? -(int)a : a); }
static inline uint8_t minu(uint8_t a, uint8_t b) { return a < b ? a : b; }
static inline uint8_t maxu(uint8_t a, uint8_t b) { return a > b ? a : b; }
static inline uint8_t subsu(uint8_t a, uint8_t b) { return (uint8_t)(a > b ? a - b : 0); }
/* a if b > 0, -a if b < 0, 0 if b == 0, in bytes (ldpc.c:140-143) */
static inline int8_t vsign(int8_t a, int8_t b) { return (int8_t)(b > 0 ? a : b < 0 ? -a : 0); }
static inline int8_t smax(int8_t a, int8_t b) { return a > b ? a : b; }
static inline int8_t smin(int8_t a, int8_t b) { return a < b ? a : b; }

/* ---- code expansion: drx_ldpc_new, ldpc.c:199-260 ----------------------- */

#define OM 360
#define MAX_DEG 32

struct oracle_ldpc {
    int N, K, M, R, q, CNL, LT, kldpc;
    uint16_t* pos; /* R * CNL information links per check, layer order */
    uint8_t* cnc;  /* information links per check (original order) */
    int lanes;     /* capacity of the buffers below, in frames */
    int8_t* data;  /* N * lanes */
    int8_t* pty;   /* R * lanes, layer order */
    int8_t* bnl;   /* LT * lanes check-to-variable messages */
    int last_it;
};

static const struct lg_code_def* find_code(int short_frame, const char* rate)
{
    for (size_t i = 0; i < sizeof(LG_RATES) / sizeof(*LG_RATES); i++)
        if (LG_RATES[i].short_frame == (short_frame != 0) && strcmp(LG_RATES[i].rate, rate) == 0)
            return &LG_CODES[LG_RATES[i].code];
    return NULL;
}

oracle_ldpc* oracle_ldpc_new(int short_frame, const char* rate, int kldpc)
{
    const struct lg_code_def* t = rate ? find_code(short_frame, rate) : NULL;
    if (!t)
        return NULL;
    oracle_ldpc* d = calloc(1, sizeof(*d));
    if (!d)
        return NULL;
    d->N = t->n;
    d->K = t->k;
    d->M = OM;
    d->R = t->n - t->k;
    d->q = d->R / d->M;
    d->CNL = t->links_max_cn - 2;
    d->LT = t->links_total;
    d->kldpc = kldpc > 0 && kldpc <= t->k ? kldpc : t->k;

    uint16_t* pos = calloc((size_t)d->R * d->CNL, sizeof(uint16_t));
    d->pos = malloc(sizeof(uint16_t) * (size_t)d->R * d->CNL);
    d->cnc = calloc((size_t)d->R, 1);
    if (!pos || !d->pos || !d->cnc) {
        free(pos);
        oracle_ldpc_free(d);
        return NULL;
    }
    /* ldpc.c:238-250: information bit j of row r, group position m, reaches
     * accumulator (addr + m q) mod R for each address of the row */
    const uint16_t* row = LG_ADDR + t->addr_off;
    int j = 0;
    for (int g = 0; g < t->ngroups; g++) {
        const int deg = LG_GROUPS[t->group_off + g].deg;
        for (int r = 0; r < LG_GROUPS[t->group_off + g].rows; r++, row += deg) {
            for (int m = 0; m < d->M; m++, j++) {
                for (int n = 0; n < deg; n++) {
                    const int acc = (row[n] + m * d->q) % d->R;
                    pos[d->CNL * acc + d->cnc[acc]++] = (uint16_t)j;
                }
            }
        }
    }
    /* ldpc.c:251-255: check q*j + i is stored as M*i + j */
    for (int i = 0; i < d->q; i++)
        for (int jj = 0; jj < d->M; jj++)
            memcpy(d->pos + (size_t)d->CNL * (d->M * i + jj), pos + (size_t)d->CNL * (d->q * jj + i),
                   sizeof(uint16_t) * d->CNL);
    free(pos);
    return d;
}

void oracle_ldpc_free(oracle_ldpc* d)
{
    if (!d)
        return;
    free(d->pos);
    free(d->cnc);
    free(d->data);
    free(d->pty);
    free(d->bnl);
    free(d);
}

int oracle_ldpc_n(const oracle_ldpc* d) { return d->N; }

static int reserve(oracle_ldpc* d, int lanes)
{
    if (lanes <= d->lanes)
        return 0;
    free(d->data);
    free(d->pty);
    free(d->bnl);
    d->data = malloc((size_t)d->N * lanes);
    d->pty = malloc((size_t)d->R * lanes);
    d->bnl = malloc((size_t)d->LT * lanes);
    if (!d->data || !d->pty || !d->bnl) {
        d->lanes = 0;
        return -1;
    }
    d->lanes = lanes;
    return 0;
}

/* ---- decoder ------------------------------------------------------------ */

/* prev_parity, ldpc.c:277-284 */
static inline int prev_parity(const oracle_ldpc* d, int i, int j)
{
    if (i)
        return d->M * (i - 1) + j;
    if (j)
        return j + (d->q - 1) * d->M - 1;
    return -1;
}

/* bad, ldpc.c:286-305: some check of some lane has a zero posterior or an
 * odd number of negative ones */
static int bad(const oracle_ldpc* d, int L)
{
    for (int i = 0; i < d->q; i++) {
        const int cnt = d->cnc[i];
        for (int j = 0; j < d->M; j++) {
            const int cn = d->M * i + j;
            const int pp = prev_parity(d, i, j);
            const uint16_t* p = d->pos + (size_t)d->CNL * cn;
            for (int f = 0; f < L; f++) {
                int8_t cnv = vsign(1, d->pty[(size_t)cn * L + f]);
                if (pp >= 0)
                    cnv = vsign(cnv, d->pty[(size_t)pp * L + f]);
                for (int c = 0; c < cnt; c++)
                    cnv = vsign(cnv, d->data[(size_t)p[c] * L + f]);
                if (cnv <= 0)
                    return 1;
            }
        }
    }
    return 0;
}

/* check_node, ldpc.c:321-362, one lane: v[c] points at the deg posteriors
 * (information links, then the node's parity, then the previous parity),
 * bl[c] at its messages. */
static inline void check_node_lane(int8_t* const* v, int8_t* bl, int deg, int first)
{
    int8_t inp[MAX_DEG];
    uint8_t min0 = 255, min1 = 255;
    int8_t signs = 0;
    for (int c = 0; c < deg; c++) { /* CN_IN, ldpc.c:328-336 */
        const int8_t x = first ? *v[c] : qsub(*v[c], bl[c]);
        const uint8_t a = absu(x);
        inp[c] = x;
        min1 = minu(min1, maxu(min0, a));
        min0 = minu(min0, a);
        signs = (int8_t)(signs ^ x);
    }
    /* offset 1 and the cap at 127, ldpc.c:353 */
    const int8_t mn0 = (int8_t)subsu(minu(min0, 127), 1), mn1 = (int8_t)subsu(minu(min1, 127), 1);
    for (int c = 0; c < deg; c++) { /* CN_OUT, ldpc.c:337-344 */
        const int8_t x = inp[c];
        const int8_t eq = absu(x) == min0 ? (int8_t)-1 : 0;
        const int8_t o = vsign(smax(mn0, (int8_t)(eq & mn1)), (int8_t)((signs ^ x) | 127));
        *v[c] = qadd(x, o);
        bl[c] = smin(smax(o, -32), 31);
    }
}

void oracle_check_node(int deg, const int8_t* L, const int8_t* m, int first, int8_t* L_out,
                       int8_t* m_out)
{
    int8_t* v[MAX_DEG];
    for (int c = 0; c < deg && c < MAX_DEG; c++) {
        L_out[c] = L[c];
        m_out[c] = m[c];
        v[c] = &L_out[c];
    }
    check_node_lane(v, m_out, deg, first);
}

/* layer / update / update_first, ldpc.c:465-497: layers in order, the M
 * nodes of a layer one after the other, posteriors updated in place */
static void update(oracle_ldpc* d, int L, int first)
{
    int8_t* bl = d->bnl;
    int8_t* v[MAX_DEG];
    for (int i = 0; i < d->q; i++) {
        const int cnt = d->cnc[i];
        for (int j = 0; j < d->M; j++) {
            const int cn = d->M * i + j;
            const int pp = prev_parity(d, i, j);
            const int deg = cnt + 1 + (pp >= 0);
            const uint16_t* p = d->pos + (size_t)d->CNL * cn;
            for (int f = 0; f < L; f++) {
                for (int c = 0; c < cnt; c++)
                    v[c] = &d->data[(size_t)p[c] * L + f];
                v[cnt] = &d->pty[(size_t)cn * L + f];
                if (pp >= 0)
                    v[cnt + 1] = &d->pty[(size_t)pp * L + f];
                /* messages of this node for lane f: deg consecutive entries
                 * of the lane's message block */
                check_node_lane(v, bl + (size_t)f * d->LT, deg, first);
            }
            bl += deg;
        }
    }
}

/* load_frames + parity gather (ldpc.c:552-576, :624-626) */
static void load(oracle_ldpc* d, const int8_t* llr, int L)
{
    for (int f = 0; f < L; f++)
        for (int v = 0; v < d->N; v++)
            d->data[(size_t)v * L + f] = llr[(size_t)f * d->N + v];
    for (int i = 0; i < d->q; i++)
        for (int j = 0; j < d->M; j++)
            for (int f = 0; f < L; f++)
                d->pty[(size_t)(d->M * i + j) * L + f] = d->data[(size_t)(d->K + d->q * j + i) * L + f];
}

/* store_frames, ldpc.c:581-612: hard decision (x < 0) on the first kldpc
 * variables, packed MSB first */
static void store(const oracle_ldpc* d, uint8_t* out, int L)
{
    const int nbytes = (d->kldpc + 7) / 8;
    for (int f = 0; f < L; f++) {
        uint8_t* o = out + (size_t)f * nbytes;
        memset(o, 0, nbytes);
        for (int v = 0; v < d->kldpc; v++)
            if (d->data[(size_t)v * L + f] < 0)
                o[v / 8] |= (uint8_t)(0x80 >> (v % 8));
    }
}

/* drx_ldpc_decode, ldpc.c:614-638, for any number of lanes */
int oracle_ldpc_decode(oracle_ldpc* d, const int8_t* llr, uint8_t* out, int nframes, int max_iter)
{
    if (nframes <= 0 || reserve(d, nframes))
        return -1;
    /* messages live per lane in blocks of LT; the reference interleaves
     * them by lane, which changes nothing but the addresses */
    load(d, llr, nframes);
    int it = 0, ok;
    while (!(ok = !bad(d, nframes)) && it < max_iter) {
        update(d, nframes, it == 0);
        it++;
    }
    d->last_it = it;
    store(d, out, nframes);
    return ok ? it : -1;
}

void oracle_ldpc_decode_frames(oracle_ldpc* d, const int8_t* llr, uint8_t* out, int nframes,
                               int max_iter, int32_t* iters, uint8_t* ok)
{
    const int nbytes = (d->kldpc + 7) / 8;
    for (int f = 0; f < nframes; f++) {
        const int r = oracle_ldpc_decode(d, llr + (size_t)f * d->N, out + (size_t)f * nbytes, 1, max_iter);
        iters[f] = d->last_it;
        ok[f] = r >= 0;
    }
}

int oracle_ldpc_check(oracle_ldpc* d, const int8_t* llr)
{
    if (reserve(d, 1))
        return 0;
    load(d, llr, 1);
    return !bad(d, 1);
}
