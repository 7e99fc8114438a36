/*
 * DVB-S2/S2X/T2 frame source: BCH + LDPC encoder, mapping, AWGN, int8
 * max-log demapper.  See include/dvbs2_sim.h.
 */
#include "dvbs2_sim.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#include "dvbs2_table_types.h"
#include "ldpc_codes.h"


/* ---- random numbers: xoshiro256** seeded by splitmix64 ----------------- */

typedef struct {
    uint64_t s[4];
    int have;
    double spare;
} rng_t;

static uint64_t splitmix64(uint64_t* x)
{
    uint64_t z = (*x += 0x9E3779B97F4A7C15ULL);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

static void rng_seed(rng_t* r, uint64_t seed, uint64_t stream)
{
    uint64_t x = seed ^ (stream * 0xD1B54A32D192ED03ULL);
    for (int i = 0; i < 4; i++)
        r->s[i] = splitmix64(&x);
    r->have = 0;
}

static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

static inline uint64_t rng_next(rng_t* r)
{
    uint64_t* s = r->s;
    const uint64_t res = rotl(s[1] * 5, 7) * 9, t = s[1] << 17;
    s[2] ^= s[0];
    s[3] ^= s[1];
    s[1] ^= s[2];
    s[0] ^= s[3];
    s[2] ^= t;
    s[3] = rotl(s[3], 45);
    return res;
}

/* uniform in (0, 1) */
static inline double rng_unit(rng_t* r) { return ((double)(rng_next(r) >> 11) + 0.5) * 0x1.0p-53; }

/* standard normal, Box-Muller */
static inline double rng_gauss(rng_t* r)
{
    if (r->have) {
        r->have = 0;
        return r->spare;
    }
    const double u1 = rng_unit(r), u2 = rng_unit(r);
    const double rad = sqrt(-2.0 * log(u1)), ang = 2.0 * M_PI * u2;
    r->spare = rad * sin(ang);
    r->have = 1;
    return rad * cos(ang);
}

/* ---- BCH (EN 302 307-1 5.3.1): g(x) = product of the minimal polynomials
 * of alpha^1, alpha^3, ..., alpha^(2t-1) over GF(2^m), m = 16 (normal,
 * x^16+x^5+x^3+x^2+1) or 14 (short, x^14+x^5+x^3+x+1) ---------------------- */

#define BCH_W_MAX 192 /* t = 12, m = 16 */
#define BCH_WORDS 3

typedef struct {
    int W; /* parity bits = degree of g */
    /* table[v] = (v(x) x^W) mod g(x), left-aligned in 192 bits: the
     * coefficient of x^(W-1) is bit 63 of word 0 */
    uint64_t table[256][BCH_WORDS];
} bch_enc;

static int bch_build(bch_enc* b, int short_frame, int t)
{
    const int m = short_frame ? 14 : 16, nq = (1 << m) - 1;
    const unsigned prim = short_frame ? 0x402Bu : 0x1002Du;
    uint16_t* ex = malloc(sizeof(uint16_t) * 2 * (size_t)nq);
    int* lg = malloc(sizeof(int) * ((size_t)nq + 1));
    if (!ex || !lg) {
        free(ex);
        free(lg);
        return -1;
    }
    unsigned x = 1;
    for (int i = 0; i < nq; i++) {
        ex[i] = ex[i + nq] = (uint16_t)x;
        lg[x] = i;
        x <<= 1;
        if (x >> m)
            x ^= prim;
    }
    /* g as GF(2) coefficients, index = degree */
    uint8_t g[BCH_W_MAX + 1] = { 1 }, tmp[BCH_W_MAX + 1];
    int dg = 0;
    for (int i = 1; i < 2 * t; i += 2) {
        /* minimal polynomial of alpha^i: prod (x + alpha^e), e over the
         * cyclotomic coset of i */
        uint16_t p[33] = { 1 };
        int deg = 0, e = i % nq;
        do {
            const uint16_t beta = ex[e];
            for (int k = deg + 1; k > 0; k--) {
                uint16_t prod = 0;
                if (p[k] && beta)
                    prod = ex[lg[p[k]] + lg[beta]];
                p[k] = p[k - 1] ^ prod;
            }
            p[0] = (p[0] && beta) ? ex[lg[p[0]] + lg[beta]] : 0;
            deg++;
            e = (2 * e) % nq;
        } while (e != i % nq && deg < 32);
        if (dg + deg > BCH_W_MAX) {
            free(ex);
            free(lg);
            return -1;
        }
        memset(tmp, 0, sizeof(tmp));
        for (int a = 0; a <= dg; a++)
            if (g[a])
                for (int c = 0; c <= deg; c++)
                    tmp[a + c] ^= (uint8_t)(p[c] & 1);
        dg += deg;
        memcpy(g, tmp, sizeof(g));
    }
    free(ex);
    free(lg);
    b->W = dg;
    /* g without x^W, left-aligned: x^k at bit position (W-1-k) from the top */
    uint64_t G[BCH_WORDS] = { 0 };
    for (int k = 0; k < dg; k++)
        if (g[k]) {
            const int from_top = dg - 1 - k;
            G[from_top / 64] |= 1ULL << (63 - from_top % 64);
        }
    for (int v = 0; v < 256; v++) {
        uint64_t r[BCH_WORDS] = { 0 };
        for (int bit = 7; bit >= 0; bit--) {
            const int fb = (int)(r[0] >> 63) ^ ((v >> bit) & 1);
            for (int w = 0; w < BCH_WORDS - 1; w++)
                r[w] = (r[w] << 1) | (r[w + 1] >> 63);
            r[BCH_WORDS - 1] <<= 1;
            if (fb)
                for (int w = 0; w < BCH_WORDS; w++)
                    r[w] ^= G[w];
        }
        memcpy(b->table[v], r, sizeof(r));
    }
    return 0;
}

/* systematic: cw = info (kbch bits) then the W parity bits, the remainder
 * of info(x) x^W by g(x), highest degree first */
static void bch_encode_bits(const bch_enc* b, const uint8_t* info, int kbch, uint8_t* cw)
{
    uint64_t r[BCH_WORDS] = { 0 };
    for (int i = 0; i < kbch; i += 8) {
        int byte = 0;
        for (int k = 0; k < 8; k++)
            byte = byte << 1 | (info[i + k] & 1);
        const uint64_t* tv = b->table[(r[0] >> 56) ^ (uint64_t)byte];
        for (int w = 0; w < BCH_WORDS - 1; w++)
            r[w] = (r[w] << 8) | (r[w + 1] >> 56);
        r[BCH_WORDS - 1] <<= 8;
        for (int w = 0; w < BCH_WORDS; w++)
            r[w] ^= tv[w];
    }
    memcpy(cw, info, (size_t)kbch);
    for (int k = 0; k < b->W; k++)
        cw[kbch + k] = (uint8_t)((r[k / 64] >> (63 - k % 64)) & 1);
}

/* ---- LDPC (IRA): p_c = sum of the info bits on check c, then the
 * accumulator p_c ^= p_(c-1) (EN 302 307-1 5.3.2) ------------------------ */

static void ldpc_encode_bits(const lg_code* c, const uint8_t* info, uint8_t* cw)
{
    uint8_t* p = cw + c->k;
    memcpy(cw, info, (size_t)c->k);
    memset(p, 0, (size_t)c->r);
    for (int i = 0; i < c->q; i++) {
        for (int e = c->layer_start[i]; e < c->layer_start[i + 1]; e++) {
            const int base = (int)(c->ent[e] >> 16), sh = (int)(c->ent[e] & 0xffff);
            /* node j of layer i is check q*j + i */
            for (int j = 0; j < LG_M; j++)
                p[c->q * j + i] ^= info[base + (j - sh + LG_M) % LG_M];
        }
    }
    for (int i = 1; i < c->r; i++)
        p[i] ^= p[i - 1];
}

/* ---- lookups ------------------------------------------------------------ */

int sim_fec_lookup(int short_frame, const char* rate, int* nldpc, int* kldpc, int* kbch, int* t)
{
    int kb, nb, tt;
    if (lg_fec_lookup(short_frame, rate, &kb, &nb, &tt))
        return -1;
    if (nldpc)
        *nldpc = short_frame ? 16200 : 64800;
    if (kldpc)
        *kldpc = nb;
    if (kbch)
        *kbch = kb;
    if (t)
        *t = tt;
    return 0;
}

static const struct lg_modcod_row* modcod_row(int pls) { return lg_modcod_get(pls); }

int sim_pls_lookup(int pls, char* rate, int rate_len, int* short_frame, int* bits, char* constellation,
                   int const_len)
{
    const struct lg_modcod_row* r = modcod_row(pls);
    if (!r)
        return -1;
    if (rate && rate_len > 0) {
        strncpy(rate, r->rate, (size_t)rate_len - 1);
        rate[rate_len - 1] = 0;
    }
    if (constellation && const_len > 0) {
        strncpy(constellation, r->constellation, (size_t)const_len - 1);
        constellation[const_len - 1] = 0;
    }
    if (short_frame)
        *short_frame = r->short_frame;
    if (bits)
        *bits = r->bits;
    return 0;
}

/* EN 302 307-1 Table 13: Es/N0 (dB) for QEF on AWGN, by DVB-S2 MODCOD
 * number (PLS code >> 2). */
static const double QEF_S2_DB[29] = {
    0,     -2.35, -1.24, -0.30, 1.00,  2.23,  3.10,  4.03,  4.68,  5.18,
    6.20,  6.42,  5.50,  6.62,  7.91,  9.35,  10.69, 10.98, 8.97,  10.21,
    11.03, 11.61, 12.89, 13.13, 12.73, 13.64, 14.28, 15.69, 16.05,
};

double sim_qef_db(int pls)
{
    const struct lg_modcod_row* r = modcod_row(pls);
    if (!r)
        return NAN;
    if ((pls & 0xFE) < 0x80)
        return QEF_S2_DB[(pls & 0x7F) >> 2];
    int num = 1, den = 1;
    const char* s = r->rate;
    num = atoi(s);
    const char* u = strchr(s, '_');
    den = u ? atoi(u + 1) : 1;
    const double eta = r->bits * (double)num / (den ? den : 1);
    return 10 * log10(pow(2, eta) - 1) + 1;
}

/* ---- per-code cache (BCH generator tables) ----------------------------- */

#define MAX_FEC 64
static bch_enc* g_bch[2][MAX_FEC];
static pthread_mutex_t g_bch_lock = PTHREAD_MUTEX_INITIALIZER;

static int fec_index(int short_frame, const char* rate)
{
    const int idx = lg_fec_index(short_frame, rate);
    return idx < MAX_FEC ? idx : -1;
}

static const bch_enc* get_bch(int short_frame, const char* rate, int t)
{
    const int idx = fec_index(short_frame, rate);
    if (idx < 0)
        return NULL;
    pthread_mutex_lock(&g_bch_lock);
    bch_enc* b = g_bch[short_frame != 0][idx];
    if (!b) {
        b = malloc(sizeof(*b));
        if (b && bch_build(b, short_frame, t)) {
            free(b);
            b = NULL;
        }
        g_bch[short_frame != 0][idx] = b;
    }
    pthread_mutex_unlock(&g_bch_lock);
    return b;
}

int sim_bch_encode(int short_frame, const char* rate, const uint8_t* info, uint8_t* cw)
{
    int kb, nb, t;
    if (lg_fec_lookup(short_frame, rate, &kb, &nb, &t))
        return -1;
    const bch_enc* b = get_bch(short_frame, rate, t);
    if (!b || kb + b->W != nb)
        return -1;
    bch_encode_bits(b, info, kb, cw);
    return 0;
}

int sim_ldpc_encode(int short_frame, const char* rate, const uint8_t* info, uint8_t* cw)
{
    const lg_code* c = lg_code_get(lg_code_lookup(short_frame, rate));
    if (!c)
        return -1;
    ldpc_encode_bits(c, info, cw);
    return 0;
}

/* ---- channel + demapper -------------------------------------------------- */

static inline int8_t quantise(double v)
{
    v = v > 127.0 ? 127.0 : v < -127.0 ? -127.0 : v;
    return (int8_t)lrint(v);
}

typedef struct {
    const sim_frame_spec* spec;
    int nframes, nthreads, tid;
    uint64_t seed;
    int8_t* llr;
    uint8_t* ref;
    int err;
} job_t;

static int gen_one(const sim_frame_spec* sp, uint64_t seed, int8_t* llr_base, uint8_t* ref_base,
                   uint8_t* info, uint8_t* bch, uint8_t* cw)
{
    const struct lg_modcod_row* mc = NULL;
    int short_frame = sp->short_frame;
    const char* rate = sp->rate;
    if (sp->pls >= 0) {
        mc = modcod_row(sp->pls);
        if (!mc)
            return -1;
        short_frame = mc->short_frame;
        rate = mc->rate;
    }
    int kb, nb, t;
    if (lg_fec_lookup(short_frame, rate, &kb, &nb, &t))
        return -1;
    const lg_code* code = lg_code_get(lg_code_lookup(short_frame, rate));
    const bch_enc* be = get_bch(short_frame, rate, t);
    if (!code || !be || kb + be->W != nb || nb > code->k)
        return -1;
    const int N = code->n;

    rng_t r;
    rng_seed(&r, seed, sp->stream);
    for (int i = 0; i < kb; i += 64) {
        uint64_t w = rng_next(&r);
        for (int k = 0; k < 64 && i + k < kb; k++)
            info[i + k] = (uint8_t)((w >> (63 - k)) & 1);
    }
    bch_encode_bits(be, info, kb, bch);
    /* codes shortened below K (VLSNR): the missing info bits are zero */
    memset(bch + nb, 0, (size_t)(code->k - nb));
    ldpc_encode_bits(code, bch, cw);
    if (ref_base) {
        uint8_t* o = ref_base + sp->ref_off;
        memset(o, 0, (size_t)(nb + 7) / 8);
        for (int i = 0; i < nb; i++)
            if (bch[i])
                o[i / 8] |= (uint8_t)(0x80 >> (i % 8));
    }
    int8_t* llr = llr_base + sp->llr_off;
    const int noiseless = sp->esn0_db >= 300.0;
    const double esn0 = pow(10.0, sp->esn0_db / 10.0);

    if (!mc) {
        /* tests/test_ldpc.c:64-73 channel */
        const double sigma2 = 1.0 / esn0, sig = sqrt(sigma2);
        for (int i = 0; i < N; i++) {
            const double y = (cw[i] ? -1.0 : 1.0) + (noiseless ? 0.0 : sig * rng_gauss(&r));
            llr[i] = quantise(noiseless ? (cw[i] ? -127.0 : 127.0) : 4.0 * y / sigma2);
        }
        return 0;
    }

    /* MODCOD: bit b of symbol j is codeword bit order[b] * nsym + j
     * (demap.c:225-233), no interleaver when order[0] < 0 (2 bits a symbol
     * in codeword order); points normalised to unit average energy */
    const int nb_ = mc->bits, np = 1 << nb_;
    const int nsym = (N + nb_ - 1) / nb_;
    double pr[256], pi[256];
    double e = 0;
    for (int p = 0; p < np; p++) {
        pr[p] = lg_points()[2 * (mc->point_off + p)];
        pi[p] = lg_points()[2 * (mc->point_off + p) + 1];
        e += pr[p] * pr[p] + pi[p] * pi[p];
    }
    const double gn = 1.0 / sqrt(e / np);
    for (int p = 0; p < np; p++) {
        pr[p] *= gn;
        pi[p] *= gn;
    }
    const double n0 = 1.0 / esn0, sn = sqrt(n0 / 2), k = 2.0 / n0;
    int posb[8];
    double t_[256], m0[8], m1[8];
    for (int j = 0; j < nsym; j++) {
        int label = 0;
        for (int b = 0; b < nb_; b++) {
            posb[b] = mc->order[0] < 0 ? nb_ * j + b : mc->order[b] * nsym + j;
            const int bit = posb[b] < N ? cw[posb[b]] : 0;
            label = label << 1 | bit;
        }
        const double x = pr[label] + (noiseless ? 0.0 : sn * rng_gauss(&r));
        const double y = pi[label] + (noiseless ? 0.0 : sn * rng_gauss(&r));
        for (int p = 0; p < np; p++) {
            const double dx = x - pr[p], dy = y - pi[p];
            t_[p] = dx * dx + dy * dy;
        }
        /* per-bit minima, least significant label bit first: even/odd
         * entries split the bit, then pairs merge for the next one */
        for (int b = nb_ - 1, cur = np; b >= 0; b--, cur /= 2) {
            double lo = t_[0], hi = t_[1];
            for (int q = 2; q < cur; q += 2) {
                lo = t_[q] < lo ? t_[q] : lo;
                hi = t_[q + 1] < hi ? t_[q + 1] : hi;
            }
            m0[b] = lo;
            m1[b] = hi;
            for (int q = 0; q < cur / 2; q++)
                t_[q] = t_[2 * q] < t_[2 * q + 1] ? t_[2 * q] : t_[2 * q + 1];
        }
        for (int b = 0; b < nb_; b++)
            if (posb[b] < N)
                llr[posb[b]] = quantise(k * (m1[b] - m0[b]));
    }
    return 0;
}

static void* gen_thread(void* arg)
{
    job_t* jb = arg;
    uint8_t* info = malloc(64800);
    uint8_t* bch = malloc(64800);
    uint8_t* cw = malloc(64800);
    if (!info || !bch || !cw) {
        jb->err = -1;
    } else {
        for (int f = jb->tid; f < jb->nframes; f += jb->nthreads)
            if (gen_one(&jb->spec[f], jb->seed, jb->llr, jb->ref, info, bch, cw))
                jb->err = -1;
    }
    free(info);
    free(bch);
    free(cw);
    return NULL;
}

int sim_generate(const sim_frame_spec* spec, int nframes, uint64_t seed, int8_t* llr, uint8_t* ref,
                 int nthreads)
{
    if (nframes <= 0)
        return 0;
    if (nthreads < 1)
        nthreads = 1;
    if (nthreads > 64)
        nthreads = 64;
    if (nthreads > nframes)
        nthreads = nframes;
    job_t jobs[64];
    pthread_t th[64];
    for (int i = 0; i < nthreads; i++) {
        jobs[i] = (job_t){ spec, nframes, nthreads, i, seed, llr, ref, 0 };
        if (i)
            pthread_create(&th[i], NULL, gen_thread, &jobs[i]);
    }
    gen_thread(&jobs[0]);
    int err = jobs[0].err;
    for (int i = 1; i < nthreads; i++) {
        pthread_join(th[i], NULL);
        err |= jobs[i].err;
    }
    return err;
}
