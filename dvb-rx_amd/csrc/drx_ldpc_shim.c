/*
 * libdvbrx's drx_ldpc_* API (libdvbrx/include/dvbrx/ldpc.h) on the MI355X
 * decoder: a handle is a decoder context (own stream and scratch) plus
 * staging for one batch of up to 64 frames of one code.  The frames of a
 * call form one group, so the call has the reference's batch semantics
 * (libdvbrx/src/ldpc.c:614-638): one iteration count for the batch, -1 if
 * any frame fails its checks at the end.
 */
#include "dvbrx/ldpc.h"

#include <stdlib.h>
#include <string.h>

#include "ldpc_codes.h"
#include "ldpc_gpu.h"

#define SHIM_LANES 64

struct drx_ldpc {
    ldpc_gpu_ctx* ctx;
    int code, n, kldpc, nbytes;
    int8_t* h_llr;
    uint8_t* h_out;
    int8_t* d_llr;
    uint8_t* d_out;
    int32_t* d_it;
    uint8_t* d_synd;
};

int drx_ldpc_lanes(void) { return SHIM_LANES; }

drx_ldpc* drx_ldpc_new(const drx_fec_params* p, const char* rate)
{
    if (!p || !rate)
        return NULL;
    const int code = ldpc_gpu_code_lookup(p->short_frame, rate);
    int n = 0, k = 0;
    if (code < 0 || ldpc_gpu_code_params(code, &n, &k, NULL, NULL, NULL))
        return NULL;
    drx_ldpc* d = calloc(1, sizeof(*d));
    if (!d)
        return NULL;
    d->code = code;
    d->n = n;
    /* kldpc == K except for the shortened VLSNR codes (kldpc < K) */
    d->kldpc = p->kldpc > 0 && p->kldpc <= k ? p->kldpc : k;
    d->nbytes = (d->kldpc + 7) / 8;
    const char* dev = getenv("LDPC_GPU_DEVICE");
    if (ldpc_gpu_create(dev ? atoi(dev) : 0, &d->ctx)) {
        free(d);
        return NULL;
    }
    d->h_llr = malloc((size_t)SHIM_LANES * n);
    d->h_out = malloc((size_t)SHIM_LANES * d->nbytes);
    d->d_llr = ldpc_gpu_malloc(d->ctx, (size_t)SHIM_LANES * n);
    d->d_out = ldpc_gpu_malloc(d->ctx, (size_t)SHIM_LANES * d->nbytes);
    d->d_it = ldpc_gpu_malloc(d->ctx, sizeof(int32_t) * SHIM_LANES);
    d->d_synd = ldpc_gpu_malloc(d->ctx, SHIM_LANES);
    if (!d->h_llr || !d->h_out || !d->d_llr || !d->d_out || !d->d_it || !d->d_synd) {
        drx_ldpc_free(d);
        return NULL;
    }
    return d;
}

void drx_ldpc_free(drx_ldpc* d)
{
    if (!d)
        return;
    if (d->ctx) {
        ldpc_gpu_free(d->ctx, d->d_llr);
        ldpc_gpu_free(d->ctx, d->d_out);
        ldpc_gpu_free(d->ctx, d->d_it);
        ldpc_gpu_free(d->ctx, d->d_synd);
        ldpc_gpu_destroy(d->ctx);
    }
    free(d->h_llr);
    free(d->h_out);
    free(d);
}

int drx_ldpc_decode(drx_ldpc* d, const int8_t* const* llr, uint8_t* const* out, int nframes, int max_iter)
{
    if (!d || nframes > SHIM_LANES)
        return -1;
    /* no frame: every (padding) lane satisfies its checks at once */
    if (nframes <= 0)
        return 0;
    if (max_iter < 0)
        max_iter = 0;
    ldpc_gpu_frame fr[SHIM_LANES];
    for (int f = 0; f < nframes; f++) {
        memcpy(d->h_llr + (size_t)f * d->n, llr[f], (size_t)d->n);
        fr[f].code = d->code;
        fr[f].max_iter = max_iter;
        fr[f].nbits = d->kldpc;
        fr[f].group = 0;
        fr[f].llr_off = (uint64_t)f * d->n;
        fr[f].out_off = (uint64_t)f * d->nbytes;
    }
    int32_t it[SHIM_LANES];
    uint8_t ok[SHIM_LANES];
    if (ldpc_gpu_upload(d->ctx, d->d_llr, d->h_llr, (size_t)nframes * d->n) ||
        ldpc_gpu_decode_batch(d->ctx, fr, nframes, d->d_llr, d->d_out, d->d_it, d->d_synd) ||
        ldpc_gpu_download(d->ctx, d->h_out, d->d_out, (size_t)nframes * d->nbytes) ||
        ldpc_gpu_download(d->ctx, it, d->d_it, sizeof(int32_t) * nframes) ||
        ldpc_gpu_download(d->ctx, ok, d->d_synd, (size_t)nframes))
        return -1;
    int all = 1;
    for (int f = 0; f < nframes; f++) {
        memcpy(out[f], d->h_out + (size_t)f * d->nbytes, (size_t)d->nbytes);
        all &= ok[f];
    }
    return all ? it[0] : -1;
}

/* ---- drx_fec_lookup and drx_bch_* (libdvbrx/include/dvbrx/fec.h) ---------
 *
 * For callers built without libdvbrx's fec.h and bch.c.  Weak, so that
 * libdvbrx's own definitions win where both are linked. */

__attribute__((weak)) int drx_fec_lookup(int short_frame, const char* rate, drx_fec_params* p)
{
    int nldpc, kldpc, kbch, t;
    if (!rate || !p || ldpc_gpu_fec_lookup(short_frame, rate, &nldpc, &kldpc, &kbch, &t))
        return -1;
    p->short_frame = short_frame;
    p->nldpc = nldpc;
    p->kldpc = kldpc;
    p->kbch = kbch;
    p->t = t;
    return 0;
}

struct drx_bch {
    ldpc_gpu_ctx* ctx;
    drx_fec_params p;
    char rate[24];
    uint8_t* d_cw;
    int32_t* d_res;
};

__attribute__((weak)) drx_bch* drx_bch_new(const drx_fec_params* p)
{
    if (!p)
        return NULL;
    drx_bch* b = calloc(1, sizeof(*b));
    if (!b)
        return NULL;
    b->p = *p;
    const char* dev = getenv("LDPC_GPU_DEVICE");
    /* the BCH code of a parameter set (drx_bch_new takes no rate): the first
     * rate with its frame size, kbch, nbch and t */
    const char* rate = lg_fec_rate(p->short_frame, p->kbch, p->kldpc, p->t);
    if (!rate || strlen(rate) >= sizeof(b->rate) || ldpc_gpu_create(dev ? atoi(dev) : 0, &b->ctx)) {
        free(b);
        return NULL;
    }
    strcpy(b->rate, rate);
    b->d_cw = ldpc_gpu_malloc(b->ctx, (size_t)(p->kldpc + 7) / 8);
    b->d_res = ldpc_gpu_malloc(b->ctx, sizeof(int32_t));
    if (!b->d_cw || !b->d_res) {
        drx_bch_free(b);
        return NULL;
    }
    return b;
}

__attribute__((weak)) void drx_bch_free(drx_bch* b)
{
    if (!b)
        return;
    if (b->ctx) {
        ldpc_gpu_free(b->ctx, b->d_cw);
        ldpc_gpu_free(b->ctx, b->d_res);
        ldpc_gpu_destroy(b->ctx);
    }
    free(b);
}

__attribute__((weak)) int drx_bch_decode(drx_bch* b, uint8_t* cw)
{
    if (!b || !cw)
        return -1;
    const size_t nb = (size_t)(b->p.kldpc + 7) / 8;
    int32_t res = -1;
    if (ldpc_gpu_upload(b->ctx, b->d_cw, cw, nb) ||
        ldpc_gpu_bch_decode(b->ctx, b->p.short_frame, b->rate, 1, b->d_cw, nb, b->d_res) ||
        ldpc_gpu_download(b->ctx, &res, b->d_res, sizeof(res)))
        return -1;
    if (res > 0 && ldpc_gpu_download(b->ctx, cw, b->d_cw, nb))
        return -1;
    return res;
}
