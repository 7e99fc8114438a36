/*
 * CPU model of the HIP decoder's schedule (test infrastructure, never part
 * of the product): the same code registry (circulant layers, dependency
 * levels), the same node arithmetic (dvb-rx_amd/csrc/ldpc_node.h) and the
 * same posterior / message layout as ldpc_layered_kernel, run on one core.
 * The nodes of one level are processed in decreasing j: the kernel runs
 * them concurrently, so the result must not depend on their order; any two
 * nodes wrongly put in the same level change the result against the
 * oracle, which processes them in increasing j.
 */
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "ldpc_codes.h"
#include "ldpc_node.h"

/* Decode one frame with per-frame early stop.  Returns iterations run;
 * *ok = syndrome holds.  out: ceil(nbits/8) bytes, MSB first. */
int kmodel_decode(int code, const int8_t* llr, int max_iter, int nbits, uint8_t* out, int* ok)
{
    const lg_code* c = lg_code_get(code);
    if (!c)
        return -1;
    const int K = c->k, R = c->r, q = c->q;
    int8_t* post = malloc((size_t)c->n);
    uint32_t* w0 = calloc((size_t)R, sizeof(uint32_t));
    uint32_t* sg = calloc((size_t)R, sizeof(uint32_t));
    if (!post || !w0 || !sg) {
        free(post);
        free(w0);
        free(sg);
        return -1;
    }
    memcpy(post, llr, (size_t)K);
    for (int t = 0; t < R; t++)
        post[K + 360 * (t % q) + t / q] = llr[K + t];
    int it = 0, good;
    for (;;) {
        good = 1;
        for (int i = 0; i < q && good; i++) {
            const int st = c->layer_start[i], deg = c->layer_start[i + 1] - st;
            for (int j = 0; j < 360; j++)
                good &= lg_node_check(post, c->ent + st, deg, i, j, K, q);
        }
        if (good || it >= max_iter)
            break;
        for (int i = 0; i < q; i++) {
            const int st = c->layer_start[i], deg = c->layer_start[i + 1] - st;
            for (int l = 0; l < c->nlev[i]; l++)
                for (int j = 359; j >= 0; j--)
                    if (c->level[360 * i + j] == l)
                        lg_node_update(post, c->ent + st, deg, i, j, K, q, &w0[360 * i + j], &sg[360 * i + j],
                                       it == 0);
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
    free(w0);
    free(sg);
    return it;
}

/* The schedule's own invariants: no two nodes of one level of a layer
 * share a variable.  Returns the number of violations. */
int kmodel_check_levels(int code)
{
    const lg_code* c = lg_code_get(code);
    if (!c)
        return -1;
    int bad = 0;
    int* seen = malloc(sizeof(int) * (size_t)c->n);
    if (!seen)
        return -1;
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
    }
    free(seen);
    return bad;
}
