/*
 * Check-node arithmetic of the layered offset min-sum decoder, shared by
 * the HIP kernels and by the CPU model of the kernel schedule that the
 * tests compare with the oracle.  Bit-exact with libdvbrx/src/ldpc.c:321-362
 * (see ldpc_gpu.hip for the formulas).
 *
 * A node's messages are stored as two words: w0 = mn0 | mn1 << 8 | ix << 16
 * (the two offset magnitudes and one edge holding the smallest input) and
 * sg = the sign bits of its (up to 32) messages.
 */
#ifndef LDPC_NODE_H
#define LDPC_NODE_H

#include <stdint.h>

#if defined(__HIPCC__)
#define LG_HD __host__ __device__ __forceinline__
#else
#define LG_HD static inline
#endif

LG_HD int lg_min(int a, int b) { return a < b ? a : b; }
LG_HD int lg_max(int a, int b) { return a > b ? a : b; }
LG_HD int lg_sat8(int x) { return lg_min(lg_max(x, -128), 127); }

/* clamped message of edge e (ldpc.c:343 clamp to [-32, 31]) */
LG_HD int lg_msg_edge(uint32_t w0, uint32_t sg, int e)
{
    const int mn0 = (int)(w0 & 0xff), mn1 = (int)((w0 >> 8) & 0xff), ix = (int)((w0 >> 16) & 0xff);
    const int mag = e == ix ? mn1 : mn0;
    return ((sg >> e) & 1) ? lg_max(-mag, -32) : lg_min(mag, 31);
}

/* posterior address of edge e of node (i, j) in the frame's posterior array
 * (info bits 0..K-1, parity of check q*j+i at K + 360*i + j): the info
 * edges of the layer (base << 16 | shift), then the node's own parity, then
 * the previous one on the staircase (prev_parity, ldpc.c:277-284) */
LG_HD int lg_edge_addr(const uint32_t* ent, int deg, int e, int i, int j, int K, int q)
{
    if (e < deg) {
        const uint32_t en = ent[e];
        const int base = (int)(en >> 16), sh = (int)(en & 0xffff);
        const int d = j - sh;
        return base + (d < 0 ? d + 360 : d);
    }
    if (e == deg)
        return K + 360 * i + j;
    return i ? K + 360 * (i - 1) + j : K + 360 * (q - 1) + j - 1;
}

/* edges of node (i, j): node (0, 0) has no previous parity */
LG_HD int lg_node_degree(int deg, int i, int j) { return deg + 1 + ((i | j) != 0); }

/* Update node (i, j) in place: posteriors post[], message words (*w0, *sg)
 * read (unless first) and rewritten. */
LG_HD void lg_node_update(int8_t* post, const uint32_t* ent, int deg, int i, int j, int K, int q,
                          uint32_t* w0p, uint32_t* sgp, int first)
{
    const int dg = lg_node_degree(deg, i, j);
    const uint32_t w0 = first ? 0u : *w0p, sg = first ? 0u : *sgp;
    int min0 = 255, min1 = 255, s = 0, ix = 0;
    for (int e = 0; e < dg; e++) {
        const int L = post[lg_edge_addr(ent, deg, e, i, j, K, q)];
        const int x = first ? L : lg_sat8(L - lg_msg_edge(w0, sg, e));
        const int a = x < 0 ? -x : x;
        min1 = lg_min(min1, lg_max(min0, a));
        ix = a < min0 ? e : ix;
        min0 = lg_min(min0, a);
        s ^= x;
    }
    const int mn0 = lg_max(lg_min(min0, 127) - 1, 0), mn1 = lg_max(lg_min(min1, 127) - 1, 0);
    uint32_t nsg = 0;
    for (int e = 0; e < dg; e++) {
        const int ad = lg_edge_addr(ent, deg, e, i, j, K, q);
        const int L = post[ad];
        const int x = first ? L : lg_sat8(L - lg_msg_edge(w0, sg, e));
        const int neg = (s ^ x) < 0;
        const int a = x < 0 ? -x : x;
        const int mag = a == min0 ? mn1 : mn0;
        post[ad] = (int8_t)lg_sat8(x + (neg ? -mag : mag));
        nsg |= (uint32_t)neg << e;
    }
    *w0p = (uint32_t)mn0 | (uint32_t)mn1 << 8 | (uint32_t)ix << 16;
    *sgp = nsg;
}

/* parity check of node (i, j) holds: no zero posterior, even number of
 * negative ones (bad(), ldpc.c:286-305) */
LG_HD int lg_node_check(const int8_t* post, const uint32_t* ent, int deg, int i, int j, int K, int q)
{
    const int dg = lg_node_degree(deg, i, j);
    int par = 0, zero = 0;
    for (int e = 0; e < dg; e++) {
        const int x = post[lg_edge_addr(ent, deg, e, i, j, K, q)];
        par ^= x;
        zero |= x == 0;
    }
    return !zero && par >= 0;
}

#endif
