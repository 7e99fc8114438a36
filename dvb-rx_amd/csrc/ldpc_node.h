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

#ifdef __cplusplus
/*
 * The same two operations with the information degree bounded at compile
 * time (DI >= deg): every loop is unrolled, all posterior loads of a node
 * are issued before the first use and the extrinsic inputs stay in
 * registers between the two passes.  Edge numbering (and so the message
 * words) as above: information edges 0..deg-1, own parity deg, previous
 * parity deg+1.
 */
template <int DI>
LG_HD void lg_node_update_t(int8_t* post, const uint32_t* ent, int deg, int i, int j, int K, int q,
                            uint32_t* w0p, uint32_t* sgp, int first)
{
    int ad[DI], x[DI];
#pragma unroll
    for (int e = 0; e < DI; e++) {
        if (e < deg) {
            const uint32_t en = ent[e];
            const int d = j - (int)(en & 0xffff);
            ad[e] = (int)(en >> 16) + (d < 0 ? d + 360 : d);
        }
    }
    const int apc = K + 360 * i + j;
    const int has_pp = (i | j) != 0;
    const int app = i ? apc - 360 : K + 360 * (q - 1) + j - 1;
    int Lv[DI];
#pragma unroll
    for (int e = 0; e < DI; e++)
        if (e < deg)
            Lv[e] = post[ad[e]];
    const int Lpc = post[apc];
    const int Lpp = has_pp ? post[app] : 0;

    const uint32_t w0 = first ? 0u : *w0p, sg = first ? 0u : *sgp;
    int min0 = 255, min1 = 255, s = 0, ix = 0;
#define LG_IN(e, L, xe)                                                       \
    do {                                                                      \
        const int m_ = first ? 0 : lg_msg_edge(w0, sg, (e));                   \
        xe = lg_sat8((L) - m_);                                               \
        const int a_ = xe < 0 ? -xe : xe;                                     \
        min1 = lg_min(min1, lg_max(min0, a_));                                \
        ix = a_ < min0 ? (e) : ix;                                            \
        min0 = lg_min(min0, a_);                                              \
        s ^= xe;                                                              \
    } while (0)
#pragma unroll
    for (int e = 0; e < DI; e++)
        if (e < deg)
            LG_IN(e, Lv[e], x[e]);
    int xpc, xpp = 0;
    LG_IN(deg, Lpc, xpc);
    if (has_pp)
        LG_IN(deg + 1, Lpp, xpp);
#undef LG_IN
    const int mn0 = lg_max(lg_min(min0, 127) - 1, 0), mn1 = lg_max(lg_min(min1, 127) - 1, 0);
    uint32_t nsg = 0;
#define LG_OUT(e, xe, addr)                                                   \
    do {                                                                      \
        const int neg_ = (s ^ xe) < 0;                                        \
        const int a_ = xe < 0 ? -xe : xe;                                     \
        const int mag_ = a_ == min0 ? mn1 : mn0;                              \
        post[addr] = (int8_t)lg_sat8(xe + (neg_ ? -mag_ : mag_));             \
        nsg |= (uint32_t)neg_ << (e);                                         \
    } while (0)
#pragma unroll
    for (int e = 0; e < DI; e++)
        if (e < deg)
            LG_OUT(e, x[e], ad[e]);
    LG_OUT(deg, xpc, apc);
    if (has_pp)
        LG_OUT(deg + 1, xpp, app);
#undef LG_OUT
    *w0p = (uint32_t)mn0 | (uint32_t)mn1 << 8 | (uint32_t)ix << 16;
    *sgp = nsg;
}

template <int DI>
LG_HD int lg_node_check_t(const int8_t* post, const uint32_t* ent, int deg, int i, int j, int K, int q)
{
    int Lv[DI];
#pragma unroll
    for (int e = 0; e < DI; e++) {
        if (e < deg) {
            const uint32_t en = ent[e];
            const int d = j - (int)(en & 0xffff);
            Lv[e] = post[(int)(en >> 16) + (d < 0 ? d + 360 : d)];
        }
    }
    const int apc = K + 360 * i + j;
    int par = post[apc];
    int zero = par == 0;
    if (i | j) {
        const int pp = post[i ? apc - 360 : K + 360 * (q - 1) + j - 1];
        par ^= pp;
        zero |= pp == 0;
    }
#pragma unroll
    for (int e = 0; e < DI; e++) {
        if (e < deg) {
            par ^= Lv[e];
            zero |= Lv[e] == 0;
        }
    }
    return !zero && par >= 0;
}

/*
 * Layer processing in three phases, exact for any dependency structure
 * inside a layer (the reference visits the 360 nodes of a layer in order,
 * ldpc.c:474-480, updating posteriors in place):
 *
 *   A  every node reads the edges no earlier node of the layer touches
 *      (their value is the pre-layer one), forms x_e = sat(L_e - m_e) and
 *      folds them into its minimums and sign; a node with no earlier
 *      neighbour (level 0) is then complete and writes its outputs on the
 *      bits a later node will read;
 *   S  for level l = 1, 2, ...: the nodes of level l read their dependent
 *      bits (written by earlier levels), fold them in and write their
 *      outputs on the bits a later node reads;
 *   C  every node writes the posteriors of the bits it is the last to touch
 *      and its new messages.
 *
 * Messages: one int8 per edge, clamped to [-32, 31] (ldpc.c:343), 4 a word:
 * edge e in byte e % 4 of word e / 4 (info edges 0..DEG-1, own parity DEG,
 * previous parity DEG+1).
 */
template <int DEG>
struct LgNode {
    int ad[DEG];    /* posterior index of each information edge */
    int x[DEG + 2]; /* extrinsic inputs: information edges, own parity, previous parity */
    int apc, app, has_pp;
    int min0, min1, s;
};

#define LG_MSG_WORDS(deg) (((deg) + 2 + 3) / 4)

LG_HD int lg_msg_byte(const uint32_t* w, int e) { return (int)(int8_t)(uint8_t)(w[e >> 2] >> (8 * (e & 3))); }

template <int DEG>
LG_HD void lg_fold(LgNode<DEG>& n, int xe)
{
    const int a = xe < 0 ? -xe : xe;
    n.min1 = lg_min(n.min1, lg_max(n.min0, a));
    n.min0 = lg_min(n.min0, a);
    n.s ^= xe;
}

template <int DEG>
LG_HD void lg_phase_a(LgNode<DEG>& n, const int8_t* post, const uint32_t* ent, int i, int j, int K, int q,
                      uint32_t dep, const uint32_t* mw, int first)
{
#pragma unroll
    for (int e = 0; e < DEG; e++) {
        const uint32_t en = ent[e];
        const int d = j - (int)(en & 0xffff);
        n.ad[e] = (int)(en >> 16) + (d < 0 ? d + 360 : d);
    }
    n.apc = K + 360 * i + j;
    n.has_pp = (i | j) != 0;
    n.app = i ? n.apc - 360 : K + 360 * (q - 1) + j - 1;
    n.min0 = 255;
    n.min1 = 255;
    n.s = 0;
#pragma unroll
    for (int e = 0; e < DEG; e++) {
        if (!((dep >> e) & 1)) {
            const int L = post[n.ad[e]];
            n.x[e] = first ? L : lg_sat8(L - lg_msg_byte(mw, e));
            lg_fold(n, n.x[e]);
        }
    }
    {
        const int L = post[n.apc];
        n.x[DEG] = first ? L : lg_sat8(L - lg_msg_byte(mw, DEG));
        lg_fold(n, n.x[DEG]);
    }
    if (n.has_pp) {
        const int L = post[n.app];
        n.x[DEG + 1] = first ? L : lg_sat8(L - lg_msg_byte(mw, DEG + 1));
        lg_fold(n, n.x[DEG + 1]);
    } else {
        n.x[DEG + 1] = 0;
    }
}

/* fold in the dependent edges (their earlier neighbours have written them) */
template <int DEG>
LG_HD void lg_step(LgNode<DEG>& n, const int8_t* post, uint32_t dep, const uint32_t* mw, int first)
{
#pragma unroll
    for (int e = 0; e < DEG; e++) {
        if ((dep >> e) & 1) {
            const int L = post[n.ad[e]];
            n.x[e] = first ? L : lg_sat8(L - lg_msg_byte(mw, e));
            lg_fold(n, n.x[e]);
        }
    }
}

/* check-to-variable value o_e of edge e once all edges are folded in */
template <int DEG>
LG_HD int lg_out(const LgNode<DEG>& n, int e)
{
    const int mn0 = lg_max(lg_min(n.min0, 127) - 1, 0), mn1 = lg_max(lg_min(n.min1, 127) - 1, 0);
    const int xe = n.x[e];
    const int a = xe < 0 ? -xe : xe;
    const int mag = a == n.min0 ? mn1 : mn0;
    return (n.s ^ xe) < 0 ? -mag : mag;
}

/* outputs on the bits a later node of the layer reads */
template <int DEG>
LG_HD void lg_write_succ(const LgNode<DEG>& n, int8_t* post, uint32_t succ)
{
#pragma unroll
    for (int e = 0; e < DEG; e++)
        if ((succ >> e) & 1)
            post[n.ad[e]] = (int8_t)lg_sat8(n.x[e] + lg_out(n, e));
}

/* final posteriors (bits this node touches last) and the new messages */
template <int DEG>
LG_HD void lg_phase_c(const LgNode<DEG>& n, int8_t* post, uint32_t succ, uint32_t* mw_out)
{
    uint32_t w[LG_MSG_WORDS(DEG)];
#pragma unroll
    for (int k = 0; k < LG_MSG_WORDS(DEG); k++)
        w[k] = 0;
#pragma unroll
    for (int e = 0; e < DEG + 2; e++) {
        if (e == DEG + 1 && !n.has_pp)
            continue;
        const int o = lg_out(n, e);
        const int ad = e < DEG ? n.ad[e] : e == DEG ? n.apc : n.app;
        if (e >= DEG || !((succ >> e) & 1))
            post[ad] = (int8_t)lg_sat8(n.x[e] + o);
        w[e >> 2] |= (uint32_t)(uint8_t)(int8_t)lg_max(lg_min(o, 31), -32) << (8 * (e & 3));
    }
#pragma unroll
    for (int k = 0; k < LG_MSG_WORDS(DEG); k++)
        mw_out[k] = w[k];
}
#endif

#endif
