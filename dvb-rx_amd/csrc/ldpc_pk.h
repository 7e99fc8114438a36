/*
 * Two frames per instruction: the check-node arithmetic of ldpc_node.h on
 * packed 16-bit lanes (gfx950 v_pk_* instructions), shared by the HIP
 * kernel and the CPU model of it (built with clang, same code).
 *
 * Representation.  A workgroup holds two frames ("slots") of the same code.
 * The LDS posterior of variable v is two bytes: slot 0 at 2v, slot 1 at
 * 2v + 1 (one ds_read_u16 serves both).  In registers an int8 value y of
 * slot k sits in 16-bit lane k as y * 256: the bytes [b0, b1] of a
 * posterior become [0, b0, 0, b1] with one v_perm_b32.  At that scale the
 * int16 saturation of v_pk_add/sub_i16 (clamp) is exactly the int8
 * saturation of the reference (qadd/qsub, libdvbrx/src/ldpc.c:128-129):
 * -128 * 256 = -32768, and +127 saturates to 32767 whose high byte is 127;
 * only the high byte of a lane is ever stored.  The low byte is zero
 * except after a positive saturation (0xff); such values only occur at
 * |y| = 127 where they cannot change a result (see pk_out).
 *
 * Edges of a node by position: 0 its own parity bit, 1 the previous parity
 * bit (the IRA staircase, ldpc.c:277-284), 2 + e information edge e of the
 * layer.  Messages: one int8 per edge per slot, two positions a word
 * [p0 slot0, p0 slot1, p1 slot0, p1 slot1], position p in word p / 2.
 *
 * The node routines are templates on DM, the largest information degree
 * a kernel instance handles; a layer's own degree D <= DM is a run-time
 * value, the same for all nodes of the layer (edges e >= D are skipped by
 * uniform branches, so they cost no vector instructions).
 */
#ifndef LDPC_PK_H
#define LDPC_PK_H

#include <stdint.h>

#include "ldpc_codes.h"
#include "ldpc_node.h"

typedef short lg_pk __attribute__((ext_vector_type(2)));
typedef unsigned short lg_pku __attribute__((ext_vector_type(2)));

#if defined(__HIPCC__)
#define LG_PK_HD __host__ __device__ __forceinline__
#else
#define LG_PK_HD static inline
#endif

LG_PK_HD lg_pk pk_of(uint32_t v) { return __builtin_bit_cast(lg_pk, v); }
LG_PK_HD uint32_t pk_bits(lg_pk v) { return __builtin_bit_cast(uint32_t, v); }
LG_PK_HD lg_pku pku(lg_pk v) { return __builtin_bit_cast(lg_pku, v); }
LG_PK_HD lg_pk pks(lg_pku v) { return __builtin_bit_cast(lg_pk, v); }
LG_PK_HD lg_pk pk_dup(int v) { return lg_pk{ (short)v, (short)v }; }

/* saturating int16 add / sub (v_pk_add_i16 / v_pk_sub_i16 with clamp) */
LG_PK_HD lg_pk pk_adds(lg_pk a, lg_pk b) { return __builtin_elementwise_add_sat(a, b); }
LG_PK_HD lg_pk pk_subs(lg_pk a, lg_pk b) { return __builtin_elementwise_sub_sat(a, b); }
LG_PK_HD lg_pk pk_max(lg_pk a, lg_pk b) { return __builtin_elementwise_max(a, b); }
LG_PK_HD lg_pk pk_min(lg_pk a, lg_pk b) { return __builtin_elementwise_min(a, b); }
LG_PK_HD lg_pk pk_minu(lg_pk a, lg_pk b) { return pks(__builtin_elementwise_min(pku(a), pku(b))); }
LG_PK_HD lg_pk pk_maxu(lg_pk a, lg_pk b) { return pks(__builtin_elementwise_max(pku(a), pku(b))); }

/* v_perm_b32: byte k of the result is byte sel[k] of {s0 (bytes 4-7), s1
 * (bytes 0-3)}; selector 12 gives 0 (the only special value used here) */
LG_PK_HD uint32_t lg_perm(uint32_t s0, uint32_t s1, uint32_t sel)
{
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_perm(s0, s1, sel);
#else
    const uint64_t v = (uint64_t)s0 << 32 | s1;
    uint32_t r = 0;
    for (int k = 0; k < 4; k++) {
        const uint32_t b = (sel >> (8 * k)) & 0xff;
        const uint32_t byte = b < 8 ? (uint32_t)(v >> (8 * b)) & 0xff : b == 12 ? 0u : 0xffu;
        r |= byte << (8 * k);
    }
    return r;
#endif
}

#define PK_UNPACK_LO 0x010C000Cu /* [b0, b1] -> [0, b0, 0, b1] */
#define PK_UNPACK_HI 0x030C020Cu /* [_, _, b2, b3] -> [0, b2, 0, b3] */
#define PK_PACK_POST 0x0C0C0301u /* lanes -> [hi0, hi1, 0, 0] */
#define PK_PACK_MSG 0x07050301u  /* (odd edge, even edge) -> [e0s0, e0s1, e1s0, e1s1] */

/* message unpack selectors of the even and odd edge of a word; a slot in
 * its first iteration reads zero messages (ldpc.c:491-497) */
LG_PK_HD uint32_t pk_msg_sel(int odd, int first0, int first1)
{
    const uint32_t b0 = first0 ? 0x0Cu : odd ? 0x02u : 0x00u;
    const uint32_t b1 = first1 ? 0x0Cu : odd ? 0x03u : 0x01u;
    return b1 << 24 | 0x0Cu << 16 | b0 << 8 | 0x0Cu;
}

#define LG_PK_WORDS(deg) (((deg) + 2 + 1) / 2)

template <int DM>
struct PkNode {
    int ad[DM];       /* posterior index (variable) of each information edge */
    lg_pk x[DM + 2];  /* extrinsic inputs by position, both slots */
    lg_pk a[DM + 2];  /* their magnitudes */
    int apc, app;
    uint32_t nopp;    /* all ones for node (0, 0), which has no previous parity */
    lg_pk min0, min1, s;
};

/* posterior index of node j's edge on table entry en (base << 16 | shift
 * in the low 9 bits): base + (j - shift) mod 360, the wrap as an unsigned
 * minimum */
LG_PK_HD int pk_addr(uint32_t en, int j)
{
    const uint32_t t = (uint32_t)j - (en & 0x1ffu);
    const uint32_t u = t + 360u;
    return (int)(en >> 16) + (int)(t < u ? t : u);
}

LG_PK_HD lg_pk pk_abs(lg_pk x) { return pk_max(x, pk_subs(pk_dup(0), x)); }

template <int DM>
LG_PK_HD void pk_fold(PkNode<DM>& n, int p, lg_pk x)
{
    const lg_pk a = pk_abs(x);
    n.x[p] = x;
    n.a[p] = a;
    n.min1 = pk_minu(n.min1, pk_maxu(n.min0, a));
    n.min0 = pk_minu(n.min0, a);
    n.s = n.s ^ x;
}

/* pk_fold of the lanes where skip is 0; lanes where it is all ones leave
 * the minimums and the sign as they were (|x| | 0xffff is above any
 * minimum).  Branch-free, so the posterior reads of all edges can be in
 * flight together. */
template <int DM>
LG_PK_HD void pk_fold_skip(PkNode<DM>& n, int p, lg_pk x, uint32_t skip)
{
    const lg_pk a = pk_of(pk_bits(pk_abs(x)) | skip);
    n.x[p] = x;
    n.a[p] = a;
    n.min1 = pk_minu(n.min1, pk_maxu(n.min0, a));
    n.min0 = pk_minu(n.min0, a);
    n.s = pk_of(pk_bits(n.s) ^ (pk_bits(x) & ~skip));
}

/* x = sat(L - m) of position p from the two posterior bytes w */
LG_PK_HD lg_pk pk_in_w(uint32_t w, const uint32_t* mw, int p, uint32_t sel_even, uint32_t sel_odd)
{
    const lg_pk L = pk_of(lg_perm(0, w, PK_UNPACK_LO));
    const lg_pk M = pk_of(lg_perm(0, mw[p >> 1], (p & 1) ? sel_odd : sel_even));
    return pk_subs(L, M);
}

LG_PK_HD lg_pk pk_in(const uint16_t* post, int v, const uint32_t* mw, int p, uint32_t sel_even, uint32_t sel_odd)
{
    return pk_in_w(post[v], mw, p, sel_even, sel_odd);
}

/* Phase A of node (i, j) in a layer of D information edges: all edges
 * read, minimums and sign formed.  TAIL (the layer has nodes sharing
 * bits): the edges in dep (read after an earlier node of the layer) stay
 * out of the minimums; the hand-over rounds fold them. */
/* information-edge posterior indices of node j in a layer from its table
 * entries (the kernel reads them precomputed, pk_addr_words) */
template <int DM>
LG_PK_HD void pk_addrs(int* ad, const uint32_t* ent, int D, int j)
{
#pragma unroll
    for (int e = 0; e < DM; e++) {
        if (e >= D)
            continue;
        ad[e] = pk_addr(ent[e], j);
    }
}

/* the same from the precomputed rows: two indices a word, edge 2w in the
 * low half of word w */
template <int DM>
LG_PK_HD void pk_addr_words(int* ad, const uint32_t* aw, int D)
{
#pragma unroll
    for (int e = 0; e < DM; e++) {
        if (e >= D)
            continue;
        ad[e] = (int)((e & 1) ? aw[e >> 1] >> 16 : aw[e >> 1] & 0xffffu);
    }
}

template <int DM, bool TAIL>
LG_PK_HD void pk_phase_a(PkNode<DM>& n, const uint16_t* post, int D, int i, int j, int K, int q, uint32_t dep,
                         const uint32_t* mw, uint32_t sel_even, uint32_t sel_odd)
{
    n.apc = K + 360 * i + j;
    n.nopp = (i | j) ? 0u : ~0u;
    /* node (0, 0) reads its own parity bit as a stand-in, never used */
    n.app = i ? n.apc - 360 : j ? K + 360 * (q - 1) + j - 1 : n.apc;
    n.min0 = pk_dup(-1);
    n.min1 = pk_dup(-1);
    n.s = pk_dup(0);
    pk_fold(n, 0, pk_in(post, n.apc, mw, 0, sel_even, sel_odd));
    pk_fold_skip(n, 1, pk_in(post, n.app, mw, 1, sel_even, sel_odd), n.nopp);
#pragma unroll
    for (int e = 0; e < DM; e++) {
        if (!(e >= D)) {
            const lg_pk x = pk_in(post, n.ad[e], mw, e + 2, sel_even, sel_odd);
            if (TAIL)
                pk_fold_skip(n, e + 2, x, 0u - ((dep >> e) & 1));
            else
                pk_fold(n, e + 2, x);
        }
    }
}

/* The offset magnitudes of a complete node, for both slots: max(min(m, 127)
 * - 1, 0) in units of 256 (ldpc.c:353).  min(m, 127*256) also removes a
 * 0xff low byte, so mn0 / mn1 are exact multiples of 256. */
struct PkOut {
    lg_pk mn1, dmn;
};

template <int DM>
LG_PK_HD PkOut pk_mags(const PkNode<DM>& n)
{
    const lg_pk mn0 = pk_max(pk_minu(n.min0, pk_dup(0x7F00)) - pk_dup(0x0100), pk_dup(0));
    const lg_pk mn1 = pk_max(pk_minu(n.min1, pk_dup(0x7F00)) - pk_dup(0x0100), pk_dup(0));
    return PkOut{ mn1, mn0 - mn1 };
}

/* Learned magnitude of one minimum, per lane: clamp((min(m, 127) * S - B)
 * >> 4, 0, 127) in units of 256, S (0..63) and B (-1024..1023) in the
 * lanes of ps / pb, so |min * S - B| < 2^15.  S = 16, B = 16 gives
 * pk_mags' values. */
LG_PK_HD lg_pk pk_nms_mag(lg_pk m, lg_pk ps, lg_pk pb)
{
    const lg_pku v = pku(pk_minu(m, pk_dup(0x7F00))) >> (unsigned short)8;
    const lg_pk p = (pks(v * pku(ps)) - pb) >> (short)4;
    return pk_min(pk_max(p, pk_dup(0)), pk_dup(127)) << (short)8;
}

template <int DM, bool NMS>
LG_PK_HD PkOut pk_mags_p(const PkNode<DM>& n, lg_pk ps, lg_pk pb)
{
    if constexpr (NMS) {
        const lg_pk mn0 = pk_nms_mag(n.min0, ps, pb), mn1 = pk_nms_mag(n.min1, ps, pb);
        return PkOut{ mn1, mn0 - mn1 };
    } else {
        (void)ps;
        (void)pb;
        return pk_mags(n);
    }
}

/* min(d, 1) * dmn + mn1 per lane as two instructions: written out so that
 * the compiler does not turn them into per-lane compares and selects */
LG_PK_HD lg_pk pk_sel_mag(lg_pk d, lg_pk dmn, lg_pk mn1)
{
#if defined(__HIP_DEVICE_COMPILE__)
    uint32_t nz, r;
    asm("v_pk_min_u16 %0, %1, %2" : "=v"(nz) : "v"(pk_bits(d)), "s"(0x00010001u));
    asm("v_pk_mad_u16 %0, %1, %2, %3" : "=v"(r) : "v"(nz), "v"(pk_bits(dmn)), "v"(pk_bits(mn1)));
    return pk_of(r);
#else
    const lg_pku nz = __builtin_elementwise_min(pku(d), lg_pku{ 1, 1 });
    return pks(nz * pku(dmn) + pku(mn1));
#endif
}

/* o_p: |o| = mn1 if |x_p| is the smallest input, else mn0; sign = product
 * of the other signs.  a >= min0 always, so a - min0 is 0 exactly for the
 * smallest input.  Where a 0xff low byte makes |x_p| differ from min0 by
 * less than 256, both are 127 and mn0 == mn1 (both minimums >= 127). */
template <int DM>
LG_PK_HD lg_pk pk_out(const PkNode<DM>& n, const PkOut& m, int p)
{
    const lg_pk mag = pk_sel_mag(n.a[p] - n.min0, m.dmn, m.mn1);
    const lg_pk t = (n.s ^ n.x[p]) >> 15;
    return (mag ^ t) - t;
}

LG_PK_HD uint16_t pk_post_bytes(lg_pk y) { return (uint16_t)lg_perm(0, pk_bits(y), PK_PACK_POST); }

LG_PK_HD lg_pk pk_msg_clamp(lg_pk o) { return pk_min(pk_max(o, pk_dup(-32 * 256)), pk_dup(31 * 256)); }

/* Phase C: final posteriors (the bits this node touches last; TAIL: the
 * edges in succ are written by a later node of the layer) and the new
 * messages, clamped to [-32, 31] (ldpc.c:343): word w of the node's
 * LG_PK_WORDS(D) goes to mw_out[stride * w] as soon as both its positions
 * are known */
template <int DM, bool TAIL, bool NMS>
LG_PK_HD void pk_phase_c(const PkNode<DM>& n, uint16_t* post, int D, uint32_t succ, uint32_t* mw_out, int stride,
                         lg_pk ps, lg_pk pb)
{
    const PkOut m = pk_mags_p<DM, NMS>(n, ps, pb);
    {
        const lg_pk o0 = pk_out(n, m, 0);
        post[n.apc] = pk_post_bytes(pk_adds(n.x[0], o0));
        const lg_pk o1 = pk_out(n, m, 1);
        if (!n.nopp)
            post[n.app] = pk_post_bytes(pk_adds(n.x[1], o1));
        mw_out[0] = lg_perm(pk_bits(pk_msg_clamp(o1)) & ~n.nopp, pk_bits(pk_msg_clamp(o0)), PK_PACK_MSG);
    }
    uint32_t even = 0;
#pragma unroll
    for (int e = 0; e < DM; e++) {
        if (e >= D)
            continue;
        const lg_pk o = pk_out(n, m, e + 2);
        if (!TAIL || !((succ >> e) & 1))
            post[n.ad[e]] = pk_post_bytes(pk_adds(n.x[e + 2], o));
        const uint32_t c = pk_bits(pk_msg_clamp(o));
        if (!(e & 1)) {
            even = c;
            if (e + 1 == D)
                mw_out[stride * (e / 2 + 1)] = lg_perm(0, even, PK_PACK_MSG);
        } else {
            mw_out[stride * (e / 2 + 1)] = lg_perm(c, even, PK_PACK_MSG);
        }
    }
}

/* parity check of node (i, j) in both slots: bit k set if slot k holds it
 * (no zero posterior, even number of negative ones; bad(), ldpc.c:286-305).
 * A byte is non-zero iff bit 7 of ((b & 0x7f) + 0x7f) | b is set. */
template <int DM>
LG_PK_HD uint32_t pk_check(const uint16_t* post, const uint32_t* ent, int D, int i, int j, int K, int q)
{
    uint32_t par = 0, nz = 0x8080;
#pragma unroll
    for (int e = 0; e < DM; e++) {
        if (e >= D)
            continue;
        const uint32_t r = post[pk_addr(ent[e], j)];
        par ^= r;
        nz &= ((r & 0x7F7F) + 0x7F7F) | r;
    }
    const int apc = K + 360 * i + j;
    uint32_t r = post[apc];
    par ^= r;
    nz &= ((r & 0x7F7F) + 0x7F7F) | r;
    if (i | j) {
        r = post[i ? apc - 360 : K + 360 * (q - 1) + j - 1];
        par ^= r;
        nz &= ((r & 0x7F7F) + 0x7F7F) | r;
    }
    const uint32_t good = nz & ~par & 0x8080;
    return ((good >> 7) & 1) | ((good >> 14) & 2);
}

/* ---- layers with nodes sharing a bit: rounds to a fixed point ----------
 *
 * The reference updates the nodes of a layer one after another
 * (ldpc.c:474-480): a node that shares an information bit with earlier
 * nodes of the layer reads it as the latest of them left it.  Here all
 * nodes of such a layer run at once, in rounds.  Round 0 takes every bit
 * as the layer found it; in round r a node reads each shared bit as its
 * predecessor on that bit wrote it in round r - 1 (hand-over slots,
 * double-buffered) and writes what it hands on to its successors.  A node
 * whose inputs are exact after round r - 1 is exact after round r, so all
 * are exact after (levels - 1) rounds; and once a round reads the same
 * inputs as the round before, every later round would too: that state is
 * the sequential result (the hand-over graph has no cycle, so its fixed
 * point is unique).  The rounds stop at the first of the two.
 *
 * Hand-over slot of edge e at rank r (its node is the r-th of the layer
 * touching the bit, 0-based): the bit is handed from rank r to r + 1
 * through slot (t0 + r) * 360 + m, m the bit's index in its group. */
struct PkPart {
    lg_pk min0, min1, s;
};

template <int DM>
LG_PK_HD PkPart pk_part(const PkNode<DM>& n)
{
    return PkPart{ n.min0, n.min1, n.s };
}

template <int DM>
LG_PK_HD void pk_restore(PkNode<DM>& n, const PkPart& p)
{
    n.min0 = p.min0;
    n.min1 = p.min1;
    n.s = p.s;
}

/* round 0: the dependent edges with the values phase A read */
template <int DM>
LG_PK_HD void pk_fold_dep0(PkNode<DM>& n, int D, uint32_t dep)
{
#pragma unroll
    for (int e = 0; e < DM; e++) {
        if (e >= D)
            continue;
        if ((dep >> e) & 1)
            pk_fold(n, e + 2, n.x[e + 2]);
    }
}

LG_PK_HD int pk_slot(uint32_t xe, int ad, int r) { return (LG_XENT_T0(xe) + r) * 360 + ad - (int)(xe >> 16); }

/* write the values handed to the successors (edges in succ) */
template <int DM, bool NMS>
LG_PK_HD void pk_put_hand(const PkNode<DM>& n, uint16_t* hb, const uint32_t* xent, int D, uint32_t dupm, uint32_t succ,
                          uint64_t pos, lg_pk ps, lg_pk pb)
{
    const PkOut m = pk_mags_p<DM, NMS>(n, ps, pb);
#pragma unroll
    for (int e = 0; e < DM; e++) {
        if (e >= D)
            continue;
        if (!((dupm >> e) & 1)) /* the same for every node of the layer */
            continue;
        if ((succ >> e) & 1)
            hb[pk_slot(xent[e], n.ad[e], (int)((pos >> (2 * e)) & 3))] =
                pk_post_bytes(pk_adds(n.x[e + 2], pk_out(n, m, e + 2)));
    }
}

/* take the values handed over by the predecessors (edges in dep), fold
 * them into the minimums restored from the other edges; returns non-zero
 * if an input differs from the previous round's */
template <int DM>
LG_PK_HD uint32_t pk_take_hand(PkNode<DM>& n, const uint16_t* hb, const uint32_t* xent, int D, uint32_t dupm,
                               uint32_t dep, uint64_t pos, const uint32_t* mw, uint32_t sel_even, uint32_t sel_odd)
{
    uint32_t chg = 0;
#pragma unroll
    for (int e = 0; e < DM; e++) {
        if (e >= D)
            continue;
        if (!((dupm >> e) & 1))
            continue;
        if ((dep >> e) & 1) {
            const int r = (int)((pos >> (2 * e)) & 3);
            const lg_pk x = pk_in_w(hb[pk_slot(xent[e], n.ad[e], r - 1)], mw, e + 2, sel_even, sel_odd);
            chg |= pk_bits(x) ^ pk_bits(n.x[e + 2]);
            pk_fold(n, e + 2, x);
        }
    }
    return chg;
}

#endif
