/*
 * Check nodes split over two lanes: the arithmetic of ldpc_pk.h (two
 * frames in the 16-bit halves of each register) with each node's edges
 * shared by a lane pair, shared by the HIP kernel (ldpc_sk_kern.hip) and
 * the CPU model of it (built with clang, same code).
 *
 * Node j of a layer is served by lanes 2j (half h = 0) and 2j + 1 (h = 1).
 * The node's edges by position p: 0 its own parity bit, 1 the previous
 * parity bit (the IRA staircase, ldpc.c:277-284), then the information
 * edges, in layers with nodes sharing bits those of the shared groups
 * first (slots 1 .. kde - 1, the only ones the hand-over rounds touch);
 * half h takes the positions p = 2k + h, its "slots" k = 0..KS-1.  A
 * layer of T = degree + 2 positions leaves slot k of half h empty when
 * 2k + h >= T (only slots k >= KA can be, KA a template bound), and node
 * (0, 0) has no previous parity (slot 0 of half 1 in layer 0).  Empty
 * slots point at a spare posterior cell; they are left out of the
 * minimums, the sign and the syndrome by all-ones skip masks, and what
 * they write is never read.
 *
 * Each half forms the minimums and the sign of its own slots; the pair
 * exchanges them (sk_combine; the two smallest of a union are the two
 * smallest of the two halves' two smallest) and both continue with the
 * node's: the results are those of the whole node, bit for bit.
 *
 * Messages: slot k in word k / 2 of the half, bytes [k even: 0, 1; odd:
 * 2, 3] (slot 0 of the two frames, then slot 1), as the selectors of
 * pk_msg_sel give them.
 *
 * Magnitudes are kept in ones' complement: a = x ^ t with t = x >> 15
 * (the lane's sign mask), which is |x| for x >= 0 and |x| - 1 for x < 0.
 * It orders the edges as |x| does, ties broken by sign, so the two
 * smallest a are edges with the two smallest |x|, and the magnitude of an
 * a is (a + 0x80) >> 8 (in units of 256); the sign mask comes back as a ^
 * x where the output needs it.  The node's sign s is the XOR of the masks.
 *
 * Slot addresses are LDS byte offsets (the code registry's address
 * records): the posterior of bit v at SK_POST_OFF + 2v from the start of
 * the kernel's LDS, the functions below take that start ("lds").
 */
#ifndef LDPC_SK_H
#define LDPC_SK_H

#include <stdint.h>

#include "ldpc_pk.h"

/* the two posterior bytes at byte offset ad of the LDS.  On the device
 * the kernel's LDS is all dynamic and starts at address 0 (the kernel
 * checks it), so ad is the LDS address itself. */
#if defined(__HIP_DEVICE_COMPILE__)
typedef __attribute__((address_space(3))) uint16_t sk_lds_u16;
LG_PK_HD uint32_t sk_ld(const uint16_t* lds, uint32_t ad)
{
    (void)lds;
    return *(const sk_lds_u16*)(size_t)ad;
}
LG_PK_HD void sk_st(uint16_t* lds, uint32_t ad, uint16_t v)
{
    (void)lds;
    *(sk_lds_u16*)(size_t)ad = v;
}
#else
LG_PK_HD uint32_t sk_ld(const uint16_t* lds, uint32_t ad) { return *(const uint16_t*)((const uint8_t*)lds + ad); }
LG_PK_HD void sk_st(uint16_t* lds, uint32_t ad, uint16_t v) { *(uint16_t*)((uint8_t*)lds + ad) = v; }
#endif

template <int KS>
struct SkHalf {
    uint32_t ad[KS]; /* posterior byte offset of each slot */
    lg_pk x[KS];     /* extrinsic inputs, both frames */
    lg_pk a[KS];     /* their magnitudes, ones' complement */
    lg_pk min0, min1, s;
};

/* slot posterior offsets from the lane's address record */
template <int KS>
LG_PK_HD void sk_addrs(SkHalf<KS>& n, const uint32_t* aw)
{
#pragma unroll
    for (int k = 0; k < KS; k++)
        n.ad[k] = aw[k];
}

/* skip mask of slot k: all ones if it is empty (2k + h >= T, or node
 * (0, 0)'s missing previous parity: skip0 on slot 0) */
template <int KA>
LG_PK_HD uint32_t sk_skip(int k, int T, int h, uint32_t skip0)
{
    uint32_t m = k == 0 ? skip0 : 0u;
    if (k >= KA)
        m |= 2 * k + h >= T ? ~0u : 0u;
    return m;
}

/* the posteriors of a half's slots (two bytes, both frames), all reads
 * issued before any is used */
template <int KS>
LG_PK_HD void sk_read(uint32_t* raw, const uint16_t* post, const SkHalf<KS>& n)
{
#pragma unroll
    for (int k = 0; k < KS; k++)
        raw[k] = sk_ld(post, n.ad[k]);
}

/* Phase A of a half from its posteriors raw: minimums and sign of the
 * present slots.  TAIL: the slots in dep (read after an earlier node of
 * the layer) stay out too; the hand-over rounds fold them. */
template <int KS, int KA, bool TAIL>
LG_PK_HD void sk_phase_a(SkHalf<KS>& n, const uint32_t* raw, int T, int h, uint32_t skip0, uint32_t dep,
                         const uint32_t* mw, uint32_t sel_even, uint32_t sel_odd)
{
    n.min0 = pk_dup(-1);
    n.min1 = pk_dup(-1);
    n.s = pk_dup(0);
#pragma unroll
    for (int k = 0; k < KS; k++) {
        const lg_pk x = pk_in_w(raw[k], mw, k, sel_even, sel_odd);
        uint32_t skip = sk_skip<KA>(k, T, h, skip0);
        if (TAIL)
            skip |= 0u - ((dep >> k) & 1);
        const lg_pk t = x >> (short)15;
        const lg_pk a = pk_of(pk_bits(x ^ t) | skip);
        n.x[k] = x;
        n.a[k] = a;
        n.min1 = pk_minu(n.min1, pk_maxu(n.min0, a));
        n.min0 = pk_minu(n.min0, a);
        n.s = pk_of(pk_bits(n.s) ^ (pk_bits(t) & ~skip));
    }
}

/* fold slot k with input x into the half's minimums and sign */
template <int KS>
LG_PK_HD void sk_fold(SkHalf<KS>& n, int k, lg_pk x)
{
    const lg_pk t = x >> (short)15;
    const lg_pk a = x ^ t;
    n.x[k] = x;
    n.a[k] = a;
    n.min1 = pk_minu(n.min1, pk_maxu(n.min0, a));
    n.min0 = pk_minu(n.min0, a);
    n.s = n.s ^ t;
}

/* merge the partner half's minimums and sign (p0, p1, ps) */
template <int KS>
LG_PK_HD void sk_combine(SkHalf<KS>& n, lg_pk p0, lg_pk p1, lg_pk ps)
{
    const lg_pk hi = pk_maxu(n.min0, p0);
    n.min0 = pk_minu(n.min0, p0);
    n.min1 = pk_minu(pk_minu(n.min1, p1), hi);
    n.s = n.s ^ ps;
}

/* NMS: the learned magnitudes with scales ps and negated offsets nb */
template <int KS, bool NMS>
LG_PK_HD PkOut sk_mags(const SkHalf<KS>& n, lg_pk ps, lg_pk nb)
{
    if constexpr (NMS) {
        const lg_pk mn0 = pk_nms_mag_a(n.min0, ps, nb), mn1 = pk_nms_mag_a(n.min1, ps, nb);
        return PkOut{ mn1, mn0 - mn1 };
    } else {
        (void)ps;
        (void)nb;
        /* max(min(m, 127) - 1, 0) of m = (a + 0x80) >> 8, in units of 256:
         * min((a -sat 0x80) & 0xff00, 126 * 256) */
        const lg_pk mn0 = pk_minu(pk_of(pk_bits(pk_subu_sat(n.min0, 0x80)) & 0xFF00FF00u), pk_dup(0x7E00));
        const lg_pk mn1 = pk_minu(pk_of(pk_bits(pk_subu_sat(n.min1, 0x80)) & 0xFF00FF00u), pk_dup(0x7E00));
        return PkOut{ mn1, mn0 - mn1 };
    }
}

/* o_k: |o| = mn1 if x_k is the node's smallest input, else mn0; sign =
 * product of the other signs (the node's sign mask s with x_k's, a_k ^ x_k,
 * taken out).  a >= min0 always, so a - min0 is 0 exactly for the input
 * that is the smallest; inputs of equal magnitude give mn0 == mn1, and so
 * does a 0xff low byte (|x| = 127: both minimums at 127). */
/* the output of slot k from its magnitude: the sign applied */
template <int KS>
LG_PK_HD lg_pk sk_sign(const SkHalf<KS>& n, int k, lg_pk mag)
{
    const lg_pk t = n.s ^ n.a[k] ^ n.x[k];
    return (mag ^ t) - t;
}

template <int KS>
LG_PK_HD lg_pk sk_out(const SkHalf<KS>& n, const PkOut& m, int k)
{
    return sk_sign(n, k, pk_sel_mag(n.a[k] - n.min0, m.dmn, m.mn1));
}

/* Phase C of a half: final posteriors (TAIL: the slots in succ are written
 * by a later node of the layer) and the new messages, clamped to [-32,
 * 31] (ldpc.c:343), word w to mo[w] */
template <int KS, bool TAIL, bool NMS>
LG_PK_HD void sk_phase_c(const SkHalf<KS>& n, uint16_t* post, uint32_t succ, uint32_t* mo, lg_pk ps, lg_pk nb)
{
    const PkOut m = sk_mags<KS, NMS>(n, ps, nb);
    /* the slots' magnitudes first, two at a time */
    lg_pk mag[KS];
#pragma unroll
    for (int k = 0; k + 1 < KS; k += 2)
        pk_sel_mag2(n.a[k] - n.min0, n.a[k + 1] - n.min0, m.dmn, m.mn1, mag[k], mag[k + 1]);
    if (KS & 1)
        mag[KS - 1] = pk_sel_mag(n.a[KS - 1] - n.min0, m.dmn, m.mn1);
    uint32_t even = 0;
#pragma unroll
    for (int k = 0; k < KS; k++) {
        const lg_pk o = sk_sign(n, k, mag[k]);
        if (!TAIL || !((succ >> k) & 1))
            sk_st(post, n.ad[k], pk_post_bytes(pk_adds(n.x[k], o)));
        const uint32_t c = pk_bits(pk_msg_clamp(o));
        if (!(k & 1)) {
            even = c;
            if (k + 1 == KS)
                mo[k / 2] = lg_perm(0, even, PK_PACK_MSG);
        } else {
            mo[k / 2] = lg_perm(c, even, PK_PACK_MSG);
        }
    }
}

/* parity and non-zero flags of a half's present slots (bad(),
 * ldpc.c:286-305): par collects the sign bits (bit 7 / 15 of the two
 * frames), nz bit 7 / 15 stays set while no posterior is zero (a byte b is
 * non-zero iff bit 7 of ((b & 0x7f) + 0x7f) | b is set) */
template <int KS, int KA>
LG_PK_HD void sk_check(const uint16_t* post, const uint32_t* ad, int T, int h, uint32_t skip0, uint32_t& par,
                       uint32_t& nz)
{
    par = 0;
    nz = 0x8080;
#pragma unroll
    for (int k = 0; k < KS; k++) {
        const uint32_t r = sk_ld(post, ad[k]);
        const uint32_t skip = sk_skip<KA>(k, T, h, skip0);
        par ^= r & ~skip;
        nz &= ((r & 0x7F7F) + 0x7F7F) | r | skip;
    }
}

/* node result of the pair's flags: bit f set if frame f holds the check */
LG_PK_HD uint32_t sk_check_bits(uint32_t par, uint32_t nz)
{
    const uint32_t good = nz & ~par & 0x8080;
    return ((good >> 7) & 1) | ((good >> 14) & 2);
}

/* ---- layers with nodes sharing a bit: hand-over rounds -------------------
 *
 * The reference updates the nodes of a layer one after another
 * (ldpc.c:474-480): a node that shares an information bit with earlier
 * nodes of the layer reads it as the latest of them left it.  Here all
 * nodes of such a layer run at once, in rounds.  Round 0 takes every bit
 * as the layer found it; in round r a node reads each shared bit as its
 * predecessor on that bit handed it over in round r - 1 (hand-over cells,
 * double-buffered) and hands on what it would write to its successors.  A
 * node whose inputs are exact after round r - 1 is exact after round r, so
 * all are exact after (levels - 1) rounds; and once a round hands over the
 * same values as the round before, the next round would read the inputs
 * of this one, and so would every later round: that state is the
 * sequential result (the hand-over graph has no cycle, so its fixed point
 * is unique).  The rounds stop at the first of the two.
 *
 * hw: the slots' hand-over cells, two u16 a word (the tail record after
 * its mask word); a slot
 * at rank r on its bit (its node is the r-th of the layer touching the
 * bit, 0-based) hands over through cell c = (t0 + r) * 360 + m, m the
 * bit's index in its group, and reads its predecessor's at c - 360. */
LG_PK_HD int sk_cell(const uint32_t* hw, int k) { return (int)((k & 1) ? hw[k >> 1] >> 16 : hw[k >> 1] & 0xffffu); }

template <int KS>
LG_PK_HD PkPart sk_part(const SkHalf<KS>& n)
{
    return PkPart{ n.min0, n.min1, n.s };
}

template <int KS>
LG_PK_HD void sk_restore(SkHalf<KS>& n, const PkPart& p)
{
    n.min0 = p.min0;
    n.min1 = p.min1;
    n.s = p.s;
}

/* round 0: the dependent slots with the values phase A read */
template <int KS>
LG_PK_HD void sk_fold_dep0(SkHalf<KS>& n, uint32_t dep, int kde)
{
#pragma unroll
    for (int k = 1; k <= SK_KD(KS); k++) {
        if (k >= kde)
            continue;
        if ((dep >> k) & 1)
            sk_fold(n, k, n.x[k]);
    }
}

/* write the values handed to the successors (slots in succ); n holds the
 * node's combined minimums */
template <int KS, bool NMS>
LG_PK_HD void sk_put_hand(const SkHalf<KS>& n, uint16_t* hb, const uint32_t* hw, uint32_t succ, int kde, lg_pk ps,
                          lg_pk nb)
{
    const PkOut m = sk_mags<KS, NMS>(n, ps, nb);
#pragma unroll
    for (int k = 1; k <= SK_KD(KS); k++) {
        if (k >= kde)
            continue;
        if ((succ >> k) & 1)
            hb[sk_cell(hw, k)] = pk_post_bytes(pk_adds(n.x[k], sk_out(n, m, k)));
    }
}

/* the values this half handed over in the previous round (slots in succ,
 * own cells of that round's buffer hb), read before the round writes */
template <int KS>
LG_PK_HD void sk_prev_hand(uint32_t* old, const uint16_t* hb, const uint32_t* hw, uint32_t succ, int kde)
{
#pragma unroll
    for (int k = 1; k <= SK_KD(KS); k++) {
        if (k >= kde)
            continue;
        old[k] = ((succ >> k) & 1) ? hb[sk_cell(hw, k)] : 0u;
    }
}

/* sk_put_hand, and non-zero if a value differs from old (the previous
 * round's): when no value handed over in the workgroup changed, the next
 * round would see the inputs of this one, and the layer is at its fixed
 * point */
template <int KS, bool NMS>
LG_PK_HD uint32_t sk_put_hand_chg(const SkHalf<KS>& n, uint16_t* hb, const uint32_t* hw, uint32_t succ, int kde,
                                  const uint32_t* old, lg_pk ps, lg_pk nb)
{
    const PkOut m = sk_mags<KS, NMS>(n, ps, nb);
    uint32_t chg = 0;
#pragma unroll
    for (int k = 1; k <= SK_KD(KS); k++) {
        if (k >= kde)
            continue;
        if ((succ >> k) & 1) {
            const uint16_t v = pk_post_bytes(pk_adds(n.x[k], sk_out(n, m, k)));
            chg |= v ^ old[k];
            hb[sk_cell(hw, k)] = v;
        }
    }
    return chg;
}

/* take the values handed over by the predecessors (slots in dep) into the
 * half's minimums */
template <int KS>
LG_PK_HD void sk_take_hand(SkHalf<KS>& n, const uint16_t* hb, const uint32_t* hw, uint32_t dep, int kde,
                           const uint32_t* mw, uint32_t sel_even, uint32_t sel_odd)
{
#pragma unroll
    for (int k = 1; k <= SK_KD(KS); k++) {
        if (k >= kde)
            continue;
        if ((dep >> k) & 1)
            sk_fold(n, k, pk_in_w(hb[sk_cell(hw, k) - 360], mw, k, sel_even, sel_odd));
    }
}

#endif
