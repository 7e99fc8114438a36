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
 * Messages: one int8 per edge per slot, two edges a word
 * [e0 slot0, e0 slot1, e1 slot0, e1 slot1], edge e in word e / 2.
 */
#ifndef LDPC_PK_H
#define LDPC_PK_H

#include <stdint.h>

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

template <int DEG>
struct PkNode {
    int ad[DEG];       /* posterior index (variable) of each information edge */
    lg_pk x[DEG + 2];  /* extrinsic inputs, both slots */
    lg_pk a[DEG + 2];  /* their magnitudes */
    int apc, app, has_pp;
    lg_pk min0, min1, s;
};

/* posterior index of node j's edge on table entry en (base << 16 | shift):
 * base + (j - shift) mod 360, the wrap as an unsigned minimum */
LG_PK_HD int pk_addr(uint32_t en, int j)
{
    const uint32_t t = (uint32_t)j - (en & 0xffff);
    const uint32_t u = t + 360u;
    return (int)(en >> 16) + (int)(t < u ? t : u);
}

LG_PK_HD lg_pk pk_abs(lg_pk x) { return pk_max(x, pk_subs(pk_dup(0), x)); }

template <int DEG>
LG_PK_HD void pk_fold(PkNode<DEG>& n, int e, lg_pk x)
{
    const lg_pk a = pk_abs(x);
    n.x[e] = x;
    n.a[e] = a;
    n.min1 = pk_minu(n.min1, pk_maxu(n.min0, a));
    n.min0 = pk_minu(n.min0, a);
    n.s = n.s ^ x;
}

/* pk_fold of the lanes where skip is 0; lanes where it is all ones leave
 * the minimums and the sign as they were (|x| | 0xffff is above any
 * minimum).  Branch-free, so the posterior reads of all edges can be in
 * flight together. */
template <int DEG>
LG_PK_HD void pk_fold_skip(PkNode<DEG>& n, int e, lg_pk x, uint32_t skip)
{
    const lg_pk a = pk_of(pk_bits(pk_abs(x)) | skip);
    n.x[e] = x;
    n.a[e] = a;
    n.min1 = pk_minu(n.min1, pk_maxu(n.min0, a));
    n.min0 = pk_minu(n.min0, a);
    n.s = pk_of(pk_bits(n.s) ^ (pk_bits(x) & ~skip));
}

LG_PK_HD lg_pk pk_in(const uint16_t* post, int v, const uint32_t* mw, int e, uint32_t sel_even, uint32_t sel_odd)
{
    const lg_pk L = pk_of(lg_perm(0, post[v], PK_UNPACK_LO));
    const lg_pk M = pk_of(lg_perm(0, mw[e >> 1], (e & 1) ? sel_odd : sel_even));
    return pk_subs(L, M);
}

/* TAIL: the layer has dependent nodes (dep masks); without, every edge is
 * read in phase A and no compile-time-unknown mask is tested */
template <int DEG, bool TAIL>
LG_PK_HD void pk_phase_a(PkNode<DEG>& n, const uint16_t* post, const uint32_t* ent, int i, int j, int K, int q,
                         uint32_t dep, const uint32_t* mw, uint32_t sel_even, uint32_t sel_odd)
{
#pragma unroll
    for (int e = 0; e < DEG; e++)
        n.ad[e] = pk_addr(ent[e], j);
    n.apc = K + 360 * i + j;
    n.has_pp = (i | j) != 0;
    n.app = i ? n.apc - 360 : K + 360 * (q - 1) + j - 1;
    n.min0 = pk_dup(-1);
    n.min1 = pk_dup(-1);
    n.s = pk_dup(0);
#pragma unroll
    for (int e = 0; e < DEG; e++) {
        const lg_pk x = pk_in(post, n.ad[e], mw, e, sel_even, sel_odd);
        if (TAIL) /* dependent edges are folded by pk_step */
            pk_fold_skip(n, e, x, 0u - ((dep >> e) & 1));
        else
            pk_fold(n, e, x);
    }
    pk_fold(n, DEG, pk_in(post, n.apc, mw, DEG, sel_even, sel_odd));
    if (n.has_pp) {
        pk_fold(n, DEG + 1, pk_in(post, n.app, mw, DEG + 1, sel_even, sel_odd));
    } else {
        n.x[DEG + 1] = pk_dup(0);
        n.a[DEG + 1] = pk_dup(0);
    }
}

template <int DEG>
LG_PK_HD void pk_step(PkNode<DEG>& n, const uint16_t* post, uint32_t dep, const uint32_t* mw, uint32_t sel_even,
                      uint32_t sel_odd)
{
#pragma unroll
    for (int e = 0; e < DEG; e++)
        if ((dep >> e) & 1)
            pk_fold(n, e, pk_in(post, n.ad[e], mw, e, sel_even, sel_odd));
}

/* The offset magnitudes of a complete node, for both slots: max(min(m, 127)
 * - 1, 0) in units of 256 (ldpc.c:353).  min(m, 127*256) also removes a
 * 0xff low byte, so mn0 / mn1 are exact multiples of 256. */
template <int DEG>
struct PkOut {
    lg_pk mn1, dmn;
};

template <int DEG>
LG_PK_HD PkOut<DEG> pk_mags(const PkNode<DEG>& n)
{
    const lg_pk mn0 = pk_max(pk_minu(n.min0, pk_dup(0x7F00)) - pk_dup(0x0100), pk_dup(0));
    const lg_pk mn1 = pk_max(pk_minu(n.min1, pk_dup(0x7F00)) - pk_dup(0x0100), pk_dup(0));
    return PkOut<DEG>{ mn1, mn0 - mn1 };
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

template <int DEG, bool NMS>
LG_PK_HD PkOut<DEG> pk_mags_p(const PkNode<DEG>& n, lg_pk ps, lg_pk pb)
{
    if constexpr (NMS) {
        const lg_pk mn0 = pk_nms_mag(n.min0, ps, pb), mn1 = pk_nms_mag(n.min1, ps, pb);
        return PkOut<DEG>{ mn1, mn0 - mn1 };
    } else {
        (void)ps;
        (void)pb;
        return pk_mags(n);
    }
}

/* o_e: |o| = mn1 if |x_e| is the smallest input, else mn0; sign = product
 * of the other signs.  a >= min0 always, so a - min0 is 0 exactly for the
 * smallest input.  Where a 0xff low byte makes |x_e| differ from min0 by
 * less than 256, both are 127 and mn0 == mn1 (both minimums >= 127). */
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

template <int DEG>
LG_PK_HD lg_pk pk_out(const PkNode<DEG>& n, const PkOut<DEG>& m, int e)
{
    const lg_pk mag = pk_sel_mag(n.a[e] - n.min0, m.dmn, m.mn1);
    const lg_pk t = (n.s ^ n.x[e]) >> 15;
    return (mag ^ t) - t;
}

LG_PK_HD uint16_t pk_post_bytes(lg_pk y) { return (uint16_t)lg_perm(0, pk_bits(y), PK_PACK_POST); }

template <int DEG, bool NMS = false>
LG_PK_HD void pk_write_succ(const PkNode<DEG>& n, uint16_t* post, uint32_t succ, lg_pk ps = lg_pk{ 16, 16 },
                            lg_pk pb = lg_pk{ 16, 16 })
{
    const PkOut<DEG> m = pk_mags_p<DEG, NMS>(n, ps, pb);
#pragma unroll
    for (int e = 0; e < DEG; e++)
        if ((succ >> e) & 1)
            post[n.ad[e]] = pk_post_bytes(pk_adds(n.x[e], pk_out(n, m, e)));
}

/* final posteriors (bits this node touches last) and the new messages,
 * clamped to [-32, 31] (ldpc.c:343) */
template <int DEG, bool TAIL, bool NMS = false>
LG_PK_HD void pk_phase_c(const PkNode<DEG>& n, uint16_t* post, uint32_t succ, uint32_t* mw_out,
                         lg_pk ps = lg_pk{ 16, 16 }, lg_pk pb = lg_pk{ 16, 16 })
{
    const PkOut<DEG> m = pk_mags_p<DEG, NMS>(n, ps, pb);
    lg_pk mo[DEG + 2];
#pragma unroll
    for (int e = 0; e < DEG + 2; e++) {
        if (e == DEG + 1 && !n.has_pp) {
            mo[e] = pk_dup(0);
            continue;
        }
        const lg_pk o = pk_out(n, m, e);
        if (e >= DEG || !TAIL || !((succ >> e) & 1)) {
            const int v = e < DEG ? n.ad[e] : e == DEG ? n.apc : n.app;
            post[v] = pk_post_bytes(pk_adds(n.x[e], o));
        }
        mo[e] = pk_min(pk_max(o, pk_dup(-32 * 256)), pk_dup(31 * 256));
    }
#pragma unroll
    for (int w = 0; w < LG_PK_WORDS(DEG); w++) {
        const lg_pk odd = 2 * w + 1 < DEG + 2 ? mo[2 * w + 1] : pk_dup(0);
        mw_out[w] = lg_perm(pk_bits(odd), pk_bits(mo[2 * w]), PK_PACK_MSG);
    }
}

/* parity check of node (i, j) in both slots: bit k set if slot k holds it
 * (no zero posterior, even number of negative ones; bad(), ldpc.c:286-305).
 * A byte is non-zero iff bit 7 of ((b & 0x7f) + 0x7f) | b is set. */
template <int DEG>
LG_PK_HD uint32_t pk_check(const uint16_t* post, const uint32_t* ent, int i, int j, int K, int q)
{
    uint32_t par = 0, nz = 0x8080;
#pragma unroll
    for (int e = 0; e < DEG; e++) {
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

#endif
