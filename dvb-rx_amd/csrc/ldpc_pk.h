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
 * |y| = 127 where they cannot change a result (see sk_out, ldpc_sk.h).
 *
 * Edges of a node by position: 0 its own parity bit, 1 the previous parity
 * bit (the IRA staircase, ldpc.c:277-284), 2 + e information edge e of the
 * layer.  Messages: one int8 per edge per slot, two positions a word
 * [p0 slot0, p0 slot1, p1 slot0, p1 slot1], position p in word p / 2.
 *
 */
#ifndef LDPC_PK_H
#define LDPC_PK_H

#include <stdint.h>

#include "ldpc_codes.h"

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

/* unsigned saturating subtract of c from both lanes (v_pk_sub_u16 clamp) */
LG_PK_HD lg_pk pk_subu_sat(lg_pk a, unsigned short c)
{
    return pks(__builtin_elementwise_sub_sat(pku(a), lg_pku{ c, c }));
}

/* the magnitude (units of 256, low byte 0) of a ones' complement
 * magnitude a = x ^ (x >> 15): (a + 0x80) & 0xff00, saturating */
LG_PK_HD lg_pk pk_mag1(lg_pk a)
{
    return pk_of(pk_bits(pks(__builtin_elementwise_add_sat(pku(a), lg_pku{ 0x80, 0x80 }))) & 0xFF00FF00u);
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

/* The offset magnitudes of a complete node, for both slots: max(min(m, 127)
 * - 1, 0) in units of 256 (ldpc.c:353).  min(m, 127*256) also removes a
 * 0xff low byte, so mn0 / mn1 are exact multiples of 256. */
struct PkOut {
    lg_pk mn1, dmn;
};

/* Learned magnitude of one minimum m (units of 256, low byte 0), per
 * lane: clamp((min(m, 127) * S - B) >> 4, 0, 127) in units of 256, S
 * (0..63) and B (-1024..1023) in the lanes of ps / pb, so |min * S - B| <
 * 2^15.  S = 16, B = 16 gives the plain decoder's values. */
LG_PK_HD lg_pk pk_nms_mag(lg_pk m, lg_pk ps, lg_pk pb)
{
    const lg_pku v = pku(pk_minu(m, pk_dup(0x7F00))) >> (unsigned short)8;
    const lg_pk p = (pks(v * pku(ps)) - pb) >> (short)4;
    return pk_min(pk_max(p, pk_dup(0)), pk_dup(127)) << (short)8;
}

/* pk_nms_mag(pk_mag1(a), ps, pb) from a ones' complement magnitude a, with
 * nb = -pb: min((a +sat 0x80) >> 8, 127) is min(m, 127) (the shift drops
 * the low byte that pk_mag1 masks), and v * S + nb a multiply-add; no
 * intermediate leaves int16 (|v * S - B| < 2^15) */
LG_PK_HD lg_pk pk_nms_mag_a(lg_pk a, lg_pk ps, lg_pk nb)
{
    const lg_pku v = __builtin_elementwise_min(__builtin_elementwise_add_sat(pku(a), lg_pku{ 0x80, 0x80 }) >>
                                                   (unsigned short)8,
                                               lg_pku{ 127, 127 });
    const lg_pk p = pks(v * pku(ps) + pku(nb)) >> (short)4;
    return pk_min(pk_max(p, pk_dup(0)), pk_dup(127)) << (short)8;
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

/* pk_sel_mag of two inputs, the instructions of the two interleaved so
 * that no instruction reads the result of the one just before it (a
 * packed result is read one instruction later at the earliest) */
LG_PK_HD void pk_sel_mag2(lg_pk d0, lg_pk d1, lg_pk dmn, lg_pk mn1, lg_pk& r0, lg_pk& r1)
{
#if defined(__HIP_DEVICE_COMPILE__)
    uint32_t a, b;
    asm("v_pk_min_u16 %0, %2, %4\n\t"
        "v_pk_min_u16 %1, %3, %4\n\t"
        "v_pk_mad_u16 %0, %0, %5, %6\n\t"
        "v_pk_mad_u16 %1, %1, %5, %6"
        : "=&v"(a), "=&v"(b)
        : "v"(pk_bits(d0)), "v"(pk_bits(d1)), "s"(0x00010001u), "v"(pk_bits(dmn)), "v"(pk_bits(mn1)));
    r0 = pk_of(a);
    r1 = pk_of(b);
#else
    r0 = pk_sel_mag(d0, dmn, mn1);
    r1 = pk_sel_mag(d1, dmn, mn1);
#endif
}

LG_PK_HD uint16_t pk_post_bytes(lg_pk y) { return (uint16_t)lg_perm(0, pk_bits(y), PK_PACK_POST); }

LG_PK_HD lg_pk pk_msg_clamp(lg_pk o) { return pk_min(pk_max(o, pk_dup(-32 * 256)), pk_dup(31 * 256)); }

/* a node's minimums and sign without the edges that rounds take again */
struct PkPart {
    lg_pk min0, min1, s;
};

#endif
