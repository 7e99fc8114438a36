/*
 * LDPC code registry and layer-schedule compiler (host C).
 *
 * Every DVB-S2/S2X/T2 code is quasi-cyclic with circulant size 360.  The
 * reference expands the address tables into one index list per check node
 * (libdvbrx/src/ldpc.c:235-255) and visits the R check nodes in "layer
 * order": check q*j+i is node j of layer i (ldpc.c:251-255).  Here the same
 * graph is kept in its circulant form: within layer i every node j has the
 * same info degree, and its e-th information neighbour is
 *
 *     bit = base_e + ((j - shift_e) mod 360),
 *
 * for one (base_e = 360 * table row, shift_e = address div q) pair per
 * table address congruent to i mod q.  A layer is thus described by its
 * degree and a handful of (base, shift) pairs instead of 360 index lists.
 *
 * The reference updates the 360 nodes of a layer one after another with
 * the posteriors updated in place (ldpc.c:474-480), so a node that shares
 * an information bit with an earlier node of the same layer sees that
 * node's update.  The schedule compiler assigns each node the level
 * 1 + max(level of the earlier nodes it shares a bit with); running the
 * levels of a layer in order, with the nodes of one level in parallel,
 * reproduces the sequential result exactly.
 */
#ifndef LDPC_CODES_H
#define LDPC_CODES_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LG_M 360          /* circulant size */
#define LG_MAX_CN_DEG 32  /* check node degree bound (tables: <= 30) */

typedef struct lg_code {
    int id;           /* index in the registry */
    const char* name; /* table name, e.g. "S2_B4" */
    int n, k, r, q;   /* codeword bits, info bits, parity bits, layers */
    int short_frame;
    int links_total, links_max_cn;
    int nent;             /* (base, shift) pairs, all layers */
    int max_layer_deg;    /* largest info degree of a layer */
    int level_steps;      /* sum over layers of their level count */
    int max_levels;       /* largest level count of one layer */
    int dep_nodes;        /* nodes at level > 0 */
    uint16_t* layer_start; /* q + 1 offsets into ent */
    uint32_t* ent;         /* base << 16 | shift */
    uint8_t* level;        /* r: level of node j of layer i at i*360+j */
    uint8_t* nlev;         /* q: level count of each layer */
    /* r: per node, bits 0-7 its level, bits 8-35 the information edges it
     * shares with an earlier node of its layer (read after that node's
     * update: "dependent"), bits 36-63 those it shares with a later node
     * (whose final value that later node writes) */
    uint64_t* meta;
    int msg_words;         /* 32-bit message words per frame: 360 * sum of ceil((deg_i+2)/4) */
    int* msg_off;          /* q: first message word row of each layer (rows of 360 words) */
    int pk_msg_words;      /* the same for two frames, two edges a word: 360 * sum of ceil((deg_i+2)/2) */
    int* pk_msg_off;       /* q: first row of each layer in that layout */
    /* Hand-over of shared bits between the nodes of a layer (the two-frame
     * kernel's rounds).  A bit touched by k > 1 nodes of a layer (a group
     * that appears k times in the layer) passes through k - 1 hand-overs,
     * node order r -> r + 1; hand-over t of a layer uses the slots
     * t * 360 + m (m: the bit's index in its group). */
    uint32_t* xent;        /* nent: ent | (k - 1) << 9 | t0 << 11, t0 the
                              group's first hand-over index (k = 1: 0) */
    uint64_t* pos;         /* r: 2 bits per information edge, the node's
                              rank r among the k nodes touching the bit */
    uint32_t* dupm;        /* q: information edges of the layer in a group
                              with k > 1 */
    int max_xfer;          /* largest number of hand-overs of one layer */
    /* Tables of the split kernel (a lane pair per node, ldpc_sk.h): lane
     * t = 2j + h of LG_LANES takes the node's positions p = 2k + h (0 own
     * parity, 1 previous parity, 2 + r the layer's r-th information edge,
     * those of shared groups first), k < sk_ks.  Per layer and lane an
     * address record of SK_RA(sk_ks) words: word k the LDS byte offset of
     * slot k's posterior (2 x its index; an empty slot: 2n, a spare cell);
     * for layers with nodes sharing bits a tail record of SK_RB(sk_ks)
     * words: the slot masks (dep in bits 0-15, succ in 16-31), then the
     * hand-over cells of slots 0 .. SK_KD, two u16 a word (slot 2w in the
     * low half).  A layer's records are consecutive in lane order. */
    int sk_ks;             /* slots per lane: (largest degree + 2) / 2 rounded up */
    int sk_wide;           /* 1 if a layer has fewer slots than sk_ks - 1 */
    uint32_t* sk_lay;      /* q x 4: positions | levels << 8 | barrier << 16 (the
                              layer waits for the workgroup's previous layers)
                              | kde << 24 (slots 1 .. kde - 1 hold the edges
                              of groups shared by nodes of the layer, the
                              first positions after the parity bits); first
                              message word (of a frame pair's block: per layer
                              and lane SK_WORDS(sk_ks) words); first word of
                              the address records; first word of the tail
                              records (0: none) */
    uint32_t* sk_tab;      /* sk_words words */
    int sk_words;
    int sk_msg_words;      /* message words per frame pair */
} lg_code;

#define LG_MAX_XFER 8      /* hand-overs per layer the kernels provide for */
#define LG_LANES 720       /* lanes of the split kernel: a pair per check node */
/* slots per lane that may hold shared edges (1 .. SK_KD) in a code whose
 * split kernel has ks slots per lane: kde - 1 <= SK_KD(ks) */
#define SK_KD(ks) ((ks) <= 4 ? 1 : (ks) <= 7 ? 3 : (ks) <= 11 ? 5 : 6)
/* words per lane and layer: messages (two slots of both frames a word),
 * address record (padded to whole 16-byte loads), tail record (masks and
 * the cells of slots 0 .. SK_KD, padded likewise) */
#define SK_WORDS(ks) (((ks) + 1) / 2)
/* byte offset of the posteriors in the split kernel's LDS (its control
 * words come first); the address records hold SK_POST_OFF + 2 x index */
#define SK_POST_OFF 32
#define SK_RA(ks) ((ks) <= 2 ? 2 : ((ks) + 3) & ~3)
#define SK_RB(ks) ((1 + SK_KD(ks) / 2 + 1 + 3) & ~3)
#define LG_XENT_SHIFT(e) ((int)((e) & 0x1ffu))
#define LG_XENT_KM1(e) ((int)(((e) >> 9) & 3u))
#define LG_XENT_T0(e) ((int)(((e) >> 11) & 15u))

#define LG_META_LEVEL(m) ((int)((m) & 0xff))
#define LG_META_DEP(m) ((uint32_t)(((m) >> 8) & 0xfffffffu))
#define LG_META_SUCC(m) ((uint32_t)(((m) >> 36) & 0xfffffffu))

int lg_code_count(void);
/* Expanded code, or NULL for an out-of-range id.  Built once, immutable. */
const lg_code* lg_code_get(int id);
/* Code used for (frame size, rate name) as libdvbrx's find_table
 * (ldpc.c:188-197): "1_2", "2_3_t2", "11_45_VLSNR_SF2", ... -1 if none. */
int lg_code_lookup(int short_frame, const char* rate);
/* BCH parameters of drx_fec_lookup (libdvbrx/src/bch.c:25-41); the T2
 * codes "2_3_t2" / "3_5_t2" take those of their DVB-S2 rate. */
int lg_fec_lookup(int short_frame, const char* rate, int* kbch, int* nbch, int* t);
/* A rate name of the BCH row with these parameters (the first), NULL if
 * none. */
const char* lg_fec_rate(int short_frame, int kbch, int nbch, int t);
/* Position of a rate among the BCH rows of its frame size, -1 if none. */
int lg_fec_index(int short_frame, const char* rate);

struct lg_modcod_row;
/* MODCOD row of a PLS code (pilot bit ignored), NULL if unknown. */
const struct lg_modcod_row* lg_modcod_get(int pls);
/* Constellation points, (re, im) pairs, indexed by lg_modcod_row.point_off. */
const float* lg_points(void);

#ifdef __cplusplus
}
#endif

#endif
