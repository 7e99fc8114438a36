/*
 * Layout of the standard data in dvbs2_tables.inc (generated by
 * tables/gen_tables.py).  Shared by the library, the simulator and the
 * oracle; the data itself is EN 302 307-1/-2 and EN 302 755.
 */
#ifndef DVBS2_TABLE_TYPES_H
#define DVBS2_TABLE_TYPES_H

#include <stdint.h>

/* Rows of one degree class of an address table: each row lists the parity
 * accumulators of the first bit of a 360-bit information group. */
struct lg_group {
    uint16_t deg, rows;
};

/* One LDPC code (EN 302 307-1 Annex B/C etc.).  Mirrors the fields of the
 * reference's struct ldpc_table (libdvbrx/src/ldpc.c:20-23). */
struct lg_code_def {
    const char* name; /* e.g. "S2_B4" (rate 1/2 normal) */
    int n, k;         /* codeword / information bits */
    int links_max_cn; /* largest check node degree */
    int links_total;  /* all edges, the 2R-1 parity edges included */
    int deg_max;      /* largest variable node degree */
    int ngroups;      /* degree classes, offset group_off in LG_GROUPS */
    int group_off;
    int addr_off; /* first address in LG_ADDR */
};

/* Rate name (gr-dvbs2rx naming, as drx_fec_lookup) -> code. */
struct lg_rate_map {
    const char* rate;
    int short_frame;
    int code; /* index in LG_CODES */
};

/* BCH parameters (EN 302 307-1 Table 5a/5b and S2X). */
struct lg_fec_row {
    const char* rate;
    int short_frame;
    int kbch, nbch, t;
};

/* MODCOD: PLS code (pilot bit cleared), constellation, bit interleaver
 * column order (order[0] < 0: none) and points by bit label, MSB first. */
struct lg_modcod_row {
    int pls;
    const char* constellation;
    const char* rate;
    int short_frame, bits;
    int order[8];
    int point_off; /* first point (pairs of floats) in LG_POINTS */
};

#endif
