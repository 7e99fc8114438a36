/*
 * Types, bounds and launch entry points shared by the decoder kernels
 * (ldpc_layered_kern.hip, ldpc_pk_kern.hip: one instance per translation
 * unit, compiled in parallel) and the host code of ldpc_gpu.hip.  Not part
 * of the C-ABI.
 */
#ifndef LDPC_KERN_H
#define LDPC_KERN_H

#include <hip/hip_runtime.h>

#include <stddef.h>
#include <stdint.h>

#define NT 384 /* threads per workgroup of the one-frame kernel: 360 check nodes of a layer */
/* the split two-frame kernel: a lane pair per check node */
#define SK_NT 720
#define NCLASS 2 /* degree classes of the one-frame kernel: one instance each (registers grow with the degree) */
/* instances of the split kernel: (slots per lane, wide) */
#define SK_NINST 15
#define SK_INSTS { { 2, 0 }, { 3, 0 }, { 4, 0 }, { 4, 1 }, { 5, 0 }, { 6, 0 }, { 7, 0 }, { 7, 1 }, { 8, 0 }, { 9, 0 }, \
                   { 10, 0 }, { 10, 1 }, { 11, 0 }, { 14, 0 }, { 15, 0 } }
#define NLAUNCH (2 * SK_NINST) /* launches of one decode: kernel instance x frame size */
/* forked streams a context holds for the concurrent launches of a decode
 * (one hardware queue each beside the context's own stream at
 * GPU_MAX_HW_QUEUES 4), and how many a decode uses by default.  Three
 * measured best on the mixed S2 batch (18 launches): 23.5-23.8 ms, against
 * 24.5 with four (two then share a queue), 25.6 with the context's stream
 * as a fourth lane and 28.0 with two. */
#define NAUX 4
#define NAUX_DEFAULT 3

namespace lgk {


struct DCode {
    int n, k, r, q;
    int ent_off;   /* into ent: base << 16 | shift */
    int lay_off;   /* into lay (see lay_deg / lay_nlev / lay_start) */
    int pkl_off;   /* into pkl: the code's q layer descriptors, then two copies of the last */
    int meta_off;  /* into meta: level and dependency masks per node */
    int nent;      /* (base, shift) pairs */
    int dmax;      /* largest information degree of a layer */
    int dmin;      /* smallest */
    int msg_words; /* 32-bit message words per frame */
    int pk_msg_words; /* message words of a frame pair in the split kernel's layout */
    int code_id;   /* index in the registry */
    int max_xfer;  /* hand-overs of its busiest layer */
    int sk_inst;   /* split kernel instance (SK_INSTS index) */
};

struct DFrame {
    int code, max_iter, nbits;
    int min_iter; /* the two-frame kernel checks the syndrome from this iteration on (groups' second pass) */
    unsigned long long llr_off, out_off;
};

struct KArgs {
    const DCode* codes;
    const uint32_t* ent;
    const uint2* lay;
    const uint64_t* meta;
    const DFrame* frames;
    const int* list;   /* frames to run (nullptr: 0..nrun-1) */
    int nrun;
    const int8_t* llr;
    uint8_t* out;
    int32_t* iters;
    uint8_t* synd;
    uint32_t* msg;     /* messages: per workgroup or per frame */
    size_t msg_stride; /* words per message block */
    int msg_per_frame;
    int8_t* state;     /* group mode: posteriors saved per frame (N_max bytes) */
    size_t state_stride;
    int32_t* it_state; /* group mode: iterations done per frame */
    const int32_t* target; /* resume mode: iteration to reach per frame */
    int resume;        /* 0: decode from the LLRs with per-frame early stop;
                          1: continue from the saved state up to target */
    int save;          /* save posteriors + iteration count at the end */
    unsigned* ticket;
};

/* largest table of one code staged in LDS: (base, shift) pairs and layers */
#define MAX_ENT 1024
#define MAX_LAYERS 192
#define SYN_CHUNK 4  /* layers checked between two early-exit votes */
#define WMAX 8       /* message words of a node: (28 + 2 + 3) / 4 */
#define LOW_MAX_DEG 12 /* class 0: layers of information degree 1..12 */
#define HIGH_MIN_DEG 9 /* class 1: 9..28 */

/* layer table entry: x = first (base, shift) pair | degree << 16 | levels << 24,
 * y = first message row (rows of 360 words) */
__device__ __forceinline__ int lay_deg(uint2 ly) { return (ly.x >> 16) & 0xff; }
__device__ __forceinline__ int lay_nlev(uint2 ly) { return ly.x >> 24; }
__device__ __forceinline__ int lay_start(uint2 ly) { return ly.x & 0xffff; }

#define LOW_DEGS(X) X(1) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12)
#define HIGH_DEGS(X)                                                                                        \
    X(9) X(10) X(11) X(12) X(13) X(14) X(15) X(16) X(17) X(18) X(19) X(20) X(21) X(22) X(23) X(24) X(25) X(26) \
        X(27) X(28)

/* ---- two frames per workgroup: packed 16-bit arithmetic (ldpc_pk.h) ----- */


struct PkArgs {
    const DCode* codes;
    size_t hoff;           /* byte offset of the hand-over buffers in LDS */
    int hstride;           /* u16 cells per hand-over buffer */
    const uint4* pkl;      /* per layer: positions | levels << 8 | ..., byte offsets of its messages,
                              address records and tail records (ldpc_sk_kern.hip) */
    const uint32_t* adt;   /* lane records of every code (lg_code.sk_tab) */
    const DFrame* frames;
    const int* list;       /* frame indices, grouped by code into segments */
    const int2* seg;       /* per segment: first list entry, length */
    int nseg;
    unsigned* cnt;         /* per segment: frames claimed so far */
    const int8_t* llr;
    uint8_t* out;
    int32_t* iters;
    uint8_t* synd;
    uint32_t* msg;         /* message block per workgroup */
    size_t msg_stride;
    /* learned mode: parameter rows (row 0: the plain decoder's S = 16,
     * B = 1), per code its first row and row count, plain fallback cap,
     * per-frame fallback flag (may be NULL) */
    const uint8_t* nS;
    const int16_t* nB;
    const int* nrow0;
    const int* nT;
    int fallback;
    uint8_t* fell;
};

/* parameter rows of the two slots for one iteration (learned mode) */
struct PkRowPtrs {
    const uint8_t* S0;
    const uint8_t* S1;
    const int16_t* B0;
    const int16_t* B1;
};

#define NCOUNTERS 256 /* work counters: tickets and per-segment claims */

/* launches (on stream st) and occupancy queries of the kernel instances */
hipError_t layered_launch(int cls, int grid, size_t lds, hipStream_t st, const KArgs& a);
hipError_t layered_occupancy(int cls, size_t lds, int* per_cu);
hipError_t reset_counters(hipStream_t st, unsigned* cnt);
/* the split kernel instance k (SK_INSTS[k]), plain or learned, built for
 * occ (1 or 2) workgroups per CU; occ 2 exists for the plain instances
 * SK_ALL2 lists (sk_has_occ2) */
hipError_t pk_launch(int k, bool nms, int occ, int grid, size_t lds, hipStream_t st, const PkArgs& a);
hipError_t pk_occupancy(int k, bool nms, int occ, size_t lds, int* per_cu);
bool sk_has_occ2(int k, bool nms);
#define SK_DECL(KS, W, N)                                                                          \
    hipError_t sk_launch_##KS##_##W##_##N(int grid, size_t lds, hipStream_t st, const PkArgs& a);   \
    hipError_t sk_occupancy_##KS##_##W##_##N(size_t lds, int* per_cu);
#define SK_ALL(X)                                                                                   \
    X(2, 0, 0) X(3, 0, 0) X(4, 0, 0) X(4, 1, 0) X(5, 0, 0) X(6, 0, 0) X(7, 0, 0) X(7, 1, 0) X(8, 0, 0) X(9, 0, 0)  \
    X(10, 0, 0) X(10, 1, 0) X(11, 0, 0) X(14, 0, 0) X(15, 0, 0) X(2, 0, 1) X(3, 0, 1) X(4, 0, 1) X(4, 1, 1)       \
    X(5, 0, 1) X(6, 0, 1) X(7, 0, 1) X(7, 1, 1) X(8, 0, 1) X(9, 0, 1) X(10, 0, 1) X(10, 1, 1) X(11, 0, 1)         \
    X(14, 0, 1) X(15, 0, 1)
SK_ALL(SK_DECL)
/* the instances also built for two workgroups per CU (short frames; their
 * registers fit 6 waves per SIMD without spilling) */
#define SK_DECL2(KS, W, N)                                                                         \
    hipError_t sk_launch_##KS##_##W##_##N##_o2(int grid, size_t lds, hipStream_t st, const PkArgs& a); \
    hipError_t sk_occupancy_##KS##_##W##_##N##_o2(size_t lds, int* per_cu);
#define SK_ALL2(X) X(2, 0, 0) X(3, 0, 0) X(4, 0, 0) X(4, 1, 0)
SK_ALL2(SK_DECL2)

} // namespace lgk

#endif
