/*
 * MI355X (gfx950) layered offset min-sum LDPC decoder: kernels and the
 * C-ABI of include/ldpc_gpu.h.
 *
 * Arithmetic (bit-exact with libdvbrx/src/ldpc.c:321-362): for a check node
 * with edges e, posteriors L_e and stored messages m_e,
 *   x_e  = sat8(L_e - m_e)            (first iteration: x_e = L_e)
 *   a_e  = |x_e| in 0..128, min0 <= min1 the two smallest a_e (with
 *          multiplicity), s = xor of the signs of x_e
 *   mn0  = max(min(min0, 127) - 1, 0), mn1 likewise       (offset 1)
 *   o_e  = (s ^ sign x_e) ? -mag : mag, mag = a_e == min0 ? mn1 : mn0
 *   L_e <- sat8(x_e + o_e),  m_e <- clamp(o_e, -32, 31)
 * Messages are stored clamped to [-32, 31] as the reference stores them
 * (one int8 per edge; the compressed store of ldpc.c:376-462 holds the
 * same values).
 *
 * Two kernels, both persistent grids over a frame table:
 *
 * ldpc_layered_kernel: one frame per workgroup.  The frame's N posteriors
 * live in LDS: information bits 0..K-1, then the R parity bits in layer
 * order (parity of check q*j+i at K + 360*i + j, ldpc.c:624-626).  Thread
 * j processes node j of every layer; the messages of node (i, j) are only
 * touched by thread j and live in global memory at rows of 360 words
 * (4 edges per word), coalesced across the wavefronts.  Within a layer the
 * nodes sharing a variable run in dependency levels (ldpc_codes.h) so the
 * result equals the reference's sequential node order.  Serves batches
 * with a shared early stop (groups) and resumable passes.
 *
 * ldpc_pk_kernel: two frames per workgroup in the two bytes of 16-bit
 * lanes (ldpc_pk.h: value y held as y*256, so 16-bit saturating packed
 * arithmetic is the reference's int8 saturation), each slot stopping on
 * its own syndrome and refilled from the frames of the same code while
 * the other slot runs.  Frames with their own early stop take this path.
 */
#include <hip/hip_runtime.h>

#include <type_traits>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <vector>

#include "ldpc_codes.h"
#include "ldpc_gpu.h"
#include "ldpc_nms.h"
#include "ldpc_node.h"
#include "ldpc_pk.h"

#define NT 384 /* threads per workgroup of the one-frame kernel: 360 check nodes of a layer */
/* the two-slot kernel: exactly one thread per check node, so no thread of
 * a layer is idle and the node code needs no lane test (the sixth wave has
 * 40 lanes) */
#define PK_NT 360
#define NCLASS 2 /* degree classes: one kernel instance each (registers grow with the degree) */
#define NLAUNCH 4 /* launches of one decode: degree class x frame size */
static const int CLASS_DI[NCLASS] = { 12, 28 }; /* largest layer degree of each class */

namespace {

struct DCode {
    int n, k, r, q;
    int ent_off;   /* into ent: base << 16 | shift */
    int lay_off;   /* into lay (see lay_deg / lay_nlev / lay_start) */
    int meta_off;  /* into meta: level and dependency masks per node */
    int nent;      /* (base, shift) pairs */
    int dmax;      /* largest information degree of a layer */
    int dmin;      /* smallest */
    int msg_words; /* 32-bit message words per frame */
    int pk_msg_words; /* the same for the two-frame layout */
    int code_id;   /* index in the registry */
};

struct DFrame {
    int code, max_iter, nbits, pad;
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

/* Syndrome (bad(), ldpc.c:286-305) over chunks of layers, stopping at the
 * first chunk where some check fails: an unconverged frame usually fails
 * in the first chunk. */
template <int CLS>
__device__ __forceinline__ int syndrome_ok(const int8_t* sm, const uint32_t* s_ent, const uint2* s_lay, int j, int K,
                                           int q)
{
    for (int i0 = 0; i0 < q; i0 += SYN_CHUNK) {
        int ok = 1;
        if (j < 360) {
            const int i1 = min(i0 + SYN_CHUNK, q);
            for (int i = i0; i < i1; i++) {
                const uint2 ly = s_lay[i];
                const uint32_t* e = s_ent + lay_start(ly);
#define X(D)                                                   \
    case D:                                                    \
        ok &= lg_node_check_t<D>(sm, e, D, i, j, K, q);        \
        break;
                if constexpr (CLS == 0) {
                    switch (lay_deg(ly)) {
                        LOW_DEGS(X)
                    default:
                        ok = 0;
                    }
                } else {
                    switch (lay_deg(ly)) {
                        HIGH_DEGS(X)
                    default:
                        ok = 0;
                    }
                }
#undef X
            }
        }
        if (!__syncthreads_and(ok))
            return 0;
    }
    return 1;
}

/* One layer by the workgroup: phases A, S (per level), C of ldpc_node.h.
 * mw: the node's message words (zero in the first iteration). */
template <int DEG>
__device__ __forceinline__ void layer_t(int8_t* sm, const uint32_t* ent, int i, int j, int K, int q, int nlev,
                                        const uint32_t* mw, uint64_t meta, uint32_t* mrow, int first)
{
    const bool act = j < 360;
    const uint32_t dep = LG_META_DEP(meta), succ = LG_META_SUCC(meta);
    const int lvl = LG_META_LEVEL(meta);
    LgNode<DEG> n;
    if (act) {
        lg_phase_a<DEG>(n, sm, ent, i, j, K, q, dep, mw, first);
        if (nlev > 1 && lvl == 0)
            lg_write_succ<DEG>(n, sm, succ);
    }
    for (int l = 1; l < nlev; l++) {
        __syncthreads();
        if (act && lvl == l) {
            lg_step<DEG>(n, sm, dep, mw, first);
            lg_write_succ<DEG>(n, sm, succ);
        }
    }
    if (act) {
        uint32_t o[LG_MSG_WORDS(DEG)];
        lg_phase_c<DEG>(n, sm, succ, o);
#pragma unroll
        for (int w = 0; w < LG_MSG_WORDS(DEG); w++)
            mrow[360 * w + j] = o[w];
    }
}

/*
 * One decode (from the LLRs or resumed) of one frame by the workgroup.
 * Returns the syndrome flag; *itp holds the iterations run.  The messages
 * and dependency masks of layer i+1 are fetched while layer i runs.
 */
template <int CLS>
__device__ __forceinline__ int decode_frame(int8_t* sm, const uint32_t* s_ent, const uint2* s_lay,
                                            const uint64_t* meta, uint32_t* msg, int K, int q, int* itp, int target,
                                            int early_stop)
{
    const int j = threadIdx.x;
    const bool act = j < 360;
    int it = *itp, ok;
    for (;;) {
        if (early_stop || it >= target) {
            ok = syndrome_ok<CLS>(sm, s_ent, s_lay, j, K, q);
            if (ok || it >= target)
                break;
        }
        const int first = it == 0;
        uint32_t nx[WMAX];
        uint64_t nmeta = 0;
#pragma unroll
        for (int w = 0; w < WMAX; w++)
            nx[w] = 0;
        {
            const uint2 ly = s_lay[0];
            const int W = LG_MSG_WORDS(lay_deg(ly));
            if (act && !first) {
#pragma unroll
                for (int w = 0; w < WMAX; w++)
                    if (w < W)
                        nx[w] = msg[360 * (ly.y + w) + j];
            }
            if (act && lay_nlev(ly) > 1)
                nmeta = meta[j];
        }
        for (int i = 0; i < q; i++) {
            uint32_t mw[WMAX];
#pragma unroll
            for (int w = 0; w < WMAX; w++)
                mw[w] = nx[w];
            const uint64_t mt = nmeta;
            const uint2 ly = s_lay[i];
            if (i + 1 < q) {
                const uint2 ln = s_lay[i + 1];
                const int W = LG_MSG_WORDS(lay_deg(ln));
                if (act && !first) {
#pragma unroll
                    for (int w = 0; w < WMAX; w++)
                        if (w < W)
                            nx[w] = msg[360 * (ln.y + w) + j];
                }
                nmeta = act && lay_nlev(ln) > 1 ? meta[360 * (i + 1) + j] : 0;
            }
            const uint32_t* e = s_ent + lay_start(ly);
            uint32_t* mrow = msg + 360 * ly.y;
            const int nlev = lay_nlev(ly);
#define X(D)                                                            \
    case D:                                                             \
        layer_t<D>(sm, e, i, j, K, q, nlev, mw, mt, mrow, first);       \
        break;
            if constexpr (CLS == 0) {
                switch (lay_deg(ly)) {
                    LOW_DEGS(X)
                }
            } else {
                switch (lay_deg(ly)) {
                    HIGH_DEGS(X)
                }
            }
#undef X
            __syncthreads();
        }
        it++;
    }
    *itp = it;
    return ok;
}

#define NCOUNTERS 256 /* work counters: tickets and per-segment claims */

__global__ void __launch_bounds__(256) reset_ticket(unsigned* t)
{
    t[threadIdx.x] = 0;
}

template <int CLS>
__global__ void __launch_bounds__(NT) ldpc_layered_kernel(KArgs a)
{
    extern __shared__ __align__(16) int8_t sm[];
    __shared__ int s_run;
    __shared__ uint32_t s_ent[MAX_ENT];
    __shared__ uint2 s_lay[MAX_LAYERS];
    const int tid = threadIdx.x;

    for (;;) {
        if (tid == 0)
            s_run = (int)atomicAdd(a.ticket, 1u);
        __syncthreads();
        const int run = s_run;
        __syncthreads();
        if (run >= a.nrun)
            break;
        const int f = a.list ? a.list[run] : run;
        const DFrame fr = a.frames[f];
        const DCode cd = a.codes[fr.code];
        const uint32_t* ent = a.ent + cd.ent_off;
        const uint2* lay = a.lay + cd.lay_off;
        const uint64_t* meta = a.meta + cd.meta_off;
        const int K = cd.k, N = cd.n, R = cd.r, q = cd.q;
        uint32_t* msg = a.msg + a.msg_stride * (size_t)(a.msg_per_frame ? f : blockIdx.x);

        int it;
        int target;
        if (!a.resume) {
            /* info bits straight, parity transposed to layer order */
            const uint2* src = (const uint2*)(a.llr + fr.llr_off);
            for (int v = tid; v < K / 8; v += NT)
                ((uint2*)sm)[v] = src[v];
            const int8_t* par = a.llr + fr.llr_off + K;
            for (int t = tid; t < R; t += NT) {
                const int jj = t / q, ii = t - jj * q;
                sm[K + 360 * ii + jj] = par[t];
            }
            it = 0;
            target = fr.max_iter;
        } else {
            const uint2* src = (const uint2*)(a.state + a.state_stride * (size_t)f);
            for (int v = tid; v < N / 8; v += NT)
                ((uint2*)sm)[v] = src[v];
            it = a.it_state[f];
            target = a.target[f];
        }
        /* the code's layer table and (base, shift) pairs, read by every
         * node of the frame */
        for (int t = tid; t < cd.nent; t += NT)
            s_ent[t] = ent[t];
        for (int t = tid; t < q; t += NT)
            s_lay[t] = lay[t];
        __syncthreads();

        const int ok = decode_frame<CLS>(sm, s_ent, s_lay, meta, msg, K, q, &it, target, !a.resume);

        /* hard decisions of the first nbits bits, MSB first, 4 bytes a thread */
        const int nbytes = (fr.nbits + 7) / 8;
        uint8_t* o = a.out + fr.out_off;
        for (int b4 = 4 * tid; b4 < nbytes; b4 += 4 * NT) {
            uint32_t v = 0;
            for (int k = 0; k < 32; k++) {
                const int bit = 8 * b4 + k;
                v = v << 1 | (uint32_t)(bit < fr.nbits && sm[bit] < 0);
            }
            for (int k = 0; k < 4 && b4 + k < nbytes; k++)
                o[b4 + k] = (uint8_t)(v >> (24 - 8 * k));
        }
        if (a.save) {
            uint2* dst = (uint2*)(a.state + a.state_stride * (size_t)f);
            for (int v = tid; v < N / 8; v += NT)
                dst[v] = ((const uint2*)sm)[v];
        }
        if (tid == 0) {
            a.iters[f] = it;
            a.synd[f] = (uint8_t)ok;
            if (a.save)
                a.it_state[f] = it;
        }
        __syncthreads();
    }
}

/* ---- two frames per workgroup: packed 16-bit arithmetic (ldpc_pk.h) ----- */

#define PK_WMAX 15 /* message words of a node: (28 + 2 + 1) / 2 */

struct PkArgs {
    const DCode* codes;
    const uint32_t* ent;
    const uint2* lay;
    const uint32_t* pkoff; /* per layer: first message row in the two-frame layout */
    const uint64_t* meta;
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

/* AND of a 2-bit value over the workgroup, one barrier: three vote words
 * in turn, the next one reset before the barrier of the current vote */
__device__ __forceinline__ uint32_t wg_and2(uint32_t v, unsigned* s_vote, int& n)
{
    const int slot = n % 3;
    n++;
    const uint64_t f0 = __ballot(!(v & 1)), f1 = __ballot(!(v & 2));
    if ((threadIdx.x & 63) == 0)
        atomicAnd(&s_vote[slot], (f0 ? 0u : 1u) | (f1 ? 0u : 2u));
    if (threadIdx.x == 0)
        s_vote[(slot + 1) % 3] = 3;
    __syncthreads();
    return s_vote[slot];
}

/* syndrome of both slots, chunks of layers with an early exit once every
 * slot that is still decoding has failed a check; bit k: slot k holds */
template <int CLS>
__device__ __forceinline__ uint32_t pk_syndrome(const uint16_t* sm, const uint32_t* ent, const uint2* lay, int j,
                                                int K, int q, uint32_t live, unsigned* s_vote, int& nvote)
{
    uint32_t good = 3;
    for (int i0 = 0; i0 < q; i0 += SYN_CHUNK) {
        uint32_t ok = 3;
        if (j < 360) {
            const int i1 = min(i0 + SYN_CHUNK, q);
            for (int i = i0; i < i1; i++) {
                const uint2 ly = lay[i];
                const uint32_t* e = ent + lay_start(ly);
#define X(D)                                             \
    case D:                                              \
        ok &= pk_check<D>(sm, e, i, j, K, q);            \
        break;
                if constexpr (CLS == 0) {
                    switch (lay_deg(ly)) {
                        LOW_DEGS(X)
                    default:
                        ok = 0;
                    }
                } else {
                    switch (lay_deg(ly)) {
                        HIGH_DEGS(X)
                    default:
                        ok = 0;
                    }
                }
#undef X
            }
        }
        good &= wg_and2(ok, s_vote, nvote);
        if (!(good & live))
            break;
    }
    return good;
}

template <int DEG, bool TAIL, bool NMS>
__device__ __forceinline__ void pk_layer_t(uint16_t* sm, const uint32_t* ent, int i, int j, int K, int q, int nlev,
                                           const uint32_t* mw, uint64_t meta, uint32_t* mrow, uint32_t se, uint32_t so,
                                           lg_pk ps, lg_pk pb)
{
    const bool act = j < 360;
    const uint32_t dep = TAIL ? LG_META_DEP(meta) : 0, succ = TAIL ? LG_META_SUCC(meta) : 0;
    const int lvl = LG_META_LEVEL(meta);
    PkNode<DEG> n;
    if (act) {
        pk_phase_a<DEG, TAIL>(n, sm, ent, i, j, K, q, dep, mw, se, so);
        if (TAIL && lvl == 0)
            pk_write_succ<DEG, NMS>(n, sm, succ, ps, pb);
    }
    if (TAIL) {
        for (int l = 1; l < nlev; l++) {
            __syncthreads();
            if (act && lvl == l) {
                pk_step<DEG>(n, sm, dep, mw, se, so);
                pk_write_succ<DEG, NMS>(n, sm, succ, ps, pb);
            }
        }
    }
    if (act) {
        uint32_t o[LG_PK_WORDS(DEG)];
        pk_phase_c<DEG, TAIL, NMS>(n, sm, succ, o, ps, pb);
#pragma unroll
        for (int w = 0; w < LG_PK_WORDS(DEG); w++)
            mrow[360 * w + j] = o[w];
    }
}

/* one iteration (all layers) on both slots; the code's tables are read
 * through uniform pointers (scalar loads) */
template <int CLS, bool NMS>
__device__ __forceinline__ void pk_iteration(uint16_t* sm, const uint32_t* ent, const uint2* lay,
                                             const uint32_t* pkoff, const uint64_t* meta, uint32_t* msg, int K,
                                             int q, int first0, int first1, PkRowPtrs rows)
{
    const int j = threadIdx.x;
    const bool act = j < 360;
    const uint32_t se = pk_msg_sel(0, first0, first1), so = pk_msg_sel(1, first0, first1);
    const bool load = !(first0 && first1);
    constexpr int WM = CLS == 0 ? LG_PK_WORDS(LOW_MAX_DEG) : PK_WMAX;
    /* messages and masks of layer i+1 are fetched while layer i runs */
    uint32_t nx[WM];
    uint64_t nm = 0;
#pragma unroll
    for (int w = 0; w < WM; w++)
        nx[w] = 0;
    auto fetch = [&](int l) {
        if (l < q && act) {
            const uint2 ln = lay[l];
            const int W = LG_PK_WORDS(lay_deg(ln));
            if (load) {
                const uint32_t* row = msg + 360 * pkoff[l] + j;
#pragma unroll
                for (int w = 0; w < WM; w++)
                    if (w < W)
                        nx[w] = row[360 * w];
            }
            nm = lay_nlev(ln) > 1 ? meta[360 * l + j] : 0;
        }
    };
    fetch(0);
    /* one layer with its degree D fixed */
    auto layer = [&](auto dtag, int i) {
        constexpr int D = decltype(dtag)::value;
        uint32_t mw[WM];
#pragma unroll
        for (int w = 0; w < WM; w++)
            mw[w] = nx[w];
        const uint64_t mt = nm;
        fetch(i + 1);
        const uint2 ly = lay[i];
        const uint32_t* e = ent + lay_start(ly);
        uint32_t* mrow = msg + 360 * pkoff[i];
        const int nlev = lay_nlev(ly);
        /* learned scale / offset of the layer's check degree, per slot;
         * node (0, 0) has no previous parity, one edge less */
        lg_pk ps = pk_dup(16), pb = pk_dup(16);
        if constexpr (NMS) {
            const int dc = D + 2 - (i == 0 && j == 0);
            ps = lg_pk{ (short)rows.S0[dc], (short)rows.S1[dc] };
            pb = lg_pk{ (short)rows.B0[dc], (short)rows.B1[dc] };
        }
        if (nlev > 1)
            pk_layer_t<D, true, NMS>(sm, e, i, j, K, q, nlev, mw, mt, mrow, se, so, ps, pb);
        else
            pk_layer_t<D, false, NMS>(sm, e, i, j, K, q, nlev, mw, mt, mrow, se, so, ps, pb);
        __syncthreads();
    };
    /* a layer of no degree of the class has no node to run */
    auto skip = [&](int i) {
        fetch(i + 1);
        __syncthreads();
    };
    if constexpr (CLS == 0) {
        /* low degrees: layers in runs of one degree, dispatched once per
         * run (the per-layer dispatch is a large share of these layers'
         * scalar work) */
        for (int i = 0; i < q;) {
#define X(D)                                                  \
    case D:                                                   \
        do                                                    \
            layer(std::integral_constant<int, D>{}, i++);     \
        while (i < q && lay_deg(lay[i]) == D);                \
        break;
            switch (lay_deg(lay[i])) {
                LOW_DEGS(X)
            default:
                skip(i++);
            }
#undef X
        }
    } else {
        /* high degrees: one dispatch per layer (runs would spill) */
        for (int i = 0; i < q; i++) {
#define X(D)                                          \
    case D:                                           \
        layer(std::integral_constant<int, D>{}, i);   \
        break;
            switch (lay_deg(lay[i])) {
                HIGH_DEGS(X)
            default:
                skip(i);
            }
#undef X
        }
    }
}

/*
 * Persistent two-slot decoder.  Frames of the launch are grouped by code
 * into segments; a workgroup works on one segment at a time, and whenever
 * a slot's frame is done (syndrome holds or max_iter reached) the slot is
 * refilled with the next frame of the same segment while the other slot
 * continues.  When the segment is exhausted and both slots are empty, the
 * workgroup moves to another segment with frames left, or exits.
 */
template <int CLS, bool NMS>
__global__ void __launch_bounds__(PK_NT) ldpc_pk_kernel(PkArgs a)
{
    extern __shared__ __align__(16) uint16_t sm16[];
    __shared__ unsigned s_vote[3];
    __shared__ int s_claim[2], s_seg, s_newseg;
    const int tid = threadIdx.x;
    if (tid < 3)
        s_vote[tid] = 3;
    int nvote = 0;
    int seg = -1;                 /* current segment (uniform) */
    int frame[2] = { -1, -1 };    /* frame in each slot, -1: empty */
    int it[2] = { 0, 0 }, maxit[2] = { 0, 0 };
    int phase[2] = { 0, 0 }, lit[2] = { 0, 0 }; /* learned mode: 1 = plain fallback; learned iterations */
    uint32_t* msg = a.msg + a.msg_stride * (size_t)blockIdx.x;
    DCode cd;
    const uint64_t* meta = nullptr;

    for (;;) {
        /* ---- claim frames for the empty slots ---- */
        if (tid == 0) {
            int cl[2] = { -1, -1 }, sg = seg, newseg = 0;
            for (int s = 0; s < 2; s++) {
                if (frame[s] >= 0 || sg < 0)
                    continue;
                const unsigned t = atomicAdd(&a.cnt[sg], 1u);
                if ((int)t < a.seg[sg].y)
                    cl[s] = a.list[a.seg[sg].x + t];
            }
            if (frame[0] < 0 && frame[1] < 0 && cl[0] < 0 && cl[1] < 0) {
                /* both slots empty, segment done: take another one */
                const int start = sg < 0 ? (int)(blockIdx.x % (unsigned)a.nseg) : sg + 1;
                for (int k = 0; k < a.nseg && cl[0] < 0; k++) {
                    const int g = (start + k) % a.nseg;
                    if (__hip_atomic_load(&a.cnt[g], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >= (unsigned)a.seg[g].y)
                        continue;
                    for (int s = 0; s < 2; s++) {
                        const unsigned t = atomicAdd(&a.cnt[g], 1u);
                        if ((int)t < a.seg[g].y) {
                            cl[s] = a.list[a.seg[g].x + t];
                            sg = g;
                            newseg = 1;
                        }
                    }
                }
            }
            s_claim[0] = cl[0];
            s_claim[1] = cl[1];
            s_seg = sg;
            s_newseg = newseg;
        }
        __syncthreads();
        const int cl0 = __builtin_amdgcn_readfirstlane(s_claim[0]);
        const int cl1 = __builtin_amdgcn_readfirstlane(s_claim[1]);
        const int newseg = __builtin_amdgcn_readfirstlane(s_newseg);
        seg = __builtin_amdgcn_readfirstlane(s_seg);
        __syncthreads(); /* s_claim is rewritten only after this */
        if (cl0 >= 0)
            frame[0] = cl0;
        if (cl1 >= 0)
            frame[1] = cl1;
        if (frame[0] < 0 && frame[1] < 0)
            break;
        const int fany = frame[0] >= 0 ? frame[0] : frame[1];
        if (newseg) {
            cd = a.codes[a.frames[fany].code];
            meta = a.meta + cd.meta_off;
        }
        const uint32_t* ent = a.ent + cd.ent_off;
        const uint2* lay = a.lay + cd.lay_off;
        const uint32_t* pkoff = a.pkoff + cd.lay_off;
        const int K = cd.k, R = cd.r, q = cd.q;
        /* a frame's LLRs into its slot bytes: information bits in order,
         * parity bit K + q*jj + ii at K + 360*ii + jj (layer order) */
        auto load_slot = [&](int s, int f) {
            const DFrame fr = a.frames[f];
            const uint32_t* src = (const uint32_t*)(a.llr + fr.llr_off);
            uint8_t* sb = (uint8_t*)sm16 + s;
            for (int v4 = tid; v4 < K / 4; v4 += PK_NT) {
                const uint32_t w = src[v4];
                for (int b = 0; b < 4; b++)
                    sb[2 * (4 * v4 + b)] = (uint8_t)(w >> (8 * b));
            }
            const uint8_t* par = (const uint8_t*)(a.llr + fr.llr_off + K);
            for (int t = tid; t < R; t += PK_NT) {
                const int jj = t / q, ii = t - jj * q;
                sb[2 * (K + 360 * ii + jj)] = par[t];
            }
        };
        /* ---- load the new frames into their slot bytes ---- */
        for (int s = 0; s < 2; s++) {
            const int f = s ? cl1 : cl0;
            if (f < 0)
                continue;
            it[s] = 0;
            maxit[s] = a.frames[f].max_iter;
            phase[s] = 0;
            lit[s] = 0;
            load_slot(s, f);
        }
        __syncthreads();

        /* ---- iterate until a slot is done ---- */
        for (;;) {
            const uint32_t live = (frame[0] >= 0 ? 1u : 0u) | (frame[1] >= 0 ? 2u : 0u);
            const uint32_t good = pk_syndrome<CLS>(sm16, ent, lay, tid, K, q, live, s_vote, nvote);
            uint32_t done = 0;
            for (int s = 0; s < 2; s++) {
                const int cap = NMS && phase[s] ? a.fallback : maxit[s];
                if (frame[s] >= 0 && (((good >> s) & 1) || it[s] >= cap))
                    done |= 1u << s;
            }
            if (done) {
                uint32_t reload = 0;
                for (int s = 0; s < 2; s++) {
                    if (!((done >> s) & 1))
                        continue;
                    if constexpr (NMS) {
                        /* learned iterations failed: decode the frame again
                         * from its LLRs with the plain arithmetic */
                        if (!((good >> s) & 1) && phase[s] == 0 && a.fallback > 0) {
                            reload |= 1u << s;
                            continue;
                        }
                    }
                    const DFrame fr = a.frames[frame[s]];
                    const int nbytes = (fr.nbits + 7) / 8;
                    uint8_t* o = a.out + fr.out_off;
                    const uint8_t* sb = (const uint8_t*)sm16 + s;
                    for (int b4 = 4 * tid; b4 < nbytes; b4 += 4 * PK_NT) {
                        uint32_t v = 0;
                        for (int k = 0; k < 32; k++) {
                            const int bit = 8 * b4 + k;
                            v = v << 1 | (uint32_t)(bit < fr.nbits && (int8_t)sb[2 * bit] < 0);
                        }
                        for (int k = 0; k < 4 && b4 + k < nbytes; k++)
                            o[b4 + k] = (uint8_t)(v >> (24 - 8 * k));
                    }
                    if (tid == 0) {
                        a.iters[frame[s]] = lit[s] + it[s];
                        a.synd[frame[s]] = (uint8_t)((good >> s) & 1);
                        if (NMS && a.fell)
                            a.fell[frame[s]] = (uint8_t)phase[s];
                    }
                    frame[s] = -1;
                }
                __syncthreads(); /* slot bytes read before a refill overwrites them */
                if constexpr (NMS) {
                    if (reload) {
                        for (int s = 0; s < 2; s++) {
                            if (!((reload >> s) & 1))
                                continue;
                            lit[s] = it[s];
                            it[s] = 0;
                            phase[s] = 1;
                            load_slot(s, frame[s]);
                        }
                        __syncthreads();
                    }
                    if (reload == done)
                        continue; /* no slot emptied: keep iterating */
                }
                break;
            }
            PkRowPtrs rows;
            if constexpr (NMS) {
                const int r0c = a.nrow0[cd.code_id], tc = max(a.nT[cd.code_id], 1);
                int row[2];
                for (int s = 0; s < 2; s++)
                    row[s] = frame[s] >= 0 && phase[s] == 0 ? r0c + min(it[s], tc - 1) : 0;
                rows.S0 = a.nS + LDPC_NMS_DEGS * row[0];
                rows.S1 = a.nS + LDPC_NMS_DEGS * row[1];
                rows.B0 = a.nB + LDPC_NMS_DEGS * row[0];
                rows.B1 = a.nB + LDPC_NMS_DEGS * row[1];
            } else {
                rows = PkRowPtrs{ nullptr, nullptr, nullptr, nullptr };
            }
            pk_iteration<CLS, NMS>(sm16, ent, lay, pkoff, meta, msg, K, q, frame[0] < 0 || it[0] == 0,
                                   frame[1] < 0 || it[1] == 0, rows);
            for (int s = 0; s < 2; s++)
                it[s] += frame[s] >= 0;
        }
    }
}

/* ---- device-resident code tables (one copy per device) ------------------ */

struct DeviceTables {
    int ready = 0;
    DCode* codes = nullptr;
    uint32_t* ent = nullptr;
    uint2* lay = nullptr;
    uint32_t* pkoff = nullptr;
    uint64_t* meta = nullptr;
    std::vector<DCode> h_codes;
};

std::mutex g_tables_lock;
DeviceTables g_tables[64];

} // namespace

struct ldpc_gpu_ctx {
    int device;
    hipStream_t stream;
    hipStream_t aux[NLAUNCH]; /* forked streams of concurrent launches */
    char err[256];
    DeviceTables* tab;
    int ncu;
    /* scratch */
    unsigned* ticket;
    DFrame* d_frames;
    size_t frames_bytes;
    int* d_list;
    size_t list_bytes;
    uint32_t* msg;
    size_t msg_bytes;
    int8_t* state;
    size_t state_bytes;
    int32_t* it_state;
    size_t it_state_bytes;
    int32_t* target;
    size_t target_bytes;
    int2* d_seg;
    size_t seg_bytes;
    hipEvent_t ev0, ev1;
    hipEvent_t ev_cls[2 * NLAUNCH];
    int cls_frames[NLAUNCH]; /* frames of each launch of the last decode */
    int cls_di[NLAUNCH];     /* its degree class bound */
    int timed;
    int force_scalar;        /* per-frame batches through the one-frame kernel */
    int serial;              /* launches of a decode one after the other on the context's stream */
    /* learned tables: row 0 plain (S = 16, B = 1), then per code its rows */
    uint8_t* h_nS;
    int16_t* h_nB;
    int* h_nrow0;
    int* h_nT;
    uint8_t* d_nS;
    int16_t* d_nB;
    int* d_nrow0;
    int* d_nT;
    int nms_dirty;
};

/* A prepared batch (ldpc_gpu_plan_*): its own copies of the frame table
 * and launch layout, its own scratch, and the launches captured as one
 * graph; variants of the launch shape for the autotuner. */
struct ldpc_gpu_plan {
    ldpc_gpu_ctx* c;
    int nframes;
    std::vector<DFrame> hf;
    const int8_t* d_llr;
    uint8_t* d_out;
    int32_t* d_iters;
    uint8_t* d_synd;
    DFrame* d_frames;
    int* d_list;
    int2* d_seg;
    unsigned* d_cnt;
    uint32_t* d_msg;
    size_t list_cap, seg_cap, msg_cap;
    hipStream_t aux[NLAUNCH];
    hipEvent_t fork, join[NLAUNCH], t0, t1;
    hipGraph_t graph;
    hipGraphExec_t exec;
    int variant;
    float variant_ms[8];
    int launched;
};

static int set_err(ldpc_gpu_ctx* c, int code, const char* fmt, ...)
{
    if (c) {
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(c->err, sizeof(c->err), fmt, ap);
        va_end(ap);
    }
    return code;
}

/* for the library's other translation units (ldpc_internal.h) */
__attribute__((visibility("hidden"))) int lg_ctx_device(ldpc_gpu_ctx* c) { return c->device; }
__attribute__((visibility("hidden"))) int lg_ctx_fail(ldpc_gpu_ctx* c, int code, const char* msg)
{
    return set_err(c, code, "%s", msg);
}

#define HIPCHK(ctx, call)                                                                       \
    do {                                                                                        \
        hipError_t e_ = (call);                                                                 \
        if (e_ != hipSuccess)                                                                   \
            return set_err((ctx), LDPC_GPU_EDEVICE, "%s: %s", #call, hipGetErrorString(e_));   \
    } while (0)

static int upload_tables(ldpc_gpu_ctx* c)
{
    std::lock_guard<std::mutex> lk(g_tables_lock);
    if (c->device < 0 || c->device >= 64)
        return set_err(c, LDPC_GPU_EINVAL, "device %d out of range", c->device);
    DeviceTables& t = g_tables[c->device];
    c->tab = &t;
    if (t.ready)
        return LDPC_GPU_OK;
    const int nc = lg_code_count();
    std::vector<DCode> codes(nc);
    std::vector<uint32_t> ent;
    std::vector<uint2> lay;
    std::vector<uint32_t> pkoff;
    std::vector<uint64_t> meta;
    for (int id = 0; id < nc; id++) {
        const lg_code* cd = lg_code_get(id);
        if (!cd)
            return set_err(c, LDPC_GPU_EINVAL, "code %d failed to build", id);
        DCode& d = codes[id];
        d.n = cd->n;
        d.k = cd->k;
        d.r = cd->r;
        d.q = cd->q;
        d.ent_off = (int)ent.size();
        d.lay_off = (int)lay.size();
        d.meta_off = (int)meta.size();
        d.nent = cd->nent;
        d.dmax = cd->max_layer_deg;
        d.dmin = 255;
        d.msg_words = cd->msg_words;
        d.pk_msg_words = cd->pk_msg_words;
        d.code_id = id;
        ent.insert(ent.end(), cd->ent, cd->ent + cd->nent);
        for (int i = 0; i < cd->q; i++) {
            const int st = cd->layer_start[i], deg = cd->layer_start[i + 1] - st;
            d.dmin = std::min(d.dmin, deg);
            lay.push_back(make_uint2((uint32_t)st | (uint32_t)deg << 16 | (uint32_t)cd->nlev[i] << 24,
                                     (uint32_t)cd->msg_off[i]));
            pkoff.push_back((uint32_t)cd->pk_msg_off[i]);
        }
        meta.insert(meta.end(), cd->meta, cd->meta + cd->r);
        /* every layer degree must be one the code's class kernel handles */
        const int cls = d.dmax <= LOW_MAX_DEG ? 0 : 1;
        if (cd->nent > MAX_ENT || cd->q > MAX_LAYERS || d.dmax > 28 || (cls == 1 && d.dmin < HIGH_MIN_DEG) ||
            d.dmin < 1)
            return set_err(c, LDPC_GPU_EINVAL, "code %d exceeds the kernel's table bounds", id);
    }
    HIPCHK(c, hipMalloc((void**)&t.codes, sizeof(DCode) * codes.size()));
    HIPCHK(c, hipMalloc((void**)&t.ent, sizeof(uint32_t) * ent.size()));
    HIPCHK(c, hipMalloc((void**)&t.lay, sizeof(uint2) * lay.size()));
    HIPCHK(c, hipMalloc((void**)&t.meta, sizeof(uint64_t) * meta.size()));
    HIPCHK(c, hipMalloc((void**)&t.pkoff, sizeof(uint32_t) * pkoff.size()));
    HIPCHK(c, hipMemcpy(t.pkoff, pkoff.data(), sizeof(uint32_t) * pkoff.size(), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(t.codes, codes.data(), sizeof(DCode) * codes.size(), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(t.ent, ent.data(), sizeof(uint32_t) * ent.size(), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(t.lay, lay.data(), sizeof(uint2) * lay.size(), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(t.meta, meta.data(), sizeof(uint64_t) * meta.size(), hipMemcpyHostToDevice));
    t.h_codes = codes;
    t.ready = 1;
    return LDPC_GPU_OK;
}

extern "C" {

int ldpc_gpu_create(int device, ldpc_gpu_ctx** out)
{
    if (!out)
        return LDPC_GPU_EINVAL;
    *out = nullptr;
    ldpc_gpu_ctx* c = (ldpc_gpu_ctx*)calloc(1, sizeof(ldpc_gpu_ctx));
    if (!c)
        return LDPC_GPU_ENOMEM;
    c->device = device;
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || device < 0 || device >= ndev) {
        free(c);
        return LDPC_GPU_EDEVICE;
    }
    int rc;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        [&] {
            for (int k = 0; k < NLAUNCH; k++)
                if (hipStreamCreateWithFlags(&c->aux[k], hipStreamNonBlocking) != hipSuccess)
                    return true;
            return false;
        }() ||
        hipEventCreate(&c->ev0) != hipSuccess || hipEventCreate(&c->ev1) != hipSuccess ||
        [&] {
            for (int k = 0; k < 2 * NLAUNCH; k++)
                if (hipEventCreate(&c->ev_cls[k]) != hipSuccess)
                    return true;
            return false;
        }()) {
        free(c);
        return LDPC_GPU_EDEVICE;
    }
    if (hipDeviceGetAttribute(&c->ncu, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess)
        c->ncu = 256;
    if ((rc = upload_tables(c)) != LDPC_GPU_OK || hipMalloc((void**)&c->ticket, NCOUNTERS * sizeof(unsigned)) != hipSuccess) {
        ldpc_gpu_destroy(c);
        return rc ? rc : LDPC_GPU_EDEVICE;
    }
    *out = c;
    return LDPC_GPU_OK;
}

void ldpc_gpu_destroy(ldpc_gpu_ctx* c)
{
    if (!c)
        return;
    (void)hipSetDevice(c->device);
    if (c->stream)
        (void)hipStreamSynchronize(c->stream);
    (void)hipFree(c->ticket);
    (void)hipFree(c->d_frames);
    (void)hipFree(c->d_list);
    (void)hipFree(c->msg);
    (void)hipFree(c->state);
    (void)hipFree(c->it_state);
    (void)hipFree(c->target);
    if (c->ev0)
        (void)hipEventDestroy(c->ev0);
    if (c->ev1)
        (void)hipEventDestroy(c->ev1);
    for (int k = 0; k < 2 * NLAUNCH; k++)
        if (c->ev_cls[k])
            (void)hipEventDestroy(c->ev_cls[k]);
    (void)hipFree(c->d_seg);
    (void)hipFree(c->d_nS);
    (void)hipFree(c->d_nB);
    (void)hipFree(c->d_nrow0);
    (void)hipFree(c->d_nT);
    free(c->h_nS);
    free(c->h_nB);
    free(c->h_nrow0);
    free(c->h_nT);
    for (int k = 0; k < NLAUNCH; k++)
        if (c->aux[k]) {
            (void)hipStreamSynchronize(c->aux[k]);
            (void)hipStreamDestroy(c->aux[k]);
        }
    if (c->stream)
        (void)hipStreamDestroy(c->stream);
    free(c);
}

void* ldpc_gpu_stream(ldpc_gpu_ctx* c) { return c ? (void*)c->stream : nullptr; }
const char* ldpc_gpu_last_error(ldpc_gpu_ctx* c) { return c ? c->err : "no context"; }

int ldpc_gpu_code_count(void) { return lg_code_count(); }
int ldpc_gpu_code_lookup(int short_frame, const char* rate) { return lg_code_lookup(short_frame, rate); }

int ldpc_gpu_code_params(int code, int* n, int* k, int* q, int* links_total, int* level_steps)
{
    const lg_code* c = lg_code_get(code);
    if (!c)
        return LDPC_GPU_EINVAL;
    if (n)
        *n = c->n;
    if (k)
        *k = c->k;
    if (q)
        *q = c->q;
    if (links_total)
        *links_total = c->links_total;
    if (level_steps)
        *level_steps = c->level_steps;
    return LDPC_GPU_OK;
}

int ldpc_gpu_code_max_degree(int code)
{
    const lg_code* c = lg_code_get(code);
    return c ? c->max_layer_deg : LDPC_GPU_EINVAL;
}

int ldpc_gpu_fec_lookup(int short_frame, const char* rate, int* nldpc, int* kldpc, int* kbch, int* t)
{
    int kb, nb, tt;
    if (lg_fec_lookup(short_frame, rate, &kb, &nb, &tt))
        return LDPC_GPU_EINVAL;
    if (nldpc)
        *nldpc = short_frame ? 16200 : 64800;
    if (kldpc)
        *kldpc = nb;
    if (kbch)
        *kbch = kb;
    if (t)
        *t = tt;
    return LDPC_GPU_OK;
}

void* ldpc_gpu_malloc(ldpc_gpu_ctx* c, size_t bytes)
{
    void* p = nullptr;
    if (!c || hipSetDevice(c->device) != hipSuccess || hipMalloc(&p, bytes ? bytes : 1) != hipSuccess)
        return nullptr;
    return p;
}

int ldpc_gpu_free(ldpc_gpu_ctx* c, void* p)
{
    if (!c)
        return LDPC_GPU_EINVAL;
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipFree(p));
    return LDPC_GPU_OK;
}

int ldpc_gpu_upload(ldpc_gpu_ctx* c, void* d_dst, const void* src, size_t bytes)
{
    if (!c)
        return LDPC_GPU_EINVAL;
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipMemcpyAsync(d_dst, src, bytes, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return LDPC_GPU_OK;
}

int ldpc_gpu_download(ldpc_gpu_ctx* c, void* dst, const void* d_src, size_t bytes)
{
    if (!c)
        return LDPC_GPU_EINVAL;
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipMemcpyAsync(dst, d_src, bytes, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return LDPC_GPU_OK;
}

int ldpc_gpu_synchronize(ldpc_gpu_ctx* c)
{
    if (!c)
        return LDPC_GPU_EINVAL;
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return LDPC_GPU_OK;
}

int ldpc_gpu_query(ldpc_gpu_ctx* c)
{
    if (!c)
        return LDPC_GPU_EINVAL;
    HIPCHK(c, hipSetDevice(c->device));
    const hipError_t e = hipStreamQuery(c->stream);
    if (e == hipErrorNotReady)
        return 0;
    HIPCHK(c, e);
    return 1;
}

int ldpc_gpu_set_concurrent(ldpc_gpu_ctx* c, int concurrent)
{
    if (!c)
        return LDPC_GPU_EINVAL;
    c->serial = !concurrent;
    return LDPC_GPU_OK;
}

int ldpc_gpu_set_kernel(ldpc_gpu_ctx* c, int kernel)
{
    if (!c || (kernel != LDPC_GPU_KERNEL_AUTO && kernel != LDPC_GPU_KERNEL_SINGLE))
        return LDPC_GPU_EINVAL;
    c->force_scalar = kernel == LDPC_GPU_KERNEL_SINGLE;
    return LDPC_GPU_OK;
}

int ldpc_gpu_last_launches(ldpc_gpu_ctx* c, int* di, int* nframes, float* ms, int max)
{
    if (!c || !c->timed)
        return 0;
    int n = 0;
    for (int k = 0; k < NLAUNCH && n < max; k++) {
        if (!c->cls_frames[k])
            continue;
        float t = -1.0f;
        if (hipEventSynchronize(c->ev_cls[2 * k + 1]) != hipSuccess ||
            hipEventElapsedTime(&t, c->ev_cls[2 * k], c->ev_cls[2 * k + 1]) != hipSuccess)
            t = -1.0f;
        if (di)
            di[n] = c->cls_di[k];
        if (nframes)
            nframes[n] = c->cls_frames[k];
        if (ms)
            ms[n] = t;
        n++;
    }
    return n;
}

float ldpc_gpu_last_kernel_ms(ldpc_gpu_ctx* c)
{
    if (!c || !c->timed)
        return -1.0f;
    float ms = -1.0f;
    if (hipEventSynchronize(c->ev1) != hipSuccess || hipEventElapsedTime(&ms, c->ev0, c->ev1) != hipSuccess)
        return -1.0f;
    return ms;
}

} // extern "C"

template <typename T>
static int grow(ldpc_gpu_ctx* c, T** p, size_t* cap_bytes, size_t need)
{
    if (need <= *cap_bytes)
        return LDPC_GPU_OK;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipFree(*p));
    *p = nullptr;
    *cap_bytes = 0;
    HIPCHK(c, hipMalloc((void**)p, need));
    *cap_bytes = need;
    return LDPC_GPU_OK;
}


static int class_of(const DCode& cd) { return cd.dmax <= LOW_MAX_DEG ? 0 : 1; }

template <int DI>
static int class_grid(ldpc_gpu_ctx* c, size_t lds, int nrun, int* grid)
{
    int per_cu = 0;
    HIPCHK(c, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, ldpc_layered_kernel<DI>, NT, lds));
    if (per_cu < 1)
        return set_err(c, LDPC_GPU_EDEVICE, "kernel does not fit (LDS %zu)", lds);
    *grid = std::min(nrun, per_cu * c->ncu);
    return LDPC_GPU_OK;
}

template <int DI>
static int launch_class(ldpc_gpu_ctx* c, const KArgs& a, size_t lds, int grid, int cls)
{
    if (!a.msg_per_frame && (size_t)grid * a.msg_stride * sizeof(uint32_t) > c->msg_bytes)
        return set_err(c, LDPC_GPU_EINVAL, "message scratch too small");
    HIPCHK(c, hipEventRecord(c->ev_cls[2 * cls], c->stream));
    hipLaunchKernelGGL(ldpc_layered_kernel<DI>, dim3(grid), dim3(NT), lds, c->stream, a);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipEventRecord(c->ev_cls[2 * cls + 1], c->stream));
    return LDPC_GPU_OK;
}

/* Decode the frames of `run` (indices into the frame table): frames are
 * split by degree class, each class is one persistent launch over its
 * frames, normal frames before short ones. */
static int launch(ldpc_gpu_ctx* c, const DFrame* d_frames, const std::vector<DFrame>& hf, const std::vector<int>& run,
                  int maxw, const int8_t* d_llr, uint8_t* d_out, int32_t* d_iters, uint8_t* d_synd, int per_frame_msg,
                  int nframes_total, int resume, int save, int timed)
{
    if (run.empty())
        return LDPC_GPU_OK;
    const std::vector<DCode>& hc = c->tab->h_codes;
    std::vector<int> lists[NCLASS];
    int maxn[NCLASS] = { 0 };
    for (int f : run) {
        const DCode& cd = hc[hf[f].code];
        const int k = class_of(cd);
        lists[k].push_back(f);
        maxn[k] = std::max(maxn[k], cd.n);
    }
    std::vector<int> all;
    int off[NCLASS];
    for (int k = 0; k < NCLASS; k++) {
        std::stable_sort(lists[k].begin(), lists[k].end(),
                         [&](int x, int y) { return hc[hf[x].code].n > hc[hf[y].code].n; });
        off[k] = (int)all.size();
        all.insert(all.end(), lists[k].begin(), lists[k].end());
    }
    int rc = grow(c, &c->d_list, &c->list_bytes, sizeof(int) * all.size());
    if (rc)
        return rc;
    /* message blocks: one per workgroup of the largest grid, or per frame */
    int grid[NCLASS] = { 0 }, maxgrid = 0;
    for (int k = 0; k < NCLASS; k++) {
        if (lists[k].empty())
            continue;
        const size_t lds = (size_t)maxn[k];
        const int nr = (int)lists[k].size();
        rc = k == 0 ? class_grid<0>(c, lds, nr, &grid[k]) : class_grid<1>(c, lds, nr, &grid[k]);
        if (rc)
            return rc;
        maxgrid = std::max(maxgrid, grid[k]);
    }
    const size_t msg_stride = (size_t)maxw;
    const size_t blocks = per_frame_msg ? (size_t)nframes_total : (size_t)maxgrid;
    if ((rc = grow(c, &c->msg, &c->msg_bytes, blocks * msg_stride * sizeof(uint32_t))))
        return rc;
    HIPCHK(c, hipMemcpyAsync(c->d_list, all.data(), sizeof(int) * all.size(), hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(reset_ticket, dim3(1), dim3(NCOUNTERS), 0, c->stream, c->ticket);
    KArgs a;
    memset(&a, 0, sizeof(a));
    a.codes = c->tab->codes;
    a.ent = c->tab->ent;
    a.lay = c->tab->lay;
    a.meta = c->tab->meta;
    a.frames = d_frames;
    a.llr = d_llr;
    a.out = d_out;
    a.iters = d_iters;
    a.synd = d_synd;
    a.msg = c->msg;
    a.msg_stride = msg_stride;
    a.msg_per_frame = per_frame_msg;
    a.state = c->state;
    a.state_stride = 64800;
    a.it_state = c->it_state;
    a.target = c->target;
    a.resume = resume;
    a.save = save;
    if (timed)
        HIPCHK(c, hipEventRecord(c->ev0, c->stream));
    for (int k = 0; k < NLAUNCH; k++)
        c->cls_frames[k] = 0;
    for (int k = 0; k < NCLASS; k++) {
        c->cls_frames[k] = (int)lists[k].size();
        c->cls_di[k] = CLASS_DI[k];
        if (lists[k].empty())
            continue;
        a.list = c->d_list + off[k];
        a.nrun = (int)lists[k].size();
        a.ticket = c->ticket + k;
        const size_t lds = (size_t)maxn[k];
        rc = k == 0 ? launch_class<0>(c, a, lds, grid[k], k) : launch_class<1>(c, a, lds, grid[k], k);
        if (rc)
            return rc;
    }
    if (timed) {
        HIPCHK(c, hipEventRecord(c->ev1, c->stream));
        c->timed = 1;
    }
    return LDPC_GPU_OK;
}

template <int CLS, bool NMS>
static int pk_grid(ldpc_gpu_ctx* c, size_t lds, int nframes, int* grid)
{
    int per_cu = 0;
    HIPCHK(c, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, ldpc_pk_kernel<CLS, NMS>, PK_NT, lds));
    if (per_cu < 1)
        return set_err(c, LDPC_GPU_EDEVICE, "two-slot kernel does not fit (LDS %zu)", lds);
    *grid = std::min((nframes + 1) / 2, per_cu * c->ncu);
    return LDPC_GPU_OK;
}

/* Launch layout of frames with their own early stop (two per workgroup,
 * ldpc_pk_kernel): one launch per (degree class, frame size), frames
 * grouped by code into segments of one list, persistent grid per launch
 * (occupancy x CUs, scaled by grid_scale). */
struct PkLayout {
    std::vector<int> list;
    std::vector<int2> segs;
    int seg0[NLAUNCH], nseg[NLAUNCH], nfr[NLAUNCH], grid[NLAUNCH], maxn[NLAUNCH];
    int maxw;
    size_t blocks; /* message blocks: sum of the grids */
};

static int pk_layout(ldpc_gpu_ctx* c, const std::vector<DFrame>& hf, bool nms, float grid_scale, PkLayout& L)
{
    const std::vector<DCode>& hc = c->tab->h_codes;
    std::vector<int> gcodes[NLAUNCH];
    std::vector<std::vector<int>> gseg[NLAUNCH];
    L.maxw = 0;
    for (int k = 0; k < NLAUNCH; k++)
        L.maxn[k] = 0;
    for (int f = 0; f < (int)hf.size(); f++) {
        const DCode& cd = hc[hf[f].code];
        const int k = class_of(cd) * 2 + (cd.n < 64800);
        auto it = std::find(gcodes[k].begin(), gcodes[k].end(), hf[f].code);
        size_t g = it - gcodes[k].begin();
        if (it == gcodes[k].end()) {
            gcodes[k].push_back(hf[f].code);
            gseg[k].emplace_back();
        }
        gseg[k][g].push_back(f);
        L.maxn[k] = std::max(L.maxn[k], cd.n);
        L.maxw = std::max(L.maxw, cd.pk_msg_words);
    }
    L.list.clear();
    L.segs.clear();
    L.blocks = 0;
    for (int k = 0; k < NLAUNCH; k++) {
        L.seg0[k] = (int)L.segs.size();
        L.nseg[k] = (int)gseg[k].size();
        L.nfr[k] = 0;
        L.grid[k] = 0;
        for (auto& sg : gseg[k]) {
            L.segs.push_back(make_int2((int)L.list.size(), (int)sg.size()));
            L.list.insert(L.list.end(), sg.begin(), sg.end());
            L.nfr[k] += (int)sg.size();
        }
        if (!L.nfr[k])
            continue;
        const size_t lds = 2 * (size_t)L.maxn[k];
        const int rc = nms ? (k < 2 ? pk_grid<0, true>(c, lds, L.nfr[k], &L.grid[k])
                                    : pk_grid<1, true>(c, lds, L.nfr[k], &L.grid[k]))
                           : (k < 2 ? pk_grid<0, false>(c, lds, L.nfr[k], &L.grid[k])
                                    : pk_grid<1, false>(c, lds, L.nfr[k], &L.grid[k]));
        if (rc)
            return rc;
        L.grid[k] = std::max(1, (int)((float)L.grid[k] * grid_scale));
        L.blocks += (size_t)L.grid[k];
    }
    if (L.segs.size() > NCOUNTERS)
        return set_err(c, LDPC_GPU_EINVAL, "too many codes in one batch");
    return LDPC_GPU_OK;
}

/* Enqueue the launches of a layout: the counter reset (a kernel) and one
 * persistent launch per (class, size), each on its own forked stream when
 * concurrent (fork / join through events), else one after the other on
 * the origin stream.  Kernels and stream dependencies only, no copies or
 * allocations, so the sequence can be captured in a graph.  tev (may be
 * NULL): start / end timing events per launch. */
static int pk_enqueue(ldpc_gpu_ctx* c, const PkLayout& L, PkArgs a, const int2* d_seg, unsigned* d_cnt,
                      uint32_t* d_msg, hipStream_t origin, const hipStream_t* aux, hipEvent_t fork,
                      const hipEvent_t* join, const hipEvent_t* tev, bool concurrent, bool nms)
{
    hipLaunchKernelGGL(reset_ticket, dim3(1), dim3(NCOUNTERS), 0, origin, d_cnt);
    HIPCHK(c, hipGetLastError());
    if (concurrent)
        HIPCHK(c, hipEventRecord(fork, origin));
    a.msg_stride = (size_t)L.maxw;
    size_t moff = 0;
    for (int k = 0; k < NLAUNCH; k++) {
        if (!L.nfr[k])
            continue;
        a.seg = d_seg + L.seg0[k];
        a.nseg = L.nseg[k];
        a.cnt = d_cnt + L.seg0[k];
        a.msg = d_msg + moff;
        moff += (size_t)L.grid[k] * L.maxw;
        const size_t lds = 2 * (size_t)L.maxn[k];
        hipStream_t st = concurrent ? aux[k] : origin;
        if (concurrent)
            HIPCHK(c, hipStreamWaitEvent(st, fork, 0));
        if (tev)
            HIPCHK(c, hipEventRecord(tev[2 * k], st));
        if (nms) {
            if (k < 2)
                hipLaunchKernelGGL((ldpc_pk_kernel<0, true>), dim3(L.grid[k]), dim3(PK_NT), lds, st, a);
            else
                hipLaunchKernelGGL((ldpc_pk_kernel<1, true>), dim3(L.grid[k]), dim3(PK_NT), lds, st, a);
        } else {
            if (k < 2)
                hipLaunchKernelGGL((ldpc_pk_kernel<0, false>), dim3(L.grid[k]), dim3(PK_NT), lds, st, a);
            else
                hipLaunchKernelGGL((ldpc_pk_kernel<1, false>), dim3(L.grid[k]), dim3(PK_NT), lds, st, a);
        }
        HIPCHK(c, hipGetLastError());
        if (tev)
            HIPCHK(c, hipEventRecord(tev[2 * k + 1], st));
        if (concurrent) {
            HIPCHK(c, hipEventRecord(join[k], st));
            HIPCHK(c, hipStreamWaitEvent(origin, join[k], 0));
        }
    }
    return LDPC_GPU_OK;
}

static PkArgs pk_args(ldpc_gpu_ctx* c, const DFrame* d_frames, const int* d_list, const int8_t* d_llr, uint8_t* d_out,
                      int32_t* d_iters, uint8_t* d_synd, bool nms, int fallback, uint8_t* d_fell)
{
    PkArgs a;
    memset(&a, 0, sizeof(a));
    a.codes = c->tab->codes;
    a.ent = c->tab->ent;
    a.lay = c->tab->lay;
    a.pkoff = c->tab->pkoff;
    a.meta = c->tab->meta;
    a.frames = d_frames;
    a.list = d_list;
    a.llr = d_llr;
    a.out = d_out;
    a.iters = d_iters;
    a.synd = d_synd;
    if (nms) {
        a.nS = c->d_nS;
        a.nB = c->d_nB;
        a.nrow0 = c->d_nrow0;
        a.nT = c->d_nT;
        a.fallback = fallback;
        a.fell = d_fell;
    }
    return a;
}

/* Frames with their own early stop, decoded now: the layout's tables go to
 * the context's scratch, the launches run concurrently on the forked
 * streams (the last frames of one launch overlap the others' work), each
 * with its own message blocks, timed per launch. */
static int launch_pk(ldpc_gpu_ctx* c, const DFrame* d_frames, const std::vector<DFrame>& hf, const int8_t* d_llr,
                     uint8_t* d_out, int32_t* d_iters, uint8_t* d_synd, bool nms = false, int fallback = 0,
                     uint8_t* d_fell = nullptr)
{
    PkLayout L;
    int rc = pk_layout(c, hf, nms, 1.0f, L);
    if (rc)
        return rc;
    if ((rc = grow(c, &c->d_list, &c->list_bytes, sizeof(int) * L.list.size())) ||
        (rc = grow(c, &c->d_seg, &c->seg_bytes, sizeof(int2) * L.segs.size())) ||
        (rc = grow(c, &c->msg, &c->msg_bytes, L.blocks * L.maxw * sizeof(uint32_t))))
        return rc;
    HIPCHK(c, hipMemcpyAsync(c->d_list, L.list.data(), sizeof(int) * L.list.size(), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->d_seg, L.segs.data(), sizeof(int2) * L.segs.size(), hipMemcpyHostToDevice, c->stream));
    const PkArgs a = pk_args(c, d_frames, c->d_list, d_llr, d_out, d_iters, d_synd, nms, fallback, d_fell);
    hipEvent_t join[NLAUNCH];
    for (int k = 0; k < NLAUNCH; k++) {
        join[k] = c->ev_cls[2 * k + 1];
        c->cls_frames[k] = L.nfr[k];
        c->cls_di[k] = CLASS_DI[k / 2];
    }
    if ((rc = pk_enqueue(c, L, a, c->d_seg, c->ticket, c->msg, c->stream, c->aux, c->ev0, join, c->ev_cls, !c->serial,
                         nms)))
        return rc;
    HIPCHK(c, hipEventRecord(c->ev1, c->stream));
    c->timed = 1;
    return LDPC_GPU_OK;
}

/* ---- learned offset min-sum tables -------------------------------------- */

#define NMS_ROWS (1 + 64 * LDPC_NMS_MAX_ITERS) /* row 0 plain, then per code LDPC_NMS_MAX_ITERS rows */

extern "C" int ldpc_gpu_set_nms(ldpc_gpu_ctx* c, int code, int T, const uint8_t* S, const int16_t* B)
{
    if (!c)
        return LDPC_GPU_EINVAL;
    const int nc = (int)c->tab->h_codes.size();
    if (code < 0 || code >= nc || nc > 64 || T < 0 || T > LDPC_NMS_MAX_ITERS || (T > 0 && (!S || !B)))
        return set_err(c, LDPC_GPU_EINVAL, "set_nms: bad arguments (code %d, T %d)", code, T);
    if (!c->h_nS) {
        c->h_nS = (uint8_t*)malloc(NMS_ROWS * LDPC_NMS_DEGS);
        c->h_nB = (int16_t*)malloc(NMS_ROWS * LDPC_NMS_DEGS * sizeof(int16_t));
        c->h_nrow0 = (int*)calloc(64, sizeof(int));
        c->h_nT = (int*)calloc(64, sizeof(int));
        if (!c->h_nS || !c->h_nB || !c->h_nrow0 || !c->h_nT)
            return set_err(c, LDPC_GPU_ENOMEM, "set_nms: host tables");
        memset(c->h_nS, 16, NMS_ROWS * LDPC_NMS_DEGS);
        for (int k = 0; k < NMS_ROWS * LDPC_NMS_DEGS; k++)
            c->h_nB[k] = 16;
        HIPCHK(c, hipSetDevice(c->device));
        HIPCHK(c, hipMalloc((void**)&c->d_nS, NMS_ROWS * LDPC_NMS_DEGS));
        HIPCHK(c, hipMalloc((void**)&c->d_nB, NMS_ROWS * LDPC_NMS_DEGS * sizeof(int16_t)));
        HIPCHK(c, hipMalloc((void**)&c->d_nrow0, 64 * sizeof(int)));
        HIPCHK(c, hipMalloc((void**)&c->d_nT, 64 * sizeof(int)));
    }
    const int r0 = 1 + code * LDPC_NMS_MAX_ITERS;
    c->h_nrow0[code] = r0;
    c->h_nT[code] = T;
    if (T > 0) {
        memcpy(c->h_nS + (size_t)r0 * LDPC_NMS_DEGS, S, (size_t)T * LDPC_NMS_DEGS);
        for (int k = 0; k < T * LDPC_NMS_DEGS; k++) {
            if (S[k] > 63 || B[k] < -1024 || B[k] > 1023)
                return set_err(c, LDPC_GPU_EINVAL, "set_nms: entry %d out of range (S 0..63, B -1024..1023)", k);
        }
        memcpy(c->h_nB + (size_t)r0 * LDPC_NMS_DEGS, B, (size_t)T * LDPC_NMS_DEGS * sizeof(int16_t));
    }
    c->nms_dirty = 1;
    return LDPC_GPU_OK;
}

/* frame table checks shared by the decode entry points */
static int check_frames(ldpc_gpu_ctx* c, const ldpc_gpu_frame* frames, int nframes, std::vector<DFrame>& df,
                        int* maxw, int* ngroup)
{
    const std::vector<DCode>& hc = c->tab->h_codes;
    df.resize(nframes);
    *maxw = 0;
    *ngroup = 0;
    for (int f = 0; f < nframes; f++) {
        const ldpc_gpu_frame& s = frames[f];
        if (s.code < 0 || s.code >= (int)hc.size())
            return set_err(c, LDPC_GPU_EINVAL, "frame %d: unknown code %d", f, s.code);
        const DCode& cd = hc[s.code];
        if (s.max_iter < 0 || s.nbits < 0 || s.nbits > cd.k || (s.llr_off & 7))
            return set_err(c, LDPC_GPU_EINVAL, "frame %d: bad max_iter/nbits/llr_off", f);
        df[f].code = s.code;
        df[f].max_iter = s.max_iter;
        df[f].nbits = s.nbits ? s.nbits : cd.k;
        df[f].pad = 0;
        df[f].llr_off = s.llr_off;
        df[f].out_off = s.out_off;
        *maxw = std::max(*maxw, cd.msg_words);
        *ngroup += s.group >= 0;
    }
    return LDPC_GPU_OK;
}

extern "C" int ldpc_gpu_decode_batch_nms(ldpc_gpu_ctx* c, const ldpc_gpu_frame* frames, int nframes,
                                         const int8_t* d_llr, uint8_t* d_out, int32_t* d_iters, uint8_t* d_synd,
                                         int fallback_max_iter, uint8_t* d_fallback)
{
    if (!c)
        return LDPC_GPU_EINVAL;
    if (nframes == 0)
        return LDPC_GPU_OK;
    if (nframes < 0 || !frames || !d_llr || !d_out || !d_iters || !d_synd || fallback_max_iter < 0)
        return set_err(c, LDPC_GPU_EINVAL, "bad arguments");
    HIPCHK(c, hipSetDevice(c->device));
    std::vector<DFrame> df;
    int maxw, ngroup;
    int rc = check_frames(c, frames, nframes, df, &maxw, &ngroup);
    if (rc)
        return rc;
    if (ngroup)
        return set_err(c, LDPC_GPU_EINVAL, "learned decoding takes frames with their own early stop (group -1)");
    for (int f = 0; f < nframes; f++)
        if (!c->h_nT || c->h_nT[df[f].code] < 1)
            return set_err(c, LDPC_GPU_EINVAL, "frame %d: no learned table for code %d", f, df[f].code);
    if (c->nms_dirty) {
        HIPCHK(c, hipMemcpyAsync(c->d_nS, c->h_nS, NMS_ROWS * LDPC_NMS_DEGS, hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipMemcpyAsync(c->d_nB, c->h_nB, NMS_ROWS * LDPC_NMS_DEGS * sizeof(int16_t), hipMemcpyHostToDevice,
                                 c->stream));
        HIPCHK(c, hipMemcpyAsync(c->d_nrow0, c->h_nrow0, 64 * sizeof(int), hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipMemcpyAsync(c->d_nT, c->h_nT, 64 * sizeof(int), hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        c->nms_dirty = 0;
    }
    if ((rc = grow(c, &c->d_frames, &c->frames_bytes, sizeof(DFrame) * nframes)))
        return rc;
    HIPCHK(c, hipMemcpyAsync(c->d_frames, df.data(), sizeof(DFrame) * nframes, hipMemcpyHostToDevice, c->stream));
    return launch_pk(c, c->d_frames, df, d_llr, d_out, d_iters, d_synd, true, fallback_max_iter, d_fallback);
}

extern "C" int ldpc_gpu_decode_batch(ldpc_gpu_ctx* c, const ldpc_gpu_frame* frames, int nframes, const int8_t* d_llr,
                                     uint8_t* d_out, int32_t* d_iters, uint8_t* d_synd)
{
    if (!c)
        return LDPC_GPU_EINVAL;
    if (nframes == 0)
        return LDPC_GPU_OK;
    if (nframes < 0 || !frames || !d_llr || !d_out || !d_iters || !d_synd)
        return set_err(c, LDPC_GPU_EINVAL, "bad arguments");
    HIPCHK(c, hipSetDevice(c->device));
    std::vector<DFrame> df;
    int maxw, ngroup;
    int rc = check_frames(c, frames, nframes, df, &maxw, &ngroup);
    if (rc)
        return rc;
    if ((rc = grow(c, &c->d_frames, &c->frames_bytes, sizeof(DFrame) * nframes)))
        return rc;
    HIPCHK(c, hipMemcpyAsync(c->d_frames, df.data(), sizeof(DFrame) * nframes, hipMemcpyHostToDevice, c->stream));
    std::vector<int> every(nframes);
    for (int f = 0; f < nframes; f++)
        every[f] = f;
    if (!ngroup && !c->force_scalar)
        return launch_pk(c, c->d_frames, df, d_llr, d_out, d_iters, d_synd);
    if (!ngroup)
        return launch(c, c->d_frames, df, every, maxw, d_llr, d_out, d_iters, d_synd, 0, nframes, 0, 0, 1);

    /* Groups (the reference's batch semantics): decode every frame with its
     * own early stop and keep its state; a group's count T is the largest
     * of its frames' counts (max_iter if one never converged); the others
     * continue to T; while some frame fails at T < max_iter, the whole
     * group goes one more iteration.  Frames are deterministic, so this
     * equals iterating the group in lock step. */
    if ((rc = grow(c, &c->state, &c->state_bytes, (size_t)nframes * 64800)) ||
        (rc = grow(c, &c->it_state, &c->it_state_bytes, sizeof(int32_t) * nframes)) ||
        (rc = grow(c, &c->target, &c->target_bytes, sizeof(int32_t) * nframes)))
        return rc;
    rc = launch(c, c->d_frames, df, every, maxw, d_llr, d_out, d_iters, d_synd, 1, nframes, 0, 1, 1);
    if (rc)
        return rc;
    std::vector<int32_t> it(nframes), tgt(nframes);
    std::vector<uint8_t> ok(nframes);
    HIPCHK(c, hipMemcpyAsync(it.data(), d_iters, sizeof(int32_t) * nframes, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(ok.data(), d_synd, nframes, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    /* group -> (T, all ok, max_iter) */
    std::vector<int> gid(nframes);
    std::vector<int> gT, gmax, gall;
    {
        std::vector<int> keys;
        for (int f = 0; f < nframes; f++) {
            if (frames[f].group < 0) {
                gid[f] = -1;
                continue;
            }
            auto pos = std::find(keys.begin(), keys.end(), frames[f].group);
            int g;
            if (pos == keys.end()) {
                g = (int)keys.size();
                keys.push_back(frames[f].group);
                gT.push_back(0);
                gmax.push_back(frames[f].max_iter);
                gall.push_back(1);
            } else {
                g = (int)(pos - keys.begin());
            }
            gid[f] = g;
            if (frames[f].max_iter != gmax[g])
                return set_err(c, LDPC_GPU_EINVAL, "frames of group %d differ in max_iter", frames[f].group);
            gT[g] = std::max(gT[g], (int)it[f]);
            gall[g] &= ok[f];
        }
    }
    for (size_t g = 0; g < gT.size(); g++)
        if (!gall[g])
            gT[g] = gmax[g];
    for (;;) {
        std::vector<int> list;
        for (int f = 0; f < nframes; f++) {
            if (gid[f] < 0)
                continue;
            tgt[f] = gT[gid[f]];
            if (it[f] < tgt[f])
                list.push_back(f);
        }
        if (list.empty())
            break;
        HIPCHK(c, hipMemcpyAsync(c->target, tgt.data(), sizeof(int32_t) * nframes, hipMemcpyHostToDevice, c->stream));
        rc = launch(c, c->d_frames, df, list, maxw, d_llr, d_out, d_iters, d_synd, 1, nframes, 1, 1, 0);
        if (rc)
            return rc;
        HIPCHK(c, hipMemcpyAsync(it.data(), d_iters, sizeof(int32_t) * nframes, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipMemcpyAsync(ok.data(), d_synd, nframes, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        /* groups still failing below their cap go one more iteration */
        std::vector<int> all(gT.size(), 1);
        for (int f = 0; f < nframes; f++)
            if (gid[f] >= 0)
                all[gid[f]] &= ok[f];
        int more = 0;
        for (size_t g = 0; g < gT.size(); g++) {
            if (!all[g] && gT[g] < gmax[g]) {
                gT[g]++;
                more = 1;
            }
        }
        if (!more)
            break;
    }
    return LDPC_GPU_OK;
}

/* ---- prepared batches: graph capture and the launch-shape autotuner ---- */

/* launch shapes the autotuner chooses from: grid scale (persistent
 * workgroups per launch relative to occupancy x CUs, i.e. how many frames
 * run at once) and whether the (class, size) launches run concurrently */
static const struct {
    float scale;
    bool concurrent;
} PLAN_VARIANTS[] = { { 1.0f, true }, { 1.0f, false }, { 0.75f, true }, { 0.5f, true }, { 0.5f, false } };
#define PLAN_NVARIANTS ((int)(sizeof(PLAN_VARIANTS) / sizeof(PLAN_VARIANTS[0])))

static int plan_build(ldpc_gpu_plan* p, int variant)
{
    ldpc_gpu_ctx* c = p->c;
    PkLayout L;
    int rc = pk_layout(c, p->hf, false, PLAN_VARIANTS[variant].scale, L);
    if (rc)
        return rc;
    if (L.list.size() > p->list_cap || L.segs.size() > p->seg_cap || L.blocks * L.maxw > p->msg_cap)
        return set_err(c, LDPC_GPU_EINVAL, "plan scratch too small");
    HIPCHK(c, hipMemcpy(p->d_list, L.list.data(), sizeof(int) * L.list.size(), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(p->d_seg, L.segs.data(), sizeof(int2) * L.segs.size(), hipMemcpyHostToDevice));
    if (p->exec) {
        (void)hipGraphExecDestroy(p->exec);
        p->exec = nullptr;
    }
    if (p->graph) {
        (void)hipGraphDestroy(p->graph);
        p->graph = nullptr;
    }
    const PkArgs a = pk_args(c, p->d_frames, p->d_list, p->d_llr, p->d_out, p->d_iters, p->d_synd, false, 0, nullptr);
    HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    rc = pk_enqueue(c, L, a, p->d_seg, p->d_cnt, p->d_msg, c->stream, p->aux, p->fork, p->join, nullptr,
                    PLAN_VARIANTS[variant].concurrent, false);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(c->stream, &g);
    if (rc || e != hipSuccess) {
        if (g)
            (void)hipGraphDestroy(g);
        return rc ? rc : set_err(c, LDPC_GPU_EDEVICE, "graph capture: %s", hipGetErrorString(e));
    }
    p->graph = g;
    HIPCHK(c, hipGraphInstantiate(&p->exec, g, nullptr, nullptr, 0));
    p->variant = variant;
    return LDPC_GPU_OK;
}

extern "C" void ldpc_gpu_plan_destroy(ldpc_gpu_plan* p)
{
    if (!p)
        return;
    ldpc_gpu_ctx* c = p->c;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    if (p->exec)
        (void)hipGraphExecDestroy(p->exec);
    if (p->graph)
        (void)hipGraphDestroy(p->graph);
    (void)hipFree(p->d_frames);
    (void)hipFree(p->d_list);
    (void)hipFree(p->d_seg);
    (void)hipFree(p->d_cnt);
    (void)hipFree(p->d_msg);
    for (int k = 0; k < NLAUNCH; k++) {
        if (p->aux[k])
            (void)hipStreamDestroy(p->aux[k]);
        if (p->join[k])
            (void)hipEventDestroy(p->join[k]);
    }
    if (p->fork)
        (void)hipEventDestroy(p->fork);
    if (p->t0)
        (void)hipEventDestroy(p->t0);
    if (p->t1)
        (void)hipEventDestroy(p->t1);
    delete p;
}

extern "C" int ldpc_gpu_plan_create(ldpc_gpu_ctx* c, const ldpc_gpu_frame* frames, int nframes, const int8_t* d_llr,
                                    uint8_t* d_out, int32_t* d_iters, uint8_t* d_synd, ldpc_gpu_plan** out)
{
    if (!c || !out)
        return LDPC_GPU_EINVAL;
    *out = nullptr;
    if (nframes < 1 || !frames || !d_llr || !d_out || !d_iters || !d_synd)
        return set_err(c, LDPC_GPU_EINVAL, "bad arguments");
    HIPCHK(c, hipSetDevice(c->device));
    std::vector<DFrame> df;
    int maxw, ngroup;
    int rc = check_frames(c, frames, nframes, df, &maxw, &ngroup);
    if (rc)
        return rc;
    if (ngroup)
        return set_err(c, LDPC_GPU_EINVAL, "plans take frames with their own early stop (group -1)");
    ldpc_gpu_plan* p = new (std::nothrow) ldpc_gpu_plan();
    if (!p)
        return LDPC_GPU_ENOMEM;
    p->c = c;
    p->nframes = nframes;
    p->hf = df;
    p->d_llr = d_llr;
    p->d_out = d_out;
    p->d_iters = d_iters;
    p->d_synd = d_synd;
    p->variant = -1;
    for (int v = 0; v < 8; v++)
        p->variant_ms[v] = -1.0f;
    /* scratch sized for the largest variant (full grid) */
    PkLayout L;
    if ((rc = pk_layout(c, df, false, 1.0f, L))) {
        delete p;
        return rc;
    }
    p->list_cap = L.list.size();
    p->seg_cap = L.segs.size();
    p->msg_cap = L.blocks * L.maxw;
    auto fail = [&](int r) {
        ldpc_gpu_plan_destroy(p);
        return r;
    };
    if (hipMalloc((void**)&p->d_frames, sizeof(DFrame) * nframes) != hipSuccess ||
        hipMalloc((void**)&p->d_list, sizeof(int) * p->list_cap) != hipSuccess ||
        hipMalloc((void**)&p->d_seg, sizeof(int2) * p->seg_cap) != hipSuccess ||
        hipMalloc((void**)&p->d_cnt, sizeof(unsigned) * NCOUNTERS) != hipSuccess ||
        hipMalloc((void**)&p->d_msg, sizeof(uint32_t) * p->msg_cap) != hipSuccess)
        return fail(set_err(c, LDPC_GPU_ENOMEM, "plan scratch"));
    if (hipEventCreateWithFlags(&p->fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreate(&p->t0) != hipSuccess || hipEventCreate(&p->t1) != hipSuccess)
        return fail(LDPC_GPU_EDEVICE);
    for (int k = 0; k < NLAUNCH; k++)
        if (hipStreamCreateWithFlags(&p->aux[k], hipStreamNonBlocking) != hipSuccess ||
            hipEventCreateWithFlags(&p->join[k], hipEventDisableTiming) != hipSuccess)
            return fail(LDPC_GPU_EDEVICE);
    if (hipMemcpy(p->d_frames, df.data(), sizeof(DFrame) * nframes, hipMemcpyHostToDevice) != hipSuccess)
        return fail(LDPC_GPU_EDEVICE);
    if ((rc = plan_build(p, 0)))
        return fail(rc);
    *out = p;
    return LDPC_GPU_OK;
}

extern "C" int ldpc_gpu_plan_launch(ldpc_gpu_plan* p)
{
    if (!p)
        return LDPC_GPU_EINVAL;
    ldpc_gpu_ctx* c = p->c;
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipEventRecord(p->t0, c->stream));
    HIPCHK(c, hipGraphLaunch(p->exec, c->stream));
    HIPCHK(c, hipEventRecord(p->t1, c->stream));
    p->launched = 1;
    return LDPC_GPU_OK;
}

extern "C" float ldpc_gpu_plan_last_ms(ldpc_gpu_plan* p)
{
    if (!p || !p->launched)
        return -1.0f;
    float ms = -1.0f;
    if (hipEventSynchronize(p->t1) != hipSuccess || hipEventElapsedTime(&ms, p->t0, p->t1) != hipSuccess)
        return -1.0f;
    return ms;
}

extern "C" int ldpc_gpu_plan_tune(ldpc_gpu_plan* p, int reps)
{
    if (!p || reps < 1)
        return LDPC_GPU_EINVAL;
    ldpc_gpu_ctx* c = p->c;
    int best = 0, rc;
    for (int v = 0; v < PLAN_NVARIANTS; v++) {
        if ((rc = plan_build(p, v)))
            return rc;
        if ((rc = ldpc_gpu_plan_launch(p)))
            return rc;
        HIPCHK(c, hipStreamSynchronize(c->stream));
        float tot = 0.0f;
        for (int r = 0; r < reps; r++) {
            if ((rc = ldpc_gpu_plan_launch(p)))
                return rc;
            const float ms = ldpc_gpu_plan_last_ms(p);
            if (ms < 0.0f)
                return set_err(c, LDPC_GPU_EDEVICE, "plan timing failed");
            tot += ms;
        }
        p->variant_ms[v] = tot / (float)reps;
        if (p->variant_ms[v] < p->variant_ms[best])
            best = v;
    }
    return plan_build(p, best);
}

extern "C" int ldpc_gpu_plan_select(ldpc_gpu_plan* p, int variant)
{
    if (!p)
        return LDPC_GPU_EINVAL;
    if (variant < 0 || variant >= PLAN_NVARIANTS)
        return set_err(p->c, LDPC_GPU_EINVAL, "plan variant %d out of range", variant);
    HIPCHK(p->c, hipSetDevice(p->c->device));
    HIPCHK(p->c, hipStreamSynchronize(p->c->stream));
    return plan_build(p, variant);
}

extern "C" int ldpc_gpu_plan_info(ldpc_gpu_plan* p, int* variant, float* scale, int* concurrent, float* ms, int max)
{
    if (!p)
        return LDPC_GPU_EINVAL;
    if (variant)
        *variant = p->variant;
    for (int v = 0; v < PLAN_NVARIANTS && v < max; v++) {
        if (scale)
            scale[v] = PLAN_VARIANTS[v].scale;
        if (concurrent)
            concurrent[v] = PLAN_VARIANTS[v].concurrent;
        if (ms)
            ms[v] = p->variant_ms[v];
    }
    return PLAN_NVARIANTS;
}

extern "C" int ldpc_gpu_decode_host(ldpc_gpu_ctx* c, const ldpc_gpu_frame* frames, int nframes, const int8_t* llr,
                                    size_t llr_bytes, uint8_t* out, size_t out_bytes, int32_t* iters, uint8_t* synd)
{
    if (!c || nframes < 0 || (nframes && (!frames || !llr || !out || !iters || !synd)))
        return LDPC_GPU_EINVAL;
    if (nframes == 0)
        return LDPC_GPU_OK;
    /* every frame's LLRs and output bytes inside the buffers given */
    for (int f = 0; f < nframes; f++) {
        int n = 0, k = 0;
        if (ldpc_gpu_code_params(frames[f].code, &n, &k, nullptr, nullptr, nullptr))
            return set_err(c, LDPC_GPU_EINVAL, "frame %d: unknown code %d", f, frames[f].code);
        const uint64_t nb = ((uint64_t)(frames[f].nbits ? frames[f].nbits : k) + 7) / 8;
        if (frames[f].llr_off + (uint64_t)n > llr_bytes || frames[f].out_off + nb > out_bytes)
            return set_err(c, LDPC_GPU_EINVAL, "frame %d: LLRs or output outside the buffers", f);
    }
    HIPCHK(c, hipSetDevice(c->device));
    int8_t* d_llr = nullptr;
    uint8_t *d_out = nullptr, *d_synd = nullptr;
    int32_t* d_it = nullptr;
    int rc = LDPC_GPU_OK;
    if (hipMalloc((void**)&d_llr, llr_bytes) != hipSuccess || hipMalloc((void**)&d_out, out_bytes ? out_bytes : 1) != hipSuccess ||
        hipMalloc((void**)&d_it, sizeof(int32_t) * nframes) != hipSuccess || hipMalloc((void**)&d_synd, nframes) != hipSuccess) {
        rc = set_err(c, LDPC_GPU_ENOMEM, "device allocation failed");
    }
    if (!rc && hipMemcpyAsync(d_llr, llr, llr_bytes, hipMemcpyHostToDevice, c->stream) != hipSuccess)
        rc = set_err(c, LDPC_GPU_EDEVICE, "upload failed");
    if (!rc && out_bytes && hipMemcpyAsync(d_out, out, out_bytes, hipMemcpyHostToDevice, c->stream) != hipSuccess)
        rc = set_err(c, LDPC_GPU_EDEVICE, "upload failed");
    if (!rc)
        rc = ldpc_gpu_decode_batch(c, frames, nframes, d_llr, d_out, d_it, d_synd);
    if (!rc && (hipMemcpyAsync(out, d_out, out_bytes, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
                hipMemcpyAsync(iters, d_it, sizeof(int32_t) * nframes, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
                hipMemcpyAsync(synd, d_synd, nframes, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
                hipStreamSynchronize(c->stream) != hipSuccess))
        rc = set_err(c, LDPC_GPU_EDEVICE, "download failed");
    (void)hipStreamSynchronize(c->stream);
    (void)hipFree(d_llr);
    (void)hipFree(d_out);
    (void)hipFree(d_it);
    (void)hipFree(d_synd);
    return rc;
}
