/*
 * The split decoder kernel (ldpc_sk_kernel): two frames per workgroup in
 * the 16-bit halves of each register (ldpc_pk.h), each check node on a
 * lane pair (ldpc_sk.h), 720 threads = 360 nodes x 2.  One instance per
 * build of this file, chosen by SK_KS (slots per lane: the code's largest
 * layer degree + 2, halved, rounded up), SK_WIDE (1: layers of the code
 * differ by more than one slot, every slot may be empty) and SK_NMS (0
 * plain, 1 learned), with its launch entry points.
 *
 * Tables (PkArgs.adt, lane records as the code registry builds them,
 * ldpc_codes.h): per layer a record of SK_RA(KS) words per lane (the LDS
 * byte offsets of its slot posteriors) and, for layers with nodes sharing
 * bits, one of SK_RB(KS) words (the slots' dependency masks: dep in bits
 * 0-15, succ in 16-31, then the hand-over cells).  Messages: per layer and
 * lane SK_WORDS(KS) words in the workgroup's message block.  Layer
 * descriptor (PkArgs.pkl, per code from DCode.pkl_off, two copies of the
 * last after it): x = positions T | level count << 8 | barrier before <<
 * 16 | barrier after << 17 | kde << 24 (slots 1 .. kde - 1 hold the shared
 * edges), y = byte
 * offset of the layer's messages in the block, z = byte offset of its
 * address records in adt, w = of its tail records.  A lane reads and
 * writes its words with buffer instructions: the layer's offset in a
 * scalar register, the lane's in a fixed vector register, so no address
 * arithmetic runs per layer on the vector ALUs.
 */
#include <hip/hip_runtime.h>

#include "ldpc_codes.h"
#include "ldpc_kern.h"
#include "ldpc_nms.h"
#include "ldpc_sk.h"

#if !defined(SK_KS) || !defined(SK_WIDE) || !defined(SK_NMS)
#error "build with -DSK_KS=2..15 -DSK_WIDE=0|1 -DSK_NMS=0|1 [-DSK_OCC=2]"
#endif
/* workgroups per CU the instance is built for: 1, or 2 (short frames:
 * 6 waves per SIMD, registers bounded to fit two workgroups) */
#ifndef SK_OCC
#define SK_OCC 1
#endif

namespace lgk {

namespace {

constexpr int KS = SK_KS;
constexpr int KA = SK_WIDE ? 1 : KS - 1; /* slots k >= KA may be empty */
constexpr bool NMS = SK_NMS != 0;
constexpr int WA = SK_WORDS(KS);   /* message words per lane and layer */
constexpr int RA = SK_RA(KS);      /* address record words */
constexpr int RB = SK_RB(KS);      /* tail record words */
constexpr int CW = SK_KD(KS) / 2 + 1; /* hand-over cell words (slots 0 .. SK_KD) */

typedef __amdgpu_buffer_rsrc_t rsrc_t;

/* buffer resource of a range of device memory (offsets up to 2 GiB) */
__device__ __forceinline__ rsrc_t mk_rsrc(const void* p)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(p), (short)0, 0x7fffffff, 0x00020000);
}

/* N consecutive words at byte offset voff (lane) + soff (uniform); the
 * constant steps go with voff, where they fold into the instructions'
 * offset field */
template <int N>
__device__ __forceinline__ void ld_words(uint32_t* w, rsrc_t r, uint32_t voff, uint32_t soff)
{
    int k = 0;
#pragma unroll
    for (; k + 4 <= N; k += 4) {
        const auto v = __builtin_amdgcn_raw_buffer_load_b128(r, voff + 4 * k, soff, 0);
        w[k] = v[0];
        w[k + 1] = v[1];
        w[k + 2] = v[2];
        w[k + 3] = v[3];
    }
    if constexpr (N % 4 >= 2) {
        const auto v = __builtin_amdgcn_raw_buffer_load_b64(r, voff + 4 * (N & ~3), soff, 0);
        w[N & ~3] = v[0];
        w[(N & ~3) + 1] = v[1];
    }
    if constexpr (N % 2)
        w[N - 1] = __builtin_amdgcn_raw_buffer_load_b32(r, voff + 4 * (N - 1), soff, 0);
}

template <int N>
__device__ __forceinline__ void st_words(const uint32_t* w, rsrc_t r, uint32_t voff, uint32_t soff)
{
    int k = 0;
#pragma unroll
    for (; k + 4 <= N; k += 4) {
        typedef unsigned int v4u __attribute__((ext_vector_type(4)));
        const v4u v = { w[k], w[k + 1], w[k + 2], w[k + 3] };
        __builtin_amdgcn_raw_buffer_store_b128(v, r, voff + 4 * k, soff, 0);
    }
    if constexpr (N % 4 >= 2) {
        typedef unsigned int v2u __attribute__((ext_vector_type(2)));
        const v2u v = { w[N & ~3], w[(N & ~3) + 1] };
        __builtin_amdgcn_raw_buffer_store_b64(v, r, voff + 4 * (N & ~3), soff, 0);
    }
    if constexpr (N % 2)
        __builtin_amdgcn_raw_buffer_store_b32(w[N - 1], r, voff + 4 * (N - 1), soff, 0);
}

/* a lane's fixed offsets: its address record, tail record and message
 * words within a layer's */
struct SkLane {
    uint32_t va, vb, vm;
};

/* layer descriptors through the scalar cache: their index is uniform over
 * the workgroup, so a constant-address-space view of the table makes them
 * scalar loads (SGPRs, waited on with the LDS counter) instead of vector
 * loads whose wait would also wait for the message stores */
typedef const __attribute__((address_space(4))) uint32_t sl_cu32;
__device__ __forceinline__ uint4 sl_desc(const uint4* skl, int i)
{
    const sl_cu32* p = (const sl_cu32*)(skl + i);
    return make_uint4(p[0], p[1], p[2], p[3]);
}

__device__ __forceinline__ int sl_T(uint32_t x) { return (int)(x & 0xff); }

/* entry k (uniform) of a learned table row, by a scalar load of the
 * aligned dword that holds it: S rows are bytes, B rows 16-bit (even
 * offsets); the device tables carry a dword of padding after their last
 * row */
__device__ __forceinline__ int nms_u8(const uint8_t* row, int k)
{
    const uintptr_t ad = (uintptr_t)(row + k);
    const sl_cu32* w = (const sl_cu32*)(ad & ~(uintptr_t)3);
    return (int)((w[0] >> (8 * (ad & 3))) & 0xffu);
}
__device__ __forceinline__ int nms_i16(const int16_t* row, int k)
{
    const uintptr_t ad = (uintptr_t)(row + k);
    const sl_cu32* w = (const sl_cu32*)(ad & ~(uintptr_t)3);
    return (int)(int16_t)(w[0] >> (8 * (ad & 3)));
}
__device__ __forceinline__ int sl_nlev(uint32_t x) { return (int)((x >> 8) & 0xff); }
__device__ __forceinline__ bool sl_bar_after(uint32_t x) { return (x >> 17) & 1; }
__device__ __forceinline__ int sl_kde(uint32_t x) { return (int)(x >> 24); }

/* the partner lane's value (lanes 2j and 2j + 1 swap: quad permutation
 * [1, 0, 3, 2]) */
__device__ __forceinline__ uint32_t lane_swap(uint32_t v) { return (uint32_t)__builtin_amdgcn_mov_dpp((int)v, 0xB1, 0xF, 0xF, false); }

template <int N>
__device__ __forceinline__ void sk_pair(SkHalf<N>& n)
{
    const lg_pk p0 = pk_of(lane_swap(pk_bits(n.min0))), p1 = pk_of(lane_swap(pk_bits(n.min1)));
    const lg_pk ps = pk_of(lane_swap(pk_bits(n.s)));
    sk_combine(n, p0, p1, ps);
}

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

/* syndrome of both frames, chunks of layers with an early exit once every
 * slot that is still decoding has failed a check; bit k: slot k holds */
__device__ __forceinline__ uint32_t sk_syndrome(const uint16_t* lds, rsrc_t rt, const uint4* skl, int t, int q,
                                                uint32_t va, uint32_t live, unsigned* s_vote, int& nvote)
{
    const int h = t & 1;
    uint32_t good = 3;
    for (int i0 = 0; i0 < q; i0 += SYN_CHUNK) {
        uint32_t ok = 3;
        const int i1 = min(i0 + SYN_CHUNK, q);
        uint32_t aw[SYN_CHUNK][KS];
        int Ts[SYN_CHUNK];
#pragma unroll
        for (int c = 0; c < SYN_CHUNK; c++) {
            if (i0 + c >= i1)
                continue;
            const uint4 ld = sl_desc(skl, i0 + c);
            Ts[c] = sl_T(ld.x);
            ld_words<KS>(aw[c], rt, va, ld.z);
        }
#pragma unroll
        for (int c = 0; c < SYN_CHUNK; c++) {
            if (i0 + c >= i1)
                continue;
            uint32_t par, nz;
            sk_check<KS, KA>(lds, aw[c], Ts[c], h, (i0 + c == 0 && t == 1) ? ~0u : 0u, par, nz);
            par ^= lane_swap(par);
            nz &= lane_swap(nz);
            ok &= sk_check_bits(par, nz);
        }
        good &= wg_and2(ok, s_vote, nvote);
        if (!(good & live))
            break;
    }
    return good;
}

/* One layer by the workgroup.  TAIL (the layer has nodes sharing bits):
 * after phase A the pairs run hand-over rounds (ldpc_sk.h) until what they
 * hand over stops changing or the layer's level count is reached; hb: the
 * two hand-over buffers, hstride u16 cells apart. */
template <bool TAIL>
__device__ __forceinline__ void sk_layer(SkHalf<KS>& n, const uint32_t* raw, uint16_t* lds, int T, int h,
                                         uint32_t skip0, int nlev, int kde, const uint32_t* mw, uint32_t tm,
                                         const uint32_t* hw, rsrc_t rm, uint32_t vm, uint32_t moff, uint32_t se,
                                         uint32_t so, lg_pk ps, lg_pk nb, uint16_t* hb, int hstride,
                                         unsigned* s_vote, int& nvote)
{
    const uint32_t dep = TAIL ? (tm & 0xffffu) : 0, succ = TAIL ? (tm >> 16) : 0;
    sk_phase_a<KS, KA, TAIL>(n, raw, T, h, skip0, dep, mw, se, so);
    if (TAIL) {
        const PkPart part = sk_part(n);
        sk_fold_dep0(n, dep, kde);
        sk_pair(n);
        sk_put_hand<KS, NMS>(n, hb, hw, succ, kde, ps, nb);
        __syncthreads();
#pragma unroll 1
        for (int r = 1; r < nlev; r++) {
            const uint16_t* prv = hb + ((r - 1) & 1) * hstride;
            uint32_t old[KS];
            sk_restore(n, part);
            sk_take_hand(n, prv, hw, dep, kde, mw, se, so);
            sk_prev_hand<KS>(old, prv, hw, succ, kde);
            sk_pair(n);
            if (r == nlev - 1)
                break;
            const uint32_t chg = sk_put_hand_chg<KS, NMS>(n, hb + (r & 1) * hstride, hw, succ, kde, old, ps, nb);
            if (wg_and2(chg ? 0u : 3u, s_vote, nvote) & 1)
                break;
        }
    } else {
        sk_pair(n);
    }
    uint32_t mo[WA];
    sk_phase_c<KS, TAIL, NMS>(n, lds, succ, mo, ps, nb);
    st_words<WA>(mo, rm, vm, moff);
}

/* one iteration (all layers) on both frames; the address records, masks
 * and messages of layer i+1 are fetched while layer i runs, the descriptor
 * of layer i+2 too */
__device__ __forceinline__ void sk_iteration(uint16_t* lds, rsrc_t rt, const uint4* skl, rsrc_t rm, SkLane ln, int q,
                                             int first0, int first1, PkRowPtrs rows, uint16_t* hb, int hstride,
                                             unsigned* s_vote, int& nvote)
{
    const int t = threadIdx.x, h = t & 1, j = t >> 1;
    const uint32_t se = pk_msg_sel(0, first0, first1), so = pk_msg_sel(1, first0, first1);
    const bool load = !(first0 && first1);
    uint32_t nx[WA], na[KS], nt[1 + CW];
#pragma unroll
    for (int w = 0; w < WA; w++)
        nx[w] = 0;
#pragma unroll
    for (int w = 0; w < 1 + CW; w++)
        nt[w] = 0;
    auto fetch = [&](const uint4& ld) {
        ld_words<KS>(na, rt, ln.va, ld.z);
        if (load)
            ld_words<WA>(nx, rm, ln.vm, ld.y);
        if (sl_nlev(ld.x) > 1)
            ld_words<1 + CW>(nt, rt, ln.vb, ld.w);
    };
    /* the descriptors: layer i's in ldn at the top of layer i, layer i +
     * 1's in ldf, then layer i + 2's from dp (the table holds two copies of
     * the last descriptor after it) */
    uint4 ldn = sl_desc(skl, 0);
    fetch(ldn);
    uint4 ldf = sl_desc(skl, 1);
    const uint4* dp = skl + 2;
    /* layer 0's words are needed at once; waiting for them here also tells
     * the compiler that nothing is outstanding when the loop starts, so the
     * loop's own waits need not cover the message stores of the layer
     * before (s_waitcnt 0: every counter) */
    __builtin_amdgcn_s_waitcnt(0);
    for (int i = 0; i < q; i++) {
        const uint4 ld = ldn;
        uint32_t mw[WA], tw[1 + CW];
#pragma unroll
        for (int w = 0; w < WA; w++)
            mw[w] = nx[w];
#pragma unroll
        for (int w = 0; w < 1 + CW; w++)
            tw[w] = nt[w];
        SkHalf<KS> n;
        sk_addrs(n, na);
        /* the next layer's fetch is issued first: the posterior reads then
         * sit in the block that uses them */
        if (i + 1 < q) {
            ldn = ldf;
            fetch(ldn);
            ldf = sl_desc(dp, 0);
            dp++;
        }
        uint32_t raw[KS];
        sk_read<KS>(raw, lds, n);
        const int T = sl_T(ld.x), nlev = sl_nlev(ld.x);
        const uint32_t skip0 = (i == 0 && t == 1) ? ~0u : 0u;
        /* learned scale / negated offset of the layer's check degree, per frame;
         * node (0, 0) has no previous parity, one edge less */
        lg_pk ps = pk_dup(16), nb = pk_dup(-16);
        if constexpr (NMS) {
            int s0 = nms_u8(rows.S0, T), s1 = nms_u8(rows.S1, T);
            int b0 = nms_i16(rows.B0, T), b1 = nms_i16(rows.B1, T);
            if (i == 0) {
                const int e0 = nms_u8(rows.S0, T - 1), e1 = nms_u8(rows.S1, T - 1);
                const int f0 = nms_i16(rows.B0, T - 1), f1 = nms_i16(rows.B1, T - 1);
                if (j == 0) {
                    s0 = e0;
                    s1 = e1;
                    b0 = f0;
                    b1 = f1;
                }
            }
            ps = lg_pk{ (short)s0, (short)s1 };
            nb = lg_pk{ (short)-b0, (short)-b1 };
        }
        if (nlev > 1)
            sk_layer<true>(n, raw, lds, T, h, skip0, nlev, sl_kde(ld.x), mw, tw[0], tw + 1, rm, ln.vm, ld.y, se, so, ps,
                           nb, hb, hstride, s_vote, nvote);
        else
            sk_layer<false>(n, raw, lds, T, h, skip0, nlev, 0, mw, 0, tw + 1, rm, ln.vm, ld.y, se, so, ps, nb, hb,
                            hstride, s_vote, nvote);
        /* the next layer waits for this one where it reads what another
         * wave wrote (its barrier flag; the end of the iteration always) */
        if (sl_bar_after(ld.x))
            __syncthreads();
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
/* the template arguments name the instance (slots per lane, wide,
 * learned, workgroups per CU) in profiles; the body uses this file's build
 * constants.  The launch bound's second argument is the waves per SIMD the
 * registers must allow: 720 threads are 12 waves, 3 per SIMD per
 * workgroup. */
template <int KS_, int WIDE_, int NMS_, int OCC_>
__global__ void __launch_bounds__(SK_NT, 3 * SK_OCC) ldpc_sk_kernel(PkArgs a)
{
    static_assert(KS_ == KS && WIDE_ == SK_WIDE && NMS_ == SK_NMS && OCC_ == SK_OCC, "instance of this build");
    /* all of the LDS is dynamic, so it starts at address 0 and the slot
     * offsets of the address records are LDS addresses as they stand:
     * control words, then the posteriors at SK_POST_OFF, then the
     * hand-over buffers at a.hoff */
    extern __shared__ __align__(16) uint16_t lds16[];
    unsigned* s_vote = (unsigned*)lds16;           /* [3] */
    int* s_claim = (int*)lds16 + 3;                /* [2] */
    int& s_seg = ((int*)lds16)[5];
    int& s_newseg = ((int*)lds16)[6];
    uint16_t* const sm16 = lds16 + SK_POST_OFF / 2; /* posteriors */
    /* the slot offsets are LDS addresses only if the LDS starts at 0 (no
     * static LDS in this kernel); otherwise decode nothing, and the frames'
     * results stay unwritten */
    if (__builtin_amdgcn_readfirstlane((uint32_t)(size_t)lds16) != 0u)
        return;
    const int tid = threadIdx.x;
    if (tid < 3)
        s_vote[tid] = 3;
    int nvote = 0;
    int seg = -1;                 /* current segment (uniform) */
    int frame[2] = { -1, -1 };    /* frame in each slot, -1: empty */
    int it[2] = { 0, 0 }, maxit[2] = { 0, 0 };
    int minit[2] = { 0, 0 }; /* iterations before the syndrome counts (DFrame.min_iter) */
    int phase[2] = { 0, 0 }, lit[2] = { 0, 0 }; /* learned mode: 1 = plain fallback; learned iterations */
    const rsrc_t rt = mk_rsrc(a.adt);
    const rsrc_t rm = mk_rsrc(a.msg + a.msg_stride * (size_t)blockIdx.x);
    const SkLane ln = { (uint32_t)(4 * RA * tid), (uint32_t)(4 * RB * tid), (uint32_t)(4 * WA * tid) };
    uint16_t* hb = (uint16_t*)((uint8_t*)lds16 + a.hoff);
    DCode cd;

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
        }
        const uint4* skl = a.pkl + cd.pkl_off;
        const int K = cd.k, R = cd.r, q = cd.q;
        /* a frame's LLRs into its slot bytes: information bits in order,
         * parity bit K + q*jj + ii at K + 360*ii + jj (layer order) */
        auto load_slot = [&](int s, int f) {
            const DFrame fr = a.frames[f];
            const uint32_t* src = (const uint32_t*)(a.llr + fr.llr_off);
            uint8_t* sb = (uint8_t*)sm16 + s;
            for (int v4 = tid; v4 < K / 4; v4 += SK_NT) {
                const uint32_t w = src[v4];
                for (int b = 0; b < 4; b++)
                    sb[2 * (4 * v4 + b)] = (uint8_t)(w >> (8 * b));
            }
            const uint8_t* par = (const uint8_t*)(a.llr + fr.llr_off + K);
            for (int t = tid; t < R; t += SK_NT) {
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
            minit[s] = a.frames[f].min_iter;
            phase[s] = 0;
            lit[s] = 0;
            load_slot(s, f);
        }
        __syncthreads();

        /* ---- iterate until a slot is done ---- */
        for (;;) {
            /* the slots whose syndrome counts at this iteration; none: no
             * syndrome pass at all */
            uint32_t live = 0;
            for (int s = 0; s < 2; s++)
                live |= (frame[s] >= 0 && it[s] >= minit[s]) ? 1u << s : 0u;
            const uint32_t good = live ? sk_syndrome(lds16, rt, skl, tid, q, ln.va, live, s_vote, nvote) & live : 0u;
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
                    for (int b4 = 4 * tid; b4 < nbytes; b4 += 4 * SK_NT) {
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
            sk_iteration(lds16, rt, skl, rm, ln, q, frame[0] < 0 || it[0] == 0, frame[1] < 0 || it[1] == 0, rows, hb,
                         a.hstride, s_vote, nvote);
            for (int s = 0; s < 2; s++)
                it[s] += frame[s] >= 0;
        }
    }
}


} // namespace

#if SK_OCC == 1
#define SK_SUF
#else
#define SK_SUF _o2
#endif
#define SK_CAT(a, b, c, d, e) a##b##_##c##_##d##e
#define SK_NAME(a, b, c, d, e) SK_CAT(a, b, c, d, e)

hipError_t SK_NAME(sk_launch_, SK_KS, SK_WIDE, SK_NMS, SK_SUF)(int grid, size_t lds, hipStream_t st, const PkArgs& a)
{
    hipLaunchKernelGGL((ldpc_sk_kernel<SK_KS, SK_WIDE, SK_NMS, SK_OCC>), dim3(grid), dim3(SK_NT), lds, st, a);
    return hipGetLastError();
}

hipError_t SK_NAME(sk_occupancy_, SK_KS, SK_WIDE, SK_NMS, SK_SUF)(size_t lds, int* per_cu)
{
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, ldpc_sk_kernel<SK_KS, SK_WIDE, SK_NMS, SK_OCC>, SK_NT,
                                                        lds);
}

} // namespace lgk
