/*
 * The two-slot decoder kernel (ldpc_pk_kernel, see ldpc_gpu.hip): one
 * instance per build of this file, chosen by LGK_DM (the largest layer
 * information degree the instance handles: 4, 8, 12, 16, 20 or 28) and
 * LGK_NMS (0 plain, 1 learned), with its launch entry points.  One layer
 * body of degree <= LGK_DM serves every layer of the codes sent to it.
 */
#include <hip/hip_runtime.h>

#include "ldpc_codes.h"
#include "ldpc_kern.h"
#include "ldpc_nms.h"
#include "ldpc_pk.h"

#if !defined(LGK_DM) || !defined(LGK_NMS)
#error "build with -DLGK_DM=4|8|12|16|20|28 -DLGK_NMS=0|1"
#endif

namespace lgk {

namespace {

constexpr int DM = LGK_DM;
constexpr bool NMS = LGK_NMS != 0;
constexpr int WM = LG_PK_WORDS(DM);

/* layer descriptor (PkArgs.pkl): x = first table entry | degree << 16 |
 * level count << 24, y = first message row, z = edges on shared bits */
__device__ __forceinline__ int ld_start(uint32_t x) { return (int)(x & 0xffff); }
__device__ __forceinline__ int ld_deg(uint32_t x) { return (int)((x >> 16) & 0xff); }
__device__ __forceinline__ int ld_nlev(uint32_t x) { return (int)(x >> 24); }

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
__device__ __forceinline__ uint32_t pk_syndrome(const uint16_t* sm, const uint32_t* ent, const uint4* pkl, int j, int K,
                                                int q, uint32_t live, unsigned* s_vote, int& nvote)
{
    uint32_t good = 3;
    for (int i0 = 0; i0 < q; i0 += SYN_CHUNK) {
        uint32_t ok = 3;
        const int i1 = min(i0 + SYN_CHUNK, q);
        for (int i = i0; i < i1; i++) {
            const uint32_t lx = pkl[i].x;
            ok &= pk_check<DM>(sm, ent + ld_start(lx), ld_deg(lx), i, j, K, q);
        }
        good &= wg_and2(ok, s_vote, nvote);
        if (!(good & live))
            break;
    }
    return good;
}

/* One layer of D <= DM information edges by the workgroup.  TAIL (the
 * layer has nodes sharing bits): after phase A the nodes run hand-over
 * rounds (ldpc_pk.h) until their inputs stop changing or the layer's level
 * count is reached; hb: the two hand-over buffers, hstride u16 cells
 * apart. */
template <bool TAIL>
__device__ __forceinline__ void pk_layer(uint16_t* sm, const uint32_t* ent, const uint32_t* aw, int D, int i, int j,
                                         int K, int q, int nlev, const uint32_t* mw, uint64_t meta, uint64_t pos,
                                         uint32_t dupm, uint32_t* mrow, uint32_t se, uint32_t so, lg_pk ps, lg_pk pb,
                                         uint16_t* hb, int hstride, unsigned* s_vote, int& nvote)
{
    const uint32_t dep = TAIL ? LG_META_DEP(meta) : 0, succ = TAIL ? LG_META_SUCC(meta) : 0;
    PkNode<DM> n;
    pk_addr_words<DM>(n.ad, aw, D);
    pk_phase_a<DM, TAIL>(n, sm, D, i, j, K, q, dep, mw, se, so);
    if (TAIL) {
        const PkPart part = pk_part(n);
        pk_fold_dep0<DM>(n, D, dep);
        pk_put_hand<DM, NMS>(n, hb, ent, D, dupm, succ, pos, ps, pb);
        __syncthreads();
#pragma unroll 1
        for (int r = 1; r < nlev; r++) {
            pk_restore(n, part);
            const uint32_t chg = pk_take_hand<DM>(n, hb + ((r - 1) & 1) * hstride, ent, D, dupm, dep, pos, mw, se, so);
            if (r == nlev - 1)
                break;
            pk_put_hand<DM, NMS>(n, hb + (r & 1) * hstride, ent, D, dupm, succ, pos, ps, pb);
            if (wg_and2(chg ? 0u : 3u, s_vote, nvote) & 1)
                break;
        }
    }
    pk_phase_c<DM, TAIL, NMS>(n, sm, D, succ, mrow + j, 360, ps, pb);
}

/* one iteration (all layers) on both slots; the code's tables are read
 * through uniform pointers (scalar loads); the messages, node masks and
 * descriptor of layer i+1 are fetched while layer i runs */
__device__ __forceinline__ void pk_iteration(uint16_t* sm, const uint32_t* ent, const uint4* pkl, const uint32_t* adt,
                                             const uint64_t* meta, const uint64_t* pos, uint32_t* msg, int K, int q,
                                             int first0, int first1, PkRowPtrs rows, uint16_t* hb, int hstride,
                                             unsigned* s_vote, int& nvote)
{
    const int j = threadIdx.x;
    const uint32_t se = pk_msg_sel(0, first0, first1), so = pk_msg_sel(1, first0, first1);
    const bool load = !(first0 && first1);
    constexpr int WA = (DM + 1) / 2;
    uint32_t nx[WM], na[WA];
    uint64_t nm = 0, np = 0;
#pragma unroll
    for (int w = 0; w < WM; w++)
        nx[w] = 0;
    auto fetch = [&](uint32_t lx, uint32_t row0, uint32_t arow, int l) {
        const int W = LG_PK_WORDS(ld_deg(lx)), A = (ld_deg(lx) + 1) / 2;
        const uint32_t* ar = adt + (size_t)360 * arow + j;
#pragma unroll
        for (int w = 0; w < WA; w++) {
            if (w >= A)
                continue;
            na[w] = ar[360 * w];
        }
        if (load) {
            const uint32_t* row = msg + 360 * row0 + j;
#pragma unroll
            for (int w = 0; w < WM; w++) {
                if (w >= W)
                    continue;
                nx[w] = row[360 * w];
            }
        }
        if (ld_nlev(lx) > 1) {
            nm = meta[360 * l + j];
            np = pos[360 * l + j];
        }
    };
    uint4 ldn = pkl[0];
    fetch(ldn.x, ldn.y, ldn.w, 0);
    for (int i = 0; i < q; i++) {
        const uint4 ld = ldn;
        uint32_t mw[WM], aw[WA];
#pragma unroll
        for (int w = 0; w < WM; w++)
            mw[w] = nx[w];
#pragma unroll
        for (int w = 0; w < WA; w++)
            aw[w] = na[w];
        const uint64_t mt = nm, pt = np;
        if (i + 1 < q) {
            ldn = pkl[i + 1];
            fetch(ldn.x, ldn.y, ldn.w, i + 1);
        }
        const int D = ld_deg(ld.x), nlev = ld_nlev(ld.x);
        const uint32_t* e = ent + ld_start(ld.x);
        uint32_t* mrow = msg + 360 * ld.y;
        /* learned scale / offset of the layer's check degree, per slot;
         * node (0, 0) has no previous parity, one edge less */
        lg_pk ps = pk_dup(16), pb = pk_dup(16);
        if constexpr (NMS) {
            const int dc = D + 2 - (i == 0 && j == 0);
            ps = lg_pk{ (short)rows.S0[dc], (short)rows.S1[dc] };
            pb = lg_pk{ (short)rows.B0[dc], (short)rows.B1[dc] };
        }
        if (nlev > 1)
            pk_layer<true>(sm, e, aw, D, i, j, K, q, nlev, mw, mt, pt, ld.z, mrow, se, so, ps, pb, hb, hstride,
                           s_vote, nvote);
        else
            pk_layer<false>(sm, e, aw, D, i, j, K, q, nlev, mw, mt, pt, 0, mrow, se, so, ps, pb, hb, hstride, s_vote,
                            nvote);
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
    uint16_t* hb = (uint16_t*)((uint8_t*)sm16 + a.hoff);
    DCode cd;
    const uint64_t* meta = nullptr;
    const uint64_t* pos = nullptr;

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
            pos = a.pos + cd.meta_off;
        }
        const uint32_t* ent = a.ent + cd.ent_off;
        const uint4* pkl = a.pkl + cd.lay_off;
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
            const uint32_t good = pk_syndrome(sm16, ent, pkl, tid, K, q, live, s_vote, nvote);
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
            pk_iteration(sm16, ent, pkl, a.adt, meta, pos, msg, K, q, frame[0] < 0 || it[0] == 0, frame[1] < 0 || it[1] == 0,
                         rows, hb, a.hstride, s_vote, nvote);
            for (int s = 0; s < 2; s++)
                it[s] += frame[s] >= 0;
        }
    }
}


} // namespace

#define LGK_CAT(a, b, c, d) a##b##_##c##d
#define LGK_NAME(a, b, c, d) LGK_CAT(a, b, c, d)

hipError_t LGK_NAME(pk_launch_, LGK_DM, LGK_NMS, )(int grid, size_t lds, hipStream_t st, const PkArgs& a)
{
    hipLaunchKernelGGL(ldpc_pk_kernel, dim3(grid), dim3(PK_NT), lds, st, a);
    return hipGetLastError();
}

hipError_t LGK_NAME(pk_occupancy_, LGK_DM, LGK_NMS, )(size_t lds, int* per_cu)
{
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, ldpc_pk_kernel, PK_NT, lds);
}

} // namespace lgk
