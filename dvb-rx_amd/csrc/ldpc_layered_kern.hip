/*
 * The one-frame decoder kernel (ldpc_layered_kernel, see ldpc_gpu.hip) and
 * the counter reset, with their launch entry points.
 */
#include <hip/hip_runtime.h>

#include "ldpc_codes.h"
#include "ldpc_kern.h"
#include "ldpc_node.h"

namespace lgk {

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


hipError_t layered_launch(int cls, int grid, size_t lds, hipStream_t st, const KArgs& a)
{
    if (cls == 0)
        hipLaunchKernelGGL(ldpc_layered_kernel<0>, dim3(grid), dim3(NT), lds, st, a);
    else
        hipLaunchKernelGGL(ldpc_layered_kernel<1>, dim3(grid), dim3(NT), lds, st, a);
    return hipGetLastError();
}

hipError_t layered_occupancy(int cls, size_t lds, int* per_cu)
{
    return cls == 0 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, ldpc_layered_kernel<0>, NT, lds)
                    : hipOccupancyMaxActiveBlocksPerMultiprocessor(per_cu, ldpc_layered_kernel<1>, NT, lds);
}

hipError_t reset_counters(hipStream_t st, unsigned* cnt)
{
    hipLaunchKernelGGL(reset_ticket, dim3(1), dim3(NCOUNTERS), 0, st, cnt);
    return hipGetLastError();
}

} // namespace lgk
