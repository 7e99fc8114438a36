/*
 * DVB-T2 FEC blocks on the device: the front end of the reference's FEC
 * thread (libdvbrx/src/dvbt2.c decode_job, :2188-2400, with the maps of
 * setup_plp_fec, :1953-2082) for a whole PLP at once, then the LDPC and BCH
 * decoders of this library.
 *
 * t2_demap_kernel: one workgroup per FEC block.
 *   1. Each block's cyclic shift P of the cell deinterleaver (the
 *      bit-reversed counter of its TI-block), by one thread.
 *   2. Cell deinterleaving, derotation and per-axis noise weights of every
 *      cell, and the MER sums.  The sums are float and must come out as
 *      the reference's: it accumulates cell i in lane i % 8 of an 8-wide
 *      vector, then adds the 8 lanes in order.  The workgroup computes the
 *      terms of 2048 cells into LDS, then 8 threads (16 with the signal
 *      sum) add their lane's terms in cell order.
 *   3. The noise estimate n0 = err / ncells + 1e-6 sets the soft-value
 *      scale k = 8 / n0; every cell is recomputed (its loads hit L2) and
 *      its m soft values go to LDS by (cell, plane).
 *   4. The bit deinterleaver as a gather: codeword bit n takes LDS entry
 *      src[n] (a per-PLP-configuration table of uint16), written to HBM in
 *      coalesced 32-bit words.
 * Floating point follows the reference's x86-64 FMA build exactly: the
 * fused products are fmaf() calls and everything else is kept unfused
 * (contraction off).
 *
 * HBM per block: 12 bytes a cell in (twice; the second pass mostly from
 * L2), 4 bytes a cell of the permutation, 3 bytes a codeword bit (table and
 * output).  The kernel is latency-bound on the ordered sums, not on HBM.
 */
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <vector>

#include "ldpc_codes.h"
#include "ldpc_gpu.h"
#include "ldpc_internal.h"

namespace {

constexpr int T2_MAX_DEV = 16;
constexpr int T2_BLOCK = 256;
constexpr int T2_CHUNK = 2048; /* cells per pass of the ordered sums */
constexpr int T2_LDS = 64800;  /* one block's soft values (the sums' terms alias it) */
static_assert(2 * T2_CHUNK * 4 <= T2_LDS, "sum terms fit the soft-value area");
static_assert(T2_LDS + 2 * T2_CHUNK * 4 + 256 <= 160 * 1024 / 2, "two workgroups per CU with staged cells");

/* Column twist (EN 302 755 table 8; libdvbrx/src/dvbt2.c:188-193),
 * 16/64/256-QAM, normal then short frames. */
const int TW_N[3][16] = { { 0, 0, 2, 4, 4, 5, 7, 7 },
                          { 0, 0, 2, 2, 3, 4, 4, 5, 5, 7, 8, 9 },
                          { 0, 2, 2, 2, 2, 3, 7, 15, 16, 20, 22, 22, 27, 27, 28, 32 } };
const int TW_S[3][16] = { { 0, 0, 0, 1, 7, 20, 20, 21 },
                          { 0, 0, 0, 2, 2, 2, 3, 3, 3, 6, 7, 7 },
                          { 0, 0, 0, 1, 7, 20, 20, 21 } };
/* Demultiplexer: bit interleaver column c carries word bit mux[c]
 * (EN 302 755 tables 13a/13b; libdvbrx/src/dvbt2.c:194-200). */
const int MUX16[8] = { 7, 1, 4, 2, 5, 3, 6, 0 };
const int MUX64[12] = { 11, 7, 3, 10, 6, 2, 9, 5, 1, 8, 4, 0 };
const int MUX256[16] = { 15, 1, 13, 3, 8, 11, 9, 5, 10, 6, 4, 7, 12, 2, 14, 0 };
const int MUX16_35[8] = { 0, 5, 1, 2, 4, 7, 3, 6 };
const int MUX64_35[12] = { 2, 7, 6, 9, 0, 3, 1, 8, 4, 11, 5, 10 };
const int MUX256_35[16] = { 2, 11, 3, 4, 0, 9, 1, 8, 10, 13, 7, 14, 6, 15, 5, 12 };
const int MUX256_23[16] = { 7, 2, 9, 0, 4, 6, 13, 3, 14, 10, 15, 5, 8, 12, 11, 1 };
const int MUX256S[8] = { 7, 3, 1, 5, 2, 6, 4, 0 };
/* constellation rotation in degrees (EN 302 755 table 9) */
const double ROT_DEG[4] = { 29.0, 16.8, 8.6, 3.576334375 };
const char* const RATES[6] = { "1_2", "3_5", "2_3", "3_4", "4_5", "5_6" };

int nbits_for(int ncells)
{
    int nd = 0;
    while ((1 << nd) < ncells)
        nd++;
    return nd;
}

/* L0, the cell interleaver's basic permutation (EN 302 755 clause 6.5;
 * dvbt2.c:2052-2081): an nd-bit PRBS with a toggling top bit, values at or
 * above ncells skipped. */
bool build_cellperm(int ncells, std::vector<int32_t>& perm)
{
    /* feedback taps of R' for nd = 11..15 */
    static const int taps[5][6] = { { 0, 3 }, { 0, 2 }, { 0, 1, 4, 6 }, { 0, 1, 4, 5, 9, 11 }, { 0, 1, 2, 12 } };
    static const int ntaps[5] = { 2, 2, 4, 6, 4 };
    const int nd = nbits_for(ncells);
    if (nd < 11 || nd > 15)
        return false;
    perm.assign(ncells, 0);
    int r = 0, q = 0;
    for (int i = 0; i < (1 << nd) && q < ncells; i++) {
        if (i >= 3) {
            int fb = 0;
            for (int t = 0; t < ntaps[nd - 11]; t++)
                fb ^= (r >> taps[nd - 11][t]) & 1;
            r = (r >> 1) | (fb << (nd - 2));
        } else {
            r = i == 2;
        }
        const int v = r | ((i & 1) << (nd - 1));
        if (v < ncells)
            perm[q++] = v;
    }
    return q == ncells;
}

/* Codeword bit n -> (cell, plane) as cell * m + plane, plane 2j = bit j of
 * I, 2j + 1 = bit j of Q: parity interleaving, column twist and the
 * demultiplexer undone (dvbt2.c:1969-2050, :2364-2385).  QPSK has no bit
 * interleaver (bit n is plane n % 2 of cell n / 2). */
bool build_bitsrc(int cod, int mod, bool sf, int N, int K, std::vector<uint16_t>& src)
{
    const int m = 2 * (mod + 1);
    src.assign(N, 0);
    if (m == 2) {
        for (int n = 0; n < N; n++)
            src[n] = (uint16_t)n;
        return true;
    }
    const int Q = (N - K) / 360;
    const int nc = (m == 8 && sf) ? 8 : 2 * m, nr = N / nc;
    const int* tw = sf ? TW_S[mod - 1] : TW_N[mod - 1];
    const int* mux = m == 4   ? (cod == 1 && !sf ? MUX16_35 : MUX16)
                     : m == 6 ? (cod == 1 && !sf ? MUX64_35 : MUX64)
                              : (sf ? MUX256S : cod == 1 ? MUX256_35 : cod == 2 ? MUX256_23 : MUX256);
    const bool pair = !(m == 8 && sf); /* two cells per demultiplexer word */
    for (int n = 0; n < N; n++) {
        const int x = n < K ? n : K + 360 * ((n - K) % Q) + (n - K) / Q;
        const int c = x / nr, row = (tw[c] + x % nr) % nr;
        const int cell = pair ? 2 * row + mux[c] / m : row;
        src[n] = (uint16_t)(cell * m + mux[c] % m);
    }
    return true;
}

struct T2Maps {
    int32_t* d_perm[2][4] = {};     /* by (fec_type, mod): L0, then its inverse (2 ncells) */
    uint16_t* d_src[2][6][4] = {};  /* by (fec_type, cod, mod) */
};
T2Maps g_maps[T2_MAX_DEV];
std::mutex g_maps_lock;

/* The PLP configuration's device tables, built at first use per device. */
int get_maps(int dev, int cod, int mod, int fec_type, int N, int K, const int32_t** perm, const uint16_t** src)
{
    std::lock_guard<std::mutex> lk(g_maps_lock);
    T2Maps& t = g_maps[dev];
    int32_t*& dp = t.d_perm[fec_type][mod];
    uint16_t*& ds = t.d_src[fec_type][cod][mod];
    if (!dp) {
        std::vector<int32_t> h;
        const int nc = N / (2 * (mod + 1));
        if (!build_cellperm(nc, h))
            return LDPC_GPU_EINVAL;
        h.resize(2 * (size_t)nc);
        for (int i = 0; i < nc; i++)
            h[nc + h[i]] = i;
        int32_t* p = nullptr;
        if (hipMalloc((void**)&p, sizeof(int32_t) * h.size()) != hipSuccess)
            return LDPC_GPU_ENOMEM;
        if (hipMemcpy(p, h.data(), sizeof(int32_t) * h.size(), hipMemcpyHostToDevice) != hipSuccess) {
            (void)hipFree(p);
            return LDPC_GPU_EDEVICE;
        }
        dp = p;
    }
    if (!ds) {
        std::vector<uint16_t> h;
        if (!build_bitsrc(cod, mod, fec_type == 0, N, K, h))
            return LDPC_GPU_EINVAL;
        uint16_t* p = nullptr;
        if (hipMalloc((void**)&p, sizeof(uint16_t) * h.size()) != hipSuccess)
            return LDPC_GPU_ENOMEM;
        if (hipMemcpy(p, h.data(), sizeof(uint16_t) * h.size(), hipMemcpyHostToDevice) != hipSuccess) {
            (void)hipFree(p);
            return LDPC_GPU_EDEVICE;
        }
        ds = p;
    }
    *perm = dp;
    *src = ds;
    return LDPC_GPU_OK;
}

struct T2Args {
    const float2* cells;
    const float* csi;
    const int32_t* perm;
    const int32_t* inv; /* inverse of perm */
    const uint16_t* src;
    int8_t* llr;
    float* err;
    float* sig;
    int ncells, N, nblocks, nti, nd, rot;
    float ct, sn, norm, ct2, sn2, s2, amax;
};

/* Shift P of block b (dvbt2.c:2230-2254): blocks are dealt to the nti
 * TI-blocks in order, the last nblocks % nti taking one more; the r-th
 * block of a TI-block takes the r-th nd-bit bit-reversed counter value
 * below ncells.  The whole workgroup scans 256 counter values per step
 * (wave ballots give each valid value its rank); every thread returns P.
 * The host keeps r below the number of valid values. */
__device__ int block_shift(int b, int nblocks, int nti, int ncells, int nd, int* s_P, int* s_wcount)
{
    if (!nti)
        return 0;
    int cum = 0;
    for (int s = 0; s < nti; s++) {
        const int ns = nblocks / nti + (s >= nti - nblocks % nti);
        if (b < cum + ns)
            break;
        cum += ns;
    }
    const int r = b - cum;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (tid == 0)
        *s_P = 0;
    int seen = 0;
    for (unsigned k0 = 0; k0 < (1u << nd); k0 += T2_BLOCK) {
        const unsigned k = k0 + tid;
        const int v = (int)(__brev(k) >> (32 - nd));
        const bool valid = k < (1u << nd) && v < ncells;
        const unsigned long long mask = __ballot(valid);
        if (lane == 0)
            s_wcount[wave] = __popcll(mask);
        __syncthreads();
        int off = seen, total = 0;
        for (int w = 0; w < T2_BLOCK / 64; w++) {
            off += w < wave ? s_wcount[w] : 0;
            total += s_wcount[w];
        }
        if (valid && off + __popcll(mask & ((1ull << lane) - 1)) == r)
            *s_P = v;
        __syncthreads();
        if (seen + total > r)
            break;
        seen += total;
    }
    return *s_P;
}

constexpr int T2_U = T2_CHUNK / T2_BLOCK; /* cells a thread takes per pass */

/* Derotated points of cells i0 + u * T2_BLOCK (u < T2_U) and their
 * per-axis weights (dvbt2.c:2257-2292): I from deinterleaved cell i, Q
 * from cell i + 1 (cyclically) when the constellation is rotated.  All
 * loads of the group are issued before any arithmetic (the cell gathers
 * depend on the permutation loads; a cell at a time would leave one chain
 * of two HBM latencies per thread in flight).  Cells past the end are
 * computed on cell ncells - 1 and left for the caller to drop. */
__device__ __forceinline__ void t2_points(const T2Args& a, const float2* in, const float* ic, int P, int i0,
                                          float* wr, float* wi, float* kr, float* ki)
{
#pragma clang fp contract(off)
    int s[T2_U], s1[T2_U];
#pragma unroll
    for (int u = 0; u < T2_U; u++) {
        const int i = min(i0 + u * T2_BLOCK, a.ncells - 1);
        s[u] = a.perm[i];
        s1[u] = a.perm[i + 1 == a.ncells ? 0 : i + 1];
    }
    float2 x[T2_U];
    float c1[T2_U], v[T2_U], c2[T2_U];
#pragma unroll
    for (int u = 0; u < T2_U; u++) {
        int t = s[u] + P;
        t -= t >= a.ncells ? a.ncells : 0;
        x[u] = in[t];
        c1[u] = ic[t];
        if (a.rot) {
            int t1 = s1[u] + P;
            t1 -= t1 >= a.ncells ? a.ncells : 0;
            v[u] = in[t1].y;
            c2[u] = ic[t1];
        }
    }
#pragma unroll
    for (int u = 0; u < T2_U; u++) {
        const float e1 = c1[u] + 1e-12f;
        const float q = a.rot ? v[u] : x[u].y;
        const float e2 = a.rot ? c2[u] + 1e-12f : e1;
        const float p = x[u].x;
        wr[u] = fmaf(a.ct, p, a.sn * q) * a.norm;
        wi[u] = fmaf(a.ct, q, -(a.sn * p)) * a.norm;
        kr[u] = ki[u] = e1;
        if (a.rot) {
            const float c12 = e1 * e2;
            kr[u] = c12 / fmaf(a.sn2, e1, a.ct2 * e2);
            ki[u] = c12 / fmaf(a.ct2, e1, a.sn2 * e2);
        }
    }
}

/* The same points with the block's cells already in LDS in deinterleaved
 * order (stage[i] = cell i); only the channel state is gathered. */
__device__ __forceinline__ void t2_points_staged(const T2Args& a, const float2* stage, const float* ic, int P,
                                                 int i0, float* wr, float* wi, float* kr, float* ki)
{
#pragma clang fp contract(off)
    int s[T2_U], s1[T2_U], ii[T2_U], ii1[T2_U];
#pragma unroll
    for (int u = 0; u < T2_U; u++) {
        const int i = min(i0 + u * T2_BLOCK, a.ncells - 1);
        ii[u] = i;
        ii1[u] = i + 1 == a.ncells ? 0 : i + 1;
        s[u] = a.perm[i];
        s1[u] = a.perm[ii1[u]];
    }
    float c1[T2_U], c2[T2_U];
#pragma unroll
    for (int u = 0; u < T2_U; u++) {
        int t = s[u] + P;
        t -= t >= a.ncells ? a.ncells : 0;
        c1[u] = ic[t];
        if (a.rot) {
            int t1 = s1[u] + P;
            t1 -= t1 >= a.ncells ? a.ncells : 0;
            c2[u] = ic[t1];
        }
    }
#pragma unroll
    for (int u = 0; u < T2_U; u++) {
        const float2 x = stage[ii[u]];
        const float e1 = c1[u] + 1e-12f;
        const float q = a.rot ? stage[ii1[u]].y : x.y;
        const float e2 = a.rot ? c2[u] + 1e-12f : e1;
        const float p = x.x;
        wr[u] = fmaf(a.ct, p, a.sn * q) * a.norm;
        wi[u] = fmaf(a.ct, q, -(a.sn * p)) * a.norm;
        kr[u] = ki[u] = e1;
        if (a.rot) {
            const float c12 = e1 * e2;
            kr[u] = c12 / fmaf(a.sn2, e1, a.ct2 * e2);
            ki[u] = c12 / fmaf(a.ct2, e1, a.sn2 * e2);
        }
    }
}

__device__ __forceinline__ int8_t t2_llr(float v)
{
    v = fminf(fmaxf(v, -127.0f), 127.0f);
    return (int8_t)(int)rintf(v);
}

/* MER terms of one chunk of cells into LDS, then the ordered sums: thread
 * l < 8 adds err lane l, thread 8 + l sig lane l, in cell order. */
__device__ __forceinline__ void chunk_sums(const T2Args& a, int base, const float* wr, const float* wi,
                                           const float* kr, const float* ki, float* tE, float* tS, float& acc)
{
#pragma clang fp contract(off)
    const int tid = threadIdx.x;
#pragma unroll
    for (int u = 0; u < T2_U; u++) {
        const int j = tid + u * T2_BLOCK;
        float te = 0.0f, ts = 0.0f;
        if (base + j < a.ncells) {
            const float awr = fabsf(wr[u]), awi = fabsf(wi[u]);
            const float ar = 2.0f * truncf(fminf(awr * 0.5f, a.amax)) + 1.0f;
            const float ai = 2.0f * truncf(fminf(awi * 0.5f, a.amax)) + 1.0f;
            const float er = awr - ar, ei = awi - ai;
            te = fmaf(er * er, kr[u], ei * ei * ki[u]);
            ts = a.s2 * (kr[u] + ki[u]);
        }
        tE[j] = te;
        tS[j] = ts;
    }
    __syncthreads();
    if (tid < 16) {
        const float* t = (tid < 8 ? tE : tS) + (tid & 7);
        const int cnt = min(T2_CHUNK, a.ncells - base);
#pragma unroll 8
        for (int q = 0; 8 * q < cnt; q++)
            acc = acc + t[8 * q];
    }
    __syncthreads();
}

/* Soft values of one chunk of cells by (cell, plane) (dvbt2.c:2333-2359). */
template <int M>
__device__ __forceinline__ void chunk_soft(const T2Args& a, int base, const float* wr, const float* wi,
                                           const float* kr, const float* ki, float k, int8_t* soft)
{
#pragma clang fp contract(off)
    constexpr int NB = M / 2;
#pragma unroll
    for (int u = 0; u < T2_U; u++) {
        const int i = base + threadIdx.x + u * T2_BLOCK;
        if (i >= a.ncells)
            continue;
        float ur = wr[u], ui = wi[u];
        int8_t v[M];
#pragma unroll
        for (int j = 0; j < NB; j++) {
            if (j) {
                const float t = (float)(1 << (NB - j));
                ur = fabsf(ur) - t;
                ui = fabsf(ui) - t;
            }
            v[2 * j] = t2_llr((ur * kr[u]) * k);
            v[2 * j + 1] = t2_llr((ui * ki[u]) * k);
        }
#pragma unroll
        for (int p = 0; p < M; p++)
            soft[i * M + p] = v[p];
    }
}

/* M bits per cell.  NCH > 0: the block has at most NCH chunks of cells
 * and each thread keeps its derotated points in registers between the MER
 * pass and the soft-value pass; NCH = 0: the second pass gathers the cells
 * again.  STAGE (with NCH > 0): the block's cells are first read in input
 * order into LDS, so the deinterleaving gathers hit LDS (for blocks whose
 * input would not stay in the L2 between chunks). */
template <int M, int NCH, bool STAGE>
__global__ __launch_bounds__(T2_BLOCK) void t2_demap_kernel(T2Args a)
{
#pragma clang fp contract(off)
    __shared__ __attribute__((aligned(16))) int8_t soft[T2_LDS];
    /* STAGE: the sum terms get their own LDS, and soft[] first holds the
     * block's cells in deinterleaved order (8 bytes a cell, at most
     * T2_LDS); it turns into the soft values once every thread holds its
     * points in registers */
    static_assert(!STAGE || NCH > 0, "staging needs the points in registers");
    __shared__ __attribute__((aligned(16))) float tsum[STAGE ? 2 * T2_CHUNK : 1];
    __shared__ float lane_sum[16];
    __shared__ int s_P, s_wcount[T2_BLOCK / 64];
    __shared__ float s_k;
    float* tE = STAGE ? tsum : (float*)soft;
    float* tS = tE + T2_CHUNK;
    const int tid = threadIdx.x, b = blockIdx.x;
    const int P = block_shift(b, a.nblocks, a.nti, a.ncells, a.nd, &s_P, s_wcount);
    const float2* in = a.cells + (size_t)b * a.ncells;
    const float* ic = a.csi + (size_t)b * a.ncells;
    float2* stage = (float2*)soft;
    if constexpr (STAGE) {
        /* input order, coalesced: input cell x is deinterleaved cell
         * inv[(x - P) mod ncells] */
        for (int x = tid; x < a.ncells; x += T2_BLOCK) {
            int y = x - P;
            y += y < 0 ? a.ncells : 0;
            stage[a.inv[y]] = in[x];
        }
        __syncthreads();
    }

    /* MER sums in the reference's order */
    float acc = 0.0f;
    constexpr int NR = NCH > 0 ? NCH : 1;
    float cwr[NR][T2_U], cwi[NR][T2_U], ckr[NR][T2_U], cki[NR][T2_U];
    if constexpr (NCH > 0) {
#pragma unroll
        for (int c = 0; c < NCH; c++) {
            const int base = c * T2_CHUNK;
            if (base < a.ncells) {
                if constexpr (STAGE)
                    t2_points_staged(a, stage, ic, P, base + tid, cwr[c], cwi[c], ckr[c], cki[c]);
                else
                    t2_points(a, in, ic, P, base + tid, cwr[c], cwi[c], ckr[c], cki[c]);
                chunk_sums(a, base, cwr[c], cwi[c], ckr[c], cki[c], tE, tS, acc);
            }
        }
    } else {
        for (int base = 0; base < a.ncells; base += T2_CHUNK) {
            t2_points(a, in, ic, P, base + tid, cwr[0], cwi[0], ckr[0], cki[0]);
            chunk_sums(a, base, cwr[0], cwi[0], ckr[0], cki[0], tE, tS, acc);
        }
    }
    if (tid < 16)
        lane_sum[tid] = acc;
    __syncthreads();
    if (tid == 0) {
        float err = 0.0f, sig = 0.0f;
        for (int l = 0; l < 8; l++) {
            err = err + lane_sum[l];
            sig = sig + lane_sum[8 + l];
        }
        a.err[b] = err;
        a.sig[b] = sig;
        const float n0u = err / (float)a.ncells + 1e-6f;
        s_k = 8.0f / n0u;
    }
    __syncthreads();
    const float k = s_k;

    /* soft values */
    if constexpr (NCH > 0) {
#pragma unroll
        for (int c = 0; c < NCH; c++)
            if (c * T2_CHUNK < a.ncells)
                chunk_soft<M>(a, c * T2_CHUNK, cwr[c], cwi[c], ckr[c], cki[c], k, soft);
    } else {
        for (int base = 0; base < a.ncells; base += T2_CHUNK) {
            t2_points(a, in, ic, P, base + tid, cwr[0], cwi[0], ckr[0], cki[0]);
            chunk_soft<M>(a, base, cwr[0], cwi[0], ckr[0], cki[0], k, soft);
        }
    }
    __syncthreads();

    /* bit deinterleaving as a gather, 4 codeword bits a thread and word;
     * the table words of 8 steps are loaded before their gathers */
    uint32_t* out = (uint32_t*)(a.llr + (size_t)b * a.N);
    const uint2* src = (const uint2*)a.src;
    const int nw = a.N / 4;
    for (int w0 = tid; w0 < nw; w0 += 8 * T2_BLOCK) {
        uint2 sv[8];
#pragma unroll
        for (int u = 0; u < 8; u++)
            sv[u] = src[min(w0 + u * T2_BLOCK, nw - 1)];
#pragma unroll
        for (int u = 0; u < 8; u++) {
            const int w = w0 + u * T2_BLOCK;
            if (w < nw) {
                const uint32_t b0 = (uint8_t)soft[sv[u].x & 0xFFFF], b1 = (uint8_t)soft[sv[u].x >> 16];
                const uint32_t b2 = (uint8_t)soft[sv[u].y & 0xFFFF], b3 = (uint8_t)soft[sv[u].y >> 16];
                out[w] = b0 | (b1 << 8) | (b2 << 16) | (b3 << 24);
            }
        }
    }
}

struct PlpGeom {
    int N, K, kbch, ncells, m, code;
    char rate[16];
};

int plp_geom(const ldpc_gpu_t2_plp* q, PlpGeom& g)
{
    if (!q || q->cod < 0 || q->cod > 5 || q->mod < 0 || q->mod > 3 || (q->fec_type != 0 && q->fec_type != 1))
        return LDPC_GPU_EINVAL;
    const bool sf = q->fec_type == 0;
    /* dvbt2.c:1966: T2's own 2/3 normal and 3/5 short codes */
    const char* name = !sf && q->cod == 2 ? "2_3_t2" : sf && q->cod == 1 ? "3_5_t2" : RATES[q->cod];
    std::strncpy(g.rate, name, sizeof g.rate - 1);
    g.rate[sizeof g.rate - 1] = 0;
    g.code = lg_code_lookup(sf, name);
    int nbch = 0, t = 0;
    if (g.code < 0 || lg_fec_lookup(sf, RATES[q->cod], &g.kbch, &nbch, &t))
        return LDPC_GPU_EINVAL;
    int qq, links, steps;
    if (ldpc_gpu_code_params(g.code, &g.N, &g.K, &qq, &links, &steps))
        return LDPC_GPU_EINVAL;
    g.m = 2 * (q->mod + 1);
    g.ncells = g.N / g.m;
    return LDPC_GPU_OK;
}

} // namespace

extern "C" int ldpc_gpu_t2_geometry(const ldpc_gpu_t2_plp* plp, int* nldpc, int* kldpc, int* kbch, int* ncells)
{
    PlpGeom g;
    if (plp_geom(plp, g))
        return LDPC_GPU_EINVAL;
    if (nldpc)
        *nldpc = g.N;
    if (kldpc)
        *kldpc = g.K;
    if (kbch)
        *kbch = g.kbch;
    if (ncells)
        *ncells = g.ncells;
    return LDPC_GPU_OK;
}

extern "C" int ldpc_gpu_t2_maps(const ldpc_gpu_t2_plp* plp, int nblocks, int32_t* perm, uint16_t* bitsrc,
                                int32_t* shift)
{
    PlpGeom g;
    if (plp_geom(plp, g) || nblocks < 0 || plp->til_length < 0 || plp->til_length > 255)
        return LDPC_GPU_EINVAL;
    if (perm) {
        std::vector<int32_t> h;
        if (!build_cellperm(g.ncells, h))
            return LDPC_GPU_EINVAL;
        std::memcpy(perm, h.data(), sizeof(int32_t) * h.size());
    }
    if (bitsrc) {
        std::vector<uint16_t> h;
        if (!build_bitsrc(plp->cod, plp->mod, plp->fec_type == 0, g.N, g.K, h))
            return LDPC_GPU_EINVAL;
        std::memcpy(bitsrc, h.data(), sizeof(uint16_t) * h.size());
    }
    if (shift) {
        /* the kernel's block_shift, on the host */
        const int nti = plp->til_length && plp->til_type == 0 ? plp->til_length : 0;
        const int nd = nbits_for(g.ncells);
        int s = 0, left = 0;
        unsigned k = 0;
        for (int b = 0; b < nblocks; b++) {
            shift[b] = 0;
            if (!nti)
                continue;
            while (left == 0 && s < nti) {
                left = nblocks / nti + (s >= nti - nblocks % nti);
                s++;
                k = 0;
            }
            for (; k < (1u << nd); k++) {
                unsigned v = 0;
                for (int j = 0; j < nd; j++)
                    v |= ((k >> j) & 1u) << (nd - 1 - j);
                if ((int)v < g.ncells) {
                    shift[b] = (int32_t)v;
                    k++;
                    break;
                }
            }
            left--;
        }
    }
    return LDPC_GPU_OK;
}

extern "C" int ldpc_gpu_t2_demap(ldpc_gpu_ctx* ctx, const ldpc_gpu_t2_plp* plp, int nblocks, const float* d_cells,
                                 const float* d_csi, int8_t* d_llr, float* d_err, float* d_sig)
{
    if (!ctx)
        return LDPC_GPU_EINVAL;
    PlpGeom g;
    if (plp_geom(plp, g))
        return lg_ctx_fail(ctx, LDPC_GPU_EINVAL, "t2_demap: unsupported PLP (cod 0..5, mod 0..3, fec_type 0/1)");
    if (nblocks < 0 || (nblocks > 0 && (!d_cells || !d_csi || !d_llr || !d_err || !d_sig)) ||
        plp->til_length < 0 || plp->til_length > 255)
        return lg_ctx_fail(ctx, LDPC_GPU_EINVAL, "t2_demap: bad arguments");
    if (nblocks == 0)
        return LDPC_GPU_OK;
    if (nblocks > 65535)
        return lg_ctx_fail(ctx, LDPC_GPU_EINVAL, "t2_demap: at most 65535 blocks per call");
    if (((uintptr_t)d_cells & 7) || ((uintptr_t)d_llr & 3))
        return lg_ctx_fail(ctx, LDPC_GPU_EINVAL, "t2_demap: d_cells must be 8-byte and d_llr 4-byte aligned");
    const int nti = plp->til_length && plp->til_type == 0 ? plp->til_length : 0;
    /* the counter of a TI-block runs up to twice its block count */
    if (nti && (nblocks / nti + 1) * 2 > (1 << nbits_for(g.ncells)))
        return lg_ctx_fail(ctx, LDPC_GPU_EINVAL, "t2_demap: more blocks per TI-block than the cell interleaver has shifts");
    const int dev = lg_ctx_device(ctx);
    if (dev < 0 || dev >= T2_MAX_DEV)
        return lg_ctx_fail(ctx, LDPC_GPU_EINVAL, "t2_demap: device index");
    if (hipSetDevice(dev) != hipSuccess)
        return lg_ctx_fail(ctx, LDPC_GPU_EDEVICE, "t2_demap: hipSetDevice");
    const int32_t* perm = nullptr;
    const uint16_t* src = nullptr;
    int rc = get_maps(dev, plp->cod, plp->mod, plp->fec_type, g.N, g.K, &perm, &src);
    if (rc)
        return lg_ctx_fail(ctx, rc, "t2_demap: deinterleaver tables");
    T2Args a;
    a.cells = (const float2*)d_cells;
    a.csi = d_csi;
    a.perm = perm;
    a.inv = perm + g.ncells;
    a.src = src;
    a.llr = d_llr;
    a.err = d_err;
    a.sig = d_sig;
    a.ncells = g.ncells;
    a.N = g.N;
    a.nblocks = nblocks;
    a.nti = nti;
    a.nd = nbits_for(g.ncells);
    a.rot = plp->rotation != 0;
    /* constants as dvbt2.c:2215-2230 forms them, in float */
    const double th = a.rot ? ROT_DEG[plp->mod] * M_PI / 180 : 0;
    a.ct = (float)std::cos(th);
    a.sn = (float)std::sin(th);
    a.norm = g.m == 2 ? (float)M_SQRT2 : g.m == 4 ? std::sqrt(10.0f) : g.m == 6 ? std::sqrt(42.0f) : std::sqrt(170.0f);
    a.ct2 = a.ct * a.ct;
    a.sn2 = a.sn * a.sn;
    a.s2 = a.norm * a.norm * 0.5f;
    a.amax = (float)((1 << (g.m / 2 - 1)) - 1);
    const hipStream_t st = (hipStream_t)ldpc_gpu_stream(ctx);
    /* points kept in registers when the block has at most 4 chunks; the
     * cells staged in LDS as well when the block's input is over 64 KB
     * (measured: 0.71 -> 0.58 ms for 4096 normal 256-QAM blocks, while
     * 2700-cell blocks, whose input stays in the L2, got slower) */
    const int nch = (g.ncells + T2_CHUNK - 1) / T2_CHUNK;
    const bool stage = nch <= 4 && (size_t)12 * g.ncells > 65536;
    if (stage && (size_t)8 * g.ncells > (size_t)T2_LDS)
        return lg_ctx_fail(ctx, LDPC_GPU_EINVAL, "t2_demap: block too large to stage");
    const dim3 grid(nblocks), blk(T2_BLOCK);
    switch (g.m) {
    case 2:
        /* QPSK: 8100 cells (short, staged) or 32400 */
        if (stage)
            hipLaunchKernelGGL((t2_demap_kernel<2, 4, true>), grid, blk, 0, st, a);
        else
            hipLaunchKernelGGL((t2_demap_kernel<2, 0, false>), grid, blk, 0, st, a);
        break;
    case 4:
        if (nch <= 2)
            hipLaunchKernelGGL((t2_demap_kernel<4, 2, false>), grid, blk, 0, st, a);
        else
            hipLaunchKernelGGL((t2_demap_kernel<4, 0, false>), grid, blk, 0, st, a);
        break;
    case 6:
        if (nch <= 2)
            hipLaunchKernelGGL((t2_demap_kernel<6, 2, false>), grid, blk, 0, st, a);
        else
            hipLaunchKernelGGL((t2_demap_kernel<6, 0, false>), grid, blk, 0, st, a);
        break;
    default:
        /* 256-QAM: 2025 cells (short) or 8100 (normal, staged) */
        if (nch <= 1)
            hipLaunchKernelGGL((t2_demap_kernel<8, 1, false>), grid, blk, 0, st, a);
        else if (stage)
            hipLaunchKernelGGL((t2_demap_kernel<8, 4, true>), grid, blk, 0, st, a);
        else
            hipLaunchKernelGGL((t2_demap_kernel<8, 0, false>), grid, blk, 0, st, a);
        break;
    }
    if (hipGetLastError() != hipSuccess)
        return lg_ctx_fail(ctx, LDPC_GPU_EDEVICE, "t2_demap: launch failed");
    return LDPC_GPU_OK;
}

extern "C" int ldpc_gpu_t2_decode(ldpc_gpu_ctx* ctx, const ldpc_gpu_t2_plp* plp, int nblocks, const float* d_cells,
                                  const float* d_csi, int max_iter, int lanes, int8_t* d_llr, uint8_t* d_cw,
                                  int32_t* d_iters, uint8_t* d_synd, int32_t* d_bch, float* d_err, float* d_sig)
{
    if (!ctx)
        return LDPC_GPU_EINVAL;
    PlpGeom g;
    if (plp_geom(plp, g))
        return lg_ctx_fail(ctx, LDPC_GPU_EINVAL, "t2_decode: unsupported PLP (cod 0..5, mod 0..3, fec_type 0/1)");
    if (max_iter < 0 || (nblocks > 0 && (!d_cw || !d_iters || !d_synd || !d_bch)))
        return lg_ctx_fail(ctx, LDPC_GPU_EINVAL, "t2_decode: bad arguments");
    if (((uintptr_t)d_llr & 7))
        return lg_ctx_fail(ctx, LDPC_GPU_EINVAL, "t2_decode: d_llr must be 8-byte aligned");
    int rc = ldpc_gpu_t2_demap(ctx, plp, nblocks, d_cells, d_csi, d_llr, d_err, d_sig);
    if (rc || nblocks == 0)
        return rc;
    const size_t stride = (size_t)g.K / 8;
    std::vector<ldpc_gpu_frame> fr(nblocks);
    for (int b = 0; b < nblocks; b++) {
        fr[b].code = g.code;
        fr[b].max_iter = max_iter;
        fr[b].nbits = g.K;
        /* lanes > 0: the blocks of one reference decoder call stop
         * together (dvbt2.c:2387-2396 decodes lanes blocks per call) */
        fr[b].group = lanes > 0 ? b / lanes : -1;
        fr[b].llr_off = (uint64_t)b * g.N;
        fr[b].out_off = (uint64_t)b * stride;
    }
    rc = ldpc_gpu_decode_batch(ctx, fr.data(), nblocks, d_llr, d_cw, d_iters, d_synd);
    if (rc)
        return rc;
    return ldpc_gpu_bch_decode(ctx, plp->fec_type == 0, RATES[plp->cod], nblocks, d_cw, stride, d_bch);
}

extern "C" int ldpc_gpu_t2_decode_host(ldpc_gpu_ctx* ctx, const ldpc_gpu_t2_plp* plp, int nblocks, const float* cells,
                                       const float* csi, int max_iter, int lanes, uint8_t* bb, uint8_t* ok,
                                       float* err, float* sig)
{
    if (!ctx)
        return LDPC_GPU_EINVAL;
    PlpGeom g;
    if (plp_geom(plp, g))
        return lg_ctx_fail(ctx, LDPC_GPU_EINVAL, "t2_decode_host: unsupported PLP (cod 0..5, mod 0..3, fec_type 0/1)");
    if (nblocks < 0 || (nblocks > 0 && (!cells || !csi || !bb || !ok || !err || !sig)))
        return lg_ctx_fail(ctx, LDPC_GPU_EINVAL, "t2_decode_host: bad arguments");
    if (nblocks == 0)
        return LDPC_GPU_OK;
    const int dev = lg_ctx_device(ctx);
    if (hipSetDevice(dev) != hipSuccess)
        return lg_ctx_fail(ctx, LDPC_GPU_EDEVICE, "t2_decode_host: hipSetDevice");
    const size_t nc = (size_t)nblocks * g.ncells, kb = (size_t)g.K / 8, bbb = (size_t)g.kbch / 8;
    float *d_cells = nullptr, *d_csi = nullptr, *d_err = nullptr, *d_sig = nullptr;
    int8_t* d_llr = nullptr;
    uint8_t *d_cw = nullptr, *d_syn = nullptr;
    int32_t *d_it = nullptr, *d_bch = nullptr;
    int rc = LDPC_GPU_OK;
    if (hipMalloc((void**)&d_cells, 8 * nc) != hipSuccess || hipMalloc((void**)&d_csi, 4 * nc) != hipSuccess ||
        hipMalloc((void**)&d_llr, (size_t)nblocks * g.N) != hipSuccess ||
        hipMalloc((void**)&d_cw, (size_t)nblocks * kb) != hipSuccess ||
        hipMalloc((void**)&d_it, sizeof(int32_t) * nblocks) != hipSuccess ||
        hipMalloc((void**)&d_syn, nblocks) != hipSuccess ||
        hipMalloc((void**)&d_bch, sizeof(int32_t) * nblocks) != hipSuccess ||
        hipMalloc((void**)&d_err, sizeof(float) * nblocks) != hipSuccess ||
        hipMalloc((void**)&d_sig, sizeof(float) * nblocks) != hipSuccess)
        rc = lg_ctx_fail(ctx, LDPC_GPU_ENOMEM, "t2_decode_host: device allocation failed");
    const hipStream_t st = (hipStream_t)ldpc_gpu_stream(ctx);
    if (!rc && (hipMemcpyAsync(d_cells, cells, 8 * nc, hipMemcpyHostToDevice, st) != hipSuccess ||
                hipMemcpyAsync(d_csi, csi, 4 * nc, hipMemcpyHostToDevice, st) != hipSuccess))
        rc = lg_ctx_fail(ctx, LDPC_GPU_EDEVICE, "t2_decode_host: upload failed");
    if (!rc)
        rc = ldpc_gpu_t2_decode(ctx, plp, nblocks, d_cells, d_csi, max_iter, lanes, d_llr, d_cw, d_it, d_syn, d_bch,
                                d_err, d_sig);
    std::vector<uint8_t> cw;
    std::vector<int32_t> res;
    if (!rc) {
        cw.resize((size_t)nblocks * kb);
        res.resize(nblocks);
        if (hipMemcpyAsync(cw.data(), d_cw, cw.size(), hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipMemcpyAsync(res.data(), d_bch, sizeof(int32_t) * nblocks, hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipMemcpyAsync(err, d_err, sizeof(float) * nblocks, hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipMemcpyAsync(sig, d_sig, sizeof(float) * nblocks, hipMemcpyDeviceToHost, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess)
            rc = lg_ctx_fail(ctx, LDPC_GPU_EDEVICE, "t2_decode_host: download failed");
    }
    (void)hipStreamSynchronize(st);
    (void)hipFree(d_cells);
    (void)hipFree(d_csi);
    (void)hipFree(d_llr);
    (void)hipFree(d_cw);
    (void)hipFree(d_it);
    (void)hipFree(d_syn);
    (void)hipFree(d_bch);
    (void)hipFree(d_err);
    (void)hipFree(d_sig);
    if (rc)
        return rc;
    /* the BBFRAME is the first kbch bits of the corrected BCH codeword
     * (dvbt2.c:2393-2394) */
    for (int b = 0; b < nblocks; b++) {
        std::memcpy(bb + (size_t)b * bbb, cw.data() + (size_t)b * kb, bbb);
        ok[b] = res[b] >= 0;
    }
    return LDPC_GPU_OK;
}
