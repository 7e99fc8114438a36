/*
 * Batched BCH decoding of the LDPC decoder's output (C-ABI
 * ldpc_gpu_bch_decode in include/ldpc_gpu.h).
 *
 * Same results as libdvbrx's drx_bch_decode (libdvbrx/src/bch.c:189-290)
 * for every DVB-S2/S2X/T2 FEC:
 *   - GF(2^16) for normal frames (x^16+x^5+x^3+x^2+1), GF(2^14) for short
 *     ones (x^14+x^5+x^3+x+1), as drx_bch_new (bch.c:106-128);
 *   - the codeword bit at index i (MSB first) has degree nbch - 1 - i; the
 *     syndromes S_j = c(alpha^j), j = 1 .. 2t, are all zero exactly when
 *     the generator (the minimal polynomials of alpha^1, alpha^3, ...,
 *     alpha^(2t-1)) divides c, the reference's remainder test; the even
 *     ones are squares of the odd ones;
 *   - the error locator by the inversionless Berlekamp-Massey algorithm
 *     (same register length and roots as bch.c:229-254's); more than t
 *     errors: -1;
 *   - Chien search over the nbch positions: a root at alpha^-p is an error
 *     at degree p (bch.c:256-289); a root count other than the locator's
 *     degree: -1, the codeword unchanged; else the bits are flipped and
 *     their count is the result.
 *
 * Work layout: one workgroup of 256 threads per codeword.  Odd syndromes:
 * each thread takes bytes strided by 256; a byte of lowest degree p0 adds
 * alpha^(j p0) times its own polynomial at alpha^j, whose log comes from a
 * 256-entry table per j in LDS (exponents reduced mod 2^m - 1 by folding),
 * then a workgroup XOR reduction.  Codewords with a
 * non-zero syndrome (few, after LDPC decoding) find the locator on one
 * thread and the Chien search on all, positions strided by 256 with
 * incremental exponents.  The field tables (exp, log) are read from
 * device memory (256 KB for GF(2^16), L2-resident).
 */
#include <hip/hip_runtime.h>

#include <cstring>
#include <mutex>
#include <vector>

#include "dvbs2_table_types.h"
#include "ldpc_codes.h"
#include "ldpc_gpu.h"
#include "ldpc_internal.h"

namespace {

constexpr int BCH_NT = 256;
constexpr int BCH_MAX_T = 12;
constexpr int BCH_MAX_DEV = 64;

/* exp (2q entries) and log (q + 1) of GF(2^m), m = 14 or 16, per device;
 * and the byte table: for odd j = 2k + 1 (k < 12) and byte value c, the
 * log of sum over set bits b of c of alpha^(j b), i.e. the byte's
 * polynomial at alpha^j (never 0 for c != 0: alpha^j's minimal polynomial
 * has degree m > 7); 0xFFFF for c = 0 */
struct GfTables {
    uint16_t* d_exp[2] = { nullptr, nullptr };
    uint16_t* d_log[2] = { nullptr, nullptr };
    uint16_t* d_blog[2] = { nullptr, nullptr };
};
static GfTables g_gf[BCH_MAX_DEV];
static std::mutex g_gf_lock;

static int gf_tables(int dev, int shortf, const uint16_t** exp, const uint16_t** log, const uint16_t** blog)
{
    std::lock_guard<std::mutex> lk(g_gf_lock);
    GfTables& g = g_gf[dev];
    if (!g.d_exp[shortf]) {
        const int m = shortf ? 14 : 16, q = (1 << m) - 1;
        const unsigned prim = shortf ? 0x402Bu : 0x1002Du;
        std::vector<uint16_t> e(2 * (size_t)q), l((size_t)q + 1, 0);
        unsigned x = 1;
        for (int i = 0; i < q; i++) {
            e[i] = e[i + q] = (uint16_t)x;
            l[x] = (uint16_t)i;
            x <<= 1;
            if (x >> m)
                x ^= prim;
        }
        std::vector<uint16_t> bl((size_t)BCH_MAX_T * 256, 0xFFFF);
        for (int k = 0; k < BCH_MAX_T; k++)
            for (int c = 1; c < 256; c++) {
                unsigned v = 0;
                for (int b = 0; b < 8; b++)
                    if (c >> b & 1)
                        v ^= e[((2 * k + 1) * b) % q];
                bl[(size_t)k * 256 + c] = v ? l[v] : 0xFFFF;
            }
        uint16_t *de = nullptr, *dl = nullptr, *db = nullptr;
        if (hipMalloc((void**)&de, sizeof(uint16_t) * e.size()) != hipSuccess)
            return LDPC_GPU_ENOMEM;
        if (hipMalloc((void**)&dl, sizeof(uint16_t) * l.size()) != hipSuccess) {
            (void)hipFree(de);
            return LDPC_GPU_ENOMEM;
        }
        if (hipMalloc((void**)&db, sizeof(uint16_t) * bl.size()) != hipSuccess) {
            (void)hipFree(de);
            (void)hipFree(dl);
            return LDPC_GPU_ENOMEM;
        }
        if (hipMemcpy(de, e.data(), sizeof(uint16_t) * e.size(), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(dl, l.data(), sizeof(uint16_t) * l.size(), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(db, bl.data(), sizeof(uint16_t) * bl.size(), hipMemcpyHostToDevice) != hipSuccess) {
            (void)hipFree(de);
            (void)hipFree(dl);
            (void)hipFree(db);
            return LDPC_GPU_EDEVICE;
        }
        g.d_exp[shortf] = de;
        g.d_log[shortf] = dl;
        g.d_blog[shortf] = db;
    }
    *exp = g.d_exp[shortf];
    *log = g.d_log[shortf];
    *blog = g.d_blog[shortf];
    return LDPC_GPU_OK;
}

struct BchArgs {
    uint8_t* cw;
    size_t stride;
    int32_t* result;
    const uint16_t* exp;
    const uint16_t* log;
    const uint16_t* blog; /* byte table, BCH_MAX_T x 256 */
    int nbch, t, m, q;
};

/* x mod q for q = 2^m - 1, x < 2^40 */
__device__ __forceinline__ uint32_t mod_q(uint64_t x, int m, uint32_t q)
{
    uint64_t r = (x & q) + (x >> m);
    r = (r & q) + (r >> m);
    r = (r & q) + (r >> m);
    return (uint32_t)(r >= q ? r - q : r);
}

__device__ __forceinline__ uint16_t gf_mul(const uint16_t* exp, const uint16_t* log, uint16_t x, uint16_t y)
{
    return x && y ? exp[(uint32_t)log[x] + log[y]] : 0;
}

__global__ void __launch_bounds__(BCH_NT) bch_kernel(BchArgs a)
{
    __shared__ uint32_t s_syn[BCH_MAX_T];
    __shared__ int s_len, s_nt, s_nroots, s_pos[BCH_MAX_T];
    __shared__ uint32_t s_lg[BCH_MAX_T + 1], s_step[BCH_MAX_T + 1], s_L0;
    const int tid = threadIdx.x, t = a.t, m = a.m, nbch = a.nbch;
    const uint32_t q = (uint32_t)a.q;
    uint8_t* cw = a.cw + a.stride * blockIdx.x;
    if (tid < BCH_MAX_T)
        s_syn[tid] = 0;
    if (tid == 0)
        s_nroots = 0;
    __syncthreads();

    /* ---- odd syndromes: bytes strided over the threads.  The byte at
     * index B holds degrees p0 .. p0 + 7 (p0 = nbch - 8 - 8 B, its LSB the
     * lowest), so it adds alpha^(j p0) times its polynomial at alpha^j:
     * one exponent sum with the byte table's log ---- */
    __shared__ uint16_t s_blog[BCH_MAX_T * 256];
    for (int i = tid; i < t * 256; i += BCH_NT)
        s_blog[i] = a.blog[i];
    __syncthreads();
    uint32_t syn[BCH_MAX_T];
#pragma unroll
    for (int k = 0; k < BCH_MAX_T; k++)
        syn[k] = 0;
    const int nbytes = nbch / 8;
    for (int B = tid; B < nbytes; B += BCH_NT) {
        const uint32_t c = cw[B];
        if (!c)
            continue;
        const uint32_t p0 = (uint32_t)(nbch - 8 - 8 * B);
#pragma unroll
        for (int k = 0; k < BCH_MAX_T; k++) {
            if (k < t) {
                const uint32_t lg = s_blog[k * 256 + c];
                syn[k] ^= a.exp[mod_q((uint64_t)(2 * k + 1) * p0 + lg, m, q)];
            }
        }
    }
    /* XOR over the workgroup: within each wave, then one LDS atomic per
     * wave and syndrome */
#pragma unroll
    for (int k = 0; k < BCH_MAX_T; k++) {
        uint32_t s = syn[k];
        for (int off = 32; off > 0; off >>= 1)
            s ^= __shfl_xor(s, off);
        if ((tid & 63) == 0 && k < t && s)
            atomicXor(&s_syn[k], s);
    }
    __syncthreads();

    /* ---- error locator on one thread: the inversionless
     * Berlekamp-Massey algorithm (Reed, Shih and Truong, IEE Proc. E 138,
     * 1991; the "iBM" recursion of Sarwate and Shanbhag, IEEE Trans. VLSI
     * 9(5), 2001).  With S_1 .. S_2t, for r = 0 .. 2t - 1:
     *   delta = sum_i lam_i S_(r+1-i)
     *   lam'  = gam lam + delta x bb
     *   if delta != 0 and kk >= 0: bb = lam, gam = delta, kk = -kk - 1
     *   else:                      bb = x bb,            kk = kk + 1
     * from lam = bb = 1, gam = 1, kk = 0.  kk = r - 2 L (L the register
     * length), so L = (2t - kk) / 2 at the end; lam is a non-zero multiple
     * of the classical locator, with the same roots.  L > t: more errors
     * than the code corrects. ---- */
    if (tid == 0) {
        constexpr int PL = 2 * BCH_MAX_T + 2; /* polynomial coefficients kept */
        uint16_t S[2 * BCH_MAX_T + 1];
        uint32_t any = 0;
        for (int k = 0; k < t; k++) {
            S[2 * k + 1] = (uint16_t)s_syn[k];
            any |= s_syn[k];
        }
        s_len = 0;
        if (any) {
            for (int j = 2; j <= 2 * t; j += 2)
                S[j] = gf_mul(a.exp, a.log, S[j / 2], S[j / 2]);
            uint16_t lam[PL], bb[PL];
            for (int i = 0; i < PL; i++)
                lam[i] = bb[i] = 0;
            lam[0] = bb[0] = 1;
            uint16_t gam = 1;
            int kk = 0;
            for (int r = 0; r < 2 * t; r++) {
                uint16_t delta = 0;
                for (int i = 0; i <= r && i < PL; i++)
                    delta ^= gf_mul(a.exp, a.log, lam[i], S[r + 1 - i]);
                uint16_t nl[PL];
                nl[0] = gf_mul(a.exp, a.log, gam, lam[0]);
                for (int i = 1; i < PL; i++)
                    nl[i] = gf_mul(a.exp, a.log, gam, lam[i]) ^ gf_mul(a.exp, a.log, delta, bb[i - 1]);
                if (delta && kk >= 0) {
                    for (int i = 0; i < PL; i++)
                        bb[i] = lam[i];
                    gam = delta;
                    kk = -kk - 1;
                } else {
                    for (int i = PL - 1; i > 0; i--)
                        bb[i] = bb[i - 1];
                    bb[0] = 0;
                    kk++;
                }
                for (int i = 0; i < PL; i++)
                    lam[i] = nl[i];
            }
            const int len = (2 * t - kk) / 2;
            if (len > t) {
                s_len = -1;
            } else {
                /* Chien terms: lam_k alpha^(-k p) as exponents log lam_k + (q - k) p */
                s_L0 = lam[0];
                int nt = 0;
                for (int k = 1; k <= len; k++)
                    if (lam[k]) {
                        s_lg[nt] = a.log[lam[k]];
                        s_step[nt] = q - (uint32_t)k;
                        nt++;
                    }
                s_len = len;
                s_nt = nt;
            }
        }
    }
    __syncthreads();
    const int len = s_len;
    if (len == 0) { /* a codeword */
        if (tid == 0)
            a.result[blockIdx.x] = 0;
        return;
    }
    if (len < 0) {
        if (tid == 0)
            a.result[blockIdx.x] = -1;
        return;
    }
    const int nt = s_nt;

    /* ---- Chien search: positions tid, tid + 256, ... ---- */
    uint32_t e[BCH_MAX_T], inc[BCH_MAX_T];
#pragma unroll
    for (int k = 0; k < BCH_MAX_T; k++) {
        if (k < nt) {
            e[k] = mod_q((uint64_t)s_lg[k] + (uint64_t)s_step[k] * (uint32_t)tid, m, q);
            inc[k] = mod_q((uint64_t)s_step[k] * BCH_NT, m, q);
        } else {
            e[k] = inc[k] = 0;
        }
    }
    const uint16_t L0 = (uint16_t)s_L0;
    for (int p = tid; p < nbch; p += BCH_NT) {
        uint32_t sum = L0;
#pragma unroll
        for (int k = 0; k < BCH_MAX_T; k++) {
            if (k < nt) {
                sum ^= a.exp[e[k]];
                e[k] += inc[k];
                if (e[k] >= q)
                    e[k] -= q;
            }
        }
        if (!sum) {
            const int idx = atomicAdd(&s_nroots, 1);
            if (idx < BCH_MAX_T)
                s_pos[idx] = p;
        }
    }
    __syncthreads();
    if (tid == 0) {
        const int nr = s_nroots;
        if (nr != len) {
            a.result[blockIdx.x] = -1;
        } else {
            for (int k = 0; k < nr; k++) {
                const int bit = nbch - 1 - s_pos[k];
                cw[bit >> 3] ^= (uint8_t)(0x80 >> (bit & 7));
            }
            a.result[blockIdx.x] = nr;
        }
    }
}

} // namespace

extern "C" int ldpc_gpu_bch_decode(ldpc_gpu_ctx* ctx, int short_frame, const char* rate, int nframes, uint8_t* d_cw,
                                   size_t stride, int32_t* d_result)
{
    if (!ctx)
        return LDPC_GPU_EINVAL;
    int kbch = 0, nbch = 0, t = 0;
    if (!rate || lg_fec_lookup(short_frame, rate, &kbch, &nbch, &t))
        return lg_ctx_fail(ctx, LDPC_GPU_EINVAL, "bch: unknown rate");
    if (t < 1 || t > BCH_MAX_T)
        return lg_ctx_fail(ctx, LDPC_GPU_EINVAL, "bch: t out of range");
    if (nframes < 0 || (nframes > 0 && (!d_cw || !d_result)) || stride < (size_t)(nbch + 7) / 8)
        return lg_ctx_fail(ctx, LDPC_GPU_EINVAL, "bch: bad arguments (stride below the codeword's bytes?)");
    if (nframes == 0)
        return LDPC_GPU_OK;
    if (nbch % 8)
        return lg_ctx_fail(ctx, LDPC_GPU_EINVAL, "bch: codeword not a whole number of bytes");
    const int dev = lg_ctx_device(ctx);
    if (dev < 0 || dev >= BCH_MAX_DEV)
        return lg_ctx_fail(ctx, LDPC_GPU_EINVAL, "bch: device index");
    if (hipSetDevice(dev) != hipSuccess)
        return lg_ctx_fail(ctx, LDPC_GPU_EDEVICE, "bch: hipSetDevice");
    const int shortf = short_frame ? 1 : 0;
    BchArgs a;
    int rc = gf_tables(dev, shortf, &a.exp, &a.log, &a.blog);
    if (rc)
        return lg_ctx_fail(ctx, rc, "bch: field tables");
    a.cw = d_cw;
    a.stride = stride;
    a.result = d_result;
    a.nbch = nbch;
    a.t = t;
    a.m = shortf ? 14 : 16;
    a.q = (1 << a.m) - 1;
    hipLaunchKernelGGL(bch_kernel, dim3(nframes), dim3(BCH_NT), 0, (hipStream_t)ldpc_gpu_stream(ctx), a);
    if (hipGetLastError() != hipSuccess)
        return lg_ctx_fail(ctx, LDPC_GPU_EDEVICE, "bch: launch failed");
    return LDPC_GPU_OK;
}
