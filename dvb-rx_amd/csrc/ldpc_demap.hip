/*
 * Max-log demapper: received symbols -> int8 LLRs in codeword order, the
 * decoder's input (C-ABI ldpc_gpu_demap in include/ldpc_gpu.h).
 *
 * Same results as libdvbrx's drx_demap_frame (libdvbrx/src/demap.c:491-532)
 * in its AVX2/FMA build, for every DVB-S2/S2X MODCOD:
 *   - constellation points at unit average energy (drx_demap_new,
 *     demap.c:184-196, the normalisation in double);
 *   - per point the metric x * ar + (y * ai + e), ar = -2 re, ai = -2 im,
 *     e = |p|^2, both multiply-adds fused as minima_avx / minima_grp
 *     (demap.c:318-425); the metric drops the |y|^2 all points share;
 *   - the constellations of 8 rings of 32 phases take the ring form of
 *     minima_ring (demap.c:457-489): projections c_j = x rc_j + y rs_j,
 *     phase minima over the rings of c_j (-2 R_r) + R_r^2, ring minima at
 *     the largest projection;
 *   - per label bit b (MSB first) the minima over the points whose bit b is
 *     0 and 1, LLR = rint(clamp(k (m1 - m0), +-127)) with k = 2 / n0
 *     (LLR_SCALE), written at order[b] * nsym + j (the bit interleaver),
 *     positions past the codeword dropped (128APSK pads 6 bits);
 *   - QPSK has no interleaver: LLR = rint(clamp(k 4 sqrt(1/2) x)) per axis.
 *
 * Work layout: one thread per symbol, a block of 256 symbols of one frame;
 * the MODCOD's point tables (3 KB) are staged in LDS per block; the full
 * search goes by blocks of 16 points (minima_full).  The minima are exact,
 * so only the metric's arithmetic (written with explicit fused
 * multiply-adds) has to follow the reference.
 */
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <vector>

#include "dvbs2_table_types.h"
#include "ldpc_codes.h"
#include "ldpc_gpu.h"
#include "ldpc_internal.h"

namespace {

constexpr int DM_BLOCK = 256;
constexpr int DM_MAX_DEV = 64;

/* point tables of one MODCOD, as the reference's drx_demap fields */
struct DemapTab {
    float ar[256], ai[256], e[256];
    float rc[32], rs[32], rm2[8], rr2[8];
    int bits, np, nsym, nldpc, ring, qpsk;
    int order[8];
};

/* drx_demap_new (demap.c:164-203) and detect_rings (:139-162) */
static bool build_tab(const lg_modcod_row* row, DemapTab& d)
{
    memset(&d, 0, sizeof(d));
    d.bits = row->bits;
    d.np = 1 << row->bits;
    d.nldpc = row->short_frame ? 16200 : 64800;
    d.nsym = (d.nldpc + d.bits - 1) / d.bits;
    for (int b = 0; b < 8; b++)
        d.order[b] = row->order[b];
    d.qpsk = row->order[0] < 0;
    const float* pt = lg_points() + 2 * row->point_off;
    double en = 0;
    for (int p = 0; p < d.np; p++) {
        const float a = pt[2 * p] * pt[2 * p], b = pt[2 * p + 1] * pt[2 * p + 1];
        en += (double)(a + b);
    }
    const float g = (float)(1.0 / std::sqrt(en / d.np));
    float re[256], im[256];
    for (int p = 0; p < d.np; p++) {
        re[p] = pt[2 * p] * g;
        im[p] = pt[2 * p + 1] * g;
        d.ar[p] = -2 * re[p];
        d.ai[p] = -2 * im[p];
        d.e[p] = std::fma(re[p], re[p], im[p] * im[p]);
    }
    if (d.bits == 8) {
        for (int r = 0; r < 8; r++) {
            const float R = hypotf(re[r << 5], im[r << 5]);
            d.rm2[r] = -2 * R;
            d.rr2[r] = R * R;
        }
        const float R0 = hypotf(re[0], im[0]);
        for (int j = 0; j < 32; j++) {
            d.rc[j] = re[j] / R0;
            d.rs[j] = im[j] / R0;
        }
        d.ring = 1;
        for (int p = 0; p < d.np && d.ring; p++) {
            const float R = -0.5f * d.rm2[p >> 5];
            if (std::fabs(R * d.rc[p & 31] - re[p]) > 1e-5f || std::fabs(R * d.rs[p & 31] - im[p]) > 1e-5f)
                d.ring = 0;
        }
    }
    return true;
}

/* all MODCODs' tables on one device, built at first use; index by PLS/2 */
struct DemapTables {
    DemapTab* d_tab = nullptr;
    int index[128];
    std::vector<DemapTab> host;
};
static DemapTables g_tabs[DM_MAX_DEV];
static std::mutex g_tabs_lock;

static int get_tables(int dev, DemapTables** out)
{
    std::lock_guard<std::mutex> lk(g_tabs_lock);
    DemapTables& t = g_tabs[dev];
    if (!t.d_tab) {
        for (int k = 0; k < 128; k++) {
            t.index[k] = -1;
            const lg_modcod_row* row = lg_modcod_get(2 * k);
            if (!row || row->pls != 2 * k)
                continue;
            DemapTab d;
            if (!build_tab(row, d))
                continue;
            t.index[k] = (int)t.host.size();
            t.host.push_back(d);
        }
        DemapTab* p = nullptr;
        if (hipMalloc((void**)&p, sizeof(DemapTab) * t.host.size()) != hipSuccess)
            return LDPC_GPU_ENOMEM;
        if (hipMemcpy(p, t.host.data(), sizeof(DemapTab) * t.host.size(), hipMemcpyHostToDevice) != hipSuccess) {
            (void)hipFree(p);
            return LDPC_GPU_EDEVICE;
        }
        t.d_tab = p;
    }
    *out = &t;
    return LDPC_GPU_OK;
}

__device__ __forceinline__ int8_t quantise(float v)
{
    v = v > 127.0f ? 127.0f : v < -127.0f ? -127.0f : v;
    return (int8_t)__builtin_rintf(v);
}

/* per-bit minima of the full search over NB-bit labels.  From 16 points
 * on, blocks of 16 consecutive labels: the 4 low label bits are resolved
 * inside a block by a fixed tree of minima (pairs, quads, halves), the
 * high bits by the block's minimum; minima are exact, so the grouping does
 * not change them. */
typedef const __attribute__((address_space(4))) float dm_cf32;

/* the point tables are read at uniform indices: from the global table
 * through the constant address space they become scalar loads (16 points
 * a block in three 64-byte loads) whose values feed the multiply-adds
 * directly */
template <int NB>
__device__ __forceinline__ void minima_full(const dm_cf32* ar, const dm_cf32* ai, const dm_cf32* e, float x, float y,
                                            float* m0, float* m1)
{
#pragma unroll
    for (int b = 0; b < NB; b++)
        m0[b] = m1[b] = INFINITY;
    if constexpr (NB < 4) {
#pragma unroll
        for (int p = 0; p < (1 << NB); p++) {
            const float t = fmaf(x, ar[p], fmaf(y, ai[p], e[p]));
#pragma unroll
            for (int b = 0; b < NB; b++) {
                if ((p >> (NB - 1 - b)) & 1)
                    m1[b] = fminf(m1[b], t);
                else
                    m0[b] = fminf(m0[b], t);
            }
        }
    } else {
#pragma unroll 1
        for (int g = 0; g < (1 << (NB - 4)); g++) {
            float v[16];
#pragma unroll
            for (int u = 0; u < 16; u++) {
                const int p = 16 * g + u;
                v[u] = fmaf(x, ar[p], fmaf(y, ai[p], e[p]));
            }
            /* label bit NB-1-L is bit L of the index within the block */
            float pr[8], qd[4];
#pragma unroll
            for (int k = 0; k < 8; k++)
                pr[k] = fminf(v[2 * k], v[2 * k + 1]);
#pragma unroll
            for (int k = 0; k < 4; k++)
                qd[k] = fminf(pr[2 * k], pr[2 * k + 1]);
            float l0 = v[0], h0 = v[1];
#pragma unroll
            for (int k = 1; k < 8; k++) {
                l0 = fminf(l0, v[2 * k]);
                h0 = fminf(h0, v[2 * k + 1]);
            }
            const float l1 = fminf(fminf(pr[0], pr[2]), fminf(pr[4], pr[6]));
            const float h1 = fminf(fminf(pr[1], pr[3]), fminf(pr[5], pr[7]));
            const float l2 = fminf(qd[0], qd[2]), h2 = fminf(qd[1], qd[3]);
            const float l3 = fminf(qd[0], qd[1]), h3 = fminf(qd[2], qd[3]);
            m0[NB - 1] = fminf(m0[NB - 1], l0);
            m1[NB - 1] = fminf(m1[NB - 1], h0);
            m0[NB - 2] = fminf(m0[NB - 2], l1);
            m1[NB - 2] = fminf(m1[NB - 2], h1);
            m0[NB - 3] = fminf(m0[NB - 3], l2);
            m1[NB - 3] = fminf(m1[NB - 3], h2);
            m0[NB - 4] = fminf(m0[NB - 4], l3);
            m1[NB - 4] = fminf(m1[NB - 4], h3);
            const float gmin = fminf(l3, h3);
            /* high label bits NB-1-L, L >= 4: bit L - 4 of the block index */
#pragma unroll
            for (int L = 4; L < NB; L++) {
                if ((g >> (L - 4)) & 1)
                    m1[NB - 1 - L] = fminf(m1[NB - 1 - L], gmin);
                else
                    m0[NB - 1 - L] = fminf(m0[NB - 1 - L], gmin);
            }
        }
    }
}

/* 8 rings x 32 phases (minima_ring) */
__device__ __forceinline__ void minima_ring(const DemapTab& d, float x, float y, float* m0, float* m1)
{
#pragma unroll
    for (int b = 0; b < 8; b++)
        m0[b] = m1[b] = INFINITY;
    float cmax = -INFINITY;
    for (int q = 0; q < 32; q++) {
        const float c = fmaf(x, d.rc[q], y * d.rs[q]);
        cmax = fmaxf(cmax, c);
        float pm = INFINITY;
#pragma unroll
        for (int r = 0; r < 8; r++)
            pm = fminf(pm, fmaf(c, d.rm2[r], d.rr2[r]));
        /* bit L of the phase index is label bit 7 - L */
#pragma unroll
        for (int L = 0; L < 5; L++) {
            if ((q >> L) & 1)
                m1[7 - L] = fminf(m1[7 - L], pm);
            else
                m0[7 - L] = fminf(m0[7 - L], pm);
        }
    }
#pragma unroll
    for (int r = 0; r < 8; r++) {
        const float t = fmaf(cmax, d.rm2[r], d.rr2[r]);
#pragma unroll
        for (int L = 0; L < 3; L++) {
            if ((r >> L) & 1)
                m1[2 - L] = fminf(m1[2 - L], t);
            else
                m0[2 - L] = fminf(m0[2 - L], t);
        }
    }
}

template <int NB>
__global__ void __launch_bounds__(DM_BLOCK) demap_kernel(const DemapTab* tab, const float2* sym, const float* n0,
                                                         int8_t* llr)
{
    __shared__ DemapTab s;
    /* stage the tables (ar, ai, e and the rest) in LDS */
    {
        const uint32_t* src = (const uint32_t*)tab;
        uint32_t* dst = (uint32_t*)&s;
        for (int k = threadIdx.x; k < (int)(sizeof(DemapTab) / 4); k += DM_BLOCK)
            dst[k] = src[k];
    }
    __syncthreads();
    const int f = blockIdx.y, j = blockIdx.x * DM_BLOCK + threadIdx.x;
    const int nsym = s.nsym, nldpc = s.nldpc;
    if (j >= nsym)
        return;
    const float2 v = sym[(size_t)f * nsym + j];
    const float k = 2.0f / n0[f];
    int8_t* o = llr + (size_t)f * nldpc;
    if (NB == 2) {
        const float kq = k * 4.0f * 0.70710678118654752440f;
        o[2 * j] = quantise(kq * v.x);
        o[2 * j + 1] = quantise(kq * v.y);
        return;
    }
    float m0[8], m1[8];
    if (NB == 8 && s.ring)
        minima_ring(s, v.x, v.y, m0, m1);
    else
        minima_full<NB>((const dm_cf32*)tab->ar, (const dm_cf32*)tab->ai, (const dm_cf32*)tab->e, v.x, v.y, m0, m1);
#pragma unroll
    for (int b = 0; b < NB; b++) {
        const int pos = s.order[b] * nsym + j;
        if (pos < nldpc)
            o[pos] = quantise(k * (m1[b] - m0[b]));
    }
}

} // namespace

extern "C" int ldpc_gpu_demap_symbols(int pls)
{
    const lg_modcod_row* row = lg_modcod_get(pls);
    if (!row)
        return -1;
    const int nldpc = row->short_frame ? 16200 : 64800;
    return (nldpc + row->bits - 1) / row->bits;
}

/* the demap kernel of MODCOD pls over nframes frames on stream st of
 * device dev (shared by ldpc_gpu_demap and the GPU channel) */
int lg_demap_launch(int dev, int pls, int nframes, const float* d_sym, const float* d_n0, int8_t* d_llr,
                    hipStream_t st)
{
    if (dev < 0 || dev >= DM_MAX_DEV)
        return LDPC_GPU_EINVAL;
    DemapTables* t = nullptr;
    int rc = get_tables(dev, &t);
    if (rc)
        return rc;
    const int idx = t->index[(pls & 0xFE) >> 1];
    if (idx < 0)
        return LDPC_GPU_EINVAL;
    const DemapTab& h = t->host[idx];
    const dim3 grid((h.nsym + DM_BLOCK - 1) / DM_BLOCK, nframes);
    const DemapTab* tab = t->d_tab + idx;
    const float2* sym = (const float2*)d_sym;
    switch (h.qpsk ? 2 : h.bits) {
    case 2:
        hipLaunchKernelGGL(demap_kernel<2>, grid, dim3(DM_BLOCK), 0, st, tab, sym, d_n0, d_llr);
        break;
    case 3:
        hipLaunchKernelGGL(demap_kernel<3>, grid, dim3(DM_BLOCK), 0, st, tab, sym, d_n0, d_llr);
        break;
    case 4:
        hipLaunchKernelGGL(demap_kernel<4>, grid, dim3(DM_BLOCK), 0, st, tab, sym, d_n0, d_llr);
        break;
    case 5:
        hipLaunchKernelGGL(demap_kernel<5>, grid, dim3(DM_BLOCK), 0, st, tab, sym, d_n0, d_llr);
        break;
    case 6:
        hipLaunchKernelGGL(demap_kernel<6>, grid, dim3(DM_BLOCK), 0, st, tab, sym, d_n0, d_llr);
        break;
    case 7:
        hipLaunchKernelGGL(demap_kernel<7>, grid, dim3(DM_BLOCK), 0, st, tab, sym, d_n0, d_llr);
        break;
    case 8:
        hipLaunchKernelGGL(demap_kernel<8>, grid, dim3(DM_BLOCK), 0, st, tab, sym, d_n0, d_llr);
        break;
    default:
        return LDPC_GPU_EINVAL;
    }
    return hipGetLastError() == hipSuccess ? LDPC_GPU_OK : LDPC_GPU_EDEVICE;
}

extern "C" int ldpc_gpu_demap(ldpc_gpu_ctx* ctx, int pls, int nframes, const float* d_sym, const float* d_n0,
                              int8_t* d_llr)
{
    if (!ctx)
        return LDPC_GPU_EINVAL;
    const lg_modcod_row* row = lg_modcod_get(pls);
    if (!row)
        return lg_ctx_fail(ctx, LDPC_GPU_EINVAL, "demap: unknown PLS code");
    if (nframes < 0 || (nframes > 0 && (!d_sym || !d_n0 || !d_llr)))
        return lg_ctx_fail(ctx, LDPC_GPU_EINVAL, "demap: bad arguments");
    if (nframes == 0)
        return LDPC_GPU_OK;
    if (nframes > 65535)
        return lg_ctx_fail(ctx, LDPC_GPU_EINVAL, "demap: at most 65535 frames per call");
    const int dev = lg_ctx_device(ctx);
    if (hipSetDevice(dev) != hipSuccess)
        return lg_ctx_fail(ctx, LDPC_GPU_EDEVICE, "demap: hipSetDevice");
    const int rc = lg_demap_launch(dev, pls, nframes, d_sym, d_n0, d_llr, (hipStream_t)ldpc_gpu_stream(ctx));
    if (rc)
        return lg_ctx_fail(ctx, rc, rc == LDPC_GPU_EDEVICE ? "demap: launch failed" : "demap: no constellation for this PLS code");
    return LDPC_GPU_OK;
}
