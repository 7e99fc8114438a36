/*
 * Learned (neural) offset min-sum: training kernels and the GPU channel
 * (C-ABI of include/ldpc_nms.h).
 *
 * Training model.  The forward pass is the integer layered decoder of the
 * reference (libdvbrx/src/ldpc.c:321-362, :465-497) with the check-node
 * magnitude replaced by clamp((min(m, 127) * S - B) >> 4, 0, 127), run
 * for T iterations without early stop.  The backward pass differentiates
 * the float model mag = clamp(alpha * min(m, 127) - beta, 0, 127) at the
 * forward's values (straight-through): the gradient of o_e reaches
 * alpha, beta and the one input that sets min over the other edges; the
 * posterior and message saturations pass it only inside their ranges.
 *
 * Work layout: one frame per workgroup (persistent over the batch).  The
 * int8 posteriors of the frame sit in LDS in the decoder's variable order
 * (information bits, then parity in layer order, bit K + q*j + i at
 * K + 360*i + j).  Node j of a layer is thread j; the nodes of a layer run
 * in dependency levels (ldpc_codes.h), which reproduces the reference's
 * node order; the backward pass visits the same levels in reverse.  Per
 * edge "slot" (layer i, edge c, node j at slot_base_i + 360*c + j) the
 * forward saves the raw input L - m as int16 for every iteration; the
 * posterior gradient (one float per variable) and message gradient (one
 * per slot) live in the workgroup's block of device memory.
 */
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "dvbs2_table_types.h"
#include "ldpc_internal.h"
#include "ldpc_codes.h"
#include "ldpc_gpu.h"
#include "ldpc_nms.h"

#define NT 384
#define DEGS LDPC_NMS_DEGS
#define NO_VAR 0xffffffffu

namespace {

struct NmsArgs {
    const uint4* lay;     /* per layer: first slot, information degree, level count */
    const uint32_t* vidx; /* per slot: variable in decoder order */
    const uint8_t* level; /* per node i*360+j */
    int N, K, q, T, lts;
    const int8_t* llr;
    const uint8_t* cw;
    const int32_t* row;
    int nframes;
    const float* params; /* alpha rows then beta rows */
    const float* w;
    float temp, inv;
    int16_t* raw; /* per workgroup: T x lts */
    int8_t* msg;  /* per workgroup: lts */
    int8_t* Ls;   /* per workgroup: T x N posteriors after each iteration */
    float* gL;    /* per workgroup: N */
    float* gm;    /* per workgroup: lts */
    float* part;  /* per workgroup: 2 T DEGS gradients + loss */
    int8_t* post; /* optional: nframes x N final posteriors, codeword order */
};

__device__ __forceinline__ int sat8(int v) { return v > 127 ? 127 : v < -128 ? -128 : v; }

__device__ __forceinline__ int cw_pos(int v, int K, int q)
{
    if (v < K)
        return v;
    const int u = v - K;
    return K + q * (u % 360) + u / 360;
}

__device__ __forceinline__ int nms_mag(int m, int S, int B)
{
    const int v = (min(m, 127) * S - B) >> 4;
    return v < 0 ? 0 : v > 127 ? 127 : v;
}

/* smallest and second smallest |x| (first index on ties), the signs of
 * those two inputs (-1 / +1, zero counts positive), negative count */
template <int DM>
__device__ __forceinline__ void node_minima(const int* x, int deg, int& m0, int& m1, int& i0, int& i1, float& s0,
                                            float& s1, int& nneg)
{
    m0 = 1 << 20;
    m1 = 1 << 20;
    i0 = -1;
    i1 = -1;
    s0 = 1.0f;
    s1 = 1.0f;
    nneg = 0;
#pragma unroll
    for (int c = 0; c < DM; c++) {
        if (c < deg) {
            const int ab = x[c] < 0 ? -x[c] : x[c];
            const float sc = x[c] < 0 ? -1.0f : 1.0f;
            if (ab < m0) {
                m1 = m0;
                i1 = i0;
                s1 = s0;
                m0 = ab;
                i0 = c;
                s0 = sc;
            } else if (ab < m1) {
                m1 = ab;
                i1 = c;
                s1 = sc;
            }
            nneg += x[c] < 0;
        }
    }
}

template <int DM>
__device__ void fwd_node(const NmsArgs& a, int8_t* post, uint4 ly, int i, int j, int first, int16_t* raw_t,
                         int8_t* msg, const int* sS, const int* sB)
{
    const int cnt = (int)ly.y, sb = (int)ly.x;
    const int deg = cnt + 1 + (i > 0 || j > 0);
    int x[DM], v[DM];
#pragma unroll
    for (int c = 0; c < DM; c++) {
        if (c < deg) {
            const int slot = sb + c * 360 + j;
            v[c] = (int)a.vidx[slot];
            const int L = post[v[c]];
            const int rw = first ? L : L - msg[slot];
            raw_t[slot] = (int16_t)rw;
            x[c] = sat8(rw);
        }
    }
    int m0, m1, i0, i1, nneg;
    float s0, s1;
    node_minima<DM>(x, deg, m0, m1, i0, i1, s0, s1, nneg);
    const int mn0 = nms_mag(m0, sS[deg], sB[deg]), mn1 = nms_mag(m1, sS[deg], sB[deg]);
#pragma unroll
    for (int c = 0; c < DM; c++) {
        if (c < deg) {
            const int mag = c == i0 ? mn1 : mn0;
            const int o = ((nneg - (x[c] < 0)) & 1) ? -mag : mag;
            post[v[c]] = (int8_t)sat8(x[c] + o);
            msg[sb + c * 360 + j] = (int8_t)(o < -32 ? -32 : o > 31 ? 31 : o);
        }
    }
}

template <int DM>
__device__ void bwd_node(const NmsArgs& a, float* gL, float* gm, const int16_t* raw_t, uint4 ly, int i, int j, int t,
                         const int* sS, const int* sB, const float* sA, const float* sBe, float* sg)
{
    const int cnt = (int)ly.y, sb = (int)ly.x;
    const int deg = cnt + 1 + (i > 0 || j > 0);
    int x[DM], v[DM], rw[DM];
#pragma unroll
    for (int c = 0; c < DM; c++) {
        if (c < deg) {
            const int slot = sb + c * 360 + j;
            v[c] = (int)a.vidx[slot];
            rw[c] = raw_t[slot];
            x[c] = sat8(rw[c]);
        }
    }
    int m0, m1, i0, i1, nneg;
    float s0, s1;
    node_minima<DM>(x, deg, m0, m1, i0, i1, s0, s1, nneg);
    const int S = sS[deg], B = sB[deg];
    const float al = sA[deg], be = sBe[deg];
    float gLn[DM];
    float g0 = 0.0f, g1 = 0.0f, dga = 0.0f, dgb = 0.0f; /* gradient to inputs i0 / i1 */
#pragma unroll
    for (int c = 0; c < DM; c++) {
        if (c < deg) {
            const int slot = sb + c * 360 + j;
            const int mu = c == i0 ? m1 : m0;
            const int mag = nms_mag(mu, S, B);
            const bool neg = (nneg - (x[c] < 0)) & 1;
            const int o = neg ? -mag : mag;
            const int Ln = x[c] + o;
            const float gl = Ln >= -128 && Ln <= 127 ? gL[v[c]] : 0.0f;
            const float go = gl + (o >= -32 && o <= 31 ? gm[slot] : 0.0f);
            gLn[c] = gl;
            const float mup = (float)min(mu, 127);
            const float pre = __fsub_rn(__fmul_rn(al, mup), be);
            if (pre > 0.0f && pre < 127.0f) {
                const float dmag = neg ? -go : go;
                dga += dmag * mup;
                dgb -= dmag;
                if (mu < 127) {
                    if (c == i0)
                        g1 += dmag * al * s1;
                    else
                        g0 += dmag * al * s0;
                }
            }
        }
    }
#pragma unroll
    for (int c = 0; c < DM; c++) {
        if (c < deg) {
            const int slot = sb + c * 360 + j;
            const float gx = gLn[c] + (c == i0 ? g0 : 0.0f) + (c == i1 ? g1 : 0.0f);
            const float g = rw[c] >= -128 && rw[c] <= 127 ? gx : 0.0f;
            gL[v[c]] = g;
            gm[slot] = t ? -g : 0.0f;
        }
    }
    atomicAdd(&sg[t * DEGS + deg], dga);
    atomicAdd(&sg[(a.T + t) * DEGS + deg], dgb);
}

template <int DM>
__global__ void __launch_bounds__(NT) nms_train_kernel(NmsArgs a)
{
    extern __shared__ __align__(16) int8_t lds[];
    const int T = a.T, np = T * DEGS;
    int* sS = (int*)lds;
    int* sB = sS + np;
    float* sA = (float*)(sB + np);
    float* sBe = sA + np;
    float* sg = sBe + np;   /* 2 np gradients + loss */
    int8_t* post = (int8_t*)(sg + 2 * np + 1);
    const int tid = threadIdx.x;
    for (int k = tid; k < np; k += NT) {
        const float al = a.params[k], be = a.params[np + k];
        const float rs = rintf(al * 16.0f), rb = rintf(be * 16.0f);
        sS[k] = rs < 0.0f ? 0 : rs > 63.0f ? 63 : (int)rs;
        sB[k] = rb < -1024.0f ? -1024 : rb > 1023.0f ? 1023 : (int)rb;
        sA[k] = al;
        sBe[k] = be;
    }
    for (int k = tid; k < 2 * np + 1; k += NT)
        sg[k] = 0.0f;
    const size_t wg = blockIdx.x;
    int16_t* raw = a.raw + wg * (size_t)T * a.lts;
    int8_t* msg = a.msg + wg * (size_t)a.lts;
    int8_t* Ls = a.Ls + wg * (size_t)T * a.N;
    float* gL = a.gL + wg * (size_t)a.N;
    float* gm = a.gm + wg * (size_t)a.lts;
    const int N = a.N, K = a.K, q = a.q;
    float loss = 0.0f;
    __syncthreads();
    for (int f = blockIdx.x; f < a.nframes; f += gridDim.x) {
        const int8_t* fl = a.llr + (size_t)f * N;
        const uint8_t* bits = a.cw + (size_t)a.row[f] * N;
        for (int v = tid; v < N; v += NT)
            post[v] = fl[cw_pos(v, K, q)];
        __syncthreads();
        /* ---- forward ---- */
        for (int t = 0; t < T; t++) {
            int16_t* raw_t = raw + (size_t)t * a.lts;
            for (int i = 0; i < q; i++) {
                const uint4 ly = a.lay[i];
                for (int l = 0; l < (int)ly.z; l++) {
                    if (tid < 360 && a.level[i * 360 + tid] == l)
                        fwd_node<DM>(a, post, ly, i, tid, t == 0, raw_t, msg, sS + t * DEGS, sB + t * DEGS);
                    __syncthreads();
                }
            }
            int8_t* Lt = Ls + (size_t)t * N;
            const float wt = a.w[t];
            for (int v = tid; v < N; v += NT) {
                const int L = post[v];
                Lt[v] = (int8_t)L;
                if (wt != 0.0f) {
                    const float z = -a.temp * (float)L;
                    const float sp = fmaxf(z, 0.0f) + log1pf(expf(-fabsf(z)));
                    loss += wt * (sp - (float)bits[cw_pos(v, K, q)] * z);
                }
            }
            __syncthreads();
        }
        if (a.post)
            for (int v = tid; v < N; v += NT)
                a.post[(size_t)f * N + cw_pos(v, K, q)] = post[v];
        /* ---- backward ---- */
        for (int v = tid; v < N; v += NT)
            gL[v] = 0.0f;
        for (int s = tid; s < a.lts; s += NT)
            gm[s] = 0.0f;
        __syncthreads();
        for (int t = T - 1; t >= 0; t--) {
            const float wt = a.w[t];
            if (wt != 0.0f) {
                const int8_t* Lt = Ls + (size_t)t * N;
                for (int v = tid; v < N; v += NT) {
                    const float s = 1.0f / (1.0f + expf(a.temp * (float)Lt[v]));
                    gL[v] += wt * a.temp * ((float)bits[cw_pos(v, K, q)] - s) * a.inv;
                }
            }
            __syncthreads();
            const int16_t* raw_t = raw + (size_t)t * a.lts;
            for (int i = q - 1; i >= 0; i--) {
                const uint4 ly = a.lay[i];
                for (int l = (int)ly.z - 1; l >= 0; l--) {
                    if (tid < 360 && a.level[i * 360 + tid] == l)
                        bwd_node<DM>(a, gL, gm, raw_t, ly, i, tid, t, sS + t * DEGS, sB + t * DEGS, sA + t * DEGS,
                                     sBe + t * DEGS, sg);
                    __syncthreads();
                }
            }
        }
    }
    /* loss of this workgroup: wave sums, then one LDS atomic per wave */
    for (int off = 32; off > 0; off >>= 1)
        loss += __shfl_down(loss, off);
    if ((tid & 63) == 0)
        atomicAdd(&sg[2 * np], loss * a.inv);
    __syncthreads();
    float* part = a.part + wg * (size_t)(2 * np + 1);
    for (int k = tid; k < 2 * np + 1; k += NT)
        part[k] = sg[k];
}

/* fixed-order sum of the per-workgroup partials */
__global__ void nms_reduce_kernel(const float* part, int nwg, int n, float* grad, float* loss)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k > n)
        return;
    float s = 0.0f;
    for (int b = 0; b < nwg; b++)
        s += part[(size_t)b * (n + 1) + k];
    if (k < n)
        grad[k] = s;
    else
        *loss = s;
}

/* Adam, then the projection onto the parameter box (alpha: first half) */
__global__ void nms_adam_kernel(float* p, const float* g, float* m, float* v, int n, float lr, float b1, float b2,
                                float eps, float c1, float c2)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n)
        return;
    const float gk = g[k];
    const float mk = b1 * m[k] + (1.0f - b1) * gk;
    const float vk = b2 * v[k] + (1.0f - b2) * gk * gk;
    m[k] = mk;
    v[k] = vk;
    const float q = p[k] - lr * (mk / c1) / (sqrtf(vk / c2) + eps);
    p[k] = k < n / 2 ? fminf(fmaxf(q, 0.25f), 1.5f) : fminf(fmaxf(q, -8.0f), 16.0f);
}

/* ---- channel ------------------------------------------------------------ */

struct ChanArgs {
    int n, nframes, bits, np, nsym, unit;
    int order[8];
    float sn_scale; /* unused placeholder for alignment */
    float pts[512]; /* unit-energy points, (re, im) by label */
};

__device__ __forceinline__ uint64_t mix64(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}

/* two standard normals from (seed, frame, index), Box-Muller */
__device__ __forceinline__ float2 gauss2(uint64_t seed, uint64_t f, uint64_t k)
{
    const uint64_t h1 = mix64(seed ^ mix64(f * 0xD1B54A32D192ED03ull ^ k));
    const uint64_t h2 = mix64(h1 ^ 0xA0761D6478BD642Full);
    const float u1 = ((float)(h1 >> 40) + 0.5f) * (1.0f / 16777216.0f);
    const float u2 = ((float)(h2 >> 40) + 0.5f) * (1.0f / 16777216.0f);
    const float r = sqrtf(-2.0f * logf(u1));
    float s, c;
    sincosf(6.283185307179586f * u2, &s, &c);
    return make_float2(r * c, r * s);
}

__device__ __forceinline__ int8_t quant8(float v)
{
    v = fminf(fmaxf(v, -127.0f), 127.0f);
    return (int8_t)rintf(v);
}

/* the unit channel of tests/test_ldpc.c:64-73: +-1 per bit, AWGN of
 * variance 1 / Es/N0, llr = sat127(rint(4 y / sigma^2)) */
__global__ void channel_kernel(ChanArgs a, const uint8_t* cw, const int32_t* row, const float* esn0_db, uint64_t seed,
                               int8_t* llr)
{
    const size_t per = (size_t)a.n;
    const size_t total = per * a.nframes;
    for (size_t g = blockIdx.x * (size_t)blockDim.x + threadIdx.x; g < total; g += (size_t)gridDim.x * blockDim.x) {
        const int f = (int)(g / per), j = (int)(g % per);
        const uint8_t* bits = cw + (size_t)row[f] * a.n;
        const float s2 = 1.0f / exp10f(esn0_db[f] * 0.1f);
        const float2 z = gauss2(seed, (uint64_t)f, (uint64_t)j);
        const float y = (bits[j] ? -1.0f : 1.0f) + sqrtf(s2) * z.x;
        llr[(size_t)f * a.n + j] = quant8(4.0f * y / s2);
    }
}

/* the channel up to the receiver's symbols: frame f's bits mapped (bit
 * interleaver as the demapper undoes it) to unit-energy points plus
 * complex AWGN of variance N0 = 1 / Es/N0; n0[f] = N0 */
__global__ void symbol_kernel(ChanArgs a, const uint8_t* cw, const int32_t* row, const float* esn0_db, uint64_t seed,
                              int f0, float2* sym, float* n0)
{
    const size_t per = (size_t)a.nsym, total = per * a.nframes;
    for (size_t g = blockIdx.x * (size_t)blockDim.x + threadIdx.x; g < total; g += (size_t)gridDim.x * blockDim.x) {
        const int fl = (int)(g / per), j = (int)(g % per), f = f0 + fl;
        const uint8_t* bits = cw + (size_t)row[f] * a.n;
        const float n0f = 1.0f / exp10f(esn0_db[f] * 0.1f), sn = sqrtf(n0f * 0.5f);
        const float2 z = gauss2(seed, (uint64_t)f, (uint64_t)j);
        int label = 0;
        for (int b = 0; b < a.bits; b++) {
            const int pos = a.order[0] < 0 ? a.bits * j + b : a.order[b] * a.nsym + j;
            label = label << 1 | (pos < a.n ? bits[pos] : 0);
        }
        sym[g] = make_float2(a.pts[2 * label] + sn * z.x, a.pts[2 * label + 1] + sn * z.y);
        if (j == 0)
            n0[fl] = n0f;
    }
}

} // namespace

struct ldpc_nms_trainer {
    int device, code, T, max_frames, grid, N, K, q, lts, dm;
    size_t lds;
    hipStream_t stream;
    char err[256];
    uint4* lay;
    uint32_t* vidx;
    uint8_t* level;
    int16_t* raw;
    int8_t* msg;
    int8_t* Ls;
    float* gL;
    float* gm;
    float* part;
};

static int terr(ldpc_nms_trainer* t, int code, const char* fmt, ...)
{
    if (t) {
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(t->err, sizeof(t->err), fmt, ap);
        va_end(ap);
    }
    return code;
}

#define TCHK(t, call)                                                                          \
    do {                                                                                       \
        hipError_t e_ = (call);                                                                \
        if (e_ != hipSuccess)                                                                  \
            return terr((t), LDPC_GPU_EDEVICE, "%s: %s", #call, hipGetErrorString(e_));       \
    } while (0)

extern "C" {

void ldpc_nms_quantize(const float* alpha, const float* beta, int T, uint8_t* S, int16_t* B)
{
    for (int k = 0; k < T * DEGS; k++) {
        const float rs = rintf(alpha[k] * 16.0f), rb = rintf(beta[k] * 16.0f);
        S[k] = (uint8_t)(rs < 0.0f ? 0 : rs > 63.0f ? 63 : (int)rs);
        B[k] = (int16_t)(rb < -1024.0f ? -1024 : rb > 1023.0f ? 1023 : (int)rb);
    }
}

void ldpc_nms_trainer_destroy(ldpc_nms_trainer* t)
{
    if (!t)
        return;
    (void)hipSetDevice(t->device);
    if (t->stream)
        (void)hipStreamSynchronize(t->stream);
    (void)hipFree(t->lay);
    (void)hipFree(t->vidx);
    (void)hipFree(t->level);
    (void)hipFree(t->raw);
    (void)hipFree(t->msg);
    (void)hipFree(t->Ls);
    (void)hipFree(t->gL);
    (void)hipFree(t->gm);
    (void)hipFree(t->part);
    if (t->stream)
        (void)hipStreamDestroy(t->stream);
    free(t);
}

int ldpc_nms_trainer_create(int device, int code, int T, int max_frames, ldpc_nms_trainer** out)
{
    if (!out)
        return LDPC_GPU_EINVAL;
    *out = nullptr;
    const lg_code* c = lg_code_get(code);
    if (!c || T < 1 || T > LDPC_NMS_MAX_ITERS || max_frames < 1)
        return LDPC_GPU_EINVAL;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev)
        return LDPC_GPU_EDEVICE;
    ldpc_nms_trainer* t = (ldpc_nms_trainer*)calloc(1, sizeof(ldpc_nms_trainer));
    if (!t)
        return LDPC_GPU_ENOMEM;
    t->device = device;
    t->code = code;
    t->T = T;
    t->max_frames = max_frames;
    t->N = c->n;
    t->K = c->k;
    t->q = c->q;
    /* slots: per layer (info degree + 2) x 360, the node-(0,0) previous-
     * parity slot unused */
    std::vector<uint4> lay(c->q);
    int sb = 0, dm = 0;
    for (int i = 0; i < c->q; i++) {
        const int cnt = c->layer_start[i + 1] - c->layer_start[i];
        lay[i] = make_uint4((unsigned)sb, (unsigned)cnt, (unsigned)c->nlev[i], (unsigned)i);
        sb += (cnt + 2) * 360;
        dm = std::max(dm, cnt + 2);
    }
    t->lts = sb;
    t->dm = dm <= 16 ? 16 : 32;
    std::vector<uint32_t> vidx((size_t)t->lts, NO_VAR);
    std::vector<uint32_t> tmp;
    for (int i = 0; i < c->q; i++) {
        const int cnt = (int)lay[i].y, base = (int)lay[i].x;
        for (int j = 0; j < 360; j++) {
            tmp.clear();
            for (int e = 0; e < cnt; e++) {
                const uint32_t en = c->ent[c->layer_start[i] + e];
                const int d = j - (int)(en & 0xffff);
                tmp.push_back((en >> 16) + (uint32_t)(d < 0 ? d + 360 : d));
            }
            /* information edges in increasing bit order, as the reference
             * lists them (ldpc.c:238-250) */
            std::sort(tmp.begin(), tmp.end());
            for (int e = 0; e < cnt; e++)
                vidx[(size_t)base + (size_t)e * 360 + j] = tmp[e];
            vidx[(size_t)base + (size_t)cnt * 360 + j] = (uint32_t)(c->k + 360 * i + j);
            /* previous parity (ldpc.c:277-284) */
            const int pp = i ? 360 * (i - 1) + j : j ? j + (c->q - 1) * 360 - 1 : -1;
            if (pp >= 0)
                vidx[(size_t)base + (size_t)(cnt + 1) * 360 + j] = (uint32_t)(c->k + pp);
        }
    }
    int rc = LDPC_GPU_OK;
    auto fail = [&](int r) {
        ldpc_nms_trainer_destroy(t);
        return r;
    };
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&t->stream, hipStreamNonBlocking) != hipSuccess)
        return fail(LDPC_GPU_EDEVICE);
    const int np = T * DEGS;
    t->lds = (size_t)np * 16 + (size_t)(2 * np + 1) * 4 + (size_t)t->N;
    t->lds = (t->lds + 15) & ~(size_t)15;
    int per_cu = 0, ncu = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess)
        ncu = 256;
    hipError_t e = t->dm == 16 ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, nms_train_kernel<16>, NT, t->lds)
                               : hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, nms_train_kernel<32>, NT, t->lds);
    if (e != hipSuccess || per_cu < 1)
        return fail(LDPC_GPU_EDEVICE);
    t->grid = std::min(max_frames, per_cu * ncu);
    const size_t g = (size_t)t->grid;
    if (hipMalloc((void**)&t->lay, sizeof(uint4) * lay.size()) != hipSuccess ||
        hipMalloc((void**)&t->vidx, sizeof(uint32_t) * vidx.size()) != hipSuccess ||
        hipMalloc((void**)&t->level, (size_t)c->r) != hipSuccess ||
        hipMalloc((void**)&t->raw, g * T * t->lts * sizeof(int16_t)) != hipSuccess ||
        hipMalloc((void**)&t->msg, g * t->lts) != hipSuccess ||
        hipMalloc((void**)&t->Ls, g * T * (size_t)t->N) != hipSuccess ||
        hipMalloc((void**)&t->gL, g * t->N * sizeof(float)) != hipSuccess ||
        hipMalloc((void**)&t->gm, g * t->lts * sizeof(float)) != hipSuccess ||
        hipMalloc((void**)&t->part, g * (2 * np + 1) * sizeof(float)) != hipSuccess)
        return fail(LDPC_GPU_ENOMEM);
    if (hipMemcpy(t->lay, lay.data(), sizeof(uint4) * lay.size(), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(t->vidx, vidx.data(), sizeof(uint32_t) * vidx.size(), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(t->level, c->level, (size_t)c->r, hipMemcpyHostToDevice) != hipSuccess)
        return fail(LDPC_GPU_EDEVICE);
    (void)rc;
    *out = t;
    return LDPC_GPU_OK;
}

void* ldpc_nms_trainer_stream(ldpc_nms_trainer* t) { return t ? (void*)t->stream : nullptr; }

const char* ldpc_nms_trainer_error(ldpc_nms_trainer* t) { return t ? t->err : "no trainer"; }

int ldpc_nms_trainer_synchronize(ldpc_nms_trainer* t)
{
    if (!t)
        return LDPC_GPU_EINVAL;
    TCHK(t, hipSetDevice(t->device));
    TCHK(t, hipStreamSynchronize(t->stream));
    return LDPC_GPU_OK;
}

int ldpc_nms_train_step(ldpc_nms_trainer* t, const int8_t* d_llr, const uint8_t* d_cw, const int32_t* d_row,
                        int nframes, const float* d_params, const float* d_w, float temp, float* d_loss,
                        float* d_grad, int8_t* d_post)
{
    if (!t)
        return LDPC_GPU_EINVAL;
    if (nframes < 1 || nframes > t->max_frames || !d_llr || !d_cw || !d_row || !d_params || !d_w || !d_loss ||
        !d_grad)
        return terr(t, LDPC_GPU_EINVAL, "bad arguments (nframes %d, max %d)", nframes, t->max_frames);
    TCHK(t, hipSetDevice(t->device));
    NmsArgs a;
    a.lay = t->lay;
    a.vidx = t->vidx;
    a.level = t->level;
    a.N = t->N;
    a.K = t->K;
    a.q = t->q;
    a.T = t->T;
    a.lts = t->lts;
    a.llr = d_llr;
    a.cw = d_cw;
    a.row = d_row;
    a.nframes = nframes;
    a.params = d_params;
    a.w = d_w;
    a.temp = temp;
    a.inv = 1.0f / ((float)t->N * (float)nframes);
    a.raw = t->raw;
    a.msg = t->msg;
    a.Ls = t->Ls;
    a.gL = t->gL;
    a.gm = t->gm;
    a.part = t->part;
    a.post = d_post;
    const int grid = std::min(t->grid, nframes);
    if (t->dm == 16)
        hipLaunchKernelGGL(nms_train_kernel<16>, dim3(grid), dim3(NT), t->lds, t->stream, a);
    else
        hipLaunchKernelGGL(nms_train_kernel<32>, dim3(grid), dim3(NT), t->lds, t->stream, a);
    TCHK(t, hipGetLastError());
    const int n = 2 * t->T * DEGS;
    hipLaunchKernelGGL(nms_reduce_kernel, dim3((n + 1 + 255) / 256), dim3(256), 0, t->stream, t->part, grid, n,
                       d_grad, d_loss);
    TCHK(t, hipGetLastError());
    return LDPC_GPU_OK;
}

int ldpc_nms_adam(ldpc_nms_trainer* t, float* d_param, const float* d_grad, float* d_m, float* d_v, int n,
                  float lr, float beta1, float beta2, float eps, int step)
{
    if (!t)
        return LDPC_GPU_EINVAL;
    if (n < 1 || step < 1 || !d_param || !d_grad || !d_m || !d_v)
        return terr(t, LDPC_GPU_EINVAL, "bad arguments");
    TCHK(t, hipSetDevice(t->device));
    const float c1 = 1.0f - powf(beta1, (float)step), c2 = 1.0f - powf(beta2, (float)step);
    hipLaunchKernelGGL(nms_adam_kernel, dim3((n + 255) / 256), dim3(256), 0, t->stream, d_param, d_grad, d_m, d_v, n,
                       lr, beta1, beta2, eps, c1, c2);
    TCHK(t, hipGetLastError());
    return LDPC_GPU_OK;
}

/* channel arguments of MODCOD pls (pls < 0: the unit channel) */
static int chan_args(int pls, int n, int nframes, ChanArgs& a)
{
    memset(&a, 0, sizeof(a));
    a.n = n;
    a.nframes = nframes;
    if (pls < 0) {
        a.unit = 1;
        a.bits = 1;
        return LDPC_GPU_OK;
    }
    const lg_modcod_row* m = lg_modcod_get(pls);
    if (!m || m->bits < 1 || m->bits > 8)
        return LDPC_GPU_EINVAL;
    a.bits = m->bits;
    a.np = 1 << m->bits;
    a.nsym = (n + a.bits - 1) / a.bits;
    for (int b = 0; b < 8; b++)
        a.order[b] = m->order[b];
    const float* pts = lg_points() + 2 * m->point_off;
    double e = 0;
    for (int p = 0; p < a.np; p++)
        e += (double)pts[2 * p] * pts[2 * p] + (double)pts[2 * p + 1] * pts[2 * p + 1];
    const double gn = 1.0 / sqrt(e / a.np);
    for (int p = 0; p < 2 * a.np; p++)
        a.pts[p] = (float)(pts[p] * gn);
    return LDPC_GPU_OK;
}

int ldpc_gpu_channel_symbols(int pls, int n, const uint8_t* d_cw, const int32_t* d_row, int nframes,
                             const float* d_esn0_db, uint64_t seed, float* d_sym, float* d_n0, void* stream)
{
    if (pls < 0 || n < 1 || nframes < 1 || !d_cw || !d_row || !d_esn0_db || !d_sym || !d_n0)
        return LDPC_GPU_EINVAL;
    ChanArgs a;
    if (chan_args(pls, n, nframes, a))
        return LDPC_GPU_EINVAL;
    const size_t total = (size_t)a.nsym * nframes;
    const int grid = (int)std::min<size_t>((total + 255) / 256, 65536);
    hipLaunchKernelGGL(symbol_kernel, dim3(grid), dim3(256), 0, (hipStream_t)stream, a, d_cw, d_row, d_esn0_db, seed, 0,
                       (float2*)d_sym, d_n0);
    return hipGetLastError() == hipSuccess ? LDPC_GPU_OK : LDPC_GPU_EDEVICE;
}

int ldpc_gpu_channel(int pls, int n, const uint8_t* d_cw, const int32_t* d_row, int nframes, const float* d_esn0_db,
                     uint64_t seed, int8_t* d_llr, void* stream)
{
    if (n < 1 || nframes < 1 || !d_cw || !d_row || !d_esn0_db || !d_llr)
        return LDPC_GPU_EINVAL;
    ChanArgs a;
    if (chan_args(pls, n, nframes, a))
        return LDPC_GPU_EINVAL;
    const hipStream_t st = (hipStream_t)stream;
    if (a.unit) {
        const size_t total = (size_t)n * nframes;
        const int grid = (int)std::min<size_t>((total + 255) / 256, 65536);
        hipLaunchKernelGGL(channel_kernel, dim3(grid), dim3(256), 0, st, a, d_cw, d_row, d_esn0_db, seed, d_llr);
        return hipGetLastError() == hipSuccess ? LDPC_GPU_OK : LDPC_GPU_EDEVICE;
    }
    /* MODCOD frames: symbols, then the library's demapper (ldpc_gpu_demap's
     * kernel, the reference's LLRs for these symbols), in chunks */
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess)
        return LDPC_GPU_EDEVICE;
    const int chunk = std::min(nframes, 2048);
    float2* sym = nullptr;
    float* n0 = nullptr;
    if (hipMalloc((void**)&sym, sizeof(float2) * (size_t)a.nsym * chunk) != hipSuccess ||
        hipMalloc((void**)&n0, sizeof(float) * (size_t)chunk) != hipSuccess) {
        (void)hipFree(sym);
        return LDPC_GPU_ENOMEM;
    }
    int rc = LDPC_GPU_OK;
    for (int f0 = 0; f0 < nframes && !rc; f0 += chunk) {
        a.nframes = std::min(chunk, nframes - f0);
        const size_t total = (size_t)a.nsym * a.nframes;
        const int grid = (int)std::min<size_t>((total + 255) / 256, 65536);
        hipLaunchKernelGGL(symbol_kernel, dim3(grid), dim3(256), 0, st, a, d_cw, d_row, d_esn0_db, seed, f0, sym, n0);
        if (hipGetLastError() != hipSuccess)
            rc = LDPC_GPU_EDEVICE;
        else
            rc = lg_demap_launch(dev, pls, a.nframes, (const float*)sym, n0, d_llr + (size_t)f0 * n, st);
    }
    if (hipStreamSynchronize(st) != hipSuccess && !rc)
        rc = LDPC_GPU_EDEVICE;
    (void)hipFree(sym);
    (void)hipFree(n0);
    return rc;
}

} // extern "C"
