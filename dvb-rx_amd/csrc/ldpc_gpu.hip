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
#include "ldpc_kern.h"
#include "ldpc_nms.h"

using namespace lgk;

static const int CLASS_DI[NCLASS] = { 12, 28 }; /* largest layer degree of each class (one-frame kernel) */
static const int SK_INST[SK_NINST][2] = SK_INSTS; /* split kernel instances: slots per lane, wide */

/* split kernel instance of a code (lg_code.sk_ks, sk_wide), -1 if none */
static int sk_instance(const lg_code* cd)
{
    for (int k = 0; k < SK_NINST; k++)
        if (SK_INST[k][0] == cd->sk_ks && SK_INST[k][1] == cd->sk_wide)
            return k;
    return -1;
}

namespace {

/* ---- device-resident code tables (one copy per device) ------------------ */

struct DeviceTables {
    int ready = 0;
    DCode* codes = nullptr;
    uint32_t* ent = nullptr;
    uint2* lay = nullptr;
    uint4* pkl = nullptr;    /* per layer of every code: the split kernel's descriptor */
    uint32_t* adt = nullptr; /* its table rows (lg_code.sk_tab of every code) */
    uint64_t* meta = nullptr;
    std::vector<DCode> h_codes;
};

std::mutex g_tables_lock;
DeviceTables g_tables[64];

} // namespace

/* Shape of one persistent launch: workgroups per CU its kernel instance is
 * built for (1, or 2 where sk_has_occ2) and the share of the resident
 * workgroups it launches (0 < scale <= 1; 0 reads as 1). */
struct PkShape {
    int occ;
    float scale;
};

struct ldpc_gpu_ctx {
    int device;
    hipStream_t stream;
    hipStream_t aux[NAUX]; /* forked streams of concurrent launches */
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
    int cls_di[NLAUNCH];     /* its kernel instance's layer degree bound */
    int cls_occ[NLAUNCH];    /* its workgroups per CU and grid */
    int cls_grid[NLAUNCH];
    int timed;
    int force_scalar;        /* per-frame batches through the one-frame kernel */
    int lanes;               /* forked streams a decode's launches use; 0: one after the other on the context's stream */
    int group_fallbacks;     /* calls with groups that the two-pass path left to the exact one */
    PkShape shape[NLAUNCH];  /* launch shape per (instance, size); zero: one workgroup per CU, full grid */
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
    hipStream_t aux[NAUX];
    hipEvent_t fork, join[NLAUNCH], t0, t1;
    hipGraph_t graph;
    hipGraphExec_t exec;
    int variant;
    float variant_ms[8];
    int launched;
    PkShape shape[NLAUNCH]; /* the launch shapes the graph was built with */
    int lanes; /* forked streams of the launches, 0: one after the other */
    int nfr[NLAUNCH], grid[NLAUNCH], occ[NLAUNCH];
    float tuned_ms; /* ms per replay of the tuned shape, < 0 before ldpc_gpu_plan_tune */
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
    std::vector<uint4> pkl;
    std::vector<uint32_t> adt;
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
        d.pk_msg_words = cd->sk_msg_words;
        d.code_id = id;
        d.max_xfer = cd->max_xfer;
        d.sk_inst = sk_instance(cd);
        if (d.sk_inst < 0)
            return set_err(c, LDPC_GPU_EINVAL, "code %d: no split kernel for %d slots per lane", id, cd->sk_ks);
        /* the split kernel's descriptors hold byte offsets: of the layer's
         * messages in a message block, of its records in adt */
        const size_t base = adt.size();
        if (4 * (base + (size_t)cd->sk_words) > 0x7fffffffu)
            return set_err(c, LDPC_GPU_EINVAL, "split kernel tables exceed 2 GiB");
        adt.insert(adt.end(), cd->sk_tab, cd->sk_tab + cd->sk_words);
        ent.insert(ent.end(), cd->ent, cd->ent + cd->nent);
        d.pkl_off = (int)pkl.size();
        for (int i = 0; i < cd->q; i++) {
            const int st = cd->layer_start[i], deg = cd->layer_start[i + 1] - st;
            d.dmin = std::min(d.dmin, deg);
            lay.push_back(make_uint2((uint32_t)st | (uint32_t)deg << 16 | (uint32_t)cd->nlev[i] << 24,
                                     (uint32_t)cd->msg_off[i]));
            const uint32_t* sl = cd->sk_lay + 4 * i;
            /* bit 17: the workgroup waits after this layer (the next one
             * waits, bit 16 of its descriptor, or this is the last) */
            const uint32_t after = i + 1 == cd->q || ((cd->sk_lay[4 * (i + 1)] >> 16) & 1);
            pkl.push_back(make_uint4(sl[0] | after << 17, 4 * sl[1], (uint32_t)(4 * (base + sl[2])),
                                     cd->nlev[i] > 1 ? (uint32_t)(4 * (base + sl[3])) : 0u));
        }
        /* the kernel reads descriptors two layers ahead without a bound */
        pkl.push_back(pkl.back());
        pkl.push_back(pkl.back());
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
    HIPCHK(c, hipMalloc((void**)&t.pkl, sizeof(uint4) * pkl.size()));
    HIPCHK(c, hipMalloc((void**)&t.adt, sizeof(uint32_t) * adt.size()));
    HIPCHK(c, hipMemcpy(t.adt, adt.data(), sizeof(uint32_t) * adt.size(), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(t.pkl, pkl.data(), sizeof(uint4) * pkl.size(), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(t.codes, codes.data(), sizeof(DCode) * codes.size(), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(t.ent, ent.data(), sizeof(uint32_t) * ent.size(), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(t.lay, lay.data(), sizeof(uint2) * lay.size(), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(t.meta, meta.data(), sizeof(uint64_t) * meta.size(), hipMemcpyHostToDevice));
    t.h_codes = codes;
    t.ready = 1;
    return LDPC_GPU_OK;
}

static int static_init(ldpc_gpu_ctx* c);

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
            for (int k = 0; k < NAUX; k++)
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
        /* releases the streams and events made so far (the rest are NULL) */
        ldpc_gpu_destroy(c);
        return LDPC_GPU_EDEVICE;
    }
    if (hipDeviceGetAttribute(&c->ncu, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess)
        c->ncu = 256;
    c->lanes = NAUX_DEFAULT;
    /* short frames on the instances built for it: two workgroups per CU
     * (18-23 % less time per frame-iteration once a launch fills the CUs,
     * the same below that) */
    for (int k = 0; k < SK_NINST; k++)
        c->shape[2 * k + 1] = PkShape{ sk_has_occ2(k, false) ? 2 : 1, 1.0f };
    if ((rc = upload_tables(c)) != LDPC_GPU_OK || (rc = static_init(c)) != LDPC_GPU_OK ||
        hipMalloc((void**)&c->ticket, NCOUNTERS * sizeof(unsigned)) != hipSuccess) {
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
    for (int k = 0; k < NAUX; k++)
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

int ldpc_gpu_code_kernel_degree(int code)
{
    const lg_code* c = lg_code_get(code);
    const int k = c ? sk_instance(c) : -1;
    return k >= 0 ? 2 * SK_INST[k][0] - 2 + 100 * SK_INST[k][1] : LDPC_GPU_EINVAL;
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

void* ldpc_gpu_host_alloc(ldpc_gpu_ctx* c, size_t bytes)
{
    if (!c || !bytes || hipSetDevice(c->device) != hipSuccess)
        return nullptr;
    void* p = nullptr;
    if (hipHostMalloc(&p, bytes, hipHostMallocDefault) != hipSuccess) {
        set_err(c, LDPC_GPU_ENOMEM, "pinned host allocation of %zu bytes failed", bytes);
        return nullptr;
    }
    return p;
}

int ldpc_gpu_host_free(ldpc_gpu_ctx* c, void* h)
{
    if (!c)
        return LDPC_GPU_EINVAL;
    if (h)
        HIPCHK(c, hipHostFree(h));
    return LDPC_GPU_OK;
}

int ldpc_gpu_upload_async(ldpc_gpu_ctx* c, void* d_dst, const void* src, size_t bytes)
{
    if (!c)
        return LDPC_GPU_EINVAL;
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipMemcpyAsync(d_dst, src, bytes, hipMemcpyHostToDevice, c->stream));
    return LDPC_GPU_OK;
}

int ldpc_gpu_download_async(ldpc_gpu_ctx* c, void* dst, const void* d_src, size_t bytes)
{
    if (!c)
        return LDPC_GPU_EINVAL;
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipMemcpyAsync(dst, d_src, bytes, hipMemcpyDeviceToHost, c->stream));
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
    if (concurrent < 0 || concurrent > NAUX)
        return set_err(c, LDPC_GPU_EINVAL, "set_concurrent: %d (0 .. %d)", concurrent, NAUX);
    c->lanes = concurrent == 1 ? NAUX_DEFAULT : concurrent;
    return LDPC_GPU_OK;
}

int ldpc_gpu_group_fallbacks(ldpc_gpu_ctx* c) { return c ? c->group_fallbacks : -1; }

int ldpc_gpu_set_kernel(ldpc_gpu_ctx* c, int kernel)
{
    if (!c || (kernel != LDPC_GPU_KERNEL_AUTO && kernel != LDPC_GPU_KERNEL_SINGLE))
        return LDPC_GPU_EINVAL;
    c->force_scalar = kernel == LDPC_GPU_KERNEL_SINGLE;
    return LDPC_GPU_OK;
}

/* the layer degree bound that names kernel instance k (as
 * ldpc_gpu_last_launches reports it) */
static int sk_di(int k) { return 2 * SK_INST[k][0] - 2 + 100 * SK_INST[k][1]; }

int ldpc_gpu_set_launch_shape(ldpc_gpu_ctx* c, int di, int short_frames, int occ, float scale)
{
    if (!c)
        return LDPC_GPU_EINVAL;
    if ((short_frames != 0 && short_frames != 1) || (occ != 1 && occ != 2) || !(scale > 0.0f && scale <= 1.0f))
        return set_err(c, LDPC_GPU_EINVAL, "launch shape: bad arguments");
    for (int k = 0; k < SK_NINST; k++) {
        if (sk_di(k) != di)
            continue;
        if (occ == 2 && !sk_has_occ2(k, false))
            return set_err(c, LDPC_GPU_EINVAL, "launch shape: kernel instance %d has no two-workgroup build", di);
        c->shape[2 * k + short_frames] = PkShape{ occ, scale };
        return LDPC_GPU_OK;
    }
    return set_err(c, LDPC_GPU_EINVAL, "launch shape: no kernel instance %d", di);
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

int ldpc_gpu_last_launch_shapes(ldpc_gpu_ctx* c, int* occ, int* grid, int max)
{
    if (!c || !c->timed)
        return 0;
    int n = 0;
    for (int k = 0; k < NLAUNCH && n < max; k++) {
        if (!c->cls_frames[k])
            continue;
        if (occ)
            occ[n] = c->cls_occ[k] ? c->cls_occ[k] : 1;
        if (grid)
            grid[n] = c->cls_grid[k];
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

namespace lgk {
#define SK_CASE_L(KS, W, N)                                                                       \
    if (SK_INST[k][0] == KS && SK_INST[k][1] == W && nms == (N != 0))                                 \
        return sk_launch_##KS##_##W##_##N(grid, lds, st, a);
#define SK_CASE_O(KS, W, N)                                                                       \
    if (SK_INST[k][0] == KS && SK_INST[k][1] == W && nms == (N != 0))                                 \
        return sk_occupancy_##KS##_##W##_##N(lds, per_cu);
#define SK_CASE_L2(KS, W, N)                                                                      \
    if (SK_INST[k][0] == KS && SK_INST[k][1] == W && nms == (N != 0))                                 \
        return sk_launch_##KS##_##W##_##N##_o2(grid, lds, st, a);
#define SK_CASE_O2(KS, W, N)                                                                      \
    if (SK_INST[k][0] == KS && SK_INST[k][1] == W && nms == (N != 0))                                 \
        return sk_occupancy_##KS##_##W##_##N##_o2(lds, per_cu);
#define SK_CASE_H2(KS, W, N)                                                                      \
    if (SK_INST[k][0] == KS && SK_INST[k][1] == W && nms == (N != 0))                                 \
        return true;

bool sk_has_occ2(int k, bool nms)
{
    if (k >= 0 && k < SK_NINST) {
        SK_ALL2(SK_CASE_H2)
    }
    return false;
}

hipError_t pk_launch(int k, bool nms, int occ, int grid, size_t lds, hipStream_t st, const PkArgs& a)
{
    if (k >= 0 && k < SK_NINST) {
        if (occ == 2) {
            SK_ALL2(SK_CASE_L2)
        } else {
            SK_ALL(SK_CASE_L)
        }
    }
    return hipErrorInvalidValue;
}

hipError_t pk_occupancy(int k, bool nms, int occ, size_t lds, int* per_cu)
{
    if (k >= 0 && k < SK_NINST) {
        if (occ == 2) {
            SK_ALL2(SK_CASE_O2)
        } else {
            SK_ALL(SK_CASE_O)
        }
    }
    return hipErrorInvalidValue;
}
} // namespace lgk

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
    HIPCHK(c, layered_occupancy(DI, lds, &per_cu));
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
    HIPCHK(c, layered_launch(DI, grid, lds, c->stream, a));
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
    HIPCHK(c, reset_counters(c->stream, c->ticket));
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
        c->cls_occ[k] = 1;
        c->cls_grid[k] = grid[k];
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

static int pk_grid(ldpc_gpu_ctx* c, int inst, bool nms, int occ, size_t lds, int nframes, int* grid)
{
    int per_cu = 0;
    HIPCHK(c, pk_occupancy(inst, nms, occ, lds, &per_cu));
    if (per_cu < 1)
        return set_err(c, LDPC_GPU_EDEVICE, "two-slot kernel does not fit (LDS %zu)", lds);
    *grid = std::min((nframes + 1) / 2, per_cu * c->ncu);
    return LDPC_GPU_OK;
}

/* Launch layout of frames with their own early stop (two per workgroup,
 * ldpc_sk_kernel): one launch per (kernel instance, frame size), frames
 * grouped by code into segments of one list, persistent grid per launch
 * (occupancy x CUs, scaled by the launch's shape). */
struct PkLayout {
    std::vector<int> list;
    std::vector<int2> segs;
    int seg0[NLAUNCH], nseg[NLAUNCH], nfr[NLAUNCH], grid[NLAUNCH], maxn[NLAUNCH], maxx[NLAUNCH], occ[NLAUNCH];
    double cost[NLAUNCH]; /* expected work: layers x slots per lane, summed over the frames */
    int maxw;
    size_t blocks; /* message blocks: sum of the grids */
};

/* LDS of a launch: the kernel's control words, both slots' posteriors (2
 * bytes per bit of its longest code, and a spare cell) from SK_POST_OFF,
 * then two hand-over buffers of maxx * 360 cells */
static size_t pk_hoff(const PkLayout& L, int k) { return (SK_POST_OFF + 2 * ((size_t)L.maxn[k] + 1) + 15) & ~(size_t)15; }
static size_t pk_lds(const PkLayout& L, int k) { return pk_hoff(L, k) + 2 * sizeof(uint16_t) * 360 * (size_t)L.maxx[k]; }

/* the workgroups per CU launch k runs with under shape sh: 2 only where
 * that instance is built for it */
static int pk_occ_of(int k, bool nms, const PkShape& sh) { return sh.occ == 2 && sk_has_occ2(k / 2, nms) ? 2 : 1; }
static float pk_scale_of(const PkShape& sh) { return sh.scale > 0.0f ? std::min(sh.scale, 1.0f) : 1.0f; }

static int pk_layout(ldpc_gpu_ctx* c, const std::vector<DFrame>& hf, bool nms, const PkShape* shape, PkLayout& L)
{
    const std::vector<DCode>& hc = c->tab->h_codes;
    std::vector<int> gcodes[NLAUNCH];
    std::vector<std::vector<int>> gseg[NLAUNCH];
    L.maxw = 0;
    for (int k = 0; k < NLAUNCH; k++) {
        L.maxn[k] = L.maxx[k] = 0;
        L.cost[k] = 0.0;
    }
    for (int f = 0; f < (int)hf.size(); f++) {
        if (hf[f].max_iter < 0)
            continue; /* left out of this decode: its outputs stay as they are */
        const DCode& cd = hc[hf[f].code];
        const int k = cd.sk_inst * 2 + (cd.n < 64800);
        L.cost[k] += (double)cd.q * SK_INST[cd.sk_inst][0];
        auto it = std::find(gcodes[k].begin(), gcodes[k].end(), hf[f].code);
        size_t g = it - gcodes[k].begin();
        if (it == gcodes[k].end()) {
            gcodes[k].push_back(hf[f].code);
            gseg[k].emplace_back();
        }
        gseg[k][g].push_back(f);
        L.maxn[k] = std::max(L.maxn[k], cd.n);
        L.maxx[k] = std::max(L.maxx[k], cd.max_xfer);
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
        L.occ[k] = pk_occ_of(k, nms, shape[k]);
        for (auto& sg : gseg[k]) {
            L.segs.push_back(make_int2((int)L.list.size(), (int)sg.size()));
            L.list.insert(L.list.end(), sg.begin(), sg.end());
            L.nfr[k] += (int)sg.size();
        }
        if (!L.nfr[k])
            continue;
        const size_t lds = pk_lds(L, k);
        const int rc = pk_grid(c, k / 2, nms, L.occ[k], lds, L.nfr[k], &L.grid[k]);
        if (rc)
            return rc;
        L.grid[k] = std::max(1, (int)((float)L.grid[k] * pk_scale_of(shape[k])));
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
                      const hipEvent_t* join, const hipEvent_t* tev, int lanes, bool nms)
{
    /* the fork event also starts the decode's timing (launch_pk), so it is
     * recorded in both modes */
    HIPCHK(c, hipEventRecord(fork, origin));
    HIPCHK(c, reset_counters(origin, d_cnt));
    a.msg_stride = (size_t)L.maxw;
    /* message blocks and enqueue order by launch index.  lanes > 0: each
     * launch onto one of the first `lanes` forked streams, whose launches
     * run one after another; taken in order of expected work, the most
     * first, each goes to the stream with the least so far */
    const bool concurrent = lanes > 0;
    size_t moff[NLAUNCH];
    int order[NLAUNCH], sidx[NLAUNCH], nl = 0;
    double load[NAUX] = { 0.0 };
    for (int k = 0, m = 0; k < NLAUNCH; k++) {
        moff[k] = (size_t)m * L.maxw;
        m += L.grid[k];
        if (L.nfr[k])
            order[nl++] = k;
    }
    std::stable_sort(order, order + nl, [&](int x, int y) { return L.cost[x] > L.cost[y]; });
    for (int i = 0; i < nl; i++) {
        int b = 0;
        for (int s2 = 1; s2 < lanes; s2++)
            if (load[s2] < load[b])
                b = s2;
        sidx[order[i]] = b;
        load[b] += L.cost[order[i]];
    }
    for (int k = 0; k < NLAUNCH; k++) {
        if (!L.nfr[k])
            continue;
        a.seg = d_seg + L.seg0[k];
        a.nseg = L.nseg[k];
        a.cnt = d_cnt + L.seg0[k];
        a.msg = d_msg + moff[k];
        a.hoff = pk_hoff(L, k);
        a.hstride = 360 * L.maxx[k];
        const size_t lds = pk_lds(L, k);
        hipStream_t st = concurrent ? aux[sidx[k]] : origin;
        if (concurrent)
            HIPCHK(c, hipStreamWaitEvent(st, fork, 0));
        if (tev)
            HIPCHK(c, hipEventRecord(tev[2 * k], st));
        HIPCHK(c, pk_launch(k / 2, nms, L.occ[k], L.grid[k], lds, st, a));
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
    a.pkl = c->tab->pkl;
    a.adt = c->tab->adt;
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
    int rc = pk_layout(c, hf, nms, c->shape, L);
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
        c->cls_di[k] = sk_di(k / 2);
        c->cls_occ[k] = L.occ[k];
        c->cls_grid[k] = L.grid[k];
    }
    if ((rc = pk_enqueue(c, L, a, c->d_seg, c->ticket, c->msg, c->stream, c->aux, c->ev0, join, c->ev_cls, c->lanes,
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
        /* a dword of padding each: the kernel reads entries as aligned
         * dwords (nms_u8, nms_i16) */
        HIPCHK(c, hipMalloc((void**)&c->d_nS, NMS_ROWS * LDPC_NMS_DEGS + 4));
        HIPCHK(c, hipMalloc((void**)&c->d_nB, NMS_ROWS * LDPC_NMS_DEGS * sizeof(int16_t) + 4));
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
        df[f].min_iter = 0;
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

/* a decode's iteration counts and syndrome flags, to the host */
static int group_fetch(ldpc_gpu_ctx* c, int nframes, const int32_t* d_iters, const uint8_t* d_synd,
                       std::vector<int32_t>& it, std::vector<uint8_t>& ok)
{
    HIPCHK(c, hipMemcpyAsync(it.data(), d_iters, sizeof(int32_t) * nframes, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(ok.data(), d_synd, nframes, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return LDPC_GPU_OK;
}

/* the groups of a call from its frames' own counts: gid[f] (-1: no group),
 * per group its count T (the largest of its frames' counts; max_iter where
 * a frame never converged) and max_iter (one per group) */
static int group_targets(ldpc_gpu_ctx* c, const ldpc_gpu_frame* frames, int nframes, const std::vector<int32_t>& it,
                         const std::vector<uint8_t>& ok, std::vector<int>& gid, std::vector<int>& gT,
                         std::vector<int>& gmax)
{
    std::vector<int> keys, gall;
    gid.assign(nframes, -1);
    gT.clear();
    gmax.clear();
    for (int f = 0; f < nframes; f++) {
        if (frames[f].group < 0)
            continue;
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
    for (size_t g = 0; g < gT.size(); g++)
        if (!gall[g])
            gT[g] = gmax[g];
    return LDPC_GPU_OK;
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

    /* Groups (the reference's batch semantics, ldpc.c:614-638): a group
     * iterates in lock step and stops at the first iteration where all its
     * frames hold their checks, or at max_iter.  No group can stop before
     * the largest of its frames' own first such iterations, so every frame
     * is first decoded with its own early stop; a group's count T is the
     * largest of its frames' counts (max_iter if one never converged). */
    std::vector<int32_t> it(nframes), tgt(nframes);
    std::vector<uint8_t> ok(nframes);
    std::vector<int> gid, gT, gmax;
    if (!c->force_scalar) {
        /* Two-frame kernel: then every frame again from its LLRs for
         * exactly its target (the syndrome looked at only there): a group's
         * frames T, a frame of no group its own count.  Where each group's
         * frames all hold their checks at T, or T is max_iter, that is the
         * lock-step result; otherwise (a frame that had converged lost a
         * check by T, rare) the call is decoded again by the exact path. */
        if ((rc = launch_pk(c, c->d_frames, df, d_llr, d_out, d_iters, d_synd)) ||
            (rc = group_fetch(c, nframes, d_iters, d_synd, it, ok)) ||
            (rc = group_targets(c, frames, nframes, it, ok, gid, gT, gmax)))
            return rc;
        /* a frame whose own count is already its target ran exactly that
         * many iterations in the first pass: same state, same outputs, so
         * it sits the second pass out (max_iter -1), as does every frame
         * of no group */
        std::vector<DFrame> df2 = df;
        int again = 0;
        for (int f = 0; f < nframes; f++) {
            const int T = gid[f] >= 0 ? gT[gid[f]] : it[f];
            df2[f].max_iter = df2[f].min_iter = T == it[f] ? -1 : T;
            again += T != it[f];
        }
        if (again) {
            HIPCHK(c, hipMemcpyAsync(c->d_frames, df2.data(), sizeof(DFrame) * nframes, hipMemcpyHostToDevice,
                                     c->stream));
            if ((rc = launch_pk(c, c->d_frames, df2, d_llr, d_out, d_iters, d_synd)) ||
                (rc = group_fetch(c, nframes, d_iters, d_synd, it, ok)))
                return rc;
        }
        bool exact = true;
        for (int f = 0; f < nframes; f++)
            if (gid[f] >= 0 && gT[gid[f]] < gmax[gid[f]] && !ok[f])
                exact = false;
        c->group_fallbacks += !exact;
        if (exact)
            return LDPC_GPU_OK;
        HIPCHK(c, hipMemcpyAsync(c->d_frames, df.data(), sizeof(DFrame) * nframes, hipMemcpyHostToDevice,
                                 c->stream));
    }
    /* One-frame kernel, exact in every case: the first pass keeps each
     * frame's state; frames below their group's T continue to T; while some
     * frame fails at T < max_iter, the whole group goes one more iteration.
     * Frames are deterministic, so this equals iterating the group in lock
     * step. */
    if ((rc = grow(c, &c->state, &c->state_bytes, (size_t)nframes * 64800)) ||
        (rc = grow(c, &c->it_state, &c->it_state_bytes, sizeof(int32_t) * nframes)) ||
        (rc = grow(c, &c->target, &c->target_bytes, sizeof(int32_t) * nframes)))
        return rc;
    rc = launch(c, c->d_frames, df, every, maxw, d_llr, d_out, d_iters, d_synd, 1, nframes, 0, 1, 1);
    if (rc || (rc = group_fetch(c, nframes, d_iters, d_synd, it, ok)) ||
        (rc = group_targets(c, frames, nframes, it, ok, gid, gT, gmax)))
        return rc;
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

/* ---- decoding without allocation (ldpc_gpu_decode_batch_static) --------
 *
 * All device scratch of the call (frame table, frame list, segments, claim
 * counters, message blocks) lives in `scratch`, a fixed array of this
 * library on every device it runs on, and the host layout in fixed host
 * arrays; the call makes no allocation.  The array's device address is
 * taken once per device when a context is created (static_init) and used
 * for every launch.  Calls on one device share the array (and the host
 * staging, under g_static_lock): they are serialised on the host, and the
 * launches of one call run one after the other on the context's stream. */
#define STATIC_SCRATCH_BYTES ((size_t)96 << 20)
#define STATIC_MAX_FRAMES 65536

static __device__ __align__(256) uint8_t scratch[STATIC_SCRATCH_BYTES];

static std::mutex g_static_lock;
static uint8_t* g_static_ptr[64];          /* scratch's address on each device */
static DFrame g_static_frames[STATIC_MAX_FRAMES];
static int g_static_list[STATIC_MAX_FRAMES];
static int2 g_static_segs[NCOUNTERS];

static int static_init(ldpc_gpu_ctx* c)
{
    std::lock_guard<std::mutex> lk(g_static_lock);
    if (g_static_ptr[c->device])
        return LDPC_GPU_OK;
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p)
        return set_err(c, LDPC_GPU_EDEVICE, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
    g_static_ptr[c->device] = (uint8_t*)p;
    return LDPC_GPU_OK;
}

extern "C" int ldpc_gpu_decode_batch_static(ldpc_gpu_ctx* c, const ldpc_gpu_frame* frames, int nframes,
                                            const int8_t* d_llr, uint8_t* d_out, int32_t* d_iters, uint8_t* d_synd)
{
    if (!c)
        return LDPC_GPU_EINVAL;
    if (nframes == 0)
        return LDPC_GPU_OK;
    if (nframes < 0 || !frames || !d_llr || !d_out || !d_iters || !d_synd)
        return set_err(c, LDPC_GPU_EINVAL, "bad arguments");
    if (nframes > STATIC_MAX_FRAMES)
        return set_err(c, LDPC_GPU_EINVAL, "static decode: %d frames, at most %d", nframes, STATIC_MAX_FRAMES);
    HIPCHK(c, hipSetDevice(c->device));
    std::lock_guard<std::mutex> lk(g_static_lock);
    uint8_t* base = g_static_ptr[c->device];
    if (!base)
        return set_err(c, LDPC_GPU_EDEVICE, "static scratch not set up on device %d", c->device);
    const DCode* hc = c->tab->h_codes.data();
    const int ncode = (int)c->tab->h_codes.size();
    /* frame table (as check_frames) and the launch / code of each frame */
    int maxw = 0, maxn[NLAUNCH] = { 0 }, maxx[NLAUNCH] = { 0 }, nfr[NLAUNCH] = { 0 };
    for (int f = 0; f < nframes; f++) {
        const ldpc_gpu_frame& s = frames[f];
        if (s.code < 0 || s.code >= ncode)
            return set_err(c, LDPC_GPU_EINVAL, "frame %d: unknown code %d", f, s.code);
        const DCode& cd = hc[s.code];
        if (s.max_iter < 0 || s.nbits < 0 || s.nbits > cd.k || (s.llr_off & 7))
            return set_err(c, LDPC_GPU_EINVAL, "frame %d: bad max_iter/nbits/llr_off", f);
        if (s.group >= 0)
            return set_err(c, LDPC_GPU_EINVAL, "static decode takes frames with their own early stop (group -1)");
        DFrame& d = g_static_frames[f];
        d.code = s.code;
        d.max_iter = s.max_iter;
        d.nbits = s.nbits ? s.nbits : cd.k;
        d.min_iter = 0;
        d.llr_off = s.llr_off;
        d.out_off = s.out_off;
        const int k = cd.sk_inst * 2 + (cd.n < 64800);
        nfr[k]++;
        maxn[k] = std::max(maxn[k], cd.n);
        maxx[k] = std::max(maxx[k], cd.max_xfer);
        maxw = std::max(maxw, cd.pk_msg_words);
    }
    /* frame list grouped by launch, then by code (one segment per code) */
    int nseg = 0, seg0[NLAUNCH], nsegk[NLAUNCH], pos = 0;
    for (int k = 0; k < NLAUNCH; k++) {
        seg0[k] = nseg;
        nsegk[k] = 0;
        if (!nfr[k])
            continue;
        for (int code = 0; code < ncode; code++) {
            if (hc[code].sk_inst * 2 + (hc[code].n < 64800) != k)
                continue;
            const int start = pos;
            for (int f = 0; f < nframes; f++)
                if (g_static_frames[f].code == code)
                    g_static_list[pos++] = f;
            if (pos == start)
                continue;
            if (nseg >= NCOUNTERS)
                return set_err(c, LDPC_GPU_EINVAL, "too many codes in one batch");
            g_static_segs[nseg++] = make_int2(start, pos - start);
            nsegk[k]++;
        }
    }
    /* scratch: frames | list | segments | counters | message blocks */
    auto up = [](size_t v) { return (v + 255) & ~(size_t)255; };
    const size_t o_list = up(sizeof(DFrame) * (size_t)nframes);
    const size_t o_seg = o_list + up(sizeof(int) * (size_t)nframes);
    const size_t o_cnt = o_seg + up(sizeof(int2) * NCOUNTERS);
    const size_t o_msg = o_cnt + up(sizeof(unsigned) * NCOUNTERS);
    const size_t msg_blocks = (STATIC_SCRATCH_BYTES - o_msg) / (sizeof(uint32_t) * (size_t)maxw);
    if (msg_blocks < 1)
        return set_err(c, LDPC_GPU_EINVAL, "static decode: scratch of %zu bytes too small", STATIC_SCRATCH_BYTES);
    DFrame* d_frames = (DFrame*)base;
    int* d_list = (int*)(base + o_list);
    int2* d_seg = (int2*)(base + o_seg);
    unsigned* d_cnt = (unsigned*)(base + o_cnt);
    uint32_t* d_msg = (uint32_t*)(base + o_msg);
    /* the host staging is reused by the next call: the copies complete
     * before this call returns */
    HIPCHK(c, hipMemcpyAsync(d_frames, g_static_frames, sizeof(DFrame) * (size_t)nframes, hipMemcpyHostToDevice,
                             c->stream));
    HIPCHK(c, hipMemcpyAsync(d_list, g_static_list, sizeof(int) * (size_t)nframes, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(d_seg, g_static_segs, sizeof(int2) * (size_t)nseg, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, reset_counters(c->stream, d_cnt));
    PkArgs a = pk_args(c, d_frames, d_list, d_llr, d_out, d_iters, d_synd, false, 0, nullptr);
    a.msg = d_msg;
    a.msg_stride = (size_t)maxw;
    c->timed = 0;
    for (int k = 0; k < NLAUNCH; k++) {
        if (!nfr[k])
            continue;
        PkLayout L;
        L.maxn[k] = maxn[k];
        L.maxx[k] = maxx[k];
        const size_t lds = pk_lds(L, k);
        int grid = 0;
        const int occ = pk_occ_of(k, false, c->shape[k]);
        int rc = pk_grid(c, k / 2, false, occ, lds, nfr[k], &grid);
        if (rc)
            return rc;
        grid = std::max(1, (int)((float)grid * pk_scale_of(c->shape[k])));
        grid = (int)std::min<size_t>((size_t)grid, msg_blocks);
        a.seg = d_seg + seg0[k];
        a.nseg = nsegk[k];
        a.cnt = d_cnt + seg0[k];
        a.hoff = pk_hoff(L, k);
        a.hstride = 360 * maxx[k];
        HIPCHK(c, pk_launch(k / 2, false, occ, grid, lds, c->stream, a));
    }
    return LDPC_GPU_OK;
}

/* ---- prepared batches: graph capture and the launch-shape autotuner ---- */

/* launch shapes the autotuner starts from, each for all launches of the
 * plan: grid scale (persistent workgroups per launch relative to occupancy
 * x CUs, i.e. how many frames run at once) and whether the (instance,
 * size) launches run concurrently; the context's workgroups per CU */
static const struct {
    float scale;
    int lanes; /* forked streams; -1: the context's count (default 3), 0: one after the other */
} PLAN_VARIANTS[] = { { 1.0f, -1 }, { 1.0f, 0 }, { 0.75f, -1 }, { 0.5f, -1 }, { 0.5f, 0 }, { 1.0f, 2 }, { 1.0f, 4 } };
#define PLAN_NVARIANTS ((int)(sizeof(PLAN_VARIANTS) / sizeof(PLAN_VARIANTS[0])))

static int plan_lanes(const ldpc_gpu_plan* p, int v)
{
    return PLAN_VARIANTS[v].lanes >= 0 ? PLAN_VARIANTS[v].lanes : (p->c->lanes ? p->c->lanes : NAUX_DEFAULT);
}

static void plan_variant_shape(ldpc_gpu_plan* p, int variant)
{
    for (int k = 0; k < NLAUNCH; k++) {
        p->shape[k].occ = p->c->shape[k].occ;
        p->shape[k].scale = PLAN_VARIANTS[variant].scale;
    }
    p->lanes = plan_lanes(p, variant);
}

/* (re)build the plan's graph with its current shape (p->shape,
 * p->lanes); variant: the start variant it came from, or -1 once the
 * per-launch search changed it */
static int plan_build(ldpc_gpu_plan* p, int variant)
{
    ldpc_gpu_ctx* c = p->c;
    PkLayout L;
    int rc = pk_layout(c, p->hf, false, p->shape, L);
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
                    p->lanes, false);
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
    for (int k = 0; k < NLAUNCH; k++) {
        p->nfr[k] = L.nfr[k];
        p->grid[k] = L.grid[k];
        p->occ[k] = L.occ[k];
    }
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
    for (int k = 0; k < NAUX; k++)
        if (p->aux[k])
            (void)hipStreamDestroy(p->aux[k]);
    for (int k = 0; k < NLAUNCH; k++)
        if (p->join[k])
            (void)hipEventDestroy(p->join[k]);
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
    p->tuned_ms = -1.0f;
    for (int v = 0; v < 8; v++)
        p->variant_ms[v] = -1.0f;
    /* scratch sized for the largest shape (full grid, two workgroups per
     * CU where built) */
    PkShape big[NLAUNCH];
    for (int k = 0; k < NLAUNCH; k++)
        big[k] = PkShape{ 2, 1.0f };
    PkLayout L;
    if ((rc = pk_layout(c, df, false, big, L))) {
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
    for (int k = 0; k < NAUX; k++)
        if (hipStreamCreateWithFlags(&p->aux[k], hipStreamNonBlocking) != hipSuccess)
            return fail(LDPC_GPU_EDEVICE);
    for (int k = 0; k < NLAUNCH; k++)
        if (hipEventCreateWithFlags(&p->join[k], hipEventDisableTiming) != hipSuccess)
            return fail(LDPC_GPU_EDEVICE);
    if (hipMemcpy(p->d_frames, df.data(), sizeof(DFrame) * nframes, hipMemcpyHostToDevice) != hipSuccess)
        return fail(LDPC_GPU_EDEVICE);
    plan_variant_shape(p, 0);
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

/* mean ms of reps replays of the plan's graph (after one untimed) */
static int plan_time(ldpc_gpu_plan* p, int reps, float* ms)
{
    ldpc_gpu_ctx* c = p->c;
    int rc;
    if ((rc = ldpc_gpu_plan_launch(p)))
        return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    float tot = 0.0f;
    for (int r = 0; r < reps; r++) {
        if ((rc = ldpc_gpu_plan_launch(p)))
            return rc;
        const float t = ldpc_gpu_plan_last_ms(p);
        if (t < 0.0f)
            return set_err(c, LDPC_GPU_EDEVICE, "plan timing failed");
        tot += t;
    }
    *ms = tot / (float)reps;
    return LDPC_GPU_OK;
}

extern "C" int ldpc_gpu_plan_tune(ldpc_gpu_plan* p, int reps)
{
    if (!p || reps < 1)
        return LDPC_GPU_EINVAL;
    ldpc_gpu_ctx* c = p->c;
    HIPCHK(c, hipSetDevice(c->device));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    int best = 0, rc;
    /* the shapes common to all launches */
    for (int v = 0; v < PLAN_NVARIANTS; v++) {
        plan_variant_shape(p, v);
        if ((rc = plan_build(p, v)) || (rc = plan_time(p, reps, &p->variant_ms[v])))
            return rc;
        if (p->variant_ms[v] < p->variant_ms[best])
            best = v;
    }
    plan_variant_shape(p, best);
    if ((rc = plan_build(p, best)))
        return rc;
    /* then one pass of coordinate descent over the launches, longest
     * first: each launch's workgroups per CU and grid scale with the
     * others held, kept where the whole plan gets faster by more than
     * 0.5 % (timing noise) */
    float cur = p->variant_ms[best];
    int order[NLAUNCH], nl = 0;
    for (int k = 0; k < NLAUNCH; k++)
        if (p->nfr[k])
            order[nl++] = k;
    /* bits decoded by the launch: a normal frame is four short ones (odd
     * launches hold short frames) */
    auto work = [&](int k) { return (size_t)p->nfr[k] * (k & 1 ? 1 : 4); };
    std::stable_sort(order, order + nl, [&](int x, int y) { return work(x) > work(y); });
    static const float SCALES[] = { 1.0f, 0.75f, 0.5f };
    for (int i = 0; i < nl; i++) {
        const int k = order[i];
        const PkShape keep = p->shape[k];
        PkShape pick = keep;
        /* two workgroups per CU only where they fit: short-frame launches
         * (odd k) of the instances built for it; a normal frame pair
         * holds most of a CU's LDS */
        const int max_occ = (k & 1) && sk_has_occ2(k / 2, false) ? 2 : 1;
        for (int occ = 1; occ <= max_occ; occ++)
            for (float sc : SCALES) {
                const PkShape t = PkShape{ occ, sc };
                if (pk_occ_of(k, false, t) == pk_occ_of(k, false, keep) && pk_scale_of(t) == pk_scale_of(keep))
                    continue;
                p->shape[k] = t;
                float ms = 0.0f;
                if ((rc = plan_build(p, -1)) || (rc = plan_time(p, reps, &ms)))
                    return rc;
                if (ms < cur * 0.995f) {
                    cur = ms;
                    pick = t;
                }
            }
        p->shape[k] = pick;
    }
    p->tuned_ms = cur;
    const bool changed = [&] {
        for (int k = 0; k < NLAUNCH; k++)
            if (p->nfr[k] && (pk_occ_of(k, false, p->shape[k]) != pk_occ_of(k, false, c->shape[k]) ||
                              pk_scale_of(p->shape[k]) != PLAN_VARIANTS[best].scale))
                return true;
        return false;
    }();
    return plan_build(p, changed ? -1 : best);
}

extern "C" int ldpc_gpu_plan_select(ldpc_gpu_plan* p, int variant)
{
    if (!p)
        return LDPC_GPU_EINVAL;
    if (variant < 0 || variant >= PLAN_NVARIANTS)
        return set_err(p->c, LDPC_GPU_EINVAL, "plan variant %d out of range", variant);
    HIPCHK(p->c, hipSetDevice(p->c->device));
    HIPCHK(p->c, hipStreamSynchronize(p->c->stream));
    plan_variant_shape(p, variant);
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
            concurrent[v] = plan_lanes(p, v);
        if (ms)
            ms[v] = p->variant_ms[v];
    }
    return PLAN_NVARIANTS;
}

extern "C" int ldpc_gpu_plan_shape(ldpc_gpu_plan* p, int* di, int* short_frames, int* occ, float* scale, int* grid,
                                   int* frames, int max)
{
    if (!p)
        return LDPC_GPU_EINVAL;
    int n = 0;
    for (int k = 0; k < NLAUNCH && n < max; k++) {
        if (!p->nfr[k])
            continue;
        if (di)
            di[n] = sk_di(k / 2);
        if (short_frames)
            short_frames[n] = k & 1;
        if (occ)
            occ[n] = p->occ[k];
        if (scale)
            scale[n] = pk_scale_of(p->shape[k]);
        if (grid)
            grid[n] = p->grid[k];
        if (frames)
            frames[n] = p->nfr[k];
        n++;
    }
    return n;
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
