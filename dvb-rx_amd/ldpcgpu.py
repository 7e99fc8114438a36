"""ctypes bindings of the MI355X LDPC decoder (lib/libldpcgpu.so, C-ABI in
include/ldpc_gpu.h) and of the frame simulator (lib/libdvbs2sim.so,
include/dvbs2_sim.h).

The decoder runs only through the HIP library: if it is missing, loading
raises; there is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBDIR = os.path.join(HERE, "lib")

FRAME_DTYPE = np.dtype(
    [
        ("code", "<i4"),
        ("max_iter", "<i4"),
        ("nbits", "<i4"),
        ("group", "<i4"),
        ("llr_off", "<u8"),
        ("out_off", "<u8"),
    ],
    align=True,
)
assert FRAME_DTYPE.itemsize == 32


class SimSpec(C.Structure):
    _fields_ = [
        ("short_frame", C.c_int),
        ("rate", C.c_char * 24),
        ("pls", C.c_int),
        ("esn0_db", C.c_double),
        ("llr_off", C.c_uint64),
        ("ref_off", C.c_uint64),
        ("stream", C.c_uint64),
    ]


def _load(name):
    path = os.path.join(LIBDIR, name)
    if not os.path.exists(path):
        raise OSError(f"{path} is not built: run `make -C dvb-rx_amd` (or __graft_entry__.build())")
    return C.CDLL(path)


_gpu = None
_sim = None


def gpu_lib():
    """The HIP decoder library, with argument types declared."""
    global _gpu
    if _gpu is None:
        L = _load("libldpcgpu.so")
        vp, i32, sz = C.c_void_p, C.c_int, C.c_size_t
        L.ldpc_gpu_create.argtypes = [i32, C.POINTER(vp)]
        L.ldpc_gpu_destroy.argtypes = [vp]
        L.ldpc_gpu_stream.argtypes = [vp]
        L.ldpc_gpu_stream.restype = vp
        L.ldpc_gpu_last_error.argtypes = [vp]
        L.ldpc_gpu_last_error.restype = C.c_char_p
        L.ldpc_gpu_code_count.argtypes = []
        L.ldpc_gpu_code_lookup.argtypes = [i32, C.c_char_p]
        ip = C.POINTER(C.c_int)
        L.ldpc_gpu_code_params.argtypes = [i32, ip, ip, ip, ip, ip]
        L.ldpc_gpu_fec_lookup.argtypes = [i32, C.c_char_p, ip, ip, ip, ip]
        L.ldpc_gpu_malloc.argtypes = [vp, sz]
        L.ldpc_gpu_malloc.restype = vp
        L.ldpc_gpu_free.argtypes = [vp, vp]
        L.ldpc_gpu_upload.argtypes = [vp, vp, vp, sz]
        L.ldpc_gpu_download.argtypes = [vp, vp, vp, sz]
        L.ldpc_gpu_synchronize.argtypes = [vp]
        L.ldpc_gpu_host_alloc.argtypes = [vp, sz]
        L.ldpc_gpu_host_alloc.restype = vp
        L.ldpc_gpu_host_free.argtypes = [vp, vp]
        L.ldpc_gpu_upload_async.argtypes = [vp, vp, vp, sz]
        L.ldpc_gpu_download_async.argtypes = [vp, vp, vp, sz]
        L.ldpc_gpu_query.argtypes = [vp]
        L.ldpc_gpu_set_concurrent.argtypes = [vp, C.c_int]
        L.ldpc_gpu_demap_symbols.argtypes = [C.c_int]
        L.ldpc_gpu_demap.argtypes = [vp, C.c_int, C.c_int, vp, vp, vp]
        L.ldpc_gpu_bch_decode.argtypes = [vp, C.c_int, C.c_char_p, C.c_int, vp, C.c_size_t, vp]
        L.ldpc_gpu_t2_geometry.argtypes = [vp, ip, ip, ip, ip]
        L.ldpc_gpu_t2_demap.argtypes = [vp, vp, i32, vp, vp, vp, vp, vp]
        L.ldpc_gpu_t2_maps.argtypes = [vp, i32, vp, vp, vp]
        L.ldpc_gpu_t2_decode_host.argtypes = [vp, vp, i32, vp, vp, i32, i32, vp, vp, vp, vp]
        L.ldpc_gpu_t2_decode.argtypes = [vp, vp, i32, vp, vp, i32, i32, vp, vp, vp, vp, vp, vp, vp]
        L.ldpc_gpu_decode_batch.argtypes = [vp, vp, i32, vp, vp, vp, vp]
        L.ldpc_gpu_decode_batch_static.argtypes = [vp, vp, i32, vp, vp, vp, vp]
        L.ldpc_gpu_decode_host.argtypes = [vp, vp, i32, vp, sz, vp, sz, vp, vp]
        L.ldpc_gpu_last_kernel_ms.argtypes = [vp]
        L.ldpc_gpu_last_kernel_ms.restype = C.c_float
        L.ldpc_gpu_last_launches.argtypes = [vp, ip, ip, C.POINTER(C.c_float), i32]
        L.ldpc_gpu_last_launch_shapes.argtypes = [vp, ip, ip, i32]
        L.ldpc_gpu_code_max_degree.argtypes = [i32]
        L.ldpc_gpu_code_kernel_degree.argtypes = [i32]
        L.ldpc_gpu_set_kernel.argtypes = [vp, i32]
        L.ldpc_gpu_group_fallbacks.argtypes = [vp]
        L.ldpc_gpu_set_nms.argtypes = [vp, i32, i32, vp, vp]
        L.ldpc_gpu_plan_create.argtypes = [vp, vp, i32, vp, vp, vp, vp, C.POINTER(vp)]
        L.ldpc_gpu_plan_launch.argtypes = [vp]
        L.ldpc_gpu_plan_tune.argtypes = [vp, i32]
        L.ldpc_gpu_plan_last_ms.argtypes = [vp]
        L.ldpc_gpu_plan_last_ms.restype = C.c_float
        L.ldpc_gpu_plan_info.argtypes = [vp, ip, vp, vp, vp, i32]
        L.ldpc_gpu_plan_select.argtypes = [vp, i32]
        L.ldpc_gpu_plan_shape.argtypes = [vp, vp, vp, vp, vp, vp, vp, i32]
        L.ldpc_gpu_set_launch_shape.argtypes = [vp, i32, i32, i32, C.c_float]
        L.ldpc_gpu_plan_destroy.argtypes = [vp]
        L.ldpc_gpu_decode_batch_nms.argtypes = [vp, vp, i32, vp, vp, vp, vp, i32, vp]
        L.drx_ldpc_lanes.argtypes = []
        L.drx_ldpc_new.argtypes = [vp, C.c_char_p]
        L.drx_ldpc_new.restype = vp
        L.drx_ldpc_free.argtypes = [vp]
        L.drx_ldpc_decode.argtypes = [vp, vp, vp, i32, i32]
        f32 = C.c_float
        L.ldpc_nms_trainer_create.argtypes = [i32, i32, i32, i32, C.POINTER(vp)]
        L.ldpc_nms_trainer_destroy.argtypes = [vp]
        L.ldpc_nms_trainer_stream.argtypes = [vp]
        L.ldpc_nms_trainer_stream.restype = vp
        L.ldpc_nms_trainer_error.argtypes = [vp]
        L.ldpc_nms_trainer_error.restype = C.c_char_p
        L.ldpc_nms_trainer_synchronize.argtypes = [vp]
        L.ldpc_nms_train_step.argtypes = [vp, vp, vp, vp, i32, vp, vp, f32, vp, vp, vp]
        L.ldpc_nms_adam.argtypes = [vp, vp, vp, vp, vp, i32, f32, f32, f32, f32, i32]
        L.ldpc_nms_quantize.argtypes = [vp, vp, i32, vp, vp]
        L.ldpc_gpu_channel.argtypes = [i32, i32, vp, vp, i32, vp, C.c_uint64, vp, vp]
        L.ldpc_gpu_channel_symbols.argtypes = [i32, i32, vp, vp, i32, vp, C.c_uint64, vp, vp, vp]
        _gpu = L
    return _gpu


def sim_lib():
    global _sim
    if _sim is None:
        L = _load("libdvbs2sim.so")
        L.sim_generate.argtypes = [C.c_void_p, C.c_int, C.c_uint64, C.c_void_p, C.c_void_p, C.c_int]
        ip = C.POINTER(C.c_int)
        L.sim_fec_lookup.argtypes = [C.c_int, C.c_char_p, ip, ip, ip, ip]
        L.sim_pls_lookup.argtypes = [C.c_int, C.c_char_p, C.c_int, ip, ip, C.c_char_p, C.c_int]
        L.sim_qef_db.argtypes = [C.c_int]
        L.sim_qef_db.restype = C.c_double
        L.sim_bch_encode.argtypes = [C.c_int, C.c_char_p, C.c_void_p, C.c_void_p]
        L.sim_ldpc_encode.argtypes = [C.c_int, C.c_char_p, C.c_void_p, C.c_void_p]
        _sim = L
    return _sim


# ---- codes -----------------------------------------------------------------


class FecParams:
    def __init__(self, short_frame, nldpc, kldpc, kbch, t):
        self.short_frame, self.nldpc, self.kldpc, self.kbch, self.t = short_frame, nldpc, kldpc, kbch, t

    def as_struct(self):
        """drx_fec_params layout (libdvbrx/include/dvbrx/fec.h:13-19)."""
        return (C.c_int * 5)(self.short_frame, self.nldpc, self.kldpc, self.kbch, self.t)


def fec_lookup(short_frame, rate):
    """drx_fec_lookup: BCH/LDPC sizes of a rate; the T2 codes ("2_3_t2",
    "3_5_t2") use the DVB-S2 parameters of their rate."""
    v = [C.c_int() for _ in range(4)]
    if gpu_lib().ldpc_gpu_fec_lookup(int(short_frame), rate.encode(), *[C.byref(x) for x in v]):
        raise KeyError(f"unknown rate {rate!r} (short={short_frame})")
    return FecParams(int(short_frame), *[x.value for x in v])


def code_lookup(short_frame, rate):
    c = gpu_lib().ldpc_gpu_code_lookup(int(short_frame), rate.encode())
    if c < 0:
        raise KeyError(f"no LDPC code for rate {rate!r} (short={short_frame})")
    return c


def code_params(code):
    v = [C.c_int() for _ in range(5)]
    if gpu_lib().ldpc_gpu_code_params(code, *[C.byref(x) for x in v]):
        raise KeyError(code)
    n, k, q, lt, steps = [x.value for x in v]
    dmax = gpu_lib().ldpc_gpu_code_max_degree(code)
    return dict(n=n, k=k, q=q, r=n - k, links_total=lt, level_steps=steps, max_degree=dmax,
                degree_class=gpu_lib().ldpc_gpu_code_kernel_degree(code))


def pls_lookup(pls):
    rate = C.create_string_buffer(24)
    const = C.create_string_buffer(24)
    sf, bits = C.c_int(), C.c_int()
    if sim_lib().sim_pls_lookup(pls, rate, 24, C.byref(sf), C.byref(bits), const, 24):
        raise KeyError(hex(pls))
    return dict(rate=rate.value.decode(), short_frame=sf.value, bits=bits.value, constellation=const.value.decode())


def qef_db(pls):
    return sim_lib().sim_qef_db(pls)


# ---- frame source ------------------------------------------------------------


def generate(specs, seed, nthreads=None, with_ref=True):
    """specs: list of dicts (short_frame, rate) or (pls), esn0_db, stream.
    Returns (llr int8 buffer, ref uint8 buffer, layout list of
    (code, n, kldpc, llr_off, ref_off))."""
    n = len(specs)
    arr = (SimSpec * n)()
    layout = []
    llr_off = ref_off = 0
    for i, s in enumerate(specs):
        if s.get("pls", -1) >= 0:
            m = pls_lookup(s["pls"])
            short, rate = m["short_frame"], m["rate"]
        else:
            short, rate = s["short_frame"], s["rate"]
        fec = fec_lookup(short, rate)
        code = code_lookup(short, rate)
        arr[i].short_frame = short
        arr[i].rate = rate.encode()
        arr[i].pls = s.get("pls", -1)
        arr[i].esn0_db = s["esn0_db"]
        arr[i].llr_off = llr_off
        arr[i].ref_off = ref_off
        arr[i].stream = s.get("stream", i)
        layout.append((code, fec.nldpc, fec.kldpc, llr_off, ref_off))
        llr_off += fec.nldpc
        ref_off += (fec.kldpc + 7) // 8
    llr = np.zeros(llr_off, np.int8)
    ref = np.zeros(ref_off, np.uint8) if with_ref else None
    rc = sim_lib().sim_generate(
        arr, n, C.c_uint64(seed), llr.ctypes.data, ref.ctypes.data if with_ref else None,
        nthreads or min(16, os.cpu_count() or 1),
    )
    if rc:
        raise RuntimeError("sim_generate failed")
    return llr, ref, layout


def frames_for(layout, max_iter, group=-1, out_offsets=None):
    """Frame table for a generated layout (output offsets = reference offsets)."""
    fr = np.zeros(len(layout), FRAME_DTYPE)
    for i, (code, n, kldpc, lo, ro) in enumerate(layout):
        fr[i] = (code, max_iter, kldpc, group if np.isscalar(group) else group[i], lo, ro if out_offsets is None else out_offsets[i])
    return fr


# ---- decoder ----------------------------------------------------------------


class Decoder:
    """A decoder context (device, stream, scratch) of the HIP library."""

    KERNELS = {"auto": 0, "single": 1}

    def __init__(self, device=0, kernel="auto"):
        self.L = gpu_lib()
        self.ctx = C.c_void_p()
        rc = self.L.ldpc_gpu_create(device, C.byref(self.ctx))
        if rc:
            raise RuntimeError(f"ldpc_gpu_create({device}) failed: {rc}")
        self._check(self.L.ldpc_gpu_set_kernel(self.ctx, self.KERNELS[kernel]), "set_kernel")

    def close(self):
        if self.ctx:
            self.L.ldpc_gpu_destroy(self.ctx)
            self.ctx = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def stream(self):
        """The context's hipStream_t (non-blocking: order other work on it)."""
        return self.L.ldpc_gpu_stream(self.ctx)

    def error(self):
        return self.L.ldpc_gpu_last_error(self.ctx).decode()

    def _check(self, rc, what):
        if rc:
            raise RuntimeError(f"{what} failed ({rc}): {self.error()}")

    # device memory
    def malloc(self, nbytes):
        p = self.L.ldpc_gpu_malloc(self.ctx, nbytes)
        if not p:
            raise MemoryError(nbytes)
        return p

    def free(self, p):
        self._check(self.L.ldpc_gpu_free(self.ctx, p), "free")

    def upload(self, dptr, arr):
        arr = np.ascontiguousarray(arr)
        self._check(self.L.ldpc_gpu_upload(self.ctx, dptr, arr.ctypes.data, arr.nbytes), "upload")

    def download(self, arr, dptr):
        self._check(self.L.ldpc_gpu_download(self.ctx, arr.ctypes.data, dptr, arr.nbytes), "download")

    def host_array(self, shape, dtype):
        """A numpy array in page-locked host memory (free with
        host_free(arr)); its transfers can be enqueued without waiting."""
        dtype = np.dtype(dtype)
        nbytes = max(1, int(np.prod(shape)) * dtype.itemsize)
        ptr = self.L.ldpc_gpu_host_alloc(self.ctx, nbytes)
        if not ptr:
            raise MemoryError(f"host_alloc({nbytes}): {self.error()}")
        buf = (C.c_uint8 * nbytes).from_address(ptr)
        return np.frombuffer(buf, np.uint8, int(np.prod(shape)) * dtype.itemsize).view(dtype).reshape(shape)

    def host_free(self, arr):
        self._check(self.L.ldpc_gpu_host_free(self.ctx, arr.ctypes.data), "host_free")

    def upload_async(self, dptr, src_ptr, nbytes):
        """Enqueue a copy of nbytes from host address src_ptr (pinned for
        an asynchronous transfer) to dptr on the context's stream."""
        self._check(self.L.ldpc_gpu_upload_async(self.ctx, dptr, src_ptr, nbytes), "upload_async")

    def download_async(self, dst_ptr, dptr, nbytes):
        self._check(self.L.ldpc_gpu_download_async(self.ctx, dst_ptr, dptr, nbytes), "download_async")

    def synchronize(self):
        self._check(self.L.ldpc_gpu_synchronize(self.ctx), "synchronize")

    def set_concurrent(self, on):
        """Launches of one decode on forked streams (True, the default: 3;
        or an int, 2 .. 4, that many) or one after the other on the
        context's stream (False / 0)."""
        self._check(self.L.ldpc_gpu_set_concurrent(self.ctx, int(on)), "set_concurrent")

    def set_launch_shape(self, di, short_frames, occ, scale=1.0):
        """Workgroups per CU (1, or 2 where built) and grid scale of the
        launch of kernel instance di over short or normal frames."""
        self._check(self.L.ldpc_gpu_set_launch_shape(self.ctx, int(di), int(short_frames), int(occ),
                                                     float(scale)), "set_launch_shape")

    def query(self):
        """True when everything enqueued on the context has finished
        (non-blocking)."""
        r = self.L.ldpc_gpu_query(self.ctx)
        self._check(min(r, 0), "query")
        return r == 1

    def decode_batch(self, frames, d_llr, d_out, d_iters, d_synd):
        frames = np.ascontiguousarray(frames, FRAME_DTYPE)
        self._check(
            self.L.ldpc_gpu_decode_batch(self.ctx, frames.ctypes.data, len(frames), d_llr, d_out, d_iters, d_synd),
            "decode_batch",
        )

    def set_nms(self, code, S, B):
        """Learned table of a code (T x NMS_DEGS scales S and offsets B);
        S = None removes it."""
        if S is None:
            self._check(self.L.ldpc_gpu_set_nms(self.ctx, code, 0, None, None), "set_nms")
            return
        S = np.ascontiguousarray(S, np.uint8)
        B = np.ascontiguousarray(B, np.int16)
        assert S.shape == B.shape and S.shape[1] == NMS_DEGS
        self._check(self.L.ldpc_gpu_set_nms(self.ctx, code, S.shape[0], S.ctypes.data, B.ctypes.data), "set_nms")

    def decode_batch_nms(self, frames, d_llr, d_out, d_iters, d_synd, fallback, d_fell=None):
        frames = np.ascontiguousarray(frames, FRAME_DTYPE)
        self._check(self.L.ldpc_gpu_decode_batch_nms(self.ctx, frames.ctypes.data, len(frames), d_llr, d_out, d_iters,
                                                     d_synd, fallback, d_fell), "decode_batch_nms")

    def group_fallbacks(self):
        """Calls whose groups needed the exact one-frame path after the
        two-pass attempt."""
        return self.L.ldpc_gpu_group_fallbacks(self.ctx)

    def last_kernel_ms(self):
        return self.L.ldpc_gpu_last_kernel_ms(self.ctx)

    def last_launches(self):
        """[(kernel instance code, frames, ms)] of the last decode's launches
        (instance code: 2 * slots per lane - 2 + 100 * wide)."""
        cap = 64  # more than the launches of one decode (instances x frame sizes)
        di = (C.c_int * cap)()
        nf = (C.c_int * cap)()
        ms = (C.c_float * cap)()
        n = self.L.ldpc_gpu_last_launches(self.ctx, di, nf, ms, cap)
        return [(di[k], nf[k], ms[k]) for k in range(n)]

    def last_launch_shapes(self):
        """[(workgroups per CU, workgroups launched)] of the same launches."""
        cap = 64
        occ = (C.c_int * cap)()
        grid = (C.c_int * cap)()
        n = self.L.ldpc_gpu_last_launch_shapes(self.ctx, occ, grid, cap)
        return [(occ[k], grid[k]) for k in range(n)]

    def decode_host(self, frames, llr, out_bytes):
        """Synchronous decode of host buffers: returns (out, iters, synd)."""
        frames = np.ascontiguousarray(frames, FRAME_DTYPE)
        llr = np.ascontiguousarray(llr, np.int8)
        n = len(frames)
        out = np.zeros(out_bytes, np.uint8)
        iters = np.zeros(n, np.int32)
        synd = np.zeros(n, np.uint8)
        self._check(
            self.L.ldpc_gpu_decode_host(
                self.ctx, frames.ctypes.data, n, llr.ctypes.data, llr.nbytes, out.ctypes.data, out.nbytes,
                iters.ctypes.data, synd.ctypes.data,
            ),
            "decode_host",
        )
        return out, iters, synd


class Plan:
    """A prepared batch replayed as one graph (ldpc_gpu_plan_*)."""

    def __init__(self, dec, frames, d_llr, d_out, d_iters, d_synd):
        self.dec = dec
        self.L = dec.L
        self.frames = np.ascontiguousarray(frames, FRAME_DTYPE)
        self.p = C.c_void_p()
        dec._check(self.L.ldpc_gpu_plan_create(dec.ctx, self.frames.ctypes.data, len(self.frames), d_llr, d_out,
                                               d_iters, d_synd, C.byref(self.p)), "plan_create")

    def close(self):
        if self.p:
            self.L.ldpc_gpu_plan_destroy(self.p)
            self.p = C.c_void_p()

    def launch(self):
        self.dec._check(self.L.ldpc_gpu_plan_launch(self.p), "plan_launch")

    def tune(self, reps=3):
        self.dec._check(self.L.ldpc_gpu_plan_tune(self.p, reps), "plan_tune")

    def select(self, variant):
        self.dec._check(self.L.ldpc_gpu_plan_select(self.p, variant), "plan_select")

    def last_ms(self):
        return self.L.ldpc_gpu_plan_last_ms(self.p)

    def info(self):
        v = C.c_int()
        sc = np.zeros(8, np.float32)
        cc = np.zeros(8, np.int32)
        ms = np.zeros(8, np.float32)
        n = self.L.ldpc_gpu_plan_info(self.p, C.byref(v), sc.ctypes.data, cc.ctypes.data, ms.ctypes.data, 8)
        return {"variant": v.value, "variants": [{"grid_scale": float(sc[k]), "streams": int(cc[k]),
                                                  "ms": round(float(ms[k]), 4)} for k in range(n)],
                "launches": self.shape()}

    def shape(self):
        """Per launch: kernel instance, frame size, workgroups per CU, grid
        scale, workgroups launched, frames and frames in flight per CU."""
        cap = 64
        a = {k: np.zeros(cap, np.int32) for k in ("di", "short", "occ", "grid", "frames")}
        sc = np.zeros(cap, np.float32)
        n = self.L.ldpc_gpu_plan_shape(self.p, a["di"].ctypes.data, a["short"].ctypes.data, a["occ"].ctypes.data,
                                       sc.ctypes.data, a["grid"].ctypes.data, a["frames"].ctypes.data, cap)
        self.dec._check(min(n, 0), "plan_shape")
        ncu = 256
        return [{"di": int(a["di"][k]), "short": bool(a["short"][k]), "occ": int(a["occ"][k]),
                 "grid_scale": round(float(sc[k]), 3), "grid": int(a["grid"][k]), "frames": int(a["frames"][k]),
                 "frames_per_cu": round(2 * min(int(a["grid"][k]), (int(a["frames"][k]) + 1) // 2) / ncu, 2)}
                for k in range(n)]


NMS_DEGS = 33


def nms_quantize(alpha, beta):
    """Integer table (S, B) of float learned parameters (T x NMS_DEGS)."""
    alpha = np.ascontiguousarray(alpha, np.float32)
    beta = np.ascontiguousarray(beta, np.float32)
    assert alpha.shape == beta.shape and alpha.shape[1] == NMS_DEGS
    S = np.zeros(alpha.shape, np.uint8)
    B = np.zeros(alpha.shape, np.int16)
    gpu_lib().ldpc_nms_quantize(alpha.ctypes.data, beta.ctypes.data, alpha.shape[0], S.ctypes.data, B.ctypes.data)
    return S, B


class NmsTrainer:
    """Learned offset min-sum trainer for one code (include/ldpc_nms.h)."""

    def __init__(self, code, T, max_frames, device=0):
        self.L = gpu_lib()
        self.t = C.c_void_p()
        rc = self.L.ldpc_nms_trainer_create(device, code, T, max_frames, C.byref(self.t))
        if rc:
            raise RuntimeError(f"ldpc_nms_trainer_create: error {rc}")
        self.code, self.T, self.max_frames = code, T, max_frames

    def close(self):
        if self.t:
            self.L.ldpc_nms_trainer_destroy(self.t)
            self.t = C.c_void_p()

    def _check(self, rc, what):
        if rc:
            raise RuntimeError(f"{what}: {self.L.ldpc_nms_trainer_error(self.t).decode()}")

    @property
    def stream(self):
        return self.L.ldpc_nms_trainer_stream(self.t)

    def synchronize(self):
        self._check(self.L.ldpc_nms_trainer_synchronize(self.t), "synchronize")

    def step(self, d_llr, d_cw, d_row, nframes, d_params, d_w, temp, d_loss, d_grad, d_post=None):
        self._check(self.L.ldpc_nms_train_step(self.t, d_llr, d_cw, d_row, nframes, d_params, d_w, temp, d_loss,
                                               d_grad, d_post), "train_step")

    def adam(self, d_param, d_grad, d_m, d_v, n, lr, step, beta1=0.9, beta2=0.999, eps=1e-8):
        self._check(self.L.ldpc_nms_adam(self.t, d_param, d_grad, d_m, d_v, n, lr, beta1, beta2, eps, step), "adam")


def channel(pls, n, d_cw, d_row, nframes, d_esn0_db, seed, d_llr, stream=None):
    """GPU channel: codeword rows -> MODCOD -> AWGN -> int8 max-log LLRs."""
    rc = gpu_lib().ldpc_gpu_channel(pls, n, d_cw, d_row, nframes, d_esn0_db, seed, d_llr, stream)
    if rc:
        raise RuntimeError(f"ldpc_gpu_channel: error {rc}")


def demap_symbols(pls):
    """Symbols per frame of a MODCOD (ldpc_gpu_demap_symbols), -1 unknown."""
    return gpu_lib().ldpc_gpu_demap_symbols(pls)


def demap(dec, pls, nframes, d_sym, d_n0, d_llr):
    """ldpc_gpu_demap on the decoder context's stream (enqueue only)."""
    dec._check(gpu_lib().ldpc_gpu_demap(dec.ctx, pls, nframes, d_sym, d_n0, d_llr), "demap")


def bch_decode(dec, short_frame, rate, nframes, d_cw, stride, d_result):
    """ldpc_gpu_bch_decode on the decoder context's stream (enqueue only)."""
    dec._check(gpu_lib().ldpc_gpu_bch_decode(dec.ctx, short_frame, rate.encode(), nframes, d_cw, stride, d_result),
               "bch_decode")


class T2Plp(C.Structure):
    """ldpc_gpu_t2_plp: a DVB-T2 PLP's FEC parameters as L1-post signals
    them (cod 0..5, mod 0..3, fec_type 1 normal / 0 short, rotation, time
    interleaver length and type)."""
    _fields_ = [(n, C.c_int32) for n in ("cod", "mod", "fec_type", "rotation", "til_length", "til_type")]


def t2_geometry(plp):
    """(nldpc, kldpc, kbch, ncells) of a PLP configuration; ValueError if
    unsupported."""
    v = [C.c_int() for _ in range(4)]
    if gpu_lib().ldpc_gpu_t2_geometry(C.byref(plp), *[C.byref(x) for x in v]):
        raise ValueError(f"unsupported DVB-T2 PLP {[getattr(plp, f) for f, _ in plp._fields_]}")
    return tuple(x.value for x in v)


def t2_maps(plp, nblocks=0):
    """(perm, bitsrc, shift) of ldpc_gpu_t2_maps as numpy arrays (host
    only: no device call)."""
    N, K, kbch, nc = t2_geometry(plp)
    perm = np.zeros(nc, np.int32)
    src = np.zeros(N, np.uint16)
    shift = np.zeros(max(nblocks, 1), np.int32)
    if gpu_lib().ldpc_gpu_t2_maps(C.byref(plp), nblocks, perm.ctypes.data, src.ctypes.data, shift.ctypes.data):
        raise ValueError("ldpc_gpu_t2_maps failed")
    return perm, src, shift[:nblocks]


def t2_demap(dec, plp, nblocks, d_cells, d_csi, d_llr, d_err, d_sig):
    """ldpc_gpu_t2_demap on the decoder context's stream (enqueue only)."""
    dec._check(gpu_lib().ldpc_gpu_t2_demap(dec.ctx, C.byref(plp), nblocks, d_cells, d_csi, d_llr, d_err, d_sig),
               "t2_demap")


def t2_decode(dec, plp, nblocks, d_cells, d_csi, max_iter, lanes, d_llr, d_cw, d_iters, d_synd, d_bch, d_err, d_sig):
    """ldpc_gpu_t2_decode: front end, LDPC and BCH of nblocks FEC blocks."""
    dec._check(gpu_lib().ldpc_gpu_t2_decode(dec.ctx, C.byref(plp), nblocks, d_cells, d_csi, max_iter, lanes, d_llr,
                                            d_cw, d_iters, d_synd, d_bch, d_err, d_sig), "t2_decode")


def t2_decode_host(dec, plp, cells, csi, max_iter=30, lanes=0):
    """ldpc_gpu_t2_decode_host on host arrays: cells (nblocks, ncells, 2)
    float32, csi (nblocks, ncells) float32 -> (BBFRAMEs (nblocks, kbch/8),
    ok, err, sig)."""
    N, K, kbch, nc = t2_geometry(plp)
    cells = np.ascontiguousarray(cells, np.float32)
    csi = np.ascontiguousarray(csi, np.float32)
    nb = cells.shape[0]
    if cells.shape != (nb, nc, 2) or csi.shape != (nb, nc):
        raise ValueError("cells must be (nblocks, ncells, 2) and csi (nblocks, ncells)")
    bb = np.zeros((nb, kbch // 8), np.uint8)
    ok = np.zeros(nb, np.uint8)
    err = np.zeros(nb, np.float32)
    sig = np.zeros(nb, np.float32)
    dec._check(gpu_lib().ldpc_gpu_t2_decode_host(dec.ctx, C.byref(plp), nb, cells.ctypes.data, csi.ctypes.data,
                                                 max_iter, lanes, bb.ctypes.data, ok.ctypes.data, err.ctypes.data,
                                                 sig.ctypes.data), "t2_decode_host")
    return bb, ok, err, sig


class DrxLdpc:
    """The drop-in drx_ldpc_* API of libldpcgpu.so (include/dvbrx/ldpc.h)."""

    def __init__(self, fec, rate):
        self.L = gpu_lib()
        self.fec = fec
        self._p = fec.as_struct()
        self.h = self.L.drx_ldpc_new(C.byref(self._p), rate.encode())
        if not self.h:
            raise RuntimeError(f"drx_ldpc_new({rate}) failed")

    def close(self):
        if self.h:
            self.L.drx_ldpc_free(self.h)
            self.h = None

    def decode(self, llr, max_iter):
        """llr: (nframes, nldpc) int8.  Returns (iterations or -1, bits)."""
        llr = np.ascontiguousarray(llr, np.int8)
        n = llr.shape[0]
        nb = (self.fec.kldpc + 7) // 8
        out = np.zeros((n, nb), np.uint8)
        ins = (C.c_void_p * max(n, 1))(*[llr[f].ctypes.data for f in range(n)])
        outs = (C.c_void_p * max(n, 1))(*[out[f].ctypes.data for f in range(n)])
        it = self.L.drx_ldpc_decode(self.h, ins, outs, n, max_iter)
        return it, out
